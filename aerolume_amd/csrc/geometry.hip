// Geometry set-up on the device: the exact signed distance of grid points to a closed polygon,
// the input of the 2D body masks (ns2d/geometry.py:229-345 builds them from the same distance).
//
// Per point, in float64: the nearest segment (first minimum of the squared distance over the
// segments, in order) and the even-odd crossing count of a ray towards +x; then, in the run's
// dtype, the distance to that segment from the polygon's vertices rounded to the dtype (its
// square root taken in float64 and rounded once), with the sign of the crossing test
// (negative inside). Every operation is the single IEEE operation
// the host formula names (floating-point contraction is off in this file), so the result does
// not depend on where it is computed.
//
// One thread per point; the polygon is staged through LDS in chunks of kChunk vertices.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>

#include "../../include/aerolume_amd.h"
#include "common.h"

#pragma clang fp contract(off)

namespace aero {
namespace geo {

constexpr int kThreads = 256;
constexpr int kChunk = 1024;   // vertices per LDS chunk (16 KB of float64 pairs)

// squared distance from (px, py) to the segment a + t e, t in [0, 1] (host formula
// _seg_d2: the projection parameter clamped, the squared length floored at the smallest
// normal number of the type)
template <typename T>
__device__ __forceinline__ T seg_d2(T px, T py, T ax, T ay, T ex, T ey, T tiny) {
  const T wx = px - ax;
  const T wy = py - ay;
  T ee = ex * ex + ey * ey;
  ee = ee > tiny ? ee : tiny;
  T t = (wx * ex + wy * ey) / ee;
  t = t < (T)0 ? (T)0 : (t > (T)1 ? (T)1 : t);
  const T dx = wx - t * ex;
  const T dy = wy - t * ey;
  return dx * dx + dy * dy;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_polygon_sdf(const double *__restrict__ poly, int n,
                                                          const T *__restrict__ xs, const T *__restrict__ ys,
                                                          long long npts, T *__restrict__ out) {
  __shared__ double V[kChunk + 1][2];
  const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
  const bool live = k < npts;
  const T xt = live ? xs[k] : (T)0, yt = live ? ys[k] : (T)0;
  const double px = (double)xt, py = (double)yt;
  double best = INFINITY;
  int arg = 0;
  bool inside = false;
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    const int m = n - c0 < kChunk ? n - c0 : kChunk;
    __syncthreads();
    // vertices c0 .. c0 + m (the last one closes the chunk's final segment, wrapping to 0)
    for (int q = threadIdx.x; q <= m; q += kThreads) {
      const int src = (c0 + q) % n;
      V[q][0] = poly[2 * src];
      V[q][1] = poly[2 * src + 1];
    }
    __syncthreads();
    if (!live) continue;
    for (int q = 0; q < m; ++q) {
      const double ax = V[q][0], ay = V[q][1], bx = V[q + 1][0], by = V[q + 1][1];
      const double ex = bx - ax, ey = by - ay;
      const double d2 = seg_d2<double>(px, py, ax, ay, ex, ey, DBL_MIN);
      if (d2 < best) {
        best = d2;
        arg = c0 + q;
      }
      // ray towards +x crosses the segment when its end points straddle py
      if ((ay > py) != (by > py)) {
        const double den = by != ay ? by - ay : 1e-30;
        const double xc = ax + (py - ay) * ex / den;
        if (px < xc) inside = !inside;
      }
    }
  }
  if (!live) return;
  const int b = arg + 1 == n ? 0 : arg + 1;
  const T ax = (T)poly[2 * arg], ay = (T)poly[2 * arg + 1];
  const T bx = (T)poly[2 * b], by = (T)poly[2 * b + 1];
  const T tiny = sizeof(T) == 8 ? (T)DBL_MIN : (T)FLT_MIN;
  T d2 = seg_d2<T>(xt, yt, ax, ay, bx - ax, by - ay, tiny);
  d2 = d2 > tiny ? d2 : tiny;
  const T d = (T)sqrt((double)d2);   // correctly rounded to T (the float64 root rounded once)
  out[k] = inside ? -d : d;
}

}  // namespace geo
}  // namespace aero

extern "C" int aero_polygon_sdf(void *stream, int32_t dtype, const double *poly, int32_t n, const void *x,
                                const void *y, int64_t npts, void *out) {
  using namespace aero;
  if (!poly || !x || !y || !out || n < 3 || npts < 0 || (dtype != 0 && dtype != 1)) {
    set_error("aero_polygon_sdf: needs a polygon of at least 3 vertices, point and output arrays, dtype 0 or 1");
    return AERO_EINVAL;
  }
  if (npts == 0) return AERO_OK;
  const long long blocks = (npts + geo::kThreads - 1) / geo::kThreads;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 1)
    hipLaunchKernelGGL(geo::k_polygon_sdf<double>, dim3((unsigned)blocks), dim3(geo::kThreads), 0, st, poly, n,
                       (const double *)x, (const double *)y, (long long)npts, (double *)out);
  else
    hipLaunchKernelGGL(geo::k_polygon_sdf<float>, dim3((unsigned)blocks), dim3(geo::kThreads), 0, st, poly, n,
                       (const float *)x, (const float *)y, (long long)npts, (float *)out);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}
