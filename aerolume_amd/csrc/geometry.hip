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

#include <algorithm>
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

// Solid fractions of analytic 3D bodies on one staggered grid of the mask window
// (ns3d/geometry.py:385-495 builds the masks from the signed distance the same way): per sample
// point (x0 + i h, y0 + j h, z0 + k h) and body, smooth_step(sdf + shift, width) =
// 1/2 - 1/2 sin(pi/2 clamp((sdf + shift) / width, -1, 1)), all in float64 with the host
// formulas' operation order (contraction off): a rounded box (centre, half extents less the
// edge radius, radius) or a cylinder along y (centre, radius, half width).
constexpr int kMaxShapes = 16;
struct Shapes3 {
  int n;
  int kind[kMaxShapes];
  double p[kMaxShapes][7];
};

__device__ __forceinline__ double shape_sdf(int kind, const double *p, double x, double y, double z) {
  if (kind == AERO3D_SHAPE_ROUNDED_BOX) {
    const double qx = fabs(x - p[0]) - p[3], qy = fabs(y - p[1]) - p[4], qz = fabs(z - p[2]) - p[5];
    const double ox = qx > 0.0 ? qx : 0.0, oy = qy > 0.0 ? qy : 0.0, oz = qz > 0.0 ? qz : 0.0;
    const double out = sqrt(ox * ox + oy * oy + oz * oz);
    double m = qx > qy ? qx : qy;
    m = m > qz ? m : qz;
    const double inside = m < 0.0 ? m : 0.0;
    return out + inside - p[6];
  }
  const double dx = x - p[0], dz = z - p[2];
  const double dr = sqrt(dx * dx + dz * dz) - p[3];
  const double da = fabs(y - p[1]) - p[4];
  const double orr = dr > 0.0 ? dr : 0.0, oa = da > 0.0 ? da : 0.0;
  const double out = sqrt(orr * orr + oa * oa);
  const double m = dr > da ? dr : da;
  return out + (m < 0.0 ? m : 0.0);
}

__global__ void __launch_bounds__(kThreads) k_analytic_fraction(Shapes3 sh, double x0, double y0, double z0, double h,
                                                                int nx, int ny, int nz, double shift, double width,
                                                                double *__restrict__ out) {
  const long long per = (long long)nx * ny * nz, total = per * sh.n;
  const double half_pi = 0.5 * M_PI;
  for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < total;
       e += (long long)gridDim.x * kThreads) {
    const int b = (int)(e / per);
    const long long r = e - (long long)b * per;
    const int k = (int)(r / ((long long)nx * ny)), j = (int)((r / nx) % ny), i = (int)(r % nx);
    const double x = x0 + (double)i * h, y = y0 + (double)j * h, z = z0 + (double)k * h;
    double s = (shape_sdf(sh.kind[b], sh.p[b], x, y, z) + shift) / width;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
    out[e] = 0.5 - 0.5 * sin(half_pi * s);
  }
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

extern "C" int aero3d_analytic_fractions(void *stream, const aero3d_shape *shapes, int32_t nshapes,
                                         const double *origin, double h, int32_t nx, int32_t ny, int32_t nz,
                                         double shift, double width, void *out) {
  using namespace aero;
  if (!shapes || !origin || !out || nshapes < 1 || nshapes > geo::kMaxShapes || nx < 1 || ny < 1 || nz < 1 ||
      !(h > 0) || !(width > 0)) {
    set_error("aero3d_analytic_fractions: 1 .. 16 shapes, a positive spacing and width, a non-empty grid");
    return AERO_EINVAL;
  }
  geo::Shapes3 sh;
  sh.n = nshapes;
  for (int b = 0; b < nshapes; ++b) {
    const int k = shapes[b].kind;
    if (k != AERO3D_SHAPE_ROUNDED_BOX && k != AERO3D_SHAPE_CYLINDER_Y) {
      set_error("aero3d_analytic_fractions: unknown shape kind");
      return AERO_EINVAL;
    }
    sh.kind[b] = k;
    for (int q = 0; q < 7; ++q) sh.p[b][q] = shapes[b].p[q];
  }
  const long long total = (long long)nx * ny * nz * nshapes;
  const long long blocks = std::min<long long>((total + geo::kThreads - 1) / geo::kThreads, 65536);
  hipLaunchKernelGGL(geo::k_analytic_fraction, dim3((unsigned)blocks), dim3(geo::kThreads), 0, (hipStream_t)stream, sh,
                     origin[0], origin[1], origin[2], h, (int)nx, (int)ny, (int)nz, shift, width, (double *)out);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}
