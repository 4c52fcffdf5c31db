// Tracer particles of the 2D solver (ns2d/particles.py:122-173): a fixed pool advected with the
// end-of-step velocity, killed when they leave the domain or enter a solid, and fed by
// particle emitters (continuous release clocks and pulsed batches) into per-nozzle rings.
//
//   k_advect   one thread per pool slot: midpoint (RK2) step with bilinear samples of u, v where
//              they live on the staggered faces; kill outside the box or where solid > 1/2
//   k_release  one block: the emitters' clocks (accumulated rate x dt, pulse windows), then per
//              nozzle up to kMaxRelease releases into its ring, placed along the local velocity
//              by the time since the threshold crossing
//
// The two launches are ordered on the caller's stream and allocate nothing, so a captured step
// carries them. Floating-point contraction is off in this file: every product and sum is the
// single rounding the host formula names.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/aerolume_amd.h"
#include "common.h"

#pragma clang fp contract(off)

namespace aero {
namespace ptc {

constexpr int kMaxRelease = 4;    // releases per continuous nozzle and step (MAX_RELEASE)
constexpr int kAdvectThreads = 256;
constexpr int kReleaseThreads = 1024;
constexpr int kMaxEmitters = 256;

template <typename T>
struct Grid {
  int nx, ny;
  T x0, y0, h;
};

// bilinear sample of f (ny, nx) whose value [j, i] sits at (x0 + (i + xo) h, y0 + (j + yo) h),
// the point clamped to the grid (ns2d/particles.py:45-56)
template <typename T>
__device__ __forceinline__ T bilinear(const T *__restrict__ f, int ny, int nx, const Grid<T> &g, T x, T y, T xo, T yo) {
  T fx = (x - g.x0) / g.h - xo;
  T fy = (y - g.y0) / g.h - yo;
  fx = fx < (T)0 ? (T)0 : (fx > (T)(nx - 1) ? (T)(nx - 1) : fx);
  fy = fy < (T)0 ? (T)0 : (fy > (T)(ny - 1) ? (T)(ny - 1) : fy);
  int i0 = (int)floor(fx), j0 = (int)floor(fy);
  i0 = i0 < nx - 2 ? i0 : nx - 2;
  j0 = j0 < ny - 2 ? j0 : ny - 2;
  const T tx = fx - (T)i0, ty = fy - (T)j0;
  const int k = j0 * nx + i0;
  const T f00 = f[k], f01 = f[k + 1], f10 = f[k + nx], f11 = f[k + nx + 1];
  return ((T)1 - ty) * (((T)1 - tx) * f00 + tx * f01) + ty * (((T)1 - tx) * f10 + tx * f11);
}

template <typename T>
__device__ __forceinline__ void velocity(const T *u, const T *v, const Grid<T> &g, T x, T y, T &uo, T &vo) {
  uo = bilinear(u, g.ny, g.nx + 1, g, x, y, (T)0, (T)0.5);
  vo = bilinear(v, g.ny + 1, g.nx, g, x, y, (T)0.5, (T)0);
}

template <typename T>
__global__ void __launch_bounds__(kAdvectThreads) k_advect(Grid<T> g, const T *__restrict__ u, const T *__restrict__ v,
                                                          const T *__restrict__ dtp, const T *__restrict__ solid,
                                                          T *__restrict__ pos, T *__restrict__ alive, int cap) {
  const int k = blockIdx.x * kAdvectThreads + threadIdx.x;
  if (k >= cap) return;
  const T dt = *dtp;
  const T x = pos[2 * k], y = pos[2 * k + 1];
  T u1, v1, u2, v2;
  velocity(u, v, g, x, y, u1, v1);
  const T half = (T)0.5 * dt;
  const T xm = x + half * u1, ym = y + half * v1;
  velocity(u, v, g, xm, ym, u2, v2);
  const T xn = x + dt * u2, yn = y + dt * v2;
  const T x1 = (T)((double)g.x0 + (double)g.nx * (double)g.h), y1 = (T)((double)g.y0 + (double)g.ny * (double)g.h);
  bool inside = xn > g.x0 && xn < x1 && yn > g.y0 && yn < y1;
  if (solid) {
    T ci = (xn - g.x0) / g.h, cj = (yn - g.y0) / g.h;
    ci = ci < (T)0 ? (T)0 : (ci > (T)(g.nx - 1) ? (T)(g.nx - 1) : ci);
    cj = cj < (T)0 ? (T)0 : (cj > (T)(g.ny - 1) ? (T)(g.ny - 1) : cj);
    inside = inside && solid[(int)cj * g.nx + (int)ci] <= (T)0.5;
  }
  const T a = alive[k] * (inside ? (T)1 : (T)0);
  if (a > (T)0) {
    pos[2 * k] = xn;
    pos[2 * k + 1] = yn;
  }
  alive[k] = a;
}

template <typename T>
__global__ void __launch_bounds__(kReleaseThreads) k_release(Grid<T> g, const T *__restrict__ u, const T *__restrict__ v,
                                                            const T *__restrict__ dtp, const T *__restrict__ tp,
                                                            aero2d_particles P) {
  __shared__ T s_acc[kMaxEmitters];     // accumulated release count before its fraction is kept
  __shared__ T s_ncont[kMaxEmitters];   // continuous releases this step
  __shared__ int s_relp[kMaxEmitters];  // pulsed batch this step
  const T dt = *dtp, t = *tp;
  T *acc = (T *)P.acc, *last = (T *)P.last;
  const T *rate = (const T *)P.rate, *timing = (const T *)P.timing, *pulsed = (const T *)P.pulsed;
  for (int e = threadIdx.x; e < P.n_emitters; e += blockDim.x) {
    const T start = timing[4 * e], stop = timing[4 * e + 1], period = timing[4 * e + 2], duty = timing[4 * e + 3];
    const bool active = t >= start && t < stop;
    const T safe = period > (T)0 ? period : (T)1;
    const T kk = floor((t - start) / safe);
    const T phase = (t - start) / safe - kk;
    const T on_cont = (active ? (T)1 : (T)0) * ((T)1 - pulsed[e]);
    const T a = acc[e] + dt * rate[e] * on_cont;
    const T fl = floor(a);
    s_acc[e] = a;
    s_ncont[e] = fl < (T)kMaxRelease ? fl : (T)kMaxRelease;
    const bool relp = pulsed[e] > (T)0 && active && kk > last[e] && phase < duty;
    s_relp[e] = relp ? 1 : 0;
    acc[e] = a - fl;
    if (relp) last[e] = kk;
  }
  __syncthreads();
  T *pos = (T *)P.pos, *birth = (T *)P.birth, *alive = (T *)P.alive;
  long long *head = (long long *)P.head;
  const T *noz = (const T *)P.noz_pos;
  const long long *owner = (const long long *)P.noz_owner;
  for (int n = threadIdx.x; n < P.n_nozzles; n += blockDim.x) {
    const int e = (int)owner[n];
    const bool is_pulsed = pulsed[e] > (T)0;
    const T nxp = noz[2 * n], nyp = noz[2 * n + 1];
    T un, vn;
    velocity(u, v, g, nxp, nyp, un, vn);
    const T r_e = rate[e] > (T)1e-30 ? rate[e] : (T)1e-30;
    long long hd = head[n];
    for (int r = 0; r < kMaxRelease; ++r) {
      const bool rel = r == 0 ? (is_pulsed ? s_relp[e] != 0 : s_ncont[e] > (T)0)
                              : (!is_pulsed && s_ncont[e] > (T)r);
      T age = (T)0;
      if (!is_pulsed) {
        age = (s_acc[e] - (T)(r + 1.0)) / r_e;
        age = age > (T)0 ? age : (T)0;
      }
      age = age < dt ? age : dt;
      if (rel) {
        const long long slot = (long long)n * P.ring + hd;
        pos[2 * slot] = nxp + un * age;
        pos[2 * slot + 1] = nyp + vn * age;
        birth[slot] = t - age;
        alive[slot] = (T)1;
        hd = hd + 1 == P.ring ? 0 : hd + 1;
      }
    }
    head[n] = hd;
  }
}

}  // namespace ptc
}  // namespace aero

extern "C" int aero2d_particles_step(void *stream, const aero2d_particles_grid *grid, const void *u, const void *v,
                                     const void *dt, const void *t, const void *solid, aero2d_particles *p) {
  using namespace aero;
  if (!grid || !p || !u || !v || !dt || !t || !p->pos || !p->birth || !p->alive || !p->head || !p->acc ||
      !p->last || !p->noz_pos || !p->noz_owner || !p->rate || !p->timing || !p->pulsed) {
    set_error("aero2d_particles_step: missing pointer");
    return AERO_EINVAL;
  }
  if (grid->dtype != 0 && grid->dtype != 1) { set_error("aero2d_particles_step: dtype must be 0 or 1"); return AERO_EINVAL; }
  if (grid->nx < 2 || grid->ny < 2 || p->n_nozzles < 1 || p->ring < 1 || p->n_emitters < 1 ||
      p->n_emitters > ptc::kMaxEmitters) {
    set_error("aero2d_particles_step: needs nx, ny >= 2, 1..256 emitters, nozzles and a ring");
    return AERO_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int cap = p->n_nozzles * p->ring;
  const int blocks = (cap + ptc::kAdvectThreads - 1) / ptc::kAdvectThreads;
  if (grid->dtype == 1) {
    ptc::Grid<double> g{grid->nx, grid->ny, grid->x0, grid->y0, grid->h};
    hipLaunchKernelGGL(ptc::k_advect<double>, dim3(blocks), dim3(ptc::kAdvectThreads), 0, st, g, (const double *)u,
                       (const double *)v, (const double *)dt, (const double *)solid, (double *)p->pos,
                       (double *)p->alive, cap);
    hipLaunchKernelGGL(ptc::k_release<double>, dim3(1), dim3(ptc::kReleaseThreads), 0, st, g, (const double *)u,
                       (const double *)v, (const double *)dt, (const double *)t, *p);
  } else {
    ptc::Grid<float> g{grid->nx, grid->ny, (float)grid->x0, (float)grid->y0, (float)grid->h};
    hipLaunchKernelGGL(ptc::k_advect<float>, dim3(blocks), dim3(ptc::kAdvectThreads), 0, st, g, (const float *)u,
                       (const float *)v, (const float *)dt, (const float *)solid, (float *)p->pos, (float *)p->alive,
                       cap);
    hipLaunchKernelGGL(ptc::k_release<float>, dim3(1), dim3(ptc::kReleaseThreads), 0, st, g, (const float *)u,
                       (const float *)v, (const float *)dt, (const float *)t, *p);
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}
