// MI355X (gfx950) kernels of the 2D incompressible Navier-Stokes step and its C ABI.
//
// One SSP-RK3 step (reference solvers/ns2d/core.py:815-969) is a fixed sequence of
// launches on the caller's stream, with no host synchronisation and no allocation:
//
//   [nut0]  reduce0            CFL / thermal dt limits from the state (core.py:685-695,
//                              physics.py:851-888); eddy viscosity max when LES is on
//   3 x  [nut] stage           upwind3/central advection + diffusion + LES stress +
//                              RK update + inlet/outlet + penalisation + buoyancy +
//                              temperature stage (core.py:571-614, 855-935)
//        poisA                 divergence (outlet mass balance folded in) -> DCT/DHT along x
//        poisB                 DCT/DHT along y, x 1/lambda (FP64-built), inverse along y
//        poisC                 inverse along x -> psi, step-averaged pressure
//        correct               u -= grad psi (core.py:645-660)
//   dye                        smoke transport with the end-of-step velocity
//   finalize                   forces, heat rates, t, dt, history rings (one block)
//
// All arithmetic is in the plan's dtype T (float or double); the Poisson eigenvalue
// inverse is formed in double, and the heat-balance sums accumulate in double.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolume_amd.h"
#include "common.h"

namespace aero {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

namespace ns2d {

constexpr int kRedBlocks = 1024;      // blocks of the state reduction
constexpr int kFftThreads = 256;
constexpr double kSsp[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
constexpr double kForceW[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0};

// Reduction slots shared by the kernels of one plan (device memory, zeroed at creation;
// every step leaves the max slots and tickets at zero again).
template <typename T>
struct Red {
  T umax, vmax, thmax, nutmax;   // max |u0|, |v0|, |theta0|, nu_t (order-free atomics)
  T dt;                          // this step's dt
  T corr;                        // outlet correction of the current stage
  unsigned int ticket;           // last-block detection of reduce0
  unsigned int pad;
  double th_old;                 // sum of theta at step start
};

// Everything a kernel needs, passed by value.
template <typename T>
struct Dev {
  int nx, ny;
  int periodic, closed, upwind, turb;
  int bcb, bct, bcl, bcr;
  int j0, j1, i0, i1, wx, wy, win;
  int nbody, thermal, buoyant, n_temp, has_heat, insulated;
  int wm[4];
  int dye_mode, n_dye, n_groups;
  int gch[AERO2D_MAX_GROUPS];
  int fixed_dt, les;
  double hd, cfl, force_ref, prt, diff_coef, les_c, gx, gy;
  T h, inv_h, kadv, inv_h2, ground;
  // geometry
  const T *chi_u, *chi_v, *chi_c, *a_u, *a_v, *k_u, *k_v;
  const T *chi_t, *a_t, *w_t, *vol_u, *vol_v, *ins_u, *ins_v;
  // params
  const T *nu, *eta, *u_inlet, *dt_fixed, *kappa, *ri, *eta_t, *theta_in, *wall_values, *heat, *heat_max;
  const T *dye_inflow, *g_weight, *g_target, *g_timing;
  // scratch
  T *nut, *F, *psi, *acc_u, *acc_v, *acc_t, *outcol;
  double *partial;     // kRedBlocks doubles (reduce0)
  double *partial2;    // per-block theta sums of the last stage
  Red<T> *red;
  const cplx<T> *tw_x, *tw_y, *cs_x, *cs_y;
  const double *lam_x, *lam_y;
  FftPlan fx, fy;
  int dct_x, dct_y;    // 1: Neumann (DCT-II), 0: periodic (DHT)
};

// Per-stage pointers and coefficients.
template <typename T>
struct Stage {
  int s;
  double a, b, w;
  const T *up, *vp, *thp;   // stage input
  const T *u0, *v0, *th0;   // step start
  T *uh, *vh, *thh;         // stage output (before projection)
};

// ------------------------------------------------------------------ ghost-aware reads
template <typename T>
__device__ __forceinline__ T wall_ghost(int kind, T val, T speed) {
  if (kind == AERO_BC_SLIP) return val;
  if (kind == AERO_BC_NOSLIP) return -val;
  return (T)2 * speed - val;
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// u at row j (-2..ny+1), face i (-2..nx+2) with the ghost rules of core.py:287-305.
template <typename T>
__device__ __forceinline__ T ug(const Dev<T> &d, const T *__restrict__ u, int j, int i) {
  const int nx = d.nx, ny = d.ny, ld = nx + 1;
  if (d.periodic) return u[wrap(j, ny) * ld + wrap(i, nx)];
  T sx = (T)1;
  if (i < 0) {
    if (d.closed) { i = -i; sx = (T)-1; } else i = 0;
  } else if (i > nx) {
    if (d.closed) { i = 2 * nx - i; sx = (T)-1; } else i = nx;
  }
  if (j < 0) return wall_ghost(d.bcb, sx * u[(-1 - j) * ld + i], d.ground);
  if (j >= ny) return wall_ghost(d.bct, sx * u[(2 * ny - 1 - j) * ld + i], (T)0);
  return sx * u[j * ld + i];
}

// v at face row j (-2..ny+2), column i (-2..nx+1) with the rules of core.py:308-325.
template <typename T>
__device__ __forceinline__ T vg(const Dev<T> &d, const T *__restrict__ v, int j, int i) {
  const int nx = d.nx, ny = d.ny;
  if (d.periodic) return v[wrap(j, ny) * nx + wrap(i, nx)];
  T sy = (T)1, sx = (T)1;
  if (j < 0) { j = -j; sy = (T)-1; } else if (j > ny) { j = 2 * ny - j; sy = (T)-1; }
  if (i < 0) {
    i = -1 - i;
    sx = (d.closed && d.bcl != AERO_BC_NOSLIP) ? (T)1 : (T)-1;
  } else if (i >= nx) {
    if (d.closed) { i = 2 * nx - 1 - i; sx = d.bcr != AERO_BC_NOSLIP ? (T)1 : (T)-1; } else i = nx - 1;
  }
  return sx * (sy * v[j * nx + i]);
}

// eddy viscosity at cell (j, c), j, c in -1..ny / -1..nx (written by k_nut)
template <typename T>
__device__ __forceinline__ T nut_at(const Dev<T> &d, int j, int c) {
  return d.nut[(j + 1) * (d.nx + 2) + (c + 1)];
}

// ------------------------------------------------------------------ stencil pieces
// Scaled face flux of core.py:364-388 (central: vsum (b+c); upwind3: 24/7 x flux).
template <typename T>
__device__ __forceinline__ T face_flux(T vsum, T a, T b, T c, T dd, int upwind) {
  const T s = b + c;
  if (!upwind) return vsum * s;
  const T cen = s + (T)(-1.0 / 7.0) * (a + dd);
  const T dis = (b - c) + (T)(-1.0 / 3.0) * (a - dd);
  return vsum * cen + (T)(3.0 / 7.0) * fabs(vsum) * dis;
}

// u-momentum x-flux at cell centre c (between faces c and c+1) of row j
template <typename T>
__device__ __forceinline__ T flux_ux(const Dev<T> &d, const T *u, int j, int c) {
  const T a = ug(d, u, j, c - 1), b = ug(d, u, j, c), cc = ug(d, u, j, c + 1), e = ug(d, u, j, c + 2);
  return face_flux(b + cc, a, b, cc, e, d.upwind);
}
// u-momentum y-flux at corner (jc, i)
template <typename T>
__device__ __forceinline__ T flux_uy(const Dev<T> &d, const T *u, const T *v, int jc, int i) {
  const T vs = vg(d, v, jc, i - 1) + vg(d, v, jc, i);
  return face_flux(vs, ug(d, u, jc - 2, i), ug(d, u, jc - 1, i), ug(d, u, jc, i), ug(d, u, jc + 1, i), d.upwind);
}
// v-momentum x-flux at corner (j, ic)
template <typename T>
__device__ __forceinline__ T flux_vx(const Dev<T> &d, const T *u, const T *v, int j, int ic) {
  const T us = ug(d, u, j - 1, ic) + ug(d, u, j, ic);
  return face_flux(us, vg(d, v, j, ic - 2), vg(d, v, j, ic - 1), vg(d, v, j, ic), vg(d, v, j, ic + 1), d.upwind);
}
// v-momentum y-flux at cell centre row cm (between v faces cm and cm+1)
template <typename T>
__device__ __forceinline__ T flux_vy(const Dev<T> &d, const T *v, int cm, int i) {
  const T a = vg(d, v, cm - 1, i), b = vg(d, v, cm, i), c = vg(d, v, cm + 1, i), e = vg(d, v, cm + 2, i);
  return face_flux(b + c, a, b, c, e, d.upwind);
}

// shear du/dy + dv/dx at corner (jc, ic)
template <typename T>
__device__ __forceinline__ T corner_shear(const Dev<T> &d, const T *u, const T *v, int jc, int ic) {
  return (ug(d, u, jc, ic) - ug(d, u, jc - 1, ic)) * d.inv_h + (vg(d, v, jc, ic) - vg(d, v, jc, ic - 1)) * d.inv_h;
}
template <typename T>
__device__ __forceinline__ T corner_nut(const Dev<T> &d, int jc, int ic) {
  return (((nut_at(d, jc - 1, ic - 1) + nut_at(d, jc, ic - 1)) + nut_at(d, jc - 1, ic)) + nut_at(d, jc, ic)) *
         (T)0.25;
}

// Eddy viscosity of core.py:444-478 from the gradient tensor.
template <typename T>
__device__ __forceinline__ T eddy(const Dev<T> &d, T g11, T g12, T g21, T g22) {
  const T delta2 = d.h * d.h;
  const T c = (T)d.les_c;
  const T s12 = (T)0.5 * (g12 + g21);
  const T ss = (g11 * g11 + g22 * g22) + (T)2 * s12 * s12;
  if (d.turb == 1) return (T)(d.les_c * d.les_c * d.hd * d.hd) * sqrt((T)2 * ss + (T)1e-30);
  if (d.turb == 2) {
    const T q11 = g11 * g11 + g12 * g21;
    const T q22 = g21 * g12 + g22 * g22;
    const T q12 = (T)0.5 * (((g11 * g12 + g12 * g22) + g21 * g11) + g22 * g21);
    const T tr3 = (q11 + q22) * (T)(1.0 / 3.0);
    const T sd = (((q11 - tr3) * (q11 - tr3) + (q22 - tr3) * (q22 - tr3)) + (T)2 * q12 * q12) + tr3 * tr3;
    const T sdp = sd > (T)0 ? sd : (T)0;
    const T ssp = ss > (T)0 ? ss : (T)0;
    const T num = pow(sdp, (T)1.5);
    const T den = (pow(ssp, (T)2.5) + pow(sdp, (T)1.25)) + (T)1e-30;
    return (T)(d.les_c * d.les_c * d.hd * d.hd) * num / den;
  }
  const T b11 = g11 * g11 + g12 * g12;
  const T b22 = g21 * g21 + g22 * g22;
  const T b12 = g11 * g21 + g12 * g22;
  T bb = b11 * b22 - b12 * b12;
  bb = bb > (T)0 ? bb : (T)0;
  const T aa = ((g11 * g11 + g12 * g12) + g21 * g21) + g22 * g22;
  (void)delta2;
  (void)c;
  return (T)(d.les_c * d.hd * d.hd) * sqrt(bb / (aa + (T)1e-30) + (T)1e-30);
}

// torch.lerp(start, end, w)
template <typename T>
__device__ __forceinline__ T lerp_t(T start, T end, T w) {
  return w < (T)0.5 ? start + w * (end - start) : end - (end - start) * ((T)1 - w);
}

// ------------------------------------------------------------------ scalar transport pieces
// cell value with the TVD ghost rules of scalar.py:42-60 (x: inflow / replicate / wrap)
template <typename T>
__device__ __forceinline__ T cx_ghost(const Dev<T> &d, const T *__restrict__ c, int j, int i, T inflow, int use_inflow) {
  const int nx = d.nx;
  if (d.periodic) return c[j * nx + wrap(i, nx)];
  if (i < 0) return (use_inflow == 2 || (use_inflow && !d.closed)) ? inflow : c[j * nx];
  if (i >= nx) return c[j * nx + nx - 1];
  return c[j * nx + i];
}
template <typename T>
__device__ __forceinline__ T cy_ghost(const Dev<T> &d, const T *__restrict__ c, int j, int i) {
  const int nx = d.nx, ny = d.ny;
  if (d.periodic) return c[wrap(j, ny) * nx + i];
  if (j < 0) j = 0;
  if (j >= ny) j = ny - 1;
  return c[j * nx + i];
}
template <typename T>
__device__ __forceinline__ T van_leer(T a, T b) {
  const T ab = a * b;
  return ab > (T)0 ? (T)2 * ab / (a + b) : (T)0;
}
// MUSCL / van Leer advective flux at x-face f of row j (scalar.py:63-83)
template <typename T>
__device__ __forceinline__ T tvd_fx(const Dev<T> &d, const T *c, const T *u, int j, int f, T inflow, int use_inflow) {
  const T cm2 = cx_ghost(d, c, j, f - 2, inflow, use_inflow), cm1 = cx_ghost(d, c, j, f - 1, inflow, use_inflow);
  const T c0 = cx_ghost(d, c, j, f, inflow, use_inflow), c1 = cx_ghost(d, c, j, f + 1, inflow, use_inflow);
  const T sl = van_leer(cm1 - cm2, c0 - cm1);
  const T sr = van_leer(c0 - cm1, c1 - c0);
  const T left = cm1 + (T)0.5 * sl;
  const T right = c0 - (T)0.5 * sr;
  const T uf = u[j * (d.nx + 1) + f];
  return uf * (uf > (T)0 ? left : right);
}
template <typename T>
__device__ __forceinline__ T tvd_fy(const Dev<T> &d, const T *c, const T *v, int f, int i) {
  const T cm2 = cy_ghost(d, c, f - 2, i), cm1 = cy_ghost(d, c, f - 1, i);
  const T c0 = cy_ghost(d, c, f, i), c1 = cy_ghost(d, c, f + 1, i);
  const T sl = van_leer(cm1 - cm2, c0 - cm1);
  const T sr = van_leer(c0 - cm1, c1 - c0);
  const T low = cm1 + (T)0.5 * sl;
  const T high = c0 - (T)0.5 * sr;
  const T vf = v[f * d.nx + i];
  return vf * (vf > (T)0 ? low : high);
}

// temperature diffusivity at x-face f of row j / y-face f of column i (physics.py:713-724)
template <typename T>
__device__ __forceinline__ T kappa_x(const Dev<T> &d, T kap, int j, int f) {
  T k = kap;
  if (d.les) {
    const T ip = (T)(1.0 / d.prt);
    T fa;
    if (d.periodic && (f == 0 || f == d.nx))
      fa = (T)0.5 * (nut_at(d, j, 0) * ip + nut_at(d, j, d.nx - 1) * ip);
    else if (f == 0)
      fa = nut_at(d, j, 0) * ip;
    else if (f == d.nx)
      fa = nut_at(d, j, d.nx - 1) * ip;
    else
      fa = (T)0.5 * (nut_at(d, j, f) * ip + nut_at(d, j, f - 1) * ip);
    k = kap + fa;
  }
  if (d.insulated && d.win && j >= d.j0 && j < d.j1 && f >= d.i0 && f <= d.i1)
    k = k * d.ins_u[(j - d.j0) * (d.wx + 1) + (f - d.i0)];
  return k;
}
template <typename T>
__device__ __forceinline__ T kappa_y(const Dev<T> &d, T kap, int f, int i) {
  T k = kap;
  if (d.les) {
    const T ip = (T)(1.0 / d.prt);
    T fa;
    if (d.periodic && (f == 0 || f == d.ny))
      fa = (T)0.5 * (nut_at(d, 0, i) * ip + nut_at(d, d.ny - 1, i) * ip);
    else if (f == 0)
      fa = nut_at(d, 0, i) * ip;
    else if (f == d.ny)
      fa = nut_at(d, d.ny - 1, i) * ip;
    else
      fa = (T)0.5 * (nut_at(d, f, i) * ip + nut_at(d, f - 1, i) * ip);
    k = kap + fa;
  }
  if (d.insulated && d.win && f >= d.j0 && f <= d.j1 && i >= d.i0 && i < d.i1)
    k = k * d.ins_v[(f - d.j0) * d.wx + (i - d.i0)];
  return k;
}

// boundary diffusive flux of scalar.py:126-139; sign +1 low side, -1 high side
template <typename T>
__device__ __forceinline__ T wall_dflux(const Dev<T> &d, int mode, T k, T inner, T value, double sign) {
  if (mode == AERO_TH_TEMPERATURE || mode == AERO_TH_INLET) return ((T)(-sign * 2.0 / d.hd) * k) * (inner - value);
  if (mode == AERO_TH_HEAT_FLUX) return ((T)sign * k) * value;
  return (T)0;
}

// diffusive flux -kappa dtheta/dx at x-face f of row j (scalar.py:99-141)
template <typename T>
__device__ __forceinline__ T diff_fx(const Dev<T> &d, const T *th, T kap, int j, int f) {
  const int nx = d.nx;
  if (f > 0 && f < nx) return -(kappa_x(d, kap, j, f) * (th[j * nx + f] - th[j * nx + f - 1])) * d.inv_h;
  if (d.periodic) return -(kappa_x(d, kap, j, 0) * (th[j * nx] - th[j * nx + nx - 1])) * d.inv_h;
  if (f == 0) return wall_dflux(d, d.wm[0], kappa_x(d, kap, j, 0), th[j * nx], d.wall_values[0], 1.0);
  return wall_dflux(d, d.wm[1], kappa_x(d, kap, j, nx), th[j * nx + nx - 1], d.wall_values[1], -1.0);
}
template <typename T>
__device__ __forceinline__ T diff_fy(const Dev<T> &d, const T *th, T kap, int f, int i) {
  const int nx = d.nx, ny = d.ny;
  if (f > 0 && f < ny) return -(kappa_y(d, kap, f, i) * (th[f * nx + i] - th[(f - 1) * nx + i])) * d.inv_h;
  if (d.periodic) return -(kappa_y(d, kap, 0, i) * (th[i] - th[(ny - 1) * nx + i])) * d.inv_h;
  if (f == 0) return wall_dflux(d, d.wm[2], kappa_y(d, kap, 0, i), th[i], d.wall_values[2], 1.0);
  return wall_dflux(d, d.wm[3], kappa_y(d, kap, ny, i), th[(ny - 1) * nx + i], d.wall_values[3], -1.0);
}

// Boussinesq force per unit Ri*(-g) component: face average of theta, zero on boundary faces
template <typename T>
__device__ __forceinline__ T theta_face_x(const Dev<T> &d, const T *th, int j, int f) {
  const int nx = d.nx;
  if (f == 0 || f == nx) {
    if (!d.periodic) return (T)0;
    return (T)0.5 * (th[j * nx] + th[j * nx + nx - 1]);
  }
  return (T)0.5 * (th[j * nx + f] + th[j * nx + f - 1]);
}
template <typename T>
__device__ __forceinline__ T theta_face_y(const Dev<T> &d, const T *th, int f, int i) {
  const int nx = d.nx, ny = d.ny;
  if (f == 0 || f == ny) {
    if (!d.periodic) return (T)0;
    return (T)0.5 * (th[i] + th[(ny - 1) * nx + i]);
  }
  return (T)0.5 * (th[f * nx + i] + th[(f - 1) * nx + i]);
}

// ================================================================== kernels
// Eddy viscosity at cells -1..ny x -1..nx of the stage velocity, switched off inside
// solids (core.py:421-503); the max goes to red->nutmax on stage 0.
template <typename T>
__global__ void __launch_bounds__(256) k_nut(Dev<T> d, const T *__restrict__ u, const T *__restrict__ v, int want_max) {
  __shared__ T lds[256];
  const int c = blockIdx.x * blockDim.x + threadIdx.x - 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y - 1;
  T val = (T)0;
  if (c <= d.nx && j <= d.ny) {
    const T ih = d.inv_h;
    const T g11 = (ug(d, u, j, c + 1) - ug(d, u, j, c)) * ih;
    const T g22 = (vg(d, v, j + 1, c) - vg(d, v, j, c)) * ih;
    const T du00 = (ug(d, u, j, c) - ug(d, u, j - 1, c)) * ih;
    const T du10 = (ug(d, u, j + 1, c) - ug(d, u, j, c)) * ih;
    const T du01 = (ug(d, u, j, c + 1) - ug(d, u, j - 1, c + 1)) * ih;
    const T du11 = (ug(d, u, j + 1, c + 1) - ug(d, u, j, c + 1)) * ih;
    const T dv00 = (vg(d, v, j, c) - vg(d, v, j, c - 1)) * ih;
    const T dv10 = (vg(d, v, j + 1, c) - vg(d, v, j + 1, c - 1)) * ih;
    const T dv01 = (vg(d, v, j, c + 1) - vg(d, v, j, c)) * ih;
    const T dv11 = (vg(d, v, j + 1, c + 1) - vg(d, v, j + 1, c)) * ih;
    const T g12 = (((du00 + du10) + du01) + du11) * (T)0.25;
    const T g21 = (((dv00 + dv10) + dv01) + dv11) * (T)0.25;
    val = eddy(d, g11, g12, g21, g22);
    if (d.win && j >= d.j0 && j < d.j1 && c >= d.i0 && c < d.i1)
      val = val * ((T)1 - d.chi_c[(j - d.j0) * d.wx + (c - d.i0)]);
    d.nut[(j + 1) * (d.nx + 2) + (c + 1)] = val;
  }
  if (want_max) {
    const T m = block_max(val, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_nonneg(&d.red->nutmax, m);
  }
}

// dt of the step (core.py:685-695 + physics.py:851-888) from the maxima of the state.
template <typename T>
__device__ T compute_dt(const Dev<T> &d, T umax, T vmax, T thmax, T nutmax) {
  if (d.fixed_dt) return *d.dt_fixed;
  const T nu = *d.nu;
  T us = umax + vmax;
  const T uc = us > (T)1e-6 ? us : (T)1e-6;
  T dt = (T)(d.cfl * d.hd) / uc;
  const T nue = d.les ? nu + nutmax : nu;
  const T vis = (T)(0.25 * d.hd * d.hd) / nue;
  dt = vis < dt ? vis : dt;
  if (d.thermal) {
    const T kap = d.les ? *d.kappa + nutmax * (T)(1.0 / d.prt) : *d.kappa;
    const T speed = us;   // u_floor is 0 without ports and fans
    const T rate = speed * (T)(2.0 / d.hd) + kap * (T)(d.diff_coef / (d.hd * d.hd));
    const T lim = (T)1 / rate;
    dt = lim < dt ? lim : dt;
    T acc = fabs(*d.ri) * thmax;
    acc = acc > (T)1e-12 ? acc : (T)1e-12;
    const T ff = (T)d.cfl * sqrt((T)d.hd / acc);
    dt = ff < dt ? ff : dt;
    if (d.has_heat) {
      T smax = *d.heat_max;
      smax = smax > (T)1e-12 ? smax : (T)1e-12;
      const T l1 = (T)1 / smax;
      dt = l1 < dt ? l1 : dt;
      T q = fabs(*d.ri) * smax;
      q = q > (T)1e-12 ? q : (T)1e-12;
      const T l2 = pow((T)(6.0 * d.cfl * d.hd) / q, (T)(1.0 / 3.0));
      dt = l2 < dt ? l2 : dt;
    }
  }
  return dt;
}

// max |u0|, |v0|, |theta0| and sum theta0 over the state; the last block sets dt.
template <typename T>
__global__ void __launch_bounds__(256) k_reduce0(Dev<T> d, const T *__restrict__ u, const T *__restrict__ v,
                                                 const T *__restrict__ th) {
  __shared__ T lds[256];
  __shared__ double ldsd[256];
  __shared__ bool last;
  const int nu_ = d.ny * (d.nx + 1), nv_ = (d.ny + 1) * d.nx, nc = d.ny * d.nx;
  const int n = nu_ > nv_ ? nu_ : nv_;
  T mu = 0, mv = 0, mt = 0;
  double st = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    if (e < nu_) { T a = fabs(u[e]); mu = a > mu ? a : mu; }
    if (e < nv_) { T a = fabs(v[e]); mv = a > mv ? a : mv; }
    if (d.thermal && e < nc) { T a = fabs(th[e]); mt = a > mt ? a : mt; st += (double)th[e]; }
  }
  mu = block_max(mu, lds);
  mv = block_max(mv, lds);
  if (d.thermal) {
    mt = block_max(mt, lds);
    st = block_sum(st, ldsd);
  }
  if (threadIdx.x == 0) {
    atomic_max_nonneg(&d.red->umax, mu);
    atomic_max_nonneg(&d.red->vmax, mv);
    if (d.thermal) atomic_max_nonneg(&d.red->thmax, mt);
    d.partial[blockIdx.x] = st;
    __threadfence();
    const unsigned t = atomicAdd(&d.red->ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // deterministic sum of the per-block partials
  double s = 0.0;
  for (int b = threadIdx.x; b < gridDim.x; b += blockDim.x) s += load_relaxed(&d.partial[b]);
  s = block_sum(s, ldsd);
  if (threadIdx.x == 0) {
    Red<T> *r = d.red;
    const T umax = load_relaxed(&r->umax), vmax = load_relaxed(&r->vmax);
    const T thmax = load_relaxed(&r->thmax), nutmax = load_relaxed(&r->nutmax);
    r->dt = compute_dt(d, umax, vmax, thmax, nutmax);
    r->th_old = s;
    r->umax = (T)0;
    r->vmax = (T)0;
    r->thmax = (T)0;
    r->nutmax = (T)0;
    r->ticket = 0u;
  }
}

// The explicit stage (core.py:855-935 per face / cell): tendencies, RK update,
// inlet/outlet, penalisation with force increments, buoyancy, temperature stage.
template <typename T>
__global__ void __launch_bounds__(256) k_stage(Dev<T> d, Stage<T> s) {
  __shared__ double ldsd[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = d.nx, ny = d.ny;
  const T dt = d.red->dt;
  const T bdt = dt * (T)s.b;
  const T a = (T)s.a;
  const T wgt = (T)s.w;
  const T nu_h2 = *d.nu * d.inv_h2;
  const T kpen = bdt / *d.eta;
  const T *up = s.up, *vp = s.vp;
  const bool first = s.s == 0;
  T ri_g = 0;
  if (d.thermal && d.buoyant) ri_g = *d.ri;

  // ---------------- u face (j, i)
  if (j < ny && i <= nx) {
    const bool bnd = !d.periodic && (i == 0 || i == nx);
    const int ie = (d.periodic && i == nx) ? 0 : i;   // periodic duplicate face computes as face 0
    T lu;
    if (bnd) {
      lu = (d.closed || i == 0) ? (T)0 : (up[j * (nx + 1) + nx - 1] - up[j * (nx + 1) + nx]) * d.inv_h;
    } else {
      const T fxl = flux_ux(d, up, j, ie - 1), fxr = flux_ux(d, up, j, ie);
      const T gyb = flux_uy(d, up, vp, j, ie), gyt = flux_uy(d, up, vp, j + 1, ie);
      const T au = (((fxl - fxr) + gyb) - gyt) * d.kadv;
      const T uc = ug(d, up, j, ie);
      const T ring = ((ug(d, up, j, ie + 1) + ug(d, up, j, ie - 1)) + ug(d, up, j + 1, ie)) + ug(d, up, j - 1, ie);
      lu = au + (ring + (T)-4 * uc) * nu_h2;
      if (d.les) {
        const T txx_r = (T)2 * nut_at(d, j, ie) * ((ug(d, up, j, ie + 1) - ug(d, up, j, ie)) * d.inv_h);
        const T txx_l = (T)2 * nut_at(d, j, ie - 1) * ((ug(d, up, j, ie) - ug(d, up, j, ie - 1)) * d.inv_h);
        const T txy_t = corner_nut(d, j + 1, ie) * corner_shear(d, up, vp, j + 1, ie);
        const T txy_b = corner_nut(d, j, ie) * corner_shear(d, up, vp, j, ie);
        lu = lu + (((txx_r - txx_l) + txy_t) - txy_b) * d.inv_h;
      }
    }
    const int o = j * (nx + 1) + i;
    const T upv = up[o];
    T uh = (s.a == 0.0 ? upv : lerp_t(upv, s.u0[o], a)) + lu * bdt;
    if (!d.periodic) {
      if (d.closed) {
        if (i == 0 || i == nx) uh = (T)0;
      } else if (i == 0) {
        uh = d.u_inlet[j];
      } else if (i == nx) {
        d.outcol[j] = uh;     // outlet mass balance applied by the projection passes
      }
    }
    T fu = 0;
    if (ri_g != (T)0 && d.gx != 0.0) fu = theta_face_x(d, s.thp, j, i) * (ri_g * (T)(-d.gx));
    const bool inwin = d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i <= d.i1;
    T ustore = uh;
    if (inwin) {
      const int wo = (j - d.j0) * (d.wx + 1) + (i - d.i0);
      const T un = (uh + kpen * d.a_u[wo]) / ((T)1 + kpen * d.chi_u[wo]);
      T acc = (un - uh) * wgt;
      if (!first) acc = d.acc_u[wo] + acc;
      if (fu != (T)0) {
        const T vu = d.vol_u ? d.vol_u[wo] : d.chi_u[wo];
        acc = acc + (fu * vu) * (bdt * wgt);
      }
      d.acc_u[wo] = acc;
      ustore = un;
    }
    if (d.periodic && i == nx) {
      // duplicate face: the value of face 0 (penalised there), core.py:912-913
      ustore = uh;
      if (d.win && j >= d.j0 && j < d.j1 && d.i0 == 0) {
        const int wo0 = (j - d.j0) * (d.wx + 1);
        ustore = (uh + kpen * d.a_u[wo0]) / ((T)1 + kpen * d.chi_u[wo0]);
      }
    }
    if (fu != (T)0) ustore = ustore + bdt * fu;
    s.uh[o] = ustore;
  }

  // ---------------- v face (j, i)
  if (j <= ny && i < nx) {
    const bool bnd = !d.periodic && (j == 0 || j == ny);
    const int je = (d.periodic && j == ny) ? 0 : j;
    T lv;
    if (bnd) {
      lv = (T)0;
    } else {
      const T fxl = flux_vx(d, up, vp, je, i), fxr = flux_vx(d, up, vp, je, i + 1);
      const T gyb = flux_vy(d, vp, je - 1, i), gyt = flux_vy(d, vp, je, i);
      const T av = (((fxl - fxr) + gyb) - gyt) * d.kadv;
      const T vc = vg(d, vp, je, i);
      const T ring = ((vg(d, vp, je, i + 1) + vg(d, vp, je, i - 1)) + vg(d, vp, je + 1, i)) + vg(d, vp, je - 1, i);
      lv = av + (ring + (T)-4 * vc) * nu_h2;
      if (d.les) {
        const T txy_r = corner_nut(d, je, i + 1) * corner_shear(d, up, vp, je, i + 1);
        const T txy_l = corner_nut(d, je, i) * corner_shear(d, up, vp, je, i);
        const T tyy_t = (T)2 * nut_at(d, je, i) * ((vg(d, vp, je + 1, i) - vg(d, vp, je, i)) * d.inv_h);
        const T tyy_b = (T)2 * nut_at(d, je - 1, i) * ((vg(d, vp, je, i) - vg(d, vp, je - 1, i)) * d.inv_h);
        lv = lv + (((txy_r - txy_l) + tyy_t) - tyy_b) * d.inv_h;
      }
    }
    const int o = j * nx + i;
    const T vpv = vp[o];
    T vh = (s.a == 0.0 ? vpv : lerp_t(vpv, s.v0[o], a)) + lv * bdt;
    T fv = 0;
    if (ri_g != (T)0 && d.gy != 0.0) fv = theta_face_y(d, s.thp, j, i) * (ri_g * (T)(-d.gy));
    const bool inwin = d.win && j >= d.j0 && j <= d.j1 && i >= d.i0 && i < d.i1;
    T vstore = vh;
    if (inwin) {
      const int wo = (j - d.j0) * d.wx + (i - d.i0);
      const T vn = (vh + kpen * d.a_v[wo]) / ((T)1 + kpen * d.chi_v[wo]);
      T acc = (vn - vh) * wgt;
      if (!first) acc = d.acc_v[wo] + acc;
      if (fv != (T)0) {
        const T vv = d.vol_v ? d.vol_v[wo] : d.chi_v[wo];
        acc = acc + (fv * vv) * (bdt * wgt);
      }
      d.acc_v[wo] = acc;
      vstore = vn;
    }
    if (d.periodic && j == ny) {
      vstore = vh;
      if (d.win && d.j0 == 0 && i >= d.i0 && i < d.i1) {
        const int wo0 = i - d.i0;
        vstore = (vh + kpen * d.a_v[wo0]) / ((T)1 + kpen * d.chi_v[wo0]);
      }
    }
    if (fv != (T)0) vstore = vstore + bdt * fv;
    s.vh[o] = vstore;
  }

  // ---------------- temperature cell (j, i) (physics.py:696-757)
  if (d.thermal) {
    double part = 0.0;
    if (j < ny && i < nx) {
      const T kap = *d.kappa;
      const T tin = *d.theta_in;
      const T *th = s.thp;
      T lt = (((tvd_fx(d, th, up, j, i, tin, 1) - tvd_fx(d, th, up, j, i + 1, tin, 1)) + tvd_fy(d, th, vp, j, i)) -
              tvd_fy(d, th, vp, j + 1, i)) * d.inv_h;
      lt = lt + (((diff_fx(d, th, kap, j, i) - diff_fx(d, th, kap, j, i + 1)) + diff_fy(d, th, kap, j, i)) -
                 diff_fy(d, th, kap, j + 1, i)) * d.inv_h;
      if (d.has_heat) lt = lt + d.heat[j * nx + i];
      const int o = j * nx + i;
      const T tp = th[o];
      T thh = (s.a == 0.0 ? tp : lerp_t(tp, s.th0[o], a)) + lt * bdt;
      if (d.n_temp > 0 && d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1) {
        const int wo = (j - d.j0) * d.wx + (i - d.i0);
        const T kt = bdt / *d.eta_t;
        const T tn = (thh + kt * d.a_t[wo]) / ((T)1 + kt * d.chi_t[wo]);
        T acc = (tn - thh) * wgt;
        if (!first) acc = d.acc_t[wo] + acc;
        d.acc_t[wo] = acc;
        thh = tn;
      }
      s.thh[o] = thh;
      part = (double)thh;
    }
    if (s.s == 2) {
      // per-block partial of sum theta_new (heat storage), fixed order
      part = block_sum(part, ldsd);
      if (threadIdx.x == 0 && threadIdx.y == 0) d.partial2[blockIdx.y * gridDim.x + blockIdx.x] = part;
    }
  }
}

// natural index of reordered position n (inverse of makhoul_pos)
__device__ __forceinline__ int makhoul_src(int n, int len) { return n < (len + 1) / 2 ? 2 * n : 2 * (len - 1 - n) + 1; }

// Outlet correction of core.py:337-357, computed identically by every block.
template <typename T>
__device__ T outlet_corr(const Dev<T> &d, T *red) {
  T si = 0, so = 0;
  const int tid = threadIdx.x;
  for (int j = tid; j < d.ny; j += blockDim.x) {
    si += d.u_inlet[j];
    so += d.outcol[j];
  }
  si = block_sum(si, red);
  so = block_sum(so, red);
  return (si - so) / (T)d.ny;
}

// divergence at cell (j, i) with the outlet face shifted by corr (core.py:621-623)
template <typename T>
__device__ __forceinline__ T div_at(const Dev<T> &d, const T *u, const T *v, int j, int i, T corr) {
  const int nx = d.nx;
  T ur = u[j * (nx + 1) + i + 1];
  if (i + 1 == nx) ur = ur + corr;
  return (((ur - u[j * (nx + 1) + i]) + v[(j + 1) * nx + i]) - v[j * nx + i]) * d.inv_h;
}

// Pass A: divergence of rows (2b, 2b+1) -> forward transform along x -> F rows.
template <typename T>
__global__ void __launch_bounds__(kFftThreads) k_poisA(Dev<T> d, const T *__restrict__ u, const T *__restrict__ v,
                                                       const T *__restrict__ f_in, int mode) {
  // mode 0: f_in is the right-hand side itself; 1: divergence of (u, v); 2: + outlet correction
  extern __shared__ unsigned char smem[];
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + d.nx;
  T *red = reinterpret_cast<T *>(B + d.nx);
  const int nx = d.nx;
  const int ja = 2 * blockIdx.x, jb = ja + 1;
  const bool hasb = jb < d.ny;
  T corr = 0;
  if (mode == 2) {
    corr = outlet_corr(d, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) d.red->corr = corr;
  }
  for (int n = threadIdx.x; n < nx; n += blockDim.x) {
    const int q = d.dct_x ? makhoul_src(n, nx) : n;
    T xa, xb = 0;
    if (mode == 0) {
      xa = f_in[ja * nx + q];
      if (hasb) xb = f_in[jb * nx + q];
    } else {
      xa = div_at(d, u, v, ja, q, corr);
      if (hasb) xb = div_at(d, u, v, jb, q, corr);
    }
    A[n] = {xa, xb};
  }
  __syncthreads();
  cplx<T> *R = fft_lds(A, B, d.fx, d.tw_x, 1, nx);
  for (int k = threadIdx.x; k < nx; k += blockDim.x) {
    cplx<T> va, vb;
    split_pair(R, nx, k, va, vb);
    T xa, xb;
    if (d.dct_x) {
      const cplx<T> cs = d.cs_x[k];
      xa = cs.x * va.x + cs.y * va.y;
      xb = cs.x * vb.x + cs.y * vb.y;
    } else {
      xa = va.x - va.y;
      xb = vb.x - vb.y;
    }
    d.F[ja * nx + k] = xa;
    if (hasb) d.F[jb * nx + k] = xb;
  }
}

// Inverse real transform of two spectra held in LDS real arrays xa, xb (length n) into
// the complex buffer Zbuf (packed for the FFT); after fft_lds the caller reads
// v_a[m] = R[m].x * inv_n, v_b[m] = -R[m].y * inv_n (DCT: at natural index makhoul_src).
template <typename T>
__device__ __forceinline__ void pack_inverse(const T *xa, const T *xb, cplx<T> *Z, int n, int dct,
                                             const cplx<T> *__restrict__ cs, int k) {
  if (dct) {
    const cplx<T> w = cs[k];
    const T xa_k = xa[k], xb_k = xb[k];
    const T xa_r = k == 0 ? (T)0 : xa[n - k];
    const T xb_r = k == 0 ? (T)0 : xb[n - k];
    // V = e^{+i pi k / 2n} (X_k - i X_{n-k})
    const cplx<T> va = {w.x * xa_k + w.y * xa_r, w.y * xa_k - w.x * xa_r};
    const cplx<T> vb = {w.x * xb_k + w.y * xb_r, w.y * xb_k - w.x * xb_r};
    // Z = va + i vb, stored conjugated (inverse FFT via the forward one)
    Z[k] = {va.x - vb.y, -(va.y + vb.x)};
  } else {
    // DHT is its own inverse: pack the two real spectra and transform forward
    Z[k] = {xa[k], xb[k]};
  }
}

// Pass B: columns (2c, 2c+1) of `nslot` column pairs per block: forward along y,
// multiply by the inverse eigenvalues (FP64-built), inverse along y, in place in F.
template <typename T>
__global__ void __launch_bounds__(kFftThreads) k_poisB(Dev<T> d, int nslot) {
  extern __shared__ unsigned char smem[];
  const int ny = d.ny, nx = d.nx;
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + nslot * ny;
  T *X = reinterpret_cast<T *>(B + nslot * ny);   // 2 * nslot * ny reals
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int c0 = 2 * nslot * blockIdx.x;
  for (int w = tid; w < nslot * ny; w += nthr) {
    const int f = w / ny, n = w - f * ny;
    const int ca = c0 + 2 * f, cb = ca + 1;
    const int q = d.dct_y ? makhoul_src(n, ny) : n;
    const T xa = ca < nx ? d.F[q * nx + ca] : (T)0;
    const T xb = cb < nx ? d.F[q * nx + cb] : (T)0;
    A[f * ny + n] = {xa, xb};
  }
  __syncthreads();
  cplx<T> *R = fft_lds(A, B, d.fy, d.tw_y, nslot, ny);
  for (int w = tid; w < nslot * ny; w += nthr) {
    const int f = w / ny, k = w - f * ny;
    const int ca = c0 + 2 * f, cb = ca + 1;
    cplx<T> va, vb;
    split_pair(R + f * ny, ny, k, va, vb);
    T xa, xb;
    if (d.dct_y) {
      const cplx<T> cs = d.cs_y[k];
      xa = cs.x * va.x + cs.y * va.y;
      xb = cs.x * vb.x + cs.y * vb.y;
    } else {
      xa = va.x - va.y;
      xb = vb.x - vb.y;
    }
    // 1 / (lambda_x + lambda_y) in double, zero for the constant mode (poisson.py:119-126)
    const double ly = d.lam_y[k];
    double la = d.lam_x[ca < nx ? ca : 0] + ly;
    double lb = d.lam_x[cb < nx ? cb : 0] + ly;
    const double ia = (la != 0.0 && !(ca == 0 && k == 0)) ? 1.0 / la : 0.0;
    const double ib = (lb != 0.0) ? 1.0 / lb : 0.0;
    X[(2 * f) * ny + k] = xa * (T)ia;
    X[(2 * f + 1) * ny + k] = xb * (T)ib;
  }
  __syncthreads();
  cplx<T> *Z = (R == A) ? B : A;
  for (int w = tid; w < nslot * ny; w += nthr) {
    const int f = w / ny, k = w - f * ny;
    pack_inverse(X + (2 * f) * ny, X + (2 * f + 1) * ny, Z + f * ny, ny, d.dct_y, d.cs_y, k);
  }
  __syncthreads();
  cplx<T> *R2 = fft_lds(Z, (Z == A) ? B : A, d.fy, d.tw_y, nslot, ny);
  const T inv_n = (T)(1.0 / (double)ny);
  for (int w = tid; w < nslot * ny; w += nthr) {
    const int f = w / ny, m = w - f * ny;
    const int ca = c0 + 2 * f, cb = ca + 1;
    T va, vb;
    if (d.dct_y) {
      va = R2[f * ny + m].x * inv_n;
      vb = -R2[f * ny + m].y * inv_n;
    } else {
      cplx<T> ha, hb;
      split_pair(R2 + f * ny, ny, m, ha, hb);
      va = (ha.x - ha.y) * inv_n;
      vb = (hb.x - hb.y) * inv_n;
    }
    const int q = d.dct_y ? makhoul_src(m, ny) : m;
    if (ca < nx) d.F[q * nx + ca] = va;
    if (cb < nx) d.F[q * nx + cb] = vb;
  }
}

// Pass C: inverse along x of rows (2b, 2b+1) -> psi; accumulate the step-averaged
// pressure p (core.py:935-942): stage 0 p = w psi, stage 1 p += w psi, stage 2
// p = (p + psi) / dt.
template <typename T>
__global__ void __launch_bounds__(kFftThreads) k_poisC(Dev<T> d, T *__restrict__ psi, T *__restrict__ p, int stage) {
  extern __shared__ unsigned char smem[];
  const int nx = d.nx;
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + nx;
  T *X = reinterpret_cast<T *>(B + nx);
  const int ja = 2 * blockIdx.x, jb = ja + 1;
  const bool hasb = jb < d.ny;
  for (int k = threadIdx.x; k < nx; k += blockDim.x) {
    X[k] = d.F[ja * nx + k];
    X[nx + k] = hasb ? d.F[jb * nx + k] : (T)0;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nx; k += blockDim.x) pack_inverse(X, X + nx, A, nx, d.dct_x, d.cs_x, k);
  __syncthreads();
  cplx<T> *R = fft_lds(A, B, d.fx, d.tw_x, 1, nx);
  const T inv_n = (T)(1.0 / (double)nx);
  T rdt = 0;
  if (p != nullptr && stage == 2) rdt = (T)1 / d.red->dt;
  for (int m = threadIdx.x; m < nx; m += blockDim.x) {
    T va, vb;
    if (d.dct_x) {
      va = R[m].x * inv_n;
      vb = -R[m].y * inv_n;
    } else {
      cplx<T> ha, hb;
      split_pair(R, nx, m, ha, hb);
      va = (ha.x - ha.y) * inv_n;
      vb = (hb.x - hb.y) * inv_n;
    }
    const int q = d.dct_x ? makhoul_src(m, nx) : m;
    psi[ja * nx + q] = va;
    if (hasb) psi[jb * nx + q] = vb;
    if (p != nullptr) {
      const T w = (T)kForceW[stage];
      if (stage == 0) {
        p[ja * nx + q] = va * w;
        if (hasb) p[jb * nx + q] = vb * w;
      } else if (stage == 1) {
        p[ja * nx + q] = p[ja * nx + q] + w * va;
        if (hasb) p[jb * nx + q] = p[jb * nx + q] + w * vb;
      } else {
        p[ja * nx + q] = (p[ja * nx + q] + va) * rdt;
        if (hasb) p[jb * nx + q] = (p[jb * nx + q] + vb) * rdt;
      }
    }
  }
}

// u -= grad psi on interior faces (core.py:645-660); the outlet gets the mass correction.
template <typename T>
__global__ void __launch_bounds__(256) k_correct(Dev<T> d, T *__restrict__ u, T *__restrict__ v,
                                                 const T *__restrict__ psi, int use_corr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = d.nx, ny = d.ny;
  if (j < ny && i <= nx) {
    const int o = j * (nx + 1) + i;
    if (d.periodic) {
      const int ie = i == nx ? 0 : i;
      u[o] = u[o] - (psi[j * nx + ie] - psi[j * nx + wrap(ie - 1, nx)]) * d.inv_h;
    } else if (i > 0 && i < nx) {
      u[o] = u[o] - (psi[j * nx + i] - psi[j * nx + i - 1]) * d.inv_h;
    } else if (i == nx && use_corr && !d.closed) {
      u[o] = u[o] + d.red->corr;
    }
  }
  if (j <= ny && i < nx) {
    const int o = j * nx + i;
    if (d.periodic) {
      const int je = j == ny ? 0 : j;
      v[o] = v[o] - (psi[je * nx + i] - psi[wrap(je - 1, ny) * nx + i]) * d.inv_h;
    } else if (j > 0 && j < ny) {
      v[o] = v[o] - (psi[j * nx + i] - psi[(j - 1) * nx + i]) * d.inv_h;
    }
  }
}

// emitter clock (physics.py:798-802)
template <typename T>
__device__ __forceinline__ T emitter_on(T t, const T *tm) {
  const T start = tm[0], stop = tm[1], period = tm[2], duty = tm[3];
  const bool active = (t >= start) && (t < stop);
  const T safe = period > (T)0 ? period : (T)1;
  const T x = t - start;
  T r = fmod(x, safe);
  if (r != (T)0 && ((r < (T)0) != (safe < (T)0))) r += safe;
  const T phase = r / safe;
  const bool on = active && ((period <= (T)0) || (phase < duty));
  return on ? (T)1 : (T)0;
}

// Smoke step with the end-of-step velocity: channels + emitters (physics.py:805-830)
// or the inlet comb alone (core.py:709-733, 758-771).
template <typename T>
__global__ void __launch_bounds__(256) k_dye(Dev<T> d, const T *__restrict__ cin, T *__restrict__ cout,
                                             const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ t_in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int ch = blockIdx.z;
  const int nx = d.nx, ny = d.ny;
  if (j >= ny || i >= nx) return;
  const T dt = d.red->dt;
  const T *c = cin + (size_t)ch * ny * nx;
  const int o = j * nx + i;
  const T inflow = d.dye_inflow ? d.dye_inflow[ch * ny + j] : (T)0;
  const int mode_in = d.dye_mode == 1 ? 2 : 1;   // the comb path always uses its inlet column
  const T fxl = tvd_fx(d, c, u, j, i, inflow, mode_in), fxr = tvd_fx(d, c, u, j, i + 1, inflow, mode_in);
  const T fyb = tvd_fy(d, c, v, j, i), fyt = tvd_fy(d, c, v, j + 1, i);
  T val;
  if (d.dye_mode == 2) {
    val = c[o] + dt * ((((fxl - fxr) + fyb) - fyt) * d.inv_h);
    const T tnew = *t_in + dt;
    for (int g = 0; g < d.n_groups; ++g) {
      if (d.gch[g] != ch) continue;
      const T wg = d.g_weight[(size_t)g * ny * nx + o] * emitter_on(tnew, d.g_timing + 4 * g);
      val = val + wg * (d.g_target[(size_t)g * ny * nx + o] - val);
    }
    if (d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1)
      val = val * (d.chi_c[(j - d.j0) * d.wx + (i - d.i0)] <= (T)0.5 ? (T)1 : (T)0);
    val = val > (T)1e-6 ? val : (T)0;
    val = val < (T)1 ? val : (T)1;
  } else {
    val = c[o] - (dt * d.inv_h) * (((fxr - fxl) + fyt) - fyb);
    val = val > (T)1e-6 ? val : (T)0;
    val = val < (T)1 ? val : (T)1;
    if (d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1)
      val = val * (d.chi_c[(j - d.j0) * d.wx + (i - d.i0)] <= (T)0.5 ? (T)1 : (T)0);
  }
  cout[(size_t)ch * ny * nx + o] = val;
}

template <typename T>
__global__ void k_copy(T *__restrict__ dst, const T *__restrict__ src, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// Per-step outputs (core.py:942-1004, physics.py:760-795): forces, heat rates, t, dt,
// history rings. One block; sums in double, fixed order.
template <typename T>
__global__ void __launch_bounds__(1024) k_finalize(Dev<T> d, const T *__restrict__ t_in, T *__restrict__ t_out,
                                                   T *__restrict__ dt_out, const T *__restrict__ u_new,
                                                   const T *__restrict__ th_new, aero2d_outputs o, int n_stage_blocks) {
  __shared__ double ldsd[1024];
  __shared__ double res[64];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const T dt = d.red->dt;
  const T tnew = *t_in + dt;
  const int wy = d.wy, wx = d.wx;
  const int nuw = wy * (wx + 1), nvw = (wy + 1) * wx;
  // forces
  T sumc[3] = {0, 0, 0};
  for (int b = 0; b < d.nbody; ++b) {
    double r[4] = {0, 0, 0, 0};
    if (d.win) {
      const T *ku = d.k_u + (size_t)b * 2 * nuw;
      const T *kv = d.k_v + (size_t)b * 2 * nvw;
      for (int e = tid; e < nuw; e += nthr) {
        r[0] += (double)ku[e] * (double)d.acc_u[e];
        r[1] += (double)ku[nuw + e] * (double)d.acc_u[e];
      }
      for (int e = tid; e < nvw; e += nthr) {
        r[2] += (double)kv[e] * (double)d.acc_v[e];
        r[3] += (double)kv[nvw + e] * (double)d.acc_v[e];
      }
    }
    for (int q = 0; q < 4; ++q) r[q] = block_sum(r[q], ldsd);
    const T scale = (T)(-(d.hd * d.hd)) / ((dt * (T)0.5) * (T)d.force_ref);
    const T cd = (T)r[0] * scale, cl = (T)r[2] * scale, cm = ((T)r[1] + (T)r[3]) * scale;
    if (tid == 0 && o.coeffs) {
      T *cf = reinterpret_cast<T *>(o.coeffs);
      cf[3 * b] = cd;
      cf[3 * b + 1] = cl;
      cf[3 * b + 2] = cm;
    }
    sumc[0] += cd;
    sumc[1] += cl;
    sumc[2] += cm;
  }
  // heat outputs
  const int nxp = d.nx + 1, nx = d.nx, ny = d.ny;
  const T h2dt = (T)(d.hd * d.hd) / dt;
  if (d.thermal) {
    for (int b = 0; b < d.n_temp; ++b) {
      double r = 0;
      if (d.win)
        for (int e = tid; e < wy * wx; e += nthr) r += (double)d.w_t[(size_t)b * wy * wx + e] * (double)d.acc_t[e];
      r = block_sum(r, ldsd);
      if (tid == 0) res[b] = (double)((T)r * h2dt);
    }
    // wall heat (left, right, bottom, top), inflow / outflow heat, storage
    const T kap = *d.kappa;
    double wsum[4] = {0, 0, 0, 0}, fl[2] = {0, 0}, st = 0;
    for (int jj = tid; jj < ny; jj += nthr) {
      wsum[0] += (double)(d.wall_values[0] - th_new[jj * nx]);
      wsum[1] += (double)(d.wall_values[1] - th_new[jj * nx + nx - 1]);
      fl[0] += (double)(u_new[jj * nxp] * *d.theta_in);
      fl[1] += (double)(u_new[jj * nxp + nx] * th_new[jj * nx + nx - 1]);
    }
    for (int ii = tid; ii < nx; ii += nthr) {
      wsum[2] += (double)(d.wall_values[2] - th_new[ii]);
      wsum[3] += (double)(d.wall_values[3] - th_new[(ny - 1) * nx + ii]);
    }
    for (int b = tid; b < n_stage_blocks; b += nthr) st += d.partial2[b];
    for (int q = 0; q < 4; ++q) wsum[q] = block_sum(wsum[q], ldsd);
    fl[0] = block_sum(fl[0], ldsd);
    fl[1] = block_sum(fl[1], ldsd);
    st = block_sum(st, ldsd);
    if (tid == 0) {
      const int nt = d.n_temp;
      const int len[4] = {ny, ny, nx, nx};
      for (int q = 0; q < 4; ++q) {
        const int m = d.wm[q];
        double val = 0;
        if (m == AERO_TH_TEMPERATURE || m == AERO_TH_INLET) val = (double)(((T)2 * kap) * (T)wsum[q]);
        else if (m == AERO_TH_HEAT_FLUX) val = (double)(kap * d.wall_values[q] * (T)(len[q] * d.hd));
        res[nt + q] = val;
      }
      res[nt + 4] = (double)((T)(st - d.red->th_old) * h2dt);
      const bool tunnel = !d.periodic && !d.closed;
      res[nt + 5] = tunnel ? (double)((T)fl[0] * (T)d.hd) : 0.0;
      res[nt + 6] = tunnel ? (double)((T)fl[1] * (T)d.hd) : 0.0;
    }
  }
  __syncthreads();
  if (tid == 0) {
    *t_out = tnew;
    *dt_out = dt;
    if (o.hist) {
      long long *cnt = reinterpret_cast<long long *>(o.hist_count);
      const long long row = *cnt % o.hist_cap;
      T *hrow = reinterpret_cast<T *>(o.hist) + row * 4;
      hrow[0] = tnew;
      hrow[1] = sumc[0];
      hrow[2] = sumc[1];
      hrow[3] = sumc[2];
      *cnt = *cnt + 1;
    }
    if (d.thermal) {
      const int nt = d.n_temp;
      if (o.heat) for (int b = 0; b < nt; ++b) reinterpret_cast<T *>(o.heat)[b] = (T)res[b];
      if (o.wall_heat) for (int q = 0; q < 4; ++q) reinterpret_cast<T *>(o.wall_heat)[q] = (T)res[nt + q];
      if (o.heat_storage) reinterpret_cast<T *>(o.heat_storage)[0] = (T)res[nt + 4];
      if (o.flow_heat) {
        reinterpret_cast<T *>(o.flow_heat)[0] = (T)res[nt + 5];
        reinterpret_cast<T *>(o.flow_heat)[1] = (T)res[nt + 6];
      }
      if (o.aux) {
        long long *cnt = reinterpret_cast<long long *>(o.aux_count);
        const long long row = *cnt % o.aux_cap;
        const int ncol = 1 + nt + 7;
        T *arow = reinterpret_cast<T *>(o.aux) + row * ncol;
        arow[0] = tnew;
        for (int q = 0; q < nt + 7; ++q) arow[1 + q] = (T)res[q];
        *cnt = *cnt + 1;
      }
    }
  }
}

}  // namespace ns2d
}  // namespace aero

// ====================================================================== host side
using namespace aero;
using namespace aero::ns2d;

namespace {

bool factor_radices(int n, FftPlan &pl) {
  pl.n = n;
  pl.nrad = 0;
  int m = n;
  while (m % 4 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 4; m /= 4; }
  while (m % 2 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 2; m /= 2; }
  while (m % 3 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 3; m /= 3; }
  while (m % 5 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 5; m /= 5; }
  return m == 1;
}

}  // namespace

struct aero2d_plan {
  aero2d_desc desc;
  int dbl;
  size_t esize;
  size_t bytes = 0;
  std::vector<void *> allocs;
  // scratch
  void *uA, *vA, *uB, *vB, *thA, *thB, *nut, *F, *psi, *acc_u, *acc_v, *acc_t, *outcol, *dye_tmp;
  void *tw_x, *tw_y, *cs_x, *cs_y;
  double *lam_x, *lam_y, *partial, *partial2;
  void *red;
  FftPlan fx, fy;
  int stage_blocks;
  int nslotB;
};

namespace {

int alloc(aero2d_plan *p, void **ptr, size_t n) {
  if (n == 0) n = 16;
  hipError_t e = hipMalloc(ptr, n);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return AERO_EHIP;
  }
  p->allocs.push_back(*ptr);
  p->bytes += n;
  e = hipMemset(*ptr, 0, n);
  if (e != hipSuccess) {
    set_error(std::string("hipMemset: ") + hipGetErrorString(e));
    return AERO_EHIP;
  }
  return AERO_OK;
}

template <typename T>
int upload_tables(aero2d_plan *p) {
  const aero2d_desc &d = p->desc;
  auto put = [&](void *dst, const void *src, size_t n) -> int {
    AERO_HIP_CHECK(hipMemcpy(dst, src, n, hipMemcpyHostToDevice));
    return AERO_OK;
  };
  for (int axis = 0; axis < 2; ++axis) {
    const int n = axis == 0 ? d.nx : d.ny;
    std::vector<cplx<T>> tw(n), cs(n);
    std::vector<double> lam(n);
    for (int m = 0; m < n; ++m) {
      const double ang = -2.0 * M_PI * (double)m / (double)n;
      tw[m] = {(T)std::cos(ang), (T)std::sin(ang)};
      const double th = M_PI * (double)m / (2.0 * (double)n);
      cs[m] = {(T)std::cos(th), (T)std::sin(th)};
      const double hh = d.h * d.h;
      lam[m] = d.periodic ? (2.0 * std::cos(2.0 * M_PI * (double)m / (double)n) - 2.0) / hh
                          : (2.0 * std::cos(M_PI * (double)m / (double)n) - 2.0) / hh;
    }
    int rc;
    if ((rc = put(axis == 0 ? p->tw_x : p->tw_y, tw.data(), n * sizeof(cplx<T>)))) return rc;
    if ((rc = put(axis == 0 ? p->cs_x : p->cs_y, cs.data(), n * sizeof(cplx<T>)))) return rc;
    if ((rc = put(axis == 0 ? (void *)p->lam_x : (void *)p->lam_y, lam.data(), n * sizeof(double)))) return rc;
  }
  return AERO_OK;
}

template <typename T>
Dev<T> make_dev(const aero2d_plan *p, const aero2d_geom *g, const aero2d_params *pr) {
  const aero2d_desc &s = p->desc;
  Dev<T> d;
  std::memset(&d, 0, sizeof(d));
  d.nx = s.nx;
  d.ny = s.ny;
  d.periodic = s.periodic;
  d.closed = s.closed;
  d.upwind = s.advection == 1;
  d.turb = s.turbulence;
  d.les = s.turbulence != 0;
  d.bcb = s.bc_bottom;
  d.bct = s.bc_top;
  d.bcl = s.bc_left;
  d.bcr = s.bc_right;
  d.j0 = s.win_j0;
  d.j1 = s.win_j1;
  d.i0 = s.win_i0;
  d.i1 = s.win_i1;
  d.win = (s.win_j1 > s.win_j0 && s.win_i1 > s.win_i0) ? 1 : 0;
  d.wy = d.win ? s.win_j1 - s.win_j0 : 0;
  d.wx = d.win ? s.win_i1 - s.win_i0 : 0;
  d.nbody = s.nbody;
  d.thermal = s.thermal;
  d.buoyant = s.buoyant;
  d.n_temp = s.n_temp;
  d.has_heat = s.has_heat;
  d.insulated = s.insulated;
  for (int q = 0; q < 4; ++q) d.wm[q] = s.wall_modes[q];
  d.dye_mode = s.dye_mode;
  d.n_dye = s.n_dye;
  d.n_groups = s.n_groups;
  for (int q = 0; q < AERO2D_MAX_GROUPS; ++q) d.gch[q] = s.group_channel[q];
  d.fixed_dt = s.fixed_dt;
  d.hd = s.h;
  d.cfl = s.cfl;
  d.force_ref = s.force_ref;
  d.prt = s.prt;
  d.diff_coef = s.diff_coef;
  d.les_c = s.les_c;
  d.gx = s.gravity_x;
  d.gy = s.gravity_y;
  d.h = (T)s.h;
  d.inv_h = (T)(1.0 / s.h);
  d.kadv = (T)((s.advection == 1 ? 7.0 / 24.0 : 0.25) / s.h);
  d.inv_h2 = (T)(1.0 / (s.h * s.h));
  d.ground = (T)s.ground_speed;
  if (g) {
    d.chi_u = (const T *)g->chi_u; d.chi_v = (const T *)g->chi_v; d.chi_c = (const T *)g->chi_c;
    d.a_u = (const T *)g->a_u; d.a_v = (const T *)g->a_v; d.k_u = (const T *)g->k_u; d.k_v = (const T *)g->k_v;
    d.chi_t = (const T *)g->chi_t; d.a_t = (const T *)g->a_t; d.w_t = (const T *)g->w_t;
    d.vol_u = (const T *)g->vol_u; d.vol_v = (const T *)g->vol_v;
    d.ins_u = (const T *)g->ins_u; d.ins_v = (const T *)g->ins_v;
  }
  if (pr) {
    d.nu = (const T *)pr->nu; d.eta = (const T *)pr->eta; d.u_inlet = (const T *)pr->u_inlet;
    d.dt_fixed = (const T *)pr->dt_fixed; d.kappa = (const T *)pr->kappa; d.ri = (const T *)pr->ri;
    d.eta_t = (const T *)pr->eta_t; d.theta_in = (const T *)pr->theta_in;
    d.wall_values = (const T *)pr->wall_values; d.heat = (const T *)pr->heat; d.heat_max = (const T *)pr->heat_max;
    d.dye_inflow = (const T *)pr->dye_inflow; d.g_weight = (const T *)pr->group_weight;
    d.g_target = (const T *)pr->group_target; d.g_timing = (const T *)pr->group_timing;
  }
  d.nut = (T *)p->nut; d.F = (T *)p->F; d.psi = (T *)p->psi;
  d.acc_u = (T *)p->acc_u; d.acc_v = (T *)p->acc_v; d.acc_t = (T *)p->acc_t; d.outcol = (T *)p->outcol;
  d.partial = p->partial; d.partial2 = p->partial2;
  d.red = (Red<T> *)p->red;
  d.tw_x = (const cplx<T> *)p->tw_x; d.tw_y = (const cplx<T> *)p->tw_y;
  d.cs_x = (const cplx<T> *)p->cs_x; d.cs_y = (const cplx<T> *)p->cs_y;
  d.lam_x = p->lam_x; d.lam_y = p->lam_y;
  d.fx = p->fx; d.fy = p->fy;
  d.dct_x = s.periodic ? 0 : 1;
  d.dct_y = s.periodic ? 0 : 1;
  return d;
}

template <typename T>
size_t lds_A(const aero2d_plan *p) { return 2 * p->desc.nx * sizeof(cplx<T>) + 2 * p->desc.nx * sizeof(T) + 256 * sizeof(T); }
template <typename T>
size_t lds_B(const aero2d_plan *p) { return p->nslotB * p->desc.ny * (2 * sizeof(cplx<T>) + 2 * sizeof(T)); }

// divergence / rhs -> psi (passes A, B, C) on `stream`
template <typename T>
int solve_passes(aero2d_plan *p, hipStream_t st, Dev<T> &d, const T *u, const T *v, const T *f, int mode, T *psi,
                 T *pacc, int stage) {
  const int ny = p->desc.ny, nx = p->desc.nx;
  hipLaunchKernelGGL(k_poisA<T>, dim3((ny + 1) / 2), dim3(kFftThreads), lds_A<T>(p), st, d, u, v, f, mode);
  const int cols_per_block = 2 * p->nslotB;
  hipLaunchKernelGGL(k_poisB<T>, dim3((nx + cols_per_block - 1) / cols_per_block), dim3(kFftThreads), lds_B<T>(p), st,
                     d, p->nslotB);
  hipLaunchKernelGGL(k_poisC<T>, dim3((ny + 1) / 2), dim3(kFftThreads), lds_A<T>(p), st, d, psi, pacc, stage);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int step_impl(aero2d_plan *p, hipStream_t st, const aero2d_fields *in, aero2d_fields *out, const aero2d_geom *g,
              const aero2d_params *pr, aero2d_outputs *o) {
  const aero2d_desc &s = p->desc;
  Dev<T> d = make_dev<T>(p, g, pr);
  const int nx = s.nx, ny = s.ny;
  const dim3 blk(64, 4);
  const dim3 grid_faces((nx + 1 + 63) / 64, (ny + 1 + 3) / 4);
  const dim3 grid_nut((nx + 2 + 63) / 64, (ny + 2 + 3) / 4);
  const T *u0 = (const T *)in->u, *v0 = (const T *)in->v, *th0 = (const T *)in->theta;
  // state reductions -> dt
  if (d.les) hipLaunchKernelGGL(k_nut<T>, grid_nut, blk, 0, st, d, u0, v0, 1);
  const int nmax = std::max(ny * (nx + 1), (ny + 1) * nx);
  const int rb = std::min(kRedBlocks, (nmax + 255) / 256);
  hipLaunchKernelGGL(k_reduce0<T>, dim3(rb), dim3(256), 0, st, d, u0, v0, th0);
  T *ubuf[3] = {(T *)p->uA, (T *)p->uB, (T *)out->u};
  T *vbuf[3] = {(T *)p->vA, (T *)p->vB, (T *)out->v};
  T *tbuf[3] = {(T *)p->thA, (T *)p->thB, (T *)out->theta};
  const T *up = u0, *vp = v0, *thp = th0;
  for (int k = 0; k < 3; ++k) {
    if (d.les && k > 0) hipLaunchKernelGGL(k_nut<T>, grid_nut, blk, 0, st, d, up, vp, 0);
    Stage<T> sg;
    sg.s = k;
    sg.a = kSsp[k][0];
    sg.b = kSsp[k][1];
    sg.w = kForceW[k];
    sg.up = up; sg.vp = vp; sg.thp = thp;
    sg.u0 = u0; sg.v0 = v0; sg.th0 = th0;
    sg.uh = ubuf[k]; sg.vh = vbuf[k]; sg.thh = s.thermal ? tbuf[k] : nullptr;
    hipLaunchKernelGGL(k_stage<T>, grid_faces, blk, 0, st, d, sg);
    int rc = solve_passes<T>(p, st, d, sg.uh, sg.vh, nullptr, (s.periodic || s.closed) ? 1 : 2, d.psi, (T *)out->p, k);
    if (rc) return rc;
    hipLaunchKernelGGL(k_correct<T>, grid_faces, blk, 0, st, d, sg.uh, sg.vh, (const T *)d.psi,
                       (s.periodic || s.closed) ? 0 : 1);
    up = sg.uh; vp = sg.vh; thp = sg.thh;
  }
  // smoke with the end-of-step velocity
  if (s.dye_mode != 0 && in->dye && out->dye) {
    const size_t n = (size_t)s.n_dye * ny * nx;
    const bool inplace = in->dye == out->dye;
    T *dst = inplace ? (T *)p->dye_tmp : (T *)out->dye;
    hipLaunchKernelGGL(k_dye<T>, dim3((nx + 63) / 64, (ny + 3) / 4, s.n_dye), blk, 0, st, d, (const T *)in->dye, dst,
                       (const T *)out->u, (const T *)out->v, (const T *)in->t);
    if (inplace) {
      const int cb = (int)std::min<size_t>(4096, (n + 255) / 256);
      hipLaunchKernelGGL(k_copy<T>, dim3(cb), dim3(256), 0, st, (T *)out->dye, (const T *)dst, n);
    }
  }
  aero2d_outputs oo;
  std::memset(&oo, 0, sizeof(oo));
  if (o) oo = *o;
  hipLaunchKernelGGL(k_finalize<T>, dim3(1), dim3(1024), 0, st, d, (const T *)in->t, (T *)out->t, (T *)out->dt,
                     (const T *)out->u, (const T *)out->theta, oo, (int)(grid_faces.x * grid_faces.y));
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int project_impl(aero2d_plan *p, hipStream_t st, void *u, void *v, void *psi_out) {
  Dev<T> d = make_dev<T>(p, nullptr, nullptr);
  T *psi = psi_out ? (T *)psi_out : d.psi;
  int rc = solve_passes<T>(p, st, d, (const T *)u, (const T *)v, nullptr, 1, psi, nullptr, 0);
  if (rc) return rc;
  const dim3 blk(64, 4);
  const dim3 grid_faces((p->desc.nx + 1 + 63) / 64, (p->desc.ny + 1 + 3) / 4);
  hipLaunchKernelGGL(k_correct<T>, grid_faces, blk, 0, st, d, (T *)u, (T *)v, (const T *)psi, 0);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int poisson_impl(aero2d_plan *p, hipStream_t st, const void *f, void *psi) {
  Dev<T> d = make_dev<T>(p, nullptr, nullptr);
  return solve_passes<T>(p, st, d, nullptr, nullptr, (const T *)f, 0, (T *)psi, nullptr, 0);
}

int check_desc(const aero2d_desc *d) {
  if (!d) { set_error("null descriptor"); return AERO_EINVAL; }
  if (d->nx < 4 || d->ny < 4) { set_error("grid needs at least 4 x 4 cells"); return AERO_EINVAL; }
  if (d->dtype != 0 && d->dtype != 1) { set_error("dtype must be 0 (float32) or 1 (float64)"); return AERO_EINVAL; }
  if (!(d->h > 0)) { set_error("h must be positive"); return AERO_EINVAL; }
  if (d->turbulence < 0 || d->turbulence > 3) { set_error("unknown turbulence model"); return AERO_EINVAL; }
  if (d->advection != 0 && d->advection != 1) { set_error("unknown advection scheme"); return AERO_EINVAL; }
  if (d->periodic && d->closed) { set_error("periodic and closed are exclusive"); return AERO_EINVAL; }
  const size_t es = d->dtype ? 8 : 4;
  const size_t ldsx = 2 * d->nx * 2 * es + 2 * d->nx * es + 256 * es;
  const size_t ldsy = d->ny * (2 * 2 * es + 2 * es);
  if (ldsx > 65536 || ldsy > 65536) { set_error("grid too large for the LDS-resident transforms"); return AERO_EUNSUPPORTED; }
  if (d->n_groups > AERO2D_MAX_GROUPS || d->n_groups < 0) { set_error("too many emitter groups"); return AERO_EUNSUPPORTED; }
  if (d->dye_mode < 0 || d->dye_mode > 2) { set_error("unknown dye mode"); return AERO_EINVAL; }
  const bool win = d->win_j1 > d->win_j0 && d->win_i1 > d->win_i0;
  if (win && (d->win_j0 < 0 || d->win_i0 < 0 || d->win_j1 > d->ny || d->win_i1 > d->nx)) {
    set_error("body window outside the grid");
    return AERO_EINVAL;
  }
  if (d->n_temp > 58) { set_error("too many temperature bodies"); return AERO_EUNSUPPORTED; }
  return AERO_OK;
}

}  // namespace

extern "C" {

const char *aero_last_error(void) { return g_error.c_str(); }
int aero_version(void) { return 1; }

int aero2d_plan_create(const aero2d_desc *desc, aero2d_plan **out) {
  if (!out) { set_error("null plan pointer"); return AERO_EINVAL; }
  *out = nullptr;
  int rc = check_desc(desc);
  if (rc) return rc;
  aero2d_plan *p = new aero2d_plan();
  p->desc = *desc;
  p->dbl = desc->dtype == 1;
  p->esize = p->dbl ? 8 : 4;
  if (!factor_radices(desc->nx, p->fx) || !factor_radices(desc->ny, p->fy)) {
    delete p;
    set_error("nx and ny must be products of 2, 3 and 5");
    return AERO_EUNSUPPORTED;
  }
  const size_t es = p->esize;
  const size_t nx = desc->nx, ny = desc->ny;
  const size_t nu_ = ny * (nx + 1), nv_ = (ny + 1) * nx, nc = ny * nx;
  const bool win = desc->win_j1 > desc->win_j0 && desc->win_i1 > desc->win_i0;
  const size_t wy = win ? desc->win_j1 - desc->win_j0 : 0, wx = win ? desc->win_i1 - desc->win_i0 : 0;
  const size_t blocks = ((nx + 1 + 63) / 64) * ((ny + 1 + 3) / 4);
  p->stage_blocks = (int)blocks;
  p->nslotB = 2;
  while (p->nslotB > 1 && ny * p->nslotB * (2 * 2 * es + 2 * es) > 65536) p->nslotB--;
  struct { void **ptr; size_t n; } tab[] = {
      {&p->uA, nu_ * es}, {&p->vA, nv_ * es}, {&p->uB, nu_ * es}, {&p->vB, nv_ * es},
      {&p->thA, desc->thermal ? nc * es : 0}, {&p->thB, desc->thermal ? nc * es : 0},
      {&p->nut, desc->turbulence ? (ny + 2) * (nx + 2) * es : 0}, {&p->F, nc * es}, {&p->psi, nc * es},
      {&p->acc_u, wy * (wx + 1) * es}, {&p->acc_v, (wy + 1) * wx * es}, {&p->acc_t, wy * wx * es},
      {&p->outcol, ny * es}, {&p->dye_tmp, desc->dye_mode ? (size_t)desc->n_dye * nc * es : 0},
      {&p->tw_x, nx * 2 * es}, {&p->tw_y, ny * 2 * es}, {&p->cs_x, nx * 2 * es}, {&p->cs_y, ny * 2 * es},
      {(void **)&p->lam_x, nx * 8}, {(void **)&p->lam_y, ny * 8}, {(void **)&p->partial, kRedBlocks * 8},
      {(void **)&p->partial2, blocks * 8}, {&p->red, 256}};
  for (auto &t : tab) {
    if ((rc = alloc(p, t.ptr, t.n))) {
      aero2d_plan_destroy(p);
      return rc;
    }
  }
  rc = p->dbl ? upload_tables<double>(p) : upload_tables<float>(p);
  if (rc) {
    aero2d_plan_destroy(p);
    return rc;
  }
  *out = p;
  return AERO_OK;
}

int aero2d_plan_destroy(aero2d_plan *p) {
  if (!p) return AERO_OK;
  for (void *a : p->allocs) (void)hipFree(a);
  delete p;
  return AERO_OK;
}

size_t aero2d_plan_workspace_bytes(const aero2d_plan *p) { return p ? p->bytes : 0; }

int aero2d_step(aero2d_plan *p, void *stream, const aero2d_fields *in, aero2d_fields *out, const aero2d_geom *g,
                const aero2d_params *pr, aero2d_outputs *o) {
  if (!p || !in || !out || !pr) { set_error("null argument"); return AERO_EINVAL; }
  if (!in->u || !in->v || !out->u || !out->v || !out->p || !in->t || !out->t || !out->dt || !pr->nu || !pr->eta) {
    set_error("missing field or parameter pointer");
    return AERO_EINVAL;
  }
  const aero2d_desc &s = p->desc;
  if (!s.periodic && !s.closed && !pr->u_inlet) { set_error("tunnel needs u_inlet"); return AERO_EINVAL; }
  if (s.fixed_dt && !pr->dt_fixed) { set_error("fixed dt needs dt_fixed"); return AERO_EINVAL; }
  const bool win = s.win_j1 > s.win_j0 && s.win_i1 > s.win_i0;
  if (win && (!g || !g->chi_u || !g->chi_v || !g->chi_c || !g->a_u || !g->a_v || !g->k_u || !g->k_v)) {
    set_error("body window set but geometry pointers missing");
    return AERO_EINVAL;
  }
  if (s.thermal && (!in->theta || !out->theta || !pr->kappa || !pr->ri || !pr->eta_t || !pr->theta_in ||
                    !pr->wall_values)) {
    set_error("thermal solve needs theta and thermal parameters");
    return AERO_EINVAL;
  }
  if (s.thermal && win && s.n_temp > 0 && (!g->chi_t || !g->a_t || !g->w_t)) {
    set_error("temperature bodies need chi_t, a_t, w_t");
    return AERO_EINVAL;
  }
  if (s.thermal && s.insulated && win && (!g->ins_u || !g->ins_v)) {
    set_error("insulated window needs ins_u, ins_v");
    return AERO_EINVAL;
  }
  if (s.has_heat && (!pr->heat || !pr->heat_max)) { set_error("heaters need heat and heat_max"); return AERO_EINVAL; }
  if (s.dye_mode == 2 && s.n_groups > 0 && (!pr->group_weight || !pr->group_target || !pr->group_timing)) {
    set_error("emitter groups need weight, target and timing");
    return AERO_EINVAL;
  }
  if (o && o->hist && (!o->hist_count || o->hist_cap <= 0)) { set_error("history ring needs count and capacity"); return AERO_EINVAL; }
  if (o && o->aux && (!o->aux_count || o->aux_cap <= 0)) { set_error("aux ring needs count and capacity"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o) : step_impl<float>(p, st, in, out, g, pr, o);
}

int aero2d_project(aero2d_plan *p, void *stream, void *u, void *v, void *psi) {
  if (!p || !u || !v) { set_error("null argument"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? project_impl<double>(p, st, u, v, psi) : project_impl<float>(p, st, u, v, psi);
}

int aero2d_poisson_solve(aero2d_plan *p, void *stream, const void *f, void *psi) {
  if (!p || !f || !psi) { set_error("null argument"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? poisson_impl<double>(p, st, f, psi) : poisson_impl<float>(p, st, f, psi);
}

}  // extern "C"
