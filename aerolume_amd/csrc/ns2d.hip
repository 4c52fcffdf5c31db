// MI355X (gfx950) kernels of the 2D incompressible Navier-Stokes step and its C ABI.
//
// One SSP-RK3 step (reference solvers/ns2d/core.py:815-969) is a fixed sequence of
// launches on the caller's stream, with no host synchronisation and no allocation:
//
//   [pre, dt]              dt inputs from the state: max |u|, |v|, |theta|, nu_t and
//                          sum theta (skipped in chained mode: the previous step's last two
//                          kernels left them) - core.py:685-695, physics.py:851-888
//   3 x stage              64x8-cell LDS tiles: the previous stage's projection applied on
//                          the fly (u = u^ - grad psi, outlet mass balance), eddy viscosity
//                          on the tile + 1 ring, upwind3/central advection, diffusion, LES
//                          stress, RK update, inlet/outlet, window penalisation with force
//                          increments, buoyancy, temperature stage (core.py:571-614, 855-935)
//       poisA              divergence (outlet correction folded in) -> DCT/DHT along x
//       poisB              DCT/DHT along y, x 1/lambda (FP64-built), inverse along y
//       poisC              inverse along x -> psi; step-averaged pressure (core.py:935-942)
//   final                  u = u^ - grad psi written to the state, smoke transport with it,
//                          per-block partials of the next step's dt inputs, forces, heat
//   finalize               one block: forces, heat rates, t, dt, next dt, history rings
//
// All arithmetic is in the plan's dtype T (float or double); the Poisson eigenvalue
// inverse is formed in double, and the heat-balance sums accumulate in double. No
// atomics: every global reduction goes through per-block partials summed in fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/aerolume_amd.h"
#include "common.h"

namespace aero {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

namespace ns2d {

constexpr int kFftThreads = 256;          // threads of the batched transform passes
constexpr int kMaxLen = 4096;              // longest transform
constexpr int kRowFill = 1024;             // row elements per batched row-pass block when filled
constexpr int TX = 64;                      // tile width: one wave across (blocks of 64 x 4 threads)
constexpr int HW = TX + 4, NW = TX + 2;     // halo-2 rows of u, v, scalar; eddy-viscosity rows (halo 1)
constexpr int kDefaultRows = 4;             // tile height unless a plan or an ensemble picks another
// A tile of R rows (R = 4, 8, 16; each thread owns R / 4 rows): halo-2 tiles of R + 4 rows,
// eddy-viscosity tile of R + 2 rows, tile elements per thread of the loaders.
template <int R>
struct TR {
  static constexpr int HH = R + 4, NH = R + 2, NE = (HH * HW + 255) / 256, RPT = R / 4;
};
constexpr int kMaxWin = 16;                 // force/heat partial slots per block (4 per body + temps)
constexpr double kSsp[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
constexpr double kForceW[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0};

// Reduction slots of one plan (device memory, zeroed at creation).
template <typename T>
struct Red {
  T dt;          // this step's dt
  T corr;        // outlet correction of the current stage
  double th_old; // sum of theta at step start
  T ctrl[2];     // PID step: the controller memory after this step's dt (the state's after the step)
};

// Everything a kernel needs, passed by value.
template <typename T>
struct Dev {
  int nx, ny, nbx, nby;
  int ty;              // tile rows of the tile kernels (pmax / pforce block index = by * nbx + bx)
  int periodic, closed, upwind, turb, les;
  int bcb, bct, bcl, bcr;
  int j0, j1, i0, i1, wx, wy, win;
  int nbody, thermal, buoyant, n_temp, has_heat, insulated, n_flux;
  int wm[4];
  int dye_mode, n_dye, n_groups;
  int gch[AERO2D_MAX_GROUPS];
  int fixed_dt;
  int wallm;         // Werner-Wengle wall shear on
  int pid;           // adaptive_dt = "pid"
  gptr<const T> dt_in;    // the step-start state's dt and controller memory (PID step)
  gptr<const T> ctrl_in;
  gptr<T> ctrl_out;       // the end-of-step state's controller memory
  double hd, cfl, force_ref, prt, diff_coef, les_c, gx, gy;
  T h, inv_h, kadv, inv_h2, ground;
  // double-built constants of the per-cell paths, rounded once (the values the kernels formed
  // in double and cast): 1/Pr_t, 2/h, Cs^2 h^2 (Smagorinsky, WALE), Cs h^2 (Vreman)
  T inv_prt, two_h, les_k2, les_kv;
  T wall_y;          // distance of the first face row from a wall, h / 2
  // geometry
  gptr<const T> chi_u, chi_v, chi_c, a_u, a_v, k_u, k_v;
  gptr<const T> chi_t, a_t, w_t, vol_u, vol_v, ins_u, ins_v;
  gptr<const T> q_skin, w_f;   // heat-flux bodies: skin source per unit kappa, skin weights
  // params
  gptr<const T> nu, eta, u_inlet, dt_fixed, kappa, ri, eta_t, theta_in, wall_values, heat, heat_max;
  gptr<const T> dye_inflow, g_weight, g_target, g_timing;
  // scratch
  gptr<T> F, psi, acc_u, acc_v, acc_t, outcol;
  gptr<double> pmax;    // per block: max |u|, |v|, |theta|, nu_t, sum theta
  gptr<double> pforce;  // per block: 4 force sums per body, heat per temperature body
  gptr<Red<T>> red;
  gptr<const cplx<T>> tw_x, tw_y, cs_x, cs_y;
  gptr<const double> lam_x, lam_y;
  gptr<const T> inv_lam;   // (nx, ny), y fastest: 1 / (lam_x + lam_y) formed in float64, 0 at the zero mode
  FftPlan fx, fy;
  int dct_x, dct_y;   // 1: Neumann (DCT-II), 0: periodic (DHT)
};


// Compile-time views of a plan for the tile kernels. DevK<T, LES, TH, PLAIN> is a Dev<T> whose
// static members hide some of the plan's run-time flags: LES fixes the eddy-viscosity model
// (0 none, 1 Smagorinsky, 2 WALE, 3 Vreman), TH heat transfer (0 off, 1 on), and PLAIN = 1 a
// wind tunnel (inflow / outflow columns) with upwind3 advection and none of the rarer features
// (wall model, heat-flux skins, heaters); LES = -1, TH = -1 or PLAIN = 0 leave
// the flag to the plan. The tile kernels are written once against the members of a Dev, so
// each view compiles to a kernel with only the code, plan fields and registers of its case.
template <typename T, int LES>
struct DevL : Dev<T> {
  static constexpr int turb = LES, les = LES != 0;
};
template <typename T>
struct DevL<T, -1> : Dev<T> {};
template <typename T, int LES, int TH>
struct DevLT : DevL<T, LES> {
  static constexpr int thermal = TH;
};
template <typename T, int LES>
struct DevLT<T, LES, -1> : DevL<T, LES> {};
template <typename T, int LES, int TH, int PLAIN>
struct DevK : DevLT<T, LES, TH> {
  static constexpr int periodic = 0, closed = 0, upwind = 1, wallm = 0, n_flux = 0, has_heat = 0;
};
template <typename T, int LES, int TH>
struct DevK<T, LES, TH, 0> : DevLT<T, LES, TH> {};

// a plan view of d (the same values; the view's static members replace some flags)
template <typename D, typename T>
__host__ __device__ D view_of(const Dev<T> &d) {
  D k;
  static_cast<Dev<T> &>(k) = d;
  return k;
}

template <typename T>
struct Stage {
  int s;
  double a, b, w;
  gptr<const T> up, vp, psi_p;  // stage input velocity (+ the previous stage's potential)
  gptr<const T> thp;              // stage input temperature
  gptr<const T> u0, v0, th0;    // step start
  gptr<T> uh, vh, thh;          // stage output (before projection)
};

template <typename T, int R>
struct Lds {
  T u[R + 4][HW];
  T v[R + 4][HW];
  T c[R + 4][HW];
  T nu[R + 2][NW];
  double red[4 * kMaxWin];
};

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

template <typename T>
__device__ __forceinline__ T wall_ghost(int kind, T val, T speed) {
  if (kind == AERO_BC_SLIP) return val;
  if (kind == AERO_BC_NOSLIP) return -val;
  return (T)2 * speed - val;
}

// ------------------------------------------------------------------ in-domain reads
// u at face (j, i) of the stage input: u^ - grad psi when psi is given (core.py:645-660),
// with the outlet face shifted by the mass correction (core.py:337-357).
template <typename T>
__device__ __forceinline__ T u_in(const Dev<T> &d, const nd<T> *__restrict__ u, const nd<T> *__restrict__ psi, T corr, int j,
                                  int i) {
  const int nx = d.nx;
  T val = u[j * (nx + 1) + i];
  if (psi) {
    if (d.periodic) {
      const int ie = i == nx ? 0 : i;
      val = val - (psi[j * nx + ie] - psi[j * nx + wrap(ie - 1, nx)]) * d.inv_h;
    } else if (i > 0 && i < nx) {
      val = val - (psi[j * nx + i] - psi[j * nx + i - 1]) * d.inv_h;
    } else if (i == nx && !d.closed) {
      val = val + corr;
    }
  }
  return val;
}
template <typename T>
__device__ __forceinline__ T v_in(const Dev<T> &d, const nd<T> *__restrict__ v, const nd<T> *__restrict__ psi, int j, int i) {
  const int nx = d.nx, ny = d.ny;
  T val = v[j * nx + i];
  if (psi) {
    if (d.periodic) {
      const int je = j == ny ? 0 : j;
      val = val - (psi[je * nx + i] - psi[wrap(je - 1, ny) * nx + i]) * d.inv_h;
    } else if (j > 0 && j < ny) {
      val = val - (psi[j * nx + i] - psi[(j - 1) * nx + i]) * d.inv_h;
    }
  }
  return val;
}

// ------------------------------------------------------------------ tile loads
// The tile loaders map every tile position to an in-domain address and a transform (the
// ghost rules of core.py:287-325 and scalar.py:42-60), load, then apply the transform.

// flag bits: 1 gradient correction, 2 outlet correction, 4 negate (mirror), 8/16/24 wall
// kind slip/noslip/moving (u rows outside the domain)
template <typename T>
__device__ __forceinline__ T finish_u(const Dev<T> &d, T x, T q1, T q0, T corr, int flag, int is_bottom) {
  if (flag & 1) x = x - (q1 - q0) * d.inv_h;
  else if (flag & 2) x = x + corr;
  if (flag & 4) x = -x;
  const int wk = flag >> 3;   // 0 interior row, else kind + 1
  if (wk == 2) x = -x;
  else if (wk == 3) x = (T)2 * (is_bottom ? d.ground : (T)0) - x;
  return x;
}
template <typename T>
__device__ __forceinline__ T finish_v(const Dev<T> &d, T x, T q1, T q0, int flag) {
  if (flag & 1) x = x - (q1 - q0) * d.inv_h;
  if (flag & 4) x = -x;
  return x;
}

// U, V tiles of the stage-input velocity around the tile at (tj0, ti0). The ghost rules split
// into a column part (inlet / outlet columns, closed-box mirrors) and a row part (wall rows,
// periodic wrap): each lane works out the column part of its tile columns once (lane c and
// c + 64), each wave the row part of its rows, and every element then costs its loads and
// finish_u / finish_v.
template <typename T, int R, typename D>
__device__ __forceinline__ void load_uv(const D &d, Lds<T, R> &L, int tj0, int ti0, const nd<T> *u, const nd<T> *v, const nd<T> *psi, T corr) {
  const int nx = d.nx, ny = d.ny;
  const bool hp = psi != nullptr;
  const int lane = threadIdx.x, wave = threadIdx.y;
  constexpr int NC = (HW + 63) / 64;   // 64-wide passes over a tile row
  // column part, per lane and pass: u - source column, psi columns, flags (1 gradient, 2 outlet
  // shift, 4 negate); v - source column, negate flag
  int ucol[NC], upc1[NC], upc0[NC], uflag[NC], vcol[NC], vneg[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int i = ti0 - 2 + lane + 64 * q;
    int ii, f = 0;
    if (d.periodic) {
      ii = clampi(wrap(i, nx), 0, nx);
      const int ie = ii == nx ? 0 : ii;
      upc1[q] = ie;
      upc0[q] = wrap(ie - 1, nx);
      if (hp) f |= 1;
    } else {
      if (i < 0) {
        if (d.closed) { ii = -i; f |= 4; } else ii = 0;
      } else if (i > nx) {
        if (d.closed) { ii = 2 * nx - i; f |= 4; } else ii = nx;
      } else {
        ii = i;
      }
      ii = clampi(ii, 0, nx);
      upc1[q] = ii < nx ? ii : nx - 1;
      upc0[q] = ii > 0 ? upc1[q] - 1 : upc1[q];
      if (hp) {
        if (ii > 0 && ii < nx) f |= 1;
        else if (ii == nx && !d.closed) f |= 2;
      }
    }
    ucol[q] = ii;
    uflag[q] = f;
    int vi;
    bool neg = false;
    if (d.periodic) {
      vi = clampi(wrap(i, nx), 0, nx - 1);
    } else {
      if (i < 0) {
        vi = -1 - i;
        if (!(d.closed && d.bcl != AERO_BC_NOSLIP)) neg = !neg;
      } else if (i >= nx) {
        if (d.closed) {
          vi = 2 * nx - 1 - i;
          if (d.bcr == AERO_BC_NOSLIP) neg = !neg;
        } else {
          vi = nx - 1;
        }
      } else {
        vi = i;
      }
      vi = clampi(vi, 0, nx - 1);
    }
    vcol[q] = vi;
    vneg[q] = neg ? 1 : 0;
  }
  // all loads of this thread first (independent, one memory round trip), then the transforms
  constexpr int NRW = (TR<R>::HH + 3) / 4;   // tile rows per wave
  T xu[NRW][NC], xv[NRW][NC], a1[NRW][NC], a0[NRW][NC], b1[NRW][NC], b0[NRW][NC];
  int uwall[NRW];
  bool vgrad[NRW], vnr[NRW];
#pragma unroll
  for (int rr = 0; rr < NRW; ++rr) {
    const int r = wave + 4 * rr;
    const int j = tj0 - 2 + r;
    // row part of u: source row, wall kind of ghost rows (flag bits 8 / 16 / 24)
    int uj;
    uwall[rr] = 0;
    if (d.periodic) {
      uj = clampi(wrap(j, ny), 0, ny - 1);
    } else if (j < 0) {
      uj = clampi(-1 - j, 0, ny - 1);
      uwall[rr] = 8 * (d.bcb + 1);
    } else if (j >= ny) {
      uj = clampi(2 * ny - 1 - j, 0, ny - 1);
      uwall[rr] = 8 * (d.bct + 1);
    } else {
      uj = j;
    }
    // row part of v: source row, psi rows, gradient flag, negation of mirrored rows
    int vj, vp1, vp0;
    vgrad[rr] = false;
    vnr[rr] = false;
    if (d.periodic) {
      vj = clampi(wrap(j, ny), 0, ny);
      const int je = vj == ny ? 0 : vj;
      vp1 = je;
      vp0 = wrap(je - 1, ny);
      vgrad[rr] = hp;
    } else {
      if (j < 0) { vj = -j; vnr[rr] = true; } else if (j > ny) { vj = 2 * ny - j; vnr[rr] = true; } else vj = j;
      vj = clampi(vj, 0, ny);
      vp1 = vj < ny ? vj : ny - 1;
      vp0 = vj > 0 ? vp1 - 1 : vp1;
      vgrad[rr] = hp && vj > 0 && vj < ny;
    }
    const T *urow = u + uj * (nx + 1);
    const T *vrow = v + vj * nx;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const bool live = r < TR<R>::HH && lane + 64 * q < HW;
      xu[rr][q] = live ? urow[ucol[q]] : (T)0;
      xv[rr][q] = live ? vrow[vcol[q]] : (T)0;
      a1[rr][q] = a0[rr][q] = b1[rr][q] = b0[rr][q] = (T)0;
      if (hp && live) {
        a1[rr][q] = psi[uj * nx + upc1[q]];
        a0[rr][q] = psi[uj * nx + upc0[q]];
        b1[rr][q] = psi[vp1 * nx + vcol[q]];
        b0[rr][q] = psi[vp0 * nx + vcol[q]];
      }
    }
  }
#pragma unroll
  for (int rr = 0; rr < NRW; ++rr) {
    const int r = wave + 4 * rr;
    if (r >= TR<R>::HH) continue;
    const int j = tj0 - 2 + r;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int c = lane + 64 * q;
      if (c >= HW) continue;
      L.u[r][c] = finish_u(d, xu[rr][q], a1[rr][q], a0[rr][q], corr, uflag[q] | uwall[rr], j < 0);
      const int vf = (vgrad[rr] ? 1 : 0) | ((vneg[q] != 0) != vnr[rr] ? 4 : 0);
      L.v[r][c] = finish_v(d, xv[rr][q], b1[rr][q], b0[rr][q], vf);
    }
  }
}
// scalar tile with the TVD ghost rules of scalar.py:42-60: inflow (or replicate) on the left,
// replicate elsewhere, wrap when periodic; use_inflow 2 is the smoke comb's rule (inflow also
// in a closed box, core.py:718-723)
template <typename T, int R, typename D>
__device__ __forceinline__ void load_c(const D &d, Lds<T, R> &L, int tj0, int ti0, const nd<T> *c, const nd<T> *inflow_col, T inflow0,
                       int use_inflow) {
  const int nx = d.nx, ny = d.ny;
  const int lane = threadIdx.x, wave = threadIdx.y;
  constexpr int NC = (HW + 63) / 64;
  int col[NC];
  bool inf[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int i = ti0 - 2 + lane + 64 * q;
    if (d.periodic) {
      col[q] = clampi(wrap(i, nx), 0, nx - 1);
      inf[q] = false;
    } else {
      col[q] = clampi(i, 0, nx - 1);
      inf[q] = i < 0 && (use_inflow == 2 || (use_inflow == 1 && !d.closed));
    }
  }
  constexpr int NRW = (TR<R>::HH + 3) / 4;
  T x[NRW][NC], fl[NRW];
#pragma unroll
  for (int rr = 0; rr < NRW; ++rr) {
    const int r = wave + 4 * rr;
    const int j = tj0 - 2 + r;
    const int jr = clampi(j, 0, ny - 1);
    const int row = d.periodic ? clampi(wrap(j, ny), 0, ny - 1) : jr;
    fl[rr] = inflow_col ? inflow_col[jr] : inflow0;
    const T *crow = c + row * nx;
#pragma unroll
    for (int q = 0; q < NC; ++q) x[rr][q] = (r < TR<R>::HH && lane + 64 * q < HW) ? crow[col[q]] : (T)0;
  }
#pragma unroll
  for (int rr = 0; rr < NRW; ++rr) {
    const int r = wave + 4 * rr;
    if (r >= TR<R>::HH) continue;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int cc = lane + 64 * q;
      if (cc < HW) L.c[r][cc] = inf[q] ? fl[rr] : x[rr][q];
    }
  }
}

// Eddy viscosity of core.py:444-478 from the gradient tensor.
template <typename T, typename D>
__device__ __forceinline__ T eddy(const D &d, T g11, T g12, T g21, T g22) {
  const T s12 = (T)0.5 * (g12 + g21);
  const T ss = (g11 * g11 + g22 * g22) + (T)2 * s12 * s12;
  if (d.turb == 1) return d.les_k2 * sqrt((T)2 * ss + (T)1e-30);
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
    return d.les_k2 * num / den;
  }
  const T b11 = g11 * g11 + g12 * g12;
  const T b22 = g21 * g21 + g22 * g22;
  const T b12 = g11 * g21 + g12 * g22;
  T bb = b11 * b22 - b12 * b12;
  bb = bb > (T)0 ? bb : (T)0;
  const T aa = ((g11 * g11 + g12 * g12) + g21 * g21) + g22 * g22;
  return d.les_kv * sqrt(bb / (aa + (T)1e-30) + (T)1e-30);
}

// nu_t on the tile's cells and their ring (core.py:421-503), switched off in solids.
template <bool WIN = true, typename T, int R, typename D>
__device__ __forceinline__ void tile_nut(const D &d, Lds<T, R> &L, int tj0, int ti0) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const T ih = d.inv_h;
  for (int e = tid; e < TR<R>::NH * NW; e += 256) {
    const int r = e / NW, c = e - r * NW;
    const int j = tj0 - 1 + r, cc = ti0 - 1 + c;   // cell
    const int ur = r + 1, uc = c + 1;               // u/v tile index of (j, cc)
    const T g11 = (L.u[ur][uc + 1] - L.u[ur][uc]) * ih;
    const T g22 = (L.v[ur + 1][uc] - L.v[ur][uc]) * ih;
    const T du00 = (L.u[ur][uc] - L.u[ur - 1][uc]) * ih;
    const T du10 = (L.u[ur + 1][uc] - L.u[ur][uc]) * ih;
    const T du01 = (L.u[ur][uc + 1] - L.u[ur - 1][uc + 1]) * ih;
    const T du11 = (L.u[ur + 1][uc + 1] - L.u[ur][uc + 1]) * ih;
    const T dv00 = (L.v[ur][uc] - L.v[ur][uc - 1]) * ih;
    const T dv10 = (L.v[ur + 1][uc] - L.v[ur + 1][uc - 1]) * ih;
    const T dv01 = (L.v[ur][uc + 1] - L.v[ur][uc]) * ih;
    const T dv11 = (L.v[ur + 1][uc + 1] - L.v[ur + 1][uc]) * ih;
    const T g12 = (((du00 + du10) + du01) + du11) * (T)0.25;
    const T g21 = (((dv00 + dv10) + dv01) + dv11) * (T)0.25;
    T val = eddy(d, g11, g12, g21, g22);
    if (WIN && d.win && j >= d.j0 && j < d.j1 && cc >= d.i0 && cc < d.i1)
      val = val * ((T)1 - d.chi_c[(j - d.j0) * d.wx + (cc - d.i0)]);
    L.nu[r][c] = val;
  }
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

// torch.lerp(start, end, w)
template <typename T>
__device__ __forceinline__ T lerp_t(T start, T end, T w) {
  return w < (T)0.5 ? start + w * (end - start) : end - (end - start) * ((T)1 - w);
}

template <typename T>
__device__ __forceinline__ T van_leer(T a, T b) {
  const T ab = a * b;
  return ab > (T)0 ? (T)2 * ab / (a + b) : (T)0;
}

// MUSCL / van Leer flux at the x-face between tile cells (r, c-1) and (r, c) (scalar.py:63-83)
// Only the upwind side's reconstruction is formed: the slopes are picked first, so each face
// costs one limiter (one division) and gives the value the two-sided form would select.
template <typename T>
__device__ __forceinline__ T tvd_face(T cm2, T cm1, T c0, T c1, T vel) {
  const bool up = vel > (T)0;
  const T base = up ? cm1 : c0;
  const T lim = van_leer(up ? cm1 - cm2 : c0 - cm1, up ? c0 - cm1 : c1 - c0);
  return vel * (up ? base + (T)0.5 * lim : base - (T)0.5 * lim);
}
template <typename T, int R>
__device__ __forceinline__ T tvd_x(const Lds<T, R> &L, int r, int c, T uf) {
  return tvd_face(L.c[r][c - 2], L.c[r][c - 1], L.c[r][c], L.c[r][c + 1], uf);
}
template <typename T, int R>
__device__ __forceinline__ T tvd_y(const Lds<T, R> &L, int r, int c, T vf) {
  return tvd_face(L.c[r - 2][c], L.c[r - 1][c], L.c[r][c], L.c[r + 1][c], vf);
}

// temperature diffusivity at x-face f of row j / y-face f of column i (physics.py:713-724)
template <bool WIN = true, typename T, int R, typename D>
__device__ __forceinline__ T kappa_x(const D &d, const Lds<T, R> &L, int tj0, int ti0, T kap, int j, int f) {
  T k = kap;
  if (d.les) {
    const T ip = d.inv_prt;
    const int r = j - tj0 + 1;
    auto N = [&](int c) { return L.nu[r][c - ti0 + 1]; };
    T fa;
    if (d.periodic && (f == 0 || f == d.nx))
      fa = (T)0.5 * (N(f == 0 ? 0 : d.nx) * ip + N(f == 0 ? -1 : d.nx - 1) * ip);
    else if (f == 0)
      fa = N(0) * ip;
    else if (f == d.nx)
      fa = N(d.nx - 1) * ip;
    else
      fa = (T)0.5 * (N(f) * ip + N(f - 1) * ip);
    k = kap + fa;
  }
  if (WIN && d.insulated && d.win && j >= d.j0 && j < d.j1 && f >= d.i0 && f <= d.i1)
    k = k * d.ins_u[(j - d.j0) * (d.wx + 1) + (f - d.i0)];
  return k;
}
template <bool WIN = true, typename T, int R, typename D>
__device__ __forceinline__ T kappa_y(const D &d, const Lds<T, R> &L, int tj0, int ti0, T kap, int f, int i) {
  T k = kap;
  if (d.les) {
    const T ip = d.inv_prt;
    const int c = i - ti0 + 1;
    auto N = [&](int j) { return L.nu[j - tj0 + 1][c]; };
    T fa;
    if (d.periodic && (f == 0 || f == d.ny))
      fa = (T)0.5 * (N(f == 0 ? 0 : d.ny) * ip + N(f == 0 ? -1 : d.ny - 1) * ip);
    else if (f == 0)
      fa = N(0) * ip;
    else if (f == d.ny)
      fa = N(d.ny - 1) * ip;
    else
      fa = (T)0.5 * (N(f) * ip + N(f - 1) * ip);
    k = kap + fa;
  }
  if (WIN && d.insulated && d.win && f >= d.j0 && f <= d.j1 && i >= d.i0 && i < d.i1)
    k = k * d.ins_v[(f - d.j0) * d.wx + (i - d.i0)];
  return k;
}

// boundary diffusive flux of scalar.py:126-139; sign +1 low side, -1 high side
template <typename T>
__device__ __forceinline__ T wall_dflux(const Dev<T> &d, int mode, T k, T inner, T value, double sign) {
  if (mode == AERO_TH_TEMPERATURE || mode == AERO_TH_INLET) return ((sign > 0 ? -d.two_h : d.two_h) * k) * (inner - value);
  if (mode == AERO_TH_HEAT_FLUX) return ((T)sign * k) * value;
  return (T)0;
}

// emitter clock (physics.py:798-802)
template <typename T>
__device__ __forceinline__ T emitter_on(T t, const nd<T> *tm) {
  const T start = tm[0], stop = tm[1], period = tm[2], duty = tm[3];
  const bool active = (t >= start) && (t < stop);
  const T safe = period > (T)0 ? period : (T)1;
  T r = fmod(t - start, safe);
  if (r != (T)0 && ((r < (T)0) != (safe < (T)0))) r += safe;
  const T phase = r / safe;
  const bool on = active && ((period <= (T)0) || (phase < duty));
  return on ? (T)1 : (T)0;
}

// Live scalars the dt rule reads, loaded once (all loads issued together).
template <typename T>
struct DtIn {
  T nu, kappa, ri, dt_fixed, heat_max;
};
template <typename T>
__device__ __forceinline__ DtIn<T> load_dt_in(const Dev<T> &d) {
  DtIn<T> r;
  r.nu = *d.nu;
  r.dt_fixed = d.fixed_dt ? *d.dt_fixed : (T)0;
  r.kappa = d.thermal ? *d.kappa : (T)0;
  r.ri = d.thermal ? *d.ri : (T)0;
  r.heat_max = d.has_heat ? *d.heat_max : (T)0;
  return r;
}

// dt of the step (core.py:685-695 + physics.py:851-888) from the maxima of the state.
template <typename T>
__device__ T compute_dt(const Dev<T> &d, const DtIn<T> &in, T umax, T vmax, T thmax, T nutmax) {
  if (d.fixed_dt) return in.dt_fixed;
  const T nu = in.nu;
  const T us = umax + vmax;
  const T uc = us > (T)1e-6 ? us : (T)1e-6;
  T dt = (T)(d.cfl * d.hd) / uc;
  const T nue = d.les ? nu + nutmax : nu;
  const T vis = (T)(0.25 * d.hd * d.hd) / nue;
  dt = vis < dt ? vis : dt;
  if (d.thermal) {
    const T kap = d.les ? in.kappa + nutmax * (T)(1.0 / d.prt) : in.kappa;
    const T rate = us * (T)(2.0 / d.hd) + kap * (T)(d.diff_coef / (d.hd * d.hd));
    const T lim = (T)1 / rate;
    dt = lim < dt ? lim : dt;
    T acc = fabs(in.ri) * thmax;
    acc = acc > (T)1e-12 ? acc : (T)1e-12;
    const T ff = (T)d.cfl * sqrt((T)d.hd / acc);
    dt = ff < dt ? ff : dt;
    if (d.has_heat) {
      T smax = in.heat_max;
      smax = smax > (T)1e-12 ? smax : (T)1e-12;
      const T l1 = (T)1 / smax;
      dt = l1 < dt ? l1 : dt;
      T q = fabs(in.ri) * smax;
      q = q > (T)1e-12 ? q : (T)1e-12;
      const T l2 = pow((T)(6.0 * d.cfl * d.hd) / q, (T)(1.0 / 3.0));
      dt = l2 < dt ? l2 : dt;
    }
  }
  return dt;
}

// PID-controlled step (ns2d/diagnostics.py:104-123): steers dt toward 0.9 x the limit in log
// space with gains (kP, kI, kD) = (0.10, 0.25, 0.02), never above the limit and never growing
// more than 3 % per step; the first step (dt_prev = 0) takes the target. ctrl_prev / ctrl_new
// hold the last two errors.
template <typename T>
__device__ T pid_step(T limit, T dt_prev, const nd<T> *ctrl_prev, nd<T> *ctrl_new) {
  const T target = (T)0.9 * limit;
  const bool first = dt_prev <= (T)0;
  const T safe_prev = first ? target : dt_prev;
  const T e = log(target / safe_prev);
  const T e1 = ctrl_prev[0], e2 = ctrl_prev[1];
  const T step = ((T)0.25 * e + (T)0.10 * (e - e1)) + (T)0.02 * ((e - (T)2 * e1) + e2);
  const T cap = (T)0.02955880224154443;   // log(1.03) in double, rounded to T
  const T grow = exp(step < cap ? step : cap);
  const T cand = safe_prev * grow;
  ctrl_new[0] = first ? (T)0 : e;
  ctrl_new[1] = e1;
  return first ? target : (cand < limit ? cand : limit);
}

// ================================================================== kernels
// The explicit stage on one 64x8 tile (core.py:855-935 per face / cell).
// Werner-Wengle wall model (core.py:509-568): the power law u+ = A (y+)^B inverted for the
// wall stress, tau = (|us| (nu / y)^B / A)^(2 / (1 + B)); what it adds to the resolved no-slip
// stress nu |us| / y (zero inside the viscous sublayer) acts on the face row next to the wall as
// a force per unit volume -sign(us) excess / h. us: the face velocity relative to the wall.
constexpr double kWwA = 8.3, kWwB = 1.0 / 7.0;
// x^e for x >= 0: the library power in float64; in float32 the hardware base-2 logarithm and
// exponential (a few ulp), which keeps the library routine's code out of the stage kernel.
template <typename T>
__device__ __forceinline__ T wall_pow(T x, T e) { return pow(x, e); }
template <>
__device__ __forceinline__ float wall_pow<float>(float x, float e) {
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}
template <typename T>
__device__ __forceinline__ T wall_force(T us, T nu_y, T inv_h) {
  const T mag = us < (T)0 ? -us : us;
  const T lam = mag * nu_y;
  const T ww = wall_pow((mag * wall_pow(nu_y, (T)kWwB)) * (T)(1.0 / kWwA), (T)(2.0 / (1.0 + kWwB)));
  T excess = ww - lam;
  excess = excess > (T)0 ? excess : (T)0;
  const T sgn = us > (T)0 ? (T)1 : (us < (T)0 ? (T)-1 : (T)0);
  return -(sgn * excess) * inv_h;
}

template <typename T, int R, bool WIN = true, typename D>
__device__ __forceinline__ void stage_body(const D &d, const Stage<T> &s, int bx, int by) {
  __shared__ Lds<T, R> L;
  const int tj0 = by * R, ti0 = bx * TX;
  const int nx = d.nx, ny = d.ny;
  const T corr = s.psi_p ? d.red->corr : (T)0;
  load_uv(d, L, tj0, ti0, s.up, s.vp, s.psi_p, corr);
  if (d.thermal) load_c(d, L, tj0, ti0, s.thp, (const T *)nullptr, *d.theta_in, 1);
  __syncthreads();
  if (d.les) {
    tile_nut<WIN>(d, L, tj0, ti0);
    __syncthreads();
  }
  const T dt = d.red->dt;
  const T bdt = dt * (T)s.b;
  const T a = (T)s.a;
  const T wgt = (T)s.w;
  const T nu_h2 = *d.nu * d.inv_h2;
  const T kpen = bdt / *d.eta;
  const bool first = s.s == 0;
  const T ih = d.inv_h;
  const T ri_g = (d.thermal && d.buoyant) ? *d.ri : (T)0;
  const T kap = d.thermal ? *d.kappa : (T)0;
  // tile accessors: absolute (j, i) -> LDS
  auto U = [&](int j, int i) { return L.u[j - tj0 + 2][i - ti0 + 2]; };
  auto V = [&](int j, int i) { return L.v[j - tj0 + 2][i - ti0 + 2]; };
  auto N = [&](int j, int c) { return L.nu[j - tj0 + 1][c - ti0 + 1]; };
  auto C = [&](int j, int i) { return L.c[j - tj0 + 2][i - ti0 + 2]; };
  auto flux_ux = [&](int j, int c) {   // u-momentum x-flux at cell centre c
    const T a_ = U(j, c - 1), b_ = U(j, c), c_ = U(j, c + 1), e_ = U(j, c + 2);
    return face_flux(b_ + c_, a_, b_, c_, e_, d.upwind);
  };
  auto flux_uy = [&](int jc, int i) {  // u-momentum y-flux at corner (jc, i)
    return face_flux(V(jc, i - 1) + V(jc, i), U(jc - 2, i), U(jc - 1, i), U(jc, i), U(jc + 1, i), d.upwind);
  };
  auto flux_vx = [&](int j, int ic) {  // v-momentum x-flux at corner (j, ic)
    return face_flux(U(j - 1, ic) + U(j, ic), V(j, ic - 2), V(j, ic - 1), V(j, ic), V(j, ic + 1), d.upwind);
  };
  auto flux_vy = [&](int cm, int i) {  // v-momentum y-flux at cell row cm
    const T a_ = V(cm - 1, i), b_ = V(cm, i), c_ = V(cm + 1, i), e_ = V(cm + 2, i);
    return face_flux(b_ + c_, a_, b_, c_, e_, d.upwind);
  };
  auto shear = [&](int jc, int ic) {
    return (U(jc, ic) - U(jc - 1, ic)) * ih + (V(jc, ic) - V(jc, ic - 1)) * ih;
  };
  auto nutk = [&](int jc, int ic) {
    return (((N(jc - 1, ic - 1) + N(jc, ic - 1)) + N(jc - 1, ic)) + N(jc, ic)) * (T)0.25;
  };
  // penalise a u-face value in the window, book the increment (+ buoyancy share)
  auto pen_u = [&](int j, int i, T uh, T fu) -> T {
    const int wo = (j - d.j0) * (d.wx + 1) + (i - d.i0);
    const T un = (uh + kpen * d.a_u[wo]) / ((T)1 + kpen * d.chi_u[wo]);
    T acc = (un - uh) * wgt;
    if (!first) acc = d.acc_u[wo] + acc;
    if (fu != (T)0) acc = acc + (fu * (d.vol_u ? d.vol_u[wo] : d.chi_u[wo])) * (bdt * wgt);
    d.acc_u[wo] = acc;
    return un;
  };
  auto pen_v = [&](int j, int i, T vh, T fv) -> T {
    const int wo = (j - d.j0) * d.wx + (i - d.i0);
    const T vn = (vh + kpen * d.a_v[wo]) / ((T)1 + kpen * d.chi_v[wo]);
    T acc = (vn - vh) * wgt;
    if (!first) acc = d.acc_v[wo] + acc;
    if (fv != (T)0) acc = acc + (fv * (d.vol_v ? d.vol_v[wo] : d.chi_v[wo])) * (bdt * wgt);
    d.acc_v[wo] = acc;
    return vn;
  };
  auto in_win_u = [&](int j, int i) { return WIN && d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i <= d.i1; };
  auto in_win_v = [&](int j, int i) { return WIN && d.win && j >= d.j0 && j <= d.j1 && i >= d.i0 && i < d.i1; };
  auto theta_fx = [&](int j, int f) -> T {   // face average of theta (zero on boundary faces)
    if (f == 0 || f == nx) return d.periodic ? (T)0.5 * (C(j, 0) + C(j, -1)) : (T)0;
    return (T)0.5 * (C(j, f) + C(j, f - 1));
  };
  auto theta_fy = [&](int f, int i) -> T {
    if (f == 0 || f == ny) return d.periodic ? (T)0.5 * (C(0, i) + C(-1, i)) : (T)0;
    return (T)0.5 * (C(f, i) + C(f - 1, i));
  };

  // u faces, v faces and temperature cells in separate passes over the tile's two rows per
  // thread: each pass keeps only its own pointers and scalars live (fewer registers spilled)
  const int i = ti0 + threadIdx.x;
#pragma unroll 1
  for (int k = 0; k < TR<R>::RPT; ++k) {
    const int j = tj0 + threadIdx.y + 4 * k;
    if (j >= ny || i >= nx) continue;
    // ---------------- u face (j, i) [+ (j, nx) by the last column]
    {
      const bool bnd = !d.periodic && i == 0;
      T lu;
      if (bnd) {
        lu = (T)0;
      } else {
        const T au = (((flux_ux(j, i - 1) - flux_ux(j, i)) + flux_uy(j, i)) - flux_uy(j + 1, i)) * d.kadv;
        const T ring = ((U(j, i + 1) + U(j, i - 1)) + U(j + 1, i)) + U(j - 1, i);
        lu = au + (ring + (T)-4 * U(j, i)) * nu_h2;
        if (d.les) {
          const T txx_r = (T)2 * N(j, i) * ((U(j, i + 1) - U(j, i)) * ih);
          const T txx_l = (T)2 * N(j, i - 1) * ((U(j, i) - U(j, i - 1)) * ih);
          lu = lu + (((txx_r - txx_l) + nutk(j + 1, i) * shear(j + 1, i)) - nutk(j, i) * shear(j, i)) * ih;
        }
        if (d.wallm) {
          if (j == 0 && (d.bcb == AERO_BC_NOSLIP || d.bcb == AERO_BC_MOVING))
            lu = lu + wall_force(d.bcb == AERO_BC_MOVING ? U(j, i) - d.ground : U(j, i), *d.nu / d.wall_y, ih);
          if (j == ny - 1 && d.bct == AERO_BC_NOSLIP) lu = lu + wall_force(U(j, i), *d.nu / d.wall_y, ih);
        }
      }
      const int o = j * (nx + 1) + i;
      const T upv = U(j, i);
      T uh = (s.a == 0.0 ? upv : lerp_t(upv, s.u0[o], a)) + lu * bdt;
      if (!d.periodic && i == 0) uh = d.closed ? (T)0 : d.u_inlet[j];
      T fu = 0;
      if (ri_g != (T)0 && d.gx != 0.0) fu = theta_fx(j, i) * (ri_g * (T)(-d.gx));
      T ust = in_win_u(j, i) ? pen_u(j, i, uh, fu) : uh;
      if (fu != (T)0) ust = ust + bdt * fu;
      s.uh[o] = ust;
      if (d.periodic && i == 0) {
        // duplicate face nx: face 0's value (core.py:912-913); its own mask books its increment
        if (in_win_u(j, nx)) (void)pen_u(j, nx, uh, fu);
        s.uh[j * (nx + 1) + nx] = ust;
      }
      if (!d.periodic && i == nx - 1) {
        // outlet face: convective condition u_t + u_x = 0 (closed: wall)
        const int on = j * (nx + 1) + nx;
        const T lo = d.closed ? (T)0 : (U(j, nx - 1) - U(j, nx)) * ih;
        const T upo = U(j, nx);
        T uo = (s.a == 0.0 ? upo : lerp_t(upo, s.u0[on], a)) + lo * bdt;
        if (d.closed) uo = (T)0;
        else d.outcol[j] = uo;   // shifted by the mass correction in the projection passes
        if (in_win_u(j, nx)) uo = pen_u(j, nx, uo, (T)0);
        s.uh[on] = uo;
      }
    }
  }
#pragma unroll 1
  for (int k = 0; k < TR<R>::RPT; ++k) {
    const int j = tj0 + threadIdx.y + 4 * k;
    if (j >= ny || i >= nx) continue;
    // ---------------- v face (j, i) [+ (ny, i) by the last row]
    {
      const bool bnd = !d.periodic && j == 0;
      T lv;
      if (bnd) {
        lv = (T)0;
      } else {
        const T av = (((flux_vx(j, i) - flux_vx(j, i + 1)) + flux_vy(j - 1, i)) - flux_vy(j, i)) * d.kadv;
        const T ring = ((V(j, i + 1) + V(j, i - 1)) + V(j + 1, i)) + V(j - 1, i);
        lv = av + (ring + (T)-4 * V(j, i)) * nu_h2;
        if (d.les) {
          const T tyy_t = (T)2 * N(j, i) * ((V(j + 1, i) - V(j, i)) * ih);
          const T tyy_b = (T)2 * N(j - 1, i) * ((V(j, i) - V(j - 1, i)) * ih);
          lv = lv + (((nutk(j, i + 1) * shear(j, i + 1) - nutk(j, i) * shear(j, i)) + tyy_t) - tyy_b) * ih;
        }
        if (d.wallm && d.closed && (d.bcl == AERO_BC_NOSLIP || d.bcr == AERO_BC_NOSLIP)) {
          if (i == 0 && d.bcl == AERO_BC_NOSLIP) lv = lv + wall_force(V(j, i), *d.nu / d.wall_y, ih);
          if (i == nx - 1 && d.bcr == AERO_BC_NOSLIP) lv = lv + wall_force(V(j, i), *d.nu / d.wall_y, ih);
        }
      }
      const int o = j * nx + i;
      const T vpv = V(j, i);
      T vh = (s.a == 0.0 ? vpv : lerp_t(vpv, s.v0[o], a)) + lv * bdt;
      T fv = 0;
      if (ri_g != (T)0 && d.gy != 0.0) fv = theta_fy(j, i) * (ri_g * (T)(-d.gy));
      T vst = in_win_v(j, i) ? pen_v(j, i, vh, fv) : vh;
      if (fv != (T)0) vst = vst + bdt * fv;
      s.vh[o] = vst;
      if (d.periodic && j == 0) {
        if (in_win_v(ny, i)) (void)pen_v(ny, i, vh, fv);
        s.vh[ny * nx + i] = vst;
      }
      if (!d.periodic && j == ny - 1) {
        const int on = ny * nx + i;
        const T vpo = V(ny, i);
        T vo = (s.a == 0.0 ? vpo : lerp_t(vpo, s.v0[on], a));
        if (in_win_v(ny, i)) vo = pen_v(ny, i, vo, (T)0);
        s.vh[on] = vo;
      }
    }
  }
#pragma unroll 1
  for (int k = 0; k < TR<R>::RPT; ++k) {
    const int j = tj0 + threadIdx.y + 4 * k;
    if (j >= ny || i >= nx) continue;
    // ---------------- temperature cell (j, i) (physics.py:696-757)
    if (d.thermal) {
      const int r = j - tj0 + 2, c = i - ti0 + 2;
      const T fxl = tvd_x(L, r, c, U(j, i)), fxr = tvd_x(L, r, c + 1, U(j, i + 1));
      const T fyb = tvd_y(L, r, c, V(j, i)), fyt = tvd_y(L, r + 1, c, V(j + 1, i));
      T lt = (((fxl - fxr) + fyb) - fyt) * ih;
      auto dfx = [&](int f) -> T {
        // periodic boundary faces: the wrapped difference (scalar.py:133-134) read from the tile
        if ((f > 0 && f < nx) || d.periodic)
          return -(kappa_x<WIN>(d, L, tj0, ti0, kap, j, f) * (C(j, f) - C(j, f - 1))) * ih;
        if (f == 0) return wall_dflux(d, d.wm[0], kappa_x<WIN>(d, L, tj0, ti0, kap, j, 0), C(j, 0), d.wall_values[0], 1.0);
        return wall_dflux(d, d.wm[1], kappa_x<WIN>(d, L, tj0, ti0, kap, j, nx), C(j, nx - 1), d.wall_values[1], -1.0);
      };
      auto dfy = [&](int f) -> T {
        if ((f > 0 && f < ny) || d.periodic)
          return -(kappa_y<WIN>(d, L, tj0, ti0, kap, f, i) * (C(f, i) - C(f - 1, i))) * ih;
        if (f == 0) return wall_dflux(d, d.wm[2], kappa_y<WIN>(d, L, tj0, ti0, kap, 0, i), C(0, i), d.wall_values[2], 1.0);
        return wall_dflux(d, d.wm[3], kappa_y<WIN>(d, L, tj0, ti0, kap, ny, i), C(ny - 1, i), d.wall_values[3], -1.0);
      };
      lt = lt + (((dfx(i) - dfx(i + 1)) + dfy(j)) - dfy(j + 1)) * ih;
      const int o = j * nx + i;
      if (d.has_heat) lt = lt + d.heat[o];
      // heat-flux bodies release kappa G P over their skins (physics.py:726-727)
      if (WIN && d.n_flux > 0 && d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1)
        lt = lt + kap * d.q_skin[(j - d.j0) * d.wx + (i - d.i0)];
      const T tp = C(j, i);
      T thh = (s.a == 0.0 ? tp : lerp_t(tp, s.th0[o], a)) + lt * bdt;
      if (WIN && d.n_temp > 0 && d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1) {
        const int wo = (j - d.j0) * d.wx + (i - d.i0);
        const T kt = bdt / *d.eta_t;
        const T tn = (thh + kt * d.a_t[wo]) / ((T)1 + kt * d.chi_t[wo]);
        T acc = (tn - thh) * wgt;
        if (!first) acc = d.acc_t[wo] + acc;
        d.acc_t[wo] = acc;
        thh = tn;
      }
      s.thh[o] = thh;
    }
  }
}
template <typename T, int R, typename D = Dev<T>>
__global__ void __launch_bounds__(256) k_stage(D d, Stage<T> s) {
  // tiles clear of the body window (2 cells around it) run without any window code
  const int tj0 = blockIdx.y * R, ti0 = blockIdx.x * TX;
  const bool touch = d.win && (d.periodic || (tj0 - 2 <= d.j1 && tj0 + R + 2 >= d.j0 && ti0 - 2 <= d.i1 &&
                                              ti0 + TX + 2 >= d.i0));
  if (touch) stage_body<T, R, true>(d, s, blockIdx.x, blockIdx.y);
  else stage_body<T, R, false>(d, s, blockIdx.x, blockIdx.y);
}

// Per-block maxima of a velocity tile (+ theta, nu_t) and sum theta: the dt inputs.
// Each tile also covers the ghost ring of nu_t cells at the domain edge (core.py:503).
template <typename T, int R, typename D>
__device__ __forceinline__ void tile_maxima(const D &d, Lds<T, R> &L, int tj0, int ti0, const T *th) {
  const int nx = d.nx, ny = d.ny;
  double mu = 0, mv = 0, mt = 0, mn = 0, st = 0;
  const int i = ti0 + threadIdx.x;
  for (int k = 0; k < TR<R>::RPT; ++k) {
    const int j = tj0 + threadIdx.y + 4 * k;
    if (j >= ny || i >= nx) continue;
    const int r = j - tj0 + 2, c = i - ti0 + 2;
    mu = fmax(mu, (double)fabs(L.u[r][c]));
    mv = fmax(mv, (double)fabs(L.v[r][c]));
    if (i == nx - 1) mu = fmax(mu, (double)fabs(L.u[r][c + 1]));
    if (j == ny - 1) mv = fmax(mv, (double)fabs(L.v[r + 1][c]));
    if (d.thermal) {
      const T tv = th[j * nx + i];
      mt = fmax(mt, (double)fabs(tv));
      st += (double)tv;
    }
  }
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (d.les) {
    for (int e = tid; e < TR<R>::NH * NW; e += 256) {
      const int r = e / NW, c = e - r * NW;
      const int j = tj0 - 1 + r, cc = ti0 - 1 + c;
      if (j < -1 || j > ny || cc < -1 || cc > nx) continue;
      mn = fmax(mn, (double)L.nu[r][c]);
    }
  }
  mu = wave_max(mu);
  mv = wave_max(mv);
  mt = wave_max(mt);
  mn = wave_max(mn);
  st = wave_sum(st);
  const int w = tid >> 6;
  __syncthreads();
  if ((tid & 63) == 0) {
    L.red[w * 5 + 0] = mu;
    L.red[w * 5 + 1] = mv;
    L.red[w * 5 + 2] = mt;
    L.red[w * 5 + 3] = mn;
    L.red[w * 5 + 4] = st;
  }
  __syncthreads();
  if (tid == 0) {
    double *p = d.pmax + 5 * ((tj0 / R) * d.nbx + ti0 / TX);
    for (int q = 0; q < 4; ++q)
      p[q] = fmax(fmax(L.red[q], L.red[5 + q]), fmax(L.red[10 + q], L.red[15 + q]));
    p[4] = ((L.red[4] + L.red[9]) + L.red[14]) + L.red[19];
  }
  __syncthreads();
}

// dt inputs of the step-start state (non-chained steps).
template <typename T, int R, typename D>
__device__ __forceinline__ void pre_body(const D &d, int bx, int by, const T *u0, const T *v0, const T *th0) {
  __shared__ Lds<T, R> L;
  const int tj0 = by * R, ti0 = bx * TX;
  load_uv(d, L, tj0, ti0, u0, v0, (const T *)nullptr, (T)0);
  __syncthreads();
  if (d.les) {
    tile_nut(d, L, tj0, ti0);
    __syncthreads();
  }
  tile_maxima(d, L, tj0, ti0, th0);
}
template <typename T, int R, typename D = Dev<T>>
__global__ void __launch_bounds__(256) k_pre(D d, const T *u0, const T *v0, const T *th0) {
  pre_body<T, R>(d, blockIdx.x, blockIdx.y, u0, v0, th0);
}

// Reduce the per-block dt inputs in a fixed order (one block; every thread gets out[]).
template <typename T>
__device__ void reduce_pmax(const Dev<T> &d, int nblocks, double *lds, double out[5]) {
  const int tid = threadIdx.x;
  double m[4] = {0, 0, 0, 0}, s = 0;
#pragma unroll 4
  for (int b = tid; b < nblocks; b += blockDim.x) {
    const double *p = d.pmax + 5 * b;
    for (int q = 0; q < 4; ++q) m[q] = fmax(m[q], p[q]);
    s += p[4];
  }
  for (int q = 0; q < 4; ++q) m[q] = wave_max(m[q]);
  s = wave_sum(s);
  const int w = tid >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((tid & 63) == 0) {
    for (int q = 0; q < 4; ++q) lds[w * 5 + q] = m[q];
    lds[w * 5 + 4] = s;
  }
  __syncthreads();
  for (int q = 0; q < 5; ++q) out[q] = 0;
  for (int ww = 0; ww < nw; ++ww) {
    for (int q = 0; q < 4; ++q) out[q] = fmax(out[q], lds[ww * 5 + q]);
    out[4] += lds[ww * 5 + 4];
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void dt_body(const Dev<T> &d) {
  __shared__ double lds[16 * 5];
  double r[5];
  DtIn<T> din;
  if (threadIdx.x == 0) din = load_dt_in(d);
  reduce_pmax(d, d.nbx * d.nby, lds, r);
  if (threadIdx.x == 0) {
    T dt = compute_dt(d, din, (T)r[0], (T)r[1], (T)r[2], (T)r[3]);
    if (d.pid && !d.fixed_dt) dt = pid_step(dt, *d.dt_in, d.ctrl_in, d.red->ctrl);
    d.red->dt = dt;
    d.red->th_old = r[4];
  }
}
template <typename T>
__global__ void __launch_bounds__(1024) k_dt(Dev<T> d) {
  dt_body<T>(d);
}

// natural index of reordered position n (inverse of makhoul_pos)
__device__ __forceinline__ int makhoul_src(int n, int len) { return n < (len + 1) / 2 ? 2 * n : 2 * (len - 1 - n) + 1; }

// Outlet correction of core.py:337-357, computed identically by every block.
template <typename T>
__device__ __forceinline__ T outlet_corr(const Dev<T> &d, double *lds) {
  double v[2] = {0, 0};
  const int tid = threadIdx.x;
  const auto u_inlet = d.u_inlet;
  const auto outcol = d.outcol;
  for (int j = tid; j < d.ny; j += blockDim.x) {
    v[0] += (double)u_inlet[j];
    v[1] += (double)outcol[j];
  }
  block_sum4<2>(v, lds);
  __syncthreads();
  if (tid == 0) {
    lds[8] = v[0];
    lds[9] = v[1];
  }
  __syncthreads();
  const T r = ((T)lds[8] - (T)lds[9]) / (T)d.ny;
  __syncthreads();
  return r;
}

// divergence at cell (j, i) with the outlet face shifted by corr (core.py:621-623)
template <typename T, typename P>
__device__ __forceinline__ T div_at(const Dev<T> &d, P u, P v, int j, int i, T corr) {
  const int nx = d.nx;
  T ur = u[j * (nx + 1) + i + 1];
  if (i + 1 == nx) ur = ur + corr;
  return (((ur - u[j * (nx + 1) + i]) + v[(j + 1) * nx + i]) - v[j * nx + i]) * d.inv_h;
}

template <typename T, typename P>
__device__ __forceinline__ void load_twiddles(cplx<T> *dst, P src, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) dst[m] = src[m];
}

// Pass A: divergence of the nslot row pairs (2p, 2p+1), p = b nslot .. b nslot + nslot - 1 ->
// forward transform along x -> F rows. The block's rows sit one after another in LDS, each pair
// transformed as one complex sequence; pairs past the last row are skipped. ONE: nslot is 1
// (single runs), known at compile time.
template <typename T, int NT, bool ONE = false>
__device__ __forceinline__ void poisA_body(const Dev<T> &d, int bx, int nslot, const T *__restrict__ u,
                                           const T *__restrict__ v, const T *__restrict__ f_in, int mode) {
  constexpr int MP = kMaxLen / NT;   // row elements per thread
  // mode 0: f_in is the right-hand side itself; 1: divergence of (u, v); 2: + outlet correction
  extern __shared__ unsigned char smem[];
  if (ONE) nslot = 1;
  const int nx = d.nx, ny = d.ny, lx = ONE ? 0 : log2_or_neg(nx);
  const int nel = nslot * nx;        // elements of the block's row pairs
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + nel;
  cplx<T> *TW = B + nel;
  double *red = reinterpret_cast<double *>(TW + nx);
  const int j0 = 2 * bx * nslot;     // first row of the block
  const auto F = d.F;
  const auto cs_x = d.cs_x;
  const auto gu = global_ptr(u), gv = global_ptr(v), gf = global_ptr(f_in);
  load_twiddles(TW, d.tw_x, nx);
  // every load of this thread is issued before the outlet correction's reduction waits:
  // the row elements (the face values of the divergence, or the right-hand side) and the DCT
  // twiddles of output frequency k = q, which this thread finishes below
  // (the first PF elements of each thread; longer rows on smaller blocks load the rest after)
  constexpr int PF = MP < 4 ? MP : 4;
  cplx<T> csr[MP];
  T ur[PF][2], ul[PF][2], vt[PF][2], vb[PF][2];
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int w = threadIdx.x + NT * r;
    if (w < nel && d.dct_x) csr[r] = cs_x[w - (ONE ? 0 : fdiv(w, nx, lx)) * nx];
  }
#pragma unroll
  for (int r = 0; r < PF; ++r) {
    const int w = threadIdx.x + NT * r;
    if (w < nel) {
      const int f = (ONE ? 0 : fdiv(w, nx, lx)), q = w - f * nx;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = j0 + 2 * f + h;
        ur[r][h] = ul[r][h] = vt[r][h] = vb[r][h] = (T)0;
        if (j >= ny) continue;
        if (mode == 0) {
          ur[r][h] = gf[j * nx + q];
        } else {
          ur[r][h] = gu[j * (nx + 1) + q + 1];
          ul[r][h] = gu[j * (nx + 1) + q];
          vt[r][h] = gv[(j + 1) * nx + q];
          vb[r][h] = gv[j * nx + q];
        }
      }
    }
  }
  T corr = 0;
  if (mode == 2) {
    corr = outlet_corr(d, red);
    if (bx == 0 && threadIdx.x == 0) d.red->corr = corr;
  }
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    // natural column q of pair f into transform position m; the divergence as div_at forms it
    const int w = threadIdx.x + NT * r;
    if (w < nel) {
      const int f = (ONE ? 0 : fdiv(w, nx, lx)), q = w - f * nx;
      const int ja = j0 + 2 * f;
      const int m = d.dct_x ? makhoul_pos(q, nx) : q;
      T x[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (ja + h >= ny) {
          x[h] = (T)0;
        } else if (r >= PF) {
          x[h] = mode == 0 ? gf[(ja + h) * nx + q] : div_at(d, gu, gv, ja + h, q, corr);
        } else if (mode == 0) {
          x[h] = ur[r < PF ? r : 0][h];
        } else {
          const int rp = r < PF ? r : 0;
          const T right = q + 1 == nx ? ur[rp][h] + corr : ur[rp][h];
          x[h] = (((right - ul[rp][h]) + vt[rp][h]) - vb[rp][h]) * d.inv_h;
        }
      }
      A[f * nx + m] = {x[0], x[1]};
    }
  }
  __syncthreads();
  cplx<T> *R = fft_run(A, B, d.fx, TW, nslot, nx);
#pragma unroll
  for (int rr = 0; rr < MP; ++rr) {
    const int w = threadIdx.x + NT * rr;
    if (w >= nel) continue;
    const int f = (ONE ? 0 : fdiv(w, nx, lx)), k = w - f * nx;
    const int ja = j0 + 2 * f;
    if (ja >= ny) continue;
    cplx<T> va, vb;
    split_pair(R + f * nx, nx, k, va, vb);
    T xa, xb;
    if (d.dct_x) {
      const cplx<T> cs = csr[rr];
      xa = cs.x * va.x + cs.y * va.y;
      xb = cs.x * vb.x + cs.y * vb.y;
    } else {
      xa = va.x - va.y;
      xb = vb.x - vb.y;
    }
    F[ja * nx + k] = xa;
    if (ja + 1 < ny) F[(ja + 1) * nx + k] = xb;
  }
}
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_poisA(Dev<T> d, const T *__restrict__ u, const T *__restrict__ v,
                                              const T *__restrict__ f_in, int mode) {
  poisA_body<T, NT, true>(d, blockIdx.x, 1, u, v, f_in, mode);
}

// Inverse real transform packing (see common.h): conj(Z) for the DCT-III, plain pair for DHT.
template <typename T>
__device__ __forceinline__ void pack_inverse(const T *xa, const T *xb, cplx<T> *Z, int n, int dct, cplx<T> w, int k) {
  if (dct) {
    const T xa_k = xa[k], xb_k = xb[k];
    const T xa_r = k == 0 ? (T)0 : xa[n - k];
    const T xb_r = k == 0 ? (T)0 : xb[n - k];
    const cplx<T> va = {w.x * xa_k + w.y * xa_r, w.y * xa_k - w.x * xa_r};
    const cplx<T> vb = {w.x * xb_k + w.y * xb_r, w.y * xb_k - w.x * xb_r};
    Z[k] = {va.x - vb.y, -(va.y + vb.x)};
  } else {
    Z[k] = {xa[k], xb[k]};
  }
}

// Pass B: column pairs, nslot per block: forward along y, x 1/lambda (FP64), inverse along y.
// The global loads and stores walk the block's column pairs at one row (2 nslot consecutive
// values of a row per segment); the spectral side keeps one column's elements per run of lanes.
// Each column sits one element apart from the next in LDS (slot stride ny + 1), so the
// row-walking accesses fall in different banks.
template <typename T, int NT>
__device__ __forceinline__ void poisB_body(const Dev<T> &d, int bx, int nslot) {
  constexpr int MP = kMaxLen / NT;   // column elements per thread
  extern __shared__ unsigned char smem[];
  const int ny = d.ny, nx = d.nx, st = ny + 1, nel = nslot * ny;
  const int ly = log2_or_neg(ny), ls = log2_or_neg(nslot);
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + nslot * st;
  cplx<T> *TW = B + nslot * st;
  T *X = reinterpret_cast<T *>(TW + ny);   // 2 nslot rows of st reals
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int c0 = 2 * nslot * bx;
  const auto F = d.F;
  const auto inv_lam = d.inv_lam;
  const auto tw_y = d.tw_y, cs_y = d.cs_y;
  cplx<T> csr[MP];
  T iar[MP], ibr[MP];
  // twiddles, column elements, DCT twiddles and spectral scalings in one pass of loads
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int w = tid + nthr * r;
    if (w < ny) TW[w] = tw_y[w];
    if (w < nel) {
      const int n = fdiv(w, nslot, ls), f = w - n * nslot;   // row-walking: pair f of row n
      const int ca = c0 + 2 * f, cb = ca + 1;
      const int q = d.dct_y ? makhoul_src(n, ny) : n;
      const T xa = ca < nx ? F[q * nx + ca] : (T)0;
      const T xb = cb < nx ? F[q * nx + cb] : (T)0;
      A[f * st + n] = {xa, xb};
      const int g = fdiv(w, ny, ly), k = w - g * ny;          // column-walking: element k of pair g
      const int ga = c0 + 2 * g, gb = ga + 1;
      csr[r] = cs_y[k];
      iar[r] = ga < nx ? inv_lam[(size_t)ga * ny + k] : (T)0;
      ibr[r] = gb < nx ? inv_lam[(size_t)gb * ny + k] : (T)0;
    }
  }
  __syncthreads();
  cplx<T> *R = fft_run(A, B, d.fy, TW, nslot, st);
#pragma unroll
  for (int rr = 0; rr < MP; ++rr) {
    const int w = tid + nthr * rr;
    if (w >= nel) continue;
    const int f = fdiv(w, ny, ly), k = w - f * ny;
    cplx<T> va, vb;
    split_pair(R + f * st, ny, k, va, vb);
    T xa, xb;
    if (d.dct_y) {
      const cplx<T> cs = csr[rr];
      xa = cs.x * va.x + cs.y * va.y;
      xb = cs.x * vb.x + cs.y * vb.y;
    } else {
      xa = va.x - va.y;
      xb = vb.x - vb.y;
    }
    X[(2 * f) * st + k] = xa * iar[rr];
    X[(2 * f + 1) * st + k] = xb * ibr[rr];
  }
  __syncthreads();
  cplx<T> *Z = (R == A) ? B : A;
#pragma unroll
  for (int rr = 0; rr < MP; ++rr) {
    const int w = tid + nthr * rr;
    if (w >= nel) continue;
    const int f = fdiv(w, ny, ly), k = w - f * ny;
    pack_inverse(X + (2 * f) * st, X + (2 * f + 1) * st, Z + f * st, ny, d.dct_y, csr[rr], k);
  }
  __syncthreads();
  cplx<T> *R2 = fft_run(Z, (Z == A) ? B : A, d.fy, TW, nslot, st);
  const T inv_n = (T)(1.0 / (double)ny);
  for (int w = tid; w < nel; w += nthr) {
    const int m = fdiv(w, nslot, ls), f = w - m * nslot;    // row-walking stores
    const int ca = c0 + 2 * f, cb = ca + 1;
    const cplx<T> *Rf = R2 + f * st;
    T va, vb;
    if (d.dct_y) {
      va = Rf[m].x * inv_n;
      vb = -Rf[m].y * inv_n;
    } else {
      cplx<T> ha, hb;
      split_pair(Rf, ny, m, ha, hb);
      va = (ha.x - ha.y) * inv_n;
      vb = (hb.x - hb.y) * inv_n;
    }
    const int q = d.dct_y ? makhoul_src(m, ny) : m;
    if (ca < nx) F[q * nx + ca] = va;
    if (cb < nx) F[q * nx + cb] = vb;
  }
}
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_poisB(Dev<T> d, int nslot) {
  // column groups dealt to the 8 XCDs in contiguous runs: the groups that share cache lines
  // of a row (nslot pairs of columns each) read them through one L2
  const int nb = gridDim.x, b = blockIdx.x;
  const int g = (nb % 8 == 0) ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  poisB_body<T, NT>(d, g, nslot);
}

// Pass C: inverse along x of the nslot row pairs of block b (as pass A) -> psi; accumulate the
// step-averaged pressure p (core.py:935-942): stage 0 p = w psi, stage 1 p += w psi, stage 2
// p = (p + psi) / dt.
template <typename T, int NT, bool ONE = false>
__device__ __forceinline__ void poisC_body(const Dev<T> &d, int bx, int nslot, T *__restrict__ psi, T *__restrict__ p,
                                           int stage) {
  constexpr int MP = kMaxLen / NT;   // row elements per thread
  extern __shared__ unsigned char smem[];
  if (ONE) nslot = 1;
  const int nx = d.nx, ny = d.ny, lx = ONE ? 0 : log2_or_neg(nx);
  const int nel = nslot * nx;        // elements of the block's row pairs
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + nel;
  cplx<T> *TW = B + nel;
  T *X = reinterpret_cast<T *>(TW + nx);   // 2 nx reals per pair: row 2p, then row 2p + 1
  const int j0 = 2 * bx * nslot;     // first row of the block
  cplx<T> csr[MP];
  // previous pressure accumulator values (stages 1, 2), at the positions this thread writes
  T pa[MP], pb[MP];
  const bool acc_p = p != nullptr && stage > 0;
  const auto F = d.F;
  const auto tw_x = d.tw_x, cs_x = d.cs_x;
  const auto gpsi = global_ptr(psi), gp = global_ptr(p);
  // twiddles, row elements, DCT twiddles and the accumulator in one pass of loads
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int w = threadIdx.x + NT * r;
    if (w < nel) {
      const int f = (ONE ? 0 : fdiv(w, nx, lx)), k = w - f * nx;
      const int ja = j0 + 2 * f;
      if (w < nx) TW[w] = tw_x[w];
      X[2 * f * nx + k] = ja < ny ? F[ja * nx + k] : (T)0;
      X[2 * f * nx + nx + k] = ja + 1 < ny ? F[(ja + 1) * nx + k] : (T)0;
      csr[r] = cplx<T>{0, 0};
      if (d.dct_x) csr[r] = cs_x[k];
      if (acc_p) {
        pa[r] = ja < ny ? gp[ja * nx + k] : (T)0;
        pb[r] = ja + 1 < ny ? gp[(ja + 1) * nx + k] : (T)0;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int w = threadIdx.x + NT * r;
    if (w < nel) {
      const int f = (ONE ? 0 : fdiv(w, nx, lx)), k = w - f * nx;
      pack_inverse(X + 2 * f * nx, X + 2 * f * nx + nx, A + f * nx, nx, d.dct_x, csr[r], k);
    }
  }
  __syncthreads();
  cplx<T> *R = fft_run(A, B, d.fx, TW, nslot, nx);
  const T inv_n = (T)(1.0 / (double)nx);
  T rdt = 0;
  if (p != nullptr && stage == 2) rdt = (T)1 / d.red->dt;
  const T w = (T)kForceW[stage];
#pragma unroll
  for (int rr = 0; rr < MP; ++rr) {
    // natural column q (coalesced stores) from transform position m
    const int e = threadIdx.x + NT * rr;
    if (e >= nel) continue;
    const int f = (ONE ? 0 : fdiv(e, nx, lx)), q = e - f * nx;
    const int ja = j0 + 2 * f, jb = ja + 1;
    if (ja >= ny) continue;
    const bool hasb = jb < ny;
    const cplx<T> *Rf = R + f * nx;
    const int m = d.dct_x ? makhoul_pos(q, nx) : q;
    T va, vb;
    if (d.dct_x) {
      va = Rf[m].x * inv_n;
      vb = -Rf[m].y * inv_n;
    } else {
      cplx<T> ha, hb;
      split_pair(Rf, nx, m, ha, hb);
      va = (ha.x - ha.y) * inv_n;
      vb = (hb.x - hb.y) * inv_n;
    }
    gpsi[ja * nx + q] = va;
    if (hasb) gpsi[jb * nx + q] = vb;
    if (p != nullptr) {
      if (stage == 0) {
        gp[ja * nx + q] = va * w;
        if (hasb) gp[jb * nx + q] = vb * w;
      } else if (stage == 1) {
        gp[ja * nx + q] = pa[rr] + w * va;
        if (hasb) gp[jb * nx + q] = pb[rr] + w * vb;
      } else {
        gp[ja * nx + q] = (pa[rr] + va) * rdt;
        if (hasb) gp[jb * nx + q] = (pb[rr] + vb) * rdt;
      }
    }
  }
}
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_poisC(Dev<T> d, T *__restrict__ psi, T *__restrict__ p, int stage) {
  poisC_body<T, NT, true>(d, blockIdx.x, 1, psi, p, stage);
}

// u -= grad psi (standalone projection, core.py:626-660).
template <typename T>
__global__ void __launch_bounds__(256) k_correct(Dev<T> d, T *__restrict__ u, T *__restrict__ v, const T *__restrict__ psi) {
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

// End of step on one tile: write u = u^ - grad psi, smoke transport with it (physics.py:
// 805-830 / core.py:709-771), next-step dt inputs and the window force / heat partials.
template <typename T, int R, bool WIN = true, typename D>
__device__ __forceinline__ void final_body(const D &d, int bx, int by, const T *uh, const T *vh, T *u_out, T *v_out, const T *th_new, const T *cin, T *cout, const T *t_in) {
  __shared__ Lds<T, R> L;
  const int tj0 = by * R, ti0 = bx * TX;
  const int nx = d.nx, ny = d.ny;
  const T corr = d.red->corr;
  load_uv(d, L, tj0, ti0, uh, vh, (const T *)d.psi, corr);
  __syncthreads();
  const int i = ti0 + threadIdx.x;
  for (int k = 0; k < TR<R>::RPT; ++k) {
    const int j = tj0 + threadIdx.y + 4 * k;
    if (j >= ny || i >= nx) continue;
    const int r = j - tj0 + 2, c = i - ti0 + 2;
    u_out[j * (nx + 1) + i] = L.u[r][c];
    v_out[j * nx + i] = L.v[r][c];
    if (i == nx - 1) u_out[j * (nx + 1) + nx] = L.u[r][c + 1];
    if (j == ny - 1) v_out[ny * nx + i] = L.v[r + 1][c];
  }
  if (d.les) {
    tile_nut<WIN>(d, L, tj0, ti0);
    __syncthreads();
  }
  tile_maxima(d, L, tj0, ti0, th_new);
  // force / heat partials of the window faces this tile owns
  const bool touch = WIN && d.win && ti0 <= d.i1 && (ti0 + TX > d.i0 || d.periodic) && tj0 <= d.j1 &&
                     (tj0 + R > d.j0 || d.periodic);
  if (touch) {
    double acc[kMaxWin];
    for (int q = 0; q < kMaxWin; ++q) acc[q] = 0;
    const int nuw = d.wy * (d.wx + 1), nvw = (d.wy + 1) * d.wx;
    auto add_u = [&](int j, int fi) {
      if (j >= d.j0 && j < d.j1 && fi >= d.i0 && fi <= d.i1) {
        const int wo = (j - d.j0) * (d.wx + 1) + (fi - d.i0);
        const double a = (double)d.acc_u[wo];
#pragma unroll
        for (int b = 0; b < kMaxWin / 4; ++b)
          if (b < d.nbody) {
            acc[4 * b] += (double)d.k_u[(size_t)b * 2 * nuw + wo] * a;
            acc[4 * b + 1] += (double)d.k_u[(size_t)b * 2 * nuw + nuw + wo] * a;
          }
      }
    };
    auto add_v = [&](int fj, int i_) {
      if (fj >= d.j0 && fj <= d.j1 && i_ >= d.i0 && i_ < d.i1) {
        const int wo = (fj - d.j0) * d.wx + (i_ - d.i0);
        const double a = (double)d.acc_v[wo];
#pragma unroll
        for (int b = 0; b < kMaxWin / 4; ++b)
          if (b < d.nbody) {
            acc[4 * b + 2] += (double)d.k_v[(size_t)b * 2 * nvw + wo] * a;
            acc[4 * b + 3] += (double)d.k_v[(size_t)b * 2 * nvw + nvw + wo] * a;
          }
      }
    };
    for (int k = 0; k < TR<R>::RPT; ++k) {
      const int j = tj0 + threadIdx.y + 4 * k;
      if (j >= ny || i >= nx) continue;
      add_u(j, i);
      add_v(j, i);
      // the extra faces nx / ny belong to the last column / row (periodic: to face 0)
      if (i == nx - 1 && !d.periodic) add_u(j, nx);
      if (j == ny - 1 && !d.periodic) add_v(ny, i);
      if (i == 0 && d.periodic) add_u(j, nx);
      if (j == 0 && d.periodic) add_v(ny, i);
      if (d.thermal && d.n_temp > 0 && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1) {
        const int wo = (j - d.j0) * d.wx + (i - d.i0);
        const double at = (double)d.acc_t[wo];
#pragma unroll
        for (int q = 0; q < kMaxWin; ++q) {
          const int b = q - 4 * d.nbody;   // heat slots follow the force slots
          if (b >= 0 && b < d.n_temp) acc[q] += (double)d.w_t[(size_t)b * d.wy * d.wx + wo] * at;
        }
      }
      if (d.n_flux > 0 && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1) {
        // skin-weighted temperature of each heat-flux body (core.py:988-990), slots after the heat
        const int wo = (j - d.j0) * d.wx + (i - d.i0);
        const double tv = (double)th_new[j * nx + i];
#pragma unroll
        for (int q = 0; q < kMaxWin; ++q) {
          const int b = q - 4 * d.nbody - d.n_temp;
          if (b >= 0 && b < d.n_flux) acc[q] += (double)d.w_f[(size_t)b * d.wy * d.wx + wo] * tv;
        }
      }
    }
    block_sum4<kMaxWin>(acc, L.red);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      double *p = d.pforce + kMaxWin * (by * d.nbx + bx);
      for (int q = 0; q < kMaxWin; ++q) p[q] = acc[q];
    }
  }
  // smoke
  if (d.dye_mode != 0 && cin != nullptr) {
    const T dt = d.red->dt;
    const T tnew = *t_in + dt;
    const T ih = d.inv_h;
    for (int ch = 0; ch < d.n_dye; ++ch) {
      const T *c = cin + (size_t)ch * ny * nx;
      __syncthreads();
      if (d.dye_mode == 2)
        load_c(d, L, tj0, ti0, c, d.dye_inflow + (size_t)ch * ny, (T)0, 1);
      else
        load_c(d, L, tj0, ti0, c, d.dye_inflow, (T)0, 2);
      __syncthreads();
      for (int k = 0; k < TR<R>::RPT; ++k) {
        const int j = tj0 + threadIdx.y + 4 * k;
        if (j >= ny || i >= nx) continue;
        const int r = j - tj0 + 2, cc = i - ti0 + 2;
        const T fxl = tvd_x(L, r, cc, L.u[r][cc]), fxr = tvd_x(L, r, cc + 1, L.u[r][cc + 1]);
        const T fyb = tvd_y(L, r, cc, L.v[r][cc]), fyt = tvd_y(L, r + 1, cc, L.v[r + 1][cc]);
        const int o = j * nx + i;
        T val;
        const bool solid = WIN && d.win && j >= d.j0 && j < d.j1 && i >= d.i0 && i < d.i1 &&
                           !(d.chi_c[(j - d.j0) * d.wx + (i - d.i0)] <= (T)0.5);
        if (d.dye_mode == 2) {
          val = L.c[r][cc] + dt * ((((fxl - fxr) + fyb) - fyt) * ih);
          for (int g = 0; g < d.n_groups; ++g) {
            if (d.gch[g] != ch) continue;
            const T wg = d.g_weight[(size_t)g * ny * nx + o] * emitter_on(tnew, d.g_timing + 4 * g);
            val = val + wg * (d.g_target[(size_t)g * ny * nx + o] - val);
          }
          if (solid) val = val * (T)0;
          val = val > (T)1e-6 ? val : (T)0;
          val = val < (T)1 ? val : (T)1;
        } else {
          val = L.c[r][cc] - (dt * ih) * (((fxr - fxl) + fyt) - fyb);
          val = val > (T)1e-6 ? val : (T)0;
          val = val < (T)1 ? val : (T)1;
          if (solid) val = val * (T)0;
        }
        cout[(size_t)ch * ny * nx + o] = val;
      }
    }
  }
}
template <typename T, int R, typename D = Dev<T>>
__global__ void __launch_bounds__(256) k_final(D d, const T *uh, const T *vh, T *u_out, T *v_out,
                                              const T *th_new, const T *cin, T *cout, const T *t_in) {
  // tiles clear of the body window (2 cells around it) run without any window code
  const int tj0 = blockIdx.y * R, ti0 = blockIdx.x * TX;
  const bool touch = d.win && (d.periodic || (tj0 - 2 <= d.j1 && tj0 + R + 2 >= d.j0 && ti0 - 2 <= d.i1 &&
                                              ti0 + TX + 2 >= d.i0));
  if (touch) final_body<T, R, true>(d, blockIdx.x, blockIdx.y, uh, vh, u_out, v_out, th_new, cin, cout, t_in);
  else final_body<T, R, false>(d, blockIdx.x, blockIdx.y, uh, vh, u_out, v_out, th_new, cin, cout, t_in);
}

// Stage trace (core.py:897-934): from a stage's u^ and potential, the velocity as the
// reference's trace holds it after penalisation (outlet face corrected), the projection's
// right-hand side and the projected velocity (u^ - grad psi, boundary-normal faces untouched).
template <typename T>
__global__ void __launch_bounds__(256) k_trace_stage(Dev<T> d, const T *__restrict__ uh, const T *__restrict__ vh,
                                                    const T *__restrict__ psi, T *pen_u, T *pen_v, T *div, T *pu,
                                                    T *pv) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  const int nx = d.nx, ny = d.ny;
  const bool tunnel = !(d.periodic || d.closed);
  const T corr = tunnel ? d.red->corr : (T)0;
  if (j < ny && i <= nx) {
    const int o = j * (nx + 1) + i;
    if (pen_u) pen_u[o] = (tunnel && i == nx) ? uh[o] + corr : uh[o];
    if (pu) pu[o] = u_in(d, uh, psi, corr, j, i);
  }
  if (j <= ny && i < nx) {
    const int o = j * nx + i;
    if (pen_v) pen_v[o] = vh[o];
    if (pv) pv[o] = v_in(d, vh, psi, j, i);
  }
  if (div && j < ny && i < nx) div[j * nx + i] = div_at(d, uh, vh, j, i, corr);
}

template <typename T>
__global__ void k_copy(T *__restrict__ dst, const T *__restrict__ src, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// Per-step outputs (core.py:942-1004, physics.py:760-795) and the next step's dt.
template <typename T>
__device__ __forceinline__ void finalize_body(const Dev<T> &d, const T *__restrict__ t_in, T *__restrict__ t_out, T *__restrict__ dt_out, const T *__restrict__ u_new, const T *__restrict__ th_new, const aero2d_outputs &o) {
  __shared__ double lds[16 * (kMaxWin + 11)];
  __shared__ double res[kMaxWin + 11];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nx = d.nx, ny = d.ny, nxp = nx + 1;
  const int w = tid >> 6, nw = nthr >> 6;
  // every load this block needs is issued here, before any reduction waits on it
  T dt = 0, tin = 0, kap = 0, tin_th = 0, wv[4] = {0, 0, 0, 0};
  double th_old = 0;
  long long hist_n = 0, aux_n = 0;
  DtIn<T> din;
  if (tid == 0) {
    dt = d.red->dt;
    tin = *t_in;
    th_old = d.red->th_old;
    din = load_dt_in(d);
    if (o.hist) hist_n = *reinterpret_cast<const long long *>(o.hist_count);
    if (d.thermal && o.aux) aux_n = *reinterpret_cast<const long long *>(o.aux_count);
    if (d.thermal) {
      kap = *d.kappa;
      tin_th = *d.theta_in;
      for (int q = 0; q < 4; ++q) wv[q] = d.wall_values[q];
    }
  }
  // the three load phases run in different waves at once (each costs a memory round trip):
  // wall sums in the first quarter of the block, window partials in the second, the per-block
  // dt inputs in the second half
  const int r1 = nthr >> 2, r2 = nthr >> 1;
  double ws[6] = {0, 0, 0, 0, 0, 0};
  if (d.thermal && tid < r1) {
    const T tin_all = *d.theta_in;
    T w0 = d.wall_values[0], w1 = d.wall_values[1], w2 = d.wall_values[2], w3 = d.wall_values[3];
#pragma unroll 2
    for (int jj = tid; jj < ny; jj += r1) {
      const T a0 = th_new[jj * nx], a1 = th_new[jj * nx + nx - 1];
      const T b0 = u_new[jj * nxp], b1 = u_new[jj * nxp + nx];
      ws[0] += (double)(w0 - a0);
      ws[1] += (double)(w1 - a1);
      ws[4] += (double)(b0 * tin_all);
      ws[5] += (double)(b1 * a1);
    }
#pragma unroll 2
    for (int ii = tid; ii < nx; ii += r1) {
      const T c0 = th_new[ii], c1 = th_new[(ny - 1) * nx + ii];
      ws[2] += (double)(w2 - c0);
      ws[3] += (double)(w3 - c1);
    }
  }
  // window partials of the blocks that own window faces, and the per-block dt inputs: all
  // loads issued before the first barrier
  double acc[kMaxWin];
  for (int q = 0; q < kMaxWin; ++q) acc[q] = 0;
  if (d.win && tid >= r1 && tid < r2) {
    int bx0 = d.i0 / TX, bx1 = (d.i1 < nx ? d.i1 : nx - 1) / TX;
    int by0 = d.j0 / d.ty, by1 = (d.j1 < ny ? d.j1 : ny - 1) / d.ty;
    if (d.periodic) { bx0 = 0; by0 = 0; }
    const int nbx = bx1 - bx0 + 1, nb = nbx * (by1 - by0 + 1);
#pragma unroll 2
    for (int e = tid - r1; e < nb; e += r2 - r1) {
      const int bx = bx0 + e % nbx, by = by0 + e / nbx;
      const double *p = d.pforce + kMaxWin * (by * d.nbx + bx);
      for (int q = 0; q < kMaxWin; ++q) acc[q] += p[q];
    }
  }
  double pm[4] = {0, 0, 0, 0}, psum = 0;
  const int nblocks = d.nbx * d.nby;
#pragma unroll 4
  for (int b = tid - r2; tid >= r2 && b < nblocks; b += nthr - r2) {
    const double *p = d.pmax + 5 * b;
    for (int q = 0; q < 4; ++q) pm[q] = fmax(pm[q], p[q]);
    psum += p[4];
  }
  // wave reductions, one barrier, then one thread per quantity adds the waves in order
  constexpr int kW = kMaxWin + 5 + 6;   // window sums, dt maxima + sum theta, wall sums
  // only the slots in use (4 per body, 1 per temperature or heat-flux body) are reduced
  const int nq = 4 * d.nbody + (d.thermal ? d.n_temp + (d.win ? d.n_flux : 0) : 0);
  // each wave reduces only the group its lanes loaded; the other groups are zero in every lane
  // of it, which is what their reduction would give
  const int t0 = tid & ~63;
  if (t0 + 63 >= r1 && t0 < r2) {
#pragma unroll
    for (int q = 0; q < kMaxWin; ++q)
      if (q < nq) acc[q] = wave_sum(acc[q]);
  }
  if (t0 + 63 >= r2) {
    for (int q = 0; q < 4; ++q) pm[q] = wave_max(pm[q]);
    psum = wave_sum(psum);
  }
  if (d.thermal && t0 < r1)
    for (int q = 0; q < 6; ++q) ws[q] = wave_sum(ws[q]);
  if ((tid & 63) == 0) {
    for (int q = 0; q < kMaxWin; ++q) lds[w * kW + q] = acc[q];
    for (int q = 0; q < 4; ++q) lds[w * kW + kMaxWin + q] = pm[q];
    lds[w * kW + kMaxWin + 4] = psum;
    for (int q = 0; q < 6; ++q) lds[w * kW + kMaxWin + 5 + q] = ws[q];
  }
  __syncthreads();
  if (tid < kW) {
    const bool is_max = tid >= kMaxWin && tid < kMaxWin + 4;
    double r = is_max ? 0.0 : lds[tid];
    if (is_max)
      for (int ww = 0; ww < nw; ++ww) r = fmax(r, lds[ww * kW + tid]);
    else
      for (int ww = 1; ww < nw; ++ww) r += lds[ww * kW + tid];
    res[tid] = r;
  }
  __syncthreads();
  double mx[5];
  for (int q = 0; q < 5; ++q) mx[q] = res[kMaxWin + q];
  if (d.thermal)
    for (int q = 0; q < 6; ++q) ws[q] = res[kMaxWin + 5 + q];
  if (tid != 0) return;
  const T tnew = tin + dt;
  *t_out = tnew;
  *dt_out = dt;
  const T scale = (T)(-(d.hd * d.hd)) / ((dt * (T)0.5) * (T)d.force_ref);
  T sumc[3] = {0, 0, 0};
  for (int b = 0; b < d.nbody; ++b) {
    const T cd = (T)res[4 * b] * scale, cl = (T)res[4 * b + 2] * scale;
    const T cm = ((T)res[4 * b + 1] + (T)res[4 * b + 3]) * scale;
    if (o.coeffs) {
      T *cf = reinterpret_cast<T *>(o.coeffs);
      cf[3 * b] = cd;
      cf[3 * b + 1] = cl;
      cf[3 * b + 2] = cm;
    }
    sumc[0] += cd;
    sumc[1] += cl;
    sumc[2] += cm;
  }
  if (o.hist) {
    long long *cnt = reinterpret_cast<long long *>(o.hist_count);
    const long long row = hist_n % o.hist_cap;
    T *hrow = reinterpret_cast<T *>(o.hist) + row * 4;
    hrow[0] = tnew;
    hrow[1] = sumc[0];
    hrow[2] = sumc[1];
    hrow[3] = sumc[2];
    *cnt = hist_n + 1;
  }
  if (d.thermal) {
    // row layout: heat of the temperature bodies, skin temperature of the heat-flux bodies,
    // wall heat (4), storage, advective in / out. Each value goes straight to its outputs (a
    // local array indexed by the body counts would live in scratch memory).
    const int nt = d.n_temp, nf = d.win ? d.n_flux : 0, w0 = nt + nf;
    const T h2dt = (T)(d.hd * d.hd) / dt;
    T *arow = nullptr;
    if (o.aux) arow = reinterpret_cast<T *>(o.aux) + (aux_n % o.aux_cap) * (1 + w0 + 7) + 1;
    auto put = [&](int col, T val, void *dst, int at) {
      if (dst) reinterpret_cast<T *>(dst)[at] = val;
      if (arow) arow[col] = val;
    };
    for (int b = 0; b < nt; ++b) put(b, (T)res[4 * d.nbody + b] * h2dt, o.heat, b);
    for (int b = 0; b < nf; ++b) put(nt + b, (T)res[4 * d.nbody + nt + b], o.theta_wall, b);
    const int len[4] = {ny, ny, nx, nx};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = d.wm[q];
      T val = 0;
      if (m == AERO_TH_TEMPERATURE || m == AERO_TH_INLET) val = ((T)2 * kap) * (T)ws[q];
      else if (m == AERO_TH_HEAT_FLUX) val = kap * wv[q] * (T)(len[q] * d.hd);
      put(w0 + q, val, o.wall_heat, q);
    }
    put(w0 + 4, (T)(mx[4] - th_old) * h2dt, o.heat_storage, 0);
    const bool tunnel = !d.periodic && !d.closed;
    put(w0 + 5, tunnel ? (T)ws[4] * (T)d.hd : (T)0, o.flow_heat, 0);
    put(w0 + 6, tunnel ? (T)ws[5] * (T)d.hd : (T)0, o.flow_heat, 1);
    if (arow) {
      arow[-1] = tnew;
      *reinterpret_cast<long long *>(o.aux_count) = aux_n + 1;
    }
  }
  // the next step (chained) starts from this state
  T dt_next = compute_dt(d, din, (T)mx[0], (T)mx[1], (T)mx[2], (T)mx[3]);
  if (d.pid && !d.fixed_dt) {
    // this step's controller memory goes to the state; the next step's is computed from it
    const T c0 = d.red->ctrl[0], c1 = d.red->ctrl[1];
    if (d.ctrl_out) {
      d.ctrl_out[0] = c0;
      d.ctrl_out[1] = c1;
    }
    const T cur[2] = {c0, c1};
    dt_next = pid_step(dt_next, dt, cur, d.red->ctrl);
  }
  d.red->dt = dt_next;
  d.red->th_old = mx[4];
}
template <typename T>
__global__ void __launch_bounds__(1024) k_finalize(Dev<T> d, const T *__restrict__ t_in, T *__restrict__ t_out,
                                                   T *__restrict__ dt_out, const T *__restrict__ u_new,
                                                   const T *__restrict__ th_new, aero2d_outputs o) {
  finalize_body<T>(d, t_in, t_out, dt_out, u_new, th_new, o);
}

// ------------------------------------------------------------------ diagnostics
// One diagnostics row of a step (ns2d/diagnostics.py:70-101): per-block partials of the
// continuity residual of the new state, of the last stage's Poisson residual L psi - f and
// right-hand side f, and of the change of u, v, p, theta over the step; then one block forms
// the float64 row. Sums and maxima accumulate in float64 like the reference's norms.
constexpr int kDiagBlocks2 = 256;
constexpr int kDV2 = 18;   // 0 sum div^2, 1 max|div|, 2 sum (L psi - f)^2, 3 sum f^2, 4 max|u|, 5 max|v| (new),
                           // 6 + 3f: sum d^2, max|d|, sum a^2 of field f (u, v, p, theta) with d = new - old
__device__ __forceinline__ bool diag2_is_max(int q) { return q == 1 || q == 4 || q == 5 || (q >= 6 && (q - 6) % 3 == 1); }

template <typename T>
struct Diag2 {
  const T *uo, *vo, *po, *to;   // state before the step
  const T *un, *vn, *pn, *tn;   // after
  const T *uh, *vh;             // last stage's u^ (its projection's right-hand side)
  const T *t_out, *dt_out;
  double *part;                 // kDiagBlocks2 x kDV2 partials
  double *row_buf;
  long long *count;
  long long cap;
};

// The discrete Laplacian of the Poisson solve at cell (j, i): Neumann walls mirror the edge
// cell, periodic boxes wrap (ns2d/poisson.py:136-144).
template <typename T>
__device__ __forceinline__ T lap_at(const Dev<T> &d, const nd<T> *__restrict__ psi, int j, int i) {
  const int nx = d.nx, ny = d.ny;
  int ie, iw, jn, js;
  if (d.periodic) {
    ie = i + 1 == nx ? 0 : i + 1;
    iw = i == 0 ? nx - 1 : i - 1;
    jn = j + 1 == ny ? 0 : j + 1;
    js = j == 0 ? ny - 1 : j - 1;
  } else {
    ie = i + 1 < nx ? i + 1 : nx - 1;
    iw = i > 0 ? i - 1 : 0;
    jn = j + 1 < ny ? j + 1 : ny - 1;
    js = j > 0 ? j - 1 : 0;
  }
  const T e = psi[j * nx + ie], w = psi[j * nx + iw], n = psi[jn * nx + i], s = psi[js * nx + i];
  return ((((e + w) + n) + s) - (T)4 * psi[j * nx + i]) / (T)(d.hd * d.hd);
}

template <typename T>
__global__ void __launch_bounds__(256) k_diag2(Dev<T> d, Diag2<T> g) {
  const int nx = d.nx, ny = d.ny;
  const size_t nc = (size_t)nx * ny;
  const size_t stride = (size_t)gridDim.x * blockDim.x, t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kDV2];
#pragma unroll
  for (int q = 0; q < kDV2; ++q) v[q] = 0.0;
  const T corr = (d.periodic || d.closed) ? (T)0 : d.red->corr;
  const bool th = d.thermal && g.tn && g.to;
  for (size_t e = t0; e < nc; e += stride) {
    const int i = (int)(e % nx), j = (int)(e / nx);
    const double dv = (double)div_at(d, g.un, g.vn, j, i, (T)0);
    v[0] += dv * dv;
    v[1] = fmax(v[1], fabs(dv));
    const T f = div_at(d, g.uh, g.vh, j, i, corr);
    const double r = (double)(lap_at(d, d.psi, j, i) - f);
    v[2] += r * r;
    v[3] += (double)f * (double)f;
    const T pn = g.pn[e];
    const double dp = (double)(pn - g.po[e]);
    v[12] += dp * dp;
    v[13] = fmax(v[13], fabs(dp));
    v[14] += (double)pn * (double)pn;
    if (th) {
      const T tn = g.tn[e];
      const double dth = (double)(tn - g.to[e]);
      v[15] += dth * dth;
      v[16] = fmax(v[16], fabs(dth));
      v[17] += (double)tn * (double)tn;
    }
  }
  const size_t nf[2] = {(size_t)ny * (nx + 1), (size_t)(ny + 1) * nx};
  const T *fo[2] = {g.uo, g.vo}, *fn[2] = {g.un, g.vn};
#pragma unroll
  for (int f = 0; f < 2; ++f)
    for (size_t e = t0; e < nf[f]; e += stride) {
      const T an = fn[f][e];
      const double dd = (double)(an - fo[f][e]);
      v[6 + 3 * f] += dd * dd;
      v[7 + 3 * f] = fmax(v[7 + 3 * f], fabs(dd));
      v[8 + 3 * f] += (double)an * (double)an;
      v[4 + f] = fmax(v[4 + f], (double)fabs(an));
    }
  __shared__ double red[kDV2 * 4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < kDV2; ++q) v[q] = diag2_is_max(q) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < kDV2; ++q) red[wv * kDV2 + q] = v[q];
  __syncthreads();
  if (threadIdx.x < kDV2) {
    const int q = threadIdx.x;
    double r = red[q];
    for (int w2 = 1; w2 < 4; ++w2) r = diag2_is_max(q) ? fmax(r, red[w2 * kDV2 + q]) : r + red[w2 * kDV2 + q];
    g.part[(size_t)blockIdx.x * kDV2 + q] = r;
  }
}

template <typename T>
__global__ void __launch_bounds__(64) k_diag2_fin(Dev<T> d, Diag2<T> g) {
  __shared__ double tot[kDV2];
  if (threadIdx.x < kDV2) {
    const int q = threadIdx.x;
    double r = 0.0;
    for (int b = 0; b < kDiagBlocks2; ++b) {
      const double x = g.part[(size_t)b * kDV2 + q];
      r = diag2_is_max(q) ? fmax(r, x) : r + x;
    }
    tot[q] = r;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int nx = d.nx, ny = d.ny;
  const double nc = (double)nx * ny;
  const double numel[4] = {(double)ny * (nx + 1), (double)(ny + 1) * nx, nc, nc};
  const double h = d.hd;
  const double dt = (double)*g.dt_out;
  const long long n = *g.count;
  double *row = g.row_buf + (size_t)(n % g.cap) * AERO2D_DIAG_COLS;
  row[0] = (double)*g.t_out;
  row[1] = dt;
  row[2] = dt * (tot[4] + tot[5]) / h;
  row[3] = sqrt(tot[0]) / sqrt(nc);
  row[4] = tot[1];
  row[5] = sqrt(tot[2]) / fmax(sqrt(tot[3]), 1e-30);
  for (int f = 0; f < 4; ++f) {
    double *o = row + 6 + 4 * f;
    if (f == 3 && !(d.thermal && g.tn && g.to)) {
      o[0] = o[1] = o[2] = o[3] = (double)NAN;
      continue;
    }
    const double nrm = sqrt(tot[6 + 3 * f]), l2 = nrm * h;
    o[0] = l2;
    o[1] = nrm / sqrt(numel[f]);
    o[2] = tot[7 + 3 * f];
    o[3] = l2 / fmax(sqrt(tot[8 + 3 * f]) * h, 1e-300);
  }
  *g.count = n + 1;
}


// ================================================================== batched ensemble
// Many independent runs of uneven size advance together: every launch covers the blocks of
// all runs through a work table of (run, block x, block y) items, and each run reads its own
// static description (Dev) and launch pointers (BRun) from device memory. The kernel bodies
// are those of the single-run launches, so each run's arithmetic is the same.
template <typename T>
struct BRun {
  Stage<T> st[3];
  gptr<const T> u0, v0, th0;        // step start (in place: also the step's output)
  gptr<T> p_out;
  gptr<const T> uC, vC;              // last stage's u^
  gptr<T> u_out, v_out;
  gptr<const T> th_new, t_in;
  gptr<T> t_out, dt_out;
  aero2d_outputs o;
  int nslotA, nslotB, mode;   // row pairs per pass A / C block, column pairs per pass B block
};

// D: the tile kernels' view of every run's plan (Dev<T> itself: generic); the run's plan is
// copied into it, and only the fields the view's code reads stay live
template <typename T, int R, typename D = Dev<T>>
__global__ void __launch_bounds__(256) kb_pre(const Dev<T> *__restrict__ devs, const BRun<T> *__restrict__ runs,
                                             const int4 *__restrict__ items) {
  const int4 it = items[blockIdx.x];
  const BRun<T> &r = runs[it.x];
  const D d = view_of<D>(devs[it.x]);
  pre_body<T, R>(d, it.y, it.z, r.u0, r.v0, r.th0);
}

template <typename T>
__global__ void __launch_bounds__(1024) kb_dt(const Dev<T> *__restrict__ devs) {
  dt_body<T>(devs[blockIdx.x]);
}

template <typename T, int R, typename D = Dev<T>>
__global__ void __launch_bounds__(256) kb_stage(const Dev<T> *__restrict__ devs, const BRun<T> *__restrict__ runs,
                                               const int4 *__restrict__ items, int stage) {
  const int4 it = items[blockIdx.x];
  const D d = view_of<D>(devs[it.x]);
  // tiles clear of the run's body window (2 cells around it) run without any window code
  const int tj0 = it.z * R, ti0 = it.y * TX;
  const bool touch = d.win && (d.periodic || (tj0 - 2 <= d.j1 && tj0 + R + 2 >= d.j0 && ti0 - 2 <= d.i1 &&
                                              ti0 + TX + 2 >= d.i0));
  if (touch) stage_body<T, R, true>(d, runs[it.x].st[stage], it.y, it.z);
  else stage_body<T, R, false>(d, runs[it.x].st[stage], it.y, it.z);
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) kb_poisA(const Dev<T> *__restrict__ devs, const BRun<T> *__restrict__ runs,
                                               const int4 *__restrict__ items, int stage) {
  const int4 it = items[blockIdx.x];
  const BRun<T> &r = runs[it.x];
  poisA_body<T, NT>(devs[it.x], it.y, r.nslotA, r.st[stage].uh, r.st[stage].vh, (const T *)nullptr, r.mode);
}

template <typename T>
__global__ void __launch_bounds__(kFftThreads) kb_poisB(const Dev<T> *__restrict__ devs,
                                                        const BRun<T> *__restrict__ runs,
                                                        const int4 *__restrict__ items) {
  const int4 it = items[blockIdx.x];
  poisB_body<T, kFftThreads>(devs[it.x], it.y, runs[it.x].nslotB);
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) kb_poisC(const Dev<T> *__restrict__ devs, const BRun<T> *__restrict__ runs,
                                               const int4 *__restrict__ items, int stage, int write_p) {
  const int4 it = items[blockIdx.x];
  const Dev<T> &d = devs[it.x];
  poisC_body<T, NT>(d, it.y, runs[it.x].nslotA, d.psi, write_p ? runs[it.x].p_out : (T *)nullptr, stage);
}

template <typename T, int R, typename D = Dev<T>>
__global__ void __launch_bounds__(256) kb_final(const Dev<T> *__restrict__ devs, const BRun<T> *__restrict__ runs,
                                               const int4 *__restrict__ items) {
  const int4 it = items[blockIdx.x];
  const BRun<T> &r = runs[it.x];
  const D d = view_of<D>(devs[it.x]);
  const int tj0 = it.z * R, ti0 = it.y * TX;
  const bool touch = d.win && (d.periodic || (tj0 - 2 <= d.j1 && tj0 + R + 2 >= d.j0 && ti0 - 2 <= d.i1 &&
                                              ti0 + TX + 2 >= d.i0));
  if (touch)
    final_body<T, R, true>(d, it.y, it.z, r.uC, r.vC, r.u_out, r.v_out, r.th_new, (const T *)nullptr, (T *)nullptr,
                           r.t_in);
  else
    final_body<T, R, false>(d, it.y, it.z, r.uC, r.vC, r.u_out, r.v_out, r.th_new, (const T *)nullptr,
                            (T *)nullptr, r.t_in);
}

template <typename T>
__global__ void __launch_bounds__(1024) kb_finalize(const Dev<T> *__restrict__ devs,
                                                    const BRun<T> *__restrict__ runs) {
  const BRun<T> &r = runs[blockIdx.x];
  finalize_body<T>(devs[blockIdx.x], r.t_in, r.t_out, r.dt_out, (const T *)r.u_out, r.th_new, r.o);
}

}  // namespace ns2d
}  // namespace aero

// ====================================================================== host side
using namespace aero;
using namespace aero::ns2d;

struct aero2d_plan {
  aero2d_desc desc;
  int dbl;
  size_t esize;
  size_t bytes = 0;
  std::vector<void *> allocs;
  void *uA, *vA, *uB, *vB, *uC, *vC, *thA, *thB, *F, *psi, *acc_u, *acc_v, *acc_t, *outcol, *dye_tmp;
  void *tw_x, *tw_y, *cs_x, *cs_y;
  double *lam_x, *lam_y, *pmax, *pforce;
  void *inv_lam;
  void *red;
  void *old_u, *old_v, *old_p, *old_t, *pdiag;   // diagnostics: the state before the step, partials
  FftPlan fx, fy;
  int nbx, nby;
  int ty;       // rows per tile (4, 8 or 16)
  int nslotB;
  int ntA, ntB; // threads of the row passes (A, C) and of the column pass (B): 256, 512 or 1024
};

namespace {

// calls f(std::integral_constant<int, R>) for the tile rows R = ty (4, 8 or 16)
template <typename F>
void by_rows(int ty, F &&f) {
  switch (ty) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
  }
}

template <typename X>
struct Tag {
  using type = X;
};

// calls f(std::integral_constant<int, R>, Tag<D>) with the tile height R and the plan view D of
// the tile kernels: at the default height a view that fixes the LES model, heat transfer and (for
// plain wind-tunnel plans) the boundary and feature flags; at other heights the generic kernel
template <typename T, typename F>
void by_tile(const aero2d_plan *p, F &&f) {
  const aero2d_desc &s = p->desc;
  if (p->ty != kDefaultRows) {
    by_rows(p->ty, [&](auto r) { f(r, Tag<Dev<T>>{}); });
    return;
  }
  const std::integral_constant<int, kDefaultRows> r{};
  const bool plain = !s.periodic && !s.closed && s.advection == 1 && !s.wall_model && !s.has_heat &&
                     !(s.thermal && s.n_flux > 0);
  auto pick = [&](auto les) {
    constexpr int L = decltype(les)::value;
    if (plain) {
      if (s.thermal) f(r, Tag<DevK<T, L, 1, 1>>{});
      else f(r, Tag<DevK<T, L, 0, 1>>{});
    } else {
      if (s.thermal) f(r, Tag<DevK<T, L, 1, 0>>{});
      else f(r, Tag<DevK<T, L, 0, 0>>{});
    }
  };
  switch (s.turbulence) {
    case 1: pick(std::integral_constant<int, 1>{}); break;
    case 2: pick(std::integral_constant<int, 2>{}); break;
    case 3: pick(std::integral_constant<int, 3>{}); break;
    default: pick(std::integral_constant<int, 0>{}); break;
  }
}


// calls f(std::integral_constant<int, NT>) for the transform block size NT = nt (256, 512, 1024)
template <typename F>
void by_threads(int nt, F &&f) {
  switch (nt) {
    case 1024: f(std::integral_constant<int, 1024>{}); break;
    case 512: f(std::integral_constant<int, 512>{}); break;
    default: f(std::integral_constant<int, 256>{}); break;
  }
}

// One transform element per thread where the length allows it: the row and column passes run
// one block per row (column) pair, so a block of about n threads keeps every wave busy.
int transform_threads(size_t n) { return n <= 256 ? 256 : (n <= 512 ? 512 : 1024); }

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
  std::vector<double> lamx, lamy;
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
    std::vector<cplx<T>> tws(n);
    stage_twiddles(axis == 0 ? p->fx : p->fy, tw.data(), tws.data());
    if ((rc = put(axis == 0 ? p->tw_x : p->tw_y, tws.data(), n * sizeof(cplx<T>)))) return rc;
    if ((rc = put(axis == 0 ? p->cs_x : p->cs_y, cs.data(), n * sizeof(cplx<T>)))) return rc;
    if ((rc = put(axis == 0 ? (void *)p->lam_x : (void *)p->lam_y, lam.data(), n * sizeof(double)))) return rc;
    (axis == 0 ? lamx : lamy) = lam;
  }
  // the spectral scaling 1 / (lam_x + lam_y), formed in float64 and rounded once (ns2d/poisson.py:115-125)
  std::vector<T> inv((size_t)d.nx * d.ny);
  for (int i = 0; i < d.nx; ++i)
    for (int k = 0; k < d.ny; ++k) {
      const double la = lamx[i] + lamy[k];
      inv[(size_t)i * d.ny + k] = (T)((la != 0.0 && !(i == 0 && k == 0)) ? 1.0 / la : 0.0);
    }
  return put(p->inv_lam, inv.data(), inv.size() * sizeof(T));
}

template <typename T>
size_t lds_A(const aero2d_plan *p);
template <typename T>
size_t lds_B(const aero2d_plan *p);

// Rows of more than ~1000 float64 values take more than 64 KB of LDS in the x passes: the
// transform kernels are then allowed the device's whole per-block LDS (one value for every
// plan, so that plans of different sizes never lower each other's limit).
template <typename T>
int lds_limits(const aero2d_plan *p) {
  const size_t a = lds_A<T>(p), b = lds_B<T>(p);
  if (a <= 65536 && b <= 65536) return AERO_OK;
  int dev = 0, max_lds = 0;
  AERO_HIP_CHECK(hipGetDevice(&dev));
  AERO_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (a > (size_t)max_lds || b > (size_t)max_lds) {
    set_error("the device offers " + std::to_string(max_lds) + " B of LDS per block; the transforms need " +
              std::to_string(std::max(a, b)));
    return AERO_EUNSUPPORTED;
  }
  const void *fns[] = {(const void *)k_poisA<T, 256>, (const void *)k_poisB<T, 256>, (const void *)k_poisC<T, 256>,
                       (const void *)k_poisA<T, 512>, (const void *)k_poisB<T, 512>, (const void *)k_poisC<T, 512>,
                       (const void *)k_poisA<T, 1024>, (const void *)k_poisB<T, 1024>, (const void *)k_poisC<T, 1024>,
                       (const void *)kb_poisA<T, 256>, (const void *)kb_poisB<T>, (const void *)kb_poisC<T, 256>,
                       (const void *)kb_poisA<T, 512>, (const void *)kb_poisC<T, 512>};
  for (const void *f : fns)
    AERO_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  return AERO_OK;
}

template <typename T>
Dev<T> make_dev(const aero2d_plan *p, const aero2d_geom *g, const aero2d_params *pr) {
  const aero2d_desc &s = p->desc;
  Dev<T> d;
  std::memset(&d, 0, sizeof(d));
  d.nx = s.nx;
  d.ny = s.ny;
  d.nbx = p->nbx;
  d.nby = p->nby;
  d.ty = p->ty;
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
  d.n_flux = s.thermal ? s.n_flux : 0;
  for (int q = 0; q < 4; ++q) d.wm[q] = s.wall_modes[q];
  d.dye_mode = s.dye_mode;
  d.n_dye = s.n_dye;
  d.n_groups = s.n_groups;
  for (int q = 0; q < AERO2D_MAX_GROUPS; ++q) d.gch[q] = s.group_channel[q];
  d.fixed_dt = s.fixed_dt;
  d.wallm = s.wall_model;
  d.wall_y = (T)(0.5 * s.h);
  d.pid = s.pid_dt;
  d.hd = s.h;
  d.cfl = s.cfl;
  d.force_ref = s.force_ref;
  d.prt = s.prt;
  d.inv_prt = (T)(1.0 / s.prt);
  d.two_h = (T)(2.0 / s.h);
  d.les_k2 = (T)(s.les_c * s.les_c * s.h * s.h);
  d.les_kv = (T)(s.les_c * s.h * s.h);
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
    d.chi_u = global_ptr((const T *)g->chi_u); d.chi_v = global_ptr((const T *)g->chi_v); d.chi_c = global_ptr((const T *)g->chi_c);
    d.a_u = global_ptr((const T *)g->a_u); d.a_v = global_ptr((const T *)g->a_v); d.k_u = global_ptr((const T *)g->k_u); d.k_v = global_ptr((const T *)g->k_v);
    d.chi_t = global_ptr((const T *)g->chi_t); d.a_t = global_ptr((const T *)g->a_t); d.w_t = global_ptr((const T *)g->w_t);
    d.vol_u = global_ptr((const T *)g->vol_u); d.vol_v = global_ptr((const T *)g->vol_v);
    d.ins_u = global_ptr((const T *)g->ins_u); d.ins_v = global_ptr((const T *)g->ins_v);
    d.q_skin = global_ptr((const T *)g->q_skin); d.w_f = global_ptr((const T *)g->w_f);
  }
  if (pr) {
    d.nu = global_ptr((const T *)pr->nu); d.eta = global_ptr((const T *)pr->eta); d.u_inlet = global_ptr((const T *)pr->u_inlet);
    d.dt_fixed = global_ptr((const T *)pr->dt_fixed); d.kappa = global_ptr((const T *)pr->kappa); d.ri = global_ptr((const T *)pr->ri);
    d.eta_t = global_ptr((const T *)pr->eta_t); d.theta_in = global_ptr((const T *)pr->theta_in);
    d.wall_values = global_ptr((const T *)pr->wall_values); d.heat = global_ptr((const T *)pr->heat); d.heat_max = global_ptr((const T *)pr->heat_max);
    d.dye_inflow = global_ptr((const T *)pr->dye_inflow); d.g_weight = global_ptr((const T *)pr->group_weight);
    d.g_target = global_ptr((const T *)pr->group_target); d.g_timing = global_ptr((const T *)pr->group_timing);
  }
  d.F = global_ptr((T *)p->F); d.psi = global_ptr((T *)p->psi);
  d.acc_u = global_ptr((T *)p->acc_u); d.acc_v = global_ptr((T *)p->acc_v); d.acc_t = global_ptr((T *)p->acc_t); d.outcol = global_ptr((T *)p->outcol);
  d.pmax = global_ptr(p->pmax); d.pforce = global_ptr(p->pforce);
  d.red = global_ptr((Red<T> *)p->red);
  d.tw_x = global_ptr((const cplx<T> *)p->tw_x); d.tw_y = global_ptr((const cplx<T> *)p->tw_y);
  d.cs_x = global_ptr((const cplx<T> *)p->cs_x); d.cs_y = global_ptr((const cplx<T> *)p->cs_y);
  d.lam_x = global_ptr(p->lam_x); d.lam_y = global_ptr(p->lam_y);
  d.inv_lam = global_ptr((const T *)p->inv_lam);
  d.fx = p->fx; d.fy = p->fy;
  d.dct_x = s.periodic ? 0 : 1;
  d.dct_y = s.periodic ? 0 : 1;
  return d;
}

// LDS of a pass A / C block over nslot row pairs of length nx: two work buffers of the pairs, the
// twiddles, then pass C's real rows (pass A's reduction scratch fits in their place)
template <typename T>
size_t lds_rows(size_t nx, size_t nslot) {
  return (2 * nslot + 1) * nx * sizeof(cplx<T>) + 2 * nslot * nx * sizeof(T) + 64 * 8;
}
template <typename T>
size_t lds_A(const aero2d_plan *p) { return lds_rows<T>(p->desc.nx, 1); }
// LDS of a pass B block over nslot column pairs of length ny: two work buffers and 2 nslot real
// rows at slot stride ny + 1, the twiddles
template <typename T>
size_t lds_cols(size_t ny, size_t nslot) {
  return nslot * (ny + 1) * (2 * sizeof(cplx<T>) + 2 * sizeof(T)) + ny * sizeof(cplx<T>);
}
template <typename T>
size_t lds_B(const aero2d_plan *p) { return lds_cols<T>(p->desc.ny, p->nslotB); }

// right-hand side -> psi (passes A, B, C) on `stream`
template <typename T>
void solve_passes(aero2d_plan *p, hipStream_t st, Dev<T> &d, const T *u, const T *v, const T *f, int mode, T *psi,
                  T *pacc, int stage) {
  const int ny = p->desc.ny, nx = p->desc.nx;
  by_threads(p->ntA, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    hipLaunchKernelGGL((k_poisA<T, NT>), dim3((ny + 1) / 2), dim3(NT), lds_A<T>(p), st, d, u, v, f, mode);
  });
  const int cols_per_block = 2 * p->nslotB;
  by_threads(p->ntB, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    hipLaunchKernelGGL((k_poisB<T, NT>), dim3((nx + cols_per_block - 1) / cols_per_block), dim3(NT), lds_B<T>(p),
                       st, d, p->nslotB);
  });
  by_threads(p->ntA, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    hipLaunchKernelGGL((k_poisC<T, NT>), dim3((ny + 1) / 2), dim3(NT), lds_A<T>(p), st, d, psi, pacc, stage);
  });
}

// one step; chained: the dt inputs come from the previous step of this plan
template <typename T>
int step_impl(aero2d_plan *p, hipStream_t st, const aero2d_fields *in, aero2d_fields *out, const aero2d_geom *g,
              const aero2d_params *pr, aero2d_outputs *o, bool chained, const aero2d_trace *tr = nullptr) {
  const aero2d_desc &s = p->desc;
  Dev<T> d = make_dev<T>(p, g, pr);
  d.dt_in = global_ptr((const T *)in->dt);
  d.ctrl_in = global_ptr((const T *)in->ctrl);
  d.ctrl_out = global_ptr((T *)out->ctrl);
  const int nx = s.nx, ny = s.ny;
  const dim3 blk(64, 4);
  const dim3 tiles(p->nbx, p->nby);
  const T *u0 = (const T *)in->u, *v0 = (const T *)in->v, *th0 = (const T *)in->theta;
  const bool diag = o && o->diag;
  if (diag) {
    // the state before the step (the step may run in place)
    const size_t nf[4] = {(size_t)ny * (nx + 1), (size_t)(ny + 1) * nx, (size_t)ny * nx, (size_t)ny * nx};
    void *dst[4] = {p->old_u, p->old_v, p->old_p, p->old_t};
    const void *src[4] = {in->u, in->v, in->p, in->theta};
    for (int f = 0; f < (s.thermal ? 4 : 3); ++f) {
      const int nb = (int)std::min<size_t>(4096, (nf[f] + 255) / 256);
      hipLaunchKernelGGL(k_copy<T>, dim3(nb), dim3(256), 0, st, (T *)dst[f], (const T *)src[f], nf[f]);
    }
  }
  if (!chained) {
    by_tile<T>(p, [&](auto r, auto tag) {
      using D = typename decltype(tag)::type;
      hipLaunchKernelGGL((k_pre<T, decltype(r)::value, D>), tiles, blk, 0, st, view_of<D>(d), u0, v0, th0);
    });
    hipLaunchKernelGGL(k_dt<T>, dim3(1), dim3(1024), 0, st, d);
  }
  T *ubuf[3] = {(T *)p->uA, (T *)p->uB, (T *)p->uC};
  T *vbuf[3] = {(T *)p->vA, (T *)p->vB, (T *)p->vC};
  T *tbuf[3] = {(T *)p->thA, (T *)p->thB, (T *)out->theta};
  const T *up = u0, *vp = v0, *thp = th0, *psi_p = nullptr;
  const int mode = (s.periodic || s.closed) ? 1 : 2;
  for (int k = 0; k < 3; ++k) {
    Stage<T> sg;
    sg.s = k;
    sg.a = kSsp[k][0];
    sg.b = kSsp[k][1];
    sg.w = kForceW[k];
    sg.up = global_ptr(up); sg.vp = global_ptr(vp); sg.psi_p = global_ptr(psi_p); sg.thp = global_ptr(thp);
    sg.u0 = global_ptr(u0); sg.v0 = global_ptr(v0); sg.th0 = global_ptr(th0);
    sg.uh = global_ptr(ubuf[k]); sg.vh = global_ptr(vbuf[k]); sg.thh = global_ptr(s.thermal ? tbuf[k] : nullptr);
    by_tile<T>(p, [&](auto r, auto tag) {
      using D = typename decltype(tag)::type;
      hipLaunchKernelGGL((k_stage<T, decltype(r)::value, D>), tiles, blk, 0, st, view_of<D>(d), sg);
    });
    solve_passes<T>(p, st, d, sg.uh, sg.vh, nullptr, mode, d.psi, (T *)out->p, k);
    if (tr) {
      const dim3 gf((nx + 1 + 63) / 64, (ny + 1 + 3) / 4);
      hipLaunchKernelGGL(k_trace_stage<T>, gf, blk, 0, st, d, (const T *)sg.uh, (const T *)sg.vh, (const T *)d.psi,
                         (T *)tr->pen_u[k], (T *)tr->pen_v[k], (T *)tr->div[k], (T *)tr->proj_u[k],
                         (T *)tr->proj_v[k]);
      const size_t nc = (size_t)nx * ny;
      const int cb = (int)std::min<size_t>(4096, (nc + 255) / 256);
      if (tr->psi[k]) hipLaunchKernelGGL(k_copy<T>, dim3(cb), dim3(256), 0, st, (T *)tr->psi[k], (const T *)d.psi, nc);
      if (tr->theta[k] && sg.thh)
        hipLaunchKernelGGL(k_copy<T>, dim3(cb), dim3(256), 0, st, (T *)tr->theta[k], (const T *)sg.thh, nc);
    }
    up = sg.uh; vp = sg.vh; thp = sg.thh; psi_p = d.psi;
  }
  const bool dye = s.dye_mode != 0 && in->dye && out->dye;
  const bool inplace = dye && in->dye == out->dye;
  T *dye_dst = dye ? (inplace ? (T *)p->dye_tmp : (T *)out->dye) : nullptr;
  by_tile<T>(p, [&](auto r, auto tag) {
    using D = typename decltype(tag)::type;
    hipLaunchKernelGGL((k_final<T, decltype(r)::value, D>), tiles, blk, 0, st, view_of<D>(d), (const T *)p->uC, (const T *)p->vC,
                       (T *)out->u, (T *)out->v, (const T *)out->theta, dye ? (const T *)in->dye : (const T *)nullptr,
                       dye_dst, (const T *)in->t);
  });
  if (inplace) {
    const size_t n = (size_t)s.n_dye * ny * nx;
    const int cb = (int)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_copy<T>, dim3(cb), dim3(256), 0, st, (T *)out->dye, (const T *)dye_dst, n);
  }
  aero2d_outputs oo;
  std::memset(&oo, 0, sizeof(oo));
  if (o) oo = *o;
  hipLaunchKernelGGL(k_finalize<T>, dim3(1), dim3(1024), 0, st, d, (const T *)in->t, (T *)out->t, (T *)out->dt,
                     (const T *)out->u, (const T *)out->theta, oo);
  if (diag) {
    Diag2<T> dg;
    dg.uo = (const T *)p->old_u; dg.vo = (const T *)p->old_v; dg.po = (const T *)p->old_p;
    dg.to = s.thermal ? (const T *)p->old_t : nullptr;
    dg.un = (const T *)out->u; dg.vn = (const T *)out->v; dg.pn = (const T *)out->p;
    dg.tn = s.thermal ? (const T *)out->theta : nullptr;
    dg.uh = (const T *)p->uC; dg.vh = (const T *)p->vC;
    dg.t_out = (const T *)out->t; dg.dt_out = (const T *)out->dt;
    dg.part = (double *)p->pdiag;
    dg.row_buf = (double *)o->diag;
    dg.count = (long long *)o->diag_count;
    dg.cap = o->diag_cap;
    hipLaunchKernelGGL(k_diag2<T>, dim3(kDiagBlocks2), dim3(256), 0, st, d, dg);
    hipLaunchKernelGGL(k_diag2_fin<T>, dim3(1), dim3(64), 0, st, d, dg);
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int project_impl(aero2d_plan *p, hipStream_t st, void *u, void *v, void *psi_out) {
  Dev<T> d = make_dev<T>(p, nullptr, nullptr);
  T *psi = psi_out ? (T *)psi_out : d.psi;
  solve_passes<T>(p, st, d, (const T *)u, (const T *)v, nullptr, 1, psi, nullptr, 0);
  const dim3 blk(64, 4);
  const dim3 grid_faces((p->desc.nx + 1 + 63) / 64, (p->desc.ny + 1 + 3) / 4);
  hipLaunchKernelGGL(k_correct<T>, grid_faces, blk, 0, st, d, (T *)u, (T *)v, (const T *)psi);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int poisson_impl(aero2d_plan *p, hipStream_t st, const void *f, void *psi) {
  Dev<T> d = make_dev<T>(p, nullptr, nullptr);
  solve_passes<T>(p, st, d, nullptr, nullptr, (const T *)f, 0, (T *)psi, nullptr, 0);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int time_stage_impl(aero2d_plan *p, hipStream_t st, const aero2d_fields *f, const aero2d_geom *g,
                    const aero2d_params *pr, int stage, int iters, double *avg_ms) {
  const aero2d_desc &s = p->desc;
  Dev<T> d = make_dev<T>(p, g, pr);
  Stage<T> sg;
  sg.s = stage;
  sg.a = kSsp[stage][0];
  sg.b = kSsp[stage][1];
  sg.w = kForceW[stage];
  // stage input = the previous stage's unprojected buffer and potential (as in a step)
  sg.up = global_ptr(stage > 0 ? (const T *)p->uB : (const T *)f->u);
  sg.vp = global_ptr(stage > 0 ? (const T *)p->vB : (const T *)f->v);
  sg.psi_p = global_ptr(stage > 0 ? (const T *)p->psi : nullptr);
  sg.thp = global_ptr((const T *)f->theta);
  sg.u0 = global_ptr((const T *)f->u); sg.v0 = global_ptr((const T *)f->v); sg.th0 = global_ptr((const T *)f->theta);
  sg.uh = global_ptr((T *)p->uA); sg.vh = global_ptr((T *)p->vA); sg.thh = global_ptr(s.thermal ? (T *)p->thA : nullptr);
  const dim3 blk(64, 4);
  const dim3 tiles(p->nbx, p->nby);
  hipEvent_t e0, e1;
  AERO_HIP_CHECK(hipEventCreate(&e0));
  AERO_HIP_CHECK(hipEventCreate(&e1));
  by_tile<T>(p, [&](auto r, auto tag) {
    constexpr int R = decltype(r)::value;
    using D = typename decltype(tag)::type;
    const D dk = view_of<D>(d);
    hipLaunchKernelGGL((k_stage<T, R, D>), tiles, blk, 0, st, dk, sg);   // warm
    (void)hipEventRecord(e0, st);
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL((k_stage<T, R, D>), tiles, blk, 0, st, dk, sg);
  });
  AERO_HIP_CHECK(hipEventRecord(e1, st));
  AERO_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  AERO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / (double)iters;
  return AERO_OK;
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
  const size_t ldsx = 3 * d->nx * 2 * es + 2 * d->nx * es + 512;
  const size_t ldsy = d->ny * (2 * 2 * es + 2 * es) + d->ny * 2 * es;
  if (d->nx > kMaxLen || d->ny > kMaxLen || ldsx > kLdsBlockMax || ldsy > kLdsBlockMax) {
    set_error("grid too large for the LDS-resident transforms");
    return AERO_EUNSUPPORTED;
  }
  if (d->n_groups > AERO2D_MAX_GROUPS || d->n_groups < 0) { set_error("too many emitter groups"); return AERO_EUNSUPPORTED; }
  if (d->dye_mode < 0 || d->dye_mode > 2) { set_error("unknown dye mode"); return AERO_EINVAL; }
  if (d->tile_rows != 0 && d->tile_rows != 4 && d->tile_rows != 8 && d->tile_rows != 16) {
    set_error("tile_rows must be 0 (default), 4, 8 or 16");
    return AERO_EINVAL;
  }
  const bool win = d->win_j1 > d->win_j0 && d->win_i1 > d->win_i0;
  if (win && (d->win_j0 < 0 || d->win_i0 < 0 || d->win_j1 > d->ny || d->win_i1 > d->nx)) {
    set_error("body window outside the grid");
    return AERO_EINVAL;
  }
  if (4 * d->nbody + (d->thermal ? d->n_temp + d->n_flux : 0) > kMaxWin) {
    set_error("too many bodies for the per-block window partials (4 per body + 1 per temperature or heat-flux "
              "body <= 16)");
    return AERO_EUNSUPPORTED;
  }
  if (d->n_flux < 0 || (d->n_flux > 0 && !d->thermal) || d->n_flux + d->n_temp > d->nbody) {
    set_error("bad heat-flux body count");
    return AERO_EINVAL;
  }
  return AERO_OK;
}

int check_step_args(aero2d_plan *p, const aero2d_fields *in, aero2d_fields *out, const aero2d_geom *g,
                    const aero2d_params *pr, aero2d_outputs *o) {
  if (!p || !in || !out || !pr) { set_error("null argument"); return AERO_EINVAL; }
  if (!in->u || !in->v || !out->u || !out->v || !out->p || !in->t || !out->t || !out->dt || !pr->nu || !pr->eta) {
    set_error("missing field or parameter pointer");
    return AERO_EINVAL;
  }
  const aero2d_desc &s = p->desc;
  if (!s.periodic && !s.closed && !pr->u_inlet) { set_error("tunnel needs u_inlet"); return AERO_EINVAL; }
  if (s.fixed_dt && !pr->dt_fixed) { set_error("fixed dt needs dt_fixed"); return AERO_EINVAL; }
  if (s.pid_dt && !s.fixed_dt && (!in->ctrl || !out->ctrl || !in->dt)) {
    set_error("the PID step needs the state's dt and controller memory (fields ctrl, dt)");
    return AERO_EINVAL;
  }
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
  if (s.thermal && s.n_flux > 0 && win && (!g->q_skin || !g->w_f)) {
    set_error("heat-flux bodies need q_skin and w_f");
    return AERO_EINVAL;
  }
  if (s.has_heat && (!pr->heat || !pr->heat_max)) { set_error("heaters need heat and heat_max"); return AERO_EINVAL; }
  if (s.dye_mode != 0 && in->dye && !pr->dye_inflow) { set_error("smoke needs dye_inflow"); return AERO_EINVAL; }
  if (s.dye_mode == 2 && s.n_groups > 0 && (!pr->group_weight || !pr->group_target || !pr->group_timing)) {
    set_error("emitter groups need weight, target and timing");
    return AERO_EINVAL;
  }
  if (o && o->hist && (!o->hist_count || o->hist_cap <= 0)) { set_error("history ring needs count and capacity"); return AERO_EINVAL; }
  if (o && o->aux && (!o->aux_count || o->aux_cap <= 0)) { set_error("aux ring needs count and capacity"); return AERO_EINVAL; }
  if (o && o->diag && (!s.diagnostics || !o->diag_count || o->diag_cap < 1 || !in->p)) {
    set_error("diagnostics rows need a plan made with desc.diagnostics, a row counter and fields.p");
    return AERO_EINVAL;
  }
  return AERO_OK;
}

}  // namespace

extern "C" {

const char *aero_last_error(void) { return g_error.c_str(); }
int aero_version(void) { return 2; }

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
  p->nbx = (int)((nx + TX - 1) / TX);
  p->ty = desc->tile_rows ? desc->tile_rows : kDefaultRows;
  p->nby = (int)((ny + p->ty - 1) / p->ty);
  // partial maxima and forces sized for the smallest tile (4 rows), so that a batched
  // launch may run this plan with any tile shape
  const size_t blocks = (size_t)p->nbx * ((ny + 3) / 4);
  // pass B: two column pairs per block when that still leaves a block for each of 256 CUs (at
  // 1024x512, with 1024-thread blocks: step 0.157 -> 0.152 ms; 512 threads gain nothing and four
  // pairs lose, 0.160 ms), one otherwise
  p->nslotB = nx >= 4 * 256 ? 2 : 1;
  p->ntA = transform_threads(nx);
  while (p->nslotB > 1 &&
         ((ny + 1) * p->nslotB * (2 * 2 * es + 2 * es) + ny * 2 * es > 65536 || ny * p->nslotB > (size_t)kMaxLen))
    p->nslotB--;
  p->ntB = transform_threads(p->nslotB * ny);
  struct { void **ptr; size_t n; } tab[] = {
      {&p->uA, nu_ * es}, {&p->vA, nv_ * es}, {&p->uB, nu_ * es}, {&p->vB, nv_ * es}, {&p->uC, nu_ * es},
      {&p->vC, nv_ * es}, {&p->thA, desc->thermal ? nc * es : 0}, {&p->thB, desc->thermal ? nc * es : 0},
      {&p->F, nc * es}, {&p->psi, nc * es},
      {&p->acc_u, wy * (wx + 1) * es}, {&p->acc_v, (wy + 1) * wx * es}, {&p->acc_t, wy * wx * es},
      {&p->outcol, ny * es}, {&p->dye_tmp, desc->dye_mode ? (size_t)desc->n_dye * nc * es : 0},
      {&p->tw_x, nx * 2 * es}, {&p->tw_y, ny * 2 * es}, {&p->cs_x, nx * 2 * es}, {&p->cs_y, ny * 2 * es},
      {(void **)&p->lam_x, nx * 8}, {(void **)&p->lam_y, ny * 8}, {&p->inv_lam, nc * es},
      {(void **)&p->pmax, blocks * 5 * 8},
      {(void **)&p->pforce, blocks * kMaxWin * 8}, {&p->red, 256},
      {&p->old_u, desc->diagnostics ? nu_ * es : 0}, {&p->old_v, desc->diagnostics ? nv_ * es : 0},
      {&p->old_p, desc->diagnostics ? nc * es : 0}, {&p->old_t, desc->diagnostics && desc->thermal ? nc * es : 0},
      {&p->pdiag, desc->diagnostics ? (size_t)kDiagBlocks2 * kDV2 * 8 : 0}};
  for (auto &t : tab) {
    if ((rc = alloc(p, t.ptr, t.n))) {
      aero2d_plan_destroy(p);
      return rc;
    }
  }
  rc = p->dbl ? upload_tables<double>(p) : upload_tables<float>(p);
  if (!rc) rc = p->dbl ? lds_limits<double>(p) : lds_limits<float>(p);
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

int aero2d_step_trace(aero2d_plan *p, void *stream, const aero2d_fields *in, aero2d_fields *out,
                      const aero2d_geom *g, const aero2d_params *pr, aero2d_outputs *o, const aero2d_trace *tr) {
  int rc = check_step_args(p, in, out, g, pr, o);
  if (rc) return rc;
  if (!tr) { set_error("aero2d_step_trace: null trace"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o, false, tr)
                : step_impl<float>(p, st, in, out, g, pr, o, false, tr);
}

int aero2d_step(aero2d_plan *p, void *stream, const aero2d_fields *in, aero2d_fields *out, const aero2d_geom *g,
                const aero2d_params *pr, aero2d_outputs *o) {
  int rc = check_step_args(p, in, out, g, pr, o);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o, false) : step_impl<float>(p, st, in, out, g, pr, o, false);
}

int aero2d_step_chained(aero2d_plan *p, void *stream, const aero2d_fields *in, aero2d_fields *out,
                        const aero2d_geom *g, const aero2d_params *pr, aero2d_outputs *o) {
  int rc = check_step_args(p, in, out, g, pr, o);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o, true) : step_impl<float>(p, st, in, out, g, pr, o, true);
}

int aero2d_time_stage(aero2d_plan *p, void *stream, const aero2d_fields *f, const aero2d_geom *g,
                      const aero2d_params *pr, int stage, int iters, double *avg_ms) {
  if (!p || !f || !pr || !avg_ms || !f->u || !f->v || iters <= 0 || stage < 0 || stage > 2) {
    set_error("bad argument");
    return AERO_EINVAL;
  }
  if (p->desc.thermal && !f->theta) { set_error("thermal plan needs theta"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? time_stage_impl<double>(p, st, f, g, pr, stage, iters, avg_ms)
                : time_stage_impl<float>(p, st, f, g, pr, stage, iters, avg_ms);
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

// ---------------------------------------------------------------- ensemble (host)
struct aero2d_ensemble {
  int n = 0, dbl = 0;
  int ty = kDefaultRows;   // rows per tile of the batched cell launches
  int nslot = 2;           // column pairs per block of the batched pass-B launch (capped per run;
                           // 0: as many as fill kRowFill elements)
  int ntrow = kFftThreads; // threads per block of the batched row passes (A, C): 256 or 512
  int rowfill = 1;         // 1: row passes take as many row pairs per block as fill kRowFill elements
  bool plain_laminar = false;   // every run a plain wind tunnel without LES or heat transfer
  std::vector<aero2d_plan *> plans;
  std::vector<aero2d_fields> fields;
  std::vector<aero2d_geom> geoms;
  std::vector<aero2d_params> params;
  std::vector<aero2d_outputs> outputs;   // empty: no outputs
  std::vector<void *> allocs;
  void *devs = nullptr, *runs = nullptr, *tiles = nullptr, *rows = nullptr, *cols = nullptr;
  int ntiles = 0, nrows = 0, ncols = 0;
  size_t ldsA = 0, ldsB = 0;
};

namespace {

int ens_alloc(aero2d_ensemble *e, void **ptr, const void *src, size_t n) {
  AERO_HIP_CHECK(hipMalloc(ptr, n));
  e->allocs.push_back(*ptr);
  AERO_HIP_CHECK(hipMemcpy(*ptr, src, n, hipMemcpyHostToDevice));
  return AERO_OK;
}

void ens_release(aero2d_ensemble *e) {
  for (void *a : e->allocs) (void)hipFree(a);
  e->allocs.clear();
  e->devs = e->runs = e->tiles = e->rows = e->cols = nullptr;
  e->ntiles = e->nrows = e->ncols = 0;
  e->ldsA = e->ldsB = 0;
}

// column pairs per pass-B block for run p: the largest of want, want / 2, ..., 1 whose
// columns fit the per-thread element count and 64 KB of LDS; want = 0: the most that keep the
// block's columns within kRowFill elements (a power of two)
template <typename T>
int run_slots(const aero2d_plan *p, int want) {
  const size_t ny = p->desc.ny;
  int s = want;
  if (want == 0) {
    s = 1;
    while ((size_t)(2 * s) * ny <= (size_t)kRowFill) s *= 2;
  }
  while (s > 1 && ((size_t)s * ny > (size_t)kMaxLen ||
                   lds_cols<T>(ny, s) > 65536))
    s /= 2;
  return s;
}

// row pairs per batched pass A / C block: one, or (fill) the most that keep the block's rows
// within kRowFill elements (a power of two, at least one)
template <typename T>
int run_row_slots(const aero2d_plan *p, int fill) {
  const size_t nx = p->desc.nx;
  int s = 1;
  while (fill && (size_t)(2 * s) * nx <= (size_t)kRowFill && lds_rows<T>(nx, 2 * s) <= 65536) s *= 2;
  return s;
}

// calls f(Tag<D>) with the view of the ensemble's tile kernels: the plain laminar view when every
// run is a plain wind tunnel without LES or heat transfer, the generic Dev<T> otherwise
template <typename T, typename F>
void by_ens_view(const aero2d_ensemble *e, F &&f) {
  if (e->plain_laminar) f(Tag<DevK<T, 0, 0, 1>>{});
  else f(Tag<Dev<T>>{});
}

bool plain_laminar(const aero2d_desc &s) {
  return s.turbulence == 0 && !s.thermal && !s.periodic && !s.closed && s.advection == 1 && !s.wall_model &&
         !s.has_heat;
}

// (re)builds the device run tables for the tile rows e->ty and column slots e->nslot
template <typename T>
int ensemble_build(aero2d_ensemble *e) {
  ens_release(e);
  const aero2d_fields *fields = e->fields.data();
  const aero2d_geom *geoms = e->geoms.data();
  const aero2d_params *params = e->params.data();
  const aero2d_outputs *outputs = e->outputs.empty() ? nullptr : e->outputs.data();
  std::vector<Dev<T>> devs(e->n);
  std::vector<BRun<T>> runs(e->n);
  std::vector<int4> tiles, rows, cols;
  for (int r = 0; r < e->n; ++r) {
    aero2d_plan *p = e->plans[r];
    const aero2d_desc &s = p->desc;
    devs[r] = make_dev<T>(p, &geoms[r], &params[r]);
    devs[r].dt_in = global_ptr((const T *)fields[r].dt);
    devs[r].ctrl_in = global_ptr((const T *)fields[r].ctrl);
    devs[r].ctrl_out = global_ptr((T *)fields[r].ctrl);
    const int nby = (s.ny + e->ty - 1) / e->ty;
    devs[r].ty = e->ty;
    devs[r].nby = nby;
    BRun<T> &b = runs[r];
    std::memset(&b, 0, sizeof(b));
    const aero2d_fields &f = fields[r];
    const T *u0 = (const T *)f.u, *v0 = (const T *)f.v, *th0 = (const T *)f.theta;
    T *ubuf[3] = {(T *)p->uA, (T *)p->uB, (T *)p->uC};
    T *vbuf[3] = {(T *)p->vA, (T *)p->vB, (T *)p->vC};
    T *tbuf[3] = {(T *)p->thA, (T *)p->thB, (T *)f.theta};
    const T *up = u0, *vp = v0, *thp = th0, *psi_p = nullptr;
    for (int k = 0; k < 3; ++k) {
      Stage<T> &sg = b.st[k];
      sg.s = k;
      sg.a = kSsp[k][0];
      sg.b = kSsp[k][1];
      sg.w = kForceW[k];
      sg.up = global_ptr(up); sg.vp = global_ptr(vp); sg.psi_p = global_ptr(psi_p); sg.thp = global_ptr(thp);
      sg.u0 = global_ptr(u0); sg.v0 = global_ptr(v0); sg.th0 = global_ptr(th0);
      sg.uh = global_ptr(ubuf[k]); sg.vh = global_ptr(vbuf[k]); sg.thh = global_ptr(s.thermal ? tbuf[k] : nullptr);
      up = sg.uh; vp = sg.vh; thp = sg.thh; psi_p = (const T *)p->psi;
    }
    b.u0 = global_ptr(u0); b.v0 = global_ptr(v0); b.th0 = global_ptr(th0);
    b.p_out = global_ptr((T *)f.p);
    b.uC = global_ptr((const T *)p->uC); b.vC = global_ptr((const T *)p->vC);
    b.u_out = global_ptr((T *)f.u); b.v_out = global_ptr((T *)f.v);
    b.th_new = global_ptr((const T *)f.theta);
    b.t_in = global_ptr((const T *)f.t); b.t_out = global_ptr((T *)f.t); b.dt_out = global_ptr((T *)f.dt);
    if (outputs) b.o = outputs[r];
    b.nslotB = run_slots<T>(p, e->nslot);
    b.nslotA = run_row_slots<T>(p, e->rowfill);
    b.mode = (s.periodic || s.closed) ? 1 : 2;
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < p->nbx; ++bx) tiles.push_back(make_int4(r, bx, by, 0));
    const int npairs = (s.ny + 1) / 2;
    for (int bx = 0; bx < (npairs + b.nslotA - 1) / b.nslotA; ++bx) rows.push_back(make_int4(r, bx, 0, 0));
    const int cpb = 2 * b.nslotB;
    for (int bx = 0; bx < (s.nx + cpb - 1) / cpb; ++bx) cols.push_back(make_int4(r, bx, 0, 0));
    e->ldsA = std::max(e->ldsA, lds_rows<T>(s.nx, b.nslotA));
    e->ldsB = std::max(e->ldsB, lds_cols<T>(s.ny, b.nslotB));
  }
  e->ntiles = (int)tiles.size();
  e->nrows = (int)rows.size();
  e->ncols = (int)cols.size();
  int rc;
  if ((rc = ens_alloc(e, &e->devs, devs.data(), devs.size() * sizeof(Dev<T>)))) return rc;
  if ((rc = ens_alloc(e, &e->runs, runs.data(), runs.size() * sizeof(BRun<T>)))) return rc;
  if ((rc = ens_alloc(e, &e->tiles, tiles.data(), tiles.size() * sizeof(int4)))) return rc;
  if ((rc = ens_alloc(e, &e->rows, rows.data(), rows.size() * sizeof(int4)))) return rc;
  if ((rc = ens_alloc(e, &e->cols, cols.data(), cols.size() * sizeof(int4)))) return rc;
  return AERO_OK;
}

template <typename T>
int ensemble_step(aero2d_ensemble *e, hipStream_t st, bool chained) {
  const Dev<T> *devs = (const Dev<T> *)e->devs;
  const BRun<T> *runs = (const BRun<T> *)e->runs;
  const int4 *tiles = (const int4 *)e->tiles, *rows = (const int4 *)e->rows, *cols = (const int4 *)e->cols;
  const dim3 blk(64, 4);
  by_ens_view<T>(e, [&](auto tag) {
  using D = typename decltype(tag)::type;
  by_rows(e->ty, [&](auto rr) {
    constexpr int R = decltype(rr)::value;
    if (!chained) {
      hipLaunchKernelGGL((kb_pre<T, R, D>), dim3(e->ntiles), blk, 0, st, devs, runs, tiles);
      hipLaunchKernelGGL(kb_dt<T>, dim3(e->n), dim3(1024), 0, st, devs);
    }
    for (int k = 0; k < 3; ++k) {
      hipLaunchKernelGGL((kb_stage<T, R, D>), dim3(e->ntiles), blk, 0, st, devs, runs, tiles, k);
      by_threads(e->ntrow, [&](auto nt) {
        constexpr int NT = decltype(nt)::value;
        hipLaunchKernelGGL((kb_poisA<T, NT>), dim3(e->nrows), dim3(NT), e->ldsA, st, devs, runs, rows, k);
      });
      hipLaunchKernelGGL(kb_poisB<T>, dim3(e->ncols), dim3(kFftThreads), e->ldsB, st, devs, runs, cols);
      by_threads(e->ntrow, [&](auto nt) {
        constexpr int NT = decltype(nt)::value;
        hipLaunchKernelGGL((kb_poisC<T, NT>), dim3(e->nrows), dim3(NT), e->ldsA, st, devs, runs, rows, k, 1);
      });
    }
    hipLaunchKernelGGL((kb_final<T, R, D>), dim3(e->ntiles), blk, 0, st, devs, runs, tiles);
  });
  });
  hipLaunchKernelGGL(kb_finalize<T>, dim3(e->n), dim3(1024), 0, st, devs, runs);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

// Launches RK stages 0 and 1 of every run (stage kernel and the three Poisson passes, the
// pressure accumulator left alone) `iters` times between two events; these write only plan
// scratch (stage buffers, right-hand side, potential), so the runs' state is untouched.
template <typename T>
int ensemble_time(aero2d_ensemble *e, hipStream_t st, int iters, float *ms) {
  const Dev<T> *devs = (const Dev<T> *)e->devs;
  const BRun<T> *runs = (const BRun<T> *)e->runs;
  const int4 *tiles = (const int4 *)e->tiles, *rows = (const int4 *)e->rows, *cols = (const int4 *)e->cols;
  const dim3 blk(64, 4);
  hipEvent_t e0, e1;
  AERO_HIP_CHECK(hipEventCreate(&e0));
  AERO_HIP_CHECK(hipEventCreate(&e1));
  by_ens_view<T>(e, [&](auto tag) {
  using D = typename decltype(tag)::type;
  by_rows(e->ty, [&](auto rr) {
    constexpr int R = decltype(rr)::value;
    for (int it = 0; it <= iters; ++it) {
      if (it == 1) (void)hipEventRecord(e0, st);   // the first round warms up
      for (int k = 0; k < 2; ++k) {
        hipLaunchKernelGGL((kb_stage<T, R, D>), dim3(e->ntiles), blk, 0, st, devs, runs, tiles, k);
        by_threads(e->ntrow, [&](auto nt) {
          constexpr int NT = decltype(nt)::value;
          hipLaunchKernelGGL((kb_poisA<T, NT>), dim3(e->nrows), dim3(NT), e->ldsA, st, devs, runs, rows, k);
        });
        hipLaunchKernelGGL(kb_poisB<T>, dim3(e->ncols), dim3(kFftThreads), e->ldsB, st, devs, runs, cols);
        by_threads(e->ntrow, [&](auto nt) {
          constexpr int NT = decltype(nt)::value;
          hipLaunchKernelGGL((kb_poisC<T, NT>), dim3(e->nrows), dim3(NT), e->ldsA, st, devs, runs, rows, k, 0);
        });
      }
    }
  });
  });
  (void)hipEventRecord(e1, st);
  hipError_t err = hipEventSynchronize(e1);
  if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) {
    set_error(std::string("ensemble timing: ") + hipGetErrorString(err));
    return AERO_EHIP;
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

// Times every (tile rows, column slots, row fill, row threads) candidate and keeps the fastest;
// choice[0..1] receives the rows and slots chosen, choice[2] the candidates timed, choice[3] the
// row-pass threads, choice[4] the row fill (1: row pairs per block fill kRowFill elements).
template <typename T>
int ensemble_tune(aero2d_ensemble *e, hipStream_t st, int iters, int32_t *choice) {
  static const int kRows[] = {4, 8, 16};
  static const int kSlots[] = {1, 2, 4, 0};   // 0: filled to kRowFill elements
  static const int kFill[] = {0, 1};
  static const int kRowThreads[] = {256, 512};
  float best = 0.f;
  int best_ty = e->ty, best_slot = e->nslot, best_fill = e->rowfill, best_nt = e->ntrow, tried = 0;
  for (int ty : kRows)
    for (int sl : kSlots)
      for (int fill : kFill) {
        e->ty = ty;
        e->nslot = sl;
        e->rowfill = fill;
        int rc = ensemble_build<T>(e);
        for (int nt : kRowThreads) {
          e->ntrow = nt;   // a launch shape only: the work tables do not depend on it
          float ms = 0.f;
          if (!rc) rc = ensemble_time<T>(e, st, iters, &ms);
          if (rc) return rc;
          ++tried;
          if (tried == 1 || ms < best) {
            best = ms;
            best_ty = ty;
            best_slot = sl;
            best_fill = fill;
            best_nt = nt;
          }
        }
      }
  e->ty = best_ty;
  e->nslot = best_slot;
  e->rowfill = best_fill;
  e->ntrow = best_nt;
  const int rc = ensemble_build<T>(e);
  if (choice) {
    choice[0] = best_ty;
    choice[1] = best_slot;
    choice[2] = tried;
    choice[3] = best_nt;
    choice[4] = best_fill;
  }
  return rc;
}

}  // namespace

extern "C" {

int aero2d_ensemble_create(int32_t n, aero2d_plan *const *plans, const aero2d_fields *fields,
                           const aero2d_geom *geoms, const aero2d_params *params, const aero2d_outputs *outputs,
                           aero2d_ensemble **out) {
  if (!out || n <= 0 || !plans || !fields || !geoms || !params) { set_error("bad ensemble arguments"); return AERO_EINVAL; }
  *out = nullptr;
  const int dbl = plans[0] ? plans[0]->dbl : 0;
  for (int r = 0; r < n; ++r) {
    aero2d_plan *p = plans[r];
    if (!p) { set_error("null plan in ensemble"); return AERO_EINVAL; }
    if (p->dbl != dbl) { set_error("ensemble plans must share one dtype"); return AERO_EINVAL; }
    if (p->desc.dye_mode != 0) { set_error("ensemble runs carry no smoke"); return AERO_EUNSUPPORTED; }
    if (outputs && outputs[r].diag) { set_error("ensemble runs record no diagnostics rows"); return AERO_EUNSUPPORTED; }
    aero2d_fields f = fields[r];
    int rc = check_step_args(p, &fields[r], &f, &geoms[r], &params[r],
                             outputs ? const_cast<aero2d_outputs *>(&outputs[r]) : nullptr);
    if (rc) return rc;
  }
  aero2d_ensemble *e = new aero2d_ensemble();
  e->n = n;
  e->dbl = dbl;
  e->plans.assign(plans, plans + n);
  e->fields.assign(fields, fields + n);
  e->geoms.assign(geoms, geoms + n);
  e->params.assign(params, params + n);
  if (outputs) e->outputs.assign(outputs, outputs + n);
  e->plain_laminar = true;
  for (int r = 0; r < n; ++r) e->plain_laminar = e->plain_laminar && plain_laminar(plans[r]->desc);
  int rc = dbl ? ensemble_build<double>(e) : ensemble_build<float>(e);
  if (rc) {
    aero2d_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return AERO_OK;
}

int aero2d_ensemble_step(aero2d_ensemble *e, void *stream, int32_t chained) {
  if (!e) { set_error("null ensemble"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return e->dbl ? ensemble_step<double>(e, st, chained != 0) : ensemble_step<float>(e, st, chained != 0);
}

int aero2d_ensemble_tune(aero2d_ensemble *e, void *stream, int32_t iters, int32_t *choice) {
  if (!e) { set_error("null ensemble"); return AERO_EINVAL; }
  if (iters < 1 || iters > 1000) { set_error("iters must be in [1, 1000]"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return e->dbl ? ensemble_tune<double>(e, st, iters, choice) : ensemble_tune<float>(e, st, iters, choice);
}

int aero2d_ensemble_destroy(aero2d_ensemble *e) {
  if (!e) return AERO_OK;
  ens_release(e);
  delete e;
  return AERO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free Poisson solve
// aero2d_poisson_solve_static: the solve of aero2d_poisson_solve without a plan, for callers
// that may not allocate device memory (code under graph capture, serving loops). Its spectral
// work array and transform tables live in `scratch`, a static device array each GPU holds
// until the process ends; every call rebuilds the tables there (two small kernels) before the
// three transform passes, which are the plan path's own kernels. Calls that overlap on one
// GPU share that array and race.
namespace aero {
namespace ns2d {

constexpr size_t kStaticCells = size_t(1) << 20;   // largest nx * ny of the static solve
constexpr size_t kStaticAlign = 256;
// float64 layout at the largest sizes: F and 1/lambda (cells each), twiddles and DCT
// twiddles (complex, per axis), eigenvalues (float64, per axis)
constexpr size_t kStaticBytes = 2 * kStaticCells * 8 + 2 * 2 * kMaxLen * 16 + 2 * kMaxLen * 8 + 8 * kStaticAlign;

__device__ __align__(256) unsigned char scratch[kStaticBytes];

// Entry o of the stage-ordered twiddle table of an n-point transform (the layout
// stage_twiddles writes on the host for a plan): stage of span p and radix R holds
// w^(r i n / (p R)) at (p - 1) + (r - 1) p + i; the last entry is w^0.
template <typename T>
__device__ cplx<T> stage_twiddle_at(const FftPlan &pl, int o) {
  const int n = pl.n;
  int p = 1;
  for (int s = 0; s < pl.nrad; ++s) {
    const int R = pl.radix[s];
    if (o < (p - 1) + (R - 1) * p) {
      const int q = o - (p - 1);
      const int m = (q / p + 1) * (q % p) * (n / (p * R));
      const double ang = -2.0 * M_PI * (double)m / (double)n;
      return {(T)cos(ang), (T)sin(ang)};
    }
    p *= R;
  }
  return {(T)1, (T)0};
}

// Per-axis tables of an n-point axis (the values upload_tables forms on the host): stage
// twiddles, quarter-wave DCT twiddles and the Laplacian eigenvalues in float64.
template <typename T>
__global__ void __launch_bounds__(256) k_static_axis(FftPlan pl, double hh, int periodic, cplx<T> *__restrict__ tw,
                                                     cplx<T> *__restrict__ cs, double *__restrict__ lam) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = pl.n;
  if (m >= n) return;
  tw[m] = stage_twiddle_at<T>(pl, m);
  const double th = M_PI * (double)m / (2.0 * (double)n);
  cs[m] = {(T)cos(th), (T)sin(th)};
  lam[m] = periodic ? (2.0 * cos(2.0 * M_PI * (double)m / (double)n) - 2.0) / hh
                    : (2.0 * cos(M_PI * (double)m / (double)n) - 2.0) / hh;
}

// 1 / (lam_x + lam_y) formed in float64 and rounded once, the zero mode 0 (ns2d/poisson.py:115-125).
template <typename T>
__global__ void __launch_bounds__(256) k_static_invlam(int nx, int ny, const double *__restrict__ lx,
                                                       const double *__restrict__ ly, T *__restrict__ inv) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nx * ny) return;
  const int i = (int)(e / ny), k = (int)(e % ny);
  const double la = lx[i] + ly[k];
  inv[e] = (T)((la != 0.0 && !(i == 0 && k == 0)) ? 1.0 / la : 0.0);
}

}  // namespace ns2d
}  // namespace aero

namespace {

constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mu;
unsigned char *g_static_ptr[kStaticMaxDevices] = {};   // device address of `scratch`, per device

// Device address of `scratch` on the current device, looked up once per device (with that
// device current) together with the transform kernels' LDS limit.
int static_scratch(unsigned char **out) {
  int dev = 0;
  AERO_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kStaticMaxDevices) {
    set_error("aero2d_poisson_solve_static: device index out of range");
    return AERO_EUNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void *ptr = nullptr;
    AERO_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(aero::ns2d::scratch)));
    int max_lds = 0;
    AERO_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const void *fns[] = {(const void *)k_poisA<float, 256>,   (const void *)k_poisB<float, 256>,
                         (const void *)k_poisC<float, 256>,   (const void *)k_poisA<float, 512>,
                         (const void *)k_poisB<float, 512>,   (const void *)k_poisC<float, 512>,
                         (const void *)k_poisA<float, 1024>,  (const void *)k_poisB<float, 1024>,
                         (const void *)k_poisC<float, 1024>,  (const void *)k_poisA<double, 256>,
                         (const void *)k_poisB<double, 256>,  (const void *)k_poisC<double, 256>,
                         (const void *)k_poisA<double, 512>,  (const void *)k_poisB<double, 512>,
                         (const void *)k_poisC<double, 512>,  (const void *)k_poisA<double, 1024>,
                         (const void *)k_poisB<double, 1024>, (const void *)k_poisC<double, 1024>};
    for (const void *f : fns)
      AERO_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    g_static_ptr[dev] = (unsigned char *)ptr;
  }
  *out = g_static_ptr[dev];
  return AERO_OK;
}

size_t align_up(size_t n) { return (n + kStaticAlign - 1) / kStaticAlign * kStaticAlign; }

template <typename T>
int static_solve(unsigned char *base, hipStream_t st, int nx, int ny, double h, int periodic, const T *f, T *psi) {
  const size_t cells = (size_t)nx * ny;
  Dev<T> d;
  std::memset(&d, 0, sizeof(d));
  d.nx = nx;
  d.ny = ny;
  d.hd = h;
  d.h = (T)h;
  d.periodic = periodic;
  d.dct_x = periodic ? 0 : 1;
  d.dct_y = periodic ? 0 : 1;
  if (!factor_radices(nx, d.fx) || !factor_radices(ny, d.fy)) {
    set_error("aero2d_poisson_solve_static: nx and ny must be products of 2, 3 and 5");
    return AERO_EUNSUPPORTED;
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char *q = base + off;
    off += align_up(bytes);
    return q;
  };
  d.F = global_ptr((T *)take(cells * sizeof(T)));
  T *inv = (T *)take(cells * sizeof(T));
  cplx<T> *tw_x = (cplx<T> *)take(nx * sizeof(cplx<T>)), *tw_y = (cplx<T> *)take(ny * sizeof(cplx<T>));
  cplx<T> *cs_x = (cplx<T> *)take(nx * sizeof(cplx<T>)), *cs_y = (cplx<T> *)take(ny * sizeof(cplx<T>));
  double *lam_x = (double *)take(nx * sizeof(double)), *lam_y = (double *)take(ny * sizeof(double));
  if (off > kStaticBytes) {
    set_error("aero2d_poisson_solve_static: scratch layout exceeds the static array");
    return AERO_EUNSUPPORTED;
  }
  d.inv_lam = global_ptr(inv);
  d.tw_x = global_ptr(tw_x); d.tw_y = global_ptr(tw_y); d.cs_x = global_ptr(cs_x); d.cs_y = global_ptr(cs_y); d.lam_x = global_ptr(lam_x); d.lam_y = global_ptr(lam_y);
  const double hh = h * h;
  hipLaunchKernelGGL(k_static_axis<T>, dim3((nx + 255) / 256), dim3(256), 0, st, d.fx, hh, periodic, tw_x, cs_x, lam_x);
  hipLaunchKernelGGL(k_static_axis<T>, dim3((ny + 255) / 256), dim3(256), 0, st, d.fy, hh, periodic, tw_y, cs_y, lam_y);
  hipLaunchKernelGGL(k_static_invlam<T>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, nx, ny,
                     (const double *)lam_x, (const double *)lam_y, inv);
  const int ntA = transform_threads(nx), ntB = transform_threads(ny);
  const size_t ldsA = 3 * nx * sizeof(cplx<T>) + 2 * nx * sizeof(T) + 64 * 8;
  const size_t ldsB = lds_cols<T>(ny, 1);
  by_threads(ntA, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    hipLaunchKernelGGL((k_poisA<T, NT>), dim3((ny + 1) / 2), dim3(NT), ldsA, st, d, (const T *)nullptr,
                       (const T *)nullptr, f, 0);
  });
  by_threads(ntB, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    hipLaunchKernelGGL((k_poisB<T, NT>), dim3((nx + 1) / 2), dim3(NT), ldsB, st, d, 1);
  });
  by_threads(ntA, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    hipLaunchKernelGGL((k_poisC<T, NT>), dim3((ny + 1) / 2), dim3(NT), ldsA, st, d, psi, (T *)nullptr, 0);
  });
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

}  // namespace

extern "C" int aero2d_poisson_solve_static(void *stream, int32_t dtype, int32_t nx, int32_t ny, double h,
                                           int32_t periodic, const void *f, void *psi) {
  if (!f || !psi) { set_error("aero2d_poisson_solve_static: null argument"); return AERO_EINVAL; }
  if (dtype != 0 && dtype != 1) { set_error("aero2d_poisson_solve_static: dtype must be 0 or 1"); return AERO_EINVAL; }
  if (nx < 4 || ny < 4 || !(h > 0)) {
    set_error("aero2d_poisson_solve_static: needs at least 4 x 4 cells and h > 0");
    return AERO_EINVAL;
  }
  const size_t es = dtype ? 8 : 4;
  const size_t ldsx = 3 * (size_t)nx * 2 * es + 2 * (size_t)nx * es + 512;
  const size_t ldsy = (size_t)ny * (2 * 2 * es + 2 * es) + (size_t)ny * 2 * es;
  if ((size_t)nx * ny > kStaticCells || nx > kMaxLen || ny > kMaxLen || ldsx > kLdsBlockMax || ldsy > kLdsBlockMax) {
    set_error("aero2d_poisson_solve_static: grid larger than the static scratch allows (nx * ny <= " +
              std::to_string(kStaticCells) + ", nx, ny <= " + std::to_string(kMaxLen) +
              ", rows within one block's LDS)");
    return AERO_EUNSUPPORTED;
  }
  unsigned char *base = nullptr;
  int rc = static_scratch(&base);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return dtype ? static_solve<double>(base, st, nx, ny, h, periodic != 0, (const double *)f, (double *)psi)
               : static_solve<float>(base, st, nx, ny, h, periodic != 0, (const float *)f, (float *)psi);
}
