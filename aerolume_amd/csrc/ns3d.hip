// MI355X (gfx950) kernels of the 3D incompressible Navier-Stokes step and its C ABI.
//
// One SSP-RK3 step (reference solvers/ns3d/core.py:753-883 without placed physics) is a
// fixed sequence of launches on the caller's stream, with no host synchronisation and no
// allocation:
//
//   [max, dt]   dt inputs from the state: max |u|, |v|, |w| and max nu_t over the cells
//               -1..n (skipped in chained mode: the previous step's last two kernels left
//               them) - core.py:637-646
//   3 x stage   LDS tiles with a 2-cell halo: the previous stage's projection applied on
//               the fly (u = u^ - grad psi, outlet mass balance), ghost rules per
//               component and axis, eddy viscosity on the tile + 1 ring, upwind3 / central
//               advection, 7-point diffusion, LES stress divergence, RK update,
//               inlet / outlet, window penalisation with force increments (core.py:331-634)
//       p1      divergence (outlet correction folded in) -> DCT / DHT along x, row pairs
//       p2      DCT / DHT along y, column chunks
//       p3      DCT / DHT along z, x 1/lambda (FP64-built), inverse along z
//       p4      inverse along y
//       p5      inverse along x -> psi; step-averaged pressure (core.py:874-875)
//   (theta)     with heat transfer, after each stage kernel: the heat-equation stage with the
//               same stage velocity (MUSCL / van Leer advection, flux diffusion with the
//               turbulent part, adiabatic bodies, sources, held-temperature penalisation);
//               the stage kernel adds the buoyancy Ri theta to w after its penalisation
//               (ns3d/physics.py:647-791)
//   final       u = u^ - grad psi written to the state; per-block partials of the next
//               step's dt inputs and of the force sums
//   finalize    one block: six coefficients per body, t, dt, next dt, history ring, heat
//               rates and wall temperatures of the bodies
//
// Arithmetic in the plan's dtype T (float or double); the Poisson eigenvalue inverse is
// formed in double; the outlet and force sums accumulate in double. No atomics: every
// global reduction goes through per-block partials summed in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolume_amd.h"
#include "common.h"

namespace aero {
namespace ns3d {

constexpr int kThreads = 256;        // Poisson passes
constexpr int kP5Threads = 512;     // the inverse x pass (8 row pairs per block)
constexpr int kPlaneThreads = 1024;  // the fused (z, y) plane pass
constexpr int kPlaneMax = 20480;     // longest plane (ny nz) of the fused pass: 20 values per thread
constexpr int kTileThreads = 1024;   // stage / final tiles (16 waves)
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kFinThreads = 1024;
constexpr int kMaxBodies = 8;
constexpr int kMaxQ = 9 * kMaxBodies;      // force sums per body: u (Fx, My, Mz), v (Fy, Mx, Mz), w (Fz, Mx, My)
constexpr double kSsp[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
constexpr double kForceW[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0};
constexpr double kViscousDt = 1.0 / 6.0;   // ns3d/core.py:51
constexpr int kPm = 5;                     // per-tile maxima: |u|, |v|, |w|, nu_t, |theta|
constexpr int kMaxT = 8;                   // held-temperature bodies

// cells per stage tile (x, y, z), 1024 threads per tile: 80 KB (float) of LDS per tile, so two
// tiles (32 waves) fit a CU's 160 KB
template <typename T>
struct Tile;
template <>
struct Tile<float> {
  static constexpr int X = 32, Y = 8, Z = 8;
  static constexpr int WAVES_PER_SIMD = 8;    // two tiles per CU: registers capped to fit
};
template <>
struct Tile<double> {
  static constexpr int X = 16, Y = 8, Z = 8;
  static constexpr int WAVES_PER_SIMD = 4;    // one tile per CU (LDS-bound)
};

template <typename T>
struct Red3 {
  T dt;          // dt of the step being taken
  T corr[3];     // outlet correction of each stage's u^
  T ctrl[3];     // PID step: the controller memory after this dt (the state's after the step)
};

// Static description of a plan, resident in device memory.
template <typename T>
struct Dev3 {
  int nx, ny, nz;
  int per[3];
  int closed, outlet;
  int upwind, turb, fixed_dt;
  int pid;                   // adaptive_dt = "pid": dt from the controller (pid_target x the rule's dt)
  T pid_target;
  int nbx, nby, nbz, nblk, xcd_chunk;
  int win, wk0, wk1, wj0, wj1, wi0, wi1, wz, wy, wx;
  int wbx0, wby0, wbz0, nwbx, nwby, nwbz, nwt;
  int nbody, symmetry, nq;
  int nut_model;
  T nut_trust;
  int dye, dk0, dk1, dj0, dj1, di0, di1, dwx, dwy;
  int rule[18];
  T rval2[18];
  T h, inv_h, inv_h2, kadv, les_k, cfl_h, visc_h2, corr_scale;
  T neg_h3, q_a, q_a_span, q_a_len, dy;
  T *F, *psi, *acc_u, *acc_v, *acc_w, *pmax;
  double *pout, *pforce;
  Red3<T> *red;
  const cplx<T> *tw[3];
  const cplx<T> *cs[3];
  const double *lam[3];
  const T *inv_lam;          // 1 / (lam_x + lam_y + lam_z) formed in float64, z fastest: (ny, nx, nz),
                             // or (nx, ny, nz) with the plane pass
  FftPlan fp[3];
  int dct[3];
  int ns1, ns2, ns3, ns5;
  int planes;                // 1: y and z transforms in one pass over (z, y) planes, spectrum (nx, nz, ny)
  int plane_buf;             // bytes of each of the plane pass's two LDS work buffers
  // heat transfer (ns3d/physics.py:494-830)
  int theta, buoy, tside[6], tqk0, tqk1, tqj0, tqj1, tqi0, tqi1, tqwx, tqwy, tnt, tbody[kMaxT], tadia, tnskin;
  T ri, half_ri, ri_w[3], abs_ri, prt, kc, cadv, cfl, c6, h3;
  T *th_a, *th_b, *nut_c;
  T *nu_pad;                 // nu_t of the last state the final kernel saw, cells -1..n per axis
  double *pheat;             // [3][window tiles][kMaxT] heat increments of held-temperature bodies
  double *pdiag;             // diagnostics partials
};

// Per-launch pointers (by value).
template <typename T>
struct Call3 {
  const T *u0, *v0, *w0;     // step start
  const T *up, *vp, *wp;     // stage input: the previous stage's u^ (with psi) or a velocity
  const T *psi;              // previous stage's psi; null: up is used as it is
  T *uo, *vo, *wo;           // outputs
  const T *chi_u, *chi_v, *chi_w, *chi_c, *a_u, *a_v, *a_w, *k_u, *k_v, *k_w;
  const T *nu, *eta, *u_inlet, *dt_fixed;
  const T *nut_field;        // learned-closure eddy viscosity (cells), when d.nut_model
  const T *dye_weight, *dye_pulse;
  const T *th0, *thp;        // theta at the step start and of the previous stage
  T *tho;                    // theta of this stage
  const T *chi_t, *a_t, *w_t, *adia_u, *adia_v, *adia_w, *vol_w, *tq;
  const int *skin_idx, *skin_off;
  const T *skin_w, *body_theta, *flux_rate;
  const T *th_sc, *th_side;  // [kappa, kappa / h, eta_t, kappa max|q|, heat scale, kappa / (h^2/6)]; side values
  int stage;
  int corr_stage;            // red->corr slot added at the outlet face of up (-1: none)
  int nu_in;                 // stage input is the state the final kernel last saw: nu_t from d.nu_pad
  const T *ctrl_in;          // PID step: controller memory [dt_prev, e1, e2] of the step start
  T *ctrl_out;               // and of the step's end
  T *tr_lu, *tr_lv, *tr_lw;  // traced stage: tendencies (u, v, w faces)
  T *tr_su, *tr_sv, *tr_sw;  // and values before penalisation
};

template <typename T>
struct Lds3 {
  static constexpr int TX = Tile<T>::X, TY = Tile<T>::Y, TZ = Tile<T>::Z;
  static constexpr int UX = TX + 5, UY = TY + 4, UZ = TZ + 4;
  static constexpr int VX = TX + 4, VY = TY + 5, VZ = TZ + 4;
  static constexpr int WX = TX + 4, WY = TY + 4, WZ = TZ + 5;
  static constexpr int NX = TX + 2, NY = TY + 2, NZ = TZ + 2;
  T u[UZ][UY][UX];
  T v[VZ][VY][VX];
  T w[WZ][WY][WX];
  T nu[NZ][NY][NX];
  double red[kTileWaves * 16];
};

// local (tile-relative) accessors: faces / cells from -2 (halo) on
#define LU(k, j, i) L.u[(k) + 2][(j) + 2][(i) + 2]
#define LV(k, j, i) L.v[(k) + 2][(j) + 2][(i) + 2]
#define LW(k, j, i) L.w[(k) + 2][(j) + 2][(i) + 2]
#define LN(k, j, i) L.nu[(k) + 1][(j) + 1][(i) + 1]

template <typename T>
__device__ __forceinline__ T step_dt(const Dev3<T> &d, const Call3<T> &c) {
  return d.fixed_dt ? *c.dt_fixed : d.red->dt;
}

// Tile of block b, consecutive tiles of a chunk on one XCD (blocks are dealt to the 8 XCDs
// round-robin, so the chunk's halos are shared through that XCD's L2).
template <typename T>
__device__ __forceinline__ void tile_of(const Dev3<T> &d, int &bx, int &by, int &bz, int &tile) {
  const int b = blockIdx.x;
  tile = d.xcd_chunk ? (b & 7) * d.xcd_chunk + (b >> 3) : b;
  bx = tile % d.nbx;
  const int t = tile / d.nbx;
  by = t % d.nby;
  bz = t / d.nby;
}

// ------------------------------------------------------------------ ghost rules
// One axis of a ghost rule (ns3d/core.py:223-272): index g (outside 0..n-1) of an axis of
// extent n -> in-range source; op 0 copy, 1 negate, 2 affine (2 value - x).
__device__ __forceinline__ int ghost_axis(int code, int g, int n, int &op) {
  const bool lo = g < 0;
  switch (code) {
    case 0: op = 0; return lo ? 0 : n - 1;                       // edge
    case 1: op = 0; return lo ? -1 - g : 2 * n - 1 - g;          // even
    case 2: op = 1; return lo ? -1 - g : 2 * n - 1 - g;          // odd
    case 3: op = 2; return lo ? -1 - g : 2 * n - 1 - g;          // affine
    case 4: op = 1; return lo ? -g : 2 * n - 2 - g;              // face_odd
    case 5: op = 0; return lo ? g + n : g - n;                   // periodic_cell
    default: op = 0; return lo ? g + n - 1 : g - n + 1;          // periodic_face
  }
}

// Block sum of N values per thread over all waves of the block (fixed order); thread 0 gets
// the result. lds holds (waves) x N doubles.
template <int N>
__device__ __forceinline__ void block_sum_w(double (&v)[N], double *lds) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) lds[w * N + q] = v[q];
  __syncthreads();
  if (tid == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double r = lds[q];
      for (int i = 1; i < nw; ++i) r += lds[i * N + q];
      v[q] = r;
    }
}

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Source index of position g along an axis of array extent ext (n cells): periodic axes wrap
// (a face axis carries the duplicate face n, so it wraps by n), others clamp; positions the
// clamp moves are ghost positions, filled afterwards by ghost_fill.
__device__ __forceinline__ int src_axis(int g, int ext, int n, int per) {
  if (per) g = g < 0 ? g + n : (g >= ext ? g - n : g);
  return g < 0 ? 0 : (g >= ext ? ext - 1 : g);
}

// Pass 1 of the tile load for component C: the stage input at every LDS position (clamped
// or wrapped source), u^ - grad psi with the outlet shift when c.psi is set (the previous
// stage's projected field, core.py:281-297, 582-612). The tile is walked row by row: each
// half-wave takes one (z, y) row of the tile, its lanes the row's first 32 x positions, so the
// source column of each lane and the source row of each half-wave are worked out once; the
// few positions past 32 of every row are loaded together, one per thread. Element offsets are
// 32-bit, as in the rest of the tile code.
template <typename T, int C, int EX, int EY, int EZ>
__device__ __forceinline__ void load_comp(const Dev3<T> &d, const Call3<T> &c, T *lds, int k0, int j0, int i0, T corr) {
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int ex = nx + (C == 0), ey = ny + (C == 1), ez = nz + (C == 2);
  const T *src = C == 0 ? c.up : (C == 1 ? c.vp : c.wp);
  const T *psi = c.psi;
  const int pn = C == 0 ? nx : (C == 1 ? ny : nz);          // cells along the normal axis
  const int pper = d.per[C];
  const bool shift = c.corr_stage >= 0 && C == 0 && d.outlet;
  const T inv_h = d.inv_h;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xl = lane & 31;
  // column part of tile position x: source column, psi columns, gradient and outlet flags
  struct Col {
    int i, fc, fm;
    bool pv, sh;
  };
  auto col_of = [&](int x) {
    Col cl;
    cl.i = src_axis(i0 - 2 + x, ex, nx, d.per[0]);
    cl.fc = cl.i == pn ? 0 : cl.i;
    cl.fm = cl.fc == 0 ? pn - 1 : cl.fc - 1;
    cl.pv = C == 0 && (pper || (cl.i > 0 && cl.i < pn));
    cl.sh = shift && cl.i == nx;
    return cl;
  };
  // row part of tile row (y, z): source row and plane, and for v (w) the row-uniform normal face
  struct Row {
    int j, k, fc, fm;
    bool pv;
  };
  auto row_of = [&](int y, int z) {
    Row rw;
    rw.j = src_axis(j0 - 2 + y, ey, ny, d.per[1]);
    rw.k = src_axis(k0 - 2 + z, ez, nz, d.per[2]);
    const int f = C == 1 ? rw.j : rw.k;
    rw.pv = C != 0 && (pper || (f > 0 && f < pn));
    rw.fc = f == pn ? 0 : f;
    rw.fm = rw.fc == 0 ? pn - 1 : rw.fc - 1;
    return rw;
  };
  // one tile element: the stage input, u^ - grad psi when psi is set, the outlet shift
  auto value = [&](const Col &cl, const Row &rw) {
    T val = src[(unsigned)((rw.k * ey + rw.j) * ex + cl.i)];
    if (psi) {
      if (C == 0) {
        if (cl.pv) {
          const T *pb = psi + (unsigned)((rw.k * ny + rw.j) * nx);
          val = val - (pb[cl.fc] - pb[cl.fm]) * inv_h;
        }
      } else if (rw.pv) {
        const unsigned pstride = C == 1 ? (unsigned)nx : (unsigned)(nx * ny);
        const T *pb = psi + (C == 1 ? (unsigned)(rw.k * ny * nx) : (unsigned)(rw.j * nx)) + cl.i;
        val = val - (pb[rw.fc * pstride] - pb[rw.fm * pstride]) * inv_h;
      }
    }
    if (cl.sh) val = val + corr;
    return val;
  };
  // positions x < 32 of every row: one half-wave per row, the lane's column worked out once
  constexpr int XM = EX < 32 ? EX : 32, XR = EX - XM;
  const bool xin = xl < XM;
  const Col cm = col_of(xin ? xl : 0);
  for (int row = 2 * wave + (lane >> 5); row < EY * EZ; row += 2 * kTileWaves) {
    const int y = row % EY, z = row / EY;
    const Row rw = row_of(y, z);
    if (xin) lds[(z * EY + y) * EX + xl] = value(cm, rw);
  }
  // the remaining XR columns of all rows together (so no half-wave pass runs mostly idle)
  if constexpr (XR > 0) {
    for (int e = threadIdx.x; e < EY * EZ * XR; e += kTileThreads) {
      const int xr = e % XR, row = e / XR;
      const int y = row % EY, z = row / EY;
      lds[(z * EY + y) * EX + XM + xr] = value(col_of(XM + xr), row_of(y, z));
    }
  }
}

// Pass 2 along one axis: ghost positions of a non-periodic side from their mirror source in
// LDS (ns3d/core.py:223-272), for every position of the other two axes. Runs after the
// lower axes' pass, so corner values compose as pad3 does (x rule, then y, then z).
template <typename T, int AX, int EX, int EY, int EZ>
__device__ __forceinline__ void ghost_fill(const Dev3<T> &d, T *lds, int comp, int a0, int ext) {
  const int E[3] = {EX, EY, EZ};
  const int ea = E[AX];
  const int other = (EX * EY * EZ) / ea;
  for (int side = 0; side < 2; ++side) {
    // ghost positions g of this side inside the tile: lo g = -2, -1 (tile at the start), hi g = ext, ext + 1
    const int g0 = side == 0 ? -2 : ext;
    const int p0 = g0 - (a0 - 2);
    if (side == 0 && a0 != 0) continue;
    if (side == 1 && (p0 >= ea || p0 + 1 < 0)) continue;
    const int code = d.rule[comp * 6 + AX * 2 + side];
    const T rv = d.rval2[comp * 6 + AX * 2 + side];
    for (int e = threadIdx.x; e < 2 * other; e += kTileThreads) {
      const int q = e >> 1, gg = g0 + (e & 1);
      const int p = gg - (a0 - 2);
      if (p < 0 || p >= ea) continue;
      int op;
      const int sg = ghost_axis(code, gg, ext, op);
      const int ps = sg - (a0 - 2);
      int idx, sidx;
      if (AX == 0) {
        idx = q * EX + p;
        sidx = q * EX + ps;
      } else if (AX == 1) {
        const int xx = q % EX, zz = q / EX;
        idx = (zz * EY + p) * EX + xx;
        sidx = (zz * EY + ps) * EX + xx;
      } else {
        idx = p * EX * EY + q;
        sidx = ps * EX * EY + q;
      }
      const T m = lds[sidx];
      lds[idx] = op == 0 ? m : (op == 1 ? -m : rv - m);
    }
  }
}

template <typename T>
__device__ __forceinline__ void load_tile(const Dev3<T> &d, const Call3<T> &c, Lds3<T> &L, int k0, int j0, int i0, T corr) {
  using S = Lds3<T>;
  load_comp<T, 0, S::UX, S::UY, S::UZ>(d, c, &L.u[0][0][0], k0, j0, i0, corr);
  load_comp<T, 1, S::VX, S::VY, S::VZ>(d, c, &L.v[0][0][0], k0, j0, i0, corr);
  load_comp<T, 2, S::WX, S::WY, S::WZ>(d, c, &L.w[0][0][0], k0, j0, i0, corr);
  __syncthreads();
  // boundary tiles: ghost layers of non-periodic sides, axis by axis
  const bool bx = !d.per[0] && (i0 == 0 || i0 + S::TX + 3 > d.nx);
  const bool by = !d.per[1] && (j0 == 0 || j0 + S::TY + 3 > d.ny);
  const bool bz = !d.per[2] && (k0 == 0 || k0 + S::TZ + 3 > d.nz);
  if (bx) {
    ghost_fill<T, 0, S::UX, S::UY, S::UZ>(d, &L.u[0][0][0], 0, i0, d.nx + 1);
    ghost_fill<T, 0, S::VX, S::VY, S::VZ>(d, &L.v[0][0][0], 1, i0, d.nx);
    ghost_fill<T, 0, S::WX, S::WY, S::WZ>(d, &L.w[0][0][0], 2, i0, d.nx);
    __syncthreads();
  }
  if (by) {
    ghost_fill<T, 1, S::UX, S::UY, S::UZ>(d, &L.u[0][0][0], 0, j0, d.ny);
    ghost_fill<T, 1, S::VX, S::VY, S::VZ>(d, &L.v[0][0][0], 1, j0, d.ny + 1);
    ghost_fill<T, 1, S::WX, S::WY, S::WZ>(d, &L.w[0][0][0], 2, j0, d.ny);
    __syncthreads();
  }
  if (bz) {
    ghost_fill<T, 2, S::UX, S::UY, S::UZ>(d, &L.u[0][0][0], 0, k0, d.nz);
    ghost_fill<T, 2, S::VX, S::VY, S::VZ>(d, &L.v[0][0][0], 1, k0, d.nz);
    ghost_fill<T, 2, S::WX, S::WY, S::WZ>(d, &L.w[0][0][0], 2, k0, d.nz + 1);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ eddy viscosity
// Cell-centred gradient g[i][j] = du_i/dx_j at local cell (k, j, i): diagonal from the
// cell's faces, off-diagonal as averages of the four surrounding edge values
// (core.py:400-434).
template <typename T>
__device__ __forceinline__ void cell_grad(const Lds3<T> &L, int k, int j, int i, T inv, T g[3][3]) {
  g[0][0] = (LU(k, j, i + 1) - LU(k, j, i)) * inv;
  g[1][1] = (LV(k, j + 1, i) - LV(k, j, i)) * inv;
  g[2][2] = (LW(k + 1, j, i) - LW(k, j, i)) * inv;
#define DUDY(kk, jj, ii) ((LU(kk, jj, ii) - LU(kk, (jj) - 1, ii)) * inv)
#define DVDX(kk, jj, ii) ((LV(kk, jj, ii) - LV(kk, jj, (ii) - 1)) * inv)
#define DUDZ(kk, jj, ii) ((LU(kk, jj, ii) - LU((kk) - 1, jj, ii)) * inv)
#define DWDX(kk, jj, ii) ((LW(kk, jj, ii) - LW(kk, jj, (ii) - 1)) * inv)
#define DVDZ(kk, jj, ii) ((LV(kk, jj, ii) - LV((kk) - 1, jj, ii)) * inv)
#define DWDY(kk, jj, ii) ((LW(kk, jj, ii) - LW(kk, (jj) - 1, ii)) * inv)
  const T q = (T)0.25;
  g[0][1] = (((DUDY(k, j, i) + DUDY(k, j + 1, i)) + DUDY(k, j, i + 1)) + DUDY(k, j + 1, i + 1)) * q;
  g[0][2] = (((DUDZ(k, j, i) + DUDZ(k + 1, j, i)) + DUDZ(k, j, i + 1)) + DUDZ(k + 1, j, i + 1)) * q;
  g[1][0] = (((DVDX(k, j, i) + DVDX(k, j + 1, i)) + DVDX(k, j, i + 1)) + DVDX(k, j + 1, i + 1)) * q;
  g[1][2] = (((DVDZ(k, j, i) + DVDZ(k + 1, j, i)) + DVDZ(k, j + 1, i)) + DVDZ(k + 1, j + 1, i)) * q;
  g[2][0] = (((DWDX(k, j, i) + DWDX(k + 1, j, i)) + DWDX(k, j, i + 1)) + DWDX(k + 1, j, i + 1)) * q;
  g[2][1] = (((DWDY(k, j, i) + DWDY(k + 1, j, i)) + DWDY(k, j + 1, i)) + DWDY(k + 1, j + 1, i)) * q;
}

// Smagorinsky / WALE / Vreman on the traceless gradient (core.py:437-482).
template <typename T>
__device__ __forceinline__ T eddy3(int turb, T les_k, T g[3][3]) {
  g[2][2] = -(g[0][0] + g[1][1]);
  T s[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) s[a][b] = (T)0.5 * (g[a][b] + g[b][a]);
  T ss = s[0][0] * s[0][0];
#pragma unroll
  for (int q = 1; q < 9; ++q) ss = ss + s[q / 3][q % 3] * s[q / 3][q % 3];
  if (turb == 1) return les_k * sqrt((T)2 * ss + (T)1e-30);
  if (turb == 2) {
    T g2[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) g2[a][b] = (g[a][0] * g[0][b] + g[a][1] * g[1][b]) + g[a][2] * g[2][b];
    const T tr3 = ((g2[0][0] + g2[1][1]) + g2[2][2]) * (T)(1.0 / 3.0);
    T sd = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        T x = (T)0.5 * (g2[a][b] + g2[b][a]);
        if (a == b) x = x - tr3;
        sd = sd + x * x;
      }
    const T sdc = sd > (T)0 ? sd : (T)0, ssc = ss > (T)0 ? ss : (T)0;
    const T rs = sqrt(sdc);
    const T num = sdc * rs;                                     // sd^1.5
    const T den = (ssc * ssc * sqrt(ssc) + sdc * sqrt(rs)) + (T)1e-30;   // ss^2.5 + sd^1.25
    return (les_k * num) / den;
  }
  T b[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) b[a][c] = (g[a][0] * g[c][0] + g[a][1] * g[c][1]) + g[a][2] * g[c][2];
  T bb = ((((b[0][0] * b[1][1] - b[0][1] * b[0][1]) + b[0][0] * b[2][2]) - b[0][2] * b[0][2]) + b[1][1] * b[2][2]) -
         b[1][2] * b[1][2];
  bb = bb > (T)0 ? bb : (T)0;
  T aa = g[0][0] * g[0][0];
#pragma unroll
  for (int q = 1; q < 9; ++q) aa = aa + g[q / 3][q % 3] * g[q / 3][q % 3];
  return les_k * sqrt(bb / (aa + (T)1e-30) + (T)1e-30);
}

// nu_t on the tile's cells and their one-cell ring, zero inside solids (x (1 - chi_c) on
// the window, core.py:491-503); returns this thread's max over the cells the tile owns
// (its own cells, plus the ghost ring at domain boundaries).
// own_only: only the cells whose maximum the tile reports (the dt input), L.nu is not filled.
template <typename T>
__device__ __forceinline__ T nu_tile(const Dev3<T> &d, const Call3<T> &c, Lds3<T> &L, int k0, int j0, int i0, bool store = false,
                     bool own_only = false) {
  using S = Lds3<T>;
  T mx = 0;
  const int ie = min(i0 + S::TX, d.nx), je = min(j0 + S::TY, d.ny), ke = min(k0 + S::TZ, d.nz);
  for (int e = threadIdx.x; e < S::NZ * S::NY * S::NX; e += kTileThreads) {
    const int x = e % S::NX, t = e / S::NX, y = t % S::NY, z = t / S::NY;
    const int i = x - 1, j = y - 1, k = z - 1;
    if (own_only) {
      const int gi = i0 + i, gj = j0 + j, gk = k0 + k;
      const bool oi = (gi >= i0 && gi < ie) || gi == -1 || gi == d.nx;
      const bool oj = (gj >= j0 && gj < je) || gj == -1 || gj == d.ny;
      const bool ok = (gk >= k0 && gk < ke) || gk == -1 || gk == d.nz;
      if (!(oi && oj && ok && gi <= ie && gj <= je && gk <= ke)) continue;
    }
    T g[3][3];
    cell_grad(L, k, j, i, d.inv_h, g);
    T nu = eddy3(d.turb, d.les_k, g);
    const int gi = i0 + i, gj = j0 + j, gk = k0 + k;
    if (d.nut_model) {
      // learned closure where its value is trusted (cells of the domain; periodic ghosts wrap)
      const bool in_i = (gi >= 0 && gi < d.nx) || d.per[0], in_j = (gj >= 0 && gj < d.ny) || d.per[1],
                 in_k = (gk >= 0 && gk < d.nz) || d.per[2];
      if (in_i && in_j && in_k && gi <= d.nx && gj <= d.ny && gk <= d.nz) {
        const int wi = wrapi(gi, d.nx), wj = wrapi(gj, d.ny), wk = wrapi(gk, d.nz);
        const T f = c.nut_field[((size_t)wk * d.ny + wj) * d.nx + wi];
        if (f >= (T)0 && f <= d.nut_trust) nu = f;
      }
    }
    if (d.win && gk >= d.wk0 && gk < d.wk1 && gj >= d.wj0 && gj < d.wj1 && gi >= d.wi0 && gi < d.wi1)
      nu = nu * ((T)1 - c.chi_c[((size_t)(gk - d.wk0) * d.wy + (gj - d.wj0)) * d.wx + (gi - d.wi0)]);
    if (!own_only) L.nu[z][y][x] = nu;
    // the heat equation's turbulent diffusivity reads the stage's nu_t of the cells
    if (store && gi >= i0 && gi < ie && gj >= j0 && gj < je && gk >= k0 && gk < ke)
      d.nut_c[((size_t)gk * d.ny + gj) * d.nx + gi] = nu;
    const bool oi = (gi >= i0 && gi < ie) || gi == -1 || gi == d.nx;
    const bool oj = (gj >= j0 && gj < je) || gj == -1 || gj == d.ny;
    const bool ok = (gk >= k0 && gk < ke) || gk == -1 || gk == d.nz;
    if (oi && oj && ok && gi <= ie && gj <= je && gk <= ke) {
      mx = nu > mx ? nu : mx;
      // the next step's first stage starts from this state: it reads these values back
      if (own_only && d.nu_pad)
        d.nu_pad[((size_t)(gk + 1) * (d.ny + 2) + (gj + 1)) * (d.nx + 2) + (gi + 1)] = nu;
    }
  }
  __syncthreads();
  return mx;
}

// L.nu of a tile from d.nu_pad (the values nu_tile computes for the state the final kernel
// last processed, tile by tile, ghost ring included); positions past the ring are unused.
template <typename T>
__device__ __forceinline__ void nu_from_pad(const Dev3<T> &d, Lds3<T> &L, int k0, int j0, int i0, bool store) {
  using S = Lds3<T>;
  const int ie = min(i0 + S::TX, d.nx), je = min(j0 + S::TY, d.ny), ke = min(k0 + S::TZ, d.nz);
  for (int e = threadIdx.x; e < S::NZ * S::NY * S::NX; e += kTileThreads) {
    const int x = e % S::NX, t = e / S::NX, y = t % S::NY, z = t / S::NY;
    const int gi = i0 + x - 1, gj = j0 + y - 1, gk = k0 + z - 1;
    T nu = 0;
    if (gi <= d.nx && gj <= d.ny && gk <= d.nz)
      nu = d.nu_pad[((size_t)(gk + 1) * (d.ny + 2) + (gj + 1)) * (d.nx + 2) + (gi + 1)];
    L.nu[z][y][x] = nu;
    if (store && gi >= i0 && gi < ie && gj >= j0 && gj < je && gk >= k0 && gk < ke)
      d.nut_c[((size_t)gk * d.ny + gj) * d.nx + gi] = nu;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ right-hand side
// upwind3 / central face flux (core.py:331-346).
template <typename T>
__device__ __forceinline__ T flux(T vs, T a, T b, T c, T dd, int upwind) {
  const T s = b + c;
  if (!upwind) return vs * s;
  const T cen = s + (T)(-1.0 / 7.0) * (a + dd);
  const T dis = (b - c) + (T)(-1.0 / 3.0) * (a - dd);
  return vs * cen + ((T)(3.0 / 7.0) * fabs(vs)) * dis;
}

template <typename T>
__device__ __forceinline__ T lerp_t(T start, T end, T w) {
  return w < (T)0.5 ? start + w * (end - start) : end - (end - start) * ((T)1 - w);
}

// SGS stress components on edges / cells (core.py:505-524), from LDS
#define NXY(k, j, i) ((((LN(k, (j) - 1, (i) - 1) + LN(k, j, (i) - 1)) + LN(k, (j) - 1, i)) + LN(k, j, i)) * (T)0.25)
#define NXZ(k, j, i) ((((LN((k) - 1, j, (i) - 1) + LN(k, j, (i) - 1)) + LN((k) - 1, j, i)) + LN(k, j, i)) * (T)0.25)
#define NYZ(k, j, i) ((((LN((k) - 1, (j) - 1, i) + LN(k, (j) - 1, i)) + LN((k) - 1, j, i)) + LN(k, j, i)) * (T)0.25)
#define T12(k, j, i) (NXY(k, j, i) * (DUDY(k, j, i) + DVDX(k, j, i)))
#define T13(k, j, i) (NXZ(k, j, i) * (DUDZ(k, j, i) + DWDX(k, j, i)))
#define T23(k, j, i) (NYZ(k, j, i) * (DVDZ(k, j, i) + DWDY(k, j, i)))
#define T11(k, j, i) (((T)2 * LN(k, j, i)) * ((LU(k, j, (i) + 1) - LU(k, j, i)) * inv))
#define T22(k, j, i) (((T)2 * LN(k, j, i)) * ((LV(k, (j) + 1, i) - LV(k, j, i)) * inv))
#define T33(k, j, i) (((T)2 * LN(k, j, i)) * ((LW((k) + 1, j, i) - LW(k, j, i)) * inv))

// explicit tendency at a u-face (local k, j, i): advection + diffusion (+ LES)
template <typename T>
__device__ __forceinline__ T rhs_u(const Dev3<T> &d, const Lds3<T> &L, int k, int j, int i, T nu_h2) {
  const int up = d.upwind;
  const T fl = flux(LU(k, j, i - 1) + LU(k, j, i), LU(k, j, i - 2), LU(k, j, i - 1), LU(k, j, i), LU(k, j, i + 1), up);
  const T fr = flux(LU(k, j, i) + LU(k, j, i + 1), LU(k, j, i - 1), LU(k, j, i), LU(k, j, i + 1), LU(k, j, i + 2), up);
  const T gd = flux(LV(k, j, i - 1) + LV(k, j, i), LU(k, j - 2, i), LU(k, j - 1, i), LU(k, j, i), LU(k, j + 1, i), up);
  const T gu =
      flux(LV(k, j + 1, i - 1) + LV(k, j + 1, i), LU(k, j - 1, i), LU(k, j, i), LU(k, j + 1, i), LU(k, j + 2, i), up);
  const T zd = flux(LW(k, j, i - 1) + LW(k, j, i), LU(k - 2, j, i), LU(k - 1, j, i), LU(k, j, i), LU(k + 1, j, i), up);
  const T zu =
      flux(LW(k + 1, j, i - 1) + LW(k + 1, j, i), LU(k - 1, j, i), LU(k, j, i), LU(k + 1, j, i), LU(k + 2, j, i), up);
  const T a = (((fl - fr) + (gd - gu)) + (zd - zu)) * d.kadv;
  const T ring = ((LU(k, j, i + 1) + LU(k, j, i - 1)) + (LU(k, j + 1, i) + LU(k, j - 1, i))) +
                 (LU(k + 1, j, i) + LU(k - 1, j, i));
  T r = a + (ring - (T)6 * LU(k, j, i)) * nu_h2;
  if (d.turb) {
    const T inv = d.inv_h;
    r = r + (((((T11(k, j, i) - T11(k, j, i - 1)) + T12(k, j + 1, i)) - T12(k, j, i)) + T13(k + 1, j, i)) -
             T13(k, j, i)) * inv;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T rhs_v(const Dev3<T> &d, const Lds3<T> &L, int k, int j, int i, T nu_h2) {
  const int up = d.upwind;
  const T gl = flux(LU(k, j - 1, i) + LU(k, j, i), LV(k, j, i - 2), LV(k, j, i - 1), LV(k, j, i), LV(k, j, i + 1), up);
  const T gr =
      flux(LU(k, j - 1, i + 1) + LU(k, j, i + 1), LV(k, j, i - 1), LV(k, j, i), LV(k, j, i + 1), LV(k, j, i + 2), up);
  const T fd = flux(LV(k, j - 1, i) + LV(k, j, i), LV(k, j - 2, i), LV(k, j - 1, i), LV(k, j, i), LV(k, j + 1, i), up);
  const T fu = flux(LV(k, j, i) + LV(k, j + 1, i), LV(k, j - 1, i), LV(k, j, i), LV(k, j + 1, i), LV(k, j + 2, i), up);
  const T zd = flux(LW(k, j - 1, i) + LW(k, j, i), LV(k - 2, j, i), LV(k - 1, j, i), LV(k, j, i), LV(k + 1, j, i), up);
  const T zu =
      flux(LW(k + 1, j - 1, i) + LW(k + 1, j, i), LV(k - 1, j, i), LV(k, j, i), LV(k + 1, j, i), LV(k + 2, j, i), up);
  const T a = (((gl - gr) + (fd - fu)) + (zd - zu)) * d.kadv;
  const T ring = ((LV(k, j, i + 1) + LV(k, j, i - 1)) + (LV(k, j + 1, i) + LV(k, j - 1, i))) +
                 (LV(k + 1, j, i) + LV(k - 1, j, i));
  T r = a + (ring - (T)6 * LV(k, j, i)) * nu_h2;
  if (d.turb) {
    const T inv = d.inv_h;
    r = r + (((((T12(k, j, i + 1) - T12(k, j, i)) + T22(k, j, i)) - T22(k, j - 1, i)) + T23(k + 1, j, i)) -
             T23(k, j, i)) * inv;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T rhs_w(const Dev3<T> &d, const Lds3<T> &L, int k, int j, int i, T nu_h2) {
  const int up = d.upwind;
  const T gl = flux(LU(k - 1, j, i) + LU(k, j, i), LW(k, j, i - 2), LW(k, j, i - 1), LW(k, j, i), LW(k, j, i + 1), up);
  const T gr =
      flux(LU(k - 1, j, i + 1) + LU(k, j, i + 1), LW(k, j, i - 1), LW(k, j, i), LW(k, j, i + 1), LW(k, j, i + 2), up);
  const T gd = flux(LV(k - 1, j, i) + LV(k, j, i), LW(k, j - 2, i), LW(k, j - 1, i), LW(k, j, i), LW(k, j + 1, i), up);
  const T gu =
      flux(LV(k - 1, j + 1, i) + LV(k, j + 1, i), LW(k, j - 1, i), LW(k, j, i), LW(k, j + 1, i), LW(k, j + 2, i), up);
  const T fd = flux(LW(k - 1, j, i) + LW(k, j, i), LW(k - 2, j, i), LW(k - 1, j, i), LW(k, j, i), LW(k + 1, j, i), up);
  const T fu = flux(LW(k, j, i) + LW(k + 1, j, i), LW(k - 1, j, i), LW(k, j, i), LW(k + 1, j, i), LW(k + 2, j, i), up);
  const T a = (((gl - gr) + (gd - gu)) + (fd - fu)) * d.kadv;
  const T ring = ((LW(k, j, i + 1) + LW(k, j, i - 1)) + (LW(k, j + 1, i) + LW(k, j - 1, i))) +
                 (LW(k + 1, j, i) + LW(k - 1, j, i));
  T r = a + (ring - (T)6 * LW(k, j, i)) * nu_h2;
  if (d.turb) {
    const T inv = d.inv_h;
    r = r + (((((T13(k, j, i + 1) - T13(k, j, i)) + T23(k, j + 1, i)) - T23(k, j, i)) + T33(k, j, i)) -
             T33(k - 1, j, i)) * inv;
  }
  return r;
}

// faces a tile writes along one axis: [c0, end); the last tile also writes the last face
__device__ __forceinline__ int face_end(int c0, int t, int n) { return c0 + t >= n ? n + 1 : c0 + t; }

// ------------------------------------------------------------------ stage kernel
// NU_IN: nu_t read from d.nu_pad (the first stage of a step whose start the final kernel
// processed) instead of computed; a separate instantiation keeps the common path's registers.
// TRACE: also store each face's tendency and its value before the penalisation (c.tr_l, c.tr_s;
// null pointers skipped), for the traced step only.
template <typename T, bool NU_IN, bool TRACE = false>
__global__ void __launch_bounds__(kTileThreads, Tile<T>::WAVES_PER_SIMD) k_stage(const Dev3<T> *__restrict__ dp, Call3<T> c) {
  using S = Lds3<T>;
  extern __shared__ unsigned char smem[];
  S &L = *reinterpret_cast<S *>(smem);
  const Dev3<T> &d = *dp;
  int bx, by, bz, tile;
  tile_of(d, bx, by, bz, tile);
  const int i0 = bx * S::TX, j0 = by * S::TY, k0 = bz * S::TZ;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int s = c.stage;
  const T corr = c.corr_stage >= 0 ? d.red->corr[c.corr_stage] : (T)0;
  load_tile(d, c, L, k0, j0, i0, corr);
  if (d.turb) {
    if constexpr (NU_IN) nu_from_pad(d, L, k0, j0, i0, d.theta != 0);
    else (void)nu_tile(d, c, L, k0, j0, i0, d.theta != 0);
  }
  const T dt = step_dt(d, c);
  const T a = (T)kSsp[s][0];
  const T bdt = dt * (T)kSsp[s][1];
  const T ws = (T)kForceW[s];
  const T nu_h2 = *c.nu * d.inv_h2;
  const T kpen = d.win ? bdt / *c.eta : (T)0;
  const int ie = min(i0 + S::TX, nx), je = min(j0 + S::TY, ny), ke = min(k0 + S::TZ, nz);
  double osum = 0.0, isum = 0.0;
  // Faces of the tile: every thread takes the cells (li, lj, lk) of the tile that its index
  // gives by shifts and masks (no runtime division); the last tile along an axis also writes
  // the domain's closing face along it (the faces a tile writes: [c0, end), face_end).
  constexpr int TX = S::TX, TY = S::TY, TZ = S::TZ;
  constexpr int NCELL = TX * TY * TZ, NR = (NCELL + kTileThreads - 1) / kTileThreads;
  const int ncx = ie - i0, ncy = je - j0, ncz = ke - k0;
  auto u_face = [&](int li, int lj, int lk) {
    const int gi = i0 + li, gj = j0 + lj, gk = k0 + lk;
    const size_t idx = ((size_t)gk * ny + gj) * (nx + 1) + gi;
    T lu;
    const bool xb = !d.per[0] && (gi == 0 || gi == nx);
    if (xb) lu = (gi == nx && d.outlet) ? (LU(lk, lj, li - 1) - LU(lk, lj, li)) * d.inv_h : (T)0;
    else lu = rhs_u(d, L, lk, lj, li, nu_h2);
    const T upv = LU(lk, lj, li);
    T uh = (s == 0 ? upv : lerp_t(upv, c.u0[idx], a)) + lu * bdt;
    if (xb && gi == 0) uh = d.outlet ? c.u_inlet[(size_t)gk * ny + gj] : (T)0;
    if (xb && gi == nx && d.closed) uh = (T)0;
    if constexpr (TRACE) {
      if (c.tr_lu) c.tr_lu[idx] = lu;
      if (c.tr_su) c.tr_su[idx] = uh;
    }
    if (d.win && gk >= d.wk0 && gk < d.wk1 && gj >= d.wj0 && gj < d.wj1 && gi >= d.wi0 && gi <= d.wi1) {
      const size_t wi = ((size_t)(gk - d.wk0) * d.wy + (gj - d.wj0)) * (d.wx + 1) + (gi - d.wi0);
      const T un = (uh + kpen * c.a_u[wi]) / ((T)1 + kpen * c.chi_u[wi]);
      const T du = un - uh;
      d.acc_u[wi] = s == 0 ? du * ws : d.acc_u[wi] + du * ws;
      uh = un;
    }
    c.uo[idx] = uh;
    if (gi == nx && d.outlet) {
      osum += (double)uh;
      isum += (double)c.u_inlet[(size_t)gk * ny + gj];
    }
  };
  auto v_face = [&](int li, int lj, int lk) {
    const int gi = i0 + li, gj = j0 + lj, gk = k0 + lk;
    const size_t idx = ((size_t)gk * (ny + 1) + gj) * nx + gi;
    const bool yb = !d.per[1] && (gj == 0 || gj == ny);
    const T lv = yb ? (T)0 : rhs_v(d, L, lk, lj, li, nu_h2);
    const T vpv = LV(lk, lj, li);
    T vh = (s == 0 ? vpv : lerp_t(vpv, c.v0[idx], a)) + lv * bdt;
    if constexpr (TRACE) {
      if (c.tr_lv) c.tr_lv[idx] = lv;
      if (c.tr_sv) c.tr_sv[idx] = vh;
    }
    if (d.win && gk >= d.wk0 && gk < d.wk1 && gj >= d.wj0 && gj <= d.wj1 && gi >= d.wi0 && gi < d.wi1) {
      const size_t wi = ((size_t)(gk - d.wk0) * (d.wy + 1) + (gj - d.wj0)) * d.wx + (gi - d.wi0);
      const T vn = (vh + kpen * c.a_v[wi]) / ((T)1 + kpen * c.chi_v[wi]);
      const T dv = vn - vh;
      d.acc_v[wi] = s == 0 ? dv * ws : d.acc_v[wi] + dv * ws;
      vh = vn;
    }
    c.vo[idx] = vh;
  };
  auto w_face = [&](int li, int lj, int lk) {
    const int gi = i0 + li, gj = j0 + lj, gk = k0 + lk;
    const size_t idx = ((size_t)gk * ny + gj) * nx + gi;
    const bool zb = !d.per[2] && (gk == 0 || gk == nz);
    const T lw = zb ? (T)0 : rhs_w(d, L, lk, lj, li, nu_h2);
    const T wpv = LW(lk, lj, li);
    T wh = (s == 0 ? wpv : lerp_t(wpv, c.w0[idx], a)) + lw * bdt;
    if constexpr (TRACE) {
      if (c.tr_lw) c.tr_lw[idx] = lw;
      if (c.tr_sw) c.tr_sw[idx] = wh;
    }
    const bool inw = d.win && gk >= d.wk0 && gk <= d.wk1 && gj >= d.wj0 && gj < d.wj1 && gi >= d.wi0 && gi < d.wi1;
    const size_t wi = inw ? ((size_t)(gk - d.wk0) * d.wy + (gj - d.wj0)) * d.wx + (gi - d.wi0) : 0;
    if (inw) {
      const T wn = (wh + kpen * c.a_w[wi]) / ((T)1 + kpen * c.chi_w[wi]);
      const T dw = wn - wh;
      d.acc_w[wi] = s == 0 ? dw * ws : d.acc_w[wi] + dw * ws;
      wh = wn;
    }
    if (d.buoy) {
      // b dt Ri theta on the w-faces after the penalisation, everywhere; the solid's share
      // (Ri theta_f vol_w) leaves the body force (physics.py:744-791)
      const size_t pl = (size_t)ny * nx, col = (size_t)gj * nx + gi;
      const int zn = nz - 1;
      const T tk = c.thp[(size_t)(gk < nz ? gk : (d.per[2] ? 0 : zn)) * pl + col];
      const T tm = c.thp[(size_t)(gk > 0 ? gk - 1 : (d.per[2] ? zn : 0)) * pl + col];
      if (gk > 0 && gk < nz) wh = wh + ((T)0.5 * (tk + tm)) * (bdt * d.ri);
      else if (d.per[2]) wh = wh + ((tk + tm) * d.half_ri) * bdt;
      if (inw) d.acc_w[wi] = d.acc_w[wi] + (((T)0.5 * (tk + tm)) * c.vol_w[wi]) * (bdt * d.ri_w[s]);
    }
    c.wo[idx] = wh;
  };
  // one call site per face kind: slots 0 .. NR - 1 are the tile's cells, slot NR the closing face
  const int tid = threadIdx.x;
#pragma unroll 1
  for (int r = 0; r <= NR; ++r) {
    const int e = tid + kTileThreads * r;
    int li = e % TX, lj = (e / TX) % TY, lk = e / (TX * TY);
    bool ok = e < NCELL;
    if (r == NR) {
      li = ncx, lj = tid % TY, lk = tid / TY;
      ok = i0 + TX >= nx && tid < TY * TZ;
    }
    if (ok && li <= ncx && lj < ncy && lk < ncz && (li < ncx || r == NR)) u_face(li, lj, lk);
  }
#pragma unroll 1
  for (int r = 0; r <= NR; ++r) {
    const int e = tid + kTileThreads * r;
    int li = e % TX, lj = (e / TX) % TY, lk = e / (TX * TY);
    bool ok = e < NCELL;
    if (r == NR) {
      li = tid % TX, lj = ncy, lk = tid / TX;
      ok = j0 + TY >= ny && tid < TX * TZ;
    }
    if (ok && li < ncx && lj <= ncy && lk < ncz && (lj < ncy || r == NR)) v_face(li, lj, lk);
  }
#pragma unroll 1
  for (int r = 0; r <= NR; ++r) {
    const int e = tid + kTileThreads * r;
    int li = e % TX, lj = (e / TX) % TY, lk = e / (TX * TY);
    bool ok = e < NCELL;
    if (r == NR) {
      li = tid % TX, lj = tid / TX, lk = ncz;
      ok = k0 + TZ >= nz && tid < TX * TY;
    }
    if (ok && li < ncx && lj < ncy && lk <= ncz && (lk < ncz || r == NR)) w_face(li, lj, lk);
  }
  // outlet plane sums of this tile (the outlet correction needs them all)
  if (d.outlet && i0 + S::TX >= nx) {
    double v[2] = {isum, osum};
    block_sum_w<2>(v, L.red);
    if (threadIdx.x == 0) {
      d.pout[2 * (bz * d.nby + by)] = v[0];
      d.pout[2 * (bz * d.nby + by) + 1] = v[1];
    }
  }
}

// ------------------------------------------------------------------ maxima / final
// mode 0: dt inputs of a state (max |u|, |v|, |w|, nu_t), no writes.
// mode 1: the last stage's projection applied (out = u^ - grad psi), written to the state,
//         plus the dt inputs of that state and the force sums of window tiles.
template <typename T>
__global__ void __launch_bounds__(kTileThreads, Tile<T>::WAVES_PER_SIMD) k_final(const Dev3<T> *__restrict__ dp, Call3<T> c, int mode) {
  using S = Lds3<T>;
  extern __shared__ unsigned char smem[];
  S &L = *reinterpret_cast<S *>(smem);
  const Dev3<T> &d = *dp;
  int bx, by, bz, tile;
  tile_of(d, bx, by, bz, tile);
  const int i0 = bx * S::TX, j0 = by * S::TY, k0 = bz * S::TZ;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const T corr = c.corr_stage >= 0 ? d.red->corr[c.corr_stage] : (T)0;
  load_tile(d, c, L, k0, j0, i0, corr);
  T mnu = 0;
  if (d.turb) mnu = nu_tile(d, c, L, k0, j0, i0, false, true);
  const int ie = min(i0 + S::TX, nx), je = min(j0 + S::TY, ny), ke = min(k0 + S::TZ, nz);
  T mu = 0, mv = 0, mw = 0, mt = 0;
  if (mode == 0 && d.theta) {
    // max |theta| of the state (the free-fall limit of the dt rule)
    const int nfx = ie - i0, nfy = je - j0, nfz = ke - k0;
    for (int e = threadIdx.x; e < nfx * nfy * nfz; e += kTileThreads) {
      const int li = e % nfx, t = e / nfx, lj = t % nfy, lk = t / nfy;
      const T x = fabs(c.th0[((size_t)(k0 + lk) * ny + (j0 + lj)) * nx + i0 + li]);
      mt = x > mt ? x : mt;
    }
  }
  {
    const int fe = face_end(i0, S::TX, nx), nfx = fe - i0, nfy = je - j0, nfz = ke - k0;
    for (int e = threadIdx.x; e < nfx * nfy * nfz; e += kTileThreads) {
      const int li = e % nfx, t = e / nfx, lj = t % nfy, lk = t / nfy;
      const T x = LU(lk, lj, li);
      if (mode) c.uo[((size_t)(k0 + lk) * ny + (j0 + lj)) * (nx + 1) + i0 + li] = x;
      mu = fabs(x) > mu ? fabs(x) : mu;
    }
  }
  {
    const int fe = face_end(j0, S::TY, ny), nfx = ie - i0, nfy = fe - j0, nfz = ke - k0;
    for (int e = threadIdx.x; e < nfx * nfy * nfz; e += kTileThreads) {
      const int li = e % nfx, t = e / nfx, lj = t % nfy, lk = t / nfy;
      const T x = LV(lk, lj, li);
      if (mode) c.vo[((size_t)(k0 + lk) * (ny + 1) + (j0 + lj)) * nx + i0 + li] = x;
      mv = fabs(x) > mv ? fabs(x) : mv;
    }
  }
  {
    const int fe = face_end(k0, S::TZ, nz), nfx = ie - i0, nfy = je - j0, nfz = fe - k0;
    for (int e = threadIdx.x; e < nfx * nfy * nfz; e += kTileThreads) {
      const int li = e % nfx, t = e / nfx, lj = t % nfy, lk = t / nfy;
      const T x = LW(lk, lj, li);
      if (mode) c.wo[((size_t)(k0 + lk) * ny + (j0 + lj)) * nx + i0 + li] = x;
      mw = fabs(x) > mw ? fabs(x) : mw;
    }
  }
  mu = wave_max(mu);
  mv = wave_max(mv);
  mw = wave_max(mw);
  mnu = wave_max(mnu);
  mt = wave_max(mt);
  T *red = reinterpret_cast<T *>(L.red);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
    red[wv * kPm + 0] = mu;
    red[wv * kPm + 1] = mv;
    red[wv * kPm + 2] = mw;
    red[wv * kPm + 3] = mnu;
    red[wv * kPm + 4] = mt;
  }
  __syncthreads();
  // slot 4 (|theta|) of the end-of-step state comes from the last heat-equation stage
  const int nslot = (mode == 0 && d.theta) ? kPm : 4;
  if (threadIdx.x < nslot) {
    T m = red[threadIdx.x];
    for (int q = 1; q < kTileWaves; ++q) m = red[q * kPm + threadIdx.x] > m ? red[q * kPm + threadIdx.x] : m;
    d.pmax[(size_t)tile * kPm + threadIdx.x] = m;
  }
  // force sums of this tile's window faces: sum_faces k[b, m] * acc (core.py:653-660), one
  // pass over the faces per (component, body) for its three weights; per-wave sums go to the
  // (no longer needed) u tile, then waves are added in order
  if (mode && d.win && bx >= d.wbx0 && bx < d.wbx0 + d.nwbx && by >= d.wby0 && by < d.wby0 + d.nwby &&
      bz >= d.wbz0 && bz < d.wbz0 + d.nwbz) {
    const int wt = ((bz - d.wbz0) * d.nwby + (by - d.wby0)) * d.nwbx + (bx - d.wbx0);
    __syncthreads();
    static_assert(sizeof(L.u) >= sizeof(double) * kTileWaves * kMaxQ, "force partials exceed the u tile");
    double *fr = reinterpret_cast<double *>(&L.u[0][0][0]);
    for (int comp = 0; comp < 3; ++comp) {
      // window faces of component comp inside this tile
      const int f0[3] = {max(i0, d.wi0), max(j0, d.wj0), max(k0, d.wk0)};
      const int f1[3] = {min(comp == 0 ? face_end(i0, S::TX, nx) : ie, d.wi1 + (comp == 0)),
                         min(comp == 1 ? face_end(j0, S::TY, ny) : je, d.wj1 + (comp == 1)),
                         min(comp == 2 ? face_end(k0, S::TZ, nz) : ke, d.wk1 + (comp == 2))};
      const int n0 = f1[0] - f0[0], n1 = f1[1] - f0[1], n2 = f1[2] - f0[2];
      const int nf = (n0 > 0 && n1 > 0 && n2 > 0) ? n0 * n1 * n2 : 0;
      const int ex = d.wx + (comp == 0), ey = d.wy + (comp == 1), ez = d.wz + (comp == 2);
      const size_t wn = (size_t)ex * ey * ez;
      const T *ac = comp == 0 ? d.acc_u : comp == 1 ? d.acc_v : d.acc_w;
      const T *kb = comp == 0 ? c.k_u : comp == 1 ? c.k_v : c.k_w;
      for (int b = 0; b < d.nbody; ++b) {
        const T *k0w = kb + (size_t)b * 3 * wn;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int e = threadIdx.x; e < nf; e += kTileThreads) {
          const int x = e % n0, t = e / n0, y = t % n1, z = t / n1;
          const size_t wi = ((size_t)(f0[2] + z - d.wk0) * ey + (f0[1] + y - d.wj0)) * ex + (f0[0] + x - d.wi0);
          const double av = (double)ac[wi];
          a0 += (double)k0w[wi] * av;
          a1 += (double)k0w[wn + wi] * av;
          a2 += (double)k0w[2 * wn + wi] * av;
        }
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        a2 = wave_sum(a2);
        if (lane == 0) {
          double *o = fr + wv * kMaxQ + b * 9 + comp * 3;
          o[0] = a0;
          o[1] = a1;
          o[2] = a2;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < d.nq) {
      double r = fr[threadIdx.x];
      for (int w2 = 1; w2 < kTileWaves; ++w2) r += fr[w2 * kMaxQ + threadIdx.x];
      d.pforce[(size_t)wt * d.nq + threadIdx.x] = r;
    }
  }
}

// Cell-centred velocity gradient of a state, times scale: out[3 i + j] = du_i/dx_j.
template <typename T>
__global__ void __launch_bounds__(kTileThreads, Tile<T>::WAVES_PER_SIMD) k_grad(const Dev3<T> *__restrict__ dp,
                                                                                Call3<T> c, T *__restrict__ out,
                                                                                T scale) {
  using S = Lds3<T>;
  extern __shared__ unsigned char smem[];
  S &L = *reinterpret_cast<S *>(smem);
  const Dev3<T> &d = *dp;
  int bx, by, bz, tile;
  tile_of(d, bx, by, bz, tile);
  const int i0 = bx * S::TX, j0 = by * S::TY, k0 = bz * S::TZ;
  load_tile(d, c, L, k0, j0, i0, (T)0);
  const int ie = min(i0 + S::TX, d.nx), je = min(j0 + S::TY, d.ny), ke = min(k0 + S::TZ, d.nz);
  const int nfx = ie - i0, nfy = je - j0, nfz = ke - k0;
  const size_t nc = (size_t)d.nx * d.ny * d.nz;
  for (int e = threadIdx.x; e < nfx * nfy * nfz; e += kTileThreads) {
    const int li = e % nfx, t = e / nfx, lj = t % nfy, lk = t / nfy;
    T g[3][3];
    cell_grad(L, lk, lj, li, d.inv_h, g);
    const size_t p = ((size_t)(k0 + lk) * d.ny + (j0 + lj)) * d.nx + (i0 + li);
#pragma unroll
    for (int q = 0; q < 9; ++q) out[q * nc + p] = g[q / 3][q % 3] * scale;
  }
}

// dt rule: min(cfl h / (max|u| + max|v| + max|w|), h^2 / 6 / (nu + nu_t,max)) (core.py:637-646)
// With heat transfer also (physics.py:1230-1262, 723-742): the combined advection-diffusion
// bound of theta 1 / (umax / (h/2) + kappa_eff / (h^2/6)), the free-fall limit
// cfl sqrt(h / (|Ri| max|theta|)) and the heaters' bounds 1 / S, (6 cfl h / (|Ri| S))^(1/3).
template <typename T>
__device__ T cfl_dt(const Dev3<T> &d, const Call3<T> &c, const T m[kPm]) {
  T umax = (m[0] + m[1]) + m[2];
  umax = umax > (T)1e-6 ? umax : (T)1e-6;
  const T dta = d.cfl_h / umax;
  const T nu = d.turb ? *c.nu + m[3] : *c.nu;
  const T dtb = d.visc_h2 / nu;
  T dt = dta < dtb ? dta : dtb;
  if (d.theta) {
    const T *sc = c.th_sc;
    const T diff = d.turb ? (sc[0] + m[3] / d.prt) / d.visc_h2 : sc[5];
    const T lim = (T)1 / (umax * d.cadv + diff);
    dt = lim < dt ? lim : dt;
    if (d.buoy) {
      T acc = d.abs_ri * m[4];
      acc = acc > (T)1e-12 ? acc : (T)1e-12;
      const T ff = d.cfl * sqrt(d.h / acc);
      dt = ff < dt ? ff : dt;
    }
    if (sc[3] > (T)0) {
      const T smax = sc[3] > (T)1e-12 ? sc[3] : (T)1e-12;
      const T a = (T)1 / smax;
      dt = a < dt ? a : dt;
      if (d.buoy) {
        const T b = pow(d.c6 / (d.abs_ri * smax), (T)(1.0 / 3.0));
        dt = b < dt ? b : dt;
      }
    }
  }
  return dt;
}

// Max of the per-tile partials (fixed order), by all threads of a 1024-thread block.
template <typename T>
__device__ void reduce_pmax(const Dev3<T> &d, T out[kPm], T *lds) {
  const int nq = d.theta ? kPm : 4;
  T m[kPm] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < d.nblk; b += blockDim.x)
    for (int q = 0; q < nq; ++q) m[q] = d.pmax[(size_t)b * kPm + q] > m[q] ? d.pmax[(size_t)b * kPm + q] : m[q];
#pragma unroll
  for (int q = 0; q < kPm; ++q) m[q] = wave_max(m[q]);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < kPm; ++q) lds[wv * kPm + q] = m[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPm; ++q) {
    T r = lds[q];
    for (int w = 1; w < nw; ++w) r = lds[w * kPm + q] > r ? lds[w * kPm + q] : r;
    out[q] = r;
  }
  __syncthreads();
}

// PID step (physics.py:1144-1177): from the rule's dt `cap` and the memory mem = [dt_prev, e1,
// e2], e = log(target cap / dt_prev) and dt = dt_prev exp(kI e + kP (e - e1) + kD (e - 2 e1 + e2))
// with gains (0.3, 0.4, 0.05), the factor clamped to [0.5, 1.2] and dt never above cap; the first
// step (dt_prev <= 0) takes target cap. Writes the memory after this dt to out (may alias mem).
template <typename T>
__device__ T pid3(const Dev3<T> &d, T cap, const T *mem, T *out) {
  const T want = cap * d.pid_target;
  const T prev_dt = mem[0], e1 = mem[1], e2 = mem[2];
  const bool first = prev_dt <= (T)0;
  const T prev = first ? want : prev_dt;
  const T e = log(want / prev);
  T fac = exp(((T)0.3 * e + (T)0.4 * (e - e1)) + (T)0.05 * ((e - (T)2 * e1) + e2));
  fac = fac < (T)0.5 ? (T)0.5 : (fac > (T)1.2 ? (T)1.2 : fac);
  const T cand = first ? want : prev * fac;
  const T dt = (cand < cap || cand != cand) ? cand : cap;
  out[0] = dt;
  out[1] = first ? (T)0 : e;
  out[2] = e1;
  return dt;
}

template <typename T>
__global__ void __launch_bounds__(kFinThreads) k_dt(const Dev3<T> *__restrict__ dp, Call3<T> c) {
  __shared__ T lds[kPm * (kFinThreads / 64)];
  const Dev3<T> &d = *dp;
  T m[kPm];
  reduce_pmax(d, m, lds);
  if (threadIdx.x == 0) {
    T dt = cfl_dt(d, c, m);
    if (d.pid) dt = pid3(d, dt, c.ctrl_in, d.red->ctrl);
    d.red->dt = dt;
  }
}

// Outputs of the finalize kernel.
template <typename T>
struct Fin3 {
  const T *t_in;
  T *t_out, *dt_out, *coeffs, *hist;
  long long *hist_count;
  long long hist_cap;
  int next_dt;
  const T *theta;                   // end-of-step temperature (heat transfer)
  T *heat_rate, *theta_wall, *phist;
  int phist_walls;
};

template <typename T>
__global__ void __launch_bounds__(kFinThreads) k_finalize(const Dev3<T> *__restrict__ dp, Call3<T> c, Fin3<T> f) {
  __shared__ T lds[kPm * (kFinThreads / 64)];
  __shared__ double fs[kMaxQ];
  __shared__ double part[kFinThreads];
  __shared__ double hsum[kMaxBodies], wsum[kMaxBodies];
  const Dev3<T> &d = *dp;
  const T dt = step_dt(d, c);
  const bool heat = d.theta && d.nbody > 0 && (f.heat_rate || f.theta_wall || f.phist);
  if (heat) {
    // heat increments of the held-temperature bodies over the stages and window tiles
    const int nt = d.tnt;
    if (threadIdx.x < kMaxBodies) hsum[threadIdx.x] = wsum[threadIdx.x] = 0.0;
    __syncthreads();
    if (nt > 0) {
      const int P = kFinThreads / nt, q = threadIdx.x % nt, sl = threadIdx.x / nt;
      double a = 0.0;
      if (sl < P)
        for (int t = sl; t < 3 * d.nwt; t += P) a += d.pheat[(size_t)t * kMaxT + q];
      part[threadIdx.x] = a;
      __syncthreads();
      if (threadIdx.x < nt) {
        double r = 0.0;
        for (int s2 = 0; s2 < P; ++s2) r += part[s2 * nt + threadIdx.x];
        hsum[d.tbody[threadIdx.x]] = r;
      }
      __syncthreads();
    }
    // skin-weighted wall temperature of the heat-flux bodies (physics.py:420-441)
    if (d.tnskin > 0)
      for (int b = 0; b < d.nbody; ++b) {
        const int o0 = c.skin_off[b], o1 = c.skin_off[b + 1];
        double a = 0.0;
        for (int e = o0 + (int)threadIdx.x; e < o1; e += kFinThreads) a += (double)(f.theta[c.skin_idx[e]] * c.skin_w[e]);
        a = wave_sum(a);
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0) part[wv] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
          double r = 0.0;
          for (int w2 = 0; w2 < kFinThreads / 64; ++w2) r += part[w2];
          wsum[b] = r;
        }
        __syncthreads();
      }
  }
  // force sums over window tiles: thread (q, slice) sums tiles slice, slice + P, ... in order
  const int nq = d.nq;
  if (nq > 0) {
    const int P = kFinThreads / nq;
    const int q = threadIdx.x % nq, sl = threadIdx.x / nq;
    double a = 0.0;
    if (sl < P)
      for (int t = sl; t < d.nwt; t += P) a += d.pforce[(size_t)t * nq + q];
    part[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x < nq) {
      double r = 0.0;
      for (int s2 = 0; s2 < P; ++s2) r += part[s2 * nq + threadIdx.x];
      fs[threadIdx.x] = r;
    }
    __syncthreads();
  }
  T mx[kPm] = {0, 0, 0, 0, 0};
  if (f.next_dt && !d.fixed_dt) reduce_pmax(d, mx, lds);
  if (threadIdx.x != 0) return;
  const T scale = d.neg_h3 / dt;
  T tot[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < d.nbody; ++b) {
    T ru[3], rv[3], rw[3];
    for (int m = 0; m < 3; ++m) {
      ru[m] = (T)fs[b * 9 + m] * scale;
      rv[m] = (T)fs[b * 9 + 3 + m] * scale;
      rw[m] = (T)fs[b * 9 + 6 + m] * scale;
    }
    T fx = ru[0], fy = rv[0], fz = rw[0];
    T mxm = rv[1] + rw[1], my = ru[1] + rw[2], mz = ru[2] + rv[2];
    if (d.symmetry) {
      fx = (T)2 * fx;
      fz = (T)2 * fz;
      my = (T)2 * my;
      fy = (T)0;
      mxm = -d.dy * fz;
      mz = d.dy * fx;
    }
    const T cf[6] = {fx / d.q_a, fy / d.q_a, fz / d.q_a, mxm / d.q_a_span, my / d.q_a_len, mz / d.q_a_span};
    for (int m = 0; m < 6; ++m) {
      if (f.coeffs) f.coeffs[b * 6 + m] = cf[m];
      tot[m] = tot[m] + cf[m];
    }
  }
  const T tnew = *f.t_in + dt;
  if (f.t_out) *f.t_out = tnew;
  if (f.dt_out) *f.dt_out = dt;
  if (heat) {
    // heat rate in conduction units: increments x (sym h^3 Pe / dt) + released flux (physics.py:713-721)
    const T scale = c.th_sc[4] / dt;
    const T sym = d.symmetry ? (T)2 : (T)1;
    T *prow = nullptr;
    if (f.phist && f.hist_count && f.hist_cap > 0) {
      const int cols = 1 + d.nbody * (f.phist_walls ? 2 : 1);
      prow = f.phist + (size_t)(*f.hist_count % f.hist_cap) * cols;
      prow[0] = tnew;
    }
    for (int b = 0; b < d.nbody; ++b) {
      const T q = (T)hsum[b] * scale + c.flux_rate[b] * sym;
      const T tw = c.body_theta[b] + (T)wsum[b];
      if (f.heat_rate) f.heat_rate[b] = q;
      if (f.theta_wall) f.theta_wall[b] = tw;
      if (prow) {
        prow[1 + b] = q;
        if (f.phist_walls) prow[1 + d.nbody + b] = tw;
      }
    }
  }
  if (f.hist && f.hist_count && f.hist_cap > 0) {
    long long n = *f.hist_count;
    T *row = f.hist + (size_t)(n % f.hist_cap) * 7;
    row[0] = tnew;
    for (int m = 0; m < 6; ++m) row[1 + m] = tot[m];
    *f.hist_count = n + 1;
  }
  if (d.pid && c.ctrl_out) {
    // commit the memory of this step's dt; fixed steps leave it as it was
    for (int q = 0; q < 3; ++q) c.ctrl_out[q] = d.fixed_dt ? c.ctrl_in[q] : d.red->ctrl[q];
  }
  if (f.next_dt && !d.fixed_dt) {
    T dt = cfl_dt(d, c, mx);
    if (d.pid) {
      T mem[3] = {d.red->ctrl[0], d.red->ctrl[1], d.red->ctrl[2]};
      dt = pid3(d, dt, mem, d.red->ctrl);
    }
    d.red->dt = dt;
  }
}


// ------------------------------------------------------------------ smoke
// Smoke value at a cell one axis away from the domain (ns3d/dye.py:232-263): x clean inflow
// / zero-gradient outlet, y and z mirrored, periodic axes wrapped.
template <typename T>
__device__ __forceinline__ T dye_at(const Dev3<T> &d, const T *__restrict__ c, int k, int j, int i) {
  if (i < 0 || i >= d.nx) {
    if (d.per[0]) i = wrapi(i, d.nx);
    else if (i < 0) return (T)0;
    else i = d.nx - 1;
  }
  if (j < 0 || j >= d.ny) j = d.per[1] ? wrapi(j, d.ny) : (j < 0 ? -1 - j : 2 * d.ny - 1 - j);
  if (k < 0 || k >= d.nz) k = d.per[2] ? wrapi(k, d.nz) : (k < 0 ? -1 - k : 2 * d.nz - 1 - k);
  return c[((size_t)k * d.ny + j) * d.nx + i];
}

// upwind MUSCL flux with the van Leer slope through the face between b and c (dye.py:267-278)
template <typename T>
__device__ __forceinline__ T dye_flux(T a, T b, T c, T dd, T vel) {
  T p, q, r;
  if (vel > (T)0) {
    p = b - a;
    q = c - b;
    const T pq = p * q;
    r = b + (T)0.5 * (pq > (T)0 ? (T)2 * pq / (p + q) : (T)0);
  } else {
    p = c - b;
    q = dd - c;
    const T pq = p * q;
    r = c - (T)0.5 * (pq > (T)0 ? (T)2 * pq / (p + q) : (T)0);
  }
  return vel * r;
}

// One smoke step with the end-of-step velocity: forward Euler, clipped at 1, floored; zero in
// solid cells (chi_c > 1/2); relaxed to 1 in the nozzles (dye.py:281-312).
template <typename T>
__global__ void __launch_bounds__(256) k_dye(const Dev3<T> *__restrict__ dp, Call3<T> c, const T *__restrict__ cin,
                                            T *__restrict__ cout, const T *__restrict__ t_in, T floor_v) {
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const size_t n = (size_t)nx * ny * nz;
  const T dt = step_dt(d, c);
  const T kk = dt * d.inv_h;
  const T t = *t_in + dt;
  T on = 0;
  if (d.dye && d.dk1 > d.dk0) {
    const T tm = fmod(t, c.dye_pulse[1]);
    on = tm < c.dye_pulse[0] ? (T)1 : (T)0;
  }
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nx), j = (int)((e / nx) % ny), k = (int)(e / ((size_t)nx * ny));
    const T c0 = cin[e];
    const T *uu = c.up + ((size_t)k * ny + j) * (nx + 1) + i;
    const T *vv = c.vp + ((size_t)k * (ny + 1) + j) * nx + i;
    const T *ww = c.wp + ((size_t)k * ny + j) * nx + i;
    const T fx0 = dye_flux(dye_at(d, cin, k, j, i - 2), dye_at(d, cin, k, j, i - 1), c0, dye_at(d, cin, k, j, i + 1), uu[0]);
    const T fx1 = dye_flux(dye_at(d, cin, k, j, i - 1), c0, dye_at(d, cin, k, j, i + 1), dye_at(d, cin, k, j, i + 2), uu[1]);
    const T fy0 = dye_flux(dye_at(d, cin, k, j - 2, i), dye_at(d, cin, k, j - 1, i), c0, dye_at(d, cin, k, j + 1, i), vv[0]);
    const T fy1 = dye_flux(dye_at(d, cin, k, j - 1, i), c0, dye_at(d, cin, k, j + 1, i), dye_at(d, cin, k, j + 2, i), vv[nx]);
    const T fz0 = dye_flux(dye_at(d, cin, k - 2, j, i), dye_at(d, cin, k - 1, j, i), c0, dye_at(d, cin, k + 1, j, i), ww[0]);
    const T fz1 = dye_flux(dye_at(d, cin, k - 1, j, i), c0, dye_at(d, cin, k + 1, j, i), dye_at(d, cin, k + 2, j, i),
                           ww[(size_t)ny * nx]);
    T r = c0 - kk * (((fx1 - fx0) + (fy1 - fy0)) + (fz1 - fz0));
    r = r > (T)1 ? (T)1 : r;
    r = r < floor_v ? (T)0 : r;
    if (d.win && k >= d.wk0 && k < d.wk1 && j >= d.wj0 && j < d.wj1 && i >= d.wi0 && i < d.wi1 &&
        c.chi_c[((size_t)(k - d.wk0) * d.wy + (j - d.wj0)) * d.wx + (i - d.wi0)] > (T)0.5)
      r = (T)0;
    if (k >= d.dk0 && k < d.dk1 && j >= d.dj0 && j < d.dj1 && i >= d.di0 && i < d.di1) {
      const T wgt = c.dye_weight[((size_t)(k - d.dk0) * d.dwy + (j - d.dj0)) * d.dwx + (i - d.di0)];
      r = r + wgt * on * ((T)1 - r);
    }
    cout[e] = r;
  }
}

// ------------------------------------------------------------------ heat transfer
// theta tile of a stage: the tile's cells with two ghost layers (cross-shaped use: edge and
// corner positions are never read).
template <typename T>
struct LdsT {
  static constexpr int TX = Tile<T>::X, TY = Tile<T>::Y, TZ = Tile<T>::Z;
  static constexpr int X = TX + 4, Y = TY + 4, Z = TZ + 4;
  T c[Z][Y][X];
  double red[kTileWaves * kMaxT];
};
#define LT(k, j, i) L.c[(k) + 2][(j) + 2][(i) + 2]

// theta at position g (outside 0..n-1) of a non-periodic side (physics.py:589-606): outflow
// zero gradient, adiabatic mirror, inflow value, temperature 2 theta_w - mirror, heat flux
// mirror + G h (first layer) / 3 G h (second).
template <typename T>
__device__ __forceinline__ T theta_ghost(const Dev3<T> &d, const Call3<T> &c, const T *th, int side, int g, int n,
                                         size_t base, size_t stride) {
  const bool lo = g < 0;
  const int m = lo ? -1 - g : 2 * n - 1 - g;
  switch (d.tside[side]) {
    case AERO3D_THETA_OUTFLOW: return th[base + (size_t)(lo ? 0 : n - 1) * stride];
    case AERO3D_THETA_INFLOW: return c.th_side[side];
    case AERO3D_THETA_TEMPERATURE: return (T)2 * c.th_side[side] - th[base + (size_t)m * stride];
    case AERO3D_THETA_HEAT_FLUX: {
      const T wf = (g == -1 || g == n) ? d.h : d.h3;
      return th[base + (size_t)m * stride] + c.th_side[side] * wf;
    }
    default: return th[base + (size_t)m * stride];   // adiabatic
  }
}

// Stage input of velocity component C at face (k, j, i): the previous stage's u^ with its
// projection and outlet shift applied exactly as the stage kernel's tile load does.
template <typename T, int C>
__device__ __forceinline__ T face_vel(const Dev3<T> &d, const Call3<T> &c, int k, int j, int i, T corr) {
  const int nx = d.nx, ny = d.ny;
  const int ex = nx + (C == 0), ey = ny + (C == 1);
  const T *src = C == 0 ? c.up : (C == 1 ? c.vp : c.wp);
  T val = src[((size_t)k * ey + j) * ex + i];
  if (c.psi) {
    const int pn = C == 0 ? nx : (C == 1 ? ny : d.nz);
    const int f = C == 0 ? i : (C == 1 ? j : k);
    if (d.per[C] || (f > 0 && f < pn)) {
      const int fc = f == pn ? 0 : f;
      const int fm = fc == 0 ? pn - 1 : fc - 1;
      const size_t ps = C == 0 ? 1 : (C == 1 ? (size_t)nx : (size_t)nx * ny);
      const T *pb = c.psi + ((size_t)(C == 2 ? 0 : k) * ny + (C == 1 ? 0 : j)) * nx + (C == 0 ? 0 : i);
      val = val - (pb[(size_t)fc * ps] - pb[(size_t)fm * ps]) * d.inv_h;
    }
  }
  if (C == 0 && c.corr_stage >= 0 && d.outlet && i == nx) val = val + corr;
  return val;
}

// Total theta flux through the low face (along axis AX) of local cell (lk, lj, li) = global
// (gk, gj, gi): MUSCL advection minus kappa_f (c - b), conduction removed through adiabatic
// bodies (physics.py:623-676).
template <typename T, int AX>
__device__ __forceinline__ T theta_face(const Dev3<T> &d, const Call3<T> &c, const LdsT<T> &L, int lk, int lj, int li,
                                        int gk, int gj, int gi, T kh, T corr) {
  T a, b, cc, dd;
  if (AX == 0) {
    a = LT(lk, lj, li - 2); b = LT(lk, lj, li - 1); cc = LT(lk, lj, li); dd = LT(lk, lj, li + 1);
  } else if (AX == 1) {
    a = LT(lk, lj - 2, li); b = LT(lk, lj - 1, li); cc = LT(lk, lj, li); dd = LT(lk, lj + 1, li);
  } else {
    a = LT(lk - 2, lj, li); b = LT(lk - 1, lj, li); cc = LT(lk, lj, li); dd = LT(lk + 1, lj, li);
  }
  const T vel = face_vel<T, AX>(d, c, gk, gj, gi, corr);
  const T diff = cc - b;
  const int g = AX == 0 ? gi : (AX == 1 ? gj : gk);
  const int n = AX == 0 ? d.nx : (AX == 1 ? d.ny : d.nz);
  T kf = kh;
  if (d.turb && (d.per[AX] || (g > 0 && g < n))) {
    const int gc = g == n ? 0 : g, gm = gc == 0 ? n - 1 : gc - 1;
    const size_t col = AX == 0 ? ((size_t)gk * d.ny + gj) * d.nx : (AX == 1 ? (size_t)gk * d.ny * d.nx + gi
                                                                            : (size_t)gj * d.nx + gi);
    const size_t st = AX == 0 ? 1 : (AX == 1 ? (size_t)d.nx : (size_t)d.nx * d.ny);
    kf = kh + d.kc * (d.nut_c[col + (size_t)gm * st] + d.nut_c[col + (size_t)gc * st]);
  }
  T f = dye_flux(a, b, cc, dd, vel) - kf * diff;
  if (d.tadia) {
    const int ek = AX == 2, ej = AX == 1, ei = AX == 0;
    if (gk >= d.wk0 && gk < d.wk1 + ek && gj >= d.wj0 && gj < d.wj1 + ej && gi >= d.wi0 && gi < d.wi1 + ei) {
      const T *chi = AX == 0 ? c.adia_u : (AX == 1 ? c.adia_v : c.adia_w);
      const size_t wi = ((size_t)(gk - d.wk0) * (d.wy + ej) + (gj - d.wj0)) * (d.wx + ei) + (gi - d.wi0);
      f = f + (kf * chi[wi]) * diff;
    }
  }
  return f;
}

// One heat-equation stage (physics.py:678-708) on a tile: theta^ = lerp(theta_prev, theta0, a)
// - b dt div F + b dt kappa q, then the held-temperature bodies' implicit penalisation with
// their heat increments (per-tile partials); the last stage also leaves max |theta|.
template <typename T>
__global__ void __launch_bounds__(kTileThreads) k_theta(const Dev3<T> *__restrict__ dp, Call3<T> c) {
  using S = LdsT<T>;
  extern __shared__ unsigned char smem[];
  S &L = *reinterpret_cast<S *>(smem);
  const Dev3<T> &d = *dp;
  int bx, by, bz, tile;
  tile_of(d, bx, by, bz, tile);
  const int i0 = bx * S::TX, j0 = by * S::TY, k0 = bz * S::TZ;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int ie = min(i0 + S::TX, nx), je = min(j0 + S::TY, ny), ke = min(k0 + S::TZ, nz);
  const T *th = c.thp;
  for (int e = threadIdx.x; e < S::X * S::Y * S::Z; e += kTileThreads) {
    const int x = e % S::X, t = e / S::X, y = t % S::Y, z = t / S::Y;
    int gi = i0 - 2 + x, gj = j0 - 2 + y, gk = k0 - 2 + z;
    if (gi >= ie + 2 || gj >= je + 2 || gk >= ke + 2) continue;     // beyond the tile's stencils
    if (d.per[0]) gi = wrapi(gi, nx);
    if (d.per[1]) gj = wrapi(gj, ny);
    if (d.per[2]) gk = wrapi(gk, nz);
    const int ox = gi < 0 || gi >= nx, oy = gj < 0 || gj >= ny, oz = gk < 0 || gk >= nz;
    if (ox + oy + oz > 1) continue;                                   // edges / corners: never read
    T val;
    if (ox) val = theta_ghost(d, c, th, gi < 0 ? 0 : 1, gi, nx, ((size_t)gk * ny + gj) * nx, 1);
    else if (oy) val = theta_ghost(d, c, th, gj < 0 ? 2 : 3, gj, ny, (size_t)gk * ny * nx + gi, (size_t)nx);
    else if (oz) val = theta_ghost(d, c, th, gk < 0 ? 4 : 5, gk, nz, (size_t)gj * nx + gi, (size_t)nx * ny);
    else val = th[((size_t)gk * ny + gj) * nx + gi];
    L.c[z][y][x] = val;
  }
  __syncthreads();
  const int s = c.stage;
  const T dt = step_dt(d, c);
  const T a = (T)kSsp[s][0];
  const T bdt = dt * (T)kSsp[s][1];
  const T wgt = (T)kForceW[s];
  const T kh = c.th_sc[1];
  const T bk = bdt * c.th_sc[0];
  const T kpen = d.tnt ? bdt / c.th_sc[2] : (T)0;
  const T corr = c.corr_stage >= 0 ? d.red->corr[c.corr_stage] : (T)0;
  const int nfx = ie - i0, nfy = je - j0, nfz = ke - k0;
  double hb[kMaxT];
#pragma unroll
  for (int b = 0; b < kMaxT; ++b) hb[b] = 0.0;
  T mt = 0;
  for (int e = threadIdx.x; e < nfx * nfy * nfz; e += kTileThreads) {
    const int li = e % nfx, t = e / nfx, lj = t % nfy, lk = t / nfy;
    const int gi = i0 + li, gj = j0 + lj, gk = k0 + lk;
    const T fx0 = theta_face<T, 0>(d, c, L, lk, lj, li, gk, gj, gi, kh, corr);
    const T fx1 = theta_face<T, 0>(d, c, L, lk, lj, li + 1, gk, gj, gi + 1, kh, corr);
    const T fy0 = theta_face<T, 1>(d, c, L, lk, lj, li, gk, gj, gi, kh, corr);
    const T fy1 = theta_face<T, 1>(d, c, L, lk, lj + 1, li, gk, gj + 1, gi, kh, corr);
    const T fz0 = theta_face<T, 2>(d, c, L, lk, lj, li, gk, gj, gi, kh, corr);
    const T fz1 = theta_face<T, 2>(d, c, L, lk + 1, lj, li, gk + 1, gj, gi, kh, corr);
    const T D = (((fx1 - fx0) + (fy1 - fy0)) + (fz1 - fz0)) * d.inv_h;
    const size_t idx = ((size_t)gk * ny + gj) * nx + gi;
    const T tp = LT(lk, lj, li);
    T ch = (s == 0 ? tp : lerp_t(tp, c.th0[idx], a)) - bdt * D;
    if (gk >= d.tqk0 && gk < d.tqk1 && gj >= d.tqj0 && gj < d.tqj1 && gi >= d.tqi0 && gi < d.tqi1)
      ch = ch + bk * c.tq[((size_t)(gk - d.tqk0) * d.tqwy + (gj - d.tqj0)) * d.tqwx + (gi - d.tqi0)];
    if (d.tnt && gk >= d.wk0 && gk < d.wk1 && gj >= d.wj0 && gj < d.wj1 && gi >= d.wi0 && gi < d.wi1) {
      const size_t wn = (size_t)d.wz * d.wy * d.wx;
      const size_t wi = ((size_t)(gk - d.wk0) * d.wy + (gj - d.wj0)) * d.wx + (gi - d.wi0);
      const T nw = (ch + kpen * c.a_t[wi]) / ((T)1 + kpen * c.chi_t[wi]);
      const T inc = (nw - ch) * wgt;
#pragma unroll
      for (int b = 0; b < kMaxT; ++b)
        if (b < d.tnt) hb[b] += (double)(c.w_t[(size_t)b * wn + wi] * inc);
      ch = nw;
    }
    c.tho[idx] = ch;
    mt = fabs(ch) > mt ? fabs(ch) : mt;
  }
  if (s == 2) {
    mt = wave_max(mt);
    T *red = reinterpret_cast<T *>(L.red);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[wv] = mt;
    __syncthreads();
    if (threadIdx.x == 0) {
      T m = red[0];
      for (int q = 1; q < kTileWaves; ++q) m = red[q] > m ? red[q] : m;
      d.pmax[(size_t)tile * kPm + 4] = m;
    }
  }
  if (d.tnt && bx >= d.wbx0 && bx < d.wbx0 + d.nwbx && by >= d.wby0 && by < d.wby0 + d.nwby && bz >= d.wbz0 &&
      bz < d.wbz0 + d.nwbz) {
    const int wt = ((bz - d.wbz0) * d.nwby + (by - d.wby0)) * d.nwbx + (bx - d.wbx0);
    __syncthreads();
    block_sum_w<kMaxT>(hb, L.red);
    if (threadIdx.x == 0)
      for (int b = 0; b < kMaxT; ++b) d.pheat[((size_t)s * d.nwt + wt) * kMaxT + b] = b < d.tnt ? hb[b] : 0.0;
  }
}

template <typename T>
__global__ void k_copy3(T *__restrict__ dst, const T *__restrict__ src, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// ------------------------------------------------------------------ Poisson passes
__device__ __forceinline__ int makhoul_src(int n, int len) { return n < (len + 1) / 2 ? 2 * n : 2 * (len - 1 - n) + 1; }

// Inverse real-transform packing: conj-weighted pair for the DCT-III, plain pair for the DHT.
template <typename T>
__device__ __forceinline__ cplx<T> pack_inv(const T *xa, const T *xb, int n, int dct, cplx<T> w, int k) {
  if (dct) {
    const T xa_k = xa[k], xb_k = xb[k];
    const T xa_r = k == 0 ? (T)0 : xa[n - k];
    const T xb_r = k == 0 ? (T)0 : xb[n - k];
    const cplx<T> va = {w.x * xa_k + w.y * xa_r, w.y * xa_k - w.x * xa_r};
    const cplx<T> vb = {w.x * xb_k + w.y * xb_r, w.y * xb_k - w.x * xb_r};
    return {va.x - vb.y, -(va.y + vb.x)};
  }
  return {xa[k], xb[k]};
}

// forward real transform values of slot f at spectral index k (after fft of packed pairs)
template <typename T>
__device__ __forceinline__ void fwd_split(const cplx<T> *R, int n, int k, int dct, cplx<T> cs, T &xa, T &xb) {
  cplx<T> va, vb;
  split_pair(R, n, k, va, vb);
  if (dct) {
    xa = cs.x * va.x + cs.y * va.y;
    xb = cs.x * vb.x + cs.y * vb.y;
  } else {
    xa = va.x - va.y;
    xb = vb.x - vb.y;
  }
}

// inverse real transform values at natural position q(m) of slot f (after fft of pack_inv)
template <typename T>
__device__ __forceinline__ void inv_unpack(const cplx<T> *R, int n, int m, int dct, T inv_n, T &va, T &vb) {
  if (dct) {
    va = R[m].x * inv_n;
    vb = -R[m].y * inv_n;
  } else {
    cplx<T> ha, hb;
    split_pair(R, n, m, ha, hb);
    va = (ha.x - ha.y) * inv_n;
    vb = (hb.x - hb.y) * inv_n;
  }
}

// divergence at cell (k, j, i) of (u^, v^, w^) with the outlet face shifted by corr
// (core.py:577-579)
template <typename T>
__device__ __forceinline__ T div_at(const Dev3<T> &d, const T *u, const T *v, const T *w, int k, int j, int i, T corr) {
  // element offsets in 32 bits (the face arrays hold fewer than 2^32 values)
  const unsigned nx = d.nx, ny = d.ny, kk = k, jj = j, ii = i;
  const T *ur = u + (kk * ny + jj) * (nx + 1);
  T u1 = ur[ii + 1];
  if (i + 1 == d.nx) u1 = u1 + corr;
  const T *vr = v + (kk * (ny + 1) + jj) * nx + ii;
  const unsigned pl = ny * nx;
  const T *wr = w + (kk * ny + jj) * nx + ii;
  return (((u1 - ur[ii]) + (vr[nx] - vr[0])) + (wr[pl] - wr[0])) * d.inv_h;
}

// ------------------------------------------------------------------ diagnostics
// One diagnostics row of a step (ns3d/solver.py:747-783): per-block partials of the
// continuity residual of the new state, of the last stage's projection right-hand side, and
// of the change of u, v, w, p, theta over the step; then one block forms the row.
constexpr int kDiagBlocks = 256;
constexpr int kDV = 25;    // 0 sum div^2, 1 max|div|, 2 sum rhs^2, 3-5 max|u,v,w| old, 6-8 new,
                           // 9 + 3f: sum d^2, max|d|, sum a^2 of field f (u, v, w, p, theta), 24 sum d theta
__device__ __forceinline__ bool diag_is_max(int q) {
  return q == 1 || (q >= 3 && q <= 8) || (q >= 9 && q <= 23 && (q - 9) % 3 == 1);
}

template <typename T>
struct Diag3 {
  const T *uo, *vo, *wo, *po, *to;     // state before the step
  const T *un, *vn, *wn, *pn, *tn;     // after
  const T *uh, *vh, *wh;               // last stage's u^ (its projection's right-hand side)
  const T *t_out, *dt_out;
  T *row_buf;
  long long *count;
  long long cap;
};

template <typename T>
__global__ void __launch_bounds__(256) k_diag(const Dev3<T> *__restrict__ dp, Diag3<T> g) {
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const size_t nc = (size_t)nx * ny * nz;
  const size_t stride = (size_t)gridDim.x * blockDim.x, t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kDV];
#pragma unroll
  for (int q = 0; q < kDV; ++q) v[q] = 0.0;
  const T corr = d.outlet ? d.red->corr[2] : (T)0;
  for (size_t e = t0; e < nc; e += stride) {
    const int i = (int)(e % nx), j = (int)((e / nx) % ny), k = (int)(e / ((size_t)nx * ny));
    const double dv = (double)div_at(d, g.un, g.vn, g.wn, k, j, i, (T)0);
    v[0] += dv * dv;
    v[1] = fmax(v[1], fabs(dv));
    const double dr = (double)div_at(d, g.uh, g.vh, g.wh, k, j, i, corr);
    v[2] += dr * dr;
    const double a = (double)g.pn[e], dd = (double)(g.pn[e] - g.po[e]);
    v[18] += dd * dd;
    v[19] = fmax(v[19], fabs(dd));
    v[20] += a * a;
    if (d.theta) {
      const double at = (double)g.tn[e], dt_ = (double)(g.tn[e] - g.to[e]);
      v[21] += dt_ * dt_;
      v[22] = fmax(v[22], fabs(dt_));
      v[23] += at * at;
      v[24] += dt_;
    }
  }
  const size_t nf[3] = {(size_t)nz * ny * (nx + 1), (size_t)nz * (ny + 1) * nx, (size_t)(nz + 1) * ny * nx};
  const T *fo[3] = {g.uo, g.vo, g.wo}, *fn[3] = {g.un, g.vn, g.wn};
#pragma unroll
  for (int f = 0; f < 3; ++f)
    for (size_t e = t0; e < nf[f]; e += stride) {
      const T an = fn[f][e], ao = fo[f][e];
      const double dd = (double)(an - ao);
      v[9 + 3 * f] += dd * dd;
      v[10 + 3 * f] = fmax(v[10 + 3 * f], fabs(dd));
      v[11 + 3 * f] += (double)an * (double)an;
      v[3 + f] = fmax(v[3 + f], (double)fabs(ao));
      v[6 + f] = fmax(v[6 + f], (double)fabs(an));
    }
  __shared__ double red[kDV * 4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < kDV; ++q) v[q] = diag_is_max(q) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < kDV; ++q) red[wv * kDV + q] = v[q];
  __syncthreads();
  if (threadIdx.x < kDV) {
    const int q = threadIdx.x;
    double r = red[q];
    for (int w2 = 1; w2 < 4; ++w2) r = diag_is_max(q) ? fmax(r, red[w2 * kDV + q]) : r + red[w2 * kDV + q];
    d.pdiag[(size_t)blockIdx.x * kDV + q] = r;
  }
}

template <typename T>
__global__ void __launch_bounds__(64) k_diag_fin(const Dev3<T> *__restrict__ dp, Call3<T> c, Diag3<T> g) {
  const Dev3<T> &d = *dp;
  __shared__ double tot[kDV];
  if (threadIdx.x < kDV) {
    const int q = threadIdx.x;
    double r = 0.0;
    for (int b = 0; b < kDiagBlocks; ++b) {
      const double x = d.pdiag[(size_t)b * kDV + q];
      r = diag_is_max(q) ? fmax(r, x) : r + x;
    }
    tot[q] = r;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const double nc = (double)nx * ny * nz;
  const double numel[5] = {(double)nz * ny * (nx + 1), (double)nz * (ny + 1) * nx, (double)(nz + 1) * ny * nx, nc, nc};
  const double tiny = (double)std::numeric_limits<T>::min();
  const double h = (double)d.h, dV = h * h * h;
  const T dt = *g.dt_out;
  const T umax_old = ((T)tot[3] + (T)tot[4]) + (T)tot[5];
  const T umax_new = ((T)tot[6] + (T)tot[7]) + (T)tot[8];
  T *row = g.row_buf + (size_t)(*g.count % g.cap) * AERO3D_DIAG_COLS;
  const T nan = (T)NAN;
  row[0] = *g.t_out;
  row[1] = dt;
  row[2] = dt * umax_old / d.h;
  row[3] = umax_new;
  row[4] = (T)sqrt(tot[0] / nc);
  row[5] = (T)tot[1];
  row[6] = (T)sqrt(tot[0] / fmax(tot[2], tiny));
  for (int f = 0; f < 5; ++f) {
    T *o = row + 7 + 4 * f;
    if (f == 4 && !d.theta) {
      o[0] = o[1] = o[2] = o[3] = nan;
      continue;
    }
    const double s2 = tot[9 + 3 * f], a2 = tot[11 + 3 * f];
    o[0] = (T)sqrt(s2 * dV);
    o[1] = (T)sqrt(s2 / numel[f]);
    o[2] = (T)tot[10 + 3 * f];
    o[3] = (T)sqrt(s2 / fmax(a2, tiny));
  }
  row[27] = d.theta ? (T)tot[24] * c.th_sc[4] / dt : nan;
  *g.count = *g.count + 1;
}

// P1: rows (pairs) along x, forward. mode 0: f is the right-hand side; 1: divergence of
// (u, v, w); 2: divergence with the outlet correction of stage `stage` (computed here).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_p1(const Dev3<T> *__restrict__ dp, const T *__restrict__ u,
                                                 const T *__restrict__ v, const T *__restrict__ w,
                                                 const T *__restrict__ f, int mode, int stage) {
  extern __shared__ unsigned char smem[];
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ns = d.ns1, nrows = d.ny * d.nz, st = nx + 1;
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + ns * st;
  cplx<T> *TW = B + ns * st;
  double *red = reinterpret_cast<double *>(TW + nx);
  for (int m = threadIdx.x; m < nx; m += kThreads) TW[m] = d.tw[0][m];
  T corr = 0;
  if (mode == 2) {
    double v2[2] = {0.0, 0.0};
    const int np = d.nby * d.nbz;
    for (int q = threadIdx.x; q < np; q += kThreads) {
      v2[0] += d.pout[2 * q];
      v2[1] += d.pout[2 * q + 1];
    }
    block_sum4<2>(v2, red);
    __syncthreads();
    if (threadIdx.x == 0) {
      red[8] = v2[0];
      red[9] = v2[1];
    }
    __syncthreads();
    corr = ((T)red[8] - (T)red[9]) * d.corr_scale;
    if (blockIdx.x == 0 && threadIdx.x == 0) d.red->corr[stage] = corr;
  }
  const int r0 = 2 * ns * blockIdx.x;
  for (int e = threadIdx.x; e < ns * nx; e += kThreads) {
    // natural column q (coalesced loads) into transform position n
    const int sl = fdiv(e, nx, log2_or_neg(nx)), q = e - sl * nx;
    const int ra = r0 + 2 * sl, rb = ra + 1;
    const int n = d.dct[0] ? makhoul_pos(q, nx) : q;
    T xa = 0, xb = 0;
    if (mode == 0) {
      if (ra < nrows) xa = f[(size_t)ra * nx + q];
      if (rb < nrows) xb = f[(size_t)rb * nx + q];
    } else {
      const int ly = log2_or_neg(d.ny);
      const int ka = fdiv(ra, d.ny, ly), kb = fdiv(rb, d.ny, ly);
      if (ra < nrows) xa = div_at(d, u, v, w, ka, ra - ka * d.ny, q, corr);
      if (rb < nrows) xb = div_at(d, u, v, w, kb, rb - kb * d.ny, q, corr);
    }
    A[sl * st + n] = {xa, xb};
  }
  __syncthreads();
  const cplx<T> *R = fft_run(A, B, d.fp[0], TW, ns, st);
  if (d.planes) {
    // spectrum (nx, rows): wavenumber k of row r at k nrows + r; lanes walk the block's rows
    // first (runs of 2 ns consecutive rows per wavenumber)
    const int lns = log2_or_neg(ns);
    for (int e = threadIdx.x; e < ns * nx; e += kThreads) {
      const int k = fdiv(e, ns, lns), sl = e - k * ns;
      const int ra = r0 + 2 * sl, rb = ra + 1;
      T xa, xb;
      fwd_split(R + sl * st, nx, k, d.dct[0], d.dct[0] ? d.cs[0][k] : cplx<T>{0, 0}, xa, xb);
      if (ra < nrows) d.F[(size_t)k * nrows + ra] = xa;
      if (rb < nrows) d.F[(size_t)k * nrows + rb] = xb;
    }
    return;
  }
  for (int e = threadIdx.x; e < ns * nx; e += kThreads) {
    const int sl = fdiv(e, nx, log2_or_neg(nx)), k = e - sl * nx;
    const int ra = r0 + 2 * sl, rb = ra + 1;
    T xa, xb;
    fwd_split(R + sl * st, nx, k, d.dct[0], d.dct[0] ? d.cs[0][k] : cplx<T>{0, 0}, xa, xb);
    if (ra < nrows) d.F[(size_t)ra * nx + k] = xa;
    if (rb < nrows) d.F[(size_t)rb * nx + k] = xb;
  }
}

// P2 / P4: column chunks along y (stride nx) of layer blockIdx.y, forward or inverse.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_py(const Dev3<T> *__restrict__ dp, int inverse) {
  extern __shared__ unsigned char smem[];
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ny = d.ny, ns = d.ns2, ncol = 2 * ns, st = ny + 1;
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + ns * st;
  cplx<T> *TW = B + ns * st;
  T *X = reinterpret_cast<T *>(TW + ny);   // ncol rows of st reals
  for (int m = threadIdx.x; m < ny; m += kThreads) TW[m] = d.tw[1][m];
  const int c0 = ncol * blockIdx.x, k = blockIdx.y;
  T *base = d.F + (size_t)k * ny * nx + c0;
  const int dct = d.dct[1];
  if (!inverse) {
    for (int e = threadIdx.x; e < ncol * ny; e += kThreads) {
      const int n = fdiv(e, ncol, log2_or_neg(ncol)), col = e - n * ncol;
      const int q = dct ? makhoul_src(n, ny) : n;
      const T val = c0 + col < nx ? base[(size_t)q * nx + col] : (T)0;
      reinterpret_cast<T *>(&A[(col >> 1) * st + n])[col & 1] = val;
    }
    __syncthreads();
    const cplx<T> *R = fft_run(A, B, d.fp[1], TW, ns, st);
    for (int e = threadIdx.x; e < ns * ny; e += kThreads) {
      const int sl = fdiv(e, ny, log2_or_neg(ny)), kk = e - sl * ny;
      T xa, xb;
      fwd_split(R + sl * st, ny, kk, dct, dct ? d.cs[1][kk] : cplx<T>{0, 0}, xa, xb);
      X[(2 * sl) * st + kk] = xa;
      X[(2 * sl + 1) * st + kk] = xb;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ncol * ny; e += kThreads) {
      const int kk = fdiv(e, ncol, log2_or_neg(ncol)), col = e - kk * ncol;
      if (c0 + col < nx) base[(size_t)kk * nx + col] = X[col * st + kk];
    }
    return;
  }
  for (int e = threadIdx.x; e < ncol * ny; e += kThreads) {
    const int kk = fdiv(e, ncol, log2_or_neg(ncol)), col = e - kk * ncol;
    X[col * st + kk] = c0 + col < nx ? base[(size_t)kk * nx + col] : (T)0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ns * ny; e += kThreads) {
    const int sl = fdiv(e, ny, log2_or_neg(ny)), kk = e - sl * ny;
    A[sl * st + kk] = pack_inv(X + (2 * sl) * st, X + (2 * sl + 1) * st, ny, dct, dct ? d.cs[1][kk] : cplx<T>{0, 0}, kk);
  }
  __syncthreads();
  const cplx<T> *R = fft_run(A, B, d.fp[1], TW, ns, st);
  const T inv_n = (T)(1.0 / (double)ny);
  for (int e = threadIdx.x; e < ns * ny; e += kThreads) {
    const int sl = fdiv(e, ny, log2_or_neg(ny)), m = e - sl * ny;
    T va, vb;
    inv_unpack(R + sl * st, ny, m, dct, inv_n, va, vb);
    const int q = dct ? makhoul_src(m, ny) : m;
    X[(2 * sl) * st + q] = va;
    X[(2 * sl + 1) * st + q] = vb;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ncol * ny; e += kThreads) {
    const int q = fdiv(e, ncol, log2_or_neg(ncol)), col = e - q * ncol;
    if (c0 + col < nx) base[(size_t)q * nx + col] = X[col * st + q];
  }
}

// P3: column chunks along z (stride ny nx) of row blockIdx.y: forward, x 1/lambda, inverse.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pz(const Dev3<T> *__restrict__ dp) {
  extern __shared__ unsigned char smem[];
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ny = d.ny, nz = d.nz, ns = d.ns3, ncol = 2 * ns, st = nz + 1;
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + ns * st;
  cplx<T> *TW = B + ns * st;
  T *X = reinterpret_cast<T *>(TW + nz);
  for (int m = threadIdx.x; m < nz; m += kThreads) TW[m] = d.tw[2][m];
  const int c0 = ncol * blockIdx.x, j = blockIdx.y;
  const size_t pl = (size_t)ny * nx;
  T *base = d.F + (size_t)j * nx + c0;
  const int dct = d.dct[2];
  for (int e = threadIdx.x; e < ncol * nz; e += kThreads) {
    const int n = fdiv(e, ncol, log2_or_neg(ncol)), col = e - n * ncol;
    const int q = dct ? makhoul_src(n, nz) : n;
    const T val = c0 + col < nx ? base[(size_t)q * pl + col] : (T)0;
    reinterpret_cast<T *>(&A[(col >> 1) * st + n])[col & 1] = val;
  }
  __syncthreads();
  const cplx<T> *R = fft_run(A, B, d.fp[2], TW, ns, st);
  const T *inv = d.inv_lam + (size_t)j * nx * nz;
  for (int e = threadIdx.x; e < ns * nz; e += kThreads) {
    const int sl = fdiv(e, nz, log2_or_neg(nz)), kz = e - sl * nz;
    T xa, xb;
    fwd_split(R + sl * st, nz, kz, dct, dct ? d.cs[2][kz] : cplx<T>{0, 0}, xa, xb);
    const int ca = c0 + 2 * sl, cb = ca + 1;
    const T ia = ca < nx ? inv[(size_t)ca * nz + kz] : (T)0;
    const T ib = cb < nx ? inv[(size_t)cb * nz + kz] : (T)0;
    X[(2 * sl) * st + kz] = xa * ia;
    X[(2 * sl + 1) * st + kz] = xb * ib;
  }
  __syncthreads();
  cplx<T> *Z = (R == A) ? B : A;
  for (int e = threadIdx.x; e < ns * nz; e += kThreads) {
    const int sl = fdiv(e, nz, log2_or_neg(nz)), kz = e - sl * nz;
    Z[sl * st + kz] = pack_inv(X + (2 * sl) * st, X + (2 * sl + 1) * st, nz, dct, dct ? d.cs[2][kz] : cplx<T>{0, 0}, kz);
  }
  __syncthreads();
  const cplx<T> *R2 = fft_run(Z, (Z == A) ? B : A, d.fp[2], TW, ns, st);
  const T inv_n = (T)(1.0 / (double)nz);
  for (int e = threadIdx.x; e < ns * nz; e += kThreads) {
    const int sl = fdiv(e, nz, log2_or_neg(nz)), m = e - sl * nz;
    T va, vb;
    inv_unpack(R2 + sl * st, nz, m, dct, inv_n, va, vb);
    const int q = dct ? makhoul_src(m, nz) : m;
    X[(2 * sl) * st + q] = va;
    X[(2 * sl + 1) * st + q] = vb;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ncol * nz; e += kThreads) {
    const int q = fdiv(e, ncol, log2_or_neg(ncol)), col = e - q * ncol;
    if (c0 + col < nx) base[(size_t)q * pl + col] = X[col * st + q];
  }
}

// P2-P4 in one pass (planes that fit one block's LDS): the (z, y) plane of x wavenumber
// blockIdx.x, contiguous in the spectrum layout (nx, nz, ny) that P1 writes. Forward along y
// (pairs of z rows as one complex sequence), forward along z (pairs of y columns), x 1/lambda,
// inverse along z, inverse along y, all in LDS: one read and one write of the plane instead of
// three of each over the whole field. Two work buffers take turns: the complex sequences of a
// transform, its Stockham partner, then the real rows the next transform packs.
template <typename T>
__global__ void __launch_bounds__(kPlaneThreads) k_planes(const Dev3<T> *__restrict__ dp) {
  constexpr int MP = kPlaneMax / kPlaneThreads;   // plane values per thread
  extern __shared__ unsigned char smem[];
  const Dev3<T> &d = *dp;
  const int ny = d.ny, nz = d.nz, sy = ny + 1, sz = nz + 1, hz = nz / 2, hy = ny / 2, np = ny * nz;
  const int ly = log2_or_neg(ny), lz = log2_or_neg(nz), tid = threadIdx.x;
  cplx<T> *const W0 = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *const W1 = reinterpret_cast<cplx<T> *>(smem + d.plane_buf);
  cplx<T> *TWy = reinterpret_cast<cplx<T> *>(smem + 2 * (size_t)d.plane_buf);
  cplx<T> *TWz = TWy + ny;
  cplx<T> *CSy = TWz + nz;
  cplx<T> *CSz = CSy + ny;
  auto other = [&](const void *b) { return b == (const void *)W0 ? W1 : W0; };
  const int dy = d.dct[1], dz = d.dct[2];
  T *plane = d.F + (size_t)blockIdx.x * np;
  const T *inv = d.inv_lam + (size_t)blockIdx.x * np;   // [ky][kz]
  // every global load of the pass is issued here, before anything waits on one: the plane,
  // the spectral scalings this thread applies, the tables
  T pv[MP], sa[MP / 2], sb[MP / 2];
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int e = tid + kPlaneThreads * r;
    if (e < np) pv[r] = plane[e];
  }
#pragma unroll
  for (int r = 0; r < MP / 2; ++r) {
    const int e = tid + kPlaneThreads * r;
    if (e < hy * nz) {
      const int t = fdiv(e, nz, lz), kz = e - t * nz;
      sa[r] = inv[(size_t)(2 * t) * nz + kz];
      sb[r] = inv[(size_t)(2 * t + 1) * nz + kz];
    }
  }
  for (int m = tid; m < ny; m += kPlaneThreads) {
    TWy[m] = d.tw[1][m];
    CSy[m] = dy ? d.cs[1][m] : cplx<T>{0, 0};
  }
  for (int m = tid; m < nz; m += kPlaneThreads) {
    TWz[m] = d.tw[2][m];
    CSz[m] = dz ? d.cs[2][m] : cplx<T>{0, 0};
  }
  // z row k, y position j -> sequence k / 2 (component k & 1) at the transform position of j
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int e = tid + kPlaneThreads * r;
    if (e < np) {
      const int k = fdiv(e, ny, ly), j = e - k * ny;
      const int n = dy ? makhoul_pos(j, ny) : j;
      reinterpret_cast<T *>(&W0[(k >> 1) * sy + n])[k & 1] = pv[r];
    }
  }
  __syncthreads();
  cplx<T> *R = fft_run(W0, W1, d.fp[1], TWy, hz, sy);
  // y spectra of z rows 2s, 2s + 1 at ky -> z sequences: y column pair ky / 2 (component
  // ky & 1), at the transform position of z row k
  cplx<T> *Z = other(R);
  for (int e = tid; e < hz * ny; e += kPlaneThreads) {
    const int sl = fdiv(e, ny, ly), ky = e - sl * ny;
    T xa, xb;
    fwd_split(R + sl * sy, ny, ky, dy, CSy[ky], xa, xb);
    const int ka = 2 * sl, kb = ka + 1;
    T *zs = reinterpret_cast<T *>(Z + (ky >> 1) * sz);
    zs[2 * (dz ? makhoul_pos(ka, nz) : ka) + (ky & 1)] = xa;
    zs[2 * (dz ? makhoul_pos(kb, nz) : kb) + (ky & 1)] = xb;
  }
  __syncthreads();
  R = fft_run(Z, other(Z), d.fp[2], TWz, hy, sz);
  // z spectra of y columns 2t, 2t + 1, x 1/lambda -> real rows X[ky][kz]
  T *X = reinterpret_cast<T *>(other(R));
#pragma unroll
  for (int r = 0; r < MP / 2; ++r) {
    const int e = tid + kPlaneThreads * r;
    if (e < hy * nz) {
      const int t = fdiv(e, nz, lz), kz = e - t * nz;
      T xa, xb;
      fwd_split(R + t * sz, nz, kz, dz, CSz[kz], xa, xb);
      X[(2 * t) * sz + kz] = xa * sa[r];
      X[(2 * t + 1) * sz + kz] = xb * sb[r];
    }
  }
  __syncthreads();
  Z = R;
  for (int e = tid; e < hy * nz; e += kPlaneThreads) {
    const int t = fdiv(e, nz, lz), kz = e - t * nz;
    Z[t * sz + kz] = pack_inv(X + (2 * t) * sz, X + (2 * t + 1) * sz, nz, dz, CSz[kz], kz);
  }
  __syncthreads();
  R = fft_run(Z, other(Z), d.fp[2], TWz, hy, sz);
  // values at z row q of y columns 2t, 2t + 1 -> real z rows X[q][ky]
  X = reinterpret_cast<T *>(other(R));
  const T inv_nz = (T)(1.0 / (double)nz), inv_ny = (T)(1.0 / (double)ny);
  for (int e = tid; e < hy * nz; e += kPlaneThreads) {
    const int t = fdiv(e, nz, lz), m = e - t * nz;
    T va, vb;
    inv_unpack(R + t * sz, nz, m, dz, inv_nz, va, vb);
    const int q = dz ? makhoul_src(m, nz) : m;
    X[q * sy + 2 * t] = va;
    X[q * sy + 2 * t + 1] = vb;
  }
  __syncthreads();
  Z = R;
  for (int e = tid; e < hz * ny; e += kPlaneThreads) {
    const int sl = fdiv(e, ny, ly), ky = e - sl * ny;
    Z[sl * sy + ky] = pack_inv(X + (2 * sl) * sy, X + (2 * sl + 1) * sy, ny, dy, CSy[ky], ky);
  }
  __syncthreads();
  R = fft_run(Z, other(Z), d.fp[1], TWy, hz, sy);
  // values at y position q of z rows 2s, 2s + 1 -> the plane, then stored in natural order
  X = reinterpret_cast<T *>(other(R));
  for (int e = tid; e < hz * ny; e += kPlaneThreads) {
    const int sl = fdiv(e, ny, ly), m = e - sl * ny;
    T va, vb;
    inv_unpack(R + sl * sy, ny, m, dy, inv_ny, va, vb);
    const int q = dy ? makhoul_src(m, ny) : m;
    X[(2 * sl) * sy + q] = va;
    X[(2 * sl + 1) * sy + q] = vb;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MP; ++r) {
    const int e = tid + kPlaneThreads * r;
    if (e < np) {
      const int k = fdiv(e, ny, ly), j = e - k * ny;
      plane[e] = X[k * sy + j];
    }
  }
}

// P5: rows (pairs) along x, inverse -> psi; the step-averaged pressure p (core.py:874-875):
// stage 0 p = w psi, stage 1 p = p + w psi, stage 2 p = (p + psi) (1 / dt).
template <typename T>
__global__ void __launch_bounds__(kP5Threads) k_p5(const Dev3<T> *__restrict__ dp, Call3<T> c, T *__restrict__ psi,
                                                 T *__restrict__ p, int stage) {
  extern __shared__ unsigned char smem[];
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ns = d.ns5, nrows = d.ny * d.nz, st = nx + 1;
  cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
  cplx<T> *B = A + ns * st;
  cplx<T> *TW = B + ns * st;
  // the real rows (2 ns of st values) share the transform's second buffer, which is free before
  // the transform; after it they move to whichever buffer does not hold the result
  T *X = reinterpret_cast<T *>(B);
  for (int m = threadIdx.x; m < nx; m += kP5Threads) TW[m] = d.tw[0][m];
  const int r0 = 2 * ns * blockIdx.x;
  const int dct = d.dct[0];
  if (d.planes) {
    // spectrum (nx, rows): lanes walk the block's rows first
    const int l2 = log2_or_neg(2 * ns);
    for (int e = threadIdx.x; e < 2 * ns * nx; e += kP5Threads) {
      const int k = fdiv(e, 2 * ns, l2), rr = e - k * 2 * ns;
      X[rr * st + k] = r0 + rr < nrows ? d.F[(size_t)k * nrows + r0 + rr] : (T)0;
    }
  } else {
    for (int e = threadIdx.x; e < 2 * ns * nx; e += kP5Threads) {
      const int rr = fdiv(e, nx, log2_or_neg(nx)), k = e - rr * nx;
      X[rr * st + k] = r0 + rr < nrows ? d.F[(size_t)(r0 + rr) * nx + k] : (T)0;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ns * nx; e += kP5Threads) {
    const int sl = fdiv(e, nx, log2_or_neg(nx)), k = e - sl * nx;
    A[sl * st + k] = pack_inv(X + (2 * sl) * st, X + (2 * sl + 1) * st, nx, dct, dct ? d.cs[0][k] : cplx<T>{0, 0}, k);
  }
  __syncthreads();
  const cplx<T> *R = fft_run(A, B, d.fp[0], TW, ns, st);
  X = reinterpret_cast<T *>(R == A ? B : A);
  const T inv_n = (T)(1.0 / (double)nx);
  T rdt = 0;
  if (p && stage == 2) rdt = (T)1 / step_dt(d, c);
  const T w = (T)kForceW[stage];
  for (int e = threadIdx.x; e < ns * nx; e += kP5Threads) {
    const int sl = fdiv(e, nx, log2_or_neg(nx)), m = e - sl * nx;
    T va, vb;
    inv_unpack(R + sl * st, nx, m, dct, inv_n, va, vb);
    const int q = dct ? makhoul_src(m, nx) : m;
    X[(2 * sl) * st + q] = va;
    X[(2 * sl + 1) * st + q] = vb;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * ns * nx; e += kP5Threads) {
    const int rr = fdiv(e, nx, log2_or_neg(nx)), q = e - rr * nx;
    if (r0 + rr >= nrows) continue;
    const size_t idx = (size_t)(r0 + rr) * nx + q;
    const T ps = X[rr * st + q];
    psi[idx] = ps;
    if (p) {
      if (stage == 0) p[idx] = ps * w;
      else if (stage == 1) p[idx] = p[idx] + w * ps;
      else p[idx] = (p[idx] + w * ps) * rdt;
    }
  }
}

// u -= grad psi in place, boundary-normal faces untouched (core.py:582-612).
template <typename T>
__global__ void k_correct(const Dev3<T> *__restrict__ dp, T *u, T *v, T *w, const T *__restrict__ psi) {
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const size_t nu = (size_t)nz * ny * (nx + 1), nv = (size_t)nz * (ny + 1) * nx, nw = (size_t)(nz + 1) * ny * nx;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nu + nv + nw; e += (size_t)gridDim.x * blockDim.x) {
    if (e < nu) {
      const int i = e % (nx + 1);
      const size_t r = e / (nx + 1);
      const T *ps = psi + r * nx;
      if (d.per[0]) {
        const int ii = i == nx ? 0 : i, im = ii == 0 ? nx - 1 : ii - 1;
        u[e] = u[e] - (ps[ii] - ps[im]) * d.inv_h;
      } else if (i > 0 && i < nx) {
        u[e] = u[e] - (ps[i] - ps[i - 1]) * d.inv_h;
      }
    } else if (e < nu + nv) {
      const size_t f = e - nu;
      const int i = f % nx;
      const int j = (f / nx) % (ny + 1);
      const size_t k = f / ((size_t)nx * (ny + 1));
      const T *ps = psi + k * ny * nx + i;
      if (d.per[1]) {
        const int jj = j == ny ? 0 : j, jm = jj == 0 ? ny - 1 : jj - 1;
        v[f] = v[f] - (ps[(size_t)jj * nx] - ps[(size_t)jm * nx]) * d.inv_h;
      } else if (j > 0 && j < ny) {
        v[f] = v[f] - (ps[(size_t)j * nx] - ps[(size_t)(j - 1) * nx]) * d.inv_h;
      }
    } else {
      const size_t f = e - nu - nv;
      const size_t pl = (size_t)ny * nx;
      const int k = (int)(f / pl);
      const size_t r = f % pl;
      const T *ps = psi + r;
      if (d.per[2]) {
        const int kk = k == nz ? 0 : k, km = kk == 0 ? nz - 1 : kk - 1;
        w[f] = w[f] - (ps[kk * pl] - ps[km * pl]) * d.inv_h;
      } else if (k > 0 && k < nz) {
        w[f] = w[f] - (ps[k * pl] - ps[(k - 1) * pl]) * d.inv_h;
      }
    }
  }
}

// Traced stage (core.py:793-862 trace records), after the stage's potential is solved: the
// outlet correction of the stage added on the outlet face of the value before penalisation and
// of the penalised value, the projection's right-hand side and the projected velocity as the
// next stage loads it (u^ - grad psi on interior and periodic faces, the outlet shift on the
// outlet face). Null outputs are skipped.
template <typename T>
__global__ void k_trace3(const Dev3<T> *__restrict__ dp, int stage, const T *__restrict__ uh, const T *__restrict__ vh,
                         const T *__restrict__ wh, const T *__restrict__ psi, T *star_u, T *pen_u, T *pen_v,
                         T *pen_w, T *div, T *pu, T *pv, T *pw) {
  const Dev3<T> &d = *dp;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const T corr = d.outlet ? d.red->corr[stage] : (T)0;
  const size_t nu = (size_t)nz * ny * (nx + 1), nv = (size_t)nz * (ny + 1) * nx, nw = (size_t)(nz + 1) * ny * nx;
  const size_t pl = (size_t)ny * nx, nc = pl * nz;
  const T inv_h = d.inv_h;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nu + nv + nw + nc;
       e += (size_t)gridDim.x * blockDim.x) {
    if (e < nu) {
      const int i = (int)(e % (nx + 1));
      const size_t r = e / (nx + 1);
      const bool out_face = d.outlet && i == nx;
      T val = uh[e];
      if (out_face) {
        val = val + corr;
        if (star_u) star_u[e] = star_u[e] + corr;
      }
      if (pen_u) pen_u[e] = val;
      if (d.per[0] || (i > 0 && i < nx)) {
        const int fc = i == nx ? 0 : i, fm = fc == 0 ? nx - 1 : fc - 1;
        val = val - (psi[r * nx + fc] - psi[r * nx + fm]) * inv_h;
      }
      if (pu) pu[e] = val;
    } else if (e < nu + nv) {
      const size_t f = e - nu;
      const int i = (int)(f % nx), j = (int)((f / nx) % (ny + 1));
      const size_t k = f / ((size_t)nx * (ny + 1));
      T val = vh[f];
      if (pen_v) pen_v[f] = val;
      if (d.per[1] || (j > 0 && j < ny)) {
        const int fc = j == ny ? 0 : j, fm = fc == 0 ? ny - 1 : fc - 1;
        const T *pb = psi + k * pl + i;
        val = val - (pb[(size_t)fc * nx] - pb[(size_t)fm * nx]) * inv_h;
      }
      if (pv) pv[f] = val;
    } else if (e < nu + nv + nw) {
      const size_t f = e - nu - nv;
      const int k = (int)(f / pl);
      const size_t r = f % pl;
      T val = wh[f];
      if (pen_w) pen_w[f] = val;
      if (d.per[2] || (k > 0 && k < nz)) {
        const int fc = k == nz ? 0 : k, fm = fc == 0 ? nz - 1 : fc - 1;
        val = val - (psi[(size_t)fc * pl + r] - psi[(size_t)fm * pl + r]) * inv_h;
      }
      if (pw) pw[f] = val;
    } else if (div) {
      const size_t q = e - nu - nv - nw;
      const int k = (int)(q / pl), j = (int)((q % pl) / nx), i = (int)(q % nx);
      div[q] = div_at(d, uh, vh, wh, k, j, i, corr);
    }
  }
}

}  // namespace ns3d
}  // namespace aero

// ====================================================================== host side
using namespace aero;
using namespace aero::ns3d;

struct aero3d_plan {
  aero3d_desc desc;
  int dbl;
  size_t esize;
  size_t bytes = 0;
  std::vector<void *> allocs;
  void *dev = nullptr;
  void *uA, *vA, *wA, *uB, *vB, *wB, *F, *psi, *acc_u, *acc_v, *acc_w, *pmax, *pout, *pforce, *red, *dye_tmp;
  void *th_a, *th_b, *nut_c, *pheat;
  void *nu_pad;
  void *old_u, *old_v, *old_w, *old_p, *old_t, *pdiag;
  void *tw[3], *cs[3];
  double *lam[3];
  void *inv_lam;
  FftPlan fp[3];
  int nbx, nby, nbz, nblk, nwt, nq;
  int ns1, ns2, ns3, ns5;
  size_t lds1, lds2, lds3, lds5;
  int planes;           // the fused (z, y) plane pass (FP32 planes that fit one block's LDS)
  size_t plane_buf, lds_planes;
  int nwt_alloc;
};

namespace {

int alloc3(aero3d_plan *p, void **ptr, size_t n) {
  *ptr = nullptr;
  if (n == 0) return AERO_OK;
  AERO_HIP_CHECK(hipMalloc(ptr, n));
  p->allocs.push_back(*ptr);
  p->bytes += n;
  AERO_HIP_CHECK(hipMemset(*ptr, 0, n));
  return AERO_OK;
}

template <typename T>
int tile_dims(int &tx, int &ty, int &tz) {
  tx = Tile<T>::X;
  ty = Tile<T>::Y;
  tz = Tile<T>::Z;
  return 0;
}

template <typename T>
int upload3(aero3d_plan *p) {
  const aero3d_desc &s = p->desc;
  const int n3[3] = {s.nx, s.ny, s.nz};
  const int per[3] = {s.periodic_x, s.periodic_y, s.periodic_z};
  std::vector<double> lams[3];
  for (int a = 0; a < 3; ++a) {
    const int n = n3[a];
    std::vector<cplx<T>> tw(n), cs(n);
    std::vector<double> lam(n);
    const double hh = s.h * s.h;
    for (int m = 0; m < n; ++m) {
      const double ang = -2.0 * M_PI * (double)m / (double)n;
      tw[m] = {(T)std::cos(ang), (T)std::sin(ang)};
      const double th = M_PI * (double)m / (2.0 * (double)n);
      cs[m] = {(T)std::cos(th), (T)std::sin(th)};
      lam[m] = per[a] ? (2.0 * std::cos(2.0 * M_PI * (double)m / (double)n) - 2.0) / hh
                      : (2.0 * std::cos(M_PI * (double)m / (double)n) - 2.0) / hh;
    }
    std::vector<cplx<T>> tws(n);
    stage_twiddles(p->fp[a], tw.data(), tws.data());
    AERO_HIP_CHECK(hipMemcpy(p->tw[a], tws.data(), n * sizeof(cplx<T>), hipMemcpyHostToDevice));
    AERO_HIP_CHECK(hipMemcpy(p->cs[a], cs.data(), n * sizeof(cplx<T>), hipMemcpyHostToDevice));
    AERO_HIP_CHECK(hipMemcpy(p->lam[a], lam.data(), n * sizeof(double), hipMemcpyHostToDevice));
    lams[a] = lam;
  }
  {
    // the spectral scaling 1 / (lam_x + lam_y + lam_z), formed in float64 in the order of
    // ns3d/poisson.py:115-125 and rounded once; the zero mode is 0
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    std::vector<T> inv((size_t)nx * ny * nz);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
          const double la = (lams[2][k] + lams[1][j]) + lams[0][i];
          const size_t at = p->planes ? ((size_t)i * ny + j) * nz + k : ((size_t)j * nx + i) * nz + k;
          inv[at] = (T)((la != 0.0 && !(i == 0 && j == 0 && k == 0)) ? 1.0 / la : 0.0);
        }
    AERO_HIP_CHECK(hipMemcpy(p->inv_lam, inv.data(), inv.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  Dev3<T> d;
  std::memset(&d, 0, sizeof(d));
  d.nx = s.nx;
  d.ny = s.ny;
  d.nz = s.nz;
  for (int a = 0; a < 3; ++a) d.per[a] = per[a];
  d.closed = s.closed;
  d.outlet = !s.periodic_x && !s.closed;
  d.upwind = s.advection == 1;
  d.turb = s.turbulence;
  d.fixed_dt = s.fixed_dt;
  d.pid = s.pid_dt;
  d.pid_target = (T)s.pid_target;
  d.nbx = p->nbx;
  d.nby = p->nby;
  d.nbz = p->nbz;
  d.nblk = p->nblk;
  d.xcd_chunk = (p->nblk % 8 == 0 && p->nblk >= 64) ? p->nblk / 8 : 0;
  d.win = s.win_k1 > s.win_k0 && s.win_j1 > s.win_j0 && s.win_i1 > s.win_i0 && s.nbody > 0;
  d.wk0 = s.win_k0;
  d.wk1 = s.win_k1;
  d.wj0 = s.win_j0;
  d.wj1 = s.win_j1;
  d.wi0 = s.win_i0;
  d.wi1 = s.win_i1;
  d.wz = d.win ? s.win_k1 - s.win_k0 : 0;
  d.wy = d.win ? s.win_j1 - s.win_j0 : 0;
  d.wx = d.win ? s.win_i1 - s.win_i0 : 0;
  int tx, ty, tz;
  tile_dims<T>(tx, ty, tz);
  if (d.win) {
    d.wbx0 = s.win_i0 / tx;
    d.wby0 = s.win_j0 / ty;
    d.wbz0 = s.win_k0 / tz;
    d.nwbx = std::min(s.win_i1 / tx, p->nbx - 1) - d.wbx0 + 1;
    d.nwby = std::min(s.win_j1 / ty, p->nby - 1) - d.wby0 + 1;
    d.nwbz = std::min(s.win_k1 / tz, p->nbz - 1) - d.wbz0 + 1;
    d.nwt = d.nwbx * d.nwby * d.nwbz;
    d.nq = 9 * s.nbody;
  }
  d.nbody = d.win ? s.nbody : 0;
  d.symmetry = s.symmetry;
  d.nut_model = s.nut_model;
  d.nut_trust = (T)s.nut_trust;
  d.dye = s.dye;
  if (s.dye && s.dye_k1 > s.dye_k0 && s.dye_j1 > s.dye_j0 && s.dye_i1 > s.dye_i0) {
    d.dk0 = s.dye_k0; d.dk1 = s.dye_k1; d.dj0 = s.dye_j0; d.dj1 = s.dye_j1; d.di0 = s.dye_i0; d.di1 = s.dye_i1;
    d.dwx = s.dye_i1 - s.dye_i0;
    d.dwy = s.dye_j1 - s.dye_j0;
  }
  for (int q = 0; q < 18; ++q) {
    d.rule[q] = s.ghost_kind[q];
    d.rval2[q] = (T)(2.0 * s.ghost_value[q]);
  }
  const double h = s.h;
  d.h = (T)h;
  d.inv_h = (T)(1.0 / h);
  d.inv_h2 = (T)(1.0 / (h * h));
  d.kadv = (T)((s.advection == 1 ? 7.0 / 24.0 : 0.25) / h);
  const double delta2 = h * h;
  d.les_k = (T)(s.turbulence == 3 ? s.les_c * delta2 : s.les_c * s.les_c * delta2);
  d.cfl_h = (T)(s.cfl * h);
  d.visc_h2 = (T)(kViscousDt * h * h);
  d.corr_scale = (T)(1.0 / ((double)s.nz * (double)s.ny));
  d.neg_h3 = (T)(-(h * h * h));
  const double q_a = 0.5 * s.ref_area;
  d.q_a = (T)q_a;
  d.q_a_span = (T)(q_a * s.span_ref);
  d.q_a_len = (T)(q_a * s.length_ref);
  d.dy = (T)(s.moment_ref[1] - s.y0);
  d.F = (T *)p->F;
  d.psi = (T *)p->psi;
  d.acc_u = (T *)p->acc_u;
  d.acc_v = (T *)p->acc_v;
  d.acc_w = (T *)p->acc_w;
  d.pmax = (T *)p->pmax;
  d.pout = (double *)p->pout;
  d.pforce = (double *)p->pforce;
  d.red = (Red3<T> *)p->red;
  for (int a = 0; a < 3; ++a) {
    d.tw[a] = (const cplx<T> *)p->tw[a];
    d.cs[a] = (const cplx<T> *)p->cs[a];
    d.lam[a] = p->lam[a];
    d.inv_lam = (const T *)p->inv_lam;
    d.fp[a] = p->fp[a];
    d.dct[a] = !per[a];
  }
  d.ns1 = p->ns1;
  d.ns2 = p->ns2;
  d.ns3 = p->ns3;
  d.ns5 = p->ns5;
  d.planes = p->planes;
  d.plane_buf = (int)p->plane_buf;
  if (s.thermal) {
    d.theta = 1;
    d.buoy = s.richardson != 0.0;
    for (int q = 0; q < 6; ++q) d.tside[q] = s.theta_side[q];
    if (s.theta_q_k1 > s.theta_q_k0 && s.theta_q_j1 > s.theta_q_j0 && s.theta_q_i1 > s.theta_q_i0) {
      d.tqk0 = s.theta_q_k0; d.tqk1 = s.theta_q_k1; d.tqj0 = s.theta_q_j0; d.tqj1 = s.theta_q_j1;
      d.tqi0 = s.theta_q_i0; d.tqi1 = s.theta_q_i1;
      d.tqwx = s.theta_q_i1 - s.theta_q_i0;
      d.tqwy = s.theta_q_j1 - s.theta_q_j0;
    }
    d.tnt = d.win ? s.theta_nt : 0;
    for (int q = 0; q < kMaxT; ++q) d.tbody[q] = s.theta_tbody[q];
    d.tadia = d.win ? s.theta_adia : 0;
    d.tnskin = d.win ? s.theta_nskin : 0;
    const double ri = s.richardson, ws[3] = {kForceW[0], kForceW[1], kForceW[2]};
    d.ri = (T)ri;
    d.half_ri = (T)(0.5 * ri);
    for (int q = 0; q < 3; ++q) d.ri_w[q] = (T)(ri * ws[q]);
    d.abs_ri = (T)std::fabs(ri);
    d.prt = (T)s.prandtl_turbulent;
    d.kc = (T)(0.5 / (s.prandtl_turbulent * h));
    d.cadv = (T)(1.0 / (0.5 * h));
    d.cfl = (T)s.cfl;
    d.c6 = (T)(6.0 * s.cfl * h);
    d.h3 = (T)3 * (T)h;
    d.th_a = (T *)p->th_a;
    d.th_b = (T *)p->th_b;
    d.nut_c = (T *)p->nut_c;
    d.pheat = (double *)p->pheat;
  }
  d.pdiag = (double *)p->pdiag;
  d.nu_pad = (T *)p->nu_pad;
  AERO_HIP_CHECK(hipMemcpy(p->dev, &d, sizeof(d), hipMemcpyHostToDevice));
  // the tile kernels take more than the default 64 KB of dynamic LDS
  int dev_id = 0, max_lds = 0;
  AERO_HIP_CHECK(hipGetDevice(&dev_id));
  AERO_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id));
  if ((size_t)max_lds < sizeof(Lds3<T>)) {
    set_error("the device offers " + std::to_string(max_lds) + " B of LDS per block; the 3D tiles need " +
              std::to_string(sizeof(Lds3<T>)));
    return AERO_EUNSUPPORTED;
  }
  AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_stage<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(Lds3<T>)));
  AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_stage<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(Lds3<T>)));
  AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_final<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(Lds3<T>)));
  AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_grad<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(Lds3<T>)));
  if (sizeof(LdsT<T>) > 64 * 1024)
    AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_theta<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sizeof(LdsT<T>)));
  if (p->planes)
    AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_planes<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)p->lds_planes));
  return AERO_OK;
}

// Largest power-of-two slot count (<= cap) whose LDS use stays within 64 KB: two columns per
// slot. The y and z passes are capped at 8 slots (16 columns, 64-byte row segments): more,
// smaller blocks in flight beat whole 128-byte lines per block (k_pz 50.5 -> 39.9 us,
// k_py 30.6 -> 26.6 us at 256x128x128).
size_t pick_slots(int n, size_t es, int extra_reals, int cap, int &ns) {
  const size_t c = 2 * es;   // sizeof(cplx<T>)
  ns = cap;
  auto need = [&](int k) { return 2 * (size_t)k * (n + 1) * c + (size_t)n * c + (size_t)extra_reals * k * (n + 1) * es + 128; };
  while (ns > 1 && need(ns) > 64 * 1024) ns >>= 1;
  return need(ns);
}

template <typename T>
Call3<T> make_call(const aero3d_fields *in, const aero3d_geom *g, const aero3d_params *pr) {
  Call3<T> c;
  std::memset(&c, 0, sizeof(c));
  c.u0 = (const T *)in->u;
  c.v0 = (const T *)in->v;
  c.w0 = (const T *)in->w;
  c.up = c.u0;
  c.vp = c.v0;
  c.wp = c.w0;
  if (g) {
    c.chi_u = (const T *)g->chi_u;
    c.chi_v = (const T *)g->chi_v;
    c.chi_w = (const T *)g->chi_w;
    c.chi_c = (const T *)g->chi_c;
    c.a_u = (const T *)g->a_u;
    c.a_v = (const T *)g->a_v;
    c.a_w = (const T *)g->a_w;
    c.k_u = (const T *)g->k_u;
    c.k_v = (const T *)g->k_v;
    c.k_w = (const T *)g->k_w;
    c.chi_t = (const T *)g->chi_t;
    c.a_t = (const T *)g->a_t;
    c.w_t = (const T *)g->w_t;
    c.adia_u = (const T *)g->adia_u;
    c.adia_v = (const T *)g->adia_v;
    c.adia_w = (const T *)g->adia_w;
    c.vol_w = (const T *)g->vol_w;
    c.tq = (const T *)g->theta_q;
    c.skin_idx = (const int *)g->skin_idx;
    c.skin_off = (const int *)g->skin_off;
    c.skin_w = (const T *)g->skin_w;
    c.body_theta = (const T *)g->body_theta;
    c.flux_rate = (const T *)g->flux_rate;
  }
  c.th0 = (const T *)in->theta;
  c.thp = c.th0;
  if (pr) {
    c.nu = (const T *)pr->nu;
    c.eta = (const T *)pr->eta;
    c.u_inlet = (const T *)pr->u_inlet;
    c.dt_fixed = (const T *)pr->dt_fixed;
    c.nut_field = (const T *)pr->nut_field;
    c.dye_weight = (const T *)pr->dye_weight;
    c.dye_pulse = (const T *)pr->dye_pulse;
    c.th_sc = (const T *)pr->theta_scalars;
    c.th_side = (const T *)pr->theta_side_values;
  }
  c.corr_stage = -1;
  return c;
}

template <typename T>
int poisson_passes(aero3d_plan *p, hipStream_t st, const Dev3<T> *dev, Call3<T> c, const T *f, int mode, int stage,
                   T *psi, T *pr) {
  const aero3d_desc &s = p->desc;
  const int rows = s.ny * s.nz;
  const int g1 = (rows + 2 * p->ns1 - 1) / (2 * p->ns1);
  hipLaunchKernelGGL(k_p1<T>, dim3(g1), dim3(kThreads), p->lds1, st, dev, c.uo, c.vo, c.wo, f, mode, stage);
  if (p->planes) {
    hipLaunchKernelGGL(k_planes<T>, dim3(s.nx), dim3(kPlaneThreads), p->lds_planes, st, dev);
  } else {
    const dim3 g2((s.nx + 2 * p->ns2 - 1) / (2 * p->ns2), s.nz);
    hipLaunchKernelGGL(k_py<T>, g2, dim3(kThreads), p->lds2, st, dev, 0);
    const dim3 g3((s.nx + 2 * p->ns3 - 1) / (2 * p->ns3), s.ny);
    hipLaunchKernelGGL(k_pz<T>, g3, dim3(kThreads), p->lds3, st, dev);
    hipLaunchKernelGGL(k_py<T>, g2, dim3(kThreads), p->lds2, st, dev, 1);
  }
  const int g5 = (rows + 2 * p->ns5 - 1) / (2 * p->ns5);
  hipLaunchKernelGGL(k_p5<T>, dim3(g5), dim3(kP5Threads), p->lds5, st, dev, c, psi, pr, stage);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int step_impl(aero3d_plan *p, hipStream_t st, const aero3d_fields *in, aero3d_fields *out, const aero3d_geom *g,
              const aero3d_params *pr, aero3d_outputs *o, bool chained, const aero3d_trace *tr = nullptr) {
  const aero3d_desc &s = p->desc;
  const Dev3<T> *dev = (const Dev3<T> *)p->dev;
  Call3<T> c = make_call<T>(in, g, pr);
  c.ctrl_in = (const T *)in->ctrl;
  c.ctrl_out = (T *)out->ctrl;
  const dim3 grid(p->nblk), blk(kTileThreads);
  const size_t lds = sizeof(Lds3<T>);
  const bool diag = o && o->diag;
  const size_t ncell = (size_t)s.nx * s.ny * s.nz;
  if (diag) {
    // the state before the step (the step may run in place)
    const size_t n3[5] = {(size_t)s.nz * s.ny * (s.nx + 1), (size_t)s.nz * (s.ny + 1) * s.nx,
                          (size_t)(s.nz + 1) * s.ny * s.nx, ncell, ncell};
    void *dst[5] = {p->old_u, p->old_v, p->old_w, p->old_p, p->old_t};
    const void *src[5] = {in->u, in->v, in->w, in->p, in->theta};
    for (int f = 0; f < (s.thermal ? 5 : 4); ++f) {
      const int nb = (int)std::min<size_t>(4096, (n3[f] + 255) / 256);
      hipLaunchKernelGGL(k_copy3<T>, dim3(nb), dim3(256), 0, st, (T *)dst[f], (const T *)src[f], n3[f]);
    }
  }
  if (!s.fixed_dt && !chained) {
    hipLaunchKernelGGL(k_final<T>, grid, blk, lds, st, dev, c, 0);
    hipLaunchKernelGGL(k_dt<T>, dim3(1), dim3(kFinThreads), 0, st, dev, c);
  }
  T *ub[2] = {(T *)p->uA, (T *)p->uB}, *vb[2] = {(T *)p->vA, (T *)p->vB}, *wb[2] = {(T *)p->wA, (T *)p->wB};
  for (int stage = 0; stage < 3; ++stage) {
    c.stage = stage;
    c.up = stage == 0 ? (const T *)in->u : ub[(stage - 1) & 1];
    c.vp = stage == 0 ? (const T *)in->v : vb[(stage - 1) & 1];
    c.wp = stage == 0 ? (const T *)in->w : wb[(stage - 1) & 1];
    c.psi = stage == 0 ? nullptr : (const T *)p->psi;
    c.corr_stage = stage == 0 ? -1 : stage - 1;
    // the step start's nu_t was saved by this step's dt pass or, chained, by the previous
    // step's final kernel
    c.nu_in = (stage == 0 && p->nu_pad && (chained || !s.fixed_dt)) ? 1 : 0;
    c.uo = ub[stage & 1];
    c.vo = vb[stage & 1];
    c.wo = wb[stage & 1];
    if (s.thermal) {
      c.thp = stage == 0 ? (const T *)in->theta : (stage == 1 ? (const T *)p->th_a : (const T *)p->th_b);
      c.tho = stage == 0 ? (T *)p->th_a : (stage == 1 ? (T *)p->th_b : (T *)out->theta);
    }
    if (tr) {
      // the traced step forms the step start's nu_t in the stage kernel (the same values)
      c.nu_in = 0;
      c.tr_lu = (T *)tr->tend_u[stage];
      c.tr_lv = (T *)tr->tend_v[stage];
      c.tr_lw = (T *)tr->tend_w[stage];
      c.tr_su = (T *)tr->star_u[stage];
      c.tr_sv = (T *)tr->star_v[stage];
      c.tr_sw = (T *)tr->star_w[stage];
      hipLaunchKernelGGL((k_stage<T, false, true>), grid, blk, lds, st, dev, c);
    } else if (c.nu_in) {
      hipLaunchKernelGGL((k_stage<T, true>), grid, blk, lds, st, dev, c);
    } else {
      hipLaunchKernelGGL((k_stage<T, false>), grid, blk, lds, st, dev, c);
    }
    if (s.thermal) hipLaunchKernelGGL(k_theta<T>, grid, blk, sizeof(LdsT<T>), st, dev, c);
    const int mode = (!s.periodic_x && !s.closed) ? 2 : 1;
    int rc = poisson_passes<T>(p, st, dev, c, nullptr, mode, stage, (T *)p->psi, (T *)out->p);
    if (rc) return rc;
    if (tr) {
      const size_t nf = (size_t)s.nz * s.ny * (s.nx + 1) + (size_t)s.nz * (s.ny + 1) * s.nx +
                        (size_t)(s.nz + 1) * s.ny * s.nx + ncell;
      const int nb = (int)std::min<size_t>(8192, (nf + 255) / 256);
      hipLaunchKernelGGL(k_trace3<T>, dim3(nb), dim3(256), 0, st, dev, stage, (const T *)c.uo, (const T *)c.vo,
                         (const T *)c.wo, (const T *)p->psi, (T *)tr->star_u[stage], (T *)tr->pen_u[stage],
                         (T *)tr->pen_v[stage], (T *)tr->pen_w[stage], (T *)tr->div[stage], (T *)tr->proj_u[stage],
                         (T *)tr->proj_v[stage], (T *)tr->proj_w[stage]);
      const int cb = (int)std::min<size_t>(8192, (ncell + 255) / 256);
      if (tr->psi[stage])
        hipLaunchKernelGGL(k_copy3<T>, dim3(cb), dim3(256), 0, st, (T *)tr->psi[stage], (const T *)p->psi, ncell);
      if (tr->theta[stage] && s.thermal)
        hipLaunchKernelGGL(k_copy3<T>, dim3(cb), dim3(256), 0, st, (T *)tr->theta[stage], (const T *)c.tho, ncell);
      c.tr_lu = c.tr_lv = c.tr_lw = c.tr_su = c.tr_sv = c.tr_sw = nullptr;
    }
  }
  c.stage = 3;
  c.nu_in = 0;
  c.up = ub[0];
  c.vp = vb[0];
  c.wp = wb[0];
  c.psi = (const T *)p->psi;
  c.corr_stage = (!s.periodic_x && !s.closed) ? 2 : -1;
  c.uo = (T *)out->u;
  c.vo = (T *)out->v;
  c.wo = (T *)out->w;
  hipLaunchKernelGGL(k_final<T>, grid, blk, lds, st, dev, c, 1);
  if (s.dye && in->dye && out->dye) {
    // smoke with the end-of-step velocity, before the finalize advances t
    Call3<T> cd = c;
    cd.up = (const T *)out->u;
    cd.vp = (const T *)out->v;
    cd.wp = (const T *)out->w;
    const size_t n = (size_t)s.nx * s.ny * s.nz;
    const int nb = (int)std::min<size_t>(8192, (n + 255) / 256);
    T *dst = in->dye == out->dye ? (T *)p->dye_tmp : (T *)out->dye;
    hipLaunchKernelGGL(k_dye<T>, dim3(nb), dim3(256), 0, st, dev, cd, (const T *)in->dye, dst, (const T *)in->t,
                       (T)1e-8);
    if (dst != (T *)out->dye) hipLaunchKernelGGL(k_copy3<T>, dim3(nb), dim3(256), 0, st, (T *)out->dye, (const T *)dst, n);
  }
  Fin3<T> f;
  std::memset(&f, 0, sizeof(f));
  f.t_in = (const T *)in->t;
  f.t_out = (T *)out->t;
  f.dt_out = (T *)out->dt;
  if (o) {
    f.coeffs = (T *)o->coeffs;
    f.hist = (T *)o->hist;
    f.hist_count = (long long *)o->hist_count;
    f.hist_cap = o->hist_cap;
  }
  f.next_dt = 1;
  if (s.thermal) {
    f.theta = (const T *)out->theta;
    if (o) {
      f.heat_rate = (T *)o->heat_rate;
      f.theta_wall = (T *)o->theta_wall;
      f.phist = (T *)o->phist;
      f.phist_walls = o->phist_walls;
    }
  }
  hipLaunchKernelGGL(k_finalize<T>, dim3(1), dim3(kFinThreads), 0, st, dev, c, f);
  if (diag) {
    Diag3<T> dg;
    dg.uo = (const T *)p->old_u; dg.vo = (const T *)p->old_v; dg.wo = (const T *)p->old_w;
    dg.po = (const T *)p->old_p; dg.to = (const T *)p->old_t;
    dg.un = (const T *)out->u; dg.vn = (const T *)out->v; dg.wn = (const T *)out->w;
    dg.pn = (const T *)out->p; dg.tn = (const T *)out->theta;
    dg.uh = ub[0]; dg.vh = vb[0]; dg.wh = wb[0];
    dg.t_out = (const T *)out->t; dg.dt_out = (const T *)out->dt;
    dg.row_buf = (T *)o->diag;
    dg.count = (long long *)o->diag_count;
    dg.cap = o->diag_cap;
    hipLaunchKernelGGL(k_diag<T>, dim3(kDiagBlocks), dim3(256), 0, st, dev, dg);
    hipLaunchKernelGGL(k_diag_fin<T>, dim3(1), dim3(64), 0, st, dev, c, dg);
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int project_impl(aero3d_plan *p, hipStream_t st, void *u, void *v, void *w, void *psi_out) {
  const Dev3<T> *dev = (const Dev3<T> *)p->dev;
  Call3<T> c;
  std::memset(&c, 0, sizeof(c));
  c.uo = (T *)u;
  c.vo = (T *)v;
  c.wo = (T *)w;
  c.corr_stage = -1;
  int rc = poisson_passes<T>(p, st, dev, c, nullptr, 1, 0, (T *)p->psi, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_correct<T>, dim3(1024), dim3(256), 0, st, dev, (T *)u, (T *)v, (T *)w, (const T *)p->psi);
  if (psi_out) {
    const aero3d_desc &s = p->desc;
    AERO_HIP_CHECK(hipMemcpyAsync(psi_out, p->psi, (size_t)s.nx * s.ny * s.nz * sizeof(T), hipMemcpyDeviceToDevice, st));
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int poisson_impl(aero3d_plan *p, hipStream_t st, const void *f, void *psi) {
  const Dev3<T> *dev = (const Dev3<T> *)p->dev;
  Call3<T> c;
  std::memset(&c, 0, sizeof(c));
  c.corr_stage = -1;
  return poisson_passes<T>(p, st, dev, c, (const T *)f, 0, 0, (T *)psi, nullptr);
}

template <typename T>
int grad_impl(aero3d_plan *p, hipStream_t st, const void *u, const void *v, const void *w, void *out, double scale) {
  const Dev3<T> *dev = (const Dev3<T> *)p->dev;
  Call3<T> c;
  std::memset(&c, 0, sizeof(c));
  c.up = (const T *)u;
  c.vp = (const T *)v;
  c.wp = (const T *)w;
  c.corr_stage = -1;
  hipLaunchKernelGGL(k_grad<T>, dim3(p->nblk), dim3(kTileThreads), sizeof(Lds3<T>), st, dev, c, (T *)out, (T)scale);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

template <typename T>
int time_stage_impl(aero3d_plan *p, hipStream_t st, const aero3d_fields *fl, const aero3d_geom *g,
                    const aero3d_params *pr, int stage, int iters, double *avg_ms) {
  const Dev3<T> *dev = (const Dev3<T> *)p->dev;
  Call3<T> c = make_call<T>(fl, g, pr);
  c.stage = stage;
  c.psi = stage == 0 ? nullptr : (const T *)p->psi;
  c.corr_stage = stage == 0 ? -1 : stage - 1;
  c.uo = (T *)p->uB;
  c.vo = (T *)p->vB;
  c.wo = (T *)p->wB;
  if (p->desc.thermal && !fl->theta) {
    set_error("a heat-transfer plan needs fields.theta");
    return AERO_EINVAL;
  }
  hipEvent_t e0, e1;
  AERO_HIP_CHECK(hipEventCreate(&e0));
  AERO_HIP_CHECK(hipEventCreate(&e1));
  const size_t lds = sizeof(Lds3<T>);
  hipLaunchKernelGGL((k_stage<T, false>), dim3(p->nblk), dim3(kTileThreads), lds, st, dev, c);
  AERO_HIP_CHECK(hipEventRecord(e0, st));
  for (int it = 0; it < iters; ++it)
    hipLaunchKernelGGL((k_stage<T, false>), dim3(p->nblk), dim3(kTileThreads), lds, st, dev, c);
  AERO_HIP_CHECK(hipEventRecord(e1, st));
  AERO_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  AERO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = (double)ms / (double)iters;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

int check3_desc(const aero3d_desc *d) {
  if (!d) { set_error("null descriptor"); return AERO_EINVAL; }
  if (d->nx < 4 || d->ny < 4 || d->nz < 4) { set_error("the grid needs at least 4 cells per axis"); return AERO_EINVAL; }
  if (d->nx > 4096 || d->ny > 4096 || d->nz > 4096) { set_error("at most 4096 cells per axis"); return AERO_EUNSUPPORTED; }
  if (d->dtype != 0 && d->dtype != 1) { set_error("dtype must be 0 (float32) or 1 (float64)"); return AERO_EINVAL; }
  if (!(d->h > 0)) { set_error("h must be positive"); return AERO_EINVAL; }
  if (d->turbulence < 0 || d->turbulence > 3 || d->advection < 0 || d->advection > 1) {
    set_error("unknown advection or turbulence code");
    return AERO_EINVAL;
  }
  if (d->closed && d->periodic_x) { set_error("a closed domain cannot be periodic in x"); return AERO_EINVAL; }
  if (d->nut_model && d->turbulence == 0) {
    set_error("a learned closure needs a turbulence model to fall back to");
    return AERO_EINVAL;
  }
  for (int q = 0; q < 18; ++q)
    if (d->ghost_kind[q] < 0 || d->ghost_kind[q] > 6) { set_error("unknown ghost rule"); return AERO_EINVAL; }
  if (d->nbody < 0 || d->nbody > kMaxBodies) { set_error("at most 8 bodies"); return AERO_EUNSUPPORTED; }
  const bool win = d->win_k1 > d->win_k0 && d->win_j1 > d->win_j0 && d->win_i1 > d->win_i0;
  if (win) {
    if (d->win_i0 < 0 || d->win_j0 < 0 || d->win_k0 < 0 || d->win_i1 > d->nx || d->win_j1 > d->ny || d->win_k1 > d->nz) {
      set_error("body window outside the grid");
      return AERO_EINVAL;
    }
    if ((d->periodic_x && (d->win_i0 == 0 || d->win_i1 == d->nx)) ||
        (d->periodic_y && (d->win_j0 == 0 || d->win_j1 == d->ny)) ||
        (d->periodic_z && (d->win_k0 == 0 || d->win_k1 == d->nz))) {
      set_error("the body window reaches a periodic boundary");
      return AERO_EUNSUPPORTED;
    }
  }
  if (d->thermal) {
    const int per[3] = {d->periodic_x, d->periodic_y, d->periodic_z};
    for (int q = 0; q < 6; ++q) {
      const int k = d->theta_side[q];
      if (k < AERO3D_THETA_PERIODIC || k > AERO3D_THETA_HEAT_FLUX || ((k == AERO3D_THETA_PERIODIC) != (per[q / 2] != 0))) {
        set_error("theta side kinds must be periodic exactly on the periodic axes");
        return AERO_EINVAL;
      }
    }
    if (d->theta_nt < 0 || d->theta_nt > kMaxT || d->theta_nt > d->nbody || d->theta_nskin < 0) {
      set_error("bad heat-transfer body counts");
      return AERO_EINVAL;
    }
    for (int q = 0; q < d->theta_nt; ++q)
      if (d->theta_tbody[q] < 0 || d->theta_tbody[q] >= d->nbody) { set_error("bad held-temperature body index"); return AERO_EINVAL; }
    const bool qw = d->theta_q_k1 > d->theta_q_k0 && d->theta_q_j1 > d->theta_q_j0 && d->theta_q_i1 > d->theta_q_i0;
    if (qw && (d->theta_q_k0 < 0 || d->theta_q_j0 < 0 || d->theta_q_i0 < 0 || d->theta_q_k1 > d->nz ||
               d->theta_q_j1 > d->ny || d->theta_q_i1 > d->nx)) {
      set_error("heat source window outside the grid");
      return AERO_EINVAL;
    }
    if (!(d->prandtl_turbulent > 0) || !std::isfinite(d->richardson)) {
      set_error("the turbulent Prandtl number must be positive and Ri finite");
      return AERO_EINVAL;
    }
  }
  if (d->pid_dt && !(d->pid_target >= 0.1 && d->pid_target <= 1.0)) {
    set_error("pid_target must be in [0.1, 1]");
    return AERO_EINVAL;
  }
  return AERO_OK;
}

int check3_step(const aero3d_plan *p, const aero3d_fields *in, const aero3d_fields *out, const aero3d_geom *g,
                const aero3d_params *pr, const aero3d_outputs *o) {
  if (!p || !in || !out || !pr) { set_error("null plan or argument"); return AERO_EINVAL; }
  if (!in->u || !in->v || !in->w || !in->t || !out->u || !out->v || !out->w || !out->p || !out->t || !out->dt) {
    set_error("missing field pointer");
    return AERO_EINVAL;
  }
  if (p->desc.dye && p->desc.dye_k1 > p->desc.dye_k0 && (!pr->dye_weight || !pr->dye_pulse)) {
    set_error("smoke nozzles need params.dye_weight and dye_pulse");
    return AERO_EINVAL;
  }
  if (p->desc.pid_dt && (!in->ctrl || !out->ctrl)) {
    set_error("the PID step needs the controller memory (fields ctrl, 3 values)");
    return AERO_EINVAL;
  }
  if (p->desc.nut_model && !pr->nut_field) {
    set_error("a learned-closure plan needs params.nut_field");
    return AERO_EINVAL;
  }
  if (!pr->nu || !pr->eta || !pr->u_inlet || (p->desc.fixed_dt && !pr->dt_fixed)) {
    set_error("missing parameter pointer");
    return AERO_EINVAL;
  }
  if (o && o->diag && (!p->desc.diagnostics || !o->diag_count || o->diag_cap < 1 || !in->p)) {
    set_error("diagnostics rows need a plan made with desc.diagnostics, a row counter and fields.p");
    return AERO_EINVAL;
  }
  const aero3d_desc &s = p->desc;
  const bool win = s.win_k1 > s.win_k0 && s.win_j1 > s.win_j0 && s.win_i1 > s.win_i0 && s.nbody > 0;
  if (win && (!g || !g->chi_u || !g->chi_v || !g->chi_w || !g->chi_c || !g->a_u || !g->a_v || !g->a_w || !g->k_u ||
              !g->k_v || !g->k_w)) {
    set_error("missing geometry pointer");
    return AERO_EINVAL;
  }
  if (s.thermal) {
    if (!in->theta || !out->theta || !pr->theta_scalars || !pr->theta_side_values) {
      set_error("a heat-transfer plan needs fields.theta and params.theta_scalars / theta_side_values");
      return AERO_EINVAL;
    }
    const bool qw = s.theta_q_k1 > s.theta_q_k0 && s.theta_q_j1 > s.theta_q_j0 && s.theta_q_i1 > s.theta_q_i0;
    if (qw && (!g || !g->theta_q)) { set_error("missing geom.theta_q"); return AERO_EINVAL; }
    if (win && (!g->body_theta || !g->flux_rate || (s.theta_nt && (!g->chi_t || !g->a_t || !g->w_t)) ||
                (s.theta_adia && (!g->adia_u || !g->adia_v || !g->adia_w)) || (s.richardson != 0.0 && !g->vol_w) ||
                (s.theta_nskin && (!g->skin_idx || !g->skin_w || !g->skin_off)))) {
      set_error("missing heat-transfer geometry pointer");
      return AERO_EINVAL;
    }
    if (in->theta == out->theta && in->dye == out->theta) { set_error("theta aliases dye"); return AERO_EINVAL; }
  }
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero3d_plan_create(const aero3d_desc *desc, aero3d_plan **out) {
  if (!out) { set_error("null output pointer"); return AERO_EINVAL; }
  *out = nullptr;
  int rc = check3_desc(desc);
  if (rc) return rc;
  aero3d_plan *p = new aero3d_plan();
  p->desc = *desc;
  p->dbl = desc->dtype == 1;
  p->esize = p->dbl ? 8 : 4;
  if (!factor_radices(desc->nx, p->fp[0]) || !factor_radices(desc->ny, p->fp[1]) || !factor_radices(desc->nz, p->fp[2])) {
    delete p;
    set_error("nx, ny and nz must be products of 2, 3 and 5");
    return AERO_EUNSUPPORTED;
  }
  const size_t es = p->esize;
  const size_t nx = desc->nx, ny = desc->ny, nz = desc->nz;
  int tx, ty, tz;
  if (p->dbl) tile_dims<double>(tx, ty, tz);
  else tile_dims<float>(tx, ty, tz);
  p->nbx = (int)((nx + tx - 1) / tx);
  p->nby = (int)((ny + ty - 1) / ty);
  p->nbz = (int)((nz + tz - 1) / tz);
  p->nblk = p->nbx * p->nby * p->nbz;
  const bool win = desc->win_k1 > desc->win_k0 && desc->win_j1 > desc->win_j0 && desc->win_i1 > desc->win_i0 &&
                   desc->nbody > 0;
  const size_t wz = win ? desc->win_k1 - desc->win_k0 : 0, wy = win ? desc->win_j1 - desc->win_j0 : 0,
               wx = win ? desc->win_i1 - desc->win_i0 : 0;
  p->nwt = 0;
  p->nq = win ? 9 * desc->nbody : 0;
  if (win) {
    const int nwbx = std::min(desc->win_i1 / tx, p->nbx - 1) - desc->win_i0 / tx + 1;
    const int nwby = std::min(desc->win_j1 / ty, p->nby - 1) - desc->win_j0 / ty + 1;
    const int nwbz = std::min(desc->win_k1 / tz, p->nbz - 1) - desc->win_k0 / tz + 1;
    p->nwt = nwbx * nwby * nwbz;
  }
  // the fused plane pass: FP32, even ny and nz, two work buffers and the twiddles in one
  // block's LDS
  {
    const size_t c = 2 * es;
    p->plane_buf = std::max((nz / 2) * (ny + 1), (ny / 2) * (nz + 1)) * c;
    p->lds_planes = 2 * p->plane_buf + 2 * (ny + nz) * c;
    int dev_id = 0, max_lds = 0;
    if (hipGetDevice(&dev_id) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id) != hipSuccess)
      max_lds = 64 * 1024;
    p->planes = !p->dbl && ny % 2 == 0 && nz % 2 == 0 && ny * nz <= (size_t)kPlaneMax &&
                p->lds_planes <= (size_t)max_lds;
  }
  // the forward x pass takes 8 row pairs per 256-thread block with the plane pass (64-byte runs of
  // each wavenumber in the (x, z, y) spectrum: 29.4 -> 26.8 us at 256x128x128), 4 otherwise; the
  // inverse pass 8 per 512-thread block (31.8 -> 27.4 us; on 256 threads 8 pairs lose, 40.8 us)
  p->lds1 = pick_slots((int)nx, es, 0, p->planes ? 8 : 4, p->ns1);
  p->lds2 = pick_slots((int)ny, es, 2, 8, p->ns2);
  p->lds3 = pick_slots((int)nz, es, 2, 8, p->ns3);
  p->lds5 = pick_slots((int)nx, es, 0, 8, p->ns5);   // its real rows share a transform buffer
  const size_t nu_ = nz * ny * (nx + 1), nv_ = nz * (ny + 1) * nx, nw_ = (nz + 1) * ny * nx, nc = nz * ny * nx;
  struct { void **ptr; size_t n; } tab[] = {
      {&p->dev, sizeof(Dev3<double>)},
      {&p->uA, nu_ * es}, {&p->vA, nv_ * es}, {&p->wA, nw_ * es}, {&p->uB, nu_ * es}, {&p->vB, nv_ * es},
      {&p->wB, nw_ * es}, {&p->F, nc * es}, {&p->psi, nc * es},
      {&p->acc_u, wz * wy * (wx + 1) * es}, {&p->acc_v, wz * (wy + 1) * wx * es}, {&p->acc_w, (wz + 1) * wy * wx * es},
      {&p->pmax, (size_t)p->nblk * kPm * es}, {&p->pout, (size_t)p->nby * p->nbz * 2 * 8},
      {&p->pforce, (size_t)std::max(1, p->nwt) * std::max(1, p->nq) * 8}, {&p->red, 256},
      {&p->dye_tmp, desc->dye ? nc * es : 0},
      {&p->th_a, desc->thermal ? nc * es : 0}, {&p->th_b, desc->thermal ? nc * es : 0},
      {&p->nut_c, desc->thermal && desc->turbulence ? nc * es : 0},
      {&p->nu_pad, desc->turbulence && !desc->nut_model ? (size_t)(desc->nx + 2) * (desc->ny + 2) * (desc->nz + 2) * es : 0},
      {&p->pheat, desc->thermal ? (size_t)3 * std::max(1, p->nwt) * kMaxT * 8 : 0},
      {&p->old_u, desc->diagnostics ? nu_ * es : 0}, {&p->old_v, desc->diagnostics ? nv_ * es : 0},
      {&p->old_w, desc->diagnostics ? nw_ * es : 0}, {&p->old_p, desc->diagnostics ? nc * es : 0},
      {&p->old_t, desc->diagnostics && desc->thermal ? nc * es : 0},
      {&p->pdiag, desc->diagnostics ? (size_t)kDiagBlocks * kDV * 8 : 0},
      {&p->tw[0], nx * 2 * es}, {&p->tw[1], ny * 2 * es}, {&p->tw[2], nz * 2 * es},
      {&p->cs[0], nx * 2 * es}, {&p->cs[1], ny * 2 * es}, {&p->cs[2], nz * 2 * es},
      {(void **)&p->lam[0], nx * 8}, {(void **)&p->lam[1], ny * 8}, {(void **)&p->lam[2], nz * 8},
      {&p->inv_lam, nc * es}};
  for (auto &t : tab) {
    if ((rc = alloc3(p, t.ptr, t.n))) {
      aero3d_plan_destroy(p);
      return rc;
    }
  }
  rc = p->dbl ? upload3<double>(p) : upload3<float>(p);
  if (rc) {
    aero3d_plan_destroy(p);
    return rc;
  }
  *out = p;
  return AERO_OK;
}

int aero3d_plan_destroy(aero3d_plan *p) {
  if (!p) return AERO_OK;
  for (void *a : p->allocs) (void)hipFree(a);
  delete p;
  return AERO_OK;
}

size_t aero3d_plan_workspace_bytes(const aero3d_plan *p) { return p ? p->bytes : 0; }

int aero3d_step(aero3d_plan *p, void *stream, const aero3d_fields *in, aero3d_fields *out, const aero3d_geom *g,
                const aero3d_params *pr, aero3d_outputs *o) {
  int rc = check3_step(p, in, out, g, pr, o);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o, false) : step_impl<float>(p, st, in, out, g, pr, o, false);
}

int aero3d_step_trace(aero3d_plan *p, void *stream, const aero3d_fields *in, aero3d_fields *out,
                      const aero3d_geom *g, const aero3d_params *pr, aero3d_outputs *o, const aero3d_trace *tr) {
  if (!tr) { set_error("aero3d_step_trace: null trace"); return AERO_EINVAL; }
  int rc = check3_step(p, in, out, g, pr, o);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o, false, tr)
                : step_impl<float>(p, st, in, out, g, pr, o, false, tr);
}

int aero3d_step_chained(aero3d_plan *p, void *stream, const aero3d_fields *in, aero3d_fields *out,
                        const aero3d_geom *g, const aero3d_params *pr, aero3d_outputs *o) {
  int rc = check3_step(p, in, out, g, pr, o);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? step_impl<double>(p, st, in, out, g, pr, o, true) : step_impl<float>(p, st, in, out, g, pr, o, true);
}

int aero3d_project(aero3d_plan *p, void *stream, void *u, void *v, void *w, void *psi) {
  if (!p || !u || !v || !w) { set_error("null plan or field"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? project_impl<double>(p, st, u, v, w, psi) : project_impl<float>(p, st, u, v, w, psi);
}

int aero3d_poisson_solve(aero3d_plan *p, void *stream, const void *f, void *psi) {
  if (!p || !f || !psi) { set_error("null plan or field"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? poisson_impl<double>(p, st, f, psi) : poisson_impl<float>(p, st, f, psi);
}

int aero3d_cell_gradient(aero3d_plan *p, void *stream, const void *u, const void *v, const void *w, void *out,
                         double scale) {
  if (!p || !u || !v || !w || !out) { set_error("null plan or field"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? grad_impl<double>(p, st, u, v, w, out, scale) : grad_impl<float>(p, st, u, v, w, out, scale);
}

int aero3d_time_stage(aero3d_plan *p, void *stream, const aero3d_fields *fields, const aero3d_geom *g,
                      const aero3d_params *pr, int stage, int iters, double *avg_ms) {
  if (!p || !fields || !pr || !avg_ms || stage < 0 || stage > 2 || iters < 1) {
    set_error("bad aero3d_time_stage arguments");
    return AERO_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  return p->dbl ? time_stage_impl<double>(p, st, fields, g, pr, stage, iters, avg_ms)
                : time_stage_impl<float>(p, st, fields, g, pr, stage, iters, avg_ms);
}

}  // extern "C"
