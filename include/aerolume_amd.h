/*
 * aerolume_amd - C ABI of the MI355X Navier-Stokes step (libaerolume_amd.so).
 *
 * Plain C types only: device pointers are `void *`, the HIP stream is passed as
 * `void *` (a hipStream_t; NULL = the null stream). The caller owns every field
 * buffer; a plan owns only its scratch (stage buffers, spectral work arrays,
 * twiddle and eigenvalue tables, reduction slots), allocated once at creation.
 * Nothing inside a step allocates, frees or synchronises with the host, so a
 * step can be captured in a hipGraph. Return codes: 0 = OK, negative = error;
 * aero_last_error() gives the message of the calling thread's last failure.
 *
 * Reference interfaces these entry points replace (TanmayPasula/Aerolume,
 * src/aerolume/solvers/...):
 *   aero2d_plan_create   - ns2d/core.py:179-202 make_operators (static operators,
 *                          PoissonSolver construction ns2d/poisson.py:107-126)
 *   aero2d_step          - ns2d/core.py:815-969 step_full (one SSP-RK3 step with
 *                          projection, penalisation, forces, temperature, smoke),
 *                          called as ns2d/solver.py:384-400 _step_static does
 *   aero2d_project       - ns2d/core.py:626-642 project
 *   aero2d_poisson_solve - ns2d/poisson.py:128-134 PoissonSolver.solve
 *   aero2d_ensemble_*    - no counterpart (batched sweep; the reference loops over solves;
 *                          aero2d_ensemble_tune picks its launch shapes)
 *   aero3d_plan_create   - ns3d/core.py:184-203 make_operators3d (static operators,
 *                          ghost-rule table core.py:147-181, PoissonSolver3D
 *                          ns3d/poisson.py:107-126)
 *   aero3d_step          - ns3d/core.py:753-762 step3d_full (one SSP-RK3 step with
 *                          projection, penalisation and the six coefficients), called
 *                          as ns3d/solver.py:583-595 _step_fn dispatches the fused path,
 *                          then the smoke of ns3d/solver.py:646-680 (ns3d/dye.py:281-312)
 *   aero3d_project       - ns3d/core.py:604-612 project3d
 *   aero3d_poisson_solve - ns3d/poisson.py:128-157 PoissonSolver3D.solve
 */
#ifndef AEROLUME_AMD_H
#define AEROLUME_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AERO_OK 0
#define AERO_EINVAL (-1)        /* bad descriptor or argument */
#define AERO_EUNSUPPORTED (-2)  /* feature combination the step does not carry */
#define AERO_EHIP (-3)          /* a HIP runtime call failed */

/* boundary kinds */
#define AERO_BC_SLIP 0
#define AERO_BC_NOSLIP 1
#define AERO_BC_MOVING 2
/* temperature boundary modes (left, right, bottom, top) */
#define AERO_TH_ADIABATIC 0
#define AERO_TH_TEMPERATURE 1
#define AERO_TH_HEAT_FLUX 2
#define AERO_TH_INLET 3
#define AERO_TH_OUTLET 4
#define AERO_TH_PERIODIC 5

#define AERO2D_MAX_GROUPS 8

typedef struct aero2d_plan aero2d_plan;

/* Static description of a 2D solve (ns2d/config.py:369-516 after set-up). */
typedef struct {
  int32_t nx, ny;          /* cells */
  int32_t dtype;           /* 0 = float32, 1 = float64 */
  int32_t periodic;        /* doubly periodic box */
  int32_t closed;          /* walls on all four sides */
  int32_t advection;       /* 0 = central, 1 = upwind3 */
  int32_t turbulence;      /* 0 none, 1 smagorinsky, 2 wale, 3 vreman */
  int32_t bc_bottom, bc_top, bc_left, bc_right;   /* AERO_BC_* */
  int32_t fixed_dt;        /* 1: dt from params.dt_fixed, else CFL rule */
  int32_t win_j0, win_j1, win_i0, win_i1;         /* body cell window (empty: j1 <= j0) */
  int32_t nbody;
  int32_t thermal;         /* temperature carried */
  int32_t buoyant;         /* Ri != 0 */
  int32_t wall_modes[4];   /* AERO_TH_* for left, right, bottom, top */
  int32_t n_temp;          /* temperature bodies */
  int32_t has_heat;        /* volumetric heater field present */
  int32_t insulated;       /* thermal masks on the window (diffusivity x ins) */
  int32_t dye_mode;        /* 0 none, 1 inlet comb (single channel), 2 channels + emitters */
  int32_t n_dye;           /* smoke channels */
  int32_t n_groups;        /* emitter groups */
  int32_t group_channel[AERO2D_MAX_GROUPS];
  double h;
  double ground_speed;     /* moving bottom wall speed */
  double cfl;
  double force_ref;
  double les_c;            /* model constant */
  double gravity_x, gravity_y;
  double prt;              /* turbulent Prandtl number */
  double diff_coef;        /* largest temperature diffusion coefficient / (kappa / h^2) */
  int32_t tile_rows;       /* rows per 64-wide tile of the cell kernels: 0 (default, 4), 4, 8 or 16 */
  int32_t pid_dt;          /* adaptive_dt = "pid" (ns2d/diagnostics.py:104-123): needs fields ctrl, dt */
  int32_t diagnostics;     /* 1: the plan keeps scratch for outputs.diag rows */
  int32_t wall_model;      /* 1: Werner-Wengle shear on the faces next to no-slip domain walls
                              (ns2d/core.py:509-568): u faces of no-slip / moving bottom and
                              no-slip top walls, v faces of no-slip side walls when closed */
  int32_t n_flux;          /* heat-flux bodies: skin sources kappa q_skin on the window and their
                              mean skin temperatures (ns2d/physics.py:726-727, core.py:988-990) */
  int32_t reserved;
} aero2d_desc;

/* Evolving fields, caller-owned device buffers of the plan's dtype.
 * u (ny, nx+1), v (ny+1, nx), p (ny, nx), theta (ny, nx), dye (n_dye, ny, nx),
 * t and dt 0-d. in == out (in place) is allowed for every field. */
typedef struct {
  void *u, *v, *p, *theta, *dye, *t, *dt;
  void *ctrl;   /* PID step's controller memory (2 values), read from in, written to out */
} aero2d_fields;

/* Window masks (ns2d/geometry.py:160-197, 120-157); wy = j1-j0, wx = i1-i0.
 * chi_u, a_u (wy, wx+1); chi_v, a_v (wy+1, wx); chi_c (wy, wx);
 * k_u (nbody, 2, wy, wx+1); k_v (nbody, 2, wy+1, wx);
 * chi_t, a_t (wy, wx); w_t (n_temp, wy, wx); vol_u (wy, wx+1); vol_v (wy+1, wx);
 * ins_u (wy, wx+1); ins_v (wy+1, wx); q_skin (wy, wx): the heat-flux bodies' skin source per
 * unit kappa (ns2d/geometry.py:403-411); w_f (n_flux, wy, wx): their normalised skin weights. */
typedef struct {
  const void *chi_u, *chi_v, *chi_c, *a_u, *a_v, *k_u, *k_v;
  const void *chi_t, *a_t, *w_t, *vol_u, *vol_v, *ins_u, *ins_v;
  const void *q_skin, *w_f;
} aero2d_geom;

/* Live values as device scalars / arrays (edited in place between steps, so a
 * captured graph stays valid): nu, eta, dt_fixed, kappa, ri, eta_t, theta_in,
 * heat_max are 0-d; u_inlet (ny); wall_values (4); heat (ny, nx);
 * dye_inflow (n_dye, ny); group_weight, group_target (n_groups, ny, nx);
 * group_timing (n_groups, 4) = [start, stop, period, duty]. */
typedef struct {
  const void *nu, *eta, *u_inlet, *dt_fixed;
  const void *kappa, *ri, *eta_t, *theta_in, *wall_values, *heat, *heat_max;
  const void *dye_inflow, *group_weight, *group_target, *group_timing;
} aero2d_params;

/* Per-step outputs (device; any may be NULL): coeffs (nbody, 3) = [Cd, Cl, Cm];
 * heat (n_temp); wall_heat (4); heat_storage (1); flow_heat (2); theta_wall (n_flux).
 * hist: ring of hist_cap rows [t, Cd, Cl, Cm] (summed over bodies) written at
 * row (*hist_count mod hist_cap), then *hist_count += 1 (int64 device counter).
 * aux: ring of aux_cap rows [t, heat.., theta_wall.., wall_heat(4), heat_storage,
 * flow_heat(2)] with counter aux_count. */
typedef struct {
  void *coeffs, *heat, *wall_heat, *heat_storage, *flow_heat;
  void *hist, *hist_count;
  int64_t hist_cap;
  void *aux, *aux_count;
  int64_t aux_cap;
  /* diagnostics (plans made with desc.diagnostics): when diag is set the step writes one
   * float64 row of AERO2D_DIAG_COLS values at (*diag_count mod diag_cap), then
   * *diag_count += 1: t, dt, CFL, div_rms, div_max, Poisson residual of the last stage, then
   * per field u, v, p, theta the change over the step (L2, RMS, max, relative L2; NaN
   * without temperature) (ns2d/diagnostics.py:70-101; the "step" column is the caller's) */
  void *diag, *diag_count;
  int64_t diag_cap;
  void *theta_wall;
} aero2d_outputs;

#define AERO2D_DIAG_COLS 22

int aero2d_plan_create(const aero2d_desc *desc, aero2d_plan **plan);
int aero2d_plan_destroy(aero2d_plan *plan);
size_t aero2d_plan_workspace_bytes(const aero2d_plan *plan);
int aero2d_step(aero2d_plan *plan, void *stream, const aero2d_fields *in, aero2d_fields *out,
                const aero2d_geom *geom, const aero2d_params *params, aero2d_outputs *outputs);
/* Same step as aero2d_step, without re-reading the state for its dt inputs: it reuses the
 * maxima (|u|, |v|, |theta|, nu_t) and sum theta that the previous aero2d_step /
 * aero2d_step_chained call of this plan computed from its output state. Valid only when
 * `in` is that output, unmodified (the stateful solver's hot loop); saves two launches. */
int aero2d_step_chained(aero2d_plan *plan, void *stream, const aero2d_fields *in, aero2d_fields *out,
                        const aero2d_geom *geom, const aero2d_params *params, aero2d_outputs *outputs);
/* Stage trace of one step (core.py:897-934 with a trace dict): per RK stage s (0..2) caller
 * buffers of the plan's dtype, any NULL: pen_u (ny, nx+1) / pen_v (ny+1, nx) the stage velocity
 * after penalisation and buoyancy (outlet face corrected), div (ny, nx) its projection's
 * right-hand side, psi (ny, nx) the potential, proj_u / proj_v the projected velocity, theta
 * (ny, nx) the stage temperature. The velocity before penalisation is not formed by the fused
 * stage kernel and is not traced. */
typedef struct {
  void *pen_u[3], *pen_v[3], *div[3], *psi[3], *proj_u[3], *proj_v[3], *theta[3];
} aero2d_trace;
/* aero2d_step that also fills `trace` (a debugging step: extra copies after every stage). */
int aero2d_step_trace(aero2d_plan *plan, void *stream, const aero2d_fields *in, aero2d_fields *out,
                      const aero2d_geom *geom, const aero2d_params *params, aero2d_outputs *outputs,
                      const aero2d_trace *trace);
/* u, v in place: u - grad psi with lap psi = div u (no inlet/outlet change); psi (ny, nx) optional. */
int aero2d_project(aero2d_plan *plan, void *stream, void *u, void *v, void *psi);
/* psi = L^-1 f with zero mean (f, psi (ny, nx); may alias). */
int aero2d_poisson_solve(aero2d_plan *plan, void *stream, const void *f, void *psi);
/* No counterpart in the reference: aero2d_poisson_solve without a plan, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops). The
 * grid is described by the arguments (dtype 0 float32 / 1 float64, nx x ny cells of size h,
 * periodic 0: Neumann DCT solve, 1: periodic); f and psi (ny, nx) are the caller's and may
 * alias. Nothing is allocated: the spectral work array and the transform tables live in a
 * static device array of 17.1 MB (16.3 MiB) that each GPU holds until the process ends, rebuilt by
 * every call. Limits: nx * ny <= 1048576 cells, nx, ny <= 4096 and 2, 3, 5-smooth (larger
 * grids return AERO_EUNSUPPORTED). Calls that overlap on one GPU share the array and race. */
int aero2d_poisson_solve_static(void *stream, int32_t dtype, int32_t nx, int32_t ny, double h, int32_t periodic,
                                const void *f, void *psi);

/* Timing helper (no reference counterpart; used by bench.py to price the dominant kernel
 * against the HBM roofline): launches the fused stage kernel of RK stage `stage` (0..2)
 * `iters` times on `stream`, reading `fields` as both stage input and step start and
 * writing only plan scratch, between two HIP events; returns the average launch time in ms.
 * Synchronises the stream. Call after at least one aero2d_step (it reuses that step's dt). */
int aero2d_time_stage(aero2d_plan *plan, void *stream, const aero2d_fields *fields, const aero2d_geom *geom,
                      const aero2d_params *params, int stage, int iters, double *avg_ms);

/* ------------------------------------------------------------------ 3D */
/* ghost-rule kinds (ns3d/core.py:147-181, 223-272) */
#define AERO3D_GHOST_EDGE 0          /* repeat the boundary value (inlet, outlet) */
#define AERO3D_GHOST_EVEN 1          /* mirror across a wall half a cell away (zero gradient) */
#define AERO3D_GHOST_ODD 2           /* mirror with sign change (zero value) */
#define AERO3D_GHOST_AFFINE 3        /* mirror to a prescribed value: 2 value - x */
#define AERO3D_GHOST_FACE_ODD 4      /* wall-normal component across a wall on the face */
#define AERO3D_GHOST_PERIODIC_CELL 5
#define AERO3D_GHOST_PERIODIC_FACE 6 /* face arrays carry a duplicate last face */
#define AERO3D_MAX_BODIES 8

typedef struct aero3d_plan aero3d_plan;

/* Static description of a 3D solve (ns3d/config.py:528-731 after set-up). */
typedef struct {
  int32_t nx, ny, nz;       /* cells; each a product of 2, 3 and 5, at most 4096 */
  int32_t dtype;            /* 0 = float32, 1 = float64 */
  int32_t periodic_x, periodic_y, periodic_z;
  int32_t closed;           /* walls at x0 and x1 (no inflow) */
  int32_t advection;        /* 0 = central, 1 = upwind3 */
  int32_t turbulence;       /* 0 none, 1 smagorinsky, 2 wale, 3 vreman */
  int32_t fixed_dt;         /* 1: dt from params.dt_fixed, else the CFL / viscous rule */
  int32_t win_k0, win_k1, win_j0, win_j1, win_i0, win_i1;   /* body cell window (empty: k1 <= k0) */
  int32_t nbody;
  int32_t symmetry;         /* y = y0 symmetry plane: full-body coefficients */
  int32_t ghost_kind[18];   /* AERO3D_GHOST_* per component (u, v, w), axis (x, y, z), side (lo, hi) */
  double ghost_value[18];   /* value of the AFFINE rule (moving floor speed) */
  double h, y0;
  double cfl, les_c;
  double ref_area, length_ref, span_ref;
  double moment_ref[3];
  int32_t nut_model;        /* 1: eddy viscosity from params.nut_field where it lies in
                               [0, nut_trust], the turbulence model elsewhere (learned closure) */
  int32_t reserved;
  double nut_trust;
  int32_t dye;              /* 1: passive smoke in fields.dye, advected after every step */
  int32_t dye_k0, dye_k1, dye_j0, dye_j1, dye_i0, dye_i1;   /* nozzle cell window (empty: none) */
  int32_t reserved2;
  /* heat transfer (ns3d/physics.py:494-830): theta in fields.theta advanced in the RK stages */
  int32_t thermal;
  int32_t theta_side[6];    /* AERO3D_THETA_* of the sides x0, x1, y0, y1, z0, z1 */
  int32_t theta_q_k0, theta_q_k1, theta_q_j0, theta_q_j1, theta_q_i0, theta_q_i1;   /* source window */
  int32_t theta_nt;         /* held-temperature bodies (0..8), body indices in theta_tbody */
  int32_t theta_tbody[8];
  int32_t theta_adia;       /* 1: adiabatic-body face fractions in geom.adia_* */
  int32_t theta_nskin;      /* entries of the heat-flux skin list (geom.skin_*) */
  double richardson;        /* buoyancy Ri theta along +z (0: passive temperature) */
  double prandtl_turbulent;
  int32_t diagnostics;      /* 1: the plan keeps scratch for outputs.diag rows */
  int32_t pid_dt;           /* adaptive_dt = "pid" (ns3d/physics.py:1144-1177): needs fields ctrl */
  double pid_target;        /* the controller's target as a fraction of the rule's dt, in [0.1, 1] */
} aero3d_desc;

/* heat-equation side conditions (ns3d/physics.py:589-606) */
enum {
  AERO3D_THETA_PERIODIC = 0, AERO3D_THETA_OUTFLOW = 1, AERO3D_THETA_ADIABATIC = 2, AERO3D_THETA_INFLOW = 3,
  AERO3D_THETA_TEMPERATURE = 4, AERO3D_THETA_HEAT_FLUX = 5
};

/* Caller-owned device buffers of the plan's dtype: u (nz, ny, nx+1), v (nz, ny+1, nx),
 * w (nz+1, ny, nx), p (nz, ny, nx), t and dt 0-d. in == out is allowed. */
typedef struct {
  void *u, *v, *w, *p, *t, *dt;
  void *dye;                /* (nz, ny, nx) smoke concentration in [0, 1] when desc.dye */
  void *theta;              /* (nz, ny, nx) temperature when desc.thermal */
  void *ctrl;               /* PID step: controller memory [dt_prev, e1, e2], read from in, written to out */
} aero3d_fields;

/* Window arrays (ns3d/geometry.py:275-303), wz = k1-k0, wy = j1-j0, wx = i1-i0:
 * chi_u, a_u (wz, wy, wx+1); chi_v, a_v (wz, wy+1, wx); chi_w, a_w (wz+1, wy, wx);
 * chi_c (wz, wy, wx); k_u (nbody, 3, wz, wy, wx+1) = [Fx, My, Mz] weights;
 * k_v (nbody, 3, ...) = [Fy, Mx, Mz]; k_w (nbody, 3, ...) = [Fz, Mx, My]. */
typedef struct {
  const void *chi_u, *chi_v, *chi_w, *chi_c, *a_u, *a_v, *a_w, *k_u, *k_v, *k_w;
  /* heat transfer on the same window: chi_t, a_t (wz, wy, wx) held-temperature fraction and
   * sum chi theta_b; w_t (theta_nt, wz, wy, wx) each such body's share; adia_u/v/w face
   * fractions of adiabatic bodies (shapes of chi_u/v/w); vol_w (wz+1, wy, wx) unshifted volume
   * fraction; theta_q source density on the desc.theta_q window; skin list: skin_idx (int32
   * flat cell index), skin_w (weight), skin_off (int32, nbody + 1 offsets of each body's
   * entries); body_theta (nbody) held values; flux_rate (nbody) heat a heat-flux body
   * releases in the domain */
  const void *chi_t, *a_t, *w_t, *adia_u, *adia_v, *adia_w, *vol_w, *theta_q;
  const void *skin_idx, *skin_w, *skin_off, *body_theta, *flux_rate;
} aero3d_geom;

/* Live device scalars / arrays: nu, eta, dt_fixed 0-d; u_inlet (nz, ny); nut_field (nz, ny, nx)
 * cell-centred eddy viscosity of a learned closure (read when desc.nut_model is set). */
typedef struct {
  const void *nu, *eta, *u_inlet, *dt_fixed;
  const void *nut_field;
  const void *dye_weight;   /* nozzle weights on the nozzle window (wz, wy, wx) */
  const void *dye_pulse;    /* [t_on, period]: nozzles emit while t mod period < t_on */
  /* heat transfer, live: theta_scalars = [kappa, kappa / h, eta_t, kappa max|q|, heat-rate
   * scale (symmetry factor h^3 Pe), kappa / (h^2 / 6)]; theta_side_values (6) */
  const void *theta_scalars, *theta_side_values;
} aero3d_params;

/* Per-step outputs (device; any may be NULL): coeffs (nbody, 6) = [CD, CS, CL, C_roll, Cm, Cn];
 * hist: ring of hist_cap rows [t, CD, CS, CL, C_roll, Cm, Cn] (summed over bodies) written at
 * row (*hist_count mod hist_cap), then *hist_count += 1 (int64 device counter). */
typedef struct {
  void *coeffs;
  void *hist, *hist_count;
  int64_t hist_cap;
  /* heat transfer: heat_rate (nbody) conduction units, full body; theta_wall (nbody) held or
   * skin-measured wall temperature; phist ring of hist_cap rows [t, heat_rate (nbody),
   * theta_wall (nbody) when phist_walls] at the history row */
  void *heat_rate, *theta_wall, *phist;
  int32_t phist_walls, reserved;
  /* diagnostics (plans made with desc.diagnostics): when diag is set the step writes one row
   * of AERO3D_DIAG_COLS values at (*diag_count mod diag_cap), then *diag_count += 1:
   * t, dt, CFL, umax, div_rms, div_max, Poisson residual, then per field u, v, w, p, theta
   * the change over the step (L2, RMS, max, relative L2), then the heat storage rate
   * (NaN without heat transfer) (ns3d/solver.py:747-783) */
  void *diag, *diag_count;
  int64_t diag_cap;
} aero3d_outputs;

#define AERO3D_DIAG_COLS 28

int aero3d_plan_create(const aero3d_desc *desc, aero3d_plan **plan);
int aero3d_plan_destroy(aero3d_plan *plan);
size_t aero3d_plan_workspace_bytes(const aero3d_plan *plan);
int aero3d_step(aero3d_plan *plan, void *stream, const aero3d_fields *in, aero3d_fields *out,
                const aero3d_geom *geom, const aero3d_params *params, aero3d_outputs *outputs);
/* Same step, reusing the dt inputs (max |u|, |v|, |w|, nu_t) that the previous step of this
 * plan computed from its output state; valid only when `in` is that output, unmodified. */
int aero3d_step_chained(aero3d_plan *plan, void *stream, const aero3d_fields *in, aero3d_fields *out,
                        const aero3d_geom *geom, const aero3d_params *params, aero3d_outputs *outputs);
/* Stage fields of one traced step (reference FlowSolver3D.trace / trace_step,
 * ns3d/solver.py:832-919, recorded as core.py:822-862 records them), per RK stage, staggered
 * like the state; null pointers are skipped. tend: the explicit tendency (boundary rules
 * applied); star: the stage value before penalisation (inlet held, outlet balanced);
 * pen: after the penalisation and the buoyancy impulse; div: the projection's right-hand side
 * (nz, ny, nx); psi: the potential; proj: the projected velocity; theta: the stage's temperature. */
typedef struct {
  void *tend_u[3], *tend_v[3], *tend_w[3];
  void *star_u[3], *star_v[3], *star_w[3];
  void *pen_u[3], *pen_v[3], *pen_w[3];
  void *div[3], *psi[3];
  void *proj_u[3], *proj_v[3], *proj_w[3];
  void *theta[3];
} aero3d_trace;
/* aero3d_step that also fills `trace` (a debugging step: unchained, extra stores and copies
 * per stage; its state, outputs and history row are those of aero3d_step). */
int aero3d_step_trace(aero3d_plan *plan, void *stream, const aero3d_fields *in, aero3d_fields *out,
                      const aero3d_geom *geom, const aero3d_params *params, aero3d_outputs *outputs,
                      const aero3d_trace *trace);
/* u, v, w in place: u - grad psi with lap psi = div u; psi (nz, ny, nx) optional. */
int aero3d_project(aero3d_plan *plan, void *stream, void *u, void *v, void *w, void *psi);
/* psi = L^-1 f with zero mean (f, psi (nz, ny, nx); may alias). */
int aero3d_poisson_solve(aero3d_plan *plan, void *stream, const void *f, void *psi);
/* Cell-centred velocity gradient g[i][j] = du_i/dx_j (9, nz, ny, nx), times `scale`, with the
 * plan's ghost rules (ns3d/core.py:400-434): the learned closure's input features. */
int aero3d_cell_gradient(aero3d_plan *plan, void *stream, const void *u, const void *v, const void *w, void *out,
                         double scale);
/* Timing helper (no reference counterpart; bench.py prices the dominant kernel with it):
 * launches the stage kernel of RK stage `stage` `iters` times on `stream`, reading `fields`
 * as stage input and step start and writing only plan scratch, between two HIP events;
 * returns the average launch time in ms. Synchronises the stream. Call after one step. */
int aero3d_time_stage(aero3d_plan *plan, void *stream, const aero3d_fields *fields, const aero3d_geom *geom,
                      const aero3d_params *params, int stage, int iters, double *avg_ms);

/* ------------------------------------------------------------------ geometry set-up */
/* Exact signed distance (negative inside) of npts points (x, y arrays of the dtype: 0 float32,
 * 1 float64) to the closed polygon poly (n float64 (x, y) pairs: the vertices as rounded to
 * the dtype), into out (npts values of the dtype), on `stream`. Replaces the brute-force
 * distance of the 2D mask builder (ns2d/geometry.py:229-345 via geometry/sdf.py:322-392):
 * nearest segment in float64 (first minimum), even-odd crossing sign, distance in the dtype. */
int aero_polygon_sdf(void *stream, int32_t dtype, const double *poly, int32_t n, const void *x, const void *y,
                     int64_t npts, void *out);

/* An analytic 3D body (ns3d/geometry.py:385-495 samples such signed distances): kind 0 a
 * rounded box, p = centre (3), half extents less the edge radius (3), edge radius; kind 1 a
 * cylinder with its axis along y, p = centre (3), radius, half width (p[5], p[6] unused). */
#define AERO3D_SHAPE_ROUNDED_BOX 0
#define AERO3D_SHAPE_CYLINDER_Y 1
typedef struct {
  int32_t kind, reserved;
  double p[7];
} aero3d_shape;
/* Solid fractions of nshapes (1 .. 16) analytic bodies on one staggered grid, on `stream`:
 * out (nshapes, nz, ny, nx) float64 with out[b, k, j, i] = smooth_step(sdf_b(origin + (i, j, k)
 * h) + shift, width), smooth_step(d, w) = 1/2 - 1/2 sin(pi/2 clamp(d / w, -1, 1)), evaluated in
 * float64 in the host builder's operation order. Replaces the per-body meshgrid evaluation of
 * the 3D mask builder (ns3d/geometry.py:385-495) for analytic bodies. */
int aero3d_analytic_fractions(void *stream, const aero3d_shape *shapes, int32_t nshapes, const double *origin,
                              double h, int32_t nx, int32_t ny, int32_t nz, double shift, double width, void *out);

/* ------------------------------------------------------------------ 2D tracer particles */
/* The grid the velocities live on: u (ny, nx+1), v (ny+1, nx), origin (x0, y0), cell h. */
typedef struct {
  int32_t dtype, nx, ny, reserved;
  double x0, y0, h;
} aero2d_particles_grid;

/* A fixed particle pool (ns2d/particles.py:59-93), device buffers of the dtype unless noted:
 * pos (capacity, 2), birth, alive (capacity) with capacity = n_nozzles * ring, nozzle n owning
 * slots [n ring, (n + 1) ring); head (n_nozzles, int64) its next slot; per emitter acc (release
 * accumulator), last (last pulse index), rate (per nozzle and unit time), timing (4: start,
 * stop, period, duty), pulsed (1: pulsed batches); noz_pos (n_nozzles, 2), noz_owner
 * (n_nozzles, int64) the emitter of each nozzle. */
typedef struct {
  void *pos, *birth, *alive, *head, *acc, *last;
  const void *noz_pos, *noz_owner, *rate, *timing, *pulsed;
  int32_t n_nozzles, n_emitters, ring, reserved;
} aero2d_particles;

/* One particle step after a flow step (ns2d/particles.py:122-173): advect every slot with the
 * end-of-step (u, v) over dt (midpoint rule, bilinear samples on the staggered faces), kill
 * slots that leave the box or land where solid (ny, nx; may be NULL) exceeds 1/2, then run the
 * emitters' release clocks at time t (the step's end) and release into the nozzles' rings.
 * dt and t are device scalars; two launches on `stream`, nothing allocated (graph-safe). */
int aero2d_particles_step(void *stream, const aero2d_particles_grid *grid, const void *u, const void *v,
                          const void *dt, const void *t, const void *solid, aero2d_particles *particles);

/* ------------------------------------------------------------------ 2D ensemble */
/* Many independent 2D runs of uneven size (angle / Reynolds sweeps, polars) advancing together:
 * each launch of a step covers the blocks of every run through a device work table, so an
 * ensemble step costs the same ~15 launches whatever the number of runs. No reference
 * counterpart: the reference runs such sweeps one solve after another (api.py:4877-4930).
 * Each run is bound at creation to a plan and its in-place fields (fields[r] is both the
 * step's input and output), geometry, parameters and outputs; runs carry no smoke and record
 * no diagnostics rows (outputs[r].diag must be NULL). */
typedef struct aero2d_ensemble aero2d_ensemble;
int aero2d_ensemble_create(int32_t n, aero2d_plan *const *plans, const aero2d_fields *fields,
                           const aero2d_geom *geoms, const aero2d_params *params,
                           const aero2d_outputs *outputs, aero2d_ensemble **ensemble);
/* One step of every run; chained as for aero2d_step_chained (each run's dt inputs from its
 * previous ensemble step). */
int aero2d_ensemble_step(aero2d_ensemble *ensemble, void *stream, int32_t chained);
/* Autotuner of the batched launches (no reference counterpart): times every candidate tile
 * shape (4, 8 or 16 rows per 64-wide cell tile) x pass-B column slots (1, 2 or 4 column pairs
 * per block, capped per run by LDS, or 0: as many as fill 1024 elements) x row pairs per row-transform block (one, or as many as
 * fill 1024 elements) x threads of the row-transform blocks (256 or 512) on RK stages 0 and 1
 * of every run, `iters` rounds each, and rebuilds the work tables with the fastest. The timed
 * launches write only plan scratch,
 * so the runs' state and outputs are untouched (the potential buffer aside); fields, u and v
 * of later steps do not depend on the choice, force sums round differently with the tile
 * rows. choice[0..4] (may be NULL) receives the rows, the slots, the candidates timed, the
 * row-transform threads and the row fill (0: one row pair per block, 1: filled).
 * Synchronises the stream; call outside graph capture. */
int aero2d_ensemble_tune(aero2d_ensemble *ensemble, void *stream, int32_t iters, int32_t *choice);
int aero2d_ensemble_destroy(aero2d_ensemble *ensemble);

/* ------------------------------------------------------------------ learned closure (3D CNN) */
/* A small convolutional network on a periodic (nz, ny, nx) grid: n_layers 3x3x3 convolutions,
 * channels[0] inputs -> channels[n_layers] outputs, BF16 weights with FP32 accumulation; its
 * forward, backward and Adam kernels train it on the GPU. No reference counterpart (the
 * learned eddy-viscosity closure is new). Fields are channel-major (C, nz, ny, nx) float32
 * device arrays; parameters are float32, per layer w[co][ci][3][3][3] then b[co]. */
typedef struct aero_cnn aero_cnn;
typedef struct {
  int32_t nz, ny, nx;
  int32_t n_layers;        /* 1..8 */
  int32_t channels[9];     /* inputs 1..32; outputs of each layer 1, 8, 9, 16 or 32 */
  int32_t act[8];          /* per layer: 0 identity, 1 relu, 2 softplus */
} aero_cnn_desc;
int aero_cnn_create(const aero_cnn_desc *desc, aero_cnn **cnn);
int aero_cnn_destroy(aero_cnn *cnn);
int aero_cnn_param_count(const aero_cnn *cnn);
size_t aero_cnn_workspace_bytes(const aero_cnn *cnn);
/* params: device float32 array of aero_cnn_param_count values; also resets the Adam moments */
int aero_cnn_set_params(aero_cnn *cnn, void *stream, const float *params);
int aero_cnn_get_params(aero_cnn *cnn, void *stream, float *params);
/* training forward: keeps every layer's activations for aero_cnn_backward */
int aero_cnn_forward(aero_cnn *cnn, void *stream, const float *x, float *y);
/* inference forward (keeps nothing; may run while a training forward's activations are kept) */
int aero_cnn_infer(aero_cnn *cnn, void *stream, const float *x, float *y);
/* parameter gradients of sum(dy * y) of the last aero_cnn_forward into the plan; dx optional */
int aero_cnn_backward(aero_cnn *cnn, void *stream, const float *dy, float *dx);
int aero_cnn_get_grads(aero_cnn *cnn, void *stream, float *grads);
/* one Adam step (t >= 1) on the float32 parameters, then their BF16 copies */
int aero_cnn_adam_step(aero_cnn *cnn, void *stream, float lr, float beta1, float beta2, float eps, int t);

const char *aero_last_error(void);
int aero_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AEROLUME_AMD_H */
