"""Set-up of the 3D solver: static operators, ghost-rule table, body masks.

Follows the reference's set-up (``solvers/ns3d/core.py:61-203``, ``geometry.py:265-508``):
``Ops3D`` holds the grid, the boundary rules and the live device scalars (nu, eta);
``Geometry3D`` holds the window arrays of the immersed bodies - solid fractions on the
u, v, w faces and cell centres, the rigid-rotation solid velocity pre-multiplied by the
fraction, and the force / moment weights. Everything here runs once per solver on the host
(float64) and is copied to the device in the solver's dtype.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np
import torch

from .config import AnalyticBody, DyeRake, ImmersedMesh, SolverConfig3D

SSP_RK3 = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))
FORCE_WEIGHTS = (1.0 / 6.0, 2.0 / 3.0, 1.0)
LES_CONSTANTS = {"smagorinsky": 0.17, "wale": 0.5, "vreman": 0.07}
ETA_PER_H = 0.05
STIFFNESS_MIN = 60.0
MASK_SHIFT = 0.32
VISCOUS_DT = 1.0 / 6.0
COEFF_NAMES = ("cd", "cs", "cl", "c_roll", "cm", "cn")
DYE_FLOOR = 1e-8                 # smoke below this is set to exactly zero
NOZZLE_RADIUS_CELLS = 1.25
NOZZLE_DEPTH_CELLS = 1.0

# ghost-rule kinds of the C ABI (one per side of every axis of every component)
RULE_CODES = {"edge": 0, "even": 1, "odd": 2, "affine": 3, "face_odd": 4, "periodic_cell": 5,
              "periodic_face": 6}


def auto_eta(h: float, re: float) -> float:
    """Brinkman permeability rule: min(0.05 h, h^2 Re / 60) (reference ``ns3d/core.py:135-137``)."""
    return min(ETA_PER_H * h, h * h * re / STIFFNESS_MIN)


@dataclass
class Ops3D:
    """Static operators of one grid (reference ``ns3d/core.py:61-108``): Python scalars are
    static, the 0-d tensors ``nu`` and ``eta`` live on the device and may be edited in place."""
    cfg: SolverConfig3D
    nx: int
    ny: int
    nz: int
    h: float
    x0: float
    y0: float
    z0: float
    device: torch.device
    dtype: torch.dtype
    periodic: bool
    periodic_axes: tuple
    symmetry: bool
    advection: str
    turbulence: str
    bc_x: str
    bc_y: str
    bc_zmin: str
    bc_zmax: str
    ground_speed: float
    cfl: float
    dt_fixed: float | None
    ref_area: float
    length_ref: float
    span_ref: float
    moment_ref: tuple
    mask_width: float
    mask_shift: float
    eta_fixed: float | None
    les_c: float
    nu: torch.Tensor
    eta: torch.Tensor
    nu_value: float
    u_inlet: torch.Tensor            # (nz, ny, 1)
    ghost: dict
    closed: bool = False
    extras: dict = field(default_factory=dict)


def resolve_device_dtype(cfg: SolverConfig3D, device=None) -> tuple[torch.device, torch.dtype]:
    dev = device if device is not None else cfg.device
    if dev in ("auto", None, ""):
        dev = "cuda" if torch.cuda.is_available() else "cpu"
    dev = torch.device(dev)
    if dev.type == "cuda" and dev.index is None and torch.cuda.is_available():
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev, (torch.float64 if cfg.dtype == "float64" else torch.float32)


def check_supported(cfg: SolverConfig3D) -> None:
    """Raise ``NotImplementedError`` for configuration parts the MI355X 3D step does not carry."""
    missing = []
    if cfg.heat_sources and cfg.thermal is None:
        raise ValueError("heat sources need heat transfer (cfg.thermal)")
    if cfg.fans or cfg.ports:
        missing.append("fans / ports")
    if cfg.emitters:
        missing.append("emitters")
    if cfg.particle_capacity:
        missing.append("tracer particles")
    if cfg.dye and cfg.emitters is not None:
        missing.append("Emitter3D smoke sources (the rake of dye_rake is carried)")
    if missing:
        raise NotImplementedError("the MI355X 3D step does not carry: " + ", ".join(missing))


def ghost_rules(cfg: SolverConfig3D) -> dict:
    """Ghost rule per velocity component and axis: ``{comp: (x, y, z)}``, each a (lo, hi) pair
    of (kind, value) (reference ``ns3d/core.py:147-181``). ``edge`` repeats the boundary value
    (inlet / zero-gradient outlet); ``even`` / ``odd`` mirror a cell-centred quantity across a
    wall half a cell away (zero gradient / zero value); ``affine`` mirrors to a prescribed
    value (moving floor); ``face_odd`` mirrors a wall-normal component across a wall on that
    face; ``periodic_cell`` / ``periodic_face`` wrap (face arrays carry a duplicate last face)."""
    px, py, pz = cfg.periodic_axes
    pc, pf = ("periodic_cell", 0.0), ("periodic_face", 0.0)
    even, odd, edge, fodd = ("even", 0.0), ("odd", 0.0), ("edge", 0.0), ("face_odd", 0.0)
    zlo_tan = {"slip": even, "noslip": odd, "moving": ("affine", float(cfg.ground_speed))}[cfg.bc_zmin]
    zlo_n = {"slip": even, "noslip": odd, "moving": odd}[cfg.bc_zmin]
    zhi = {"slip": even, "noslip": odd}[cfg.bc_zmax]
    y_tan = odd if cfg.bc_y == "noslip" else even
    if cfg.closed:
        x_tan = odd if cfg.bc_x == "noslip" else even
        ux, vx = (fodd, fodd), (x_tan, x_tan)
    else:
        ux, vx = (edge, edge), (odd, edge)
    u = ((pf, pf) if px else ux, (pc, pc) if py else (y_tan, y_tan), (pc, pc) if pz else (zlo_tan, zhi))
    v = ((pc, pc) if px else vx, (pf, pf) if py else (fodd, fodd), (pc, pc) if pz else (zlo_n, zhi))
    w = ((pc, pc) if px else vx, (pc, pc) if py else (y_tan, y_tan), (pf, pf) if pz else (fodd, fodd))
    return {"u": u, "v": v, "w": w}


def rule_table(ghost: dict) -> tuple[list[int], list[float]]:
    """Flatten the ghost rules to 18 codes and 18 values, ordered comp (u, v, w), axis
    (x, y, z), side (lo, hi)."""
    codes, vals = [], []
    for comp in ("u", "v", "w"):
        for axis in ghost[comp]:
            for kind, val in axis:
                codes.append(RULE_CODES[kind])
                vals.append(float(val))
    return codes, vals


def make_operators3d(cfg: SolverConfig3D, device=None) -> Ops3D:
    """Static operators for a config (reference ``ns3d/core.py:184-203``)."""
    dev, dtype = resolve_device_dtype(cfg, device)
    nu = torch.tensor(1.0 / cfg.re, device=dev, dtype=dtype)
    eta = torch.tensor(cfg.eta if cfg.eta is not None else auto_eta(cfg.h, cfg.re), device=dev, dtype=dtype)
    u_in = torch.ones(cfg.nz, cfg.ny, 1, device=dev, dtype=dtype)
    les_c = cfg.les_constant if cfg.les_constant is not None else LES_CONSTANTS.get(cfg.turbulence, 0.0)
    ops = Ops3D(cfg=cfg, nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, h=cfg.h, x0=cfg.x0, y0=cfg.y0, z0=cfg.z0, device=dev,
                dtype=dtype, periodic=cfg.periodic, periodic_axes=cfg.periodic_axes, symmetry=cfg.symmetry,
                advection=cfg.advection, turbulence=cfg.turbulence, bc_x=cfg.bc_x, bc_y=cfg.bc_y,
                bc_zmin=cfg.bc_zmin, bc_zmax=cfg.bc_zmax, ground_speed=float(cfg.ground_speed), cfl=cfg.cfl,
                dt_fixed=cfg.dt, ref_area=cfg.ref_area, length_ref=cfg.length_ref, span_ref=cfg.span_ref,
                moment_ref=cfg.moment_ref, mask_width=cfg.mask_width, mask_shift=MASK_SHIFT, eta_fixed=cfg.eta,
                les_c=les_c, nu=nu, eta=eta, nu_value=1.0 / cfg.re, u_inlet=u_in, ghost=ghost_rules(cfg),
                closed=cfg.closed)
    if cfg.dt is not None:
        ops.extras["dt_fixed"] = torch.tensor(float(cfg.dt), device=dev, dtype=dtype)
    ops.extras["re"] = float(cfg.re)
    return ops


def set_reynolds(ops: Ops3D, re: float) -> None:
    """Change Re in place (nu, eta and the launch value), reference ``ns3d/core.py:140-144``."""
    ops.nu.fill_(1.0 / re)
    ops.nu_value = 1.0 / re
    ops.extras["re"] = float(re)
    ops.eta.fill_(ops.eta_fixed if ops.eta_fixed is not None else auto_eta(ops.h, re))


# ---------------------------------------------------------------------------
# signed distances
# ---------------------------------------------------------------------------

def smooth_step(d: torch.Tensor, w: float) -> torch.Tensor:
    """C1 solid fraction of a signed distance: 1 inside beyond w, 0 outside beyond w."""
    s = (d / w).clamp(-1.0, 1.0)
    return 0.5 - 0.5 * torch.sin(0.5 * math.pi * s)


def rounded_box_sdf(center, size, radius):
    """Signed distance of a box of full extents ``size`` whose edges are rounded by ``radius``."""
    cx, cy, cz = center
    hx, hy, hz = (0.5 * s - radius for s in size)

    def sdf(x, y, z):
        qx, qy, qz = (x - cx).abs() - hx, (y - cy).abs() - hy, (z - cz).abs() - hz
        out = torch.sqrt(qx.clamp_min(0) ** 2 + qy.clamp_min(0) ** 2 + qz.clamp_min(0) ** 2)
        inside = torch.maximum(torch.maximum(qx, qy), qz).clamp_max(0.0)
        return out + inside - radius
    sdf.shape = ("rounded_box", (float(cx), float(cy), float(cz), float(hx), float(hy), float(hz), float(radius)))
    return sdf


def cylinder_y_sdf(center, radius, width):
    """Signed distance of a cylinder of ``radius`` with its axis along y, ``width`` long."""
    cx, cy, cz = center

    def sdf(x, y, z):
        dr = torch.sqrt((x - cx) ** 2 + (z - cz) ** 2) - radius
        da = (y - cy).abs() - 0.5 * width
        out = torch.sqrt(dr.clamp_min(0) ** 2 + da.clamp_min(0) ** 2)
        return out + torch.maximum(dr, da).clamp_max(0.0)
    sdf.shape = ("cylinder_y", (float(cx), float(cy), float(cz), float(radius), float(0.5 * width), 0.0, 0.0))
    return sdf


def car_bodies(center=(0.0, 0.0, 0.0), length=4.0, width=1.6, height=1.2, corner=0.2, wheel_radius=0.3,
               wheel_width=0.3, ground_speed=1.0, clearance=0.2):
    """A car-like body on the floor z = center[2]: a rounded box (length x width x height)
    above four wheel cylinders that roll at ``ground_speed`` (omega = -U / r about +y)."""
    cx, cy, z0 = center
    zc = z0 + clearance + 0.5 * height
    box = AnalyticBody(rounded_box_sdf((cx, cy, zc), (length, width, height), corner),
                       ((cx - 0.5 * length, cy - 0.5 * width, zc - 0.5 * height),
                        (cx + 0.5 * length, cy + 0.5 * width, zc + 0.5 * height)), name="body")
    out = [box]
    zw = z0 + wheel_radius
    xs = (cx - 0.3 * length, cx + 0.3 * length)
    ys = (cy - 0.5 * width + 0.5 * wheel_width, cy + 0.5 * width - 0.5 * wheel_width)
    for i, x in enumerate(xs):
        for j, y in enumerate(ys):
            c = (x, y, zw)
            out.append(AnalyticBody(cylinder_y_sdf(c, wheel_radius, wheel_width),
                                    ((x - wheel_radius, y - 0.5 * wheel_width, zw - wheel_radius),
                                     (x + wheel_radius, y + 0.5 * wheel_width, zw + wheel_radius)),
                                    name=f"wheel{2 * i + j}", role="wheel", omega=-ground_speed / wheel_radius,
                                    center=c, axis=(0.0, 1.0, 0.0)))
    return out


def _tri_dist2(p, a, b, c):
    """Squared distance of points p (P, 3) to triangles a, b, c (T, 3): the foot of the
    perpendicular when it falls inside the triangle, else the nearest point of its edges."""
    p = p[:, None, :]
    a, b, c = a[None], b[None], c[None]
    ab, bc, ca = b - a, c - b, a - c
    n = torch.linalg.cross(ab, -ca)
    nn = (n * n).sum(-1)

    def inner(e, v0):
        return (n * torch.linalg.cross(e.expand_as(p - v0), p - v0)).sum(-1) >= 0

    inside = inner(ab, a) & inner(bc, b) & inner(ca, c) & (nn > 1e-300)
    pd = (n * (p - a)).sum(-1)
    plane = pd * pd / nn.clamp_min(1e-300)

    def seg(v0, e):
        w = p - v0
        t = ((w * e).sum(-1) / (e * e).sum(-1).clamp_min(1e-300)).clamp(0.0, 1.0)
        r = w - t[..., None] * e
        return (r * r).sum(-1)

    edge = torch.minimum(torch.minimum(seg(a, ab), seg(b, bc)), seg(c, ca))
    return torch.where(inside, plane, edge)


def _solid_angles(p, a, b, c):
    """Signed solid angle of each triangle seen from each point (van Oosterom & Strackee)."""
    ra, rb, rc = a[None] - p[:, None], b[None] - p[:, None], c[None] - p[:, None]
    la, lb, lc = (torch.sqrt((r * r).sum(-1)) for r in (ra, rb, rc))
    det = (ra * torch.linalg.cross(rb, rc, dim=-1)).sum(-1)
    den = la * lb * lc + (ra * rb).sum(-1) * lc + (rb * rc).sum(-1) * la + (rc * ra).sum(-1) * lb
    return 2.0 * torch.atan2(det, den)


def mesh_sdf(vertices: np.ndarray, faces: np.ndarray):
    """Signed distance of a triangle soup (negative where the generalised winding number
    exceeds 1/2), brute force over triangles in float64 (set-up only)."""
    v = torch.as_tensor(np.asarray(vertices, dtype=np.float64))
    f = torch.as_tensor(np.asarray(faces, dtype=np.int64))
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]

    def sdf(x, y, z):
        shape = x.shape
        pts = torch.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], 1).to(torch.float64)
        out = torch.empty(pts.shape[0], dtype=torch.float64)
        step = max(1, (1 << 21) // max(1, f.shape[0]))
        for s in range(0, pts.shape[0], step):
            p = pts[s:s + step]
            d2 = _tri_dist2(p, a, b, c).amin(1)
            wn = _solid_angles(p, a, b, c).sum(1)
            out[s:s + step] = torch.where(wn.abs() > 2.0 * math.pi, -1.0, 1.0) * torch.sqrt(d2)
        return out.reshape(shape)
    return sdf


def body_sdf(body):
    if isinstance(body, AnalyticBody):
        return body.sdf
    if isinstance(body, ImmersedMesh):
        return mesh_sdf(body.vertices, body.faces)
    raise TypeError(f"unsupported body {type(body).__name__}")


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------

class Geometry3D(NamedTuple):
    """Window arrays of all bodies (reference ``ns3d/geometry.py:275-303``).

    ``window = (k0, k1, j0, j1, i0, i1)`` is a block of cells. u-face arrays cover layers
    k0..k1-1, rows j0..j1-1, faces i0..i1 (wz, wy, wx+1); v-face arrays (wz, wy+1, wx);
    w-face arrays (wz+1, wy, wx); chi_c (wz, wy, wx). ``k_u`` (nb, 3, ...) maps u increments
    to [Fx, My, Mz] parts, ``k_v`` v increments to [Fy, Mx, Mz], ``k_w`` w increments to
    [Fz, Mx, My] (lever arms about the moment reference).
    """
    window: tuple
    chi_u: torch.Tensor
    chi_v: torch.Tensor
    chi_w: torch.Tensor
    chi_c: torch.Tensor
    a_u: torch.Tensor
    a_v: torch.Tensor
    a_w: torch.Tensor
    k_u: torch.Tensor
    k_v: torch.Tensor
    k_w: torch.Tensor
    nbody: int

    @property
    def empty(self) -> bool:
        k0, k1, j0, j1, i0, i1 = self.window
        return k1 <= k0 or j1 <= j0 or i1 <= i0

    def to(self, device=None, dtype=None) -> "Geometry3D":
        def cv(t):
            return t.to(device=device, dtype=dtype).contiguous()
        return Geometry3D(self.window, *(cv(t) for t in self[1:11]), self.nbody)


def rotation_matrix(pitch_deg: float, yaw_deg: float) -> np.ndarray:
    """Attitude: yaw about +z by -yaw, then pitch about +y by +pitch (reference
    ``ns3d/geometry.py:315-320``)."""
    a, b = math.radians(pitch_deg), math.radians(-yaw_deg)
    ry = np.array([[math.cos(a), 0.0, math.sin(a)], [0.0, 1.0, 0.0], [-math.sin(a), 0.0, math.cos(a)]])
    rz = np.array([[math.cos(b), -math.sin(b), 0.0], [math.sin(b), math.cos(b), 0.0], [0.0, 0.0, 1.0]])
    return ry @ rz


def body_window3d(vertex_sets, ops: Ops3D, margin_cells: float | None = None, *, pitch_deg: float = 0.0,
                  yaw_deg: float = 0.0) -> tuple:
    """Cell window covering every vertex plus a margin of mask_width + 3 + |shift| cells,
    clipped to the domain (reference ``ns3d/geometry.py:323-356``)."""
    clouds = [c for c in (np.asarray(v, dtype=np.float64).reshape(-1, 3) for v in vertex_sets) if c.size]
    if not clouds:
        return (0, 0, 0, 0, 0, 0)
    pts = np.concatenate(clouds)
    if pitch_deg > 0 or yaw_deg > 0:
        # the bodies at every attitude a live edit may give them: 7 samples of each swept angle
        def sweep(amp):
            return np.linspace(-amp, amp, 7) if amp > 0 else np.zeros(1)

        ref = np.asarray(ops.moment_ref, dtype=np.float64)
        arm = pts - ref
        pts = np.concatenate([arm @ rotation_matrix(a, b).T + ref for a in sweep(pitch_deg) for b in sweep(yaw_deg)])
    margin = (ops.mask_width + 3.0 + abs(ops.mask_shift)) if margin_cells is None else margin_cells
    h = ops.h
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    org = (ops.x0, ops.y0, ops.z0)
    n = (ops.nx, ops.ny, ops.nz)
    box = []
    for a in (2, 1, 0):
        c0 = max(math.floor((lo[a] - org[a]) / h - margin), 0)
        c1 = min(math.ceil((hi[a] - org[a]) / h + margin), n[a])
        if c1 <= c0:
            return (0, 0, 0, 0, 0, 0)
        box += [c0, c1]
    return tuple(box)


def empty_geometry(nb: int, device=None, dtype=torch.float64) -> Geometry3D:
    z = torch.zeros(0, 0, 0, device=device, dtype=dtype)
    k = torch.zeros(0, 3, 0, 0, 0, device=device, dtype=dtype)
    return Geometry3D((0, 0, 0, 0, 0, 0), z, z, z, z, z, z, z, k, k, k, nb)


def _analytic_shapes(bodies):
    """``(kind, params)`` of every body when all are analytic shapes the HIP fraction kernel
    knows (the rounded box and the y cylinder above), else None."""
    out = []
    for b in bodies:
        shape = getattr(b.sdf, "shape", None) if isinstance(b, AnalyticBody) else None
        if shape is None:
            return None
        out.append(shape)
    return out


def _hip_fractions(shapes, start, n, h, shift, width, device) -> torch.Tensor:
    """Solid fractions (nbody, nz, ny, nx) float64 of analytic shapes on the grid whose first
    sample is ``start`` = (z, y, x) and counts ``n`` = (nz, ny, nx), by the HIP kernel
    ``aero3d_analytic_fractions`` (the host builder's float64 formulas, at most 16 bodies a
    launch)."""
    import ctypes

    from .. import _lib
    from .._lib3d import SHAPE_KINDS, Shape3D
    lib = _lib.load()
    nz, ny, nx = (int(v) for v in n)
    out = torch.empty(len(shapes), nz, ny, nx, dtype=torch.float64, device=device)
    origin = (ctypes.c_double * 3)(float(start[2]), float(start[1]), float(start[0]))
    stream = ctypes.c_void_p(_lib.stream_handle(device))
    for b0 in range(0, len(shapes), 16):
        part = shapes[b0:b0 + 16]
        arr = (Shape3D * len(part))()
        for s, (kind, params) in zip(arr, part):
            s.kind = SHAPE_KINDS[kind]
            for q, val in enumerate(params):
                s.p[q] = float(val)
        _lib.check(lib.aero3d_analytic_fractions(stream, arr, len(part), origin, float(h), nx, ny, nz, float(shift),
                                                 float(width), out[b0].data_ptr()), "aero3d_analytic_fractions")
    return out


def build_geometry3d(bodies, ops: Ops3D, *, window=None) -> Geometry3D:
    """Masks, solid velocities and force weights on the window, built in float64 (on a GPU
    device the fractions of analytic shapes come from the HIP kernel and the rest runs there;
    otherwise on the host) and returned in the solver's dtype on its device (semantics of the reference's
    ``ns3d/geometry.py:385-508``: per-body fraction ``smooth_step(d + 0.32 h, mask_width h)``
    on each staggered grid, zero on domain-boundary faces, union clamped at 1; rigid rotation
    pre-multiplied by each body's fraction; each body's share of a face where bodies overlap)."""
    nb = len(bodies)
    if window is None:
        window = body_window3d([b.vertices_numpy() for b in bodies], ops,
                               pitch_deg=ops.cfg.window_pitch_deg, yaw_deg=ops.cfg.window_yaw_deg)
    k0, k1, j0, j1, i0, i1 = window
    wz, wy, wx = k1 - k0, j1 - j0, i1 - i0
    if nb == 0 or wz <= 0 or wy <= 0 or wx <= 0:
        return empty_geometry(nb, ops.device, ops.dtype)
    f64 = torch.float64
    h = ops.h
    w = ops.mask_width * h
    shift = ops.mask_shift * h
    X0, Y0, Z0 = ops.x0 + i0 * h, ops.y0 + j0 * h, ops.z0 + k0 * h
    shapes = _analytic_shapes(bodies) if ops.device.type == "cuda" else None
    dev = ops.device if shapes is not None else torch.device("cpu")

    # first sample (z, y, x) and point counts (z, y, x) of each staggered grid
    starts = {
        "u": ((Z0 + 0.5 * h, Y0 + 0.5 * h, X0), (wz, wy, wx + 1)),
        "v": ((Z0 + 0.5 * h, Y0, X0 + 0.5 * h), (wz, wy + 1, wx)),
        "w": ((Z0, Y0 + 0.5 * h, X0 + 0.5 * h), (wz + 1, wy, wx)),
        "c": ((Z0 + 0.5 * h, Y0 + 0.5 * h, X0 + 0.5 * h), (wz, wy, wx)),
    }
    grids = {k: tuple(s0 + torch.arange(n, dtype=f64, device=dev) * h for s0, n in zip(*v))
             for k, v in starts.items()}
    if shapes is not None:
        chis = {k: _hip_fractions(shapes, start, n, h, shift, w, dev) for k, (start, n) in starts.items()}
    else:
        chis = {k: [] for k in grids}
        for b in bodies:
            sdf = body_sdf(b)
            for key, (zs, ys, xs) in grids.items():
                zz, yy, xx = torch.meshgrid(zs, ys, xs, indexing="ij")
                chis[key].append(smooth_step(sdf(xx, yy, zz) + shift, w))
        chis = {k: torch.stack(v) for k, v in chis.items()}
    cu, cv, cw, cc = (chis[k] for k in ("u", "v", "w", "c"))
    px, py, pz = ops.periodic_axes
    if not px:
        if i0 == 0:
            cu[..., 0] = 0.0
        if i1 == ops.nx:
            cu[..., -1] = 0.0
    if not py:
        if j0 == 0:
            cv[:, :, 0] = 0.0
        if j1 == ops.ny:
            cv[:, :, -1] = 0.0
    if not pz:
        if k0 == 0:
            cw[:, 0] = 0.0
        if k1 == ops.nz:
            cw[:, -1] = 0.0
    chi_u, chi_v, chi_w, chi_c = (c.sum(0).clamp(max=1.0) for c in (cu, cv, cw, cc))

    zu, yu = grids["u"][0][:, None, None], grids["u"][1][None, :, None]
    zv, xv = grids["v"][0][:, None, None], grids["v"][2][None, None, :]
    yw, xw = grids["w"][1][None, :, None], grids["w"][2][None, None, :]
    a_u, a_v, a_w = torch.zeros_like(chi_u), torch.zeros_like(chi_v), torch.zeros_like(chi_w)
    for i, b in enumerate(bodies):
        om = b.omega
        if om == 0.0:
            continue
        (cx, cy, cz), (ax, ay, az) = b.center, b.axis
        a_u = a_u + cu[i] * (om * (ay * (zu - cz) - az * (yu - cy)))
        a_v = a_v + cv[i] * (om * (az * (xv - cx) - ax * (zv - cz)))
        a_w = a_w + cw[i] * (om * (ax * (yw - cy) - ay * (xw - cx)))
    # each body's share of a face's penalisation force: its mask over the summed masks (a lone
    # body takes the whole force)
    def share(c):
        return torch.ones_like(c) if nb == 1 else c / (c.sum(0, keepdim=True) + 1e-30)

    wu, wv, ww = share(cu), share(cv), share(cw)
    xr, yr, zr = ops.moment_ref
    # M = r x F: Fx at u-faces gives My += (z - zr) Fx, Mz -= (y - yr) Fx; Fy at v-faces gives
    # Mx -= (z - zr) Fy, Mz += (x - xr) Fy; Fz at w-faces gives Mx += (y - yr) Fz, My -= (x - xr) Fz
    k_u = torch.stack([wu, wu * (zu - zr), -wu * (yu - yr)], dim=1)
    k_v = torch.stack([wv, -wv * (zv - zr), wv * (xv - xr)], dim=1)
    k_w = torch.stack([ww, ww * (yw - yr), -ww * (xw - xr)], dim=1)
    g = Geometry3D(tuple(int(c) for c in window), chi_u, chi_v, chi_w, chi_c, a_u, a_v, a_w, k_u, k_v, k_w, nb)
    return g.to(ops.device, ops.dtype)


# ---------------------------------------------------------------------------
# smoke sources (reference ns3d/dye.py:84-250)
# ---------------------------------------------------------------------------

def default_dye_rake(bounds, extent, h: float, *, kind: str = "rake", symmetry: bool = False, ground: bool = False,
                     n: int | None = None) -> DyeRake:
    """Smoke nozzles placed from the bodies' bounding box (the reference's placement,
    ``ns3d/dye.py:84-150``). Release plane: 0.4 of the body's smaller size (frontal or
    streamwise) upstream of the nose, kept four cells off the inlet and two off the nose.
    Cross-stream: the body's span widened by 20 % each side and its height shifted up by
    10 %, both kept a nozzle radius and a cell inside the domain (on a road, two cells above
    it). "rake": one row across the span at 40 % of the body height; "grid": span x height;
    "plane": a band between two corners; nozzles about six radii apart. With a symmetry plane
    only the solved side's nozzles are kept."""
    (bx0, bx1, by0, by1, bz0, bz1), (ex0, _, ey0, ey1, ez0, ez1) = ([float(v) for v in q] for q in (bounds, extent))
    span, height = max(by1 - by0, h), max(bz1 - bz0, h)
    frontal = max(span, height)
    radius = NOZZLE_RADIUS_CELLS * h
    inward = lambda e: e - radius - h  # noqa: E731   (a nozzle's extent plus a cell inside a far side)
    outward = lambda e: e + radius + h  # noqa: E731
    upstream = max(0.4 * min(frontal, max(bx1 - bx0, h)), 4.0 * h)
    x = min(max(bx0 - upstream, ex0 + 4.0 * h), bx0 - 2.0 * h)
    # the mirrored body of a half-domain run spans both sides of the plane, like a full run
    y_floor = 2.0 * ey0 - inward(ey1) if symmetry else outward(ey0)
    y_lo, y_hi = max(by0 - 0.2 * span, y_floor), min(by1 + 0.2 * span, inward(ey1))
    z_lo, z_hi = bz0 + 0.1 * height, bz1 + 0.1 * height
    if ground:
        z_lo = max(z_lo, ez0 + 2.0 * h + radius)
    z_lo, z_hi = max(z_lo, outward(ez0)), min(z_hi, inward(ez1))
    if kind == "plane":
        return DyeRake("plane", [(x, max(y_lo, ey0) if symmetry else y_lo, z_lo), (x, y_hi, z_hi)], radius=radius)
    gap = max(6.0 * radius, frontal / 7.0)
    count = lambda extent_, lo, hi: max(lo, min(hi, round(extent_ / gap) + 1))  # noqa: E731
    kept = lambda ys: ys[ys >= ey0 - 1e-9] if symmetry else ys  # noqa: E731
    if kind == "rake":
        ys = kept(np.linspace(y_lo, y_hi, int(n) if n else count(y_hi - y_lo, 3, 17)))
        z = min(max(bz0 + 0.40 * height, ez0 + radius + 2 * h), inward(ez1))
        return DyeRake("rake", [(x, y, z) for y in ys], radius=radius)
    if n:
        ky = max(2, round(math.sqrt(n * (y_hi - y_lo) / max(z_hi - z_lo, 1e-9))))
        kz = max(2, round(n / ky))
    else:
        ky, kz = count(y_hi - y_lo, 3, 11), count(z_hi - z_lo, 2, 9)
    ys, zs = kept(np.linspace(y_lo, y_hi, ky)), np.linspace(z_lo, z_hi, kz)
    return DyeRake("grid", [(x, y, z) for z in zs for y in ys], radius=radius)


def _ramp(d, edge: float, h: float):
    """1 inside edge - h/2, 0 beyond edge + h/2, linear between (anti-aliased edge)."""
    return np.clip((edge + 0.5 * h - d) / h, 0.0, 1.0)


def dye_source_weights(rake: DyeRake | None, ops: Ops3D):
    """(window, weights (numpy)) of a rake's nozzles on the grid cells (reference
    ``ns3d/dye.py:171-226``): radial ramp of the nozzle radius times an axial ramp of the
    nozzle depth, the maximum over nozzles; a "plane" band is a product of ramps."""
    empty = (0, 0, 0, 0, 0, 0), np.zeros((0, 0, 0))
    if rake is None:
        return empty
    pos = np.asarray(rake.positions, dtype=np.float64).reshape(-1, 3)
    if not len(pos):
        return empty
    h = ops.h
    r = rake.radius if rake.radius > 0 else NOZZLE_RADIUS_CELLS * h
    depth = NOZZLE_DEPTH_CELLS * h
    if rake.kind == "plane":
        lo, hi = pos.min(axis=0) - np.array([depth, 0.0, 0.0]), pos.max(axis=0) + np.array([depth, 0.0, 0.0])
    else:
        lo, hi = pos.min(axis=0) - np.array([depth, r, r]), pos.max(axis=0) + np.array([depth, r, r])
    lo, hi = lo - h, hi + h
    org, n = (ops.x0, ops.y0, ops.z0), (ops.nx, ops.ny, ops.nz)
    c0 = [max(math.floor((lo[a] - org[a]) / h), 0) for a in range(3)]
    c1 = [min(math.ceil((hi[a] - org[a]) / h), n[a]) for a in range(3)]
    if any(c1[a] <= c0[a] for a in range(3)):
        return empty
    xc, yc, zc = (org[a] + (np.arange(c0[a], c1[a]) + 0.5) * h for a in range(3))
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    if rake.kind == "plane":
        (_, ya, za), (_, yb, zb) = pos[0], pos[1]
        ya, yb = min(ya, yb), max(ya, yb)
        za, zb = min(za, zb), max(za, zb)
        w = (_ramp(np.abs(X - 0.5 * (pos[0, 0] + pos[1, 0])), depth, h) * _ramp(np.maximum(ya - Y, Y - yb), 0.0, h)
             * _ramp(np.maximum(za - Z, Z - zb), 0.0, h))
    else:
        w = np.zeros_like(X)
        for px, py, pz in pos:
            rr = np.sqrt((Y - py) ** 2 + (Z - pz) ** 2)
            w = np.maximum(w, _ramp(rr, r, h) * _ramp(np.abs(X - px), depth, h))
    if not w.any():
        return empty
    return (c0[2], c1[2], c0[1], c1[1], c0[0], c1[0]), w


def dye_pulse(rake: DyeRake | None) -> list[float]:
    """[t_on, period]: a continuous rake gets t_on > period (always on)."""
    if rake is not None and rake.pulsed:
        return [rake.pulse_on, rake.pulse_on + rake.pulse_off]
    return [2.0, 1.0]


def auto_dye_rake(cfg: SolverConfig3D, ops: Ops3D, kind: str = "rake") -> DyeRake:
    """The automatic rake of a configuration (reference ``FlowSolver3D.default_dye_rake``):
    from the bodies' bounding box, or across the middle of the inlet third without bodies."""
    if cfg.bodies:
        v = np.vstack([b.vertices_numpy() for b in cfg.bodies])
        lo, hi = v.min(axis=0), v.max(axis=0)
        bounds = (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    else:
        ex = cfg.extent
        cx, cy, cz = ex[0] + 0.3 * (ex[1] - ex[0]), 0.5 * (ex[2] + ex[3]), 0.5 * (ex[4] + ex[5])
        q = 0.15 * min(ex[3] - ex[2], ex[5] - ex[4])
        bounds = (cx, cx + q, cy - q, cy + q, cz - q, cz + q)
    return default_dye_rake(bounds, cfg.extent, ops.h, kind=kind, symmetry=ops.symmetry,
                            ground=cfg.bc_zmin in ("noslip", "moving"))
