"""The C-ABI library loads and exports every function ``include/aerolume_amd.h`` declares;
plan creation validates its descriptor (no GPU needed for either)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for hdr in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(aero[a-z0-9_]*)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    for n in ("aero2d_plan_create", "aero2d_step", "aero2d_project", "aero2d_poisson_solve", "aero_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from aerolume_amd import _lib
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.aero_version() >= 1


def test_plan_create_rejects_bad_descriptors():
    from aerolume_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    d = _lib.Desc2D()
    d.nx, d.ny, d.h = 2, 2, 0.1
    assert lib.aero2d_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.AERO_EINVAL
    assert b"4 x 4" in lib.aero_last_error()
    d.nx, d.ny = 14, 8   # 14 = 2 * 7: not 5-smooth
    rc = lib.aero2d_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _lib.AERO_EUNSUPPORTED
    with pytest.raises(NotImplementedError):
        _lib.check(rc, "plan")


def test_plan3d_create_rejects_bad_descriptors():
    from aerolume_amd import _lib
    from aerolume_amd._lib3d import Desc3D
    lib = _lib.load()
    h = ctypes.c_void_p()
    d = Desc3D()
    d.nx, d.ny, d.nz, d.h = 8, 8, 2, 0.1
    assert lib.aero3d_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.AERO_EINVAL
    d.nz = 14                # 2 * 7: not 5-smooth
    assert lib.aero3d_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.AERO_EUNSUPPORTED
    d.nz = 8
    d.ghost_kind[3] = 9      # unknown ghost rule
    assert lib.aero3d_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.AERO_EINVAL
    d.ghost_kind[3] = 1
    d.periodic_y = 1         # a body window on the periodic boundary
    d.win_k0, d.win_k1, d.win_j0, d.win_j1, d.win_i0, d.win_i1, d.nbody = 1, 4, 0, 4, 1, 4, 1
    assert lib.aero3d_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.AERO_EUNSUPPORTED
    assert b"periodic" in lib.aero_last_error()


def test_solver_refuses_cpu_device():
    from aerolume_amd.ns2d import solver as S
    from tests import cases
    with pytest.raises(RuntimeError, match="GPU"):
        S.make_operators(cases.trace_cfg("float64", device="cpu"))
    from aerolume_amd.ns3d import solver as S3
    with pytest.raises(RuntimeError, match="GPU"):
        S3.make_operators3d(cases.cfg3d("laminar", device="cpu"))


def test_3d_unsupported_features_fail_loudly():
    from aerolume_amd.ns3d import setup as su
    from aerolume_amd.ns3d.config import SolverConfig3D
    su.check_supported(SolverConfig3D(nx=8, ny=8, nz=8, h=0.1, x0=0, y0=0, z0=0, re=100.0))   # smoke is carried
    cfg = SolverConfig3D(nx=8, ny=8, nz=8, h=0.1, x0=0, y0=0, z0=0, re=100.0, particle_capacity=16)
    with pytest.raises(NotImplementedError, match="particle"):
        su.check_supported(cfg)


def test_ctypes_structures_match_the_header(tmp_path):
    """Every ctypes mirror has the C structure's size and field offsets (gcc on the header)."""
    import shutil
    import subprocess
    from aerolume_amd import _lib, _lib3d
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    pairs = {"aero2d_desc": _lib.Desc2D, "aero2d_fields": _lib.Fields2D, "aero2d_geom": _lib.Geom2D,
             "aero2d_params": _lib.Params2D, "aero2d_outputs": _lib.Outputs2D, "aero3d_desc": _lib3d.Desc3D,
             "aero3d_fields": _lib3d.Fields3D, "aero3d_geom": _lib3d.Geom3D, "aero3d_params": _lib3d.Params3D,
             "aero3d_outputs": _lib3d.Outputs3D, "aero2d_particles_grid": _lib.ParticlesGrid2D,
             "aero2d_particles": _lib.Particles2D, "aero2d_trace": _lib.Trace2D}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "aerolume_amd.h"', "int main(void) {"]
    for cname, cls in pairs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out if ln}
    for cname, cls in pairs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
