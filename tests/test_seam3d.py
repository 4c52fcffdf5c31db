"""CPU checks of the 3D kernel seam (``aerolume_amd.ns3d.seam``): the reference's types carry the
attributes the HIP step reads (``tests/golden/ref_interface_fields.json``, their field names
parsed from the reference's sources), and ``load_kernels`` refuses where the step cannot run,
with the ``RuntimeError`` the reference falls back on."""
from __future__ import annotations

import json
from pathlib import Path

import pytest
import torch

GOLDEN = Path(__file__).resolve().parent / "golden"


def _fields():
    path = GOLDEN / "ref_interface_fields.json"
    if not path.exists():
        pytest.skip("missing fixture ref_interface_fields.json")
    return json.loads(path.read_text())


def test_state_and_geometry_fields_match_the_reference():
    from aerolume_amd.ns3d import setup as su
    from aerolume_amd.ns3d.solver import State3D
    f = _fields()
    assert list(State3D._fields) == f["State3D"]["fields"]
    geom_fields = list(su.Geometry3D._fields)
    for name in f["Geometry3D"]["fields"]:
        assert name in geom_fields, name
    assert "empty" in f["Geometry3D"]["methods"]


def test_ops_attributes_the_step_reads_exist_on_the_reference_ops():
    """Every Ops3D attribute the plan and the functional step read is a field of the
    reference's Ops3D (so its objects can be handed over as they are)."""
    read = {"advection", "cfl", "closed", "device", "dtype", "eta", "extras", "ghost", "h", "length_ref", "les_c",
            "moment_ref", "nu", "nx", "ny", "nz", "periodic_axes", "ref_area", "span_ref", "symmetry", "turbulence",
            "u_inlet", "y0", "dt_fixed"}
    ref = set(_fields()["Ops3D"]["fields"])
    assert read <= ref, sorted(read - ref)


def test_load_kernels_refuses_without_a_gpu():
    from aerolume_amd.ns3d import seam
    with pytest.raises(RuntimeError):
        seam.load_kernels("upwind3", torch.device("cpu"), torch.float32)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            seam.load_kernels("upwind3", torch.device("cuda"), torch.float32)
