"""The stage with the z face fluxes formed in the residual kernel (k_residual_zfd, an option that
$NSGPU_ZFD=1 turns on: no z sweep, fz never stored but for the first face of each row segment)
against the oracle, and against the default z sweep + residual path in every math mode, on row
segments of 64, 128 and 256 lanes with partial segments, partial row groups, walls and periodic
z, and several segments per row.  The residual's z term is (fz(k + 1/2) - fz(k - 1/2)) / dz of weno-5.cpp:60-66 (x, then
y, then z per cell); the same face_flux on the same split values gives it bitwise.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    return package()


WALLS_Z = ["periodic", "periodic", "periodic", "periodic", "neumann", "neumann"]
MIXED = ["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"]

# (nx, ny, nz), bc, convection, Sutherland, lengths: segments of 128 (one partial, 2 rows per
# block over an odd ny), 256 (two per row, the second partial; walls in z) and 64 (4 rows per
# block, ny not a multiple of 4), square and non-square dx, dy
GEOMETRIES = [
    ((19, 13, 70), ["periodic"] * 6, "weno5cw", False, (1.0, 0.8, 0.7)),
    ((10, 9, 300), WALLS_Z, "weno5", True, (1.0, 0.9, 3.0)),
    ((21, 6, 33), MIXED, "weno5cw", True, (1.0, 1.0, 1.5)),
    ((17, 11, 256), ["periodic"] * 6, "weno5", False, (1.0, 1.0, 1.0)),
]


def _case(C, dims, bc, conv, suth, L):
    case = dict(C.base_case(*dims), conv=conv, bc=bc, mu=3e-3, L=L)
    if suth:
        C.sutherland(case, case["mu"], 2.2 ** 2)
    return case


def _solver(ns, case, math, zfd, monkeypatch):
    # the option is read when the context is created
    monkeypatch.setenv("NSGPU_ZFD", "1" if zfd else "0")
    s = ns.Solver(case, math=math)
    monkeypatch.delenv("NSGPU_ZFD")
    return s


def _run(ns, case, W0, steps, monkeypatch, math="exact", zfd=True):
    s = _solver(ns, case, math, zfd, monkeypatch)
    s.set_state(W0, "primitive")
    t, dt = s.step(steps)
    out = (po.interior(s.get_state("primitive")), po.interior(s.get_state("conservative")), dt)
    s.close()
    return out


@pytest.mark.parametrize("geo", GEOMETRIES, ids=["seg128", "seg256x2_walls", "seg64_mixed",
                                                 "seg256"])
def test_zfd_steps_match_oracle_bitwise(ns, oracle, geo, monkeypatch):
    C = cases_module()
    case = _case(C, *geo)
    W0 = C.random_smooth_W(case, seed=17)
    print(f"[gpu] zfd vs oracle {geo[0]} {geo[2]}", flush=True)
    W, U, dt = _run(ns, case, W0, 3, monkeypatch)
    Wo, Uo, dto = oracle.step(case, W0, 3)
    assert dt == dto[-1]
    np.testing.assert_array_equal(W, po.interior(Wo))
    np.testing.assert_array_equal(U, po.interior(Uo))


@pytest.mark.parametrize("math", ["exact", "fast", "fp32"])
def test_zfd_equals_z_sweep_path(ns, math, monkeypatch):
    # the two stage paths are bitwise the same in every math mode (FAST and FP32 have no bitwise
    # oracle); a 48 x 40 x 260 block: two segments per row
    C = cases_module()
    case = _case(C, (48, 40, 260), MIXED, "weno5cw", True, (1.0, 1.0, 2.0))
    W0 = C.random_smooth_W(case, seed=5)
    print(f"[gpu] zfd vs z sweep, {math}", flush=True)
    a = _run(ns, case, W0, 3, monkeypatch, math)
    b = _run(ns, case, W0, 3, monkeypatch, math, zfd=False)
    assert a[2] == b[2]
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_zfd_replaces_the_z_sweep(ns, monkeypatch):
    # the timed stage launches no z sweep where the residual forms the z faces, and does by
    # default and where the configuration is not covered (CD6)
    C = cases_module()
    case = _case(C, (16, 16, 64), ["periodic"] * 6, "weno5cw", False, (1.0, 1.0, 1.0))
    W0 = C.random_smooth_W(case, seed=3)
    counts = {}
    for label, c, on in (("zfd", case, True), ("off", case, False),
                         ("cd6", dict(case, diff_order=6), True)):
        s = _solver(ns, c, "exact", on, monkeypatch)
        s.set_state(W0, "primitive")
        s.timing(True)
        s.step(1)
        counts[label] = s.timers()["face_flux_z"][1]
        s.close()
    assert counts["zfd"] == 0 and counts["off"] > 0 and counts["cd6"] > 0, counts
