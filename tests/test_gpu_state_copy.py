"""The single-precision step's face sweeps read float copies of the stage state (U and its
primitives) that the stage kernel writes beside its FP64 output; the halo fill covers the copies,
their z halos only for the copies.  The sweeps convert every loaded value to float either way, so
the steps must be bitwise those of the sweeps reading the double arrays ($NSGPU_STATE_COPY=0):
characteristic and component-wise WENO-5, periodic and walled faces, both viscosity laws, the
learned weights, calls of one step (prologue kernel only) and of several (the prologue formed by
the last stage), and ensembles of uneven blocks (whose batched kernels use the members' copies).
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package
from test_gpu_ensemble import _members
from test_mlp_oracle import random_params

pytestmark = pytest.mark.gpu

WALLS = ["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"]


@pytest.fixture(scope="module")
def ns():
    return package()


def _run(ns, case, W, monkeypatch, copy, calls, params=None):
    monkeypatch.setenv("NSGPU_STATE_COPY", copy)
    s = ns.Solver(case, math="fp32")
    if params is not None:
        s.set_mlp(params)
    s.set_state(W, "primitive")
    for n in calls:
        s.step(n)
    out = (po.interior(s.get_state("primitive")), po.interior(s.get_state("conservative")))
    s.close()
    return out


@pytest.mark.parametrize("conv", ["weno5cw", "weno5"])
@pytest.mark.parametrize("bc", ["periodic", "walls"])
@pytest.mark.parametrize("viscosity", ["constant", "sutherland"])
def test_state_copies_give_the_same_steps(ns, monkeypatch, conv, bc, viscosity):
    print(f"[gpu] state copies {conv} {bc} {viscosity}", flush=True)
    C = cases_module()
    case = dict(C.base_case(40, 24, 36), conv=conv, viscosity=viscosity,
                bc=(["periodic"] * 6 if bc == "periodic" else WALLS))
    W = C.random_smooth_W(case, seed=17)
    a = _run(ns, case, W, monkeypatch, "0", (1, 3))
    b = _run(ns, case, W, monkeypatch, "1", (1, 3))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert np.all(np.isfinite(a[0]))


@pytest.mark.parametrize("conv", ["weno5cw", "weno5"])
def test_state_copies_odd_block(ns, monkeypatch, conv):
    # extents that leave partial k segments, row tiles and sweep tiles in every kernel
    print(f"[gpu] state copies odd block {conv}", flush=True)
    C = cases_module()
    case = dict(C.base_case(67, 45, 130), conv=conv, viscosity="sutherland", L=(1.3, 0.9, 2.1))
    W = C.random_smooth_W(case, seed=31)
    a = _run(ns, case, W, monkeypatch, "0", (3,))
    b = _run(ns, case, W, monkeypatch, "1", (3,))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("conv,alpha,physics", [("weno5", "local_max", "reference"),
                                                ("weno5cw", "local_max", "reference"),
                                                ("weno5cw", "const3", "corrected_convection")])
def test_sweeps_with_options_keep_the_double_halos(ns, monkeypatch, conv, alpha, physics):
    # the per-face alpha and the corrected projection read the double arrays: their halos
    # (U, and the z faces) must still be filled when the context holds copies
    print(f"[gpu] state copies with options {conv} {alpha} {physics}", flush=True)
    C = cases_module()
    case = dict(C.base_case(19, 13, 11), conv=conv, alpha_mode=alpha, physics=physics,
                L=(1.0, 0.8, 0.7), bc=WALLS)
    W = C.random_smooth_W(case, seed=29)
    a = _run(ns, case, W, monkeypatch, "0", (1, 3))
    b = _run(ns, case, W, monkeypatch, "1", (1, 3))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    # and the same steps in EXACT math agree to the single-precision round-off
    monkeypatch.setenv("NSGPU_STATE_COPY", "1")
    s = ns.Solver(case, math="exact")
    s.set_state(W, "primitive")
    s.step(1)
    s.step(3)
    e = po.interior(s.get_state("primitive"))
    s.close()
    d = np.abs(e - b[0]).reshape(-1, 5).max(axis=0) / np.abs(e).reshape(-1, 5).max(axis=0)
    assert d.max() < 1e-4, d


@pytest.mark.parametrize("nz", [1, 2])
def test_state_copies_two_dimensional(ns, monkeypatch, nz):
    # a 2-D block (no z sweep; its z halos are thinner than the halo for nz < 3)
    print(f"[gpu] state copies 2-D nz={nz}", flush=True)
    C = cases_module()
    case = dict(C.base_case(32, 24, nz), dim=2, L=(1.0, 1.0, nz / 32.0), conv="weno5cw",
                bc=WALLS)
    W = C.random_smooth_W(case, seed=23)
    a = _run(ns, case, W, monkeypatch, "0", (1, 3))
    b = _run(ns, case, W, monkeypatch, "1", (1, 3))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_state_copies_learned_weights(ns, monkeypatch):
    print("[gpu] state copies learned", flush=True)
    C = cases_module()
    case = dict(C.base_case(40, 24, 36), conv="weno5", weights="learned")
    W = C.random_smooth_W(case, seed=19)
    params = random_params(41)
    a = _run(ns, case, W, monkeypatch, "0", (2, 2), params)
    b = _run(ns, case, W, monkeypatch, "1", (2, 2), params)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("conv,alpha", [("weno5cw", "const3"), ("weno5", "const3"),
                                        ("weno5cw", "local_max")])
def test_state_copies_in_ensembles(ns, monkeypatch, conv, alpha):
    print(f"[gpu] state copies ensemble {conv} {alpha}", flush=True)
    members = [(dict(case, alpha_mode=alpha), W) for case, W in _members(conv=conv)]
    out = {}
    for copy in ("0", "1"):
        monkeypatch.setenv("NSGPU_STATE_COPY", copy)
        solvers = []
        for case, W in members:
            s = ns.Solver(case, math="fp32")
            s.set_state(W, "primitive")
            solvers.append(s)
        ens = ns.Ensemble(solvers)
        ens.step(2)
        ens.step(1, graph=False)
        out[copy] = [(po.interior(s.get_state("primitive")),
                      po.interior(s.get_state("conservative"))) for s in solvers]
        ens.close()
        for s in solvers:
            s.close()
    for a, b in zip(out["0"], out["1"]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
