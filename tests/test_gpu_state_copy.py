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


@pytest.mark.parametrize("conv", ["weno5cw", "weno5"])
def test_state_copies_in_ensembles(ns, monkeypatch, conv):
    print(f"[gpu] state copies ensemble {conv}", flush=True)
    members = _members(conv=conv)
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
