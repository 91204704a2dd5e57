"""Within one nsgpu_step call of several steps, the last stage of every step but the last forms
the next step's prologue (its CFL minimum, U = p2c(W) and the stage-1 primitives) from the
primitives it holds, in place of the prologue kernel.  The same arithmetic on the same values:
a call of n steps must equal n calls of one step bitwise (state, time and dt) in every math
mode, for both WENO forms, walls and periodic faces, both viscosity laws and the learned weights.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package
from test_mlp_oracle import random_params

pytestmark = pytest.mark.gpu

WALLS = ["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"]


@pytest.fixture(scope="module")
def ns():
    return package()


def _steps(ns, case, W, math, calls, params=None):
    s = ns.Solver(case, math=math)
    if params is not None:
        s.set_mlp(params)
    s.set_state(W, "primitive")
    t = 0.0
    for n in calls:
        t, dt = s.step(n, time=t)
    out = (po.interior(s.get_state("primitive")), po.interior(s.get_state("conservative")), t, dt)
    s.close()
    return out


@pytest.mark.parametrize("math", ["exact", "fast", "fp32"])
@pytest.mark.parametrize("conv,bc,viscosity", [("weno5cw", "periodic", "sutherland"),
                                               ("weno5", "walls", "constant")])
def test_one_call_equals_single_steps(ns, math, conv, bc, viscosity):
    print(f"[gpu] prologue formed by the last stage: {math} {conv} {bc} {viscosity}", flush=True)
    C = cases_module()
    case = dict(C.base_case(40, 24, 36), conv=conv, viscosity=viscosity,
                bc=(["periodic"] * 6 if bc == "periodic" else WALLS))
    W = C.random_smooth_W(case, seed=37)
    a = _steps(ns, case, W, math, (4,))
    b = _steps(ns, case, W, math, (1, 1, 1, 1))
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2] and a[3] == b[3]


@pytest.mark.parametrize("math", ["exact", "fp32"])
def test_one_call_equals_single_steps_learned(ns, math):
    print(f"[gpu] prologue formed by the last stage: learned {math}", flush=True)
    C = cases_module()
    case = dict(C.base_case(40, 24, 36), conv="weno5", weights="learned")
    W = C.random_smooth_W(case, seed=43)
    params = random_params(47)
    a = _steps(ns, case, W, math, (3,), params)
    b = _steps(ns, case, W, math, (1, 2), params)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2] and a[3] == b[3]
