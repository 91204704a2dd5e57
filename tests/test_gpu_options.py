"""GPU parity of the scheme options without a reference counterpart: the per-face
Lax-Friedrichs constant (alpha_mode "local_max"), the corrected characteristic projection
(physics "corrected_convection") and the full 3-D viscous terms (physics "corrected_viscous").
The oracle is pinned for all three in test_options_oracle.py;
here the kernels must match it bitwise in EXACT math (RHS and whole SSP-RK3 steps, tile and
last-plane kernels, walls and periodic axes), to round-off in FAST / FP32, and ensemble members
must equal the same blocks stepped alone.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package

pytestmark = pytest.mark.gpu

OPTIONS = [
    ("weno5", "local_max", "reference"),
    ("weno5cw", "local_max", "reference"),
    ("weno5cw", "const3", "corrected_convection"),
    ("weno5cw", "local_max", "corrected_convection"),
]


@pytest.fixture(scope="module")
def ns():
    return package()


def option_case(conv, alpha, physics, bc=None, order=4, weights="shu"):
    C = cases_module()
    case = dict(C.base_case(19, 13, 11), conv=conv, weights=weights, diff_order=order,
                L=(1.0, 0.8, 0.7), alpha_mode=alpha, physics=physics)
    if bc is not None:
        case["bc"] = bc
    return case


def rel_err(a, b):
    a = np.asarray(a).reshape(-1, 5)
    b = np.asarray(b).reshape(-1, 5)
    den = np.abs(b).max(axis=0)
    den[den == 0] = 1.0
    return np.abs(a - b).max(axis=0) / den


@pytest.mark.parametrize("conv,alpha,physics", OPTIONS)
@pytest.mark.parametrize("bc", [["periodic"] * 6, ["neumann"] * 6])
def test_option_rhs_matches_oracle_bitwise(ns, oracle, conv, alpha, physics, bc):
    case = option_case(conv, alpha, physics, bc)
    C = cases_module()
    W = C.random_smooth_W(case, 41)
    U = oracle.p2c(case, W)
    print(f"[gpu] option rhs {conv} {alpha} {physics} {bc[0]}", flush=True)
    s = ns.Solver(case, math="exact")
    for which in (1, 3):
        _, _, R = s.rhs(U, which)
        _, _, Ro = oracle.rhs(case, U, which)
        np.testing.assert_array_equal(po.interior(R), po.interior(Ro))
    s.close()


@pytest.mark.parametrize("conv,alpha,physics", OPTIONS)
def test_option_steps_match_oracle_bitwise(ns, oracle, conv, alpha, physics):
    case = option_case(conv, alpha, physics,
                       bc=["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"])
    C = cases_module()
    W0 = C.random_smooth_W(case, 42)
    print(f"[gpu] option steps {conv} {alpha} {physics}", flush=True)
    Wo, _, dto = oracle.step(case, W0, 5)
    s = ns.Solver(case, math="exact")
    s.set_state(W0, "primitive")
    t, dt = s.step(5)
    assert dt == dto[-1]
    np.testing.assert_array_equal(po.interior(s.get_state("primitive")), po.interior(Wo))
    _, _, flags, _ = s.reduce()
    if alpha == "local_max":
        assert not (flags & 4)  # no alpha flag with the per-face constant
    s.close()


@pytest.mark.parametrize("math,tol", [("fast", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("conv,alpha,physics", OPTIONS)
def test_option_rhs_other_maths(ns, oracle, math, tol, conv, alpha, physics):
    case = option_case(conv, alpha, physics)
    C = cases_module()
    W = C.random_smooth_W(case, 43)
    U = oracle.p2c(case, W)
    print(f"[gpu] option rhs {math} {conv} {alpha} {physics}", flush=True)
    s = ns.Solver(case, math=math)
    _, _, R = s.rhs(U, 3)
    _, _, Ro = oracle.rhs(case, U, 3)
    e = rel_err(po.interior(R), po.interior(Ro))
    assert e.max() < tol, e
    s.close()


def test_option_ensemble_members_equal_solo_runs(ns):
    C = cases_module()
    members = []
    for b, (n, M) in enumerate([((16, 24, 32), 0.3), ((24, 16, 40), 0.5), ((20, 36, 24), 0.4)]):
        case = dict(C.base_case(*n), conv="weno5cw", weights="shu", diff_order=4,
                    alpha_mode="local_max", physics="corrected_convection")
        members.append((case, C.shear_W(case, M=M, amp=0.05, seed=b + 1)))
    print("[gpu] option ensemble", flush=True)
    solo = []
    for case, W in members:
        s = ns.Solver(case, math="exact")
        s.set_state(W, "primitive")
        s.step(3)
        solo.append(po.interior(s.get_state("primitive")))
        s.close()
    solvers = []
    for case, W in members:
        s = ns.Solver(case, math="exact")
        s.set_state(W, "primitive")
        solvers.append(s)
    ens = ns.Ensemble(solvers)
    ens.tiles = np.array([[32, 16, 8], [16, 32, 3], [64, 64, 16]])
    ens.step(3)
    for s, ref in zip(solvers, solo):
        np.testing.assert_array_equal(po.interior(s.get_state("primitive")), ref)
    ens.close()
    for s in solvers:
        s.close()
    # members must share the options
    a = ns.Solver(members[0][0], math="exact")
    b = ns.Solver(dict(members[1][0], physics="reference"), math="exact")
    with pytest.raises(ns.NsgpuError):
        ns.Ensemble([a, b])
    a.close()
    b.close()


def test_local_alpha_runs_unit_sound_speed_shear_layer(ns, oracle):
    # the case the fixed alpha = 3 blows up (3 * CFL * alpha / max|u +- c| near 4) runs clean
    # with the per-face constant, and equals the oracle bitwise
    C = cases_module()
    case = dict(C.base_case(32, 32, 32), conv="weno5", weights="shu", diff_order=4,
                mu=0.3 / 1000.0, alpha_mode="local_max")
    W0 = C.shear_W(case, M=0.3, amp=0.05, seed=1, c0=1.0)
    print("[gpu] local alpha c0=1 shear layer", flush=True)
    Wo, _, _ = oracle.step(case, W0, 40)
    s = ns.Solver(case, math="exact")
    s.set_state(W0, "primitive")
    s.step(40)
    np.testing.assert_array_equal(po.interior(s.get_state("primitive")), po.interior(Wo))
    _, _, flags, rc = s.reduce()
    assert rc == 0 and flags == 0
    s.close()


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("viscosity", ["constant", "sutherland"])
@pytest.mark.parametrize("bc", [["periodic"] * 6,
                                ["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"]])
def test_full_viscous_rhs_matches_oracle_bitwise(ns, oracle, order, viscosity, bc):
    C = cases_module()
    case = option_case("weno5cw", "const3", "corrected_viscous", bc=bc, order=order)
    case["mu"] = 0.02
    if viscosity == "sutherland":
        case = C.sutherland(case, 0.02, 1.0)
    W = C.random_smooth_W(case, 44)
    U = oracle.p2c(case, W)
    print(f"[gpu] full viscous rhs cd{order} {viscosity} {bc[0]}/{bc[2]}", flush=True)
    s = ns.Solver(case, math="exact")
    for which in (2, 3):
        _, _, R = s.rhs(U, which)
        _, _, Ro = oracle.rhs(case, U, which)
        np.testing.assert_array_equal(po.interior(R), po.interior(Ro))
    s.close()


@pytest.mark.parametrize("math,tol", [("exact", 0.0), ("fast", 1e-10), ("fp32", 1e-4)])
def test_corrected_physics_steps_match_oracle(ns, oracle, math, tol):
    # every correction and the per-face constant together, walls on x and z
    C = cases_module()
    case = option_case("weno5cw", "local_max", "corrected",
                       bc=["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"])
    case["mu"] = 0.01
    W0 = C.random_smooth_W(case, 45)
    print(f"[gpu] corrected physics steps {math}", flush=True)
    Wo, _, dto = oracle.step(case, W0, 5)
    s = ns.Solver(case, math=math)
    s.set_state(W0, "primitive")
    s.step(5)
    Wg = po.interior(s.get_state("primitive"))
    if tol == 0.0:
        np.testing.assert_array_equal(Wg, po.interior(Wo))
    else:
        e = rel_err(Wg, po.interior(Wo))
        assert e.max() < tol, e
    s.close()


def test_full_viscous_refused_by_ensembles(ns):
    C = cases_module()
    case = option_case("weno5", "const3", "corrected_viscous")
    a = ns.Solver(case, math="exact")
    b = ns.Solver(case, math="exact")
    with pytest.raises(ns.NsgpuError):
        ns.Ensemble([a, b])
    a.close()
    b.close()
