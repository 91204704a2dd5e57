"""Contracts of the C ABI beyond parity: the step's health status, the prescribed-dt step, the
operator-level calls leaving the resident state alone, ensemble membership and the autotuner
restoring the members.  Failure states are planted from the host (a NaN, a negative density)
on a block with no convection and no diffusion, whose stages copy the state unchanged, so the
planted value reaches the stage's health check while every kernel runs its normal path.
"""

import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    return package()


def _still_case(C, n=12):
    """A periodic block without convection or diffusion: every stage returns its input."""
    return dict(C.base_case(n, n + 2, n + 4), conv="none", diff_order=0)


def test_step_reports_nonfinite_and_positivity(ns):
    C = cases_module()
    case = _still_case(C)
    W0 = C.tgv_W(case, M=0.1, seed=2)
    print("[gpu] step health: planted NaN / negative density", flush=True)
    s = ns.Solver(case)
    s.set_state(W0, "primitive")
    s.step(2)  # healthy: no error
    W = W0.copy()
    W[5, 6, 7, 0] = -0.5
    s.set_state(W, "primitive")
    with pytest.raises(ns.NsgpuError) as e:
        s.step(1)
    assert e.value.code == -5
    W[6, 7, 8, 3] = np.nan
    s.set_state(W, "primitive")
    with pytest.raises(ns.NsgpuError) as e:
        s.step(1)
    assert e.value.code == -4
    # the flags belong to the call: a healthy state steps cleanly again
    s.set_state(W0, "primitive")
    s.step(1)
    s.close()


def test_step_health_of_a_real_scheme(ns):
    # a NaN planted in a characteristic WENO-5 + CD4 run spreads through the stencils and the
    # step reports it; the same block without it steps cleanly
    C = cases_module()
    case, W0 = C.config2(16)
    s = ns.Solver(case)
    W = W0.copy()
    W[8, 8, 8, 1] = np.nan
    s.set_state(W, "primitive")
    with pytest.raises(ns.NsgpuError) as e:
        s.step(2)
    assert e.value.code == -4
    s.set_state(W0, "primitive")
    s.step(2)
    s.close()


def test_ensemble_step_reports_health(ns):
    C = cases_module()
    cases = [_still_case(C, 12), _still_case(C, 16)]
    Ws = [C.tgv_W(c, M=0.1, seed=2 + b) for b, c in enumerate(cases)]
    solvers = []
    for c, W in zip(cases, Ws):
        s = ns.Solver(c)
        s.set_state(W, "primitive")
        solvers.append(s)
    ens = ns.Ensemble(solvers)
    ens.step(2)
    bad = Ws[1].copy()
    bad[7, 7, 7, 0] = -1.0
    solvers[1].set_state(bad, "primitive")
    with pytest.raises(ns.NsgpuError) as e:
        ens.step(1)
    assert e.value.code == -5
    # nsgpu_reduce names the member
    assert solvers[0].reduce()[3] in (0, -6) and solvers[1].reduce()[3] == -5
    ens.close()
    for s in solvers:
        s.close()


def test_step_fixed_uses_the_prescribed_dt(ns):
    # every stage and the time use dt itself: the time is n * dt summed in order, and the run
    # does not depend on the CFL of the configuration
    C = cases_module()
    case, W0 = C.config2(16)
    dt = 0.0123456789
    runs = []
    for cfl in (0.5, 0.3):
        s = ns.Solver(dict(case, cfl=cfl))
        s.set_state(W0, "primitive")
        t = s.step_fixed(5, dt)
        ref = 0.0
        for _ in range(5):
            ref = ref + dt
        assert t == ref
        assert s.dt == dt
        runs.append(s.get_state("primitive"))
        s.close()
    np.testing.assert_array_equal(po.interior(runs[0]), po.interior(runs[1]))
    # a CFL step after it reports the CFL dt again
    s = ns.Solver(case)
    s.set_state(W0, "primitive")
    s.step_fixed(1, dt)
    _, dt_cfl = s.step(1)
    assert dt_cfl != dt and dt_cfl == s.dt
    s.close()


def test_operator_calls_leave_the_resident_state(ns):
    C = cases_module()
    case, W0 = C.config2(16)
    s = ns.Solver(case)
    s.set_state(W0, "primitive")
    s.step(1)
    W1 = s.get_state("primitive")
    U1 = s.get_state("conservative")
    sums1 = s.reduce()[0]
    other = C.tgv_W(case, M=0.1, seed=7)
    U_other, _, _ = s.rhs(po.Oracle().p2c(case, other))
    s.convection_solve(other, U_other, np.zeros_like(other))
    s.diffusion_solve(other, np.zeros_like(other))
    s.apply_boundary(other)
    np.testing.assert_array_equal(s.get_state("primitive"), W1)
    np.testing.assert_array_equal(po.interior(s.get_state("conservative")), po.interior(U1))
    np.testing.assert_array_equal(s.reduce()[0], sums1)
    # and the next step advances the resident state, as if the calls had not happened
    s.step(1)
    W2 = s.get_state("primitive")
    r = ns.Solver(case)
    r.set_state(W0, "primitive")
    r.step(2)
    np.testing.assert_array_equal(po.interior(W2), po.interior(r.get_state("primitive")))
    r.close()
    s.close()


def test_enrolled_member_is_not_destroyed(ns):
    C = cases_module()
    members = []
    for n in (12, 16):
        case, W = C.config2(n)
        s = ns.Solver(case)
        s.set_state(W, "primitive")
        members.append(s)
    ens = ns.Ensemble(members)
    with pytest.raises(RuntimeError):
        members[0].close()
    # the C ABI refuses as well (prints an error, frees nothing): the ensemble still runs
    ns.load_library().nsgpu_destroy(members[1].h)
    ens.step(2)
    assert np.isfinite(po.interior(members[1].get_state("primitive"))).all()
    ens.close()
    for s in members:
        s.close()  # allowed once the ensemble is gone


def test_tune_leaves_members_as_they_were(ns):
    C = cases_module()
    members = []
    for n in (12, 16, 20):
        case, W = C.config2(n)
        s = ns.Solver(case)
        s.set_state(W, "primitive")
        members.append(s)
    before = []
    for s in members:
        t, dt = s.step(1)
        before.append((s.get_state("primitive"), po.interior(s.get_state("conservative")), s.dt))
    ens = ns.Ensemble(members)
    ens.tune(steps=1)
    for s, (W, U, dt) in zip(members, before):
        np.testing.assert_array_equal(s.get_state("primitive"), W)
        np.testing.assert_array_equal(po.interior(s.get_state("conservative")), U)
        assert s.dt == dt
    ens.close()
    for s in members:
        s.close()
