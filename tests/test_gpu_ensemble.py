"""Ensembles of uneven blocks (batched launches, one hipGraph per SSP-RK3 step, autotuned tiles).

Each member's result must be bitwise the result of the same block stepped alone through
Solver.step, with the same kernels (the single-block launches), for every tile shape the
autotuner may choose, with and without the graph.  The ensemble path itself is pinned to the
oracle through that identity: Solver.step is bitwise the oracle in EXACT math (test_gpu_parity).
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    return package()


def _members(conv="weno5cw", diff_order=4):
    """Uneven blocks with different physics and boundary kinds, one case each."""
    C = cases_module()
    specs = [
        ((16, 24, 32), ["periodic"] * 6, 0.3, 1),
        ((24, 16, 40), ["neumann"] * 6, 0.5, 2),
        ((32, 32, 16), ["periodic", "periodic", "neumann", "neumann", "periodic", "periodic"], 0.2, 3),
        ((20, 36, 24), ["periodic"] * 6, 0.4, 4),
        ((8, 12, 72), ["periodic", "periodic", "periodic", "periodic", "neumann", "neumann"], 0.6, 5),
    ]
    out = []
    for (nx, ny, nz), bc, M, seed in specs:
        case = dict(C.base_case(nx, ny, nz), conv=conv, diff_order=diff_order, bc=bc,
                    mu=2e-3 * seed, L=(1.0, 1.0 + 0.25 * seed, 1.5))
        W = C.shear_W(case, M=M, amp=0.05, seed=seed)
        out.append((case, W))
    return out


def _solo(ns, members, math, nsteps):
    res = []
    for case, W in members:
        s = ns.Solver(case, math=math)
        s.set_state(W, "primitive")
        t, dt = s.step(nsteps)
        res.append((po.interior(s.get_state("primitive")), po.interior(s.get_state("conservative")),
                    t, dt))
        s.close()
    return res


def _run_ensemble(ns, members, math, nsteps, tiles=None, graph=True, split=None):
    solvers = []
    for case, W in members:
        s = ns.Solver(case, math=math)
        s.set_state(W, "primitive")
        solvers.append(s)
    ens = ns.Ensemble(solvers)
    if tiles is not None:
        ens.tiles = tiles
    for chunk in (split or [nsteps]):
        ens.step(chunk, graph=graph)
    out = []
    for s in solvers:
        out.append((po.interior(s.get_state("primitive")), po.interior(s.get_state("conservative")),
                    s.dt))
    return ens, solvers, out


@pytest.mark.parametrize("math", ["exact", "fast", "fp32"])
@pytest.mark.parametrize("tiles", ["default", "narrow", "mixed"])
def test_ensemble_members_match_solo_runs_bitwise(ns, math, tiles):
    print(f"[gpu] ensemble vs solo math={math} tiles={tiles}", flush=True)
    members = _members()
    n = len(members)
    t = None
    if tiles == "narrow":
        t = np.tile([16, 16, 4], (n, 1))
    elif tiles == "mixed":
        t = np.array([[32, 16, 8], [16, 32, 3], [64, 32, 32], [32, 64, 1], [16, 16, 16]])
    solo = _solo(ns, members, math, 3)
    ens, solvers, got = _run_ensemble(ns, members, math, 3, tiles=t, graph=True, split=[1, 2])
    for b, ((W1, U1, _, dt1), (W2, U2, dt2)) in enumerate(zip(solo, got)):
        np.testing.assert_array_equal(W2, W1, err_msg=f"member {b} primitive")
        np.testing.assert_array_equal(U2, U1, err_msg=f"member {b} conservative")
        assert dt2 == dt1, (b, dt1, dt2)
    ens.close()
    for s in solvers:
        s.close()


def test_ensemble_without_graph_and_component_wise(ns):
    print("[gpu] ensemble weno5 no graph", flush=True)
    members = _members(conv="weno5", diff_order=6)
    solo = _solo(ns, members, "exact", 2)
    ens, solvers, got = _run_ensemble(ns, members, "exact", 2, graph=False)
    for (W1, U1, _, _), (W2, U2, _) in zip(solo, got):
        np.testing.assert_array_equal(W2, W1)
        np.testing.assert_array_equal(U2, U1)
    ens.close()
    for s in solvers:
        s.close()


def test_ensemble_interleaved_with_solo_steps(ns):
    """A member stepped alone between ensemble calls (which swaps its state buffers) is picked up
    by the next ensemble call."""
    print("[gpu] ensemble interleaved", flush=True)
    members = _members()[:3]
    solo = _solo(ns, members, "exact", 4)
    solvers = []
    for case, W in members:
        s = ns.Solver(case, math="exact")
        s.set_state(W, "primitive")
        solvers.append(s)
    ens = ns.Ensemble(solvers)
    ens.step(1)                         # every member: 1 step
    solvers[1].step(1)                  # member 1: 2 (its state buffers swap)
    ens.step(1)                         # 2, 3, 2
    solvers[0].step(1)
    solvers[2].step(1)                  # 3, 3, 3
    ens.step(1)                         # 4, 4, 4
    for b, s in enumerate(solvers):
        np.testing.assert_array_equal(po.interior(s.get_state("primitive")), solo[b][0])
        np.testing.assert_array_equal(po.interior(s.get_state("conservative")), solo[b][1])
    ens.close()
    for s in solvers:
        s.close()


def test_autotuner_restores_state_and_picks_valid_tiles(ns):
    print("[gpu] ensemble autotuner", flush=True)
    members = _members()
    solo = _solo(ns, members, "fp32", 2)
    solvers = []
    for case, W in members:
        s = ns.Solver(case, math="fp32")
        s.set_state(W, "primitive")
        solvers.append(s)
    ens = ns.Ensemble(solvers)
    ms0, ms1 = ens.tune(steps=2)
    assert ms0 > 0 and 0 < ms1 <= ms0 * 1.0001
    t = ens.tiles
    assert set(t[:, 0]) <= {16, 32, 64} and set(t[:, 1]) <= {16, 32, 64}
    assert (t[:, 2] >= 1).all()
    # tuning left the states where they were: two more steps give the solo result
    ens.step(2)
    for (W1, U1, _, _), s in zip(solo, solvers):
        np.testing.assert_array_equal(po.interior(s.get_state("primitive")), W1)
        np.testing.assert_array_equal(po.interior(s.get_state("conservative")), U1)
    ens.close()
    for s in solvers:
        s.close()


def test_ensemble_rejects_mixed_options(ns):
    C = cases_module()
    a = ns.Solver(C.base_case(8, 8, 8), math="exact")
    b = ns.Solver(C.base_case(8, 8, 8), math="fp32")
    c = ns.Solver(dict(C.base_case(8, 8, 8), diff_order=2), math="exact")
    for pair in ([a, b], [a, c], [a, a]):
        with pytest.raises(ns.NsgpuError):
            ns.Ensemble(pair)
    e = ns.Ensemble([a])
    with pytest.raises(ns.NsgpuError):
        e.tiles = [[48, 64, 16]]
    with pytest.raises(ns.NsgpuError):
        e.tiles = [[64, 64, 0]]
    e.close()
    for s in (a, b, c):
        s.close()
