"""GPU checks of the learned WENO-5 weights: the BF16-MFMA network kernels against the FP64
oracle (oracle/nsmlp.c), with the oracle rounding the operands of every matrix product to
bfloat16 where the kernels do (so the remaining differences are FP32 accumulation and the FP32
GELU / softmax), and against the plain FP64 network within bfloat16 tolerance.

The face kernels' learned reconstruction is checked through nsgpu_mlp_recon, which runs the same
device function: smooth and flagged stencils must give the classic reconstruction bitwise (EXACT
math).  Learned solver steps are compared with the oracle's learned RHS, and learned ensemble
members with the same blocks stepped alone.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package
from test_mlp_oracle import random_params, stencils

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    return package()


@pytest.fixture(scope="module")
def samples(oracle):
    V = stencils(21, 4096)
    X = np.array([oracle.mlp_features(v) for v in V])
    Q = np.array([oracle.mlp_candidates(x) for x in X])
    # target: the linear fifth-order upwind value (reachable with the ideal weights), normalised
    a, b, c, d, e = V.T
    t = (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60.0
    m, M = V.min(axis=1), V.max(axis=1)
    r = np.where(M > m, M - m, 1.0)
    T = (t - m) / r
    return V, X, Q, T


def test_forward_matches_oracle(ns, oracle, samples):
    print("[gpu] mlp forward", flush=True)
    _, X, _, _ = samples
    params = random_params(31)
    w = ns.mlp_forward(params, X).astype(np.float64)
    bf = np.array([oracle.mlp_forward(params, x, bf16=True) for x in X])
    f64 = np.array([oracle.mlp_forward(params, x, bf16=False) for x in X])
    d = np.abs(w - bf)
    assert d.max() < 5e-3 and d.mean() < 2e-4, (d.max(), d.mean())
    assert np.abs(w - f64).max() < 3e-2
    np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-6)


def test_gradient_matches_oracle(ns, oracle, samples):
    print("[gpu] mlp gradient", flush=True)
    _, X, Q, T = samples
    params = random_params(32)
    loss, g = ns.mlp_loss_grad(params, X, Q, T)
    lbf, gbf = oracle.mlp_loss_grad(params, X, Q, T, bf16=True)
    l64, g64 = oracle.mlp_loss_grad(params, X, Q, T, bf16=False)
    assert loss == pytest.approx(lbf, rel=2e-3)
    assert loss == pytest.approx(l64, rel=2e-2)
    g = g.astype(np.float64)
    assert np.linalg.norm(g - gbf) / np.linalg.norm(gbf) < 2e-2
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2
    cos = g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64))
    assert cos > 0.999, cos


def test_training_reduces_loss(ns, samples):
    print("[gpu] mlp training", flush=True)
    _, X, Q, T = samples
    params = random_params(33, scale=0.3)
    p, hist = ns.mlp_train(params, X, Q, T, steps=400, batch=4096, lr=3e-3)
    assert np.isfinite(hist).all()
    loss0, _ = ns.mlp_loss_grad(params, X, Q, T)
    loss1, _ = ns.mlp_loss_grad(p, X, Q, T)
    assert hist[0] == pytest.approx(loss0, rel=1e-4)
    assert loss1 < 0.1 * loss0, (loss0, loss1)


@pytest.mark.parametrize("math", ["exact", "fp32"])
def test_kernel_reconstruction_matches_oracle(ns, oracle, math):
    print(f"[gpu] learned reconstruction {math}", flush=True)
    V = stencils(41, 4096)
    params = random_params(34)
    for side in (0, 1):
        out, rg = ns.mlp_recon(params, V, side, math=math)
        ref = [oracle.weno_learned(params, v, side, bf16=True) for v in V]
        rv = np.array([r[0] for r in ref])
        rr = np.array([r[1] for r in ref])
        same = rg == rr
        assert same.mean() > 0.995
        classic = same & (rr != po.REGIMES["mlp"])
        net = same & (rr == po.REGIMES["mlp"])
        assert classic.sum() > 500 and net.sum() > 500
        scale = V.max(axis=1) - V.min(axis=1)
        if math == "exact":
            np.testing.assert_array_equal(out[classic], rv[classic])
        else:
            assert (np.abs(out - rv)[classic] <= 1e-5 * (np.abs(V).max(axis=1)[classic] + 1)).all()
        err = np.abs(out - rv)[net] / scale[net]
        assert err.max() < 2e-2 and err.mean() < 1e-3, (err.max(), err.mean())


def _learned_case(conv):
    C = cases_module()
    case = dict(C.base_case(12, 10, 8), conv=conv, weights="learned", diff_order=4)
    W = C.random_smooth_W(case, seed=7)
    return case, W


@pytest.mark.parametrize("conv", ["weno5", "weno5cw"])
def test_learned_rhs_matches_oracle(ns, oracle, conv):
    print(f"[gpu] learned rhs {conv}", flush=True)
    case, W = _learned_case(conv)
    params = random_params(35)
    s = ns.Solver(case, math="exact")
    with pytest.raises(ns.NsgpuError):
        s.step(1)  # learned weights without a network
    s.set_mlp(params)
    U = oracle.p2c(case, W)
    _, _, R = s.rhs(U, 1)
    oracle.set_mlp(np.asarray(params, dtype=np.float32).astype(np.float64), bf16=True)
    _, _, Rref = oracle.rhs(case, U, 1)
    shu = ns.Solver(dict(case, weights="shu"), math="exact")
    _, _, Rshu = shu.rhs(U, 1)
    a, b, c = po.interior(R), po.interior(Rref), po.interior(Rshu)
    for q in range(5):
        den = np.abs(b[..., q]).max()
        assert np.abs(a[..., q] - b[..., q]).max() <= 2e-2 * den, q
    # the network changes the result
    assert np.abs(a - c).max() > 1e-6 * np.abs(c).max()
    s.close()
    shu.close()


@pytest.mark.parametrize("conv", ["weno5"])
def test_noise_level_momentum_stays_classic(ns, oracle, conv):
    # a spanwise velocity at round-off level (1e-12 of the flow's): every stencil of the rho w
    # split fluxes is constant against the face's momentum scale, so the learned RHS of rho w is
    # the classic one bitwise (EXACT math), as in the oracle, while the other components still
    # take the network
    # (a plane block: in a z sweep the rho w flux carries the pressure and is no noise)
    print(f"[gpu] learned rhs, noise-level w ({conv})", flush=True)
    C = cases_module()
    case = dict(C.base_case(12, 10, 1), dim=2, conv=conv, weights="learned", diff_order=4)
    W = C.random_smooth_W(case, seed=7)
    W[..., 4] = 1e-12 * np.random.default_rng(5).standard_normal(W.shape[:3])
    params = random_params(35)
    U = oracle.p2c(case, W)
    s = ns.Solver(case, math="exact")
    s.set_mlp(params)
    _, _, R = s.rhs(U, 1)
    shu = ns.Solver(dict(case, weights="shu"), math="exact")
    _, _, Rshu = shu.rhs(U, 1)
    oracle.set_mlp(np.asarray(params, dtype=np.float32).astype(np.float64), bf16=True)
    _, _, Rref = oracle.rhs(case, U, 1)
    a, b, c = po.interior(R), po.interior(Rref), po.interior(Rshu)
    np.testing.assert_array_equal(a[..., 4], c[..., 4])
    np.testing.assert_array_equal(b[..., 4], c[..., 4])
    assert np.abs(a[..., 2] - c[..., 2]).max() > 1e-6 * np.abs(c[..., 2]).max()
    for q in range(4):
        den = np.abs(b[..., q]).max()
        assert np.abs(a[..., q] - b[..., q]).max() <= 2e-2 * den, q
    s.close()
    shu.close()


@pytest.mark.parametrize("math", ["exact", "fp32"])
def test_network_pass_equals_in_kernel_network(ns, math, monkeypatch):
    # a single block's learned component-wise sweeps hand their network stencils to a separate
    # network pass; with NSGPU_LEARN_SINK=0 the face kernels evaluate the network themselves.
    # Both must give the same steps bitwise
    print(f"[gpu] network pass vs in-kernel network ({math})", flush=True)
    C = cases_module()
    case = dict(C.base_case(40, 24, 36), conv="weno5", weights="learned", diff_order=4)
    W = C.random_smooth_W(case, seed=9)
    params = random_params(38)
    out = {}
    for sink in ("0", "1"):
        monkeypatch.setenv("NSGPU_LEARN_SINK", sink)
        s = ns.Solver(case, math=math)
        s.set_mlp(params)
        s.set_state(W, "primitive")
        s.step(3)
        out[sink] = s.get_state("primitive")
        s.close()
    np.testing.assert_array_equal(po.interior(out["0"]), po.interior(out["1"]))


@pytest.mark.parametrize("math", ["exact", "fp32"])
def test_learned_ensemble_matches_solo(ns, math):
    print(f"[gpu] learned ensemble {math}", flush=True)
    C = cases_module()
    params = random_params(36)
    members = []
    for b, (n, M) in enumerate([((16, 24, 32), 0.3), ((24, 16, 40), 0.5), ((20, 36, 24), 0.4)]):
        case = dict(C.base_case(*n), conv="weno5", weights="learned", diff_order=4)
        members.append((case, C.shear_W(case, M=M, amp=0.05, seed=b + 1)))
    solo = []
    for case, W in members:
        s = ns.Solver(case, math=math)
        s.set_mlp(params)
        s.set_state(W, "primitive")
        s.step(3)
        solo.append(po.interior(s.get_state("primitive")))
        s.close()
    solvers = []
    for case, W in members:
        s = ns.Solver(case, math=math)
        s.set_mlp(params)
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


def test_trained_network_beats_classic_weno(ns):
    """Ablation: the shipped trained network against the method without it (classic Shu
    weights), both FP32 reconstruction, on the two evaluation problems (64^3, 100 steps with a
    prescribed dt) against a 128^3 FP64 Shu run at the same time.  Measured on MI355X: shear
    layer rho 1.25e-3 -> 0.84e-3, rhoE 1.70e-3 -> 1.43e-3; Sod-like jump rho 6.0e-2 -> 4.4e-2,
    rhoE 3.9e-2 -> 2.9e-2.  The bounds keep the gain at no less than 5 %."""
    print("[gpu] learned ablation", flush=True)
    import importlib
    lr = importlib.import_module("3dnssolver_amd.learned")
    res = lr.evaluate(lr.load(), n=64, steps=100, maths=("fp32",))
    for prob in ("shear", "jump"):
        r = res[prob]
        for q in ("rho", "rhoE"):
            assert r["fp32_learned"][q] < 0.95 * r["fp32_shu"][q], (prob, q, r)
