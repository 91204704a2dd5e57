"""CPU checks of the learned-weights oracle (oracle/nsmlp.c).

Pinned independently of its own formulas: torch (CPU, float64) for the forward pass and
autograd for the gradient, central finite differences (h = 1e-6) for the backward pass, torch's
float32 -> bfloat16 conversion for the rounding helper, the textbook linear fifth-order upwind
reconstruction (2a - 13b + 47c + 27d - 3e) / 60 for a network that outputs the ideal weights,
and the classic WENO-5 reconstruction (bitwise) on smooth and flagged stencils.
"""
import numpy as np
import pytest

import pyoracle as po

torch = pytest.importorskip("torch")
F = torch.nn.functional

NP = po.MLP_NPARAM


def random_params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(NP) * scale
    return p


def stencils(seed, n):
    """Mixed stencils: smooth waves of random frequency, steps, kinks and noise."""
    rng = np.random.default_rng(seed)
    out = []
    x = np.arange(5.0)
    for i in range(n):
        kind = i % 4
        if kind == 0:
            k = rng.uniform(0.05, 1.5)
            v = np.sin(k * x + rng.uniform(0, 6.3)) * rng.uniform(0.1, 10) + rng.uniform(-5, 5)
        elif kind == 1:
            j = rng.integers(1, 5)
            v = np.where(x < j, 0.0, 1.0) * rng.uniform(-3, 3) + rng.uniform(-1, 1)
        elif kind == 2:
            v = np.abs(x - rng.uniform(0, 4)) * rng.uniform(0.1, 2)
        else:
            v = rng.standard_normal(5)
        out.append(v)
    return np.array(out)


def torch_net(params, X):
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    W1, b1 = p[0:256].view(32, 8), p[256:288]
    W2, b2 = p[288:1312].view(32, 32), p[1312:1344]
    W3, b3 = p[1344:1440].view(3, 32), p[1440:1443]
    x = torch.tensor(X, dtype=torch.float64)
    h1 = F.gelu(x @ W1.T + b1, approximate="tanh")
    h2 = F.gelu(h1 @ W2.T + b2, approximate="tanh")
    return p, torch.softmax(h2 @ W3.T + b3, dim=1)


@pytest.fixture(scope="module")
def data(oracle):
    V = stencils(11, 256)
    X = np.array([oracle.mlp_features(v) for v in V])
    Q = np.array([oracle.mlp_candidates(x) for x in X])
    T = np.random.default_rng(12).uniform(-0.2, 1.2, len(V))
    return V, X, Q, T


def test_forward_matches_torch(oracle, data):
    _, X, _, _ = data
    params = random_params(1)
    _, w = torch_net(params, X)
    ours = np.array([oracle.mlp_forward(params, x) for x in X])
    np.testing.assert_allclose(ours, w.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_gradient_matches_torch_autograd(oracle, data):
    _, X, Q, T = data
    params = random_params(2)
    p, w = torch_net(params, X)
    loss = (((w * torch.tensor(Q)).sum(1) - torch.tensor(T)) ** 2).mean()
    loss.backward()
    ours_loss, g = oracle.mlp_loss_grad(params, X, Q, T)
    assert ours_loss == pytest.approx(loss.item(), rel=1e-12)
    np.testing.assert_allclose(g, p.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_gradient_central_finite_differences(oracle, data):
    _, X, Q, T = data
    X, Q, T = X[:64], Q[:64], T[:64]
    params = random_params(3)
    _, g = oracle.mlp_loss_grad(params, X, Q, T)
    h = 1e-6
    idx = np.random.default_rng(4).choice(NP, 300, replace=False)
    idx = np.union1d(idx, [0, 255, 256, 287, 288, 1311, 1312, 1343, 1344, 1439, 1440, 1442])
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.mlp_loss_grad(pp, X, Q, T, grad=False)
              - oracle.mlp_loss_grad(pm, X, Q, T, grad=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max() + 1e-9, (i, fd, g[i])


def test_bf16_rounding_matches_torch(oracle):
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0]])
    ref = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    ours = np.array([oracle.bf16(x) for x in v])
    np.testing.assert_array_equal(ours, ref)


def test_softmax_weights_positive_and_consistent(oracle, data):
    _, X, _, _ = data
    params = random_params(6, scale=2.0)
    for x in X:
        for bf in (False, True):
            w = oracle.mlp_forward(params, x, bf16=bf)
            assert (w >= 0).all()
            assert abs(w.sum() - 1.0) < 1e-15


def test_regimes(oracle):
    lin = oracle.mlp_features(np.array([1.0, 2.0, 3.0, 4.0, 5.0]))
    assert oracle.mlp_regime(lin) == po.REGIMES["smooth"]
    const = oracle.mlp_features(np.full(5, 7.0))
    assert (const == 0).all() and oracle.mlp_regime(const) == po.REGIMES["smooth"]
    step = oracle.mlp_features(np.array([0.0, 0.0, 0.0, 1.0, 1.0]))
    assert oracle.mlp_regime(step) == po.REGIMES["flagged"]
    assert oracle.mlp_regime(oracle.mlp_features(np.sin(0.6 * np.arange(5.0)))) == 0  # resolved
    peak = oracle.mlp_features(np.sin(1.5 * np.arange(5.0)))  # ~4 points per wavelength
    assert oracle.mlp_regime(peak) == po.REGIMES["mlp"]
    # scale and offset invariance of the features
    v = np.array([0.3, -1.2, 2.5, 0.7, 1.1])
    np.testing.assert_allclose(oracle.mlp_features(5.0 * v + 3.0), oracle.mlp_features(v),
                               rtol=1e-12, atol=1e-14)


def test_learned_falls_back_to_classic_bitwise(oracle, data):
    V = data[0]
    params = random_params(7)
    seen = {0: 0, 2: 0}
    for v in V:
        for side, classic in ((0, oracle.weno_left), (1, oracle.weno_right)):
            val, rg = oracle.weno_learned(params, v, side)
            if rg != po.REGIMES["mlp"]:
                seen[rg] += 1
                assert val == classic(v), (v, side)
    assert seen[0] > 10 and seen[2] > 10


def test_right_biased_is_mirrored_left(oracle, data):
    params = random_params(8)
    n = 0
    for v in data[0]:
        a, ra = oracle.weno_learned(params, v, 1)
        b, rb = oracle.weno_learned(params, v[::-1].copy(), 0)
        assert ra == rb
        if ra == po.REGIMES["mlp"]:  # (the classic parts differ in operation order)
            assert a == pytest.approx(b, rel=1e-13, abs=1e-13 * np.abs(v).max())
            n += 1
    assert n > 20


def test_blend_is_continuous_between_regimes(oracle):
    b = oracle.mlp_blend
    assert b(2.0) == 0.0 and b(3.0) == 0.0 and b(4.5) == 0.5 and b(6.0) == 1.0
    assert b(250.0) == 1.0 and b(625.0) == 0.5 and b(1000.0) == 0.0 and b(5000.0) == 0.0
    r = np.concatenate([np.linspace(2.9, 7, 2000), np.linspace(240, 1010, 2000)])
    v = np.array([b(x) for x in r])
    assert np.abs(np.diff(v)).max() < 2e-3  # no jump


def test_ideal_weight_network_gives_linear_upwind_scheme(oracle, data):
    # W = 0 and b3 = log(d) make the softmax output the ideal weights (0.1, 0.6, 0.3): where the
    # network has the whole share the reconstruction is the linear fifth-order upwind formula,
    # elsewhere in its regime the blend of that formula and the classic value
    params = np.zeros(NP)
    params[1440:1443] = np.log([0.1, 0.6, 0.3])
    full = part = 0
    for v in data[0]:
        val, rg = oracle.weno_learned(params, v, 0, bf16=False)
        if rg != po.REGIMES["mlp"]:
            continue
        a, b, c, d, e = v
        lin = (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60.0
        lam = oracle.mlp_blend(oracle.mlp_ratio(oracle.mlp_features(v)))
        tol = 1e-13 * np.abs(v).max()
        if lam == 1.0:
            assert val == pytest.approx(lin, rel=1e-13, abs=tol)
            full += 1
        else:
            cl = oracle.weno_left(v)
            assert val == pytest.approx(cl + lam * (lin - cl), rel=1e-13, abs=tol)
            part += 1
    assert full > 20 and part > 5


def test_roundoff_level_stencils_stay_classic(oracle):
    # a stencil that varies by less than 2^-13 of its magnitude (round-off, seeded noise) counts as
    # constant: smooth regime and the classic reconstruction bitwise, whatever shape the
    # normalised noise has; just above the floor the shape decides again
    rng = np.random.default_rng(61)
    params = random_params(37)
    oracle.set_mlp(np.asarray(params, dtype=np.float32).astype(np.float64), bf16=True)
    for scale in (1.0, 3.7, -250.0):
        for _ in range(50):
            v = scale * (1.0 + 1e-5 * rng.standard_normal(5))
            x = oracle.mlp_features(v)
            assert oracle.mlp_regime(x) == po.REGIMES["smooth"]
            for side in (0, 1):
                val, rg = oracle.weno_learned(params, v, side, bf16=True)
                assert rg == po.REGIMES["smooth"]
                classic = oracle.weno_left(v, "shu") if side == 0 else oracle.weno_right(v, "shu")
                assert val == classic
    # above the floor the normalised shape decides: a network-regime shape riding on a large
    # constant at 1e-3 relative variation keeps its regime and features
    shape = np.array([0.0, 0.1, 0.3, 0.35, 1.0])
    assert oracle.mlp_regime(oracle.mlp_features(shape)) == po.REGIMES["mlp"]
    v = 2.0 + 2e-3 * shape
    x = oracle.mlp_features(v)
    assert oracle.mlp_regime(x) == po.REGIMES["mlp"]
    np.testing.assert_allclose(x[:5], shape, atol=1e-9)


def test_momentum_scale_holds_noise_stencils_constant(oracle):
    # a momentum component that vanishes in exact arithmetic holds only round-off: its stencils,
    # measured against the face's momentum scale, count as constant whatever shape the noise
    # has; measured against their own magnitude alone they would look like structure
    rng = np.random.default_rng(67)
    shape = np.array([0.0, 0.1, 0.3, 0.35, 1.0])  # a network-regime shape
    for mag in (1e-12, 1e-9, 1e-6):
        v = mag * shape
        assert oracle.mlp_regime(oracle.mlp_features(v)) == po.REGIMES["mlp"]
        x = oracle.mlp_features_s(v, 1.0)
        assert not x.any() and oracle.mlp_regime(x) == po.REGIMES["smooth"]
    # the scale only ever adds the constant case: below the stencil's own magnitude it changes
    # nothing, and a range above 2^-13 of the scale keeps the stencil's own features
    for _ in range(200):
        v = rng.standard_normal(5) * 10.0 ** rng.uniform(-3, 3)
        own = np.abs(v).max()
        np.testing.assert_array_equal(oracle.mlp_features_s(v, own * rng.uniform(0, 1)),
                                      oracle.mlp_features(v))
        r = v.max() - v.min()
        x = oracle.mlp_features_s(v, r * 8192.0 * 0.999)
        np.testing.assert_array_equal(x, oracle.mlp_features(v))
        assert not oracle.mlp_features_s(v, r * 8192.0 * 1.001).any()
    # the face scale: the largest |split value| of the momentum components (positions 2..4) at
    # the face's two cells (fl cells 2, 3 and fr cells 1, 2 of the stencils)
    fl = rng.standard_normal((5, 5))
    fr = rng.standard_normal((5, 5))
    want = max(np.abs(fl[2:, 2:4]).max(), np.abs(fr[2:, 1:3]).max())
    assert oracle.mlp_face_scale(fl, fr) == want
    fl2, fr2 = fl.copy(), fr.copy()
    fl2[:2] = 1e6  # mass and energy do not enter
    fl2[2:, [0, 1, 4]] = 1e6  # nor the outer stencil cells
    fr2[2:, [0, 3, 4]] = 1e6
    assert oracle.mlp_face_scale(fl2, fr2) == want
