"""CPU checks of the training-data side of the learned weights (3dnssolver_amd/learned.py):
features and regimes against the oracle, the face target against the closed form of the function
whose cell averages are the point values, the split fluxes against the Euler flux definition, and
the initial network against the linear fifth-order scheme."""
import importlib

import numpy as np
import pytest

import pyoracle as po
from test_mlp_oracle import stencils


@pytest.fixture(scope="module")
def lr():
    return importlib.import_module("3dnssolver_amd.learned")


def test_features_and_regimes_match_oracle(lr, oracle):
    V = stencils(51, 2000)
    X, Q, rg, m, r = lr.features(V)
    Xo = np.array([oracle.mlp_features(v) for v in V])
    np.testing.assert_allclose(X, Xo, rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(rg, [oracle.mlp_regime(x) for x in Xo])
    np.testing.assert_allclose(Q, [oracle.mlp_candidates(x) for x in Xo], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("m", [2, 3, 5])
def test_face_target_is_the_primitive_reconstruction(lr, m):
    # f = sin(kx) at fine points; the target at fine point p is h(x_p) with h the function whose
    # averages over coarse cells (width D = 2d) are f: h = f (kD/2) / sin(kD/2)
    for n in (64, 128):
        d = 1.0 / n
        x = (np.arange(n) + 0.5) * d
        k = 2 * np.pi * m
        f = np.sin(k * x)
        D = 2 * d
        h = f * (k * D / 2) / np.sin(k * D / 2)
        err = np.abs(lr.face_target(f, 0) - h).max()
        # sixth order: error / (k D)^6 bounded
        assert err < 0.05 * (k * D) ** 6 + 1e-14, (n, err)


def test_split_fluxes_definition(lr):
    rng = np.random.default_rng(3)
    Wi = np.stack([1 + 0.1 * rng.random((4, 5, 6)), 2 + 0.1 * rng.random((4, 5, 6)),
                   rng.standard_normal((4, 5, 6)) * 0.2, rng.standard_normal((4, 5, 6)) * 0.2,
                   rng.standard_normal((4, 5, 6)) * 0.2], axis=-1)
    g = 1.4
    (fLx, fRx), (fLy, fRy), _ = lr.split_fluxes(Wi, g)
    rho, p, u, v, w = (Wi[..., q] for q in range(5))
    E = p / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)
    # F_x = [rho u, u (E + p), rho u^2 + p, rho u v, rho u w]; fL - fR = 3 U
    np.testing.assert_allclose(fLx + fRx, [rho * u, u * (E + p), rho * u * u + p, rho * u * v,
                                           rho * u * w], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(fLy - fRy, 3.0 * np.stack([rho, E, rho * u, rho * v, rho * w]),
                               rtol=1e-13)


def test_initial_network_is_the_linear_scheme(lr, oracle):
    p = lr.init_params(0).astype(np.float64)
    p0 = p.copy()
    p0[1344:1440] = 0.0  # without the small output layer weights: exactly the ideal weights
    x = oracle.mlp_features(np.sin(1.5 * np.arange(5.0)))
    np.testing.assert_allclose(oracle.mlp_forward(p0, x), [0.1, 0.6, 0.3], rtol=1e-6)  # fp32 logs
    w = oracle.mlp_forward(p, x)
    assert np.abs(w - [0.1, 0.6, 0.3]).max() < 0.1


def test_blend_and_classic_match_oracle(lr, oracle):
    V = stencils(52, 3000) * np.random.default_rng(53).uniform(0.01, 100, (3000, 1))
    X, Q, rg, m, r = lr.features(V)
    ratio = lr.ratio_of(X)
    np.testing.assert_allclose(ratio, [oracle.mlp_ratio(x) for x in X], rtol=1e-12)
    np.testing.assert_allclose(lr.blend(ratio), [oracle.mlp_blend(q) for q in ratio],
                               rtol=1e-12, atol=1e-15)
    cn = lr.classic_normalised(X, Q, r)
    ref = (np.array([oracle.weno_left(v) for v in V]) - m) / r
    np.testing.assert_allclose(cn, ref, rtol=1e-9, atol=1e-9)
