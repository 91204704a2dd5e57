"""Pins the oracle's two scheme options that have no counterpart in the reference's code path:

- alpha_mode "local_max": the Lax-Friedrichs constant of a face is the largest eigenvalue
  max(|q_n - c|, |q_n + c|) over the cells i-2 .. i+2 (the value convection_flux.cpp:235-238
  computes and leaves unused; get_max_alpha over the same stencil, weno-5cw.cpp:60-64);
- physics "corrected_convection": the characteristic projection with the true sound speed, the
  eigenvectors ordered like the stored conserved vector, and the Roe-averaged face state.

What fixes them: the eigen system must diagonalise the flux Jacobian (computed here by complex
step from the Euler flux), the Roe state must satisfy Roe's property A (U_R - U_L) = F_R - F_L,
the face constant is checked by brute force, and whole RHS evaluations are restated in numpy
from those pinned pieces.  A behavioural check: with the local constant the unit-sound-speed
shear layer that the constant 3.0 blows up stays bounded.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module

GAMMA = 1.4


def stored_flux(Us, d, g=GAMMA):
    """Euler flux along axis d of stored-order conserved vectors (last axis), complex-safe."""
    rho, rE, mom = Us[..., 0], Us[..., 1], Us[..., 2:5]
    with np.errstate(invalid="ignore", divide="ignore"):  # never-read edge cells hold zeros
        vel = mom / rho[..., None]
    p = (g - 1.0) * (rE - 0.5 * rho * (vel * vel).sum(-1))
    q = vel[..., d]
    with np.errstate(invalid="ignore", divide="ignore"):
        H = (rE + p) / rho
    F = np.empty_like(Us)
    F[..., 0] = rho * q
    F[..., 1] = rho * q * H
    for a in range(3):
        F[..., 2 + a] = rho * q * vel[..., a] + (p if a == d else 0.0)
    return F


def prim_to_stored(W, g=GAMMA):
    rho, p, u = W[0], W[1], np.asarray(W[2:5])
    return np.array([rho, p / (g - 1.0) + 0.5 * rho * (u @ u), rho * u[0], rho * u[1], rho * u[2]])


def jacobian(Ws, d, g=GAMMA):
    U = prim_to_stored(Ws, g).astype(complex)
    J = np.zeros((5, 5))
    h = 1e-30
    for k in range(5):
        e = U.copy()
        e[k] += 1j * h
        J[:, k] = stored_flux(e, d, g).imag / h
    return J


def random_states(rng, n):
    out = []
    for _ in range(n):
        rho = rng.uniform(0.5, 2.0)
        p = rng.uniform(0.3, 3.0)
        u = rng.uniform(-0.8, 0.8, 3)
        out.append(np.array([rho, p, *u]))
    return out


def case_with(**kw):
    C = cases_module()
    case = dict(C.base_case(6, 5, 7), conv="weno5cw", weights="shu", diff_order=0)
    case.update(kw)
    return case


def test_corrected_eigensystem_diagonalises_the_flux_jacobian(oracle):
    rng = np.random.default_rng(11)
    case = case_with(physics="corrected_convection")
    for Wl, Wr in zip(random_states(rng, 6), random_states(rng, 6)):
        for d in range(3):
            Wf, L, R = oracle.face_eigensystem(case, Wl, Wr, d)
            np.testing.assert_allclose(L @ R, np.eye(5), atol=1e-13)
            c = np.sqrt(GAMMA * Wf[1] / Wf[0])
            q = Wf[2 + d]
            lam = np.array([q - c, q, q, q + c, q])
            np.testing.assert_allclose(L @ jacobian(Wf, d) @ R, np.diag(lam), atol=1e-12)


def test_reference_eigensystem_is_not_the_jacobians(oracle):
    # the reference's eigenvectors (c^2 = gamma / (p rho), conserved order applied to the stored
    # order) do not diagonalise the Jacobian: the check above would catch either defect
    rng = np.random.default_rng(12)
    case = case_with()
    Wl, Wr = random_states(rng, 2)
    Wf, L, R = oracle.face_eigensystem(case, Wl, Wr, 0)
    np.testing.assert_allclose(Wf, 0.5 * (Wl + Wr), rtol=0, atol=0)
    Lr, Rr = oracle.eigenvectors(0.5 * (Wl + Wr), GAMMA, 0)
    np.testing.assert_array_equal(L, Lr)
    np.testing.assert_array_equal(R, Rr)
    np.testing.assert_allclose(L @ R, np.eye(5), atol=1e-13)  # still a basis
    M = L @ jacobian(Wf, 0) @ R
    assert np.abs(M - np.diag(np.diag(M))).max() > 1e-2


def test_roe_state_satisfies_roes_property(oracle):
    rng = np.random.default_rng(13)
    case = case_with(physics="corrected_convection")
    for Wl, Wr in zip(random_states(rng, 8), random_states(rng, 8)):
        for d in range(3):
            Wf, _, _ = oracle.face_eigensystem(case, Wl, Wr, d)
            Ul, Ur = prim_to_stored(Wl), prim_to_stored(Wr)
            lhs = jacobian(Wf, d) @ (Ur - Ul)
            rhs = stored_flux(Ur, d) - stored_flux(Ul, d)
            np.testing.assert_allclose(lhs, rhs, atol=1e-12 * np.abs(rhs).max())
    # equal states: the Roe state is the state itself
    W = random_states(rng, 1)[0]
    Wf, _, _ = oracle.face_eigensystem(case, W, W, 1)
    np.testing.assert_allclose(Wf, W, rtol=1e-14)


def filled_state(oracle, case, seed):
    C = cases_module()
    W = C.random_smooth_W(case, seed=seed)
    U0 = oracle.p2c(case, W)
    U, Wf, _ = oracle.rhs(case, U0, which=1)
    return U, Wf


@pytest.mark.parametrize("mode,expect3", [("const3", True), ("local_max", False)])
def test_face_alpha_brute_force(oracle, mode, expect3):
    case = case_with(alpha_mode=mode)
    _, W = filled_state(oracle, case, 21)
    rng = np.random.default_rng(5)
    for _ in range(40):
        d = int(rng.integers(3))
        idx = [int(rng.integers(3, 3 + n)) for n in (case["nx"], case["ny"], case["nz"])]
        idx[d] = int(rng.integers(2, 3 + (case["nx"], case["ny"], case["nz"])[d]))
        a = oracle.face_alpha(case, W, d, *idx)
        if expect3:
            assert a == 3.0
            continue
        eig = []
        for o in range(-2, 3):
            t = list(idx)
            t[d] += o
            w = W[tuple(t)]
            eig.append(abs(w[2 + d]) + np.sqrt(GAMMA * w[1] / w[0]))
        assert a == pytest.approx(max(eig), rel=1e-15)


# --- numpy restatement of the convection RHS from the pinned pieces -------------------------

def weno(a, b, c, d, e, right):
    if right:
        a, b, c, d, e = e, d, c, b, a  # the right-biased value is the mirrored left-biased one
    f0 = (2 * a - 7 * b + 11 * c) / 6
    f1 = (-b + 5 * c + 2 * d) / 6
    f2 = (2 * c + 5 * d - e) / 6
    b0 = 13 / 12 * (a - 2 * b + c) ** 2 + 0.25 * (a - 4 * b + 3 * c) ** 2
    b1 = 13 / 12 * (b - 2 * c + d) ** 2 + 0.25 * (b - d) ** 2
    b2 = 13 / 12 * (c - 2 * d + e) ** 2 + 0.25 * (3 * c - 4 * d + e) ** 2
    w = [0.1 / (1e-6 + b0) ** 2, 0.6 / (1e-6 + b1) ** 2, 0.3 / (1e-6 + b2) ** 2]
    s = w[0] + w[1] + w[2]
    return (w[0] * f0 + w[1] * f1 + w[2] * f2) / s


def numpy_convection(oracle, case, U, W):
    n = (case["nx"], case["ny"], case["nz"])
    dxs = [case["L"][a] / n[a] for a in range(3)]
    R = np.zeros_like(U)
    for d in range(3):
        f = np.zeros_like(U)
        F = stored_flux(U, d)
        ranges = [range(3, 3 + m) for m in n]
        ranges[d] = range(2, 3 + n[d])
        for i in ranges[0]:
            for j in ranges[1]:
                for k in ranges[2]:
                    base = [i, j, k]
                    cells = []
                    for o in range(-2, 4):
                        t = list(base)
                        t[d] += o
                        cells.append(tuple(t))
                    Fs = np.array([F[c] for c in cells])  # (6, 5)
                    Us = np.array([U[c] for c in cells])
                    al = oracle.face_alpha(case, W, d, i, j, k)
                    fl = (Fs[:5] + al * Us[:5]) * 0.5
                    fr = (Fs[1:] - al * Us[1:]) * 0.5
                    if case["conv"] == "weno5":
                        out = np.array([weno(*fl[:, q], False) + weno(*fr[:, q], True)
                                        for q in range(5)])
                    else:
                        _, L, Rm = oracle.face_eigensystem(case, W[cells[2]], W[cells[3]], d)
                        a, b = fl @ L.T, fr @ L.T  # (5 stencil, 5 characteristic)
                        ft = np.array([weno(*a[:, q], False) + weno(*b[:, q], True)
                                       for q in range(5)])
                        out = Rm @ ft
                    f[i, j, k] = out
        s = [0, 0, 0]
        s[d] = 1
        I = tuple(slice(3, 3 + m) for m in n)
        Im = tuple(slice(3 - s[a], 3 + n[a] - s[a]) for a in range(3))
        R[I] += (f[I] - f[Im]) / dxs[d]
    return R


@pytest.mark.parametrize("conv,alpha,physics", [
    ("weno5", "local_max", "reference"),
    ("weno5cw", "const3", "corrected_convection"),
    ("weno5cw", "local_max", "corrected_convection"),
    ("weno5cw", "local_max", "reference"),
])
def test_convection_rhs_matches_numpy_restatement(oracle, conv, alpha, physics):
    case = case_with(conv=conv, alpha_mode=alpha, physics=physics)
    U, W = filled_state(oracle, case, 22)
    _, _, Ro = oracle.rhs(case, U, which=1)
    Rn = numpy_convection(oracle, case, U, W)
    a, b = po.interior(Ro), po.interior(Rn)
    assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()
    # and the option changes the result
    _, _, Rref = oracle.rhs(case_with(conv=conv), U, which=1)
    assert np.abs(po.interior(Rref) - a).max() > 1e-6 * np.abs(a).max()


def test_local_alpha_stabilises_unit_sound_speed_shear_layer(oracle):
    # with c0 = 1 the fixed alpha = 3 makes 3 * CFL * alpha / max|u +- c| about 4 at CFL 0.5:
    # the reference scheme blows up within a few steps; the local constant keeps it near 1.5
    C = cases_module()
    case = dict(C.base_case(16, 16, 16), conv="weno5", weights="shu", diff_order=4,
                mu=0.3 * 1.0 / 1000.0)
    W0 = C.shear_W(case, M=0.3, amp=0.05, seed=1, c0=1.0)
    Wc, _, _ = oracle.step(case, W0, 30)
    Wl, _, dts = oracle.step(dict(case, alpha_mode="local_max"), W0, 30)
    bad = ~np.isfinite(po.interior(Wc)).all() or po.interior(Wc)[..., 1].min() <= 0 or \
        np.abs(po.interior(Wc)[..., 2]).max() > 10.0
    assert bad
    I = po.interior(Wl)
    assert np.isfinite(I).all() and I[..., 0].min() > 0 and I[..., 1].min() > 0
    assert np.abs(I[..., 2]).max() < 0.5
    U0, U1 = oracle.p2c(case, W0), oracle.p2c(case, Wl)
    np.testing.assert_allclose(po.interior(U1)[..., 0].sum(), po.interior(U0)[..., 0].sum(),
                               rtol=1e-12)


# --- full 3-D viscous terms (physics "corrected_viscous") ------------------------------------

def smooth_periodic_W(case):
    """Band-limited 3-D fields on the periodic block (cell centres), halos left to the oracle."""
    C = cases_module()
    n = (case["nx"], case["ny"], case["nz"])
    L = case["L"]
    x, y, z = (2 * np.pi * (np.arange(n[a]) + 0.5) / n[a] for a in range(3))
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    W = np.zeros(po.padded_shape(case))
    I = (slice(3, -3),) * 3
    W[I + (0,)] = 1.0 + 0.2 * np.sin(X + Y) + 0.1 * np.cos(Z - X)
    W[I + (1,)] = 1.5 + 0.3 * np.cos(Y + 2 * Z) + 0.1 * np.sin(X)
    W[I + (2,)] = 0.3 * np.sin(Y) * np.cos(Z) + 0.1 * np.cos(X + Z)
    W[I + (3,)] = 0.2 * np.cos(X) * np.sin(Z) - 0.1 * np.sin(2 * Y)
    W[I + (4,)] = 0.25 * np.sin(X - Y) + 0.1 * np.cos(Z)
    return W


def spectral_viscous(case, W):
    """The viscous and heat-flux terms in conservative form, every derivative spectral (exact
    for these resolved periodic fields up to round-off): the independent statement of what the
    corrected operator discretises.  Returns R (interior, stored order) with the reference's
    sign convention (U_new = U - dt R)."""
    n = (case["nx"], case["ny"], case["nz"])
    L = case["L"]
    g, Pr = case["gamma"], case["Pr"]
    I = (slice(3, -3),) * 3
    rho, p = W[I + (0,)], W[I + (1,)]
    vel = [W[I + (2 + a,)] for a in range(3)]
    k = [np.fft.fftfreq(n[a], L[a] / n[a]) * 2 * np.pi for a in range(3)]
    K = np.meshgrid(*k, indexing="ij")

    def d(f, a):
        return np.real(np.fft.ifftn(1j * K[a] * np.fft.fftn(f)))

    T = p / rho
    if case.get("viscosity", "constant") == "sutherland":
        Tt = g * p / rho
        C = case["suth_C"]
        mu = case["mu"] * (1 + C) / (Tt + C) * Tt ** 1.5
    else:
        mu = case["mu"] * np.ones_like(rho)
    grad = [[d(vel[a], b) for b in range(3)] for a in range(3)]
    div = grad[0][0] + grad[1][1] + grad[2][2]
    tau = [[mu * (grad[a][b] + grad[b][a]) + (a == b) * (-2.0 / 3.0) * mu * div for b in range(3)]
           for a in range(3)]
    heat = g / (Pr * (g - 1.0))
    R = np.zeros(rho.shape + (5,))
    for a in range(3):
        R[..., 2 + a] = -sum(d(tau[a][b], b) for b in range(3))
    work = sum(d(sum(vel[a] * tau[a][b] for a in range(3)), b) for b in range(3))
    qdiv = sum(d(-heat * mu * d(T, b), b) for b in range(3))
    R[..., 1] = -work + qdiv
    return R


def viscous_case(n, order, viscosity="constant", physics="corrected_viscous"):
    C = cases_module()
    case = dict(C.base_case(*n), L=(1.0, 0.9, 0.8), conv="none", weights="shu",
                diff_order=order, mu=0.05, physics=physics)
    if viscosity == "sutherland":
        case = C.sutherland(case, 0.05, 1.4 * 1.5)
    return case


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("viscosity", ["constant", "sutherland"])
def test_corrected_viscous_converges_to_spectral_operator(oracle, order, viscosity):
    # per component (energy, three momenta): max error over max value; halving the cells
    # divides it by 2^order (measured 3.9-4.0, 15.1-15.9, 51-63 for CD2, CD4, CD6)
    errs = []
    for n in ((24, 20, 28), (48, 40, 56)):
        case = viscous_case(n, order, viscosity)
        W = smooth_periodic_W(case)
        _, _, R = oracle.rhs(case, oracle.p2c(case, W), which=2)
        ref = spectral_viscous(case, W)
        d = np.abs(po.interior(R) - ref)
        errs.append(np.array([d[..., q].max() / np.abs(ref[..., q]).max() for q in range(1, 5)]))
    ratio = errs[0] / errs[1]
    assert (ratio > 0.75 * 2 ** order).all(), (errs, ratio)
    assert errs[1].max() < {2: 8e-3, 4: 1.5e-4, 6: 5e-6}[order], errs
    # mass gets no viscous term
    assert np.abs(po.interior(R)[..., 0]).max() == 0.0


def test_reference_viscous_terms_do_not_converge_in_3d(oracle):
    # the reference's 2-D subset with its quirks misses the 3-D operator by O(1): the test above
    # has teeth
    errs = []
    for n in ((12, 10, 14), (24, 20, 28)):
        case = viscous_case(n, 4, physics="reference")
        W = smooth_periodic_W(case)
        _, _, R = oracle.rhs(case, oracle.p2c(case, W), which=2)
        ref = spectral_viscous(case, W)
        errs.append(np.abs(po.interior(R) - ref)[..., 1:].max() / np.abs(ref[..., 1:]).max())
    assert min(errs) > 0.1, errs


def test_corrected_viscous_conserves_momentum_on_periodic_block(oracle):
    # the momentum terms are divergences of tau: their sum over a periodic block is round-off
    # for the exactly differentiated part; the expanded discrete form leaves an O(h^order) sum
    case = viscous_case((24, 20, 28), 4)
    W = smooth_periodic_W(case)
    _, _, R = oracle.rhs(case, oracle.p2c(case, W), which=2)
    Ri = po.interior(R)
    for a in range(3):
        assert abs(Ri[..., 2 + a].sum()) <= 1e-4 * np.abs(Ri[..., 2 + a]).sum()
