"""Pins the CPU oracle (oracle/nsoracle.c) before it is trusted as the checker of the GPU path.

1. Golden fixtures written by the reference solver itself (tests/golden/make_golden.py, strict
   build of the reference's own translation units): bitwise equality.
2. Live comparison with the reference build when oracle/_ref exists in this container.
3. What the mathematics fixes, chosen so that a plausible slip (a wrong linear weight or
   smoothness coefficient, a sign, a transposed eigenvector, a wrong stencil) fails a check.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import GOLDEN_RHS, cases_module, load_golden

I3 = (slice(3, -3),) * 3


# ---------------------------------------------------------------------------------------------
# 1. golden fixtures from the reference
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", GOLDEN_RHS)
def test_oracle_matches_reference_golden_rhs(oracle, name):
    case, g = load_golden(name)
    Uf, W, Rc = oracle.rhs(case, g["U_in"], which=1)
    _, _, Rd = oracle.rhs(case, g["U_in"], which=2)
    _, _, R = oracle.rhs(case, g["U_in"], which=3)
    m = po.faces_mask(case)
    np.testing.assert_array_equal(Uf[m], g["U_filled"][m])
    np.testing.assert_array_equal(W[m], g["W"][m])
    np.testing.assert_array_equal(po.interior(Rc), po.interior(g["R_conv"]))
    np.testing.assert_array_equal(po.interior(Rd), po.interior(g["R_diff"]))
    np.testing.assert_array_equal(po.interior(R), po.interior(g["R"]))


def test_oracle_matches_reference_golden_steps(oracle):
    case, g = load_golden("G7_steps_weno5cw_tgv16")
    W = g["W_in"]
    for s in range(g["W_steps"].shape[0]):
        W, _, dt = oracle.step(case, W, 1)
        assert dt[0] == g["dts"][s]
        np.testing.assert_array_equal(po.interior(W), g["W_steps"][s])


def test_oracle_multi_step_call_equals_single_steps(oracle):
    case, g = load_golden("G7_steps_weno5cw_tgv16")
    W, _, dts = oracle.step(case, g["W_in"], 4)
    np.testing.assert_array_equal(dts, g["dts"][:4])
    np.testing.assert_array_equal(po.interior(W), g["W_steps"][3])


# ---------------------------------------------------------------------------------------------
# 2. live reference (only in the build container, where oracle/_ref was compiled)
# ---------------------------------------------------------------------------------------------

@pytest.mark.skipif(not po.reference_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("conv", ["weno5", "weno5cw"])
@pytest.mark.parametrize("bc", [["periodic"] * 6, ["neumann"] * 6,
                                ["neumann", "neumann", "periodic", "periodic", "neumann", "neumann"]])
def test_oracle_matches_reference_live(oracle, conv, bc):
    C = cases_module()
    ref = po.Reference("strict")
    case = dict(C.base_case(9, 13, 7), conv=conv, weights="borges", diff_order=6, bc=bc)
    U = oracle.p2c(case, C.random_smooth_W(case, 5))
    Uo, Wo, Ro = oracle.rhs(case, U)
    Ur, Wr, Rr = ref.rhs(case, U)
    np.testing.assert_array_equal(po.interior(Ro), po.interior(Rr))
    W0 = C.random_smooth_W(case, 6)
    Wo, _, do = oracle.step(case, W0, 3)
    Wr, _, dr = ref.step(case, W0, 3)
    np.testing.assert_array_equal(do, dr)
    np.testing.assert_array_equal(po.interior(Wo), po.interior(Wr))


# ---------------------------------------------------------------------------------------------
# 3. mathematics
# ---------------------------------------------------------------------------------------------

def test_weno_linear_limit_is_textbook_upwind(oracle):
    # equal smoothness indicators (linear data) -> ideal weights -> 5th-order upwind formula
    f = np.array([0.3, 1.3, 2.3, 3.3, 4.3])
    a, b, c, d, e = f
    left = (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60
    right = (-3 * a + 27 * b + 47 * c - 13 * d + 2 * e) / 60
    for w in ("shu", "borges"):
        assert abs(oracle.weno_left(f, w) - left) < 1e-14
        assert abs(oracle.weno_right(f, w) - right) < 1e-14
    # small smooth data (beta << eps): Shu weights -> ideal weights
    x = np.arange(5) * 0.01
    g = 1e-6 * np.sin(x + 0.3)
    a, b, c, d, e = g
    assert abs(oracle.weno_left(g) - (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60) < 1e-18


def test_weno_right_is_mirrored_left(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        f = rng.standard_normal(5)
        for w in ("shu", "borges"):
            assert abs(oracle.weno_right(f, w) - oracle.weno_left(f[::-1], w)) <= 1e-14 * (1 + abs(f).max())


def test_weno_fifth_order_on_smooth_data(oracle):
    # reconstruction of the cell-average-free point values of sin at x_{i+1/2}: the WENO point
    # scheme is 5th order on smooth data once beta >> eps (large amplitude)
    errs = []
    for n in (40, 80, 160):
        h = 2 * np.pi / n
        e = 0.0
        for i in range(0, n, 7):
            xs = (i + np.arange(-2, 3)) * h
            val = oracle.weno_left(np.sin(xs) * 1e3, "shu") / 1e3
            # point-value WENO reconstructs the flux H with H' = f: compare with the exact
            # interface value of the 5th-order upwind primitive-function reconstruction
            a, b, c, d, ee = np.sin(xs)
            ref = (2 * a - 13 * b + 47 * c + 27 * d - 3 * ee) / 60
            e = max(e, abs(val - ref))
        errs.append(e)
    # nonlinear weights deviate from ideal by O(h^3) in smooth regions -> the deviation from the
    # linear scheme converges at >= 4th order
    order = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert min(order) > 3.5, (errs, order)


def test_weno_discontinuity_picks_smooth_stencil(oracle):
    # step between c and d: the left-biased reconstruction must stay on the left plateau
    f = np.array([1.0, 1.0, 1.0, 0.0, 0.0])
    assert abs(oracle.weno_left(f) - 1.0) < 1e-9
    assert abs(oracle.weno_left(f, "borges") - 1.0) < 1e-9


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_eigenvectors_are_inverse_pair(oracle, direction):
    rng = np.random.default_rng(direction)
    for _ in range(50):
        Wavg = np.array([1 + 0.5 * rng.random(), 0.5 + rng.random(), *rng.standard_normal(3) * 0.5])
        L, R = oracle.eigenvectors(Wavg, 1.4, direction)
        np.testing.assert_allclose(L @ R, np.eye(5), atol=1e-13)
        np.testing.assert_allclose(R @ L, np.eye(5), atol=1e-13)


def test_eigenvector_tangent_basis(oracle):
    # l, m of ConvectionFlux::orthogonal_unit_vectors for the axis normals (rows 1 and 4 of L
    # carry m and l: L[1] = [-qm, m], L[4] = [-ql, l])
    Wavg = np.array([1.0, 1.0, 0.0, 0.0, 0.0])
    expect = {0: ((0, -1, 0), (0, 0, 1)), 1: ((1, 0, 0), (0, 0, 1)), 2: ((1, 0, 0), (0, -1, 0))}
    for d, (m, l) in expect.items():
        L, _ = oracle.eigenvectors(Wavg, 1.4, d)
        np.testing.assert_array_equal(L[1, 1:4] + 0.0, np.array(m, float))
        np.testing.assert_array_equal(L[4, 1:4] + 0.0, np.array(l, float))


def test_eigenvector_sound_speed_quirk(oracle):
    # the reference uses c^2 = gamma / (p rho) (convection_flux.cpp:139): entry L[2][4] = -K/c2
    rho, p, g = 1.3, 0.7, 1.4
    L, R = oracle.eigenvectors(np.array([rho, p, 0.1, 0.2, 0.3]), g, 0)
    c2 = g / (p * rho)
    assert L[2, 4] == -(g - 1.0) / c2
    assert R[1, 0] == 0.1 - np.sqrt(c2)


def _case(**kw):
    C = cases_module()
    return dict(C.base_case(10, 9, 8), **kw)


@pytest.mark.parametrize("conv", ["weno5", "weno5cw"])
def test_free_stream_preservation(oracle, conv):
    case = _case(conv=conv, diff_order=4)
    W = np.zeros(po.padded_shape(case))
    W[I3] = [1.2, 2.5, 0.3, -0.2, 0.1]
    U = oracle.p2c(case, W)
    _, _, Rc = oracle.rhs(case, U, which=1)
    assert np.all(po.interior(Rc) == 0.0)
    _, _, Rd = oracle.rhs(case, U, which=2)
    assert np.abs(po.interior(Rd)).max() < 1e-12


@pytest.mark.parametrize("conv", ["weno5", "weno5cw"])
def test_periodic_conservation(oracle, conv):
    C = cases_module()
    case = _case(conv=conv, diff_order=4)
    U = oracle.p2c(case, C.random_smooth_W(case, 3))
    _, _, R = oracle.rhs(case, U)
    Ri = po.interior(R)
    scale = np.abs(Ri).sum(axis=(0, 1, 2))
    sums = Ri.sum(axis=(0, 1, 2))
    for q in (0, 2, 3, 4):  # mass and momentum telescope; energy does not (non-conservative form)
        assert abs(sums[q]) < 1e-12 * scale[q] + 1e-300


def test_axis_permutation_symmetry(oracle):
    # component-wise scheme without diffusion: swapping x<->y (and u<->v) permutes R
    C = cases_module()
    case = _case(conv="weno5", diff_order=0, L=(1.0, 1.0, 1.0))
    case["nx"] = case["ny"] = 10
    W = C.random_smooth_W(case, 4)
    Wt = W.transpose(1, 0, 2, 3)[..., [0, 1, 3, 2, 4]].copy()
    _, _, R = oracle.rhs(case, oracle.p2c(case, W))
    _, _, Rt = oracle.rhs(case, oracle.p2c(case, Wt))
    Rp = Rt.transpose(1, 0, 2, 3)[..., [0, 1, 3, 2, 4]]
    np.testing.assert_allclose(po.interior(R), po.interior(Rp), rtol=0, atol=1e-12 * np.abs(R).max())


@pytest.mark.parametrize("order", [2, 4, 6])
def test_fd_exact_on_polynomials(oracle, order):
    case = _case(diff_order=order, L=(1.0, 1.0, 1.0))
    NX, NY, NZ, _ = po.padded_shape(case)
    h = 1.0 / case["nx"]
    x = (np.arange(NX) - 3 + 0.5) * h
    f = np.zeros(po.padded_shape(case))
    p = np.polynomial.Polynomial(np.arange(1, order + 2) * 0.1)  # degree = order
    f[..., 2] = p(x)[:, None, None]
    df = oracle.fd(case, f, 0, 0, h)
    np.testing.assert_allclose(po.interior(df)[..., 2], np.broadcast_to(p.deriv()(x[3:-3])[:, None, None], po.interior(df)[..., 2].shape), rtol=1e-9, atol=1e-9)
    ddf = oracle.fd(case, f, 0, 1, h * h)
    np.testing.assert_allclose(po.interior(ddf)[..., 2], np.broadcast_to(p.deriv(2)(x[3:-3])[:, None, None], po.interior(ddf)[..., 2].shape), rtol=1e-7, atol=1e-7)


def test_second_derivative_y_uses_x_stencil(oracle):
    # quirk: ddf_dydy runs the x stencil (cd-n.cpp:74-79); data varying only in y gives 0
    case = _case(diff_order=4)
    f = np.zeros(po.padded_shape(case))
    f[..., 0] = np.sin(np.arange(f.shape[1]))[None, :, None]
    # (zero up to the rounding of the coefficient sum) while the true d2/dy2 is O(1)
    assert np.abs(po.interior(oracle.fd(case, f, 0, 1, 1.0))[..., 0]).max() < 1e-15


def test_eos_round_trip_and_dt(oracle):
    C = cases_module()
    case = _case()
    W = C.random_smooth_W(case, 9)
    U = oracle.p2c(case, W)
    W2 = oracle.c2p(case, U)
    np.testing.assert_allclose(po.interior(W2), po.interior(W), rtol=1e-13, atol=1e-15)
    Wi = po.interior(W)
    c = np.sqrt(case["gamma"] * Wi[..., 1] / Wi[..., 0])
    d = [case["L"][a] / case[n] for a, n in enumerate(("nx", "ny", "nz"))]
    dt = min((d[a] / np.maximum(np.abs(Wi[..., 2 + a] + c), np.abs(Wi[..., 2 + a] - c))).min() for a in range(3))
    assert oracle.dt(case, W) == pytest.approx(dt * case["cfl"], rel=1e-14)


def test_halo_semantics(oracle):
    case = _case(bc=["periodic", "periodic", "neumann", "neumann", "periodic", "periodic"])
    rng = np.random.default_rng(1)
    A = np.full(po.padded_shape(case), np.nan)
    A[I3] = rng.standard_normal(po.interior(A).shape)
    B = oracle.fill_halos(case, A)
    nx, ny, nz = case["nx"], case["ny"], case["nz"]
    J = slice(3, 3 + ny)
    K = slice(3, 3 + nz)
    Ii = slice(3, 3 + nx)
    np.testing.assert_array_equal(B[0:3, J, K], A[nx:nx + 3, J, K])          # periodic x low
    np.testing.assert_array_equal(B[3 + nx:6 + nx, J, K], A[3:6, J, K])      # periodic x high
    for h in range(3):                                                       # Neumann y
        np.testing.assert_array_equal(B[Ii, h, K], A[Ii, 3, K])
        np.testing.assert_array_equal(B[Ii, 3 + ny + h, K], A[Ii, 2 + ny, K])
    assert np.isnan(B[0, 0, 4]).all() and np.isnan(B[0, 4, 0]).all()        # edges untouched
    # edges are never read: NaN-poisoned edges leave R finite
    C = cases_module()
    W = C.random_smooth_W(case, 2)
    U = oracle.p2c(case, W)
    U[~po.faces_mask(case)] = np.nan
    _, _, R = oracle.rhs(case, U)
    assert np.isfinite(po.interior(R)).all()


def test_shear_and_tgv_stable(oracle):
    # the nondimensionalisation (c0 = 3/(1+M)) keeps the reference scheme stable
    C = cases_module()
    case, W = C.config1(16)
    W2, _, dts = oracle.step(case, W, 20)
    assert np.isfinite(po.interior(W2)).all()
    assert dts.max() / dts.min() < 1.05


# ---------------------------------------------------------------------------------------------
# Sutherland viscosity (extension: the reference defines the law, viscosity.cpp:5-11, but never
# calls it; these checks pin the oracle's use of it)
# ---------------------------------------------------------------------------------------------

def test_sutherland_law_closed_form(oracle):
    mu0, C, g = 2e-3, 110.4 / 300.0, 1.4
    # T = gamma p / rho = 1 (the reference temperature) -> mu0
    assert oracle.mu_sutherland(mu0, C, g, 1.4, 1.0) == pytest.approx(mu0, rel=1e-15)
    for rho, p in [(1.0, 2.0), (0.7, 0.3), (2.0, 5.0)]:
        T = g * p / rho
        assert oracle.mu_sutherland(mu0, C, g, rho, p) == pytest.approx(
            mu0 * (1 + C) / (T + C) * T ** 1.5, rel=1e-14)


def test_sutherland_uniform_temperature_equals_constant_mu(oracle):
    # rho varies but p / rho does not: mu(T) is uniform, dmu = 0, so the residual equals the
    # constant-viscosity residual with mu = mu(T)
    C_ = cases_module()
    case = _case(conv="none", diff_order=4, mu=1e-2)
    W = C_.random_smooth_W(case, 21)
    W[I3 + (1,)] = 0.9 * W[I3 + (0,)]
    U = oracle.p2c(case, W)
    _, _, Rs = oracle.rhs(dict(case, viscosity="sutherland"), U, which=2)
    mu_T = oracle.mu_sutherland(case["mu"], po.SUTH_C, case["gamma"], 1.0, 0.9)
    _, _, Rc = oracle.rhs(dict(case, mu=mu_T), U, which=2)
    np.testing.assert_allclose(po.interior(Rs), po.interior(Rc), rtol=0,
                               atol=1e-12 * np.abs(po.interior(Rc)).max())


def test_sutherland_viscosity_gradient_term_converges(oracle):
    # u = v = 0, uniform p, rho(x): only the heat-flux terms act on energy.  With T = gamma p/rho,
    # the residual as written is -k [ mu' (-p rho'/rho^2) + mu (p rho'^2/rho^3 - p rho''/rho^2) ]
    # (k = gamma/(Pr (gamma-1))) plus the y term; compare with exact derivatives: the error must
    # fall at 4th order
    errs = []
    for n in (16, 32):
        case = dict(_case(conv="none", diff_order=4, mu=1e-2, L=(1.0, 1.0, 1.0)),
                    nx=n, ny=4, nz=4, viscosity="sutherland")
        x = (np.arange(n) + 0.5) / n
        rho = 1.0 + 0.3 * np.sin(2 * np.pi * x)
        r1 = 0.3 * 2 * np.pi * np.cos(2 * np.pi * x)
        r2 = -0.3 * (2 * np.pi) ** 2 * np.sin(2 * np.pi * x)
        p, g, C, mu0 = 1.0, case["gamma"], po.SUTH_C, case["mu"]
        W = np.zeros(po.padded_shape(case))
        W[I3 + (0,)] = rho[:, None, None]
        W[I3 + (1,)] = p
        _, _, R = oracle.rhs(case, oracle.p2c(case, W), which=2)
        T = g * p / rho
        mu = mu0 * (1 + C) / (T + C) * T ** 1.5
        dmu_dT = mu * (1.5 / T - 1.0 / (T + C))
        mu_x = dmu_dT * (-g * p * r1 / rho ** 2)
        k = g / (case["Pr"] * (g - 1))
        exact = -k * (mu_x * (-p * r1 / rho ** 2) + mu * (p * r1 ** 2 / rho ** 3 - p * r2 / rho ** 2))
        # quirk: "d2rho/dy2" is the x stencil divided by dy^2, i.e. rho'' (dx/dy)^2
        ratio = (case["L"][0] / n) ** 2 / (case["L"][1] / case["ny"]) ** 2
        exact += -k * mu * (-p * r2 * ratio / rho ** 2)
        errs.append(np.abs(po.interior(R)[:, 1, 1, 1] - exact).max())
    assert errs[0] / errs[1] > 12, errs
