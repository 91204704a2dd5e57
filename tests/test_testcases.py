"""The reference's initial conditions (testcases.cpp): the generators of cases.py against the
reference's own functions, compiled from its sources into the reference harness (bitwise), and
the walled Riemann problems on the oracle (no GPU needed)."""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module

C = cases_module()


@pytest.fixture(scope="module")
def ref():
    if not po.reference_available("strict"):
        pytest.skip("reference build (oracle/_ref) not present")
    return po.Reference("strict")


@pytest.mark.parametrize("name", list(C.RIEMANN))
def test_initial_conditions_match_reference_bitwise(ref, name):
    for shape in ((16, 12, 4), (13, 9, 5)):
        case = dict(C.base_case(*shape), L=(1.0, 1.3, 0.7))
        W = C.riemann_W(case, name)
        Wr = ref.testcase(case, C.RIEMANN_IDS[name])
        I = po.interior(W)
        Ir = po.interior(Wr)
        if name == "case17":
            # SW's w is read past the end of a 4-element vector by the reference
            X, Y, _ = C.reference_coords(case)
            sw = (X <= 0.5) & (Y > 1.3 / 2.0)
            np.testing.assert_array_equal(I[..., :4], Ir[..., :4])
            np.testing.assert_array_equal(I[~sw][:, 4], Ir[~sw][:, 4])
        else:
            np.testing.assert_array_equal(I, Ir)


def test_sod_problem_on_oracle_moves_the_shock_right(oracle):
    # a Sod tube along x between zero-gradient (Neumann) faces: after 10 steps only the numerical
    # tails of the waves (the stencils widen the domain of dependence by 9 cells a step) reach the
    # faces, so mass changes by < 1e-8 relative; the density between the initial levels shows
    # the waves leaving x = 0.5
    case, W = C.riemann_case("sod_x", n=32, nz=4)
    W1, U1, _ = oracle.step(case, W, 10)
    U0 = oracle.p2c(case, W)
    m0 = po.interior(U0)[..., 0].sum()
    m1 = po.interior(U1)[..., 0].sum()
    assert abs(m1 - m0) < 1e-8 * m0
    I = po.interior(W1)
    assert np.isfinite(I).all()
    rho = I[:, 5, 2, 0]
    assert rho[20] > 0.125 + 1e-2 and rho[12] < 1.0 - 1e-2
    # uniform in y and z (a 1-D problem on a 3-D block)
    np.testing.assert_array_equal(I[:, :, :, 0], np.broadcast_to(I[:, :1, :1, 0], I.shape[:3]))


def test_sutherland_case_hits_reference_viscosity():
    # the library's law mu(T) = mu (1 + C) / (T + C) T^1.5 with T = gamma p / rho: after
    # cases.sutherland the viscosity at the reference temperature is the requested one, and the
    # ratio mu(T) / mu(T_ref) is air's Sutherland curve in T / T_ref
    C = cases_module()
    case, _ = C.config2(8, viscosity="sutherland")
    const, _ = C.config2(8)
    T_ref = (C.TGV_SPEED_FRAC * C.sound_speed(0.1)) ** 2
    law = lambda T: case["mu"] * (1 + case["suth_C"]) / (T + case["suth_C"]) * T ** 1.5
    assert law(T_ref) == pytest.approx(const["mu"], rel=1e-14)
    for r in (0.5, 2.0):
        air = (300.0 + 110.4) / (300.0 * r + 110.4) * r ** 1.5
        assert law(r * T_ref) / law(T_ref) == pytest.approx(air, rel=1e-13)
    with pytest.raises(ValueError):
        C.config2(8, viscosity="power")


def test_crand_is_the_c_library_sequence():
    # known answer of srand(1): the first draws of the C library's rand(), and the library itself
    import ctypes
    C = cases_module()
    g = C.CRand(1)
    first = (g.take(5) * C.RAND_MAX).astype(np.int64).tolist()
    assert first == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    ref = np.array([libc.rand() for _ in range(20000)], dtype=np.float64) / C.RAND_MAX
    np.testing.assert_array_equal(C.CRand(1).take(20000), ref)


@pytest.mark.skipif(not po.reference_available("strict"), reason="reference build absent")
@pytest.mark.parametrize("name,ident,shape,L", [
    ("shear_xy", 9, (20, 16, 1), (1.0, 1.0, 1.0)),
    ("shear_xy", 9, (12, 10, 6), (2.0, 1.5, 0.5)),
    ("shear_3d", 10, (12, 10, 6), (1.0, 1.0, 1.0)),
])
def test_shear_layers_match_reference(name, ident, shape, L):
    # the reference's shear layers (testcases.cpp:250-344, rand() perturbations) bitwise
    C = cases_module()
    case = dict(C.base_case(*shape), L=L, dim=3 if shape[2] > 1 else 2)
    W = C.reference_initial_condition(case, name)
    Wr = po.Reference("strict").testcase(case, ident)
    np.testing.assert_array_equal(W, Wr)
    assert np.array_equal(np.signbit(W), np.signbit(Wr))  # the -0.0 of 0.0 * sin(x < 0) too


def test_block_initial_conditions_tile_the_domain():
    # a decomposed run's blocks: coordinates from the block's start (grid.cpp:24-57), so the
    # Riemann states of the blocks tile the single block's exactly
    C = cases_module()
    case = dict(C.base_case(20, 16, 4), L=(1.0, 1.0, 0.2))
    whole = C.reference_initial_condition(case, "case6")
    for start, count in (((0, 0, 0), (10, 8, 4)), ((10, 8, 0), (10, 8, 4)), ((10, 0, 0), (10, 8, 4))):
        blk = C.reference_initial_condition(case, "case6", start, count)
        sl = tuple(slice(3 + s, 3 + s + c) for s, c in zip(start, count))
        np.testing.assert_array_equal(blk[3:-3, 3:-3, 3:-3], whole[sl])
    with pytest.raises(ValueError):
        C.reference_initial_condition(case, "jet")
