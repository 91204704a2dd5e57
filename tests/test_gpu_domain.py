"""Decomposed blocks on one GPU (nsgpu_domain): the reference's Cartesian decomposition with the
remainder to block 0 (decomposition.cpp:26-57), the face exchange between blocks
(communicator.cpp:16-144) and the derivative exchange (centered_diffusion.cpp:22-24) as halo
kernels, one dt minimum over the blocks.  Every decomposition must give bitwise the fields and
time steps of the same cells advanced as one block (Solver.step, itself bitwise the oracle).
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package

pytestmark = pytest.mark.gpu

PERIODIC = ["periodic"] * 6
WALLS = ["neumann"] * 6
MIXED = ["periodic", "periodic", "neumann", "neumann", "periodic", "periodic"]


@pytest.fixture(scope="module")
def ns():
    return package()


def _split(n, p, b):
    """decomposition.cpp:41-52 and :84-90: n / p cells, the remainder to block 0."""
    local, rem = divmod(n, p)
    return b * local + (rem if b > 0 else 0), local + (rem if b == 0 else 0)


def _case(C, bc, conv="weno5cw", viscosity="constant", shape=(20, 17, 15)):
    case = dict(C.base_case(*shape), conv=conv, bc=bc, mu=3e-3, L=(1.0, 1.2, 0.9))
    if viscosity == "sutherland":
        C.sutherland(case, case["mu"], 2.2 ** 2)
    return case, C.random_smooth_W(case, seed=11, c0=2.2)


def _single(ns, case, W0, steps, math):
    s = ns.Solver(case, math=math)
    s.set_state(W0, "primitive")
    t, dts = 0.0, []
    for _ in range(steps):
        t, dt = s.step(1, time=t)
        dts.append(dt)
    out = s.get_state("primitive"), s.get_state("conservative"), t, dts
    s.close()
    return out


# uneven splits on every axis (20, 17, 15 cells): blocks of different extents, strides and planes
@pytest.mark.parametrize("nproc", [(2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1), (3, 2, 2), (1, 3, 3)])
@pytest.mark.parametrize("bc", [PERIODIC, WALLS, MIXED], ids=["periodic", "walls", "mixed"])
def test_domain_equals_single_block(ns, nproc, bc):
    C = cases_module()
    case, W0 = _case(C, bc)
    print(f"[gpu] domain {nproc} {bc[0]}/{bc[2]}", flush=True)
    steps = 5
    Ws, Us, ts, dts = _single(ns, case, W0, steps, "exact")
    d = ns.Domain(case, nproc)
    assert len(d.blocks) == int(np.prod(nproc))
    for r, b in enumerate(d.blocks):
        i, j, k = r % nproc[0], (r // nproc[0]) % nproc[1], r // (nproc[0] * nproc[1])
        for a, (n, p, c) in enumerate(zip((case["nx"], case["ny"], case["nz"]), nproc, (i, j, k))):
            assert (b.start[a], b.count[a]) == _split(n, p, c)
    d.set_state(W0, "primitive")
    t, dl = 0.0, []
    for _ in range(steps):
        t, dt = d.step(1, time=t)
        dl.append(dt)
    np.testing.assert_array_equal(po.interior(d.get_state("primitive")), po.interior(Ws))
    np.testing.assert_array_equal(po.interior(d.get_state("conservative")), po.interior(Us))
    assert dl == dts and t == ts
    sums, _, flags, rc = d.reduce()
    assert rc in (0, -6) and not (flags & 3)
    d.close()


@pytest.mark.parametrize("math", ["exact", "fast", "fp32"])
@pytest.mark.parametrize("conv,viscosity", [("weno5cw", "sutherland"), ("weno5", "constant")])
def test_domain_schemes_and_math(ns, math, conv, viscosity):
    C = cases_module()
    case, W0 = _case(C, MIXED, conv=conv, viscosity=viscosity)
    Ws, Us, ts, dts = _single(ns, case, W0, 4, math)
    d = ns.Domain(case, (2, 2, 2), math=math)
    d.set_state(W0, "primitive")
    t, dt = d.step(4)
    np.testing.assert_array_equal(po.interior(d.get_state("primitive")), po.interior(Ws))
    assert dt == dts[-1] and t == ts
    d.close()


def test_domain_thin_2d_blocks(ns):
    # a 2-D input (Nz = 1, the z sweep skipped) split in x and y, as the reference's 2-D
    # tutorials are run on 2 x 2 ranks
    C = cases_module()
    case = dict(C.base_case(24, 18, 1), dim=2, conv="weno5", diff_order=6, mu=1e-3,
                L=(1.0, 1.0, 1.0 / 24))
    W0 = C.random_smooth_W(case, seed=5, c0=2.2)
    Ws, _, ts, dts = _single(ns, case, W0, 6, "exact")
    d = ns.Domain(case, (2, 2, 1))
    d.set_state(W0, "primitive")
    t, dt = d.step(6)
    np.testing.assert_array_equal(po.interior(d.get_state("primitive")), po.interior(Ws))
    assert dt == dts[-1]
    d.close()


def test_domain_blocks_are_owned(ns):
    # the blocks belong to the domain: single-block stepping and operator calls are refused,
    # and an invalid split (a block thinner than the halo) is rejected
    C = cases_module()
    case, W0 = _case(C, PERIODIC)
    d = ns.Domain(case, (2, 1, 1))
    b = d.blocks[0]
    assert ns.load_library().nsgpu_step(b.h, 1, None, None) == -1
    d.close()
    with pytest.raises(ns.NsgpuError):
        ns.Domain(case, (8, 1, 1))  # 20 / 8 = 2 < 3
