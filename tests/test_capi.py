"""CPU-side checks of the C ABI library: it loads, exports every entry point that include/nsgpu.h
declares, and rejects invalid configurations before touching a device (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, cases_module, package


def header_symbols():
    text = open(os.path.join(ROOT, "include", "nsgpu.h")).read()
    return sorted(set(re.findall(r"\b(nsgpu_[a-z_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    ns = package()
    lib = ns.load_library()
    syms = header_symbols()
    assert len(syms) >= 15
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(ns.EXPORTS)


def test_library_is_gfx950_hip_code():
    path = package().LIB_PATH
    blob = open(path, "rb").read()
    assert b"gfx950" in blob


def _create(case, **over):
    ns = package()
    cfg = ns.make_config(case)
    for k, v in over.items():
        if k == "bc":
            for i, b in enumerate(v):
                cfg.bc[i] = b
        else:
            setattr(cfg, k, v)
    h = ctypes.c_void_p()
    return ns.load_library().nsgpu_create(ctypes.byref(cfg), 0, ctypes.byref(h))


@pytest.mark.parametrize("over", [
    dict(ngc=2), dict(nx=2), dict(dim=1), dict(diff_order=3), dict(conv=3), dict(weights=3),
    dict(math=5), dict(gamma=1.0), dict(Lx=0.0), dict(alpha_mode=2), dict(physics=4),
    dict(physics=2, dim=2),  # the full 3-D viscous terms need a 3-D block
    dict(bc=[1, 0, 0, 0, 0, 0]),   # x not periodic (West Neumann) but East periodic
    dict(bc=[0, 0, 0, 0, 1, 0]),   # z not periodic (Front Neumann) but Back periodic
])
def test_invalid_configs_rejected(over):
    C = cases_module()
    assert _create(C.base_case(8, 8, 8), **over) == -1


def test_learned_weights_with_scheme_options_unsupported():
    # validated before any device call: runs without a GPU
    C = cases_module()
    case = dict(C.base_case(8, 8, 8), weights="learned")
    assert _create(case, alpha_mode=1) == -7
    assert _create(case, physics=1) == -7


def test_strerror_and_null_args():
    ns = package()
    L = ns.load_library()
    assert L.nsgpu_strerror(-6).decode().startswith("max |u|+c")
    assert L.nsgpu_step(None, 1, None, None) == -1
    assert L.nsgpu_array_len(None) == 0


def test_ensemble_rejects_bad_arguments_before_touching_a_device():
    L = package().load_library()
    h = ctypes.c_void_p()
    none = (ctypes.c_void_p * 1)(None)
    assert L.nsgpu_ensemble_create(None, 1, ctypes.byref(h)) == -1
    assert L.nsgpu_ensemble_create(none, 0, ctypes.byref(h)) == -1
    assert L.nsgpu_ensemble_create(none, 1, ctypes.byref(h)) == -1
    assert L.nsgpu_ensemble_create(none, 1, None) == -1
    assert L.nsgpu_ensemble_step(None, 1, 1) == -1
    assert L.nsgpu_ensemble_tune(None, 1, None, None) == -1
    assert L.nsgpu_ensemble_set_tiles(None, None) == -1
