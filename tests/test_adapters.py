"""The reference-side adapters (3dnssolver_amd/adapters/gpu_schemes.hpp) compile against the
reference's own headers: a translation unit that includes them, instantiates the three adapter
classes and calls every method the reference's driver would call is checked with g++
-fsyntax-only.  Runs only where the reference tree and the MPI headers it includes are present
(this container); nothing of the reference is copied or executed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
MPI_INC = "/opt/conda/include"

TU = r'''
#include "gpu_schemes.hpp"
void use(const Params& p, Domain& d, const Grid& g, const Communicator& comm,
         Array4d& W, Array4d& U, Array4d& R) {
    nsgpu_adapters::GpuConvection conv(p, W);
    nsgpu_adapters::GpuDiffusion diff(p, W);
    ConvectionScheme* cs = &conv;
    DiffusionScheme*  ds = &diff;
    cs->solve(g, W, U, R);
    ds->solve(g, comm, W, R);
    nsgpu_adapters::GpuTimeScheme ts(p, d);
    ts.step(d, g, comm, cs, ds);
    double dt = ts.get_dt();
    (void)dt;
    ts.sync_to_host(d);
}
'''


@pytest.mark.skipif(not (os.path.isdir(REF_SRC) and os.path.exists(os.path.join(MPI_INC, "mpi.h"))
                         and shutil.which("g++")),
                    reason="reference tree or MPI headers not present")
def test_adapters_compile_against_reference_headers(tmp_path):
    src = tmp_path / "use_adapters.cpp"
    src.write_text(TU)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-I", REF_SRC, "-I", MPI_INC,
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "3dnssolver_amd", "adapters"),
           str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


MAKE_CONFIG_TU = r'''
#include <cstdio>
#include "gpu_schemes.hpp"
int main() {
    Params p;
    p.dim = 3, p.nx = 8, p.ny = 8, p.nz = 8, p.Lx = p.Ly = p.Lz = 1.0;
    p.mu = 1e-3, p.Pr = 0.72, p.gamma = 1.4, p.CFL_max = 0.5;
    p.weight_type = "Shu", p.diffusion_scheme = "Centered", p.diffusion_scheme_order = 4;
    p.convection_scheme_order = 5, p.convection_in_characteristic_space = 1;
    const char* names[6] = {"West", "East", "North", "South", "Front", "Back"};
    for (auto n : names) p.boundary_conditions[n] = "Periodic";
    Array4d like(8, 8, 8, 5, 3);
    int fails = 0;
    p.convection_scheme = "Weno";
    if (nsgpu_adapters::make_config(p, like, NSGPU_MATH_EXACT).conv != NSGPU_CONV_WENO5CW) fails++;
    p.convection_scheme = "none";
    if (nsgpu_adapters::make_config(p, like, NSGPU_MATH_EXACT).conv != NSGPU_CONV_NONE) fails++;
    p.convection_scheme = "Centered";
    bool threw = false;
    try {
        nsgpu_adapters::make_config(p, like, NSGPU_MATH_EXACT);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    if (!threw) fails++;
    std::printf("fails=%d\n", fails);
    return fails;
}
'''


@pytest.mark.skipif(not (os.path.isdir(REF_SRC) and os.path.exists(os.path.join(MPI_INC, "mpi.h"))
                         and shutil.which("g++")),
                    reason="reference tree or MPI headers not present")
def test_make_config_refuses_centered_convection(tmp_path):
    # Driver::create_convection_operator (driver.cpp:74-94) builds CenteredConvection for
    # "Centered"; the adapters have no GPU counterpart and must refuse it, not run without
    # convection.  make_config only touches Params and Array4d (containers/array_4d.cpp).
    src = tmp_path / "make_config.cpp"
    src.write_text(MAKE_CONFIG_TU)
    exe = tmp_path / "make_config"
    cmd = ["g++", "-std=c++17", "-O0", "-I", REF_SRC, "-I", MPI_INC,
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "3dnssolver_amd", "adapters"),
           str(src), os.path.join(REF_SRC, "containers", "array_4d.cpp"), "-o", str(exe),
           # the MPI library by path (its directory also holds an older libstdc++)
           os.path.join(os.path.dirname(MPI_INC), "lib", "libmpi.so")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # the system C++ runtime ahead of the MPI directory's older one
    env = dict(os.environ, LD_LIBRARY_PATH="/usr/lib/x86_64-linux-gnu:" +
               os.path.join(os.path.dirname(MPI_INC), "lib"))
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0 and "fails=0" in r.stdout, (r.stdout, r.stderr)


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libnsref_adapters.so")),
                    reason="adapter harness not built (needs the reference tree)")
def test_adapter_library_make_config():
    # the adapter harness (the reference's TUs + gpu_schemes.hpp, linked to libnsgpu.so) loads
    # without a GPU, and its make_config refuses the Centered convection scheme
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    C = __import__("importlib").import_module("3dnssolver_amd.cases")
    a = pyoracle.Adapters()
    case = C.base_case(8, 8, 8)
    assert a.make_config(case, "Weno") == 0
    assert a.make_config(case, "none") == 0
    assert a.make_config(case, "Centered") == 1
