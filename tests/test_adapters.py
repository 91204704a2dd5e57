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
