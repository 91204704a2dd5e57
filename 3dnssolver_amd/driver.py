"""A run from the reference's input file on the GPU: Driver::run (driver.cpp:3-72).

  python -m <package>.driver input.xml [--max-steps N] [--math exact|fast|fp32] [--write DIR]
                                       [--checkpoint-every N --checkpoint-dir DIR]
                                       [--restart FILE]

Reads input.xml (io/reader.cpp:3-172, Params::sanity_check), decomposes the block over
NprocX x NprocY x NprocZ linked blocks on the one GPU when more than one is asked for
(decomposition.cpp, communicator.cpp: nsgpu_domain), sets the initial condition of every block as
Driver::set_initial_condition does on each rank (driver.cpp:107-136; every rank draws its own
rand() sequence), and steps until the flow time reaches EndT (driver.cpp:43-68), printing after
every step the reference's line

  Flow time: <t>\tCPU time (step): <s>\tdt: <dt>\tCFL: <CFL>

with std::cout's default formatting (six significant digits).  The reference's output call is
commented out (driver.cpp:49); --write DIR performs it: every OutputInterval iterations before the
step, the HDF5 file results/primitive_fields_3d.h5 (FileFormat Hdf5) or the per-rank binary files
(FileFormat Binary).  Checkpoints (raw W + time) and --restart resume a run bitwise.
"""
import importlib
import os
import sys
import time as _time

import numpy as np

from . import inputs, output

_pkg = importlib.import_module(__package__)
_cases = importlib.import_module(__package__ + ".cases")


def cout(x):
    """A double as `std::cout << x` prints it by default (%g, six significant digits)."""
    return "%g" % x


def step_line(time, cpu, dt, cfl):
    """The per-step line of driver.cpp:61-65."""
    return f"Flow time: {cout(time)}\tCPU time (step): {cout(cpu)}\tdt: {cout(dt)}\tCFL: {cout(cfl)}"


class Run:
    """The block(s) of one input file on the GPU: a Solver, or a Domain of linked blocks."""

    def __init__(self, path, math="exact", device=0):
        self.case, self.extra = inputs.read_input_xml(path)
        self.nproc = tuple(self.extra["nproc"])
        if int(np.prod(self.nproc)) > 1:
            self.dom = _pkg.Domain(self.case, self.nproc, math=math, device=device)
            self.blocks = self.dom.blocks
        else:
            self.dom = None
            self.solver = _pkg.Solver(self.case, math=math, device=device)
        self.time, self.iter, self.dt = 0.0, 0, 0.0

    def close(self):
        if self.dom is not None:
            self.dom.close()
        else:
            self.solver.close()

    def set_initial_condition(self):
        """Every rank's block gets the reference's initial condition of its cells."""
        name = self.extra["testcase"]
        if self.dom is None:
            self.solver.set_state(_cases.reference_initial_condition(self.case, name), "primitive")
            return
        for b in self.blocks:
            b.set_state(_cases.reference_initial_condition(self.case, name, b.start, b.count),
                        "primitive")

    def set_state(self, W, time=0.0, it=0):
        (self.dom or self.solver).set_state(W, "primitive")
        self.time, self.iter = time, it

    def get_state(self, kind="primitive"):
        return (self.dom or self.solver).get_state(kind)

    def step(self):
        t, dt = (self.dom or self.solver).step(1, time=self.time)
        self.time, self.dt = t, dt
        self.iter += 1
        return t, dt

    def block_states(self):
        """(rank, start, count, padded primitive state) of every block."""
        if self.dom is None:
            c = (self.case["nx"], self.case["ny"], self.case["nz"])
            return [(0, (0, 0, 0), c, self.solver.get_state("primitive"))]
        return [(r, b.start, b.count, b.get_state("primitive")) for r, b in enumerate(self.blocks)]


def run(path, math="exact", max_steps=None, stream=None, write_dir=None, checkpoint_every=0,
        checkpoint_dir=None, restart=None, device=0):
    """Runs the input file to EndT (or max_steps more steps); returns the Run (closed) with its
    final primitive state in .W."""
    stream = stream if stream is not None else sys.stdout
    r = Run(path, math=math, device=device)
    try:
        if restart:
            W, t, it, _ = output.read_checkpoint(restart, r.case)
            r.set_state(W, t, it)
        else:
            r.set_initial_condition()
        end_t, cfl = r.extra["end_t"], r.case["cfl"]
        interval = max(r.extra["io_interval"], 1)
        h5 = None
        if write_dir and r.extra["output_file_format"] == "Hdf5":
            os.makedirs(os.path.join(write_dir, "results"), exist_ok=True)
            h5 = os.path.join(write_dir, "results", "primitive_fields_3d.h5")
            output.hdf5_write_mesh(h5, r.case)
        elif write_dir:
            for rank, st, ct, _ in r.block_states():
                output.write_mesh(r.case, write_dir, rank, st, ct)
        n_ckpt, done = 0, 0
        while r.time < end_t and (max_steps is None or done < max_steps):
            if write_dir and r.iter % interval == 0:  # OutputInfo::is_write_time
                if h5:
                    output.hdf5_write_checkpoint(h5, n_ckpt, r.get_state(), r.time)
                    n_ckpt += 1
                else:
                    for rank, _, _, Wb in r.block_states():
                        output.write_fields(Wb, write_dir, r.iter, rank)
            t0 = _time.perf_counter()
            t, dt = r.step()
            cpu = _time.perf_counter() - t0
            done += 1
            print(step_line(t, cpu, dt, cfl), file=stream, flush=True)
            if checkpoint_every and checkpoint_dir and r.iter % checkpoint_every == 0:
                output.write_checkpoint(checkpoint_dir, r.case, r.get_state(), r.time, r.iter,
                                        r.dt)
        r.W = r.get_state()
    finally:
        r.close()
    return r


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input")
    ap.add_argument("--math", choices=["exact", "fast", "fp32"], default="exact")
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--write", default=None, help="directory for the reference's output files")
    ap.add_argument("--checkpoint-every", type=int, default=0)
    ap.add_argument("--checkpoint-dir", default=None)
    ap.add_argument("--restart", default=None, help="raw checkpoint (.dat) to resume from")
    a = ap.parse_args(argv)
    run(a.input, math=a.math, max_steps=a.max_steps, write_dir=a.write,
        checkpoint_every=a.checkpoint_every, checkpoint_dir=a.checkpoint_dir, restart=a.restart)


if __name__ == "__main__":
    main()
