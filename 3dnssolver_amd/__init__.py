"""MI355X-native RHS + SSP-RK3 path of the compressible Navier-Stokes block solver.

Python view of the C ABI in include/nsgpu.h (lib/libnsgpu.so, HIP for gfx950), shaped like the
reference's operator interface (paths relative to the reference's src/):

  Solver.apply_boundary(A)          Domain::apply_boundary_conditions + Communicator::communicate
                                    (domain.cpp:18-28, communicator.cpp:6-144)
  Solver.convection_solve(W, U, R)  ConvectionScheme::solve (convection_scheme.hpp:26): R += conv
  Solver.diffusion_solve(W, R)      DiffusionScheme::solve (diffusion_scheme.hpp:14-17): R += diff
  Solver.rhs(U, which)              one stage body of TimeScheme::step (time_scheme.cpp:33-41)
  Solver.step(n)                    TimeScheme::step (time_scheme.cpp:3-70), state on the GPU
  Solver.dt                         TimeScheme::get_dt (time_scheme.hpp:31)

  Ensemble(solvers)                 many blocks advanced together: batched launches, one hipGraph
                                    per SSP-RK3 step, tile shapes from an autotuner (no reference
                                    counterpart; each member's result is its own Solver.step)

Arrays are numpy float64 in the reference Array4d layout (NX, NY, NZ, 5), NX = nx + 6.
There is no CPU fallback: without the HIP library or a GPU every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnsgpu.so")

BC = {"periodic": 0, "neumann": 1}
CONV = {"none": 0, "weno5": 1, "weno5cw": 2}
WEIGHTS = {"shu": 0, "borges": 1, "learned": 2}
MLP_NPARAM = 32 * 8 + 32 + 32 * 32 + 32 + 3 * 32 + 3
MATH = {"exact": 0, "fast": 1, "fp32": 2}
STATE = {"conservative": 0, "primitive": 1}
VISCOSITY = {"constant": 0, "sutherland": 1}
ALPHA_MODE = {"const3": 0, "local_max": 1}
PHYSICS = {"reference": 0, "corrected_convection": 1, "corrected_viscous": 2, "corrected": 3}
SUTH_C = 110.4 / 300.0  # Sutherland constant of air over a 300 K reference temperature

EXPORTS = ["nsgpu_create", "nsgpu_destroy", "nsgpu_strerror", "nsgpu_array_len", "nsgpu_set_state",
           "nsgpu_get_state", "nsgpu_apply_boundary", "nsgpu_convection_solve",
           "nsgpu_diffusion_solve", "nsgpu_rhs", "nsgpu_step", "nsgpu_get_dt", "nsgpu_reduce",
           "nsgpu_stream", "nsgpu_selftest", "nsgpu_timing", "nsgpu_timer_count", "nsgpu_timer_name",
           "nsgpu_timer_read", "nsgpu_ensemble_create", "nsgpu_ensemble_destroy",
           "nsgpu_ensemble_set_tiles", "nsgpu_ensemble_get_tiles", "nsgpu_ensemble_step",
           "nsgpu_ensemble_tune", "nsgpu_set_mlp", "nsgpu_mlp_forward", "nsgpu_mlp_loss_grad",
           "nsgpu_mlp_train", "nsgpu_mlp_recon", "nsgpu_step_fixed", "nsgpu_mlp_loss_grad_noalloc",
           "nsgpu_domain_create", "nsgpu_domain_destroy", "nsgpu_domain_size", "nsgpu_domain_block",
           "nsgpu_domain_step"]


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ngc", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("gamma", ctypes.c_double), ("mu", ctypes.c_double), ("Pr", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("conv", ctypes.c_int), ("weights", ctypes.c_int), ("diff_order", ctypes.c_int),
        ("bc", ctypes.c_int * 6), ("math", ctypes.c_int),
        ("viscosity", ctypes.c_int), ("suth_C", ctypes.c_double),
        ("alpha_mode", ctypes.c_int), ("physics", ctypes.c_int),
    ]


class NsgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nsgpu error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path=None):
    """Loads lib/libnsgpu.so, or the library named by $NSGPU_LIB (a tuning build); raises if it
    was not built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("NSGPU_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python 3dnssolver_amd/build.py`")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    L.nsgpu_create.argtypes = [ctypes.POINTER(Config), ctypes.c_int, ctypes.POINTER(P)]
    L.nsgpu_destroy.argtypes = [P]
    L.nsgpu_destroy.restype = None
    L.nsgpu_strerror.restype = ctypes.c_char_p
    L.nsgpu_array_len.argtypes = [P]
    L.nsgpu_array_len.restype = ctypes.c_size_t
    L.nsgpu_set_state.argtypes = [P, D, ctypes.c_int]
    L.nsgpu_get_state.argtypes = [P, D, ctypes.c_int]
    L.nsgpu_apply_boundary.argtypes = [P, D]
    L.nsgpu_convection_solve.argtypes = [P, D, D, D]
    L.nsgpu_diffusion_solve.argtypes = [P, D, D]
    L.nsgpu_rhs.argtypes = [P, D, D, D, ctypes.c_int]
    L.nsgpu_step.argtypes = [P, ctypes.c_int, D, D]
    L.nsgpu_step_fixed.argtypes = [P, ctypes.c_int, ctypes.c_double, D]
    L.nsgpu_get_dt.argtypes = [P, D]
    L.nsgpu_reduce.argtypes = [P, D, D, ctypes.POINTER(ctypes.c_uint)]
    L.nsgpu_stream.argtypes = [P]
    L.nsgpu_stream.restype = P
    L.nsgpu_selftest.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_ulonglong,
                                 ctypes.POINTER(ctypes.c_long)]
    L.nsgpu_timing.argtypes = [P, ctypes.c_int]
    L.nsgpu_timer_name.restype = ctypes.c_char_p
    L.nsgpu_timer_read.argtypes = [P, ctypes.c_int, D, ctypes.POINTER(ctypes.c_long)]
    I = ctypes.POINTER(ctypes.c_int)
    L.nsgpu_ensemble_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(P)]
    L.nsgpu_ensemble_destroy.argtypes = [P]
    L.nsgpu_ensemble_destroy.restype = None
    L.nsgpu_ensemble_set_tiles.argtypes = [P, I]
    L.nsgpu_ensemble_get_tiles.argtypes = [P, I]
    L.nsgpu_ensemble_step.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.nsgpu_ensemble_tune.argtypes = [P, ctypes.c_int, D, D]
    Fp = ctypes.POINTER(ctypes.c_float)
    L.nsgpu_set_mlp.argtypes = [P, Fp]
    L.nsgpu_mlp_forward.argtypes = [ctypes.c_int, Fp, ctypes.c_long, Fp, Fp]
    L.nsgpu_mlp_loss_grad.argtypes = [ctypes.c_int, Fp, ctypes.c_long, Fp, Fp, Fp, D, Fp]
    L.nsgpu_mlp_train.argtypes = [ctypes.c_int, Fp, ctypes.c_long, Fp, Fp, Fp, ctypes.c_int,
                                  ctypes.c_long, ctypes.c_float, Fp]
    L.nsgpu_mlp_recon.argtypes = [ctypes.c_int, ctypes.c_int, Fp, ctypes.c_long, D, ctypes.c_int, D,
                                  I]
    # device pointers and a hipStream_t (allocation-free gradient)
    L.nsgpu_mlp_loss_grad_noalloc.argtypes = [ctypes.c_int, P, ctypes.c_long, P, P, P, P, P]
    I3 = ctypes.c_int * 3
    L.nsgpu_domain_create.argtypes = [ctypes.POINTER(Config), I3, ctypes.c_int, ctypes.POINTER(P)]
    L.nsgpu_domain_destroy.argtypes = [P]
    L.nsgpu_domain_destroy.restype = None
    L.nsgpu_domain_size.argtypes = [P]
    L.nsgpu_domain_block.argtypes = [P, ctypes.c_int, I3, I3]
    L.nsgpu_domain_block.restype = P
    L.nsgpu_domain_step.argtypes = [P, ctypes.c_int, D, D]
    _lib = L
    return L


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(case, math="exact"):
    c = Config()
    c.nx, c.ny, c.nz = case["nx"], case["ny"], case["nz"]
    c.ngc = 3
    c.dim = case.get("dim", 3)
    c.Lx, c.Ly, c.Lz = case["L"]
    c.gamma, c.mu, c.Pr, c.cfl = case["gamma"], case["mu"], case["Pr"], case["cfl"]
    c.conv = CONV[case["conv"]]
    c.weights = WEIGHTS[case["weights"]]
    c.diff_order = case["diff_order"]
    for f, name in enumerate(case["bc"]):
        c.bc[f] = BC[name]
    c.math = MATH[math]
    c.viscosity = VISCOSITY[case.get("viscosity", "constant")]
    c.suth_C = case.get("suth_C", SUTH_C)
    c.alpha_mode = ALPHA_MODE[case.get("alpha_mode", "const3")]
    c.physics = PHYSICS[case.get("physics", "reference")]
    return c


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"array shape {a.shape} != {shape}")
    return a


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _call(rc):
    if rc != 0:
        raise NsgpuError(rc, load_library().nsgpu_strerror(rc).decode())


def mlp_forward(params, X, device=0):
    """Softmax weights (n, 3) of the learned-weights network for features X (n, 8), on the GPU."""
    L = load_library()
    p = _f32(params, (MLP_NPARAM,))
    X = _f32(X)
    W = np.zeros((len(X), 3), dtype=np.float32)
    _call(L.nsgpu_mlp_forward(device, _fp(p), len(X), _fp(X), _fp(W)))
    return W


def mlp_loss_grad(params, X, Q, T, device=0):
    """(mean squared reconstruction error, gradient) on the GPU."""
    L = load_library()
    p = _f32(params, (MLP_NPARAM,))
    X, Q, T = _f32(X), _f32(Q), _f32(T)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = np.zeros(1)
    _call(L.nsgpu_mlp_loss_grad(device, _fp(p), len(T), _fp(X), _fp(Q), _fp(T), _ptr(loss), _fp(g)))
    return float(loss[0]), g


def mlp_train(params, X, Q, T, steps, batch=65536, lr=1e-3, device=0):
    """Adam training on the GPU; returns (new parameters, loss per step)."""
    L = load_library()
    p = _f32(params, (MLP_NPARAM,)).copy()
    X, Q, T = _f32(X), _f32(Q), _f32(T)
    hist = np.zeros(max(steps, 1), dtype=np.float32)
    _call(L.nsgpu_mlp_train(device, _fp(p), len(T), _fp(X), _fp(Q), _fp(T), steps, batch, lr,
                            _fp(hist)))
    return p, hist[:steps]


def mlp_recon(params, S, side, math="exact", device=0):
    """The face kernels' learned reconstruction of stencils S (n, 5): (values, regimes)."""
    L = load_library()
    p = _f32(params, (MLP_NPARAM,))
    S = np.ascontiguousarray(S, dtype=np.float64)
    out = np.zeros(len(S))
    rg = np.zeros(len(S), dtype=np.int32)
    _call(L.nsgpu_mlp_recon(device, MATH[math], _fp(p), len(S), _ptr(S), side, _ptr(out),
                            rg.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return out, rg


def selftest(n=1 << 26, seed=1, device=0):
    """Number of random operand pairs where the kernels' division differs from IEEE `/`."""
    L = load_library()
    out = ctypes.c_long(-1)
    rc = L.nsgpu_selftest(device, n, seed, ctypes.byref(out))
    if rc != 0:
        raise NsgpuError(rc, L.nsgpu_strerror(rc).decode())
    return out.value


class Solver:
    """One block on one GPU (one nsgpu context)."""

    def __init__(self, case, math="exact", device=0):
        self.L = load_library()
        self.case = dict(case)
        self.cfg = make_config(case, math)
        self.shape = (case["nx"] + 6, case["ny"] + 6, case["nz"] + 6, 5)
        h = ctypes.c_void_p()
        self._check(self.L.nsgpu_create(ctypes.byref(self.cfg), device, ctypes.byref(h)))
        self.h = h
        self._ensembles = 0  # live Ensembles holding this context

    def _check(self, rc):
        if rc != 0:
            raise NsgpuError(rc, self.L.nsgpu_strerror(rc).decode())

    def close(self):
        """Frees the context.  Refused (RuntimeError) while a live Ensemble holds it: the
        ensemble's block table and step graph point at the context's device buffers."""
        if getattr(self, "_ensembles", 0) > 0:
            raise RuntimeError("Solver.close: the context is a member of a live Ensemble; "
                               "close the ensemble first")
        if getattr(self, "h", None):
            self.L.nsgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"array shape {a.shape} != {self.shape}")
        return a

    def set_state(self, A, kind="primitive"):
        A = self._arr(A)
        self._check(self.L.nsgpu_set_state(self.h, _ptr(A), STATE[kind]))

    def get_state(self, kind="primitive"):
        A = np.zeros(self.shape)
        self._check(self.L.nsgpu_get_state(self.h, _ptr(A), STATE[kind]))
        return A

    def set_mlp(self, params):
        """Network of the learned weights (contexts with weights = "learned")."""
        p = _f32(params, (MLP_NPARAM,))
        self._check(self.L.nsgpu_set_mlp(self.h, _fp(p)))

    def apply_boundary(self, A):
        A = self._arr(A).copy()
        self._check(self.L.nsgpu_apply_boundary(self.h, _ptr(A)))
        return A

    def convection_solve(self, W, U, R):
        W, U, R = self._arr(W), self._arr(U), self._arr(R).copy()
        self._check(self.L.nsgpu_convection_solve(self.h, _ptr(W), _ptr(U), _ptr(R)))
        return R

    def diffusion_solve(self, W, R):
        W, R = self._arr(W), self._arr(R).copy()
        self._check(self.L.nsgpu_diffusion_solve(self.h, _ptr(W), _ptr(R)))
        return R

    def rhs(self, U, which=3):
        """Returns (U with halos filled, W = c2p(U), R)."""
        U = self._arr(U).copy()
        W = np.zeros(self.shape)
        R = np.zeros(self.shape)
        self._check(self.L.nsgpu_rhs(self.h, _ptr(U), _ptr(W), _ptr(R), which))
        return U, W, R

    def step(self, nsteps=1, time=0.0):
        """Advances the resident primitive state; returns (time, last dt).  Raises NsgpuError
        (code -4 / -5) when a stage produced a non-finite or non-positive state."""
        t = np.array([time], dtype=np.float64)
        dt = np.zeros(1)
        self._check(self.L.nsgpu_step(self.h, nsteps, _ptr(t), _ptr(dt)))
        return float(t[0]), float(dt[0])

    def step_fixed(self, nsteps, dt, time=0.0):
        """nsteps steps with the prescribed dt; returns the time."""
        t = np.array([time], dtype=np.float64)
        self._check(self.L.nsgpu_step_fixed(self.h, nsteps, dt, _ptr(t)))
        return float(t[0])

    @property
    def dt(self):
        dt = np.zeros(1)
        self._check(self.L.nsgpu_get_dt(self.h, _ptr(dt)))
        return float(dt[0])

    def reduce(self):
        """Returns (conservation sums[5], CFL dt, flags, status code)."""
        s = np.zeros(5)
        dt = np.zeros(1)
        fl = ctypes.c_uint(0)
        rc = self.L.nsgpu_reduce(self.h, _ptr(s), _ptr(dt), ctypes.byref(fl))
        if rc not in (0, -4, -5, -6):
            self._check(rc)
        return s, float(dt[0]), fl.value, rc

    def stream(self):
        return self.L.nsgpu_stream(self.h)

    def timing(self, enable=True):
        self._check(self.L.nsgpu_timing(self.h, 1 if enable else 0))

    def timers(self):
        out = {}
        for i in range(self.L.nsgpu_timer_count()):
            ms = ctypes.c_double()
            n = ctypes.c_long()
            self._check(self.L.nsgpu_timer_read(self.h, i, ctypes.byref(ms), ctypes.byref(n)))
            out[self.L.nsgpu_timer_name(i).decode()] = (ms.value, n.value)
        return out


class Ensemble:
    """Independent blocks (Solver objects on one device, same math mode and scheme options)
    advanced together.  Members stay usable as Solvers between ensemble calls."""

    def __init__(self, members):
        self.members = list(members)
        self.L = load_library()
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        rc = self.L.nsgpu_ensemble_create(arr, len(self.members), ctypes.byref(h))
        if rc != 0:
            raise NsgpuError(rc, self.L.nsgpu_strerror(rc).decode())
        self.h = h
        for m in self.members:
            m._ensembles += 1

    def _check(self, rc):
        if rc != 0:
            raise NsgpuError(rc, self.L.nsgpu_strerror(rc).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.nsgpu_ensemble_destroy(self.h)
            self.h = None
            for m in self.members:
                m._ensembles -= 1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tiles(self):
        """(n, 3) int array: kw_flux, kw_cell, march per member."""
        t = np.zeros((len(self.members), 3), dtype=np.int32)
        self._check(self.L.nsgpu_ensemble_get_tiles(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return t

    @tiles.setter
    def tiles(self, t):
        t = np.ascontiguousarray(t, dtype=np.int32).reshape(len(self.members), 3)
        self._check(self.L.nsgpu_ensemble_set_tiles(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def step(self, nsteps=1, graph=True):
        self._check(self.L.nsgpu_ensemble_step(self.h, nsteps, 1 if graph else 0))

    def tune(self, steps=3):
        """Runs the autotuner; returns (ms per step with default tiles, with the chosen tiles)."""
        a = np.zeros(1)
        b = np.zeros(1)
        self._check(self.L.nsgpu_ensemble_tune(self.h, steps, _ptr(a), _ptr(b)))
        return float(a[0]), float(b[0])


class _Block:
    """A block of a Domain: the state calls of a Solver on a context the domain owns."""

    def __init__(self, L, h, shape, start, count):
        self.L, self.h, self.shape, self.start, self.count = L, h, shape, start, count

    def _check(self, rc):
        if rc != 0:
            raise NsgpuError(rc, self.L.nsgpu_strerror(rc).decode())

    set_state = Solver.set_state
    get_state = Solver.get_state
    reduce = Solver.reduce
    _arr = Solver._arr


class Domain:
    """One block split into nproc = (px, py, pz) linked blocks on one GPU, the reference's
    decomposition of a multi-rank run (decomposition.cpp:26-57, communicator.cpp:6-144).  Global
    arrays are the padded (nx+6, ny+6, nz+6, 5) layout of the undivided block; set_state scatters
    their interiors to the blocks, get_state gathers them back (halos zero).  Each block's own
    padded arrays are available through .blocks (what one rank of the reference holds)."""

    def __init__(self, case, nproc, math="exact", device=0):
        self.L = load_library()
        self.case = dict(case)
        self.cfg = make_config(case, math)
        self.nproc = tuple(int(p) for p in nproc)
        self.shape = (case["nx"] + 6, case["ny"] + 6, case["nz"] + 6, 5)
        h = ctypes.c_void_p()
        rc = self.L.nsgpu_domain_create(ctypes.byref(self.cfg), (ctypes.c_int * 3)(*self.nproc),
                                        device, ctypes.byref(h))
        if rc != 0:
            raise NsgpuError(rc, self.L.nsgpu_strerror(rc).decode())
        self.h = h
        self.blocks = []
        for r in range(self.L.nsgpu_domain_size(h)):
            st, ct = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
            bh = self.L.nsgpu_domain_block(h, r, st, ct)
            st, ct = tuple(st), tuple(ct)
            self.blocks.append(_Block(self.L, ctypes.c_void_p(bh),
                                      (ct[0] + 6, ct[1] + 6, ct[2] + 6, 5), st, ct))

    def close(self):
        if getattr(self, "h", None):
            self.L.nsgpu_domain_destroy(self.h)
            self.h = None
            self.blocks = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _slices(self, b):
        return tuple(slice(3 + s, 3 + s + c) for s, c in zip(b.start, b.count))

    def set_state(self, A, kind="primitive"):
        A = np.ascontiguousarray(A, dtype=np.float64)
        if A.shape != self.shape:
            raise ValueError(f"array shape {A.shape} != {self.shape}")
        for b in self.blocks:
            local = np.zeros(b.shape)
            local[3:-3, 3:-3, 3:-3] = A[self._slices(b)]
            b.set_state(local, kind)

    def get_state(self, kind="primitive"):
        out = np.zeros(self.shape)
        for b in self.blocks:
            out[self._slices(b)] = b.get_state(kind)[3:-3, 3:-3, 3:-3]
        return out

    def step(self, nsteps=1, time=0.0):
        """Advances every block; returns (time, last dt); raises on a failed health check."""
        t = np.array([time], dtype=np.float64)
        dt = np.zeros(1)
        rc = self.L.nsgpu_domain_step(self.h, nsteps, _ptr(t), _ptr(dt))
        if rc != 0:
            raise NsgpuError(rc, self.L.nsgpu_strerror(rc).decode())
        return float(t[0]), float(dt[0])

    def reduce(self):
        """Conservation sums over all blocks, the CFL dt of the whole domain, the OR of the
        blocks' health flags and the worst status."""
        sums, dts, flags = np.zeros(5), [], 0
        for b in self.blocks:
            s, dt, fl, _ = b.reduce()
            sums += s
            dts.append(dt)
            flags |= fl
        status = -4 if flags & 1 else (-5 if flags & 2 else (-6 if flags & 4 else 0))
        return sums, min(dts), flags, status
