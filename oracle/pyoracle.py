"""ctypes loader for the CPU oracle (lib/libnsoracle.so) and, where it was built, the reference
harness (_ref/libnsref_strict.so, _ref/libnsref_fast.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product path never imports this module.

All arrays are numpy float64 in the reference Array4d layout: shape (NX, NY, NZ, 5) with
NX = nx + 6 (containers/array_4d.hpp:116-124, array_4d_dimensions.hpp:40-49).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NGC = 3

BC = {"periodic": 0, "neumann": 1}
CONV = {"none": 0, "weno5": 1, "weno5cw": 2}
WEIGHTS = {"shu": 0, "borges": 1, "learned": 2}
MLP_NPARAM = 32 * 8 + 32 + 32 * 32 + 32 + 3 * 32 + 3
REGIMES = {"smooth": 0, "mlp": 1, "flagged": 2}
VISCOSITY = {"constant": 0, "sutherland": 1}
SUTH_C = 110.4 / 300.0  # Sutherland constant of air over a 300 K reference temperature
ALPHA_MODE = {"const3": 0, "local_max": 1}
PHYSICS = {"reference": 0, "corrected_convection": 1, "corrected_viscous": 2, "corrected": 3}


class NsoCfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dim", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("gamma", ctypes.c_double), ("mu", ctypes.c_double), ("Pr", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("conv", ctypes.c_int), ("weights", ctypes.c_int), ("diff_order", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("viscosity", ctypes.c_int), ("suth_C", ctypes.c_double),
        ("alpha_mode", ctypes.c_int), ("physics", ctypes.c_int),
    ]


def make_cfg(case):
    """case: dict with nx, ny, nz, dim, L (3-tuple), gamma, mu, Pr, cfl, conv, weights,
    diff_order, bc (6 names, W,E,N,S,F,B)."""
    c = NsoCfg()
    c.nx, c.ny, c.nz = case["nx"], case["ny"], case["nz"]
    c.dim = case.get("dim", 3)
    c.Lx, c.Ly, c.Lz = case["L"]
    c.gamma, c.mu, c.Pr, c.cfl = case["gamma"], case["mu"], case["Pr"], case["cfl"]
    c.conv = CONV[case["conv"]]
    c.weights = WEIGHTS[case["weights"]]
    c.diff_order = case["diff_order"]
    for f, name in enumerate(case["bc"]):
        c.bc[f] = BC[name]
    c.viscosity = VISCOSITY[case.get("viscosity", "constant")]
    c.suth_C = case.get("suth_C", SUTH_C)
    c.alpha_mode = ALPHA_MODE[case.get("alpha_mode", "const3")]
    c.physics = PHYSICS[case.get("physics", "reference")]
    return c


def padded_shape(case):
    return (case["nx"] + 2 * NGC, case["ny"] + 2 * NGC, case["nz"] + 2 * NGC, 5)


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _Lib:
    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        for name in ("rhs", "step"):
            fn = getattr(self.lib, prefix + name)
            fn.restype = ctypes.c_int
        nt = getattr(self.lib, prefix + "num_threads")
        nt.restype = ctypes.c_int

    def num_threads(self):
        return getattr(self.lib, self.prefix + "num_threads")()

    def rhs(self, case, U, which=3):
        """One stage RHS from conservative U (padded, interior used).  Returns (U_filled, W, R)."""
        cfg = make_cfg(case)
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        W = np.zeros_like(U)
        R = np.zeros_like(U)
        rc = getattr(self.lib, self.prefix + "rhs")(ctypes.byref(cfg), _ptr(U), _ptr(W), _ptr(R),
                                                    ctypes.c_int(which))
        if rc != 0:
            raise ValueError("rhs rejected the configuration")
        return U, W, R

    def step(self, case, W, nsteps):
        """nsteps SSP-RK3 steps from primitive W.  Returns (W, U, dts)."""
        cfg = make_cfg(case)
        W = np.ascontiguousarray(W, dtype=np.float64).copy()
        U = np.zeros_like(W)
        dts = np.zeros(max(nsteps, 1), dtype=np.float64)
        rc = getattr(self.lib, self.prefix + "step")(ctypes.byref(cfg), _ptr(W), _ptr(U),
                                                     ctypes.c_int(nsteps), _ptr(dts))
        if rc != 0:
            raise ValueError("step rejected the configuration")
        return W, U, dts[:nsteps]


class Oracle(_Lib):
    def __init__(self):
        path = os.path.join(HERE, "lib", "libnsoracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` (or build())")
        super().__init__(path, "nso_")
        L = self.lib
        L.nso_dt.restype = ctypes.c_double
        L.nso_weno5_left.restype = None
        L.nso_weno5_right.restype = None
        L.nso_mu_sutherland.restype = ctypes.c_double
        L.nso_mu_sutherland.argtypes = [ctypes.c_double] * 5
        L.nso_bf16.restype = ctypes.c_double
        L.nso_bf16.argtypes = [ctypes.c_double]
        L.nso_mlp_regime.restype = ctypes.c_int
        L.nso_mlp_ratio.restype = ctypes.c_double
        L.nso_mlp_blend.restype = ctypes.c_double
        L.nso_mlp_blend.argtypes = [ctypes.c_double]
        L.nso_mlp_loss_grad.restype = ctypes.c_double
        L.nso_mlp_loss_grad.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_long] + \
            [ctypes.POINTER(ctypes.c_double)] * 3 + [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]

    def dt(self, case, W):
        cfg = make_cfg(case)
        W = np.ascontiguousarray(W, dtype=np.float64)
        return self.lib.nso_dt(ctypes.byref(cfg), _ptr(W))

    def mu_sutherland(self, mu0, C, gamma, rho, p):
        return self.lib.nso_mu_sutherland(mu0, C, gamma, rho, p)

    def p2c(self, case, W):
        cfg = make_cfg(case)
        W = np.ascontiguousarray(W, dtype=np.float64)
        U = np.zeros_like(W)
        self.lib.nso_p2c(ctypes.byref(cfg), _ptr(W), _ptr(U))
        return U

    def c2p(self, case, U):
        cfg = make_cfg(case)
        U = np.ascontiguousarray(U, dtype=np.float64)
        W = np.zeros_like(U)
        self.lib.nso_c2p(ctypes.byref(cfg), _ptr(U), _ptr(W))
        return W

    def fill_halos(self, case, A):
        cfg = make_cfg(case)
        A = np.ascontiguousarray(A, dtype=np.float64).copy()
        self.lib.nso_fill_halos(ctypes.byref(cfg), _ptr(A))
        return A

    def weno_left(self, f, weights="shu"):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(1)
        self.lib.nso_weno5_left(_ptr(f), ctypes.c_int(WEIGHTS[weights]), _ptr(out))
        return out[0]

    def weno_right(self, f, weights="shu"):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(1)
        self.lib.nso_weno5_right(_ptr(f), ctypes.c_int(WEIGHTS[weights]), _ptr(out))
        return out[0]

    def eigenvectors(self, Wavg, gamma, direction):
        Wavg = np.ascontiguousarray(Wavg, dtype=np.float64)
        L = np.zeros(25)
        R = np.zeros(25)
        self.lib.nso_eigenvectors(_ptr(Wavg), ctypes.c_double(gamma), ctypes.c_int(direction),
                                  _ptr(L), _ptr(R))
        return L.reshape(5, 5), R.reshape(5, 5)

    def face_eigensystem(self, case, Wl, Wr, direction):
        """(face state, L, R) of the characteristic sweep between two cells; L and R act on
        vectors in the stored order [rho, rho E, rho u, rho v, rho w]."""
        cfg = make_cfg(case)
        Wl = np.ascontiguousarray(Wl, dtype=np.float64)
        Wr = np.ascontiguousarray(Wr, dtype=np.float64)
        Wf, L, R = np.zeros(5), np.zeros(25), np.zeros(25)
        self.lib.nso_face_eigensystem(ctypes.byref(cfg), _ptr(Wl), _ptr(Wr), ctypes.c_int(direction),
                                      _ptr(Wf), _ptr(L), _ptr(R))
        return Wf, L.reshape(5, 5), R.reshape(5, 5)

    def face_alpha(self, case, W, direction, i, j, k):
        """Lax-Friedrichs constant of face i+1/2 (padded cell indices) along `direction`."""
        cfg = make_cfg(case)
        W = np.ascontiguousarray(W, dtype=np.float64)
        f = self.lib.nso_face_alpha
        f.restype = ctypes.c_double
        return f(ctypes.byref(cfg), _ptr(W), ctypes.c_int(direction), ctypes.c_long(i),
                 ctypes.c_long(j), ctypes.c_long(k))

    # ---- learned weights ------------------------------------------------------------------
    def bf16(self, v):
        return self.lib.nso_bf16(float(v))

    def mlp_features(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        x = np.zeros(8)
        self.lib.nso_mlp_features(_ptr(v), _ptr(x))
        return x

    def mlp_features_s(self, v, scale):
        """mlp_features with the range also measured against `scale` (the momentum scale of a
        component-wise face)"""
        v = np.ascontiguousarray(v, dtype=np.float64)
        x = np.zeros(8)
        self.lib.nso_mlp_features_s(_ptr(v), ctypes.c_double(float(scale)), _ptr(x))
        return x

    def mlp_face_scale(self, fl, fr):
        """momentum scale of a component-wise face from its split stencils fl, fr [5][5]"""
        f = self.lib.nso_mlp_face_scale
        f.restype = ctypes.c_double
        return f(_ptr(np.ascontiguousarray(fl, dtype=np.float64)),
                 _ptr(np.ascontiguousarray(fr, dtype=np.float64)))

    def mlp_regime(self, x):
        return self.lib.nso_mlp_regime(_ptr(np.ascontiguousarray(x, dtype=np.float64)))

    def mlp_ratio(self, x):
        return self.lib.nso_mlp_ratio(_ptr(np.ascontiguousarray(x, dtype=np.float64)))

    def mlp_blend(self, ratio):
        return self.lib.nso_mlp_blend(float(ratio))

    def mlp_candidates(self, x):
        q = np.zeros(3)
        self.lib.nso_mlp_candidates(_ptr(np.ascontiguousarray(x, dtype=np.float64)), _ptr(q))
        return q

    def mlp_forward(self, params, x, bf16=False):
        w = np.zeros(3)
        self.lib.nso_mlp_forward(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                 _ptr(np.ascontiguousarray(x, dtype=np.float64)),
                                 ctypes.c_int(int(bf16)), _ptr(w))
        return w

    def mlp_loss_grad(self, params, X, Q, T, bf16=False, grad=True):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        g = np.zeros(MLP_NPARAM) if grad else None
        loss = self.lib.nso_mlp_loss_grad(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                          len(T), _ptr(X), _ptr(Q), _ptr(T), int(bf16),
                                          _ptr(g) if grad else None)
        return (loss, g) if grad else loss

    def weno_learned(self, params, f, side, bf16=True):
        """(value, regime) of the learned reconstruction; side 0 = left-, 1 = right-biased."""
        out = np.zeros(1)
        rg = ctypes.c_int(-1)
        self.lib.nso_weno5_learned(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                                   ctypes.c_int(int(bf16)),
                                   _ptr(np.ascontiguousarray(f, dtype=np.float64)),
                                   ctypes.c_int(side), _ptr(out), ctypes.byref(rg))
        return out[0], rg.value

    def set_mlp(self, params, bf16=True):
        self.lib.nso_set_mlp(_ptr(np.ascontiguousarray(params, dtype=np.float64)),
                             ctypes.c_int(int(bf16)))

    def fd(self, case, f, direction, second, d):
        cfg = make_cfg(case)
        f = np.ascontiguousarray(f, dtype=np.float64)
        df = np.zeros_like(f)
        self.lib.nso_fd(ctypes.byref(cfg), _ptr(f), _ptr(df), ctypes.c_int(direction),
                        ctypes.c_int(second), ctypes.c_double(d))
        return df


class Reference(_Lib):
    """The reference's own classes (strict or fast build), if oracle/_ref was built here."""

    def __init__(self, flavour="strict"):
        path = os.path.join(HERE, "_ref", f"libnsref_{flavour}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        super().__init__(path, "ref_")
        self.lib.ref_testcase.restype = ctypes.c_int

    def write_binary(self, case, W, it, directory):
        """The reference's BinaryWriter output (mesh and fields files) of W into directory/results."""
        cfg = make_cfg(case)
        W = np.ascontiguousarray(W, dtype=np.float64)
        if self.lib.ref_write_binary(ctypes.byref(cfg), _ptr(W), ctypes.c_int(it),
                                     str(directory).encode()) != 0:
            raise OSError("reference writer failed")

    def testcase(self, case, ident):
        """The reference's initial condition number `ident` (testcases.cpp) on the block."""
        cfg = make_cfg(case)
        W = np.zeros(padded_shape(case))
        if self.lib.ref_testcase(ctypes.byref(cfg), ctypes.c_int(ident), _ptr(W)) != 0:
            raise ValueError("unknown test case")
        return W


class Adapters:
    """The reference's time loop driving the GPU adapters (oracle/_ref/libnsref_adapters.so,
    adapter_harness.cpp): with GpuConvection / GpuDiffusion as its schemes, or GpuTimeScheme."""

    def __init__(self):
        path = os.path.join(HERE, "_ref", "libnsref_adapters.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle adapters`")
        self.lib = ctypes.CDLL(path)
        self.lib.adapt_make_config.argtypes = [ctypes.c_void_p, ctypes.c_char_p]

    def _step(self, fn, case, W, nsteps):
        cfg = make_cfg(case)
        W = np.ascontiguousarray(W, dtype=np.float64).copy()
        dts = np.zeros(max(nsteps, 1))
        rc = fn(ctypes.byref(cfg), _ptr(W), ctypes.c_int(nsteps), _ptr(dts))
        if rc != 0:
            raise RuntimeError(f"adapter run failed ({rc})")
        return W, dts[:nsteps]

    def step_schemes(self, case, W, nsteps):
        """(W, dts) after nsteps of TimeScheme::step with GpuConvection / GpuDiffusion."""
        return self._step(self.lib.adapt_step_schemes, case, W, nsteps)

    def step_timescheme(self, case, W, nsteps):
        """(W, dts) after nsteps of GpuTimeScheme::step."""
        return self._step(self.lib.adapt_step_timescheme, case, W, nsteps)

    def make_config(self, case, scheme):
        """0: make_config accepts convection_scheme `scheme`; 1: it throws invalid_argument."""
        cfg = make_cfg(case)
        return self.lib.adapt_make_config(ctypes.byref(cfg), scheme.encode())


def adapters_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libnsref_adapters.so"))


def reference_available(flavour="strict"):
    return os.path.exists(os.path.join(HERE, "_ref", f"libnsref_{flavour}.so"))


def interior(A):
    return A[NGC:-NGC, NGC:-NGC, NGC:-NGC, :]


def faces_mask(case):
    """Boolean mask (NX,NY,NZ) of interior + face halos (no edges / corners)."""
    NX, NY, NZ, _ = padded_shape(case)
    ii = np.zeros(NX, bool); ii[NGC:-NGC] = True
    jj = np.zeros(NY, bool); jj[NGC:-NGC] = True
    kk = np.zeros(NZ, bool); kk[NGC:-NGC] = True
    I, J, K = np.meshgrid(ii, jj, kk, indexing="ij")
    n_in = I.astype(int) + J.astype(int) + K.astype(int)
    return n_in >= 2
