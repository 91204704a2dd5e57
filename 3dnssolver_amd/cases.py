"""Seeded synthetic workloads for the RHS + SSP-RK3 path.

A case is a plain dict (the fields of the reference's Params that the path reads, params.hpp:20-43):
  nx, ny, nz      interior cells of the block
  dim             3, or 2 (z sweep skipped, weno-5.cpp:28)
  L               (Lx, Ly, Lz); dx = Lx / nx (grid.cpp:14-16)
  gamma, mu, Pr   fluid properties; cfl = CFL_max
  conv            'none' | 'weno5' | 'weno5cw'
  weights         'shu' | 'borges'
  diff_order      0 (no diffusion) | 2 | 4 | 6
  bc              six of 'periodic' | 'neumann', faces West, East, North, South, Front, Back

Primitive fields W are numpy float64 arrays in the reference Array4d layout (NX, NY, NZ, 5),
NX = nx + 6, variables [rho, p, u, v, w]; halos are left zero like a fresh Domain array.

Nondimensionalisation.  The reference's Lax-Friedrichs split uses alpha = 3.0 everywhere
(convection_flux.cpp:241).  Its 3-D scheme is only stable at CFL 0.5 when the largest wave speed is
close to alpha, so every case sets rho0 = 1 and the sound speed c0 = 3 / (1 + M): the initial
max |u| + c is then 3.  Velocities scale with U0 = M * c0, pressure p0 = rho0 * c0^2 / gamma and
mu = rho0 * U0 * Lref / Re.
"""
import math

import numpy as np

NGC = 3
ALPHA = 3.0


def base_case(nx, ny, nz):
    return dict(nx=nx, ny=ny, nz=nz, dim=3, L=(1.0, 1.0, 1.0), gamma=1.4, mu=1e-3, Pr=0.72,
                cfl=0.5, conv="weno5cw", weights="shu", diff_order=4, bc=["periodic"] * 6)


def padded_shape(case):
    return (case["nx"] + 2 * NGC, case["ny"] + 2 * NGC, case["nz"] + 2 * NGC, 5)


def _centres(case):
    Lx, Ly, Lz = case["L"]
    x = (np.arange(case["nx"]) + 0.5) * Lx / case["nx"]
    y = (np.arange(case["ny"]) + 0.5) * Ly / case["ny"]
    z = (np.arange(case["nz"]) + 0.5) * Lz / case["nz"]
    return np.meshgrid(x, y, z, indexing="ij")


def _empty(case):
    return np.zeros(padded_shape(case), dtype=np.float64)


def sound_speed(M):
    return ALPHA / (1.0 + M)


def random_smooth_W(case, seed, c0=2.2):
    """Smooth waves, a steep tanh front and seeded noise in every variable: exercises the
    nonlinear WENO weights and all viscous terms.  max|u| + c stays below alpha = 3."""
    rng = np.random.default_rng(seed)
    X, Y, Z = _centres(case)
    Lx, Ly, Lz = case["L"]
    sh = X.shape
    x, y, z = X / Lx, Y / Ly, Z / Lz
    W = _empty(case)
    I = (slice(NGC, -NGC),) * 3
    front = np.tanh((x - 0.5) / 0.05)
    W[I + (0,)] = 1.0 + 0.2 * np.sin(2 * np.pi * (x + y)) + 0.1 * front + 0.02 * rng.standard_normal(sh)
    W[I + (1,)] = c0 * c0 / case["gamma"] * (1.0 + 0.1 * np.cos(2 * np.pi * (y + z)) + 0.05 * front
                                             + 0.01 * rng.standard_normal(sh))
    W[I + (2,)] = 0.3 * np.sin(2 * np.pi * z) + 0.03 * rng.standard_normal(sh)
    W[I + (3,)] = 0.2 * np.cos(2 * np.pi * x) + 0.03 * rng.standard_normal(sh)
    W[I + (4,)] = 0.1 * np.sin(2 * np.pi * y) + 0.03 * rng.standard_normal(sh)
    return W


def tgv_W(case, M=0.1, seed=2, c0=None):
    """Taylor-Green vortex on [0, 2pi]^3 with a seeded 1e-6 uniform density perturbation."""
    c0 = sound_speed(M) if c0 is None else c0
    V = M * c0
    rng = np.random.default_rng(seed)
    X, Y, Z = _centres(case)
    W = _empty(case)
    I = (slice(NGC, -NGC),) * 3
    W[I + (0,)] = 1.0 + 1e-6 * (2.0 * rng.random(X.shape) - 1.0)
    W[I + (1,)] = c0 * c0 / case["gamma"] + (V * V / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    W[I + (2,)] = V * np.sin(X) * np.cos(Y) * np.cos(Z)
    W[I + (3,)] = -V * np.cos(X) * np.sin(Y) * np.cos(Z)
    return W


def shear_W(case, M=0.3, amp=0.05, seed=1, c0=None, smooth_phase=False):
    """3-D periodic double shear layer on [0,1]^3: u = U0 tanh((|y-0.5|-0.25)/0.02),
    v = amp * U0 * sin(2 pi x + phi(z)) with a seeded phase: drawn independently per z plane,
    or with smooth_phase a seeded smooth periodic function of z (three Fourier modes), which
    describes the same field at every resolution (for comparisons across grids)."""
    c0 = sound_speed(M) if c0 is None else c0
    V = M * c0
    rng = np.random.default_rng(seed)
    if smooth_phase:
        z = (np.arange(case["nz"]) + 0.5) / case["nz"]
        a, b = rng.random(3), rng.random(3) * 2.0 * np.pi
        phase = sum(np.pi * a[m] / (m + 1) * np.sin(2 * np.pi * (m + 1) * z + b[m]) for m in range(3))
    else:
        phase = rng.random(case["nz"]) * 2.0 * np.pi
    X, Y, Z = _centres(case)
    Ly = case["L"][1]
    Lx = case["L"][0]
    W = _empty(case)
    I = (slice(NGC, -NGC),) * 3
    W[I + (0,)] = 1.0
    W[I + (1,)] = c0 * c0 / case["gamma"]
    W[I + (2,)] = V * np.tanh((np.abs(Y / Ly - 0.5) - 0.25) / 0.02)
    W[I + (3,)] = amp * V * np.sin(2 * np.pi * X / Lx + phase[None, None, :])
    return W


def config1(n=64):
    """Plain method, CPU-runnable: double shear layer, 64^3, [0,1]^3, component-wise WENO-5 (Shu)
    with LF splitting, CD4 diffusion, Re = 1000, Mach 0.3, CFL 0.5; run for 100 steps."""
    M, Re = 0.3, 1000.0
    c0 = sound_speed(M)
    case = dict(base_case(n, n, n), conv="weno5", weights="shu", diff_order=4,
                mu=1.0 * M * c0 * 1.0 / Re)
    return case, shear_W(case, M=M, amp=0.05, seed=1)


def sutherland(case, mu_ref, T_ref, C_rel=110.4 / 300.0):
    """Switches a case to the Sutherland law with mu = mu_ref at the temperature T_ref (the
    library's temperature is T = gamma p / rho, the square of the sound speed) and a Sutherland
    constant C_rel * T_ref (air: 110.4 K over a 300 K reference).  The library evaluates
    mu(T) = mu (1 + C) / (T + C) * T^1.5, so mu and C are rescaled to put the reference point at
    T_ref."""
    C = C_rel * T_ref
    case.update(viscosity="sutherland", suth_C=C,
                mu=mu_ref * (T_ref + C) / ((1.0 + C) * T_ref * math.sqrt(T_ref)))
    return case


TGV_SPEED_FRAC = 0.97


def config2(n=256, viscosity="constant"):
    """Plain method, one GPU subdomain: Taylor-Green vortex, 256^3, [0,2pi]^3, characteristic
    WENO-5 (Shu) + CD4 viscous terms, Mach 0.1, Re = 1600, CFL 0.5; 200 steps.  viscosity
    "constant" is the reference's path (mu fixed); "sutherland" gives mu = rho0 U0 L / Re at the
    initial temperature c0^2 and lets it follow the local temperature."""
    M, Re = 0.1, 1600.0
    # 3 % below the LF constant: the vortex's |u| + c then stays below alpha = 3 over the run
    # (at c0 = 3 / (1 + M) it starts exactly at alpha and the split is non-upwind in some cells)
    c0 = TGV_SPEED_FRAC * sound_speed(M)
    case = dict(base_case(n, n, n), L=(2 * math.pi,) * 3, conv="weno5cw", weights="shu",
                diff_order=4, mu=1.0 * M * c0 * 1.0 / Re)
    if viscosity == "sutherland":
        sutherland(case, case["mu"], c0 * c0)
    elif viscosity != "constant":
        raise ValueError(viscosity)
    return case, tgv_W(case, M=M, seed=2, c0=c0)


def isotropic_W(case, Mt=0.5, k0=4.0, seed=3, c0=None, speed_frac=1.0):
    """Decaying isotropic turbulence on [0, 2pi]^3: seeded random-phase solenoidal velocity with
    energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2), uniform density and pressure, rms velocity
    Mt * c0.  Returns (W, u_prime, taylor_lambda) with u' the rms of one component and
    lambda = u' / rms(du/dx) (spectral derivative)."""
    n = case["nx"]
    assert case["ny"] == n and case["nz"] == n
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kr = np.fft.rfftfreq(n, 1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, kr, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K = np.sqrt(K2)
    K2[0, 0, 0] = 1.0
    amp = np.sqrt(K ** 4 * np.exp(-2.0 * (K / k0) ** 2) / (4.0 * np.pi * K2))
    amp[0, 0, 0] = 0.0
    comps = []
    for _ in range(3):
        ph = rng.random(K.shape) * 2.0 * np.pi
        comps.append(amp * np.exp(1j * ph) * rng.standard_normal(K.shape))
    # solenoidal projection: u_hat -= k (k . u_hat) / k^2
    dot = (KX * comps[0] + KY * comps[1] + KZ * comps[2]) / K2
    comps = [comps[0] - KX * dot, comps[1] - KY * dot, comps[2] - KZ * dot]
    vel = [np.fft.irfftn(cc, s=(n, n, n), axes=(0, 1, 2)) for cc in comps]
    urms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    if c0 is None:
        # c0 with max |u_i| + c0 = speed_frac * alpha (alpha = 3, the LF split constant)
        vmax = max(np.abs(v).max() for v in vel) / urms
        c0 = speed_frac * ALPHA / (1.0 + Mt * vmax)
    scale = Mt * c0 / urms
    vel = [v * scale for v in vel]
    dudx = np.fft.irfftn(1j * KX * np.fft.rfftn(vel[0]), s=(n, n, n), axes=(0, 1, 2))
    uprime = np.sqrt(np.mean(vel[0] ** 2))
    lam = uprime / np.sqrt(np.mean(dudx ** 2))
    W = _empty(case)
    I = (slice(NGC, -NGC),) * 3
    W[I + (0,)] = 1.0
    W[I + (1,)] = c0 * c0 / case["gamma"]
    for a in range(3):
        W[I + (2 + a,)] = vel[a]
    return W, uprime, lam


def config3(n=512, speed_frac=0.85, cfl=0.3):
    """Precision form: decaying isotropic turbulence, n^3 periodic on [0, 2pi]^3, turbulent Mach
    0.5, Re_lambda = 100, characteristic WENO-5 (Shu) + CD4; mu = rho u' lambda / 100.

    The time step is the reference's (CFL x the minimum over cells and axes of dx / max|u +- c|,
    time_scheme.cpp:98-136) while the flux split uses the fixed alpha = 3 on every axis, so the
    linear stability of WENO-5 + SSP-RK3 needs 3 * CFL * alpha / max|u +- c| below about 1.4.
    With max|u_i| + c0 = 0.85 alpha at the start (headroom for the shocklets' rise in |u| + c),
    CFL 0.3 keeps that number near 1.1; at CFL 0.5 this flow blows up within 150 steps in FP64
    with either WENO form, with or without viscosity (measured at 256^3 and 512^3)."""
    case = dict(base_case(n, n, n), L=(2 * math.pi,) * 3, conv="weno5cw", weights="shu",
                diff_order=4, cfl=cfl)
    W, uprime, lam = isotropic_W(case, Mt=0.5, k0=4.0, seed=3, speed_frac=speed_frac)
    case["mu"] = 1.0 * uprime * lam / 100.0
    return case, W


ENSEMBLE_EDGES = (32, 48, 64, 96, 128)


def config5(n_blocks=64, seed=5, edges=ENSEMBLE_EDGES):
    """Batched form: n_blocks independent periodic double shear layers on [0,1]^3, each edge
    length drawn uniformly from `edges` (independently per axis), Mach uniform in [0.2, 0.6],
    Re uniform in [500, 5000], perturbation amplitude uniform in [0.01, 0.1]; component-wise
    WENO-5 (Shu) + CD4.  Returns a list of (case, W, params).

    Sound speed c0 = 0.9 alpha / (1 + M), so max|u| + c starts 10 % below the LF constant, and
    CFL 0.3: with the reference's time step (minimum over axes, not a sum) the stability number
    3 * CFL * alpha / max|u +- c| reaches 2 on blocks whose finest axis is not the flow axis, and
    such blocks amplify round-off at CFL 0.5; at 0.3 every block stays below 1.2."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_blocks):
        nx, ny, nz = (int(v) for v in rng.choice(edges, size=3))
        M = float(rng.uniform(0.2, 0.6))
        Re = float(rng.uniform(500.0, 5000.0))
        amp = float(rng.uniform(0.01, 0.1))
        c0 = 0.9 * sound_speed(M)
        case = dict(base_case(nx, ny, nz), conv="weno5", weights="shu", diff_order=4,
                    mu=1.0 * M * c0 * 1.0 / Re, cfl=0.3)
        W = shear_W(case, M=M, amp=amp, seed=int(rng.integers(1 << 30)), c0=c0)
        out.append((case, W, dict(M=M, Re=Re, amp=amp)))
    return out


def jump_W(case, seed=4, rho_ratio=8.0, p_ratio=10.0, width=0.02):
    """Sod-like planar density / pressure jump at a seeded random orientation, periodic: the
    slabs where sin(2 pi k.x + phi) > 0 hold the high state (rho 1, p 1), the others the low one
    (rho 1/rho_ratio, p 1/p_ratio); k is a random integer lattice vector (so the planes are
    periodic in [0,1]^3) and the jump is smoothed over `width` (a tanh of the phase)."""
    rng = np.random.default_rng(seed)
    while True:
        k = rng.integers(-2, 3, size=3)
        if (k != 0).sum() >= 2:
            break
    phi = rng.random() * 2.0 * np.pi
    X, Y, Z = _centres(case)
    Lx, Ly, Lz = case["L"]
    s = np.sin(2 * np.pi * (k[0] * X / Lx + k[1] * Y / Ly + k[2] * Z / Lz) + phi)
    h = 0.5 * (1.0 + np.tanh(s / width))
    W = _empty(case)
    I = (slice(NGC, -NGC),) * 3
    W[I + (0,)] = 1.0 / rho_ratio + (1.0 - 1.0 / rho_ratio) * h
    W[I + (1,)] = 1.0 / p_ratio + (1.0 - 1.0 / p_ratio) * h
    return W


def config4(n=256, problem="shear"):
    """Learned form evaluation problems on [0,1]^3: the double shear layer of config1 at n^3
    (component-wise WENO-5 + CD4, Mach 0.3, Re 1000; the perturbation phase a smooth seeded
    function of z so that grids of different resolution hold the same field) or the Sod-like
    planar jump (density 8:1, pressure 10:1, seeded orientation) with the same scheme."""
    case, W = config1(n)
    if problem == "jump":
        case = dict(case, mu=1e-4)
        W = jump_W(case)
    else:
        W = shear_W(case, M=0.3, amp=0.05, seed=1, smooth_phase=True)
    return case, W


# ---- the reference's initial conditions (testcases.cpp:5-180) ---------------------------------
# states [rho, p, u, v, w] of the quadrants NW (x <= Lx/2, y <= Ly/2), NE, SW, SE in the x-y
# plane (the z-direction Sod uses the z-y plane).  North is the j-low side, as in the
# reference's boundary naming.  Configuration 17 of the reference lists only four values for SW
# and reads a fifth past the end of its vector (testcases.cpp:113); w = 0 stands in here.
_SOD = ((1.0, 1.0, 0.0, 0.0, 0.0), (0.125, 0.1, 0.0, 0.0, 0.0))
RIEMANN = {
    "sod_x": (_SOD[0], _SOD[1], _SOD[0], _SOD[1], "xy"),
    "sod_y": (_SOD[0], _SOD[0], _SOD[1], _SOD[1], "xy"),
    "sod_z": (_SOD[0], _SOD[1], _SOD[0], _SOD[1], "zy"),
    "case3": ((0.5323, 0.3, 1.206, 0.0, 0.0), (1.5, 1.5, 0.0, 0.0, 0.0),
              (0.138, 0.029, 1.206, -1.206, 0.0), (0.5323, 0.3, 0.0, -1.206, 0.0), "xy"),
    "case4": ((0.5065, 0.35, 0.8939, 0.0, 0.0), (1.1, 1.1, 0.0, 0.0, 0.0),
              (1.1, 1.1, 0.8939, -0.8938, 0.0), (0.5065, 0.35, 0.0, -0.8939, 0.0), "xy"),
    "case6": ((2.0, 1.0, 0.75, -0.5, 0.0), (1.0, 1.0, 0.75, 0.5, 0.0),
              (1.0, 1.0, -0.75, -0.5, 0.0), (3.0, 1.0, -0.75, 0.5, 0.0), "xy"),
    "case12": ((1.0, 1.0, 0.7276, 0.0, 0.0), (0.5313, 0.4, 0.0, 0.0, 0.0),
               (0.8, 1.0, 0.0, 0.0, 0.0), (1.0, 1.0, 0.0, -0.7276, 0.0), "xy"),
    "case15": ((0.5197, 0.4, -0.6259, 0.3, 0.0), (1.0, 1.0, 0.1, 0.3, 0.0),
               (0.8, 0.4, 0.1, 0.3, 0.0), (0.5313, 0.4, 0.1, -0.4276, 0.0), "xy"),
    "case17": ((2.0, 1.0, 0.0, 0.3, 0.0), (1.0, 1.0, 0.0, 0.4, 0.0),
               (1.0625, 0.4, 0.0, -0.2145, 0.0), (0.5197, 0.4, 0.0, 1.1259, 0.0), "xy"),
}
RIEMANN_IDS = {name: i for i, name in enumerate(RIEMANN)}  # the reference harness's numbering


def reference_coords(case):
    """Cell centres as Grid computes them (grid.cpp:40-56): x = dx * 0.5 + i * dx, block origin 0."""
    out = []
    for n, L in zip((case["nx"], case["ny"], case["nz"]), case["L"]):
        d = L / n
        out.append(d * 0.5 + np.arange(n) * d)
    return np.meshgrid(*out, indexing="ij")


def riemann_W(case, name):
    """The reference's quadrant initial condition `name` (a key of RIEMANN) on the block."""
    NW, NE, SW, SE, plane = RIEMANN[name]
    X, Y, Z = reference_coords(case)
    Lx, Ly, Lz = case["L"]
    a, La = (X, Lx) if plane == "xy" else (Z, Lz)
    lo_a, lo_y = a <= La / 2.0, Y <= Ly / 2.0
    W = _empty(case)
    I = (slice(NGC, -NGC),) * 3
    for q in range(5):
        W[I + (q,)] = np.where(lo_a & lo_y, NW[q], np.where(~lo_a & lo_y, NE[q],
                                                         np.where(lo_a & ~lo_y, SW[q], SE[q])))
    return W


def riemann_case(name, n=64, nz=4):
    """A walled block for the reference's Riemann problems: [0,1]^2 x [0, nz/n], Neumann faces in
    x and y, periodic in z (the 2-D configurations are uniform in z), component-wise WENO-5,
    inviscid, CFL 0.5."""
    case = dict(base_case(n, n, nz), L=(1.0, 1.0, nz / n), conv="weno5", diff_order=0, mu=0.0,
                bc=["neumann", "neumann", "neumann", "neumann", "periodic", "periodic"])
    if name == "sod_z":
        case = dict(base_case(nz, n, n), L=(nz / n, 1.0, 1.0), conv="weno5", diff_order=0,
                    mu=0.0, bc=["periodic", "periodic", "neumann", "neumann", "neumann", "neumann"])
    return case, riemann_W(case, name)


# ---- the reference's shear layers (testcases.cpp:250-344) ---------------------------------------
# They draw their perturbations from the C library's rand() (unseeded, so seed 1, RAND_MAX =
# 2^31 - 1) once or twice per cell in i, j, k order over each rank's block: every rank of a
# decomposed run starts the sequence afresh.  CRand restates that generator (the additive
# feedback generator of the GNU C library, its default TYPE_3 state: r[0] = seed,
# r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31, r[i] = r[i-31] for i = 31..33,
# r[i] = r[i-3] + r[i-31] mod 2^32 after that, outputs r[i] >> 1 from i = 344 on).
RAND_MAX = 2147483647


class CRand:
    """The C library's rand() sequence from srand(seed) (pure Python: ~0.1 us per value)."""

    def __init__(self, seed=1):
        seed = seed or 1
        r = [seed]
        for i in range(1, 31):
            hi, lo = divmod(r[i - 1], 127773)
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r.append(word)
        r += r[:3]
        for i in range(34, 344):
            r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
        self.r = r[-31:]  # the last 31 values: the lags the recurrence reads

    def take(self, n):
        """The next n outputs as a float64 array divided by RAND_MAX, as (double)rand() /
        (RAND_MAX) evaluates them."""
        r = self.r
        out = np.empty(n, dtype=np.float64)
        for k in range(n):
            v = (r[-31] + r[-3]) & 0xFFFFFFFF
            r.append(v)
            out[k] = (v >> 1)
        self.r = r[-31:]
        return out / float(RAND_MAX)


def block_coords(case, start=(0, 0, 0), count=None):
    """Cell centres of a block as Grid computes them (grid.cpp:24-57): x = (x0 + dx * 0.5) +
    i * dx with x0 = dx * start_i and dx = L / n of the whole domain."""
    count = count or (case["nx"], case["ny"], case["nz"])
    out = []
    for n, L, s, c in zip((case["nx"], case["ny"], case["nz"]), case["L"], start, count):
        d = L / n
        x0 = d * s
        out.append((x0 + d * 0.5) + np.arange(c) * d)
    return np.meshgrid(*out, indexing="ij")


def shear_layer_W(case, name, start=(0, 0, 0), count=None, seed=1):
    """The reference's "shear_xy" (testcases.cpp:250-296) or "shear_3d" (:298-344) primitive state
    on the block at `start` with `count` cells (padded array of the block), the perturbations from
    a rand() sequence started at `seed` (the reference: a fresh sequence per rank)."""
    count = tuple(count or (case["nx"], case["ny"], case["nz"]))
    X, Y, _ = block_coords(case, start, count)
    Lx, Ly = case["L"][0], case["L"][1]
    x = X - 0.5 * Lx
    y = Y - 0.5 * Ly
    pi = 3.14159265358979323846
    # sin of the C library, once per distinct x
    xs = np.unique(x)
    sx = dict(zip(xs.tolist(), (math.sin(2.0 * pi * v / Lx) for v in xs.tolist())))
    s = np.vectorize(sx.__getitem__, otypes=[np.float64])(x)
    ncell = int(np.prod(count))
    g = CRand(seed)
    W = np.zeros((count[0] + 6, count[1] + 6, count[2] + 6, 5))
    I = (slice(NGC, -NGC),) * 3
    if name == "shear_xy":
        ypert = (0.01 * g.take(ncell)).reshape(count)
        inside = np.abs(y) <= 0.1 * Ly
        W[I + (0,)] = 1.0
        W[I + (1,)] = 2.5
        W[I + (2,)] = np.where(inside, -0.5, 0.5)
        W[I + (3,)] = np.where(inside, ypert, 0.0 * s)
        W[I + (4,)] = 0.0
    elif name == "shear_3d":
        pert = (0.02 * g.take(2 * ncell)).reshape(count + (2,))  # z then y, per cell
        zp, yp = pert[..., 0], pert[..., 1]
        inside = np.abs(y) <= 0.25 * Ly
        W[I + (0,)] = np.where(inside, 2.0, 1.0)
        W[I + (1,)] = 2.5
        W[I + (2,)] = np.where(inside, -0.5, 0.5)
        W[I + (3,)] = np.where(inside, 0.01 * s + yp, 0.01 * s)
        W[I + (4,)] = zp
    else:
        raise ValueError(name)
    return W


def reference_initial_condition(case, name, start=(0, 0, 0), count=None):
    """The state Driver::set_initial_condition (driver.cpp:107-136) sets on one rank's block:
    the Riemann configurations and the two shear layers.  Other names leave the reference's
    state zero (no initial condition): ValueError."""
    count = tuple(count or (case["nx"], case["ny"], case["nz"]))
    if name in ("shear_xy", "shear_3d"):
        return shear_layer_W(case, name, start, count)
    if name in RIEMANN:
        NW, NE, SW, SE, plane = RIEMANN[name]
        X, Y, Z = block_coords(case, start, count)
        Lx, Ly, Lz = case["L"]
        a, La = (X, Lx) if plane == "xy" else (Z, Lz)
        lo_a, lo_y = a <= La / 2.0, Y <= Ly / 2.0
        W = np.zeros((count[0] + 6, count[1] + 6, count[2] + 6, 5))
        I = (slice(NGC, -NGC),) * 3
        for q in range(5):
            W[I + (q,)] = np.where(lo_a & lo_y, NW[q], np.where(~lo_a & lo_y, NE[q],
                                                             np.where(lo_a & ~lo_y, SW[q], SE[q])))
        return W
    raise ValueError(f"test case {name!r} sets no initial condition in the reference")
