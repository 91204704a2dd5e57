"""Learned WENO-5 weights: training data from 2x-resolution FP64 runs, GPU training, and the
trained network shipped with the package (data/weno_mlp.npz).

Samples.  A 2x-resolution solution (spacing d) holds, along every grid line, point values f_j
of the split fluxes fL = (F + 3U)/2 and fR = (F - 3U)/2 of each conservative variable
(the component-wise Lax-Friedrichs split of the solver).  Every other fine point is a point of a
grid of spacing D = 2d, so the five values f[p-5], f[p-3], f[p-1], f[p+1], f[p+3] are the stencil
of the left-biased reconstruction at the coarse face that lies on fine point p.  Finite-
difference WENO reconstructs at a face the function h whose cell averages are the point values
(Shu 1998, sec. 2.3), h = f - D^2/24 f'' + 7 D^4/5760 f'''' + O(D^6); with fourth- and
second-order fine differences for f'' and f'''' the target is the fixed combination

    t = f_p - (-f[p-2] + 16 f[p-1] - 30 f_p + 16 f[p+1] - f[p+2]) / 72
            + 7/360 (f[p-2] - 4 f[p-1] + 6 f_p - 4 f[p+1] + f[p+2]).

fR uses the right-biased reconstruction, which is the left-biased one of the reversed line.
Features, candidates and the target are normalised by the stencil's range; only stencils in the
network's regime (neither smooth nor flagged) are kept, since only those use the network.
"""
import importlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "data", "weno_mlp.npz")
NPARAM = 32 * 8 + 32 + 32 * 32 + 32 + 3 * 32 + 3
ALPHA = 3.0
EPS, SMOOTH, FLAG = 1e-6, 3.0, 1e3  # regime thresholds of the kernels (ns_mlp.h)
RANGE = 1.0 / 8192.0  # relative stencil range below which a stencil counts as constant


def init_params(seed=0):
    """Glorot-uniform weights, zero hidden biases, output bias log of the ideal weights (the
    untrained network starts at the linear fifth-order scheme)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)

    def glorot(n_out, n_in):
        lim = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-lim, lim, (n_out, n_in))

    p[0:256] = glorot(32, 8).ravel()
    p[288:1312] = glorot(32, 32).ravel()
    p[1344:1440] = 0.1 * glorot(3, 32).ravel()
    p[1440:1443] = np.log([0.1, 0.6, 0.3])
    return p


def features(V):
    """(X, Q, regime, m, r) of left-biased stencils V (n, 5): X = normalised stencil + smoothness
    indicators, Q = the three ENO candidates of the normalised stencil."""
    m = V.min(axis=1)
    M = V.max(axis=1)
    r = M - m
    vary = r > RANGE * np.maximum(np.abs(m), np.abs(M))  # else constant up to round-off
    rs = np.where(vary, r, 1.0)
    x = np.where(vary[:, None], (V - m[:, None]) / rs[:, None], 0.0)
    a, b, c, d, e = x.T
    b0 = 13.0 / 12.0 * (a - 2 * b + c) ** 2 + 0.25 * (a - 4 * b + 3 * c) ** 2
    b1 = 13.0 / 12.0 * (b - 2 * c + d) ** 2 + 0.25 * (b - d) ** 2
    b2 = 13.0 / 12.0 * (c - 2 * d + e) ** 2 + 0.25 * (3 * c - 4 * d + e) ** 2
    B = np.stack([b0, b1, b2], axis=1)
    X = np.concatenate([x, B], axis=1)
    ratio = (B.max(axis=1) + EPS) / (B.min(axis=1) + EPS)
    regime = np.where(ratio < SMOOTH, 0, np.where(ratio > FLAG, 2, 1))
    Q = np.stack([(2 * a - 7 * b + 11 * c) / 6.0, (-b + 5 * c + 2 * d) / 6.0,
                  (2 * c + 5 * d - e) / 6.0], axis=1)
    return X, Q, regime, m, rs


def blend(ratio):
    """Share of the network in the reconstruction (the kernels' mlp_blend)."""
    up = np.minimum(1.0, (ratio - SMOOTH) / SMOOTH)
    down = np.minimum(1.0, (FLAG - ratio) / (0.75 * FLAG))
    return np.maximum(0.0, up) * np.maximum(0.0, down)


def ratio_of(X):
    B = X[:, 5:8]
    return (B.max(axis=1) + EPS) / (B.min(axis=1) + EPS)


def classic_normalised(X, Q, r):
    """The classic Shu reconstruction in the normalised units of X, Q: the Shu weights see the
    raw smoothness indicators r^2 beta (their epsilon is absolute)."""
    a = np.array([0.1, 0.6, 0.3]) / (1e-6 + (r[:, None] ** 2) * X[:, 5:8]) ** 2
    w = a / a.sum(axis=1, keepdims=True)
    return (w * Q).sum(axis=1)


def face_target(f, axis, parts=False):
    """Target face value at every fine point of the periodic field f along `axis` (the coarse
    face on that point, coarse spacing twice the fine one); with parts, also the magnitudes of
    its second- and fourth-derivative corrections."""
    sh = lambda k: np.roll(f, -k, axis=axis)  # noqa: E731  f[p + k]
    fm2, fm1, f0, fp1, fp2 = (sh(k) for k in (-2, -1, 0, 1, 2))
    c2 = (-fm2 + 16 * fm1 - 30 * f0 + 16 * fp1 - fp2) / 72.0
    c4 = 7.0 / 360.0 * (fm2 - 4 * fm1 + 6 * f0 - 4 * fp1 + fp2)
    t = f0 - c2 + c4
    return (t, np.abs(c2), np.abs(c4)) if parts else t


def coarse_stencils(f, axis):
    """(npoints, 5) left-biased coarse stencils f[p-5], f[p-3], f[p-1], f[p+1], f[p+3]."""
    sh = lambda k: np.roll(f, -k, axis=axis)  # noqa: E731
    return np.stack([sh(-5), sh(-3), sh(-1), sh(1), sh(3)], axis=-1).reshape(-1, 5)


def line_samples(f, axis, rng, keep=1.0):
    """Samples of every point of the periodic field f along `axis` (both orientations are
    handled by the caller): the network-regime stencils whose range is at least 1e-4 of the
    field's magnitude (not round-off or seeded noise).  Where the fourth-derivative correction
    exceeds a quarter of the second-derivative one the expansion of h does not hold (an
    under-resolved front): there the target is the classic reconstruction, so the network
    learns to keep the classic behaviour where the fine solution cannot say better.  Returns
    (X, Q, T, S, Cn)."""
    V = coarse_stencils(f, axis)
    t, c2, c4 = (a.ravel() for a in face_target(f, axis, parts=True))
    if keep < 1.0:
        sel = rng.random(len(t)) < keep
        V, t, c2, c4 = V[sel], t[sel], c2[sel], c4[sel]
    X, Q, rg, m, r = features(V)
    T = (t - m) / r
    mag = np.abs(f).max()
    ok = (rg == 1) & (r >= 1e-4 * mag)
    reliable = (c4 <= 0.25 * c2 + 1e-12 * mag) & (np.abs(T) < 3.0)
    X, Q, T, r, reliable = X[ok], Q[ok], T[ok], r[ok], reliable[ok]
    Cn = classic_normalised(X, Q, r)
    T = np.where(reliable, T, Cn)
    return X, Q, T, r / mag, Cn


def split_fluxes(Wi, gamma):
    """Component-wise LF split of the Euler flux along each axis: list over axes of
    (fL, fR), each (5, nx, ny, nz), from the interior primitive state Wi (nx, ny, nz, 5)."""
    rho, p, u, v, w = (Wi[..., q] for q in range(5))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    U = np.stack([rho, E, rho * u, rho * v, rho * w])
    H = (E + p) / rho
    out = []
    for a, qn in enumerate((u, v, w)):
        m = rho * qn
        F = np.stack([m, m * H, m * u + (p if a == 0 else 0.0), m * v + (p if a == 1 else 0.0),
                      m * w + (p if a == 2 else 0.0)])
        out.append((0.5 * (F + ALPHA * U), 0.5 * (F - ALPHA * U)))
    return out


def snapshot_samples(Wi, gamma, rng, keep):
    """(X, Q, T, S, Cn) of one field: features, normalised candidates and target, the scale
    S = stencil range / field magnitude and the normalised classic reconstruction."""
    parts = [[], [], [], [], []]
    for a, (fL, fR) in enumerate(split_fluxes(Wi, gamma)):
        for q in range(5):
            for f in (fL[q], np.flip(fR[q], axis=a)):
                for lst, arr in zip(parts, line_samples(f, a, rng, keep)):
                    lst.append(arr)
    return tuple(np.concatenate(p) for p in parts)


def fine_runs(n_coarse=64):
    """Yields interior primitive states of 2x-resolution FP64 plain-WENO runs (GPU, EXACT math):
    double shear layers in the style of the 64^3 shear configuration (component-wise WENO-5,
    three seeded variants of Mach number and perturbation, smooth perturbation phase) and the
    Taylor-Green vortex
    (characteristic WENO-5), each at two times, with their gamma."""
    ns = importlib.import_module("3dnssolver_amd")
    C = importlib.import_module("3dnssolver_amd.cases")
    n = 2 * n_coarse
    runs = []
    for M, amp, seed in ((0.3, 0.05, 1), (0.5, 0.1, 2), (0.2, 0.02, 3)):
        case, _ = C.config1(n)
        c0 = C.sound_speed(M)
        case = dict(case, mu=M * c0 / 1000.0)
        runs.append((case, C.shear_W(case, M=M, amp=amp, seed=seed, smooth_phase=True),
                     (60, 120)))
    case, W = C.config2(n)
    runs.append((case, W, (200, 400)))
    for case, W, steps in runs:
        s = ns.Solver(case, math="exact")
        s.set_state(W, "primitive")
        done = 0
        for k in steps:
            s.step(k - done)
            done = k
            yield s.get_state("primitive")[3:-3, 3:-3, 3:-3].copy(), case["gamma"]
        s.close()


def training_set(n_samples=20_000_000, n_coarse=64, seed=0, keep=0.5):
    """(X, Qb, Tb) float32 training arrays of n_samples shuffled samples: the reconstruction the
    solver forms, classic + blend (network - classic), weighted by the sample's scale S, is
    (w . Qb - Tb) / 1 with Qb = S blend Q and Tb = S (T - (1 - blend) classic) -- the network's
    loss on (X, Qb, Tb) is the scale-weighted squared error of the blended reconstruction."""
    rng = np.random.default_rng(seed)
    parts = [[], [], [], [], []]
    for Wi, gamma in fine_runs(n_coarse):
        for lst, arr in zip(parts, snapshot_samples(Wi, gamma, rng, keep)):
            lst.append(arr)
    X, Q, T, S, Cn = (np.concatenate(p) for p in parts)
    perm = rng.permutation(len(T))[:n_samples]
    X, Q, T, S, Cn = X[perm], Q[perm], T[perm], S[perm], Cn[perm]
    lam = blend(ratio_of(X))
    Qb = (S * lam)[:, None] * Q
    Tb = S * (T - (1.0 - lam) * Cn)
    return X.astype(np.float32), Qb.astype(np.float32), Tb.astype(np.float32)


def restrict(Wf):
    """Point values at the centres of the grid of twice the spacing: the coarse centre lies
    midway between fine points 2i and 2i+1 of every axis; fourth-order midpoint interpolation
    (-f[2i-1] + 9 f[2i] + 9 f[2i+1] - f[2i+2]) / 16 along each axis in turn, periodic."""
    out = Wf
    for a in range(3):
        r = lambda k: np.roll(out, -k, axis=a)  # noqa: E731
        mid = (-r(-1) + 9.0 * out + 9.0 * r(1) - r(2)) / 16.0
        out = np.take(mid, np.arange(0, out.shape[a], 2), axis=a)
    return out


def l2_rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def reference(case, W, fcase, fW, steps):
    """Prescribed coarse dt (0.8 CFL dx / alpha: the Lax-Friedrichs split dissipates at
    alpha = 3 whatever the wave speeds) and the restricted conservative state of the
    2x-resolution FP64 Shu run to the time of `steps` coarse steps."""
    ns = importlib.import_module("3dnssolver_amd")
    d = min(case["L"][0] / case["nx"], case["L"][1] / case["ny"], case["L"][2] / case["nz"])
    dt = 0.8 * case["cfl"] * d / ALPHA
    f = ns.Solver(fcase, math="exact")
    f.set_state(fW, "primitive")
    t_end = f.step_fixed(2 * steps, 0.5 * dt)
    ref = restrict(f.get_state("conservative")[3:-3, 3:-3, 3:-3])
    f.close()
    return dt, t_end, ref


def coarse_error(case, W, dt, steps, ref, math="exact", weights="shu", params=None):
    """Relative L2 errors of density and total energy of a coarse run against `ref`."""
    ns = importlib.import_module("3dnssolver_amd")
    s = ns.Solver(dict(case, weights=weights), math=math)
    if weights == "learned":
        s.set_mlp(params)
    s.set_state(W, "primitive")
    s.step_fixed(steps, dt)
    U = s.get_state("conservative")[3:-3, 3:-3, 3:-3]
    out = {"rho": l2_rel(U[..., 0], ref[..., 0]), "rhoE": l2_rel(U[..., 1], ref[..., 1]),
           "health": s.reduce()[3]}
    s.close()
    return out


def evaluate(params, n=64, steps=100, problems=("shear", "jump"), maths=("exact", "fp32")):
    """Ablation on the evaluation problems: coarse n^3 runs with the classic Shu weights and
    with the learned weights (each math mode) against a 2n^3 FP64 Shu run at the same time.
    Returns, per problem, the relative L2 errors of density and total energy."""
    C = importlib.import_module("3dnssolver_amd.cases")
    out = {}
    for prob in problems:
        case, W = C.config4(n, prob)
        fcase, fW = C.config4(2 * n, prob)
        dt, t_end, ref = reference(case, W, fcase, fW, steps)
        res = {"t_end": t_end, "dt": dt, "steps": steps, "n": n}
        for math in maths:
            for weights in ("shu", "learned"):
                res[f"{math}_{weights}"] = coarse_error(case, W, dt, steps, ref, math, weights,
                                                        params)
        out[prob] = res
    return out


def train(X, Q, T, steps=3000, batch=65536, lr=3e-3, seed=0, params=None):
    """GPU training (Adam, the project's BF16-MFMA forward and backward kernels) from
    init_params(seed) or `params` on the arrays of training_set."""
    ns = importlib.import_module("3dnssolver_amd")
    p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
    return ns.mlp_train(p, X, Q, T, steps=steps, batch=batch, lr=lr)


def save(params, path=WEIGHTS_PATH, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path=WEIGHTS_PATH):
    """The trained network shipped with the package (float32, NPARAM)."""
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


if __name__ == "__main__":
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description="train the learned WENO-5 weights on the GPU")
    ap.add_argument("--samples", type=int, default=20_000_000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--n-coarse", type=int, default=64)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args()
    t0 = time.time()
    X, Q, T = training_set(a.samples, a.n_coarse)
    t1 = time.time()
    print(f"[learned] {len(T)} samples in {t1 - t0:.1f} s", flush=True)
    p, hist = train(X, Q, T, a.steps, a.batch, a.lr)
    t2 = time.time()
    n_eval = min(len(T), 1_000_000)
    ns = importlib.import_module("3dnssolver_amd")
    l0, _ = ns.mlp_loss_grad(init_params(0), X[:n_eval], Q[:n_eval], T[:n_eval])
    l1, _ = ns.mlp_loss_grad(p, X[:n_eval], Q[:n_eval], T[:n_eval])
    info = {"samples": int(len(T)), "steps": a.steps, "batch": a.batch, "lr": a.lr,
            "train_s": round(t2 - t1, 2), "loss_init_net": float(l0), "loss_trained": float(l1)}
    print(json.dumps(info), flush=True)
    save(p, a.out, info=json.dumps(info), loss_hist=hist)
