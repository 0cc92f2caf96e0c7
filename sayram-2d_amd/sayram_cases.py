"""Seeded synthetic Sayram-2D problems (the ``Seeded_LogNormal`` case).

The reference ships one physical case (``source/Cases/Albert_Young.{h,cc}``),
which needs its HDF5 diffusion table.  The workloads here use a seeded,
resolution-independent stand-in with the same shape of inputs:

* grid: uniform in (alpha0 [rad], ln E [MeV]) built exactly as
  ``make_uniform`` (source/main.cc:20-37) with the degree->radian and ln E
  conversions of source/Parameters.h:35-36 and source/Parameters.cc:50-51;
* D tensor: log10 Daa and log10 DEE are smooth seeded fields (8 cosine modes,
  wavenumbers <= 3 per axis) mapped onto [-6, -2] (1/s), DaE = rho*sqrt(Daa*DEE)
  with |rho| <= 0.9, all times an amplitude scale and converted to 1/day
  (the reference also converts seconds to days, source/Cases/Albert_Young.cc:118);
* Jacobian G = sin(a0) cos(a0) p^2 with p = e2p(E) (source/utils.h:12-14);
* f0 = exp(-(E-0.2)/0.1) sin(a0);
* boundaries: XMIN (loss cone) Dirichlet 0, XMAX (90 deg) zero flux, YMIN and
  YMAX Dirichlet with f0 evaluated on the boundary vertices.

Every cell array is returned in the reference's ``Xtensor2d`` layout: shape
(nx, ny), element (i, j) at ``i*ny + j``.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

E0 = 0.511875          # MeV, source/common.h:42
GPI = 3.141592653589793238462   # source/common.h:39
SECONDS_PER_DAY = 3600 * 24

DIRICHLET, ZEROFLUX = 0, 1      # source/BCTypes.h:16
XMIN, XMAX, YMIN, YMAX = 0, 1, 2, 3   # source/BCTypes.h:13


def e2p(E, E0=E0):
    """Energy -> momentum, source/utils.h:12-14 (gC = 1)."""
    return np.sqrt(E * (E + 2 * E0))


def uniform_edges(n: int, lo: float, hi: float) -> np.ndarray:
    """Cell edges as make_uniform builds them, source/main.cc:25-34."""
    d = (hi - lo) / float(n)
    return np.array([lo + d * float(i) for i in range(n + 1)], dtype=np.float64)


def smooth_field(rng: np.random.Generator, u: np.ndarray, v: np.ndarray,
                 nmodes: int = 8, kmax: int = 3) -> np.ndarray:
    """Sum of seeded separable cosine modes, normalised into [-1, 1]."""
    kx = rng.integers(0, kmax + 1, nmodes)
    ky = rng.integers(0, kmax + 1, nmodes)
    amp = rng.uniform(0.5, 1.0, nmodes)
    px = rng.uniform(0.0, 2 * math.pi, nmodes)
    py = rng.uniform(0.0, 2 * math.pi, nmodes)
    s = np.zeros(np.broadcast(u, v).shape)
    for m in range(nmodes):
        s += amp[m] * np.cos(math.pi * kx[m] * u + px[m]) * np.cos(math.pi * ky[m] * v + py[m])
    return s / amp.sum()


@dataclass
class Problem:
    """One PPFV instance: grid, coefficients, boundary data, initial state."""
    nx: int
    ny: int
    xe: np.ndarray
    ye: np.ndarray
    G: np.ndarray
    Dxx: np.ndarray
    Dxy: np.ndarray
    Dyy: np.ndarray
    inv_tau: np.ndarray
    bc_type: list
    bc_val: list
    f0: np.ndarray
    dt: float
    meta: dict = field(default_factory=dict)

    @property
    def ncells(self) -> int:
        return self.nx * self.ny


def seeded_case(nx: int, ny: int, seed: int = 1, scale: float = 1.0, inv_tau: float = 0.0,
                a0min_deg: float = 5.0, a0max_deg: float = 90.0, Emin: float = 0.2,
                Emax: float = 5.0, dt: float = 1e-3) -> Problem:
    if nx < 1 or ny < 1:
        raise ValueError("nx and ny must be >= 1")
    xe = uniform_edges(nx, a0min_deg * GPI / 180, a0max_deg * GPI / 180)
    ye = uniform_edges(ny, math.log(Emin), math.log(Emax))
    x = 0.5 * (xe[:-1] + xe[1:])
    y = 0.5 * (ye[:-1] + ye[1:])
    u = ((x - xe[0]) / (xe[-1] - xe[0]))[:, None]
    v = ((y - ye[0]) / (ye[-1] - ye[0]))[None, :]
    rng = np.random.default_rng(seed)
    s_aa = smooth_field(rng, u, v)
    s_ee = smooth_field(rng, u, v)
    s_rho = smooth_field(rng, u, v)
    daa = 10.0 ** (-4.0 + 2.0 * s_aa) * scale * SECONDS_PER_DAY
    dee = 10.0 ** (-4.0 + 2.0 * s_ee) * scale * SECONDS_PER_DAY
    dae = 0.9 * np.tanh(2.0 * s_rho) * np.sqrt(daa * dee)
    X = np.broadcast_to(x[:, None], (nx, ny))
    E = np.broadcast_to(np.exp(y)[None, :], (nx, ny))
    p = e2p(E)
    G = np.sin(X) * np.cos(X) * p * p
    f0 = np.exp(-(E - 0.2) / 0.1) * np.sin(X)
    bc_type = [DIRICHLET, ZEROFLUX, DIRICHLET, DIRICHLET]
    bc_val = [np.zeros(ny + 1),
              np.zeros(ny + 1),
              math.exp(-(math.exp(ye[0]) - 0.2) / 0.1) * np.sin(xe),
              math.exp(-(math.exp(ye[-1]) - 0.2) / 0.1) * np.sin(xe)]
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return Problem(nx, ny, c(xe), c(ye), c(G), c(daa), c(dae), c(dee),
                   c(np.full((nx, ny), float(inv_tau))), bc_type, [c(b) for b in bc_val],
                   c(f0), float(dt), dict(seed=seed, scale=scale))


def batch_specs(count: int, lo: int, hi: int, seed: int):
    """Uneven batch: dims uniform in [lo, hi], per-instance D seed and a
    log-uniform amplitude scale in [0.1, 10]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nx = int(rng.integers(lo, hi + 1))
        ny = int(rng.integers(lo, hi + 1))
        s = float(10.0 ** rng.uniform(-1.0, 1.0))
        out.append(dict(nx=nx, ny=ny, seed=int(seed * 100003 + k), scale=s))
    return out


def batch_problems(count: int = 512, lo: int = 32, hi: int = 512, big: int = 4096,
                   seed: int = 2026):
    """The batched workloads: `count` instances of batch_specs, plus one big x
    big grid (D seed derived from `seed`, amplitude 1) when big > 0. With the
    defaults: 512 instances with dims uniform in 32..512 and one 4096^2 grid,
    55.0 M cells."""
    probs = [seeded_case(s["nx"], s["ny"], seed=s["seed"], scale=s["scale"])
             for s in batch_specs(count, lo, hi, seed)]
    if big:
        probs.append(seeded_case(big, big, seed=seed * 7 + 1, scale=1.0))
    return probs


def write_problem(p: Problem, path: str) -> None:
    """Binary problem file read by the C++ driver (host/sayram_gpu_run.cc):
    magic 'SGPROB1\\0', int32 nx, ny, then G, Dxx, Dxy, Dyy, inv_tau, f0 as
    (nx, ny) float64 arrays, int32 bc types[4], and the four Dirichlet vertex
    arrays (ny+1, ny+1, nx+1, nx+1 values; zeros for zero-flux sides). A
    time-dependent problem (p.meta["rates"], see time_dependent) appends magic
    'SGTIME1\\0' and its four rates as float64."""
    with open(path, "wb") as fh:
        fh.write(b"SGPROB1\0")
        fh.write(np.array([p.nx, p.ny], dtype=np.int32).tobytes())
        for a in (p.G, p.Dxx, p.Dxy, p.Dyy, p.inv_tau, p.f0):
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        fh.write(np.array(p.bc_type, dtype=np.int32).tobytes())
        for s, n in zip(range(4), (p.ny + 1, p.ny + 1, p.nx + 1, p.nx + 1)):
            v = p.bc_val[s] if p.bc_val[s] is not None else np.zeros(n)
            fh.write(np.ascontiguousarray(v, dtype=np.float64).tobytes())
        if p.meta.get("rates") is not None:
            fh.write(b"SGTIME1\0")
            fh.write(np.array(p.meta["rates"], dtype=np.float64).tobytes())


def time_dependent(p: Problem, rG: float = 0.0, rD: float = 0.0, rtau: float = 0.0,
                   rbc: float = 0.0) -> Problem:
    """Marks p time-dependent: at time t the Equation holds G (1 + rG t),
    D (1 + rD t), 1/tau (1 + rtau t) and Dirichlet values (1 + rbc t) of the
    base arrays (linear in t, so the C++ driver's Equation and problem_at
    round every value identically)."""
    q = copy.copy(p)
    q.meta = dict(p.meta, rates=(float(rG), float(rD), float(rtau), float(rbc)))
    return q


def problem_at(p: Problem, t: float) -> Problem:
    """The inputs a time-dependent problem's Equation holds after update(t)
    (source/Solver.cc:287-289): the step from t uses them."""
    rates = p.meta.get("rates")
    if rates is None:
        return p
    rG, rD, rtau, rbc = rates
    q = copy.copy(p)
    q.G = p.G * (1.0 + rG * t)
    q.Dxx, q.Dxy, q.Dyy = (a * (1.0 + rD * t) for a in (p.Dxx, p.Dxy, p.Dyy))
    q.inv_tau = p.inv_tau * (1.0 + rtau * t)
    q.bc_val = [None if v is None else v * (1.0 + rbc * t) for v in p.bc_val]
    return q


# ---------------------------------------------------------------------------
# The reference's Albert & Young (2005) chorus cases (source/Cases/
# Albert_Young.{h,cc}, Albert_Young_LC.{h,cc}) from the diffusion table
# D/<dID>.h5 (source/Cases/Albert_Young_IO.cc:17-37). Every value is evaluated
# per cell with the math module (the platform libm) in the reference's
# operation order, like host/albert_young.h, so both give identical inputs.
# ---------------------------------------------------------------------------
RE_EARTH = 6371000.0      # m, source/common.h:44
EPS = 2.0 ** -52          # gEPS, source/common.h:38


def _e2p(E):              # source/utils.h:17-19, gC = 1
    return math.sqrt(E * (E + 2 * E0)) / 1.0


def _p2e(p):              # source/utils.h:12-14
    return math.sqrt(p * p * 1.0 * 1.0 + E0 * E0) - E0


def _dlogE_dp(logE):      # source/utils.h:21-24
    E = math.exp(logE)
    return _e2p(E) * 1.0 * 1.0 / (E * (E + E0))


def _cal_weight(pos, n):
    """calWeight (source/utils.h:41-51) with the reference's size_t index: a
    position below the table becomes an out-of-range unsigned index on x86-64
    and is clamped to the last interval with weight 0."""
    fl = math.floor(pos)
    i = n if fl < 0 else int(fl)
    if i < n:
        return i, 1.0 - (pos - float(i))
    return n - 1, 0.0


def _interp2d(raw, i0, j0, wi, wj):   # interp2D, source/utils.h:25-38
    return (raw[i0][j0] * wi * wj + raw[i0 + 1][j0] * (1 - wi) * wj +
            raw[i0 + 1][j0 + 1] * (1 - wi) * (1 - wj) + raw[i0][j0 + 1] * wi * (1 - wj))


def albert_young_case(nalpha0, nE, alpha0min, alpha0max, Emin, Emax, dt, table, lc=False):
    """Problem of the Albert_Young (lc=False) or Albert_Young_LC (lc=True)
    Equation on make_uniform's grid. table: dict with 'alpha0' (degrees), 'E'
    (MeV) and 'Daa', 'Dap', 'Dpp' of shape (n_alpha0, n_E), as stored in
    D/<dID>.h5."""
    xe = uniform_edges(nalpha0, alpha0min * GPI / 180, alpha0max * GPI / 180)
    logEmin, logEmax = math.log(Emin), math.log(Emax)
    ye = uniform_edges(nE, logEmin, logEmax)
    x = [0.5 * (float(xe[i]) + float(xe[i + 1])) for i in range(nalpha0)]
    y = [0.5 * (float(ye[j]) + float(ye[j + 1])) for j in range(nE)]
    tx = [float(v) * GPI / 180 for v in np.asarray(table["alpha0"], dtype=np.float64)]
    ty = [float(v) for v in np.asarray(table["E"], dtype=np.float64)]
    raw = {k: np.asarray(table[k], dtype=np.float64).tolist() for k in ("Daa", "Dap", "Dpp")}
    xmin, xmax, ymin, ymax = tx[0], tx[-1], ty[0], ty[-1]

    def init_f(a, logE):     # Albert_Young.h:37-40, Albert_Young_LC.h:36-39
        p = _e2p(math.exp(logE))
        if lc:
            return math.exp(-(math.exp(logE) - 0.2) / 0.1) * math.sin(a) / (p * p) + EPS
        return (math.exp(-(math.exp(logE) - 0.2) / 0.1) * (math.sin(a) - math.sin(5 * GPI / 180)) /
                (p * p) + EPS)

    def calc_G(alpha, logE):   # Albert_Young.h:42-45
        t = 1.30 - 0.56 * math.sin(alpha)
        return (math.pow(_e2p(math.exp(logE)), 2) * t * math.sin(alpha) * math.cos(alpha) /
                _dlogE_dp(logE))

    G = np.empty((nalpha0, nE)); f0 = np.empty((nalpha0, nE))
    Dxx = np.empty((nalpha0, nE)); Dxy = np.empty((nalpha0, nE)); Dyy = np.empty((nalpha0, nE))
    itau = np.zeros((nalpha0, nE))
    denorm = (E0 / (1.0 * 1.0)) * (E0 / (1.0 * 1.0)) * 1.0 * 1.0
    s2d = float(3600 * 24)
    for i in range(nalpha0):
        for j in range(nE):
            G[i, j] = calc_G(x[i], y[j])
            f0[i, j] = init_f(x[i], y[j])
    if lc:                     # Albert_Young_LC.cc:42-53, bounce_period :150-160
        L = 4.5
        alpha0lc = math.asin(math.pow(math.pow(L, 5) * (4 * L - 3), -0.25))
        T0, T1 = 1.3802, 0.7405
        for i in range(nalpha0):
            a0 = x[i]
            if a0 < alpha0lc:
                for j in range(nE):
                    p = _e2p(math.exp(y[j]))
                    s = math.sin(a0)
                    Ty = T0 - 0.5 * (T0 - T1) * (s + math.sqrt(s))
                    tb = 4 * L * RE_EARTH * ((E0 + _p2e(p)) / (1.0 * 1.0)) / p * Ty / (3e8 * 3600 * 24)
                    itau[i, j] = 4.0 / tb
    for i in range(nalpha0):   # constructD, Albert_Young.cc:113-134
        a0 = x[i]
        for j in range(nE):
            logE = y[j]
            p = _e2p(math.exp(logE))
            pos_x = (a0 - xmin) / (xmax - xmin) * float(len(tx) - 1)
            pos_y = (logE - math.log(ymin)) / (math.log(ymax) - math.log(ymin)) * float(len(ty) - 1)
            i0, wi = _cal_weight(pos_x, len(tx) - 1)
            j0, wj = _cal_weight(pos_y, len(ty) - 1)
            dl = _dlogE_dp(logE)
            Dxx[i, j] = _interp2d(raw["Daa"], i0, j0, wi, wj) * denorm * s2d / (p * p)
            Dxy[i, j] = _interp2d(raw["Dap"], i0, j0, wi, wj) * denorm * s2d * dl / p
            Dyy[i, j] = _interp2d(raw["Dpp"], i0, j0, wi, wj) * denorm * s2d * math.pow(dl, 2)
    # boundaries (Albert_Young.cc:42-92): XMIN Dirichlet 0 (zero flux for LC),
    # XMAX zero flux, YMIN Dirichlet init_f(x_edge, log Emin), YMAX Dirichlet 0
    bc_type = [ZEROFLUX if lc else DIRICHLET, ZEROFLUX, DIRICHLET, DIRICHLET]
    bc_val = [None if lc else np.zeros(nE + 1), None,
              np.array([init_f(float(v), logEmin) for v in xe]), np.zeros(nalpha0 + 1)]
    return Problem(nalpha0, nE, xe, ye, G, Dxx, Dxy, Dyy, itau, bc_type, bc_val, f0, float(dt),
                   dict(case="albert_young_lc" if lc else "albert_young"))
