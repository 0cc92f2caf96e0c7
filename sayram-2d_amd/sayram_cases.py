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


def write_problem(p: Problem, path: str) -> None:
    """Binary problem file read by the C++ driver (host/sayram_gpu_run.cc):
    magic 'SGPROB1\\0', int32 nx, ny, then G, Dxx, Dxy, Dyy, inv_tau, f0 as
    (nx, ny) float64 arrays, int32 bc types[4], and the four Dirichlet vertex
    arrays (ny+1, ny+1, nx+1, nx+1 values; zeros for zero-flux sides)."""
    with open(path, "wb") as fh:
        fh.write(b"SGPROB1\0")
        fh.write(np.array([p.nx, p.ny], dtype=np.int32).tobytes())
        for a in (p.G, p.Dxx, p.Dxy, p.Dyy, p.inv_tau, p.f0):
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        fh.write(np.array(p.bc_type, dtype=np.int32).tobytes())
        for s, n in zip(range(4), (p.ny + 1, p.ny + 1, p.nx + 1, p.nx + 1)):
            v = p.bc_val[s] if p.bc_val[s] is not None else np.zeros(n)
            fh.write(np.ascontiguousarray(v, dtype=np.float64).tobytes())
