"""Writes the golden fixtures of the PPFV step (tests/golden/*.npz).

The reference ships no tests, fixtures or recorded output, and cannot be built
here (Eigen/xtensor/HighFive are absent), so these vectors come from an
independent NumPy/SciPy restatement of its algorithm, written separately from
the C oracle (oracle/ppfv_oracle.c) and from the HIP kernels:

* geometry and the per-face edge table follow source/Mesh.cc:36-155 and
  source/Edge.h:40-53 (vertices, unit normals, lengths as vector norms);
* vertex values follow source/Solver.cc:299-422 (corners, bilinear interior,
  linear boundary edges, Dirichlet overrides in the order XMIN, XMAX, YMIN,
  YMAX);
* the flux pairs, Dirichlet faces and the time/loss diagonal follow
  source/Solver.cc:68-267, emitted as COO triplets in the reference's loop order
  and summed as duplicates by scipy (as Eigen's setFromTriplets does);
* the linear system is solved with scipy.sparse.linalg.spsolve (SuperLU), an
  independent direct solver.

Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "sayram-2d_amd"))
import sayram_cases as sc   # noqa: E402  (input generator only)

EPS = np.finfo(np.float64).eps


class Restatement:
    def __init__(self, p):
        self.p = p
        self.nx, self.ny = p.nx, p.ny
        self.xc = 0.5 * (p.xe[:-1] + p.xe[1:])
        self.yc = 0.5 * (p.ye[:-1] + p.ye[1:])
        self.Lam = np.zeros((p.nx, p.ny, 2, 2))
        self.Lam[..., 0, 0] = p.Dxx * p.G
        self.Lam[..., 0, 1] = p.Dxy * p.G
        self.Lam[..., 1, 0] = p.Dxy * p.G
        self.Lam[..., 1, 1] = p.Dyy * p.G

    def edge(self, i, j, inbr):
        xe, ye = self.p.xe, self.p.ye
        if inbr == 0:
            A, B, va, vb, n = (xe[i], ye[j + 1]), (xe[i], ye[j]), (i, j + 1), (i, j), (-1.0, 0.0)
        elif inbr == 1:
            A, B, va, vb, n = (xe[i + 1], ye[j + 1]), (xe[i], ye[j + 1]), (i + 1, j + 1), (i, j + 1), (0.0, 1.0)
        elif inbr == 2:
            A, B, va, vb, n = (xe[i + 1], ye[j]), (xe[i + 1], ye[j + 1]), (i + 1, j), (i + 1, j + 1), (1.0, 0.0)
        else:
            A, B, va, vb, n = (xe[i], ye[j]), (xe[i + 1], ye[j]), (i, j), (i + 1, j), (0.0, -1.0)
        return np.array(A), np.array(B), va, vb, np.array(n)

    def vertex(self, f):
        nx, ny, p = self.nx, self.ny, self.p
        V = np.zeros((nx + 1, ny + 1))
        V[0, 0], V[nx, 0], V[0, ny], V[nx, ny] = f[0, 0], f[nx - 1, 0], f[0, ny - 1], f[nx - 1, ny - 1]
        for i in range(1, nx):
            for j in range(1, ny):
                xv, yv = p.xe[i], p.ye[j]
                xL, xR, yB, yT = self.xc[i - 1], self.xc[i], self.yc[j - 1], self.yc[j]
                wxL, wxR = (xR - xv) / (xR - xL), (xv - xL) / (xR - xL)
                wyB, wyT = (yT - yv) / (yT - yB), (yv - yB) / (yT - yB)
                V[i, j] = (wxL * wyB * f[i - 1, j - 1] + wxR * wyB * f[i, j - 1]
                           + wxL * wyT * f[i - 1, j] + wxR * wyT * f[i, j])

        def lerp(x, x0, x1, f0, f1):
            return (x1 - x) / (x1 - x0) * f0 + (x - x0) / (x1 - x0) * f1
        for j in range(1, ny):
            V[0, j] = lerp(p.ye[j], self.yc[j - 1], self.yc[j], f[0, j - 1], f[0, j])
            V[nx, j] = lerp(p.ye[j], self.yc[j - 1], self.yc[j], f[nx - 1, j - 1], f[nx - 1, j])
        for i in range(1, nx):
            V[i, 0] = lerp(p.xe[i], self.xc[i - 1], self.xc[i], f[i - 1, 0], f[i, 0])
            V[i, ny] = lerp(p.xe[i], self.xc[i - 1], self.xc[i], f[i - 1, ny - 1], f[i, ny - 1])
        if p.bc_type[0] == sc.DIRICHLET:
            V[0, :] = p.bc_val[0]
        if p.bc_type[1] == sc.DIRICHLET:
            V[nx, :] = p.bc_val[1]
        if p.bc_type[2] == sc.DIRICHLET:
            V[:, 0] = p.bc_val[2]
        if p.bc_type[3] == sc.DIRICHLET:
            V[:, ny] = p.bc_val[3]
        return V

    def a_sigma(self, V, i, j, inbr):
        A, B, va, vb, n = self.edge(i, j, inbr)
        K = np.array([self.xc[i], self.yc[j]])
        vka, vkb = A - K, B - K
        rvka, rvkb = np.array([vka[1], -vka[0]]), np.array([vkb[1], -vkb[0]])
        L = self.Lam[i, j]
        length = np.linalg.norm(B - A)
        aA = length * (n @ L @ rvkb) / (vka @ rvkb)
        aB = length * (n @ L @ rvka) / (vkb @ rvka)
        return aA, aB, aA * V[va] + aB * V[vb]

    def assemble(self, f):
        nx, ny, p = self.nx, self.ny, self.p
        V = self.vertex(f)
        rows, cols, vals = [], [], []
        R = np.zeros(nx * ny)
        kid = lambda i, j: j * nx + i

        def pair(i, j, inbr):
            li, lj = (i - 1, j) if inbr == 0 else (i, j - 1)
            aKA, aKB, sK = self.a_sigma(V, i, j, inbr)
            aLA, aLB, sL = self.a_sigma(V, li, lj, (inbr + 2) % 4)
            muK = (abs(sL) + EPS) / (abs(sK) + abs(sL) + 2 * EPS)
            muL = 1 - muK
            Bs = muL * sL - muK * sK
            Bp, Bm = (abs(Bs) + Bs) / 2.0, (abs(Bs) - Bs) / 2.0
            AK = muK * (aKA + aKB) + Bp / (f[i, j] + EPS)
            AL = muL * (aLA + aLB) + Bm / (f[li, lj] + EPS)
            K, L = kid(i, j), kid(li, lj)
            rows.extend([K, K, L, L]); cols.extend([K, L, L, K]); vals.extend([AK, -AL, AL, -AK])

        def dface(i, j, inbr):
            aA, aB, s = self.a_sigma(V, i, j, inbr)
            Bs = -s
            Bp, Bm = (abs(Bs) + Bs) / 2.0, (abs(Bs) - Bs) / 2.0
            rows.append(kid(i, j)); cols.append(kid(i, j)); vals.append(aA + aB + Bp / (f[i, j] + EPS))
            R[kid(i, j)] += Bm

        for i in range(1, nx):
            for j in range(ny):
                pair(i, j, 0)
        for i in range(nx):
            for j in range(1, ny):
                pair(i, j, 3)
        if p.bc_type[0] == sc.DIRICHLET:
            for j in range(ny):
                dface(0, j, 0)
        if p.bc_type[1] == sc.DIRICHLET:
            for j in range(ny):
                dface(nx - 1, j, 2)
        if p.bc_type[2] == sc.DIRICHLET:
            for i in range(nx):
                dface(i, 0, 3)
        if p.bc_type[3] == sc.DIRICHLET:
            for i in range(nx):
                dface(i, ny - 1, 1)
        dx, dy = np.diff(p.xe), np.diff(p.ye)
        for i in range(nx):
            for j in range(ny):
                UKK = p.G[i, j] * (dx[i] * dy[j] / p.dt)
                rows.append(kid(i, j)); cols.append(kid(i, j)); vals.append(UKK * (1 + p.dt * p.inv_tau[i, j]))
                R[kid(i, j)] += UKK * f[i, j]
        M = sp.coo_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny)).tocsc()
        return V, M, R

    def step(self, f):
        V, M, R = self.assemble(f)
        x = spla.spsolve(M, R)
        return x.reshape(self.ny, self.nx).T.copy()


def bands_of(M, nx, ny):
    """5 bands (d, w, e, s, n) in K order from a sparse matrix."""
    N = nx * ny
    D = M.todense() if N <= 400 else None
    out = np.zeros((5, N))
    Mc = M.tocsr()
    for K in range(N):
        i, j = K % nx, K // nx
        out[0, K] = Mc[K, K]
        if i > 0:
            out[1, K] = Mc[K, K - 1]
        if i < nx - 1:
            out[2, K] = Mc[K, K + 1]
        if j > 0:
            out[3, K] = Mc[K, K - nx]
        if j < ny - 1:
            out[4, K] = Mc[K, K + nx]
    del D
    return out


def variant(p, kind):
    if kind == "all_dirichlet":
        p.bc_type = [sc.DIRICHLET] * 4
        p.bc_val[1] = math.sin(p.xe[-1]) * np.exp(-(np.exp(p.ye) - 0.2) / 0.1)
    elif kind == "all_zeroflux":
        p.bc_type = [sc.ZEROFLUX] * 4
    elif kind == "lc_loss":
        p.bc_type = [sc.ZEROFLUX, sc.ZEROFLUX, sc.DIRICHLET, sc.DIRICHLET]
        it = np.zeros((p.nx, p.ny))
        it[: max(1, p.nx // 4), :] = 3.0
        p.inv_tau = it
    elif kind == "nonuniform":
        # strictly increasing, non-uniform edges (Grid2D allows them)
        t = np.linspace(0.0, 1.0, p.nx + 1)
        p.xe = p.xe[0] + (p.xe[-1] - p.xe[0]) * (t + 0.15 * np.sin(np.pi * t) * t)
        u = np.linspace(0.0, 1.0, p.ny + 1)
        p.ye = p.ye[0] + (p.ye[-1] - p.ye[0]) * u ** 1.3
    return p


CASES = [("c7x5_default", 7, 5, 101, 2.0, "default"),
         ("c6x9_all_dirichlet", 6, 9, 102, 0.5, "all_dirichlet"),
         ("c8x8_all_zeroflux", 8, 8, 103, 4.0, "all_zeroflux"),
         ("c9x7_lc_loss", 9, 7, 104, 1.0, "lc_loss"),
         ("c10x6_nonuniform", 10, 6, 105, 3.0, "nonuniform"),
         ("c1x6_default", 1, 6, 106, 1.0, "default"),
         ("c5x1_default", 5, 1, 107, 1.0, "default"),
         ("c16x12_default", 16, 12, 108, 8.0, "default")]


def main():
    for name, nx, ny, seed, scale, kind in CASES:
        p = variant(sc.seeded_case(nx, ny, seed=seed, scale=scale), kind)
        rs = Restatement(p)
        V0, M0, R0 = rs.assemble(p.f0)
        f = p.f0.copy()
        steps = []
        for _ in range(3):
            f = rs.step(f)
            steps.append(f)
        np.savez(os.path.join(HERE, f"{name}.npz"),
                 nx=nx, ny=ny, xe=p.xe, ye=p.ye, G=p.G, Dxx=p.Dxx, Dxy=p.Dxy, Dyy=p.Dyy,
                 inv_tau=p.inv_tau, bc_type=np.array(p.bc_type),
                 bc_val0=p.bc_val[0], bc_val1=p.bc_val[1], bc_val2=p.bc_val[2],
                 bc_val3=p.bc_val[3], f0=p.f0, dt=p.dt,
                 V0=V0, bands0=bands_of(M0, nx, ny), R0=R0, f_steps=np.array(steps))
        print("wrote", name)


if __name__ == "__main__":
    main()
