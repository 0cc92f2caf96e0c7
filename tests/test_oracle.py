"""Pins of the CPU oracle (oracle/ppfv_oracle.c), the checker of the HIP path.

The reference ships no tests or golden output and cannot be built here, so the
oracle is pinned by
  * golden vectors written by an independent NumPy/SciPy restatement
    (tests/golden/make_golden.py, SuperLU as the direct solve);
  * closed forms that the mathematics fixes: the per-face two-point weights on
    rectangles, and the special case (constant f, no cross diffusion) in which
    the nonlinear scheme reduces to the textbook harmonic-mean TPFA;
  * invariants: conservation under zero-flux boundaries, preservation of
    constants, exact decay without diffusion, transposition symmetry, linear
    exactness of the vertex rule, the Dirichlet corner precedence, the
    M-matrix structure, and the banded LU against a dense solve.
Each check is chosen so that a plausible slip (a dropped term, a wrong sign,
a swapped index or operand) fails it.
"""
import glob
import math
import os

import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    nx, ny = int(z["nx"]), int(z["ny"])
    p = sc.Problem(nx, ny, z["xe"], z["ye"], z["G"], z["Dxx"], z["Dxy"], z["Dyy"], z["inv_tau"],
                   [int(t) for t in z["bc_type"]],
                   [z["bc_val0"], z["bc_val1"], z["bc_val2"], z["bc_val3"]], z["f0"], float(z["dt"]))
    return p, z


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 8


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(g)[:-4] for g in GOLDEN])
def test_oracle_matches_golden(path):
    p, z = load_golden(path)
    op = orc.OracleProblem(p)
    V = orc.vertex(op, p.f0)
    np.testing.assert_allclose(V, z["V0"], rtol=1e-14, atol=1e-300)
    bands, R = orc.assemble(op, p.f0, V)
    # Entries agree to ~1e-14 of their row, except where B = muL aL - muK aK is a
    # small difference of large terms (f near 1e-21 at high energy): there the
    # two restatements' last-bit differences in a_A, a_B grow to ~1e-12.
    scale = np.max(np.abs(z["bands0"]), axis=0)
    assert np.all(np.abs(bands - z["bands0"]) <= 1e-11 * scale[None, :])
    np.testing.assert_allclose(R, z["R0"], rtol=1e-13, atol=1e-13 * np.max(np.abs(z["R0"])))
    f = p.f0
    for k in range(z["f_steps"].shape[0]):
        f = orc.step(op, f)
        assert rel_l2(f, z["f_steps"][k]) <= 1e-12, f"step {k + 1}"


def _nonuniform(p):
    t = np.linspace(0.0, 1.0, p.nx + 1)
    p.xe = p.xe[0] + (p.xe[-1] - p.xe[0]) * (t + 0.2 * t * (1 - t))
    u = np.linspace(0.0, 1.0, p.ny + 1)
    p.ye = p.ye[0] + (p.ye[-1] - p.ye[0]) * u ** 1.2
    return p


def test_a_sigma_closed_form_on_rectangles():
    """a_sigma_func (source/Solver.cc:68-97) on an axis-aligned cell with
    rx = dy/dx, ry = dx/dy: W and E faces give (Lxx rx - Lxy, Lxx rx + Lxy),
    N and S faces (Lyy ry + Lxy, Lyy ry - Lxy), and a_A(A-K) + a_B(B-K) equals
    |sigma| Lambda n."""
    p = _nonuniform(sc.seeded_case(7, 6, seed=21, scale=3.0))
    op = orc.OracleProblem(p)
    V = orc.vertex(op, p.f0)
    for i in range(p.nx):
        for j in range(p.ny):
            g = p.G[i, j]
            Lxx, Lxy, Lyy = p.Dxx[i, j] * g, p.Dxy[i, j] * g, p.Dyy[i, j] * g
            dx, dy = p.xe[i + 1] - p.xe[i], p.ye[j + 1] - p.ye[j]
            rx, ry = dy / dx, dx / dy
            K = np.array([0.5 * (p.xe[i] + p.xe[i + 1]), 0.5 * (p.ye[j] + p.ye[j + 1])])
            edges = {0: ((p.xe[i], p.ye[j + 1]), (p.xe[i], p.ye[j]), (-1, 0), dy),
                     1: ((p.xe[i + 1], p.ye[j + 1]), (p.xe[i], p.ye[j + 1]), (0, 1), dx),
                     2: ((p.xe[i + 1], p.ye[j]), (p.xe[i + 1], p.ye[j + 1]), (1, 0), dy),
                     3: ((p.xe[i], p.ye[j]), (p.xe[i + 1], p.ye[j]), (0, -1), dx)}
            Lam = np.array([[Lxx, Lxy], [Lxy, Lyy]])
            for inbr in range(4):
                aA, aB, _ = orc.a_sigma(op, V, i, j, inbr)
                if inbr in (0, 2):
                    cA, cB = Lxx * rx - Lxy, Lxx * rx + Lxy
                else:
                    cA, cB = Lyy * ry + Lxy, Lyy * ry - Lxy
                s = abs(Lxx * rx) + abs(Lyy * ry) + abs(Lxy)
                assert abs(aA - cA) <= 1e-14 * s and abs(aB - cB) <= 1e-14 * s
                A, B, n, length = edges[inbr]
                lhs = aA * (np.array(A) - K) + aB * (np.array(B) - K)
                rhs = length * Lam @ np.array(n, dtype=float)
                assert np.allclose(lhs, rhs, rtol=1e-13, atol=1e-13 * np.abs(rhs).max())


def test_reduces_to_harmonic_mean_tpfa():
    """Constant f > 0, no cross diffusion, zero-flux sides: every interior
    coupling is the harmonic-mean two-point flux 2 / (hK/LK + hL/LL) |sigma|
    (h: centre-to-face distances), the scheme of the textbook 5-point method."""
    p = _nonuniform(sc.seeded_case(9, 8, seed=5, scale=2.0))
    p.Dxy = np.zeros_like(p.Dxy)
    p.bc_type = [sc.ZEROFLUX] * 4
    c = 0.37
    f = np.full((p.nx, p.ny), c)
    bands, R = orc.assemble(orc.OracleProblem(p), f)
    nx = p.nx
    for i in range(p.nx):
        for j in range(p.ny):
            K = j * nx + i
            LK = np.array([p.Dxx[i, j], p.Dyy[i, j]]) * p.G[i, j]
            if i > 0:
                LL = p.Dxx[i - 1, j] * p.G[i - 1, j]
                hK, hL = (p.xe[i + 1] - p.xe[i]) / 2, (p.xe[i] - p.xe[i - 1]) / 2
                t = 2.0 / (hK / LK[0] + hL / LL) * (p.ye[j + 1] - p.ye[j]) / 2
                assert abs(-bands[1, K] - t) <= 1e-12 * t
            if j > 0:
                LL = p.Dyy[i, j - 1] * p.G[i, j - 1]
                hK, hL = (p.ye[j + 1] - p.ye[j]) / 2, (p.ye[j] - p.ye[j - 1]) / 2
                t = 2.0 / (hK / LK[1] + hL / LL) * (p.xe[i + 1] - p.xe[i]) / 2
                assert abs(-bands[3, K] - t) <= 1e-12 * t


def test_dirichlet_face_half_cell_flux():
    """Constant f = c with Dirichlet vertex value c on XMIN and no cross
    diffusion: the face adds 2 Lxx dy/dx to the diagonal and the same times c
    to R, i.e. the flux 2 Lxx (f_K - c) dy/dx over the half-cell distance."""
    p = sc.seeded_case(6, 5, seed=8)
    p.Dxy = np.zeros_like(p.Dxy)
    c = 0.8
    p.bc_type = [sc.DIRICHLET, sc.ZEROFLUX, sc.ZEROFLUX, sc.ZEROFLUX]
    p.bc_val[0] = np.full(p.ny + 1, c)
    f = np.full((p.nx, p.ny), c)
    op = orc.OracleProblem(p)
    bands, R = orc.assemble(op, f)
    p2 = sc.seeded_case(6, 5, seed=8)
    p2.Dxy = np.zeros_like(p2.Dxy)
    p2.bc_type = [sc.ZEROFLUX] * 4
    b2, R2 = orc.assemble(orc.OracleProblem(p2), f)
    for j in range(p.ny):
        K = j * p.nx
        a = 2 * p.Dxx[0, j] * p.G[0, j] * (p.ye[j + 1] - p.ye[j]) / (p.xe[1] - p.xe[0])
        assert abs((bands[0, K] - b2[0, K]) - a) <= 1e-10 * a
        assert abs((R[K] - R2[K]) - a * c) <= 1e-10 * a * c


def test_conservation_with_zero_flux_sides():
    """No loss, zero-flux sides: the flux columns of M sum to zero
    (source/Solver.cc:136-139), so sum_K G_K dx dy f_K is conserved."""
    p = sc.seeded_case(12, 10, seed=9, scale=5.0)
    p.bc_type = [sc.ZEROFLUX] * 4
    op = orc.OracleProblem(p)
    w = p.G * np.outer(np.diff(p.xe), np.diff(p.ye))
    m0 = np.sum(w * p.f0)
    f = orc.run(op, p.f0, 6)
    assert abs(np.sum(w * f) - m0) <= 1e-12 * m0


def test_constant_state_is_preserved():
    p = sc.seeded_case(8, 7, seed=10, scale=5.0)
    p.bc_type = [sc.ZEROFLUX] * 4
    f = orc.run(orc.OracleProblem(p), np.full((p.nx, p.ny), 2.5), 4)
    assert np.max(np.abs(f - 2.5)) <= 1e-12 * 2.5


def test_no_diffusion_gives_exact_decay():
    """D = 0: f^{n+1} = f^n / (1 + dt/tau) (source/Solver.cc:195-197)."""
    p = sc.seeded_case(6, 5, seed=11)
    for a in ("Dxx", "Dxy", "Dyy"):
        setattr(p, a, np.zeros_like(getattr(p, a)))
    p.inv_tau = np.linspace(0.0, 4.0, p.nx * p.ny).reshape(p.nx, p.ny)
    f = orc.step(orc.OracleProblem(p), p.f0)
    np.testing.assert_allclose(f, p.f0 / (1 + p.dt * p.inv_tau), rtol=1e-14)


def test_transposition_symmetry():
    """Transposing every input (grid axes, G, f0, D with Dxx <-> Dyy, sides
    XMIN <-> YMIN and XMAX <-> YMAX) transposes the solution."""
    p = sc.seeded_case(7, 9, seed=12, scale=3.0)
    p.bc_type = [sc.ZEROFLUX] * 4
    p.inv_tau = np.full((p.nx, p.ny), 0.5)
    q = sc.Problem(p.ny, p.nx, p.ye.copy(), p.xe.copy(), p.G.T.copy(), p.Dyy.T.copy(),
                   p.Dxy.T.copy(), p.Dxx.T.copy(), p.inv_tau.T.copy(), [sc.ZEROFLUX] * 4,
                   [np.zeros(p.nx + 1), np.zeros(p.nx + 1), np.zeros(p.ny + 1), np.zeros(p.ny + 1)],
                   p.f0.T.copy(), p.dt)
    fp = orc.run(orc.OracleProblem(p), p.f0, 3)
    fq = orc.run(orc.OracleProblem(q), q.f0, 3)
    assert rel_l2(fq.T, fp) <= 1e-12


def test_vertex_rule_linear_exact_and_corners():
    """The vertex rule (source/Solver.cc:326-381): interior vertices are
    bilinear in the four cell centres, so any linear field is exact there;
    boundary-edge vertices interpolate along the edge between the two adjacent
    cells, so a field linear along that edge is exact; corners copy the
    adjacent cell (:316-319)."""
    p = _nonuniform(sc.seeded_case(6, 5, seed=13))
    p.bc_type = [sc.ZEROFLUX] * 4
    op = orc.OracleProblem(p)
    xc = 0.5 * (p.xe[:-1] + p.xe[1:])
    yc = 0.5 * (p.ye[:-1] + p.ye[1:])
    f = 1.5 + 2.0 * xc[:, None] - 0.7 * yc[None, :]
    V = orc.vertex(op, f)
    ex = 1.5 + 2.0 * p.xe[:, None] - 0.7 * p.ye[None, :]
    np.testing.assert_allclose(V[1:-1, 1:-1], ex[1:-1, 1:-1], rtol=1e-13)
    assert V[0, 0] == f[0, 0] and V[-1, -1] == f[-1, -1]
    assert V[-1, 0] == f[-1, 0] and V[0, -1] == f[0, -1]
    fy = np.broadcast_to(1.5 - 0.7 * yc[None, :], (p.nx, p.ny))      # linear along x = const edges
    Vy = orc.vertex(op, fy)
    ey = 1.5 - 0.7 * p.ye
    np.testing.assert_allclose(Vy[0, 1:-1], ey[1:-1], rtol=1e-13)
    np.testing.assert_allclose(Vy[-1, 1:-1], ey[1:-1], rtol=1e-13)
    fx = np.broadcast_to(1.5 + 2.0 * xc[:, None], (p.nx, p.ny))      # linear along y = const edges
    Vx = orc.vertex(op, fx)
    exx = 1.5 + 2.0 * p.xe
    np.testing.assert_allclose(Vx[1:-1, 0], exx[1:-1], rtol=1e-13)
    np.testing.assert_allclose(Vx[1:-1, -1], exx[1:-1], rtol=1e-13)
    # across a boundary the rule does not extrapolate: it takes the value half
    # a cell inside
    assert abs(V[0, 2] - (1.5 + 2.0 * xc[0] - 0.7 * p.ye[2])) <= 1e-13 * abs(V[0, 2])


def test_dirichlet_corner_precedence():
    """Sides are written XMIN, XMAX, YMIN, YMAX (source/Solver.cc:398-420):
    the y sides win at the corners."""
    p = sc.seeded_case(5, 4, seed=14)
    p.bc_type = [sc.DIRICHLET] * 4
    p.bc_val = [np.full(p.ny + 1, 1.0), np.full(p.ny + 1, 2.0), np.full(p.nx + 1, 3.0),
                np.full(p.nx + 1, 4.0)]
    V = orc.vertex(orc.OracleProblem(p), p.f0)
    assert V[0, 0] == 3.0 and V[-1, 0] == 3.0 and V[0, -1] == 4.0 and V[-1, -1] == 4.0
    assert np.all(V[0, 1:-1] == 1.0) and np.all(V[-1, 1:-1] == 2.0)


def test_m_matrix_structure_and_positivity():
    """Off-diagonals <= 0 and column sums >= the time term > 0: M is a
    column diagonally dominant M-matrix, so f >= 0 is preserved."""
    p = sc.seeded_case(14, 11, seed=15, scale=8.0)
    op = orc.OracleProblem(p)
    f = p.f0
    for _ in range(3):
        bands, R = orc.assemble(op, f)
        nx, ny, N = p.nx, p.ny, p.nx * p.ny
        assert np.all(bands[1:] <= 0.0)
        cs = bands[0].copy()
        for K in range(N):   # column K collects M[L][K] from its neighbours
            i, j = K % nx, K // nx
            if i > 0:
                cs[K] += bands[2, K - 1]
            if i < nx - 1:
                cs[K] += bands[1, K + 1]
            if j > 0:
                cs[K] += bands[4, K - nx]
            if j < ny - 1:
                cs[K] += bands[3, K + nx]
        ukk = (p.G * np.outer(np.diff(p.xe), np.diff(p.ye)) / p.dt).T.ravel()
        assert np.all(cs >= ukk * (1 - 1e-12))
        f = orc.step(op, f)
        assert f.min() >= 0.0


def test_band_solver_matches_dense_solve():
    p = sc.seeded_case(9, 7, seed=16, scale=4.0)
    bands, R = orc.assemble(orc.OracleProblem(p), p.f0)
    nx, ny, N = p.nx, p.ny, p.nx * p.ny
    M = np.zeros((N, N))
    for K in range(N):
        i, j = K % nx, K // nx
        M[K, K] = bands[0, K]
        if i > 0:
            M[K, K - 1] = bands[1, K]
        if i < nx - 1:
            M[K, K + 1] = bands[2, K]
        if j > 0:
            M[K, K - nx] = bands[3, K]
        if j < ny - 1:
            M[K, K + nx] = bands[4, K]
    x = orc.band_solve(nx, ny, bands, R)
    assert rel_l2(x, np.linalg.solve(M, R)) <= 1e-13
    assert rel_l2(orc.matvec(nx, ny, bands, x), R) <= 1e-14


def test_band_solver_matches_superlu_at_moderate_size():
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    p = sc.seeded_case(40, 33, seed=17, scale=6.0)
    bands, R = orc.assemble(orc.OracleProblem(p), p.f0)
    nx, ny = p.nx, p.ny
    M = sps.diags([bands[3][nx:], bands[1][1:], bands[0], bands[2][:-1], bands[4][:-nx]],
                  [-nx, -1, 0, 1, nx], format="csc")
    assert rel_l2(orc.band_solve(nx, ny, bands, R), spla.spsolve(M, R)) <= 1e-12


def test_edge_length_identity():
    """The oracle takes |B - A| as sqrt(dx^2 + dy^2), the HIP kernel as
    |dx| + |dy|; with one of the two exactly zero both are |d| bit for bit."""
    rng = np.random.default_rng(0)
    for d in rng.uniform(1e-6, 10.0, 2000):
        assert math.sqrt(d * d + 0.0) == abs(d)
