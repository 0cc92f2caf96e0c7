/*
 * ppfv_oracle.h -- CPU FP64 restatement of Sayram-2D's implicit PPFV step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker for the HIP path: only the
 * tests, the smoke entry point and the CPU-baseline leg of the benchmark may
 * load it. The product path never links or calls it.
 *
 * It restates, function by function, the reference's Solver
 * (source/Solver.h, source/Solver.cc), geometry (source/Mesh.h, source/Mesh.cc,
 * source/Edge.h) and vertex rule, in plain C with no Eigen/xtensor.
 *
 * Layouts (same as the reference):
 *   cell arrays f, G, Dxx, Dxy, Dyy, inv_tau : shape (nx, ny), element (i,j) at
 *       i*ny + j  (xt::xtensor<double,2>, source/common.h:28)
 *   vertex table V : shape (nx+1, ny+1), vertex (i,j) at i*(ny+1) + j
 *       (vertex_f_, source/Solver.h:51)
 *   matrix bands and R : system order K = j*nx + i (source/Mesh.h:66-68);
 *       band d = M[K][K], w = M[K][K-1], e = M[K][K+1], s = M[K][K-nx],
 *       n = M[K][K+nx] (the operator is 5-point).
 *   boundary sides in the order XMIN, XMAX, YMIN, YMAX (source/BCTypes.h:13);
 *       Dirichlet vertex values for XMIN/XMAX have ny+1 entries (vertex j),
 *       for YMIN/YMAX nx+1 entries (vertex i).
 */
#ifndef PPFV_ORACLE_H
#define PPFV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DIRICHLET = 0, ORC_ZEROFLUX = 1 };   /* source/BCTypes.h:16 */

typedef struct {
  int nx, ny;
  const double *xe;        /* nx+1 cell edges in x (alpha0, rad) */
  const double *ye;        /* ny+1 cell edges in y (ln E) */
  const double *G, *Dxx, *Dxy, *Dyy, *inv_tau;   /* (nx,ny) i-major */
  int bc_type[4];
  const double *bc_val[4];
  double dt;
} orc_problem;

/* fill_vertex_from_cells + fill_vertex_from_bcs, source/Solver.cc:299-422 */
void orc_vertex(const orc_problem *p, const double *f, double *V);

/* Solver::assemble, source/Solver.cc:167-267: bands (5*N, d|w|e|s|n) and R (N) */
void orc_assemble(const orc_problem *p, const double *f, const double *V,
                  double *bands, double *R);

/* Per-face two-point weights of one cell face: a_sigma_func, source/Solver.cc:68-97.
 * inbr: 0 im (west), 1 jp (north), 2 ip (east), 3 jm (south).
 * Outputs (a_A, a_B) and a_sigma = a_A f_A + a_B f_B. */
void orc_a_sigma(const orc_problem *p, const double *V, int i, int j, int inbr,
                 double *aA, double *aB, double *asig);

/* Direct banded LU without pivoting, bandwidth nx, K order. Returns 0 on success. */
int orc_band_solve(int nx, int ny, const double *bands, const double *R, double *x);

/* Matrix-vector product with the 5 bands, K order. */
void orc_matvec(int nx, int ny, const double *bands, const double *x, double *y);

/* One Solver::update (source/Solver.cc:270-290): vertex, assemble, direct solve,
 * copy back to (nx,ny) layout.  f_in and f_out may alias.  Returns 0 on success. */
int orc_step(const orc_problem *p, const double *f_in, double *f_out);

/* nsteps consecutive steps in place. */
int orc_run(const orc_problem *p, double *f, int nsteps);

#ifdef __cplusplus
}
#endif
#endif
