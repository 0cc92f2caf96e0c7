/*
 * ppfv_oracle.c -- CPU FP64 restatement of Sayram-2D's implicit PPFV step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped or
 * measured as the product). Built with -O2 -ffp-contract=off so that every
 * floating-point operation is rounded exactly as written, in the order the
 * reference evaluates it.
 *
 * Each function cites the reference lines it follows. The reference itself
 * (Eigen + xtensor + HighFive) cannot be compiled in this image; this file is
 * a restatement written from reading it, not a copy.
 */
#include "ppfv_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* gEPS = std::numeric_limits<double>::epsilon(), source/common.h:38 */
static const double EPS = DBL_EPSILON;

#define CELL(p, a, i, j) ((a)[(size_t)(i) * (size_t)(p)->ny + (size_t)(j)])
#define VTX(p, V, i, j) ((V)[(size_t)(i) * (size_t)((p)->ny + 1) + (size_t)(j)])

/* Mesh::build_geometry_from_edges_, source/Mesh.cc:36-65 */
static double xc(const orc_problem *p, int i) { return 0.5 * (p->xe[i] + p->xe[i + 1]); }
static double yc(const orc_problem *p, int j) { return 0.5 * (p->ye[j] + p->ye[j + 1]); }
static double dxc(const orc_problem *p, int i) { return p->xe[i + 1] - p->xe[i]; }
static double dyc(const orc_problem *p, int j) { return p->ye[j + 1] - p->ye[j]; }

/* lerp lambda of fill_vertex_from_cells, source/Solver.cc:304-309 */
static double lerp(double x, double x0, double x1, double f0, double f1) {
  const double denom = (x1 - x0);
  if (denom == 0.0) return 0.5 * (f0 + f1);
  const double w0 = (x1 - x) / denom;
  const double w1 = (x - x0) / denom;
  return w0 * f0 + w1 * f1;
}

/* fill_vertex_from_cells (source/Solver.cc:299-383), then fill_vertex_from_bcs
 * (source/Solver.cc:385-422): Dirichlet sides overwrite in the order XMIN, XMAX,
 * YMIN, YMAX, so the last side written wins at the corners. */
void orc_vertex(const orc_problem *p, const double *f, double *V) {
  const int nx = p->nx, ny = p->ny;
  /* corners: the single adjacent cell (:316-319) */
  VTX(p, V, 0, 0) = CELL(p, f, 0, 0);
  VTX(p, V, nx, 0) = CELL(p, f, nx - 1, 0);
  VTX(p, V, 0, ny) = CELL(p, f, 0, ny - 1);
  VTX(p, V, nx, ny) = CELL(p, f, nx - 1, ny - 1);
  /* interior vertices, bilinear in the four surrounding cells (:326-349) */
  for (int i = 1; i < nx; ++i) {
    for (int j = 1; j < ny; ++j) {
      const double xv = p->xe[i], yv = p->ye[j];
      const double xL = xc(p, i - 1), xR = xc(p, i);
      const double yB = yc(p, j - 1), yT = yc(p, j);
      const double denom_x = (xR - xL), denom_y = (yT - yB);
      const double wxL = (xR - xv) / denom_x;
      const double wxR = (xv - xL) / denom_x;
      const double wyB = (yT - yv) / denom_y;
      const double wyT = (yv - yB) / denom_y;
      VTX(p, V, i, j) = wxL * wyB * CELL(p, f, i - 1, j - 1) + wxR * wyB * CELL(p, f, i, j - 1) +
                        wxL * wyT * CELL(p, f, i - 1, j) + wxR * wyT * CELL(p, f, i, j);
    }
  }
  /* vertical boundary edges i = 0 and i = nx, corners excluded (:355-366) */
  for (int j = 1; j < ny; ++j) {
    const double yv = p->ye[j], y0 = yc(p, j - 1), y1 = yc(p, j);
    VTX(p, V, 0, j) = lerp(yv, y0, y1, CELL(p, f, 0, j - 1), CELL(p, f, 0, j));
    VTX(p, V, nx, j) = lerp(yv, y0, y1, CELL(p, f, nx - 1, j - 1), CELL(p, f, nx - 1, j));
  }
  /* horizontal boundary edges j = 0 and j = ny, corners excluded (:371-381) */
  for (int i = 1; i < nx; ++i) {
    const double xv = p->xe[i], x0 = xc(p, i - 1), x1 = xc(p, i);
    VTX(p, V, i, 0) = lerp(xv, x0, x1, CELL(p, f, i - 1, 0), CELL(p, f, i, 0));
    VTX(p, V, i, ny) = lerp(xv, x0, x1, CELL(p, f, i - 1, ny - 1), CELL(p, f, i, ny - 1));
  }
  /* Dirichlet overrides (:385-422) */
  if (p->bc_type[0] == ORC_DIRICHLET)
    for (int j = 0; j <= ny; ++j) VTX(p, V, 0, j) = p->bc_val[0][j];
  if (p->bc_type[1] == ORC_DIRICHLET)
    for (int j = 0; j <= ny; ++j) VTX(p, V, nx, j) = p->bc_val[1][j];
  if (p->bc_type[2] == ORC_DIRICHLET)
    for (int i = 0; i <= nx; ++i) VTX(p, V, i, 0) = p->bc_val[2][i];
  if (p->bc_type[3] == ORC_DIRICHLET)
    for (int i = 0; i <= nx; ++i) VTX(p, V, i, ny) = p->bc_val[3][i];
}

/* Edge table of Mesh::build_connectivity_ (source/Mesh.cc:67-155) and the
 * normals of Edge::set (source/Edge.h:40-53). */
static void edge_of(const orc_problem *p, int i, int j, int inbr, double A[2], double B[2],
                    int va[2], int vb[2], double n[2]) {
  switch (inbr) {
    case 0: /* im, XNEG */
      A[0] = p->xe[i]; A[1] = p->ye[j + 1]; B[0] = p->xe[i]; B[1] = p->ye[j];
      va[0] = i; va[1] = j + 1; vb[0] = i; vb[1] = j; n[0] = -1.0; n[1] = 0.0; break;
    case 1: /* jp, YPOS */
      A[0] = p->xe[i + 1]; A[1] = p->ye[j + 1]; B[0] = p->xe[i]; B[1] = p->ye[j + 1];
      va[0] = i + 1; va[1] = j + 1; vb[0] = i; vb[1] = j + 1; n[0] = 0.0; n[1] = 1.0; break;
    case 2: /* ip, XPOS */
      A[0] = p->xe[i + 1]; A[1] = p->ye[j]; B[0] = p->xe[i + 1]; B[1] = p->ye[j + 1];
      va[0] = i + 1; va[1] = j; vb[0] = i + 1; vb[1] = j + 1; n[0] = 1.0; n[1] = 0.0; break;
    default: /* jm, YNEG */
      A[0] = p->xe[i]; A[1] = p->ye[j]; B[0] = p->xe[i + 1]; B[1] = p->ye[j];
      va[0] = i; va[1] = j; vb[0] = i + 1; vb[1] = j; n[0] = 0.0; n[1] = -1.0; break;
  }
}

/* Solver::a_sigma_func, source/Solver.cc:68-97, with Lambda = G*D from
 * Solver::update_Lambda, source/Solver.cc:57-65. */
void orc_a_sigma(const orc_problem *p, const double *V, int i, int j, int inbr, double *aA,
                 double *aB, double *asig) {
  double A[2], B[2], n[2];
  int va[2], vb[2];
  edge_of(p, i, j, inbr, A, B, va, vb, n);
  const double g = CELL(p, p->G, i, j);
  const double L00 = CELL(p, p->Dxx, i, j) * g, L01 = CELL(p, p->Dxy, i, j) * g;
  const double L10 = CELL(p, p->Dxy, i, j) * g, L11 = CELL(p, p->Dyy, i, j) * g;
  const double K[2] = {xc(p, i), yc(p, j)};
  const double ex = B[0] - A[0], ey = B[1] - A[1];
  const double len = sqrt(ex * ex + ey * ey);
  const double vkb[2] = {B[0] - K[0], B[1] - K[1]};
  const double vka[2] = {A[0] - K[0], A[1] - K[1]};
  const double rvkb[2] = {vkb[1], -vkb[0]};
  const double rvka[2] = {vka[1], -vka[0]};
  /* n^T Lambda (row vector) */
  const double t0 = n[0] * L00 + n[1] * L10;
  const double t1 = n[0] * L01 + n[1] * L11;
  const double numa = t0 * rvkb[0] + t1 * rvkb[1];
  const double dena = vka[0] * rvkb[0] + vka[1] * rvkb[1];
  const double numb = t0 * rvka[0] + t1 * rvka[1];
  const double denb = vkb[0] * rvka[0] + vkb[1] * rvka[1];
  *aA = len * numa / dena;
  *aB = len * numb / denb;
  *asig = (*aA) * VTX(p, V, va[0], va[1]) + (*aB) * VTX(p, V, vb[0], vb[1]);
}

/* Solver::calculate_mu, source/Solver.h:63-67 */
static void calc_mu(double aK, double aL, double *muK, double *muL) {
  const double denom = fabs(aK) + fabs(aL) + 2 * EPS;
  *muK = (fabs(aL) + EPS) / denom;
  *muL = 1 - *muK;
}

#define BAND(b, N, k) ((b) + (size_t)(k) * (size_t)(N))

/* Solver::apply_inner_face_pair, source/Solver.cc:99-141. Adds the four
 * entries (K,K,+A_K) (K,L,-A_L) (L,L,+A_L) (L,K,-A_K) into the bands. */
static void inner_pair(const orc_problem *p, const double *f, const double *V, int i, int j,
                       int inbr, double *bands) {
  const int nx = p->nx, N = p->nx * p->ny;
  double aKA, aKB, asK, aLA, aLB, asL, muK, muL;
  orc_a_sigma(p, V, i, j, inbr, &aKA, &aKB, &asK);
  const int li = (inbr == 0) ? i - 1 : i, lj = (inbr == 3) ? j - 1 : j;
  const int rinb = (inbr + 2) % 4; /* rinbr_ = {2,3,0,1}, source/Mesh.h:115 */
  orc_a_sigma(p, V, li, lj, rinb, &aLA, &aLB, &asL);
  calc_mu(asK, asL, &muK, &muL);
  const double Bs = muL * asL - muK * asK;
  const double Ba = fabs(Bs);
  const double Bp = (Ba + Bs) / 2.0, Bm = (Ba - Bs) / 2.0;
  const double AK = muK * (aKA + aKB) + Bp / (CELL(p, f, i, j) + EPS);
  const double AL = muL * (aLA + aLB) + Bm / (CELL(p, f, li, lj) + EPS);
  const int K = j * nx + i, L = lj * nx + li;
  BAND(bands, N, 0)[K] += AK;
  if (inbr == 0) {
    BAND(bands, N, 1)[K] += -AL; /* M[K][K-1] */
    BAND(bands, N, 2)[L] += -AK; /* M[L][L+1] */
  } else {
    BAND(bands, N, 3)[K] += -AL; /* M[K][K-nx] */
    BAND(bands, N, 4)[L] += -AK; /* M[L][L+nx] */
  }
  BAND(bands, N, 0)[L] += AL;
  (void)K;
}

/* Solver::apply_dirichlet_face, source/Solver.cc:143-164 */
static void dirichlet_face(const orc_problem *p, const double *f, const double *V, int i, int j,
                           int inbr, double *bands, double *R) {
  const int N = p->nx * p->ny;
  double aA, aB, as;
  orc_a_sigma(p, V, i, j, inbr, &aA, &aB, &as);
  const double Bs = -as;
  const double Ba = fabs(Bs);
  const double Bp = (Ba + Bs) / 2.0, Bm = (Ba - Bs) / 2.0;
  const double AK = aA + aB + Bp / (CELL(p, f, i, j) + EPS);
  const int K = j * p->nx + i;
  BAND(bands, N, 0)[K] += AK;
  R[K] += Bm;
}

/* Solver::assemble + apply_boundary_faces, source/Solver.cc:167-267.  Entries
 * are summed in the order the reference pushes its triplets, which is the
 * order Eigen's setFromTriplets sums duplicates in. */
void orc_assemble(const orc_problem *p, const double *f, const double *V, double *bands,
                  double *R) {
  const int nx = p->nx, ny = p->ny, N = nx * ny;
  memset(bands, 0, sizeof(double) * 5 * (size_t)N);
  memset(R, 0, sizeof(double) * (size_t)N);
  for (int i = 1; i < nx; ++i)
    for (int j = 0; j < ny; ++j) inner_pair(p, f, V, i, j, 0, bands);
  for (int i = 0; i < nx; ++i)
    for (int j = 1; j < ny; ++j) inner_pair(p, f, V, i, j, 3, bands);
  if (p->bc_type[0] == ORC_DIRICHLET)
    for (int j = 0; j < ny; ++j) dirichlet_face(p, f, V, 0, j, 0, bands, R);
  if (p->bc_type[1] == ORC_DIRICHLET)
    for (int j = 0; j < ny; ++j) dirichlet_face(p, f, V, nx - 1, j, 2, bands, R);
  if (p->bc_type[2] == ORC_DIRICHLET)
    for (int i = 0; i < nx; ++i) dirichlet_face(p, f, V, i, 0, 3, bands, R);
  if (p->bc_type[3] == ORC_DIRICHLET)
    for (int i = 0; i < nx; ++i) dirichlet_face(p, f, V, i, ny - 1, 1, bands, R);
  /* time / loss diagonal, source/Solver.cc:190-199 and Mesh::cell_area_dt,
   * source/Mesh.h:61-63 */
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      const int K = j * nx + i;
      const double UKK = CELL(p, p->G, i, j) * (dxc(p, i) * dyc(p, j) / p->dt);
      BAND(bands, N, 0)[K] += UKK * (1 + p->dt * CELL(p, p->inv_tau, i, j));
      R[K] += UKK * CELL(p, f, i, j);
    }
  }
}

void orc_matvec(int nx, int ny, const double *bands, const double *x, double *y) {
  const int N = nx * ny;
  for (int K = 0; K < N; ++K) {
    const int i = K % nx, j = K / nx;
    double s = BAND(bands, N, 0)[K] * x[K];
    if (i > 0) s += BAND(bands, N, 1)[K] * x[K - 1];
    if (i < nx - 1) s += BAND(bands, N, 2)[K] * x[K + 1];
    if (j > 0) s += BAND(bands, N, 3)[K] * x[K - nx];
    if (j < ny - 1) s += BAND(bands, N, 4)[K] * x[K + nx];
    y[K] = s;
  }
}

/* Stand-in for Eigen SparseLU (source/Solver.cc:276-278): Gaussian elimination
 * without pivoting on the band of half-width nx. M is column diagonally
 * dominant with non-positive off-diagonals, for which elimination without
 * pivoting is stable. */
int orc_band_solve(int nx, int ny, const double *bands, const double *R, double *x) {
  const long N = (long)nx * ny, W = 2L * nx + 1;
  double *A = (double *)calloc((size_t)(N * W), sizeof(double));
  if (!A) return -1;
#define AB(r, c) A[(size_t)(r) * (size_t)W + (size_t)((c) - (r) + nx)]
  for (long K = 0; K < N; ++K) {
    const long i = K % nx, j = K / nx;
    AB(K, K) = BAND(bands, N, 0)[K];
    if (i > 0) AB(K, K - 1) = BAND(bands, N, 1)[K];
    if (i < nx - 1) AB(K, K + 1) = BAND(bands, N, 2)[K];
    if (j > 0) AB(K, K - nx) = BAND(bands, N, 3)[K];
    if (j < ny - 1) AB(K, K + nx) = BAND(bands, N, 4)[K];
    x[K] = R[K];
  }
  for (long k = 0; k < N; ++k) {
    const double piv = AB(k, k);
    if (!(piv != 0.0) || !isfinite(piv)) { free(A); return -2; }
    const long rmax = (k + nx < N - 1) ? k + nx : N - 1;
    for (long r = k + 1; r <= rmax; ++r) {
      const double a = AB(r, k);
      if (a == 0.0) continue;
      const double l = a / piv;
      AB(r, k) = 0.0;
      for (long c = k + 1; c <= rmax; ++c) AB(r, c) -= l * AB(k, c);
      x[r] -= l * x[k];
    }
  }
  for (long k = N - 1; k >= 0; --k) {
    const long cmax = (k + nx < N - 1) ? k + nx : N - 1;
    double s = x[k];
    for (long c = k + 1; c <= cmax; ++c) s -= AB(k, c) * x[c];
    x[k] = s / AB(k, k);
  }
#undef AB
  free(A);
  return 0;
}

/* Solver::update, source/Solver.cc:270-290 (vertex values come first here: the
 * reference refreshes them at the end of the previous update and in init). */
int orc_step(const orc_problem *p, const double *f_in, double *f_out) {
  const int nx = p->nx, ny = p->ny;
  const size_t N = (size_t)nx * ny;
  double *V = (double *)malloc(sizeof(double) * (size_t)(nx + 1) * (size_t)(ny + 1));
  double *bands = (double *)malloc(sizeof(double) * 5 * N);
  double *R = (double *)malloc(sizeof(double) * N);
  double *x = (double *)malloc(sizeof(double) * N);
  int rc = -1;
  if (V && bands && R && x) {
    orc_vertex(p, f_in, V);
    orc_assemble(p, f_in, V, bands, R);
    rc = orc_band_solve(nx, ny, bands, R, x);
    if (rc == 0)
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) f_out[(size_t)i * ny + j] = x[(size_t)j * nx + i];
  }
  free(V); free(bands); free(R); free(x);
  return rc;
}

int orc_run(const orc_problem *p, double *f, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    const int rc = orc_step(p, f, f);
    if (rc) return rc;
  }
  return 0;
}
