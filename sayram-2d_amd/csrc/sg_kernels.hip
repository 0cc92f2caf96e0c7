// sg_kernels.hip -- hand-written gfx950 kernels of the implicit PPFV step.
//
// Every kernel is batched: a launch covers all instances of a batch through a
// block map (cell-parallel kernels, 256 cells of one instance per workgroup) or
// a line map (line solves, one row or column per lane). A plain single-instance
// solve is a batch of one.
//
// Nothing here is a dense contraction, so MFMA is not used: every kernel is
// bound by HBM bandwidth or by latency, and is written for coalesced FP64/FP32
// streams along K = j*nx + i.
#include "sg_kernels.hpp"

#include <cfloat>

namespace sg {

static constexpr double kEps = DBL_EPSILON;   // gEPS, source/common.h:38

template <typename T> __device__ __forceinline__ T *ptr(void *p) { return static_cast<T *>(p); }
template <typename T> __device__ __forceinline__ const T *cptr(const void *p) {
  return static_cast<const T *>(p);
}

// ---------------------------------------------------------------------------
// reductions: wave64 shuffles, LDS across the 4 waves, then a deterministic
// per-instance "last workgroup" pass over the partial sums.
// ---------------------------------------------------------------------------
template <typename A, int NV>
__device__ __forceinline__ void block_sum(A (&v)[NV]) {
  __shared__ A red[NV][kBlock / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    A s = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    v[k] = s;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      A s = red[k][0];
      for (int w = 1; w < kBlock / 64; ++w) s += red[k][w];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Writes this workgroup's NV partial sums, and returns true in the last
// workgroup of the instance to finish, with the instance totals in tot[] (in
// thread 0). The totals are summed in a fixed order, so results do not depend
// on scheduling or on which other instances share the launch.
template <typename A, int NV>
__device__ bool instance_reduce(const InstDev &I, int lblk, A (&v)[NV], A (&tot)[NV],
                                int which_ticket) {
  __shared__ int s_last;
  block_sum<A, NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) I.partial[(size_t)k * I.nblk0 + lblk] = (double)v[k];
    __threadfence();
    const unsigned t = atomicAdd(&I.ticket[which_ticket], 1u);
    s_last = (t == (unsigned)(I.nblk0 - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  A acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = A(0);
  for (int b = threadIdx.x; b < I.nblk0; b += kBlock) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] += (A)__builtin_nontemporal_load(&I.partial[(size_t)k * I.nblk0 + b]);
  }
  block_sum<A, NV>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = acc[k];
    I.ticket[which_ticket] = 0u;   // reusable by the next launch without a memset
  }
  return true;
}

// ---------------------------------------------------------------------------
// step start: statuses, tolerances and the active counters (a kernel rather
// than a memset, so the sequence can be captured into a graph)
// ---------------------------------------------------------------------------
__global__ void k_begin(const InstDev *__restrict__ insts, int n, const int *__restrict__ part,
                        int *__restrict__ active, int *__restrict__ oactive, double tol2,
                        double otol2, int maxit, int omaxit) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    KScal &s = *insts[k].sc;
    const int p = part[k];
    s.part = p;
    s.status = p ? ST_ACTIVE : ST_IDLE;
    s.ostatus = p ? ST_ACTIVE : ST_IDLE;
    s.tol2 = tol2; s.otol2 = otol2; s.maxit = maxit; s.omaxit = omaxit;
    s.iters = 0; s.titers = 0; s.orefs = 0; s.clamps = 0;
    mine += p;
  }
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) { *active = cnt; *oactive = cnt; }
}

__global__ void k_copy_counter(const int *__restrict__ src, int *__restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

// ---------------------------------------------------------------------------
// assembly (Solver::assemble, source/Solver.cc:167-267)
// ---------------------------------------------------------------------------

// Vertex value (fill_vertex_from_cells + fill_vertex_from_bcs,
// source/Solver.cc:299-422). Dirichlet sides were written in the order XMIN,
// XMAX, YMIN, YMAX there, so the precedence here is the reverse.
__device__ double vertex_value(const InstDev &I, const double *__restrict__ f, int vi, int vj) {
#pragma clang fp contract(off)
  const int nx = I.nx, ny = I.ny;
  if (vj == ny && (I.bcmask & 8)) return I.bcv[3][vi];
  if (vj == 0 && (I.bcmask & 4)) return I.bcv[2][vi];
  if (vi == nx && (I.bcmask & 2)) return I.bcv[1][vj];
  if (vi == 0 && (I.bcmask & 1)) return I.bcv[0][vj];
  const bool bx = (vi == 0 || vi == nx), by = (vj == 0 || vj == ny);
  if (bx && by) {   // corners copy the adjacent cell (:316-319)
    const int ci = vi == 0 ? 0 : nx - 1, cj = vj == 0 ? 0 : ny - 1;
    return f[(size_t)cj * nx + ci];
  }
  if (bx) {         // vertical boundary edge: linear in y (:355-366)
    const int ci = vi == 0 ? 0 : nx - 1;
    return I.wy0[vj] * f[(size_t)(vj - 1) * nx + ci] + I.wy1[vj] * f[(size_t)vj * nx + ci];
  }
  if (by) {         // horizontal boundary edge: linear in x (:371-381)
    const int cj = vj == 0 ? 0 : ny - 1;
    return I.wx0[vi] * f[(size_t)cj * nx + vi - 1] + I.wx1[vi] * f[(size_t)cj * nx + vi];
  }
  const double wxL = I.wx0[vi], wxR = I.wx1[vi], wyB = I.wy0[vj], wyT = I.wy1[vj];
  const size_t kb = (size_t)(vj - 1) * nx, kt = (size_t)vj * nx;
  return wxL * wyB * f[kb + vi - 1] + wxR * wyB * f[kb + vi] + wxL * wyT * f[kt + vi - 1] +
         wxR * wyT * f[kt + vi];
}

// a_sigma_func (source/Solver.cc:68-97) with the edge table of
// Mesh::build_connectivity_ (source/Mesh.cc:67-155). The operations are those
// of the reference in the same order; |B - A| is written as |dx| + |dy|, which
// equals sqrt(dx^2 + dy^2) exactly because one of the two is zero.
__device__ __forceinline__ void a_sigma(const double *__restrict__ xe, const double *__restrict__ ye,
                                        int i, int j, int inbr, double Lxx, double Lxy, double Lyy,
                                        double fA, double fB, double &aA, double &aB, double &as) {
#pragma clang fp contract(off)
  double Ax, Ay, Bx, By, n0, n1;
  if (inbr == 0) {        // im, XNEG
    Ax = xe[i]; Ay = ye[j + 1]; Bx = xe[i]; By = ye[j]; n0 = -1.0; n1 = 0.0;
  } else if (inbr == 1) { // jp, YPOS
    Ax = xe[i + 1]; Ay = ye[j + 1]; Bx = xe[i]; By = ye[j + 1]; n0 = 0.0; n1 = 1.0;
  } else if (inbr == 2) { // ip, XPOS
    Ax = xe[i + 1]; Ay = ye[j]; Bx = xe[i + 1]; By = ye[j + 1]; n0 = 1.0; n1 = 0.0;
  } else {                // jm, YNEG
    Ax = xe[i]; Ay = ye[j]; Bx = xe[i + 1]; By = ye[j]; n0 = 0.0; n1 = -1.0;
  }
  const double Kx = 0.5 * (xe[i] + xe[i + 1]);
  const double Ky = 0.5 * (ye[j] + ye[j + 1]);
  const double len = fabs(Bx - Ax) + fabs(By - Ay);
  const double vkbx = Bx - Kx, vkby = By - Ky;
  const double vkax = Ax - Kx, vkay = Ay - Ky;
  const double rvkb0 = vkby, rvkb1 = -vkbx;
  const double rvka0 = vkay, rvka1 = -vkax;
  const double t0 = n0 * Lxx + n1 * Lxy;
  const double t1 = n0 * Lxy + n1 * Lyy;
  const double numa = t0 * rvkb0 + t1 * rvkb1;
  const double dena = vkax * rvkb0 + vkay * rvkb1;
  const double numb = t0 * rvka0 + t1 * rvka1;
  const double denb = vkbx * rvka0 + vkby * rvka1;
  aA = len * numa / dena;
  aB = len * numb / denb;
  as = aA * fA + aB * fB;
}

// apply_inner_face_pair (source/Solver.cc:99-141) for the face owned by cell
// (io, jo): its west face (inbr 0) or south face (inbr 3). vA/vB are the
// owner's edge vertex values; the neighbour sees the same edge reversed.
__device__ __forceinline__ void face_pair(const double *__restrict__ xe,
                                          const double *__restrict__ ye, int io, int jo, int inbr,
                                          const double (&LK)[3], const double (&LL)[3], double fK,
                                          double fL, double vA, double vB, double &AK, double &AL) {
#pragma clang fp contract(off)
  double aKA, aKB, asK, aLA, aLB, asL;
  a_sigma(xe, ye, io, jo, inbr, LK[0], LK[1], LK[2], vA, vB, aKA, aKB, asK);
  const int li = inbr == 0 ? io - 1 : io, lj = inbr == 0 ? jo : jo - 1;
  a_sigma(xe, ye, li, lj, inbr == 0 ? 2 : 1, LL[0], LL[1], LL[2], vB, vA, aLA, aLB, asL);
  const double denom = fabs(asK) + fabs(asL) + 2 * kEps;     // calculate_mu, Solver.h:63-67
  const double muK = (fabs(asL) + kEps) / denom;
  const double muL = 1 - muK;
  const double Bs = muL * asL - muK * asK;
  const double Ba = fabs(Bs);
  const double Bp = (Ba + Bs) / 2.0, Bm = (Ba - Bs) / 2.0;
  AK = muK * (aKA + aKB) + Bp / (fK + kEps);
  AL = muL * (aLA + aLB) + Bm / (fL + kEps);
}

// apply_dirichlet_face (source/Solver.cc:143-164)
__device__ __forceinline__ void dirichlet_face(const double *__restrict__ xe,
                                               const double *__restrict__ ye, int i, int j,
                                               int inbr, const double (&LK)[3], double fK,
                                               double vA, double vB, double &AK, double &Bm) {
#pragma clang fp contract(off)
  double aA, aB, as;
  a_sigma(xe, ye, i, j, inbr, LK[0], LK[1], LK[2], vA, vB, aA, aB, as);
  const double Bs = -as;
  const double Ba = fabs(Bs);
  const double Bp = (Ba + Bs) / 2.0;
  Bm = (Ba - Bs) / 2.0;
  AK = aA + aB + Bp / (fK + kEps);
}

// One work-item per cell writes row K of M (d, w, e, s, n) and R. The diagonal
// is summed in the order the reference pushes its triplets (west pairs, south
// pairs, Dirichlet XMIN/XMAX/YMIN/YMAX, time term), which is the order Eigen's
// setFromTriplets adds duplicates; R likewise. Also seeds x = f (warm start).
__global__ void __launch_bounds__(kBlock) k_assemble(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map) {
#pragma clang fp contract(off)
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus == ST_IDLE) return;
  const int nx = I.nx, ny = I.ny;
  const int K = m.k0 + threadIdx.x;
  if (K >= nx * ny) return;
  const int j = K / nx, i = K - j * nx;
  const double *__restrict__ f = I.f;
  const double *xe = I.xe, *ye = I.ye;
  const double vSW = vertex_value(I, f, i, j), vSE = vertex_value(I, f, i + 1, j);
  const double vNW = vertex_value(I, f, i, j + 1), vNE = vertex_value(I, f, i + 1, j + 1);
  const double LK[3] = {I.Lxx[K], I.Lxy[K], I.Lyy[K]};
  const double fK = f[K];
  double d = 0.0, R = 0.0, bw = 0.0, be = 0.0, bs = 0.0, bn = 0.0, AK, AL;
  if (i > 0) {   // own west face (west-pair loop, :173-175)
    const double LL[3] = {I.Lxx[K - 1], I.Lxy[K - 1], I.Lyy[K - 1]};
    face_pair(xe, ye, i, j, 0, LK, LL, fK, f[K - 1], vNW, vSW, AK, AL);
    d += AK;
    bw = 0.0 + -AL;
  }
  if (i < nx - 1) {   // east face, owned by (i+1, j): this cell is its L
    const double LO[3] = {I.Lxx[K + 1], I.Lxy[K + 1], I.Lyy[K + 1]};
    face_pair(xe, ye, i + 1, j, 0, LO, LK, f[K + 1], fK, vNE, vSE, AK, AL);
    d += AL;
    be = 0.0 + -AK;
  }
  if (j > 0) {   // own south face (south-pair loop, :178-182)
    const double LL[3] = {I.Lxx[K - nx], I.Lxy[K - nx], I.Lyy[K - nx]};
    face_pair(xe, ye, i, j, 3, LK, LL, fK, f[K - nx], vSW, vSE, AK, AL);
    d += AK;
    bs = 0.0 + -AL;
  }
  if (j < ny - 1) {   // north face, owned by (i, j+1)
    const double LO[3] = {I.Lxx[K + nx], I.Lxy[K + nx], I.Lyy[K + nx]};
    face_pair(xe, ye, i, j + 1, 3, LO, LK, f[K + nx], fK, vNW, vNE, AK, AL);
    d += AL;
    bn = 0.0 + -AK;
  }
  double Bm;
  if (i == 0 && (I.bcmask & 1)) {        // XMIN: face im
    dirichlet_face(xe, ye, i, j, 0, LK, fK, vNW, vSW, AK, Bm);
    d += AK; R += Bm;
  }
  if (i == nx - 1 && (I.bcmask & 2)) {   // XMAX: face ip
    dirichlet_face(xe, ye, i, j, 2, LK, fK, vSE, vNE, AK, Bm);
    d += AK; R += Bm;
  }
  if (j == 0 && (I.bcmask & 4)) {        // YMIN: face jm
    dirichlet_face(xe, ye, i, j, 3, LK, fK, vSW, vSE, AK, Bm);
    d += AK; R += Bm;
  }
  if (j == ny - 1 && (I.bcmask & 8)) {   // YMAX: face jp
    dirichlet_face(xe, ye, i, j, 1, LK, fK, vNE, vNW, AK, Bm);
    d += AK; R += Bm;
  }
  // time / loss diagonal (:190-199), cell_area_dt (source/Mesh.h:61-63)
  const double dx = xe[i + 1] - xe[i], dy = ye[j + 1] - ye[j];
  const double UKK = I.G[K] * (dx * dy / I.dt);
  d += UKK * (1 + I.dt * I.itau[K]);
  R += UKK * fK;
  I.band64[BC][K] = d;
  I.band64[BW][K] = bw;
  I.band64[BE][K] = be;
  I.band64[BS][K] = bs;
  I.band64[BN][K] = bn;
  I.R[K] = R;
  I.x[K] = fK;
}

// FP64 level-0 bands -> FP32 copies (mixed precision)
__global__ void __launch_bounds__(kBlock) k_bands_to_f32(const InstDev *__restrict__ insts,
                                                        const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus == ST_IDLE) return;
  const int K = m.k0 + threadIdx.x;
  if (K >= I.nx * I.ny) return;
#pragma unroll
  for (int b = 0; b < 5; ++b) ptr<float>(I.lev[0].band[b])[K] = (float)I.band64[b][K];
}

// ---------------------------------------------------------------------------
// Black-box multigrid hierarchy (operator-dependent interpolation on a
// vertex-centred coarsening: coarse points are the even (i, j) of the level).
// P is rebuilt from the operator every step, R = P^T, and the coarse operator
// A_{l+1} = P^T A_l P is 9-point; the whole chain is FP64.
//
// Prolongation weights of fine point (i, j) refer to the coarse points
// (i>>1, j>>1), (i>>1 + 1, j>>1), (i>>1, j>>1 + 1), (i>>1 + 1, j>>1 + 1).
// ---------------------------------------------------------------------------
struct St9 { double c, w, e, s, n, sw, se, nw, ne; };

template <int NB>
__device__ __forceinline__ St9 load_st(const Lev &L, size_t K) {
  St9 a;
  a.c = L.band64[BC][K]; a.w = L.band64[BW][K]; a.e = L.band64[BE][K];
  a.s = L.band64[BS][K]; a.n = L.band64[BN][K];
  if (NB == 9) {
    a.sw = L.band64[BSW][K]; a.se = L.band64[BSE][K];
    a.nw = L.band64[BNW][K]; a.ne = L.band64[BNE][K];
  } else {
    a.sw = a.se = a.nw = a.ne = 0.0;
  }
  return a;
}

__device__ __forceinline__ double safe_den(double d, double c) {
  return (d > 0.0 && d < INFINITY) ? d : c;
}

// x-edge point (i odd, j even): interpolate from west/east coarse points with
// the stencil collapsed along y.
template <int NB>
__device__ __forceinline__ void w_xedge(const Lev &L, size_t K, int i, double &wW, double &wE) {
  const St9 a = load_st<NB>(L, K);
  const double den = safe_den(a.c + a.n + a.s, a.c);
  wW = -(a.w + a.nw + a.sw) / den;
  wE = (i + 1 < L.nx) ? -(a.e + a.ne + a.se) / den : 0.0;
}

// y-edge point (i even, j odd): from south/north, stencil collapsed along x.
template <int NB>
__device__ __forceinline__ void w_yedge(const Lev &L, size_t K, int j, double &wS, double &wN) {
  const St9 a = load_st<NB>(L, K);
  const double den = safe_den(a.c + a.w + a.e, a.c);
  wS = -(a.s + a.sw + a.se) / den;
  wN = (j + 1 < L.ny) ? -(a.n + a.nw + a.ne) / den : 0.0;
}

template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_pweights(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map, int l) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus == ST_IDLE) return;
  const Lev &L = I.lev[l];
  const int nx = L.nx, ny = L.ny;
  const int K = m.k0 + threadIdx.x;
  if (K >= nx * ny) return;
  const int j = K / nx, i = K - j * nx;
  double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
  const bool io = i & 1, jo = j & 1;
  if (!io && !jo) {
    w0 = 1.0;
  } else if (io && !jo) {
    w_xedge<NB>(L, K, i, w0, w1);
  } else if (!io && jo) {
    w_yedge<NB>(L, K, j, w0, w2);
  } else {
    const St9 a = load_st<NB>(L, K);
    double sW = 0, nW = 0, sE = 0, nE = 0, wS = 0, eS = 0, wN = 0, eN = 0;
    w_yedge<NB>(L, K - 1, j, sW, nW);
    if (i + 1 < nx) w_yedge<NB>(L, K + 1, j, sE, nE);
    w_xedge<NB>(L, K - nx, i, wS, eS);
    if (j + 1 < ny) w_xedge<NB>(L, K + nx, i, wN, eN);
    const double c = a.c;
    w0 = -(a.w * sW + a.s * wS + a.sw) / c;
    w1 = -(a.e * sE + a.s * eS + a.se) / c;
    w2 = -(a.w * nW + a.n * wN + a.nw) / c;
    w3 = -(a.e * nE + a.n * eN + a.ne) / c;
  }
  ptr<T>(L.pw[0])[K] = (T)w0;
  ptr<T>(L.pw[1])[K] = (T)w1;
  ptr<T>(L.pw[2])[K] = (T)w2;
  ptr<T>(L.pw[3])[K] = (T)w3;
}

// A_{l+1}(C, C') = sum_f sum_g P(f, C) A_l(f, g) P(g, C'); one work-item per
// coarse point, whose support is the 3x3 fine block around (2I, 2J).
template <typename T, int NB, bool COPY>
__global__ void __launch_bounds__(kBlock) k_rap(const InstDev *__restrict__ insts,
                                               const BlkMap *__restrict__ map, int l) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus == ST_IDLE) return;
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int cx = Cl.nx, cy = Cl.ny, fx = F.nx, fy = F.ny;
  const int KC = m.k0 + threadIdx.x;
  if (KC >= cx * cy) return;
  const int J = KC / cx, Ic = KC - J * cx;
  const T *__restrict__ pw0 = cptr<T>(F.pw[0]);
  const T *__restrict__ pw1 = cptr<T>(F.pw[1]);
  const T *__restrict__ pw2 = cptr<T>(F.pw[2]);
  const T *__restrict__ pw3 = cptr<T>(F.pw[3]);
  double acc[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[a][c] = 0.0;
  const int di_[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1};
  const int dj_[9] = {0, 0, 0, -1, 1, -1, -1, 1, 1};
  for (int fj = 2 * J - 1; fj <= 2 * J + 1; ++fj) {
    if (fj < 0 || fj >= fy) continue;
    const int dJ = J - (fj >> 1);
    for (int fi = 2 * Ic - 1; fi <= 2 * Ic + 1; ++fi) {
      if (fi < 0 || fi >= fx) continue;
      const int dI = Ic - (fi >> 1);
      const size_t kf = (size_t)fj * fx + fi;
      const int slot = dI + 2 * dJ;
      const double wf = (double)(slot == 0 ? pw0[kf] : slot == 1 ? pw1[kf] : slot == 2 ? pw2[kf] : pw3[kf]);
      if (wf == 0.0) continue;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int gi = fi + di_[b], gj = fj + dj_[b];
        if (gi < 0 || gi >= fx || gj < 0 || gj >= fy) continue;
        const double av = F.band64[b][kf];
        if (av == 0.0) continue;
        const double wa = wf * av;
        const size_t kg = (size_t)gj * fx + gi;
        const int oi = (gi >> 1) - Ic + 1, oj = (gj >> 1) - J + 1;
        const double wg[4] = {(double)pw0[kg], (double)pw1[kg], (double)pw2[kg], (double)pw3[kg]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (wg[q] == 0.0) continue;
          const int ai = oi + (q & 1), aj = oj + (q >> 1);
          if (ai < 0 || ai > 2 || aj < 0 || aj > 2) continue;
          acc[aj][ai] += wa * wg[q];
        }
      }
    }
  }
  const double out[9] = {acc[1][1], acc[1][0], acc[1][2], acc[0][1], acc[2][1],
                         acc[0][0], acc[0][2], acc[2][0], acc[2][2]};
#pragma unroll
  for (int b = 0; b < 9; ++b) {
    Cl.band64[b][KC] = out[b];
    if (COPY) ptr<float>(Cl.band[b])[KC] = (float)out[b];
  }
}

// ---------------------------------------------------------------------------
// multigrid cycle kernels (precision T)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void level_xb(const InstDev &I, int l, int xid, int bid, T *&x,
                                         const T *&b) {
  if (l == 0) { x = ptr<T>(I.kv[xid]); b = cptr<T>(I.kv[bid]); }
  else { x = ptr<T>(I.lev[l].x); b = cptr<T>(I.lev[l].b); }
}

// Zebra line Gauss-Seidel: every lane solves one row (DIR 0) or column (DIR 1)
// exactly with the Thomas algorithm, with the off-line couplings moved to the
// right-hand side. ZERO: the other lines are zero (first sweep of a cycle).
template <typename T, int NB, int DIR, bool ZERO>
__global__ void __launch_bounds__(kLineBlock) k_line(const InstDev *__restrict__ insts,
                                                    const LineMap *__restrict__ map, int nlines,
                                                    int l, int xid, int bid) {
  const int g = blockIdx.x * kLineBlock + threadIdx.x;
  if (g >= nlines) return;
  const LineMap m = map[g];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const Lev &L = I.lev[l];
  const int nx = L.nx, ny = L.ny;
  T *x; const T *b;
  level_xb<T>(I, l, xid, bid, x, b);
  const T *__restrict__ bc = cptr<T>(L.band[BC]);
  const T *__restrict__ bw = cptr<T>(L.band[BW]);
  const T *__restrict__ be = cptr<T>(L.band[BE]);
  const T *__restrict__ bs = cptr<T>(L.band[BS]);
  const T *__restrict__ bn = cptr<T>(L.band[BN]);
  const T *bsw = nullptr, *bse = nullptr, *bnw = nullptr, *bne = nullptr;
  if (NB == 9) {
    bsw = cptr<T>(L.band[BSW]); bse = cptr<T>(L.band[BSE]);
    bnw = cptr<T>(L.band[BNW]); bne = cptr<T>(L.band[BNE]);
  }
  T *__restrict__ cp = ptr<T>(L.cp);
  T *__restrict__ dp = ptr<T>(L.dp);
  const int n = DIR == 0 ? nx : ny;
  const size_t k0 = DIR == 0 ? (size_t)m.line * nx : (size_t)m.line;
  const size_t st = DIR == 0 ? 1 : (size_t)nx;
  const int li = m.line;   // row j (DIR 0) or column i (DIR 1)
  T cprev = T(0), dprev = T(0);
  for (int t = 0; t < n; ++t) {
    const size_t K = k0 + (size_t)t * st;
    const int i = DIR == 0 ? t : li, j = DIR == 0 ? li : t;
    T rhs = b[K];
    T a, c;
    if (DIR == 0) {
      a = i > 0 ? bw[K] : T(0);
      c = i < nx - 1 ? be[K] : T(0);
      if (!ZERO) {
        if (j > 0) {
          rhs -= bs[K] * x[K - nx];
          if (NB == 9) {
            if (i > 0) rhs -= bsw[K] * x[K - nx - 1];
            if (i < nx - 1) rhs -= bse[K] * x[K - nx + 1];
          }
        }
        if (j < ny - 1) {
          rhs -= bn[K] * x[K + nx];
          if (NB == 9) {
            if (i > 0) rhs -= bnw[K] * x[K + nx - 1];
            if (i < nx - 1) rhs -= bne[K] * x[K + nx + 1];
          }
        }
      }
    } else {
      a = j > 0 ? bs[K] : T(0);
      c = j < ny - 1 ? bn[K] : T(0);
      if (!ZERO) {
        if (i > 0) {
          rhs -= bw[K] * x[K - 1];
          if (NB == 9) {
            if (j > 0) rhs -= bsw[K] * x[K - nx - 1];
            if (j < ny - 1) rhs -= bnw[K] * x[K + nx - 1];
          }
        }
        if (i < nx - 1) {
          rhs -= be[K] * x[K + 1];
          if (NB == 9) {
            if (j > 0) rhs -= bse[K] * x[K - nx + 1];
            if (j < ny - 1) rhs -= bne[K] * x[K + nx + 1];
          }
        }
      }
    }
    const T inv = T(1) / (bc[K] - a * cprev);
    cprev = c * inv;
    dprev = (rhs - a * dprev) * inv;
    cp[K] = cprev;
    dp[K] = dprev;
  }
  T xn = dprev;
  x[k0 + (size_t)(n - 1) * st] = xn;
  for (int t = n - 2; t >= 0; --t) {
    const size_t K = k0 + (size_t)t * st;
    xn = dp[K] - cp[K] * xn;
    x[K] = xn;
  }
}

template <typename T, int NB>
__device__ __forceinline__ T stencil_apply(const Lev &L, size_t K, int i, int j, const T *x) {
  const int nx = L.nx, ny = L.ny;
  T s = cptr<T>(L.band[BC])[K] * x[K];
  if (i > 0) s += cptr<T>(L.band[BW])[K] * x[K - 1];
  if (i < nx - 1) s += cptr<T>(L.band[BE])[K] * x[K + 1];
  if (j > 0) {
    s += cptr<T>(L.band[BS])[K] * x[K - nx];
    if (NB == 9) {
      if (i > 0) s += cptr<T>(L.band[BSW])[K] * x[K - nx - 1];
      if (i < nx - 1) s += cptr<T>(L.band[BSE])[K] * x[K - nx + 1];
    }
  }
  if (j < ny - 1) {
    s += cptr<T>(L.band[BN])[K] * x[K + nx];
    if (NB == 9) {
      if (i > 0) s += cptr<T>(L.band[BNW])[K] * x[K + nx - 1];
      if (i < nx - 1) s += cptr<T>(L.band[BNE])[K] * x[K + nx + 1];
    }
  }
  return s;
}

// r_l = b_l - A_l x_l into the level's scratch (the line-solve buffer cp is
// free between the pre- and post-smoothing sweeps). ZERO: x_l = 0.
template <typename T, int NB, bool ZERO>
__global__ void __launch_bounds__(kBlock) k_residual(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map, int l,
                                                    int xid, int bid) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const Lev &F = I.lev[l];
  const int K = m.k0 + threadIdx.x;
  if (K >= F.nx * F.ny) return;
  const int j = K / F.nx, i = K - j * F.nx;
  T *x; const T *b;
  level_xb<T>(I, l, xid, bid, x, b);
  ptr<T>(F.cp)[K] = ZERO ? b[K] : b[K] - stencil_apply<T, NB>(F, K, i, j, x);
}

// b_{l+1}(C) = sum_f P(f, C) r_l(f) over the 3x3 fine block around (2I, 2J)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map, int l) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int KC = m.k0 + threadIdx.x;
  if (KC >= Cl.nx * Cl.ny) return;
  const int J = KC / Cl.nx, Ic = KC - J * Cl.nx;
  const T *__restrict__ r = cptr<T>(F.cp);
  T s = T(0);
  for (int fj = 2 * J - 1; fj <= 2 * J + 1; ++fj) {
    if (fj < 0 || fj >= F.ny) continue;
    const int dJ = J - (fj >> 1);
    for (int fi = 2 * Ic - 1; fi <= 2 * Ic + 1; ++fi) {
      if (fi < 0 || fi >= F.nx) continue;
      const int slot = (Ic - (fi >> 1)) + 2 * dJ;
      const size_t kf = (size_t)fj * F.nx + fi;
      s += cptr<T>(F.pw[slot])[kf] * r[kf];
    }
  }
  ptr<T>(Cl.b)[KC] = s;
}

// x_l += P x_{l+1}. ZERO: x_l = P x_{l+1}.
template <typename T, bool ZERO>
__global__ void __launch_bounds__(kBlock) k_prolong(const InstDev *__restrict__ insts,
                                                   const BlkMap *__restrict__ map, int l,
                                                   int xid) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int K = m.k0 + threadIdx.x;
  if (K >= F.nx * F.ny) return;
  const int j = K / F.nx, i = K - j * F.nx;
  const int I0 = i >> 1, J0 = j >> 1;
  const int I1 = I0 + 1 < Cl.nx ? I0 + 1 : I0, J1 = J0 + 1 < Cl.ny ? J0 + 1 : J0;
  const T *__restrict__ xc = cptr<T>(Cl.x);
  const size_t r0 = (size_t)J0 * Cl.nx, r1 = (size_t)J1 * Cl.nx;
  // weights of parents outside the coarse grid are zero
  const T e = cptr<T>(F.pw[0])[K] * xc[r0 + I0] + cptr<T>(F.pw[1])[K] * xc[r0 + I1] +
              cptr<T>(F.pw[2])[K] * xc[r1 + I0] + cptr<T>(F.pw[3])[K] * xc[r1 + I1];
  T *x = l == 0 ? ptr<T>(I.kv[xid]) : ptr<T>(F.x);
  x[K] = ZERO ? e : x[K] + e;
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// BiCGSTAB (right-preconditioned), precision T for vectors and sums; the
// operator is the level-0 5-band matrix of the V-cycle precision.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T spmv5(const InstDev &I, size_t K, int i, int j, const T *x) {
  return stencil_apply<T, 5>(I.lev[0], K, i, j, x);
}

// FP64 start: r = R - M x, rh = r, p = v = 0; rho = (r, r), bnorm = ||R||.
__global__ void __launch_bounds__(kBlock) k_bicg_init64(const InstDev *__restrict__ insts,
                                                       const BlkMap *__restrict__ map,
                                                       int *__restrict__ active) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus == ST_IDLE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  double v[2] = {0.0, 0.0};
  if (K < N) {
    const int j = K / I.nx, i = K - j * I.nx;
    const double r = I.R[K] - spmv5<double>(I, K, i, j, I.x);
    ptr<double>(I.kv[VR])[K] = r;
    ptr<double>(I.kv[VRH])[K] = r;
    ptr<double>(I.kv[VP])[K] = 0.0;
    ptr<double>(I.kv[VV])[K] = 0.0;
    v[0] = r * r;
    v[1] = I.R[K] * I.R[K];
  }
  double tot[2];
  if (instance_reduce<double, 2>(I, m.k0 / kBlock, v, tot, 0) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    s.rho = tot[0]; s.rho_old = 1.0; s.alpha = 1.0; s.omega = 1.0; s.beta = 0.0;
    s.rnorm2 = tot[0]; s.bnorm2 = tot[1]; s.iters = 0;
    const bool done = tot[0] <= s.tol2 * tot[1];
    s.status = done ? ST_DONE : ST_ACTIVE;
    if (done) atomicSub(active, 1);
  }
}

// p = r + beta (p - omega v)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_p(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  if (K >= I.nx * I.ny) return;
  const T beta = (T)I.sc->beta, omega = (T)I.sc->omega;
  T *p = ptr<T>(I.kv[VP]);
  p[K] = cptr<T>(I.kv[VR])[K] + beta * (p[K] - omega * cptr<T>(I.kv[VV])[K]);
}

// v = M ph; alpha = rho / (rh, v)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_v(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  T acc[1] = {T(0)};
  if (K < N) {
    const int j = K / I.nx, i = K - j * I.nx;
    const T v = spmv5<T>(I, K, i, j, cptr<T>(I.kv[VPH]));
    ptr<T>(I.kv[VV])[K] = v;
    acc[0] = cptr<T>(I.kv[VRH])[K] * v;
  }
  T tot[1];
  if (instance_reduce<T, 1>(I, m.k0 / kBlock, acc, tot, 0) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    s.alpha = (double)((T)s.rho / tot[0]);
  }
}

// s = r - alpha v; if ||s|| is small the step ends at x + alpha ph (HALF)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_s(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  T acc[1] = {T(0)};
  if (K < N) {
    const T s = cptr<T>(I.kv[VR])[K] - (T)I.sc->alpha * cptr<T>(I.kv[VV])[K];
    ptr<T>(I.kv[VS])[K] = s;
    acc[0] = s * s;
  }
  T tot[1];
  if (instance_reduce<T, 1>(I, m.k0 / kBlock, acc, tot, 0) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    if ((double)tot[0] <= s.tol2 * s.bnorm2) { s.status = ST_HALF; s.rnorm2 = (double)tot[0]; }
  }
}

// t = M sh; omega = (t, s) / (t, t)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_t(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->status != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  T acc[2] = {T(0), T(0)};
  if (K < N) {
    const int j = K / I.nx, i = K - j * I.nx;
    const T t = spmv5<T>(I, K, i, j, cptr<T>(I.kv[VSH]));
    ptr<T>(I.kv[VT])[K] = t;
    acc[0] = t * cptr<T>(I.kv[VS])[K];
    acc[1] = t * t;
  }
  T tot[2];
  if (instance_reduce<T, 2>(I, m.k0 / kBlock, acc, tot, 0) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    s.omega = tot[1] > T(0) ? (double)(tot[0] / tot[1]) : 0.0;
  }
}

// x += alpha ph + omega sh; r = s - omega t; rho = (rh, r); convergence.
// XV: slot of the solution vector (the FP64 state x, or the FP32 inner
// correction in mixed precision).
template <typename T, bool XSTATE>
__global__ void __launch_bounds__(kBlock) k_bicg_x(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map,
                                                  int *__restrict__ active) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  const int st = I.sc->status;
  if (st != ST_ACTIVE && st != ST_HALF) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  const T alpha = (T)I.sc->alpha, omega = (T)I.sc->omega;
  T acc[2] = {T(0), T(0)};
  if (K < N) {
    T *x = XSTATE ? (T *)I.x : ptr<T>(I.kv[VXK]);
    if (st == ST_HALF) {
      x[K] = x[K] + alpha * cptr<T>(I.kv[VPH])[K];
    } else {
      x[K] = x[K] + alpha * cptr<T>(I.kv[VPH])[K] + omega * cptr<T>(I.kv[VSH])[K];
      const T r = cptr<T>(I.kv[VS])[K] - omega * cptr<T>(I.kv[VT])[K];
      ptr<T>(I.kv[VR])[K] = r;
      acc[0] = cptr<T>(I.kv[VRH])[K] * r;
      acc[1] = r * r;
    }
  }
  T tot[2];
  if (instance_reduce<T, 2>(I, m.k0 / kBlock, acc, tot, 1) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    s.iters += 1;
    s.titers += 1;
    if (st == ST_HALF) {
      s.status = ST_DONE;
      atomicSub(active, 1);
      return;
    }
    s.rho_old = s.rho;
    s.rho = (double)tot[0];
    s.rnorm2 = (double)tot[1];
    s.beta = (double)(((T)s.rho / (T)s.rho_old) * (alpha / omega));
    if (s.rnorm2 <= s.tol2 * s.bnorm2) {
      s.status = ST_DONE;
      atomicSub(active, 1);
    } else if (s.iters >= s.maxit || !(s.rnorm2 == s.rnorm2) || s.rho == 0.0 || s.omega == 0.0) {
      s.status = ST_FAIL;
      atomicSub(active, 1);
    }
  }
}

// ---------------------------------------------------------------------------
// mixed precision: FP64 residual and refinement around an FP32 solve
// ---------------------------------------------------------------------------
// r64 = R - M x (FP64 bands); decides whether another refinement is needed.
__global__ void __launch_bounds__(kBlock) k_ir_residual(const InstDev *__restrict__ insts,
                                                       const BlkMap *__restrict__ map,
                                                       int *__restrict__ oactive, int first) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  double v[2] = {0.0, 0.0};
  if (K < N) {
    const int j = K / I.nx, i = K - j * I.nx;
    const Lev &L = I.lev[0];
    double s = I.band64[BC][K] * I.x[K];
    if (i > 0) s += I.band64[BW][K] * I.x[K - 1];
    if (i < L.nx - 1) s += I.band64[BE][K] * I.x[K + 1];
    if (j > 0) s += I.band64[BS][K] * I.x[K - L.nx];
    if (j < L.ny - 1) s += I.band64[BN][K] * I.x[K + L.nx];
    const double r = I.R[K] - s;
    I.r64[K] = r;
    v[0] = r * r;
    v[1] = first ? I.R[K] * I.R[K] : 0.0;
  }
  double tot[2];
  if (instance_reduce<double, 2>(I, m.k0 / kBlock, v, tot, 0) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    if (first) s.obnorm2 = tot[1];
    s.onorm2 = tot[0];
    s.scale = sqrt(tot[0]);
    // Converged; or stalled at the FP64 rounding floor of R - M x (the
    // residual stopped halving within 100x of the tolerance).
    const bool conv = tot[0] <= s.otol2 * s.obnorm2;
    const bool floor = !first && tot[0] > 0.25 * s.oprev2 && tot[0] <= 1e4 * s.otol2 * s.obnorm2;
    s.oprev2 = tot[0];
    if (conv || floor) {
      s.ostatus = ST_DONE;
      s.status = ST_DONE;
      atomicSub(oactive, 1);
    } else if (s.orefs >= s.omaxit) {
      s.ostatus = ST_FAIL;
      s.status = ST_DONE;
      atomicSub(oactive, 1);
    } else {
      s.orefs += 1;
      s.status = ST_ACTIVE;
    }
  }
}

// inner start: r = r64 / ||r64|| (FP32), rh = r, p = v = xk = 0
__global__ void __launch_bounds__(kBlock) k_ir_inner_init(const InstDev *__restrict__ insts,
                                                         const BlkMap *__restrict__ map,
                                                         int *__restrict__ active) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  float v[1] = {0.0f};
  if (K < N) {
    const float r = (float)(I.r64[K] / I.sc->scale);
    ptr<float>(I.kv[VR])[K] = r;
    ptr<float>(I.kv[VRH])[K] = r;
    ptr<float>(I.kv[VP])[K] = 0.0f;
    ptr<float>(I.kv[VV])[K] = 0.0f;
    ptr<float>(I.kv[VXK])[K] = 0.0f;
    v[0] = r * r;
  }
  float tot[1];
  if (instance_reduce<float, 1>(I, m.k0 / kBlock, v, tot, 0) && threadIdx.x == 0) {
    KScal &s = *I.sc;
    s.rho = tot[0]; s.rho_old = 1.0; s.alpha = 1.0; s.omega = 1.0; s.beta = 0.0;
    s.rnorm2 = tot[0]; s.bnorm2 = tot[0]; s.iters = 0;
    s.status = ST_ACTIVE;
    (void)active;
  }
}

// x += ||r64|| * xk (FP64)
__global__ void __launch_bounds__(kBlock) k_ir_update(const InstDev *__restrict__ insts,
                                                     const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  if (K >= I.nx * I.ny) return;
  I.x[K] += I.sc->scale * (double)ptr<float>(I.kv[VXK])[K];
}

// f = max(x, 0); true residual of the unclamped x; min f; clamp count.
__global__ void __launch_bounds__(kBlock) k_finalize(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (I.sc->ostatus == ST_IDLE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  double v[4] = {0.0, 0.0, 0.0, 0.0};   // r^2, R^2, clamps, -min
  __shared__ double smin[kBlock];
  double mn = DBL_MAX;
  if (K < N) {
    const int j = K / I.nx, i = K - j * I.nx;
    const double xv = I.x[K];
    double s = I.band64[BC][K] * xv;
    if (i > 0) s += I.band64[BW][K] * I.x[K - 1];
    if (i < I.nx - 1) s += I.band64[BE][K] * I.x[K + 1];
    if (j > 0) s += I.band64[BS][K] * I.x[K - I.nx];
    if (j < I.ny - 1) s += I.band64[BN][K] * I.x[K + I.nx];
    const double r = I.R[K] - s;
    v[0] = r * r;
    v[1] = I.R[K] * I.R[K];
    v[2] = xv < 0.0 ? 1.0 : 0.0;
    mn = xv;
  }
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
    __syncthreads();
  }
  v[3] = 0.0;
  // min is carried through the partial array as a separate slot by the last block
  double tot[3];
  double v3[3] = {v[0], v[1], v[2]};
  const double bmin = smin[0];
  if (threadIdx.x == 0) I.partial[(size_t)3 * I.nblk0 + m.k0 / kBlock] = bmin;
  if (instance_reduce<double, 3>(I, m.k0 / kBlock, v3, tot, 0)) {
    __shared__ double sm2[kBlock];
    double mm = DBL_MAX;
    for (int b = threadIdx.x; b < I.nblk0; b += kBlock)
      mm = fmin(mm, __builtin_nontemporal_load(&I.partial[(size_t)3 * I.nblk0 + b]));
    sm2[threadIdx.x] = mm;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) sm2[threadIdx.x] = fmin(sm2[threadIdx.x], sm2[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      KScal &s = *I.sc;
      s.true_rel = tot[1] > 0.0 ? sqrt(tot[0] / tot[1]) : sqrt(tot[0]);
      s.clamps = (int)tot[2];
      s.minf = sm2[0];
    }
  }
  if (K < N) {
    const double xv = I.x[K];
    I.f[K] = xv > 0.0 ? xv : 0.0;   // min f >= 0 enforced exactly
  }
}

// (nx,ny) i-major <-> K order (j*nx+i) transposes for the host boundary
__global__ void __launch_bounds__(kBlock) k_to_korder(const double *__restrict__ src,
                                                     double *__restrict__ dst, int nx, int ny) {
  const long long K = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (K >= (long long)nx * ny) return;
  const int j = (int)(K / nx), i = (int)(K - (long long)j * nx);
  dst[K] = src[(size_t)i * ny + j];
}
__global__ void __launch_bounds__(kBlock) k_to_imajor(const double *__restrict__ src,
                                                     double *__restrict__ dst, int nx, int ny) {
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (q >= (long long)nx * ny) return;
  const int i = (int)(q / ny), j = (int)(q - (long long)i * ny);
  dst[q] = src[(size_t)j * nx + i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_begin(hipStream_t s, const InstDev *d, int n, const int *part, int *active,
                  int *oactive, double tol2, double otol2, int maxit, int omaxit) {
  k_begin<<<1, 1024, 0, s>>>(d, n, part, active, oactive, tol2, otol2, maxit, omaxit);
}
void launch_copy_counter(hipStream_t s, const int *src, int *dst) {
  k_copy_counter<<<1, 64, 0, s>>>(src, dst);
}
void launch_assemble(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_assemble<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_bands_to_f32(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_bands_to_f32<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_pweights(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool mixed) {
  if (!nb) return;
  if (l == 0) {
    if (mixed) k_pweights<float, 5><<<nb, kBlock, 0, s>>>(d, m, l);
    else k_pweights<double, 5><<<nb, kBlock, 0, s>>>(d, m, l);
  } else {
    if (mixed) k_pweights<float, 9><<<nb, kBlock, 0, s>>>(d, m, l);
    else k_pweights<double, 9><<<nb, kBlock, 0, s>>>(d, m, l);
  }
}
void launch_rap(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool mixed) {
  if (!nb) return;
  if (l == 0) {
    if (mixed) k_rap<float, 5, true><<<nb, kBlock, 0, s>>>(d, m, l);
    else k_rap<double, 5, false><<<nb, kBlock, 0, s>>>(d, m, l);
  } else {
    if (mixed) k_rap<float, 9, true><<<nb, kBlock, 0, s>>>(d, m, l);
    else k_rap<double, 9, false><<<nb, kBlock, 0, s>>>(d, m, l);
  }
}

template <typename T>
void launch_line(hipStream_t s, const InstDev *d, const LineMap *m, int n, int l, int dir,
                 bool zero, int xid, int bid) {
  if (!n) return;
  const unsigned g = nblocks(n, kLineBlock);
  const bool nine = l > 0;
#define SG_LINE(NB, DIR, Z) k_line<T, NB, DIR, Z><<<g, kLineBlock, 0, s>>>(d, m, n, l, xid, bid)
  if (nine) {
    if (dir == 0) { if (zero) SG_LINE(9, 0, true); else SG_LINE(9, 0, false); }
    else { if (zero) SG_LINE(9, 1, true); else SG_LINE(9, 1, false); }
  } else {
    if (dir == 0) { if (zero) SG_LINE(5, 0, true); else SG_LINE(5, 0, false); }
    else { if (zero) SG_LINE(5, 1, true); else SG_LINE(5, 1, false); }
  }
#undef SG_LINE
}
template void launch_line<double>(hipStream_t, const InstDev *, const LineMap *, int, int, int,
                                  bool, int, int);
template void launch_line<float>(hipStream_t, const InstDev *, const LineMap *, int, int, int,
                                 bool, int, int);

template <typename T>
void launch_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool zero,
                     int xid, int bid) {
  if (!nb) return;
  if (l == 0) {
    if (zero) k_residual<T, 5, true><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
    else k_residual<T, 5, false><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
  } else {
    if (zero) k_residual<T, 9, true><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
    else k_residual<T, 9, false><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
  }
}
template void launch_residual<double>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                      int, int);
template void launch_residual<float>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                     int, int);

template <typename T>
void launch_restrict(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l) {
  if (nb) k_restrict<T><<<nb, kBlock, 0, s>>>(d, m, l);
}
template void launch_restrict<double>(hipStream_t, const InstDev *, const BlkMap *, int, int);
template void launch_restrict<float>(hipStream_t, const InstDev *, const BlkMap *, int, int);

template <typename T>
void launch_prolong(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool zero,
                    int xid) {
  if (!nb) return;
  if (zero) k_prolong<T, true><<<nb, kBlock, 0, s>>>(d, m, l, xid);
  else k_prolong<T, false><<<nb, kBlock, 0, s>>>(d, m, l, xid);
}
template void launch_prolong<double>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                     int);
template void launch_prolong<float>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                    int);

void launch_bicg_init64(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int *active) {
  if (nb) k_bicg_init64<<<nb, kBlock, 0, s>>>(d, m, active);
}
template <typename T>
void launch_bicg(hipStream_t s, int which, const InstDev *d, const BlkMap *m, int nb,
                 int *active, bool xstate) {
  if (!nb) return;
  switch (which) {
    case 0: k_bicg_p<T><<<nb, kBlock, 0, s>>>(d, m); break;
    case 1: k_bicg_v<T><<<nb, kBlock, 0, s>>>(d, m); break;
    case 2: k_bicg_s<T><<<nb, kBlock, 0, s>>>(d, m); break;
    case 3: k_bicg_t<T><<<nb, kBlock, 0, s>>>(d, m); break;
    default:
      if (xstate) k_bicg_x<T, true><<<nb, kBlock, 0, s>>>(d, m, active);
      else k_bicg_x<T, false><<<nb, kBlock, 0, s>>>(d, m, active);
  }
}
template void launch_bicg<double>(hipStream_t, int, const InstDev *, const BlkMap *, int, int *,
                                  bool);
template void launch_bicg<float>(hipStream_t, int, const InstDev *, const BlkMap *, int, int *,
                                 bool);

void launch_ir_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int *oactive,
                        int first) {
  if (nb) k_ir_residual<<<nb, kBlock, 0, s>>>(d, m, oactive, first);
}
void launch_ir_inner_init(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int *active) {
  if (nb) k_ir_inner_init<<<nb, kBlock, 0, s>>>(d, m, active);
}
void launch_ir_update(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_ir_update<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_finalize(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_finalize<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_to_korder(hipStream_t s, const double *src, double *dst, int nx, int ny) {
  k_to_korder<<<nblocks((long long)nx * ny, kBlock), kBlock, 0, s>>>(src, dst, nx, ny);
}
void launch_to_imajor(hipStream_t s, const double *src, double *dst, int nx, int ny) {
  k_to_imajor<<<nblocks((long long)nx * ny, kBlock), kBlock, 0, s>>>(src, dst, nx, ny);
}

}  // namespace sg
