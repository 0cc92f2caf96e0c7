// sg_kernels.hip -- hand-written gfx950 kernels of the implicit PPFV step.
//
// Every kernel is batched: a launch covers all instances of a batch through a
// block map (cell-parallel kernels, 256 cells of one instance per workgroup) or
// a line map (line solves, one row or column per lane). A plain single-instance
// solve is a batch of one.
//
// Nothing here is a dense contraction, so MFMA is not used: every kernel is
// bound by HBM bandwidth or by latency, and is written for coalesced FP64/FP32
// streams over the device cell order of sg_internal.hpp (rows split into their
// even and odd columns). A cell kernel's work-item K is a storage slot; it
// decodes (i, j) with cell_ij and finds its x-neighbours with xslot.
#include "sg_kernels.hpp"
#include "sayram_gpu.h"   // sg_precision values of InstDev::prec

#include <algorithm>
#include <cfloat>

namespace sg {

static constexpr double kEps = DBL_EPSILON;   // gEPS, source/common.h:38

template <typename T> __device__ __forceinline__ T *ptr(void *p) { return static_cast<T *>(p); }

// Element `slot` of an array whose base is uniform across the workgroup, through
// a 32-bit byte offset in the global address space. The pointers of an
// instance are read from its table, so without the cast the compiler must use
// flat accesses, which also count against the LDS counter: every LDS wait would
// then wait for the global loads in flight too. Cast, the access is
// base + zext(offset), the scalar-base form of the global loads and stores, and
// one offset register serves every array of the same element type. Instances
// hold at most kMaxCells cells, so the offsets of FP64 arrays stay below 2^31.
template <typename T> __device__ __forceinline__ T ld32(const T *base, uint32_t slot) {
  using GC = const __attribute__((address_space(1))) char;
  using GT = const __attribute__((address_space(1))) T;
  return *(GT *)((GC *)base + slot * (uint32_t)sizeof(T));
}
template <typename T> __device__ __forceinline__ void st32(T *base, uint32_t slot, T v) {
  using GC = __attribute__((address_space(1))) char;
  using GT = __attribute__((address_space(1))) T;
  *(GT *)((GC *)base + slot * (uint32_t)sizeof(T)) = v;
}

// LDS slot of step u of line k in a tile of S steps. A line's steps are stored
// step-mod-4 major, so when the solve threads, each owning 4 consecutive steps,
// read their e-th step, consecutive lanes read consecutive words: no bank
// conflicts (in plain row order the stride of 4 words made every solve access
// a 4- to 8-way conflict). Loads of 16 lines x 4 steps per wave stay
// conflict-free too: the row stride S + 4 shifts each line by 4 banks.
template <int S> __device__ __forceinline__ int tslot(int k, int u) {
  return k * (S + 4) + (u & 3) * (S / 4 + 1) + (u >> 2);
}

// Global-address-space view of an array pointer read from the instance table
// (plain indexing of such a pointer compiles to flat accesses).
template <typename T> __device__ __forceinline__ __attribute__((address_space(1))) T *gbl(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// Workgroup barrier for LDS data only: it waits for this thread's LDS
// accesses but not for its global loads, so tiles prefetched into registers
// stay in flight across it. Valid where threads exchange data through LDS
// only (a thread that reloads a global element is the one that stored it).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// (i, j) of storage slot K of a level with nx columns (hx = half_x(nx))
__device__ __forceinline__ void cell_ij(int K, int nx, int hx, int &i, int &j) {
  j = K / nx;
  i = xcol(K - j * nx, hx);
}
template <typename T> __device__ __forceinline__ const T *cptr(const void *p) {
  return static_cast<const T *>(p);
}

// Read-only view of an instance's scalars, for the kernels that do not write
// them (only the begin and finish kernels do). Through the constant address
// space the compiler reads these workgroup-uniform values with scalar loads
// instead of per-lane loads that every later load waits behind.
typedef const __attribute__((address_space(4))) KScal ConstKScal;
__device__ __forceinline__ ConstKScal *rsc(const InstDev &I) { return (ConstKScal *)I.sc; }

// ---------------------------------------------------------------------------
// workgroup reduction of NT threads: wave64 shuffles, then LDS across the
// NT / 64 waves
// ---------------------------------------------------------------------------
template <typename A, int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(A (&v)[NV]) {
  __shared__ A red[NV][NT / 64];
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
      for (int w = 1; w < NT / 64; ++w) s += red[k][w];
      v[k] = s;
    }
  }
  __syncthreads();
}

// partial sums of a workgroup (NV values) into slot lblk of the instance
template <typename A, int NV>
__device__ __forceinline__ void store_partials(const InstDev &I, int lblk, A (&v)[NV]) {
  block_sum<A, NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) I.partial[(size_t)k * I.nblk0 + lblk] = (double)v[k];
  }
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
    s.iters = 0; s.titers = 0; s.orefs = 0; s.clamps = 0; s.vcyc = 0;
    s.phase = 0; s.nonfin = 0;
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
// source/Solver.cc:299-422) of vertex (vi, vj) from the f of the four cells
// around it, fSW = (vi-1, vj-1), fSE = (vi, vj-1), fNW = (vi-1, vj),
// fNE = (vi, vj) (loaded up front; only cells inside the grid are used here),
// and wxL/wxR = wx0/wx1[vi], wyB/wyT = wy0/wy1[vj]. Dirichlet sides were
// written in the order XMIN, XMAX, YMIN, YMAX there, so the precedence here
// is the reverse.
__device__ __forceinline__ double vertex_value(const InstDev &I, int vi, int vj, double fSW,
                                               double fSE, double fNW, double fNE, double wxL,
                                               double wxR, double wyB, double wyT) {
#pragma clang fp contract(off)
  const int nx = I.nx, ny = I.ny;
  if (vj == ny && (I.bcmask & 8)) return I.bcv[3][vi];
  if (vj == 0 && (I.bcmask & 4)) return I.bcv[2][vi];
  if (vi == nx && (I.bcmask & 2)) return I.bcv[1][vj];
  if (vi == 0 && (I.bcmask & 1)) return I.bcv[0][vj];
  const bool bx = (vi == 0 || vi == nx), by = (vj == 0 || vj == ny);
  // the cell column inside the grid on a vertical boundary (x = 0: the
  // vertex's east cells), the row inside on a horizontal one
  const bool east = vi == 0, north = vj == 0;
  if (bx && by) {   // corners copy the adjacent cell (:316-319)
    return north ? (east ? fNE : fNW) : (east ? fSE : fSW);
  }
  if (bx) {         // vertical boundary edge: linear in y (:355-366)
    return east ? wyB * fSE + wyT * fNE : wyB * fSW + wyT * fNW;
  }
  if (by) {         // horizontal boundary edge: linear in x (:371-381)
    return north ? wxL * fNW + wxR * fNE : wxL * fSW + wxR * fSE;
  }
  return wxL * wyB * fSW + wxR * wyB * fSE + wxL * wyT * fNW + wxR * wyT * fNE;
}

// Edges of one cell (i, j): xe[i], xe[i+1], ye[j], ye[j+1].
struct CellEdges { double x0, x1, y0, y1; };

// a_sigma_func (source/Solver.cc:68-97) with the edge table of
// Mesh::build_connectivity_ (source/Mesh.cc:67-155). The operations are those
// of the reference in the same order; |B - A| is written as |dx| + |dy|, which
// equals sqrt(dx^2 + dy^2) exactly because one of the two is zero.
__device__ __forceinline__ void a_sigma(const CellEdges &c, int inbr, double Lxx, double Lxy,
                                        double Lyy, double fA, double fB, double &aA, double &aB,
                                        double &as) {
#pragma clang fp contract(off)
  double Ax, Ay, Bx, By, n0, n1;
  if (inbr == 0) {        // im, XNEG
    Ax = c.x0; Ay = c.y1; Bx = c.x0; By = c.y0; n0 = -1.0; n1 = 0.0;
  } else if (inbr == 1) { // jp, YPOS
    Ax = c.x1; Ay = c.y1; Bx = c.x0; By = c.y1; n0 = 0.0; n1 = 1.0;
  } else if (inbr == 2) { // ip, XPOS
    Ax = c.x1; Ay = c.y0; Bx = c.x1; By = c.y1; n0 = 1.0; n1 = 0.0;
  } else {                // jm, YNEG
    Ax = c.x0; Ay = c.y0; Bx = c.x1; By = c.y0; n0 = 0.0; n1 = -1.0;
  }
  const double Kx = 0.5 * (c.x0 + c.x1);
  const double Ky = 0.5 * (c.y0 + c.y1);
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
// co: its west face (inbr 0) or south face (inbr 3), with cl the neighbour
// across it. vA/vB are the owner's edge vertex values; the neighbour sees the
// same edge reversed.
__device__ __forceinline__ void face_pair(const CellEdges &co, const CellEdges &cl, int inbr,
                                          const double (&LK)[3], const double (&LL)[3], double fK,
                                          double fL, double vA, double vB, double &AK, double &AL) {
#pragma clang fp contract(off)
  double aKA, aKB, asK, aLA, aLB, asL;
  a_sigma(co, inbr, LK[0], LK[1], LK[2], vA, vB, aKA, aKB, asK);
  a_sigma(cl, inbr == 0 ? 2 : 1, LL[0], LL[1], LL[2], vB, vA, aLA, aLB, asL);
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
__device__ __forceinline__ void dirichlet_face(const CellEdges &c, int inbr, const double (&LK)[3],
                                               double fK, double vA, double vB, double &AK,
                                               double &Bm) {
#pragma clang fp contract(off)
  double aA, aB, as;
  a_sigma(c, inbr, LK[0], LK[1], LK[2], vA, vB, aA, aB, as);
  const double Bs = -as;
  const double Ba = fabs(Bs);
  const double Bp = (Ba + Bs) / 2.0;
  Bm = (Ba - Bs) / 2.0;
  AK = aA + aB + Bp / (fK + kEps);
}

// r = R - M x of the FP64 Krylov start, with the products summed as spmv64
// sums them (centre, west, east, south, north; fused multiply-adds), from
// the band values and x of the cell and its neighbours (couplings outside the
// grid are zero).
__device__ __forceinline__ double start_residual(double R, double ac, double x0, double aw,
                                                 double xw, double ae, double xe, double as,
                                                 double xs, double an, double xn) {
  // explicit fused multiply-adds: with contraction left to the compiler it
  // may fuse either product of ac x0 + aw xw, and the choice changed with
  // unrelated code around the call
  double s = ac * x0;
  s = fma(aw, xw, s);
  s = fma(ae, xe, s);
  s = fma(as, xs, s);
  s = fma(an, xn, s);
  return R - s;
}

// One work-item per cell writes row K of M (d, w, e, s, n) and R. The diagonal
// is summed in the order the reference pushes its triplets (west pairs, south
// pairs, Dirichlet XMIN/XMAX/YMIN/YMAX, time term), which is the order Eigen's
// setFromTriplets adds duplicates; R likewise. Also seeds x = f (warm start)
// and, for the FP64 Krylov forms, starts BiCGSTAB: r = rh = R - M x, p = v = 0
// and the partial sums (r, r), (R, R).
// Every array value the cell needs is loaded first, with neighbour indices
// clamped into the grid (values of neighbours outside it are loaded but never
// used), so that a work-item waits for memory once rather than behind each
// boundary test; only the Dirichlet vertex arrays, read by boundary cells, are
// loaded where they are used.
__global__ void __launch_bounds__(kBlock) k_assemble(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map) {
#pragma clang fp contract(off)
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->ostatus == ST_IDLE) return;
  const int nx = I.nx, ny = I.ny;
  const int K = m.k0 + threadIdx.x;
  const bool krylov64 = I.prec != SG_MIXED;   // uniform over the workgroup
  double part[2] = {0.0, 0.0};
  if (K < nx * ny) {
    const int hx = half_x(nx);
    int i, j;
    cell_ij(K, nx, hx, i, j);
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    const size_t r0 = (size_t)j * nx, rm = hs ? r0 - nx : r0, rp = hn ? r0 + nx : r0;
    const int sm = xslot(hw ? i - 1 : i, hx), s0 = xslot(i, hx), sp = xslot(he ? i + 1 : i, hx);
    const auto f = gbl(I.f);
    const double f3[3][3] = {{f[rm + sm], f[rm + s0], f[rm + sp]},
                             {f[r0 + sm], f[K], f[r0 + sp]},
                             {f[rp + sm], f[rp + s0], f[rp + sp]}};
    // Lambda = D * G per cell, entry by entry as update_Lambda forms it
    // (source/Solver.cc:57-65)
    const auto Dxx = gbl(I.Dxx), Dxy = gbl(I.Dxy), Dyy = gbl(I.Dyy), Gv = gbl(I.G);
    const size_t kw = r0 + sm, ke = r0 + sp, ks = rm + s0, kn = rp + s0;
    auto lam = [&](size_t q, double (&L)[3]) {
      const double g = Gv[q];
      L[0] = Dxx[q] * g; L[1] = Dxy[q] * g; L[2] = Dyy[q] * g;
    };
    double LK[3], LW[3], LE[3], LS[3], LN[3];
    lam(K, LK); lam(kw, LW); lam(ke, LE); lam(ks, LS); lam(kn, LN);
    const auto xe = gbl(I.xe), ye = gbl(I.ye);
    const double xm = xe[hw ? i - 1 : i], x0 = xe[i], x1 = xe[i + 1], x2 = xe[he ? i + 2 : i + 1];
    const double ym = ye[hs ? j - 1 : j], y0 = ye[j], y1 = ye[j + 1], y2 = ye[hn ? j + 2 : j + 1];
    const auto wx0 = gbl(I.wx0), wx1 = gbl(I.wx1), wy0 = gbl(I.wy0), wy1 = gbl(I.wy1);
    const double wxLi = wx0[i], wxRi = wx1[i], wxLe = wx0[i + 1], wxRe = wx1[i + 1];
    const double wyBj = wy0[j], wyTj = wy1[j], wyBn = wy0[j + 1], wyTn = wy1[j + 1];
    const double GK = Gv[K], itK = gbl(I.itau)[K];
    const CellEdges cK{x0, x1, y0, y1}, cW{xm, x0, y0, y1}, cE{x1, x2, y0, y1};
    const CellEdges cS{x0, x1, ym, y0}, cN{x0, x1, y1, y2};
    const double vSW = vertex_value(I, i, j, f3[0][0], f3[0][1], f3[1][0], f3[1][1], wxLi, wxRi,
                                    wyBj, wyTj);
    const double vSE = vertex_value(I, i + 1, j, f3[0][1], f3[0][2], f3[1][1], f3[1][2], wxLe, wxRe,
                                    wyBj, wyTj);
    const double vNW = vertex_value(I, i, j + 1, f3[1][0], f3[1][1], f3[2][0], f3[2][1], wxLi, wxRi,
                                    wyBn, wyTn);
    const double vNE = vertex_value(I, i + 1, j + 1, f3[1][1], f3[1][2], f3[2][1], f3[2][2], wxLe,
                                    wxRe, wyBn, wyTn);
    const double fK = f3[1][1];
    double d = 0.0, R = 0.0, bw = 0.0, be = 0.0, bs = 0.0, bn = 0.0, AK, AL;
    if (hw) {   // own west face (west-pair loop, :173-175)
      face_pair(cK, cW, 0, LK, LW, fK, f3[1][0], vNW, vSW, AK, AL);
      d += AK;
      bw = 0.0 + -AL;
    }
    if (he) {   // east face, owned by (i+1, j): this cell is its L
      face_pair(cE, cK, 0, LE, LK, f3[1][2], fK, vNE, vSE, AK, AL);
      d += AL;
      be = 0.0 + -AK;
    }
    if (hs) {   // own south face (south-pair loop, :178-182)
      face_pair(cK, cS, 3, LK, LS, fK, f3[0][1], vSW, vSE, AK, AL);
      d += AK;
      bs = 0.0 + -AL;
    }
    if (hn) {   // north face, owned by (i, j+1)
      face_pair(cN, cK, 3, LN, LK, f3[2][1], fK, vNW, vNE, AK, AL);
      d += AL;
      bn = 0.0 + -AK;
    }
    double Bm;
    if (i == 0 && (I.bcmask & 1)) {        // XMIN: face im
      dirichlet_face(cK, 0, LK, fK, vNW, vSW, AK, Bm);
      d += AK; R += Bm;
    }
    if (i == nx - 1 && (I.bcmask & 2)) {   // XMAX: face ip
      dirichlet_face(cK, 2, LK, fK, vSE, vNE, AK, Bm);
      d += AK; R += Bm;
    }
    if (j == 0 && (I.bcmask & 4)) {        // YMIN: face jm
      dirichlet_face(cK, 3, LK, fK, vSW, vSE, AK, Bm);
      d += AK; R += Bm;
    }
    if (j == ny - 1 && (I.bcmask & 8)) {   // YMAX: face jp
      dirichlet_face(cK, 1, LK, fK, vNE, vNW, AK, Bm);
      d += AK; R += Bm;
    }
    // time / loss diagonal (:190-199), cell_area_dt (source/Mesh.h:61-63)
    const double dx = x1 - x0, dy = y1 - y0;
    const double UKK = GK * (dx * dy / I.dt);
    d += UKK * (1 + I.dt * itK);
    R += UKK * fK;
    I.band64[BC][K] = d;
    I.band64[BW][K] = bw;
    I.band64[BE][K] = be;
    I.band64[BS][K] = bs;
    I.band64[BN][K] = bn;
    if (I.prec != SG_FP64) {   // FP32 copies for the FP32 V-cycle and Krylov forms
      const Lev &L0 = I.lev[0];
      ptr<float>(L0.band[BC])[K] = (float)d;
      ptr<float>(L0.band[BW])[K] = (float)bw;
      ptr<float>(L0.band[BE])[K] = (float)be;
      ptr<float>(L0.band[BS])[K] = (float)bs;
      ptr<float>(L0.band[BN])[K] = (float)bn;
    }
    // the solve starts at f + theta df, df the increment of the last
    // accepted step and theta the ratio of the last two increments (an
    // extrapolation in time), or at f
    double x0K = fK, x0W = f3[1][0], x0E = f3[1][2], x0S = f3[0][1], x0N = f3[2][1];
    const double th = rsc(I)->theta;   // uniform over the workgroup
    if (I.warm && th > 0.0) {
      const auto df = gbl(I.df);
      x0K += th * (double)df[K];
      x0W += th * (double)df[kw];
      x0E += th * (double)df[ke];
      x0S += th * (double)df[ks];
      x0N += th * (double)df[kn];
    }
    I.R[K] = R;
    I.x[K] = x0K;
    if (krylov64) {
      const double r = start_residual(R, d, x0K, bw, x0W, be, x0E, bs, x0S, bn, x0N);
      gbl(ptr<double>(I.kv[VR]))[K] = r;
      if (I.prec != SG_PC32) {   // PC32: the first k_bicg_p sets rh and p, from r alone
        gbl(ptr<double>(I.kv[VRH]))[K] = r;
        gbl(ptr<double>(I.kv[VP]))[K] = 0.0;
        gbl(ptr<double>(I.kv[VV]))[K] = 0.0;
      }
      part[0] = r * r;
      part[1] = R * R;
    }
  }
  if (krylov64) store_partials<double, 2>(I, m.k0 / kBlock, part);
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
  if (rsc(I)->ostatus == ST_IDLE) return;
  const Lev &L = I.lev[l];
  const int nx = L.nx, ny = L.ny;
  const int K = m.k0 + threadIdx.x;
  if (K >= nx * ny) return;
  const int hx = half_x(nx);
  int i, j;
  cell_ij(K, nx, hx, i, j);
  double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
  const bool io = i & 1, jo = j & 1;
  if (!io && !jo) {
    w0 = 1.0;
  } else if (io && !jo) {
    w_xedge<NB>(L, K, i, w0, w1);
  } else if (!io && jo) {
    w_yedge<NB>(L, K, j, w0, w2);
  } else {
    // The east and north neighbours' edge weights are computed at clamped
    // positions and zeroed when those neighbours lie outside the grid, so
    // that all five stencils are loaded together.
    const St9 a = load_st<NB>(L, K);
    const bool he = i + 1 < nx, hn = j + 1 < ny;
    double sW, nW, sE, nE, wS, eS, wN, eN;
    w_yedge<NB>(L, (size_t)j * nx + xslot(i - 1, hx), j, sW, nW);
    w_yedge<NB>(L, (size_t)j * nx + xslot(he ? i + 1 : i - 1, hx), j, sE, nE);
    w_xedge<NB>(L, K - nx, i, wS, eS);
    w_xedge<NB>(L, hn ? K + nx : K - nx, i, wN, eN);
    if (!he) { sE = 0; nE = 0; }
    if (!hn) { wN = 0; eN = 0; }
    const double c = a.c;
    w0 = -(a.w * sW + a.s * wS + a.sw) / c;
    w1 = -(a.e * sE + a.s * eS + a.se) / c;
    w2 = -(a.w * nW + a.n * wN + a.nw) / c;
    w3 = -(a.e * nE + a.n * eN + a.ne) / c;
  }
  // Only the weights that are neither structurally zero nor exactly one are
  // stored: an even-even point's single weight is 1 and its others are 0, and
  // the parent at I0 + 1 of an even i (at J0 + 1 of an even j) has weight 0.
  // Every reader takes those values as constants (rap_row, restrict_point,
  // prolong_point, k_prolong), so the unstored entries are never read.
  if (io || jo) ptr<T>(L.pw[0])[K] = (T)w0;
  if (io) ptr<T>(L.pw[1])[K] = (T)w1;
  if (jo) ptr<T>(L.pw[2])[K] = (T)w2;
  if (io && jo) ptr<T>(L.pw[3])[K] = (T)w3;
}

// A_{l+1}(C, C') = sum_f sum_g P(f, C) A_l(f, g) P(g, C'); one work-item per
// coarse point C = (I, J), whose support is the 3x3 fine block around (2I, 2J).
// All offsets are compile-time constants (fine offset fa, fb; stencil offset;
// parent slot q), so the 3x3 accumulator lives in registers.
// The terms of fine row 2 J + fb of the support of coarse point (Ic, J):
// acc[aj][ai] += P(f, C) A(f, g) P(g, C') for the fine points f of that row
// inside the grid, their stencil neighbours g inside it and the parents C' of
// g, in that order.
template <typename T, int NB>
__device__ __forceinline__ void rap_row(const Lev &F, const T *const (&pw)[4], int Ic, int J,
                                        int fb, int hxf, double (&acc)[3][3]) {
#pragma clang fp contract(off)
  constexpr int di_[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1};
  constexpr int dj_[9] = {0, 0, 0, -1, 1, -1, -1, 1, 1};
  const int fx = F.nx, fy = F.ny;
  const int fj = 2 * J + fb;
  if (fj < 0 || fj >= fy) return;
#pragma unroll
  for (int fa = -1; fa <= 1; ++fa) {
    const int fi = 2 * Ic + fa;
    if (fi < 0 || fi >= fx) continue;
    const size_t kf = (size_t)fj * fx + xslot(fi, hxf);
    // slot of C among the parents of f: (I - (fi >> 1)) + 2 (J - (fj >> 1));
    // an even-even f (fa = fb = 0, slot 0) is injected with weight exactly 1
    const int slot = (fa < 0 ? 1 : 0) + 2 * (fb < 0 ? 1 : 0);
    const double wf = (fa == 0 && fb == 0) ? 1.0 : (double)gbl(pw[slot])[kf];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ga = fa + di_[b], gb = fb + dj_[b];
      const int gi = 2 * Ic + ga, gj = 2 * J + gb;
      if (gi < 0 || gi >= fx || gj < 0 || gj >= fy) continue;
      const double wa = wf * gbl(F.band64[b])[kf];
      const size_t kg = (size_t)gj * fx + xslot(gi, hxf);
      const int oi = (ga >= 0 ? ga / 2 : -((1 - ga) / 2)) + 1;   // floor(ga / 2) + 1
      const int oj = (gb >= 0 ? gb / 2 : -((1 - gb) / 2)) + 1;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ai = oi + (q & 1), aj = oj + (q >> 1);
        if (ai > 2 || aj > 2) continue;   // such parents carry zero weight
        // so do the parents at I0 + 1 of an even gi and at J0 + 1 of an even
        // gj (gi, gj have the parities of ga, gb): k_pweights writes exact
        // zeros there, and the sum skips them
        if (((q & 1) && !(ga & 1)) || ((q & 2) && !(gb & 1))) continue;
        // an even-even g has the single weight 1 (q = 0 here)
        const bool unit = !(ga & 1) && !(gb & 1);
        acc[aj][ai] += wa * (unit ? 1.0 : (double)gbl(pw[q])[kg]);
      }
    }
  }
}

template <typename T, int NB, bool COPY>
__global__ void __launch_bounds__(kBlock) k_rap(const InstDev *__restrict__ insts,
                                               const BlkMap *__restrict__ map, int l) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->ostatus == ST_IDLE) return;
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int cx = Cl.nx, cy = Cl.ny, fx = F.nx;
  const int KN = m.k0 + threadIdx.x;
  if (KN >= cx * cy) return;
  const int hxf = half_x(fx);
  // Work-items take the coarse points in natural order (Ic fastest), so that
  // the fine points 2 Ic + fa of neighbouring work-items are neighbouring
  // slots of one half-row of the split layout (coalesced); the output goes to
  // the point's slot.
  const int J = KN / cx, Ic = KN - J * cx;
  const size_t KC = (size_t)J * cx + xslot(Ic, half_x(cx));
  const T *const pw[4] = {cptr<T>(F.pw[0]), cptr<T>(F.pw[1]), cptr<T>(F.pw[2]),
                          cptr<T>(F.pw[3])};
  double acc[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
  for (int fb = -1; fb <= 1; ++fb) rap_row<T, NB>(F, pw, Ic, J, fb, hxf, acc);
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
// level-0 vector by id: a Krylov slot (Vec) or an FP32 V-cycle vector (PcVec)
__device__ __forceinline__ void *vec(const InstDev &I, int id) {
  return id < kNumVec ? I.kv[id] : (void *)I.pv[id - kNumVec];
}

template <typename T>
__device__ __forceinline__ void level_xb(const InstDev &I, int l, int xid, int bid, T *&x,
                                         const T *&b) {
  if (l == 0) { x = ptr<T>(vec(I, xid)); b = cptr<T>(vec(I, bid)); }
  else { x = ptr<T>(I.lev[l].x); b = cptr<T>(I.lev[l].b); }
}

// LU factors of every line matrix of a level (rows: DIR 0, columns: DIR 1),
// once per step: lup_t = upper_t / pivot_t, linv_t = 1 / pivot_t with
// pivot_t = diag_t - lower_t lup_{t-1} (Thomas elimination without pivoting;
// the level-0 line matrices are column diagonally dominant M-matrices).
// Same tiles and layout as the sweeps below. The recurrence
// lup_t = c_t / (b_t - a_t lup_{t-1}) is a Moebius map, i.e. the 2x2 matrix
// [[0, c_t], [-a_t, b_t]] acting on (P, Q) with lup = P / Q, so the threads of
// a line compose their 4 steps' matrices (rescaled to stay in range), scan
// them across the line and carry (P, Q) between tiles; each thread then runs
// the exact recurrence over its 4 steps from the value the scan gives it.
template <typename T, int DIR, int SH>
__global__ void __launch_bounds__(tile_threads(SH)) k_line_factor(const InstDev *__restrict__ insts,
                                                                  const LineGrp *__restrict__ map,
                                                                  int l) {
  constexpr int S = tile_steps(SH), L = tile_lines(SH), NT = tile_threads(SH);
  constexpr int W = NT / L;          // threads per line in the solve, 4 steps each
  constexpr int Q = S * L / NT;      // elements per thread in the loads
  static_assert(W * 4 == S && W <= 64, "a line's threads cover its tile in one wave");
  __shared__ double s_lo[L * (S + 4)];
  __shared__ double s_d[L * (S + 4)];
  __shared__ double s_up[L * (S + 4)];
  __shared__ double s_c[L];
  const LineGrp gr = map[blockIdx.x];
  const InstDev &I = insts[gr.inst];
  if (rsc(I)->ostatus == ST_IDLE) return;
  const Lev &Lv = I.lev[l];
  const int nx = Lv.nx, ny = Lv.ny;
  const double *__restrict__ bc = Lv.band64[BC];
  const double *__restrict__ lo = Lv.band64[DIR == 0 ? BW : BS];
  const double *__restrict__ up = Lv.band64[DIR == 0 ? BE : BN];
  T *__restrict__ inv = ptr<T>(Lv.linv[DIR]);
  T *__restrict__ cu = ptr<T>(Lv.lup[DIR]);
  const int n = DIR == 0 ? nx : ny;
  const int hx = half_x(nx);
  // storage slot of step t of line li (rows: DIR 0, columns: DIR 1)
  auto at = [&](int li, int t) -> uint32_t {
    return DIR == 0 ? (uint32_t)(li * nx + xslot(t, hx)) : (uint32_t)(t * nx + xslot(li, hx));
  };
  const int cnt = gr.count;
  const int tid = threadIdx.x;
  // (line, step) of this thread's q-th element (shifts and masks: recomputed
  // where used rather than held in registers)
  auto ek = [&](int q) { const int e = q * NT + tid; return DIR == 0 ? e / S : e % L; };
  auto eu = [&](int q) {
    const int e = q * NT + tid, v = e % S;
    return DIR == 0 ? (v < S / 2 ? 2 * v : 2 * (v - S / 2) + 1) : e / L;
  };
  const int sk = tid / W, sc = tid % W;
  const bool sk_ok = sk < cnt;
  double vl[Q], vd[Q], vu[Q];
  // every load of a tile is issued at once, into registers
  auto load = [&](int t0) {
    const int tw = min(S, n - t0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = ek(q), u = eu(q);
      vl[q] = 0.0; vd[q] = 1.0; vu[q] = 0.0;
      if (k < cnt && u < tw) {
        const int t = t0 + u;
        const uint32_t K = at(gr.first + k * gr.step, t);
        vl[q] = t > 0 ? ld32(lo, K) : 0.0;
        vd[q] = ld32(bc, K);
        vu[q] = t < n - 1 ? ld32(up, K) : 0.0;
      }
    }
  };
  if (tid < L) s_c[tid] = 0.0;
  load(0);
  for (int t0 = 0; t0 < n; t0 += S) {
    const int tw = min(S, n - t0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      s_lo[tslot<S>(ek(q), eu(q))] = vl[q];
      s_d[tslot<S>(ek(q), eu(q))] = vd[q];
      s_up[tslot<S>(ek(q), eu(q))] = vu[q];
    }
    // The next tile's loads fly while this one is factored and stored: they
    // read the bands, which this kernel does not write.
    if (t0 + S < n) load(t0 + S);
    lds_barrier();
    // compose this thread's 4 maps: (P, Q) <- [[0, c], [-a, b]] (P, Q)
    double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int u = 4 * sc + e;
      if (sk_ok && u < tw) {
        const double a = s_lo[tslot<S>(sk, u)], bb = s_d[tslot<S>(sk, u)], c = s_up[tslot<S>(sk, u)];
        const double n00 = c * m10, n01 = c * m11;
        const double n10 = -a * m00 + bb * m10, n11 = -a * m01 + bb * m11;
        const double sc_ = 1.0 / fmax(fmax(fabs(n00), fabs(n01)), fmax(fabs(n10), fabs(n11)));
        m00 = n00 * sc_; m01 = n01 * sc_; m10 = n10 * sc_; m11 = n11 * sc_;
      }
    }
#pragma unroll
    for (int off = 1; off < W; off <<= 1) {
      const double p00 = __shfl_up(m00, off, W), p01 = __shfl_up(m01, off, W);
      const double p10 = __shfl_up(m10, off, W), p11 = __shfl_up(m11, off, W);
      if (sc >= off) {   // own * previous
        const double n00 = m00 * p00 + m01 * p10, n01 = m00 * p01 + m01 * p11;
        const double n10 = m10 * p00 + m11 * p10, n11 = m10 * p01 + m11 * p11;
        const double sc_ = 1.0 / fmax(fmax(fabs(n00), fabs(n01)), fmax(fabs(n10), fabs(n11)));
        m00 = n00 * sc_; m01 = n01 * sc_; m10 = n10 * sc_; m11 = n11 * sc_;
      }
    }
    double e00 = __shfl_up(m00, 1, W), e01 = __shfl_up(m01, 1, W);
    double e10 = __shfl_up(m10, 1, W), e11 = __shfl_up(m11, 1, W);
    if (sc == 0) { e00 = 1.0; e01 = 0.0; e10 = 0.0; e11 = 1.0; }
    // lup before this thread's first step: the exclusive map applied to the
    // carry (P, Q) = (lup_carry, 1)
    const double c0 = sk_ok ? s_c[sk] : 0.0;
    const double P = e00 * c0 + e01, Qv = e10 * c0 + e11;
    double cprev = Qv != 0.0 ? P / Qv : 0.0;
    double ivs[4], cus[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int u = 4 * sc + e;
      ivs[e] = 0.0; cus[e] = 0.0;
      if (sk_ok && u < tw) {
        const double iv = 1.0 / (s_d[tslot<S>(sk, u)] - s_lo[tslot<S>(sk, u)] * cprev);
        cprev = s_up[tslot<S>(sk, u)] * iv;
        ivs[e] = iv;
        cus[e] = cprev;
      }
    }
    lds_barrier();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int u = 4 * sc + e;
      if (sk_ok && u < tw) { s_d[tslot<S>(sk, u)] = ivs[e]; s_up[tslot<S>(sk, u)] = cus[e]; }
    }
    if (sk_ok && sc == W - 1) s_c[sk] = cprev;
    lds_barrier();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = ek(q), u = eu(q);
      if (k < cnt && u < tw) {
        const uint32_t K = at(gr.first + k * gr.step, t0 + u);
        st32(inv, K, (T)s_d[tslot<S>(k, u)]);
        st32(cu, K, (T)s_up[tslot<S>(k, u)]);
      }
    }
    lds_barrier();
  }
}

// Zebra line Gauss-Seidel. A workgroup solves a group of lines of one colour
// exactly with the precomputed factors, the off-line couplings moved to the
// right-hand side. The lines are walked in tiles of L lines x S steps
// (L*S = 1024; S = 16, 64 or 256 by line length) staged through LDS: the 256
// threads load a tile with coalesced global accesses (along the lines for
// rows, across the lines for columns), then solve it with the threads laid out
// along the lines: both Thomas recurrences are first-order affine maps
// (d_t = a_t d_{t-1} + b_t forward, x_t = d_t - u_t x_{t+1} backward), so
// every thread composes the maps of its 4 consecutive steps, a segmented wave
// scan combines them across the S/4 threads of a line, and the carry from the
// previous tile enters through the composed map. The eliminated right-hand
// side d of the last ring_tiles tiles of each line stays in LDS, in a ring of
// tile buffers (tile t in buffer t mod KT); when a buffer is reused, the d it
// holds is parked in x (no other line of this colour reads it). The backward
// pass turns d into the solution, reading back only the parked tiles. ZERO:
// the other lines are zero (first sweep of a cycle).
//
// LDS of a line sweep of shape SH: the ring of KT tile buffers, the tile's
// lower band (then upper factor) and inverse pivots, and the carry per line.
template <typename T, int NB, int SH> constexpr int line_lds_elems() {
  return (ring_tiles(SH, sizeof(T), NB) + 2) * tile_lines(SH) * (tile_steps(SH) + 4) +
         tile_lines(SH);
}

// The sweep of one line group gr by one workgroup of tile_threads(SH)
// threads, in the LDS lds (line_lds_elems<T, NB, SH>() values); k_line and
// k_line_shapes below are its kernels.
template <typename T, int NB, int DIR, bool ZERO, int SH>
__device__ __forceinline__ void line_sweep(const InstDev *__restrict__ insts, const LineGrp gr,
                                           int l, int xid, int bid, T *lds) {
  constexpr int S = tile_steps(SH), L = tile_lines(SH), NT = tile_threads(SH);
  constexpr int W = NT / L;          // threads per line in the solve, 4 steps each
  constexpr int Q = S * L / NT;      // elements per thread in the loads
  constexpr int KT = ring_tiles(SH, sizeof(T), NB);   // tiles of d kept in LDS
  constexpr int TB = L * (S + 4);                  // one tile buffer
  static_assert(W * 4 == S && W <= 64, "a line's threads cover its tile in one wave");
  T *const s_rb = lds;             // r, then d, then x of a tile, in a ring of KT
  T *const s_a = s_rb + KT * TB;
  T *const s_i = s_a + TB;
  T *const s_c = s_i + TB;         // L values
  const InstDev &I = insts[gr.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const Lev &Lv = I.lev[l];
  const int nx = Lv.nx, ny = Lv.ny;
  T *x; const T *b;
  level_xb<T>(I, l, xid, bid, x, b);
  const T *__restrict__ lo = cptr<T>(Lv.band[DIR == 0 ? BW : BS]);
  const T *__restrict__ om = cptr<T>(Lv.band[DIR == 0 ? BS : BW]);
  const T *__restrict__ op = cptr<T>(Lv.band[DIR == 0 ? BN : BE]);
  const T *omm = nullptr, *omp = nullptr, *opm = nullptr, *opp = nullptr;
  if (NB == 9) {   // diagonal couplings: (minus side, along -1/+1), (plus side, along -1/+1)
    omm = cptr<T>(Lv.band[BSW]);
    omp = cptr<T>(Lv.band[DIR == 0 ? BSE : BNW]);
    opm = cptr<T>(Lv.band[DIR == 0 ? BNW : BSE]);
    opp = cptr<T>(Lv.band[BNE]);
  }
  const T *__restrict__ inv = cptr<T>(Lv.linv[DIR]);
  const T *__restrict__ cu = cptr<T>(Lv.lup[DIR]);
  const int n = DIR == 0 ? nx : ny;            // line length
  const int nperp = DIR == 0 ? ny : nx;
  const int hx = half_x(nx);
  // storage slot of step t of line li (rows: DIR 0, columns: DIR 1)
  auto at = [&](int li, int t) -> uint32_t {
    return DIR == 0 ? (uint32_t)(li * nx + xslot(t, hx)) : (uint32_t)(t * nx + xslot(li, hx));
  };
  const int cnt = gr.count;
  const int tid = threadIdx.x;
  // load/store layout: element e -> (line k, step u), coalesced in memory
  // (line, step) of this thread's q-th element (shifts and masks: recomputed
  // where used rather than held in registers)
  auto ek = [&](int q) { const int e = q * NT + tid; return DIR == 0 ? e / S : e % L; };
  auto eu = [&](int q) {
    const int e = q * NT + tid, v = e % S;
    return DIR == 0 ? (v < S / 2 ? 2 * v : 2 * (v - S / 2) + 1) : e / L;
  };
  // solve layout: thread -> (line sk, steps 4*sc .. 4*sc+3)
  const int sk = tid / W, sc = tid % W;
  const bool sk_ok = sk < cnt;
  T vb[Q], vlo[Q], viv[Q], vom[Q], vop[Q], xm[Q], xp[Q];
  T dmm[Q], dmp[Q], dpm[Q], dpp[Q], xmm[Q], xmp[Q], xpm[Q], xpp[Q];
  // every load of a tile is issued at once, into registers
  auto load_fwd = [&](int t0) {
    const int tw = min(S, n - t0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = ek(q), u = eu(q);
      vb[q] = vlo[q] = viv[q] = vom[q] = vop[q] = xm[q] = xp[q] = T(0);
      dmm[q] = dmp[q] = dpm[q] = dpp[q] = xmm[q] = xmp[q] = xpm[q] = xpp[q] = T(0);
      if (k < cnt && u < tw) {
        const int li = gr.first + k * gr.step, t = t0 + u;
        const uint32_t K = at(li, t);
        vb[q] = ld32(b, K);
        vlo[q] = t > 0 ? ld32(lo, K) : T(0);
        viv[q] = ld32(inv, K);
        if (!ZERO) {
          if (li > 0) {
            vom[q] = ld32(om, K); xm[q] = ld32(x, at(li - 1, t));
            if (NB == 9) {
              if (t > 0) { dmm[q] = ld32(omm, K); xmm[q] = ld32(x, at(li - 1, t - 1)); }
              if (t < n - 1) { dmp[q] = ld32(omp, K); xmp[q] = ld32(x, at(li - 1, t + 1)); }
            }
          }
          if (li < nperp - 1) {
            vop[q] = ld32(op, K); xp[q] = ld32(x, at(li + 1, t));
            if (NB == 9) {
              if (t > 0) { dpm[q] = ld32(opm, K); xpm[q] = ld32(x, at(li + 1, t - 1)); }
              if (t < n - 1) { dpp[q] = ld32(opp, K); xpp[q] = ld32(x, at(li + 1, t + 1)); }
            }
          }
        }
      }
    }
  };
  if (tid < L) s_c[tid] = T(0);
  load_fwd(0);
  for (int t0 = 0, tt = 0; t0 < n; t0 += S, ++tt) {
    const int tw = min(S, n - t0);
    T *const s_r = s_rb + (tt % KT) * TB;
    if (tt >= KT) {
      // this buffer holds d of tile tt - KT (a full tile): park it in x. Each
      // element is read here, and rewritten below, by the thread that loads
      // it back in the backward pass.
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int k = ek(q), u = eu(q);
        if (k < cnt) st32(x, at(gr.first + k * gr.step, t0 - KT * S + u), s_r[tslot<S>(k, u)]);
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      T r = vb[q] - vom[q] * xm[q] - vop[q] * xp[q];
      if (NB == 9) r = r - dmm[q] * xmm[q] - dmp[q] * xmp[q] - dpm[q] * xpm[q] - dpp[q] * xpp[q];
      s_r[tslot<S>(ek(q), eu(q))] = r;
      s_a[tslot<S>(ek(q), eu(q))] = vlo[q];
      s_i[tslot<S>(ek(q), eu(q))] = viv[q];
    }
    // The next tile's loads fly while this one is solved and stored. They
    // read bands, b and the other colour's lines, which no thread of this
    // launch writes.
    if (t0 + S < n) load_fwd(t0 + S);
    lds_barrier();
    {   // forward: d_t = a_t d_{t-1} + b_t with a_t = -lo_t inv_t, b_t = r_t inv_t
      T al[4], be[4];
      T A = T(1), B = T(0);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int u = 4 * sc + e;
        al[e] = T(1); be[e] = T(0);
        if (sk_ok && u < tw) {
          al[e] = -s_a[tslot<S>(sk, u)] * s_i[tslot<S>(sk, u)];
          be[e] = s_r[tslot<S>(sk, u)] * s_i[tslot<S>(sk, u)];
        }
        B = al[e] * B + be[e];
        A = al[e] * A;
      }
#pragma unroll
      for (int off = 1; off < W; off <<= 1) {
        const T Ap = __shfl_up(A, off, W), Bp = __shfl_up(B, off, W);
        if (sc >= off) { B = A * Bp + B; A = A * Ap; }
      }
      T Ae = __shfl_up(A, 1, W), Be = __shfl_up(B, 1, W);
      if (sc == 0) { Ae = T(1); Be = T(0); }
      T d = Ae * (sk_ok ? s_c[sk] : T(0)) + Be;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int u = 4 * sc + e;
        d = al[e] * d + be[e];
        if (sk_ok && u < tw) s_r[tslot<S>(sk, u)] = d;
      }
      lds_barrier();
      if (sk_ok && sc == W - 1) s_c[sk] = d;
    }
    lds_barrier();
  }
  T vd[Q], vc[Q];
  auto load_bwd = [&](int t0, bool with_d) {
    const int tw = min(S, n - t0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = ek(q), u = eu(q);
      vd[q] = vc[q] = T(0);
      if (k < cnt && u < tw) {
        const uint32_t K = at(gr.first + k * gr.step, t0 + u);
        if (with_d) vd[q] = ld32(x, K);
        vc[q] = ld32(cu, K);
      }
    }
  };
  if (tid < L) s_c[tid] = T(0);
  const int ntiles = (n + S - 1) / S;
  // tiles from `resident` on are still in their buffers; the d of every
  // earlier tile was parked above by the thread that loads it here
  const int resident = ntiles > KT ? ntiles - KT : 0;
  load_bwd((ntiles - 1) * S, false);
  for (int tt = ntiles - 1; tt >= 0; --tt) {
    const int t0 = tt * S;
    const int tw = min(S, n - t0);
    T *const s_r = s_rb + (tt % KT) * TB;   // free once its resident tile was done
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (tt < resident) s_r[tslot<S>(ek(q), eu(q))] = vd[q];
      s_a[tslot<S>(ek(q), eu(q))] = vc[q];
    }
    // prefetch the tile below (this iteration stores tile tt only)
    if (tt > 0) load_bwd(t0 - S, tt - 1 < resident);
    lds_barrier();
    {   // backward: x_t = -u_t x_{t+1} + d_t, composed from the line's end
      T al[4], be[4];
      T A = T(1), B = T(0);
#pragma unroll
      for (int e = 3; e >= 0; --e) {
        const int u = 4 * sc + e;
        al[e] = T(1); be[e] = T(0);
        if (sk_ok && u < tw) {
          al[e] = -s_a[tslot<S>(sk, u)];
          be[e] = s_r[tslot<S>(sk, u)];
        }
        B = al[e] * B + be[e];
        A = al[e] * A;
      }
#pragma unroll
      for (int off = 1; off < W; off <<= 1) {
        const T An = __shfl_down(A, off, W), Bn = __shfl_down(B, off, W);
        if (sc + off < W) { B = A * Bn + B; A = A * An; }
      }
      T Ae = __shfl_down(A, 1, W), Be = __shfl_down(B, 1, W);
      if (sc == W - 1) { Ae = T(1); Be = T(0); }
      T xv = Ae * (sk_ok ? s_c[sk] : T(0)) + Be;
#pragma unroll
      for (int e = 3; e >= 0; --e) {
        const int u = 4 * sc + e;
        xv = al[e] * xv + be[e];
        if (sk_ok && u < tw) s_r[tslot<S>(sk, u)] = xv;
      }
      lds_barrier();
      if (sk_ok && sc == 0) s_c[sk] = xv;
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = ek(q), u = eu(q);
      if (k < cnt && u < tw)
        st32(x, at(gr.first + k * gr.step, t0 + u), s_r[tslot<S>(k, u)]);
    }
    lds_barrier();
  }
}

template <typename T, int NB, int DIR, bool ZERO, int SH>
__global__ void __launch_bounds__(tile_threads(SH)) k_line(const InstDev *__restrict__ insts,
                                                           const LineGrp *__restrict__ map, int l,
                                                           int xid, int bid) {
  __shared__ T lds[line_lds_elems<T, NB, SH>()];
  line_sweep<T, NB, DIR, ZERO, SH>(insts, map[blockIdx.x], l, xid, bid, lds);
}

// The tile shapes of one colour's sweep in one launch (256-thread shapes):
// map holds the groups of shape SA (the first na entries), then of SB (the
// next nb), then of SC (the rest; SC < 0: none). Each workgroup runs the
// sweep of its shape in the largest of their LDS layouts. One launch instead
// of one per shape: the shapes' workgroups overlap instead of each shape
// filling and draining the GPU in turn. The FP32 row kernel is held to 4
// waves per SIMD (128 VGPRs, no spills; without the bound its three shape
// bodies take 140 and it runs 3, as the shapes alone do not).
template <typename T, int NB, int DIR, bool ZERO, int SA, int SB, int SC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIR == 0 && sizeof(T) == 4 ? 4 : 1)))
k_line_shapes(const InstDev *__restrict__ insts,
                                                     const LineGrp *__restrict__ map, int na,
                                                     int nb, int l, int xid, int bid) {
  static_assert(tile_threads(SA) == 256 && tile_threads(SB) == 256 &&
                    (SC < 0 || tile_threads(SC) == 256),
                "one workgroup size");
  constexpr int SC_ = SC < 0 ? SB : SC;
  constexpr int EA = line_lds_elems<T, NB, SA>(), EB = line_lds_elems<T, NB, SB>();
  constexpr int EC = SC < 0 ? 0 : line_lds_elems<T, NB, SC_>();
  constexpr int E = EA > EB ? (EA > EC ? EA : EC) : (EB > EC ? EB : EC);
  __shared__ T lds[E];
  const int w = blockIdx.x;
  const LineGrp gr = map[w];
  if (w < na) line_sweep<T, NB, DIR, ZERO, SA>(insts, gr, l, xid, bid, lds);
  else if (SC < 0 || w < na + nb) line_sweep<T, NB, DIR, ZERO, SB>(insts, gr, l, xid, bid, lds);
  else line_sweep<T, NB, DIR, ZERO, SC_>(insts, gr, l, xid, bid, lds);
}

// (A x)(i, j) at slot K. Every operand is loaded before the sum, so that a
// work-item waits for memory once (with the loads behind the boundary tests
// the compiler issues one round trip per neighbour): neighbour slots are
// clamped into the grid and the coupling of a neighbour outside it is taken
// as zero (fma(0, x, s) == s). The terms are added in the stencil's order:
// centre, west, east, south (south-west, south-east), north (north-west,
// north-east).
template <typename T, int NB>
__device__ __forceinline__ T stencil_apply(const Lev &L, size_t K, int i, int j, const T *xp) {
  const int nx = L.nx, ny = L.ny, hx = half_x(nx);
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
  const size_t row = (size_t)j * nx;
  const size_t kw = row + xslot(hw ? i - 1 : i, hx), ke = row + xslot(he ? i + 1 : i, hx);
  const size_t ds = hs ? nx : 0, dn = hn ? nx : 0;
  const auto x = gbl(xp);
  auto band = [&](int b) { return gbl(cptr<T>(L.band[b]))[K]; };
  const T ac = band(BC), aw = band(BW), ae = band(BE), as = band(BS), an = band(BN);
  const T x0 = x[K], xw = x[kw], xe = x[ke], xs = x[K - ds], xn = x[K + dn];
  T asw = T(0), ase = T(0), anw = T(0), ane = T(0), xsw = T(0), xse = T(0), xnw = T(0), xne = T(0);
  if (NB == 9) {
    asw = band(BSW); ase = band(BSE); anw = band(BNW); ane = band(BNE);
    xsw = x[kw - ds]; xse = x[ke - ds]; xnw = x[kw + dn]; xne = x[ke + dn];
  }
  // explicit fused multiply-adds, so that every caller rounds alike
  T s = ac * x0;
  s = fma(hw ? aw : T(0), xw, s);
  s = fma(he ? ae : T(0), xe, s);
  s = fma(hs ? as : T(0), xs, s);
  if (NB == 9) {
    s = fma(hs && hw ? asw : T(0), xsw, s);
    s = fma(hs && he ? ase : T(0), xse, s);
  }
  s = fma(hn ? an : T(0), xn, s);
  if (NB == 9) {
    s = fma(hn && hw ? anw : T(0), xnw, s);
    s = fma(hn && he ? ane : T(0), xne, s);
  }
  return s;
}

// M x at slot K with the FP64 level-0 bands (x in FP64, or an FP32 vector
// promoted on load); loads first, as in stencil_apply
template <typename X>
__device__ __forceinline__ double spmv64(const InstDev &I, size_t K, int i, int j, const X *xp) {
  const int nx = I.nx, hx = half_x(nx);
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < I.ny - 1;
  const size_t row = (size_t)j * nx;
  const size_t kw = row + xslot(hw ? i - 1 : i, hx), ke = row + xslot(he ? i + 1 : i, hx);
  const auto x = gbl(xp);
  const double ac = gbl(I.band64[BC])[K], aw = gbl(I.band64[BW])[K], ae = gbl(I.band64[BE])[K];
  const double as = gbl(I.band64[BS])[K], an = gbl(I.band64[BN])[K];
  const double x0 = (double)x[K], xw = (double)x[kw], xe = (double)x[ke];
  const double xs = (double)x[hs ? K - nx : K], xn = (double)x[hn ? K + nx : K];
  double s = ac * x0;
  s += (hw ? aw : 0.0) * xw;
  s += (he ? ae : 0.0) * xe;
  s += (hs ? as : 0.0) * xs;
  s += (hn ? an : 0.0) * xn;
  return s;
}

// r_l = b_l - A_l x_l into the level's scratch (the line-solve buffer cp is
// free between the pre- and post-smoothing sweeps), after a pre-smoothing
// sweep: with x_l = 0 the restriction reads b_l itself, and after an x-line
// pre-sweep alone restrict_residual_point forms the residual it needs.
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_residual(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map, int l,
                                                    int xid, int bid) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const Lev &F = I.lev[l];
  const int K = m.k0 + threadIdx.x;
  if (K >= F.nx * F.ny) return;
  int i, j;
  cell_ij(K, F.nx, half_x(F.nx), i, j);
  T *x; const T *b;
  level_xb<T>(I, l, xid, bid, x, b);
  ptr<T>(F.cp)[K] = b[K] - stencil_apply<T, NB>(F, K, i, j, x);
}

// b_{l+1}(C) = sum_f P(f, C) r_l(f) over the 3x3 fine block around (2I, 2J).
// FROM_B: x_l is still zero, so r_l is the level's right-hand side itself.
// Work-items take the coarse points of a row in column order (not storage
// order): the fine columns 2I and 2I +- 1 of consecutive work-items are then
// consecutive in the fine row's even and odd halves, so every load is
// contiguous across the wave; the coarse store alternates between the two
// contiguous halves of the coarse row.
template <typename T, bool FROM_B>
__device__ __forceinline__ void restrict_point(const InstDev &I, int l, int xid, int bid, int KC) {
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int J = KC / Cl.nx, Ic = KC - J * Cl.nx;    // column order
  const int hxf = half_x(F.nx);
  T *x; const T *b;
  level_xb<T>(I, l, xid, bid, x, b);
  const auto r = gbl(FROM_B ? b : cptr<T>(F.cp));
  const auto pw0 = gbl(cptr<T>(F.pw[0])), pw1 = gbl(cptr<T>(F.pw[1]));
  const auto pw2 = gbl(cptr<T>(F.pw[2])), pw3 = gbl(cptr<T>(F.pw[3]));
  // All 18 loads are issued before the sum (fine points outside the grid are
  // clamped in and weighted by zero), then added in row-major order.
  T wv[3][3], rv[3][3];
  bool ok[3][3];
#pragma unroll
  for (int db = -1; db <= 1; ++db) {
    const int fj = 2 * J + db;
    const bool okj = fj >= 0 && fj < F.ny;
    const size_t rowf = (size_t)(okj ? fj : 2 * J) * F.nx;
#pragma unroll
    for (int da = -1; da <= 1; ++da) {
      const int fi = 2 * Ic + da;
      const bool oki = fi >= 0 && fi < F.nx;
      const size_t kf = rowf + xslot(oki ? fi : 2 * Ic, hxf);
      // slot of C among the parents of f: (I - (fi >> 1)) + 2 (J - (fj >> 1))
      const int slot = (da < 0 ? 1 : 0) + 2 * (db < 0 ? 1 : 0);
      const auto w = slot == 0 ? pw0 : slot == 1 ? pw1 : slot == 2 ? pw2 : pw3;
      // the fine point at C itself is injected with weight exactly 1 (which
      // k_pweights does not store)
      wv[db + 1][da + 1] = (da == 0 && db == 0) ? T(1) : T(w[kf]);
      rv[db + 1][da + 1] = r[kf];
      ok[db + 1][da + 1] = okj && oki;
    }
  }
  T s = T(0);
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) s = fma(ok[b][a] ? wv[b][a] : T(0), rv[b][a], s);
  gbl(ptr<T>(Cl.b))[(size_t)J * Cl.nx + xslot(Ic, half_x(Cl.nx))] = s;
}

// b_{l+1}(C) = sum_f P(f, C) r_l(f) after an x-line pre-sweep: r_l vanishes
// on the odd rows, and the three fine points of row 2 J form theirs here,
// r = b - A x as residual_point does (same operations), instead of reading a
// stored residual: one launch per level and no residual array written or
// read. The odd fine points are shared by two coarse points, so they are
// formed twice.
template <typename T, int NB>
__device__ __forceinline__ void restrict_residual_point(const InstDev &I, int l, int xid, int bid,
                                                        int KC) {
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int J = KC / Cl.nx, Ic = KC - J * Cl.nx;    // column order
  const int hxf = half_x(F.nx);
  T *x; const T *b;
  level_xb<T>(I, l, xid, bid, x, b);
  const auto pw0 = gbl(cptr<T>(F.pw[0])), pw1 = gbl(cptr<T>(F.pw[1]));
  const int fj = 2 * J;
  const size_t rowf = (size_t)fj * F.nx;
  T wv[3], rv[3];
  bool ok[3];
#pragma unroll
  for (int da = -1; da <= 1; ++da) {
    const int fi = 2 * Ic + da;
    const bool oki = fi >= 0 && fi < F.nx;
    const int ci = oki ? fi : 2 * Ic;
    const size_t kf = rowf + xslot(ci, hxf);
    // parent slot of C: 1 for the point to the west (da < 0), 0 otherwise;
    // the point at C itself is injected with weight exactly 1
    wv[da + 1] = da == 0 ? T(1) : T(da < 0 ? pw1[kf] : pw0[kf]);
    rv[da + 1] = gbl(b)[kf] - stencil_apply<T, NB>(F, kf, ci, fj, x);
    ok[da + 1] = oki;
  }
  T s = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a) s = fma(ok[a] ? wv[a] : T(0), rv[a], s);
  gbl(ptr<T>(Cl.b))[(size_t)J * Cl.nx + xslot(Ic, half_x(Cl.nx))] = s;
}

template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_restrict_residual(const InstDev *__restrict__ insts,
                                                             const BlkMap *__restrict__ map, int l,
                                                             int xid, int bid) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const int KC = m.k0 + threadIdx.x;
  if (KC >= I.lev[l + 1].nx * I.lev[l + 1].ny) return;
  restrict_residual_point<T, NB>(I, l, xid, bid, KC);
}

template <typename T, bool FROM_B>
__global__ void __launch_bounds__(kBlock) k_restrict(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map, int l,
                                                    int xid, int bid) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const int KC = m.k0 + threadIdx.x;
  if (KC >= I.lev[l + 1].nx * I.lev[l + 1].ny) return;
  restrict_point<T, FROM_B>(I, l, xid, bid, KC);
}

// x_l += P x_{l+1}. ZERO: x_l = P x_{l+1}.
// ROWS 0: every cell. ROWS 1: odd rows only; ROWS 2: odd columns only. A
// zebra line sweep solves its colour-0 lines from their neighbours and never
// reads the lines it solves, and the colour-1 lines it reads are solved next:
// after the prolongation only the colour-1 lines of the first post-smoothing
// sweep are read (rows 1, 3, .. for x-lines; columns 1, 3, .. for y-lines).
// The block map covers that compact range.
template <int ROWS> __device__ __forceinline__ int prolong_count(const Lev &F) {
  return ROWS == 1 ? F.nx * (F.ny >> 1) : ROWS == 2 ? (F.nx - half_x(F.nx)) * F.ny : F.nx * F.ny;
}

template <typename T, bool ZERO, int ROWS>
__device__ __forceinline__ void prolong_point(const InstDev &I, int l, int xid, int idx) {
  const Lev &F = I.lev[l];
  const Lev &Cl = I.lev[l + 1];
  const int hx = half_x(F.nx);
  int K, i, j;
  if (ROWS == 1) {            // rows 1, 3, ..: nx cells each
    if (idx >= F.nx * (F.ny >> 1)) return;
    const int r = idx / F.nx, p = idx - r * F.nx;
    j = 2 * r + 1; i = xcol(p, hx); K = j * F.nx + p;
  } else if (ROWS == 2) {     // columns 1, 3, ..: the second half of each row
    const int no = F.nx - hx;
    if (idx >= no * F.ny) return;
    j = idx / no;
    const int p = hx + (idx - j * no);
    i = xcol(p, hx); K = j * F.nx + p;
  } else {
    K = idx;
    if (K >= F.nx * F.ny) return;
    cell_ij(K, F.nx, hx, i, j);
  }
  const int I0 = i >> 1, J0 = j >> 1;
  const int I1 = I0 + 1 < Cl.nx ? I0 + 1 : I0, J1 = J0 + 1 < Cl.ny ? J0 + 1 : J0;
  const auto xc = gbl(cptr<T>(Cl.x));
  const int hxc = half_x(Cl.nx);
  const size_t r0 = (size_t)J0 * Cl.nx, r1 = (size_t)J1 * Cl.nx;
  const int s0 = xslot(I0, hxc), s1 = xslot(I1, hxc);
  // Weights of parents outside the coarse grid are zero, and so are the
  // weights to the parents at I0 + 1 of an even i and at J0 + 1 of an even j
  // (k_pweights writes exact zeros there): those terms are skipped, so half
  // of each row's weight runs is not read. Each parity case issues all its
  // loads before it adds (i's parity is uniform over each half-row).
  const auto w0 = gbl(cptr<T>(F.pw[0])), w1 = gbl(cptr<T>(F.pw[1]));
  const auto w2 = gbl(cptr<T>(F.pw[2])), w3 = gbl(cptr<T>(F.pw[3]));
  T e;
  if ((i & 1) && (j & 1)) {
    const T a0 = w0[K], a1 = w1[K], a2 = w2[K], a3 = w3[K];
    const T c0 = xc[r0 + s0], c1 = xc[r0 + s1], c2 = xc[r1 + s0], c3 = xc[r1 + s1];
    e = a0 * c0;
    e += a1 * c1;
    e += a2 * c2;
    e += a3 * c3;
  } else if (i & 1) {
    const T a0 = w0[K], a1 = w1[K], c0 = xc[r0 + s0], c1 = xc[r0 + s1];
    e = a0 * c0;
    e += a1 * c1;
  } else if (j & 1) {
    const T a0 = w0[K], a2 = w2[K], c0 = xc[r0 + s0], c2 = xc[r1 + s0];
    e = a0 * c0;
    e += a2 * c2;
  } else {
    e = xc[r0 + s0];   // injection: the even-even weight is exactly 1
  }
  const auto x = gbl(l == 0 ? ptr<T>(vec(I, xid)) : ptr<T>(F.x));
  x[K] = ZERO ? e : x[K] + e;
}

// A workgroup takes kEnt map entries. Each entry's chain of dependent reads
// (map entry, instance, its status, its level's pointers) and then its loads
// are issued before any entry is computed or stored, so a workgroup has
// kEnt x 256 cells' loads in flight instead of waiting out one chain per 256
// cells. The terms are those of prolong_point in its order (a zero weight of
// a parent that point skips adds fma(0, 0, e) = e).
template <typename T, bool ZERO, int ROWS>
__global__ void __launch_bounds__(kBlock) k_prolong(const InstDev *__restrict__ insts,
                                                   const BlkMap *__restrict__ map, int l,
                                                   int xid, int nmap) {
  T a0[kEnt], a1[kEnt], a2[kEnt], a3[kEnt], c0[kEnt], c1[kEnt], c2[kEnt], c3[kEnt], xo[kEnt];
  T *xs[kEnt];
  int Ks[kEnt];
  bool on[kEnt];
#pragma unroll
  for (int e = 0; e < kEnt; ++e) {
    const int bi = kEnt * blockIdx.x + e;
    on[e] = false;
    a0[e] = a1[e] = a2[e] = a3[e] = c0[e] = c1[e] = c2[e] = c3[e] = xo[e] = T(0);
    xs[e] = nullptr;
    Ks[e] = 0;
    if (bi >= nmap) continue;
    const BlkMap m = map[bi];
    const InstDev &I = insts[m.inst];
    if (rsc(I)->status != ST_ACTIVE) continue;
    const Lev &F = I.lev[l];
    const Lev &Cl = I.lev[l + 1];
    const int idx = m.k0 + threadIdx.x;
    if (idx >= prolong_count<ROWS>(F)) continue;
    const int hx = half_x(F.nx);
    int K, i, j;
    if (ROWS == 1) {            // rows 1, 3, ..: nx cells each
      const int r = idx / F.nx, p = idx - r * F.nx;
      j = 2 * r + 1; i = xcol(p, hx); K = j * F.nx + p;
    } else if (ROWS == 2) {     // columns 1, 3, ..: the second half of each row
      const int no = F.nx - hx;
      j = idx / no;
      const int p = hx + (idx - j * no);
      i = xcol(p, hx); K = j * F.nx + p;
    } else {
      K = idx;
      cell_ij(K, F.nx, hx, i, j);
    }
    const int I0 = i >> 1, J0 = j >> 1;
    const int I1 = I0 + 1 < Cl.nx ? I0 + 1 : I0, J1 = J0 + 1 < Cl.ny ? J0 + 1 : J0;
    const auto xc = gbl(cptr<T>(Cl.x));
    const int hxc = half_x(Cl.nx);
    const size_t r0 = (size_t)J0 * Cl.nx, r1 = (size_t)J1 * Cl.nx;
    const int s0 = xslot(I0, hxc), s1 = xslot(I1, hxc);
    const bool io = i & 1, jo = j & 1;
    a0[e] = (io || jo) ? gbl(cptr<T>(F.pw[0]))[K] : T(1);   // even-even: injected
    c0[e] = xc[r0 + s0];
    if (io) { a1[e] = gbl(cptr<T>(F.pw[1]))[K]; c1[e] = xc[r0 + s1]; }
    if (jo) { a2[e] = gbl(cptr<T>(F.pw[2]))[K]; c2[e] = xc[r1 + s0]; }
    if (io && jo) { a3[e] = gbl(cptr<T>(F.pw[3]))[K]; c3[e] = xc[r1 + s1]; }
    T *x = l == 0 ? ptr<T>(vec(I, xid)) : ptr<T>(F.x);
    if (!ZERO) xo[e] = gbl(x)[K];
    xs[e] = x;
    Ks[e] = K;
    on[e] = true;
  }
#pragma unroll
  for (int e = 0; e < kEnt; ++e) {
    if (!on[e]) continue;
    T v = a0[e] * c0[e];
    v += a1[e] * c1[e];
    v += a2[e] * c2[e];
    v += a3[e] * c3[e];
    gbl(xs[e])[Ks[e]] = ZERO ? v : xo[e] + v;
  }
}

// ---------------------------------------------------------------------------
// Fused coarse cycle: the levels of one instance from `lev` (its first level
// with at most kFuseCells cells) down to its coarsest and back, in one
// workgroup with barriers between the phases, instead of several launches per
// level. The cycle is the per-level one: per level on the way down an
// optional pre-smoothing sweep of x-lines and the restriction of the
// residual (of b itself without pre-smoothing), the exact line solve of the
// coarsest level, then per level the prolongation of the lines the first
// post-sweep reads and the zebra post-sweeps.
//
// A level being smoothed lives in LDS (x, and per sweep the line right-hand
// sides and factors), so only the parallel phases touch global memory: every
// thread forms the right-hand sides of some cells (b minus the couplings to
// the other colour's lines) and copies their factors, then one thread per line
// runs the two Thomas recurrences out of LDS (lines are short here). The line
// solves are the per-level ones rounded in another order.
// ---------------------------------------------------------------------------
template <typename T, int CAP>
struct FuseLds {
  T x[CAP];     // x of the level being smoothed
  T r[CAP];     // line right-hand sides, then d of the forward pass
  T lo[CAP];    // coupling to the previous cell of the line (0 at t = 0)
  T inv[CAP];   // 1 / pivot
  T up[CAP];    // upper / pivot
};

// Levels of at most SCAP cells are staged into LDS whole (b, the off-diagonal
// bands and the line factors of both directions) once per level, so that the
// four sweeps read no global memory; larger levels (up to CAP cells) and FP64
// levels read them from global memory in each sweep's parallel phase. The
// variants keep a workgroup's LDS within 80 KB, two workgroups per CU:
//   FP32: CAP 1024 with SCAP 1024 (72 KB), CAP 2048 with SCAP 512 (66 KB),
//         CAP 4096 without staging (80 KB);  FP64: CAP 1024 (40 KB).
enum FuseStageArr { FS_B, FS_W, FS_E, FS_S, FS_N, FS_SW, FS_SE, FS_NW, FS_NE, FS_I0, FS_U0, FS_I1,
                    FS_U1, FS_COUNT };
template <typename T, int SCAP> struct FuseStage { T a[FS_COUNT][SCAP > 0 ? SCAP : 1]; };

template <typename T, int NB, int SCAP>
__device__ void fused_stage(const InstDev &I, int l, int bid, FuseStage<T, SCAP> &G) {
  const Lev &Lv = I.lev[l];
  const int nc = Lv.nx * Lv.ny;
  const auto b = gbl(l == 0 ? cptr<T>(vec(I, bid)) : cptr<T>(Lv.b));
  for (int K = threadIdx.x; K < nc; K += blockDim.x) {
    G.a[FS_B][K] = b[K];
    G.a[FS_W][K] = gbl(cptr<T>(Lv.band[BW]))[K];
    G.a[FS_E][K] = gbl(cptr<T>(Lv.band[BE]))[K];
    G.a[FS_S][K] = gbl(cptr<T>(Lv.band[BS]))[K];
    G.a[FS_N][K] = gbl(cptr<T>(Lv.band[BN]))[K];
    if (NB == 9) {
      G.a[FS_SW][K] = gbl(cptr<T>(Lv.band[BSW]))[K];
      G.a[FS_SE][K] = gbl(cptr<T>(Lv.band[BSE]))[K];
      G.a[FS_NW][K] = gbl(cptr<T>(Lv.band[BNW]))[K];
      G.a[FS_NE][K] = gbl(cptr<T>(Lv.band[BNE]))[K];
    }
    G.a[FS_I0][K] = gbl(cptr<T>(Lv.linv[0]))[K];
    G.a[FS_U0][K] = gbl(cptr<T>(Lv.lup[0]))[K];
    G.a[FS_I1][K] = gbl(cptr<T>(Lv.linv[1]))[K];
    G.a[FS_U1][K] = gbl(cptr<T>(Lv.lup[1]))[K];
  }
}

// One zebra sweep (lines of one colour) of level l. STG: the level's inputs
// are in G (LDS); otherwise they are read from global memory and the line
// factors copied to S.lo, S.inv, S.up. Both read the same values and do the
// same operations.
template <typename T, int NB, int DIR, bool STG, int CAP, int SCAP>
__device__ void fused_sweep(const InstDev &I, int l, int bid, int colour, bool zero,
                            FuseLds<T, CAP> &S, FuseStage<T, SCAP> &G) {
  const Lev &Lv = I.lev[l];
  const int nx = Lv.nx, ny = Lv.ny, hx = half_x(nx);
  const T *b = l == 0 ? cptr<T>(vec(I, bid)) : cptr<T>(Lv.b);
  const T *lo = cptr<T>(Lv.band[DIR == 0 ? BW : BS]);
  const T *om = cptr<T>(Lv.band[DIR == 0 ? BS : BW]);
  const T *op = cptr<T>(Lv.band[DIR == 0 ? BN : BE]);
  const T *omm = nullptr, *omp = nullptr, *opm = nullptr, *opp = nullptr;
  if (NB == 9) {
    omm = cptr<T>(Lv.band[BSW]);
    omp = cptr<T>(Lv.band[DIR == 0 ? BSE : BNW]);
    opm = cptr<T>(Lv.band[DIR == 0 ? BNW : BSE]);
    opp = cptr<T>(Lv.band[BNE]);
  }
  const T *inv = cptr<T>(Lv.linv[DIR]);
  const T *cu = cptr<T>(Lv.lup[DIR]);
  // staged copies of the same arrays
  constexpr int gLo = DIR == 0 ? FS_W : FS_S, gOm = DIR == 0 ? FS_S : FS_W;
  constexpr int gOp = DIR == 0 ? FS_N : FS_E, gOmm = FS_SW, gOmp = DIR == 0 ? FS_SE : FS_NW;
  constexpr int gOpm = DIR == 0 ? FS_NW : FS_SE, gOpp = FS_NE;
  constexpr int gInv = DIR == 0 ? FS_I0 : FS_I1, gUp = DIR == 0 ? FS_U0 : FS_U1;
  auto val = [&](const T *p, int g, uint32_t K) -> T {
    if constexpr (STG) return G.a[g][K];
    else return ld32(p, K);
  };
  const int n = DIR == 0 ? nx : ny, nperp = DIR == 0 ? ny : nx;
  const int nl = (nperp - colour + 1) >> 1;   // lines colour, colour + 2, ..
  auto at = [&](int li, int t) -> uint32_t {
    return DIR == 0 ? (uint32_t)(li * nx + xslot(t, hx)) : (uint32_t)(t * nx + xslot(li, hx));
  };
  for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
    const int q = idx / n, t = idx - q * n, li = colour + 2 * q;
    const uint32_t K = at(li, t);
    T r = val(b, FS_B, K);
    if (!zero) {
      if (li > 0) {
        r -= val(om, gOm, K) * S.x[at(li - 1, t)];
        if (NB == 9) {
          if (t > 0) r -= val(omm, gOmm, K) * S.x[at(li - 1, t - 1)];
          if (t < n - 1) r -= val(omp, gOmp, K) * S.x[at(li - 1, t + 1)];
        }
      }
      if (li < nperp - 1) {
        r -= val(op, gOp, K) * S.x[at(li + 1, t)];
        if (NB == 9) {
          if (t > 0) r -= val(opm, gOpm, K) * S.x[at(li + 1, t - 1)];
          if (t < n - 1) r -= val(opp, gOpp, K) * S.x[at(li + 1, t + 1)];
        }
      }
    }
    S.r[K] = r;
    if constexpr (!STG) {
      S.lo[K] = t > 0 ? ld32(lo, K) : T(0);
      S.inv[K] = ld32(inv, K);
      S.up[K] = ld32(cu, K);
    }
  }
  __syncthreads();
  const T *s_lo, *s_inv, *s_up;
  if constexpr (STG) {
    s_lo = G.a[gLo]; s_inv = G.a[gInv]; s_up = G.a[gUp];
  } else {
    s_lo = S.lo; s_inv = S.inv; s_up = S.up;
  }
  for (int q = threadIdx.x; q < nl; q += blockDim.x) {
    const int li = colour + 2 * q;
    // (unrolled so that the LDS reads of several steps, which do not depend
    // on the recurrence, are issued together)
    T d = T(0);
#pragma unroll 8
    for (int t = 0; t < n; ++t) {   // forward: d_t = (r_t - lo_t d_{t-1}) / pivot_t
      const uint32_t K = at(li, t);
      const T lv = t > 0 ? s_lo[K] : T(0);
      d = (S.r[K] - lv * d) * s_inv[K];
      S.r[K] = d;
    }
    T xn = T(0);
#pragma unroll 8
    for (int t = n - 1; t >= 0; --t) {   // backward: x_t = d_t - u_t x_{t+1}
      const uint32_t K = at(li, t);
      xn = S.r[K] - s_up[K] * xn;
      S.x[K] = xn;
    }
  }
  __syncthreads();
}

// x of level l: global -> LDS (load) or LDS -> global (store); ends with a
// barrier.
template <typename T, int CAP>
__device__ void fused_x(const InstDev &I, int l, int xid, FuseLds<T, CAP> &S, bool load) {
  const auto x = gbl(l == 0 ? ptr<T>(vec(I, xid)) : ptr<T>(I.lev[l].x));
  const int nc = I.lev[l].nx * I.lev[l].ny;
  for (int K = threadIdx.x; K < nc; K += blockDim.x) {
    if (load) S.x[K] = x[K];
    else x[K] = S.x[K];
  }
  __syncthreads();
}

template <typename T, int NB, bool STG, int CAP, int SCAP>
__device__ void fused_sweeps(const InstDev &I, int l, int bid, bool post_x, bool post_y,
                             FuseLds<T, CAP> &S, FuseStage<T, SCAP> &G) {
  if (post_x) {
    fused_sweep<T, NB, 0, STG>(I, l, bid, 0, false, S, G);
    fused_sweep<T, NB, 0, STG>(I, l, bid, 1, false, S, G);
  }
  if (post_y) {
    fused_sweep<T, NB, 1, STG>(I, l, bid, 0, false, S, G);
    fused_sweep<T, NB, 1, STG>(I, l, bid, 1, false, S, G);
  }
}

template <typename T, int NB, int CAP, int SCAP>
__device__ void fused_smooth(const InstDev &I, int l, int xid, int bid, bool post_x, bool post_y,
                             FuseLds<T, CAP> &S, FuseStage<T, SCAP> &G) {
  if constexpr (SCAP > 0) {
    if (I.lev[l].nx * I.lev[l].ny <= SCAP) {   // uniform over the workgroup
      fused_stage<T, NB>(I, l, bid, G);   // the barrier of fused_x covers it
      fused_x<T>(I, l, xid, S, true);
      fused_sweeps<T, NB, true>(I, l, bid, post_x, post_y, S, G);
      fused_x<T>(I, l, xid, S, false);
      return;
    }
  }
  fused_x<T>(I, l, xid, S, true);
  fused_sweeps<T, NB, false>(I, l, bid, post_x, post_y, S, G);
  fused_x<T>(I, l, xid, S, false);
}

// smoother: the finest level's schedule in bits 0-3, the coarser levels' in
// bits 4-7 (post-smoothing bits and the x-line pre-smoothing bit)
template <typename T, int CAP, int SCAP>
__global__ void __launch_bounds__(kFuseThreads) k_coarse_cycle(const InstDev *__restrict__ insts,
                                                               const FuseMap *__restrict__ map,
                                                               int xid, int bid, int smoother) {
  __shared__ FuseLds<T, CAP> S;
  __shared__ FuseStage<T, SCAP> G;
  const FuseMap fm = map[blockIdx.x];
  const InstDev &I = insts[fm.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const int lf = fm.lev, nl = I.nlev;
  auto sched = [&](int m) { return m == 0 ? (smoother & 15) : (smoother >> 4); };
  for (int m = lf; m + 1 < nl; ++m) {
    const int nc = I.lev[m + 1].nx * I.lev[m + 1].ny;
    if (sched(m) & 1) {
      // pre-smoothing with x-lines from x_m = 0 (colour 0 from b alone), x_m
      // kept for the ascent, then b_{m+1} = R r_m with r_m = b_m - A_m x_m,
      // which is zero on the odd rows (solved last)
      if (m == 0) {
        fused_sweep<T, 5, 0, false>(I, m, bid, 0, true, S, G);
        fused_sweep<T, 5, 0, false>(I, m, bid, 1, false, S, G);
      } else {
        fused_sweep<T, 9, 0, false>(I, m, bid, 0, true, S, G);
        fused_sweep<T, 9, 0, false>(I, m, bid, 1, false, S, G);
      }
      fused_x<T>(I, m, xid, S, false);
      for (int KC = threadIdx.x; KC < nc; KC += blockDim.x) {
        if (m == 0) restrict_residual_point<T, 5>(I, m, xid, bid, KC);
        else restrict_residual_point<T, 9>(I, m, xid, bid, KC);
      }
    } else {   // b_{m+1} = R b_m (x_m = 0)
      for (int KC = threadIdx.x; KC < nc; KC += blockDim.x)
        restrict_point<T, true>(I, m, xid, bid, KC);
    }
    __syncthreads();
  }
  {   // exact solve of the coarsest level (x = 0): one row or one column
    const int c = nl - 1;
    if (I.lev[c].ny == 1) {
      if (c == 0) fused_sweep<T, 5, 0, false>(I, c, bid, 0, true, S, G);
      else fused_sweep<T, 9, 0, false>(I, c, bid, 0, true, S, G);
    } else {
      if (c == 0) fused_sweep<T, 5, 1, false>(I, c, bid, 0, true, S, G);
      else fused_sweep<T, 9, 1, false>(I, c, bid, 0, true, S, G);
    }
    fused_x<T>(I, c, xid, S, false);
  }
  for (int m = nl - 2; m >= lf; --m) {  // x_m (+)= P x_{m+1}, then post-smoothing
    const Lev &F = I.lev[m];
    // schedule of level m: the finest level's (low 4 bits) or the coarser ones'
    const int sml = sched(m);
    const bool post_x = sml & 4, post_y = sml & 8;
    const int rows = prolong_rows(sml);
    const int np = rows == 1 ? prolong_count<1>(F) : rows == 2 ? prolong_count<2>(F)
                                                                : prolong_count<0>(F);
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
      if (sml & 1) {   // pre-smoothed: the correction is added to x_m
        if (rows == 1) prolong_point<T, false, 1>(I, m, xid, k);
        else if (rows == 2) prolong_point<T, false, 2>(I, m, xid, k);
        else prolong_point<T, false, 0>(I, m, xid, k);
      } else {
        if (rows == 1) prolong_point<T, true, 1>(I, m, xid, k);
        else if (rows == 2) prolong_point<T, true, 2>(I, m, xid, k);
        else prolong_point<T, true, 0>(I, m, xid, k);
      }
    }
    __syncthreads();
    if (m == 0) fused_smooth<T, 5>(I, m, xid, bid, post_x, post_y, S, G);
    else fused_smooth<T, 9>(I, m, xid, bid, post_x, post_y, S, G);
  }
}

template <typename T>
void launch_coarse_cycle(hipStream_t s, const InstDev *d, const FuseMap *m, int n, int xid, int bid,
                         int smoother, int cap) {
  if (!n) return;
  if constexpr (sizeof(T) == 8) {
    k_coarse_cycle<T, 1024, 0><<<n, kFuseThreads, 0, s>>>(d, m, xid, bid, smoother);
  } else {
    if (cap <= 1024) k_coarse_cycle<T, 1024, 1024><<<n, kFuseThreads, 0, s>>>(d, m, xid, bid, smoother);
    else if (cap <= 2048) k_coarse_cycle<T, 2048, 512><<<n, kFuseThreads, 0, s>>>(d, m, xid, bid, smoother);
    else k_coarse_cycle<T, 4096, 0><<<n, kFuseThreads, 0, s>>>(d, m, xid, bid, smoother);
  }
}
template void launch_coarse_cycle<double>(hipStream_t, const InstDev *, const FuseMap *, int, int,
                                          int, int, int);
template void launch_coarse_cycle<float>(hipStream_t, const InstDev *, const FuseMap *, int, int,
                                         int, int, int);

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// BiCGSTAB (right-preconditioned), precision T for vectors and sums; the
// operator is the level-0 5-band matrix of the V-cycle precision.
//
// Reductions are two-phase: every workgroup stores its partial sums (plain
// stores), and a finish kernel with one workgroup per instance adds them in a
// fixed order and updates the scalars. The kernel boundary orders the two, so
// no device-wide fence or atomic is needed, and the result of an instance does
// not depend on the rest of the batch.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T spmv5(const InstDev &I, size_t K, int i, int j, const T *x) {
  return stencil_apply<T, 5>(I.lev[0], K, i, j, x);
}

// PC (FP64 Krylov, FP32 V-cycle): the preconditioned vectors ph = M^-1 p and
// sh = M^-1 s are FP32 V-cycle outputs scaled by pcs, the residual norm of
// the iteration (the V-cycle input is the vector divided by pcs, so it stays
// in FP32 range however far the residual has fallen).
template <typename T, bool PC>
__device__ __forceinline__ T precond_vec(const InstDev &I, int id, int pcid, size_t K) {
  if (PC) return (T)(rsc(I)->pcs * (double)gbl(I.pv[pcid - kNumVec])[K]);
  return gbl(cptr<T>(I.kv[id]))[K];
}

// p = r + beta (p - omega v).
// PC: p is stored only in FP32, divided by the residual norm pcs (it is the
// V-cycle input P = p / pcs). p is a search direction: x moves along
// ph = M^-1 P and r along v = A ph, both from the same ph, so the rounding of
// P changes the direction (like the FP32 V-cycle does), never the agreement
// of r with R - A x. The first iteration (p = r0) ignores the stored P and v,
// and also stores P as the shadow residual rh = fl32(r0 / ||r0||): BiCGSTAB
// admits any fixed rh with (rh, r0) != 0, so rh is r0's direction, read in
// FP32 by every later iteration, with rho0 = (rh, r0) = ||r0|| to FP32
// rounding.
//
// A workgroup takes kLightBlocks (4) map entries: the chain of dependent
// reads in front of an entry's vector traffic (map entry, instance, its
// scalars) is latency this light kernel cannot hide within one 256-cell
// block, and the next entries' chains are issued while the first one's
// vectors move (the map, the instance table and the scalars are read-only
// here).
template <typename T, bool PC>
__global__ void __launch_bounds__(kBlock) k_bicg_p(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map, int nmap) {
#pragma unroll
  for (int e = 0; e < kLightBlocks; ++e) {
    const int bi = kLightBlocks * blockIdx.x + e;
    if (bi >= nmap) break;
    const BlkMap m = map[bi];
    const InstDev &I = insts[m.inst];
    if (rsc(I)->status != ST_ACTIVE) continue;
    const int K = m.k0 + threadIdx.x;
    if (K >= I.nx * I.ny) continue;
    const T beta = (T)rsc(I)->beta, omega = (T)rsc(I)->omega;
    if (PC) {
      const double pcs = rsc(I)->pcs, pcs_p = rsc(I)->pcs_p;   // read before the stores
      const bool first = rsc(I)->iters == 0;
      float *P = I.pv[PV_P - kNumVec];
      const T r = gbl(cptr<T>(I.kv[VR]))[K], v = gbl(cptr<T>(I.kv[VV]))[K];
      const T pold = (T)(pcs_p * (double)gbl(P)[K]);
      const T pv = first ? r : r + beta * (pold - omega * v);
      const float p32 = (float)(pv / pcs);
      gbl(P)[K] = p32;
      if (first) gbl(I.pv[PV_RH - kNumVec])[K] = p32;
    } else {
      T *p = ptr<T>(I.kv[VP]);
      gbl(p)[K] = gbl(cptr<T>(I.kv[VR]))[K] + beta * (gbl(p)[K] - omega * gbl(cptr<T>(I.kv[VV]))[K]);
    }
  }
}

// the shadow residual rh at slot K (PC: FP32, see k_bicg_p)
template <typename T, bool PC>
__device__ __forceinline__ T shadow(const InstDev &I, size_t K) {
  if (PC) return (T)gbl(I.pv[PV_RH - kNumVec])[K];
  return gbl(cptr<T>(I.kv[VRH]))[K];
}

// v = M ph; partial (rh, v)
template <typename T, bool PC>
__global__ void __launch_bounds__(kBlock) k_bicg_v(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  T acc[1] = {T(0)};
  if (K < N) {
    int i, j;
    cell_ij(K, I.nx, half_x(I.nx), i, j);
    const T rh = shadow<T, PC>(I, K);   // loaded before the store below
    const T v = PC ? (T)(rsc(I)->pcs * spmv64(I, K, i, j, cptr<float>(I.pv[PV_PH - kNumVec])))
                   : spmv5<T>(I, K, i, j, cptr<T>(I.kv[VPH]));
    gbl(ptr<T>(I.kv[VV]))[K] = v;
    acc[0] = rh * v;
  }
  store_partials<T, 1>(I, m.k0 / kBlock, acc);
}

// s = r - alpha v; partial (s, s); PC: also the scaled FP32 V-cycle input
template <typename T, bool PC>
__global__ void __launch_bounds__(kBlock) k_bicg_s(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map, int nmap) {
#pragma unroll
  for (int e = 0; e < kLightBlocks; ++e) {   // kLightBlocks map entries, as k_bicg_p
    const int bi = kLightBlocks * blockIdx.x + e;
    if (bi >= nmap) break;
    const BlkMap m = map[bi];
    const InstDev &I = insts[m.inst];
    if (rsc(I)->status != ST_ACTIVE) continue;   // uniform over the workgroup
    const int K = m.k0 + threadIdx.x;
    const int N = I.nx * I.ny;
    T acc[1] = {T(0)};
    if (K < N) {
      const double pcs = PC ? rsc(I)->pcs : 1.0;   // read before the stores
      const T s = gbl(cptr<T>(I.kv[VR]))[K] - (T)rsc(I)->alpha * gbl(cptr<T>(I.kv[VV]))[K];
      gbl(ptr<T>(I.kv[VS]))[K] = s;
      if (PC) gbl(I.pv[PV_B - kNumVec])[K] = (float)(s / pcs);
      acc[0] = s * s;
    }
    store_partials<T, 1>(I, m.k0 / kBlock, acc);
  }
}

// t = M sh; partials (t, s), (t, t)
template <typename T, bool PC>
__global__ void __launch_bounds__(kBlock) k_bicg_t(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->status != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  T acc[2] = {T(0), T(0)};
  if (K < N) {
    int i, j;
    cell_ij(K, I.nx, half_x(I.nx), i, j);
    const T sv = gbl(cptr<T>(I.kv[VS]))[K];   // loaded before the store below
    const T t = PC ? (T)(rsc(I)->pcs * spmv64(I, K, i, j, cptr<float>(I.pv[PV_SH - kNumVec])))
                   : spmv5<T>(I, K, i, j, cptr<T>(I.kv[VSH]));
    gbl(ptr<T>(I.kv[VT]))[K] = t;
    acc[0] = t * sv;
    acc[1] = t * t;
  }
  store_partials<T, 2>(I, m.k0 / kBlock, acc);
}

// x += alpha ph + omega sh; r = s - omega t; partials (rh, r), (r, r).
// HALF (converged at s): x += alpha ph only. XSTATE: x is the FP64 state,
// otherwise the inner FP32 correction of mixed precision.
template <typename T, bool XSTATE, bool PC>
__global__ void __launch_bounds__(kBlock) k_bicg_x(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  const int st = rsc(I)->status;
  if (st != ST_ACTIVE && st != ST_HALF) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  const T alpha = (T)rsc(I)->alpha, omega = (T)rsc(I)->omega;
  T acc[2] = {T(0), T(0)};
  if (K < N) {
    // every load before the first store (the stores could alias the later
    // loads for the compiler, which would then wait for each in turn)
    // (an instance that converged at s reads sh, s, t and rh needlessly, once;
    // they enter both cases through selects so that their loads are issued
    // with the others)
    T *x = XSTATE ? (T *)I.x : ptr<T>(I.kv[VXK]);
    const T xv = gbl(x)[K];
    const T ph = precond_vec<T, PC>(I, VPH, PV_PH, K);
    const T sh = precond_vec<T, PC>(I, VSH, PV_SH, K);
    const T sv = gbl(cptr<T>(I.kv[VS]))[K], tv = gbl(cptr<T>(I.kv[VT]))[K];
    const T rh = shadow<T, PC>(I, K);
    const bool act = st != ST_HALF;
    const T xh = xv + alpha * ph;
    const T r = sv - omega * tv;
    gbl(x)[K] = act ? xh + omega * sh : xh;
    acc[0] = act ? rh * r : T(0);
    acc[1] = act ? r * r : T(0);
    if (act) gbl(ptr<T>(I.kv[VR]))[K] = r;
  }
  if (st == ST_ACTIVE) store_partials<T, 2>(I, m.k0 / kBlock, acc);
}

// mixed precision: r64 = R - M x with the FP64 bands; partials (r, r), (R, R)
__global__ void __launch_bounds__(kBlock) k_ir_residual(const InstDev *__restrict__ insts,
                                                       const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->ostatus != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  double v[2] = {0.0, 0.0};
  if (K < N) {
    int i, j;
    cell_ij(K, I.nx, half_x(I.nx), i, j);
    const double r = gbl(I.R)[K] - spmv64(I, K, i, j, I.x);
    gbl(I.r64)[K] = r;
    v[0] = r * r;
    v[1] = gbl(I.R)[K] * gbl(I.R)[K];
  }
  store_partials<double, 2>(I, m.k0 / kBlock, v);
}

// inner start: r = r64 / ||r64|| (FP32), rh = r, p = v = xk = 0; partial (r, r)
__global__ void __launch_bounds__(kBlock) k_ir_inner_init(const InstDev *__restrict__ insts,
                                                         const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->ostatus != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  float v[1] = {0.0f};
  if (K < N) {
    const float r = (float)(gbl(I.r64)[K] / rsc(I)->scale);
    gbl(ptr<float>(I.kv[VR]))[K] = r;
    gbl(ptr<float>(I.kv[VRH]))[K] = r;
    gbl(ptr<float>(I.kv[VP]))[K] = 0.0f;
    gbl(ptr<float>(I.kv[VV]))[K] = 0.0f;
    gbl(ptr<float>(I.kv[VXK]))[K] = 0.0f;
    v[0] = r * r;
  }
  store_partials<float, 1>(I, m.k0 / kBlock, v);
}

// x += ||r64|| * xk (FP64)
__global__ void __launch_bounds__(kBlock) k_ir_update(const InstDev *__restrict__ insts,
                                                     const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->ostatus != ST_ACTIVE) return;
  const int K = m.k0 + threadIdx.x;
  if (K >= I.nx * I.ny) return;
  gbl(I.x)[K] += rsc(I)->scale * (double)gbl(ptr<float>(I.kv[VXK]))[K];
}

// Whether an instance's solve of this step has ended: the FP64 refinement
// (mixed precision), or the Krylov solve (the other precisions).
__device__ __forceinline__ bool solve_ended(int prec, int status, int ostatus) {
  if (prec == SG_MIXED) return ostatus == ST_DONE || ostatus == ST_FAIL;
  return status == ST_DONE || status == ST_FAIL;
}

// Finalize pass over the instances whose solve has ended and that were not
// finalized in an earlier pass of this step (phase 0): partials of the true
// residual ||R - M x||, ||R||, the count of negative x (the cells f = max(x, 0)
// clamps), the count of non-finite x and min x (slot 4). The finish
// kernel decides from them whether the step is accepted; k_commit then writes
// f. Nothing here writes f.
__global__ void __launch_bounds__(kBlock) k_finalize(const InstDev *__restrict__ insts,
                                                    const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->phase != 0 || !solve_ended(I.prec, rsc(I)->status, rsc(I)->ostatus)) return;
  const int K = m.k0 + threadIdx.x;
  const int N = I.nx * I.ny;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  double mn = DBL_MAX;
  if (K < N) {
    int i, j;
    cell_ij(K, I.nx, half_x(I.nx), i, j);
    const double xv = gbl(I.x)[K];
    const double r = gbl(I.R)[K] - spmv64(I, K, i, j, I.x);
    v[0] = r * r;
    v[1] = gbl(I.R)[K] * gbl(I.R)[K];
    v[2] = xv < 0.0 ? 1.0 : 0.0;
    v[3] = isfinite(xv) ? 0.0 : 1.0;
    mn = xv;
  }
  __shared__ double smin[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mn = fmin(mn, __shfl_down(mn, off, 64));
  if ((threadIdx.x & 63) == 0) smin[threadIdx.x >> 6] = mn;
  block_sum<double, 4>(v);   // contains __syncthreads
  if (threadIdx.x == 0) {
    double b = smin[0];
    for (int w = 1; w < kBlock / 64; ++w) b = fmin(b, smin[w]);
    const int lblk = m.k0 / kBlock;
    I.partial[lblk] = v[0];
    I.partial[(size_t)I.nblk0 + lblk] = v[1];
    I.partial[(size_t)2 * I.nblk0 + lblk] = v[2];
    I.partial[(size_t)3 * I.nblk0 + lblk] = v[3];
    I.partial[(size_t)4 * I.nblk0 + lblk] = b;
  }
}

// f = max(x, 0) (or f = x with noclamp) for the instances finalized in this
// pass whose step was accepted; min f >= 0 is then exact. A rejected step
// (ostatus ST_FAIL: the solve broke down, hit its iteration cap, or left a
// residual above the accepted floor, or x has non-finite cells) leaves f as it
// was.
// With the warm start it also keeps the step's increment df = f_new - f_old
// and the partial sums (df_new, df_old), (df_old, df_old) from which F_EXTRAP
// forms the next start's extrapolation factor.
__global__ void __launch_bounds__(kBlock) k_commit(const InstDev *__restrict__ insts,
                                                  const BlkMap *__restrict__ map) {
  const BlkMap m = map[blockIdx.x];
  const InstDev &I = insts[m.inst];
  if (rsc(I)->phase != 1 || rsc(I)->ostatus != ST_DONE) return;   // uniform
  const int K = m.k0 + threadIdx.x;
  const bool hist = rsc(I)->hist;   // df holds an earlier increment (uniform)
  double v[2] = {0.0, 0.0};
  if (K < I.nx * I.ny) {
    const double xv = gbl(I.x)[K];
    const double fn = (I.noclamp || xv > 0.0) ? xv : 0.0;
    if (I.warm) {
      const double dnew = fn - gbl(I.f)[K];
      if (hist) {
        const double dold = (double)gbl(I.df)[K];
        v[0] = dnew * dold;
        v[1] = dold * dold;
      }
      gbl(I.df)[K] = (float)dnew;
    }
    gbl(I.f)[K] = fn;
  }
  if (I.warm) store_partials<double, 2>(I, m.k0 / kBlock, v);
}

// Finish of a reduction: one workgroup per instance sums the partial sums in a
// fixed order and applies the BiCGSTAB / refinement scalar updates.
enum Finish { F_INIT64 = 0, F_V, F_S, F_T, F_X, F_IR_RES, F_INNER_INIT, F_FINALIZE, F_EXTRAP };

template <typename A, int NV>
__device__ __forceinline__ void sum_partials(const InstDev &I, A (&tot)[NV]) {
  // Thread t adds partials t, t + kFinThreads, .. in that order; the loads of
  // 8 of them are issued before they are added, so a large instance (65536
  // partials at 4096^2) costs a few memory latencies, not one per partial.
  constexpr int U = 8;
  A acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = A(0);
  const int nb = I.nblk0;
  int b = threadIdx.x;
  for (; b + (U - 1) * kFinThreads < nb; b += U * kFinThreads) {
    A v[NV][U];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < NV; ++k) v[k][u] = (A)I.partial[(size_t)k * nb + b + u * kFinThreads];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += v[k][u];
  }
  for (; b < nb; b += kFinThreads) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += (A)I.partial[(size_t)k * nb + b];
  }
  block_sum<A, NV, kFinThreads>(acc);
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = acc[k];   // valid in thread 0
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kFinThreads) k_finish(const InstDev *__restrict__ insts, int n,
                                                  int *__restrict__ cnt) {
  const int k = blockIdx.x;
  if (k >= n) return;
  const InstDev &I = insts[k];
  KScal &s = *I.sc;
  const int st = s.status, ost = s.ostatus;
  if (MODE == F_IR_RES || MODE == F_INNER_INIT) {
    if (ost != ST_ACTIVE) return;
  } else if (MODE == F_EXTRAP) {
    if (s.phase != 1 || ost != ST_DONE) return;   // committed in this pass
  } else if (MODE == F_FINALIZE) {
    if (s.phase == 1) {   // committed by the previous pass
      if (threadIdx.x == 0) s.phase = 2;
      return;
    }
    if (s.phase != 0 || !solve_ended(I.prec, st, ost)) return;
  } else if (MODE == F_X) {
    if (st != ST_ACTIVE && st != ST_HALF) return;
  } else {
    if (st != ST_ACTIVE) return;
  }
  if (MODE == F_INIT64) {
    double t[2];
    sum_partials<double, 2>(I, t);
    if (threadIdx.x == 0) {
      s.pcs = t[0] > 0.0 ? sqrt(t[0]) : 1.0;
      // PC32: the shadow residual is r0 / ||r0|| (k_bicg_p), so rho0 = ||r0||
      s.rho = I.prec == SG_PC32 ? t[0] / s.pcs : t[0];
      s.rho_old = 1.0; s.alpha = 1.0; s.omega = 1.0; s.beta = 0.0;
      s.rnorm2 = t[0]; s.bnorm2 = t[1]; s.iters = 0;
      s.best2 = t[0]; s.best_it = 0;
      if (t[0] <= s.tol2 * t[1]) { s.status = ST_DONE; atomicSub(cnt, 1); }
    }
  } else if (MODE == F_V) {   // after the V-cycle of p (active instances only)
    T t[1];
    sum_partials<T, 1>(I, t);
    if (threadIdx.x == 0) {
      s.vcyc += 1;
      // Breakdown: (rh, v) zero or not finite (a V-cycle output that
      // overflowed). The solve stops with x at its last iterate, which no
      // later kernel of the step updates (they skip failed instances).
      const double alpha = (double)((T)s.rho / t[0]);
      if (t[0] == T(0) || !isfinite((double)t[0]) || !isfinite(alpha)) {
        s.status = ST_FAIL;
        atomicSub(cnt, 1);
      } else {
        s.alpha = alpha;
      }
    }
  } else if (MODE == F_S) {
    T t[1];
    sum_partials<T, 1>(I, t);
    if (threadIdx.x == 0) {
      if (!isfinite((double)t[0])) {
        s.status = ST_FAIL;
        atomicSub(cnt, 1);
      } else if ((double)t[0] <= s.tol2 * s.bnorm2) {
        s.status = ST_HALF;
        s.rnorm2 = (double)t[0];
      }
    }
  } else if (MODE == F_T) {   // after the V-cycle of s (still-active instances only)
    T t[2];
    sum_partials<T, 2>(I, t);
    if (threadIdx.x == 0) {
      s.vcyc += 1;
      const double omega = t[1] > T(0) ? (double)(t[0] / t[1]) : 0.0;
      if (!isfinite((double)t[0]) || !isfinite((double)t[1]) || !isfinite(omega)) {
        s.status = ST_FAIL;   // before x moves along a non-finite direction
        atomicSub(cnt, 1);
      } else {
        s.omega = omega;
      }
    }
  } else if (MODE == F_X) {
    T t[2] = {T(0), T(0)};
    if (st == ST_ACTIVE) sum_partials<T, 2>(I, t);
    if (threadIdx.x == 0) {
      s.iters += 1;
      s.titers += 1;
      if (st == ST_HALF) {
        s.status = ST_DONE;
        atomicSub(cnt, 1);
      } else {
        s.rho_old = s.rho;
        s.rho = (double)t[0];
        s.rnorm2 = (double)t[1];
        s.pcs_p = s.pcs;
        s.pcs = s.rnorm2 > 0.0 ? sqrt(s.rnorm2) : 1.0;
        s.beta = (double)(((T)s.rho / (T)s.rho_old) * ((T)s.alpha / (T)s.omega));
        // Converged; or stalled at the rounding floor: no halving of the
        // residual for 4 iterations while below 1e-12 relative, or for 8
        // while below 1e-10 (the per-step error budget), or 100x the
        // tolerance for the loose inner solves of mixed precision. A plateau
        // of 4 above 1e-12 is often BiCGSTAB's, not the floor: it converges
        // on a few iterations later.
        bool floor = false;
        if (s.rnorm2 < 0.25 * s.best2) { s.best2 = s.rnorm2; s.best_it = s.iters; }
        else {
          const int flat = s.iters - s.best_it;
          floor = (flat >= 4 && s.rnorm2 <= fmax(1e-24, 1e4 * s.tol2) * s.bnorm2) ||
                  (flat >= 8 && s.rnorm2 <= fmax(1e-20, 1e4 * s.tol2) * s.bnorm2);
        }
        if (s.rnorm2 <= s.tol2 * s.bnorm2 || floor) {
          s.status = ST_DONE;
          atomicSub(cnt, 1);
        } else if (s.iters >= s.maxit || !(s.rnorm2 == s.rnorm2) || s.rho == 0.0 ||
                   s.omega == 0.0) {
          s.status = ST_FAIL;
          atomicSub(cnt, 1);
        }
      }
    }
  } else if (MODE == F_IR_RES) {
    double t[2];
    sum_partials<double, 2>(I, t);
    if (threadIdx.x == 0) {
      const bool first = s.orefs == 0;
      if (first) s.obnorm2 = t[1];
      s.onorm2 = t[0];
      s.scale = sqrt(t[0]);
      // Converged; or stalled at the FP64 rounding floor of R - M x, which is
      // about cond(M) eps and what a direct solve reaches too (the residual
      // stopped halving while below 1e-10 relative).
      const bool conv = t[0] <= s.otol2 * s.obnorm2;
      const bool floor = !first && t[0] > 0.25 * s.oprev2 &&
                         t[0] <= fmax(1e-20, 1e4 * s.otol2) * s.obnorm2;
      s.oprev2 = t[0];
      if (!isfinite(t[0]) || !isfinite(t[1])) {
        s.ostatus = ST_FAIL;
        s.status = ST_DONE;
        atomicSub(cnt, 1);
      } else if (conv || floor) {
        s.ostatus = ST_DONE;
        s.status = ST_DONE;
        atomicSub(cnt, 1);
      } else if (s.orefs >= s.omaxit) {
        s.ostatus = ST_FAIL;
        s.status = ST_DONE;
        atomicSub(cnt, 1);
      } else {
        s.orefs += 1;
        s.status = ST_ACTIVE;
      }
    }
  } else if (MODE == F_INNER_INIT) {
    float t[1];
    sum_partials<float, 1>(I, t);
    if (threadIdx.x == 0) {
      s.rho = t[0]; s.rho_old = 1.0; s.alpha = 1.0; s.omega = 1.0; s.beta = 0.0;
      s.rnorm2 = t[0]; s.bnorm2 = t[0]; s.iters = 0;
      s.best2 = t[0]; s.best_it = 0;
      s.status = ST_ACTIVE;
    }
  } else if (MODE == F_EXTRAP) {
    // The next solve starts at f + theta df: theta = (df_new, df_old) /
    // (df_old, df_old), clamped to [0, 1], the decay ratio of the increments,
    // once two increments are known (a start from f before). A trajectory
    // relaxing smoothly gets theta near 1; the fast transients of the first
    // steps, whose increments shrink or turn from step to step, get a small
    // theta instead of an overshooting start.
    double t[2];
    sum_partials<double, 2>(I, t);
    if (threadIdx.x == 0) {
      const double r = t[1] > 0.0 ? t[0] / t[1] : 0.0;
      s.theta = s.hist ? fmin(fmax(r, 0.0), 1.0) : 0.0;
      if (!isfinite(s.theta)) s.theta = 0.0;
      s.hist = 1;
    }
  } else {   // F_FINALIZE
    double t[4];
    sum_partials<double, 4>(I, t);
    __shared__ double sm[kFinThreads];
    double mm = DBL_MAX;
    const double *pmin = I.partial + (size_t)4 * I.nblk0;
    int b = threadIdx.x;
    for (; b + 7 * kFinThreads < I.nblk0; b += 8 * kFinThreads) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pmin[b + u * kFinThreads];
#pragma unroll
      for (int u = 0; u < 8; ++u) mm = fmin(mm, v[u]);
    }
    for (; b < I.nblk0; b += kFinThreads) mm = fmin(mm, pmin[b]);
    sm[threadIdx.x] = mm;
    __syncthreads();
    for (int o = kFinThreads / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double nf = t[3];
      s.true_rel = t[1] > 0.0 ? sqrt(t[0] / t[1]) : sqrt(t[0]);
      s.nonfin = (int)nf;
      s.clamps = (I.noclamp || nf > 0.0) ? 0 : (int)t[2];
      s.minf = sm[0];
      // The step is accepted when the solve ended without failing, x is
      // finite and the true residual meets the tolerance, or the rounding
      // floor the stall rule accepts (1e-10 relative) if that is larger.
      const bool failed = I.prec == SG_MIXED ? ost == ST_FAIL : st == ST_FAIL;
      const bool res_ok = t[0] <= fmax(s.otol2, 1e-20) * t[1];
      s.ostatus = (!failed && nf == 0.0 && res_ok) ? ST_DONE : ST_FAIL;
      // the commit that follows an accepted step records one more increment
      s.phase = 1;
    }
  }
}

// (nx,ny) i-major (the reference's Xtensor2d) <-> device cell order, for the
// host boundary
__global__ void __launch_bounds__(kBlock) k_to_korder(const double *__restrict__ src,
                                                     double *__restrict__ dst, int nx, int ny) {
  const long long K = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (K >= (long long)nx * ny) return;
  const int j = (int)(K / nx), i = xcol((int)(K - (long long)j * nx), half_x(nx));
  dst[K] = src[(size_t)i * ny + j];
}
// Count of the non-finite values of src[0 .. n) added to *cnt (validation of
// host arrays uploaded to a staging buffer before they replace device state).
__global__ void __launch_bounds__(kBlock) k_count_nonfinite(const double *__restrict__ src,
                                                           long long n, int *__restrict__ cnt) {
  int bad = 0;
  for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < n;
       q += (long long)gridDim.x * kBlock)
    bad += isfinite(src[q]) ? 0 : 1;
  int v[1] = {bad};
  block_sum<int, 1>(v);
  if (threadIdx.x == 0 && v[0]) atomicAdd(cnt, v[0]);
}
__global__ void k_zero_int(int *__restrict__ p, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}

__global__ void __launch_bounds__(kBlock) k_to_imajor(const double *__restrict__ src,
                                                     double *__restrict__ dst, int nx, int ny) {
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (q >= (long long)nx * ny) return;
  const int i = (int)(q / ny), j = (int)(q - (long long)i * ny);
  dst[q] = src[cell_slot(i, j, nx)];
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

void launch_line_factor(hipStream_t s, const InstDev *d, const LineGrp *m, int ngroups, int l,
                        int dir, int shape, bool mixed) {
  if (!ngroups) return;
#define SG_FAC(T, DIR, SH) k_line_factor<T, DIR, SH><<<ngroups, tile_threads(SH), 0, s>>>(d, m, l)
#define SG_FAC_S(T, DIR)                                   \
  do {                                                     \
    if (shape == 0) SG_FAC(T, DIR, 0);                     \
    else if (shape == 1) SG_FAC(T, DIR, 1);                \
    else if (shape == 2) SG_FAC(T, DIR, 2);                \
    else if (shape == 3) SG_FAC(T, DIR, 3);                \
    else SG_FAC(T, DIR, 4);                                \
  } while (0)
  if (mixed) { if (dir == 0) SG_FAC_S(float, 0); else SG_FAC_S(float, 1); }
  else { if (dir == 0) SG_FAC_S(double, 0); else SG_FAC_S(double, 1); }
#undef SG_FAC_S
#undef SG_FAC
}

template <typename T, int SH>
static void launch_line_s(hipStream_t s, const InstDev *d, const LineGrp *m, int ngroups, int l,
                          int dir, bool zero, int xid, int bid) {
  const bool nine = l > 0;
#define SG_LINE(NB, DIR, Z) \
  k_line<T, NB, DIR, Z, SH><<<ngroups, tile_threads(SH), 0, s>>>(d, m, l, xid, bid)
  if (nine) {
    if (dir == 0) { if (zero) SG_LINE(9, 0, true); else SG_LINE(9, 0, false); }
    else { if (zero) SG_LINE(9, 1, true); else SG_LINE(9, 1, false); }
  } else {
    if (dir == 0) { if (zero) SG_LINE(5, 0, true); else SG_LINE(5, 0, false); }
    else { if (zero) SG_LINE(5, 1, true); else SG_LINE(5, 1, false); }
  }
#undef SG_LINE
}

template <typename T>
void launch_line(hipStream_t s, const InstDev *d, const LineGrp *m, int ngroups, int l, int dir,
                 int shape, bool zero, int xid, int bid) {
  if (!ngroups) return;
  if (shape == 0) launch_line_s<T, 0>(s, d, m, ngroups, l, dir, zero, xid, bid);
  else if (shape == 1) launch_line_s<T, 1>(s, d, m, ngroups, l, dir, zero, xid, bid);
  else if (shape == 2) launch_line_s<T, 2>(s, d, m, ngroups, l, dir, zero, xid, bid);
  else if (shape == 3) launch_line_s<T, 3>(s, d, m, ngroups, l, dir, zero, xid, bid);
  else launch_line_s<T, 4>(s, d, m, ngroups, l, dir, zero, xid, bid);
}
template void launch_line<double>(hipStream_t, const InstDev *, const LineGrp *, int, int, int,
                                  int, bool, int, int);
template void launch_line<float>(hipStream_t, const InstDev *, const LineGrp *, int, int, int,
                                 int, bool, int, int);

template <typename T>
void launch_line_shapes(hipStream_t s, const InstDev *d, const LineGrp *m, int na, int nb,
                        int ngroups, int l, int dir, bool zero, int xid, int bid) {
  if (!ngroups) return;
  if (dir == 0) {
    if (zero) k_line_shapes<T, 9, 0, true, 2, 1, 0><<<ngroups, 256, 0, s>>>(d, m, na, nb, l, xid, bid);
    else k_line_shapes<T, 9, 0, false, 2, 1, 0><<<ngroups, 256, 0, s>>>(d, m, na, nb, l, xid, bid);
  } else {
    // columns: shapes 4, 0 (FP32 V-cycles) or 1, 0 (FP64: tile_shape gives
    // their columns 16 x 64 tiles)
    constexpr int SA = sizeof(T) == 4 ? 4 : 1;
    if (zero) k_line_shapes<T, 9, 1, true, SA, 0, -1><<<ngroups, 256, 0, s>>>(d, m, na, nb, l, xid, bid);
    else k_line_shapes<T, 9, 1, false, SA, 0, -1><<<ngroups, 256, 0, s>>>(d, m, na, nb, l, xid, bid);
  }
}
template void launch_line_shapes<double>(hipStream_t, const InstDev *, const LineGrp *, int, int,
                                         int, int, int, bool, int, int);
template void launch_line_shapes<float>(hipStream_t, const InstDev *, const LineGrp *, int, int,
                                        int, int, int, bool, int, int);

template <typename T>
void launch_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, int xid,
                     int bid) {
  if (!nb) return;
  if (l == 0) k_residual<T, 5><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
  else k_residual<T, 9><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
}
template void launch_residual<double>(hipStream_t, const InstDev *, const BlkMap *, int, int, int,
                                      int);
template void launch_residual<float>(hipStream_t, const InstDev *, const BlkMap *, int, int, int,
                                     int);

template <typename T>
void launch_restrict_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l,
                              int xid, int bid) {
  if (!nb) return;
  if (l == 0) k_restrict_residual<T, 5><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
  else k_restrict_residual<T, 9><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
}
template void launch_restrict_residual<double>(hipStream_t, const InstDev *, const BlkMap *, int,
                                               int, int, int);
template void launch_restrict_residual<float>(hipStream_t, const InstDev *, const BlkMap *, int,
                                              int, int, int);

template <typename T>
void launch_restrict(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool from_b,
                     int xid, int bid) {
  if (!nb) return;
  if (from_b) k_restrict<T, true><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
  else k_restrict<T, false><<<nb, kBlock, 0, s>>>(d, m, l, xid, bid);
}
template void launch_restrict<double>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                      int, int);
template void launch_restrict<float>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                     int, int);

template <typename T>
void launch_prolong(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool zero,
                    int xid, int rows) {
  if (!nb) return;
#define SG_PRO(Z, R) k_prolong<T, Z, R><<<(nb + kEnt - 1) / kEnt, kBlock, 0, s>>>(d, m, l, xid, nb)
  if (zero) { if (rows == 1) SG_PRO(true, 1); else if (rows == 2) SG_PRO(true, 2); else SG_PRO(true, 0); }
  else { if (rows == 1) SG_PRO(false, 1); else if (rows == 2) SG_PRO(false, 2); else SG_PRO(false, 0); }
#undef SG_PRO
}
template void launch_prolong<double>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                     int, int);
template void launch_prolong<float>(hipStream_t, const InstDev *, const BlkMap *, int, int, bool,
                                    int, int);

template <typename T, bool PC>
static void launch_bicg_t(hipStream_t s, int which, const InstDev *d, const BlkMap *m, int nb,
                          bool xstate) {
  switch (which) {
    case 0: k_bicg_p<T, PC><<<(nb + kLightBlocks - 1) / kLightBlocks, kBlock, 0, s>>>(d, m, nb); break;
    case 1: k_bicg_v<T, PC><<<nb, kBlock, 0, s>>>(d, m); break;
    case 2: k_bicg_s<T, PC><<<(nb + kLightBlocks - 1) / kLightBlocks, kBlock, 0, s>>>(d, m, nb); break;
    case 3: k_bicg_t<T, PC><<<nb, kBlock, 0, s>>>(d, m); break;
    default:
      if (xstate) k_bicg_x<T, true, PC><<<nb, kBlock, 0, s>>>(d, m);
      else k_bicg_x<T, false, PC><<<nb, kBlock, 0, s>>>(d, m);
  }
}

template <typename T>
void launch_bicg(hipStream_t s, int which, const InstDev *d, const BlkMap *m, int nb, bool xstate,
                 bool pc) {
  if (!nb) return;
  if (pc) launch_bicg_t<double, true>(s, which, d, m, nb, xstate);
  else launch_bicg_t<T, false>(s, which, d, m, nb, xstate);
}
template void launch_bicg<double>(hipStream_t, int, const InstDev *, const BlkMap *, int, bool,
                                  bool);
template void launch_bicg<float>(hipStream_t, int, const InstDev *, const BlkMap *, int, bool,
                                 bool);

template <typename T>
void launch_finish(hipStream_t s, int mode, const InstDev *d, int n, int *cnt) {
  if (!n) return;
  switch (mode) {
    case F_INIT64: k_finish<T, F_INIT64><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_V: k_finish<T, F_V><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_S: k_finish<T, F_S><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_T: k_finish<T, F_T><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_X: k_finish<T, F_X><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_IR_RES: k_finish<T, F_IR_RES><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_INNER_INIT: k_finish<T, F_INNER_INIT><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    case F_FINALIZE: k_finish<T, F_FINALIZE><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
    default: k_finish<T, F_EXTRAP><<<n, kFinThreads, 0, s>>>(d, n, cnt); break;
  }
}
template void launch_finish<double>(hipStream_t, int, const InstDev *, int, int *);
template void launch_finish<float>(hipStream_t, int, const InstDev *, int, int *);

void launch_ir_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_ir_residual<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_ir_inner_init(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_ir_inner_init<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_ir_update(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_ir_update<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_finalize(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_finalize<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_commit(hipStream_t s, const InstDev *d, const BlkMap *m, int nb) {
  if (nb) k_commit<<<nb, kBlock, 0, s>>>(d, m);
}
void launch_empty(hipStream_t s, int *p) { k_zero_int<<<1, 64, 0, s>>>(p, 1); }
void launch_count_nonfinite(hipStream_t s, const double *src, long long n, int *cnt) {
  k_zero_int<<<1, 64, 0, s>>>(cnt, 1);
  const long long nb = std::min<long long>(nblocks(n, kBlock), 2048);
  if (n > 0) k_count_nonfinite<<<(unsigned)nb, kBlock, 0, s>>>(src, n, cnt);
}
void launch_to_korder(hipStream_t s, const double *src, double *dst, int nx, int ny) {
  k_to_korder<<<nblocks((long long)nx * ny, kBlock), kBlock, 0, s>>>(src, dst, nx, ny);
}
void launch_to_imajor(hipStream_t s, const double *src, double *dst, int nx, int ny) {
  k_to_imajor<<<nblocks((long long)nx * ny, kBlock), kBlock, 0, s>>>(src, dst, nx, ny);
}

}  // namespace sg
