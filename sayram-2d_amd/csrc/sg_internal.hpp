// sg_internal.hpp -- device-side data structures of the gfx950 PPFV step.
//
// One instance (one Solver in the reference, source/Solver.h:18-82) lives in
// one device slab. The 5-point operator is 5 bands and the multigrid levels are
// 9 bands, one value per cell each. Many instances of different sizes are
// processed by the same launches through per-level block/line maps.
//
// Device cell order. The reference numbers cells K = j*nx + i
// (source/Mesh.h:66-68). On the device every cell array of a level stores row
// j as its even-i cells followed by its odd-i cells:
//   slot(i, j) = j*nx + (i even ? i/2 : hx + i/2),   hx = (nx + 1) / 2.
// The zebra column sweeps solve every other column; in this order the columns
// of one colour are adjacent in memory, so their loads use whole cache lines
// (in the plain order every fetched line is half wasted). Rows stay two
// contiguous runs, and the coarse row J (the even points of fine row 2J) is
// exactly the first half of the fine row. Host arrays keep the reference's
// layouts; the conversion happens at upload and download.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sg {

constexpr int kMaxLev = 16;
// partial sums per level-0 workgroup (the finalize pass uses all five)
constexpr int kNumPartials = 5;
// Cells per instance: every per-cell array then stays below 2 GiB, so kernels
// may address it with 32-bit byte offsets from its base.
constexpr long long kMaxCells = 1LL << 28;
constexpr int kBlock = 256;     // cells per workgroup in cell-parallel kernels
constexpr int kLightBlocks = 4; // map entries per workgroup of the lightest vector kernels
constexpr int kEnt = 4;         // map entries per workgroup of the multi-entry cell kernels
// Tile shapes of the line solves: L lines x S steps staged through LDS by a
// workgroup of NT threads, each thread owning 4 consecutive steps of a line in
// the solve:
//   0: 64 x 16 (256 threads)  short lines
//   1: 16 x 64 (256 threads)  medium rows; long columns of FP64 9-point levels
//   2:  4 x 256 (256 threads) long rows (a row is contiguous in memory)
//   3: 16 x 256 (1024 threads) long columns (16 columns of one colour are 64
//      contiguous bytes per row; 4x fewer serial tiles than shape 1)
//   4: 32 x 32 (256 threads) medium columns (32 columns of one colour are 128
//      contiguous bytes per row, so rows at any offset waste less of each
//      fetched line)
constexpr int kNumTileShapes = 5;
__host__ __device__ constexpr int tile_steps(int sh) {
  return sh == 0 ? 16 : sh == 1 ? 64 : sh == 4 ? 32 : 256;
}
__host__ __device__ constexpr int tile_lines(int sh) {
  return sh == 0 ? 64 : sh == 1 ? 16 : sh == 2 ? 4 : sh == 3 ? 16 : 32;
}
__host__ __device__ constexpr int tile_threads(int sh) { return sh == 3 ? 1024 : 256; }
// Rank of a shape by the length of the lines it takes (the longest lines are
// the critical path of a sweep).
__host__ __device__ constexpr int shape_rank(int sh) { return sh == 4 ? 1 : sh; }
// Tiles of the eliminated right-hand side d that a line sweep keeps in LDS
// (the last ones of each line; earlier tiles are parked in x and read back by
// the backward pass), for values of tbytes bytes and an nb-point operator.
// Sized so that the LDS of a workgroup does not lower the occupancy its
// registers allow, except for the 5-point 32 x 32 tiles: 9 tiles (288 steps)
// at 3 workgroups per CU measured faster than 6 at 4 (fewer parked tiles of the
// 512-cell columns). The 5-point long-line shapes (2, 3) hold one tile more
// than the 9-point ones, whose registers allow fewer workgroups.
__host__ __device__ constexpr int ring_tiles(int sh, int tbytes, int nb) {
  if (tbytes != 4) return sh == 3 ? 1 : 2;
  switch (sh) {
    case 0: return 2;
    case 2: return nb == 9 ? 4 : 5;
    case 3: return nb == 9 ? 6 : 7;
    case 4: return nb == 9 ? 8 : 9;
    default: return 4;
  }
}

__host__ __device__ inline int half_x(int nx) { return (nx + 1) >> 1; }

// The prolongation fills only the lines the first post-smoothing sweep reads:
// odd rows before x-lines (1), odd columns before y-lines (2), else every
// point (0). `smoother` bits: 1/2 pre x/y, 4/8 post x/y.
__host__ __device__ inline int prolong_rows(int smoother) {
  return (smoother & 4) ? 1 : (smoother & 8) ? 2 : 0;
}
// After a pre-smoothing sweep of x-lines alone the odd rows (colour 1, solved
// last from the even rows' new values) satisfy their line equations, so the
// residual the restriction needs is that of the even rows only.
__host__ __device__ inline bool half_residual(int smoother) {
  return (smoother & 3) == 1;
}
// position of column i within its row
__host__ __device__ inline int xslot(int i, int hx) { return (i & 1) ? hx + (i >> 1) : (i >> 1); }
// column stored at position p of a row
__host__ __device__ inline int xcol(int p, int hx) { return p < hx ? 2 * p : 2 * (p - hx) + 1; }
__host__ __device__ inline size_t cell_slot(int i, int j, int nx) {
  return (size_t)j * nx + xslot(i, half_x(nx));
}

// band slots; level 0 uses the first five (the operator is 5-point), coarse
// Galerkin levels use all nine.
enum Band { BC = 0, BW = 1, BE = 2, BS = 3, BN = 4, BSW = 5, BSE = 6, BNW = 7, BNE = 8 };

// Krylov vector slots (precision of the Krylov iteration).
enum Vec { VR = 0, VRH, VP, VV, VS, VT, VPH, VSH, VXK, kNumVec };
// FP32 level-0 vectors of an FP64 Krylov solve with an FP32 V-cycle: the
// V-cycle input (p or s, scaled) and its outputs for p and s. Ids continue
// the Krylov slots, so a V-cycle addresses either kind by one index.
enum PcVec { PV_B = kNumVec, PV_P, PV_PH, PV_SH, PV_RH, kNumVecAll };

// solver status per instance
enum Status : int { ST_ACTIVE = 0, ST_HALF = 1, ST_DONE = 2, ST_FAIL = 3, ST_IDLE = 4 };

struct KScal {          // per-instance Krylov scalars, FP64 storage
  double rho, rho_old, alpha, omega, beta;
  double bnorm2, rnorm2, tol2;
  double scale;          // mixed: norm of the FP64 residual being corrected
  double pcs;            // FP32 V-cycle of an FP64 solve: scale of its input
                         // and output, the current residual norm
  double pcs_p;          // PC32: the scale the stored FP32 direction p was
                         // written with (pcs of the previous iteration)
  double onorm2, obnorm2, otol2;   // mixed: outer (FP64) residual norms
  double oprev2;         // mixed: residual norm^2 of the previous refinement
  double best2;          // smallest Krylov residual norm^2 of this solve
  int best_it;           // iteration at which it was reached
  int phase;             // step end: 0 solving, 1 finalized in this pass (f is
                         // written next), 2 finished in an earlier pass
  double true_rel;       // finalize: ||R - M x|| / ||R|| of the solution x
  double minf;           // finalize: min x (f before clamping)
  double theta;          // warm start: the next solve starts at f + theta df
  int iters, status;     // Krylov iteration count and status of this solve
  int ostatus, orefs;    // mixed: outer refinement status and count; every
                         // precision: ST_FAIL after finalize rejects the step
  int maxit, clamps;
  int omaxit, part;
  int titers;           // Krylov iterations summed over the refinements of a step
  int vcyc;             // level-0 V-cycles this instance ran in the step
  int nonfin;           // finalize: cells whose x is not finite
  int hist;             // 1: df holds the increment of an accepted step
};

struct Lev {            // one multigrid level of one instance
  int nx, ny;
  double *band64[9];    // FP64 operator (level 0: the 5 assembled bands)
  void *band[9];        // V-cycle precision copy (aliases band64 in FP64 mode)
  void *pw[4];          // prolongation weights to the next coarser level
  void *linv[2], *lup[2];   // LU factors of the row (0) and column (1) line
                            // matrices: 1/pivot and the eliminated upper entry
  void *x, *b, *cp, *dp;
};

struct InstDev {
  int nx, ny, nlev, bcmask;   // bcmask bit s set: side s Dirichlet
  int nblk0;                  // level-0 workgroups of this instance
  int prec;                   // sg_precision of the instance
  int noclamp;                // 1: f = x at the end of a step (no clamp at 0)
  int warm;                   // 1: the solve starts at f + theta df, 0: at f
  double dt;
  // geometry (source/Mesh.cc:36-65) and per-vertex interpolation weights of
  // fill_vertex_from_cells (source/Solver.cc:326-381), computed once
  const double *xe, *ye;
  const double *wx0, *wx1, *wy0, *wy1;
  const double *bcv[4];
  // Equation inputs, device cell order: G, D, 1/tau (the assembly forms
  // Lambda = G*D, source/Solver.cc:57-65, as it reads them)
  const double *G, *Dxx, *Dxy, *Dyy, *itau;
  // FP64 state
  double *f, *R, *x, *r64;
  float *df;                  // f^n - f^(n-1) of the last accepted step
                              // (valid when KScal::hist is 1)
  double *band64[5];
  void *kv[kNumVec];          // Krylov vectors (double, or float in MIXED)
  float *pv[kNumVecAll - kNumVec];   // PC32: FP32 V-cycle inputs (s, p) and outputs,
                                     // shadow residual
  double *partial;            // kNumPartials * nblk0
  unsigned int *ticket;       // 2 tickets
  KScal *sc;
  Lev lev[kMaxLev];
};

struct BlkMap { int inst; int k0; };      // cell-parallel: 256 cells of one instance
struct LineGrp { int inst; int first; int count; int step; };   // lines first + k*step, k < count
struct FuseMap { int inst; int lev; };    // an instance whose levels from lev down run fused

// The coarse end of a V-cycle runs in one workgroup per instance from the
// first level with at most fuse_cells cells (the fused coarse cycle; default
// kFuseCells).
constexpr int kFuseCells = 1024;
// the largest fused level an FP32 V-cycle can keep in LDS
constexpr int kFuseCells32 = 4096;
// Threads of the per-instance finish kernels (sums of the partials).
constexpr int kFinThreads = 1024;
constexpr int kFuseThreads = 1024;

}  // namespace sg
