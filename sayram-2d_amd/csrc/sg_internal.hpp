// sg_internal.hpp -- device-side data structures of the gfx950 PPFV step.
//
// One instance (one Solver in the reference, source/Solver.h:18-82) lives in
// one device slab. Cell arrays are stored in system order K = j*nx + i
// (source/Mesh.h:66-68), so the 5-point operator is 5 coalesced bands and the
// multigrid levels are 9 coalesced bands. Many instances of different sizes are
// processed by the same launches through per-level block/line maps.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sg {

constexpr int kMaxLev = 16;
constexpr int kBlock = 256;     // cells per workgroup in cell-parallel kernels
constexpr int kLineBlock = 64;  // lines per workgroup in line kernels (one wave)
constexpr int kLineThreads = 256;   // threads of a line-solve workgroup
constexpr int kTileElems = 1024;    // elements of one LDS tile (lines x steps), 4 per thread
constexpr int kNumTileShapes = 3;   // tile steps S = 16, 64, 256 (lines 64, 16, 4)
__host__ __device__ constexpr int tile_steps(int sc) { return sc == 0 ? 16 : sc == 1 ? 64 : 256; }

// band slots; level 0 uses the first five (the operator is 5-point), coarse
// Galerkin levels use all nine.
enum Band { BC = 0, BW = 1, BE = 2, BS = 3, BN = 4, BSW = 5, BSE = 6, BNW = 7, BNE = 8 };

// Krylov vector slots (precision of the Krylov iteration).
enum Vec { VR = 0, VRH, VP, VV, VS, VT, VPH, VSH, VXK, kNumVec };

// solver status per instance
enum Status : int { ST_ACTIVE = 0, ST_HALF = 1, ST_DONE = 2, ST_FAIL = 3, ST_IDLE = 4 };

struct KScal {          // per-instance Krylov scalars, FP64 storage
  double rho, rho_old, alpha, omega, beta;
  double bnorm2, rnorm2, tol2;
  double scale;          // mixed: norm of the FP64 residual being corrected
  double onorm2, obnorm2, otol2;   // mixed: outer (FP64) residual norms
  double oprev2;         // mixed: residual norm^2 of the previous refinement
  double best2;          // smallest Krylov residual norm^2 of this solve
  int best_it, pad2;     // iteration at which it was reached
  double true_rel;       // finalize: ||R - M f|| / ||R||
  double minf;           // finalize: min f before clamping
  int iters, status;     // Krylov iteration count and status of this solve
  int ostatus, orefs;    // mixed: outer refinement status and count
  int maxit, clamps;
  int omaxit, part;
  int titers, pad;      // Krylov iterations summed over the refinements of a step
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
  int mixed;
  double dt;
  // geometry (source/Mesh.cc:36-65) and per-vertex interpolation weights of
  // fill_vertex_from_cells (source/Solver.cc:326-381), computed once
  const double *xe, *ye;
  const double *wx0, *wx1, *wy0, *wy1;
  const double *bcv[4];
  // Equation inputs, K order: G, Lambda = G*D (source/Solver.cc:57-65), 1/tau
  const double *G, *Lxx, *Lxy, *Lyy, *itau;
  // FP64 state
  double *f, *R, *x, *r64;
  double *band64[5];
  void *kv[kNumVec];          // Krylov vectors (double, or float in MIXED)
  double *partial;            // 4 * nblk0
  unsigned int *ticket;       // 2 tickets
  KScal *sc;
  Lev lev[kMaxLev];
};

struct BlkMap { int inst; int k0; };      // cell-parallel: 256 cells of one instance
struct LineMap { int inst; int line; };   // line-parallel: one row or column
struct LineGrp { int inst; int first; int count; int step; };   // lines first + k*step, k < count

}  // namespace sg
