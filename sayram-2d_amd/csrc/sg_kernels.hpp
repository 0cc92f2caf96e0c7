// sg_kernels.hpp -- launch wrappers of the gfx950 kernels (sg_kernels.hip).
#pragma once
#include "sg_internal.hpp"

namespace sg {

void launch_begin(hipStream_t s, const InstDev *d, int n, const int *part, int *active,
                  int *oactive, double tol2, double otol2, int maxit, int omaxit);
void launch_assemble(hipStream_t s, const InstDev *d, const BlkMap *m, int nb);
void launch_pweights(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool mixed);
void launch_rap(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool mixed);
void launch_line_factor(hipStream_t s, const InstDev *d, const LineGrp *m, int ngroups, int l,
                        int dir, int shape, bool mixed);
// shape: tile shape 0..3 of sg_internal.hpp (lines x steps per workgroup)
template <typename T>
void launch_line(hipStream_t s, const InstDev *d, const LineGrp *m, int ngroups, int l, int dir,
                 int shape, bool zero, int xid, int bid);
// one colour of a 9-point sweep over several tile shapes in one launch: m
// holds ngroups line groups, rows in shapes 2, 1, 0 and columns in shapes
// 4, 0 with T = float or 1, 0 with T = double (na groups of the first shape,
// then nb of the second, then the rest)
template <typename T>
void launch_line_shapes(hipStream_t s, const InstDev *d, const LineGrp *m, int na, int nb,
                        int ngroups, int l, int dir, bool zero, int xid, int bid);
template <typename T>
void launch_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, int xid,
                     int bid);
// the residual of the even rows formed and restricted in one pass (after an
// x-line pre-sweep the residual of the odd rows vanishes)
template <typename T>
void launch_restrict_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l,
                              int xid, int bid);
// from_b: restrict the level's right-hand side (x_l = 0) instead of its
// residual
template <typename T>
void launch_restrict(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool from_b,
                     int xid, int bid);
// rows: 0 every cell, 1 odd rows only, 2 odd columns only (the cells the first
// post-smoothing sweep reads)
template <typename T>
void launch_prolong(hipStream_t s, const InstDev *d, const BlkMap *m, int nb, int l, bool zero,
                    int xid, int rows);
// the fused coarse cycle of the instances in m (one workgroup each): post-
// smoothing in x and/or y (smoother bits 4 and 8), optionally after an x-line
// pre-sweep (bit 1); a y-line pre-sweep (bit 2) is not fused
// cap: the largest level a fused instance may have (fuse_cells); it selects
// the LDS layout of the kernel
template <typename T>
void launch_coarse_cycle(hipStream_t s, const InstDev *d, const FuseMap *m, int n, int xid, int bid,
                         int smoother, int cap);
// which: 0 p-update, 1 v = M ph, 2 s-update, 3 t = M sh, 4 x/r update.
// pc: FP64 Krylov vectors with FP32 preconditioned vectors (T must be double).
template <typename T>
void launch_bicg(hipStream_t s, int which, const InstDev *d, const BlkMap *m, int nb, bool xstate,
                 bool pc);
// reduction finish modes (one workgroup per instance)
enum { FIN_INIT64 = 0, FIN_V, FIN_S, FIN_T, FIN_X, FIN_IR_RES, FIN_INNER_INIT, FIN_FINALIZE,
       FIN_EXTRAP };
template <typename T>
void launch_finish(hipStream_t s, int mode, const InstDev *d, int n, int *cnt);
void launch_ir_residual(hipStream_t s, const InstDev *d, const BlkMap *m, int nb);
void launch_ir_inner_init(hipStream_t s, const InstDev *d, const BlkMap *m, int nb);
void launch_copy_counter(hipStream_t s, const int *src, int *dst);
void launch_ir_update(hipStream_t s, const InstDev *d, const BlkMap *m, int nb);
// finalize pass over the instances whose solve ended (true residual, clamp
// and non-finite counts, min x); k_commit writes f of the accepted ones
void launch_finalize(hipStream_t s, const InstDev *d, const BlkMap *m, int nb);
void launch_commit(hipStream_t s, const InstDev *d, const BlkMap *m, int nb);
// a one-workgroup kernel that zeroes *p (the dispatch whose HIP-event bracket
// measures the timing overhead)
void launch_empty(hipStream_t s, int *p);
// *cnt = number of non-finite values in src[0 .. n)
void launch_count_nonfinite(hipStream_t s, const double *src, long long n, int *cnt);
void launch_to_korder(hipStream_t s, const double *src, double *dst, int nx, int ny);
void launch_to_imajor(hipStream_t s, const double *src, double *dst, int nx, int ny);

}  // namespace sg
