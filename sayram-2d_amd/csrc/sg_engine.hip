// sg_engine.hip -- host orchestration of the gfx950 PPFV step and the C ABI
// declared in include/sayram_gpu.h.
//
// A batch owns one HIP stream, a table of instances (one device slab each) and
// the per-level block/line maps the batched kernels walk. A single solver
// (the reference's Solver, source/Solver.h:18-82) is a batch of one.
#include "sayram_gpu.h"
#include "sg_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

namespace sg {

static thread_local std::string g_err;

// Stream capture and host-blocking HIP calls. A batch steps its two instance
// groups on two host threads, and callers may drive several batches from
// several threads, so one thread may capture a graph while another allocates,
// copies synchronously or waits on a stream. Captures hold this lock
// exclusively; host-blocking calls hold it shared, so no such call runs while
// any thread is capturing.
static std::shared_mutex g_capture_mu;
using HostLock = std::shared_lock<std::shared_mutex>;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(SG_ERR_HIP, std::string(#call) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

static size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

struct Dims { int nx, ny; };

// Coarsen both directions (ceil halving) until one of them is 1; the last
// level is then solved exactly by one line solve.
static std::vector<Dims> level_dims(int nx, int ny) {
  std::vector<Dims> v{{nx, ny}};
  while (std::min(nx, ny) > 1) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
    v.push_back({nx, ny});
  }
  return v;
}

// Assigns every device pointer of an instance inside one slab starting at
// `base` (0 to size it); returns the slab size in bytes.
static size_t carve(uintptr_t base, int nx, int ny, int prec, InstDev &I) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    void *p = reinterpret_cast<void *>(base + off);
    off += a256(bytes);
    return p;
  };
  const size_t N = (size_t)nx * ny;
  const size_t D = sizeof(double);
  const bool mixed = prec == SG_MIXED;    // FP32 Krylov inside FP64 refinement
  const bool h32 = prec != SG_FP64;       // FP32 multigrid hierarchy
  const size_t T = h32 ? sizeof(float) : sizeof(double);     // hierarchy
  const size_t KT = mixed ? sizeof(float) : sizeof(double);  // Krylov vectors
  const std::vector<Dims> dims = level_dims(nx, ny);
  I.nx = nx; I.ny = ny; I.nlev = (int)dims.size();
  I.nblk0 = (int)((N + kBlock - 1) / kBlock);
  I.prec = prec;
  I.G = (double *)take(N * D); I.Dxx = (double *)take(N * D); I.Dxy = (double *)take(N * D);
  I.Dyy = (double *)take(N * D); I.itau = (double *)take(N * D);
  I.f = (double *)take(N * D); I.R = (double *)take(N * D); I.x = (double *)take(N * D);
  for (int b = 0; b < 5; ++b) I.band64[b] = (double *)take(N * D);
  I.r64 = mixed ? (double *)take(N * D) : nullptr;
  I.df = (float *)take(N * sizeof(float));
  I.xe = (double *)take((nx + 1) * D); I.ye = (double *)take((ny + 1) * D);
  I.wx0 = (double *)take((nx + 1) * D); I.wx1 = (double *)take((nx + 1) * D);
  I.wy0 = (double *)take((ny + 1) * D); I.wy1 = (double *)take((ny + 1) * D);
  I.bcv[0] = (double *)take((ny + 1) * D); I.bcv[1] = (double *)take((ny + 1) * D);
  I.bcv[2] = (double *)take((nx + 1) * D); I.bcv[3] = (double *)take((nx + 1) * D);
  I.partial = (double *)take((size_t)kNumPartials * I.nblk0 * D);
  // PC32 keeps its shadow residual and its search direction in FP32 (PV_RH,
  // PV_P) instead of VRH and VP
  for (int v = 0; v < kNumVec; ++v)
    I.kv[v] = (v == VXK && !mixed) || ((v == VRH || v == VP) && prec == SG_PC32) ? nullptr
                                                                                 : take(N * KT);
  for (int v = 0; v < kNumVecAll - kNumVec; ++v)
    I.pv[v] = prec == SG_PC32 ? (float *)take(N * sizeof(float)) : nullptr;
  for (int l = 0; l < kMaxLev; ++l) {
    Lev &L = I.lev[l];
    std::memset(&L, 0, sizeof(L));
    if (l >= I.nlev) continue;
    L.nx = dims[l].nx; L.ny = dims[l].ny;
    const size_t Nl = (size_t)L.nx * L.ny;
    if (l == 0) {
      for (int b = 0; b < 5; ++b) {
        L.band64[b] = I.band64[b];
        L.band[b] = h32 ? take(N * T) : (void *)I.band64[b];
      }
    } else {
      for (int b = 0; b < 9; ++b) {
        L.band64[b] = (double *)take(Nl * D);
        L.band[b] = h32 ? take(Nl * T) : (void *)L.band64[b];
      }
      L.x = take(Nl * T);
      L.b = take(Nl * T);
    }
    if (l + 1 < I.nlev)
      for (int q = 0; q < 4; ++q) L.pw[q] = take(Nl * T);
    for (int d = 0; d < 2; ++d) {
      L.linv[d] = take(Nl * T);
      L.lup[d] = take(Nl * T);
    }
    L.cp = take(Nl * T);
    L.dp = nullptr;
  }
  return off;
}

struct Slot {
  bool used = false;
  int nx = 0, ny = 0;
  int prec = SG_FP64;
  int fuse_lev = 0;       // first level of the fused coarse cycle (nlev: none)
  void *slab = nullptr;
  size_t bytes = 0;
  InstDev dev{};
  int64_t steps = 0;
  int64_t target = 0;
  double dt = 0;
  sg_step_stats st{};
};

// Setup maps cover every level of every instance (assembly, Galerkin chain,
// line factors). The cycle maps leave out the levels an instance runs in the
// fused coarse cycle. The factor maps gfac hold the line groups of both
// colours (a factorisation does not depend on x, so both colours of a
// direction and tile shape go in one launch).
struct Maps {
  std::vector<BlkMap> cmap[kMaxLev], pmap[kMaxLev];
  std::vector<LineGrp> gfac[kMaxLev][2][kNumTileShapes], gcx[kMaxLev][kNumTileShapes],
      gcy[kMaxLev][kNumTileShapes];
  // cycle: restriction into level l, prolongation into level l, sweeps and
  // coarsest solves, and every instance with a fused coarse cycle
  std::vector<BlkMap> rmap[kMaxLev], qmap[kMaxLev];
  std::vector<LineGrp> sw[kMaxLev][2][2][kNumTileShapes], swx[kMaxLev][kNumTileShapes],
      swy[kMaxLev][kNumTileShapes];
  // the 256-thread shapes of one colour's sweep of a 9-point level, in one
  // map (rows: shapes 2, 1, 0; columns: 4, 0, or 1, 0 with FP64 V-cycles),
  // the first two shapes' group counts, and whether the map and the
  // 1024-thread shape 3 hold every group of the colour (else the shapes are
  // launched one by one)
  std::vector<LineGrp> swm[kMaxLev][2][2];
  int swm_na[kMaxLev][2][2], swm_nb[kMaxLev][2][2];
  bool swm_all[kMaxLev][2][2];
  std::vector<FuseMap> fmap;
};

// Tile shape of a line solve by line length and direction: short lines in
// 64-line x 16-step tiles, longer rows in 16 x 64, longer columns in 32 x 32
// (32 columns of one colour read 128 contiguous bytes per row), long rows in
// 4 x 256 (rows stay coalesced along the line), long columns in 16 x 256 with
// 1024 threads (16 columns per tile keep the accesses across the lines
// coalesced with 4x fewer serial tiles). FP64 V-cycles keep the long columns
// of the 9-point levels on 16 x 64 tiles, whose loads fit the registers of
// 1024-thread tiles only in FP32, and their medium columns too (32 x 32 FP64
// tiles need twice the registers and keep half the steps on chip). It depends
// only on the instance, its level and the batch's precision, so results do
// not depend on the rest of the batch.
static int tile_shape(const int (&lim)[4], int n, int dir, int l, bool fp64_vcycle) {
  if (n <= lim[2 * dir]) return 0;
  if (dir == 0) return n <= lim[1] ? 1 : 2;
  if (n <= lim[3]) return fp64_vcycle ? 1 : 4;
  return (fp64_vcycle && l > 0) ? 1 : 3;
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD
// b % 8), each with its own L2. Neighbouring column groups of one colour share
// the cache lines at their common edge (a group's run of each row starts at an
// arbitrary offset), so in each block of 8 C dispatch slots the chunk of C
// consecutive map entries C x .. C x + C - 1 goes to slots x, x + 8, ..: the
// same XCD, dispatched together, one fetch per shared line. The order of the
// chunks is kept. C = 4 (measured: 0.4% faster than pairs on the batched
// workload, equal to C = 8).
template <typename G> static void xcd_chunks(std::vector<G> &v) {
  constexpr int kXcd = 8, C = 4;
  std::vector<G> out(v);
  for (size_t s = 0; s + C * kXcd <= v.size(); s += C * kXcd)
    for (int x = 0; x < kXcd; ++x)
      for (int c = 0; c < C; ++c) out[s + c * kXcd + x] = v[s + C * x + c];
  v.swap(out);
}

// The cell-parallel maps (256 cells of one instance per entry, instances one
// after another, each in row order) are reordered so that XCD x takes one
// contiguous eighth of the map: workgroup w (on XCD w % 8) takes unit
// start(x) + w / 8, a unit being the `unit` consecutive entries one workgroup
// of the kernel takes. The blocks of neighbouring rows of an instance then
// run on one XCD and read their halo rows (stencils, restriction, Galerkin
// products) from its L2 instead of fetching them again. A last partial unit
// keeps its place at the end. No result depends on the order: every block
// writes its own cells and its own partial-sum slot.
static void xcd_blocks(std::vector<BlkMap> &v, int unit) {
  constexpr size_t kXcd = 8;
  const size_t nu = v.size() / unit;   // full units
  if (nu < 2 * kXcd) return;
  const size_t base = nu / kXcd, rem = nu % kXcd;
  std::vector<BlkMap> out(v);
  for (size_t w = 0; w < nu; ++w) {
    const size_t x = w % kXcd, j = w / kXcd;
    const size_t u = x * base + std::min(x, rem) + j;
    for (int e = 0; e < unit; ++e) out[w * unit + e] = v[u * unit + e];
  }
  v.swap(out);
}

struct DevMaps {   // offsets into one device buffer
  const BlkMap *cmap[kMaxLev], *pmap[kMaxLev];
  int ncmap[kMaxLev], npmap[kMaxLev];
  const LineGrp *gfac[kMaxLev][2][kNumTileShapes], *gcx[kMaxLev][kNumTileShapes],
      *gcy[kMaxLev][kNumTileShapes];
  int ngfac[kMaxLev][2][kNumTileShapes], ngcx[kMaxLev][kNumTileShapes],
      ngcy[kMaxLev][kNumTileShapes];
  const BlkMap *rmap[kMaxLev], *qmap[kMaxLev];
  int nrmap[kMaxLev], nqmap[kMaxLev];
  const LineGrp *sw[kMaxLev][2][2][kNumTileShapes], *swx[kMaxLev][kNumTileShapes],
      *swy[kMaxLev][kNumTileShapes];
  int nsw[kMaxLev][2][2][kNumTileShapes], nswx[kMaxLev][kNumTileShapes],
      nswy[kMaxLev][kNumTileShapes];
  const LineGrp *swm[kMaxLev][2][2];
  int nswm[kMaxLev][2][2], swm_na[kMaxLev][2][2], swm_nb[kMaxLev][2][2];
  bool swm_all[kMaxLev][2][2];
  const FuseMap *fmap;
  int nfmap;
  int maxlev;
};

// Kernel families timed by sg_batch_time_kernel (HIP events around every
// dispatch of the family): the assembly; the per-step set-up (prolongation
// weights, Galerkin products, line factorisations); the level-0 line sweeps;
// the line sweeps of coarser levels (and the exact solves of the coarsest);
// restriction; prolongation; the two SpMVs of BiCGSTAB; its vector updates;
// the FP64 residual of iterative refinement; the step-end finalize and
// commit; the fused coarse cycle; the per-instance reduction finishes.
static const char *kKernelNames[] = {"assemble",    "rap",         "line_l0",    "line_coarse",
                                     "restrict",    "prolong",     "spmv_v",     "spmv_t",
                                     "vec_update",  "ir_residual", "finalize",   "coarse_cycle",
                                     "finish"};
constexpr int kNumKernelIds = 13;
constexpr int kTimeAll = kNumKernelIds;   // sg_batch_time_kernel id: every family
// Krylov iterations the per-step graph holds at most (more are run in
// pipelined chunks after it)
constexpr int kHeadIters = 16;

struct Batch {
  sg_opts opts{};
  size_t budget = 0, in_use = 0;
  hipStream_t stream = nullptr;
  // Second stream (opts.streams == 2): the tile shapes of one finest-level
  // sweep solve disjoint lines, so the shape with the longest lines runs on
  // `side` beside the others on `stream`; `stream` waits for it before the
  // next sweep.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_lev[kMaxLev] = {};   // with `side`: level l's operator is built
  hipError_t fan_err = hipSuccess;   // first failure of the fork/join calls
  std::vector<Slot> slots;
  int capacity = 0;
  InstDev *d_insts = nullptr;   // compact table
  KScal *d_scal = nullptr;      // by slot
  unsigned *d_ticket = nullptr; // 2 per slot
  int *d_part = nullptr;        // by compact index
  int *d_cnt = nullptr;         // [0] active, [1] oactive
  int *h_cnt = nullptr;         // pinned
  int *h_pipe = nullptr;        // pinned, counters of the pipelined Krylov chunks
  int *h_part = nullptr;        // pinned, participation flags being uploaded
  int h_part_cap = 0;
  hipEvent_t ev_cnt[2] = {nullptr, nullptr};
  // Captured graphs by key: the per-step head (set-up, Krylov start, the first
  // iterations and a step-end pass) and chunks of iterations. They hold the map
  // and instance pointers and launch sizes of the current batch layout, so
  // every rebuild drops them.
  struct ChunkGraph { int key; hipGraphExec_t exec; };
  std::vector<ChunkGraph> graphs;
  void drop_graphs();
  int last_iters_max = 0;   // largest Krylov iteration count of the last step
  void *d_maps = nullptr;
  size_t d_maps_bytes = 0;
  DevMaps dm{};
  std::vector<int> order;       // compact index -> slot
  std::vector<int> part_host;
  bool dirty = true;
  double *d_scratch = nullptr;  // layout transposes and coefficient staging
  size_t scratch_elems = 0;
  KScal *h_scal = nullptr;      // pinned, by slot (capacity entries)
  // kernel timing: bit k of time_mask times family k
  int time_mask = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_ids;      // family of each recorded pair
  size_t ev_used = 0;
  double t_ms[kNumKernelIds] = {};
  int64_t t_launches[kNumKernelIds] = {};
  double t_bytes[kNumKernelIds] = {};   // algorithmic bytes of the timed launches
  void reset_time() {
    for (int k = 0; k < kNumKernelIds; ++k) { t_ms[k] = 0; t_launches[k] = 0; t_bytes[k] = 0; }
  }

  ~Batch();
  int init(const sg_opts *o, size_t budget);
  int ensure_capacity(int cap);
  int add(const sg_grid *g, const sg_coeffs *c, const sg_bc *bc, const double *f0, double dt,
          int *slot);
  int set_coeffs(Slot &s, const sg_coeffs *c);
  int upload_bc(Slot &s, int side, const double *vals);
  int remove(int slot);
  int rebuild();
  int fuse_level(const InstDev &I) const;
  int step(int nsteps, const int64_t *targets, int ntargets);
  int step_once();
  bool graphs_on() const { return opts.use_graph >= 0 && time_mask == 0; }
  // Runs what enqueue() enqueues on `stream`: replayed from the graph captured
  // under `key` (captured on first use), or launched eagerly.
  template <typename F> int run_graph(int key, F &&enqueue);
  void setup_kernels(bool start64);
  void end_pass();
  // KT: Krylov vectors, VT: V-cycle (KT double with VT float is SG_PC32)
  template <typename KT, typename VT> int krylov(int maxit, bool xstate, int it0, bool ends);
  template <typename KT, typename VT> int solve64();
  template <typename KT, typename VT> void iteration(bool xstate);
  void count_bytes(const std::vector<int> &part);
  template <typename T> void vcycle(int l, int xid, int bid);
  template <typename F> void fan_out(const int (&cnt)[kNumTileShapes], F &&launch);
  int read_counter(int which, int *out);
  int scratch(size_t n);
  int get_f(int slot, double *out);
  int get_f_static(int slot, double *out, double *buf);
  int set_f(int slot, const double *f, int64_t istep);
  int debug_assemble(int slot, double *bands, double *R);
  // timing helpers
  void tbegin(int id, hipStream_t s = nullptr);
  void tend(int id, hipStream_t s = nullptr);
  int collect_time();
};

Batch::~Batch() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &s : slots)
    if (s.slab) (void)hipFree(s.slab);
  (void)hipFree(d_insts); (void)hipFree(d_scal); (void)hipFree(d_ticket); (void)hipFree(d_part); (void)hipFree(d_cnt);
  (void)hipFree(d_maps); (void)hipFree(d_scratch);
  if (h_cnt) (void)hipHostFree(h_cnt);
  if (h_pipe) (void)hipHostFree(h_pipe);
  if (h_scal) (void)hipHostFree(h_scal);
  if (h_part) (void)hipHostFree(h_part);
  drop_graphs();
  for (auto ev : ev_cnt)
    if (ev) (void)hipEventDestroy(ev);
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  for (auto e : ev_lev)
    if (e) (void)hipEventDestroy(e);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

static void apply_defaults(sg_opts &o) {
  if (o.rtol <= 0) o.rtol = 1e-14;
  if (o.max_iters <= 0) o.max_iters = o.precision == SG_MIXED ? 12 : 200;
  // a loose inner solve and more refinements: on the batched workload 3e-3
  // takes 50.7 ms/step against 62.4 with 1e-5 (3.4 refinements against 2.5)
  if (o.inner_rtol <= 0) o.inner_rtol = 3e-3;
  if (o.max_inner <= 0) o.max_inner = 40;
  if (o.check_every <= 0) o.check_every = 1;
  // Default schedule: post-smoothing with x- and y-lines on the finest level,
  // and on the coarser levels an x-line pre-sweep before it. An explicit
  // smoother with coarse_smoother 0 applies to every level.
  const bool sm_default = o.smoother <= 0 || o.smoother > 15;
  if (sm_default) o.smoother = 12;
  if (o.coarse_smoother == 0) o.coarse_smoother = sm_default ? 13 : o.smoother;
  if (o.fuse_cells == 0) o.fuse_cells = kFuseCells;
  if (o.streams == 0) o.streams = 1;
  if (o.group_cells == 0) o.group_cells = 1 << 21;
  if (o.warm_start == 0) o.warm_start = 1;
  const int tiles[4] = {16, 160, 16, 768};
  for (int k = 0; k < 4; ++k) {
    if (o.line_tiles[k] <= 0) o.line_tiles[k] = tiles[k];
    if (o.coarse_tiles[k] <= 0) o.coarse_tiles[k] = tiles[k];
  }
}

int Batch::init(const sg_opts *o, size_t b) {
  if (o) opts = *o;
  if (opts.precision != SG_FP64 && opts.precision != SG_MIXED && opts.precision != SG_PC32)
    return fail(SG_ERR_ARG, "opts.precision must be SG_FP64, SG_MIXED or SG_PC32");
  apply_defaults(opts);
  if (opts.fuse_cells > (opts.precision == SG_FP64 ? kFuseCells : kFuseCells32))
    return fail(SG_ERR_ARG, opts.precision == SG_FP64
                                ? "opts.fuse_cells must be at most 1024 with FP64 V-cycles (the fused "
                                  "cycle keeps a level in LDS)"
                                : "opts.fuse_cells must be at most 4096 (the fused cycle keeps a level "
                                  "in LDS)");
  if (opts.streams != 1 && opts.streams != 2) return fail(SG_ERR_ARG, "opts.streams must be 0, 1 or 2");
  if (opts.warm_start < -1 || opts.warm_start > 1)
    return fail(SG_ERR_ARG, "opts.warm_start must be -1, 0 or 1");
  if (opts.coarse_smoother != opts.smoother && (opts.coarse_smoother & ~13 || !(opts.coarse_smoother & 12)))
    return fail(SG_ERR_ARG, "opts.coarse_smoother must be 4, 8 or 12 (post-smoothing), or one of "
                            "them plus 1 (an x-line pre-sweep)");
  budget = b;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SG_ERR_HIP, "no GPU visible: the PPFV step has no CPU fallback");
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (opts.streams == 2) {
    HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    for (auto &e : ev_lev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  HIPCHK(hipMalloc(&d_cnt, 2 * sizeof(int)));
  HIPCHK(hipHostMalloc((void **)&h_cnt, 2 * sizeof(int), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&h_pipe, 2 * sizeof(int), hipHostMallocDefault));
  return ensure_capacity(16);
}

int Batch::ensure_capacity(int cap) {
  if (cap <= capacity) return SG_OK;
  drop_graphs();
  int nc = std::max(cap, capacity * 2);
  KScal *ns = nullptr;
  unsigned *nt = nullptr;
  InstDev *ni = nullptr;
  int *np = nullptr;
  HIPCHK(hipMalloc(&ns, sizeof(KScal) * nc));
  HIPCHK(hipMalloc(&nt, sizeof(unsigned) * 2 * nc));
  HIPCHK(hipMalloc(&ni, sizeof(InstDev) * nc));
  HIPCHK(hipMalloc(&np, sizeof(int) * nc));
  HIPCHK(hipMemsetAsync(ns, 0, sizeof(KScal) * nc, stream));
  HIPCHK(hipMemsetAsync(nt, 0, sizeof(unsigned) * 2 * nc, stream));
  if (capacity) {
    HIPCHK(hipMemcpyAsync(ns, d_scal, sizeof(KScal) * capacity, hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemcpyAsync(nt, d_ticket, sizeof(unsigned) * 2 * capacity, hipMemcpyDeviceToDevice,
                          stream));
  }
  KScal *nh = nullptr;
  HIPCHK(hipHostMalloc((void **)&nh, sizeof(KScal) * nc, hipHostMallocDefault));
  std::memset(nh, 0, sizeof(KScal) * nc);
  HIPCHK(hipStreamSynchronize(stream));
  (void)hipFree(d_scal); (void)hipFree(d_ticket); (void)hipFree(d_insts); (void)hipFree(d_part);
  if (h_scal) (void)hipHostFree(h_scal);
  d_scal = ns; d_ticket = nt; d_insts = ni; d_part = np; h_scal = nh;
  capacity = nc;
  for (size_t k = 0; k < slots.size(); ++k) {
    slots[k].dev.sc = d_scal + k;
    slots[k].dev.ticket = d_ticket + 2 * k;
  }
  dirty = true;
  return SG_OK;
}

static int check_grid(const sg_grid *g) {
  if (!g || !g->x_edges || !g->y_edges) return fail(SG_ERR_ARG, "grid: null pointer");
  if (g->nx < 1 || g->ny < 1) return fail(SG_ERR_ARG, "grid: nx and ny must be >= 1");
  if ((long long)g->nx * g->ny > kMaxCells)
    return fail(SG_ERR_ARG, "grid: more than 2^28 cells in one instance");
  if (level_dims(g->nx, g->ny).size() > (size_t)kMaxLev)
    return fail(SG_ERR_ARG, "grid: too many multigrid levels");
  // Grid2D::validate, source/Grid2D.h:44-67
  for (int i = 0; i < g->nx; ++i)
    if (!(g->x_edges[i + 1] > g->x_edges[i]))
      return fail(SG_ERR_ARG, "Grid2D: x_edges must be strictly increasing at i=" + std::to_string(i));
  for (int j = 0; j < g->ny; ++j)
    if (!(g->y_edges[j + 1] > g->y_edges[j]))
      return fail(SG_ERR_ARG, "Grid2D: y_edges must be strictly increasing at j=" + std::to_string(j));
  return SG_OK;
}

// Creation needs G and D; inv_tau may be NULL (1/tau = 0,
// source/Equation.h:39-40). A refresh (sg_set_coeffs) takes any subset: NULL
// arrays keep their device values.
static int check_coeffs(const sg_coeffs *c, bool create) {
  if (!c) return fail(SG_ERR_ARG, "coeffs: null pointer");
  if (create && (!c->G || !c->Dxx || !c->Dxy || !c->Dyy))
    return fail(SG_ERR_ARG, "coeffs: null pointer");
  return SG_OK;
}

// Interpolation weights of the vertex rule (source/Solver.cc:326-381), evaluated
// exactly as written there (no fused multiply-add).
static void vertex_weights(const double *e, int n, std::vector<double> &w0, std::vector<double> &w1) {
#pragma clang fp contract(off)
  w0.assign(n + 1, 0.0);
  w1.assign(n + 1, 0.0);
  for (int i = 1; i < n; ++i) {
    const double v = e[i];
    const double cL = 0.5 * (e[i - 1] + e[i]);
    const double cR = 0.5 * (e[i] + e[i + 1]);
    const double den = (cR - cL);
    w0[i] = (cR - v) / den;
    w1[i] = (v - cL) / den;
  }
}

// The given arrays of one instance, in the (nx, ny) host layout, replace the
// device ones (device cell order). They are staged on the device and checked
// there for non-finite values before any of them is written, so a rejected
// call changes nothing; the layout change runs on the device, and the
// assembly forms Lambda = G * D (source/Solver.cc:57-65) as it reads them, so
// a refresh after Equation::update (source/Solver.cc:287-288) moves only the
// arrays that changed and runs no per-cell host loop.
int Batch::set_coeffs(Slot &s, const sg_coeffs *c) {
  const double *src[5] = {c->G, c->Dxx, c->Dxy, c->Dyy, c->inv_tau};
  double *dst[5] = {(double *)s.dev.G, (double *)s.dev.Dxx, (double *)s.dev.Dxy,
                    (double *)s.dev.Dyy, (double *)s.dev.itau};
  const size_t N = (size_t)s.nx * s.ny;
  int nsrc = 0;
  for (auto p : src) nsrc += p != nullptr;
  if (!nsrc) return SG_OK;
  int rc;
  if ((rc = scratch((size_t)nsrc * N))) return rc;
  for (int k = 0, q = 0; k < 5; ++k)
    if (src[k])
      HIPCHK(hipMemcpyAsync(d_scratch + (size_t)(q++) * N, src[k], N * 8, hipMemcpyHostToDevice,
                            stream));
  launch_count_nonfinite(stream, d_scratch, (long long)nsrc * N, d_cnt);
  HIPCHK(hipMemcpyAsync(h_cnt, d_cnt, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  if (h_cnt[0])
    return fail(SG_ERR_ARG, "coeffs: " + std::to_string(h_cnt[0]) +
                                " non-finite values of G, D or 1/tau");
  for (int k = 0, q = 0; k < 5; ++k)
    if (src[k]) launch_to_korder(stream, d_scratch + (size_t)(q++) * N, dst[k], s.nx, s.ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  return SG_OK;
}

int Batch::upload_bc(Slot &s, int side, const double *vals) {
  const int n = (side < 2) ? s.ny + 1 : s.nx + 1;
  if (!vals) return fail(SG_ERR_BC, "Dirichlet BC: missing value.");
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(vals[k])) return fail(SG_ERR_BC, "Dirichlet BC: non-finite value.");
  HIPCHK(hipMemcpyAsync((void *)s.dev.bcv[side], vals, n * 8, hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SG_OK;
}

int Batch::add(const sg_grid *g, const sg_coeffs *c, const sg_bc *bc, const double *f0, double dt,
               int *slot_out) {
  int rc;
  if ((rc = check_grid(g))) return rc;
  if ((rc = check_coeffs(c, true))) return rc;
  if (!bc || !f0 || !slot_out) return fail(SG_ERR_ARG, "bc, f0 and slot must not be null");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(SG_ERR_ARG, "dt must be positive");
  for (int s = 0; s < 4; ++s) {
    if (bc->type[s] != SG_DIRICHLET && bc->type[s] != SG_ZEROFLUX)
      return fail(SG_ERR_ARG, "bc type must be SG_DIRICHLET or SG_ZEROFLUX");
    if (bc->type[s] == SG_DIRICHLET && !bc->values[s])
      return fail(SG_ERR_BC, "Dirichlet BC: missing value.");
  }
  InstDev tmp{};
  const size_t bytes = carve(0, g->nx, g->ny, opts.precision, tmp);
  if (budget && in_use + bytes > budget)
    return fail(SG_ERR_NOMEM, "HBM budget exceeded: need " + std::to_string(bytes) +
                                  " more bytes, " + std::to_string(budget - in_use) + " free");
  int slot = -1;
  for (size_t k = 0; k < slots.size(); ++k)
    if (!slots[k].used) { slot = (int)k; break; }
  if (slot < 0) {
    slot = (int)slots.size();
    if ((rc = ensure_capacity(slot + 1))) return rc;
    slots.emplace_back();
  }
  Slot &s = slots[slot];
  void *slab = nullptr;
  if (hipMalloc(&slab, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SG_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  s = Slot{};
  s.nx = g->nx; s.ny = g->ny; s.prec = opts.precision; s.slab = slab; s.bytes = bytes; s.dt = dt;
  carve(reinterpret_cast<uintptr_t>(slab), g->nx, g->ny, opts.precision, s.dev);
  s.dev.dt = dt;
  s.dev.noclamp = opts.no_clamp ? 1 : 0;
  s.dev.warm = opts.warm_start > 0 ? 1 : 0;
  s.dev.sc = d_scal + slot;
  s.dev.ticket = d_ticket + 2 * slot;
  s.dev.bcmask = 0;
  for (int k = 0; k < 4; ++k)
    if (bc->type[k] == SG_DIRICHLET) s.dev.bcmask |= 1 << k;
  auto undo = [&](int code) {
    (void)hipFree(s.slab);
    s = Slot{};
    return code;
  };
  const int nx = g->nx, ny = g->ny;
  const size_t N = (size_t)nx * ny;
  if (hipMemcpy((void *)s.dev.xe, g->x_edges, (nx + 1) * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy((void *)s.dev.ye, g->y_edges, (ny + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
    return undo(fail(SG_ERR_HIP, "upload of the grid failed"));
  std::vector<double> w0, w1;
  vertex_weights(g->x_edges, nx, w0, w1);
  (void)hipMemcpy((void *)s.dev.wx0, w0.data(), (nx + 1) * 8, hipMemcpyHostToDevice);
  (void)hipMemcpy((void *)s.dev.wx1, w1.data(), (nx + 1) * 8, hipMemcpyHostToDevice);
  vertex_weights(g->y_edges, ny, w0, w1);
  (void)hipMemcpy((void *)s.dev.wy0, w0.data(), (ny + 1) * 8, hipMemcpyHostToDevice);
  (void)hipMemcpy((void *)s.dev.wy1, w1.data(), (ny + 1) * 8, hipMemcpyHostToDevice);
  if (!c->inv_tau && hipMemsetAsync((void *)s.dev.itau, 0, N * 8, stream) != hipSuccess)
    return undo(fail(SG_ERR_HIP, "clearing 1/tau failed"));
  if ((rc = set_coeffs(s, c))) return undo(rc);
  for (int k = 0; k < 4; ++k)
    if (bc->type[k] == SG_DIRICHLET && (rc = upload_bc(s, k, bc->values[k]))) return undo(rc);
  std::vector<double> fk(N);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double v = f0[(size_t)i * ny + j];
      if (!std::isfinite(v)) return undo(fail(SG_ERR_ARG, "f0: non-finite value"));
      fk[cell_slot(i, j, nx)] = v;
    }
  if (hipMemcpy(s.dev.f, fk.data(), N * 8, hipMemcpyHostToDevice) != hipSuccess)
    return undo(fail(SG_ERR_HIP, "upload of f0 failed"));
  // the scalars of a reused slot start from zero
  if (hipMemsetAsync(d_scal + slot, 0, sizeof(KScal), stream) != hipSuccess)
    return undo(fail(SG_ERR_HIP, "clearing the instance scalars failed"));
  s.used = true;
  in_use += bytes;
  dirty = true;
  *slot_out = slot;
  return SG_OK;
}

int Batch::remove(int slot) {
  if (slot < 0 || slot >= (int)slots.size() || !slots[slot].used)
    return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  HIPCHK(hipStreamSynchronize(stream));
  (void)hipFree(slots[slot].slab);
  in_use -= slots[slot].bytes;
  slots[slot] = Slot{};
  dirty = true;
  return SG_OK;
}

// First level of an instance's fused coarse cycle: its first level with at
// most fuse_cells cells; nlev (none) when fusion is off, when the schedule
// pre-smooths (the fused cycle implements post-smoothing) or when no level is
// that small. It depends only on the instance and the options.
int Batch::fuse_level(const InstDev &I) const {
  // the fused cycle runs post-smoothing, optionally after an x-line pre-sweep
  if (opts.fuse_cells < 0 || ((opts.smoother | opts.coarse_smoother) & 2)) return I.nlev;
  for (int l = 0; l < I.nlev; ++l)
    if ((long long)I.lev[l].nx * I.lev[l].ny <= opts.fuse_cells) return l;
  return I.nlev;
}

void Batch::drop_graphs() {
  for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
  graphs.clear();
}

int Batch::rebuild() {
  drop_graphs();
  Maps M{};
  order.clear();
  const bool f64 = opts.precision == SG_FP64;   // FP64 V-cycle
  // the lines the first post-smoothing sweep of level l reads
  auto prow_of = [&](int l) { return prolong_rows(l == 0 ? opts.smoother : opts.coarse_smoother); };
  std::vector<InstDev> table;
  int maxlev = 0;
  for (size_t k = 0; k < slots.size(); ++k) {
    if (!slots[k].used) continue;
    const int c = (int)order.size();
    order.push_back((int)k);
    const InstDev &I = slots[k].dev;
    table.push_back(I);
    maxlev = std::max(maxlev, I.nlev);
    const int lf = fuse_level(I);
    slots[k].fuse_lev = lf;
    if (lf < I.nlev) M.fmap.push_back({c, lf});
    for (int l = 0; l < I.nlev; ++l) {
      const int lx = I.lev[l].nx, ly = I.lev[l].ny, Nl = lx * ly;
      const bool coarsest = (l == I.nlev - 1);
      const int (&lim)[4] = l == 0 ? opts.line_tiles : opts.coarse_tiles;
      for (int k0 = 0; k0 < Nl; k0 += kBlock) {
        M.cmap[l].push_back({c, k0});
        if (!coarsest) M.pmap[l].push_back({c, k0});
        if (l >= 1 && l <= lf) M.rmap[l].push_back({c, k0});        // restricted into from l - 1
      }
      // prolonged into from l + 1: only the lines the first post-sweep reads
      const int prow = prow_of(l);
      const int nq = prow == 1 ? lx * (ly >> 1) : prow == 2 ? (lx - half_x(lx)) * ly : Nl;
      if (!coarsest && l < lf)
        for (int k0 = 0; k0 < nq; k0 += kBlock) M.qmap[l].push_back({c, k0});
      if (coarsest) {
        LineGrp g{c, 0, 1, 1};
        auto &set = ly == 1 ? M.gcx[l][tile_shape(lim, lx, 0, l, f64)]
                            : M.gcy[l][tile_shape(lim, ly, 1, l, f64)];
        auto &cyc = ly == 1 ? M.swx[l][tile_shape(lim, lx, 0, l, f64)]
                            : M.swy[l][tile_shape(lim, ly, 1, l, f64)];
        set.push_back(g);
        if (l < lf) cyc.push_back(g);
      } else {
        // groups of tile_lines(shape) lines of one colour
        for (int d = 0; d < 2; ++d) {
          const int nl = d == 0 ? ly : lx;          // number of lines
          const int sh = tile_shape(lim, d == 0 ? lx : ly, d, l, f64);
          const int per = tile_lines(sh);
          for (int p = 0; p < 2; ++p)
            for (int f = p; f < nl; f += 2 * per) {
              const int cntl = std::min(per, (nl - f + 1) / 2);
              M.gfac[l][d][sh].push_back({c, f, cntl, 2});
              if (l < lf) M.sw[l][d][p][sh].push_back({c, f, cntl, 2});
            }
        }
      }
    }
  }
  // Longest lines first: a group walks its lines tile by tile, so the groups
  // of the longest lines are the critical path of a sweep and must start at
  // once; the short groups of other instances fill the machine around them.
  // The column groups of one colour are neighbours in memory (one half of
  // each row of the split layout), and xcd_chunks puts 4 neighbouring groups
  // on one XCD; in the factor maps both colours of an instance follow one
  // another in the same way.
  for (int l = 0; l < kMaxLev; ++l)
    for (int sh = 0; sh < kNumTileShapes; ++sh)
      for (int d = 0; d < 2; ++d) {
        auto len = [&](const LineGrp &g) {
          return d == 0 ? table[g.inst].lev[l].nx : table[g.inst].lev[l].ny;
        };
        for (int p = 0; p < 3; ++p) {
          auto &w = p < 2 ? M.sw[l][d][p][sh] : M.gfac[l][d][sh];
          std::stable_sort(w.begin(), w.end(),
                           [&](const LineGrp &a, const LineGrp &b) { return len(a) > len(b); });
          if (d == 1) xcd_chunks(w);
        }
      }
  // merged sweep maps of the 9-point levels (the shapes keep their order:
  // each shape's groups longest first, column groups in XCD chunks, whose
  // members stay on one XCD at any offset in the map)
  const bool fp64_cycle = opts.precision == SG_FP64;
  for (int l = 1; l < kMaxLev; ++l)
    for (int d = 0; d < 2; ++d)
      for (int p = 0; p < 2; ++p) {
        // the shapes launch_line_shapes takes, in its order
        const int order[2][3] = {{2, 1, 0}, {fp64_cycle ? 1 : 4, 0, -1}};
        auto &v = M.swm[l][d][p];
        size_t all = 0;
        for (int sh = 0; sh < kNumTileShapes; ++sh) all += M.sw[l][d][p][sh].size();
        for (int k = 0; k < 3 && order[d][k] >= 0; ++k) {
          const auto &w = M.sw[l][d][p][order[d][k]];
          if (k == 0) M.swm_na[l][d][p] = (int)w.size();
          if (k == 1) M.swm_nb[l][d][p] = (int)w.size();
          v.insert(v.end(), w.begin(), w.end());
        }
        M.swm_all[l][d][p] = v.size() + M.sw[l][d][p][3].size() == all;
      }
  for (int l = 0; l < kMaxLev; ++l) {
    xcd_blocks(M.cmap[l], 1);
    xcd_blocks(M.pmap[l], 1);
    xcd_blocks(M.rmap[l], 1);
    xcd_blocks(M.qmap[l], kEnt);   // k_prolong takes kEnt entries per workgroup
  }
  // fused cycles with the most levels first: they are the critical path
  std::stable_sort(M.fmap.begin(), M.fmap.end(), [&](const FuseMap &a, const FuseMap &b) {
    return table[a.inst].nlev - a.lev > table[b.inst].nlev - b.lev;
  });
  // one device buffer for all maps
  size_t bytes = 0;
  auto sz = [&](size_t n, size_t e) { bytes += a256(n * e); };
  sz(M.fmap.size(), sizeof(FuseMap));
  for (int l = 0; l < kMaxLev; ++l) {
    sz(M.cmap[l].size(), sizeof(BlkMap)); sz(M.pmap[l].size(), sizeof(BlkMap));
    sz(M.rmap[l].size(), sizeof(BlkMap)); sz(M.qmap[l].size(), sizeof(BlkMap));
    for (int sh = 0; sh < kNumTileShapes; ++sh) {
      for (int d = 0; d < 2; ++d) {
        sz(M.gfac[l][d][sh].size(), sizeof(LineGrp));
        for (int p = 0; p < 2; ++p) sz(M.sw[l][d][p][sh].size(), sizeof(LineGrp));
      }
      sz(M.gcx[l][sh].size(), sizeof(LineGrp)); sz(M.gcy[l][sh].size(), sizeof(LineGrp));
      sz(M.swx[l][sh].size(), sizeof(LineGrp)); sz(M.swy[l][sh].size(), sizeof(LineGrp));
    }
    for (int d = 0; d < 2; ++d)
      for (int p = 0; p < 2; ++p) sz(M.swm[l][d][p].size(), sizeof(LineGrp));
  }
  HIPCHK(hipStreamSynchronize(stream));
  if (bytes > d_maps_bytes) {
    (void)hipFree(d_maps);
    d_maps = nullptr;
    HIPCHK(hipMalloc(&d_maps, bytes));
    d_maps_bytes = bytes;
  }
  std::vector<char> host(bytes ? bytes : 1);
  size_t off = 0;
  auto put = [&](const void *src, size_t n, size_t e) -> const void * {
    if (n) std::memcpy(host.data() + off, src, n * e);
    const void *p = static_cast<char *>(d_maps) + off;
    off += a256(n * e);
    return p;
  };
  DevMaps dmn{};
  dmn.fmap = (const FuseMap *)put(M.fmap.data(), M.fmap.size(), sizeof(FuseMap));
  dmn.nfmap = (int)M.fmap.size();
  for (int l = 0; l < kMaxLev; ++l) {
    dmn.cmap[l] = (const BlkMap *)put(M.cmap[l].data(), M.cmap[l].size(), sizeof(BlkMap));
    dmn.ncmap[l] = (int)M.cmap[l].size();
    dmn.pmap[l] = (const BlkMap *)put(M.pmap[l].data(), M.pmap[l].size(), sizeof(BlkMap));
    dmn.npmap[l] = (int)M.pmap[l].size();
    dmn.rmap[l] = (const BlkMap *)put(M.rmap[l].data(), M.rmap[l].size(), sizeof(BlkMap));
    dmn.nrmap[l] = (int)M.rmap[l].size();
    dmn.qmap[l] = (const BlkMap *)put(M.qmap[l].data(), M.qmap[l].size(), sizeof(BlkMap));
    dmn.nqmap[l] = (int)M.qmap[l].size();
    for (int sh = 0; sh < kNumTileShapes; ++sh) {
      for (int d = 0; d < 2; ++d) {
        auto &v = M.gfac[l][d][sh];
        dmn.gfac[l][d][sh] = (const LineGrp *)put(v.data(), v.size(), sizeof(LineGrp));
        dmn.ngfac[l][d][sh] = (int)v.size();
        for (int p = 0; p < 2; ++p) {
          auto &w = M.sw[l][d][p][sh];
          dmn.sw[l][d][p][sh] = (const LineGrp *)put(w.data(), w.size(), sizeof(LineGrp));
          dmn.nsw[l][d][p][sh] = (int)w.size();
        }
      }
      dmn.gcx[l][sh] = (const LineGrp *)put(M.gcx[l][sh].data(), M.gcx[l][sh].size(), sizeof(LineGrp));
      dmn.ngcx[l][sh] = (int)M.gcx[l][sh].size();
      dmn.gcy[l][sh] = (const LineGrp *)put(M.gcy[l][sh].data(), M.gcy[l][sh].size(), sizeof(LineGrp));
      dmn.ngcy[l][sh] = (int)M.gcy[l][sh].size();
      dmn.swx[l][sh] = (const LineGrp *)put(M.swx[l][sh].data(), M.swx[l][sh].size(), sizeof(LineGrp));
      dmn.nswx[l][sh] = (int)M.swx[l][sh].size();
      dmn.swy[l][sh] = (const LineGrp *)put(M.swy[l][sh].data(), M.swy[l][sh].size(), sizeof(LineGrp));
      dmn.nswy[l][sh] = (int)M.swy[l][sh].size();
    }
    for (int d = 0; d < 2; ++d)
      for (int p = 0; p < 2; ++p) {
        auto &v = M.swm[l][d][p];
        dmn.swm[l][d][p] = (const LineGrp *)put(v.data(), v.size(), sizeof(LineGrp));
        dmn.nswm[l][d][p] = (int)v.size();
        dmn.swm_na[l][d][p] = M.swm_na[l][d][p];
        dmn.swm_nb[l][d][p] = M.swm_nb[l][d][p];
        dmn.swm_all[l][d][p] = M.swm_all[l][d][p];
      }
  }
  dmn.maxlev = maxlev;
  if (bytes) HIPCHK(hipMemcpy(d_maps, host.data(), bytes, hipMemcpyHostToDevice));
  if (!table.empty())
    HIPCHK(hipMemcpy(d_insts, table.data(), table.size() * sizeof(InstDev), hipMemcpyHostToDevice));
  dm = dmn;
  part_host.assign(order.size(), -1);   // forces a re-upload of the participation flags
  dirty = false;
  return SG_OK;
}

int Batch::read_counter(int which, int *out) {
  HIPCHK(hipMemcpyAsync(h_cnt + which, d_cnt + which, sizeof(int), hipMemcpyDeviceToHost, stream));
  HostLock lk(g_capture_mu);
  HIPCHK(hipStreamSynchronize(stream));
  *out = h_cnt[which];
  return SG_OK;
}

void Batch::tbegin(int id, hipStream_t s) {
  if (!(time_mask >> id & 1)) return;
  if (!s) s = stream;
  if (ev_used + 2 > ev_pool.size()) {
    for (int k = 0; k < 64; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) == hipSuccess) ev_pool.push_back(e);
    }
  }
  if (ev_used + 2 <= ev_pool.size()) (void)hipEventRecord(ev_pool[ev_used], s);
}

void Batch::tend(int id, hipStream_t s) {
  if (!(time_mask >> id & 1) || ev_used + 2 > ev_pool.size()) return;
  (void)hipEventRecord(ev_pool[ev_used + 1], s ? s : stream);
  if (ev_ids.size() < ev_pool.size() / 2) ev_ids.resize(ev_pool.size() / 2);
  ev_ids[ev_used / 2] = id;
  ev_used += 2;
}

int Batch::collect_time() {
  if (!ev_used) return SG_OK;
  HostLock lk(g_capture_mu);
  HIPCHK(hipDeviceSynchronize());   // the side stream's events too
  for (size_t k = 0; k + 1 < ev_used; k += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev_pool[k], ev_pool[k + 1]));
    const int id = ev_ids[k / 2];
    t_ms[id] += ms;
    t_launches[id] += 1;
  }
  ev_used = 0;
  return SG_OK;
}

// One launch per tile shape with work (launch(stream, shape)). With a side
// stream and two or more shapes, the shape of the longest lines (its groups
// are the critical path) runs on `side`, forked from and joined back into
// `stream`; the shapes write disjoint lines, so the results do not change.
// Under stream capture the fork and join become graph edges.
template <typename F> void Batch::fan_out(const int (&cnt)[kNumTileShapes], F &&launch) {
  int nshape = 0, last = -1;
  for (int sh = 0; sh < kNumTileShapes; ++sh)
    if (cnt[sh]) {
      ++nshape;
      if (last < 0 || shape_rank(sh) >= shape_rank(last)) last = sh;
    }
  if (nshape < 2 || !side) {
    for (int sh = 0; sh < kNumTileShapes; ++sh)
      if (cnt[sh]) launch(stream, sh);
    return;
  }
  auto note = [&](hipError_t e) { if (e != hipSuccess && fan_err == hipSuccess) fan_err = e; };
  note(hipEventRecord(ev_fork, stream));
  note(hipStreamWaitEvent(side, ev_fork, 0));
  launch(side, last);
  note(hipEventRecord(ev_join, side));
  for (int sh = 0; sh < kNumTileShapes; ++sh)
    if (cnt[sh] && sh != last) launch(stream, sh);
  note(hipStreamWaitEvent(stream, ev_join, 0));
}

template <typename T> void Batch::vcycle(int l, int xid, int bid) {
  const DevMaps &m = dm;
  const int sm = l == 0 ? opts.smoother : opts.coarse_smoother;
  const bool pre_x = sm & 1, pre_y = sm & 2, post_x = sm & 4, post_y = sm & 8;
  const int lid = l == 0 ? 2 : 3;
  bool zero = true;
  auto sweep = [&](int dir) {
    for (int p = 0; p < 2; ++p) {
      if (l >= 1 && m.swm_all[l][dir][p]) {
        // a 9-point level, with one stream or two: the 1024-thread long-column
        // groups, then every other shape of the colour in one launch (so that
        // both stream modes run the same kernels and agree bit for bit)
        if (m.nsw[l][dir][p][3]) {
          tbegin(lid);
          launch_line<T>(stream, d_insts, m.sw[l][dir][p][3], m.nsw[l][dir][p][3], l, dir, 3,
                         zero, xid, bid);
          tend(lid);
        }
        if (m.nswm[l][dir][p]) {
          tbegin(lid);
          launch_line_shapes<T>(stream, d_insts, m.swm[l][dir][p], m.swm_na[l][dir][p],
                                m.swm_nb[l][dir][p], m.nswm[l][dir][p], l, dir, zero, xid, bid);
          tend(lid);
        }
      } else {
        fan_out(m.nsw[l][dir][p], [&](hipStream_t s, int sh) {
          tbegin(lid, s);   // one timed interval per kernel dispatch
          launch_line<T>(s, d_insts, m.sw[l][dir][p][sh], m.nsw[l][dir][p][sh], l, dir, sh, zero,
                         xid, bid);
          tend(lid, s);
        });
      }
      zero = false;
    }
  };
  // Level l + 1 has per-level work for some instance (`deeper`), or this is
  // the bottom of the descent. With post-smoothing only, the descent is
  // restrictions of b, so at the bottom every fused cycle has the b of its
  // first level and no per-level prolongation has started: all of them run
  // here, in one launch, one workgroup per instance.
  const bool deeper = l + 1 < m.maxlev && m.nrmap[l + 1] > 0;
  if (!deeper && m.nfmap) {
    tbegin(11);
    launch_coarse_cycle<T>(stream, d_insts, m.fmap, m.nfmap, xid, bid,
                           opts.smoother | opts.coarse_smoother << 4, opts.fuse_cells);
    tend(11);
  }
  if (pre_x) sweep(0);
  if (pre_y) sweep(1);
  const bool pre_zero = zero;
  // exact solve on the coarsest level of the instances that end here
  for (int sh = 0; sh < kNumTileShapes; ++sh) {
    if (m.nswx[l][sh]) {
      tbegin(3);
      launch_line<T>(stream, d_insts, m.swx[l][sh], m.nswx[l][sh], l, 0, sh, true, xid, bid);
      tend(3);
    }
    if (m.nswy[l][sh]) {
      tbegin(3);
      launch_line<T>(stream, d_insts, m.swy[l][sh], m.nswy[l][sh], l, 1, sh, true, xid, bid);
      tend(3);
    }
  }
  if (deeper) {
    tbegin(4);
    // after a pre-sweep of x-lines alone the residual vanishes on the odd
    // rows: the even rows' residual is formed inside the restriction
    if (!pre_zero && half_residual(sm)) {
      launch_restrict_residual<T>(stream, d_insts, m.rmap[l + 1], m.nrmap[l + 1], l, xid, bid);
    } else {
      if (!pre_zero) launch_residual<T>(stream, d_insts, m.pmap[l], m.npmap[l], l, xid, bid);
      launch_restrict<T>(stream, d_insts, m.rmap[l + 1], m.nrmap[l + 1], l, pre_zero, xid, bid);
    }
    tend(4);
    vcycle<T>(l + 1, xid, bid);
    tbegin(5);
    // only the lines the first post-smoothing sweep reads are needed
    launch_prolong<T>(stream, d_insts, m.qmap[l], m.nqmap[l], l, pre_zero, xid, prolong_rows(sm));
    tend(5);
    zero = false;
    if (post_x) sweep(0);
    if (post_y) sweep(1);
  }
}

template <typename KT, typename VT> void Batch::iteration(bool xstate) {
  const int nb = dm.ncmap[0], n = (int)order.size();
  const BlkMap *m0 = dm.cmap[0];
  constexpr bool pc = sizeof(KT) != sizeof(VT);
  tbegin(8); launch_bicg<KT>(stream, 0, d_insts, m0, nb, xstate, pc); tend(8);
  vcycle<VT>(0, pc ? PV_PH : VPH, pc ? PV_P : VP);
  tbegin(6); launch_bicg<KT>(stream, 1, d_insts, m0, nb, xstate, pc); tend(6);
  tbegin(12); launch_finish<KT>(stream, FIN_V, d_insts, n, d_cnt); tend(12);
  tbegin(8); launch_bicg<KT>(stream, 2, d_insts, m0, nb, xstate, pc); tend(8);
  tbegin(12); launch_finish<KT>(stream, FIN_S, d_insts, n, d_cnt); tend(12);
  vcycle<VT>(0, pc ? PV_SH : VSH, pc ? PV_B : VS);
  tbegin(7); launch_bicg<KT>(stream, 3, d_insts, m0, nb, xstate, pc); tend(7);
  tbegin(12); launch_finish<KT>(stream, FIN_T, d_insts, n, d_cnt); tend(12);
  tbegin(8); launch_bicg<KT>(stream, 4, d_insts, m0, nb, xstate, pc); tend(8);
  tbegin(12); launch_finish<KT>(stream, FIN_X, d_insts, n, d_cnt); tend(12);
}

template <typename F> int Batch::run_graph(int key, F &&enqueue) {
  if (!graphs_on()) {
    enqueue();
    HIPCHK(fan_err);
    return SG_OK;
  }
  hipGraphExec_t exec = nullptr;
  for (auto &g : graphs)
    if (g.key == key) exec = g.exec;
  if (!exec) {
    std::unique_lock<std::shared_mutex> lk(g_capture_mu);
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    const hipError_t ce = hipStreamEndCapture(stream, &graph);
    HIPCHK(fan_err);
    HIPCHK(ce);
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(ie);
    graphs.push_back({key, exec});
  }
  HIPCHK(hipGraphLaunch(exec, stream));
  return SG_OK;
}

// The set-up of a step: statuses and counters, assembly, the Galerkin chain of
// the multigrid hierarchy and the line factorisations; start64: also the start
// of the FP64 BiCGSTAB (the assembly wrote r0 and its partial sums).
void Batch::setup_kernels(bool start64) {
  const int n = (int)order.size(), nb0 = dm.ncmap[0];
  const bool mixed = opts.precision == SG_MIXED;
  const bool h32 = opts.precision != SG_FP64;   // FP32 hierarchy and V-cycle
  const double tol2 = mixed ? opts.inner_rtol * opts.inner_rtol : opts.rtol * opts.rtol;
  launch_begin(stream, d_insts, n, d_part, d_cnt, d_cnt + 1, tol2, opts.rtol * opts.rtol,
               mixed ? opts.max_inner : opts.max_iters, opts.max_iters);
  tbegin(0); launch_assemble(stream, d_insts, dm.cmap[0], nb0); tend(0);
  // LU factors of the line matrices of level l: smoothing lines of both
  // colours (one launch per direction and tile shape) and the exact solve on
  // the coarsest level
  auto factors = [&](int l, hipStream_t s) {
    for (int d = 0; d < 2; ++d)
      for (int sh = 0; sh < kNumTileShapes; ++sh)
        launch_line_factor(s, d_insts, dm.gfac[l][d][sh], dm.ngfac[l][d][sh], l, d, sh, h32);
    for (int sh = 0; sh < kNumTileShapes; ++sh) {
      launch_line_factor(s, d_insts, dm.gcx[l][sh], dm.ngcx[l][sh], l, 0, sh, h32);
      launch_line_factor(s, d_insts, dm.gcy[l][sh], dm.ngcy[l][sh], l, 1, sh, h32);
    }
  };
  auto note = [&](hipError_t e) { if (e != hipSuccess && fan_err == hipSuccess) fan_err = e; };
  if (side && time_mask == 0) {
    // A level's factorisations depend only on its operator, and nothing else
    // in the set-up reads them: they run on the side stream, each level's as
    // soon as its operator exists, beside the chain of prolongation weights and
    // Galerkin products that builds the next level on the main stream. The
    // main stream waits for them before the solve. Same kernels, same inputs.
    note(hipEventRecord(ev_lev[0], stream));
    for (int l = 0; l + 1 < dm.maxlev; ++l) {
      launch_pweights(stream, d_insts, dm.pmap[l], dm.npmap[l], l, h32);
      launch_rap(stream, d_insts, dm.cmap[l + 1], dm.ncmap[l + 1], l, h32);
      note(hipEventRecord(ev_lev[l + 1], stream));
    }
    for (int l = 0; l < dm.maxlev; ++l) {
      note(hipStreamWaitEvent(side, ev_lev[l], 0));
      factors(l, side);
    }
    note(hipEventRecord(ev_join, side));
    note(hipStreamWaitEvent(stream, ev_join, 0));
  } else {
    for (int l = 0; l < dm.maxlev; ++l) {
      tbegin(1);
      factors(l, stream);
      if (l + 1 < dm.maxlev) {
        launch_pweights(stream, d_insts, dm.pmap[l], dm.npmap[l], l, h32);
        launch_rap(stream, d_insts, dm.cmap[l + 1], dm.ncmap[l + 1], l, h32);
      }
      tend(1);
    }
  }
  if (start64) {
    tbegin(12); launch_finish<double>(stream, FIN_INIT64, d_insts, n, d_cnt); tend(12);
  }
}

// A step-end pass over the instances whose solve ended since the previous
// pass: true residual, counts and min of x, the acceptance decision, and f of
// the accepted ones. Instances finalized by an earlier pass are left alone, so
// passes may follow every chunk of iterations.
void Batch::end_pass() {
  const int n = (int)order.size(), nb0 = dm.ncmap[0];
  tbegin(10); launch_finalize(stream, d_insts, dm.cmap[0], nb0); tend(10);
  tbegin(12); launch_finish<double>(stream, FIN_FINALIZE, d_insts, n, d_cnt); tend(12);
  tbegin(10); launch_commit(stream, d_insts, dm.cmap[0], nb0); tend(10);
  if (opts.warm_start > 0) {
    tbegin(12); launch_finish<double>(stream, FIN_EXTRAP, d_insts, n, d_cnt); tend(12);
  }
}

static int graph_key(int kind, int n) { return (kind << 16) | n; }

// Krylov iterations from iteration it0, in chunks of check_every (each chunk
// followed by a step-end pass when `ends`), until the active-instance counter
// reads 0 or maxit. The next chunk is enqueued before the host waits for the
// counter of the current one, so the GPU never drains while the host decides;
// a chunk enqueued after the last instance converged costs only early-exiting
// launches.
template <typename KT, typename VT>
int Batch::krylov(int maxit, bool xstate, int it0, bool ends) {
  if (!ev_cnt[0]) {
    for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreateWithFlags(&ev_cnt[k], hipEventDisableTiming));
  }
  const int variant = (sizeof(KT) == 4 ? 0 : sizeof(VT) == 4 ? 1 : 2) * 4 + (xstate ? 2 : 0) +
                      (ends ? 1 : 0);
  auto chunk = [&](int slot, int &it) -> int {
    const int c = std::min(opts.check_every, maxit - it);
    int rc = run_graph(graph_key(2 + variant, c), [&] {
      for (int k = 0; k < c; ++k) iteration<KT, VT>(xstate);
      if (ends) end_pass();
    });
    if (rc) return rc;
    it += c;
    HIPCHK(hipMemcpyAsync(h_pipe + slot, d_cnt, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipEventRecord(ev_cnt[slot], stream));
    return SG_OK;
  };
  int it = it0, slot = 0, rc;
  if (it >= maxit) return SG_OK;
  if ((rc = chunk(slot, it))) return rc;
  for (;;) {
    const bool more = it < maxit;
    if (more && (rc = chunk(slot ^ 1, it))) return rc;
    {
      HostLock lk(g_capture_mu);
      HIPCHK(hipEventSynchronize(ev_cnt[slot]));
    }
    if (h_pipe[slot] == 0 || !more) break;
    slot ^= 1;
  }
  return SG_OK;
}

// FP64 BiCGSTAB (FP64 or FP32 V-cycle). The head of the step -- set-up, the
// Krylov start, as many iterations as the previous step needed (at most
// kHeadIters) and a step-end pass -- is one graph launch; the host then reads
// the active-instance counter once. Only when some instance still iterates do
// pipelined chunks of iterations, each with its own step-end pass, follow.
template <typename KT, typename VT> int Batch::solve64() {
  const int maxit = opts.max_iters;
  const int want = opts.head_iters < 0 ? 0 : opts.head_iters > 0 ? opts.head_iters : last_iters_max;
  const int k0 = std::max(0, std::min({want, maxit, kHeadIters}));
  int rc = run_graph(graph_key(1, k0), [&] {
    setup_kernels(true);
    for (int k = 0; k < k0; ++k) iteration<KT, VT>(true);
    end_pass();
  });
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h_pipe, d_cnt, sizeof(int), hipMemcpyDeviceToHost, stream));
  {
    HostLock lk(g_capture_mu);
    HIPCHK(hipStreamSynchronize(stream));
  }
  if (h_pipe[0] == 0 || k0 >= maxit) return SG_OK;
  return krylov<KT, VT>(maxit, true, k0, true);
}

// Algorithmic bytes of the timed families in one step, from the V-cycles,
// iterations and refinements each participating instance ran: every array a
// kernel reads or writes counted once per launch, T the V-cycle precision.
void Batch::count_bytes(const std::vector<int> &part) {
  const int prec = opts.precision;
  const double T = prec == SG_FP64 ? 8.0 : 4.0;   // hierarchy / V-cycle values
  const double h32 = prec != SG_FP64 ? 1.0 : 0.0;
  // BiCGSTAB per cell and iteration (SpMV v, SpMV t, and the p, s, x updates)
  const double bv = prec == SG_MIXED ? 16.0 + 4 * 3 : prec == SG_PC32 ? 40.0 + 4 + 4 + 8 : 64.0;
  const double bt = prec == SG_MIXED ? 16.0 + 4 * 3 : prec == SG_PC32 ? 40.0 + 4 + 8 + 8 : 64.0;
  const double bvec = prec == SG_MIXED ? 60.0 : prec == SG_PC32 ? 24.0 + 28.0 + 52.0 : 120.0;
  auto half_prolong = [&](int l) {
    return prolong_rows(l == 0 ? opts.smoother : opts.coarse_smoother) != 0;
  };
  for (size_t c = 0; c < part.size(); ++c) {
    if (!part[c]) continue;
    const Slot &s = slots[order[c]];
    const KScal &k = h_scal[order[c]];
    const InstDev &I = s.dev;
    const double N0 = (double)s.nx * s.ny, vc = k.vcyc, it = k.titers;
    t_bytes[0] += N0 * (48.0 + 56.0 + 20.0 * h32 + (prec == SG_MIXED ? 0.0 : 8.0) +
                        (prec == SG_FP64 ? 24.0 : 0.0) + (I.warm && k.theta > 0.0 ? 4.0 : 0.0));
    for (int l = 0; l < I.nlev; ++l) {
      const double Nl = (double)I.lev[l].nx * I.lev[l].ny;
      const double nb = l == 0 ? 5.0 : 9.0;
      // line factorisations: the 5 bands they read, 2 factors per direction written
      t_bytes[1] += Nl * (5.0 * 8.0 + 4.0 * T);
      if (l + 1 < I.nlev) {
        const double Nc = (double)I.lev[l + 1].nx * I.lev[l + 1].ny;
        // prolongation weights (bands in, the stored weights out: 0, 2, 2
        // and 4 for the even-even, odd-even, even-odd and odd-odd points, 2
        // per point on average) and the Galerkin product (bands and stored
        // weights in, 9 FP64 bands and their copies out)
        t_bytes[1] += Nl * (nb * 8.0 + 2.0 * T) + Nl * (nb * 8.0 + 2.0 * T) +
                      Nc * (72.0 + 36.0 * h32);
      }
      if (l >= s.fuse_lev) continue;   // the fused coarse cycle
      if (l == I.nlev - 1) {           // exact line solve of the coarsest level
        t_bytes[3] += vc * Nl * 5.0 * T;
        continue;
      }
      // a V-cycle sweeps every cell once per direction it smooths: 8
      // (5-point) or 12 (9-point) values per swept cell
      const int sml = l == 0 ? opts.smoother : opts.coarse_smoother;
      const double vals = l == 0 ? 8.0 : 12.0;
      const double post = ((sml & 4) ? 1.0 : 0.0) + ((sml & 8) ? 1.0 : 0.0);
      const double pre = ((sml & 1) ? 1.0 : 0.0) + ((sml & 2) ? 1.0 : 0.0);
      // the first colour of the first pre-sweep starts from x = 0: b, the
      // lower band and the two factors in, x out (5 values per swept cell)
      const double pre_vals = pre > 0 ? (5.0 + vals) / 2.0 + (pre - 1.0) * vals : 0.0;
      t_bytes[l == 0 ? 2 : 3] += vc * (post * vals + pre_vals) * T * Nl;
      const double Nc = (double)I.lev[l + 1].nx * I.lev[l + 1].ny;
      if (!(sml & 3)) {
        // restriction of b into level l + 1 (x = 0): b and the stored weights
        // (2 per fine cell on average) in, b out
        t_bytes[4] += vc * (3.0 * T * Nl + T * Nc);
      } else {
        // after an x-line pre-sweep: the even rows' residual formed inside
        // the restriction (b, the bands, x and 2 stored weights per cell in;
        // b out); otherwise the residual of every cell (b, the bands, x; r
        // out) and its restriction (r and 2 weights per cell in; b out)
        if (half_residual(sml))
          t_bytes[4] += vc * (0.5 * Nl * (nb + 4.0) * T + T * Nc);
        else
          t_bytes[4] += vc * (Nl * (nb + 3.0) * T + 3.0 * T * Nl + T * Nc);
      }
      // prolongation into the lines the first sweep reads (half the cells:
      // the odd rows or columns, 3 stored weights per cell on average), x in
      // and out, and the coarse x once; into every cell (2 weights on
      // average) when the first sweep is not a post-smoothing line sweep
      t_bytes[5] += vc * ((half_prolong(l) ? 0.5 * 5.0 : 4.0) * Nl * T + T * Nc);
    }
    t_bytes[6] += it * bv * N0;
    t_bytes[7] += it * bt * N0;
    t_bytes[8] += it * bvec * N0;
    if (prec == SG_MIXED) t_bytes[9] += (k.orefs + 1) * 64.0 * N0;
    t_bytes[10] += N0 * (56.0 + 16.0 + (I.warm ? 16.0 : 0.0));
  }
}

int Batch::step_once() {
  int rc;
  const int n = (int)order.size();
  if (!n) return SG_OK;
  std::vector<int> part(n);
  int any = 0;
  for (int c = 0; c < n; ++c) {
    const Slot &s = slots[order[c]];
    part[c] = (s.target < 0 || s.steps < s.target) ? 1 : 0;
    any += part[c];
  }
  if (!any) return SG_OK;
  if (part != part_host) {
    if (n > h_part_cap) {   // grows only between steps (pinned: the copy stays asynchronous)
      HostLock lk(g_capture_mu);
      HIPCHK(hipStreamSynchronize(stream));
      if (h_part) (void)hipHostFree(h_part);
      h_part = nullptr;
      h_part_cap = 0;
      HIPCHK(hipHostMalloc((void **)&h_part, sizeof(int) * n, hipHostMallocDefault));
      h_part_cap = n;
    }
    std::memcpy(h_part, part.data(), n * sizeof(int));
    HIPCHK(hipMemcpyAsync(d_part, h_part, n * sizeof(int), hipMemcpyHostToDevice, stream));
    part_host = part;
  }
  const bool mixed = opts.precision == SG_MIXED;
  const bool h32 = opts.precision != SG_FP64;   // FP32 hierarchy and V-cycle
  if (!mixed) {
    if ((rc = h32 ? solve64<double, float>() : solve64<double, double>())) return rc;
  } else {
    if ((rc = run_graph(graph_key(1, 0), [&] { setup_kernels(false); }))) return rc;
    const int nb0 = dm.ncmap[0];
    for (int ref = 0; ref <= opts.max_iters; ++ref) {
      tbegin(9); launch_ir_residual(stream, d_insts, dm.cmap[0], nb0); tend(9);
      tbegin(12); launch_finish<double>(stream, FIN_IR_RES, d_insts, n, d_cnt + 1); tend(12);
      int oactive = 0;
      if ((rc = read_counter(1, &oactive))) return rc;
      if (!oactive) break;
      launch_ir_inner_init(stream, d_insts, dm.cmap[0], nb0);
      tbegin(12); launch_finish<float>(stream, FIN_INNER_INIT, d_insts, n, d_cnt); tend(12);
      launch_copy_counter(stream, d_cnt + 1, d_cnt);
      if ((rc = krylov<float, float>(opts.max_inner, false, 0, false))) return rc;
      launch_ir_update(stream, d_insts, dm.cmap[0], nb0);
    }
    end_pass();
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h_scal, d_scal, sizeof(KScal) * slots.size(), hipMemcpyDeviceToHost,
                        stream));
  {
    HostLock lk(g_capture_mu);
    HIPCHK(hipStreamSynchronize(stream));
  }
  if (time_mask) {
    if ((rc = collect_time())) return rc;
    count_bytes(part);
  }
  int worst = SG_OK, nfail = 0, maxit = 0;
  for (int c = 0; c < n; ++c) {
    if (!part[c]) continue;
    Slot &s = slots[order[c]];
    const KScal &k = h_scal[order[c]];
    const bool ok = k.ostatus == ST_DONE;
    if (ok) s.steps += 1;   // a rejected step leaves f and the step count as they were
    maxit = std::max(maxit, k.iters);
    s.st.steps = s.steps;
    s.st.last_iters = k.titers;
    s.st.total_iters += k.titers;
    s.st.last_refinements = mixed ? k.orefs : 0;
    s.st.last_rel_residual = k.true_rel;
    s.st.last_min_x = k.minf;
    s.st.min_f = (opts.no_clamp || !ok) ? k.minf : std::max(k.minf, 0.0);
    s.st.last_clamps = ok ? k.clamps : 0;
    s.st.clamp_count += s.st.last_clamps;
    s.st.last_nonfinite = k.nonfin;
    s.st.status = ok ? SG_OK : SG_ERR_NOCONV;
    if (!ok) { worst = SG_ERR_NOCONV; ++nfail; }
  }
  last_iters_max = maxit;
  if (nfail)
    g_err = "step rejected for " + std::to_string(nfail) +
            " instance(s): the solve did not reach its tolerance, broke down or left non-finite "
            "values (f and the step count are unchanged)";
  return worst;
}

int Batch::step(int nsteps, const int64_t *targets, int ntargets) {
  int rc;
  if (nsteps < 0) return fail(SG_ERR_ARG, "nsteps must be >= 0");
  if (dirty) {
    HostLock lk(g_capture_mu);
    if ((rc = rebuild())) return rc;
  }
  for (size_t k = 0; k < slots.size(); ++k)
    slots[k].target = (targets && (int)k < ntargets) ? targets[k] : -1;
  int worst = SG_OK;
  for (int s = 0; s < nsteps; ++s) {
    rc = step_once();
    if (rc == SG_ERR_NOCONV) { worst = rc; continue; }
    if (rc) return rc;
  }
  return worst;
}

int Batch::scratch(size_t n) {
  if (n <= scratch_elems) return SG_OK;
  (void)hipFree(d_scratch);
  d_scratch = nullptr;
  HIPCHK(hipMalloc(&d_scratch, n * sizeof(double)));
  scratch_elems = n;
  return SG_OK;
}

int Batch::get_f(int slot, double *out) {
  if (slot < 0 || slot >= (int)slots.size() || !slots[slot].used)
    return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  if (!out) return fail(SG_ERR_ARG, "out must not be null");
  const Slot &s = slots[slot];
  int rc;
  if ((rc = scratch((size_t)s.nx * s.ny))) return rc;
  launch_to_imajor(stream, s.dev.f, d_scratch, s.nx, s.ny);
  HIPCHK(hipMemcpyAsync(out, d_scratch, (size_t)s.nx * s.ny * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SG_OK;
}

int Batch::set_f(int slot, const double *f, int64_t istep) {
  if (slot < 0 || slot >= (int)slots.size() || !slots[slot].used)
    return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  if (!f) return fail(SG_ERR_ARG, "f must not be null");
  Slot &s = slots[slot];
  const size_t N = (size_t)s.nx * s.ny;
  for (size_t q = 0; q < N; ++q)
    if (!std::isfinite(f[q])) return fail(SG_ERR_ARG, "f: non-finite value");
  int rc;
  if ((rc = scratch(N))) return rc;
  HIPCHK(hipMemcpyAsync(d_scratch, f, N * 8, hipMemcpyHostToDevice, stream));
  launch_to_korder(stream, d_scratch, s.dev.f, s.nx, s.ny);
  // the state carries no previous step: the next solve starts at f
  HIPCHK(hipMemsetAsync(&d_scal[slot].hist, 0, sizeof(int), stream));
  HIPCHK(hipMemsetAsync(&d_scal[slot].theta, 0, sizeof(double), stream));
  HIPCHK(hipStreamSynchronize(stream));
  s.steps = istep;
  s.st.steps = istep;
  return SG_OK;
}

int Batch::debug_assemble(int slot, double *bands, double *R) {
  if (slot < 0 || slot >= (int)slots.size() || !slots[slot].used)
    return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  int rc;
  if (dirty && (rc = rebuild())) return rc;
  const int n = (int)order.size();
  std::vector<int> part(n, 0);
  for (int c = 0; c < n; ++c) part[c] = order[c] == slot;
  HIPCHK(hipMemcpyAsync(d_part, part.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
  part_host = part;
  launch_begin(stream, d_insts, n, d_part, d_cnt, d_cnt + 1, 0, 0, 1, 1);
  launch_assemble(stream, d_insts, dm.cmap[0], dm.ncmap[0]);
  HIPCHK(hipGetLastError());
  const Slot &s = slots[slot];
  const size_t N = (size_t)s.nx * s.ny;
  std::vector<double> h(6 * N);
  for (int b = 0; b < 5; ++b)
    HIPCHK(hipMemcpyAsync(h.data() + b * N, s.dev.band64[b], N * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(h.data() + 5 * N, s.dev.R, N * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  // device cell order -> K = j*nx + i
  for (int j = 0; j < s.ny; ++j)
    for (int i = 0; i < s.nx; ++i) {
      const size_t K = (size_t)j * s.nx + i, d = cell_slot(i, j, s.nx);
      for (int b = 0; b < 5; ++b) bands[b * N + K] = h[b * N + d];
      R[K] = h[5 * N + d];
    }
  return SG_OK;
}

}  // namespace sg

// ---------------------------------------------------------------------------
// Static scratch of the no-allocation read-out (sg_batch_get_f_static): a
// file-scope device array, one copy per GPU for the life of the process. Host
// code never uses the array's name as an address: its device address on each
// GPU comes from hipGetSymbolAddress, taken once per device with that device
// current, when the first batch is created there.
// ---------------------------------------------------------------------------
namespace sg {

constexpr size_t kStaticScratchElems = size_t(1) << 22;   // 4 Mi doubles, 32 MiB
static __device__ __align__(256) double scratch[kStaticScratchElems];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (taken once per device).
static int static_scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(SG_ERR_HIP, "device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return SG_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SG_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
  return SG_OK;
}

static double *static_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return static_cast<double *>(g_scratch_ptr[dev]);
}

// sg_batch_get_f with the layout transpose staged in the static scratch:
// no device allocation, whatever was read out before.
int Batch::get_f_static(int slot, double *out, double *buf) {
  if (slot < 0 || slot >= (int)slots.size() || !slots[slot].used)
    return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  if (!out) return fail(SG_ERR_ARG, "out must not be null");
  const Slot &s = slots[slot];
  const size_t N = (size_t)s.nx * s.ny;
  if (N > kStaticScratchElems)
    return fail(SG_ERR_ARG, "instance of " + std::to_string(N) +
                                " cells exceeds the static read-out scratch of " +
                                std::to_string(kStaticScratchElems) + " cells");
  launch_to_imajor(stream, s.dev.f, buf, s.nx, s.ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, buf, N * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SG_OK;
}

}  // namespace sg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using sg::Batch;
using sg::fail;

// A batch keeps its instances in two groups: instances of at least
// opts.group_cells cells (whose long line solves are latency-bound) and the
// rest (bandwidth-bound). Each group is an engine Batch with its own stream;
// a step advances the two concurrently, the large group on a second host
// thread. Instances are independent, so results do not depend on the
// grouping. Slots seen by the caller are global; `where` maps them to a group
// and that group's slot.
struct sg_batch {
  Batch g[2];
  bool inited[2] = {false, false};
  sg_opts opts{};
  size_t budget = 0;
  int device = 0;
  std::vector<std::pair<int, int>> where;   // global slot -> (group, local slot); group -1: free
  int time_mask = 0;

  Batch *group_of(int slot, int *local) {
    if (slot < 0 || slot >= (int)where.size() || where[slot].first < 0) return nullptr;
    *local = where[slot].second;
    return &g[where[slot].first];
  }
  const Batch *group_of(int slot, int *local) const {
    return const_cast<sg_batch *>(this)->group_of(slot, local);
  }
  int ensure_group(int k) {
    if (inited[k]) return SG_OK;
    int rc = g[k].init(&opts, 0);
    if (rc) return rc;
    g[k].time_mask = time_mask;
    inited[k] = true;
    return SG_OK;
  }
  size_t in_use() const { return g[0].in_use + g[1].in_use; }
  int size() const {
    int n = 0;
    for (auto &w : where) n += w.first >= 0;
    return n;
  }
  // Steps both groups (targets indexed by global slot, as sg_batch_step_until).
  int step(int nsteps, const int64_t *targets, int nslots) {
    std::vector<int64_t> t[2];
    if (targets) {
      for (int k = 0; k < 2; ++k) t[k].assign(g[k].slots.size(), -1);
      for (int s = 0; s < (int)where.size() && s < nslots; ++s)
        if (where[s].first >= 0) t[where[s].first][where[s].second] = targets[s];
    }
    auto run = [&](int k) {
      return g[k].step(nsteps, targets ? t[k].data() : nullptr, targets ? (int)t[k].size() : 0);
    };
    const bool busy[2] = {inited[0] && has(0), inited[1] && has(1)};
    if (!(busy[0] && busy[1])) return busy[1] ? run(1) : busy[0] ? run(0) : SG_OK;
    // Rebuilding a group's maps frees and allocates device memory, and a free
    // synchronizes the device, which would invalidate a graph the other
    // group's thread is capturing. So both groups are rebuilt here, before
    // they step concurrently; the concurrent steps then only launch, capture
    // and wait on their own streams.
    for (int k = 0; k < 2; ++k)
      if (g[k].dirty) {
        sg::HostLock lk(sg::g_capture_mu);
        const int rc = g[k].rebuild();
        if (rc) return rc;
      }
    int rc1 = SG_OK;
    std::string err1;
    std::thread worker([&] {
      (void)hipSetDevice(device);
      rc1 = run(1);
      if (rc1) err1 = sg::g_err;
    });
    const int rc0 = run(0);
    worker.join();
    if (rc0 && rc0 != SG_ERR_NOCONV) return rc0;
    if (rc1) { sg::g_err = err1; return rc1; }
    return rc0;
  }
  bool has(int k) const {
    for (auto &w : where)
      if (w.first == k) return true;
    return false;
  }
};
struct sg_solver { sg_batch *bt; int slot; };

extern "C" {

int sg_abi_version(void) { return SG_ABI_VERSION; }
const char *sg_last_error(void) { return sg::g_err.c_str(); }
int sg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return 0; }
  return n;
}

size_t sg_instance_bytes(int nx, int ny, int precision) {
  if (nx < 1 || ny < 1) return 0;
  sg::InstDev tmp{};
  if (precision != SG_FP64 && precision != SG_MIXED && precision != SG_PC32) return 0;
  return sg::carve(0, nx, ny, precision, tmp);
}

int sg_batch_create(const sg_opts *opts, size_t budget, sg_batch **out) {
  if (!out) return fail(SG_ERR_ARG, "out must not be null");
  sg::HostLock lk(sg::g_capture_mu);
  *out = nullptr;
  std::unique_ptr<sg_batch> b(new sg_batch);
  if (opts) b->opts = *opts;
  b->budget = budget;
  // group 0 is created here, so option errors and a missing GPU show at once
  int rc = b->ensure_group(0);
  if (rc) return rc;
  b->opts = b->g[0].opts;   // with the defaults applied
  (void)hipGetDevice(&b->device);
  // the device address of the static read-out scratch on this GPU
  if ((rc = sg::static_scratch_setup(b->device))) return rc;
  *out = b.release();
  return SG_OK;
}

int sg_batch_add(sg_batch *b, const sg_grid *g, const sg_coeffs *c, const sg_bc *bc,
                 const double *f0, double dt, int *slot) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  if (!g) return fail(SG_ERR_ARG, "grid must not be null");
  const long long cells = (long long)g->nx * g->ny;
  const int k = (b->opts.group_cells > 0 && cells >= b->opts.group_cells) ? 1 : 0;
  int rc = b->ensure_group(k);
  if (rc) return rc;
  // one HBM budget for both groups
  Batch &G = b->g[k];
  const size_t other = b->g[1 - k].in_use;
  G.budget = b->budget ? (b->budget > other ? b->budget - other : 1) : 0;
  int local = -1;
  if ((rc = G.add(g, c, bc, f0, dt, &local))) return rc;
  int gs = -1;
  for (int s = 0; s < (int)b->where.size(); ++s)
    if (b->where[s].first < 0) { gs = s; break; }
  if (gs < 0) { gs = (int)b->where.size(); b->where.push_back({-1, -1}); }
  b->where[gs] = {k, local};
  if (slot) *slot = gs;
  return SG_OK;
}
int sg_batch_remove(sg_batch *b, int slot) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  int local;
  Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  int rc = G->remove(local);
  if (rc) return rc;
  b->where[slot] = {-1, -1};
  return SG_OK;
}
int sg_batch_step(sg_batch *b, int nsteps) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  if (nsteps < 0) return fail(SG_ERR_ARG, "nsteps must be >= 0");
  return b->step(nsteps, nullptr, 0);
}
int sg_batch_step_until(sg_batch *b, const int64_t *targets, int nslots, int nsteps) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  if (!targets) return fail(SG_ERR_ARG, "targets must not be null");
  if (nsteps < 0) return fail(SG_ERR_ARG, "nsteps must be >= 0");
  return b->step(nsteps, targets, nslots);
}
int sg_batch_get_f(sg_batch *b, int slot, double *out) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  int local;
  Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  return G->get_f(local, out);
}
int sg_batch_get_f_static(sg_batch *b, int slot, double *out) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  int local;
  Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  double *buf = sg::static_scratch(b->device);
  if (!buf) return fail(SG_ERR_STATE, "static scratch not set up on this device");
  return G->get_f_static(local, out, buf);
}
int sg_batch_set_f(sg_batch *b, int slot, const double *f, int64_t istep) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  int local;
  Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  return G->set_f(local, f, istep);
}
int sg_batch_set_bc_values(sg_batch *b, int slot, int side, const double *values) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  int local;
  Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot");
  if (side < 0 || side > 3) return fail(SG_ERR_ARG, "side must be 0..3");
  auto &s = G->slots[local];
  if (!(s.dev.bcmask & (1 << side))) return fail(SG_ERR_ARG, "side is not Dirichlet");
  return G->upload_bc(s, side, values);
}
int sg_batch_set_coeffs(sg_batch *b, int slot, const sg_coeffs *c) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  sg::HostLock lk(sg::g_capture_mu);
  int rc = sg::check_coeffs(c, false);
  if (rc) return rc;
  int local;
  Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot " + std::to_string(slot));
  return G->set_coeffs(G->slots[local], c);
}
int sg_batch_stats(const sg_batch *b, int slot, sg_step_stats *out) {
  if (!b || !out) return fail(SG_ERR_ARG, "null argument");
  int local;
  const Batch *G = b->group_of(slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance in slot");
  *out = G->slots[local].st;
  return SG_OK;
}
int sg_batch_size(const sg_batch *b) { return b ? b->size() : 0; }
size_t sg_batch_bytes_in_use(const sg_batch *b) { return b ? b->in_use() : 0; }
void sg_batch_destroy(sg_batch *b) {
  if (!b) return;
  sg::HostLock lk(sg::g_capture_mu);
  delete b;
}

int sg_batch_time_kernel(sg_batch *b, int kernel_id) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  if (kernel_id > sg::kTimeAll) return fail(SG_ERR_ARG, "unknown kernel id");
  const int mask = kernel_id < 0 ? 0
                   : kernel_id == sg::kTimeAll ? (1 << sg::kNumKernelIds) - 1
                                               : 1 << kernel_id;
  b->time_mask = mask;
  for (auto &G : b->g) {
    G.time_mask = mask;
    G.reset_time();
  }
  return SG_OK;
}
int sg_batch_kernel_time(sg_batch *b, double *total_ms, int64_t *launches, double *bytes) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  double ms = 0, by = 0;
  int64_t nl = 0;
  for (auto &G : b->g)
    for (int k = 0; k < sg::kNumKernelIds; ++k) {
      ms += G.t_ms[k];
      nl += G.t_launches[k];
      by += G.t_bytes[k];
    }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = nl;
  if (bytes) *bytes = by;
  return SG_OK;
}
int sg_batch_kernel_time_id(sg_batch *b, int id, double *total_ms, int64_t *launches,
                            double *bytes) {
  if (!b) return fail(SG_ERR_ARG, "null batch");
  if (id < 0 || id >= sg::kNumKernelIds) return fail(SG_ERR_ARG, "unknown kernel id");
  if (total_ms) *total_ms = b->g[0].t_ms[id] + b->g[1].t_ms[id];
  if (launches) *launches = b->g[0].t_launches[id] + b->g[1].t_launches[id];
  if (bytes) *bytes = b->g[0].t_bytes[id] + b->g[1].t_bytes[id];
  return SG_OK;
}
int sg_batch_event_overhead(sg_batch *b, double *ms) {
  if (!b || !ms) return fail(SG_ERR_ARG, "null argument");
  sg::HostLock lk(sg::g_capture_mu);
  Batch &G = b->g[0];
  constexpr int kPairs = 64;
  hipEvent_t ev[2 * kPairs];
  int made = 0;
  for (; made < 2 * kPairs; ++made)
    if (hipEventCreate(&ev[made]) != hipSuccess) break;
  int rc = made == 2 * kPairs ? SG_OK : fail(SG_ERR_HIP, "hipEventCreate failed");
  double tot = 0;
  if (!rc) {
    for (int k = 0; k < kPairs; ++k) {
      (void)hipEventRecord(ev[2 * k], G.stream);
      sg::launch_empty(G.stream, G.d_cnt);
      (void)hipEventRecord(ev[2 * k + 1], G.stream);
    }
    if (hipStreamSynchronize(G.stream) != hipSuccess) rc = fail(SG_ERR_HIP, "synchronize failed");
    for (int k = 0; !rc && k < kPairs; ++k) {
      float t = 0;
      if (hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]) != hipSuccess)
        rc = fail(SG_ERR_HIP, "hipEventElapsedTime failed");
      tot += t;
    }
  }
  for (int k = 0; k < made; ++k) (void)hipEventDestroy(ev[k]);
  if (!rc) *ms = tot / kPairs;
  return rc;
}
const char *sg_kernel_name(int id) {
  return (id >= 0 && id < sg::kNumKernelIds) ? sg::kKernelNames[id] : "";
}
int sg_kernel_count(void) { return sg::kNumKernelIds; }

int sg_create(const sg_grid *grid, const sg_coeffs *coeffs, const sg_bc *bc, const double *f0,
              double dt, const sg_opts *opts, sg_solver **out) {
  if (!out) return fail(SG_ERR_ARG, "out must not be null");
  *out = nullptr;
  sg_batch *b = nullptr;
  int rc = sg_batch_create(opts, 0, &b);
  if (rc) return rc;
  int slot = -1;
  rc = sg_batch_add(b, grid, coeffs, bc, f0, dt, &slot);
  if (rc) { sg_batch_destroy(b); return rc; }
  *out = new sg_solver{b, slot};
  return SG_OK;
}
int sg_set_bc_values(sg_solver *h, int side, const double *values) {
  if (!h) return fail(SG_ERR_ARG, "null handle");
  return sg_batch_set_bc_values(h->bt, h->slot, side, values);
}
int sg_set_coeffs(sg_solver *h, const sg_coeffs *c) {
  if (!h) return fail(SG_ERR_ARG, "null handle");
  return sg_batch_set_coeffs(h->bt, h->slot, c);
}
int sg_step(sg_solver *h, int nsteps) {
  if (!h) return fail(SG_ERR_ARG, "null handle");
  return sg_batch_step(h->bt, nsteps);
}
int sg_get_f(sg_solver *h, double *out) {
  if (!h) return fail(SG_ERR_ARG, "null handle");
  return sg_batch_get_f(h->bt, h->slot, out);
}
int sg_set_f(sg_solver *h, const double *f, int64_t istep) {
  if (!h) return fail(SG_ERR_ARG, "null handle");
  return sg_batch_set_f(h->bt, h->slot, f, istep);
}
double sg_time(const sg_solver *h) {
  if (!h) return 0.0;
  int local;
  const Batch *G = h->bt->group_of(h->slot, &local);
  if (!G) return 0.0;
  const auto &s = G->slots[local];
  return (double)s.steps * s.dt;   // Solver::t(), source/Solver.h:23
}
int sg_stats(const sg_solver *h, sg_step_stats *out) {
  if (!h) return fail(SG_ERR_ARG, "null handle");
  return sg_batch_stats(h->bt, h->slot, out);
}
void sg_destroy(sg_solver *h) {
  if (!h) return;
  sg_batch_destroy(h->bt);
  delete h;
}
int sg_debug_assemble(sg_solver *h, double *bands5, double *R) {
  if (!h || !bands5 || !R) return fail(SG_ERR_ARG, "null argument");
  sg::HostLock lk(sg::g_capture_mu);
  int local;
  Batch *G = h->bt->group_of(h->slot, &local);
  if (!G) return fail(SG_ERR_ARG, "no instance");
  return G->debug_assemble(local, bands5, R);
}
int sg_solver_batch(sg_solver *h, sg_batch **out) {
  if (!h || !out) return fail(SG_ERR_ARG, "null argument");
  *out = h->bt;
  return SG_OK;
}

}  // extern "C"
