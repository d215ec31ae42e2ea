// Host driver and C ABI of the MI355X semi-discrete OT solver (include/caustic_ot.h).
//
// A context owns one batch of problems resident in HBM.  Sites are re-ordered by a bucket grid
// (8x8-bucket tiles, row-major inside a tile) so that a block of 128 consecutive sites is a
// compact patch of the plane and its candidate reads hit the same cache lines.
//
// The solver is the L-BFGS loop of direct_ot (ot_op.cuh:289-537) with the reference's rules.
// All N-length vectors (weights, direction, gradient, history S/Y) stay on the device; the
// two-loop recursion runs in coefficient space on the host from Gram matrices of the history
// (s_a.y_b, y_a.y_b) and the dot products s_a.g, y_a.g, which the accept kernel computes in the
// same pass that writes the new history pair.  The direction kernel then forms
// d = -(c_g g + sum c_y[h] y_h + sum c_s[h] s_h) in one pass.  Problems of a batch run in
// lockstep: every round gives each unfinished problem one line-search trial.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/caustic_ot.h"
#include "cot_internal.h"
#include "cot_layout.h"

static_assert(COT_MAX_VERTS == cot::kMaxVerts, "cot_cell_polygons writes kMaxVerts slots per cell");

namespace cot {
hipError_t launch_cell(const CellArgs& a, int nblocks, bool fp32, int int_lanes,
                       int repair_blocks, const VerifyArgs* check, hipStream_t st);
hipError_t launch_prefix(const double* rho, int pm, int pn, bool fp32, DensityGroup* grp,
                         double4* tot, hipStream_t st);
hipError_t launch_save_adjacency(const BlockDesc* b, const int* gate, const uint8_t* active,
                                 const uint8_t* nedge, const int* elab, uint8_t* adjn, int* adj,
                                 int* elist, int* ecount, int ecap, int nblocks, hipStream_t st);
hipError_t launch_raise_ring(const BlockDesc* b, const int* list, const int* list_count,
                             uint8_t* ring, int rmax, int repair_blocks, hipStream_t st);
hipError_t launch_adapt_ring(const BlockDesc* b, const uint8_t* used, uint8_t* ring, int rmin,
                             int rmax, int nblocks, hipStream_t st);
hipError_t launch_reduce_eval(const BlockDesc* b, const double* mass, const double* fcell,
                              const double* area, const double* nu, const uint8_t* exist,
                              const double* dir, double* part, int nblocks, hipStream_t st);
hipError_t launch_gradient(const BlockDesc* b, const double* mass, const double* nu, double* g,
                           int nblocks, hipStream_t st);
hipError_t launch_direction(const BlockDesc* b, const uint8_t* active, const double* coef,
                            const int* hcount, const double* S, const double* Y, size_t ld,
                            const double* g, double* d, double* part, int nblocks,
                            hipStream_t st);
hipError_t launch_accept(const BlockDesc* b, const uint8_t* active, const double* step,
                         const int* slot, const int* hcount, const double* mass, const double* nu,
                         const double* d, double* w, double* g, double* S, double* Y, size_t ld,
                         double* part, int nblocks, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t st);
hipError_t launch_fill_i32(int* p, size_t n, int v, hipStream_t st);
// kind 0: gather f64 (caller -> sorted); 1: scatter f64; 2: scatter f64 pairs; 3: scatter u8
hipError_t launch_permute(int kind, const void* src, const int* perm, size_t n, void* dst,
                          hipStream_t st);
hipError_t launch_dev_count(const DevState* st, int np, int* out, hipStream_t stream);
hipError_t launch_dev_direction(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                                const SolveParams& sp, const double* S, const double* Y,
                                size_t ld, const double* g, double* d, double* part,
                                size_t pstride, int* counter, uint8_t* active, int nblk,
                                hipStream_t stream);
hipError_t launch_dev_decide(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                             const SolveParams& sp, const double* mass, const double* fcell,
                             const double* area, const double* nu, const double* g,
                             const double* d, const int* fails,
                             int* overflow, double* part, size_t pstride, int* counter,
                             uint8_t* active, int* fail_counters, int n_fail_counters, int np,
                             int nblk, hipStream_t stream);
hipError_t launch_dev_accept(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                             const SolveParams& sp, const double* mass, const double* nu,
                             const double* d, double* w, double* g, double* S, double* Y,
                             size_t ld, double* part, size_t pstride, int* counter,
                             uint8_t* active, int nblk, hipStream_t stream);
hipError_t launch_dev_accept_finish(const ProbDesc* probs, DevState* st, const SolveParams& sp,
                                    const double* part, size_t pstride, uint8_t* active, int np,
                                    hipStream_t stream);

hipError_t launch_fresnel_weight(const BlockDesc* blocks, const double2* points, const double* ins,
                                 const double* wdir, const double* height,
                                 const double* clip_area, const double2* centroid, double z_len,
                                 double z_recv, double n_idx, double* weight, double* part,
                                 int nblk, hipStream_t st);
hipError_t launch_solid_angle(const BlockDesc* blocks, const double2* vtx, const uint8_t* nedge,
                              const double light[3], double z, double* omega, int nblk,
                              hipStream_t st);
hipError_t launch_splat(const float* pts, const float* energy, int n, int ry, int rx,
                        const double dom[4], float sigma, int r, const float* G, float c,
                        float* out, hipStream_t st);
hipError_t launch_tri_render(const float* energy, const int* faces, int n_faces, const float* xy,
                             float gray, const double dom[4], int pm, int pn, float* image,
                             hipStream_t st);
hipError_t launch_tri_grad(const float* energy, const int* faces, int n_faces, const float* xy,
                           int n_points, float gray, const float* real, const double dom[4],
                           int pm, int pn, float* render, float* diff, float* xy_render_grad,
                           float* tri_render_grad, float* inte_diff, float* xy_inte_grad,
                           float* inte_result, float* sob, float* xy_sobel_grad,
                           float* tri_sobel_grad, hipStream_t st);
hipError_t launch_scale(const BlockDesc* blocks, const double* weight, double scale, double* nu,
                        int nblk, hipStream_t st);
hipError_t launch_aa_store(const BlockDesc* blocks, const double* w, const double* g, double beta,
                           int slot, int H, double* G, double* R, size_t ld, double* part,
                           int nblk, hipStream_t st);
hipError_t launch_aa_direction(const BlockDesc* blocks, const double* alpha, int H,
                               const double* G, size_t ld, const double* w, double* d, int nblk,
                               hipStream_t st);
hipError_t launch_hessian(const ProbDesc* probs, const BlockDesc* blocks, const uint8_t* active,
                          const double2* xy, const double2* vtx, const uint8_t* nedge,
                          const int* elab, const DensityGroup* grp, const int* bagg, double* h,
                          int* hj, double* diag, int* sagg, int nblk, hipStream_t st);
hipError_t launch_coarse_rows(const BlockDesc* blocks, const int* bagg, const int* sagg,
                              const uint8_t* nedge, const int* hj, const double* h,
                              CoarseEntry* out, int nblk, hipStream_t st);
hipError_t launch_cg_coarse(const int4* tasks, int ntask, const CoarseDesc* cdesc,
                            const int* agg_start, const int* agg_blk, const double* rsum,
                            const double* ainv, double* cvec, double* part_rzc, hipStream_t st);
hipError_t launch_cg_init(const BlockDesc* blocks, const uint8_t* active, const double* g,
                          const double* shift, const double* diag, const uint8_t* nedge,
                          const int* hj, const double* h, int sweeps, const double* x0, double* x,
                          double* r, double* z, double* part_rz, double* part_rr, double* part_bb,
                          double* rsum, int nblk, hipStream_t st);
hipError_t launch_cg_iteration(const CgArgs& a, const double* p_old, double* p_new, int cur,
                               int first, hipStream_t st);
hipError_t launch_cg_small(const SmallCgArgs& a, hipStream_t st);
hipError_t launch_axpy_step(const BlockDesc* blocks, const uint8_t* active, const double* step,
                            const double* d, double* w, double* carry, int nblk, hipStream_t st);
hipError_t launch_render(const ProbDesc* probs, const BlockDesc* blocks, const double2* vtx,
                         const uint8_t* nedge, const double* nu, double* image, size_t pixels,
                         int nblk, hipStream_t st);
hipError_t launch_hessian_apply(const BlockDesc* blocks, const uint8_t* nedge, const int* hj,
                                const double* h, const double* diag, const double* p, double* q,
                                double* part_pq, int nblk, hipStream_t st);
hipError_t launch_axpy(const BlockDesc* blocks, double t, const double* d, double* w, int nblk,
                       hipStream_t st);
hipError_t launch_dev_adapt_ring(const BlockDesc* blocks, const DevState* st,
                                 const int* final_fails, const uint8_t* used, uint8_t* ring,
                                 int ring_min, int ring_max, int nblk, hipStream_t stream);
constexpr int kEvalPartials = 8;
constexpr int kAcceptPartials = 5 * kMaxHist;
constexpr int kCoefWidth = 1 + 2 * kMaxHist;
}  // namespace cot

using namespace cot;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(COT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

#define CHK(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != COT_OK) return rc_; \
  } while (0)

// Per-problem evaluation summary (host).
struct EvalSum {
  double f = 0, mass = 0, area = 0, maxrel = 0, zeros = 0, g2 = 0, gd = 0, empty = 0;
  int fails = 0, overflow = 0;
};

// Per-problem L-BFGS state (host scalars).
struct LState {
  int m = 10;
  int loop_iter = 0;
  bool restart = false;
  double H0 = 1.0;
  double SY[kMaxHist][kMaxHist] = {};
  double YY[kMaxHist][kMaxHist] = {};
  double Sg[kMaxHist] = {};
  double Yg[kMaxHist] = {};
  double rho[kMaxHist] = {};
  double alpha[kMaxHist] = {};
};

}  // namespace

struct cot_context {
  hipStream_t st = nullptr;
  cot_options opt{};
  int np = 0, ntot = 0, nblk = 0;
  bool fp32 = false;
  std::vector<ProbDesc> probs;
  std::vector<BlockDesc> blocks;
  std::vector<int> perm;  // sorted global index -> caller global index
  // the entry points on device arrays (cot_*_dev): inside one of them the caller's arrays are
  // device memory in caller order, moved by gather / scatter kernels over d_perm; the events
  // order the context's stream after the caller's and the caller's after the context's
  bool io_device = false;
  int* d_perm = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::vector<void*> allocs;
  // device
  ProbDesc* d_probs = nullptr;
  BlockDesc* d_blocks = nullptr;
  double2* d_xy = nullptr;
  double* d_nu = nullptr;
  int2* d_buckets = nullptr;
  DensityGroup* d_grp = nullptr;  // density groups (compact row prefixes) of the targets
  int* d_defer = nullptr;         // sites the lane full pass hands to its wave pass
  int* d_dcount = nullptr;        //   and their count
  double4* d_tot = nullptr;       // chunk totals
  double *d_w = nullptr, *d_dir = nullptr, *d_step = nullptr, *d_mass = nullptr;
  double *d_fcell = nullptr, *d_area = nullptr, *d_g = nullptr;
  double2 *d_cen = nullptr, *d_geo = nullptr;
  uint8_t *d_exist = nullptr, *d_nedge = nullptr, *d_ring = nullptr, *d_used = nullptr;
  uint8_t* d_active = nullptr;
  int* d_redo = nullptr;         // repair flags (see CellArgs)
  int* d_list = nullptr;         // two work lists of flagged sites, N slots each
  int* d_lcount = nullptr;       // their lengths
  double2* d_org = nullptr;      // FP32 geometry: clipping frame origins
  int* d_elab = nullptr;
  float2* d_espan = nullptr;
  int* d_adj = nullptr;          // previous evaluation's edge labels (candidates)
  uint8_t* d_adjn = nullptr;     // previous evaluation's edge counts (0: none yet)
  int* d_elist = nullptr;         // cells the last evaluation left empty (see CellArgs)
  int* d_ecount = nullptr;
  uint8_t* d_prebuilt = nullptr;
  bool have_adj = false;
  double2* d_vtx = nullptr;      // cell vertices (construction -> integration)
  int2* d_kbox = nullptr;
  int* d_xcnt = nullptr;         // candidates queued by the conformity check
  int* d_xcand = nullptr;
  unsigned long long* d_timer = nullptr;  // kTimers records of kTimerWords
  unsigned long long* d_diag = nullptr;   // repair counters (kDiagWords), after the timers
  double clock_hz = 1e8;                  // wall-clock rate of the in-kernel timers
  int *d_overflow = nullptr, *d_fails = nullptr;
  double* d_part = nullptr;
  double *d_S = nullptr, *d_Y = nullptr;
  int hist_alloc = 0;
  double* d_coef = nullptr;
  int *d_slot = nullptr, *d_hcount = nullptr;
  // Newton steps: edge Hessian entries, diagonal, CG vectors and partials
  int* d_hj = nullptr;  // Newton: neighbour of each Hessian edge slot (hslot order)
  double *d_hh = nullptr, *d_hdiag = nullptr, *d_nx = nullptr, *d_nr = nullptr, *d_nz = nullptr;
  double *d_np = nullptr, *d_nq = nullptr, *d_prz = nullptr, *d_ppq = nullptr, *d_prr = nullptr;
  double* d_ncarry = nullptr;  // the part of the last Newton direction its step left
  double* d_pbb = nullptr;     // partials of |rhs|^2
  double* d_np2 = nullptr;  // the other CG direction buffer (directions alternate)
  double* d_nrz = nullptr;  // r.z of the current and the previous CG iteration
  double* d_nshift = nullptr;
  // coarse space of the CG preconditioner, built for coarse_max patches per problem
  std::vector<int2> blk_cell;  // bucket (x, y) of each block's first site
  int coarse_max = 0, coarse_total = 0, coarse_ntask = 0;
  size_t coarse_mat = 0;  // doubles of all problems' dense patch matrices
  std::vector<CoarseDesc> cdesc;
  std::vector<int> bagg;   // block -> global patch (-1: none)
  int* d_bagg = nullptr;
  int* d_sagg = nullptr;   // site -> global patch, rewritten with every Hessian
  int *d_agg_start = nullptr, *d_agg_blk = nullptr;
  int* d_agg_prob = nullptr;  // patch -> problem
  double* d_cgout = nullptr;  // the small-system CG's results (4 doubles)
  CoarseDesc* d_cdesc = nullptr;
  int4* d_ctask = nullptr;
  double *d_ainv = nullptr, *d_cvec = nullptr, *d_prc = nullptr;
  double* d_rsum = nullptr;
  CoarseEntry* d_crows = nullptr;  // per-block couplings to other patches
  CoarseEntry* h_crows = nullptr;  // pinned copy
  double* h_ac = nullptr;          // pinned: the patch matrices, then their inverses (up)
  std::vector<double> nu_sum;  // per-problem sum of the source masses
  size_t img_total = 0;        // pixels of all targets
  double* d_img = nullptr;     // diagnostic render (allocated on first use)
  // device-resident solver
  DevState* d_state = nullptr;
  int* d_cnt = nullptr;        // last-block counters: direction, decide, accept (np each)
  double* d_pdir = nullptr;    // direction partials  [1][nblk]
  double* d_peval = nullptr;   // decide partials     [6][nblk]
  double* d_pacc = nullptr;    // accept partials     [5K][nblk]
  int* d_rfails = nullptr;     // per-round conformity failures [(R+1)][np]
  int rfail_rounds = 0;
  int* d_ndone = nullptr;
  int* h_ndone = nullptr;      // pinned, 2 slots
  // pinned staging
  double* h_part = nullptr;
  int* h_counts = nullptr;  // fails[np], overflow[np]
  double* h_vec = nullptr;  // ntot doubles
  double* h_small = nullptr;  // per-problem steps
  double* h_coef = nullptr;   // per-problem direction coefficients
  int* h_ismall = nullptr;
  uint8_t* h_active = nullptr;
  cot_stats stats{};

  ~cot_context() {
    if (st) hipStreamSynchronize(st);
    for (void* p : allocs) pool_free_device(p);
    for (void* p : {(void*)h_part, (void*)h_counts, (void*)h_vec, (void*)h_ac, (void*)h_crows,
                    (void*)h_small, (void*)h_coef, (void*)h_ismall, (void*)h_active,
                    (void*)h_ndone})
      pool_free_pinned(p);
    if (ev_in) hipEventDestroy(ev_in);
    if (ev_out) hipEventDestroy(ev_out);
    if (st) hipStreamDestroy(st);
  }
};

namespace {

// Pinned host memory from the cache (hipSuccess or hipErrorOutOfMemory, as hipHostMalloc).
hipError_t pinned_alloc(void** p, size_t bytes) {
  *p = pool_alloc_pinned(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}

template <typename T>
int dalloc(cot_context* c, T** p, size_t count) {
  void* q = pool_alloc_device(std::max<size_t>(count, 1) * sizeof(T));
  if (!q) return fail(COT_ENOMEM, "device allocation failed");
  c->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return COT_OK;
}

// Device buffers that live for one call: freed, after the stream drains, when the owner ends.
struct CallBuffers {
  hipStream_t st;
  std::vector<void*> held;
  explicit CallBuffers(hipStream_t s) : st(s) {}
  template <typename T>
  int get(T** out, size_t count) {
    void* q = pool_alloc_device(std::max<size_t>(count, 1) * sizeof(T));
    if (!q) return fail(COT_ENOMEM, "device allocation failed");
    held.push_back(q);
    *out = reinterpret_cast<T*>(q);
    return COT_OK;
  }
  ~CallBuffers() {
    if (!held.empty()) hipStreamSynchronize(st);
    for (void* q : held) pool_free_device(q);
  }
};

// Sum the per-block partials of the last reduce into per-problem summaries.
void sum_partials(cot_context* c, std::vector<EvalSum>& out) {
  out.assign(c->np, EvalSum());
  for (int p = 0; p < c->np; ++p) {
    EvalSum& e = out[p];
    e.maxrel = 0.0;
    const ProbDesc& pd = c->probs[p];
    for (int b = pd.blk_off; b < pd.blk_off + pd.nblk; ++b) {
      const double* q = c->h_part + (size_t)b * kEvalPartials;
      e.f += q[0];
      e.mass += q[1];
      e.area += q[2];
      e.maxrel = std::max(e.maxrel, q[3]);
      e.zeros += q[4];
      e.g2 += q[5];
      e.gd += q[6];
      e.empty += q[7];
    }
    e.fails = c->h_counts[p];
    e.overflow = c->h_counts[c->np + p];
  }
}

// Work list k % 2 of flagged sites and its length.
inline int* list(cot_context* c, int k) { return c->d_list + (size_t)(k & 1) * c->ntot; }
inline int* list_count(cot_context* c, int k) { return c->d_lcount + (k & 1); }

// Kernel arguments shared by every cell / conformity-check launch of the context.
void base_args(cot_context* c, CellArgs& ca, VerifyArgs& va) {
  ca = CellArgs{};
  ca.probs = c->d_probs;
  ca.blocks = c->d_blocks;
  ca.xy = c->d_xy;
  ca.w = c->d_w;
  ca.nu = c->d_nu;
  ca.buckets = c->d_buckets;
  ca.grp = c->d_grp;
  ca.tot = c->d_tot;
  ca.ring = c->d_ring;
  ca.redo = c->d_redo;
  ca.active = c->d_active;
  ca.g = c->d_g;
  ca.mass = c->d_mass;
  ca.fcell = c->d_fcell;
  ca.area = c->d_area;
  ca.cen = c->d_cen;
  ca.geo = c->d_geo;
  ca.exist = c->d_exist;
  ca.nedge = c->d_nedge;
  ca.elab = c->d_elab;
  ca.espan = c->d_espan;
  ca.ring_used = c->d_used;
  ca.overflow = c->d_overflow;
  ca.org = c->d_org;
  ca.ring_max = c->opt.ring_max;
  ca.repair_lane = c->opt.repair_mode;
  ca.pass_wave = c->opt.pass_mode;
  ca.adj = c->d_adj;
  ca.adjn = c->d_adjn;
  ca.xcnt = c->d_xcnt;
  ca.xcand = c->d_xcand;
  ca.vtx = c->d_vtx;
  ca.kbox = c->d_kbox;
  ca.defer_list = c->d_defer;
  ca.defer_count = c->d_dcount;
  ca.empty_list = c->d_elist;
  ca.empty_count = c->d_ecount;
  ca.empty_cap = (int)c->ntot;
  ca.prebuilt = c->d_prebuilt;
  va = VerifyArgs{};
  va.probs = c->d_probs;
  va.blocks = c->d_blocks;
  va.xy = c->d_xy;
  va.nedge = c->d_nedge;
  va.elab = c->d_elab;
  va.espan = c->d_espan;
  va.active = c->d_active;
  va.redo = c->d_redo;
  va.fails = c->d_fails;
  va.tol_rel = c->opt.verify_tol;
  va.xcnt = c->d_xcnt;
  va.xcand = c->d_xcand;
  ca.diag = c->d_diag;
  va.diag_fail = c->d_diag + kDiagCheckFull;
}

// Candidates: the FP64 full pass takes the previous evaluation's cells' two-ring (an edge flip
// makes a neighbour's neighbour adjacent, and the edge it creates starts too short for the
// conformity check to see).  Repairs take the cells just built (larger than the true cells, so
// their neighbour lists hold more) plus the sites the check queued, then the bucket rings, grown
// 2r + 1 before each round from the ring that held the cell's last neighbours.  Without a
// previous evaluation the full pass uses the rings.
constexpr int kAdjRepairRounds = 2;
inline int repair_cand(int r) { return r <= kAdjRepairRounds ? kCandAdj2 : kCandBucket; }
// FP32 geometry's full pass takes the one-ring by default (full_pass_ring 0 or 1): an edge a
// flip creates and the check cannot yet see changes a mass by ~1e-9 relative, far inside FP32's
// own error; full_pass_ring 2 takes the two-ring, which costs more per cell and leaves fewer
// cells to the repairs where weights move far between evaluations.  FP64 keeps the two-ring.
inline int full_pass_cand(const cot_context* c) {
  return (c->fp32 && c->opt.full_pass_ring != 2) ? kCandAdj1 : kCandAdj2;
}

// Fold the in-kernel timers into the context's statistics and clear their accumulators.
int pull_timers(cot_context* c) {
  unsigned long long tm[kTimers * kTimerWords + kDiagWords];
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(tm, c->d_timer, sizeof(tm), hipMemcpyDeviceToHost));
  const double inv = 1.0 / c->clock_hz;
  constexpr int kAcc = 2 * kTimerSlots, kCnt = 2 * kTimerSlots + 1;
  c->stats.cell_seconds += tm[kTimerCell * kTimerWords + kAcc] * inv;
  c->stats.cell_launches += (long long)tm[kTimerCell * kTimerWords + kCnt];
  c->stats.repair_seconds += tm[kTimerRepair * kTimerWords + kAcc] * inv;
  c->stats.repair_launches += (long long)tm[kTimerRepair * kTimerWords + kCnt];
  c->stats.integ_seconds += tm[kTimerInteg * kTimerWords + kAcc] * inv;
  c->stats.integ_launches += (long long)tm[kTimerInteg * kTimerWords + kCnt];
  c->stats.repair_seconds += tm[kTimerIntegRepair * kTimerWords + kAcc] * inv;
  for (int k = 0; k < kTimers; ++k) {
    tm[k * kTimerWords + kAcc] = 0;
    tm[k * kTimerWords + kCnt] = 0;
  }
  const unsigned long long* dg = tm + kTimers * kTimerWords;
  c->stats.handoff_degenerate += (long long)dg[kDiagDegenerate];
  c->stats.handoff_overflow += (long long)dg[kDiagOverflow];
  c->stats.check_failures += (long long)dg[kDiagCheckFull];
  c->stats.repair_check_failures += (long long)dg[kDiagCheckRepair];
  c->stats.repaired_cells += (long long)dg[kDiagRebuilt];
  c->stats.reappeared_cells += (long long)dg[kDiagReappeared];
  std::fill(tm + kTimers * kTimerWords, tm + kTimers * kTimerWords + kDiagWords, 0ull);
  HIPCHK(hipMemcpy(c->d_timer, tm, sizeof(tm), hipMemcpyHostToDevice));
  return COT_OK;
}

// One evaluation of every active problem at w (+ step * dir when dir is set), with the
// conformity check and repairs.  Leaves mass/fcell/... on the device and summaries on the host.
int evaluate_device(cot_context* c, const double* d_w, const double* d_dir,
                    const std::vector<uint8_t>& active, std::vector<EvalSum>& sum,
                    int max_rounds = -1) {
  if (max_rounds < 0) max_rounds = c->opt.max_repairs;
  hipStream_t st = c->st;
  std::memcpy(c->h_active, active.data(), c->np);
  HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, c->np, hipMemcpyHostToDevice, st));
  HIPCHK(launch_fill_i32(c->d_fails, 2 * (size_t)c->np, 0, st));
  HIPCHK(launch_fill_i32(c->d_redo, c->ntot, 0, st));

  CellArgs ca;
  VerifyArgs va;
  base_args(c, ca, va);
  ca.w = d_w;
  ca.dir = d_dir;
  ca.step = c->d_step;
  ca.cand = c->have_adj ? full_pass_cand(c) : kCandBucket;
  ca.timer = c->d_timer + kTimerCell * kTimerWords;
  ca.timer_int = c->d_timer + kTimerInteg * kTimerWords;
  ca.coarse = c->d_fails;
  ca.reset_count = list_count(c, 0);
  va.list = list(c, 0);
  va.list_count = list_count(c, 0);

  HIPCHK(launch_cell(ca, c->nblk, c->fp32, c->opt.int_lanes, c->opt.repair_blocks, &va, st));
  HIPCHK(launch_reduce_eval(c->d_blocks, c->d_mass, c->d_fcell, c->d_area, c->d_nu, c->d_exist,
                            d_dir, c->d_part, c->nblk, st));
  HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * kEvalPartials * c->nblk,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(c->h_counts, c->d_fails, sizeof(int) * 2 * c->np, hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  c->stats.evaluations += 1;

  ca.redo_mode = 1;
  ca.coarse = nullptr;
  ca.timer = c->d_timer + kTimerRepair * kTimerWords;
  ca.timer_int = c->d_timer + kTimerIntegRepair * kTimerWords;
  va.diag_fail = c->d_diag + kDiagCheckRepair;
  int rounds = 0;
  for (;;) {
    int tot = 0;
    for (int p = 0; p < c->np; ++p) tot += active[p] ? c->h_counts[p] : 0;
    if (tot == 0 || rounds >= max_rounds) break;
    ++rounds;
    // round r rebuilds the list the previous check filled and its check fills the other one
    ca.list = list(c, rounds - 1);
    ca.list_count = list_count(c, rounds - 1);
    ca.reset_count = list_count(c, rounds);
    va.list = list(c, rounds);
    va.list_count = list_count(c, rounds);
    ca.cand = repair_cand(rounds);
    if (ca.cand == kCandBucket)
      HIPCHK(launch_raise_ring(c->d_blocks, ca.list, ca.list_count, c->d_ring, c->opt.ring_max,
                               c->opt.repair_blocks, st));
    else
      HIPCHK(launch_save_adjacency(c->d_blocks, c->d_fails, c->d_active, c->d_nedge, c->d_elab,
                                   c->d_adjn, c->d_adj, nullptr, nullptr, 0, c->nblk, st));
    HIPCHK(launch_fill_i32(c->d_fails, (size_t)c->np, 0, st));
    HIPCHK(launch_cell(ca, c->nblk, c->fp32, c->opt.int_lanes, c->opt.repair_blocks, &va, st));
    HIPCHK(launch_reduce_eval(c->d_blocks, c->d_mass, c->d_fcell, c->d_area, c->d_nu,
                              c->d_exist, d_dir, c->d_part, c->nblk, st));
    HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * kEvalPartials * c->nblk,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(c->h_counts, c->d_fails, sizeof(int) * 2 * c->np,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  c->stats.repair_rounds += rounds;
  sum_partials(c, sum);
  // the cover check: cells that passed the edge check but do not tile the target (a consistent
  // multiple cover) leave the evaluation unverified, and the next evaluation searches bucket
  // rings rather than their adjacency
  bool covered = true;
  for (int p = 0; p < c->np; ++p)
    if (active[p] && sum[p].fails == 0 && !covers_domain(c->probs[p], sum[p].area)) {
      sum[p].fails = 1;
      covered = false;
    }
  // the cells just built are the next evaluation's candidates
  HIPCHK(launch_save_adjacency(c->d_blocks, nullptr, c->d_active, c->d_nedge, c->d_elab,
                               c->d_adjn, c->d_adj, c->d_elist, c->d_ecount, (int)c->ntot,
                               c->nblk, st));
  c->have_adj = covered;
  // after a clean check the next bucket search starts from the outermost used ring plus one;
  // after an unresolved one the raised rings stay
  bool clean = true;
  for (int p = 0; p < c->np; ++p) clean = clean && (!active[p] || sum[p].fails == 0);
  if (clean)
    HIPCHK(launch_adapt_ring(c->d_blocks, c->d_used, c->d_ring, 1, c->opt.ring_max, c->nblk, st));
  return COT_OK;
}

// Scatter caller-order host values into sorted order and upload.
int upload_sorted(cot_context* c, const double* src, double* dst) {
  if (c->io_device) {
    HIPCHK(launch_permute(0, src, c->d_perm, c->ntot, dst, c->st));
    return COT_OK;
  }
  for (int r = 0; r < c->ntot; ++r) c->h_vec[r] = src[c->perm[r]];
  HIPCHK(hipMemcpyAsync(dst, c->h_vec, sizeof(double) * c->ntot, hipMemcpyHostToDevice, c->st));
  return COT_OK;
}

int download_sorted(cot_context* c, const double* src, double* dst) {
  if (c->io_device) {
    HIPCHK(launch_permute(1, src, c->d_perm, c->ntot, dst, c->st));
    return COT_OK;
  }
  HIPCHK(hipMemcpyAsync(c->h_vec, src, sizeof(double) * c->ntot, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int r = 0; r < c->ntot; ++r) dst[c->perm[r]] = c->h_vec[r];
  return COT_OK;
}

int download_sorted2(cot_context* c, const double2* src, double* dst) {
  if (c->io_device) {
    HIPCHK(launch_permute(2, src, c->d_perm, c->ntot, dst, c->st));
    return COT_OK;
  }
  std::vector<double> tmp(2 * (size_t)c->ntot);
  HIPCHK(hipMemcpyAsync(tmp.data(), src, sizeof(double2) * c->ntot, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int r = 0; r < c->ntot; ++r) {
    dst[2 * c->perm[r]] = tmp[2 * r];
    dst[2 * c->perm[r] + 1] = tmp[2 * r + 1];
  }
  return COT_OK;
}

int download_sorted_u8(cot_context* c, const uint8_t* src, unsigned char* dst) {
  if (c->io_device) {
    HIPCHK(launch_permute(3, src, c->d_perm, c->ntot, dst, c->st));
    return COT_OK;
  }
  std::vector<uint8_t> tmp(c->ntot);
  HIPCHK(hipMemcpyAsync(tmp.data(), src, c->ntot, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int r = 0; r < c->ntot; ++r) dst[c->perm[r]] = tmp[r];
  return COT_OK;
}

int ensure_history(cot_context* c, int m) {
  if (c->hist_alloc >= m) return COT_OK;
  CHK(dalloc(c, &c->d_S, (size_t)m * c->ntot));
  CHK(dalloc(c, &c->d_Y, (size_t)m * c->ntot));
  c->hist_alloc = m;
  return COT_OK;
}

// Two-loop recursion of ot_op.cuh:298-337 in coefficient space: d = dg*g + sum qy*y + sum qs*s.
// Writes coef = [cg, cy[], cs[]] for d = -(cg g + sum cy y + sum cs s); returns the history count.
int two_loop(LState& L, double* coef) {
  const int K = kMaxHist;
  std::fill(coef, coef + kCoefWidth, 0.0);
  const int maxidx = std::min(L.loop_iter, L.m);
  double dg = -1.0;
  double qy[K] = {}, qs[K] = {};
  if (!L.restart) {
    for (int i = 1; i <= maxidx; ++i) {
      const int idx = ((L.loop_iter - i) % L.m + L.m) % L.m;
      double sd = dg * L.Sg[idx];
      for (int b = 0; b < maxidx; ++b) sd += qy[b] * L.SY[idx][b];
      L.alpha[idx] = sd * L.rho[idx];
      qy[idx] -= L.alpha[idx];
    }
    dg *= L.H0;
    for (int b = 0; b < maxidx; ++b) qy[b] *= L.H0;
    for (int i = maxidx; i >= 1; --i) {
      const int idx = ((L.loop_iter - i) % L.m + L.m) % L.m;
      double yd = dg * L.Yg[idx];
      for (int b = 0; b < maxidx; ++b) yd += qy[b] * L.YY[idx][b] + qs[b] * L.SY[b][idx];
      const double beta = L.rho[idx] * yd;
      qs[idx] += L.alpha[idx] - beta;
    }
  }
  coef[0] = -dg;
  for (int b = 0; b < K; ++b) {
    coef[1 + b] = -qy[b];
    coef[1 + K + b] = -qs[b];
  }
  return L.restart ? 0 : maxidx;
}

int ensure_device_solver(cot_context* c, int repair_rounds) {
  if (!c->d_state) {
    CHK(dalloc(c, &c->d_state, c->np));
    CHK(dalloc(c, &c->d_cnt, 3 * (size_t)c->np));
    CHK(dalloc(c, &c->d_pdir, c->nblk));
    CHK(dalloc(c, &c->d_peval, 6 * (size_t)c->nblk));
    CHK(dalloc(c, &c->d_pacc, (size_t)kAcceptPartials * c->nblk));
    CHK(dalloc(c, &c->d_ndone, 2));
    if (pinned_alloc((void**)&c->h_ndone, 2 * sizeof(int)) != hipSuccess)
      return fail(COT_ENOMEM, "pinned allocation failed");
    HIPCHK(hipMemset(c->d_cnt, 0, 3 * sizeof(int) * c->np));
  }
  if (c->rfail_rounds < repair_rounds + 1) {
    CHK(dalloc(c, &c->d_rfails, (size_t)(repair_rounds + 1) * c->np));
    c->rfail_rounds = repair_rounds + 1;
  }
  HIPCHK(hipMemset(c->d_rfails, 0, sizeof(int) * c->rfail_rounds * c->np));
  return COT_OK;
}

// One device round (see cot_solver.hip).  Captured once per solve into a hipGraph; the
// previous evaluation's adjacency must exist (the solve starts with a host-driven evaluation).
int launch_round(cot_context* c, const SolveParams& sp, int R) {
  hipStream_t st = c->st;
  const size_t ld = c->ntot, ps = c->nblk;
  const int np = c->np;
  HIPCHK(launch_dev_direction(c->d_probs, c->d_blocks, c->d_state, sp, c->d_S, c->d_Y, ld, c->d_g,
                              c->d_dir, c->d_pdir, ps, c->d_cnt, c->d_active, c->nblk, st));
  CellArgs ca;
  VerifyArgs va;
  base_args(c, ca, va);
  ca.dir = c->d_dir;
  ca.st = c->d_state;
  ca.cand = full_pass_cand(c);
  ca.timer = c->d_timer + kTimerCell * kTimerWords;
  ca.timer_int = c->d_timer + kTimerInteg * kTimerWords;
  ca.coarse = c->d_rfails;
  ca.reset_count = list_count(c, 0);
  va.fails = c->d_rfails;
  va.list = list(c, 0);
  va.list_count = list_count(c, 0);
  HIPCHK(launch_cell(ca, c->nblk, c->fp32, c->opt.int_lanes, c->opt.repair_blocks, &va, st));
  ca.redo_mode = 1;
  ca.coarse = nullptr;
  ca.timer = c->d_timer + kTimerRepair * kTimerWords;
  ca.timer_int = c->d_timer + kTimerIntegRepair * kTimerWords;
  va.diag_fail = c->d_diag + kDiagCheckRepair;
  for (int r = 1; r <= R; ++r) {
    const int* gate = c->d_rfails + (size_t)(r - 1) * np;
    // round r rebuilds the list the previous check filled and its check fills the other one
    ca.list = list(c, r - 1);
    ca.list_count = list_count(c, r - 1);
    ca.reset_count = list_count(c, r);
    va.list = list(c, r);
    va.list_count = list_count(c, r);
    ca.cand = repair_cand(r);
    if (ca.cand == kCandBucket)
      HIPCHK(launch_raise_ring(c->d_blocks, ca.list, ca.list_count, c->d_ring, c->opt.ring_max,
                               c->opt.repair_blocks, st));
    else
      HIPCHK(launch_save_adjacency(c->d_blocks, gate, c->d_active, c->d_nedge, c->d_elab,
                                   c->d_adjn, c->d_adj, nullptr, nullptr, 0, c->nblk, st));
    va.gate = gate;
    va.fails = c->d_rfails + (size_t)r * np;
    HIPCHK(launch_cell(ca, c->nblk, c->fp32, c->opt.int_lanes, c->opt.repair_blocks, &va, st));
  }
  HIPCHK(launch_save_adjacency(c->d_blocks, nullptr, c->d_active, c->d_nedge, c->d_elab,
                               c->d_adjn, c->d_adj, c->d_elist, c->d_ecount, (int)c->ntot,
                               c->nblk, st));
  const int* final_fails = c->d_rfails + (size_t)R * np;
  HIPCHK(launch_dev_adapt_ring(c->d_blocks, c->d_state, final_fails, c->d_used, c->d_ring, 1,
                               c->opt.ring_max, c->nblk, st));
  HIPCHK(launch_dev_decide(c->d_probs, c->d_blocks, c->d_state, sp, c->d_mass, c->d_fcell,
                           c->d_area, c->d_nu, c->d_g, c->d_dir, final_fails, c->d_overflow,
                           c->d_peval, ps,
                           c->d_cnt + np, c->d_active, c->d_rfails, R + 1, np, c->nblk, st));
  HIPCHK(launch_dev_accept(c->d_probs, c->d_blocks, c->d_state, sp, c->d_mass, c->d_nu,
                           c->d_dir, c->d_w, c->d_g, c->d_S, c->d_Y, ld, c->d_pacc, ps,
                           c->d_cnt + 2 * np, c->d_active, c->nblk, st));
  HIPCHK(launch_dev_accept_finish(c->d_probs, c->d_state, sp, c->d_pacc, ps, c->d_active, np, st));
  return COT_OK;
}

// Device-resident solve: the host only replays the round graph and polls a finished count.
int solve_device(cot_context* c, double* w, const cot_solve_opts& o, cot_report* reports) {
  const auto t_start = std::chrono::steady_clock::now();
  const int np = c->np;
  const int R = o.repair_rounds;
  CHK(ensure_history(c, o.history));
  CHK(ensure_device_solver(c, R));
  CHK(upload_sorted(c, w, c->d_w));

  // initial evaluation (host path, with its own conformity repairs)
  std::vector<uint8_t> act(np, 1);
  std::vector<EvalSum> cur;
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  HIPCHK(launch_gradient(c->d_blocks, c->d_mass, c->d_nu, c->d_g, c->nblk, c->st));
  std::vector<DevState> hs(np);
  for (int p = 0; p < np; ++p) {
    DevState& s = hs[p];
    std::memset(&s, 0, sizeof(DevState));
    s.H0 = 1.0;
    s.t = o.initial_step;
    s.f_cur = cur[p].f;
    s.g2_cur = cur[p].g2;
    s.zeros_cur = cur[p].zeros;
    s.maxrel_cur = cur[p].maxrel;
    s.evals = 1;
    s.coef[0] = 1.0;  // first direction: -g
    s.phase = kPhaseDir;
    if (cur[p].fails) s.status = COT_EUNVERIFIED;
    if ((o.stop_rule == COT_STOP_MAXREL && cur[p].maxrel < o.tol) ||
        (o.stop_rule == COT_STOP_HANDOFF && cur[p].maxrel < o.tol && cur[p].zeros == 0.0)) {
      s.conv = 1;
      s.phase = kPhaseDone;
    }
    c->h_active[p] = 0;
  }
  HIPCHK(hipMemcpyAsync(c->d_state, hs.data(), sizeof(DevState) * np, hipMemcpyHostToDevice,
                        c->st));
  HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, np, hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_fill_i32(c->d_fails, 2 * (size_t)np, 0, c->st));

  SolveParams sp{};
  sp.m = o.history;
  sp.max_iter = o.max_iter;
  sp.stop_rule = o.stop_rule == COT_STOP_MAXREL ? 1 : (o.stop_rule == COT_STOP_HANDOFF ? 2 : 0);
  sp.mixed_ls = o.accept_rule;
  sp.ring_max = c->opt.ring_max;
  sp.max_unverified = 4;
  sp.tol = o.tol;
  sp.t0 = o.initial_step;
  sp.h0_thr = o.h0_threshold;

  // capture one round (device_loop 2: launch the same rounds directly, no graph)
  const bool eager = o.device_loop == 2;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHK(hipStreamSynchronize(c->st));
  if (!eager) {
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int rc = launch_round(c, sp, R);
    hipError_t ce = hipStreamEndCapture(c->st, &graph);
    if (rc != COT_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIPCHK(ce);
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  struct GraphGuard {
    hipGraph_t g;
    hipGraphExec_t e;
    ~GraphGuard() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } guard{graph, exec};

  // replay in batches of rounds, then read the finished count
  const int kBatch = 8;
  const long long max_rounds = (long long)(o.max_iter + 1) * 20 + 64;
  long long rounds = 0;
  auto last_print = std::chrono::steady_clock::now();
  for (;;) {
    for (int k = 0; k < kBatch; ++k) {
      if (eager) {
        CHK(launch_round(c, sp, R));
      } else {
        HIPCHK(hipGraphLaunch(exec, c->st));
      }
    }
    rounds += kBatch;
    HIPCHK(launch_dev_count(c->d_state, np, c->d_ndone, c->st));
    HIPCHK(hipMemcpyAsync(c->h_ndone, c->d_ndone, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const int left = c->h_ndone[0];
    if (left == 0 || rounds > max_rounds) break;
    if (o.verbose > 0) {
      // progress line at most every two seconds: the first unfinished problem's state
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last_print).count() >= 2.0) {
        last_print = now;
        HIPCHK(hipMemcpy(hs.data(), c->d_state, sizeof(DevState) * np, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; ++p) {
          if (hs[p].phase == kPhaseDone) continue;
          std::printf("[cot] %d left; prob %d iter %d evals %d maxrel %.3e step %g (%.1fs)\n",
                      left, p, hs[p].iters, hs[p].evals, hs[p].maxrel_cur, hs[p].t,
                      std::chrono::duration<double>(now - t_start).count());
          std::fflush(stdout);
          break;
        }
      }
    }
  }
  HIPCHK(hipMemcpy(hs.data(), c->d_state, sizeof(DevState) * np, hipMemcpyDeviceToHost));
  CHK(download_sorted(c, c->d_w, w));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  int result = COT_OK;
  for (int p = 0; p < np; ++p) {
    const DevState& s = hs[p];
    int status = s.status;
    if (!s.conv && status == COT_OK) status = COT_ENOTCONV;
    if (status != COT_OK && result == COT_OK) result = status;
    c->stats.evaluations += s.evals - 1;  // the initial evaluation is already counted
    if (reports) {
      cot_report& r = reports[p];
      r.iterations = s.iters;
      r.evaluations = s.evals;
      r.converged = s.conv;
      r.status = status;
      r.f = s.f_cur;
      r.grad_norm = std::sqrt(s.g2_cur);
      r.max_rel_err = s.maxrel_cur;
      r.seconds = secs;
    }
    if (o.verbose > 0)
      std::printf("[cot] prob %d: %s after %d iterations (%d evaluations), maxrel %.3e, "
                  "%.2f ms\n",
                  p, s.conv ? "converged" : "stopped", s.iters, s.evals, s.maxrel_cur,
                  1e3 * secs);
  }
  if (result == COT_ENOTCONV) g_err = "iteration limit reached";
  if (result == COT_EUNVERIFIED) g_err = "cells failed the conformity check at the largest ring";
  return result;
}

}  // namespace

// Accessors for the host drivers built on the C ABI (cot_ladder.cpp).
namespace cot {
const cot_options* context_options(const cot_context* c) { return &c->opt; }
int context_problems(const cot_context* c) { return c->np; }
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace cot

extern "C" {

void cot_default_options(cot_options* o) {
  o->precision = COT_PREC_FP64;
  o->sites_per_bucket = 1.0;
  o->ring_init = 2;
  o->ring_max = 32;
  o->verify_tol = 1e-4;
  o->max_repairs = 12;
  o->int_lanes = kIntLanes;
  o->repair_blocks = kRepairBlocks;
  o->full_pass_ring = 0;
  o->repair_mode = 0;
  o->pass_mode = 0;
}

void cot_default_solve_opts(cot_solve_opts* o) {
  o->history = 10;
  o->max_iter = 50001;
  o->tol = 1e-5;
  o->initial_step = 1.0;
  o->stop_rule = COT_STOP_MAXREL;
  o->h0_threshold = 1e-5;
  o->verbose = 0;
  o->device_loop = 1;
  o->accept_rule = 0;
  o->repair_rounds = 3;
  o->newton_sweeps = 2;
  o->newton_coarse = 256;
  o->small_cg = 1;
}

const char* cot_last_error(void) { return g_err.c_str(); }

int cot_total_sites(const cot_context* c) { return c ? c->ntot : 0; }

int cot_create(const cot_problem* problems, int n_problems, const cot_options* opt,
               cot_context** out) {
  if (!problems || n_problems <= 0 || !out) return fail(COT_EINVAL, "no problems");
  for (int p = 0; p < n_problems; ++p) {
    const cot_problem& pr = problems[p];
    if (pr.n_sites <= 0 || !pr.sites || !pr.nu || !pr.rho || pr.pm <= 0 || pr.pn <= 0)
      return fail(COT_EINVAL, "problem " + std::to_string(p) + ": empty or missing arrays");
    if (!(pr.proj_domain[1] > pr.proj_domain[0]) || !(pr.proj_domain[3] > pr.proj_domain[2]))
      return fail(COT_EINVAL, "problem " + std::to_string(p) + ": empty proj_domain");
  }
  cot_context* c = new cot_context();
  int rc = COT_OK;
  auto bail = [&](int code) {
    delete c;
    return code;
  };
  if (opt) {
    c->opt = *opt;
  } else {
    cot_default_options(&c->opt);
  }
  c->opt.ring_max = std::min(std::max(c->opt.ring_max, 1), 250);
  c->opt.ring_init = std::min(std::max(c->opt.ring_init, 0), c->opt.ring_max);
  if (!(c->opt.sites_per_bucket > 0)) c->opt.sites_per_bucket = 1.0;
  if (!(c->opt.verify_tol > 0)) c->opt.verify_tol = 1e-4;
  if (c->opt.max_repairs < 0) c->opt.max_repairs = 0;
  if (c->opt.int_lanes != 4 && c->opt.int_lanes != 8 && c->opt.int_lanes != 16)
    return bail(fail(COT_EINVAL, "int_lanes must be 4, 8 or 16"));
  if (c->opt.repair_blocks < 1 || c->opt.repair_blocks > 65535)
    return bail(fail(COT_EINVAL, "repair_blocks must be in [1, 65535]"));
  if (c->opt.full_pass_ring < 0 || c->opt.full_pass_ring > 2)
    return bail(fail(COT_EINVAL, "full_pass_ring must be 0, 1 or 2"));
  if (c->opt.repair_mode != 0 && c->opt.repair_mode != 1)
    return bail(fail(COT_EINVAL, "repair_mode must be 0 or 1"));
  if (c->opt.pass_mode != 0 && c->opt.pass_mode != 1)
    return bail(fail(COT_EINVAL, "pass_mode must be 0 or 1"));
  c->fp32 = c->opt.precision == COT_PREC_MIXED;
  c->np = n_problems;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(COT_EHIP, "hipStreamCreate failed"));
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0)
      return bail(fail(COT_EHIP, "cannot read the device wall-clock rate"));
    c->clock_hz = 1e3 * khz;
  }

  // layout
  std::vector<int2> all_buckets;
  std::vector<double2> xy;
  std::vector<double> nu;
  size_t img_total = 0, tot_total = 0, grp_total = 0;
  const int gsize = c->fp32 ? kGroup32 : kGroup64;
  int site_off = 0;
  for (int p = 0; p < n_problems; ++p) {
    const cot_problem& pr = problems[p];
    ProbDesc pd{};
    pd.site_off = site_off;
    pd.n = pr.n_sites;
    pd.pm = pr.pm;
    pd.pn = pr.pn;
    pd.pre_off = (int)img_total;
    pd.nchunk = (pr.pm + kChunk - 1) / kChunk;
    pd.tot_off = (int)tot_total;
    pd.gsize = gsize;
    pd.ngrp = (pr.pm + gsize - 1) / gsize;
    pd.grp_off = (int)grp_total;
    pd.x0 = pr.proj_domain[0];
    pd.x1 = pr.proj_domain[1];
    pd.y0 = pr.proj_domain[2];
    pd.y1 = pr.proj_domain[3];
    pd.dx = (pd.x1 - pd.x0) / pr.pm;
    pd.dy = (pd.y1 - pd.y0) / pr.pn;
    std::vector<int> order;
    std::vector<int2> bk;
    bucket_sort(pr, c->opt.sites_per_bucket, pd, order, bk);
    pd.bkt_off = (int)all_buckets.size();
    all_buckets.insert(all_buckets.end(), bk.begin(), bk.end());
    pd.blk_off = (int)c->blocks.size();
    for (int f = 0; f < pr.n_sites; f += kCellBlock) {
      BlockDesc b{};
      b.prob = p;
      b.first = site_off + f;
      b.count = std::min(kCellBlock, pr.n_sites - f);
      c->blocks.push_back(b);
      const int o = order[f];
      const int bx = (int)std::floor((pr.sites[2 * o] - pd.bx0) * pd.binv);
      const int by = (int)std::floor((pr.sites[2 * o + 1] - pd.by0) * pd.binv);
      c->blk_cell.push_back(make_int2(std::min(std::max(bx, 0), pd.gx - 1),
                                      std::min(std::max(by, 0), pd.gy - 1)));
    }
    pd.nblk = (int)c->blocks.size() - pd.blk_off;
    double nus = 0.0;
    for (int r = 0; r < pr.n_sites; ++r) {
      int o = order[r];
      c->perm.push_back(site_off + o);
      xy.push_back(make_double2(pr.sites[2 * o], pr.sites[2 * o + 1]));
      nu.push_back(pr.nu[o]);
      nus += pr.nu[o];
    }
    c->nu_sum.push_back(nus);
    img_total += (size_t)pr.pm * pr.pn;
    tot_total += (size_t)pd.nchunk * pr.pn;
    grp_total += (size_t)pd.ngrp * pr.pn;
    if (img_total > (size_t)INT32_MAX || grp_total > (size_t)INT32_MAX)
      return bail(fail(COT_EINVAL, "images too large"));
    c->probs.push_back(pd);
    site_off += pr.n_sites;
  }
  c->ntot = site_off;
  c->nblk = (int)c->blocks.size();
  const size_t N = c->ntot;

  if ((rc = dalloc(c, &c->d_probs, c->np)) || (rc = dalloc(c, &c->d_blocks, c->nblk)) ||
      (rc = dalloc(c, &c->d_xy, N)) || (rc = dalloc(c, &c->d_nu, N)) ||
      (rc = dalloc(c, &c->d_buckets, all_buckets.size())) ||
      (rc = dalloc(c, &c->d_w, N)) || (rc = dalloc(c, &c->d_dir, N)) ||
      (rc = dalloc(c, &c->d_step, c->np)) || (rc = dalloc(c, &c->d_mass, N)) ||
      (rc = dalloc(c, &c->d_fcell, N)) || (rc = dalloc(c, &c->d_area, N)) ||
      (rc = dalloc(c, &c->d_g, N)) || (rc = dalloc(c, &c->d_cen, N)) ||
      (rc = dalloc(c, &c->d_geo, N)) || (rc = dalloc(c, &c->d_exist, N)) ||
      (rc = dalloc(c, &c->d_nedge, N)) || (rc = dalloc(c, &c->d_ring, N)) ||
      (rc = dalloc(c, &c->d_used, N)) || (rc = dalloc(c, &c->d_redo, N)) ||
      (rc = dalloc(c, &c->d_list, 2 * N)) || (rc = dalloc(c, &c->d_lcount, 2)) ||
      (rc = dalloc(c, &c->d_active, c->np)) || (rc = dalloc(c, &c->d_elab, N * kMaxVerts)) ||
      (rc = dalloc(c, &c->d_espan, N * kMaxVerts)) ||
      (rc = dalloc(c, &c->d_overflow, c->np)) || (rc = dalloc(c, &c->d_fails, 2 * c->np)) ||
      (rc = dalloc(c, &c->d_part, (size_t)c->nblk * std::max(kEvalPartials, kAcceptPartials))) ||
      (rc = dalloc(c, &c->d_coef, (size_t)c->np * kCoefWidth)) ||
      (rc = dalloc(c, &c->d_slot, c->np)) || (rc = dalloc(c, &c->d_hcount, c->np)) ||
      (rc = dalloc(c, &c->d_org, N)) || (rc = dalloc(c, &c->d_adj, N * kMaxVerts)) ||
      (rc = dalloc(c, &c->d_adjn, N)) || (rc = dalloc(c, &c->d_elist, N)) ||
      (rc = dalloc(c, &c->d_ecount, 1)) || (rc = dalloc(c, &c->d_prebuilt, N)) ||
      (rc = dalloc(c, &c->d_xcnt, N)) ||
      (rc = dalloc(c, &c->d_vtx, N * kMaxVerts)) || (rc = dalloc(c, &c->d_kbox, N)) ||
      (rc = dalloc(c, &c->d_xcand, N * kMaxExtra)) ||
      (rc = dalloc(c, &c->d_defer, N)) || (rc = dalloc(c, &c->d_dcount, 1)) ||
      (rc = dalloc(c, &c->d_timer, (size_t)kTimers * kTimerWords + kDiagWords)))
    return bail(rc);
  c->d_diag = c->d_timer + (size_t)kTimers * kTimerWords;
  // overflow counters live right after the fail counters so one copy reads both
  c->d_overflow = c->d_fails + c->np;
  c->img_total = img_total;
  if ((rc = dalloc(c, &c->d_grp, grp_total)) || (rc = dalloc(c, &c->d_tot, tot_total)))
    return bail(rc);
  if (pinned_alloc((void**)&c->h_part,
                    sizeof(double) * c->nblk * std::max(kEvalPartials, kAcceptPartials)) ||
      pinned_alloc((void**)&c->h_counts, sizeof(int) * 2 * c->np) ||
      pinned_alloc((void**)&c->h_vec, sizeof(double) * N) ||
      pinned_alloc((void**)&c->h_small, sizeof(double) * c->np) ||
      pinned_alloc((void**)&c->h_coef, sizeof(double) * c->np * kCoefWidth) ||
      pinned_alloc((void**)&c->h_ismall, sizeof(int) * 2 * c->np) ||
      pinned_alloc((void**)&c->h_active, c->np))
    return bail(fail(COT_ENOMEM, "pinned allocation failed"));

  // uploads (synchronous: inputs are caller-owned)
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up(c->d_probs, c->probs.data(), sizeof(ProbDesc) * c->np) &&
            up(c->d_blocks, c->blocks.data(), sizeof(BlockDesc) * c->nblk) &&
            up(c->d_xy, xy.data(), sizeof(double2) * N) &&
            up(c->d_nu, nu.data(), sizeof(double) * N) &&
            up(c->d_buckets, all_buckets.data(), sizeof(int2) * all_buckets.size());
  // the images, each at its offset of one staging buffer, and their prefix tables built on the
  // device; the staging buffer goes back once the stream has drained
  if (ok) {
    double* d_rho = static_cast<double*>(pool_alloc_device(sizeof(double) * img_total));
    if (!d_rho) return bail(fail(COT_ENOMEM, "device allocation failed"));
    for (int p = 0; p < n_problems && ok; ++p) {
      const cot_problem& pr = problems[p];
      const ProbDesc& pd = c->probs[p];
      ok = up(d_rho + pd.pre_off, pr.rho, sizeof(double) * pr.pm * pr.pn) &&
           launch_prefix(d_rho + pd.pre_off, pr.pm, pr.pn, c->fp32, c->d_grp + pd.grp_off,
                         c->d_tot + pd.tot_off, c->st) == hipSuccess;
    }
    ok = (hipStreamSynchronize(c->st) == hipSuccess) && ok;
    pool_free_device(d_rho);
  }
  {
    // FP32 clipping frames start at the sites
    ok = ok && up(c->d_org, xy.data(), sizeof(double2) * N);
    // timers: every slot's start at the maximum, ends and accumulators zero
    unsigned long long tm[kTimers * kTimerWords + kDiagWords] = {};
    for (int k = 0; k < kTimers; ++k)
      for (int r = 0; r < kTimerSlots; ++r) tm[k * kTimerWords + 2 * r] = ~0ull;
    ok = ok && up(c->d_timer, tm, sizeof(tm));
  }
  if (!ok) return bail(fail(COT_EHIP, "upload failed"));
  if (launch_fill_u8(c->d_ring, N, (uint8_t)c->opt.ring_init, c->st) != hipSuccess ||
      launch_fill_u8(c->d_adjn, N, 0, c->st) != hipSuccess ||
      launch_fill_i32(c->d_ecount, 1, 0, c->st) != hipSuccess ||
      launch_fill_u8(c->d_prebuilt, N, 0, c->st) != hipSuccess ||
      launch_fill_i32(c->d_redo, N, 0, c->st) != hipSuccess ||
      launch_fill_i32(c->d_lcount, 2, 0, c->st) != hipSuccess ||
      launch_fill_i32(c->d_xcnt, N, 0, c->st) != hipSuccess ||
      launch_fill_i32(c->d_dcount, 1, 0, c->st) != hipSuccess ||
      hipMemset(c->d_fails, 0, sizeof(int) * 2 * c->np) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return bail(fail(COT_EHIP, "initialisation failed"));
  *out = c;
  return COT_OK;
}

int cot_destroy(cot_context* c) {
  delete c;
  return COT_OK;
}

int cot_evaluate(cot_context* c, const double* w, double* mass, double* f, double* centroid,
                 unsigned char* exist) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  CHK(upload_sorted(c, w, c->d_w));
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> sum;
  CHK(evaluate_device(c, c->d_w, nullptr, act, sum));
  if (mass) CHK(download_sorted(c, c->d_mass, mass));
  if (f && c->io_device) {
    for (int p = 0; p < c->np; ++p) c->h_small[p] = sum[p].f;
    HIPCHK(hipMemcpyAsync(f, c->h_small, sizeof(double) * c->np, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));  // the pinned staging is reused
  } else if (f) {
    for (int p = 0; p < c->np; ++p) f[p] = sum[p].f;
  }
  if (centroid) CHK(download_sorted2(c, c->d_cen, centroid));
  if (exist) CHK(download_sorted_u8(c, c->d_exist, exist));
  for (int p = 0; p < c->np; ++p)
    if (sum[p].fails) return fail(COT_EUNVERIFIED, "cells failed the conformity check");
  return COT_OK;
}

int cot_centroids(cot_context* c, const double* w, int weighted, double* centroid,
                  unsigned char* exist) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  CHK(upload_sorted(c, w, c->d_w));
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> sum;
  CHK(evaluate_device(c, c->d_w, nullptr, act, sum));
  if (centroid) CHK(download_sorted2(c, weighted ? c->d_cen : c->d_geo, centroid));
  if (exist) CHK(download_sorted_u8(c, c->d_exist, exist));
  return COT_OK;
}

int cot_cell_areas(cot_context* c, const double* w, double* area) {
  if (!c || !w || !area) return fail(COT_EINVAL, "null argument");
  CHK(upload_sorted(c, w, c->d_w));
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> sum;
  CHK(evaluate_device(c, c->d_w, nullptr, act, sum));
  CHK(download_sorted(c, c->d_area, area));
  return COT_OK;
}

namespace {

// Arguments common to the irradiance render entry points; every vertex index must name a point
// before a kernel dereferences it (the faces are read back once to check).
int check_mesh(const float* tri_energy, const int* faces, int n_faces, const float* xy,
               int n_points, const double* proj_domain, int pm, int pn, hipStream_t st) {
  if (n_faces < 0 || n_points < 0 || pm <= 0 || pn <= 0 || !proj_domain ||
      (n_faces > 0 && (!tri_energy || !faces || !xy)))
    return fail(COT_EINVAL, "bad argument");
  if (!(proj_domain[1] > proj_domain[0] && proj_domain[3] > proj_domain[2]))
    return fail(COT_EINVAL, "empty proj_domain");
  if (n_faces > 0) {
    std::vector<int> fh(3 * (size_t)n_faces);
    HIPCHK(hipMemcpyAsync(fh.data(), faces, sizeof(int) * fh.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int v : fh)
      if (v < 0 || v >= n_points) return fail(COT_EINVAL, "face index out of range");
  }
  return COT_OK;
}

}  // namespace

int cot_splat(const float* points, const float* energy, int n, int res_y, int res_x,
              const double* proj_domain, float sigma, int radius, const float* image_grad,
              float scale, float* out, void* stream) {
  if (n < 0 || res_y <= 0 || res_x <= 0 || radius < 0 || radius > 16 || !(sigma > 0.0f) ||
      !proj_domain || !out || (n > 0 && (!points || !energy)))
    return fail(COT_EINVAL, "bad argument");
  if (!(proj_domain[1] > proj_domain[0] && proj_domain[3] > proj_domain[2]))
    return fail(COT_EINVAL, "empty proj_domain");
  HIPCHK(launch_splat(points, energy, n, res_y, res_x, proj_domain, sigma, radius, image_grad,
                      scale, out, reinterpret_cast<hipStream_t>(stream)));
  return COT_OK;
}

int cot_irra_render(const float* tri_energy, const int* faces, int n_faces, const float* xy,
                    int n_points, float total_gray, const double* proj_domain, int pm, int pn,
                    float* image, void* stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!image) return fail(COT_EINVAL, "bad argument");
  CHK(check_mesh(tri_energy, faces, n_faces, xy, n_points, proj_domain, pm, pn, st));
  HIPCHK(launch_tri_render(tri_energy, faces, n_faces, xy, total_gray, proj_domain, pm, pn, image,
                           st));
  return COT_OK;
}

int cot_irra_render_grad(const float* tri_energy, const int* faces, int n_faces, const float* xy,
                         int n_points, float total_gray, const float* real_picture,
                         const double* proj_domain, int pm, int pn, float* render,
                         float* render_diff, float* xy_render_grad, float* tri_render_grad,
                         float* inte_diff, float* xy_inte_grad, float* inte_result,
                         int use_sobel, float* xy_sobel_grad, float* tri_sobel_grad,
                         void* stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!real_picture || !render || !render_diff ||
      (n_points > 0 && (!xy_render_grad || !xy_inte_grad || (use_sobel && !xy_sobel_grad))) ||
      (n_faces > 0 && (!tri_render_grad || !inte_diff || !inte_result ||
                       (use_sobel && !tri_sobel_grad))))
    return fail(COT_EINVAL, "bad argument");
  CHK(check_mesh(tri_energy, faces, n_faces, xy, n_points, proj_domain, pm, pn, st));
  float* sob = nullptr;  // the Sobel weight image, held for this call
  CallBuffers buf(st);
  if (use_sobel) CHK(buf.get(&sob, (size_t)pm * pn));
  HIPCHK(launch_tri_grad(tri_energy, faces, n_faces, xy, n_points, total_gray, real_picture,
                         proj_domain, pm, pn, render, render_diff, xy_render_grad,
                         tri_render_grad, inte_diff, xy_inte_grad, inte_result, sob,
                         xy_sobel_grad, tri_sobel_grad, st));
  return COT_OK;
}

int cot_cell_solid_angles(cot_context* c, const double* w, const double* light, double z,
                          double* omega) {
  if (!c || !w || !light || !omega) return fail(COT_EINVAL, "null argument");
  if (!(z != light[2])) return fail(COT_EINVAL, "the light lies in the plane of the cells");
  double* d_om = nullptr;
  CallBuffers buf(c->st);
  CHK(buf.get(&d_om, c->ntot));
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> cur;
  CHK(upload_sorted(c, w, c->d_w));
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  for (int p = 0; p < c->np; ++p)
    if (cur[p].fails) return fail(COT_EUNVERIFIED, "cells failed the conformity check");
  HIPCHK(launch_solid_angle(c->d_blocks, c->d_vtx, c->d_nedge, light, z, d_om, c->nblk, c->st));
  CHK(download_sorted(c, d_om, omega));
  return COT_OK;
}

namespace {

// Solve (A + 1e-10 I) a = 1 for a small symmetric positive semi-definite A (Cholesky), then
// normalise a to sum 1 (ot_op.cuh:1160-1163).  Returns false when the factorisation breaks down.
bool anderson_weights(const std::vector<double>& A, int H, std::vector<double>& a) {
  std::vector<double> L(H * H, 0.0);
  for (int i = 0; i < H; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = A[i * H + j] + (i == j ? 1e-10 : 0.0);
      for (int k = 0; k < j; ++k) s -= L[i * H + k] * L[j * H + k];
      if (i == j) {
        if (!(s > 0.0)) return false;
        L[i * H + i] = std::sqrt(s);
      } else {
        L[i * H + j] = s / L[j * H + j];
      }
    }
  }
  std::vector<double> y(H);
  for (int i = 0; i < H; ++i) {
    double s = 1.0;
    for (int k = 0; k < i; ++k) s -= L[i * H + k] * y[k];
    y[i] = s / L[i * H + i];
  }
  a.assign(H, 0.0);
  for (int i = H - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < H; ++k) s -= L[k * H + i] * a[k];
    a[i] = s / L[i * H + i];
  }
  double tot = 0.0;
  for (double v : a) tot += v;
  if (!(std::fabs(tot) > 0.0)) return false;
  for (double& v : a) v /= tot;
  return true;
}

}  // namespace

int cot_solve_reflection(cot_context* c, double* w, const cot_solve_opts* so,
                         const cot_reflection* r, cot_report* rep) {
  if (!c || !w || !r || !r->points || !r->inside_dir || !r->height || !r->clip_area)
    return fail(COT_EINVAL, "null argument");
  if (c->np != 1) return fail(COT_EINVAL, "the reflection solve takes one-problem contexts");
  if (r->aa_history_size < 1 || r->aa_history_size > kMaxHist)
    return fail(COT_EINVAL, "aa_history_size must be in [1, " + std::to_string(kMaxHist) + "]");
  cot_solve_opts o;
  if (so) {
    o = *so;
  } else {
    cot_default_solve_opts(&o);
  }
  o.stop_rule = COT_STOP_REFERENCE;
  const auto t_start = std::chrono::steady_clock::now();
  const int N = c->ntot, nb = c->nblk;
  const size_t ld = N;
  hipStream_t st = c->st;
  const int mA = r->aa_history_size;

  // per-site lens data in bucket order
  double2* d_pts = nullptr;
  double *d_ins = nullptr, *d_h = nullptr, *d_area = nullptr, *d_wt = nullptr, *d_G = nullptr;
  double *d_R = nullptr, *d_alpha = nullptr, *d_pp = nullptr, *d_wi = nullptr;
  CallBuffers buf(st);
  CHK(buf.get(&d_pts, N));
  CHK(buf.get(&d_ins, 3 * (size_t)N));
  if (r->incident_dir) CHK(buf.get(&d_wi, 3 * (size_t)N));
  CHK(buf.get(&d_h, N));
  CHK(buf.get(&d_area, N));
  CHK(buf.get(&d_wt, N));
  CHK(buf.get(&d_G, (size_t)mA * N));
  CHK(buf.get(&d_R, (size_t)mA * N));
  CHK(buf.get(&d_alpha, mA));
  CHK(buf.get(&d_pp, (size_t)std::max(mA, 1) * nb));
  {
    std::vector<double2> pts(N);
    std::vector<double> ins(3 * (size_t)N), wi(r->incident_dir ? 3 * (size_t)N : 0), h(N), ar(N);
    for (int q = 0; q < N; ++q) {
      const int o2 = c->perm[q];
      pts[q] = make_double2(r->points[2 * o2], r->points[2 * o2 + 1]);
      for (int k = 0; k < 3; ++k) ins[3 * q + k] = r->inside_dir[3 * o2 + k];
      if (r->incident_dir)
        for (int k = 0; k < 3; ++k) wi[3 * q + k] = r->incident_dir[3 * o2 + k];
      h[q] = r->height[o2];
      ar[q] = r->clip_area[o2];
    }
    HIPCHK(hipMemcpy(d_pts, pts.data(), sizeof(double2) * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ins, ins.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice));
    if (d_wi) HIPCHK(hipMemcpy(d_wi, wi.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_h, h.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_area, ar.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  }
  std::vector<double> hp(std::max(mA, 1) * (size_t)nb);
  // nu from the Fresnel weights at the centroids the last evaluation left in d_geo
  auto update_nu = [&]() -> int {
    HIPCHK(launch_fresnel_weight(c->d_blocks, d_pts, d_ins, d_wi, d_h, d_area, c->d_geo, r->z_of_len,
                                 r->z_of_receiver, r->refraction_index, d_wt, d_pp, nb, st));
    HIPCHK(hipMemcpyAsync(hp.data(), d_pp, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double tot = 0.0;
    for (int b = 0; b < nb; ++b) tot += hp[b];
    if (!(tot > 0.0)) return fail(COT_EINVAL, "Fresnel-weighted source mass is zero");
    HIPCHK(launch_scale(c->d_blocks, d_wt, r->area_scalar / tot, c->d_nu, nb, st));
    return COT_OK;
  };
  // |mass - nu|^2 and the empty-cell count with the current nu
  auto residual = [&](EvalSum& e) -> int {
    HIPCHK(launch_reduce_eval(c->d_blocks, c->d_mass, c->d_fcell, c->d_area, c->d_nu,
                              c->d_exist, nullptr, c->d_part, nb, st));
    HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * kEvalPartials * nb,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<EvalSum> s;
    sum_partials(c, s);
    e.g2 = s[0].g2;
    e.zeros = s[0].zeros;
    e.maxrel = s[0].maxrel;
    return COT_OK;
  };

  std::vector<uint8_t> act(1, 1);
  std::vector<EvalSum> cur;
  CHK(upload_sorted(c, w, c->d_w));
  // masses at the warm start's geometric centroids (ot_op.cuh:779-795)
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  CHK(update_nu());
  int iters_a = 0, evals_a = 0;
  if (r->if_aa_op_initial) {
    // stage A: L-BFGS with frozen masses (ot_op.cuh:823-1074)
    std::vector<double> wh(N);
    CHK(download_sorted(c, c->d_w, wh.data()));
    cot_report ra{};
    const int rc = cot_solve(c, wh.data(), &o, &ra);
    if (rc != COT_OK && rc != COT_ENOTCONV) return rc;
    iters_a = ra.iterations;
    evals_a = ra.evaluations;
    CHK(upload_sorted(c, wh.data(), c->d_w));
  }
  // re-estimate the masses at the stage-A result (ot_op.cuh:1076-1110)
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  CHK(update_nu());
  HIPCHK(launch_gradient(c->d_blocks, c->d_mass, c->d_nu, c->d_g, nb, st));
  EvalSum now;
  CHK(residual(now));

  // stage B: Anderson acceleration of G(w) = w + beta g(w) (ot_op.cuh:1112-1323)
  std::vector<double> A((size_t)mA * mA, 0.0), alpha;
  int aa_loop_iter = 0, iters_b = 0, evals_b = 0;
  bool conv = false;
  int status = COT_OK;
  for (int iter = 0; iter < o.max_iter; ++iter) {
    const int H = std::min(aa_loop_iter + 1, mA);
    const int slot = aa_loop_iter % mA;
    HIPCHK(launch_aa_store(c->d_blocks, c->d_w, c->d_g, r->aa_fixed_para, slot, H, d_G, d_R, ld,
                           d_pp, nb, st));
    HIPCHK(hipMemcpyAsync(hp.data(), d_pp, sizeof(double) * H * nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int h = 0; h < H; ++h) {
      double s = 0.0;
      for (int b = 0; b < nb; ++b) s += hp[(size_t)h * nb + b];
      A[slot * mA + h] = A[h * mA + slot] = s;
    }
    std::vector<double> Ah((size_t)H * H);
    for (int i2 = 0; i2 < H; ++i2)
      for (int j2 = 0; j2 < H; ++j2) Ah[i2 * H + j2] = A[i2 * mA + j2];
    if (!anderson_weights(Ah, H, alpha)) {
      alpha.assign(H, 0.0);
      alpha[slot] = 1.0;
    }
    HIPCHK(hipMemcpyAsync(d_alpha, alpha.data(), sizeof(double) * H, hipMemcpyHostToDevice, st));
    HIPCHK(launch_aa_direction(c->d_blocks, d_alpha, H, d_G, ld, c->d_w, c->d_dir, nb, st));
    const double zeros_start = now.zeros;
    double step = 1.0;
    for (;;) {
      c->h_small[0] = step;
      HIPCHK(hipMemcpyAsync(c->d_step, c->h_small, sizeof(double), hipMemcpyHostToDevice, st));
      std::vector<EvalSum> tr;
      CHK(evaluate_device(c, c->d_w, c->d_dir, act, tr));
      ++evals_b;
      if (tr[0].fails) status = COT_EUNVERIFIED;
      if (tr[0].overflow > 0) {
        step *= 0.5;
      } else {
        // every trial recomputes the masses from its own cells (ot_op.cuh:1212-1240)
        CHK(update_nu());
        HIPCHK(launch_gradient(c->d_blocks, c->d_mass, c->d_nu, c->d_g, nb, st));
        HIPCHK(launch_axpy(c->d_blocks, step, c->d_dir, c->d_w, nb, st));
        ++aa_loop_iter;
        CHK(residual(now));
        break;
      }
      if (step <= 1e-4 * o.initial_step) {
        aa_loop_iter = 0;
        break;
      }
    }
    ++iters_b;
    // the empty-cell count is evaluated on every 10th iteration only (ot_op.cuh:1257-1268)
    const double degenerate = (iter % 10 == 0) ? zeros_start : 0.0;
    if (o.verbose > 0 && iter % o.verbose == 0)
      std::printf("[cot] anderson iter %d |g| %.6e maxrel %.3e step %g\n", iter,
                  std::sqrt(now.g2), now.maxrel, step);
    if (std::sqrt(now.g2) < o.tol && degenerate == 0.0) {
      conv = true;
      break;
    }
  }
  CHK(download_sorted(c, c->d_w, w));
  if (rep) {
    rep->iterations = iters_a + iters_b;
    rep->evaluations = evals_a + evals_b;
    rep->converged = conv ? 1 : 0;
    rep->status = status != COT_OK ? status : (conv ? COT_OK : COT_ENOTCONV);
    rep->f = 0.0;
    rep->grad_norm = std::sqrt(now.g2);
    rep->max_rel_err = now.maxrel;
    rep->seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  if (status != COT_OK) return fail(status, "cells failed the conformity check");
  if (!conv) return fail(COT_ENOTCONV, "iteration limit reached");
  return COT_OK;
}

}  // extern "C"

namespace {

// Buffers of the Newton steps, allocated on first use.
int ensure_newton(cot_context* c) {
  if (c->d_hh) return COT_OK;
  const size_t N = c->ntot, nb = c->nblk;
  CHK(dalloc(c, &c->d_hh, nb * kMaxVerts * kCellBlock));  // hslot order
  CHK(dalloc(c, &c->d_hj, nb * kMaxVerts * kCellBlock));
  CHK(dalloc(c, &c->d_hdiag, N));
  CHK(dalloc(c, &c->d_nx, N));
  CHK(dalloc(c, &c->d_nr, N));
  CHK(dalloc(c, &c->d_nz, N));
  CHK(dalloc(c, &c->d_np, N));
  CHK(dalloc(c, &c->d_np2, N));
  CHK(dalloc(c, &c->d_nq, N));
  CHK(dalloc(c, &c->d_prz, 2 * nb));
  CHK(dalloc(c, &c->d_nrz, 2));
  CHK(dalloc(c, &c->d_ppq, nb));
  CHK(dalloc(c, &c->d_prr, nb));
  CHK(dalloc(c, &c->d_pbb, nb));
  CHK(dalloc(c, &c->d_ncarry, N));
  CHK(dalloc(c, &c->d_nshift, c->np));
  CHK(dalloc(c, &c->d_cgout, 4));
  return COT_OK;
}

// Patches of the coarse space for at most maxp per problem (0: none).  A problem's bucket grid is
// cut into rectangles of ax x ay buckets (multiples of the 8x8 tiles, doubled alternately until
// at most maxp rectangles remain); a block joins the rectangle of its first site, so a patch is a
// union of whole blocks and its residual sum is a sum of per-block sums.  A problem with a single
// patch gets none (the constants are the Hessian's null space).
int build_coarse(cot_context* c, int maxp) {
  maxp = std::min(std::max(maxp, 0), kMaxCoarse);
  if (maxp == c->coarse_max) return COT_OK;
  const int np = c->np, nb = c->nblk;
  std::vector<int> bagg(nb, -1), start, blk, aprob;
  std::vector<CoarseDesc> cd(np, CoarseDesc{0, 0, 0, 0});
  std::vector<int4> tasks;
  int total = 0;
  size_t moff = 0;
  for (int p = 0; p < np && maxp > 1; ++p) {
    const ProbDesc& pd = c->probs[p];
    int ax = 8, ay = 8;
    auto cells = [&](int a, int b) {
      return (long long)((pd.gx + a - 1) / a) * ((pd.gy + b - 1) / b);
    };
    while (cells(ax, ay) > maxp) {
      if (ax <= ay) {
        ax *= 2;
      } else {
        ay *= 2;
      }
    }
    const int nx = (pd.gx + ax - 1) / ax;
    std::vector<int> idx((size_t)cells(ax, ay), -1);
    std::vector<std::vector<int>> members;
    for (int b = pd.blk_off; b < pd.blk_off + pd.nblk; ++b) {
      const int2 bc = c->blk_cell[b];
      const size_t key = (size_t)(bc.y / ay) * nx + bc.x / ax;
      if (idx[key] < 0) {
        idx[key] = (int)members.size();
        members.emplace_back();
      }
      members[idx[key]].push_back(b);
    }
    const int na = (int)members.size();
    if (na <= 1) continue;
    if (moff + (size_t)na * na > (size_t)INT32_MAX)
      return fail(COT_EINVAL, "coarse space too large for the batch");
    cd[p] = CoarseDesc{total, na, (int)moff, 0};
    for (int a = 0; a < na; ++a) {
      start.push_back((int)blk.size());
      aprob.push_back(p);
      for (int b : members[a]) {
        blk.push_back(b);
        bagg[b] = total + a;
      }
    }
    for (int r0 = 0; r0 < na; r0 += kCoarseRows)
      tasks.push_back(make_int4(p, r0, std::min(kCoarseRows, na - r0), 0));
    total += na;
    moff += (size_t)na * na;
  }
  start.push_back((int)blk.size());
  c->coarse_max = maxp;
  c->coarse_total = total;
  c->coarse_ntask = (int)tasks.size();
  c->coarse_mat = moff;
  c->cdesc = cd;
  c->bagg = bagg;
  if (tasks.empty()) return COT_OK;
  CHK(dalloc(c, &c->d_bagg, nb));
  if (!c->d_sagg) CHK(dalloc(c, &c->d_sagg, c->ntot));
  if (!c->d_rsum) CHK(dalloc(c, &c->d_rsum, nb));
  CHK(dalloc(c, &c->d_agg_start, start.size()));
  CHK(dalloc(c, &c->d_agg_blk, blk.size()));
  CHK(dalloc(c, &c->d_agg_prob, aprob.size()));
  CHK(dalloc(c, &c->d_cdesc, np));
  CHK(dalloc(c, &c->d_ctask, tasks.size()));
  if (!c->d_crows) CHK(dalloc(c, &c->d_crows, (size_t)nb * kBlockNbr));
  CHK(dalloc(c, &c->d_ainv, moff));
  CHK(dalloc(c, &c->d_cvec, total));
  CHK(dalloc(c, &c->d_prc, 2 * tasks.size()));
  pool_free_pinned(c->h_ac);
  c->h_ac = nullptr;
  if (pinned_alloc((void**)&c->h_ac, sizeof(double) * moff) != hipSuccess)
    return fail(COT_ENOMEM, "pinned allocation failed");
  if (!c->h_crows &&
      pinned_alloc((void**)&c->h_crows, sizeof(CoarseEntry) * (size_t)nb * kBlockNbr) !=
          hipSuccess)
    return fail(COT_ENOMEM, "pinned allocation failed");
  HIPCHK(hipMemcpy(c->d_bagg, bagg.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_agg_start, start.data(), sizeof(int) * start.size(),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_agg_blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_agg_prob, aprob.data(), sizeof(int) * aprob.size(),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_cdesc, cd.data(), sizeof(CoarseDesc) * np, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_ctask, tasks.data(), sizeof(int4) * tasks.size(),
                   hipMemcpyHostToDevice));
  return COT_OK;
}


}  // namespace

extern "C" {

int cot_render(cot_context* c, const double* w, double* image) {
  if (!c || !w || !image) return fail(COT_EINVAL, "null argument");
  if (!c->d_img) CHK(dalloc(c, &c->d_img, c->img_total));
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> cur;
  CHK(upload_sorted(c, w, c->d_w));
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  for (int p = 0; p < c->np; ++p)
    if (cur[p].fails) return fail(COT_EUNVERIFIED, "cells failed the conformity check");
  HIPCHK(launch_render(c->d_probs, c->d_blocks, c->d_vtx, c->d_nedge, c->d_nu, c->d_img,
                       c->img_total, c->nblk, c->st));
  HIPCHK(hipMemcpyAsync(image, c->d_img, sizeof(double) * c->img_total, hipMemcpyDeviceToHost,
                        c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return COT_OK;
}

int cot_cell_polygons(cot_context* c, const double* w, int* counts, double* verts) {
  if (!c || !w || !counts || !verts) return fail(COT_EINVAL, "null argument");
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> cur;
  CHK(upload_sorted(c, w, c->d_w));
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  for (int p = 0; p < c->np; ++p)
    if (cur[p].fails) return fail(COT_EUNVERIFIED, "cells failed the conformity check");
  const size_t N = c->ntot;
  std::vector<uint8_t> ne(N);
  std::vector<double2> vx(N * kMaxVerts);
  HIPCHK(hipMemcpyAsync(ne.data(), c->d_nedge, N, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(vx.data(), c->d_vtx, sizeof(double2) * N * kMaxVerts,
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (size_t r = 0; r < N; ++r) {
    const int o = c->perm[r];
    const int nv = ne[r] >= 3 ? ne[r] : 0;
    counts[o] = nv;
    for (int k = 0; k < kMaxVerts; ++k) {
      const double2 v = k < nv ? vx[r * kMaxVerts + k] : make_double2(0.0, 0.0);
      verts[((size_t)o * kMaxVerts + k) * 2] = v.x;
      verts[((size_t)o * kMaxVerts + k) * 2 + 1] = v.y;
    }
  }
  return COT_OK;
}

int cot_hessian_product(cot_context* c, const double* w, const double* x, double* y) {
  if (!c || !w || !x || !y) return fail(COT_EINVAL, "null argument");
  CHK(ensure_newton(c));
  const int nb = c->nblk;
  std::vector<uint8_t> act(c->np, 1);
  std::vector<EvalSum> cur;
  CHK(upload_sorted(c, w, c->d_w));
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  for (int p = 0; p < c->np; ++p)
    if (cur[p].fails) return fail(COT_EUNVERIFIED, "cells failed the conformity check");
  std::memset(c->h_active, 1, c->np);
  HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, c->np, hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_hessian(c->d_probs, c->d_blocks, c->d_active, c->d_xy, c->d_vtx, c->d_nedge,
                        c->d_elab, c->d_grp, nullptr, c->d_hh, c->d_hj, c->d_hdiag, nullptr, nb,
                        c->st));
  CHK(upload_sorted(c, x, c->d_np));
  HIPCHK(launch_hessian_apply(c->d_blocks, c->d_nedge, c->d_hj, c->d_hh, c->d_hdiag, c->d_np,
                              c->d_nq, c->d_ppq, nb, c->st));
  CHK(download_sorted(c, c->d_nq, y));
  return COT_OK;
}

int cot_solve_newton(cot_context* c, double* w, const cot_solve_opts* so, cot_report* reports) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  cot_solve_opts o;
  if (so) {
    o = *so;
  } else {
    cot_default_solve_opts(&o);
  }
  const auto t_start = std::chrono::steady_clock::now();
  const int np = c->np, nb = c->nblk;
  hipStream_t st = c->st;
  CHK(ensure_newton(c));
  const int rounds = std::min(c->opt.max_repairs, std::max(o.repair_rounds, 1));
  CHK(upload_sorted(c, w, c->d_w));
  std::vector<uint8_t> act(np, 1), conv(np, 0), failed(np, 0);
  std::vector<EvalSum> cur, trial;
  std::vector<int> iters(np, 0), evals(np, 1);
  // stopping rule: the reference's |g|_2 < tol (Newton steps keep every cell non-empty), or
  // max-rel < tol
  auto met = [&](const EvalSum& e) {
    return o.stop_rule == COT_STOP_REFERENCE ? (std::sqrt(e.g2) < o.tol && e.zeros == 0)
                                             : e.maxrel < o.tol;
  };
  CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
  HIPCHK(launch_gradient(c->d_blocks, c->d_mass, c->d_nu, c->d_g, nb, st));
  // the first conjugate-gradient solve starts from zero
  HIPCHK(launch_fill_i32(reinterpret_cast<int*>(c->d_ncarry), 2 * c->ntot, 0, st));
  std::vector<double> shift(np, 0.0), t(np, 1.0), rr_part(nb);
  // CG stops at |r| <= eta |rhs| -- an inexact Newton step, eta = 0.1 max-rel clamped to
  // [1e-4, 0.1], loose while far from the solution -- or after kMaxCG iterations.  Iteration 0
  // is launched directly; the rest run in chunks of `check` iterations (a captured graph unless
  // device_loop == 2), with the residual checked after each chunk.
  constexpr int kMaxCG = 4000;
  const bool eager = o.device_loop == 2;
  const int sweeps = std::min(std::max(o.newton_sweeps, 0), 64);
  CHK(build_coarse(c, o.newton_coarse));
  const bool coarse = c->coarse_ntask > 0;
  // the coarse term cuts the iterations ~8x: check the residual more often
  const int check = coarse ? 16 : 32;
  CgArgs ca{};
  ca.blocks = c->d_blocks;
  ca.nblk = nb;
  ca.sweeps = sweeps;
  ca.ntask = coarse ? c->coarse_ntask : 0;
  ca.nedge = c->d_nedge;
  ca.hj = c->d_hj;
  ca.h = c->d_hh;
  ca.diag = c->d_hdiag;
  ca.x = c->d_nx;
  ca.r = c->d_nr;
  ca.z = c->d_nz;
  ca.q = c->d_nq;
  ca.part_rz = c->d_prz;
  ca.part_pq = c->d_ppq;
  ca.part_rr = c->d_prr;
  ca.part_rzc = c->d_prc;
  ca.rz = c->d_nrz;
  ca.tasks = c->d_ctask;
  ca.cdesc = c->d_cdesc;
  ca.agg_start = c->d_agg_start;
  ca.agg_blk = c->d_agg_blk;
  ca.sagg = c->d_sagg;
  ca.rsum = c->d_rsum;
  ca.ainv = c->d_ainv;
  ca.cvec = c->d_cvec;
  // the Jacobi-only arguments: the fallback when the full preconditioner's CG diverges
  CgArgs cj = ca;
  cj.ntask = 0;
  cj.sweeps = 0;
  // captured chunk graphs: [0] the full preconditioner, [1] Jacobi only
  hipGraph_t cg_graph[2] = {nullptr, nullptr};
  hipGraphExec_t cg_exec[2] = {nullptr, nullptr};
  struct CgGraphGuard {
    hipGraph_t* g;
    hipGraphExec_t* e;
    ~CgGraphGuard() {
      for (int k = 0; k < 2; ++k) {
        if (e[k]) (void)hipGraphExecDestroy(e[k]);
        if (g[k]) (void)hipGraphDestroy(g[k]);
      }
    }
  } cg_guard{cg_graph, cg_exec};
  // iteration k: r.z partials in slot k & 1 (current) and (k + 1) & 1 (previous), the
  // direction read from pbuf[k & 1] and written to pbuf[(k + 1) & 1]; a chunk runs iterations
  // k0 .. k0 + check - 1 with k0 odd, so the buffers realign from chunk to chunk
  double* pbuf[2] = {c->d_np, c->d_np2};
  auto cg_chunk = [&](const CgArgs& a) -> int {
    for (int k = 1; k <= check; ++k)
      HIPCHK(launch_cg_iteration(a, pbuf[k & 1], pbuf[(k + 1) & 1], k & 1, 0, st));
    return COT_OK;
  };
  // per problem, the sum of its blocks' partials (|r|^2 from part_rr by default)
  auto sum_rr = [&](std::vector<double>& rr, const double* part = nullptr) -> int {
    HIPCHK(hipMemcpyAsync(rr_part.data(), part ? part : c->d_prr, sizeof(double) * nb,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int p = 0; p < np; ++p) {
      const ProbDesc& pd = c->probs[p];
      double v = 0.0;
      for (int b = pd.blk_off; b < pd.blk_off + pd.nblk; ++b) v += rr_part[b];
      rr[p] = v;
    }
    return COT_OK;
  };
  for (int it = 0; it < o.max_iter; ++it) {
    bool any = false;
    for (int p = 0; p < np; ++p) {
      if (!conv[p] && !failed[p]) {
        if (met(cur[p])) {
          conv[p] = 1;
        } else if (cur[p].zeros > 0 || cur[p].fails > 0) {
          failed[p] = 1;  // Newton needs every cell non-empty and verified
        }
      }
      act[p] = !conv[p] && !failed[p];
      any |= act[p] != 0;
      shift[p] = act[p] ? (cur[p].mass - c->nu_sum[p]) / c->probs[p].n : 0.0;
    }
    if (!any) break;
    const auto t_it = std::chrono::steady_clock::now();
    std::memcpy(c->h_active, act.data(), np);
    HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, np, hipMemcpyHostToDevice, st));
    std::memcpy(c->h_small, shift.data(), sizeof(double) * np);
    HIPCHK(hipMemcpyAsync(c->d_nshift, c->h_small, sizeof(double) * np, hipMemcpyHostToDevice,
                          st));
    // Hessian at the current cells, then H x = -(g - mean g)
    HIPCHK(launch_hessian(c->d_probs, c->d_blocks, c->d_active, c->d_xy, c->d_vtx, c->d_nedge,
                          c->d_elab, c->d_grp, coarse ? c->d_bagg : nullptr, c->d_hh, c->d_hj,
                          c->d_hdiag,
                          coarse ? c->d_sagg : nullptr, nb, st));
    if (coarse) {
      // the blocks' couplings down; the patches' Laplacian summed over each patch's blocks in
      // block order and inverted (host FP64); the inverses up
      HIPCHK(launch_coarse_rows(c->d_blocks, c->d_bagg, c->d_sagg, c->d_nedge, c->d_hj,
                                c->d_hh, c->d_crows, nb, st));
      HIPCHK(hipMemcpyAsync(c->h_crows, c->d_crows, sizeof(CoarseEntry) * (size_t)nb * kBlockNbr,
                            hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      const auto t0 = std::chrono::steady_clock::now();
      for (int p = 0; p < np; ++p) {
        const CoarseDesc& cd = c->cdesc[p];
        if (cd.na == 0) continue;
        double* m = c->h_ac + cd.moff;
        std::fill(m, m + (size_t)cd.na * cd.na, 0.0);
        if (!act[p]) continue;
        const ProbDesc& pd = c->probs[p];
        for (int b = pd.blk_off; b < pd.blk_off + pd.nblk; ++b) {
          const int a = c->bagg[b] - cd.a0;
          if (a < 0) continue;
          const CoarseEntry* row = c->h_crows + (size_t)b * kBlockNbr;
          for (int k = 0; k < kBlockNbr && row[k].b >= 0; ++k)
            m[(size_t)a * cd.na + (row[k].b - cd.a0)] -= row[k].v;
        }
        coarse_inverse(m, cd.na);
      }
      c->stats.coarse_seconds +=
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      HIPCHK(hipMemcpyAsync(c->d_ainv, c->h_ac, sizeof(double) * c->coarse_mat,
                            hipMemcpyHostToDevice, st));
    }
    // each problem's CG tolerance from its own state (a batch's problems converge at their own
    // pace: one problem's loose tolerance, or the batch's summed residual, would stop the CG of
    // another far from its solution -- its Newton step then fails and L-BFGS takes over)
    // rr0: |rhs|^2, the stopping rule's reference; rs0: |r|^2 at the start (the divergence
    // test's scale, with |rhs|^2: a carried start may begin above |rhs|)
    std::vector<double> rr0(np, 0.0), rs0(np, 0.0), rr(np, 0.0), eta(np, 0.0);
    double eta_min = 0.1;
    for (int p = 0; p < np; ++p) {
      if (!act[p]) continue;
      eta[p] = std::min(0.1, std::max(1e-4, 0.1 * cur[p].maxrel));
      eta_min = std::min(eta_min, eta[p]);
    }
    // stop when every active problem has |r_p| <= eta_p |r0_p|; diverged when one grew past
    // 1e4 |r0_p| (or is not a number)
    auto cg_state = [&]() -> int {  // 1 done, 0 continue, -1 diverged
      bool done = true;
      for (int p = 0; p < np; ++p) {
        if (!act[p]) continue;
        if (!(rr[p] <= 1e8 * std::max(rr0[p], rs0[p]))) return -1;
        if (rr[p] > eta[p] * eta[p] * rr0[p]) done = false;
      }
      return done ? 1 : 0;
    };
    // CG to every problem's tolerance with the arguments a (graph slot gs); returns the
    // iterations run, negative when it diverged
    auto run_cg = [&](const CgArgs& a, int gs, int& iters) -> int {
      const bool co = a.ntask > 0;
      // from the part of the last direction its step left (d_ncarry, zero after a full step)
      HIPCHK(launch_cg_init(c->d_blocks, c->d_active, c->d_g, c->d_nshift, c->d_hdiag,
                            c->d_nedge, c->d_hj, c->d_hh, a.sweeps, c->d_ncarry, c->d_nx,
                            c->d_nr, c->d_nz, c->d_prz, c->d_prr, c->d_pbb,
                            co ? c->d_rsum : nullptr, nb, st));
      if (co)
        HIPCHK(launch_cg_coarse(c->d_ctask, c->coarse_ntask, c->d_cdesc, c->d_agg_start,
                                c->d_agg_blk, c->d_rsum, c->d_ainv, c->d_cvec, c->d_prc, st));
      CHK(sum_rr(rr0, c->d_pbb));  // |rhs|^2: the stopping rule's reference
      CHK(sum_rr(rr));
      rs0 = rr;
      iters = 0;
      if (cg_state() > 0) return COT_OK;  // the carried direction meets the tolerance
      HIPCHK(launch_cg_iteration(a, pbuf[0], pbuf[1], 0, 1, st));
      iters = 1;
      while (iters < kMaxCG) {
        if (eager) {
          CHK(cg_chunk(a));
        } else {
          if (!cg_exec[gs]) {
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = cg_chunk(a);
            const hipError_t ce = hipStreamEndCapture(st, &cg_graph[gs]);
            if (rc != COT_OK) return rc;
            HIPCHK(ce);
            HIPCHK(hipGraphInstantiate(&cg_exec[gs], cg_graph[gs], nullptr, nullptr, 0));
          }
          HIPCHK(hipGraphLaunch(cg_exec[gs], st));
        }
        iters += check;
        CHK(sum_rr(rr));
        const int state = cg_state();
        if (state < 0) {
          iters = -iters;
          return COT_OK;
        }
        if (state > 0) break;
      }
      return COT_OK;
    };
    // a small system: the whole CG in one workgroup, one launch
    auto run_cg_small = [&](const CgArgs& a, int& iters) -> int {
      SmallCgArgs sa{};
      sa.blocks = c->d_blocks;
      sa.nblk = nb;
      sa.sweeps = a.sweeps;
      sa.npatch = a.ntask > 0 ? c->coarse_total : 0;
      sa.max_iter = kMaxCG;
      sa.active = c->d_active;
      sa.g = c->d_g;
      sa.shift = c->d_nshift;
      sa.diag = c->d_hdiag;
      sa.nedge = c->d_nedge;
      sa.hj = c->d_hj;
      sa.h = c->d_hh;
      sa.sagg = a.ntask > 0 ? c->d_sagg : nullptr;
      sa.agg_start = c->d_agg_start;
      sa.agg_blk = c->d_agg_blk;
      sa.agg_prob = c->d_agg_prob;
      sa.cdesc = c->d_cdesc;
      sa.ainv = c->d_ainv;
      sa.x = c->d_nx;
      sa.eta2 = eta_min * eta_min;  // one residual for the workgroup's problems: the strictest
      sa.out = c->d_cgout;
      HIPCHK(launch_cg_small(sa, st));
      double res[4];
      HIPCHK(hipMemcpyAsync(res, c->d_cgout, sizeof(res), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      iters = (int)res[0];
      std::fill(rr0.begin(), rr0.end(), res[1]);
      std::fill(rr.begin(), rr.end(), res[2]);
      if (res[3] != 0.0) iters = -iters;
      return COT_OK;
    };
    const bool small = nb <= kSmallBlocks && c->ntot <= kSmallBlocks * kCellBlock && o.small_cg;
    int cg = 0;
    if (small) {
      CHK(run_cg_small(ca, cg));
    } else {
      CHK(run_cg(ca, 0, cg));
    }
    if (cg < 0 && (coarse || sweeps > 0)) {
      // the preconditioned iteration diverged (rounding along the Hessian's null space at
      // residuals near the FP64 floor): redo the step's CG with Jacobi alone
      if (o.verbose > 0)
        std::printf("[cot] newton CG diverged after %d iterations; Jacobi restart\n", -cg);
      c->stats.cg_iterations += -cg;
      c->stats.cg_restarts += 1;
      if (small) {
        CHK(run_cg_small(cj, cg));
      } else {
        CHK(run_cg(cj, 1, cg));
      }
    }
    if (cg < 0) cg = -cg;  // a diverged Jacobi CG: the line search rejects its direction
    const auto t_cg = std::chrono::steady_clock::now();
    int trials = 0;
    c->stats.newton_steps += 1;
    c->stats.cg_iterations += cg;
    // damped step per problem: every cell keeps mass and |g| falls by the factor (1 - t / 2).
    // Halving from 1 at every step: halving from twice the last accepted step saved 2% on the
    // N=65,536 shapes solve but left some problems of a batch creeping along a poor direction
    // for dozens of steps (they had to regrow t one doubling per step)
    std::vector<uint8_t> tr = act, acc(np, 0);
    std::fill(t.begin(), t.end(), 1.0);
    for (;;) {
      bool any_trial = false;
      for (int p = 0; p < np; ++p) {
        any_trial |= tr[p] != 0;
        c->h_small[p] = t[p];
      }
      if (!any_trial) break;
      HIPCHK(hipMemcpyAsync(c->d_step, c->h_small, sizeof(double) * np, hipMemcpyHostToDevice,
                            st));
      CHK(evaluate_device(c, c->d_w, c->d_nx, tr, trial, rounds));
      ++trials;
      for (int p = 0; p < np; ++p) {
        if (!tr[p]) continue;
        ++evals[p];
        const bool ok = trial[p].fails == 0 && trial[p].overflow == 0 && trial[p].zeros == 0 &&
                        std::sqrt(trial[p].g2) <= (1.0 - 0.5 * t[p]) * std::sqrt(cur[p].g2);
        if (o.verbose > 1 && !ok)
          std::printf("[cot] newton prob %d rejects step %g: unverified %d overflow %d empty %g "
                      "|g| ratio %.6f\n",
                      p, t[p], trial[p].fails, trial[p].overflow, trial[p].zeros,
                      std::sqrt(trial[p].g2 / std::max(cur[p].g2, 1e-300)));
        if (ok) {
          acc[p] = 1;
          tr[p] = 0;
          cur[p] = trial[p];
        } else {
          t[p] *= 0.5;
          if (t[p] < 1.0 / 4096.0) {
            tr[p] = 0;
            failed[p] = 1;
          }
        }
      }
    }
    std::memcpy(c->h_active, acc.data(), np);
    for (int p = 0; p < np; ++p) c->h_small[p] = t[p];
    HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->d_step, c->h_small, sizeof(double) * np, hipMemcpyHostToDevice, st));
    HIPCHK(launch_axpy_step(c->d_blocks, c->d_active, c->d_step, c->d_nx, c->d_w, c->d_ncarry, nb,
                            st));
    HIPCHK(launch_gradient(c->d_blocks, c->d_mass, c->d_nu, c->d_g, nb, st));
    // the pinned staging (h_active, h_small) is rewritten at the top of the next iteration:
    // the copies above must have been made
    HIPCHK(hipStreamSynchronize(st));
    for (int p = 0; p < np; ++p) iters[p] += acc[p];
    if (o.verbose > 0) {
      const auto t_end = std::chrono::steady_clock::now();
      const auto msec = [](std::chrono::steady_clock::time_point a,
                           std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
      };
      for (int p = 0; p < np; ++p)
        if (act[p])
          std::printf("[cot] newton prob %d iter %d cg %d |r|/|r0| %.2e step %g maxrel %.3e "
                      "(ms: Hessian + CG %.2f, %d trial evaluations %.2f)\n",
                      p, iters[p], cg, std::sqrt(rr[p] / std::max(rr0[p], 1e-300)), t[p],
                      cur[p].maxrel, msec(t_it, t_cg), trials, msec(t_cg, t_end));
      std::fflush(stdout);
    }
  }
  CHK(download_sorted(c, c->d_w, w));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  bool all = true;
  for (int p = 0; p < np; ++p) {
    if (!conv[p] && !failed[p] && met(cur[p])) conv[p] = 1;
    all = all && conv[p];
    if (reports) {
      cot_report& r = reports[p];
      r.iterations = iters[p];
      r.evaluations = evals[p];
      r.converged = conv[p];
      r.status = conv[p] ? COT_OK : COT_ENOTCONV;
      r.f = cur[p].f;
      r.grad_norm = std::sqrt(cur[p].g2);
      r.max_rel_err = cur[p].maxrel;
      r.seconds = secs;
    }
  }
  if (!all) return fail(COT_ENOTCONV, "Newton steps did not reach the tolerance");
  return COT_OK;
}

}  // extern "C"

namespace {

// One call of an entry point on device arrays: the caller's arrays are device memory; the
// context's stream starts after the work queued on the caller's stream and the caller's stream
// continues after the context's.
struct IoScope {
  cot_context* c;
  hipStream_t caller;
  int rc = COT_OK;
  IoScope(cot_context* ctx, void* stream) : c(ctx), caller(reinterpret_cast<hipStream_t>(stream)) {
    if (!c->d_perm) {
      if ((rc = dalloc(c, &c->d_perm, c->ntot)) != COT_OK) return;
      if (hipMemcpy(c->d_perm, c->perm.data(), sizeof(int) * c->ntot, hipMemcpyHostToDevice) !=
          hipSuccess) {
        rc = fail(COT_EHIP, "permutation upload failed");
        return;
      }
    }
    if ((!c->ev_in && hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming) != hipSuccess) ||
        (!c->ev_out && hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming) != hipSuccess) ||
        hipEventRecord(c->ev_in, caller) != hipSuccess ||
        hipStreamWaitEvent(c->st, c->ev_in, 0) != hipSuccess) {
      rc = fail(COT_EHIP, "stream ordering failed");
      return;
    }
    c->io_device = true;
  }
  ~IoScope() {
    if (!c->io_device) return;
    c->io_device = false;
    if (hipEventRecord(c->ev_out, c->st) == hipSuccess) (void)hipStreamWaitEvent(caller, c->ev_out, 0);
  }
};

}  // namespace

extern "C" {

int cot_evaluate_dev(cot_context* c, const double* w, double* mass, double* f, double* centroid,
                     unsigned char* exist, void* stream) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  IoScope io(c, stream);
  if (io.rc != COT_OK) return io.rc;
  return cot_evaluate(c, w, mass, f, centroid, exist);
}

int cot_solve_dev(cot_context* c, double* w, const cot_solve_opts* so, cot_report* reports,
                  void* stream) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  IoScope io(c, stream);
  if (io.rc != COT_OK) return io.rc;
  return cot_solve(c, w, so, reports);
}

int cot_centroids_dev(cot_context* c, const double* w, int weighted, double* centroid,
                      unsigned char* exist, void* stream) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  IoScope io(c, stream);
  if (io.rc != COT_OK) return io.rc;
  return cot_centroids(c, w, weighted, centroid, exist);
}

int cot_get_stats(cot_context* c, cot_stats* out) {
  if (!c || !out) return fail(COT_EINVAL, "null argument");
  CHK(pull_timers(c));
  *out = c->stats;
  return COT_OK;
}

int cot_reset_stats(cot_context* c) {
  if (!c) return fail(COT_EINVAL, "null context");
  CHK(pull_timers(c));
  c->stats = cot_stats{};
  return COT_OK;
}

int cot_solve(cot_context* c, double* w, const cot_solve_opts* so, cot_report* reports) {
  if (!c || !w) return fail(COT_EINVAL, "null argument");
  cot_solve_opts o;
  if (so) {
    o = *so;
  } else {
    cot_default_solve_opts(&o);
  }
  if (o.history < 1 || o.history > kMaxHist)
    return fail(COT_EINVAL, "history must be in [1, " + std::to_string(kMaxHist) + "]");
  if (o.repair_rounds < 1) o.repair_rounds = 1;
  if (o.repair_rounds > 8) o.repair_rounds = 8;
  if (o.device_loop) return solve_device(c, w, o, reports);
  const auto t_start = std::chrono::steady_clock::now();
  const int np = c->np;
  const size_t ld = c->ntot;
  CHK(ensure_history(c, o.history));
  CHK(upload_sorted(c, w, c->d_w));

  enum Phase { kDir, kTrial, kDone };
  std::vector<LState> L(np);
  std::vector<Phase> ph(np, kDir);
  std::vector<EvalSum> cur(np), trial(np);
  std::vector<double> t(np, o.initial_step), f_prev(np, 0.0);
  std::vector<int> iters(np, 0), evals(np, 1), status(np, COT_OK);
  std::vector<uint8_t> conv(np, 0);
  std::vector<double> snap_g(np, 0.0), snap_zero(np, 0.0);
  for (int p = 0; p < np; ++p) L[p].m = o.history;

  // initial evaluation
  {
    std::vector<uint8_t> act(np, 1);
    CHK(evaluate_device(c, c->d_w, nullptr, act, cur));
    HIPCHK(launch_gradient(c->d_blocks, c->d_mass, c->d_nu, c->d_g, c->nblk, c->st));
    for (int p = 0; p < np; ++p) {
      if (cur[p].fails) status[p] = COT_EUNVERIFIED;
      if ((o.stop_rule == COT_STOP_MAXREL && cur[p].maxrel < o.tol) ||
          (o.stop_rule == COT_STOP_HANDOFF && cur[p].maxrel < o.tol && cur[p].zeros == 0.0)) {
        conv[p] = 1;
        ph[p] = kDone;
      }
    }
  }

  std::vector<uint8_t> act(np, 0);
  for (;;) {
    // ---- directions for problems that start an iteration
    bool any_dir = false;
    for (int p = 0; p < np; ++p) {
      act[p] = 0;
      if (ph[p] != kDir) continue;
      if (iters[p] >= o.max_iter) {
        ph[p] = kDone;
        continue;
      }
      snap_g[p] = std::sqrt(cur[p].g2);
      snap_zero[p] = cur[p].zeros;
      const int h = two_loop(L[p], c->h_coef + (size_t)p * kCoefWidth);
      c->h_ismall[p] = h;
      act[p] = 1;
      any_dir = true;
    }
    if (any_dir) {
      std::memcpy(c->h_active, act.data(), np);
      HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, np, hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->d_coef, c->h_coef, sizeof(double) * np * kCoefWidth,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->d_hcount, c->h_ismall, sizeof(int) * np, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(launch_direction(c->d_blocks, c->d_active, c->d_coef, c->d_hcount, c->d_S, c->d_Y,
                              ld, c->d_g, c->d_dir, c->d_part, c->nblk, c->st));
      HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * c->nblk,
                            hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      // descent check (ot_op.cuh:339-351): fall back to -g and clear the history
      bool any_fix = false;
      for (int p = 0; p < np; ++p) {
        if (!act[p]) continue;
        double gd = 0.0;
        for (int b = c->probs[p].blk_off; b < c->probs[p].blk_off + c->probs[p].nblk; ++b)
          gd += c->h_part[b];
        if (gd >= 0.0) {
          L[p].loop_iter = 0;
          L[p].H0 = 1.0;
          double* cf = c->h_coef + (size_t)p * kCoefWidth;
          std::fill(cf, cf + kCoefWidth, 0.0);
          cf[0] = 1.0;
          c->h_ismall[p] = 0;
          any_fix = true;
        } else {
          act[p] = 0;
        }
      }
      if (any_fix) {
        std::memcpy(c->h_active, act.data(), np);
        HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, np, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_coef, c->h_coef, sizeof(double) * np * kCoefWidth,
                              hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_hcount, c->h_ismall, sizeof(int) * np,
                              hipMemcpyHostToDevice, c->st));
        HIPCHK(launch_direction(c->d_blocks, c->d_active, c->d_coef, c->d_hcount, c->d_S,
                                c->d_Y, ld, c->d_g, c->d_dir, c->d_part, c->nblk, c->st));
      }
      for (int p = 0; p < np; ++p) {
        if (ph[p] != kDir || iters[p] >= o.max_iter) continue;
        f_prev[p] = cur[p].f;
        t[p] = o.initial_step;
        ph[p] = kTrial;
      }
    }

    // ---- one line-search trial for every problem in the trial phase
    bool any_trial = false;
    for (int p = 0; p < np; ++p) {
      act[p] = ph[p] == kTrial;
      any_trial |= act[p] != 0;
      c->h_small[p] = t[p];
    }
    if (!any_trial) break;
    HIPCHK(hipMemcpyAsync(c->d_step, c->h_small, sizeof(double) * np, hipMemcpyHostToDevice,
                          c->st));
    CHK(evaluate_device(c, c->d_w, c->d_dir, act, trial,
                        std::min(c->opt.max_repairs, o.repair_rounds)));

    // ---- decisions (ot_op.cuh:360-431)
    std::vector<uint8_t> acc(np, 0);
    bool any_acc = false;
    for (int p = 0; p < np; ++p) {
      if (!act[p]) continue;
      ++evals[p];
      bool ok = false;
      if (trial[p].overflow > 0 || trial[p].fails > 0) {
        // over the vertex limit, or unverified after the repair rounds: reject the trial
        t[p] *= 0.5;
      } else if (trial[p].f < f_prev[p] || (o.accept_rule && trial[p].gd <= 0.0)) {
        ok = true;
      } else {
        t[p] *= 0.5;
      }
      if (ok) {
        acc[p] = 1;
        any_acc = true;
        // an accepted trial passed the conformity check: an unverified start no longer
        // describes the iterate
        if (status[p] == COT_EUNVERIFIED) status[p] = COT_OK;
        continue;
      }
      if (t[p] <= 1e-4 * o.initial_step) {
        // revert (w and g were never changed), restart with an empty history
        L[p].restart = true;
        L[p].loop_iter = 0;
        L[p].H0 = 1.0;
        ++iters[p];
        ph[p] = kDir;
      }
    }
    if (any_acc) {
      for (int p = 0; p < np; ++p) {
        c->h_small[p] = t[p];
        c->h_ismall[p] = L[p].loop_iter % L[p].m;
        c->h_ismall[np + p] = std::min(L[p].loop_iter + 1, L[p].m);
      }
      std::memcpy(c->h_active, acc.data(), np);
      HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, np, hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->d_step, c->h_small, sizeof(double) * np, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(hipMemcpyAsync(c->d_slot, c->h_ismall, sizeof(int) * np, hipMemcpyHostToDevice,
                            c->st));
      HIPCHK(hipMemcpyAsync(c->d_hcount, c->h_ismall + np, sizeof(int) * np,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(launch_accept(c->d_blocks, c->d_active, c->d_step, c->d_slot, c->d_hcount,
                           c->d_mass, c->d_nu, c->d_dir, c->d_w, c->d_g, c->d_S, c->d_Y, ld,
                           c->d_part, c->nblk, c->st));
      HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * kAcceptPartials * c->nblk,
                            hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      for (int p = 0; p < np; ++p) {
        if (!acc[p]) continue;
        LState& S = L[p];
        const int h = S.loop_iter % S.m;
        const int H = std::min(S.loop_iter + 1, S.m);
        double P[kAcceptPartials] = {};
        for (int b = c->probs[p].blk_off; b < c->probs[p].blk_off + c->probs[p].nblk; ++b)
          for (int k = 0; k < kAcceptPartials; ++k) P[k] += c->h_part[(size_t)b * kAcceptPartials + k];
        for (int b = 0; b < H; ++b) {
          S.SY[h][b] = P[b];
          S.SY[b][h] = P[kMaxHist + b];
          S.YY[h][b] = S.YY[b][h] = P[2 * kMaxHist + b];
          S.Sg[b] = P[3 * kMaxHist + b];
          S.Yg[b] = P[4 * kMaxHist + b];
        }
        S.SY[h][h] = P[h];
        const double ys = S.SY[h][h], yy = S.YY[h][h];
        S.rho[h] = 1.0 / ys;
        S.H0 = yy > o.h0_threshold ? ys / yy : 1.0;
        S.loop_iter += 1;
        S.restart = false;
        cur[p] = trial[p];
        ++iters[p];
        ph[p] = kDir;
        bool done = false;
        if (o.stop_rule == COT_STOP_MAXREL) {
          done = cur[p].maxrel < o.tol;
        } else if (o.stop_rule == COT_STOP_HANDOFF) {
          done = cur[p].maxrel < o.tol && cur[p].zeros == 0.0;
        } else {
          done = snap_g[p] < o.tol && snap_zero[p] == 0.0;
        }
        if (o.verbose > 0 && (iters[p] % o.verbose == 0 || done))
          std::printf("[cot] prob %d iter %d f %.12e |g| %.6e maxrel %.3e empty %g step %g\n", p,
                      iters[p], cur[p].f, std::sqrt(cur[p].g2), cur[p].maxrel, cur[p].zeros,
                      t[p]);
        if (done) {
          conv[p] = 1;
          ph[p] = kDone;
        }
      }
    }
  }

  CHK(download_sorted(c, c->d_w, w));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  int rc = COT_OK;
  for (int p = 0; p < np; ++p) {
    if (!conv[p] && status[p] == COT_OK) status[p] = COT_ENOTCONV;
    if (status[p] != COT_OK && rc == COT_OK) rc = status[p];
    if (reports) {
      cot_report& r = reports[p];
      r.iterations = iters[p];
      r.evaluations = evals[p];
      r.converged = conv[p];
      r.status = status[p];
      r.f = cur[p].f;
      r.grad_norm = std::sqrt(cur[p].g2);
      r.max_rel_err = cur[p].maxrel;
      r.seconds = secs;
    }
  }
  if (rc == COT_ENOTCONV) g_err = "iteration limit reached";
  if (rc == COT_EUNVERIFIED) g_err = "cells failed the conformity check at the largest ring";
  return rc;
}

}  // extern "C"
