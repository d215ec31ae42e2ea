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

namespace cot {
hipError_t launch_cell(const CellArgs& a, int nblocks, bool fp32, hipStream_t st);
hipError_t launch_verify(const VerifyArgs& a, int nblocks, hipStream_t st);
hipError_t launch_raise_ring(const BlockDesc* b, const uint8_t* redo, uint8_t* ring, int rmax,
                             float* bhalf, const uint8_t* boxhit, int nblocks, hipStream_t st);
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
hipError_t launch_dev_count(const DevState* st, int np, int* out, hipStream_t stream);
hipError_t launch_dev_direction(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                                const SolveParams& sp, const double* S, const double* Y,
                                size_t ld, const double* g, double* d, double* part,
                                size_t pstride, int* counter, uint8_t* active, int nblk,
                                hipStream_t stream);
hipError_t launch_dev_decide(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                             const SolveParams& sp, const double* mass, const double* fcell,
                             const double* nu, const double* g, const double* d, const int* fails,
                             int* overflow, double* part, size_t pstride, int* counter,
                             uint8_t* active, int* fail_counters, int n_fail_counters, int np,
                             int nblk, hipStream_t stream);
hipError_t launch_dev_accept(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                             const SolveParams& sp, const double* mass, const double* nu,
                             const double* d, double* w, double* g, double* S, double* Y,
                             size_t ld, double* part, size_t pstride, int* counter,
                             uint8_t* active, int nblk, hipStream_t stream);
hipError_t launch_dev_raise_ring(const BlockDesc* blocks, const DevState* st, const int* gate,
                                 uint8_t* redo, uint8_t* ring, int ring_max, float* bhalf,
                                 const uint8_t* boxhit, int nblk, hipStream_t stream);
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
  std::vector<void*> allocs;
  // device
  ProbDesc* d_probs = nullptr;
  BlockDesc* d_blocks = nullptr;
  double2* d_xy = nullptr;
  double* d_nu = nullptr;
  int2* d_buckets = nullptr;
  void* d_img = nullptr;
  double *d_w = nullptr, *d_dir = nullptr, *d_step = nullptr, *d_mass = nullptr;
  double *d_fcell = nullptr, *d_area = nullptr, *d_g = nullptr;
  double2 *d_cen = nullptr, *d_geo = nullptr;
  uint8_t *d_exist = nullptr, *d_nedge = nullptr, *d_ring = nullptr, *d_used = nullptr;
  uint8_t *d_redo = nullptr, *d_active = nullptr;
  double2* d_org = nullptr;      // FP32 geometry: clipping frame origins
  float* d_bhalf = nullptr;      // FP32 geometry: artificial box half-sizes
  uint8_t* d_boxhit = nullptr;   // FP32 geometry: box reached the final polygon
  int* d_elab = nullptr;
  float2* d_espan = nullptr;
  int *d_overflow = nullptr, *d_fails = nullptr;
  double* d_part = nullptr;
  double *d_S = nullptr, *d_Y = nullptr;
  int hist_alloc = 0;
  double* d_coef = nullptr;
  int *d_slot = nullptr, *d_hcount = nullptr;
  // device-resident solver
  DevState* d_state = nullptr;
  int* d_cnt = nullptr;        // last-block counters: direction, decide, accept (np each)
  double* d_pdir = nullptr;    // direction partials  [1][nblk]
  double* d_peval = nullptr;   // decide partials     [5][nblk]
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
  hipEvent_t ev[4] = {};
  cot_stats stats{};

  ~cot_context() {
    if (st) hipStreamSynchronize(st);
    for (void* p : allocs) hipFree(p);
    if (h_part) hipHostFree(h_part);
    if (h_counts) hipHostFree(h_counts);
    if (h_vec) hipHostFree(h_vec);
    if (h_small) hipHostFree(h_small);
    if (h_coef) hipHostFree(h_coef);
    if (h_ismall) hipHostFree(h_ismall);
    if (h_active) hipHostFree(h_active);
    if (h_ndone) hipHostFree(h_ndone);
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (st) hipStreamDestroy(st);
  }
};

namespace {

template <typename T>
int dalloc(cot_context* c, T** p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (hipMalloc(&q, bytes) != hipSuccess) return fail(COT_ENOMEM, "hipMalloc failed");
  c->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return COT_OK;
}

// Bucket grid and tile-ordered permutation of one problem's sites.
void bucket_sort(const cot_problem& pr, double spb, ProbDesc& pd, std::vector<int>& order,
                 std::vector<int2>& buckets) {
  const int n = pr.n_sites;
  double xmin = pr.proj_domain[0], xmax = pr.proj_domain[1];
  double ymin = pr.proj_domain[2], ymax = pr.proj_domain[3];
  for (int i = 0; i < n; ++i) {
    xmin = std::min(xmin, pr.sites[2 * i]);
    xmax = std::max(xmax, pr.sites[2 * i]);
    ymin = std::min(ymin, pr.sites[2 * i + 1]);
    ymax = std::max(ymax, pr.sites[2 * i + 1]);
  }
  const double wdt = std::max(xmax - xmin, 1e-300), hgt = std::max(ymax - ymin, 1e-300);
  double bs = std::sqrt(wdt * hgt * spb / std::max(n, 1));
  int gx = std::max(1, (int)std::ceil(wdt / bs));
  int gy = std::max(1, (int)std::ceil(hgt / bs));
  // keep the grid bounded for degenerate inputs
  while ((long long)gx * gy > 4LL * n + 64) {
    bs *= 1.25;
    gx = std::max(1, (int)std::ceil(wdt / bs));
    gy = std::max(1, (int)std::ceil(hgt / bs));
  }
  pd.gx = gx;
  pd.gy = gy;
  pd.bx0 = xmin;
  pd.by0 = ymin;
  pd.bsize = bs;
  pd.binv = 1.0 / bs;
  const int tiles_x = (gx + 7) / 8;
  std::vector<long long> key(n);
  std::vector<int> cell(n);
  for (int i = 0; i < n; ++i) {
    int cx = std::min(std::max((int)std::floor((pr.sites[2 * i] - xmin) * pd.binv), 0), gx - 1);
    int cy =
        std::min(std::max((int)std::floor((pr.sites[2 * i + 1] - ymin) * pd.binv), 0), gy - 1);
    cell[i] = cy * gx + cx;
    long long tile = (long long)(cy / 8) * tiles_x + (cx / 8);
    key[i] = (tile * 64 + (cy % 8) * 8 + (cx % 8));
  }
  order.resize(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  buckets.assign((size_t)gx * gy, make_int2(0, 0));
  for (int r = 0; r < n; ++r) {
    int b = cell[order[r]];
    if (buckets[b].y == 0) buckets[b].x = r;
    buckets[b].y += 1;
  }
}

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

// One evaluation of every active problem at w (+ step * dir when dir is set), with the
// conformity check and repairs.  Leaves mass/fcell/... on the device and summaries on the host.
int evaluate_device(cot_context* c, const double* d_w, const double* d_dir,
                    const std::vector<uint8_t>& active, std::vector<EvalSum>& sum) {
  hipStream_t st = c->st;
  std::memcpy(c->h_active, active.data(), c->np);
  HIPCHK(hipMemcpyAsync(c->d_active, c->h_active, c->np, hipMemcpyHostToDevice, st));
  HIPCHK(launch_fill_i32(c->d_fails, 2 * (size_t)c->np, 0, st));
  HIPCHK(launch_fill_u8(c->d_redo, c->ntot, 0, st));

  CellArgs ca{};
  ca.probs = c->d_probs;
  ca.blocks = c->d_blocks;
  ca.xy = c->d_xy;
  ca.w = d_w;
  ca.dir = d_dir;
  ca.step = c->d_step;
  ca.nu = c->d_nu;
  ca.buckets = c->d_buckets;
  ca.img = c->d_img;
  ca.ring = c->d_ring;
  ca.redo = c->d_redo;
  ca.redo_mode = 0;
  ca.st = nullptr;
  ca.g = c->d_g;
  ca.gate = nullptr;
  ca.active = c->d_active;
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
  ca.bhalf = c->d_bhalf;
  ca.boxhit = c->d_boxhit;

  VerifyArgs va{};
  va.probs = c->d_probs;
  va.blocks = c->d_blocks;
  va.xy = c->d_xy;
  va.nedge = c->d_nedge;
  va.elab = c->d_elab;
  va.espan = c->d_espan;
  va.active = c->d_active;
  va.gate = nullptr;
  va.redo = c->d_redo;
  va.fails = c->d_fails;
  va.tol_rel = c->opt.verify_tol;
  va.boxhit = c->fp32 ? c->d_boxhit : nullptr;

  HIPCHK(hipEventRecord(c->ev[0], st));
  HIPCHK(launch_cell(ca, c->nblk, c->fp32, st));
  HIPCHK(hipEventRecord(c->ev[1], st));
  HIPCHK(launch_verify(va, c->nblk, st));
  HIPCHK(hipEventRecord(c->ev[2], st));
  HIPCHK(launch_reduce_eval(c->d_blocks, c->d_mass, c->d_fcell, c->d_area, c->d_nu, c->d_exist,
                            d_dir, c->d_part, c->nblk, st));
  HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * kEvalPartials * c->nblk,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(c->h_counts, c->d_fails, sizeof(int) * 2 * c->np, hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->stats.cell_seconds += ms * 1e-3;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
  c->stats.verify_seconds += ms * 1e-3;
  c->stats.cell_launches += 1;
  c->stats.evaluations += 1;
  for (int p = 0; p < c->np; ++p) {
    if (!active[p]) continue;
    c->stats.sites_evaluated += c->probs[p].n;
    c->stats.pixels_covered += (long long)c->probs[p].pm * c->probs[p].pn;
  }

  int rounds = 0;
  for (;;) {
    int tot = 0;
    for (int p = 0; p < c->np; ++p) tot += active[p] ? c->h_counts[p] : 0;
    if (tot == 0 || rounds >= c->opt.max_repairs) break;
    ++rounds;
    HIPCHK(launch_raise_ring(c->d_blocks, c->d_redo, c->d_ring, c->opt.ring_max,
                             c->fp32 ? c->d_bhalf : nullptr, c->d_boxhit, c->nblk, st));
    ca.redo_mode = 1;
    HIPCHK(launch_cell(ca, c->nblk, c->fp32, st));
    HIPCHK(launch_fill_u8(c->d_redo, c->ntot, 0, st));
    HIPCHK(launch_fill_i32(c->d_fails, (size_t)c->np, 0, st));
    HIPCHK(launch_verify(va, c->nblk, st));
    HIPCHK(launch_reduce_eval(c->d_blocks, c->d_mass, c->d_fcell, c->d_area, c->d_nu,
                              c->d_exist, d_dir, c->d_part, c->nblk, st));
    HIPCHK(hipMemcpyAsync(c->h_part, c->d_part, sizeof(double) * kEvalPartials * c->nblk,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(c->h_counts, c->d_fails, sizeof(int) * 2 * c->np,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  c->stats.repair_rounds += rounds;
  HIPCHK(launch_adapt_ring(c->d_blocks, c->d_used, c->d_ring, 1, c->opt.ring_max, c->nblk, st));
  sum_partials(c, sum);
  return COT_OK;
}

// Scatter caller-order host values into sorted order and upload.
int upload_sorted(cot_context* c, const double* src, double* dst) {
  for (int r = 0; r < c->ntot; ++r) c->h_vec[r] = src[c->perm[r]];
  HIPCHK(hipMemcpyAsync(dst, c->h_vec, sizeof(double) * c->ntot, hipMemcpyHostToDevice, c->st));
  return COT_OK;
}

int download_sorted(cot_context* c, const double* src, double* dst) {
  HIPCHK(hipMemcpyAsync(c->h_vec, src, sizeof(double) * c->ntot, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int r = 0; r < c->ntot; ++r) dst[c->perm[r]] = c->h_vec[r];
  return COT_OK;
}

int download_sorted2(cot_context* c, const double2* src, double* dst) {
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
    CHK(dalloc(c, &c->d_peval, 5 * (size_t)c->nblk));
    CHK(dalloc(c, &c->d_pacc, (size_t)kAcceptPartials * c->nblk));
    CHK(dalloc(c, &c->d_ndone, 2));
    if (hipHostMalloc((void**)&c->h_ndone, 2 * sizeof(int)) != hipSuccess)
      return fail(COT_ENOMEM, "hipHostMalloc failed");
    HIPCHK(hipMemset(c->d_cnt, 0, 3 * sizeof(int) * c->np));
  }
  if (c->rfail_rounds < repair_rounds + 1) {
    CHK(dalloc(c, &c->d_rfails, (size_t)(repair_rounds + 1) * c->np));
    c->rfail_rounds = repair_rounds + 1;
  }
  HIPCHK(hipMemset(c->d_rfails, 0, sizeof(int) * c->rfail_rounds * c->np));
  return COT_OK;
}

// One device round (see cot_solver.hip).  Captured once per solve into a hipGraph.
int launch_round(cot_context* c, const SolveParams& sp, int R, hipEvent_t t0, hipEvent_t t1) {
  hipStream_t st = c->st;
  const size_t ld = c->ntot, ps = c->nblk;
  const int np = c->np;
  HIPCHK(launch_dev_direction(c->d_probs, c->d_blocks, c->d_state, sp, c->d_S, c->d_Y, ld, c->d_g,
                              c->d_dir, c->d_pdir, ps, c->d_cnt, c->d_active, c->nblk, st));
  CellArgs ca{};
  ca.probs = c->d_probs;
  ca.blocks = c->d_blocks;
  ca.xy = c->d_xy;
  ca.w = c->d_w;
  ca.dir = c->d_dir;
  ca.step = nullptr;
  ca.nu = c->d_nu;
  ca.buckets = c->d_buckets;
  ca.img = c->d_img;
  ca.ring = c->d_ring;
  ca.redo = c->d_redo;
  ca.redo_mode = 0;
  ca.active = c->d_active;
  ca.st = c->d_state;
  ca.g = c->d_g;
  ca.gate = nullptr;
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
  ca.bhalf = c->d_bhalf;
  ca.boxhit = c->d_boxhit;
  VerifyArgs va{};
  va.probs = c->d_probs;
  va.blocks = c->d_blocks;
  va.xy = c->d_xy;
  va.nedge = c->d_nedge;
  va.elab = c->d_elab;
  va.espan = c->d_espan;
  va.active = c->d_active;
  va.gate = nullptr;
  va.boxhit = c->fp32 ? c->d_boxhit : nullptr;
  va.redo = c->d_redo;
  va.fails = c->d_rfails;
  va.tol_rel = c->opt.verify_tol;
  if (t0) HIPCHK(hipEventRecord(t0, st));
  HIPCHK(launch_cell(ca, c->nblk, c->fp32, st));
  if (t1) HIPCHK(hipEventRecord(t1, st));
  HIPCHK(launch_verify(va, c->nblk, st));
  for (int r = 1; r <= R; ++r) {
    const int* gate = c->d_rfails + (size_t)(r - 1) * np;
    HIPCHK(launch_dev_raise_ring(c->d_blocks, c->d_state, gate, c->d_redo, c->d_ring,
                                 c->opt.ring_max, c->fp32 ? c->d_bhalf : nullptr, c->d_boxhit,
                                 c->nblk, st));
    ca.redo_mode = 1;
    ca.gate = gate;
    HIPCHK(launch_cell(ca, c->nblk, c->fp32, st));
    va.gate = gate;
    va.fails = c->d_rfails + (size_t)r * np;
    HIPCHK(launch_verify(va, c->nblk, st));
  }
  const int* final_fails = c->d_rfails + (size_t)R * np;
  HIPCHK(launch_dev_adapt_ring(c->d_blocks, c->d_state, final_fails, c->d_used, c->d_ring, 1,
                               c->opt.ring_max, c->nblk, st));
  HIPCHK(launch_dev_decide(c->d_probs, c->d_blocks, c->d_state, sp, c->d_mass, c->d_fcell,
                           c->d_nu, c->d_g, c->d_dir, final_fails, c->d_overflow, c->d_peval, ps,
                           c->d_cnt + np, c->d_active, c->d_rfails, R + 1, np, c->nblk, st));
  HIPCHK(launch_dev_accept(c->d_probs, c->d_blocks, c->d_state, sp, c->d_mass, c->d_nu,
                           c->d_dir, c->d_w, c->d_g, c->d_S, c->d_Y, ld, c->d_pacc, ps,
                           c->d_cnt + 2 * np, c->d_active, c->nblk, st));
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
    if (o.stop_rule == COT_STOP_MAXREL && cur[p].maxrel < o.tol) {
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
  sp.stop_rule = o.stop_rule == COT_STOP_MAXREL ? 1 : 0;
  sp.mixed_ls = o.accept_rule;
  sp.ring_max = c->opt.ring_max;
  sp.max_unverified = 4;
  sp.tol = o.tol;
  sp.t0 = o.initial_step;
  sp.h0_thr = o.h0_threshold;

  // capture one round
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int rc = launch_round(c, sp, R, c->ev[0], c->ev[1]);
  hipError_t ce = hipStreamEndCapture(c->st, &graph);
  if (rc != COT_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIPCHK(ce);
  HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
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
  for (;;) {
    for (int k = 0; k < kBatch; ++k) HIPCHK(hipGraphLaunch(exec, c->st));
    rounds += kBatch;
    HIPCHK(launch_dev_count(c->d_state, np, c->d_ndone, c->st));
    HIPCHK(hipMemcpyAsync(c->h_ndone, c->d_ndone, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const int left = c->h_ndone[0];
    // live kernel timing from the event nodes around the full cell launch of the batch's last
    // round (only while that round still had work)
    float ms = 0.f;
    if (left > 0 && hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) {
      c->stats.cell_seconds += ms * 1e-3;
      c->stats.cell_launches += 1;
    }
    if (left == 0 || rounds > max_rounds) break;
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
    c->stats.evaluations += s.evals;
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
      std::printf("[cot] prob %d: %s after %d iterations (%d evaluations), maxrel %.3e\n", p,
                  s.conv ? "converged" : "stopped", s.iters, s.evals, s.maxrel_cur);
  }
  if (result == COT_ENOTCONV) g_err = "iteration limit reached";
  if (result == COT_EUNVERIFIED) g_err = "cells failed the conformity check at the largest ring";
  return result;
}

}  // namespace

extern "C" {

void cot_default_options(cot_options* o) {
  o->precision = COT_PREC_FP64;
  o->sites_per_bucket = 1.0;
  o->ring_init = 2;
  o->ring_max = 32;
  o->verify_tol = 1e-4;
  o->max_repairs = 12;
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
  o->repair_rounds = 2;
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
  c->fp32 = c->opt.precision == COT_PREC_MIXED;
  c->np = n_problems;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(COT_EHIP, "hipStreamCreate failed"));
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return bail(fail(COT_EHIP, "hipEventCreate failed"));

  // layout
  std::vector<int2> all_buckets;
  std::vector<double2> xy;
  std::vector<double> nu;
  size_t img_total = 0;
  int site_off = 0;
  for (int p = 0; p < n_problems; ++p) {
    const cot_problem& pr = problems[p];
    ProbDesc pd{};
    pd.site_off = site_off;
    pd.n = pr.n_sites;
    pd.pm = pr.pm;
    pd.pn = pr.pn;
    pd.img_off = (int)img_total;
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
    }
    pd.nblk = (int)c->blocks.size() - pd.blk_off;
    for (int r = 0; r < pr.n_sites; ++r) {
      int o = order[r];
      c->perm.push_back(site_off + o);
      xy.push_back(make_double2(pr.sites[2 * o], pr.sites[2 * o + 1]));
      nu.push_back(pr.nu[o]);
    }
    img_total += (size_t)pr.pm * pr.pn;
    if (img_total > (size_t)INT32_MAX) return bail(fail(COT_EINVAL, "images too large"));
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
      (rc = dalloc(c, &c->d_active, c->np)) || (rc = dalloc(c, &c->d_elab, N * kMaxVerts)) ||
      (rc = dalloc(c, &c->d_espan, N * kMaxVerts)) ||
      (rc = dalloc(c, &c->d_overflow, c->np)) || (rc = dalloc(c, &c->d_fails, 2 * c->np)) ||
      (rc = dalloc(c, &c->d_part, (size_t)c->nblk * std::max(kEvalPartials, kAcceptPartials))) ||
      (rc = dalloc(c, &c->d_coef, (size_t)c->np * kCoefWidth)) ||
      (rc = dalloc(c, &c->d_slot, c->np)) || (rc = dalloc(c, &c->d_hcount, c->np)) ||
      (rc = dalloc(c, &c->d_org, N)) || (rc = dalloc(c, &c->d_bhalf, N)) ||
      (rc = dalloc(c, &c->d_boxhit, N)))
    return bail(rc);
  // overflow counters live right after the fail counters so one copy reads both
  c->d_overflow = c->d_fails + c->np;
  if (c->fp32) {
    float* im = nullptr;
    if ((rc = dalloc(c, &im, img_total))) return bail(rc);
    c->d_img = im;
  } else {
    double* im = nullptr;
    if ((rc = dalloc(c, &im, img_total))) return bail(rc);
    c->d_img = im;
  }
  if (hipHostMalloc((void**)&c->h_part,
                    sizeof(double) * c->nblk * std::max(kEvalPartials, kAcceptPartials)) ||
      hipHostMalloc((void**)&c->h_counts, sizeof(int) * 2 * c->np) ||
      hipHostMalloc((void**)&c->h_vec, sizeof(double) * N) ||
      hipHostMalloc((void**)&c->h_small, sizeof(double) * c->np) ||
      hipHostMalloc((void**)&c->h_coef, sizeof(double) * c->np * kCoefWidth) ||
      hipHostMalloc((void**)&c->h_ismall, sizeof(int) * 2 * c->np) ||
      hipHostMalloc((void**)&c->h_active, c->np))
    return bail(fail(COT_ENOMEM, "hipHostMalloc failed"));

  // uploads (synchronous: inputs are caller-owned)
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up(c->d_probs, c->probs.data(), sizeof(ProbDesc) * c->np) &&
            up(c->d_blocks, c->blocks.data(), sizeof(BlockDesc) * c->nblk) &&
            up(c->d_xy, xy.data(), sizeof(double2) * N) &&
            up(c->d_nu, nu.data(), sizeof(double) * N) &&
            up(c->d_buckets, all_buckets.data(), sizeof(int2) * all_buckets.size());
  for (int p = 0; p < n_problems && ok; ++p) {
    const cot_problem& pr = problems[p];
    const size_t P = (size_t)pr.pm * pr.pn;
    if (c->fp32) {
      std::vector<float> tmp(pr.rho, pr.rho + P);
      ok = up((float*)c->d_img + c->probs[p].img_off, tmp.data(), sizeof(float) * P);
    } else {
      ok = up((double*)c->d_img + c->probs[p].img_off, pr.rho, sizeof(double) * P);
    }
  }
  {
    // FP32 frames start at the sites with a box of two buckets
    std::vector<float> bh(N);
    for (int p = 0; p < n_problems; ++p)
      for (int r = 0; r < c->probs[p].n; ++r)
        bh[c->probs[p].site_off + r] = (float)(2.0 * c->probs[p].bsize);
    ok = ok && up(c->d_org, xy.data(), sizeof(double2) * N) &&
         up(c->d_bhalf, bh.data(), sizeof(float) * N) &&
         hipMemset(c->d_boxhit, 0, N) == hipSuccess;
  }
  if (!ok) return bail(fail(COT_EHIP, "upload failed"));
  if (launch_fill_u8(c->d_ring, N, (uint8_t)c->opt.ring_init, c->st) != hipSuccess ||
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
  if (f)
    for (int p = 0; p < c->np; ++p) f[p] = sum[p].f;
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

int cot_get_stats(const cot_context* c, cot_stats* out) {
  if (!c || !out) return fail(COT_EINVAL, "null argument");
  *out = c->stats;
  return COT_OK;
}

int cot_reset_stats(cot_context* c) {
  if (!c) return fail(COT_EINVAL, "null context");
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
      if (o.stop_rule == COT_STOP_MAXREL && cur[p].maxrel < o.tol) {
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
    CHK(evaluate_device(c, c->d_w, c->d_dir, act, trial));

    // ---- decisions (ot_op.cuh:360-431)
    std::vector<uint8_t> acc(np, 0);
    bool any_acc = false;
    for (int p = 0; p < np; ++p) {
      if (!act[p]) continue;
      ++evals[p];
      bool ok = false;
      if (trial[p].fails) status[p] = COT_EUNVERIFIED;
      if (trial[p].overflow > 0) {
        t[p] *= 0.5;
      } else if (trial[p].f < f_prev[p] || (o.accept_rule && trial[p].gd <= 0.0)) {
        ok = true;
      } else {
        t[p] *= 0.5;
      }
      if (ok) {
        acc[p] = 1;
        any_acc = true;
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
        } else {
          done = snap_g[p] < o.tol && snap_zero[p] == 0.0;
        }
        if (o.verbose > 0 && (iters[p] % o.verbose == 0 || done))
          std::printf("[cot] prob %d iter %d f %.12e |g| %.6e maxrel %.3e step %g\n", p,
                      iters[p], cur[p].f, std::sqrt(cur[p].g2), cur[p].maxrel, t[p]);
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
