// Device-resident L-BFGS for a batch of OT problems: no host round trip inside a round.
//
// A round gives every unfinished problem one line-search trial:
//   dev_direction_kernel   problems that start an iteration: d from the two-loop coefficients;
//                          last block: descent check (ot_op.cuh:339-351), trial set-up
//   cell_kernel / integrate_kernel with the conformity check / repair rounds (cot_kernels.hip)
//   dev_decide_kernel      FP64 reductions of the trial; last block: accept / halve / restart
//                          (ot_op.cuh:360-431)
//   dev_accept_kernel      w += t d, g = mass - nu, history pair and its dot products; last block:
//                          Gram update, H0 (ot_op.cuh:434-466), stopping test (:510), and the
//                          two-loop recursion (:298-337) for the next direction, run in
//                          coefficient space
// Every kernel reads the per-problem DevState, so the round is a fixed launch sequence that can
// be captured once in a hipGraph and replayed.
//
// The "last block" of a problem is found with a per-problem counter: each block publishes its
// partial sums, fences, and increments the counter; the block that sees count == nblk - 1 reads
// all partials (after an acquire fence) and resets the counter.
#include "../../include/caustic_ot.h"
#include "cot_internal.h"

namespace cot {

namespace {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Partials that one block publishes for the problem's last block: device-scope atomic stores and
// loads, which reach the level every XCD shares, so the hand-over needs no L2 write-back or
// invalidation (see arrive_last).
__device__ __forceinline__ void put_partial(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double get_partial(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum `nv` per-block partials (layout part[v * stride + b], blocks [b0, b0 + nb)) into out[v]
// using the whole block; result valid in out[] for all threads after return.
__device__ void sum_partials(const double* part, size_t stride, int b0, int nb, int nv,
                             double* lds_w, double* out) {
  constexpr int kChunk = 16;
  const int nw = blockDim.x >> 6, wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int c0 = 0; c0 < nv; c0 += kChunk) {
    double acc[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
      for (int k = 0; k < kChunk; ++k)
        if (c0 + k < nv) acc[k] += get_partial(&part[(size_t)(c0 + k) * stride + b0 + b]);
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const double s = wsum(acc[k]);
      if (ln == 0) lds_w[wv * kChunk + k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kChunk && c0 + (int)threadIdx.x < nv) {
      double s = 0.0;
      for (int q = 0; q < nw; ++q) s += lds_w[q * kChunk + threadIdx.x];
      out[c0 + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

__device__ double block_sum1(double v, double* red) {
  v = wsum(v);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  if (ln == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
  __syncthreads();
  return s;
}

__device__ double block_max1(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  if (ln == 0) red[wv] = v;
  __syncthreads();
  double s = -INFINITY;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s = fmax(s, red[k]);
  __syncthreads();
  return s;
}

// Count this block's arrival and learn whether it is the problem's last block.  Thread 0 stored
// the block's partials (put_partial) and waits for those stores to complete before it counts the
// arrival; the last block reads every partial with get_partial after its count returned.  (An
// agent-scope fence in their place writes back and invalidates the XCD's L2 in every block.)
__device__ bool arrive_last(int* counter, int nblk) {
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const int prev = atomicAdd(counter, 1);
    last = (prev == nblk - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) atomicExch(counter, 0);
  return last != 0;
}

// Two-loop recursion of ot_op.cuh:298-337 in coefficient space (d = dg g + sum qy y + sum qs s),
// stored as coef = [cg, cy[], cs[]] with d = -(cg g + sum cy y + sum cs s).  Runs on one thread
// over a shared-memory copy of the Gram matrices (L), with its work arrays in shared memory too.
struct LdsHist {
  double SY[kMaxHist][kMaxHist];
  double YY[kMaxHist][kMaxHist];
  double Sg[kMaxHist], Yg[kMaxHist], rho[kMaxHist], alpha[kMaxHist], qy[kMaxHist], qs[kMaxHist];
};

__device__ void plan_direction(LdsHist& L, int loop_iter, bool restart, double H0, int m,
                               double* coef, int& hcount) {
  constexpr int K = kMaxHist;
  for (int b = 0; b < K; ++b) L.qy[b] = L.qs[b] = 0.0;
  const int maxidx = min(loop_iter, m);
  double dg = -1.0;
  if (!restart) {
    for (int i = 1; i <= maxidx; ++i) {
      const int idx = ((loop_iter - i) % m + m) % m;
      double sd = dg * L.Sg[idx];
      for (int b = 0; b < maxidx; ++b) sd += L.qy[b] * L.SY[idx][b];
      L.alpha[idx] = sd * L.rho[idx];
      L.qy[idx] -= L.alpha[idx];
    }
    dg *= H0;
    for (int b = 0; b < maxidx; ++b) L.qy[b] *= H0;
    for (int i = maxidx; i >= 1; --i) {
      const int idx = ((loop_iter - i) % m + m) % m;
      double yd = dg * L.Yg[idx];
      for (int b = 0; b < maxidx; ++b) yd += L.qy[b] * L.YY[idx][b] + L.qs[b] * L.SY[b][idx];
      const double beta = L.rho[idx] * yd;
      L.qs[idx] += L.alpha[idx] - beta;
    }
  }
  coef[0] = -dg;
  for (int b = 0; b < K; ++b) {
    coef[1 + b] = -L.qy[b];
    coef[1 + K + b] = -L.qs[b];
  }
  hcount = restart ? 0 : maxidx;
}

}  // namespace

// ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(kCellBlock) void dev_direction_kernel(
    const ProbDesc* probs, const BlockDesc* blocks, DevState* st, SolveParams sp,
    const double* S, const double* Y, size_t ld, const double* g, double* d, double* part,
    size_t pstride, int* counter, uint8_t* active) {
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  DevState& s = st[bd.prob];
  if (s.phase != kPhaseDir) return;  // uniform per block
  const ProbDesc pd = probs[bd.prob];
  const int ln = threadIdx.x;
  const int H = s.hcount;
  double gd = 0.0;
  if (ln < bd.count) {
    const int i = bd.first + ln;
    double acc = s.coef[0] * g[i];
    for (int h = 0; h < H; ++h)
      acc += s.coef[1 + h] * Y[h * ld + i] + s.coef[1 + kMaxHist + h] * S[h * ld + i];
    const double di = -acc;
    d[i] = di;
    gd = g[i] * di;
  }
  const double v = block_sum1(gd, red);
  if (ln == 0) put_partial(&part[blockIdx.x], v);
  if (!arrive_last(&counter[bd.prob], pd.nblk)) return;
  __shared__ double tot[1];
  __shared__ double lw[4 * 16];
  sum_partials(part, pstride, pd.blk_off, pd.nblk, 1, lw, tot);
  if (ln != 0) return;
  if (s.iters >= sp.max_iter) {
    s.phase = kPhaseDone;
    active[bd.prob] = 0;
    return;
  }
  s.steep = 0;
  if (tot[0] >= 0.0) {  // not a descent direction: use -g, clear the history
    s.steep = 1;
    s.loop_iter = 0;
    s.H0 = 1.0;
  }
  s.snap_g = sqrt(s.g2_cur);
  s.snap_zero = s.zeros_cur;
  s.f_prev = s.f_cur;
  s.t = sp.t0;
  s.phase = kPhaseTrial;
  active[bd.prob] = 1;
}

// Reductions of a trial evaluation and the line-search decision.
// Partials (layout [v][block]): 0 f, 1 max rel err, 2 empty-mass count, 3 |g|^2, 4 g.d_trial,
// 5 cell area (the cover check, kCoverTol)
__global__ __launch_bounds__(kCellBlock) void dev_decide_kernel(
    const ProbDesc* probs, const BlockDesc* blocks, DevState* st, SolveParams sp,
    const double* mass, const double* fcell, const double* area, const double* nu,
    const double* g, const double* d,
    const int* fails, int* overflow, double* part, size_t pstride, int* counter,
    uint8_t* active, int* fail_counters, int n_fail_counters, int np) {
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  DevState& s = st[bd.prob];
  if (s.phase != kPhaseTrial) return;
  const ProbDesc pd = probs[bd.prob];
  const int ln = threadIdx.x;
  double f = 0, mr = 0, z = 0, g2 = 0, gd = 0, ar = 0;
  if (ln < bd.count) {
    const int i = bd.first + ln;
    const double gi = mass[i] - nu[i];
    const double di = s.steep ? -g[i] : d[i];
    f = fcell[i];
    mr = fabs(gi) / nu[i];
    z = mass[i] == 0.0 ? 1.0 : 0.0;
    g2 = gi * gi;
    gd = gi * di;
    ar = area[i];
  }
  double v;
  v = block_sum1(f, red);  if (ln == 0) put_partial(&part[0 * pstride + blockIdx.x], v);
  v = block_max1(mr, red); if (ln == 0) put_partial(&part[1 * pstride + blockIdx.x], v);
  v = block_sum1(z, red);  if (ln == 0) put_partial(&part[2 * pstride + blockIdx.x], v);
  v = block_sum1(g2, red); if (ln == 0) put_partial(&part[3 * pstride + blockIdx.x], v);
  v = block_sum1(gd, red); if (ln == 0) put_partial(&part[4 * pstride + blockIdx.x], v);
  v = block_sum1(ar, red); if (ln == 0) put_partial(&part[5 * pstride + blockIdx.x], v);
  if (!arrive_last(&counter[bd.prob], pd.nblk)) return;
  __shared__ double lw[4 * 16];
  __shared__ double tot[6];
  sum_partials(part, pstride, pd.blk_off, pd.nblk, 6, lw, tot);
  // the max relative error is a max, not a sum
  double mloc = 0.0;
  for (int b = ln; b < pd.nblk; b += blockDim.x)
    mloc = fmax(mloc, get_partial(&part[1 * pstride + pd.blk_off + b]));
  mloc = block_max1(mloc, red);
  if (ln == 0) tot[1] = mloc;
  __syncthreads();
  if (ln != 0) return;
  const int p = bd.prob;
  (void)fails;
  const int unresolved = fail_counters[(size_t)(n_fail_counters - 1) * np + p];
  const int ovf = overflow[p];
  overflow[p] = 0;
  for (int r = 0; r < n_fail_counters; ++r) fail_counters[(size_t)r * np + p] = 0;
  s.evals += 1;
  s.f_try = tot[0];
  s.maxrel_try = tot[1];
  s.zeros_try = tot[2];
  s.g2_try = tot[3];
  s.gd_try = tot[4];
  bool ok = false;
  // the cover check carries information when the edge check passed (cells still unverified
  // miss the target's area anyway, and their adjacency is the right start for the next trial):
  // a consistent multiple cover is rejected and the next trial searches bucket rings
  const bool covered = covers_domain(pd, tot[5]);
  s.fresh = (!covered && unresolved == 0 && ovf == 0) ? 1 : 0;
  if (ovf > 0 || unresolved > 0 || !covered) {
    // a cell over the vertex limit, cells still unverified after the repair rounds, or cells
    // that do not tile the target: reject the trial like the reference's "polygon too complex"
    // rule (ot_op.cuh:376-382); the raised rings stay for the next trial
    s.t *= 0.5;
  } else if (s.f_try < s.f_prev || (sp.mixed_ls && s.gd_try <= 0.0)) {
    ok = true;
  } else {
    s.t *= 0.5;
  }
  if (ok) {
    s.accept = 1;
    // an accepted trial passed the conformity check: an unverified start no longer describes
    // the solve's iterate
    if (s.status == COT_EUNVERIFIED) s.status = COT_OK;
    return;
  }
  if (s.t <= 1e-4 * sp.t0) {
    // revert (w and g were never changed); restart with an empty history
    s.restart = 1;
    s.loop_iter = 0;
    s.H0 = 1.0;
    s.iters += 1;
    s.steep = 0;
    s.coef[0] = 1.0;
    for (int k = 1; k < 1 + 2 * kMaxHist; ++k) s.coef[k] = 0.0;
    s.hcount = 0;
    s.phase = kPhaseDir;
    active[p] = 0;
  }
}

// Accept the trial: history pair, weights, gradient; then the Gram update, stopping test and the
// next direction's coefficients.  Partials (layout [v][block]), H = history count after insert:
//   [h] s_new.y_h, [K + h] s_h.y_new, [2K + h] y_new.y_h, [3K + h] g_new.s_h, [4K + h] g_new.y_h
__global__ __launch_bounds__(kCellBlock) void dev_accept_kernel(
    const ProbDesc* probs, const BlockDesc* blocks, DevState* st, SolveParams sp,
    const double* mass, const double* nu, const double* d, double* w, double* g, double* S,
    double* Y, size_t ld, double* part, size_t pstride, int* counter, uint8_t* active) {
  const BlockDesc bd = blocks[blockIdx.x];
  DevState& s = st[bd.prob];
  if (s.phase != kPhaseTrial || !s.accept) return;
  const ProbDesc pd = probs[bd.prob];
  const int ln = threadIdx.x;
  const int m = sp.m;
  const int slot = s.loop_iter % m;
  const int H = min(s.loop_iter + 1, m);
  const bool live = ln < bd.count;
  const int i = bd.first + ln;
  double sn = 0.0, yn = 0.0, gn = 0.0;
  if (live) {
    const double t = s.t;
    const double di = s.steep ? -g[i] : d[i];
    gn = mass[i] - nu[i];
    sn = t * di;
    yn = gn - g[i];
    S[slot * ld + i] = sn;
    Y[slot * ld + i] = yn;
    w[i] = w[i] + t * di;
    g[i] = gn;
  }
  // dot-product partials, three history slots (15 values) per barrier
  {
    __shared__ double wpart[kCellBlock / 64][15];
    const int wv = ln >> 6, wl = ln & 63;
    for (int h0 = 0; h0 < H; h0 += 3) {
      double val[15];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int h = h0 + k;
        double sh = 0.0, yh = 0.0;
        if (live && h < H) {
          sh = S[h * ld + i];
          yh = Y[h * ld + i];
        }
        val[5 * k + 0] = sn * yh;
        val[5 * k + 1] = sh * yn;
        val[5 * k + 2] = yn * yh;
        val[5 * k + 3] = gn * sh;
        val[5 * k + 4] = gn * yh;
      }
#pragma unroll
      for (int q = 0; q < 15; ++q) {
        const double t = wsum(val[q]);
        if (wl == 0) wpart[wv][q] = t;
      }
      __syncthreads();
      if (ln < 15) {
        const int k = ln / 5, q = ln % 5, h = h0 + k;
        if (h < H) {
          double t = 0.0;
          for (int w2 = 0; w2 < kCellBlock / 64; ++w2) t += wpart[w2][ln];
          part[(size_t)(q * kMaxHist + h) * pstride + blockIdx.x] = t;
        }
      }
      __syncthreads();
    }
  }
  (void)counter;
  (void)pd;
}

// The accepted problems' reductions of the accept partials (one block of kFinishBlock threads
// per problem: each wave sums some of the 5H values over the problem's blocks), then the Gram
// update, H0 (ot_op.cuh:434-466), stopping test (:510) and the next direction's coefficients.
constexpr int kFinishBlock = 1024;
__global__ __launch_bounds__(kFinishBlock) void dev_accept_finish_kernel(
    const ProbDesc* probs, DevState* st, SolveParams sp, const double* part, size_t pstride,
    uint8_t* active) {
  DevState& s = st[blockIdx.x];
  if (s.phase != kPhaseTrial || !s.accept) return;
  const ProbDesc pd = probs[blockIdx.x];
  const int ln = threadIdx.x, wv = ln >> 6, wl = ln & 63;
  const int m = sp.m;
  const int slot = s.loop_iter % m;
  const int H = min(s.loop_iter + 1, m);
  constexpr int K = kMaxHist;
  __shared__ double P[5 * K];
  __shared__ LdsHist L;
  for (int v = wv; v < 5 * K; v += kFinishBlock / 64) {
    if (v % K >= H) continue;
    double acc = 0.0;
    for (int b = wl; b < pd.nblk; b += 64) acc += part[(size_t)v * pstride + pd.blk_off + b];
    acc = wsum(acc);
    if (wl == 0) P[v] = acc;
  }
  __syncthreads();
  // Gram update in shared memory (the whole block), written back to the state
  for (int e = ln; e < K * K; e += blockDim.x) {
    const int r = e / K, c = e % K;
    double sy = s.SY[r][c], yy = s.YY[r][c];
    if (r == slot && c < H) sy = P[c];
    if (c == slot && r < H) sy = P[K + r];
    if (r == slot && c < H) yy = P[2 * K + c];
    if (c == slot && r < H) yy = P[2 * K + r];
    if (r == slot && c == slot) sy = P[slot];
    L.SY[r][c] = sy;
    L.YY[r][c] = yy;
    s.SY[r][c] = sy;
    s.YY[r][c] = yy;
  }
  for (int b = ln; b < K; b += blockDim.x) {
    L.Sg[b] = b < H ? P[3 * K + b] : s.Sg[b];
    L.Yg[b] = b < H ? P[4 * K + b] : s.Yg[b];
    L.rho[b] = s.rho[b];
    s.Sg[b] = L.Sg[b];
    s.Yg[b] = L.Yg[b];
  }
  __syncthreads();
  if (ln != 0) return;
  const double ys = L.SY[slot][slot], yy = L.YY[slot][slot];
  L.rho[slot] = 1.0 / ys;
  s.rho[slot] = L.rho[slot];
  s.H0 = yy > sp.h0_thr ? ys / yy : 1.0;
  s.loop_iter += 1;
  s.restart = 0;
  s.accept = 0;
  s.steep = 0;
  s.iters += 1;
  s.f_cur = s.f_try;
  s.g2_cur = s.g2_try;
  s.zeros_cur = s.zeros_try;
  s.maxrel_cur = s.maxrel_try;
  bool done;
  if (sp.stop_rule == 1) {
    done = s.maxrel_cur < sp.tol;
  } else if (sp.stop_rule == 2) {
    done = s.maxrel_cur < sp.tol && s.zeros_cur == 0.0;
  } else {
    done = s.snap_g < sp.tol && s.snap_zero == 0.0;
  }
  if (done) {
    s.conv = 1;
    s.phase = kPhaseDone;
    active[blockIdx.x] = 0;
    return;
  }
  int hc = 0;
  plan_direction(L, s.loop_iter, s.restart != 0, s.H0, m, s.coef, hc);
  s.hcount = hc;
  s.phase = kPhaseDir;
  active[blockIdx.x] = 0;
}

// Next ring of the sites of problems whose final check was clean (see adapt_ring_kernel).
__global__ void dev_adapt_ring_kernel(const BlockDesc* blocks, const DevState* st,
                                      const int* final_fails, const uint8_t* used, uint8_t* ring,
                                      int ring_min, int ring_max) {
  const BlockDesc bd = blocks[blockIdx.x];
  if (st[bd.prob].phase != kPhaseTrial || final_fails[bd.prob] != 0) return;
  const int ln = threadIdx.x;
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  ring[i] = (uint8_t)min(max((int)used[i] + 1, ring_min), ring_max);
}

// Number of problems that are not finished (one block).
__global__ void dev_count_kernel(const DevState* st, int np, int* out) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  int c = 0;
  for (int p = threadIdx.x; p < np; p += blockDim.x) c += st[p].phase != kPhaseDone ? 1 : 0;
  if (c) atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) *out = total;
}

// ---------------------------------------------------------------------------------------------

hipError_t launch_dev_count(const DevState* st, int np, int* out, hipStream_t stream) {
  hipLaunchKernelGGL(dev_count_kernel, dim3(1), dim3(256), 0, stream, st, np, out);
  return hipGetLastError();
}

hipError_t launch_dev_direction(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                                const SolveParams& sp, const double* S, const double* Y,
                                size_t ld, const double* g, double* d, double* part,
                                size_t pstride, int* counter, uint8_t* active, int nblk,
                                hipStream_t stream) {
  hipLaunchKernelGGL(dev_direction_kernel, dim3(nblk), dim3(kCellBlock), 0, stream, probs, blocks,
                     st, sp, S, Y, ld, g, d, part, pstride, counter, active);
  return hipGetLastError();
}

hipError_t launch_dev_decide(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                             const SolveParams& sp, const double* mass, const double* fcell,
                             const double* area, const double* nu, const double* g,
                             const double* d, const int* fails,
                             int* overflow, double* part, size_t pstride, int* counter,
                             uint8_t* active, int* fail_counters, int n_fail_counters, int np,
                             int nblk, hipStream_t stream) {
  hipLaunchKernelGGL(dev_decide_kernel, dim3(nblk), dim3(kCellBlock), 0, stream, probs, blocks,
                     st, sp, mass, fcell, area, nu, g, d, fails, overflow, part, pstride, counter,
                     active, fail_counters, n_fail_counters, np);
  return hipGetLastError();
}

hipError_t launch_dev_accept(const ProbDesc* probs, const BlockDesc* blocks, DevState* st,
                             const SolveParams& sp, const double* mass, const double* nu,
                             const double* d, double* w, double* g, double* S, double* Y,
                             size_t ld, double* part, size_t pstride, int* counter,
                             uint8_t* active, int nblk, hipStream_t stream) {
  hipLaunchKernelGGL(dev_accept_kernel, dim3(nblk), dim3(kCellBlock), 0, stream, probs, blocks,
                     st, sp, mass, nu, d, w, g, S, Y, ld, part, pstride, counter, active);
  return hipGetLastError();
}

hipError_t launch_dev_accept_finish(const ProbDesc* probs, DevState* st, const SolveParams& sp,
                                    const double* part, size_t pstride, uint8_t* active, int np,
                                    hipStream_t stream) {
  hipLaunchKernelGGL(dev_accept_finish_kernel, dim3(np), dim3(kFinishBlock), 0, stream, probs, st,
                     sp, part, pstride, active);
  return hipGetLastError();
}


hipError_t launch_dev_adapt_ring(const BlockDesc* blocks, const DevState* st,
                                 const int* final_fails, const uint8_t* used, uint8_t* ring,
                                 int ring_min, int ring_max, int nblk, hipStream_t stream) {
  hipLaunchKernelGGL(dev_adapt_ring_kernel, dim3(nblk), dim3(kCellBlock), 0, stream, blocks, st,
                     final_fails, used, ring, ring_min, ring_max);
  return hipGetLastError();
}

}  // namespace cot
