// Internal layout shared by the HIP kernels (cot_kernels.hip) and the host driver (cot_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cot {

// Device and pinned host memory cached by size across owners in the process (cot_pool.cpp): a
// block is released only after the stream that used it has drained.  nullptr on failure.
void* pool_alloc_device(size_t bytes);
void pool_free_device(void* p);
void* pool_alloc_pinned(size_t bytes);
void pool_free_pinned(void* p);

// Largest polygon a cell may hold (vertices == edges): the stride of the per-site edge and
// vertex records and the polygon of the FP64 repair pass.  A cell that would exceed it rejects
// the line-search trial, as the reference's 200-vertex rule does (ot_op.cuh:366-382).
constexpr int kMaxVerts = 48;
// Polygon of the full construction passes (LDS per block: kPassVerts x kCellBlock vertices, the
// limit on their occupancy).  A cell whose clipping would exceed it is handed to the repair pass,
// which rebuilds it with kMaxVerts; at the reference's own configuration (N = 160,978, 76.6 %
// zero density) cells reach ~24 vertices along an L-BFGS trajectory and intermediate polygons
// a few more.
constexpr int kPassVerts = 24;
// Threads per block of the cell kernel; one block covers kCellBlock consecutive sites of one
// problem (sites are stored in bucket order, so a block is a compact patch of the plane).
constexpr int kCellBlock = 128;
// Default blocks of the repair passes, which walk the work list of flagged sites (one per CU).
constexpr int kRepairBlocks = 256;
// Integration pass: default lanes per site (4, 8 and 16 are built) and threads per block.
constexpr int kIntLanes = 8;
constexpr int kIntBlock = 256;
// Threads per block of the vector kernels.
constexpr int kVecBlock = 256;
// Largest L-BFGS history.
constexpr int kMaxHist = 16;
// Row prefixes of the target: columns in chunks of kChunk.  The prefix at (iy, ix) is the three
// sums over the pixels k of the chunk of ix that lie left of ix, of rho_k * (1, k' + 1/2,
// k'^2 + k' + 1/3) with k' = k - (chunk start) -- the integrals of 1, U, U^2 over whole pixels in
// chunk-local coordinates; the chunk totals hold the same three sums over whole chunks.  Sums
// between two columns then need no subtraction of large prefixes.
constexpr int kChunk = 32;
// The prefixes are stored compactly: one 32-byte DensityGroup per pair of pixels of a row (FP32
// image) or per pixel (FP64 image) -- the group's densities and the prefix at its first pixel; a
// reader of the pair's second pixel adds the first pixel's terms with the fused multiply-adds that
// built the prefix, so it gets the prefix the row's sequential sum reaches there, bit for bit.
// Rows are padded to whole groups (density 0).  16 bytes per pixel of an FP32 image where a full
// prefix per pixel took 32.  (Groups of 8 FP32 pixels in 64 bytes took a quarter of the bytes but
// doubled the integration pass's time: the in-group sums and the records' registers.)
constexpr int kGroup32 = 2;
constexpr int kGroup64 = 1;
struct alignas(32) DensityGroup {
  double s0, s1, s2;  // prefix at the group's first pixel (chunk-local sums)
  union {
    float f[kGroup32];  // FP32 image
    double d;           // FP64 image
  } r;
};
static_assert(sizeof(DensityGroup) == 32, "one 32-byte record per group");
static_assert(kChunk % kGroup32 == 0, "groups never straddle chunks");
static_assert(kGroup32 == 2 && kGroup64 == 1, "readers index groups by shifts");
// The three prefix sums one pixel further: the terms of pixel k' = kl of density r.
__host__ __device__ inline void prefix_step(double& s0, double& s1, double& s2, double r,
                                            double kl) {
  s0 += r;
  s1 = fma(r, kl + 0.5, s1);
  s2 = fma(r, fma(kl, kl, kl) + 1.0 / 3.0, s2);
}

// The last step of the conformity argument: when every cell edge is matched by its neighbour's,
// the cells cover the target rectangle D a constant number of times, so sum |P_i| = |D| forces a
// tiling.  An evaluation whose cell areas sum to more than kCoverTol |D| away from |D| counts as
// unverified (FP64 vertices shared by neighbouring cells agree to rounding, so a tiling sums to
// |D| within ~1e-13 |D|; a double cover misses by |D|).
constexpr double kCoverTol = 1e-8;

// Edge labels below zero: the four sides of the target rectangle.
constexpr int kSideBottom = -1, kSideRight = -2, kSideTop = -3, kSideLeft = -4;
// Candidate sources of the cell kernel.  kCandBucket: the bucket rings 0..ring[i] around the
// site.  kCandAdj1 / kCandAdj2: the neighbours of the site in the previous evaluation's cells
// (adj), and for kCandAdj2 also their neighbours.  A site with no previous neighbours (a hidden
// cell, or no previous evaluation) uses its bucket rings, widened to 2 ring + 1 for kCandAdj2.
// After an adjacency pass ring_used[i] is the bucket ring holding all of the cell's neighbours.
enum : int { kCandBucket = 0, kCandAdj1 = 1, kCandAdj2 = 2 };
// FP32 geometry hands a cell to an FP64 rebuild (the repair pass) when its area is below
// kSliverRel bucket areas, or when a clip emptied it while a vertex lay within kMarginRel bucket
// sizes of the clipping line.
constexpr double kSliverRel = 1e-3;
constexpr double kMarginRel = 1e-3;
// Candidates the conformity check hands to a repair: when site i has an edge with site j that
// j's cell lacks, i is queued for j (at most kMaxExtra per site and round; the repair takes them
// in increasing site order, so the result does not depend on the order they were queued in
// unless more than kMaxExtra arrive).
constexpr int kMaxExtra = 24;
// Edges the conformity check reads per batch (their loads issued together).
constexpr int kVerifyBatch = 8;
// Adjacency candidates the lane construction loads per batch (their loads issued together).
constexpr int kCandBatch = 8;
// Waves of the pass that builds the sites the lane full pass deferred (a grid-stride loop over
// the list: 4 waves per SIMD of MI355X's 1,024).
constexpr int kDeferWaves = 4096;
// A lane full pass over the two-ring defers a site whose neighbours' edge counts add up to more
// than this (a typical cell: 6 neighbours of 6 edges).
constexpr int kDeferCand = 64;
// Integration lanes per site of the repair passes: few cells, often the large ones of
// low-density regions, so the latency of one cell's walk is what counts.
constexpr int kRepairLanes = 64;
// The pre-pass of the cells the last evaluation left empty (cell_reappear_kernel) runs before a
// full pass over more than kPrepassMinSites sites: on fewer, one repair round rebuilds the cells
// around those that came back for less than the pre-pass launch costs (N = 65,536 shapes
// ladder: the 16^2 to 64^2 levels 0.46 ms faster without it, the 128^2 level 0.4 ms slower).
// The rule does not depend on the pass mode, so the lane and the wave pass keep building the
// same cells from the same candidates.
constexpr int kPrepassMinSites = 4096;
__host__ __device__ inline bool use_prepass(int nblocks) {
  return nblocks * kCellBlock > kPrepassMinSites;
}
// Kernel timer (in-kernel wall clock, one record per kernel kind).  Each block that did work
// folds its start and end into one of kTimerSlots (start, end) pairs picked by its block index;
// the launch is then closed -- by the next kernel on the stream, or by the launch's own last
// block (arrival count at [2 * kTimerSlots + 2]): [2 * kTimerSlots] accumulates the ticks from
// the earliest start to the latest end, [2 * kTimerSlots + 1] counts such launches, and the
// slots are reset.
constexpr int kTimerSlots = 16;
constexpr int kTimerWords = 2 * kTimerSlots + 3;
enum : int { kTimerCell = 0, kTimerRepair = 1, kTimerInteg = 2, kTimerIntegRepair = 3, kTimers = 4 };
// Repair counters (one 64-bit word each, accumulated over every evaluation): cells a full FP32
// pass handed to the FP64 repair for a degeneracy (sliver, marginal emptying, vertex the
// refinement could not place) or for a polygon above kPassVerts vertices; sites that failed the
// conformity check after a full pass and after a repair pass; cells rebuilt by repair passes;
// cells the last evaluation left empty that came back.
enum : int {
  kDiagDegenerate = 0,
  kDiagOverflow = 1,
  kDiagCheckFull = 2,
  kDiagCheckRepair = 3,
  kDiagRebuilt = 4,
  kDiagReappeared = 5,
  kDiagWords = 8
};

// One OT problem of a batch.  All per-site arrays are indexed by global sorted site index.
struct ProbDesc {
  int site_off, n;        // sites [site_off, site_off + n) in bucket order
  int pre_off, pm, pn;    // target: pm pixels along x (fast index), pn along y; row prefixes
                          // at pre_off (pn * pm entries)
  int bkt_off, gx, gy;    // bucket grid: gx * gy buckets starting at bkt_off
  int blk_off, nblk;      // cell/vector blocks of this problem
  int tot_off, nchunk;    // chunk totals at tot_off (pn * nchunk entries)
  int grp_off, ngrp;      // density groups at grp_off (pn * ngrp records)
  int gsize;              // pixels per group: kGroup32 (FP32 image) or kGroup64
  double x0, x1, y0, y1;  // target rectangle (proj_domain)
  double dx, dy;          // pixel size
  double bx0, by0;        // bucket grid origin
  double binv;            // 1 / bucket size
  double bsize;           // bucket size
};

// Workgroups of a launch are dealt to MI355X's eight XCDs round robin (workgroup w runs on XCD
// w % 8), and each XCD has its own L2.  xcd_slot renumbers them so that XCD x takes the contiguous
// slots [x q + min(x, r), x q + min(x, r) + q + (x < r)) of the G workgroups (q = G / 8,
// r = G % 8): neighbouring cell blocks -- whose cells read each other's sites and weights and the
// same pixel rows -- then share an L2 instead of being spread over all eight.  A bijection on
// [0, G).
__host__ __device__ inline int xcd_slot(int w, int G) {
  const int x = w & 7, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + (w >> 3);
}

// Newton edge arrays (the Hessian's couplings h_ij and their neighbours' labels) in
// block-transposed order: slot k of lane ln of cell block b at (b kMaxVerts + k) kCellBlock + ln,
// so the lanes of a wave read one slot of 64 consecutive sites as one contiguous segment (the
// cells' own edge arrays keep a site's kMaxVerts slots together, which the conjugate gradients'
// lanes read 64 rows apart).
__host__ __device__ inline size_t hslot(int b, int k, int ln) {
  return ((size_t)b * kMaxVerts + k) * kCellBlock + ln;
}

// The cover check (kCoverTol): do the cells' areas add up to the target rectangle's?
__host__ __device__ inline bool covers_domain(const ProbDesc& pd, double area) {
  const double dom = (pd.x1 - pd.x0) * (pd.y1 - pd.y0);
  return fabs(area - dom) <= kCoverTol * dom;
}

// Per-block site range.
struct BlockDesc {
  int prob, first, count, pad;
};

// Coarse space of the Newton steps' preconditioner: the patches (unions of whole blocks) of a
// problem are global indices [a0, a0 + na); their na x na matrices (A_c, its inverse) are dense
// row-major at offset moff.  At most kMaxCoarse patches per problem; the coarse kernel handles
// kCoarseRows rows per block.
constexpr int kMaxCoarse = 1024;
constexpr int kCoarseRows = 16;
struct CoarseDesc {
  int a0, na, moff, pad;
};
// One coupling of a block's sites to another patch b (sum of the edge weights h); at most
// kBlockNbr per block, patch -1 ends a short list.
constexpr int kBlockNbr = 16;
struct CoarseEntry {
  int b, pad;
  double v;
};

// The whole conjugate-gradient solve of a small system in one workgroup (cot_newton.hip): at
// most kSmallBlocks blocks, one site per thread (measured on MI355X: 8 blocks, 1,024 sites, gain
// ~10% per Newton step over the multi-kernel iteration; at 32 blocks, four sites per thread on
// one compute unit lose to 32 blocks on 32 units).
constexpr int kSmallBlocks = 8;
constexpr int kSmallThreads = 1024;
struct SmallCgArgs {
  const BlockDesc* blocks;
  int nblk, sweeps, npatch, max_iter;
  const uint8_t* active;
  const double* g;
  const double* shift;
  const double* diag;
  const uint8_t* nedge;
  const int* hj;          // neighbour of each edge slot (hslot order)
  const double* h;        // h_ij of each edge slot (hslot order)
  const int* sagg;        // nullptr: no coarse term
  const int* agg_start;
  const int* agg_blk;
  const int* agg_prob;    // patch -> problem
  const CoarseDesc* cdesc;
  const double* ainv;
  double* x;              // N: the solution
  double eta2;            // stop at r.r <= eta2 r0.r0
  double* out;            // [0] iterations, [1] r0.r0, [2] r.r, [3] 1 when the residual diverged
};

// Buffers of one conjugate-gradient iteration (cot_newton.hip).  part_rz holds two slots of nblk
// r.z partials, part_rzc two slots of ntask coarse partials; ntask == 0 runs without the coarse
// term (tasks, cdesc, agg_*, rsum, ainv, cvec, sagg unused).
struct CgArgs {
  const BlockDesc* blocks;
  int nblk, sweeps, ntask, pad;
  const uint8_t* nedge;
  const int* hj;          // neighbour of each edge slot (hslot order)
  const double* h;        // h_ij of each edge slot (hslot order)
  const double* diag;
  double *x, *r, *z, *q;
  double *part_rz, *part_pq, *part_rr, *part_rzc;
  double* rz;             // two slots: r.z of the current and of the previous iteration
  const int4* tasks;
  const CoarseDesc* cdesc;
  const int* agg_start;   // patch -> its blocks (CSR over agg_blk)
  const int* agg_blk;
  const int* sagg;        // site -> global patch, -1 where the Hessian row is zero
  double* rsum;           // per-block sums of r
  const double* ainv;     // per-problem inverses of A_c (+ rank one)
  double* cvec;           // per-patch coarse correction
};

// Everything the cell kernel reads and writes.
struct CellArgs {
  const ProbDesc* probs;
  const BlockDesc* blocks;
  const double2* xy;        // site positions
  const double* w;          // base weights
  const double* dir;        // search direction (nullptr: evaluate at w)
  const double* step;       // per-problem step (used when dir != nullptr)
  const double* nu;         // source masses
  const int2* buckets;      // (start, count) per bucket
  const DensityGroup* grp;  // density groups (compact row prefixes)
  const double4* tot;       // chunk totals (x, y, z)
  uint8_t* ring;            // per-site candidate ring radius (read, raised on repair)
  int* redo;                // repair flags: 0 none, 1 listed for a repair, 2 an FP32 cell that
                            // the next check lists; a full pass clears its sites' flags, a
                            // repair construction those of the sites it rebuilds
  int redo_mode;            // 0: all sites of active problems; 1: the sites of the work list
  const int* list;          // repair pass: work list of site slots (block * kCellBlock + lane)
  const int* list_count;    //   and its length
  int* reset_count;         // construction pass: the length of the list the check that follows
                            // fills, cleared here (nullptr: none)
  const uint8_t* active;    // per-problem: evaluate only problems with active[p] != 0
  const struct DevState* st;  // device-driven solve: step, direction sign, activity per problem
  const double* g;          // gradient (steepest-descent trials)
  const int* gate;          // repair pass: skip problems whose counter is zero (nullptr: none)
  // outputs
  double* mass;
  double* fcell;
  double* area;
  double2* cen;             // rho-weighted centroid (geometric when the cell holds no mass)
  double2* geo;             // geometric centroid of the cell
  uint8_t* exist;           // cell is non-empty
  uint8_t* nedge;           // vertex count of the final cell
  int* elab;                // [site * kMaxVerts + k]: label of edge k (site or side)
  float2* espan;            // [site * kMaxVerts + k]: edge span along the canonical bisector axis
  uint8_t* ring_used;       // outermost ring that cut the cell
  int* overflow;            // per-problem count of cells that exceeded kMaxVerts
  int* coarse;              // per-problem count of FP32 cells handed to the FP64 repair (the
                            // conformity-failure counter of the check that follows)
  // FP32 geometry only: clipping frame origin (the cell's last geometric centroid)
  double2* org;
  // candidate source (kCand*) and the previous evaluation's adjacency
  int cand;
  int ring_max;
  int repair_lane;          // repair construction: 0 one wave per listed site, 1 one lane per site
  int pass_wave;            // full construction: 0 one lane per site, 1 one wave per site
  // lane full pass: sites with no previous neighbours (their previous cells were empty) would
  // scan bucket rings one candidate at a time in one lane -- the pass's longest lanes; they are
  // appended here instead (site slots) and built by a wave per site right after (nullptr: built
  // in the lane).  The integration pass clears the count.
  int* defer_list;
  int* defer_count;
  // cells the last evaluation left empty (site slots, at most empty_cap; the evaluation's last
  // save_adjacency fills the list, the integration pass clears the count): built before the full
  // pass, which then skips them, to queue those that come back with their new neighbours
  // (cell_reappear_kernel)
  int* empty_list;
  int* empty_count;
  int empty_cap;
  uint8_t* prebuilt;        // per site: built by that pre-pass (the full pass clears it)
  int prepass;              // this launch is the pre-pass
  const int* adj;           // [site * kMaxVerts + k]: edge labels of the previous cell
  const uint8_t* adjn;      // edge count of the previous cell (0: none)
  int* xcnt;                // per site: queued candidates (cleared when the site is built)
  const int* xcand;         // [site * kMaxExtra + k]
  unsigned long long* timer;  // the construction pass's timer (kTimerWords words) or nullptr
  unsigned long long* timer_int;  // the integration pass's timer, closed by the next kernel
  unsigned long long* diag;  // repair counters (kDiagWords words) or nullptr
  // construction -> integration
  double2* vtx;             // [site * kMaxVerts + k]: vertex k of the cell (absolute, FP64)
  int2* kbox;               // pixel of the cell's lower-left corner (integration origin)
};

// ---------------------------------------------------------------------------------------------
// Device-resident L-BFGS (one state per problem, updated by the last block of each kernel)
// ---------------------------------------------------------------------------------------------
enum : int { kPhaseDir = 0, kPhaseTrial = 1, kPhaseDone = 2 };

struct SolveParams {
  int m;             // history size
  int max_iter;      // iteration limit
  int stop_rule;     // 0: reference (lagged |g|_2 < tol, no empty cell), 1: max relative error,
                     // 2: max relative error and no empty cell (the hand-off to Newton steps)
  int mixed_ls;      // accept a trial on f decrease OR non-positive directional derivative
  int ring_max;
  int max_unverified;  // consecutive trials left unverified before giving up
  double tol;
  double t0;         // initial step
  double h0_thr;     // H0 = y.s / y.y only when y.y > h0_thr
};

struct DevState {
  int phase, loop_iter, restart, iters, evals, conv, status, steep, accept, hcount;
  int unverified;
  int fresh;         // the last trial's cells passed the edge check but did not tile the
                     // target (cover check): the next trial searches bucket rings instead of
                     // that trial's adjacency
  double t, f_cur, f_prev, g2_cur, zeros_cur, maxrel_cur, snap_g, snap_zero, H0;
  double f_try, g2_try, zeros_try, maxrel_try, gd_try;
  double SY[kMaxHist][kMaxHist];
  double YY[kMaxHist][kMaxHist];
  double Sg[kMaxHist], Yg[kMaxHist], rho[kMaxHist], alpha[kMaxHist];
  double coef[1 + 2 * kMaxHist];
};

struct VerifyArgs {
  const ProbDesc* probs;
  const BlockDesc* blocks;
  const double2* xy;
  const uint8_t* nedge;
  const int* elab;
  const float2* espan;
  const uint8_t* active;    // per-problem mask
  const int* gate;          // skip problems whose counter is zero (nullptr: none)
  int* redo;                // set for both cells of a mismatched edge (see CellArgs)
  int* list;                // work list the check appends newly flagged sites to (nullptr: none)
  int* list_count;
  int* fails;               // per-problem mismatch count
  double tol_rel;           // tolerance relative to the site spacing of the pair
  int* xcnt;                // candidate queues of the repair (see kMaxExtra)
  int* xcand;
  unsigned long long* diag_fail;    // counter of the sites this check fails, or nullptr
};

}  // namespace cot
