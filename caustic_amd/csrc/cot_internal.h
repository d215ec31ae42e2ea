// Internal layout shared by the HIP kernels (cot_kernels.hip) and the host driver (cot_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cot {

// Largest polygon a cell may hold while it is clipped (vertices == edges).
constexpr int kMaxVerts = 24;
// Threads per block of the cell kernel; one block covers kCellBlock consecutive sites of one
// problem (sites are stored in bucket order, so a block is a compact patch of the plane).
constexpr int kCellBlock = 128;
// Threads per block of the vector kernels.
constexpr int kVecBlock = 256;
// Largest L-BFGS history.
constexpr int kMaxHist = 16;
// Edge labels below zero: the four sides of the target rectangle.
constexpr int kSideBottom = -1, kSideRight = -2, kSideTop = -3, kSideLeft = -4;

// One OT problem of a batch.  All per-site arrays are indexed by global sorted site index.
struct ProbDesc {
  int site_off, n;        // sites [site_off, site_off + n) in bucket order
  int img_off, pm, pn;    // target image: pm pixels along x (fast index), pn along y
  int bkt_off, gx, gy;    // bucket grid: gx * gy buckets starting at bkt_off
  int blk_off, nblk;      // cell/vector blocks of this problem
  int pad0, pad1;
  double x0, x1, y0, y1;  // target rectangle (proj_domain)
  double dx, dy;          // pixel size
  double bx0, by0;        // bucket grid origin
  double binv;            // 1 / bucket size
  double bsize;           // bucket size
};

// Per-block site range.
struct BlockDesc {
  int prob, first, count, pad;
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
  const void* img;          // float or double image, per-problem offsets
  uint8_t* ring;            // per-site candidate ring radius (read, raised on repair)
  uint8_t* redo;            // repair flags; every processed site's flag is cleared
  int redo_mode;            // 0: all sites of active problems; 1: only sites with redo[i] != 0
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
  // FP32 geometry only: clipping frame origin (the cell's last geometric centroid)
  double2* org;
};

// ---------------------------------------------------------------------------------------------
// Device-resident L-BFGS (one state per problem, updated by the last block of each kernel)
// ---------------------------------------------------------------------------------------------
enum : int { kPhaseDir = 0, kPhaseTrial = 1, kPhaseDone = 2 };

struct SolveParams {
  int m;             // history size
  int max_iter;      // iteration limit
  int stop_rule;     // 0: reference (lagged |g|_2 < tol, no empty cell), 1: max relative error
  int mixed_ls;      // accept a trial on f decrease OR non-positive directional derivative
  int ring_max;
  int max_unverified;  // consecutive trials left unverified before giving up
  double tol;
  double t0;         // initial step
  double h0_thr;     // H0 = y.s / y.y only when y.y > h0_thr
};

struct DevState {
  int phase, loop_iter, restart, iters, evals, conv, status, steep, accept, hcount;
  int unverified, pad;
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
  uint8_t* redo;            // set for both cells of a mismatched edge
  int* fails;               // per-problem mismatch count
  double tol_rel;           // tolerance relative to the site spacing of the pair
};

}  // namespace cot
