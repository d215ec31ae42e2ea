// Damped Newton steps for the semi-discrete OT dual (the tail of a max-rel solve), for gfx950.
//
// The Hessian of the dual objective is the cell-adjacency matrix
//   dm_i/dw_j = -h_ij,  h_ij = int_{C_i n C_j} rho ds / (2 |s_i - s_j|)   (i != j),
//   dm_i/dw_i = sum_j h_ij,
// a weighted graph Laplacian (Kitagawa, Merigot and Thibert, "Convergence of a Newton algorithm
// for semi-discrete optimal transport", JEMS 2019).  hessian_kernel integrates the target along
// every cell edge of the last evaluation (the edge labels name the neighbour), cg_* kernels solve
// H d = -g by preconditioned conjugate gradients with H applied edge by edge (the evaluation's
// edge table is the sparse pattern), and the host driver (cot_api.cpp) damps the step so that
// every cell keeps mass and |g| decreases.
//
// Preconditioner: M = D^-1 (or block-local Jacobi sweeps in LDS) + P A_c^-1 P^T, an additive
// two-level operator.  P is the aggregation of sites into patches (unions of whole 128-site blocks,
// at most a few hundred per problem), A_c = P^T H P the patches' Laplacian: each block sums its
// couplings to the neighbouring patches here (deterministically), the host adds the blocks of a
// patch and inverts (plus a rank-one term for its constant null space).  The
// coarse term carries the smooth part of the residual across the domain in one step, which
// Jacobi alone needs O(sqrt N) iterations for: at N = 65,536 the iterations to 1e-4 fall from
// ~700 to ~90.  One iteration is three kernels -- cg_update (x, r, the Jacobi part, per-block
// sums of r), cg_coarse (patch sums, c = A_c^-1 P^T r, r.Pc) and cg_step (direction
// p = D^-1 r + P c + beta p, q = H p).
#include "cot_internal.h"

namespace cot {

namespace {

// edges of a cell read per batch by the Hessian products (their loads issued together)
constexpr int kEdgeBatch = 8;
// Smallest damped step whose remaining direction starts the next conjugate-gradient solve.
constexpr double kCarryMinStep = 0.25;

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
  __syncthreads();
  return t;
}

// Sum of n partials, the same value (same order) in every block.
__device__ __forceinline__ double sum_partials(const double* part, int n, double* red) {
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
  return block_sum(v, red);
}

// Sum of na partials a and nb partials b in one pass (all loads in flight before the one
// reduction), the same value in every block.
__device__ __forceinline__ double sum_partials2(const double* a, int na, const double* b, int nb,
                                                double* red) {
  double v = 0.0;
  for (int k = threadIdx.x; k < na; k += blockDim.x) v += a[k];
  double u = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) u += b[k];
  return block_sum(v + u, red);
}

// Sum over pieces of (parameter length) * rho(pixel) along the segment a -> b given in pixel
// units of the problem's lattice (pm x pn pixels).
__device__ double edge_rho(double ua, double va, double ub, double vb, const DensityGroup* grp,
                           int ngrp, int gsize, int pm, int pn) {
  const double du = ub - ua, dv = vb - va;
  const double iu0 = floor(ua), iv0 = floor(va);
  double tu = du > 0.0 ? (iu0 + 1.0 - ua) / du : (du < 0.0 ? (iu0 - ua) / du : 2.0);
  double tv = dv > 0.0 ? (iv0 + 1.0 - va) / dv : (dv < 0.0 ? (iv0 - va) / dv : 2.0);
  const double su = du != 0.0 ? 1.0 / fabs(du) : 0.0, sv = dv != 0.0 ? 1.0 / fabs(dv) : 0.0;
  int guard = (int)(fabs(floor(ub) - iu0) + fabs(floor(vb) - iv0)) + 3;
  double t = 0.0, acc = 0.0;
  while (t < 1.0 && guard-- > 0) {
    const double tn = fmin(fmin(tu, tv), 1.0);
    const double tm = 0.5 * (t + tn);
    const int ix = min(max((int)floor(ua + tm * du), 0), pm - 1);
    const int iy = min(max((int)floor(va + tm * dv), 0), pn - 1);
    const DensityGroup& rec = grp[(size_t)iy * ngrp + (gsize == kGroup32 ? ix >> 1 : ix)];
    const double r = gsize == kGroup32 ? (double)rec.r.f[ix & 1] : rec.r.d;
    acc += (tn - t) * r;
    t = tn;
    if (tu <= tv) {
      tu += su;
    } else {
      tv += sv;
    }
  }
  return acc;
}

// Preconditioner of the conjugate gradients, applied by a block of kCellBlock threads to the
// residual r_i of its sites (one per thread): `sweeps` Jacobi sweeps on the block's principal
// submatrix D - B of H (B: the couplings between two sites of the block),
//   z_0 = D^-1 r,  z_{k+1} = D^-1 (r + B z_k),
// i.e. the Neumann polynomial sum_{k<=sweeps} (D^-1 B)^k D^-1 r -- a fixed symmetric positive
// definite operator (D^-1/2 B D^-1/2 has spectral radius below 1: H is diagonally dominant and a
// block has neighbours outside it).  sweeps = 0 is Jacobi.  Every thread of the block must call
// this (it synchronises).
__device__ double precondition(int b, const BlockDesc& bd, int i, double ri, double dg,
                               const uint8_t* nedge, const int* hj, const double* h,
                               int sweeps, double* zl) {
  const int ln = threadIdx.x;
  const bool own = ln < bd.count && dg > 0.0;
  double z = own ? ri / dg : 0.0;
  if (sweeps <= 0) return z;
  const int nv = own ? nedge[i] : 0;
  // the first kEdgeBatch edges' in-block neighbours and couplings, loaded together once for all
  // sweeps (a cell has 6 edges on average; the rest are read in the sweep)
  int jl[kEdgeBatch];
  double hl[kEdgeBatch];
#pragma unroll
  for (int u = 0; u < kEdgeBatch; ++u) {
    jl[u] = u < nv ? hj[hslot(b, u, ln)] - bd.first : -1;
    hl[u] = u < nv ? h[hslot(b, u, ln)] : 0.0;
  }
  for (int s = 0; s < sweeps; ++s) {
    zl[ln] = z;
    __syncthreads();
    if (own) {
      double acc = ri;
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u)
        if (jl[u] >= 0 && jl[u] < bd.count) acc += hl[u] * zl[jl[u]];
      for (int k = kEdgeBatch; k < nv; ++k) {
        const int j = hj[hslot(b, k, ln)] - bd.first;
        if (j >= 0 && j < bd.count) acc += h[hslot(b, k, ln)] * zl[j];
      }
      z = acc / dg;
    }
    __syncthreads();
  }
  return z;
}

}  // namespace

// h[hslot(block, k, lane)] = h_ij of edge k of the site's cell (0 for sides), hj the same slot's
// neighbour (the edge label), diag[site] = sum_k h.  Sites of inactive problems and empty cells
// get zeros.  Slots beyond a cell's edge count are not written (readers stop at nedge).
__global__ __launch_bounds__(256) void hessian_kernel(const ProbDesc* probs,
                                                      const BlockDesc* blocks,
                                                      const uint8_t* active, const double2* xy,
                                                      const double2* vtx, const uint8_t* nedge,
                                                      const int* elab,
                                                      const DensityGroup* grp,
                                                      const int* bagg, double* h, int* hj,
                                                      double* diag, int* sagg) {
  const int blk = xcd_slot(blockIdx.x, gridDim.x);
  const BlockDesc bd = blocks[blk];
  const ProbDesc pd = probs[bd.prob];
  const bool on = active[bd.prob] != 0;
  const double inv_pix = 1.0 / (pd.dx * pd.dy);
  // slot-major: consecutive threads take consecutive lanes of one slot (coalesced stores)
  for (int e = threadIdx.x; e < kMaxVerts * kCellBlock; e += blockDim.x) {
    const int k = e / kCellBlock, ln = e % kCellBlock;
    if (ln >= bd.count) continue;
    const int i = bd.first + ln;
    const int nv = nedge[i];
    if (k >= nv) continue;  // readers stop at the cell's edge count
    double hv = 0.0;
    const int j = elab[(size_t)i * kMaxVerts + k];
    if (on && nv >= 3 && j >= 0) {
      const double2 a = vtx[(size_t)i * kMaxVerts + k];
      const double2 b = vtx[(size_t)i * kMaxVerts + (k + 1 < nv ? k + 1 : 0)];
      const double len = sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
      if (len > 0.0) {
        const double r = edge_rho((a.x - pd.x0) / pd.dx, (a.y - pd.y0) / pd.dy,
                                  (b.x - pd.x0) / pd.dx, (b.y - pd.y0) / pd.dy,
                                  grp + pd.grp_off, pd.ngrp, pd.gsize, pd.pm, pd.pn);
        const double2 si = xy[i], sj = xy[j];
        const double d = sqrt((si.x - sj.x) * (si.x - sj.x) + (si.y - sj.y) * (si.y - sj.y));
        hv = r * len * inv_pix / (2.0 * d);
      }
    }
    h[hslot(blk, k, ln)] = hv;
    hj[hslot(blk, k, ln)] = j;
  }
  __syncthreads();
  if ((int)threadIdx.x < bd.count) {
    const int ln = threadIdx.x, i = bd.first + ln;
    double s = 0.0;
    const int nv = nedge[i];
    for (int k = 0; k < nv; ++k) s += h[hslot(blk, k, ln)];
    diag[i] = s;
    if (sagg) sagg[i] = s > 0.0 ? bagg[blk] : -1;
  }
}

// Couplings of each block's sites to the other patches, the block's row of A_c = P^T H P before
// the sum over the patch's blocks: out[blk * kBlockNbr + k] = (patch b, sum of h over the edges
// from the block's sites to sites of patch b), for the block's distinct neighbouring patches in
// increasing order (patch -1 ends a short list).  Deterministic: the patches are collected,
// sorted, then each is summed with one block reduction in a fixed order.  A block with more than
// kBlockNbr neighbouring patches keeps the kBlockNbr smallest -- the couplings dropped leave a
// weaker but still valid (Laplacian, so SPD after the rank-one term) coarse matrix.
__global__ __launch_bounds__(kCellBlock) void coarse_rows_kernel(
    const BlockDesc* blocks, const int* bagg, const int* sagg, const uint8_t* nedge,
    const int* hj, const double* h, CoarseEntry* out) {
  __shared__ int ids[kBlockNbr];
  __shared__ int wmin[kCellBlock / 64];
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  const int a = bagg[blockIdx.x];
  CoarseEntry* row = out + (size_t)blockIdx.x * kBlockNbr;
  if (a < 0) {
    if (threadIdx.x < kBlockNbr) row[threadIdx.x] = CoarseEntry{-1, 0, 0.0};
    return;
  }
  const bool own = (int)threadIdx.x < bd.count;
  const int i = bd.first + threadIdx.x;
  const int nv = own ? nedge[i] : 0;
  const bool live = own && sagg[i] >= 0;
  // round k takes the smallest neighbouring patch above the one taken in round k - 1 (a block
  // minimum), so the list comes out sorted and free of repeats; n is the same in every thread
  int n = 0, last = -1;
  for (int k = 0; k < kBlockNbr; ++k) {
    int m = INT32_MAX;
    if (live)
      for (int e = 0; e < nv; ++e) {
        const int j = hj[hslot(blockIdx.x, e, threadIdx.x)];
        if (j < 0) continue;
        const int b = sagg[j];
        if (b >= 0 && b != a && b > last && b < m) m = b;
      }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    int bm = INT32_MAX;
    for (int w = 0; w < kCellBlock / 64; ++w) bm = min(bm, wmin[w]);
    __syncthreads();
    if (bm == INT32_MAX) break;
    if (threadIdx.x == 0) ids[k] = bm;
    n = k + 1;
    last = bm;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const int b = ids[k];
    double v = 0.0;
    if (live)
      for (int e = 0; e < nv; ++e) {
        const int j = hj[hslot(blockIdx.x, e, threadIdx.x)];
        if (j >= 0 && sagg[j] == b) v += h[hslot(blockIdx.x, e, threadIdx.x)];
      }
    v = block_sum(v, red);
    if (threadIdx.x == 0) row[k] = CoarseEntry{b, 0, v};
  }
  if ((int)threadIdx.x >= n && threadIdx.x < kBlockNbr) row[threadIdx.x] = CoarseEntry{-1, 0, 0.0};
}

// Coarse part of the preconditioner for kCoarseRows patches of one problem (a task): R = P^T r
// (the patch sums of the per-block sums rsum, in the patch's block-list order, the same in every
// task), c_a = sum_b Ainv[a][b] R_b for the task's rows, partial r.Pc = sum_a R_a c_a.
__global__ __launch_bounds__(256) void cg_coarse_kernel(const int4* tasks, const CoarseDesc* cdesc,
                                                        const int* agg_start, const int* agg_blk,
                                                        const double* rsum, const double* ainv,
                                                        double* cvec, double* part_rzc) {
  __shared__ double R[kMaxCoarse];
  __shared__ double red[4];
  const int4 tk = tasks[blockIdx.x];  // (problem, first row, rows, -)
  const CoarseDesc cd = cdesc[tk.x];
  for (int t = threadIdx.x; t < cd.na; t += blockDim.x) {
    double v = 0.0;
    for (int k = agg_start[cd.a0 + t]; k < agg_start[cd.a0 + t + 1]; ++k) v += rsum[agg_blk[k]];
    R[t] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0.0;
  for (int row = wave; row < tk.z; row += 4) {
    const int a = tk.y + row;
    const double* arow = ainv + (size_t)cd.moff + (size_t)a * cd.na;
    double v = 0.0;
    for (int b = lane; b < cd.na; b += 64) v += arow[b] * R[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) {
      cvec[cd.a0 + a] = v;
      acc += R[a] * v;
    }
  }
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part_rzc[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// CG start: x = 0, r = -(g - shift[prob]) on the active problems' non-empty cells (0 elsewhere;
// shift = the problem's mean of g, which puts the right-hand side in the range of the singular
// H), z = r / diag, partial r.z and r.r.
// x0 (nullptr: zero) starts the iteration: x = x0, r = b - H x0 (x0 is read, x written: two
// arrays).  part_bb: the partials of |b|^2, the reference of the stopping rule.
__global__ __launch_bounds__(kCellBlock) void cg_init_kernel(
    const BlockDesc* blocks, const uint8_t* active, const double* g, const double* shift,
    const double* diag, const uint8_t* nedge, const int* hj, const double* h, int sweeps,
    const double* x0, double* x, double* r, double* z, double* part_rz, double* part_rr,
    double* part_bb, double* rsum) {
  __shared__ double red[kCellBlock / 64];
  __shared__ double zl[kCellBlock];
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  const int i = bd.first + ln;
  double ri = 0.0, dg = 0.0, bi = 0.0, xi = 0.0;
  if (ln < bd.count) {
    dg = diag[i];
    if (x0) xi = x0[i];
    if (active[bd.prob] && dg > 0.0) {
      bi = shift[bd.prob] - g[i];
      ri = bi;
      if (x0) {
        double s = dg * xi;
        const int nv = nedge[i];
        for (int k0 = 0; k0 < nv; k0 += kEdgeBatch) {
          int jj[kEdgeBatch];
          double hh[kEdgeBatch], xx[kEdgeBatch];
#pragma unroll
          for (int u = 0; u < kEdgeBatch; ++u) {
            const bool in = k0 + u < nv;
            jj[u] = in ? hj[hslot(blockIdx.x, k0 + u, ln)] : -1;
            hh[u] = in ? h[hslot(blockIdx.x, k0 + u, ln)] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < kEdgeBatch; ++u) xx[u] = jj[u] >= 0 ? x0[jj[u]] : 0.0;
#pragma unroll
          for (int u = 0; u < kEdgeBatch; ++u)
            if (jj[u] >= 0) s -= hh[u] * xx[u];
        }
        ri = bi - s;
      }
    }
  }
  const double zi = precondition(blockIdx.x, bd, i, ri, dg, nedge, hj, h, sweeps, zl);
  double rz = 0.0, rr = 0.0, bb = 0.0;
  if (ln < bd.count) {
    x[i] = xi;
    r[i] = ri;
    z[i] = zi;
    rz = ri * zi;
    rr = ri * ri;
    bb = bi * bi;
  }
  rz = block_sum(rz, red);
  rr = block_sum(rr, red);
  bb = block_sum(bb, red);
  const double rs = rsum ? block_sum(ri, red) : 0.0;
  if (ln == 0) {
    part_rz[blockIdx.x] = rz;
    part_rr[blockIdx.x] = rr;
    part_bb[blockIdx.x] = bb;
    if (rsum) rsum[blockIdx.x] = rs;
  }
}

// One CG direction and product: r.z from the partials of the update (and coarse) kernels,
// stored by block 0 in rz_out for the update kernel and the next iteration; beta = r.z / rz_prev
// (0 on the first iteration), p_new = z + beta p_old for the block's sites, q = H p_new with the
// neighbours' p_new formed on the fly from z and p_old (which no block writes here), partial p.q.
// With the coarse term (sagg != nullptr) the preconditioned residual is z_i + cvec[sagg_i]
// (sagg_i < 0: no patch) and r.z adds the coarse partials.
__global__ __launch_bounds__(kCellBlock) void cg_step_kernel(
    const BlockDesc* blocks, int nblk, const double* part_rz, const double* part_rzc, int ntask,
    const double* rz_prev, double* rz_out, int first, const uint8_t* nedge, const int* hj,
    const double* h, const double* diag, const int* sagg, const double* cvec, const double* z,
    const double* p_old, double* p_new, double* q, double* part_pq) {
  __shared__ double red[kCellBlock / 64];
  const double rz = sum_partials2(part_rz, nblk, part_rzc, ntask, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) *rz_out = rz;
  double beta = 0.0;
  if (!first) {
    const double rz0 = *rz_prev;
    beta = rz0 != 0.0 ? rz / rz0 : 0.0;
  }
  auto zc = [&](int j) {
    double v = z[j];
    if (sagg) {
      const int a = sagg[j];
      if (a >= 0) v += cvec[a];
    }
    return v;
  };
  const BlockDesc bd = blocks[blockIdx.x];
  double pq = 0.0;
  if ((int)threadIdx.x < bd.count) {
    const int i = bd.first + threadIdx.x;
    const double pi = first ? zc(i) : zc(i) + beta * p_old[i];
    p_new[i] = pi;
    double s = diag[i] * pi;
    const int nv = nedge[i];
    // kEdgeBatch edges at a time: their labels and couplings, then the neighbours' values, then
    // their patch terms, each round's loads in flight together; the terms are added in edge
    // order as one edge at a time would add them
    for (int k0 = 0; k0 < nv; k0 += kEdgeBatch) {
      int jj[kEdgeBatch];
      double hh[kEdgeBatch];
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u) {
        const bool in = k0 + u < nv;
        jj[u] = in ? hj[hslot(blockIdx.x, k0 + u, threadIdx.x)] : -1;
        hh[u] = in ? h[hslot(blockIdx.x, k0 + u, threadIdx.x)] : 0.0;
      }
      double zz[kEdgeBatch], pp[kEdgeBatch];
      int aa[kEdgeBatch];
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u) {
        const int j = jj[u];
        zz[u] = j >= 0 ? z[j] : 0.0;
        pp[u] = (j >= 0 && !first) ? p_old[j] : 0.0;
        aa[u] = (j >= 0 && sagg) ? sagg[j] : -1;
      }
      double cc[kEdgeBatch];
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u) cc[u] = aa[u] >= 0 ? cvec[aa[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u) {
        if (jj[u] >= 0) {
          double v = zz[u];
          if (aa[u] >= 0) v += cc[u];
          const double pj = first ? v : v + beta * pp[u];
          s -= hh[u] * pj;
        }
      }
    }
    q[i] = s;
    pq = pi * s;
  }
  pq = block_sum(pq, red);
  if (threadIdx.x == 0) part_pq[blockIdx.x] = pq;
}

// q = H p (edge by edge), partial p.q.
__global__ __launch_bounds__(kCellBlock) void cg_apply_kernel(
    const BlockDesc* blocks, const uint8_t* nedge, const int* hj, const double* h,
    const double* diag, const double* p, double* q, double* part_pq) {
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  double pq = 0.0;
  if ((int)threadIdx.x < bd.count) {
    const int i = bd.first + threadIdx.x;
    const double pi = p[i];
    double s = diag[i] * pi;
    const int nv = nedge[i];
    for (int k0 = 0; k0 < nv; k0 += kEdgeBatch) {
      int jj[kEdgeBatch];
      double hh[kEdgeBatch], pp[kEdgeBatch];
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u) {
        const bool in = k0 + u < nv;
        jj[u] = in ? hj[hslot(blockIdx.x, k0 + u, threadIdx.x)] : -1;
        hh[u] = in ? h[hslot(blockIdx.x, k0 + u, threadIdx.x)] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u) pp[u] = jj[u] >= 0 ? p[jj[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < kEdgeBatch; ++u)
        if (jj[u] >= 0) s -= hh[u] * pp[u];
    }
    q[i] = s;
    pq = pi * s;
  }
  pq = block_sum(pq, red);
  if (threadIdx.x == 0) part_pq[blockIdx.x] = pq;
}

// alpha = (r.z) / (p.q) with r.z as the direction kernel stored it; x += alpha p; r -= alpha q;
// z = r / diag; partial r.z (new), r.r; with the coarse term rsum receives the block sums of r.
__global__ __launch_bounds__(kCellBlock) void cg_update_kernel(
    const BlockDesc* blocks, int nblk, const double* rz_cur, const double* part_pq,
    const double* diag, const uint8_t* nedge, const int* hj, const double* h, int sweeps,
    const double* p, const double* q, double* x, double* r, double* z, double* part_rz_new,
    double* part_rr, double* rsum) {
  __shared__ double red[kCellBlock / 64];
  __shared__ double zl[kCellBlock];
  const double rz = *rz_cur;
  const double pq = sum_partials(part_pq, nblk, red);
  const double alpha = pq > 0.0 ? rz / pq : 0.0;
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  const int i = bd.first + ln;
  double ri = 0.0, dg = 0.0;
  if (ln < bd.count) {
    x[i] += alpha * p[i];
    ri = r[i] - alpha * q[i];
    dg = diag[i];
  }
  const double zi = precondition(blockIdx.x, bd, i, ri, dg, nedge, hj, h, sweeps, zl);
  double rzn = 0.0, rr = 0.0;
  if (ln < bd.count) {
    r[i] = ri;
    z[i] = zi;
    rzn = ri * zi;
    rr = ri * ri;
  }
  rzn = block_sum(rzn, red);
  rr = block_sum(rr, red);
  const double rs = rsum ? block_sum(ri, red) : 0.0;
  if (threadIdx.x == 0) {
    part_rz_new[blockIdx.x] = rzn;
    part_rr[blockIdx.x] = rr;
    if (rsum) rsum[blockIdx.x] = rs;
  }
}

// The conjugate-gradient solve of H x = shift - g (the right-hand side of cg_init_kernel) for a
// small system, start to finish in one workgroup: the same preconditioner as the multi-kernel
// iteration (Jacobi, `sweeps` block-local sweeps, the coarse patch term), the same stopping rule,
// workgroup barriers in place of kernel boundaries and no host round trip per chunk.  Thread t
// serves lane t % 128 of blocks t / 128, t / 128 + 8, ... (at most kSmallBlocks blocks); the
// residual, the solution and q stay in registers, p and z go through LDS for the neighbours (the
// system's sites are 0 .. N - 1, N <= kSmallBlocks * 128).
// Every branch that contains a barrier depends only on workgroup-wide values.
__device__ __forceinline__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < kSmallThreads / 64; ++k) t += red[k];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(kSmallThreads) void cg_small_kernel(SmallCgArgs a) {
  constexpr int kRounds = kSmallBlocks / (kSmallThreads / kCellBlock);
  constexpr int kSites = kSmallBlocks * kCellBlock;  // sites 0 .. N - 1 with N <= kSites
  __shared__ double red[kSmallThreads / 64];
  __shared__ double wsum[2 * kSmallBlocks];
  __shared__ double Rp[kSmallBlocks];
  __shared__ double cp[kSmallBlocks];
  // z and p of every site, for the neighbours; each site's owner writes both before any thread
  // reads them (apply_m, then the p update, each followed by a barrier)
  __shared__ double zs[kSites];
  __shared__ double ps[kSites];
  const int t = threadIdx.x, l = t & (kCellBlock - 1), b0 = t / kCellBlock;
  const int half = (t >> 6) & 1, lane = t & 63;
  // per round: the site (-1: none), its block's size, diagonal, residual, solution, p and z
  int site[kRounds], bcount[kRounds];
  double dg[kRounds], r[kRounds], x[kRounds], pl[kRounds], zl[kRounds];
#pragma unroll
  for (int m = 0; m < kRounds; ++m) {
    site[m] = -1;
    bcount[m] = 0;
    dg[m] = r[m] = x[m] = pl[m] = zl[m] = 0.0;
    const int b = b0 + m * (kSmallThreads / kCellBlock);
    if (b < a.nblk) {
      const BlockDesc bd = a.blocks[b];
      bcount[m] = bd.count;
      if (l < bd.count) {
        const int i = bd.first + l;
        site[m] = i;
        dg[m] = a.diag[i];
        if (a.active[bd.prob] && dg[m] > 0.0) r[m] = a.shift[bd.prob] - a.g[i];
      }
    }
  }
  // z = M r, returns r.z
  auto apply_m = [&]() -> double {
#pragma unroll
    for (int m = 0; m < kRounds; ++m) {
      zl[m] = dg[m] > 0.0 ? r[m] / dg[m] : 0.0;
      if (site[m] >= 0) zs[site[m]] = zl[m];
    }
    for (int sw = 0; sw < a.sweeps; ++sw) {
      __syncthreads();
      double zn[kRounds];
#pragma unroll
      for (int m = 0; m < kRounds; ++m) {
        zn[m] = zl[m];
        const int i = site[m];
        if (i >= 0 && dg[m] > 0.0) {
          double acc = r[m];
          const int nv = a.nedge[i];
          const int bf = i - l;  // the block's first site
          const int b = b0 + m * (kSmallThreads / kCellBlock);
          for (int k = 0; k < nv; ++k) {
            const int j = a.hj[hslot(b, k, l)] - bf;
            if (j >= 0 && j < bcount[m]) acc += a.h[hslot(b, k, l)] * zs[bf + j];
          }
          zn[m] = acc / dg[m];
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kRounds; ++m) {
        zl[m] = zn[m];
        if (site[m] >= 0) zs[site[m]] = zl[m];
      }
    }
    if (a.sagg) {
      // patch sums of r: per block the two waves' sums, then the patch's blocks in order
#pragma unroll
      for (int m = 0; m < kRounds; ++m) {
        const int b = b0 + m * (kSmallThreads / kCellBlock);
        if (b < a.nblk) {  // uniform across the wave
          double v = site[m] >= 0 ? r[m] : 0.0;
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
          if (lane == 0) wsum[2 * b + half] = v;
        }
      }
      __syncthreads();
      for (int q = t; q < a.npatch; q += kSmallThreads) {
        double v = 0.0;
        for (int k = a.agg_start[q]; k < a.agg_start[q + 1]; ++k) {
          const int b = a.agg_blk[k];
          v += wsum[2 * b] + wsum[2 * b + 1];
        }
        Rp[q] = v;
      }
      __syncthreads();
      for (int q = t; q < a.npatch; q += kSmallThreads) {
        const CoarseDesc cd = a.cdesc[a.agg_prob[q]];
        const double* row = a.ainv + (size_t)cd.moff + (size_t)(q - cd.a0) * cd.na;
        double v = 0.0;
        for (int k = 0; k < cd.na; ++k) v += row[k] * Rp[cd.a0 + k];
        cp[q] = v;
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kRounds; ++m)
        if (site[m] >= 0) {
          const int q = a.sagg[site[m]];
          if (q >= 0) zl[m] += cp[q];
        }
    }
    double rz = 0.0;
#pragma unroll
    for (int m = 0; m < kRounds; ++m) rz += r[m] * zl[m];
    return wg_sum(rz, red);
  };

  double rr0 = 0.0;
#pragma unroll
  for (int m = 0; m < kRounds; ++m) rr0 += r[m] * r[m];
  rr0 = wg_sum(rr0, red);
  double rz = apply_m();
#pragma unroll
  for (int m = 0; m < kRounds; ++m) {
    pl[m] = zl[m];
    if (site[m] >= 0) ps[site[m]] = pl[m];
  }
  int it = 0;
  double rr = rr0;
  int diverged = 0;
  while (it < a.max_iter) {
    __syncthreads();  // p of every site written
    double q[kRounds], pq = 0.0;
#pragma unroll
    for (int m = 0; m < kRounds; ++m) {
      q[m] = 0.0;
      const int i = site[m];
      if (i < 0) continue;
      double sacc = dg[m] * pl[m];
      const int nv = a.nedge[i];
      const int b = b0 + m * (kSmallThreads / kCellBlock);
      for (int k = 0; k < nv; ++k) {
        const int j = a.hj[hslot(b, k, l)];
        if (j >= 0) sacc -= a.h[hslot(b, k, l)] * ps[j];
      }
      q[m] = sacc;
      pq += pl[m] * sacc;
    }
    pq = wg_sum(pq, red);
    const double alpha = pq > 0.0 ? rz / pq : 0.0;
    double rrl = 0.0;
#pragma unroll
    for (int m = 0; m < kRounds; ++m) {
      x[m] += alpha * pl[m];
      r[m] -= alpha * q[m];
      rrl += r[m] * r[m];
    }
    rr = wg_sum(rrl, red);
    ++it;
    if (!(rr <= 1e8 * rr0)) {
      diverged = 1;
      break;
    }
    if (!(rr > a.eta2 * rr0)) break;
    const double rzn = apply_m();
    const double beta = rz != 0.0 ? rzn / rz : 0.0;
    rz = rzn;
    __syncthreads();  // every thread's reads of p (H p above) are done
#pragma unroll
    for (int m = 0; m < kRounds; ++m) {
      pl[m] = zl[m] + beta * pl[m];
      if (site[m] >= 0) ps[site[m]] = pl[m];
    }
  }
#pragma unroll
  for (int m = 0; m < kRounds; ++m)
    if (site[m] >= 0) a.x[site[m]] = x[m];
  if (t == 0) {
    a.out[0] = it;
    a.out[1] = rr0;
    a.out[2] = rr;
    a.out[3] = diverged;
  }
}

// w += step[prob] * d for the problems with active[prob] != 0.
// w += t d for the problems that took a step (active); carry (nullptr: none) receives the part
// of the direction the step left, (1 - t) d: after a damped step the gradient has fallen by about
// the factor 1 - t, and so has the next Newton direction, which carry then approximates as the
// start of the next conjugate-gradient solve.  Zero where no step was taken, and after a step
// below kCarryMinStep: a direction the damping cut that short was a poor model of the next one.
__global__ __launch_bounds__(kCellBlock) void axpy_step_kernel(const BlockDesc* blocks,
                                                               const uint8_t* active,
                                                               const double* step,
                                                               const double* d, double* w,
                                                               double* carry) {
  const BlockDesc bd = blocks[blockIdx.x];
  if ((int)threadIdx.x >= bd.count) return;
  const int i = bd.first + threadIdx.x;
  if (!active[bd.prob]) {
    if (carry) carry[i] = 0.0;
    return;
  }
  const double t = step[bd.prob];
  w[i] += t * d[i];
  if (carry) carry[i] = t >= kCarryMinStep ? (1.0 - t) * d[i] : 0.0;
}

hipError_t launch_hessian(const ProbDesc* probs, const BlockDesc* blocks, const uint8_t* active,
                          const double2* xy, const double2* vtx, const uint8_t* nedge,
                          const int* elab, const DensityGroup* grp, const int* bagg, double* h,
                          int* hj, double* diag, int* sagg, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(hessian_kernel, dim3(nblk), dim3(256), 0, st, probs, blocks, active, xy, vtx,
                     nedge, elab, grp, bagg, h, hj, diag, sagg);
  return hipGetLastError();
}

hipError_t launch_coarse_rows(const BlockDesc* blocks, const int* bagg, const int* sagg,
                              const uint8_t* nedge, const int* hj, const double* h,
                              CoarseEntry* out, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(coarse_rows_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, bagg, sagg,
                     nedge, hj, h, out);
  return hipGetLastError();
}

hipError_t launch_cg_coarse(const int4* tasks, int ntask, const CoarseDesc* cdesc,
                            const int* agg_start, const int* agg_blk, const double* rsum,
                            const double* ainv, double* cvec, double* part_rzc, hipStream_t st) {
  hipLaunchKernelGGL(cg_coarse_kernel, dim3(ntask), dim3(256), 0, st, tasks, cdesc, agg_start,
                     agg_blk, rsum, ainv, cvec, part_rzc);
  return hipGetLastError();
}

hipError_t launch_cg_init(const BlockDesc* blocks, const uint8_t* active, const double* g,
                          const double* shift, const double* diag, const uint8_t* nedge,
                          const int* hj, const double* h, int sweeps, const double* x0, double* x,
                          double* r, double* z, double* part_rz, double* part_rr, double* part_bb,
                          double* rsum, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(cg_init_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, active, g, shift,
                     diag, nedge, hj, h, sweeps, x0, x, r, z, part_rz, part_rr, part_bb, rsum);
  return hipGetLastError();
}

// One CG iteration: direction and H p (cg_step), the update, and with the coarse term
// (cg->ntask > 0) the coarse kernel.  The r.z partials of the current residual are in slot `cur`
// of part_rz / part_rzc, and the update and the coarse kernel write the new ones into the other
// slot; the direction kernel stores the summed r.z in rz[cur] (the previous iteration's is in
// rz[other]).  p_old is the previous direction, p_new receives this one.
hipError_t launch_cg_iteration(const CgArgs& a, const double* p_old, double* p_new, int cur,
                               int first, hipStream_t st) {
  const int oth = cur ^ 1;
  const double* rzc_cur = a.part_rzc + (size_t)cur * a.ntask;
  double* rzc_oth = a.part_rzc + (size_t)oth * a.ntask;
  const bool coarse = a.ntask > 0;
  hipLaunchKernelGGL(cg_step_kernel, dim3(a.nblk), dim3(kCellBlock), 0, st, a.blocks, a.nblk,
                     a.part_rz + (size_t)cur * a.nblk, rzc_cur, a.ntask, a.rz + oth, a.rz + cur,
                     first, a.nedge, a.hj, a.h, a.diag, coarse ? a.sagg : nullptr, a.cvec, a.z,
                     p_old, p_new, a.q, a.part_pq);
  hipLaunchKernelGGL(cg_update_kernel, dim3(a.nblk), dim3(kCellBlock), 0, st, a.blocks, a.nblk,
                     a.rz + cur, a.part_pq, a.diag, a.nedge, a.hj, a.h, a.sweeps, p_new, a.q,
                     a.x, a.r, a.z, a.part_rz + (size_t)oth * a.nblk, a.part_rr,
                     coarse ? a.rsum : nullptr);
  if (coarse)
    hipLaunchKernelGGL(cg_coarse_kernel, dim3(a.ntask), dim3(256), 0, st, a.tasks, a.cdesc,
                       a.agg_start, a.agg_blk, a.rsum, a.ainv, a.cvec, rzc_oth);
  return hipGetLastError();
}

hipError_t launch_cg_small(const SmallCgArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(cg_small_kernel, dim3(1), dim3(kSmallThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_hessian_apply(const BlockDesc* blocks, const uint8_t* nedge, const int* hj,
                                const double* h, const double* diag, const double* p, double* q,
                                double* part_pq, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(cg_apply_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, nedge, hj, h,
                     diag, p, q, part_pq);
  return hipGetLastError();
}

hipError_t launch_axpy_step(const BlockDesc* blocks, const uint8_t* active, const double* step,
                            const double* d, double* w, double* carry, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(axpy_step_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, active, step,
                     d, w, carry);
  return hipGetLastError();
}

}  // namespace cot
