// Host-side layout of a problem: bucket grid, tile-ordered site permutation, block table.
// Shared by the library (cot_api.cpp) and the host build of the kernels used in tests.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/caustic_ot.h"
#include "cot_internal.h"

namespace cot {

// Bucket grid and tile-ordered permutation of one problem's sites.
inline void bucket_sort(const cot_problem& pr, double spb, ProbDesc& pd, std::vector<int>& order,
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
  const int tiles_x = (gx + 7) / 8, tiles_y = (gy + 7) / 8;
  std::vector<int> key(n);
  std::vector<int> cell(n);
  for (int i = 0; i < n; ++i) {
    int cx = std::min(std::max((int)std::floor((pr.sites[2 * i] - xmin) * pd.binv), 0), gx - 1);
    int cy =
        std::min(std::max((int)std::floor((pr.sites[2 * i + 1] - ymin) * pd.binv), 0), gy - 1);
    cell[i] = cy * gx + cx;
    const int tile = (cy / 8) * tiles_x + (cx / 8);
    key[i] = tile * 64 + (cy % 8) * 8 + (cx % 8);
  }
  // sites in increasing key, ties in index order (a stable counting sort: keys are below
  // 64 x tiles, a few times n)
  const size_t nkey = (size_t)tiles_x * tiles_y * 64;
  std::vector<int> start(nkey + 1, 0);
  for (int i = 0; i < n; ++i) ++start[(size_t)key[i] + 1];
  for (size_t k = 0; k < nkey; ++k) start[k + 1] += start[k];
  order.resize(n);
  for (int i = 0; i < n; ++i) order[start[key[i]]++] = i;
  buckets.assign((size_t)gx * gy, make_int2(0, 0));
  for (int r = 0; r < n; ++r) {
    int b = cell[order[r]];
    if (buckets[b].y == 0) buckets[b].x = r;
    buckets[b].y += 1;
  }
}

// Density groups and chunk totals of one target image (layout in cot_internal.h).  FP32
// geometry reads the image rounded to float, as its pieces use it.
inline void build_prefix(const double* rho, int pm, int pn, bool fp32, DensityGroup* grp,
                         double4* tot) {
  const int nchunk = (pm + kChunk - 1) / kChunk;
  const int gsize = fp32 ? kGroup32 : kGroup64;
  const int ngrp = (pm + gsize - 1) / gsize;
  for (int iy = 0; iy < pn; ++iy) {
    for (int c = 0; c < nchunk; ++c) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      const int k0 = c * kChunk, k1 = std::min(k0 + kChunk, pm);
      for (int g0 = k0; g0 < k1; g0 += gsize) {
        DensityGroup& rec = grp[(size_t)iy * ngrp + g0 / gsize];
        rec.s0 = s0;
        rec.s1 = s1;
        rec.s2 = s2;
        for (int u = 0; u < gsize; ++u) {
          const int k = g0 + u;
          double r = k < pm ? rho[(size_t)iy * pm + k] : 0.0;
          if (fp32) {
            r = (double)(float)r;
            rec.r.f[u] = (float)r;
          } else {
            rec.r.d = r;
          }
          if (k < k1) prefix_step(s0, s1, s2, r, (double)(k - k0));
        }
      }
      tot[(size_t)iy * nchunk + c] = make_double4(s0, s1, s2, 0.0);
    }
  }
}

// In place: a (n x n, row-major) holds the off-diagonal entries of the patches' Laplacian as the
// device accumulated them (each side of an edge adds to its own row, so a[i][j] and a[j][i] differ
// by rounding); on return it holds the pseudo-inverse of A_c, A_c symmetrised with its diagonal
// set to minus the off-diagonal row sums.  A narrow band (bandwidth <= n / 4, the usual case)
// takes the grounded banded route below (~1.5 bw n^2 multiply-adds); otherwise
// J (A_c + s 11^T)^-1 J with J = I - 11^T / n by a dense Cholesky factor, L^-1 and L^-T L^-1
// (~n^3 / 2).  A zero, disconnected or numerically indefinite matrix leaves zeros (no coarse
// correction).
inline void coarse_inverse(double* a, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double v = 0.5 * (a[(size_t)i * n + j] + a[(size_t)j * n + i]);
      a[(size_t)i * n + j] = v;
      a[(size_t)j * n + i] = v;
    }
  double tr = 0.0;
  for (int i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int j = 0; j < n; ++j)
      if (j != i) sum += a[(size_t)i * n + j];
    a[(size_t)i * n + i] = -sum;
    tr -= sum;
  }
  const size_t nn = (size_t)n * n;
  if (!(tr > 0.0)) {
    std::fill(a, a + nn, 0.0);
    return;
  }
  // bandwidth of the patch graph (patches are numbered row by row over the problem's grid of
  // rectangles, so a grid of k columns has bandwidth ~k + 1)
  int bw = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i - bw; ++j)
      if (a[(size_t)i * n + j] != 0.0) {
        bw = i - j;
        break;
      }
  if (n > 8 && 4 * bw <= n) {
    // Banded path.  Ground the last patch: A_red (the first m = n - 1 rows and columns) is
    // SPD for a connected patch graph and keeps the band.  Its banded Cholesky factor, then
    // G = A_red^-1 by two banded triangular solves with all right-hand sides (~1.5 bw n^2
    // multiply-adds), then the pseudo-inverse J G J (G padded with a zero last row and
    // column): for a connected graph Laplacian the centred grounded inverse is the
    // pseudo-inverse.
    const int m = n - 1, w = bw + 1;
    std::vector<double> L((size_t)m * w, 0.0);  // L[i][i - k] at L[i * w + k], k = 0..bw
    auto el = [&](int i, int j) -> double& { return L[(size_t)i * w + (i - j)]; };
    for (int j = 0; j < m; ++j) {
      double d = a[(size_t)j * n + j];
      for (int k = std::max(0, j - bw); k < j; ++k) d -= el(j, k) * el(j, k);
      if (!(d > 1e-14 * tr / n)) {
        std::fill(a, a + nn, 0.0);
        return;
      }
      el(j, j) = std::sqrt(d);
      const double inv = 1.0 / el(j, j);
      for (int i = j + 1; i < std::min(m, j + bw + 1); ++i) {
        double v = a[(size_t)i * n + j];
        for (int k = std::max(0, i - bw); k < j; ++k) v -= el(i, k) * el(j, k);
        el(i, j) = v * inv;
      }
    }
    // all right-hand sides at once, row by row (independent axpys over rows, which vectorise,
    // instead of one dependent reduction per column): rows of Y = L^-1 (row k is zero right of
    // column k), then rows of X = L^-T Y from the bottom up
    std::fill(a, a + nn, 0.0);
    for (int k = 0; k < m; ++k) {
      double* rk = a + (size_t)k * n;
      rk[k] = 1.0;
      for (int l = std::max(0, k - bw); l < k; ++l) {
        const double f = el(k, l);
        const double* rl = a + (size_t)l * n;
        for (int j = 0; j <= l; ++j) rk[j] -= f * rl[j];
      }
      const double inv = 1.0 / el(k, k);
      for (int j = 0; j <= k; ++j) rk[j] *= inv;
    }
    for (int k = m - 1; k >= 0; --k) {
      double* rk = a + (size_t)k * n;
      for (int l = k + 1; l < std::min(m, k + bw + 1); ++l) {
        const double f = el(l, k);
        const double* rl = a + (size_t)l * n;
        for (int j = 0; j < m; ++j) rk[j] -= f * rl[j];
      }
      const double inv = 1.0 / el(k, k);
      for (int j = 0; j < m; ++j) rk[j] *= inv;
    }
  } else {
    // Dense path: (A_c + s 11^T)^-1, s = trace / n^2 (the rank-one term makes the constant
    // vector invertible; the projection below removes it again).  Cholesky, lower triangle in
    // place
    const double sig = tr / ((double)n * n);
    for (size_t k = 0; k < nn; ++k) a[k] += sig;
    for (int j = 0; j < n; ++j) {
      double* rj = a + (size_t)j * n;
      double d = rj[j];
      for (int k = 0; k < j; ++k) d -= rj[k] * rj[k];
      if (!(d > 1e-14 * tr / n)) {
        std::fill(a, a + nn, 0.0);
        return;
      }
      rj[j] = std::sqrt(d);
      const double inv = 1.0 / rj[j];
      for (int i = j + 1; i < n; ++i) {
        double* ri = a + (size_t)i * n;
        double v = ri[j];
        for (int k = 0; k < j; ++k) v -= ri[k] * rj[k];
        ri[j] = v * inv;
      }
    }
    // M = L^-1 (lower): row i = (e_i - sum_{k<i} L[i][k] row k) / L[i][i]
    std::vector<double> m(nn, 0.0);
    for (int i = 0; i < n; ++i) {
      double* mi = m.data() + (size_t)i * n;
      const double* li = a + (size_t)i * n;
      for (int k = 0; k < i; ++k) {
        const double f = li[k];
        const double* mk = m.data() + (size_t)k * n;
        for (int j = 0; j <= k; ++j) mi[j] -= f * mk[j];
      }
      mi[i] = 1.0;
      const double inv = 1.0 / li[i];
      for (int j = 0; j <= i; ++j) mi[j] *= inv;
    }
    // a = M^T M: a[p][q] = sum_{k >= max(p, q)} M[k][p] M[k][q]
    std::fill(a, a + nn, 0.0);
    for (int k = 0; k < n; ++k) {
      const double* mk = m.data() + (size_t)k * n;
      for (int p = 0; p <= k; ++p) {
        const double f = mk[p];
        double* ap = a + (size_t)p * n;
        for (int q = 0; q <= p; ++q) ap[q] += f * mk[q];
      }
    }
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) a[(size_t)p * n + q] = a[(size_t)q * n + p];
  }
  // project out the constant vector (J a J, J = I - 11^T / n): the pseudo-inverse of A_c, so the
  // coarse term adds nothing along the Hessian's null space
  std::vector<double> mean(n, 0.0);
  double all = 0.0;
  for (int p = 0; p < n; ++p) {
    double r = 0.0;
    for (int q = 0; q < n; ++q) r += a[(size_t)p * n + q];
    mean[p] = r / n;
    all += r;
  }
  all /= (double)n * n;
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) a[(size_t)p * n + q] += all - mean[p] - mean[q];
}

}  // namespace cot
