// HIP kernels of the semi-discrete OT dual evaluation and the L-BFGS vector work, for gfx950.
//
// cell_kernel   one site per lane.  Builds the restricted power cell of the site by clipping the
//               target rectangle with the half-planes of candidate sites (the previous cells'
//               neighbours, or bucket rings), polygon held per lane in LDS; stores the cell's
//               vertices (FP64), edge records, area and centroid.
// integrate_kernel  4, 8 or 16 lanes per site (a tuning option): integrates the piecewise-constant target over the
//               cell exactly with Green's theorem -- each edge is walked across the pixel
//               lattice and every piece adds the closed-form integral of the row prefix of the
//               density.  Mass, first and second moments are summed in FP64 (mass with a
//               compensated sum), across the site's lanes with shuffles.
//               Reference semantics: polygons_in_VD + compute_inte + compute_f + compute_center
//               (power_diagram.cpp:8-134, basic_function.cuh:743-1191).
// the conformity check (verify_site, run by workgroups of the integration launch) checks that
//               every edge of every cell is matched by the reciprocal edge of the neighbouring
//               cell; a cell built from too few candidates fails this and is rebuilt with a
//               larger ring.
// reduce/accept/direction kernels: FP64 per-problem reductions and the L-BFGS vector updates.
#include "cot_internal.h"

namespace cot {

// ---------------------------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Wave-wide min / max (exact operations, so any pairing gives the same result) without the LDS
// crossbar: four DPP steps within each row of 16 lanes (lane i with i ^ 1, then i ^ 2, then the
// half-row mirror 7 - i, then the row mirror 15 - i), then the four rows' values by readlane.
// Called with every lane of the wave active; the result is the same in every lane.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror
constexpr int kDppMirror = 0x140;      // row_mirror
constexpr int kDppShr1 = 0x111, kDppShr2 = 0x112, kDppShr4 = 0x114, kDppShr8 = 0x118;

template <int C>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, C, 0xF, 0xF, true);
}
template <int C>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_i<C>((int)(unsigned)b);
  const unsigned hi = (unsigned)dpp_i<C>((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double rdl_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_d<kDppXor1>(v));
  v = fmax(v, dpp_d<kDppXor2>(v));
  v = fmax(v, dpp_d<kDppHalfMirror>(v));
  v = fmax(v, dpp_d<kDppMirror>(v));
  return fmax(fmax(rdl_d(v, 0), rdl_d(v, 16)), fmax(rdl_d(v, 32), rdl_d(v, 48)));
}
__device__ __forceinline__ double wave_min(double v) {
  v = fmin(v, dpp_d<kDppXor1>(v));
  v = fmin(v, dpp_d<kDppXor2>(v));
  v = fmin(v, dpp_d<kDppHalfMirror>(v));
  v = fmin(v, dpp_d<kDppMirror>(v));
  return fmin(fmin(rdl_d(v, 0), rdl_d(v, 16)), fmin(rdl_d(v, 32), rdl_d(v, 48)));
}
__device__ __forceinline__ int wave_max_i(int v) {
  v = max(v, dpp_i<kDppXor1>(v));
  v = max(v, dpp_i<kDppXor2>(v));
  v = max(v, dpp_i<kDppHalfMirror>(v));
  v = max(v, dpp_i<kDppMirror>(v));
  return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
// Inclusive prefix sum over the wave (every lane active): a shift-and-add scan within each row of
// 16 lanes (DPP row shifts by 1, 2, 4, 8; lanes shifted in from outside the row add 0), then the
// totals of the rows below.
__device__ __forceinline__ int wave_scan_i(int v) {
  const int ln = (int)__lane_id();
  v += dpp_i<kDppShr1>(v);
  v += dpp_i<kDppShr2>(v);
  v += dpp_i<kDppShr4>(v);
  v += dpp_i<kDppShr8>(v);
  const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31);
  const int r2 = __builtin_amdgcn_readlane(v, 47);
  return v + (ln >= 16 ? r0 : 0) + (ln >= 32 ? r1 : 0) + (ln >= 48 ? r2 : 0);
}

// Sum of v over the block (blockDim.x a multiple of 64, at most 1024); result valid in thread 0.
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) s += red[k];
  }
  return s;
}
__device__ double block_max(double v, double* red) {
  v = wave_max(v);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) red[wv] = v;
  __syncthreads();
  double s = -INFINITY;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) s = fmax(s, red[k]);
  }
  return s;
}

// Atomic increment and exchange usable from the per-site code on the device and in its host
// build.
__host__ __device__ __forceinline__ int fetch_add(int* p, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}
__host__ __device__ __forceinline__ void count_add(unsigned long long* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd(p, 1ull);
#else
  __atomic_fetch_add(p, 1ull, __ATOMIC_RELAXED);
#endif
}
__host__ __device__ __forceinline__ int exchange(int* p, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicExch(p, v);
#else
  return __atomic_exchange_n(p, v, __ATOMIC_RELAXED);
#endif
}

// Neumaier compensated accumulation.
struct KSum {
  double s = 0.0, c = 0.0;
  __host__ __device__ __forceinline__ void add(double x) {
    double t = s + x;
    c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
  }
  __host__ __device__ __forceinline__ double value() const { return s + c; }
};

template <typename T, int V>
struct PolyLDS {
  static constexpr int kVerts = V;
  T x[V][kCellBlock];
  T y[V][kCellBlock];
  int lab[V][kCellBlock];
};

// Keep the part of the lane's convex polygon with u.d <= c; the new edge gets `label`.
// Returns the new vertex count (0 when empty).  Clipping is in place: at step k the writes land
// at positions <= k + 1, and vertex k + 1 has already been read into registers.
// When the polygon is emptied, sdmin is the smallest u.d - c over its vertices.
template <typename T, int V>
__host__ __device__ int clip_halfplane(PolyLDS<T, V>& P, int ln, int n, T dx, T dy, T c,
                                       int label, bool& overflow, bool& changed, T& sdmin) {
#pragma clang fp contract(on)
  int nout = 0, trans = 0;
  {
    T sd_first = P.x[0][ln] * dx + P.y[0][ln] * dy - c;
    sdmin = sd_first;
    bool prev = sd_first > T(0);
    nout += prev;
    bool first = prev;
    for (int k = 1; k < n; ++k) {
      T sd = P.x[k][ln] * dx + P.y[k][ln] * dy - c;
      sdmin = sd < sdmin ? sd : sdmin;
      bool o = sd > T(0);
      nout += o;
      trans += (o != prev);
      prev = o;
    }
    trans += (prev != first);
  }
  changed = false;
  if (nout == 0) return n;
  if (nout == n) {
    changed = true;
    return 0;
  }
  if (trans != 2) return n;  // inconsistent signs from rounding on a degenerate polygon: keep
  if (n + 1 > V) {
    overflow = true;
    return n;
  }
  changed = true;
  T sx = P.x[0][ln], sy = P.y[0][ln];
  int sl = P.lab[0][ln];
  const T fx = sx, fy = sy;
  const int fl = sl;
  T ssd = sx * dx + sy * dy - c;
  int out = 0;
  for (int k = 0; k < n; ++k) {
    T ex, ey;
    int el;
    if (k + 1 < n) {
      ex = P.x[k + 1][ln];
      ey = P.y[k + 1][ln];
      el = P.lab[k + 1][ln];
    } else {
      ex = fx;
      ey = fy;
      el = fl;
    }
    T esd = ex * dx + ey * dy - c;
    bool sin_ = ssd <= T(0), ein = esd <= T(0);
    if (sin_) {
      P.x[out][ln] = sx;
      P.y[out][ln] = sy;
      P.lab[out][ln] = sl;
      ++out;
    }
    if (sin_ != ein) {
      T tt = ssd / (ssd - esd);
      P.x[out][ln] = sx + tt * (ex - sx);
      P.y[out][ln] = sy + tt * (ey - sy);
      P.lab[out][ln] = sin_ ? label : sl;
      ++out;
    }
    sx = ex;
    sy = ey;
    sl = el;
    ssd = esd;
  }
  return out >= 3 ? out : 0;
}

// Row-prefix lookup for the Green walk.  Columns are local to the cell (local column 0 = global
// column kox); the sums over the whole pixels of local columns 0 .. ix-1 of
// rho_k * (1, k + 1/2, k^2 + k + 1/3) (the integrals of 1, U, U^2 over those pixels, local
// coordinates) and rho at (iy, ix) come from the chunked prefixes: the part of kox's chunk right
// of kox, whole chunks between, and the part of ix's chunk left of ix, each moved to the local
// origin by at most a chunk plus the cell's width.  The chunk-local prefix at a pixel comes from
// its density group (cot_internal.h): the group's prefix plus the group's pixels left of it.  A
// lookup is split into load() -- independent 64-byte record reads, issued one piece ahead by the
// walk -- and eval().
struct PrefixTaps {
  DensityGroup t;   // group of (row, kox + ix)
  DensityGroup qk;  // group of (row, kox)            (fresh rows only)
  double4 tk;       // total of kox's chunk in row    (fresh rows only)
  bool fresh;       // the piece starts a new row
};

// Prefix (s0, s1, s2) and density r at global column g of the row, from g's group record
// (gsize pixels per group).  The second pixel of an FP32 pair adds the first one's terms with
// prefix_step, as the prefix kernel summed them.
__host__ __device__ __forceinline__ void group_tap(const DensityGroup& rec, int g, int gsize,
                                                   double& s0, double& s1, double& s2,
                                                   double& r) {
  s0 = rec.s0;
  s1 = rec.s1;
  s2 = rec.s2;
  if (gsize == kGroup32) {
    const double r0 = (double)rec.r.f[0];
    r = r0;
    if (g & 1) {
      prefix_step(s0, s1, s2, r0, (double)((g - 1) % kChunk));
      r = (double)rec.r.f[1];
    }
  } else {
    r = rec.r.d;
  }
}

struct Prefix {
  const DensityGroup* grp;  // this problem's density groups
  const double4* tot;       // this problem's chunk totals
  int ngrp, gsize, gshift, nchunk, kox, koy;  // gsize = 1 << gshift
  int lrow = -1;       // row of the last load
  double q0 = 0, q1 = 0, q2 = 0;  // prefix of kox within its chunk (chunk-local), current row
  double f0 = 0, f1 = 0, f2 = 0;  // kox .. end of its chunk, local coordinates, current row

  // moments s (coordinates k') moved to coordinates k' + d
  __host__ __device__ static __forceinline__ void shift(double& s0, double& s1, double& s2,
                                                        double d) {
    s2 += 2.0 * d * s1 + d * d * s0;
    s1 += d * s0;
  }
  // Reads for the piece at (iy, ix); the row's two references only when the row changes.
  __host__ __device__ __forceinline__ PrefixTaps load(int iy, int ix) {
    const int grow = koy + iy;
    const DensityGroup* gr = grp + (size_t)grow * ngrp;
    PrefixTaps p;
    p.t = gr[(kox + ix) >> gshift];
    p.fresh = iy != lrow;
    if (p.fresh) {
      p.qk = gr[kox >> gshift];
      p.tk = tot[(size_t)grow * nchunk + kox / kChunk];
      lrow = iy;
    }
    return p;
  }
  // Take the row references of a piece that starts a new row (every piece, in walk order, also
  // those that add nothing).
  __host__ __device__ __forceinline__ void refresh(const PrefixTaps& p) {
    if (!p.fresh) return;
    double rk;
    group_tap(p.qk, kox, gsize, q0, q1, q2, rk);
    f0 = p.tk.x - q0;
    f1 = p.tk.y - q1;
    f2 = p.tk.z - q2;
    const int ck = kox / kChunk;
    shift(f0, f1, f2, (double)(ck * kChunk - kox));
  }
  __host__ __device__ __forceinline__ void eval(const PrefixTaps& p, int iy, int ix, double& a0,
                                                double& a1, double& a2, double& r) const {
    const int g = kox + ix;
    const int ck = kox / kChunk, cg = g / kChunk;
    double t0, t1, t2;
    group_tap(p.t, g, gsize, t0, t1, t2, r);
    if (cg == ck) {
      a0 = t0 - q0;
      a1 = t1 - q1;
      a2 = t2 - q2;
      shift(a0, a1, a2, (double)(ck * kChunk - kox));
      return;
    }
    a0 = f0;
    a1 = f1;
    a2 = f2;
    const int grow = koy + iy;
    for (int c = ck + 1; c < cg; ++c) {
      const double4 tc = tot[(size_t)grow * nchunk + c];
      double b0 = tc.x, b1 = tc.y, b2 = tc.z;
      shift(b0, b1, b2, (double)(c * kChunk - kox));
      a0 += b0;
      a1 += b1;
      a2 += b2;
    }
    double b0 = t0, b1 = t1, b2 = t2;
    shift(b0, b1, b2, (double)(cg * kChunk - kox));
    a0 += b0;
    a1 += b1;
    a2 += b2;
  }
};

struct Moments {
  KSum m0;
  double m1u = 0.0, m1v = 0.0, m2u = 0.0, m2v = 0.0;
};

// Exact contribution of one edge piece inside pixel (ix, iy) (local pixel units) to
// oint F dV for F = int rho, int rho*U, int rho*U^2 along U, and V*F0, V^2*F0.
// TP = double: everything in FP64.  TP = float: the polynomial terms of the first and second
// moments are evaluated in FP32 in coordinates local to the pixel (u, v in [0, 1]); the mass
// term, the row prefixes A0..A2, the integer offsets and every accumulation stay FP64 (the
// edges' contributions to a cell's mass cancel in part -- Green's theorem -- which would
// amplify FP32 rounding of the pixel-local coordinates past 1e-6 of the cell's mass in long
// cells; four FP64 operations per piece in a latency-bound walk cost nothing measurable).
template <typename TP>
__host__ __device__ __forceinline__ void piece(double U0, double V0, double U1, double V1, int ix,
                                               int iy, Prefix& px, const PrefixTaps& taps,
                                               Moments& M) {
  px.refresh(taps);
  const double dv = V1 - V0;
  if (dv == 0.0) return;
  double a0, a1, a2, rv;
  px.eval(taps, iy, ix, a0, a1, a2, rv);
  const double fx = (double)ix, fy = (double)iy;
  if constexpr (sizeof(TP) == 4) {
    const float r = (float)rv;
    const float u0 = (float)(U0 - fx), u1 = (float)(U1 - fx);
    const float v0 = (float)(V0 - fy), v1 = (float)(V1 - fy);
    const float dvf = (float)dv;
    const float um = 0.5f * (u0 + u1), vm = 0.5f * (v0 + v1);
    const float Iu = um * dvf;                                                   // int u dV
    const float Iuu = (u0 * u0 + u0 * u1 + u1 * u1) * (1.0f / 3.0f) * dvf;        // int u^2
    const float Iuuu = (u0 * u0 * u0 + 4.0f * um * um * um + u1 * u1 * u1) * (1.0f / 6.0f) * dvf;
    const float Iv = vm * dvf;                                                   // int v
    const float Ivv = (v0 * v0 + v0 * v1 + v1 * v1) * (1.0f / 3.0f) * dvf;        // int v^2
    const float Ivu = (v0 * u0 + 4.0f * vm * um + v1 * u1) * (1.0f / 6.0f) * dvf;
    const float Ivvu = (v0 * v0 * u0 + 4.0f * vm * vm * um + v1 * v1 * u1) * (1.0f / 6.0f) * dvf;
    const double pm = a0 * dv + rv * (0.5 * ((U0 - fx) + (U1 - fx)) * dv);  // int F0 dV
    M.m0.add(pm);
    M.m1u += a1 * dv + fx * (double)(r * Iu) + 0.5 * (double)(r * Iuu);
    const double vF0 = a0 * (double)Iv + (double)(r * Ivu);  // int v F0 dV
    M.m1v += fy * pm + vF0;
    M.m2u += a2 * dv + fx * fx * (double)(r * Iu) + fx * (double)(r * Iuu) +
             (1.0 / 3.0) * (double)(r * Iuuu);
    const double v2F0 = a0 * (double)Ivv + (double)(r * Ivvu);  // int v^2 F0 dV
    M.m2v += fy * fy * pm + 2.0 * fy * vF0 + v2F0;
  } else {
    const double rc_ = rv;
    const double Um = 0.5 * (U0 + U1), Vm = 0.5 * (V0 + V1);
    // mass: int (A0 + rho (U - ix)) dV
    M.m0.add((a0 + rc_ * (Um - fx)) * dv);
    // int rho U: int (A1 + rho (U^2 - ix^2)/2) dV
    const double u2avg = (U0 * U0 + U0 * U1 + U1 * U1) * (1.0 / 3.0);
    M.m1u += (a1 + 0.5 * rc_ * (u2avg - fx * fx)) * dv;
    // int rho V: int V (A0 + rho (U - ix)) dV
    const double vu = (V0 * U0 + 4.0 * Vm * Um + V1 * U1) * (1.0 / 6.0) * dv;
    M.m1v += (a0 - rc_ * fx) * 0.5 * (V1 * V1 - V0 * V0) + rc_ * vu;
    // int rho U^2: int (A2 + rho (U^3 - ix^3)/3) dV
    const double u3 = (U0 * U0 * U0 + 4.0 * Um * Um * Um + U1 * U1 * U1) * (1.0 / 6.0) * dv;
    M.m2u += (a2 - rc_ * fx * fx * fx * (1.0 / 3.0)) * dv + rc_ * (1.0 / 3.0) * u3;
    // int rho V^2: int V^2 (A0 + rho (U - ix)) dV
    const double v2u = (V0 * V0 * U0 + 4.0 * Vm * Vm * Um + V1 * V1 * U1) * (1.0 / 6.0) * dv;
    M.m2v += (a0 - rc_ * fx) * (V1 * V1 * V1 - V0 * V0 * V0) * (1.0 / 3.0) + rc_ * v2u;
  }
}

// Walk the edge (Ua,Va)->(Ub,Vb) across the pixel lattice (local pixel units, lattice limits
// lim_u x lim_v) and add each piece.  The table reads of the next piece are issued before the
// current piece is evaluated.
struct EdgeWalker {
  double ua, va, ub, vb, du, dvv, idu, idv, tu, tv, t;
  int su, sv, nu_line, nv_line, lim_u, lim_v, steps_left;
  double U0, V0;

  __host__ __device__ __forceinline__ void init(double a_u, double a_v, double b_u, double b_v,
                                                int lu, int lv) {
    ua = a_u; va = a_v; ub = b_u; vb = b_v; lim_u = lu; lim_v = lv;
    du = ub - ua;
    dvv = vb - va;
    const int iu = (int)floor(ua), iv = (int)floor(va);
    const int iu_end = (int)floor(ub), iv_end = (int)floor(vb);
    su = du > 0.0 ? 1 : -1;
    sv = dvv > 0.0 ? 1 : -1;
    idu = du != 0.0 ? 1.0 / du : 0.0;
    idv = 1.0 / dvv;
    nu_line = du > 0.0 ? iu + 1 : iu;
    nv_line = dvv > 0.0 ? iv + 1 : iv;
    tu = du != 0.0 ? ((double)nu_line - ua) * idu : 2.0;
    tv = ((double)nv_line - va) * idv;
    steps_left = abs(iu_end - iu) + abs(iv_end - iv) + 2;
    t = 0.0;
    U0 = ua;
    V0 = va;
  }
  // The next piece: its end point and pixel; returns true when it is the edge's last piece.
  __host__ __device__ __forceinline__ bool next(double& P0u, double& P0v, double& P1u,
                                                double& P1v, int& ix, int& iy) {
    double tn = fmin(fmin(tu, tv), 1.0);
    if (tn < t) tn = t;
    const bool last = tn >= 1.0 || --steps_left < 0;
    double U1, V1;
    if (tn >= 1.0) {
      U1 = ub;
      V1 = vb;
    } else {
      U1 = ua + tn * du;
      V1 = va + tn * dvv;
    }
    ix = min(max((int)floor(0.5 * (U0 + U1)), 0), lim_u - 1);
    iy = min(max((int)floor(0.5 * (V0 + V1)), 0), lim_v - 1);
    P0u = U0;
    P0v = V0;
    P1u = U1;
    P1v = V1;
    if (!last) {
      if (tu <= tv) {
        nu_line += su;
        tu = ((double)nu_line - ua) * idu;
      } else {
        nv_line += sv;
        tv = ((double)nv_line - va) * idv;
      }
      t = tn;
      U0 = U1;
      V0 = V1;
    }
    return last;
  }
};

template <typename T>
__host__ __device__ void walk_edge(double ua, double va, double ub, double vb, int lim_u, int lim_v,
                                   Prefix& px, Moments& M) {
  if (vb - va == 0.0) return;
  EdgeWalker w;
  w.init(ua, va, ub, vb, lim_u, lim_v);
  for (;;) {
    double a0u, a0v, a1u, a1v;
    int ix, iy;
    const bool last = w.next(a0u, a0v, a1u, a1v, ix, iy);
    const PrefixTaps taps = px.load(iy, ix);
    piece<T>(a0u, a0v, a1u, a1v, ix, iy, px, taps, M);
    if (last) break;
  }
}

// ---------------------------------------------------------------------------------------------
// cell kernel
// ---------------------------------------------------------------------------------------------

// One site of the cell kernel (lane `site_ln` of block `blk`); returns -1 when there is no site
// to build, 1 when the cell exceeded kMaxVerts, 0 otherwise.  A full pass clips in a polygon of
// V = kPassVerts vertices: a cell that would exceed it is handed to the repair pass (like an
// FP32 cell FP32 cannot resolve), whose polygon holds kMaxVerts.  Shared by the kernels and the
// host build used to test this code on the CPU.
template <typename T, int V>
__host__ __device__ int cell_site(const CellArgs& a, int blk, int site_ln, PolyLDS<T, V>& P,
                                  int ln) {
  // multiply-adds fused only within one expression, so that the arithmetic is fixed by the
  // source and the wave construction below reproduces it bit for bit
#pragma clang fp contract(on)
  // site_ln: the site's lane in its cell block; ln: the LDS column of the polygon
  const BlockDesc bd = a.blocks[blk];
  if (site_ln >= bd.count) return -1;
  if (!a.active[bd.prob]) return -1;
  if (a.gate && a.gate[bd.prob] == 0) return -1;
  const int i = bd.first + site_ln;
  if (a.redo_mode && !a.redo[i]) return -1;
  if (a.redo_mode) {
    // a repair clears the flag of the site it rebuilds (the check that follows may set it again)
    a.redo[i] = 0;
    if (a.diag) count_add(&a.diag[kDiagRebuilt]);
  }
  if (!a.redo_mode && !a.prepass && a.prebuilt && a.prebuilt[i]) {
    // built by the pre-pass (cell_reappear_kernel); the candidates queued for it are dropped
    // (its cell came from the bucket rings)
    a.prebuilt[i] = 0;
    if (a.xcnt) a.xcnt[i] = 0;
    return -1;
  }
  // the full pass clears the flag (an FP32 cell may set it again below); in a repair pass the
  // integration pass that follows clears it
  if (!a.redo_mode && a.redo) a.redo[i] = 0;
  const ProbDesc pd = a.probs[bd.prob];
  // trial weights w + t * d, with d = -g for a steepest-descent trial
  double tstep = 0.0;
  bool steep = false;
  if (a.st) {
    tstep = a.st[bd.prob].t;
    steep = a.st[bd.prob].steep != 0;
  } else if (a.dir) {
    tstep = a.step[bd.prob];
  }
  const bool trial = a.dir != nullptr;
  auto wtrial = [&](int j) {
    if (!trial) return a.w[j];
    return a.w[j] + tstep * (steep ? -a.g[j] : a.dir[j]);
  };

  const double2 s = a.xy[i];
  const double wi = wtrial(i);
  constexpr bool kF32 = sizeof(T) == 4;
  // Clipping frame: the site (FP64) or the cell's last centroid (FP32, so that the final
  // vertices are small numbers).  Both start from the target rectangle, a true constraint, so the
  // clipped polygon always contains the cell.
  const double2 o = kF32 ? a.org[i] : s;
  const double osx = o.x - s.x, osy = o.y - s.y;  // frame origin relative to the site
  P.x[0][ln] = (T)(pd.x0 - o.x); P.y[0][ln] = (T)(pd.y0 - o.y); P.lab[0][ln] = kSideBottom;
  P.x[1][ln] = (T)(pd.x1 - o.x); P.y[1][ln] = (T)(pd.y0 - o.y); P.lab[1][ln] = kSideRight;
  P.x[2][ln] = (T)(pd.x1 - o.x); P.y[2][ln] = (T)(pd.y1 - o.y); P.lab[2][ln] = kSideTop;
  P.x[3][ln] = (T)(pd.x0 - o.x); P.y[3][ln] = (T)(pd.y1 - o.y); P.lab[3][ln] = kSideLeft;
  int nv = 4;
  bool overflow = false;
  // FP32 geometry: the cell is near a degeneracy FP32 cannot resolve (a sliver, a marginal
  // emptying, a vertex the FP64 refinement could not place) and is rebuilt by an FP64 pass
  bool coarse = false;

  // bounding circle of the polygon (centre cx, cy relative to the frame; radius^2 r2)
  auto bound = [&](double& cx, double& cy, double& r2) {
    double mnx = P.x[0][ln], mxx = mnx, mny = P.y[0][ln], mxy = mny;
    for (int k = 1; k < nv; ++k) {
      double xx = P.x[k][ln], yy = P.y[k][ln];
      mnx = fmin(mnx, xx); mxx = fmax(mxx, xx);
      mny = fmin(mny, yy); mxy = fmax(mxy, yy);
    }
    cx = 0.5 * (mnx + mxx);
    cy = 0.5 * (mny + mxy);
    r2 = 0.0;
    for (int k = 0; k < nv; ++k) {
      double ex = (double)P.x[k][ln] - cx, ey = (double)P.y[k][ln] - cy;
      r2 = fmax(r2, ex * ex + ey * ey);
    }
  };
  double bcx, bcy, br2;
  bound(bcx, bcy, br2);
  int used = 0;
  // clip by the bisector of site j (candidate from bucket ring r) at position sj with trial
  // weight wj; `dedupe`: j may already bound an edge (a candidate list with repeats), then
  // clipping by the same line again is skipped
  auto consider_at = [&](int j, double2 sj, double wj, int r, bool dedupe) {
    const double ddx = sj.x - s.x, ddy = sj.y - s.y;
    // half-plane u.d <= c in the clipping frame (u = x - o)
    const double c = 0.5 * (ddx * ddx + ddy * ddy + wi - wj) - (osx * ddx + osy * ddy);
    // the half-plane u.d <= c contains the bounding circle: nothing to clip
    const double slack = c - (bcx * ddx + bcy * ddy);
    if (slack >= 0.0 && slack * slack >= br2 * (ddx * ddx + ddy * ddy)) return;
    if (dedupe) {
      bool have = false;
      for (int e = 0; e < nv; ++e) have = have || P.lab[e][ln] == j;
      if (have) return;
    }
    bool changed;
    T sdmin;
    nv = clip_halfplane<T, V>(P, ln, nv, (T)ddx, (T)ddy, (T)c, j, overflow, changed, sdmin);
    if (changed) {
      used = max(used, r);
      if (nv > 0) {
        bound(bcx, bcy, br2);
      } else if (kF32 && (double)sdmin < kMarginRel * pd.bsize * sqrt(ddx * ddx + ddy * ddy)) {
        coarse = true;  // emptied by a margin FP32 cannot resolve: may be a sliver
      }
    }
  };
  auto consider = [&](int j, int r, bool dedupe) { consider_at(j, a.xy[j], wtrial(j), r, dedupe); };
  // candidates js[0 .. kCandBatch) (-1: none) in order, as consider() one after another; their
  // sites and weights are loaded together first, so the batch waits on memory once rather than
  // once per candidate (the lane walks its candidates alone: nothing else hides the latency)
  auto consider_batch = [&](const int (&js)[kCandBatch], bool dedupe) {
    double2 sj[kCandBatch];
    double wj[kCandBatch];
#pragma unroll
    for (int u = 0; u < kCandBatch; ++u) {
      sj[u] = make_double2(0.0, 0.0);
      wj[u] = 0.0;
      if (js[u] >= 0) {
        sj[u] = a.xy[js[u]];
        wj[u] = wtrial(js[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kCandBatch; ++u)
      if (js[u] >= 0 && nv > 0) consider_at(js[u], sj[u], wj[u], 0, dedupe);
  };
  const int R0 = a.ring[i];
  // the previous cells' adjacency, unless their cells failed the cover check (then fresh
  // bucket rings)
  const bool fresh = a.st && a.st[bd.prob].fresh;
  const int nadj = a.cand != kCandBucket && !fresh ? (int)a.adjn[i] : 0;
  if (a.cand != kCandBucket && !fresh && !a.redo_mode && a.defer_list) {
    // No previous neighbours means the bucket rings, and a large two-ring many candidates: a long
    // serial scan for one lane, which the whole wave waits on.  The wave pass that follows builds
    // such a cell instead (the same cell: cell_site_wave), its candidates 64 at a time.
    bool defer = nadj == 0;
    if (!defer && a.cand == kCandAdj2) {
      const int* row = a.adj + (size_t)i * kMaxVerts;
      int tot2 = 0;
      for (int k0 = 0; k0 < nadj; k0 += kCandBatch) {
        int js[kCandBatch];
#pragma unroll
        for (int u = 0; u < kCandBatch; ++u) js[u] = k0 + u < nadj ? row[k0 + u] : -1;
#pragma unroll
        for (int u = 0; u < kCandBatch; ++u) tot2 += js[u] >= 0 ? (int)a.adjn[js[u]] : 0;
      }
      defer = tot2 > kDeferCand;
    }
    if (defer) {
      a.defer_list[fetch_add(a.defer_count, 1)] = blk * kCellBlock + site_ln;
      return -1;
    }
  }
  const int bx = min(max((int)floor((s.x - pd.bx0) * pd.binv), 0), pd.gx - 1);
  const int by = min(max((int)floor((s.y - pd.by0) * pd.binv), 0), pd.gy - 1);
  // candidates queued by the conformity check, taken in increasing site order (the check appends
  // them in whatever order its threads run; sorting makes the clipping order, and so the FP64
  // vertices, independent of it)
  if (a.xcnt) {
    const int nx = min(a.xcnt[i], kMaxExtra);
    const int* xc = a.xcand + (size_t)i * kMaxExtra;
    int last = -1;
    for (int t = 0; t < nx && nv > 0; ++t) {
      int m = INT32_MAX;
      for (int k = 0; k < nx; ++k) {
        const int j = xc[k];
        if (j > last && j < m) m = j;
      }
      if (m == INT32_MAX) break;
      consider(m, 0, true);
      last = m;
    }
    a.xcnt[i] = 0;
  }
  if (nadj > 0) {
    // the previous cell's neighbours ...
    const int* row = a.adj + (size_t)i * kMaxVerts;
    const bool queued = a.xcnt != nullptr;
    for (int k0 = 0; k0 < nadj && nv > 0; k0 += kCandBatch) {
      int js[kCandBatch];
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) {
        const int j = k0 + u < nadj ? row[k0 + u] : -1;
        js[u] = j >= 0 && j != i ? j : -1;
      }
      consider_batch(js, queued);
    }
    // ... and theirs (an edge flip makes a neighbour's neighbour adjacent)
    if (a.cand == kCandAdj2) {
      for (int k = 0; k < nadj && nv > 0; ++k) {
        const int j = row[k];
        if (j < 0) continue;
        const int nj = a.adjn[j];
        const int* rj = a.adj + (size_t)j * kMaxVerts;
        for (int q0 = 0; q0 < nj && nv > 0; q0 += kCandBatch) {
          int js[kCandBatch];
#pragma unroll
          for (int u = 0; u < kCandBatch; ++u) {
            const int jj = q0 + u < nj ? rj[q0 + u] : -1;
            js[u] = jj >= 0 && jj != i ? jj : -1;
          }
          consider_batch(js, true);
        }
      }
    }
    // the bucket ring that holds every neighbour of the cell (the fallback search's start)
    for (int k = 0; k < nv; ++k) {
      const int j = P.lab[k][ln];
      if (j < 0) continue;
      const double2 sj = a.xy[j];
      const int qx = min(max((int)floor((sj.x - pd.bx0) * pd.binv), 0), pd.gx - 1);
      const int qy = min(max((int)floor((sj.y - pd.by0) * pd.binv), 0), pd.gy - 1);
      used = max(used, max(abs(qx - bx), abs(qy - by)));
    }
  } else {
    const int R = a.cand == kCandAdj2 ? min(2 * R0 + 1, a.ring_max) : R0;
    const int2* bk = a.buckets + pd.bkt_off;
    for (int r = 0; r <= R && nv > 0; ++r) {
      const int ylo = by - r, yhi = by + r;
      for (int cy = max(ylo, 0); cy <= min(yhi, pd.gy - 1) && nv > 0; ++cy) {
        const bool full = (cy == ylo) || (cy == yhi);
        const int stepx = full ? 1 : 2 * r;
        for (int cx = bx - r; cx <= bx + r && nv > 0; cx += stepx) {
          if (cx < 0 || cx >= pd.gx) continue;
          const int2 b = bk[cy * pd.gx + cx];
          for (int jj = 0; jj < b.y && nv > 0; ++jj) {
            const int j = pd.site_off + b.x + jj;
            if (j != i) consider(j, r, false);
          }
        }
      }
    }
  }
  a.ring_used[i] = (uint8_t)min(used, 255);
  a.nedge[i] = (uint8_t)nv;

  double ar = 0.0;
  double2 geo = make_double2(0.0, 0.0);
  int2 kb = make_int2(0, 0);
  if (nv >= 3) {
    // The final polygon in FP64, edge by edge.  FP64 geometry: the vertices as clipped.  FP32
    // geometry: the clipping only decides which constraint bounds each edge; every vertex is
    // then recomputed in FP64 as the intersection of its two constraint lines (frame o).
    auto line_of = [&](int lab, double& nx, double& ny, double& c) {
      if (lab >= 0) {
        const double2 sj = a.xy[lab];
        nx = sj.x - s.x;
        ny = sj.y - s.y;
        c = 0.5 * (nx * nx + ny * ny + wi - wtrial(lab)) - (osx * nx + osy * ny);
      } else if (lab == kSideBottom) {
        nx = 0.0; ny = -1.0; c = o.y - pd.y0;
      } else if (lab == kSideRight) {
        nx = 1.0; ny = 0.0; c = pd.x1 - o.x;
      } else if (lab == kSideTop) {
        nx = 0.0; ny = 1.0; c = pd.y1 - o.y;
      } else if (lab == kSideLeft) {
        nx = -1.0; ny = 0.0; c = o.x - pd.x0;
      } else {
        nx = ny = c = 0.0;
      }
    };
    // FP32: extent of the clipped polygon, for the refinement guard below
    double ext2 = 0.0;
    if (kF32) {
      double mnx = P.x[0][ln], mxx = mnx, mny = P.y[0][ln], mxy = mny;
      for (int k = 1; k < nv; ++k) {
        mnx = fmin(mnx, (double)P.x[k][ln]); mxx = fmax(mxx, (double)P.x[k][ln]);
        mny = fmin(mny, (double)P.y[k][ln]); mxy = fmax(mxy, (double)P.y[k][ln]);
      }
      ext2 = (mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny);
    }
    auto vertex = [&](int k, double ax, double ay, double ac, double bx, double by,
                      double bc) -> double2 {
      const double fx = (double)P.x[k][ln], fy = (double)P.y[k][ln];
      if (!kF32) return make_double2(fx, fy);
      const double det = ax * by - ay * bx;
      const double sc = sqrt((ax * ax + ay * ay) * (bx * bx + by * by));
      if (!(fabs(det) > 1e-7 * sc)) {
        coarse = true;
        return make_double2(fx, fy);
      }
      const double rx = (ac * by - bc * ay) / det, ry = (ax * bc - bx * ac) / det;
      // a near-parallel pair that FP32 rounding made adjacent can meet far away: keep the
      // clipped vertex when the refinement moves it by more than 1e-3 of the cell's extent
      const double ex = rx - fx, ey = ry - fy;
      if (ex * ex + ey * ey > 1e-6 * ext2) {
        coarse = true;
        return make_double2(fx, fy);
      }
      return make_double2(rx, ry);
    };
    // calls f(k, va, vb) for every edge k from vertex k to vertex k+1 (frame coordinates)
    auto for_each_edge = [&](auto&& f) {
      double px, py, pc, kx, ky, kc;
      line_of(P.lab[nv - 1][ln], px, py, pc);
      line_of(P.lab[0][ln], kx, ky, kc);
      const double2 first = vertex(0, px, py, pc, kx, ky, kc);
      double2 vk = first;
      for (int k = 0; k < nv; ++k) {
        double2 vn;
        if (k + 1 < nv) {
          double nx_, ny_, nc_;
          line_of(P.lab[k + 1][ln], nx_, ny_, nc_);
          vn = vertex(k + 1, kx, ky, kc, nx_, ny_, nc_);
          kx = nx_;
          ky = ny_;
          kc = nc_;
        } else {
          vn = first;
        }
        f(k, vk, vn);
        vk = vn;
      }
    };

    // edge records for the conformity check, geometric area/centroid, bounding corner
    double mnu = INFINITY, mnv = INFINITY;
    double A2 = 0.0, gx = 0.0, gy = 0.0;
    for_each_edge([&](int k, double2 va, double2 vb) {
      const double cr = va.x * vb.y - vb.x * va.y;
      A2 += cr;
      gx += (va.x + vb.x) * cr;
      gy += (va.y + vb.y) * cr;
      mnu = fmin(mnu, va.x);
      mnv = fmin(mnv, va.y);
      const int lab = P.lab[k][ln];
      float2 span = make_float2(0.f, 0.f);
      if (lab >= 0) {
        const int lo = min(i, lab), hi = max(i, lab);
        const double2 pl = a.xy[lo], ph = a.xy[hi];
        const double Dx = ph.x - pl.x, Dy = ph.y - pl.y;
        const double inv = 1.0 / sqrt(Dx * Dx + Dy * Dy);
        const double ex = -Dy * inv, ey = Dx * inv;
        const double mx = 0.5 * (pl.x + ph.x), my = 0.5 * (pl.y + ph.y);
        const double t0 = (o.x + va.x - mx) * ex + (o.y + va.y - my) * ey;
        const double t1 = (o.x + vb.x - mx) * ex + (o.y + vb.y - my) * ey;
        span = make_float2((float)fmin(t0, t1), (float)fmax(t0, t1));
      }
      a.elab[(size_t)i * kMaxVerts + k] = lab;
      a.espan[(size_t)i * kMaxVerts + k] = span;
      // the vertex for the integration pass, absolute coordinates
      a.vtx[(size_t)i * kMaxVerts + k] = make_double2(o.x + va.x, o.y + va.y);
    });
    ar = 0.5 * A2;
    if (A2 != 0.0) {
      geo = make_double2(o.x + gx / (3.0 * A2), o.y + gy / (3.0 * A2));
    } else {
      geo = o;
    }

    // integer pixel origin of the cell's box, for the integration pass
    int kox = (int)floor((o.x + mnu - pd.x0) / pd.dx);
    int koy = (int)floor((o.y + mnv - pd.y0) / pd.dy);
    kb = make_int2(min(max(kox, 0), pd.pm - 1), min(max(koy, 0), pd.pn - 1));
    if (kF32) a.org[i] = geo;  // next frame: this centroid
    if (kF32 && fabs(ar) < kSliverRel * pd.bsize * pd.bsize) coarse = true;
  }
  a.kbox[i] = kb;
  // a full pass whose polygon overflowed kPassVerts skipped a constraint: the repair pass
  // rebuilds the cell in its larger polygon
  const bool handoff = overflow && V < kMaxVerts && !a.redo_mode && a.coarse;
  if (handoff) overflow = false;
  if (((kF32 && coarse) || handoff) && a.coarse) {
    a.redo[i] = 2;  // the check that follows lists it
    fetch_add(&a.coarse[bd.prob], 1);
    if (a.diag) count_add(&a.diag[kF32 && coarse ? kDiagDegenerate : kDiagOverflow]);
  }
  a.area[i] = ar;
  a.geo[i] = geo;
  a.exist[i] = nv >= 3 ? 1 : 0;
  return overflow ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// integration pass: L lanes per site (4, 8 or 16); lane q walks edges q, q + L, ...
// ---------------------------------------------------------------------------------------------

// The site of integration lane `t` of block (blk, quarter) with L lanes per site; -1 when none
// (a full pass: the repair passes integrate the sites of their work list).
template <int L = kIntLanes>
__host__ __device__ __forceinline__ int integ_site(const CellArgs& a, int blk, int quarter,
                                                   int t) {
  const BlockDesc bd = a.blocks[blk];
  const int ln = quarter * (kIntBlock / L) + t / L;
  if (ln >= bd.count) return -1;
  if (!a.active[bd.prob]) return -1;
  if (a.gate && a.gate[bd.prob] == 0) return -1;
  return bd.first + ln;
}

// A point of the segment a -> b at parameter t, by one formula wherever a split point of the
// boundary is computed (the two lanes that share it get the same point).
__host__ __device__ __forceinline__ double2 split_point(double2 a, double2 b, double t) {
  return make_double2(fma(t, b.x - a.x, a.x), fma(t, b.y - a.y, a.y));
}

// Walking work of the edge a -> b (local pixel units): its L1 length plus one, about the pieces
// its walk crosses; nothing for an edge along U (no dV: the walk returns at once).
__host__ __device__ __forceinline__ double walk_weight(double2 a, double2 b) {
  return b.y == a.y ? 0.0 : fabs(b.x - a.x) + fabs(b.y - a.y) + 1.0;
}

// Moments of lane q's share of site i's boundary (local pixel units, origin kbox[i]).  The
// boundary, edge after edge, is cut into L runs of equal walking work (walk_weight) and lane q
// walks run q: a cell whose edges differ in length -- most of them, and the large cells of
// low-density regions above all -- would otherwise keep its site's lanes waiting on the lane
// with the longest edge.  The runs meet at split points inside edges (split_point); Green's
// integral is additive over them.
template <typename T, int L = kIntLanes>
__host__ __device__ void integrate_lane(const CellArgs& a, const ProbDesc& pd, int i, int q,
                                        Moments& M) {
  const int nv = a.nedge[i];
  if (nv < 3) return;
  const int2 kb = a.kbox[i];
  Prefix px;
  px.grp = a.grp + pd.grp_off;
  px.tot = a.tot + pd.tot_off;
  px.ngrp = pd.ngrp;
  px.gsize = pd.gsize;
  px.gshift = pd.gsize == kGroup32 ? 1 : 0;
  px.nchunk = pd.nchunk;
  px.kox = kb.x;
  px.koy = kb.y;
  const double idx_ = 1.0 / pd.dx, idy_ = 1.0 / pd.dy;
  const double fu0 = (double)kb.x, fv0 = (double)kb.y;
  const int lim_u = pd.pm - kb.x, lim_v = pd.pn - kb.y;
  const double2* v = a.vtx + (size_t)i * kMaxVerts;
  auto local = [&](int k) {
    const double2 p = v[k];
    return make_double2((p.x - pd.x0) * idx_ - fu0, (p.y - pd.y0) * idy_ - fv0);
  };
  const double2 first = local(0);
  double total = 0.0;
  {
    double2 pa = first;
    for (int k = 0; k < nv; ++k) {
      const double2 pb = k + 1 < nv ? local(k + 1) : first;
      total += walk_weight(pa, pb);
      pa = pb;
    }
  }
  // run q: [s0, s1) of the cumulative weight (the last run ends at the total itself)
  const double s0 = q == 0 ? 0.0 : total * (double)q / (double)L;
  const double s1 = q + 1 == L ? total : total * (double)(q + 1) / (double)L;
  if (!(s1 > s0)) return;
  // the run's first edge k (cumulative weight c0 before it, from pa to pb, weight wk)
  int k = 0;
  double c0 = 0.0;
  double2 pa = first;
  double2 pb = nv > 1 ? local(1) : first;
  double wk = walk_weight(pa, pb);
  while (c0 + wk <= s0 && k + 1 < nv) {
    c0 += wk;
    pa = pb;
    ++k;
    pb = k + 1 < nv ? local(k + 1) : first;
    wk = walk_weight(pa, pb);
  }
  // one flat loop over the run's pieces, edge after edge: the lanes of a wave step through their
  // runs together (a loop per edge would make them wait for each other edge by edge)
  EdgeWalker wlk;
  bool walking = false;
  for (;;) {
    if (!walking) {
      if (k >= nv || c0 >= s1) break;
      const double c1 = c0 + wk;
      const double lo = fmax(c0, s0), hi = fmin(c1, s1);
      if (hi > lo) {
        const double2 p0 = lo == c0 ? pa : split_point(pa, pb, (lo - c0) / wk);
        const double2 p1 = hi == c1 ? pb : split_point(pa, pb, (hi - c0) / wk);
        if (p1.y != p0.y) {  // a part too short to change V adds nothing
          wlk.init(p0.x, p0.y, p1.x, p1.y, lim_u, lim_v);
          walking = true;
        }
      }
      c0 = c1;
      pa = pb;
      ++k;
      if (k < nv) {
        pb = k + 1 < nv ? local(k + 1) : first;
        wk = walk_weight(pa, pb);
      }
      if (!walking) continue;
    }
    double a0u, a0v, a1u, a1v;
    int ix, iy;
    const bool last = wlk.next(a0u, a0v, a1u, a1v, ix, iy);
    const PrefixTaps taps = px.load(iy, ix);
    piece<T>(a0u, a0v, a1u, a1v, ix, iy, px, taps, M);
    if (last) walking = false;
  }
}

// Add moments b into a (compensated for the mass).
__host__ __device__ __forceinline__ void combine(Moments& a, const Moments& b) {
  a.m0.add(b.m0.s);
  a.m0.c += b.m0.c;
  a.m1u += b.m1u;
  a.m1v += b.m1v;
  a.m2u += b.m2u;
  a.m2v += b.m2v;
}

// Outputs of site i from its cell's total moments: mass, objective term, centroid.
__host__ __device__ void finish_site(const CellArgs& a, const ProbDesc& pd, int prob, int i,
                                     const Moments& M) {
  double tstep = 0.0;
  bool steep = false;
  if (a.st) {
    tstep = a.st[prob].t;
    steep = a.st[prob].steep != 0;
  } else if (a.dir) {
    tstep = a.step[prob];
  }
  const double wi = a.dir ? a.w[i] + tstep * (steep ? -a.g[i] : a.dir[i]) : a.w[i];
  double m0 = 0.0, fi = -wi * a.nu[i];
  double2 cen = make_double2(0.0, 0.0);
  if (a.nedge[i] >= 3) {
    const int2 kb = a.kbox[i];
    const double2 s = a.xy[i];
    m0 = M.m0.value();
    // second moment about the site, physical units
    const double ou = (s.x - pd.x0) / pd.dx - (double)kb.x;
    const double ov = (s.y - pd.y0) / pd.dy - (double)kb.y;
    const double su2 = M.m2u - 2.0 * ou * M.m1u + ou * ou * m0;
    const double sv2 = M.m2v - 2.0 * ov * M.m1v + ov * ov * m0;
    fi += wi * m0 - (pd.dx * pd.dx * su2 + pd.dy * pd.dy * sv2);
    cen = m0 != 0.0 ? make_double2(pd.x0 + pd.dx * ((double)kb.x + M.m1u / m0),
                                   pd.y0 + pd.dy * ((double)kb.y + M.m1v / m0))
                    : a.geo[i];
  }
  a.mass[i] = m0;
  a.fcell[i] = fi;
  a.cen[i] = cen;
}


// In-kernel launch timer (kTimerWords words, see cot_internal.h), read with the constant-rate
// wall clock.  timer_begin: the block's start (thread 0); timer_mark: after the block's work, a
// block in which some thread worked folds (start, now) into its slot; timer_close: thread 0 of
// block 0 of the kernel that follows on the stream accumulates the closed launch; timer_finish:
// the launch's last block to end does the same.
__device__ __forceinline__ unsigned long long timer_begin() {
  return threadIdx.x == 0 ? (unsigned long long)wall_clock64() : 0ull;
}

__device__ void timer_mark(unsigned long long* tm, bool worked, unsigned long long t0) {
  if (!tm) return;
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  if (worked) any = 1;
  __syncthreads();
  if (threadIdx.x != 0 || !any) return;
  const int r = (blockIdx.x + gridDim.x * blockIdx.y) % kTimerSlots;
  atomicMin(&tm[2 * r], t0);
  atomicMax(&tm[2 * r + 1], (unsigned long long)wall_clock64());
}

__device__ void timer_close(unsigned long long* tm) {
  if (!tm || threadIdx.x != 0 || blockIdx.x != 0 || blockIdx.y != 0) return;
  unsigned long long t0 = ~0ull, t1 = 0ull;
  for (int r = 0; r < kTimerSlots; ++r) {
    t0 = min(t0, tm[2 * r]);
    t1 = max(t1, tm[2 * r + 1]);
    tm[2 * r] = ~0ull;
    tm[2 * r + 1] = 0ull;
  }
  if (t0 != ~0ull && t1 > t0) {
    tm[2 * kTimerSlots] += t1 - t0;
    tm[2 * kTimerSlots + 1] += 1;
  }
}

// The launch's last block to reach this (every block calls it once, after its timer_mark) closes
// the launch's timer.  Every access to the record is a device-scope atomic, performed where all
// XCDs see it, so ordering needs no cache maintenance: the block waits for its mark's atomics to
// complete before it counts its arrival (an agent-scope fence here would write back and
// invalidate the XCD's L2 once per workgroup), and the last block's reads are atomics issued
// after its arrival returned.
__device__ void timer_finish(unsigned long long* tm) {
  if (!tm) return;
  __syncthreads();
  if (threadIdx.x != 0) return;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  const unsigned long long total = (unsigned long long)gridDim.x * gridDim.y;
  if (atomicAdd(&tm[2 * kTimerSlots + 2], 1ull) + 1 != total) return;
  atomicExch(&tm[2 * kTimerSlots + 2], 0ull);
  unsigned long long t0 = ~0ull, t1 = 0ull;
  for (int r = 0; r < kTimerSlots; ++r) {
    t0 = min(t0, atomicExch(&tm[2 * r], ~0ull));
    t1 = max(t1, atomicExch(&tm[2 * r + 1], 0ull));
  }
  if (t0 != ~0ull && t1 > t0) {
    atomicAdd(&tm[2 * kTimerSlots], t1 - t0);
    atomicAdd(&tm[2 * kTimerSlots + 1], 1ull);
  }
}

// Construction pass: clears the length of the work list the conformity check that follows fills
// (no kernel reads or writes that list until the check).
__device__ __forceinline__ void reset_check_list(const CellArgs& a) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && a.reset_count) *a.reset_count = 0;
}

template <typename T>
__global__ __launch_bounds__(kCellBlock) void cell_kernel(CellArgs a) {
  __shared__ PolyLDS<T, kPassVerts> P;
  reset_check_list(a);
  const unsigned long long t0 = timer_begin();
  const int blk = xcd_slot(blockIdx.x, gridDim.x);
  const int rc = cell_site<T, kPassVerts>(a, blk, threadIdx.x, P, threadIdx.x);
  if (rc == 1) atomicAdd(&a.overflow[a.blocks[blk].prob], 1);
  timer_mark(a.timer, rc >= 0, t0);
}

// The same code for a repair pass (only sites flagged in redo[]); a kernel of its own so that
// profiles and timers keep full evaluations and repairs apart.
template <typename T>
__global__ __launch_bounds__(kCellBlock) void cell_repair_kernel(CellArgs a) {
  __shared__ PolyLDS<T, kMaxVerts> P;
  reset_check_list(a);
  const unsigned long long t0 = timer_begin();
  const int n = *a.list_count;
  bool worked = false;
  // entries dealt round-robin over the blocks, so that a short list still spreads over the CUs
  for (int e = threadIdx.x * gridDim.x + blockIdx.x; e < n; e += gridDim.x * kCellBlock) {
    const int slot = a.list[e];
    const int blk = slot / kCellBlock;
    const int rc = cell_site<T, kMaxVerts>(a, blk, slot % kCellBlock, P, threadIdx.x);
    if (rc == 1) atomicAdd(&a.overflow[a.blocks[blk].prob], 1);
    worked = worked || rc >= 0;
  }
  timer_mark(a.timer, worked, t0);
}

// ---------------------------------------------------------------------------------------------
// wave-cooperative repair: one wave per listed site (FP64)
// ---------------------------------------------------------------------------------------------
//
// The repair pass rebuilds few cells, so a lane per site leaves the chip idle while each lane
// walks its candidates one after another, every candidate a chain of dependent loads.  Here a
// wave builds one cell: vertex k of the polygon lives in lane k (at most kMaxVerts), candidates
// are taken 64 at a time (one per lane, loaded together), a ballot keeps those whose half-plane
// may cut the polygon's bounding circle as it stood when the batch was loaded -- the polygon only
// shrinks, so a half-plane holding that circle holds the polygon -- and the kept ones are clipped
// in lane order, each first tested against the current circle as the lane code tests it.  The
// clip is cooperative: signed distances per lane, the inside/crossing pattern from two ballots,
// output positions from their prefix counts, the new polygon redistributed through LDS.  The
// candidate order (queued sites ascending, the neighbours, their neighbours, the bucket rings
// cell by cell), the clipping arithmetic and the order of the area and centroid sums are those of
// cell_site<double, kMaxVerts>, so both build the same cells.

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// v from lane l (l wave-uniform)
__device__ __forceinline__ int rdl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double rdl(double v, int l) { return rdl_d(v, l); }

template <typename T>
struct WaveLDS {
  T x[kMaxVerts], y[kMaxVerts];  // clip output, staged
  int lab[kMaxVerts];
  int off[64];   // per-slot first candidate of a flattened candidate batch
  int key[64];   // per-slot base (a neighbour's site, a bucket's first site)
};

// Largest slot s in [0, n) with W.off[s] <= f (W.off ascending, W.off[0] = 0).
template <typename W_>
__device__ __forceinline__ int slot_of(const W_& W, int n, int f) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (W.off[mid] <= f) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Bucket (cx, cy) of entry idx of ring r around (bx, by), in the lane code's order: row by - r
// left to right, the rows between (left cell, right cell), row by + r left to right.
__device__ __forceinline__ void ring_cell(int r, int idx, int bx, int by, int& cx, int& cy) {
  if (r == 0) {
    cx = bx;
    cy = by;
    return;
  }
  const int row = 2 * r + 1;
  if (idx < row) {
    cx = bx - r + idx;
    cy = by - r;
    return;
  }
  const int m = idx - row;
  if (m < 2 * (2 * r - 1)) {
    cy = by - r + 1 + (m >> 1);
    cx = (m & 1) ? bx + r : bx - r;
    return;
  }
  cx = bx - r + (m - 2 * (2 * r - 1));
  cy = by + r;
}

// One site built by the calling wave (lane k holds vertex k of the polygon, V at most); returns
// as cell_site<T, V>, and builds the same cell: the same candidates in the same order, the same
// clipping arithmetic in T, FP32 vertices refined in FP64 from their two constraint lines, and
// the same hand-offs to the FP64 repair.
template <typename T, int V>
__device__ int cell_site_wave(const CellArgs& a, int blk, int site_ln, WaveLDS<T>& W) {
#pragma clang fp contract(on)
  static_assert(V <= 64, "one vertex per lane");
  const int ln = lane_id();
  const BlockDesc bd = a.blocks[blk];
  if (site_ln >= bd.count) return -1;
  if (!a.active[bd.prob]) return -1;
  if (a.gate && a.gate[bd.prob] == 0) return -1;
  const int i = bd.first + site_ln;
  if (a.redo_mode && !a.redo[i]) return -1;
  // every lane has read the flag before lane 0 clears it (as cell_site)
  __syncthreads();
  if (a.redo_mode && ln == 0) {
    a.redo[i] = 0;
    if (a.diag) count_add(&a.diag[kDiagRebuilt]);
  }
  if (!a.redo_mode && !a.prepass && a.prebuilt && a.prebuilt[i]) {
    // built by the pre-pass (as cell_site)
    if (ln == 0) {
      a.prebuilt[i] = 0;
      if (a.xcnt) a.xcnt[i] = 0;
    }
    return -1;
  }
  if (!a.redo_mode && a.redo && ln == 0) a.redo[i] = 0;
  const ProbDesc pd = a.probs[bd.prob];
  double tstep = 0.0;
  bool steep = false;
  if (a.st) {
    tstep = a.st[bd.prob].t;
    steep = a.st[bd.prob].steep != 0;
  } else if (a.dir) {
    tstep = a.step[bd.prob];
  }
  const bool trial = a.dir != nullptr;
  auto wtrial = [&](int j) {
    if (!trial) return a.w[j];
    return a.w[j] + tstep * (steep ? -a.g[j] : a.dir[j]);
  };
  const double2 s = a.xy[i];
  const double wi = wtrial(i);
  constexpr bool kF32 = sizeof(T) == 4;
  // clipping frame: the site (FP64) or the cell's last centroid (FP32)
  const double2 o = kF32 ? a.org[i] : s;
  const double osx = o.x - s.x, osy = o.y - s.y;
  // the target rectangle, lanes 0..3
  T px = T(0), py = T(0);
  int pl = 0;
  if (ln == 0) { px = (T)(pd.x0 - o.x); py = (T)(pd.y0 - o.y); pl = kSideBottom; }
  if (ln == 1) { px = (T)(pd.x1 - o.x); py = (T)(pd.y0 - o.y); pl = kSideRight; }
  if (ln == 2) { px = (T)(pd.x1 - o.x); py = (T)(pd.y1 - o.y); pl = kSideTop; }
  if (ln == 3) { px = (T)(pd.x0 - o.x); py = (T)(pd.y1 - o.y); pl = kSideLeft; }
  int nv = 4;
  bool overflow = false;
  bool coarse = false;

  double bcx, bcy, br2;
  auto bound = [&]() {
    const bool on = ln < nv;
    const double dx_ = (double)px, dy_ = (double)py;
    const double mnx = wave_min(on ? dx_ : INFINITY), mxx = wave_max(on ? dx_ : -INFINITY);
    const double mny = wave_min(on ? dy_ : INFINITY), mxy = wave_max(on ? dy_ : -INFINITY);
    bcx = 0.5 * (mnx + mxx);
    bcy = 0.5 * (mny + mxy);
    const double ex = dx_ - bcx, ey = dy_ - bcy;
    br2 = wave_max(on ? ex * ex + ey * ey : 0.0);
  };
  bound();

  // u.d <= c: keep the part of the polygon inside; the new edge gets `label`.  Returns whether
  // the polygon changed; sdmin: when it was emptied, the smallest u.d - c over its vertices.
  auto clip = [&](T dx, T dy, T c, int label, double& sdmin) -> bool {
    const bool on = ln < nv;
    const T sd = px * dx + py * dy - c;
    const unsigned long long mo = __ballot(on && sd > T(0));
    const int nout = __popcll(mo);
    if (nout == 0) return false;
    if (nout == nv) {
      if (kF32) sdmin = wave_min(on ? (double)sd : INFINITY);
      nv = 0;
      return true;
    }
    // cyclic transitions of the outside pattern
    const unsigned long long rot = (mo >> 1) | ((mo & 1ull) << (nv - 1));
    if (__popcll(mo ^ rot) != 2) return false;
    if (nv + 1 > V) {
      overflow = true;
      return false;
    }
    const int nx = ln + 1 < nv ? ln + 1 : 0;
    const T ex = __shfl(px, nx, 64), ey = __shfl(py, nx, 64), esd = __shfl(sd, nx, 64);
    const bool sin_ = on && sd <= T(0);
    const bool cross = on && ((sd <= T(0)) != (esd <= T(0)));
    const unsigned long long mi = __ballot(sin_), mc = __ballot(cross);
    const unsigned long long below = (1ull << ln) - 1ull;
    int pos = __popcll(mi & below) + __popcll(mc & below);
    if (sin_) {
      W.x[pos] = px;
      W.y[pos] = py;
      W.lab[pos] = pl;
      ++pos;
    }
    if (cross) {
      const T tt = sd / (sd - esd);
      W.x[pos] = px + tt * (ex - px);
      W.y[pos] = py + tt * (ey - py);
      W.lab[pos] = sin_ ? label : pl;
    }
    const int out = __popcll(mi) + __popcll(mc);
    if (kF32 && out < 3) sdmin = wave_min(on ? (double)sd : INFINITY);
    __syncthreads();
    if (ln < out) {
      px = W.x[ln];
      py = W.y[ln];
      pl = W.lab[ln];
    }
    __syncthreads();
    nv = out >= 3 ? out : 0;
    return true;
  };

  int used = 0;
  // one batch: lane l offers candidate j (-1: none) from ring r
  auto batch = [&](int j, int r, bool dedupe) {
    double ddx = 0.0, ddy = 0.0, c = 0.0;
    bool keep = false;
    if (j >= 0) {
      const double2 sj = a.xy[j];
      const double wj = wtrial(j);
      ddx = sj.x - s.x;
      ddy = sj.y - s.y;
      c = 0.5 * (ddx * ddx + ddy * ddy + wi - wj) - (osx * ddx + osy * ddy);
      const double slack = c - (bcx * ddx + bcy * ddy);
      keep = !(slack >= 0.0 && slack * slack >= br2 * (ddx * ddx + ddy * ddy));
    }
    unsigned long long m = __ballot(keep);
    while (m != 0ull && nv > 0) {
      const int t = __ffsll((long long)m) - 1;
      m &= m - 1ull;
      const double tdx = rdl(ddx, t), tdy = rdl(ddy, t), tc = rdl(c, t);
      const int tj = rdl(j, t), tr = rdl(r, t);
      const double slack = tc - (bcx * tdx + bcy * tdy);
      if (slack >= 0.0 && slack * slack >= br2 * (tdx * tdx + tdy * tdy)) continue;
      if (dedupe && __ballot(ln < nv && pl == tj) != 0ull) continue;
      double sdmin = 0.0;
      if (clip((T)tdx, (T)tdy, (T)tc, tj, sdmin)) {
        used = max(used, tr);
        if (nv > 0) {
          bound();
        } else if (kF32 && sdmin < kMarginRel * pd.bsize * sqrt(tdx * tdx + tdy * tdy)) {
          coarse = true;  // emptied by a margin FP32 cannot resolve: may be a sliver
        }
      }
    }
  };

  const int R0 = a.ring[i];
  const bool fresh = a.st && a.st[bd.prob].fresh;
  const int nadj = a.cand != kCandBucket && !fresh ? (int)a.adjn[i] : 0;
  const int bx = min(max((int)floor((s.x - pd.bx0) * pd.binv), 0), pd.gx - 1);
  const int by = min(max((int)floor((s.y - pd.by0) * pd.binv), 0), pd.gy - 1);
  if (a.xcnt) {
    // the queued candidates, ascending and without repeats
    const int nx = __builtin_amdgcn_readfirstlane(min(a.xcnt[i], kMaxExtra));
    const int v = ln < nx ? a.xcand[(size_t)i * kMaxExtra + ln] : INT32_MAX;
    bool first = ln < nx;
    for (int q = 0; q < nx; ++q) {
      const int u = rdl(v, q);
      if (q < ln && u == v) first = false;
    }
    int rank = 0;
    for (int q = 0; q < nx; ++q) {
      const int u = rdl(v, q);
      const int fq = rdl((int)first, q);
      if (fq && u < v) ++rank;
    }
    const int nu = __popcll(__ballot(first));
    if (first) W.key[rank] = v;
    __syncthreads();
    const int j = ln < nu ? W.key[ln] : -1;
    __syncthreads();
    if (nu > 0) batch(j, 0, true);
    if (ln == 0) a.xcnt[i] = 0;
  }
  if (nadj > 0) {
    const int* row = a.adj + (size_t)i * kMaxVerts;
    const bool queued = a.xcnt != nullptr;
    const int jn = ln < nadj ? row[ln] : -1;
    if (nv > 0) batch(jn >= 0 && jn != i ? jn : -1, 0, queued);
    if (a.cand == kCandAdj2 && nv > 0) {
      // the neighbours' neighbours, flattened neighbour by neighbour
      const int cnt = jn >= 0 ? (int)a.adjn[jn] : 0;
      const int inc = wave_scan_i(cnt);
      const int total = rdl(inc, 63);
      W.off[ln] = inc - cnt;
      W.key[ln] = jn;
      __syncthreads();
      for (int base = 0; base < total && nv > 0; base += 64) {
        const int f = base + ln;
        int jj = -1;
        if (f < total) {
          const int k = slot_of(W, nadj, f);
          jj = a.adj[(size_t)W.key[k] * kMaxVerts + (f - W.off[k])];
          if (jj == i) jj = -1;
        }
        batch(jj, 0, true);
      }
      __syncthreads();
    }
    // the bucket ring that holds every neighbour of the cell
    int u = 0;
    if (ln < nv && pl >= 0) {
      const double2 sj = a.xy[pl];
      const int qx = min(max((int)floor((sj.x - pd.bx0) * pd.binv), 0), pd.gx - 1);
      const int qy = min(max((int)floor((sj.y - pd.by0) * pd.binv), 0), pd.gy - 1);
      u = max(abs(qx - bx), abs(qy - by));
    }
    used = max(used, wave_max_i(u));
  } else {
    const int R = a.cand == kCandAdj2 ? min(2 * R0 + 1, a.ring_max) : R0;
    const int2* bk = a.buckets + pd.bkt_off;
    for (int r = 0; r <= R && nv > 0; ++r) {
      const int nb = r == 0 ? 1 : 8 * r;
      for (int c0 = 0; c0 < nb && nv > 0; c0 += 64) {
        int cnt = 0, start = 0;
        if (c0 + ln < nb) {
          int cx, cy;
          ring_cell(r, c0 + ln, bx, by, cx, cy);
          if (cx >= 0 && cx < pd.gx && cy >= 0 && cy < pd.gy) {
            const int2 b = bk[cy * pd.gx + cx];
            start = b.x;
            cnt = b.y;
          }
        }
        const int inc = wave_scan_i(cnt);
        const int total = rdl(inc, 63);
        W.off[ln] = inc - cnt;
        W.key[ln] = start;
        __syncthreads();
        for (int base = 0; base < total && nv > 0; base += 64) {
          const int f = base + ln;
          int j = -1;
          if (f < total) {
            const int k = slot_of(W, 64, f);
            j = pd.site_off + W.key[k] + (f - W.off[k]);
            if (j == i) j = -1;
          }
          batch(j, r, false);
        }
        __syncthreads();
      }
    }
  }
  if (ln == 0) {
    a.ring_used[i] = (uint8_t)min(used, 255);
    a.nedge[i] = (uint8_t)nv;
  }

  double ar = 0.0;
  double2 geo = make_double2(0.0, 0.0);
  int2 kb = make_int2(0, 0);
  if (nv >= 3) {
    const bool on = ln < nv;
    // vertex k in FP64: as clipped (FP64 geometry), or the intersection of the constraint lines
    // of edges k - 1 and k (FP32 geometry; kept as clipped where that is ill-conditioned)
    double vx = (double)px, vy = (double)py;
    if (kF32) {
      double lx, ly, lc;  // the line of edge k (label pl)
      if (pl >= 0) {
        const double2 sj = a.xy[on ? pl : i];
        lx = sj.x - s.x;
        ly = sj.y - s.y;
        lc = 0.5 * (lx * lx + ly * ly + wi - wtrial(on ? pl : i)) - (osx * lx + osy * ly);
      } else if (pl == kSideBottom) {
        lx = 0.0; ly = -1.0; lc = o.y - pd.y0;
      } else if (pl == kSideRight) {
        lx = 1.0; ly = 0.0; lc = pd.x1 - o.x;
      } else if (pl == kSideTop) {
        lx = 0.0; ly = 1.0; lc = pd.y1 - o.y;
      } else if (pl == kSideLeft) {
        lx = -1.0; ly = 0.0; lc = o.x - pd.x0;
      } else {
        lx = ly = lc = 0.0;
      }
      const int pv = ln > 0 ? ln - 1 : nv - 1;
      const double ax = __shfl(lx, pv, 64), ay = __shfl(ly, pv, 64), ac = __shfl(lc, pv, 64);
      const double mnx = wave_min(on ? vx : INFINITY), mxx = wave_max(on ? vx : -INFINITY);
      const double mny = wave_min(on ? vy : INFINITY), mxy = wave_max(on ? vy : -INFINITY);
      const double ext2 = (mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny);
      bool bad = false;
      const double det = ax * ly - ay * lx;
      const double sc = sqrt((ax * ax + ay * ay) * (lx * lx + ly * ly));
      if (!(fabs(det) > 1e-7 * sc)) {
        bad = true;
      } else {
        const double rx = (ac * ly - lc * ay) / det, ry = (ax * lc - lx * ac) / det;
        const double ex = rx - vx, ey = ry - vy;
        if (ex * ex + ey * ey > 1e-6 * ext2) {
          bad = true;
        } else {
          vx = rx;
          vy = ry;
        }
      }
      if (__ballot(on && bad) != 0ull) coarse = true;
    }
    const int nx = ln + 1 < nv ? ln + 1 : 0;
    const double qx = __shfl(vx, nx, 64), qy = __shfl(vy, nx, 64);
    const double cr = vx * qy - qx * vy;
    const double sx = vx + qx, sy = vy + qy;
    // sums in edge order, each term's product fused into the running sum as in cell_site
    double A2 = 0.0, gx = 0.0, gy = 0.0;
    for (int k = 0; k < nv; ++k) {
      const double crk = rdl(cr, k);
      A2 += crk;
      gx += rdl(sx, k) * crk;
      gy += rdl(sy, k) * crk;
    }
    const double mnu = wave_min(on ? vx : INFINITY), mnv = wave_min(on ? vy : INFINITY);
    if (on) {
      float2 span = make_float2(0.f, 0.f);
      if (pl >= 0) {
        const int lo = min(i, pl), hi = max(i, pl);
        const double2 sl = a.xy[lo], sh = a.xy[hi];
        const double Dx = sh.x - sl.x, Dy = sh.y - sl.y;
        const double inv = 1.0 / sqrt(Dx * Dx + Dy * Dy);
        const double ex = -Dy * inv, ey = Dx * inv;
        const double mx = 0.5 * (sl.x + sh.x), my = 0.5 * (sl.y + sh.y);
        const double t0 = (o.x + vx - mx) * ex + (o.y + vy - my) * ey;
        const double t1 = (o.x + qx - mx) * ex + (o.y + qy - my) * ey;
        span = make_float2((float)fmin(t0, t1), (float)fmax(t0, t1));
      }
      a.elab[(size_t)i * kMaxVerts + ln] = pl;
      a.espan[(size_t)i * kMaxVerts + ln] = span;
      a.vtx[(size_t)i * kMaxVerts + ln] = make_double2(o.x + vx, o.y + vy);
    }
    ar = 0.5 * A2;
    geo = A2 != 0.0 ? make_double2(o.x + gx / (3.0 * A2), o.y + gy / (3.0 * A2)) : o;
    const int kox = (int)floor((o.x + mnu - pd.x0) / pd.dx);
    const int koy = (int)floor((o.y + mnv - pd.y0) / pd.dy);
    kb = make_int2(min(max(kox, 0), pd.pm - 1), min(max(koy, 0), pd.pn - 1));
    if (kF32 && ln == 0) a.org[i] = geo;  // next frame: this centroid
    if (kF32 && fabs(ar) < kSliverRel * pd.bsize * pd.bsize) coarse = true;
  }
  // a full pass whose polygon overflowed V skipped a constraint: the repair pass rebuilds the
  // cell in its larger polygon
  const bool handoff = overflow && V < kMaxVerts && !a.redo_mode && a.coarse;
  if (handoff) overflow = false;
  if (ln == 0) {
    a.kbox[i] = kb;
    if (((kF32 && coarse) || handoff) && a.coarse) {
      a.redo[i] = 2;  // the check that follows lists it
      atomicAdd(&a.coarse[bd.prob], 1);
      if (a.diag) count_add(&a.diag[kF32 && coarse ? kDiagDegenerate : kDiagOverflow]);
    }
    a.area[i] = ar;
    a.geo[i] = geo;
    a.exist[i] = nv >= 3 ? 1 : 0;
  }
  return overflow ? 1 : 0;
}

// Full construction pass with one wave (block) per site: block b builds site b % kCellBlock of
// cell block b / kCellBlock.
template <typename T>
__global__ __launch_bounds__(64) void cell_wave_kernel(CellArgs a) {
  __shared__ WaveLDS<T> W;
  reset_check_list(a);
  const unsigned long long t0 = timer_begin();
  const int slot = xcd_slot(blockIdx.x, gridDim.x);
  const int blk = slot / kCellBlock;
  const int rc = cell_site_wave<T, kPassVerts>(a, blk, slot % kCellBlock, W);
  if (rc == 1 && threadIdx.x == 0) atomicAdd(&a.overflow[a.blocks[blk].prob], 1);
  timer_mark(a.timer, rc >= 0, t0);
}

// The sites the lane full pass deferred (CellArgs::defer_list), one wave (block) per listed
// site; any grid size.  Folds into the construction pass's timer.
template <typename T>
__global__ __launch_bounds__(64) void cell_defer_kernel(CellArgs a) {
  __shared__ WaveLDS<T> W;
  const unsigned long long t0 = timer_begin();
  const int n = *a.defer_count;
  bool worked = false;
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int slot = a.defer_list[e];
    const int blk = slot / kCellBlock;
    const int rc = cell_site_wave<T, kPassVerts>(a, blk, slot % kCellBlock, W);
    if (rc == 1 && threadIdx.x == 0) atomicAdd(&a.overflow[a.blocks[blk].prob], 1);
    worked = worked || rc >= 0;
  }
  timer_mark(a.timer, worked, t0);
}

// The cells the last evaluation left empty (CellArgs::empty_list), built at the new weights
// before the full pass, one wave (block) per listed site; any grid size.  A cell that empties
// and comes back reappears between sites whose adjacency (the previous cells) does not hold it:
// each would miss the new edge, fail the conformity check and need a repair.  A listed cell that
// comes back is queued as a candidate of each of its new neighbours (as the conformity check
// queues a missed candidate), which the full pass then takes with their adjacency.  The listed
// cells are built as the full pass would build them (no adjacency: the bucket rings), except
// that the candidates queued for them are not taken (other listed cells queue themselves here
// while they are built); the full pass skips them (CellArgs::prebuilt).  Folds into the
// construction pass's timer.
template <typename T>
__global__ __launch_bounds__(64) void cell_reappear_kernel(CellArgs a, int* xcand) {
  __shared__ WaveLDS<T> W;
  const unsigned long long t0 = timer_begin();
  int* const xcnt = a.xcnt;
  a.xcnt = nullptr;
  a.defer_list = nullptr;
  a.prepass = 1;
  const int n = min(*a.empty_count, a.empty_cap);
  bool worked = false;
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int slot = a.empty_list[e];
    const int blk = slot / kCellBlock;
    const int rc = cell_site_wave<T, kPassVerts>(a, blk, slot % kCellBlock, W);
    __syncthreads();  // the cell's edge labels, written by their lanes
    if (rc < 0) continue;
    if (rc == 1 && threadIdx.x == 0) atomicAdd(&a.overflow[a.blocks[blk].prob], 1);
    worked = true;
    const int i = a.blocks[blk].first + slot % kCellBlock;
    const int ne = a.nedge[i];
    const int ln = lane_id();
    if (ln == 0) a.prebuilt[i] = 1;
    if (ln < ne) {
      const int j = a.elab[(size_t)i * kMaxVerts + ln];
      if (j >= 0) {
        const int q = atomicAdd(&xcnt[j], 1);
        if (q < kMaxExtra) xcand[(size_t)j * kMaxExtra + q] = i;
      }
    }
    if (ne > 0 && ln == 0 && a.diag) count_add(&a.diag[kDiagReappeared]);
  }
  timer_mark(a.timer, worked, t0);  // part of the construction pass
}

// Repair construction (FP64) with one wave (block) per listed site; any grid size.
__global__ __launch_bounds__(64) void cell_repair_wave_kernel(CellArgs a) {
  __shared__ WaveLDS<double> W;
  reset_check_list(a);
  const unsigned long long t0 = timer_begin();
  const int n = *a.list_count;
  bool worked = false;
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int slot = a.list[e];
    const int blk = slot / kCellBlock;
    const int rc = cell_site_wave<double, kMaxVerts>(a, blk, slot % kCellBlock, W);
    if (rc == 1 && threadIdx.x == 0) atomicAdd(&a.overflow[a.blocks[blk].prob], 1);
    worked = worked || rc >= 0;
  }
  timer_mark(a.timer, worked, t0);
}

// Integration pass.  Grid (cell blocks, kCellBlock * L / kIntBlock), block kIntBlock: L
// consecutive lanes per site; the lanes' moments are summed with shuffles within their aligned
// group.
template <typename T, int L>
__device__ __forceinline__ void integrate_block(const CellArgs& a, int blk, int quarter) {
  const int t = threadIdx.x;
  const int i = integ_site<L>(a, blk, quarter, t);
  const int prob = a.blocks[blk].prob;
  Moments M;
  if (i >= 0) integrate_lane<T, L>(a, a.probs[prob], i, t % L, M);
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) {
    Moments o;
    o.m0.s = __shfl_xor(M.m0.s, off, 64);
    o.m0.c = __shfl_xor(M.m0.c, off, 64);
    o.m1u = __shfl_xor(M.m1u, off, 64);
    o.m1v = __shfl_xor(M.m1v, off, 64);
    o.m2u = __shfl_xor(M.m2u, off, 64);
    o.m2v = __shfl_xor(M.m2v, off, 64);
    combine(M, o);
  }
  if (i >= 0 && t % L == 0) finish_site(a, a.probs[prob], prob, i, M);
}

// Integration pass: clears the lengths of the lists the construction consumed (the deferred
// sites; the cells the last evaluation left empty).
__device__ __forceinline__ void reset_list(const CellArgs& a) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || blockIdx.y != 0) return;
  if (a.defer_count && !a.redo_mode) *a.defer_count = 0;
  if (a.empty_count && !a.redo_mode) *a.empty_count = 0;
}

__host__ __device__ int verify_site(const VerifyArgs& a, int blk, int ln);

// The conformity check of cell blocks g * (kIntBlock / kCellBlock) ... in a kIntBlock-thread
// workgroup of an integration launch (one thread per site).
__device__ void check_group(const VerifyArgs& v, int g, int nblk) {
  const int cb = g * (kIntBlock / kCellBlock) + (int)threadIdx.x / kCellBlock;
  if (cb >= nblk) return;
  const int rc = verify_site(v, cb, (int)threadIdx.x % kCellBlock);
  if (rc == 1) {
    atomicAdd(&v.fails[v.blocks[cb].prob], 1);
    if (v.diag_fail) count_add(v.diag_fail);
  }
}

// Integration pass (1-D grid).  The last `vgroups` workgroups run the conformity check `v` of
// the cells just built (check_group; nblk cell blocks): it reads only the construction's output,
// as the integration does, so it runs in the same launch, in the integration's tail (the
// workgroups are dispatched in order: earlier, they would hold compute units the integration's
// workgroups wait for).  The others integrate: workgroup (blk, quarter) in XCD order, the
// quarters of a cell block on one XCD.  a.timer: the construction pass's timer, closed here;
// a.timer_int: this pass's, closed by its last block.
template <typename T, int L>
__global__ __launch_bounds__(kIntBlock) void integrate_kernel(CellArgs a, VerifyArgs v,
                                                              int vgroups, int nblk) {
  timer_close(a.timer);
  reset_list(a);
  const unsigned long long t0 = timer_begin();
  bool worked = false;
  const int ni = (int)gridDim.x - vgroups;
  if ((int)blockIdx.x >= ni) {
    check_group(v, (int)blockIdx.x - ni, nblk);
  } else {
    constexpr int kQuarters = kCellBlock * L / kIntBlock;
    const int slot = xcd_slot((int)blockIdx.x, ni);
    const int blk = slot / kQuarters, quarter = slot % kQuarters;
    worked = integ_site<L>(a, blk, quarter, threadIdx.x) >= 0;
    integrate_block<T, L>(a, blk, quarter);
  }
  timer_mark(a.timer_int, worked, t0);
  timer_finish(a.timer_int);
}

// Repair integration over the work list: L lanes per listed site, any grid size; the last
// `vgroups` workgroups run the check of the round (as integrate_kernel).
template <typename T, int L>
__global__ __launch_bounds__(kIntBlock) void integrate_repair_kernel(CellArgs a, VerifyArgs v,
                                                                     int vgroups, int nblk) {
  timer_close(a.timer);
  reset_list(a);
  const unsigned long long t0 = timer_begin();
  bool worked = false;
  const int nb = (int)gridDim.x - vgroups, b = (int)blockIdx.x;
  if (b >= nb) {
    check_group(v, b - nb, nblk);
    timer_mark(a.timer_int, false, t0);
    timer_finish(a.timer_int);
    return;
  }
  const int n = *a.list_count;
  constexpr int kSites = kIntBlock / L;
  const int t = threadIdx.x;
  // entries dealt round-robin over the blocks (as in cell_repair_kernel)
  for (int base = 0; base < n; base += nb * kSites) {
    const int e = base + (t / L) * nb + b;
    int i = -1, prob = 0;
    if (e < n) {
      const int slot = a.list[e];
      const BlockDesc bd = a.blocks[slot / kCellBlock];
      i = bd.first + slot % kCellBlock;
      prob = bd.prob;
    }
    Moments M;
    if (i >= 0) integrate_lane<T, L>(a, a.probs[prob], i, t % L, M);
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
      Moments o;
      o.m0.s = __shfl_xor(M.m0.s, off, 64);
      o.m0.c = __shfl_xor(M.m0.c, off, 64);
      o.m1u = __shfl_xor(M.m1u, off, 64);
      o.m1v = __shfl_xor(M.m1v, off, 64);
      o.m2u = __shfl_xor(M.m2u, off, 64);
      o.m2v = __shfl_xor(M.m2v, off, 64);
      combine(M, o);
    }
    if (i >= 0 && t % L == 0) finish_site(a, a.probs[prob], prob, i, M);
    worked = worked || i >= 0;
  }
  timer_mark(a.timer_int, worked, t0);
  timer_finish(a.timer_int);
}

// ---------------------------------------------------------------------------------------------
// conformity check
// ---------------------------------------------------------------------------------------------

// One site of the conformity check; returns -1 when the lane has no site to check, 1 when one
// of its edges is unmatched, 0 otherwise.
__host__ __device__ int verify_site(const VerifyArgs& a, int blk, int ln) {
  const BlockDesc bd = a.blocks[blk];
  if (ln >= bd.count) return -1;
  if (!a.active[bd.prob]) return -1;
  if (a.gate && a.gate[bd.prob] == 0) return -1;
  const int i = bd.first + ln;
  const int ne = a.nedge[i];
  const double2 si = a.xy[i];
  bool bad = false;
  // flag site k for a repair; the first to flag it appends it to the work list
  const ProbDesc& pd = a.probs[bd.prob];
  auto mark = [&](int k) {
    if (exchange(&a.redo[k], 1) == 1 || !a.list) return;
    const int local = k - pd.site_off;
    const int slot = (pd.blk_off + local / kCellBlock) * kCellBlock + local % kCellBlock;
    a.list[fetch_add(a.list_count, 1)] = slot;
  };
  if (a.redo[i] == 2) mark(i);  // an FP32 cell handed to the FP64 repair
  // edges kVerifyBatch at a time: their labels and spans, then the neighbours' sites and edge
  // counts, each round's loads in flight together; then, edge by edge in order, the neighbour's
  // labels kVerifyBatch at a time until the reciprocal edge turns up
  constexpr int B = kVerifyBatch;
  for (int k0 = 0; k0 < ne; k0 += B) {
    int jj[B], nj[B];
    float2 sp[B];
    double2 sj[B];
    for (int u = 0; u < B; ++u) {
      const bool in = k0 + u < ne;
      jj[u] = in ? a.elab[(size_t)i * kMaxVerts + k0 + u] : -1;
      sp[u] = in ? a.espan[(size_t)i * kMaxVerts + k0 + u] : make_float2(0.f, 0.f);
    }
    for (int u = 0; u < B; ++u) {
      const int j = jj[u];
      sj[u] = j >= 0 ? a.xy[j] : make_double2(0.0, 0.0);
      nj[u] = j >= 0 ? (int)a.nedge[j] : 0;
    }
    for (int u = 0; u < B; ++u) {
      const int j = jj[u];
      if (j < 0) continue;
      const double ddx = sj[u].x - si.x, ddy = sj[u].y - si.y;
      const double tol = a.tol_rel * sqrt(ddx * ddx + ddy * ddy);
      const double len = (double)sp[u].y - (double)sp[u].x;
      bool found = false, match = false;
      for (int q0 = 0; q0 < nj[u] && !found; q0 += B) {
        int lab[B];
        for (int v = 0; v < B; ++v)
          lab[v] = q0 + v < nj[u] ? a.elab[(size_t)j * kMaxVerts + q0 + v] : -1;
        int at = -1;
        for (int v = B - 1; v >= 0; --v)
          if (lab[v] == i) at = q0 + v;  // the first occurrence
        if (at >= 0) {
          const float2 sq = a.espan[(size_t)j * kMaxVerts + at];
          found = true;
          match = fabs((double)sp[u].x - (double)sq.x) + fabs((double)sp[u].y - (double)sq.y) <=
                  tol;
        }
      }
      if (found ? !match : (len > tol)) {
        bad = true;
        mark(j);
        if (!found && a.xcnt) {
          // j's cell lacks the edge: i is a candidate j missed
          const int q = fetch_add(&a.xcnt[j], 1);
          if (q < kMaxExtra) a.xcand[(size_t)j * kMaxExtra + q] = i;
        }
      }
    }
  }
  if (bad) mark(i);
  return bad ? 1 : 0;
}

// Keep the cells' adjacency as the candidates of the next pass (the next evaluation, or a
// repair).  gate: skip problems whose counter is zero (nullptr: none).  elist set (the
// evaluation's last save; problems in `active` only): also list the empty cells (site slots,
// at most ecap) for the next full pass's cell_reappear_kernel.
__host__ __device__ void save_adjacency_site(const BlockDesc& bd, int blk, int ln,
                                             const uint8_t* nedge, const int* elab, uint8_t* adjn,
                                             int* adj, int* elist, int* ecount, int ecap) {
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  const int ne = nedge[i];
  adjn[i] = (uint8_t)ne;
  const int* src = elab + (size_t)i * kMaxVerts;
  int* dst = adj + (size_t)i * kMaxVerts;
  for (int k = 0; k < ne; ++k) dst[k] = src[k];
  if (ne == 0 && elist) {
    const int q = fetch_add(ecount, 1);
    if (q < ecap) elist[q] = blk * kCellBlock + ln;
  }
}

__global__ void save_adjacency_kernel(const BlockDesc* blocks, const int* gate,
                                      const uint8_t* active, const uint8_t* nedge,
                                      const int* elab, uint8_t* adjn, int* adj, int* elist,
                                      int* ecount, int ecap) {
  const BlockDesc bd = blocks[blockIdx.x];
  if (gate && gate[bd.prob] == 0) return;
  if (elist && !active[bd.prob]) return;
  save_adjacency_site(bd, blockIdx.x, threadIdx.x, nedge, elab, adjn, adj, elist, ecount, ecap);
}

// Grow the ring of every cell marked for a rebuild (r -> 2r + 1).
__global__ void raise_ring_kernel(const BlockDesc* blocks, const int* list, const int* list_count,
                                  uint8_t* ring, int ring_max) {
  const int n = *list_count;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int slot = list[e];
    const int i = blocks[slot / kCellBlock].first + slot % kCellBlock;
    ring[i] = (uint8_t)min(2 * (int)ring[i] + 1, ring_max);
  }
}

// Next evaluation's ring: the outermost ring that shaped the cell plus one, at least ring_min.
__global__ void adapt_ring_kernel(const BlockDesc* blocks, const uint8_t* used, uint8_t* ring,
                                  int ring_min, int ring_max) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  ring[i] = (uint8_t)min(max((int)used[i] + 1, ring_min), ring_max);
}

// ---------------------------------------------------------------------------------------------
// reductions and L-BFGS vector kernels (one block per cell block; per-block partials)
// ---------------------------------------------------------------------------------------------

// Partials per block: [0] sum f_i, [1] sum mass, [2] sum area, [3] max |mass-nu|/nu,
// [4] count(mass == 0), [5] sum (mass - nu)^2, [6] sum (mass - nu) * dir (directional
// derivative at the trial point, dir may be null), [7] count(empty cells).
constexpr int kEvalPartials = 8;

__global__ __launch_bounds__(kCellBlock) void reduce_eval_kernel(
    const BlockDesc* blocks, const double* mass, const double* fcell, const double* area,
    const double* nu, const uint8_t* exist, const double* dir, double* part) {
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  double f = 0, m = 0, ar = 0, mr = 0, z = 0, g2 = 0, gd = 0, em = 0;
  if (ln < bd.count) {
    const int i = bd.first + ln;
    const double g = mass[i] - nu[i];
    f = fcell[i];
    m = mass[i];
    ar = area[i];
    mr = fabs(g) / nu[i];
    z = mass[i] == 0.0 ? 1.0 : 0.0;
    g2 = g * g;
    gd = dir ? g * dir[i] : 0.0;
    em = exist[i] ? 0.0 : 1.0;
  }
  double* out = part + (size_t)blockIdx.x * kEvalPartials;
  double v;
  v = block_sum(f, red);   if (ln == 0) out[0] = v;
  v = block_sum(m, red);   if (ln == 0) out[1] = v;
  v = block_sum(ar, red);  if (ln == 0) out[2] = v;
  v = block_max(mr, red);  if (ln == 0) out[3] = v;
  v = block_sum(z, red);   if (ln == 0) out[4] = v;
  v = block_sum(g2, red);  if (ln == 0) out[5] = v;
  v = block_sum(gd, red);  if (ln == 0) out[6] = v;
  v = block_sum(em, red);  if (ln == 0) out[7] = v;
}

// g = mass - nu
__global__ void gradient_kernel(const BlockDesc* blocks, const double* mass, const double* nu,
                                double* g) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  g[i] = mass[i] - nu[i];
}

// Direction d = -(cg * g + sum_h cy[h] * Y[h] + sum_h cs[h] * S[h]) for active problems, with
// per-block partial g.d.  coef layout per problem: [cg, cy[kMaxHist], cs[kMaxHist]].
__global__ __launch_bounds__(kCellBlock) void direction_kernel(
    const BlockDesc* blocks, const uint8_t* active, const double* coef, const int* hcount,
    const double* S, const double* Y, size_t ld, const double* g, double* d, double* part) {
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  if (!active[bd.prob]) return;  // uniform per block
  const int ln = threadIdx.x;
  const double* cf = coef + (size_t)bd.prob * (1 + 2 * kMaxHist);
  const int H = hcount[bd.prob];
  double gd = 0.0;
  if (ln < bd.count) {
    const int i = bd.first + ln;
    double acc = cf[0] * g[i];
    for (int h = 0; h < H; ++h) acc += cf[1 + h] * Y[h * ld + i] + cf[1 + kMaxHist + h] * S[h * ld + i];
    const double di = -acc;
    d[i] = di;
    gd = g[i] * di;
  }
  const double v = block_sum(gd, red);
  if (ln == 0) part[blockIdx.x] = v;
}

// Accept a line-search trial for active problems: with g_new = mass - nu,
//   s = t * d, y = g_new - g, stored in history slot `slot`; w += s; g = g_new.
// Per-block partials (layout per block, width 5 * kMaxHist):
//   [h]               s_slot . y_h
//   [kMaxHist + h]    s_h . y_slot
//   [2kMaxHist + h]   y_slot . y_h
//   [3kMaxHist + h]   g_new . s_h
//   [4kMaxHist + h]   g_new . y_h
// for h < hcount (the count after the insertion).
constexpr int kAcceptPartials = 5 * kMaxHist;

__global__ __launch_bounds__(kCellBlock) void accept_kernel(
    const BlockDesc* blocks, const uint8_t* active, const double* step, const int* slot,
    const int* hcount, const double* mass, const double* nu, const double* d, double* w,
    double* g, double* S, double* Y, size_t ld, double* part) {
  __shared__ double red[kCellBlock / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  if (!active[bd.prob]) return;
  const int ln = threadIdx.x;
  const int sl = slot[bd.prob], H = hcount[bd.prob];
  double sn = 0.0, yn = 0.0, gn = 0.0;
  const bool live = ln < bd.count;
  const int i = bd.first + ln;
  if (live) {
    const double t = step[bd.prob];
    gn = mass[i] - nu[i];
    sn = t * d[i];
    yn = gn - g[i];
    S[sl * ld + i] = sn;
    Y[sl * ld + i] = yn;
    w[i] = w[i] + t * d[i];
    g[i] = gn;
  }
  double* out = part + (size_t)blockIdx.x * kAcceptPartials;
  for (int h = 0; h < H; ++h) {
    double sh = 0.0, yh = 0.0;
    if (live) {
      sh = S[h * ld + i];
      yh = Y[h * ld + i];
    }
    double v;
    v = block_sum(sn * yh, red); if (ln == 0) out[h] = v;
    v = block_sum(sh * yn, red); if (ln == 0) out[kMaxHist + h] = v;
    v = block_sum(yn * yh, red); if (ln == 0) out[2 * kMaxHist + h] = v;
    v = block_sum(gn * sh, red); if (ln == 0) out[3 * kMaxHist + h] = v;
    v = block_sum(gn * yh, red); if (ln == 0) out[4 * kMaxHist + h] = v;
  }
}

// Fill a byte array (graph-safe replacement for memset).
__global__ void fill_u8_kernel(uint8_t* p, size_t n, uint8_t v) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}
// Caller order <-> the library's bucket order on the device (perm: sorted -> caller index),
// for the entry points on device arrays.
__global__ void gather_f64_kernel(const double* src, const int* perm, size_t n, double* dst) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[perm[r]];
}
__global__ void scatter_f64_kernel(const double* src, const int* perm, size_t n, double* dst) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[perm[r]] = src[r];
}
__global__ void scatter_f64x2_kernel(const double2* src, const int* perm, size_t n, double* dst) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const size_t o = (size_t)perm[r];
  dst[2 * o] = src[r].x;
  dst[2 * o + 1] = src[r].y;
}
__global__ void scatter_u8_kernel(const uint8_t* src, const int* perm, size_t n, uint8_t* dst) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[perm[r]] = src[r];
}

// Density groups and chunk totals of one target image (the layout of cot_internal.h), one thread
// per (row, chunk): the sums of build_prefix (cot_layout.h, the host build), in the same order.
// FP32 geometry reads the image rounded to float, as its pieces use it.
__global__ void prefix_kernel(const double* rho, int pm, int pn, int nchunk, int ngrp, int gsize,
                              DensityGroup* grp, double4* tot) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)pn * nchunk) return;
  const int iy = (int)(t / nchunk), c = (int)(t % nchunk);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int k0 = c * kChunk, k1 = min(k0 + kChunk, pm);
  for (int g0 = k0; g0 < k1; g0 += gsize) {
    DensityGroup* rec = grp + (size_t)iy * ngrp + g0 / gsize;
    rec->s0 = s0;
    rec->s1 = s1;
    rec->s2 = s2;
    for (int u = 0; u < gsize; ++u) {
      const int k = g0 + u;
      double r = k < pm ? rho[(size_t)iy * pm + k] : 0.0;
      if (gsize == kGroup32) {
        r = (double)(float)r;
        rec->r.f[u] = (float)r;
      } else {
        rec->r.d = r;
      }
      if (k < k1) prefix_step(s0, s1, s2, r, (double)(k - k0));
    }
  }
  tot[(size_t)iy * nchunk + c] = make_double4(s0, s1, s2, 0.0);
}

__global__ void fill_i32_kernel(int* p, size_t n, int v) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------

// Integration pass with L lanes per site: the full grid, or the repair grid over the work list.
template <int L>
void launch_integrate(const CellArgs& a, int nblocks, bool fp32, int repair_blocks,
                      const VerifyArgs* check, hipStream_t st) {
  // check workgroups: kIntBlock / kCellBlock cell blocks each
  constexpr int kPer = kIntBlock / kCellBlock;
  const int vgroups = check ? (nblocks + kPer - 1) / kPer : 0;
  const VerifyArgs v = check ? *check : VerifyArgs{};
  const int ig = vgroups + nblocks * (kCellBlock * L / kIntBlock);
  if (a.redo_mode) {
    hipLaunchKernelGGL((integrate_repair_kernel<double, kRepairLanes>),
                       dim3(vgroups + repair_blocks), dim3(kIntBlock), 0, st, a, v, vgroups,
                       nblocks);
  } else if (fp32) {
    hipLaunchKernelGGL((integrate_kernel<float, L>), dim3(ig), dim3(kIntBlock), 0, st, a, v,
                       vgroups, nblocks);
  } else {
    hipLaunchKernelGGL((integrate_kernel<double, L>), dim3(ig), dim3(kIntBlock), 0, st, a, v,
                       vgroups, nblocks);
  }
}

// Construction (one lane per site) then integration (int_lanes lanes per site: 4, 8 or 16) with
// the conformity check `check` (nullptr: none) of the cells built.  The integration pass closes
// a.timer and a.timer_int.
hipError_t launch_cell(const CellArgs& a, int nblocks, bool fp32, int int_lanes,
                       int repair_blocks, const VerifyArgs* check, hipStream_t st) {
  // waves of the passes that walk a list of sites: at most kDeferWaves
  const int list_waves = nblocks * kCellBlock < kDeferWaves ? nblocks * kCellBlock : kDeferWaves;
  if (a.redo_mode) {
    // repairs rebuild few cells: always in FP64 (with FP32 geometry, this is also the rebuild
    // of the cells the FP32 pass could not resolve), on a small grid over the work list: one
    // wave per site (4 waves per repair block's worth of lanes), or one lane per site
    if (a.repair_lane)
      hipLaunchKernelGGL((cell_repair_kernel<double>), dim3(repair_blocks), dim3(kCellBlock), 0,
                         st, a);
    else
      hipLaunchKernelGGL(cell_repair_wave_kernel, dim3(4 * repair_blocks), dim3(64), 0, st, a);
  } else {
    if (a.empty_list && a.prebuilt && a.cand != kCandBucket && use_prepass(nblocks)) {
      // the cells the last evaluation left empty, before the full pass builds their neighbours
      int* xcand = const_cast<int*>(a.xcand);
      if (fp32)
        hipLaunchKernelGGL((cell_reappear_kernel<float>), dim3(list_waves), dim3(64), 0, st, a,
                           xcand);
      else
        hipLaunchKernelGGL((cell_reappear_kernel<double>), dim3(list_waves), dim3(64), 0, st, a,
                           xcand);
    }
    if (a.pass_wave) {
      // one wave per site: many more waves in flight than sites per lane, for small problems
      const dim3 g((unsigned)nblocks * kCellBlock);
      if (fp32)
        hipLaunchKernelGGL((cell_wave_kernel<float>), g, dim3(64), 0, st, a);
      else
        hipLaunchKernelGGL((cell_wave_kernel<double>), g, dim3(64), 0, st, a);
    } else {
      if (fp32)
        hipLaunchKernelGGL((cell_kernel<float>), dim3(nblocks), dim3(kCellBlock), 0, st, a);
      else
        hipLaunchKernelGGL((cell_kernel<double>), dim3(nblocks), dim3(kCellBlock), 0, st, a);
      if (a.defer_list) {
        // the deferred sites: a grid of at most kDeferWaves waves walks the list
        if (fp32)
          hipLaunchKernelGGL((cell_defer_kernel<float>), dim3(list_waves), dim3(64), 0, st, a);
        else
          hipLaunchKernelGGL((cell_defer_kernel<double>), dim3(list_waves), dim3(64), 0, st, a);
      }
    }
  }
  switch (int_lanes) {
    case 4: launch_integrate<4>(a, nblocks, fp32, repair_blocks, check, st); break;
    case 16: launch_integrate<16>(a, nblocks, fp32, repair_blocks, check, st); break;
    default: launch_integrate<8>(a, nblocks, fp32, repair_blocks, check, st); break;
  }
  return hipGetLastError();
}

hipError_t launch_prefix(const double* rho, int pm, int pn, bool fp32, DensityGroup* grp,
                         double4* tot, hipStream_t st) {
  const int nchunk = (pm + kChunk - 1) / kChunk;
  const int gsize = fp32 ? kGroup32 : kGroup64;
  const int ngrp = (pm + gsize - 1) / gsize;
  const long long n = (long long)pn * nchunk;
  hipLaunchKernelGGL(prefix_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rho, pm,
                     pn, nchunk, ngrp, gsize, grp, tot);
  return hipGetLastError();
}


hipError_t launch_save_adjacency(const BlockDesc* b, const int* gate, const uint8_t* active,
                                 const uint8_t* nedge, const int* elab, uint8_t* adjn, int* adj,
                                 int* elist, int* ecount, int ecap, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(save_adjacency_kernel, dim3(nblocks), dim3(kCellBlock), 0, st, b, gate,
                     active, nedge, elab, adjn, adj, elist, ecount, ecap);
  return hipGetLastError();
}

// Grow the ring (r -> 2r + 1) of every site on the work list.
hipError_t launch_raise_ring(const BlockDesc* b, const int* list, const int* list_count,
                             uint8_t* ring, int rmax, int repair_blocks, hipStream_t st) {
  hipLaunchKernelGGL(raise_ring_kernel, dim3(repair_blocks), dim3(kCellBlock), 0, st, b, list,
                     list_count, ring, rmax);
  return hipGetLastError();
}

hipError_t launch_adapt_ring(const BlockDesc* b, const uint8_t* used, uint8_t* ring, int rmin,
                             int rmax, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(adapt_ring_kernel, dim3(nblocks), dim3(kCellBlock), 0, st, b, used, ring,
                     rmin, rmax);
  return hipGetLastError();
}

hipError_t launch_reduce_eval(const BlockDesc* b, const double* mass, const double* fcell,
                              const double* area, const double* nu, const uint8_t* exist,
                              const double* dir, double* part, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(reduce_eval_kernel, dim3(nblocks), dim3(kCellBlock), 0, st, b, mass, fcell,
                     area, nu, exist, dir, part);
  return hipGetLastError();
}

hipError_t launch_gradient(const BlockDesc* b, const double* mass, const double* nu, double* g,
                           int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(gradient_kernel, dim3(nblocks), dim3(kCellBlock), 0, st, b, mass, nu, g);
  return hipGetLastError();
}

hipError_t launch_direction(const BlockDesc* b, const uint8_t* active, const double* coef,
                            const int* hcount, const double* S, const double* Y, size_t ld,
                            const double* g, double* d, double* part, int nblocks,
                            hipStream_t st) {
  hipLaunchKernelGGL(direction_kernel, dim3(nblocks), dim3(kCellBlock), 0, st, b, active, coef,
                     hcount, S, Y, ld, g, d, part);
  return hipGetLastError();
}

hipError_t launch_accept(const BlockDesc* b, const uint8_t* active, const double* step,
                         const int* slot, const int* hcount, const double* mass, const double* nu,
                         const double* d, double* w, double* g, double* S, double* Y, size_t ld,
                         double* part, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(accept_kernel, dim3(nblocks), dim3(kCellBlock), 0, st, b, active, step, slot,
                     hcount, mass, nu, d, w, g, S, Y, ld, part);
  return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u8_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}
hipError_t launch_permute(int kind, const void* src, const int* perm, size_t n, void* dst,
                          hipStream_t st) {
  if (n == 0) return hipSuccess;
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
  switch (kind) {
    case 0:
      hipLaunchKernelGGL(gather_f64_kernel, g, b, 0, st, static_cast<const double*>(src), perm, n,
                         static_cast<double*>(dst));
      break;
    case 1:
      hipLaunchKernelGGL(scatter_f64_kernel, g, b, 0, st, static_cast<const double*>(src), perm, n,
                         static_cast<double*>(dst));
      break;
    case 2:
      hipLaunchKernelGGL(scatter_f64x2_kernel, g, b, 0, st, static_cast<const double2*>(src), perm,
                         n, static_cast<double*>(dst));
      break;
    default:
      hipLaunchKernelGGL(scatter_u8_kernel, g, b, 0, st, static_cast<const uint8_t*>(src), perm, n,
                         static_cast<uint8_t*>(dst));
      break;
  }
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, size_t n, int v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n,
                     v);
  return hipGetLastError();
}

}  // namespace cot
