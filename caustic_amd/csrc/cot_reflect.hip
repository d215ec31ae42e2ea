// Kernels of the reflection-aware solve (ot_time >= 1; aarl_direct_ot, ot_op.cuh:591-1385):
// per-site Fresnel transmittance of the lens at the current cell centroids, the source masses it
// weights, and the Anderson-acceleration history of the fixed-point map G(w) = w + beta * g(w).
#include "cot_internal.h"

namespace cot {

namespace {

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 unit(V3 v) {
  const double n = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  if (n < 1e-12) return V3{0.0, 0.0, 0.0};
  return V3{v.x / n, v.y / n, v.z / n};
}
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Unpolarised Fresnel transmittance 1 - (Rs + Rp)/2 for a ray crossing from index ei to et
// (basic_function.cuh:407-435): cosines clamped to [0, 1], denominators kept away from zero.
__device__ double transmittance(V3 wi, V3 wt, V3 n, double ei, double et) {
  double ci = fmin(1.0, fmax(0.0, fabs(dot3(wi, n))));
  double ct = fmin(1.0, fmax(0.0, fabs(dot3(wt, n))));
  double ds = ei * ci + et * ct, dp = ei * ct + et * ci;
  if (fabs(ds) < 1e-12) ds = 1e-12;
  if (fabs(dp) < 1e-12) dp = 1e-12;
  const double rs = (ei * ci - et * ct) / ds, rp = (ei * ct - et * ci) / dp;
  return 1.0 - 0.5 * (rs * rs + rp * rp);
}

__device__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace

// weight_i = clip_area_i * T1_i * T2_i with the receiver point at the cell's geometric centroid
// (compute_energy_remain_rate, basic_function.cuh:437-494).  The light reaching point i travels
// along wdir[3i..3i+2] (a point light, main.cu:218-231), or along +z when wdir is null (parallel
// light, main.cu:211-217).  Per-block partial sums of the weights go to part[block].
__global__ void fresnel_weight_kernel(const BlockDesc* blocks, const double2* points,
                                      const double* ins, const double* wdir, const double* height,
                                      const double* clip_area, const double2* centroid,
                                      double z_len, double z_recv, double n_idx, double* weight,
                                      double* part) {
  __shared__ double red[4];
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  double wv = 0.0;
  if (ln < bd.count) {
    const int i = bd.first + ln;
    const double2 p = points[i];
    const V3 d = unit(V3{ins[3 * i], ins[3 * i + 1], ins[3 * i + 2]});
    const double h = height[i];
    const V3 outp{p.x + h * d.x, p.y + h * d.y, z_len + h * d.z};
    const double2 c = centroid[i];
    const V3 wt = unit(V3{c.x - outp.x, c.y - outp.y, z_recv - outp.z});
    const V3 n2 = unit(V3{n_idx * d.x - wt.x, n_idx * d.y - wt.y, n_idx * d.z - wt.z});
    const V3 wi = wdir ? V3{wdir[3 * i], wdir[3 * i + 1], wdir[3 * i + 2]} : V3{0.0, 0.0, 1.0};
    const V3 n1{0.0, 0.0, 1.0};
    const double rate = transmittance(wi, d, n1, 1.0, n_idx) * transmittance(d, wt, n2, n_idx, 1.0);
    wv = clip_area[i] * rate;
    weight[i] = wv;
  }
  double s = wave_sum_d(wv);
  if ((ln & 63) == 0) red[ln >> 6] = s;
  __syncthreads();
  if (ln == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
    part[blockIdx.x] = t;
  }
}

// nu_i = weight_i * scale
__global__ void scale_kernel(const BlockDesc* blocks, const double* weight, double scale,
                             double* nu) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  nu[i] = weight[i] * scale;
}

// Anderson history slot: G[slot] = w + beta * g, R[slot] = g; partials of R[slot] . R[h] for
// h < H (layout part[h * nblk + block]).
__global__ void aa_store_kernel(const BlockDesc* blocks, const double* w, const double* g,
                                double beta, int slot, int H, double* G, double* R, size_t ld,
                                double* part, int nblk) {
  __shared__ double red[4];
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  const bool live = ln < bd.count;
  const int i = bd.first + ln;
  double gi = 0.0;
  if (live) {
    gi = g[i];
    G[slot * ld + i] = w[i] + beta * gi;
    R[slot * ld + i] = gi;
  }
  for (int h = 0; h < H; ++h) {
    const double v = live ? gi * R[h * ld + i] : 0.0;
    const double s = wave_sum_d(v);
    __syncthreads();
    if ((ln & 63) == 0) red[ln >> 6] = s;
    __syncthreads();
    if (ln == 0) {
      double t = 0.0;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
      part[(size_t)h * nblk + blockIdx.x] = t;
    }
  }
}

// d = sum_h alpha[h] G[h] - w
__global__ void aa_direction_kernel(const BlockDesc* blocks, const double* alpha, int H,
                                    const double* G, size_t ld, const double* w, double* d) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  double acc = 0.0;
  for (int h = 0; h < H; ++h) acc += alpha[h] * G[h * ld + i];
  d[i] = acc - w[i];
}

// w += t * d
__global__ void axpy_kernel(const BlockDesc* blocks, double t, const double* d, double* w) {
  const BlockDesc bd = blocks[blockIdx.x];
  const int ln = threadIdx.x;
  if (ln >= bd.count) return;
  const int i = bd.first + ln;
  w[i] = w[i] + t * d[i];
}

hipError_t launch_fresnel_weight(const BlockDesc* blocks, const double2* points, const double* ins,
                                 const double* wdir, const double* height,
                                 const double* clip_area, const double2* centroid, double z_len,
                                 double z_recv, double n_idx, double* weight, double* part,
                                 int nblk, hipStream_t st) {
  hipLaunchKernelGGL(fresnel_weight_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, points,
                     ins, wdir, height, clip_area, centroid, z_len, z_recv, n_idx, weight, part);
  return hipGetLastError();
}

hipError_t launch_scale(const BlockDesc* blocks, const double* weight, double scale, double* nu,
                        int nblk, hipStream_t st) {
  hipLaunchKernelGGL(scale_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, weight, scale, nu);
  return hipGetLastError();
}

hipError_t launch_aa_store(const BlockDesc* blocks, const double* w, const double* g, double beta,
                           int slot, int H, double* G, double* R, size_t ld, double* part,
                           int nblk, hipStream_t st) {
  hipLaunchKernelGGL(aa_store_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, w, g, beta,
                     slot, H, G, R, ld, part, nblk);
  return hipGetLastError();
}

hipError_t launch_aa_direction(const BlockDesc* blocks, const double* alpha, int H,
                               const double* G, size_t ld, const double* w, double* d, int nblk,
                               hipStream_t st) {
  hipLaunchKernelGGL(aa_direction_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, alpha, H,
                     G, ld, w, d);
  return hipGetLastError();
}

hipError_t launch_axpy(const BlockDesc* blocks, double t, const double* d, double* w, int nblk,
                       hipStream_t st) {
  hipLaunchKernelGGL(axpy_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, t, d, w);
  return hipGetLastError();
}

}  // namespace cot
