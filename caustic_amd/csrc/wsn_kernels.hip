// HIP kernels of the warm-start network (a U-Net mapping a downsampled target density to dual
// weights), for gfx950.
//
// Convolutions are implicit GEMMs on the BF16 matrix cores with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16).  Activations, gradients and master weights are FP32 in HBM; the
// GEMM operands are rounded to BF16 (nearest even) as they are loaded.
//
//   conv_gemm_kernel   out[m][n] = act(sum_k A(m, k) * Bt[n][k] + bias[n]): the forward conv
//                      (A = im2col of the input, Bt = packed weights) and the input gradient
//                      (A = im2col of the pre-activation gradient, Bt = flipped, transposed
//                      weights).  A wave owns a 32 x 32 output tile (2 x 2 MFMA tiles); a block
//                      is 4 waves along m.
//   wgrad_kernel       dW[k][n] += sum_m A(m, k) * dZ[m][n] over a chunk of pixels per block,
//                      FP32 atomics into dW.
// The rest (pooling, upsampling, channel copies, ReLU mask, bias gradient, loss, Adam) is
// elementwise or a short reduction.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

namespace wsn {

// rows of the bias gradient's partial-sum buffer (launch_bias_grad): kBiasBlocks x 256 floats
constexpr int kBiasBlocks = 1024;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  const unsigned int u = __float_as_uint(f);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct ConvArgs {
  const float* in;         // [B, H, W, Cin] NHWC: the GEMM's A source
  const uint16_t* wt;      // [Np][Kp] BF16 bits: column n of the GEMM's B, k contiguous
  const float* bias;       // [N] or nullptr
  float* out;              // [B, H, W, N]
  int B, H, W, Cin, N, K, Kp, ksize, relu;
};

// Eight consecutive k values of im2col row m = (b, y, x), starting at k (a multiple of 8).
__device__ __forceinline__ bf16x8 load_a(const ConvArgs& a, int b, int y, int x, int k) {
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (k >= a.K) return __builtin_bit_cast(bf16x8, v);
  const int pad = a.ksize >> 1;
  if ((a.Cin & 7) == 0) {
    const int tap = k / a.Cin, c = k - tap * a.Cin;
    const int ky = tap / a.ksize, kx = tap - ky * a.ksize;
    const int yy = y + ky - pad, xx = x + kx - pad;
    if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
      const float4* p =
          reinterpret_cast<const float4*>(a.in + (((size_t)b * a.H + yy) * a.W + xx) * a.Cin + c);
      const float4 lo = p[0], hi = p[1];
      v[0] = bf16_bits(lo.x); v[1] = bf16_bits(lo.y); v[2] = bf16_bits(lo.z); v[3] = bf16_bits(lo.w);
      v[4] = bf16_bits(hi.x); v[5] = bf16_bits(hi.y); v[6] = bf16_bits(hi.z); v[7] = bf16_bits(hi.w);
    }
    return __builtin_bit_cast(bf16x8, v);
  }
  for (int e = 0; e < 8; ++e) {
    const int kk = k + e;
    if (kk >= a.K) break;
    const int tap = kk / a.Cin, c = kk - tap * a.Cin;
    const int ky = tap / a.ksize, kx = tap - ky * a.ksize;
    const int yy = y + ky - pad, xx = x + kx - pad;
    if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
      v[e] = bf16_bits(a.in[(((size_t)b * a.H + yy) * a.W + xx) * a.Cin + c]);
  }
  return __builtin_bit_cast(bf16x8, v);
}

// Grid: (M / 128, Np / 32); block 256.  M = B*H*W must be a multiple of 32.
__global__ __launch_bounds__(256) void conv_gemm_kernel(ConvArgs a) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int M = a.B * a.H * a.W;
  const int m0 = (blockIdx.x * 4 + wave) * 32;
  if (m0 >= M) return;
  const int n0 = blockIdx.y * 32;
  int rb[2], ry[2], rx[2];
  for (int s = 0; s < 2; ++s) {
    const int m = m0 + 16 * s + (l & 15);
    rx[s] = m % a.W;
    const int t = m / a.W;
    ry[s] = t % a.H;
    rb[s] = t / a.H;
  }
  const int kq = 8 * (l >> 4);
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < a.Kp; k0 += 32) {
    const int k = k0 + kq;
    bf16x8 A0 = load_a(a, rb[0], ry[0], rx[0], k);
    bf16x8 A1 = load_a(a, rb[1], ry[1], rx[1], k);
    const bf16x8 B0 =
        *reinterpret_cast<const bf16x8*>(a.wt + (size_t)(n0 + (l & 15)) * a.Kp + k);
    const bf16x8 B1 =
        *reinterpret_cast<const bf16x8*>(a.wt + (size_t)(n0 + 16 + (l & 15)) * a.Kp + k);
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B1, acc[1][1], 0, 0, 0);
  }
  // D layout: lane l holds rows 4 * (l / 16) + r, column l % 16 of each 16 x 16 tile
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + 16 * j + (l & 15);
      if (n >= a.N) continue;
      const float bn = a.bias ? a.bias[n] : 0.f;
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * i + 4 * (l >> 4) + r;
        float v = acc[i][j][r] + bn;
        if (a.relu) v = fmaxf(v, 0.f);
        a.out[(size_t)m * a.N + n] = v;
      }
    }
}

struct WgradArgs {
  const float* in;   // [B, H, W, Cin]
  const float* dz;   // [B, H, W, N]
  float* dw;         // [K][N], accumulated
  int B, H, W, Cin, N, K, ksize, chunk;
};

// Grid: (ceil(K / 32), ceil(N / 32), ceil(M / chunk)); block 64 (one wave).  chunk is a
// multiple of 32 and W a multiple of 8, so each lane's 8 pixels lie in one image row.
__global__ __launch_bounds__(64) void wgrad_kernel(WgradArgs a) {
  const int l = threadIdx.x;
  const int M = a.B * a.H * a.W;
  const int k0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
  const int mbeg = blockIdx.z * a.chunk, mend = min(mbeg + a.chunk, M);
  const int pad = a.ksize >> 1;
  // the lane's two weight rows (k) and two columns (n)
  int tk[2], ky[2], kx[2], kc[2];
  bool kv[2];
  for (int s = 0; s < 2; ++s) {
    const int k = k0 + 16 * s + (l & 15);
    kv[s] = k < a.K;
    const int kk = kv[s] ? k : 0;
    tk[s] = kk / a.Cin;
    kc[s] = kk - tk[s] * a.Cin;
    ky[s] = tk[s] / a.ksize;
    kx[s] = tk[s] - ky[s] * a.ksize;
  }
  int nn[2];
  bool nv[2];
  for (int t = 0; t < 2; ++t) {
    nn[t] = n0 + 16 * t + (l & 15);
    nv[t] = nn[t] < a.N;
  }
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int mm = mbeg; mm < mend; mm += 32) {
    const int mf = mm + 8 * (l >> 4);  // first of the lane's 8 pixels
    const int x0 = mf % a.W;
    const int t = mf / a.W;
    const int y = t % a.H, b = t / a.H;
    u16x8 av[2], bv[2];
    for (int s = 0; s < 2; ++s) {
      const int yy = y + ky[s] - pad;
      const bool rowok = kv[s] && yy >= 0 && yy < a.H && mf < mend;
      for (int e = 0; e < 8; ++e) {
        const int xx = x0 + e + kx[s] - pad;
        float v = 0.f;
        if (rowok && xx >= 0 && xx < a.W)
          v = a.in[(((size_t)b * a.H + yy) * a.W + xx) * a.Cin + kc[s]];
        av[s][e] = bf16_bits(v);
      }
    }
    for (int u = 0; u < 2; ++u)
      for (int e = 0; e < 8; ++e) {
        float v = 0.f;
        if (nv[u] && mf < mend) v = a.dz[(size_t)(mf + e) * a.N + nn[u]];
        bv[u][e] = bf16_bits(v);
      }
    const bf16x8 A0 = __builtin_bit_cast(bf16x8, av[0]), A1 = __builtin_bit_cast(bf16x8, av[1]);
    const bf16x8 B0 = __builtin_bit_cast(bf16x8, bv[0]), B1 = __builtin_bit_cast(bf16x8, bv[1]);
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B1, acc[1][1], 0, 0, 0);
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + 16 * j + (l & 15);
      if (n >= a.N) continue;
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + 16 * i + 4 * (l >> 4) + r;
        if (k < a.K) atomicAdd(&a.dw[(size_t)k * a.N + n], acc[i][j][r]);
      }
    }
}

// Forward packing: wt[n][k] = bf16(w[k][n]) (zero padding to Np x Kp).
__global__ void pack_forward_kernel(const float* w, int K, int N, int Kp, int Np, uint16_t* wt) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)Np * Kp) return;
  const int n = (int)(idx / Kp), k = (int)(idx - (size_t)n * Kp);
  wt[idx] = (n < N && k < K) ? bf16_bits(w[(size_t)k * N + n]) : (uint16_t)0;
}

// Input-gradient packing for a ksize x ksize conv with Cin inputs and Cout outputs: the GEMM
// column is the input channel ci, its k index (tap', co) with tap' the flipped tap:
// wd[ci][tap' * Cout + co] = bf16(w[(T - 1 - tap') * Cin + ci][co]), T = ksize^2.
__global__ void pack_backward_kernel(const float* w, int ksize, int Cin, int Cout, int Kp, int Np,
                                     uint16_t* wd) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)Np * Kp) return;
  const int ci = (int)(idx / Kp), k = (int)(idx - (size_t)ci * Kp);
  const int T = ksize * ksize;
  uint16_t v = 0;
  if (ci < Cin && k < T * Cout) {
    const int tp = k / Cout, co = k - tp * Cout;
    v = bf16_bits(w[((size_t)(T - 1 - tp) * Cin + ci) * Cout + co]);
  }
  wd[idx] = v;
}

// dz = dy * (y > 0)
__global__ void relu_grad_kernel(const float* y, const float* dy, float* dz, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = y[i] > 0.f ? dy[i] : 0.f;
}

// Bias gradient db[n] += sum_m dz[m][n] (N <= 256) in two passes with no atomics.  Pass 1:
// block b sums rows [b * chunk, (b + 1) * chunk); its 256 threads split into R = 256 / N row
// phases of N columns each (thread t: column t % N, rows from t / N in steps of R), so a
// narrow layer (N = 1 or 32) still keeps every thread loading; the R phases of a column are
// added in LDS and the block writes part[b * N + n].
__global__ __launch_bounds__(256) void bias_part_kernel(const float* dz, int M, int N, int chunk,
                                                        float* part) {
  __shared__ float red[256];
  const int R = 256 / N;
  const int t = threadIdx.x;
  const int n = t % N, r = t / N;
  float s = 0.f;
  if (r < R) {
    const int m0 = blockIdx.x * chunk, m1 = min(m0 + chunk, M);
    for (int m = m0 + r; m < m1; m += R) s += dz[(size_t)m * N + n];
  }
  red[t] = s;
  __syncthreads();
  if (t < N) {
    float a = 0.f;
    for (int q = 0; q < R; ++q) a += red[q * N + t];
    part[(size_t)blockIdx.x * N + t] = a;
  }
}

// Pass 2: one block per column adds the nb partials (tree in LDS) into db[n].
__global__ __launch_bounds__(256) void bias_sum_kernel(const float* part, int nb, int N,
                                                       float* db) {
  __shared__ float red[256];
  const int n = blockIdx.x, t = threadIdx.x;
  float a = 0.f;
  for (int b = t; b < nb; b += 256) a += part[(size_t)b * N + n];
  red[t] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) db[n] += red[0];
}

// 2 x 2 average pool: out [B, H/2, W/2, C]
__global__ void pool_kernel(const float* in, int B, int H, int W, int C, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  if (i >= (size_t)B * Ho * Wo * C) return;
  const int c = (int)(i % C);
  size_t t = i / C;
  const int x = (int)(t % Wo);
  t /= Wo;
  const int y = (int)(t % Ho);
  const int b = (int)(t / Ho);
  const float* p = in + (((size_t)b * H + 2 * y) * W + 2 * x) * C + c;
  out[i] = 0.25f * (p[0] + p[C] + p[(size_t)W * C] + p[(size_t)W * C + C]);
}

// adjoint of the pool: din [B, H, W, C] (+)= dout [B, H/2, W/2, C] / 4
__global__ void pool_back_kernel(const float* dout, int B, int H, int W, int C, float* din,
                                 int accumulate) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)B * H * W * C) return;
  const int c = (int)(i % C);
  size_t t = i / C;
  const int x = (int)(t % W);
  t /= W;
  const int y = (int)(t % H);
  const int b = (int)(t / H);
  const float v = 0.25f * dout[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
  din[i] = accumulate ? din[i] + v : v;
}

// nearest 2x upsampling of in [B, H, W, C] into channels [c0, c0 + C) of out [B, 2H, 2W, Cd]
__global__ void up_kernel(const float* in, int B, int H, int W, int C, float* out, int Cd,
                          int c0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int H2 = 2 * H, W2 = 2 * W;
  if (i >= (size_t)B * H2 * W2 * C) return;
  const int c = (int)(i % C);
  size_t t = i / C;
  const int x = (int)(t % W2);
  t /= W2;
  const int y = (int)(t % H2);
  const int b = (int)(t / H2);
  out[(((size_t)b * H2 + y) * W2 + x) * Cd + c0 + c] =
      in[(((size_t)b * H + y / 2) * W + x / 2) * C + c];
}

// adjoint of up_kernel: din [B, H, W, C] = sum of the 2 x 2 block of channels [c0, c0 + C) of
// dout [B, 2H, 2W, Cd]
__global__ void up_back_kernel(const float* dout, int B, int H, int W, int C, int Cd, int c0,
                               float* din) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)B * H * W * C) return;
  const int c = (int)(i % C);
  size_t t = i / C;
  const int x = (int)(t % W);
  t /= W;
  const int y = (int)(t % H);
  const int b = (int)(t / H);
  const size_t W2 = 2 * (size_t)W;
  const float* p = dout + (((size_t)b * 2 * H + 2 * y) * W2 + 2 * x) * Cd + c0 + c;
  din[i] = p[0] + p[Cd] + p[W2 * Cd] + p[W2 * Cd + Cd];
}

// dst[p][d0 + c] (+)= src[p][s0 + c] for c < nc over P pixels
__global__ void copy_channels_kernel(const float* src, int Cs, int s0, float* dst, int Cd, int d0,
                                     int nc, size_t P, int accumulate) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * nc) return;
  const size_t p = i / nc;
  const int c = (int)(i - p * nc);
  const float v = src[p * Cs + s0 + c];
  float* d = dst + p * Cd + d0 + c;
  *d = accumulate ? *d + v : v;
}

// Per sample b (one block of 256): e = out - mean(out) - t over the S x S image;
// loss_b = (sum e^2 + lam * sum over neighbouring pixel pairs (e_a - e_b)^2) / (B S^2);
// dout = de - mean(de) with de_k = 2 (e_k + lam * sum over k's neighbours (e_k - e_nbr)) / (B S^2).
__global__ void loss_kernel(const float* out, const float* t, int S, int B, float lam,
                            float* loss, float* dout) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const int S2 = S * S;
  const float* o = out + (size_t)b * S2;
  const float* tt = t + (size_t)b * S2;
  auto block_sum = [&](double v) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  double so = 0.0, st = 0.0;
  for (int i = threadIdx.x; i < S2; i += blockDim.x) {
    so += o[i];
    st += tt[i];
  }
  const double mo = block_sum(so) / S2, mt = block_sum(st) / S2;
  auto err = [&](int k) { return (double)o[k] - mo - tt[k]; };
  double se = 0.0;
  for (int i = threadIdx.x; i < S2; i += blockDim.x) {
    const double e = err(i);
    se += e * e;
    const int x = i % S, y = i / S;
    if (x + 1 < S) {
      const double d = err(i + 1) - e;
      se += lam * d * d;
    }
    if (y + 1 < S) {
      const double d = err(i + S) - e;
      se += lam * d * d;
    }
  }
  const double ls = block_sum(se);
  const double sc = 2.0 / ((double)B * S2);
  if (threadIdx.x == 0) loss[b] = (float)(ls / ((double)B * S2));
  // mean over the image of de: the difference terms cancel; mean(e) = -mean(t)
  const double mean_de = sc * (-mt);
  for (int i = threadIdx.x; i < S2; i += blockDim.x) {
    const double e = err(i);
    const int x = i % S, y = i / S;
    double lap = 0.0;
    if (x > 0) lap += e - err(i - 1);
    if (x + 1 < S) lap += e - err(i + 1);
    if (y > 0) lap += e - err(i - S);
    if (y + 1 < S) lap += e - err(i + S);
    dout[(size_t)b * S2 + i] = (float)(sc * (e + lam * lap) - mean_de);
  }
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void fill_f32_kernel(float* p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------------------

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_conv(const ConvArgs& a, int Np, hipStream_t st) {
  const int M = a.B * a.H * a.W;
  hipLaunchKernelGGL(conv_gemm_kernel, dim3((M + 127) / 128, Np / 32), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_wgrad(const WgradArgs& a, hipStream_t st) {
  const int M = a.B * a.H * a.W;
  hipLaunchKernelGGL(wgrad_kernel, dim3((a.K + 31) / 32, (a.N + 31) / 32, (M + a.chunk - 1) / a.chunk),
                     dim3(64), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_pack_forward(const float* w, int K, int N, int Kp, int Np, uint16_t* wt,
                               hipStream_t st) {
  hipLaunchKernelGGL(pack_forward_kernel, dim3(nblk((size_t)Np * Kp, 256)), dim3(256), 0, st, w,
                     K, N, Kp, Np, wt);
  return hipGetLastError();
}
hipError_t launch_pack_backward(const float* w, int ksize, int Cin, int Cout, int Kp, int Np,
                                uint16_t* wd, hipStream_t st) {
  hipLaunchKernelGGL(pack_backward_kernel, dim3(nblk((size_t)Np * Kp, 256)), dim3(256), 0, st, w,
                     ksize, Cin, Cout, Kp, Np, wd);
  return hipGetLastError();
}
hipError_t launch_relu_grad(const float* y, const float* dy, float* dz, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(relu_grad_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, y, dy, dz, n);
  return hipGetLastError();
}
hipError_t launch_bias_grad(const float* dz, int M, int N, float* db, float* part,
                            hipStream_t st) {
  if (N < 1 || N > 256 || M < 1) return hipErrorInvalidValue;
  // about 16 loads per thread, at most kBiasBlocks blocks (the partial buffer's rows)
  long long nb = ((long long)M * N + 4095) / 4096;
  nb = std::max(1LL, std::min(nb, (long long)kBiasBlocks));
  const int chunk = (int)((M + nb - 1) / nb);
  const int blocks = (M + chunk - 1) / chunk;
  hipLaunchKernelGGL(bias_part_kernel, dim3(blocks), dim3(256), 0, st, dz, M, N, chunk, part);
  hipLaunchKernelGGL(bias_sum_kernel, dim3(N), dim3(256), 0, st, part, blocks, N, db);
  return hipGetLastError();
}
hipError_t launch_pool(const float* in, int B, int H, int W, int C, float* out, hipStream_t st) {
  hipLaunchKernelGGL(pool_kernel, dim3(nblk((size_t)B * (H / 2) * (W / 2) * C, 256)), dim3(256), 0,
                     st, in, B, H, W, C, out);
  return hipGetLastError();
}
hipError_t launch_pool_back(const float* dout, int B, int H, int W, int C, float* din, int acc,
                            hipStream_t st) {
  hipLaunchKernelGGL(pool_back_kernel, dim3(nblk((size_t)B * H * W * C, 256)), dim3(256), 0, st,
                     dout, B, H, W, C, din, acc);
  return hipGetLastError();
}
hipError_t launch_up(const float* in, int B, int H, int W, int C, float* out, int Cd, int c0,
                     hipStream_t st) {
  hipLaunchKernelGGL(up_kernel, dim3(nblk((size_t)B * 4 * H * W * C, 256)), dim3(256), 0, st, in,
                     B, H, W, C, out, Cd, c0);
  return hipGetLastError();
}
hipError_t launch_up_back(const float* dout, int B, int H, int W, int C, int Cd, int c0,
                          float* din, hipStream_t st) {
  hipLaunchKernelGGL(up_back_kernel, dim3(nblk((size_t)B * H * W * C, 256)), dim3(256), 0, st,
                     dout, B, H, W, C, Cd, c0, din);
  return hipGetLastError();
}
hipError_t launch_copy_channels(const float* src, int Cs, int s0, float* dst, int Cd, int d0,
                                int nc, size_t P, int acc, hipStream_t st) {
  hipLaunchKernelGGL(copy_channels_kernel, dim3(nblk(P * nc, 256)), dim3(256), 0, st, src, Cs, s0,
                     dst, Cd, d0, nc, P, acc);
  return hipGetLastError();
}
hipError_t launch_loss(const float* out, const float* t, int S, int B, float lam, float* loss,
                       float* dout, hipStream_t st) {
  hipLaunchKernelGGL(loss_kernel, dim3(B), dim3(256), 0, st, out, t, S, B, lam, loss, dout);
  return hipGetLastError();
}
hipError_t launch_adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2,
                     eps, c1, c2);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t st) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace wsn
