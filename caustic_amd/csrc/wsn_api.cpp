// Host driver and C ABI of the warm-start network (include/caustic_wsn.h).
//
// The network is a four-level U-Net (layer table below) mapping log-density images S x S to
// per-site dual weights on an S x S site grid.  Parameters live in one FP32 array (per layer:
// weights [k*k*c_in][c_out], then the bias); the GEMM kernels read BF16 copies packed after every
// update (one for the forward product, one flipped and transposed for the input gradient).
// Activations and gradients are FP32 NHWC buffers sized for the largest batch.  Training is
// Adam on the mean squared error of the gauge-free prediction (its per-sample mean removed).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/caustic_wsn.h"

namespace wsn {
struct ConvArgs {
  const float* in;
  const uint16_t* wt;
  const float* bias;
  float* out;
  int B, H, W, Cin, N, K, Kp, ksize, relu;
};
struct WgradArgs {
  const float* in;
  const float* dz;
  float* dw;
  int B, H, W, Cin, N, K, ksize, chunk;
};
hipError_t launch_conv(const ConvArgs& a, int Np, hipStream_t st);
hipError_t launch_wgrad(const WgradArgs& a, hipStream_t st);
hipError_t launch_pack_forward(const float* w, int K, int N, int Kp, int Np, uint16_t* wt,
                               hipStream_t st);
hipError_t launch_pack_backward(const float* w, int ksize, int Cin, int Cout, int Kp, int Np,
                                uint16_t* wd, hipStream_t st);
hipError_t launch_relu_grad(const float* y, const float* dy, float* dz, size_t n, hipStream_t st);
hipError_t launch_bias_grad(const float* dz, int M, int N, float* db, float* part,
                            hipStream_t st);
constexpr size_t kBiasPartFloats = 1024 * 256;  // kBiasBlocks x 256 (wsn_kernels.hip)
hipError_t launch_pool(const float* in, int B, int H, int W, int C, float* out, hipStream_t st);
hipError_t launch_pool_back(const float* dout, int B, int H, int W, int C, float* din, int acc,
                            hipStream_t st);
hipError_t launch_up(const float* in, int B, int H, int W, int C, float* out, int Cd, int c0,
                     hipStream_t st);
hipError_t launch_up_back(const float* dout, int B, int H, int W, int C, int Cd, int c0,
                          float* din, hipStream_t st);
hipError_t launch_copy_channels(const float* src, int Cs, int s0, float* dst, int Cd, int d0,
                                int nc, size_t P, int acc, hipStream_t st);
hipError_t launch_loss(const float* out, const float* t, int S, int B, float lam, float* loss,
                       float* dout, hipStream_t st);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st);
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t st);
}  // namespace wsn

using namespace wsn;

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
      return fail(WSN_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

#define CHK(expr)                  \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != WSN_OK) return rc_; \
  } while (0)

inline int rup(int v, int m) { return (v + m - 1) / m * m; }

struct LayerSpec {
  const char* name;
  int k, cin, cout, relu;
};
// Same table as the CPU oracle of the network.
const LayerSpec kLayers[] = {
    {"e0a", 3, 1, 32, 1},    {"e0b", 3, 32, 32, 1},   {"e1a", 3, 32, 64, 1},
    {"e1b", 3, 64, 64, 1},   {"e2a", 3, 64, 128, 1},  {"e2b", 3, 128, 128, 1},
    {"e3a", 3, 128, 192, 1}, {"e3b", 3, 192, 192, 1}, {"d2a", 3, 320, 128, 1},
    {"d2b", 3, 128, 128, 1}, {"d1a", 3, 192, 64, 1},  {"d1b", 3, 64, 64, 1},
    {"d0a", 3, 96, 32, 1},   {"d0b", 3, 32, 32, 1},   {"out", 1, 32, 1, 0},
};
constexpr int kNumLayers = sizeof(kLayers) / sizeof(kLayers[0]);
enum {
  E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT
};

struct Layer {
  LayerSpec s;
  size_t woff = 0, boff = 0;
  int K = 0, Kp = 0, Np = 0;    // forward GEMM: K = k*k*cin, columns cout
  int Kpd = 0, Npd = 0;         // input-gradient GEMM: K = k*k*cout, columns cin
  uint16_t* wt = nullptr;       // packed forward weights [Np][Kp]
  uint16_t* wd = nullptr;       // packed input-gradient weights [Npd][Kpd]
};

}  // namespace

struct wsn_net {
  int S = 0, maxB = 0;
  hipStream_t st = nullptr;
  Layer L[kNumLayers];
  size_t np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  long long adam_t = 0;
  std::vector<void*> allocs;
  // activations (post-activation outputs of each layer, pooled and concatenated tensors)
  float* in = nullptr;
  float* act[kNumLayers] = {};
  float *p0 = nullptr, *p1 = nullptr, *p2 = nullptr, *c2 = nullptr, *c1 = nullptr, *c0 = nullptr;
  // gradients with respect to the same tensors
  float* gact[kNumLayers] = {};
  float *gp0 = nullptr, *gp1 = nullptr, *gp2 = nullptr, *gc2 = nullptr, *gc1 = nullptr,
        *gc0 = nullptr;
  float* dz = nullptr;      // pre-activation gradient scratch
  float* bpart = nullptr;   // per-block partial sums of the bias gradients
  float* tgt = nullptr;     // targets of the batch
  float* loss = nullptr;    // per-sample loss
  // resident training set
  float *ds_x = nullptr, *ds_t = nullptr;
  int ds_n = 0;
  double train_seconds = 0.0;
  long long train_samples = 0;
  float grad_weight = 0.f;  // weight of the neighbouring-difference term of the loss

  ~wsn_net() {
    if (st) hipStreamSynchronize(st);
    for (void* q : allocs) hipFree(q);
    if (st) hipStreamDestroy(st);
  }
};

namespace {

template <typename T>
int dalloc(wsn_net* n, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
    return fail(WSN_ENOMEM, "hipMalloc failed");
  n->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return WSN_OK;
}

// spatial size and channel count of each layer's output
int level_of(int li) {
  switch (li) {
    case E0A: case E0B: case D0A: case D0B: case OUT: return 0;
    case E1A: case E1B: case D1A: case D1B: return 1;
    case E2A: case E2B: case D2A: case D2B: return 2;
    default: return 3;
  }
}

int pack_all(wsn_net* n) {
  for (int i = 0; i < kNumLayers; ++i) {
    Layer& l = n->L[i];
    HIPCHK(launch_pack_forward(n->p + l.woff, l.K, l.s.cout, l.Kp, l.Np, l.wt, n->st));
    HIPCHK(launch_pack_backward(n->p + l.woff, l.s.k, l.s.cin, l.s.cout, l.Kpd, l.Npd, l.wd,
                                n->st));
  }
  return WSN_OK;
}

int conv_fwd(wsn_net* n, int li, const float* x, int B) {
  const Layer& l = n->L[li];
  const int H = n->S >> level_of(li);
  ConvArgs a{};
  a.in = x;
  a.wt = l.wt;
  a.bias = n->p + l.boff;
  a.out = n->act[li];
  a.B = B;
  a.H = H;
  a.W = H;
  a.Cin = l.s.cin;
  a.N = l.s.cout;
  a.K = l.K;
  a.Kp = l.Kp;
  a.ksize = l.s.k;
  a.relu = l.s.relu;
  HIPCHK(launch_conv(a, l.Np, n->st));
  return WSN_OK;
}

int forward(wsn_net* n, int B) {
  const int S = n->S;
  hipStream_t st = n->st;
  CHK(conv_fwd(n, E0A, n->in, B));
  CHK(conv_fwd(n, E0B, n->act[E0A], B));
  HIPCHK(launch_pool(n->act[E0B], B, S, S, 32, n->p0, st));
  CHK(conv_fwd(n, E1A, n->p0, B));
  CHK(conv_fwd(n, E1B, n->act[E1A], B));
  HIPCHK(launch_pool(n->act[E1B], B, S / 2, S / 2, 64, n->p1, st));
  CHK(conv_fwd(n, E2A, n->p1, B));
  CHK(conv_fwd(n, E2B, n->act[E2A], B));
  HIPCHK(launch_pool(n->act[E2B], B, S / 4, S / 4, 128, n->p2, st));
  CHK(conv_fwd(n, E3A, n->p2, B));
  CHK(conv_fwd(n, E3B, n->act[E3A], B));
  HIPCHK(launch_up(n->act[E3B], B, S / 8, S / 8, 192, n->c2, 320, 0, st));
  HIPCHK(launch_copy_channels(n->act[E2B], 128, 0, n->c2, 320, 192, 128,
                              (size_t)B * (S / 4) * (S / 4), 0, st));
  CHK(conv_fwd(n, D2A, n->c2, B));
  CHK(conv_fwd(n, D2B, n->act[D2A], B));
  HIPCHK(launch_up(n->act[D2B], B, S / 4, S / 4, 128, n->c1, 192, 0, st));
  HIPCHK(launch_copy_channels(n->act[E1B], 64, 0, n->c1, 192, 128, 64,
                              (size_t)B * (S / 2) * (S / 2), 0, st));
  CHK(conv_fwd(n, D1A, n->c1, B));
  CHK(conv_fwd(n, D1B, n->act[D1A], B));
  HIPCHK(launch_up(n->act[D1B], B, S / 2, S / 2, 64, n->c0, 96, 0, st));
  HIPCHK(launch_copy_channels(n->act[E0B], 32, 0, n->c0, 96, 64, 32, (size_t)B * S * S, 0, st));
  CHK(conv_fwd(n, D0A, n->c0, B));
  CHK(conv_fwd(n, D0B, n->act[D0A], B));
  CHK(conv_fwd(n, OUT, n->act[D0B], B));
  return WSN_OK;
}

// Backward of one conv layer: parameter gradients from its input x and the gradient gy of its
// output; the input gradient into gx when gx != nullptr.
int conv_bwd(wsn_net* n, int li, const float* x, const float* gy, float* gx, int B) {
  const Layer& l = n->L[li];
  const int H = n->S >> level_of(li);
  const int M = B * H * H;
  const size_t cnt = (size_t)M * l.s.cout;
  const float* dz = gy;
  if (l.s.relu) {
    HIPCHK(launch_relu_grad(n->act[li], gy, n->dz, cnt, n->st));
    dz = n->dz;
  }
  WgradArgs w{};
  w.in = x;
  w.dz = dz;
  w.dw = n->g + l.woff;
  w.B = B;
  w.H = H;
  w.W = H;
  w.Cin = l.s.cin;
  w.N = l.s.cout;
  w.K = l.K;
  w.ksize = l.s.k;
  // pixel chunk per block: enough blocks to fill the chip, at least 32 pixels
  const int tiles = ((l.K + 31) / 32) * ((l.s.cout + 31) / 32);
  int chunk = rup(std::max(32, (int)((long long)M * tiles / 2048)), 32);
  chunk = std::min(chunk, rup(M, 32));
  w.chunk = chunk;
  HIPCHK(launch_wgrad(w, n->st));
  HIPCHK(launch_bias_grad(dz, M, l.s.cout, n->g + l.boff, n->bpart, n->st));
  if (gx) {
    ConvArgs a{};
    a.in = dz;
    a.wt = l.wd;
    a.bias = nullptr;
    a.out = gx;
    a.B = B;
    a.H = H;
    a.W = H;
    a.Cin = l.s.cout;
    a.N = l.s.cin;
    a.K = l.s.k * l.s.k * l.s.cout;
    a.Kp = l.Kpd;
    a.ksize = l.s.k;
    a.relu = 0;
    HIPCHK(launch_conv(a, l.Npd, n->st));
  }
  return WSN_OK;
}

// Gradients of every parameter into n->g given dL/d(out) in n->gact[OUT].
int backward(wsn_net* n, int B) {
  const int S = n->S;
  hipStream_t st = n->st;
  HIPCHK(launch_fill_f32(n->g, n->np, 0.f, st));
  float** A = n->act;
  float** G = n->gact;
  CHK(conv_bwd(n, OUT, A[D0B], G[OUT], G[D0B], B));
  CHK(conv_bwd(n, D0B, A[D0A], G[D0B], G[D0A], B));
  CHK(conv_bwd(n, D0A, n->c0, G[D0A], n->gc0, B));
  HIPCHK(launch_up_back(n->gc0, B, S / 2, S / 2, 64, 96, 0, G[D1B], st));
  HIPCHK(launch_copy_channels(n->gc0, 96, 64, G[E0B], 32, 0, 32, (size_t)B * S * S, 0, st));
  CHK(conv_bwd(n, D1B, A[D1A], G[D1B], G[D1A], B));
  CHK(conv_bwd(n, D1A, n->c1, G[D1A], n->gc1, B));
  HIPCHK(launch_up_back(n->gc1, B, S / 4, S / 4, 128, 192, 0, G[D2B], st));
  HIPCHK(launch_copy_channels(n->gc1, 192, 128, G[E1B], 64, 0, 64, (size_t)B * (S / 2) * (S / 2),
                              0, st));
  CHK(conv_bwd(n, D2B, A[D2A], G[D2B], G[D2A], B));
  CHK(conv_bwd(n, D2A, n->c2, G[D2A], n->gc2, B));
  HIPCHK(launch_up_back(n->gc2, B, S / 8, S / 8, 192, 320, 0, G[E3B], st));
  HIPCHK(launch_copy_channels(n->gc2, 320, 192, G[E2B], 128, 0, 128,
                              (size_t)B * (S / 4) * (S / 4), 0, st));
  CHK(conv_bwd(n, E3B, A[E3A], G[E3B], G[E3A], B));
  CHK(conv_bwd(n, E3A, n->p2, G[E3A], n->gp2, B));
  HIPCHK(launch_pool_back(n->gp2, B, S / 4, S / 4, 128, G[E2B], 1, st));
  CHK(conv_bwd(n, E2B, A[E2A], G[E2B], G[E2A], B));
  CHK(conv_bwd(n, E2A, n->p1, G[E2A], n->gp1, B));
  HIPCHK(launch_pool_back(n->gp1, B, S / 2, S / 2, 64, G[E1B], 1, st));
  CHK(conv_bwd(n, E1B, A[E1A], G[E1B], G[E1A], B));
  CHK(conv_bwd(n, E1A, n->p0, G[E1A], n->gp0, B));
  HIPCHK(launch_pool_back(n->gp0, B, S, S, 32, G[E0B], 1, st));
  CHK(conv_bwd(n, E0B, A[E0A], G[E0B], G[E0A], B));
  CHK(conv_bwd(n, E0A, n->in, G[E0A], nullptr, B));
  return WSN_OK;
}

int check_batch(wsn_net* n, int B) {
  if (B < 1 || B > n->maxB)
    return fail(WSN_EINVAL, "batch size must be in [1, " + std::to_string(n->maxB) + "]");
  return WSN_OK;
}

// forward + loss (+ gradients into n->g); in and tgt already hold the batch
int loss_and_grad(wsn_net* n, int B, float* loss_out) {
  CHK(forward(n, B));
  HIPCHK(launch_loss(n->act[OUT], n->tgt, n->S, B, n->grad_weight, n->loss, n->gact[OUT], n->st));
  CHK(backward(n, B));
  if (loss_out) {
    std::vector<float> ls(B);
    HIPCHK(hipMemcpyAsync(ls.data(), n->loss, sizeof(float) * B, hipMemcpyDeviceToHost, n->st));
    HIPCHK(hipStreamSynchronize(n->st));
    double s = 0.0;
    for (float x : ls) s += x;
    *loss_out = (float)s;
  }
  return WSN_OK;
}

int adam_step(wsn_net* n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->adam_t += 1;
  const float c1 = 1.f - std::pow(b1, (float)n->adam_t);
  const float c2 = 1.f - std::pow(b2, (float)n->adam_t);
  HIPCHK(launch_adam(n->p, n->g, n->m, n->v, n->np, lr, b1, b2, eps, c1, c2, n->st));
  CHK(pack_all(n));
  return WSN_OK;
}

}  // namespace

extern "C" {

const char* wsn_last_error(void) { return g_err.c_str(); }

int wsn_create(int S, int max_batch, wsn_net** out) {
  if (!out) return fail(WSN_EINVAL, "null output");
  *out = nullptr;
  // four levels, the coarsest at least 8 pixels wide (the weight-gradient kernel takes 8
  // consecutive pixels of one image row per lane)
  if (S < 64 || S % 64 != 0) return fail(WSN_EINVAL, "S must be a positive multiple of 64");
  if (max_batch < 1) return fail(WSN_EINVAL, "max_batch must be positive");
  wsn_net* n = new wsn_net();
  auto bail = [&](int code) {
    delete n;
    return code;
  };
  n->S = S;
  n->maxB = max_batch;
  if (hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(WSN_EHIP, "hipStreamCreate failed"));
  size_t off = 0;
  for (int i = 0; i < kNumLayers; ++i) {
    Layer& l = n->L[i];
    l.s = kLayers[i];
    l.K = l.s.k * l.s.k * l.s.cin;
    l.Kp = rup(l.K, 32);
    l.Np = rup(l.s.cout, 32);
    l.Kpd = rup(l.s.k * l.s.k * l.s.cout, 32);
    l.Npd = rup(l.s.cin, 32);
    l.woff = off;
    off += (size_t)l.K * l.s.cout;
    l.boff = off;
    off += l.s.cout;
  }
  n->np = off;
  int rc = WSN_OK;
  if ((rc = dalloc(n, &n->p, n->np)) || (rc = dalloc(n, &n->g, n->np)) ||
      (rc = dalloc(n, &n->m, n->np)) || (rc = dalloc(n, &n->v, n->np)))
    return bail(rc);
  for (int i = 0; i < kNumLayers; ++i) {
    Layer& l = n->L[i];
    if ((rc = dalloc(n, &l.wt, (size_t)l.Np * l.Kp)) ||
        (rc = dalloc(n, &l.wd, (size_t)l.Npd * l.Kpd)))
      return bail(rc);
  }
  const size_t B = max_batch, P0 = (size_t)S * S;
  auto A = [&](float** p, size_t count) { return dalloc(n, p, count); };
  if ((rc = A(&n->in, B * P0)) || (rc = A(&n->tgt, B * P0)) || (rc = A(&n->loss, B)) ||
      (rc = A(&n->dz, B * P0 * 32)) || (rc = A(&n->bpart, kBiasPartFloats)) || (rc = A(&n->p0, B * P0 / 4 * 32)) ||
      (rc = A(&n->p1, B * P0 / 16 * 64)) || (rc = A(&n->p2, B * P0 / 64 * 128)) ||
      (rc = A(&n->c2, B * P0 / 16 * 320)) || (rc = A(&n->c1, B * P0 / 4 * 192)) ||
      (rc = A(&n->c0, B * P0 * 96)) || (rc = A(&n->gp0, B * P0 / 4 * 32)) ||
      (rc = A(&n->gp1, B * P0 / 16 * 64)) || (rc = A(&n->gp2, B * P0 / 64 * 128)) ||
      (rc = A(&n->gc2, B * P0 / 16 * 320)) || (rc = A(&n->gc1, B * P0 / 4 * 192)) ||
      (rc = A(&n->gc0, B * P0 * 96)))
    return bail(rc);
  for (int i = 0; i < kNumLayers; ++i) {
    const size_t pix = B * (P0 >> (2 * level_of(i)));
    if ((rc = A(&n->act[i], pix * kLayers[i].cout)) || (rc = A(&n->gact[i], pix * kLayers[i].cout)))
      return bail(rc);
  }
  if (hipMemsetAsync(n->m, 0, sizeof(float) * n->np, n->st) != hipSuccess ||
      hipMemsetAsync(n->v, 0, sizeof(float) * n->np, n->st) != hipSuccess ||
      hipMemsetAsync(n->p, 0, sizeof(float) * n->np, n->st) != hipSuccess ||
      pack_all(n) != WSN_OK || hipStreamSynchronize(n->st) != hipSuccess)
    return bail(fail(WSN_EHIP, "initialisation failed"));
  *out = n;
  return WSN_OK;
}

int wsn_destroy(wsn_net* n) {
  delete n;
  return WSN_OK;
}

long long wsn_param_count(const wsn_net* n) { return n ? (long long)n->np : -1; }

int wsn_set_params(wsn_net* n, const float* params) {
  if (!n || !params) return fail(WSN_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(n->p, params, sizeof(float) * n->np, hipMemcpyHostToDevice, n->st));
  HIPCHK(hipMemsetAsync(n->m, 0, sizeof(float) * n->np, n->st));
  HIPCHK(hipMemsetAsync(n->v, 0, sizeof(float) * n->np, n->st));
  n->adam_t = 0;
  CHK(pack_all(n));
  HIPCHK(hipStreamSynchronize(n->st));
  return WSN_OK;
}

int wsn_get_params(wsn_net* n, float* params) {
  if (!n || !params) return fail(WSN_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(params, n->p, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->st));
  HIPCHK(hipStreamSynchronize(n->st));
  return WSN_OK;
}

int wsn_forward(wsn_net* n, const float* x, int B, float* y) {
  if (!n || !x || !y) return fail(WSN_EINVAL, "null argument");
  CHK(check_batch(n, B));
  const size_t cnt = (size_t)B * n->S * n->S;
  HIPCHK(hipMemcpyAsync(n->in, x, sizeof(float) * cnt, hipMemcpyHostToDevice, n->st));
  CHK(forward(n, B));
  HIPCHK(hipMemcpyAsync(y, n->act[OUT], sizeof(float) * cnt, hipMemcpyDeviceToHost, n->st));
  HIPCHK(hipStreamSynchronize(n->st));
  return WSN_OK;
}

int wsn_gradients(wsn_net* n, const float* x, const float* t, int B, float* grad, float* loss) {
  if (!n || !x || !t || !grad) return fail(WSN_EINVAL, "null argument");
  CHK(check_batch(n, B));
  const size_t cnt = (size_t)B * n->S * n->S;
  HIPCHK(hipMemcpyAsync(n->in, x, sizeof(float) * cnt, hipMemcpyHostToDevice, n->st));
  HIPCHK(hipMemcpyAsync(n->tgt, t, sizeof(float) * cnt, hipMemcpyHostToDevice, n->st));
  CHK(loss_and_grad(n, B, loss));
  HIPCHK(hipMemcpyAsync(grad, n->g, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->st));
  HIPCHK(hipStreamSynchronize(n->st));
  return WSN_OK;
}

int wsn_set_loss(wsn_net* n, float grad_weight) {
  if (!n || !(grad_weight >= 0.f)) return fail(WSN_EINVAL, "bad loss weight");
  n->grad_weight = grad_weight;
  return WSN_OK;
}

int wsn_set_dataset(wsn_net* n, const float* x, const float* t, int count) {
  if (!n || !x || !t || count < 1) return fail(WSN_EINVAL, "bad dataset");
  const size_t P0 = (size_t)n->S * n->S;
  if (count > n->ds_n) {
    if (n->ds_x) {
      // grow: the old buffers stay in the allocation list until the network is destroyed
      n->ds_x = n->ds_t = nullptr;
    }
    CHK(dalloc(n, &n->ds_x, (size_t)count * P0));
    CHK(dalloc(n, &n->ds_t, (size_t)count * P0));
  }
  n->ds_n = count;
  HIPCHK(hipMemcpyAsync(n->ds_x, x, sizeof(float) * count * P0, hipMemcpyHostToDevice, n->st));
  HIPCHK(hipMemcpyAsync(n->ds_t, t, sizeof(float) * count * P0, hipMemcpyHostToDevice, n->st));
  HIPCHK(hipStreamSynchronize(n->st));
  return WSN_OK;
}

int wsn_train(wsn_net* n, const int* order, int n_order, int B, float lr, float* losses) {
  if (!n || !order) return fail(WSN_EINVAL, "null argument");
  CHK(check_batch(n, B));
  if (!n->ds_x) return fail(WSN_EINVAL, "no dataset");
  for (int i = 0; i < n_order; ++i)
    if (order[i] < 0 || order[i] >= n->ds_n) return fail(WSN_EINVAL, "sample index out of range");
  const size_t P0 = (size_t)n->S * n->S;
  const auto t0 = std::chrono::steady_clock::now();
  const int steps = n_order / B;
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < B; ++b) {
      const int k = order[s * B + b];
      HIPCHK(hipMemcpyAsync(n->in + b * P0, n->ds_x + (size_t)k * P0, sizeof(float) * P0,
                            hipMemcpyDeviceToDevice, n->st));
      HIPCHK(hipMemcpyAsync(n->tgt + b * P0, n->ds_t + (size_t)k * P0, sizeof(float) * P0,
                            hipMemcpyDeviceToDevice, n->st));
    }
    float l = 0.f;
    CHK(loss_and_grad(n, B, losses ? &l : nullptr));
    if (losses) losses[s] = l;
    CHK(adam_step(n, lr));
  }
  HIPCHK(hipStreamSynchronize(n->st));
  n->train_seconds +=
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  n->train_samples += (long long)steps * B;
  return WSN_OK;
}

int wsn_train_stats(const wsn_net* n, double* seconds, long long* samples, double* flops) {
  if (!n) return fail(WSN_EINVAL, "null network");
  if (seconds) *seconds = n->train_seconds;
  if (samples) *samples = n->train_samples;
  if (flops) {
    // one sample's forward conv FLOPs (2 per multiply-add); backward is twice that
    double f = 0.0;
    for (int i = 0; i < kNumLayers; ++i) {
      const double pix = (double)(n->S >> level_of(i)) * (n->S >> level_of(i));
      f += 2.0 * pix * kLayers[i].k * kLayers[i].k * kLayers[i].cin * kLayers[i].cout;
    }
    *flops = 3.0 * f;
  }
  return WSN_OK;
}

}  // extern "C"
