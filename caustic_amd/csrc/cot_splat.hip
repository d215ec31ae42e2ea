// Gaussian splatting of ray hits onto the receiver image, forward and point gradients
// (_gaussian_splat_raw_chunk and render_gaussian_splat_two_pass_backward, utils/
// basic_function.py:674-772), for gfx950.  One point per lane: its (2r+1)^2 window of pixels
// around floor((x - x0) / sx), floor((y - y0) / sy), each weighted by
// exp(-(du^2 + dv^2) / (2 sigma^2)) with du, dv the pixel centre's offset in pixels, the weights
// normalised by their sum over the whole window (held constant in the gradient, as the
// reference detaches it); window pixels outside the image are dropped.  FP32 like the
// reference, FP32 atomics into the image.
#include <hip/hip_runtime.h>

namespace cot {

__global__ void splat_kernel(const float* pts, const float* energy, int n, int ry, int rx,
                             float x0, float y0, float sx, float sy, float sigma, int r,
                             float* image) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = (pts[2 * i] - x0) / sx, py = (pts[2 * i + 1] - y0) / sy;
  const int cx = (int)floorf(px), cy = (int)floorf(py);
  const float inv = 1.0f / (2.0f * sigma * sigma);
  float sum = 0.0f;
  for (int dy = -r; dy <= r; ++dy)
    for (int dx = -r; dx <= r; ++dx) {
      const float u = cx + dx + 0.5f - px, v = cy + dy + 0.5f - py;
      sum += expf(-(u * u + v * v) * inv);
    }
  const float scale = energy[i] / fmaxf(sum, 1e-12f);
  for (int dy = -r; dy <= r; ++dy) {
    const int y = cy + dy;
    if (y < 0 || y >= ry) continue;
    for (int dx = -r; dx <= r; ++dx) {
      const int x = cx + dx;
      if (x < 0 || x >= rx) continue;
      const float u = x + 0.5f - px, v = y + 0.5f - py;
      atomicAdd(&image[(size_t)y * rx + x], expf(-(u * u + v * v) * inv) * scale);
    }
  }
}

// d/d(point) of sum_p c * raw_p * G_p (raw the splat image of the points, G the image gradient):
// for window pixel k, d w_k / dx = w_k u_k / (sigma^2 sx) with the normalisation held fixed.
__global__ void splat_grad_kernel(const float* pts, const float* energy, int n, int ry, int rx,
                                  float x0, float y0, float sx, float sy, float sigma, int r,
                                  const float* G, float c, float* grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = (pts[2 * i] - x0) / sx, py = (pts[2 * i + 1] - y0) / sy;
  const int cx = (int)floorf(px), cy = (int)floorf(py);
  const float inv = 1.0f / (2.0f * sigma * sigma);
  float sum = 0.0f;
  for (int dy = -r; dy <= r; ++dy)
    for (int dx = -r; dx <= r; ++dx) {
      const float u = cx + dx + 0.5f - px, v = cy + dy + 0.5f - py;
      sum += expf(-(u * u + v * v) * inv);
    }
  const float scale = c * energy[i] / fmaxf(sum, 1e-12f);
  float gx = 0.0f, gy = 0.0f;
  for (int dy = -r; dy <= r; ++dy) {
    const int y = cy + dy;
    if (y < 0 || y >= ry) continue;
    for (int dx = -r; dx <= r; ++dx) {
      const int x = cx + dx;
      if (x < 0 || x >= rx) continue;
      const float u = x + 0.5f - px, v = y + 0.5f - py;
      const float t = expf(-(u * u + v * v) * inv) * scale * G[(size_t)y * rx + x];
      gx += t * u;
      gy += t * v;
    }
  }
  const float s2 = sigma * sigma;
  grad[2 * i] = gx / (s2 * sx);
  grad[2 * i + 1] = gy / (s2 * sy);
}

hipError_t launch_splat(const float* pts, const float* energy, int n, int ry, int rx,
                        const double dom[4], float sigma, int r, const float* G, float c,
                        float* out, hipStream_t st) {
  const float sx = (float)((dom[1] - dom[0]) / rx), sy = (float)((dom[3] - dom[2]) / ry);
  if (n <= 0) return hipSuccess;
  if (G)
    hipLaunchKernelGGL(splat_grad_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pts, energy, n,
                       ry, rx, (float)dom[0], (float)dom[2], sx, sy, sigma, r, G, c, out);
  else
    hipLaunchKernelGGL(splat_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pts, energy, n, ry,
                       rx, (float)dom[0], (float)dom[2], sx, sy, sigma, r, out);
  return hipGetLastError();
}

}  // namespace cot
