// Diagnostic outputs of the evaluation (the reference's `render` and `writeOBJ`,
// basic_function.cuh:660-741, 15-59), for gfx950.
//
// render_kernel spreads each site's source mass nu_i uniformly over its cell: pixel p receives
// nu_i |C_i n p| / |C_i|, the pieces found by clipping the cell (FP64 vertices of the last
// evaluation) against each pixel of its bounding box (Sutherland-Hodgman) and accumulated with
// FP64 atomics.  One site per lane; a diagnostic, not a hot path.
#include <mutex>
#include <string>

#include "../../include/caustic_ot.h"
#include "cot_internal.h"

namespace cot {

namespace {

constexpr int kClipMax = kMaxVerts + 8;

// Clip polygon (n vertices in x, y) against the half-plane given by side s of the rectangle
// [x0, x1] x [y0, y1] (0: y >= y0, 1: x <= x1, 2: y <= y1, 3: x >= x0); result in ox, oy.
__device__ int clip_side(const double* x, const double* y, int n, int s, double x0, double x1,
                         double y0, double y1, double* ox, double* oy) {
  auto inside = [&](double px, double py) {
    switch (s) {
      case 0: return py >= y0;
      case 1: return px <= x1;
      case 2: return py <= y1;
      default: return px >= x0;
    }
  };
  auto cross = [&](double ax, double ay, double bx, double by, double& cx, double& cy) {
    if (s == 0 || s == 2) {
      const double yc = s == 0 ? y0 : y1;
      const double t = (yc - ay) / (by - ay);
      cx = ax + t * (bx - ax);
      cy = yc;
    } else {
      const double xc = s == 1 ? x1 : x0;
      const double t = (xc - ax) / (bx - ax);
      cx = xc;
      cy = ay + t * (by - ay);
    }
  };
  int m = 0;
  for (int k = 0; k < n; ++k) {
    const double ax = x[k], ay = y[k];
    const double bx = x[k + 1 < n ? k + 1 : 0], by = y[k + 1 < n ? k + 1 : 0];
    const bool ia = inside(ax, ay), ib = inside(bx, by);
    if (ia) {
      if (m < kClipMax) { ox[m] = ax; oy[m] = ay; ++m; }
      if (!ib && m < kClipMax) { cross(ax, ay, bx, by, ox[m], oy[m]); ++m; }
    } else if (ib && m < kClipMax) {
      cross(ax, ay, bx, by, ox[m], oy[m]);
      ++m;
    }
  }
  return m;
}

__device__ double shoelace(const double* x, const double* y, int n) {
  double a = 0.0;
  for (int k = 0; k < n; ++k) {
    const int j = k + 1 < n ? k + 1 : 0;
    a += x[k] * y[j] - x[j] * y[k];
  }
  return 0.5 * fabs(a);
}

}  // namespace

__global__ __launch_bounds__(kCellBlock) void render_kernel(const ProbDesc* probs,
                                                            const BlockDesc* blocks,
                                                            const double2* vtx,
                                                            const uint8_t* nedge,
                                                            const double* nu, double* image) {
  const BlockDesc bd = blocks[blockIdx.x];
  if ((int)threadIdx.x >= bd.count) return;
  const ProbDesc pd = probs[bd.prob];
  const int i = bd.first + threadIdx.x;
  const int nv = nedge[i];
  if (nv < 3) return;
  double px[kMaxVerts], py[kMaxVerts];
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int k = 0; k < nv; ++k) {
    const double2 v = vtx[(size_t)i * kMaxVerts + k];
    px[k] = v.x;
    py[k] = v.y;
    mnx = fmin(mnx, v.x); mxx = fmax(mxx, v.x);
    mny = fmin(mny, v.y); mxy = fmax(mxy, v.y);
  }
  const double area = shoelace(px, py, nv);
  if (!(area > 0.0)) return;
  const int ix0 = max((int)floor((mnx - pd.x0) / pd.dx), 0);
  const int iy0 = max((int)floor((mny - pd.y0) / pd.dy), 0);
  const int ix1 = min((int)floor((mxx - pd.x0) / pd.dx) + 1, pd.pm);
  const int iy1 = min((int)floor((mxy - pd.y0) / pd.dy) + 1, pd.pn);
  const double scale = nu[i] / area;
  double* img = image + pd.pre_off;
  double ax[kClipMax], ay[kClipMax], bx[kClipMax], by[kClipMax];
  for (int iy = iy0; iy < iy1; ++iy) {
    const double y0 = pd.y0 + iy * pd.dy, y1 = pd.y0 + (iy + 1) * pd.dy;
    for (int ix = ix0; ix < ix1; ++ix) {
      const double x0 = pd.x0 + ix * pd.dx, x1 = pd.x0 + (ix + 1) * pd.dx;
      int n = clip_side(px, py, nv, 0, x0, x1, y0, y1, ax, ay);
      n = clip_side(ax, ay, n, 1, x0, x1, y0, y1, bx, by);
      n = clip_side(bx, by, n, 2, x0, x1, y0, y1, ax, ay);
      n = clip_side(ax, ay, n, 3, x0, x1, y0, y1, bx, by);
      if (n < 3) continue;
      const double a = shoelace(bx, by, n);
      if (a > 0.0) atomicAdd(&img[(size_t)iy * pd.pm + ix], scale * a);
    }
  }
}

// Irradiance image of a triangle mesh (compute_render of the irra_render extension,
// render_exe_program/render.cu:93-171): face f spreads tri_energy[f] * gray uniformly over its
// triangle, pixel (ix, iy) of the pm x pn grid over [x0, x1] x [y0, y1] receiving
// e_f |T_f n p| / |T_f|.  The image is written in the extension's output layout, row
// pm - 1 - ix, column iy of a pm x pn array.  FP32 in and out; clipping and areas in FP64.
// One face per lane.
__global__ void tri_render_kernel(const float* energy, const int* faces, int n_faces,
                                  const float* xy, float gray, double x0, double y0, double dx,
                                  double dy, int pm, int pn, float* image) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  double px[3], py[3];
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int k = 0; k < 3; ++k) {
    const int v = faces[3 * f + k];
    px[k] = xy[2 * v];
    py[k] = xy[2 * v + 1];
    mnx = fmin(mnx, px[k]); mxx = fmax(mxx, px[k]);
    mny = fmin(mny, py[k]); mxy = fmax(mxy, py[k]);
  }
  const double area = shoelace(px, py, 3);
  if (!(area > 0.0)) return;
  const int ix0 = max((int)floor((mnx - x0) / dx), 0);
  const int iy0 = max((int)floor((mny - y0) / dy), 0);
  const int ix1 = min((int)floor((mxx - x0) / dx) + 1, pm);
  const int iy1 = min((int)floor((mxy - y0) / dy) + 1, pn);
  const double scale = (double)energy[f] * gray / area;
  double ax[kClipMax], ay[kClipMax], bx[kClipMax], by[kClipMax];
  for (int ix = ix0; ix < ix1; ++ix) {
    const double cx0 = x0 + ix * dx, cx1 = x0 + (ix + 1) * dx;
    float* row = image + (size_t)(pm - 1 - ix) * pn;
    for (int iy = iy0; iy < iy1; ++iy) {
      const double cy0 = y0 + iy * dy, cy1 = y0 + (iy + 1) * dy;
      int n = clip_side(px, py, 3, 0, cx0, cx1, cy0, cy1, ax, ay);
      n = clip_side(ax, ay, n, 1, cx0, cx1, cy0, cy1, bx, by);
      n = clip_side(bx, by, n, 2, cx0, cx1, cy0, cy1, ax, ay);
      n = clip_side(ax, ay, n, 3, cx0, cx1, cy0, cy1, bx, by);
      if (n < 3) continue;
      const double a = shoelace(bx, by, n);
      if (a > 0.0) atomicAdd(&row[iy], (float)(scale * a));
    }
  }
}

// diff = render - real, render and diff in the output layout (row pm - 1 - ix, column iy), real
// in the grid layout iy * pm + ix that the extension's kernels read.
__global__ void tri_diff_kernel(const float* render, const float* real, int pm, int pn,
                                float* diff) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)pm * pn) return;
  const int ix = pm - 1 - (int)(k / pn), iy = (int)(k % pn);
  diff[k] = render[k] - real[(size_t)iy * pm + ix];
}

// Sobel weights of compute_sobel_weight (render_exe_program/render.cu:63-91), output layout:
// w_p = sum over the 4-neighbours q inside the image of (diff_p - diff_q), the graph Laplacian
// of the difference image (so that sum_p 2 w_p d render_p is the derivative of
// sum over neighbouring pairs of (diff_p - diff_q)^2).
__global__ void tri_sobel_kernel(const float* diff, int pm, int pn, float* w) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)pm * pn) return;
  const int r = (int)(k / pn), c = (int)(k % pn);
  const float d = diff[k];
  float s = 0.0f;
  if (r > 0) s += d - diff[k - pn];
  if (r + 1 < pm) s += d - diff[k + pn];
  if (c > 0) s += d - diff[k - 1];
  if (c + 1 < pn) s += d - diff[k + 1];
  w[k] = s;
}

namespace {

// Parameter range [t0, t1] of the segment a + t d (t in [0, 1]) inside the pixel
// [x0, x1) x [y0, y1) (Liang-Barsky); false when it misses.  The pixel is half-open -- closed on
// the image's last column (close_x) and row (close_y) -- so that a segment lying on a grid line
// belongs to one pixel only: the boundary transport then sums over the pixels to the
// triangle's (the derivative for moving the segment towards +x or +y).
__device__ bool segment_in_rect(double ax, double ay, double dxs, double dys, double x0, double x1,
                                double y0, double y1, bool close_x, bool close_y, double& t0,
                                double& t1) {
  t0 = 0.0;
  t1 = 1.0;
  const double p[4] = {-dxs, dxs, -dys, dys};
  const double q[4] = {ax - x0, x1 - ax, ay - y0, y1 - ay};
  const bool closed[4] = {true, close_x, true, close_y};
  for (int k = 0; k < 4; ++k) {
    if (p[k] == 0.0) {
      if (q[k] < 0.0 || (q[k] == 0.0 && !closed[k])) return false;
      continue;
    }
    const double r = q[k] / p[k];
    if (p[k] < 0.0) {
      if (r > t1) return false;
      t0 = fmax(t0, r);
    } else {
      if (r < t0) return false;
      t1 = fmin(t1, r);
    }
  }
  return t1 > t0;
}

}  // namespace

// Gradients of the extension's render_grad (render_exe_program/render.cu:173-288, 422-542,
// 593-728; sobel off) for the loss sum_p (render_p - real_p)^2 and the per-face terms
// (inte_f - e_f)^2, inte_f = sum_p |T_f n p| real_p / |p|.  One face per lane.  The
// derivative of a clipped area |T n p| with respect to vertex v_k is the transport of the
// triangle's boundary: sum over the edges at v_k of N_e * integral over the part of the edge
// inside p of the hat function of v_k (N_e the outward edge normal scaled by the edge length),
// so each pixel needs the area of one clip and three segment clips; d|T|/dv_k follows with the
// whole edges.  Vertex gradients are added with FP32 atomics (vertices are shared by faces).
__global__ void tri_grad_kernel(const float* energy, const int* faces, int n_faces,
                                const float* xy, float gray, const float* real, const float* diff,
                                double x0, double y0, double dx, double dy, int pm, int pn,
                                float* xy_render_grad, float* tri_render_grad, float* inte_diff,
                                float* xy_inte_grad, float* inte_result, const float* sob,
                                float* xy_sobel_grad, float* tri_sobel_grad) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  int vid[3];
  double px[3], py[3];
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int k = 0; k < 3; ++k) {
    vid[k] = faces[3 * f + k];
    px[k] = xy[2 * vid[k]];
    py[k] = xy[2 * vid[k] + 1];
    mnx = fmin(mnx, px[k]); mxx = fmax(mxx, px[k]);
    mny = fmin(mny, py[k]); mxy = fmax(mxy, py[k]);
  }
  const double e = (double)energy[f] * gray;
  const double signed2 = (px[1] - px[0]) * (py[2] - py[0]) - (px[2] - px[0]) * (py[1] - py[0]);
  const double A = 0.5 * fabs(signed2);
  const double orient = signed2 >= 0.0 ? 1.0 : -1.0;
  // outward edge normals times edge lengths, edge k from vertex k to vertex k + 1
  double nx[3], ny[3];
  for (int k = 0; k < 3; ++k) {
    const int j = k == 2 ? 0 : k + 1;
    nx[k] = orient * (py[j] - py[k]);
    ny[k] = -orient * (px[j] - px[k]);
  }
  double dAx[3], dAy[3];
  for (int k = 0; k < 3; ++k) {
    const int km = k == 0 ? 2 : k - 1;  // the edge ending at vertex k
    dAx[k] = 0.5 * (nx[k] + nx[km]);
    dAy[k] = 0.5 * (ny[k] + ny[km]);
  }
  const int ix0 = max((int)floor((mnx - x0) / dx), 0);
  const int iy0 = max((int)floor((mny - y0) / dy), 0);
  const int ix1 = min((int)floor((mxx - x0) / dx) + 1, pm);
  const int iy1 = min((int)floor((mxy - y0) / dy) + 1, pn);
  const double cell = dx * dy;
  double gr[6] = {0, 0, 0, 0, 0, 0}, gi[6] = {0, 0, 0, 0, 0, 0}, gs[6] = {0, 0, 0, 0, 0, 0};
  double tri_g = 0.0, tri_s = 0.0, inte = 0.0;
  double ax[kClipMax], ay[kClipMax], bx[kClipMax], by[kClipMax];
  for (int ix = ix0; ix < ix1; ++ix) {
    const double cx0 = x0 + ix * dx, cx1 = x0 + (ix + 1) * dx;
    for (int iy = iy0; iy < iy1; ++iy) {
      const double cy0 = y0 + iy * dy, cy1 = y0 + (iy + 1) * dy;
      int n = clip_side(px, py, 3, 0, cx0, cx1, cy0, cy1, ax, ay);
      n = clip_side(ax, ay, n, 1, cx0, cx1, cy0, cy1, bx, by);
      n = clip_side(bx, by, n, 2, cx0, cx1, cy0, cy1, ax, ay);
      n = clip_side(ax, ay, n, 3, cx0, cx1, cy0, cy1, bx, by);
      const double a = n >= 3 ? shoelace(bx, by, n) : 0.0;
      // d a / d v_k from the triangle's edges inside the pixel
      double dax[3] = {0, 0, 0}, day[3] = {0, 0, 0};
      bool any = false;
      for (int k = 0; k < 3; ++k) {
        const int j = k == 2 ? 0 : k + 1;
        double t0, t1;
        if (!segment_in_rect(px[k], py[k], px[j] - px[k], py[j] - py[k], cx0, cx1, cy0, cy1,
                             ix == pm - 1, iy == pn - 1, t0, t1))
          continue;
        const double i0 = t1 - t0, i1 = 0.5 * (t1 * t1 - t0 * t0);
        dax[k] += (i0 - i1) * nx[k];
        day[k] += (i0 - i1) * ny[k];
        dax[j] += i1 * nx[k];
        day[j] += i1 * ny[k];
        any = true;
      }
      if (!(a > 0.0) && !any) continue;
      // the per-face integral takes the face's orientation, as the reference's signed clipped
      // areas do (compute_grad_area, render_exe_program/clip.h:277-292): negative for a
      // clockwise face, and so are its derivatives; the render terms below use a / A, where
      // the sign cancels
      const double rp = (double)real[(size_t)iy * pm + ix] / cell;
      inte += orient * a * rp;
      for (int k = 0; k < 3; ++k) {
        gi[2 * k] += 2.0 * rp * orient * dax[k];
        gi[2 * k + 1] += 2.0 * rp * orient * day[k];
      }
      if (A > 0.0) {
        const size_t q = (size_t)(pm - 1 - ix) * pn + iy;
        const double d2 = 2.0 * (double)diff[q];
        tri_g += d2 * a / A;
        const double s = d2 * e / (A * A);
        for (int k = 0; k < 3; ++k) {
          gr[2 * k] += s * (dax[k] * A - a * dAx[k]);
          gr[2 * k + 1] += s * (day[k] * A - a * dAy[k]);
        }
        if (sob) {
          const double w2 = 2.0 * (double)sob[q];
          tri_s += w2 * a / A;
          const double sw = w2 * e / (A * A);
          for (int k = 0; k < 3; ++k) {
            gs[2 * k] += sw * (dax[k] * A - a * dAx[k]);
            gs[2 * k + 1] += sw * (day[k] * A - a * dAy[k]);
          }
        }
      }
    }
  }
  tri_render_grad[f] = (float)(tri_g * gray);
  if (sob) {
    tri_sobel_grad[f] = (float)(tri_s * gray);
    for (int k = 0; k < 3; ++k) {
      atomicAdd(&xy_sobel_grad[2 * vid[k]], (float)gs[2 * k]);
      atomicAdd(&xy_sobel_grad[2 * vid[k] + 1], (float)gs[2 * k + 1]);
    }
  }
  inte_result[f] = (float)inte;
  inte_diff[f] = (float)(inte - e);
  for (int k = 0; k < 3; ++k) {
    atomicAdd(&xy_render_grad[2 * vid[k]], (float)gr[2 * k]);
    atomicAdd(&xy_render_grad[2 * vid[k] + 1], (float)gr[2 * k + 1]);
    atomicAdd(&xy_inte_grad[2 * vid[k]], (float)((inte - e) * gi[2 * k]));
    atomicAdd(&xy_inte_grad[2 * vid[k] + 1], (float)((inte - e) * gi[2 * k + 1]));
  }
}

__global__ void fill_f32_kernel(float* p, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0f;
}

hipError_t launch_tri_render(const float* energy, const int* faces, int n_faces, const float* xy,
                             float gray, const double dom[4], int pm, int pn, float* image,
                             hipStream_t st) {
  const size_t px = (size_t)pm * pn;
  hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, st, image,
                     px);
  if (n_faces > 0)
    hipLaunchKernelGGL(tri_render_kernel, dim3((n_faces + 255) / 256), dim3(256), 0, st, energy,
                       faces, n_faces, xy, gray, dom[0], dom[2], (dom[1] - dom[0]) / pm,
                       (dom[3] - dom[2]) / pn, pm, pn, image);
  return hipGetLastError();
}

hipError_t launch_tri_grad(const float* energy, const int* faces, int n_faces, const float* xy,
                           int n_points, float gray, const float* real, const double dom[4],
                           int pm, int pn, float* render, float* diff, float* xy_render_grad,
                           float* tri_render_grad, float* inte_diff, float* xy_inte_grad,
                           float* inte_result, float* sob, float* xy_sobel_grad,
                           float* tri_sobel_grad, hipStream_t st) {
  const hipError_t e0 = launch_tri_render(energy, faces, n_faces, xy, gray, dom, pm, pn, render, st);
  if (e0 != hipSuccess) return e0;
  const size_t px = (size_t)pm * pn, nv = 2 * (size_t)n_points;
  hipLaunchKernelGGL(tri_diff_kernel, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, st, render,
                     real, pm, pn, diff);
  if (sob)
    hipLaunchKernelGGL(tri_sobel_kernel, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, st,
                       diff, pm, pn, sob);
  if (nv > 0) {
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st,
                       xy_render_grad, nv);
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st,
                       xy_inte_grad, nv);
    if (sob)
      hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st,
                         xy_sobel_grad, nv);
  }
  if (n_faces > 0)
    hipLaunchKernelGGL(tri_grad_kernel, dim3((n_faces + 255) / 256), dim3(256), 0, st, energy,
                       faces, n_faces, xy, gray, real, diff, dom[0], dom[2],
                       (dom[1] - dom[0]) / pm, (dom[3] - dom[2]) / pn, pm, pn, xy_render_grad,
                       tri_render_grad, inte_diff, xy_inte_grad, inte_result, sob, xy_sobel_grad,
                       tri_sobel_grad);
  return hipGetLastError();
}

// Solid angle under which a point light at (lx, ly, lz) sees each cell placed in the plane z:
// the cell fanned from its vertex 0 into triangles, each a spherical triangle of the unit
// vectors a, b, c towards its corners, Omega = 2 atan(a.(b x c) / (1 + a.b + b.c + c.a))
// (Van Oosterom and Strackee 1983), a triangle with |denominator| < 1e-6 skipped, and the
// magnitude of the sum taken (compute_solid_angle, basic_function.cuh:558-588).  Cells with
// fewer than 3 vertices get 0.  One site per lane.
__global__ __launch_bounds__(kCellBlock) void solid_angle_kernel(const BlockDesc* blocks,
                                                                 const double2* vtx,
                                                                 const uint8_t* nedge, double lx,
                                                                 double ly, double lz, double z,
                                                                 double* omega) {
  const BlockDesc bd = blocks[blockIdx.x];
  if ((int)threadIdx.x >= bd.count) return;
  const int i = bd.first + threadIdx.x;
  const int nv = nedge[i];
  double sum = 0.0;
  if (nv >= 3) {
    const double2* v = vtx + (size_t)i * kMaxVerts;
    const double dz = z - lz;
    auto dir = [&](double2 p, double& x, double& y, double& zz) {
      const double ex = p.x - lx, ey = p.y - ly;
      const double r = sqrt(ex * ex + ey * ey + dz * dz);
      x = ex / r;
      y = ey / r;
      zz = dz / r;
    };
    double ax, ay, az, bx, by, bz, cx, cy, cz;
    dir(v[0], ax, ay, az);
    dir(v[1], bx, by, bz);
    for (int k = 2; k < nv; ++k) {
      dir(v[k], cx, cy, cz);
      const double num = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) +
                         az * (bx * cy - by * cx);
      const double den = 1.0 + (ax * bx + ay * by + az * bz) + (bx * cx + by * cy + bz * cz) +
                         (ax * cx + ay * cy + az * cz);
      if (fabs(den) >= 1e-6) sum += 2.0 * atan(num / den);
      bx = cx;
      by = cy;
      bz = cz;
    }
  }
  omega[i] = fabs(sum);
}

hipError_t launch_solid_angle(const BlockDesc* blocks, const double2* vtx, const uint8_t* nedge,
                              const double light[3], double z, double* omega, int nblk,
                              hipStream_t st) {
  hipLaunchKernelGGL(solid_angle_kernel, dim3(nblk), dim3(kCellBlock), 0, st, blocks, vtx, nedge,
                     light[0], light[1], light[2], z, omega);
  return hipGetLastError();
}

__global__ void fill_f64_kernel(double* p, size_t n, double v) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

hipError_t launch_render(const ProbDesc* probs, const BlockDesc* blocks, const double2* vtx,
                         const uint8_t* nedge, const double* nu, double* image, size_t pixels,
                         int nblk, hipStream_t st) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, st,
                     image, pixels, 0.0);
  hipLaunchKernelGGL(render_kernel, dim3(nblk), dim3(kCellBlock), 0, st, probs, blocks, vtx, nedge,
                     nu, image);
  return hipGetLastError();
}

}  // namespace cot

// ---------------------------------------------------------------------------------------------
// cot_irra_render_grad_noalloc: render_grad with no device allocation
// ---------------------------------------------------------------------------------------------

namespace cot {

int set_error(int code, const std::string& msg);

namespace {

// The entry point's device scratch, a file-scope array that each GPU holds from the module's
// load to the end of the process: bytes [0, kScratchHead) hold the count of face indices that
// name no point, the rest the Sobel weight image (at most COT_NOALLOC_MAX_PIXELS floats).  In
// host code `scratch` is only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device (scratch_on_current_device).
constexpr size_t kScratchHead = 256;
static __device__ __align__(256) unsigned char
    scratch[kScratchHead + sizeof(float) * (size_t)COT_NOALLOC_MAX_PIXELS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

// The scratch's address on the current device (looked up on the first call there).
int scratch_on_current_device(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(COT_EHIP, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return set_error(COT_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return set_error(COT_EHIP, "hipGetSymbolAddress(scratch) failed");
    g_scratch[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[dev];
  return COT_OK;
}

__global__ void clear_count_kernel(int* count) { *count = 0; }

// count += the face indices outside [0, n_points)
__global__ void face_check_kernel(const int* faces, size_t n3, int n_points, int* count) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  const int v = faces[k];
  if (v < 0 || v >= n_points) atomicAdd(count, 1);
}

}  // namespace
}  // namespace cot

extern "C" int cot_irra_render_grad_noalloc(
    const float* tri_energy, const int* faces, int n_faces, const float* xy, int n_points,
    float total_gray, const float* real_picture, const double* proj_domain, int pm, int pn,
    float* render, float* render_diff, float* xy_render_grad, float* tri_render_grad,
    float* inte_diff, float* xy_inte_grad, float* inte_result, int use_sobel,
    float* xy_sobel_grad, float* tri_sobel_grad, void* stream) {
  using namespace cot;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!real_picture || !render || !render_diff || n_faces < 0 || n_points < 0 || pm <= 0 ||
      pn <= 0 || !proj_domain || (n_faces > 0 && (!tri_energy || !faces || !xy)) ||
      (n_points > 0 && (!xy_render_grad || !xy_inte_grad || (use_sobel && !xy_sobel_grad))) ||
      (n_faces > 0 && (!tri_render_grad || !inte_diff || !inte_result ||
                       (use_sobel && !tri_sobel_grad))))
    return set_error(COT_EINVAL, "bad argument");
  if (!(proj_domain[1] > proj_domain[0] && proj_domain[3] > proj_domain[2]))
    return set_error(COT_EINVAL, "empty proj_domain");
  if (use_sobel && (size_t)pm * (size_t)pn > (size_t)COT_NOALLOC_MAX_PIXELS)
    return set_error(COT_EINVAL, "image of " + std::to_string((size_t)pm * pn) +
                                     " pixels exceeds the allocation-free limit of " +
                                     std::to_string(COT_NOALLOC_MAX_PIXELS));
  unsigned char* base = nullptr;
  const int rc = scratch_on_current_device(&base);
  if (rc != COT_OK) return rc;
  int* count = reinterpret_cast<int*>(base);
  float* sob = use_sobel ? reinterpret_cast<float*>(base + kScratchHead) : nullptr;
  // every vertex index must name a point before a kernel dereferences it: counted on the
  // device, the count read back into a host variable
  if (n_faces > 0) {
    const size_t n3 = 3 * (size_t)n_faces;
    int bad = 0;
    hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(1), 0, st, count);
    hipLaunchKernelGGL(face_check_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st,
                       faces, n3, n_points, count);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bad, count, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return set_error(COT_EHIP, "face check failed");
    if (bad > 0) return set_error(COT_EINVAL, "face index out of range");
  }
  if (launch_tri_grad(tri_energy, faces, n_faces, xy, n_points, total_gray, real_picture,
                      proj_domain, pm, pn, render, render_diff, xy_render_grad, tri_render_grad,
                      inte_diff, xy_inte_grad, inte_result, sob, xy_sobel_grad, tri_sobel_grad,
                      st) != hipSuccess)
    return set_error(COT_EHIP, "render_grad launch failed");
  return COT_OK;
}
