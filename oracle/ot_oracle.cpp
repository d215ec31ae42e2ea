// ORACLE -- test infrastructure only.
//
// FP64 CPU restatement of the reference's semi-discrete optimal-transport evaluation
// (rmk523/caustic, Optimal_Transport_rl/).  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker: the
// product path (caustic_amd/) never calls it.
//
// What each function follows:
//   clip_rect        clip.cuh:76-149       Sutherland-Hodgman against an axis-aligned rectangle,
//                                          edge order y>ymin, x<xmax, y<ymax, x>xmin, strict tests
//   fan_area         basic_function.cuh:527-539   fan shoelace, absolute value
//   fan_centroid     basic_function.cuh:590-608   area-weighted fan-triangle centroid
//   second_moment    basic_function.cuh:826-865   integral of |x-s|^2 over a polygon (the oracle uses
//                                          the triangle-fan closed form, equal to the reference's
//                                          per-edge Green polynomial; checked in tests)
//   orc_power_cells  power_diagram.cpp:8-134 (semantics only): the cell of site i is the set of
//                    points of the bounding box whose power |x-s_i|^2-w_i is not larger than any
//                    other site's.  The reference builds it with CGAL's regular triangulation plus
//                    mirrored boundary sites; inside proj_domain both constructions give the same
//                    cells.  Here each cell is the bounding box clipped by the half-planes of its
//                    candidate sites (all sites, or a caller-supplied neighbour list).
//   orc_cell_integrals  compute_inte (basic_function.cuh:743-824), compute_f (:867-952),
//                    compute_center weighted / geometric (:1056-1191)
//   orc_clip_area    compute_clip_area (basic_function.cuh:610-633)
//   orc_render       render (basic_function.cuh:660-741), without the gray_multiple factor
//   orc_fresnel_rate compute_energy_remain_rate / fresnel_T_unpolarized (basic_function.cuh:407-494)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

using Poly = std::vector<double>;  // flat x0,y0,x1,y1,...

// rect = {xmin, ymin, xmax, ymax} as in the reference's `square`.
inline bool rect_inside(double x, double y, const double* r, int side) {
  switch (side) {
    case 0: return y > r[1];
    case 1: return x < r[2];
    case 2: return y < r[3];
    default: return x > r[0];
  }
}

inline void rect_cross(double sx, double sy, double ex, double ey, const double* r, int side,
                       double& ox, double& oy) {
  if (side == 0 || side == 2) {
    double yc = side == 0 ? r[1] : r[3];
    ox = (sx - ex) * (yc - ey) / (sy - ey) + ex;
    oy = yc;
  } else {
    double xc = side == 1 ? r[2] : r[0];
    oy = (sy - ey) * (xc - ex) / (sx - ex) + ey;
    ox = xc;
  }
}

void clip_rect(const double* in, int n, const double* r, Poly& out) {
  Poly cur(in, in + 2 * n), nxt;
  for (int side = 0; side < 4; ++side) {
    nxt.clear();
    int m = (int)cur.size() / 2;
    for (int i = 0; i < m; ++i) {
      int k = (i + 1) % m;
      double sx = cur[2 * i], sy = cur[2 * i + 1], ex = cur[2 * k], ey = cur[2 * k + 1];
      bool ein = rect_inside(ex, ey, r, side), sin_ = rect_inside(sx, sy, r, side);
      if (ein) {
        if (!sin_) {
          double ox, oy;
          rect_cross(sx, sy, ex, ey, r, side, ox, oy);
          nxt.push_back(ox);
          nxt.push_back(oy);
        }
        nxt.push_back(ex);
        nxt.push_back(ey);
      } else if (sin_) {
        double ox, oy;
        rect_cross(sx, sy, ex, ey, r, side, ox, oy);
        nxt.push_back(ox);
        nxt.push_back(oy);
      }
    }
    cur.swap(nxt);
  }
  out.swap(cur);
}

double fan_area(const double* p, int n) {
  double a = 0.0;
  for (int i = 1; i < n - 1; ++i) {
    double bx = p[2 * i] - p[0], by = p[2 * i + 1] - p[1];
    double cx = p[2 * i + 2] - p[0], cy = p[2 * i + 3] - p[1];
    a += 0.5 * (bx * cy - by * cx);
  }
  return std::fabs(a);
}

void fan_centroid(const double* p, int n, double& cx, double& cy) {
  double big = fan_area(p, n);
  cx = cy = 0.0;
  for (int i = 0; i < n - 2; ++i) {
    double t[6] = {p[0], p[1], p[2 * i + 2], p[2 * i + 3], p[2 * i + 4], p[2 * i + 5]};
    double a = fan_area(t, 3);
    cx += (t[0] + t[2] + t[4]) / 3.0 * a;
    cy += (t[1] + t[3] + t[5]) / 3.0 * a;
  }
  cx /= big;
  cy /= big;
}

// Signed integral of |x - s|^2 over polygon p (positive for counter-clockwise order).
double second_moment(const double* p, int n, double sx, double sy) {
  double acc = 0.0;
  double ax = p[0] - sx, ay = p[1] - sy;
  for (int i = 1; i < n - 1; ++i) {
    double bx = p[2 * i] - sx, by = p[2 * i + 1] - sy;
    double cx = p[2 * i + 2] - sx, cy = p[2 * i + 3] - sy;
    double at = 0.5 * ((bx - ax) * (cy - ay) - (by - ay) * (cx - ax));
    double q = ax * ax + ay * ay + bx * bx + by * by + cx * cx + cy * cy + ax * bx + ay * by +
               bx * cx + by * cy + cx * ax + cy * ay;
    acc += at * q / 6.0;
  }
  return acc;
}

// Keep the part of convex polygon `poly` (coordinates relative to the site) with u.d <= c.
void clip_halfplane(Poly& poly, double dx, double dy, double c, Poly& tmp) {
  int n = (int)poly.size() / 2;
  if (n == 0) return;
  tmp.clear();
  for (int i = 0; i < n; ++i) {
    int k = (i + 1) % n;
    double ux = poly[2 * i], uy = poly[2 * i + 1], vx = poly[2 * k], vy = poly[2 * k + 1];
    double su = ux * dx + uy * dy - c, sv = vx * dx + vy * dy - c;
    bool uin = su <= 0.0, vin = sv <= 0.0;
    if (uin) {
      tmp.push_back(ux);
      tmp.push_back(uy);
    }
    if (uin != vin) {
      double t = su / (su - sv);
      tmp.push_back(ux + t * (vx - ux));
      tmp.push_back(uy + t * (vy - uy));
    }
  }
  if (tmp.size() < 6) tmp.clear();
  poly.swap(tmp);
}

struct PixGrid {
  double x0, x1, y0, y1, dx, dy, area;
  int m, n;
  PixGrid(const double* proj, int pm, int pn) {
    x0 = proj[0];
    x1 = proj[1];
    y0 = proj[2];
    y1 = proj[3];
    m = pm;
    n = pn;
    dx = (x1 - x0) / (double)pm;
    dy = (y1 - y0) / (double)pn;
    area = dx * dy;
  }
  void bbox(const Poly& p, int& lx, int& ly, int& hx, int& hy) const {
    double mnx = p[0], mxx = p[0], mny = p[1], mxy = p[1];
    for (size_t i = 0; i < p.size(); i += 2) {
      mnx = std::min(mnx, p[i]);
      mxx = std::max(mxx, p[i]);
      mny = std::min(mny, p[i + 1]);
      mxy = std::max(mxy, p[i + 1]);
    }
    lx = (int)std::floor((mnx - x0) / dx);
    ly = (int)std::floor((mny - y0) / dy);
    hx = (int)std::floor((mxx - x0) / dx) + 1;
    hy = (int)std::floor((mxy - y0) / dy) + 1;
    lx = std::max(lx, 0);
    ly = std::max(ly, 0);
    hx = std::min(hx, m);
    hy = std::min(hy, n);
  }
  void pixel(int ix, int iy, double* r) const {
    r[0] = ix * dx + x0;
    r[1] = iy * dy + y0;
    r[2] = (ix + 1) * dx + x0;
    r[3] = (iy + 1) * dy + y0;
  }
};

// Pieces at or below this area are left out of the rho-weighted centroid
// (basic_function.cuh:1153 uses 1e-10); 0 keeps every piece of positive area.
double g_centroid_min_piece = 1e-10;

}  // namespace

extern "C" {

void orc_set_centroid_min_piece(double a) { g_centroid_min_piece = a; }

// Restricted power cells.  box = {x0, x1, y0, y1}.  If nbr_ptr is NULL every other site is a
// candidate (sorted by distance, stopped by the exact weight-aware bound); otherwise the
// candidates of site i are nbr_idx[nbr_ptr[i] .. nbr_ptr[i+1]).  Output is CCW polygons in the
// reference's flattened layout (start/len in vertices, verts as x,y pairs).  Returns the total
// vertex count, or -1 when it would exceed cap_total.
long long orc_power_cells(int n, const double* xy, const double* w, const double* box,
                          const int* nbr_ptr, const int* nbr_idx, long long cap_total, int* start,
                          int* len, double* verts) {
  std::vector<Poly> cells(n);
  double wmax = -INFINITY;
  for (int i = 0; i < n; ++i) wmax = std::max(wmax, w[i]);
#pragma omp parallel
  {
    Poly tmp;
    std::vector<std::pair<double, int>> order;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
      double sx = xy[2 * i], sy = xy[2 * i + 1];
      Poly poly = {box[0] - sx, box[2] - sy, box[1] - sx, box[2] - sy,
                   box[1] - sx, box[3] - sy, box[0] - sx, box[3] - sy};
      auto apply = [&](int j) {
        double dx = xy[2 * j] - sx, dy = xy[2 * j + 1] - sy;
        double c = 0.5 * (dx * dx + dy * dy + w[i] - w[j]);
        clip_halfplane(poly, dx, dy, c, tmp);
      };
      if (nbr_ptr) {
        // A site with no neighbour in a triangulation of several sites is hidden: empty cell.
        if (n > 1 && nbr_ptr[i] == nbr_ptr[i + 1]) poly.clear();
        for (int k = nbr_ptr[i]; k < nbr_ptr[i + 1]; ++k) {
          int j = nbr_idx[k];
          if (j != i && !poly.empty()) apply(j);
        }
      } else {
        order.clear();
        for (int j = 0; j < n; ++j) {
          if (j == i) continue;
          double dx = xy[2 * j] - sx, dy = xy[2 * j + 1] - sy;
          order.push_back({dx * dx + dy * dy, j});
        }
        std::sort(order.begin(), order.end());
        for (auto& dj : order) {
          if (poly.empty()) break;
          double R2 = 0.0;
          for (size_t k = 0; k < poly.size(); k += 2)
            R2 = std::max(R2, poly[k] * poly[k] + poly[k + 1] * poly[k + 1]);
          double R = std::sqrt(R2), D = std::sqrt(dj.first);
          // No site at distance >= D can own a point within R of s_i once
          // (D - R)^2 - wmax >= R^2 - w_i.
          if (D >= R && (D - R) * (D - R) - wmax >= R2 - w[i]) break;
          apply(dj.second);
        }
      }
      for (size_t k = 0; k < poly.size(); k += 2) {
        poly[k] += sx;
        poly[k + 1] += sy;
      }
      cells[i].swap(poly);
    }
  }
  long long total = 0;
  for (int i = 0; i < n; ++i) {
    start[i] = (int)total;
    len[i] = (int)cells[i].size() / 2;
    total += len[i];
  }
  if (total > cap_total) return -1;
  for (int i = 0; i < n; ++i)
    std::memcpy(verts + 2 * (size_t)start[i], cells[i].data(), cells[i].size() * sizeof(double));
  return total;
}

// Per-cell integrals over the piecewise-constant target rho (reference layout data[iy*m+ix],
// proj = {x0,x1,y0,y1}).  Any output pointer may be NULL.
//   mass[i]  = sum_p rho_p |C_i n p| / A_pix                 (compute_inte)
//   f[i]     = -w_i nu_i - sum_p rho_p/A_pix (int_{C_i n p}|x-s_i|^2 - w_i |C_i n p|)  (compute_f)
//   cen[2i]  = rho-weighted centroid, pieces with area > 1e-10 (compute_center, weighted)
//   exist[i] = the domain-clipped cell has >= 3 vertices
//   area[i]  = area of the domain-clipped cell
//   geo[2i]  = geometric centroid of the domain-clipped cell (compute_center, unweighted)
void orc_cell_integrals(int n, const double* xy, const double* w, const double* nu,
                        const int* start, const int* len, const double* verts, const double* rho,
                        int pm, int pn, const double* proj, double* mass, double* f, double* cen,
                        unsigned char* exist, double* area, double* geo) {
  PixGrid g(proj, pm, pn);
  const double square[4] = {proj[0], proj[2], proj[1], proj[3]};
#pragma omp parallel
  {
    Poly inside, piece;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
      double wi = w ? w[i] : 0.0, nui = nu ? nu[i] : 0.0;
      double m_acc = 0.0, f_acc = -wi * nui, tot = 0.0, cx = 0.0, cy = 0.0, a_in = 0.0;
      double gx = 0.0, gy = 0.0;
      bool ex = false;
      if (len[i] != 0) {
        clip_rect(verts + 2 * (size_t)start[i], len[i], square, inside);
        int ni = (int)inside.size() / 2;
        if (ni >= 3) {
          ex = true;
          a_in = fan_area(inside.data(), ni);
          fan_centroid(inside.data(), ni, gx, gy);
          int lx, ly, hx, hy;
          g.bbox(inside, lx, ly, hx, hy);
          double r[4];
          for (int ix = lx; ix < hx; ++ix) {
            for (int iy = ly; iy < hy; ++iy) {
              g.pixel(ix, iy, r);
              clip_rect(inside.data(), ni, r, piece);
              int np = (int)piece.size() / 2;
              double a = np >= 3 ? fan_area(piece.data(), np) : 0.0;
              double rv = rho[ix + (size_t)iy * pm];
              m_acc += rv * a / g.area;
              if (a > 0.0) {
                double sq = second_moment(piece.data(), np, xy[2 * i], xy[2 * i + 1]);
                f_acc -= rv / g.area * (sq - wi * a);
              }
              if (a > g_centroid_min_piece) {
                double lcx, lcy;
                fan_centroid(piece.data(), np, lcx, lcy);
                double contrib = rv * a;
                tot += contrib;
                cx += lcx * contrib;
                cy += lcy * contrib;
              }
            }
          }
          if (tot != 0.0) {
            cx /= tot;
            cy /= tot;
          } else {
            cx = gx;
            cy = gy;
          }
        }
      }
      if (mass) mass[i] = m_acc;
      if (f) f[i] = f_acc;
      if (cen) {
        cen[2 * i] = ex ? cx : 0.0;
        cen[2 * i + 1] = ex ? cy : 0.0;
      }
      if (exist) exist[i] = ex ? 1 : 0;
      if (area) area[i] = a_in;
      if (geo) {
        geo[2 * i] = ex ? gx : 0.0;
        geo[2 * i + 1] = ex ? gy : 0.0;
      }
    }
  }
}

// Area of each cell clipped to dom = {x0,x1,y0,y1} (source masses before normalisation).
void orc_clip_area(int n, const int* start, const int* len, const double* verts,
                   const double* dom, double* area) {
  const double square[4] = {dom[0], dom[2], dom[1], dom[3]};
#pragma omp parallel
  {
    Poly inside;
#pragma omp for
    for (int i = 0; i < n; ++i) {
      if (len[i] > 1) {
        clip_rect(verts + 2 * (size_t)start[i], len[i], square, inside);
        int ni = (int)inside.size() / 2;
        area[i] = ni >= 3 ? fan_area(inside.data(), ni) : 0.0;
      } else {
        area[i] = 0.0;
      }
    }
  }
}

// Diagnostic image: pixel p += nu_i |C_i n p| / |C_i n D|, summed over cells (serial, so the
// summation order is fixed).
void orc_render(int n, const double* nu, const int* start, const int* len, const double* verts,
                int pm, int pn, const double* proj, double* img) {
  PixGrid g(proj, pm, pn);
  const double square[4] = {proj[0], proj[2], proj[1], proj[3]};
  Poly inside, piece;
  for (int i = 0; i < n; ++i) {
    if (len[i] == 0) continue;
    clip_rect(verts + 2 * (size_t)start[i], len[i], square, inside);
    int ni = (int)inside.size() / 2;
    if (ni < 3) continue;
    double big = fan_area(inside.data(), ni);
    int lx, ly, hx, hy;
    g.bbox(inside, lx, ly, hx, hy);
    double r[4];
    for (int ix = lx; ix < hx; ++ix)
      for (int iy = ly; iy < hy; ++iy) {
        g.pixel(ix, iy, r);
        clip_rect(inside.data(), ni, r, piece);
        int np = (int)piece.size() / 2;
        double a = np >= 3 ? fan_area(piece.data(), np) : 0.0;
        if (big != 0.0) img[ix + (size_t)iy * pm] += nu[i] * a / big;
      }
  }
}

// Unpolarised Fresnel transmittance product T1*T2 per site (parallel or point light).
//   pts: 2n source points, in_dir: 3n incident directions, ins: 3n unit directions inside the
//   lens, h: n heights, tgt: 2n receiver points (cell centroids), z_len, z_recv, n_idx.
static double fresnel_T(const double* wi, const double* wt, const double* nrm, double ei,
                        double et) {
  double ci = std::fabs(wi[0] * nrm[0] + wi[1] * nrm[1] + wi[2] * nrm[2]);
  double ct = std::fabs(wt[0] * nrm[0] + wt[1] * nrm[1] + wt[2] * nrm[2]);
  ci = std::min(1.0, std::max(0.0, ci));
  ct = std::min(1.0, std::max(0.0, ct));
  double ds = ei * ci + et * ct, dp = ei * ct + et * ci;
  if (std::fabs(ds) < 1e-12) ds = 1e-12;
  if (std::fabs(dp) < 1e-12) dp = 1e-12;
  double rs = (ei * ci - et * ct) / ds, rp = (ei * ct - et * ci) / dp;
  return 1.0 - 0.5 * (rs * rs + rp * rp);
}

static void unit3(double* v) {
  double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (nv < 1e-12) {
    v[0] = v[1] = v[2] = 0.0;
  } else {
    v[0] /= nv;
    v[1] /= nv;
    v[2] /= nv;
  }
}

void orc_fresnel_rate(int n, const double* pts, const double* in_dir, const double* ins,
                      const double* h, const double* tgt, double z_len, double z_recv,
                      double n_idx, double* out) {
  for (int i = 0; i < n; ++i) {
    const double* wi = in_dir + 3 * i;
    const double* d = ins + 3 * i;
    double outp[3] = {pts[2 * i] + h[i] * d[0], pts[2 * i + 1] + h[i] * d[1], z_len + h[i] * d[2]};
    double wt[3] = {tgt[2 * i] - outp[0], tgt[2 * i + 1] - outp[1], z_recv - outp[2]};
    unit3(wt);
    double n2[3] = {n_idx * d[0] - wt[0], n_idx * d[1] - wt[1], n_idx * d[2] - wt[2]};
    unit3(n2);
    const double n1[3] = {0.0, 0.0, 1.0};
    out[i] = fresnel_T(wi, d, n1, 1.0, n_idx) * fresnel_T(d, wt, n2, n_idx, 1.0);
  }
}

// Integral of |x - s|^2 over a polygon by the reference's per-edge Green polynomial
// (Edge_ints / Polygon_Ints, basic_function.cuh:826-865), exported so tests can check it
// against the triangle-fan form used above.
double orc_edge_green_moment(const double* p, int n, double sx, double sy) {
  auto edge = [&](const double* a, const double* b) {
    double x2mx1 = b[0] - a[0], y1my2 = a[1] - b[1];
    double x1px2 = a[0] + b[0], y1py2 = a[1] + b[1];
    double xs2 = a[0] * a[0] + b[0] * b[0], ys2 = a[1] * a[1] + b[1] * b[1];
    double xs2x = xs2 + a[0] * b[0], ys2y = ys2 + a[1] * b[1];
    double cr = a[0] * b[1] - b[0] * a[1];
    double q = x2mx1 * y1py2 * ys2 / 4.0 + (cr - 2.0 * sy * x2mx1) * ys2y / 3.0 +
               (sy * sy * x2mx1 / 2.0 - sy * cr) * y1py2 + sy * sy * cr;
    double pp = y1my2 * x1px2 * xs2 / 4.0 + (cr - 2.0 * sx * y1my2) * xs2x / 3.0 +
                (sx * sx * y1my2 / 2.0 - sx * cr) * x1px2 + sx * sx * cr;
    return q + pp;
  };
  double acc = edge(p + 2 * (n - 1), p);
  for (int i = 0; i < n - 1; ++i) acc += edge(p + 2 * i, p + 2 * i + 2);
  return acc;
}

double orc_fan_moment(const double* p, int n, double sx, double sy) {
  return second_moment(p, n, sx, sy);
}

}  // extern "C"
