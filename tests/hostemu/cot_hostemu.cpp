// Host build of the cell and conformity-check kernels (test infrastructure).
//
// Runs the exact per-site code of cot_kernels.hip (cell_site, verify_site) on the CPU, one lane
// at a time, with the library's layout (cot_layout.h) and repair rounds, so that the GPU code
// path can be checked against the oracle without a GPU.  Built by tests/hostemu/build.sh.
#include <cstring>
#include <vector>

#include "../../caustic_amd/csrc/cot_kernels.hip"
#include "../../caustic_amd/csrc/cot_layout.h"

using namespace cot;

namespace {

template <typename T>
struct Emu {
  PolyLDS<T>* P = new PolyLDS<T>();
  ~Emu() { delete P; }
};

}  // namespace

extern "C" {

// Evaluate one problem at w.  Outputs in caller order; any may be NULL.
// stats[0] = conformity failures left, stats[1] = repair rounds, stats[2] = overflow cells,
// stats[3] = largest ring used.
int emu_evaluate(int n, const double* sites, const double* nu, const double* rho, int pm, int pn,
                 const double* proj, const double* w, int fp32, int ring_init, int max_repairs,
                 double verify_tol, double* mass, double* fcell, double* cen, double* geo,
                 unsigned char* exist, int* stats, int* ring_io) {
  cot_problem pr{};
  pr.n_sites = n;
  pr.sites = sites;
  pr.nu = nu;
  pr.rho = rho;
  pr.pm = pm;
  pr.pn = pn;
  for (int k = 0; k < 4; ++k) pr.proj_domain[k] = proj[k];
  ProbDesc pd{};
  pd.site_off = 0;
  pd.n = n;
  pd.pm = pm;
  pd.pn = pn;
  pd.img_off = 0;
  pd.x0 = proj[0];
  pd.x1 = proj[1];
  pd.y0 = proj[2];
  pd.y1 = proj[3];
  pd.dx = (pd.x1 - pd.x0) / pm;
  pd.dy = (pd.y1 - pd.y0) / pn;
  std::vector<int> order;
  std::vector<int2> bk;
  bucket_sort(pr, 1.0, pd, order, bk);
  pd.bkt_off = 0;
  std::vector<BlockDesc> blocks;
  for (int f = 0; f < n; f += kCellBlock) {
    BlockDesc b{};
    b.prob = 0;
    b.first = f;
    b.count = std::min(kCellBlock, n - f);
    blocks.push_back(b);
  }
  pd.blk_off = 0;
  pd.nblk = (int)blocks.size();
  const int nb = pd.nblk;
  std::vector<double2> xy(n), org(n);
  std::vector<double> vnu(n), vw(n), vmass(n), vf(n), va(n);
  std::vector<double2> vcen(n), vgeo(n);
  std::vector<uint8_t> vex(n), ne(n), ring(n, (uint8_t)ring_init), used(n), redo(n, 0), box(n, 0);
  std::vector<int> elab((size_t)n * kMaxVerts);
  std::vector<float2> espan((size_t)n * kMaxVerts);
  for (int r = 0; r < n; ++r) {
    const int o = order[r];
    xy[r] = make_double2(sites[2 * o], sites[2 * o + 1]);
    org[r] = xy[r];
    vnu[r] = nu[o];
    vw[r] = w[o];
    // ring_io: the per-site candidate rings carried between evaluations (0: start at ring_init)
    if (ring_io && ring_io[o] > 0) ring[r] = (uint8_t)ring_io[o];
  }
  std::vector<float> rho32;
  if (fp32) rho32.assign(rho, rho + (size_t)pm * pn);
  uint8_t active = 1;
  int overflow = 0;
  CellArgs a{};
  a.probs = &pd;
  a.blocks = blocks.data();
  a.xy = xy.data();
  a.w = vw.data();
  a.dir = nullptr;
  a.nu = vnu.data();
  a.buckets = bk.data();
  a.img = fp32 ? (const void*)rho32.data() : (const void*)rho;
  a.ring = ring.data();
  a.redo = redo.data();
  a.redo_mode = 0;
  a.active = &active;
  a.mass = vmass.data();
  a.fcell = vf.data();
  a.area = va.data();
  a.cen = vcen.data();
  a.geo = vgeo.data();
  a.exist = vex.data();
  a.nedge = ne.data();
  a.elab = elab.data();
  a.espan = espan.data();
  a.ring_used = used.data();
  a.overflow = &overflow;
  a.org = org.data();
  VerifyArgs v{};
  v.probs = &pd;
  v.blocks = blocks.data();
  v.xy = xy.data();
  v.nedge = ne.data();
  v.elab = elab.data();
  v.espan = espan.data();
  v.active = &active;
  v.gate = nullptr;
  v.redo = redo.data();
  v.tol_rel = verify_tol;

  auto run_cells = [&]() {
#pragma omp parallel
    {
      Emu<float> ef;
      Emu<double> ed;
#pragma omp for schedule(dynamic, 1) reduction(+ : overflow)
      for (int b = 0; b < nb; ++b)
        for (int ln = 0; ln < kCellBlock; ++ln) {
          bool o = fp32 ? cell_site<float, float>(a, b, ln, *ef.P)
                        : cell_site<double, double>(a, b, ln, *ed.P);
          overflow += o ? 1 : 0;
        }
    }
  };
  auto run_verify = [&]() {
    int fails = 0;
    for (int b = 0; b < nb; ++b)
      for (int ln = 0; ln < kCellBlock; ++ln) fails += verify_site(v, b, ln) ? 1 : 0;
    return fails;
  };
  run_cells();
  int fails = run_verify();
  int rounds = 0;
  while (fails > 0 && rounds < max_repairs) {
    ++rounds;
    for (int r = 0; r < n; ++r)
      if (redo[r]) {
        ring[r] = (uint8_t)std::min(2 * (int)ring[r] + 1, 250);
      }
    a.redo_mode = 1;
    run_cells();
    a.redo_mode = 0;
    std::fill(redo.begin(), redo.end(), 0);
    fails = run_verify();
  }
  int rmax = 0;
  for (int r = 0; r < n; ++r) rmax = std::max(rmax, (int)ring[r]);
  // as the library: after a clean check the next ring is the outermost used ring plus one
  if (fails == 0)
    for (int r = 0; r < n; ++r) ring[r] = (uint8_t)std::max((int)used[r] + 1, 1);
  if (stats) {
    stats[0] = fails;
    stats[1] = rounds;
    stats[2] = overflow;
    stats[3] = rmax;
  }
  for (int r = 0; r < n; ++r) {
    const int o = order[r];
    if (mass) mass[o] = vmass[r];
    if (fcell) fcell[o] = vf[r];
    if (cen) {
      cen[2 * o] = vcen[r].x;
      cen[2 * o + 1] = vcen[r].y;
    }
    if (geo) {
      geo[2 * o] = vgeo[r].x;
      geo[2 * o + 1] = vgeo[r].y;
    }
    if (exist) exist[o] = vex[r];
    if (ring_io) ring_io[o] = (int)ring[r];
  }
  return 0;
}

}  // extern "C"
