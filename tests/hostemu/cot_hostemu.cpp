// Host build of the cell and conformity-check kernels (test infrastructure).
//
// Runs the exact per-site code of cot_kernels.hip (cell_site, verify_site) on the CPU, one lane
// at a time, with the library's layout (cot_layout.h) and its evaluation sequence: a full pass
// from the previous evaluation's two-ring adjacency, the conformity check, and repair rounds
// (the current cells' two-ring and the check's queued candidates, then bucket rings).  A session keeps rings, adjacency and FP32
// frames between evaluations like a library context.  Built by tests/hostemu/build.sh.
#include <cstring>
#include <type_traits>
#include <vector>

#include "../../caustic_amd/csrc/cot_kernels.hip"
#include "../../caustic_amd/csrc/cot_layout.h"

using namespace cot;

namespace {

struct Session {
  int n = 0, pm = 0, pn = 0, fp32 = 0, max_repairs = 12;
  double verify_tol = 1e-4;
  ProbDesc pd{};
  std::vector<int> order;
  std::vector<int2> bk;
  std::vector<BlockDesc> blocks;
  std::vector<double2> xy, org;
  std::vector<double> nu, w, mass, fcell, area;
  std::vector<DensityGroup> grp;
  std::vector<double4> tot;
  std::vector<double2> cen, geo;
  std::vector<uint8_t> exist, ne, ring, used, adjn;
  std::vector<int> redo;
  std::vector<int> elab, adj, xcnt, xcand;
  std::vector<double2> vtx;
  std::vector<int2> kbox;
  std::vector<float2> espan;
  bool have_adj = false;
  std::vector<int> elist;  // cells the last evaluation left empty (site slots)
  int ecount = 0;
  std::vector<uint8_t> prebuilt;
  int ring_max = 32;  // the library's default
};

template <typename T, int V>
struct Lanes {
  PolyLDS<T, V>* P = new PolyLDS<T, V>();
  ~Lanes() { delete P; }
};

}  // namespace

extern "C" {

int emu_max_verts() { return kMaxVerts; }

int emu_xcd_slot(int w, int G) { return xcd_slot(w, G); }

// Pixels (of all rows) where the density groups' lookup (group_tap) differs in any bit from the
// sequential chunk-local prefix sum, or where the chunk totals differ from the sum over the chunk.
int emu_prefix_mismatches(const double* rho, int pm, int pn, int fp32) {
  const int nchunk = (pm + kChunk - 1) / kChunk;
  const int gsize = fp32 ? kGroup32 : kGroup64;
  const int ngrp = (pm + gsize - 1) / gsize;
  std::vector<DensityGroup> grp((size_t)ngrp * pn);
  std::vector<double4> tot((size_t)nchunk * pn);
  build_prefix(rho, pm, pn, fp32 != 0, grp.data(), tot.data());
  int bad = 0;
  for (int iy = 0; iy < pn; ++iy) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int g = 0; g < pm; ++g) {
      if (g % kChunk == 0) s0 = s1 = s2 = 0.0;
      double r = rho[(size_t)iy * pm + g];
      if (fp32) r = (double)(float)r;
      double t0, t1, t2, tr;
      group_tap(grp[(size_t)iy * ngrp + g / gsize], g, gsize, t0, t1, t2, tr);
      if (t0 != s0 || t1 != s1 || t2 != s2 || tr != r) ++bad;
      prefix_step(s0, s1, s2, r, (double)(g % kChunk));
      if (g % kChunk == kChunk - 1 || g == pm - 1) {
        const double4 c = tot[(size_t)iy * nchunk + g / kChunk];
        if (c.x != s0 || c.y != s1 || c.z != s2) ++bad;
      }
    }
  }
  return bad;
}

void* emu_create(int n, const double* sites, const double* nu, const double* rho, int pm, int pn,
                 const double* proj, int fp32, int ring_init, int max_repairs, double verify_tol) {
  Session* S = new Session();
  S->n = n;
  S->pm = pm;
  S->pn = pn;
  S->fp32 = fp32;
  S->max_repairs = max_repairs;
  S->verify_tol = verify_tol;
  cot_problem pr{};
  pr.n_sites = n;
  pr.sites = sites;
  pr.nu = nu;
  pr.rho = rho;
  pr.pm = pm;
  pr.pn = pn;
  for (int k = 0; k < 4; ++k) pr.proj_domain[k] = proj[k];
  ProbDesc& pd = S->pd;
  pd.n = n;
  pd.pm = pm;
  pd.pn = pn;
  pd.x0 = proj[0];
  pd.x1 = proj[1];
  pd.y0 = proj[2];
  pd.y1 = proj[3];
  pd.dx = (pd.x1 - pd.x0) / pm;
  pd.dy = (pd.y1 - pd.y0) / pn;
  bucket_sort(pr, 1.0, pd, S->order, S->bk);
  for (int f = 0; f < n; f += kCellBlock) {
    BlockDesc b{};
    b.prob = 0;
    b.first = f;
    b.count = std::min(kCellBlock, n - f);
    S->blocks.push_back(b);
  }
  pd.nblk = (int)S->blocks.size();
  S->xy.resize(n);
  S->nu.resize(n);
  for (int r = 0; r < n; ++r) {
    const int o = S->order[r];
    S->xy[r] = make_double2(sites[2 * o], sites[2 * o + 1]);
    S->nu[r] = nu[o];
  }
  S->org = S->xy;
  S->w.assign(n, 0.0);
  S->mass.assign(n, 0.0);
  S->fcell.assign(n, 0.0);
  S->area.assign(n, 0.0);
  S->cen.assign(n, make_double2(0, 0));
  S->geo.assign(n, make_double2(0, 0));
  S->exist.assign(n, 0);
  S->ne.assign(n, 0);
  S->ring.assign(n, (uint8_t)ring_init);
  S->used.assign(n, 0);
  S->redo.assign(n, 0);
  S->adjn.assign(n, 0);
  S->elist.assign(n, 0);
  S->prebuilt.assign(n, 0);
  S->elab.assign((size_t)n * kMaxVerts, 0);
  S->adj.assign((size_t)n * kMaxVerts, 0);
  S->xcnt.assign(n, 0);
  S->vtx.assign((size_t)n * kMaxVerts, make_double2(0, 0));
  S->kbox.assign(n, make_int2(0, 0));
  S->xcand.assign((size_t)n * kMaxExtra, 0);
  S->espan.assign((size_t)n * kMaxVerts, make_float2(0.f, 0.f));
  pd.nchunk = (pm + kChunk - 1) / kChunk;
  pd.gsize = fp32 ? kGroup32 : kGroup64;
  pd.ngrp = (pm + pd.gsize - 1) / pd.gsize;
  pd.grp_off = 0;
  S->grp.resize((size_t)pd.ngrp * pn);
  S->tot.resize((size_t)pd.nchunk * pn);
  build_prefix(rho, pm, pn, fp32 != 0, S->grp.data(), S->tot.data());
  return S;
}

void emu_destroy(void* h) { delete static_cast<Session*>(h); }

// Candidates of the next evaluation as if a previous evaluation had left them: counts[i]
// neighbour labels of site i at labels[i * kMaxVerts + k] (caller indices).
void emu_set_adjacency(void* h, const int* counts, const int* labels) {
  Session& S = *static_cast<Session*>(h);
  std::vector<int> sorted_of(S.n);
  for (int r = 0; r < S.n; ++r) sorted_of[S.order[r]] = r;
  for (int r = 0; r < S.n; ++r) {
    const int o = S.order[r];
    S.adjn[r] = (uint8_t)std::min(counts[o], kMaxVerts);
    for (int k = 0; k < S.adjn[r]; ++k) {
      const int j = labels[(size_t)o * kMaxVerts + k];
      S.adj[(size_t)r * kMaxVerts + k] = j >= 0 ? sorted_of[j] : j;
    }
  }
  S.have_adj = true;
}

// Per-site state after the last evaluation (caller order; any may be NULL): the conformity
// flags of the last check, edge counts, edge labels (caller indices, sides below zero) and spans
// (kMaxVerts per site), and the bucket rings.
void emu_state(void* h, unsigned char* redo, int* nedge, int* elab, float* espan, int* ring) {
  Session& S = *static_cast<Session*>(h);
  for (int r = 0; r < S.n; ++r) {
    const int o = S.order[r];
    if (redo) redo[o] = (unsigned char)(S.redo[r] != 0);
    if (nedge) nedge[o] = S.ne[r];
    if (ring) ring[o] = S.ring[r];
    for (int k = 0; k < kMaxVerts; ++k) {
      const int lab = S.elab[(size_t)r * kMaxVerts + k];
      if (elab) elab[(size_t)o * kMaxVerts + k] = lab >= 0 ? S.order[lab] : lab;
      if (espan) {
        espan[2 * ((size_t)o * kMaxVerts + k)] = S.espan[(size_t)r * kMaxVerts + k].x;
        espan[2 * ((size_t)o * kMaxVerts + k) + 1] = S.espan[(size_t)r * kMaxVerts + k].y;
      }
    }
  }
}

// Evaluate the session's problem at w (caller order).  Outputs in caller order; any may be NULL.
// stats[0] = conformity failures left (plus one when the cells do not tile the target),
// stats[1] = repair rounds, stats[2] = overflow cells, stats[3] = largest bucket ring,
// stats[4] = cells rebuilt by repairs, stats[5] = 1 when the cell areas add up to the target's.
int emu_eval(void* h, const double* w, double* mass, double* fcell, double* cen, double* geo,
             unsigned char* exist, int* stats) {
  Session& S = *static_cast<Session*>(h);
  const int n = S.n, nb = S.pd.nblk;
  for (int r = 0; r < n; ++r) S.w[r] = w[S.order[r]];
  uint8_t active = 1;
  int overflow = 0;
  CellArgs a{};
  a.probs = &S.pd;
  a.blocks = S.blocks.data();
  a.xy = S.xy.data();
  a.w = S.w.data();
  a.nu = S.nu.data();
  a.buckets = S.bk.data();
  a.grp = S.grp.data();
  a.tot = S.tot.data();
  a.ring = S.ring.data();
  a.redo = S.redo.data();
  a.active = &active;
  a.mass = S.mass.data();
  a.fcell = S.fcell.data();
  a.area = S.area.data();
  a.cen = S.cen.data();
  a.geo = S.geo.data();
  a.exist = S.exist.data();
  a.nedge = S.ne.data();
  a.elab = S.elab.data();
  a.espan = S.espan.data();
  a.ring_used = S.used.data();
  a.overflow = &overflow;
  a.org = S.org.data();
  a.ring_max = S.ring_max;
  a.adj = S.adj.data();
  a.adjn = S.adjn.data();
  a.cand = S.have_adj ? (S.fp32 ? kCandAdj1 : kCandAdj2) : kCandBucket;  // as the library
  a.xcnt = S.xcnt.data();
  a.xcand = S.xcand.data();
  a.prebuilt = S.prebuilt.data();
  a.vtx = S.vtx.data();
  a.kbox = S.kbox.data();
  VerifyArgs v{};
  v.probs = &S.pd;
  v.blocks = S.blocks.data();
  v.xy = S.xy.data();
  v.nedge = S.ne.data();
  v.elab = S.elab.data();
  v.espan = S.espan.data();
  v.active = &active;
  v.redo = S.redo.data();
  v.tol_rel = S.verify_tol;
  v.xcnt = S.xcnt.data();
  v.xcand = S.xcand.data();

  int rebuilt = 0, coarse = 0;
  a.coarse = &coarse;
  // the full pass in the context's precision, repairs in FP64 (as the library)
  std::vector<uint8_t> rebuilt_site(n, 0);
  auto run_cells = [&]() {
    int built = 0;
    const bool f32 = S.fp32 && !a.redo_mode;
    std::fill(rebuilt_site.begin(), rebuilt_site.end(), 0);
#pragma omp parallel
    {
      Lanes<float, kPassVerts> lf;
      Lanes<double, kPassVerts> ld;
      Lanes<double, kMaxVerts> lr;
#pragma omp for schedule(dynamic, 1) reduction(+ : overflow, built)
      for (int b = 0; b < nb; ++b)
        for (int ln = 0; ln < kCellBlock; ++ln) {
          // full passes clip in kPassVerts, repairs in kMaxVerts (as the kernels)
          const int rc = a.redo_mode ? cell_site<double, kMaxVerts>(a, b, ln, *lr.P, ln)
                         : f32       ? cell_site<float, kPassVerts>(a, b, ln, *lf.P, ln)
                                     : cell_site<double, kPassVerts>(a, b, ln, *ld.P, ln);
          overflow += rc == 1 ? 1 : 0;
          built += rc >= 0 ? 1 : 0;
          if (rc >= 0) rebuilt_site[S.blocks[b].first + ln] = 1;
        }
    }
    // integration pass: L lanes per site (kIntLanes in a full pass, kRepairLanes in a repair,
    // as the library), summed in the order of the GPU's butterfly
    auto integrate_site = [&](auto lanes, int i) {
      constexpr int L = decltype(lanes)::value;
      Moments M[L];
      for (int q = 0; q < L; ++q)
        f32 ? integrate_lane<float, L>(a, S.pd, i, q, M[q])
            : integrate_lane<double, L>(a, S.pd, i, q, M[q]);
      for (int off = L / 2; off > 0; off >>= 1) {
        Moments nxt[L];
        for (int q = 0; q < L; ++q) {
          nxt[q] = M[q];
          combine(nxt[q], M[q ^ off]);
        }
        for (int q = 0; q < L; ++q) M[q] = nxt[q];
      }
      finish_site(a, S.pd, 0, i, M[0]);
    };
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
      for (int quarter = 0; quarter < kCellBlock * kIntLanes / kIntBlock; ++quarter)
        for (int t0 = 0; t0 < kIntBlock; t0 += kIntLanes) {
          const int i = integ_site(a, b, quarter, t0);
          // a repair pass integrates the sites it rebuilt (the library: its work list)
          if (i < 0 || (a.redo_mode && !rebuilt_site[i])) continue;
          if (a.redo_mode)
            integrate_site(std::integral_constant<int, kRepairLanes>{}, i);
          else
            integrate_site(std::integral_constant<int, kIntLanes>{}, i);
        }
    return built;
  };
  auto run_verify = [&]() {
    int fails = 0;
    for (int b = 0; b < nb; ++b)
      for (int ln = 0; ln < kCellBlock; ++ln) fails += verify_site(v, b, ln) == 1 ? 1 : 0;
    return fails;
  };
  // as the library: the evaluation's last save also lists the empty cells for the next one
  auto save_adjacency = [&](bool last) {
    if (last) S.ecount = 0;
    for (int b = 0; b < nb; ++b)
      for (int ln = 0; ln < kCellBlock; ++ln)
        save_adjacency_site(S.blocks[b], b, ln, S.ne.data(), S.elab.data(), S.adjn.data(),
                            S.adj.data(), last ? S.elist.data() : nullptr, &S.ecount, n);
  };
  if (S.have_adj && use_prepass(nb)) {
    // as the library (cell_reappear_kernel): the cells the last evaluation left empty, built
    // before the full pass (which skips them) without their queued candidates; those that come
    // back are queued as candidates of their new neighbours
    CellArgs b = a;
    b.xcnt = nullptr;
    b.prepass = 1;
    Lanes<float, kPassVerts> lf;
    Lanes<double, kPassVerts> ld;
    const int ne_list = std::min(S.ecount, n);
    for (int e = 0; e < ne_list; ++e) {
      const int slot = S.elist[e], blk = slot / kCellBlock, ln = slot % kCellBlock;
      const int rc = S.fp32 ? cell_site<float, kPassVerts>(b, blk, ln, *lf.P, ln)
                            : cell_site<double, kPassVerts>(b, blk, ln, *ld.P, ln);
      if (rc < 0) continue;
      overflow += rc == 1 ? 1 : 0;
      const int i = S.blocks[blk].first + ln;
      S.prebuilt[i] = 1;
      for (int k = 0; k < S.ne[i]; ++k) {
        const int j = S.elab[(size_t)i * kMaxVerts + k];
        if (j < 0) continue;
        const int q = S.xcnt[j]++;
        if (q < kMaxExtra) S.xcand[(size_t)j * kMaxExtra + q] = i;
      }
    }
  }
  run_cells();
  int fails = run_verify() + coarse;
  int rounds = 0;
  a.redo_mode = 1;
  while (fails > 0 && rounds < S.max_repairs) {
    ++rounds;
    // as the library: two rounds from the current cells' two-ring and the queued candidates,
    // then bucket rings grown 2r + 1
    a.cand = rounds <= 2 ? kCandAdj2 : kCandBucket;
    if (a.cand == kCandBucket) {
      for (int r = 0; r < n; ++r)
        if (S.redo[r]) S.ring[r] = (uint8_t)std::min(2 * (int)S.ring[r] + 1, S.ring_max);
    } else {
      save_adjacency(false);
    }
    rebuilt += run_cells();
    std::fill(S.redo.begin(), S.redo.end(), 0);
    fails = run_verify();
  }
  int rmax = 0;
  for (int r = 0; r < n; ++r) rmax = std::max(rmax, (int)S.ring[r]);
  // as the library: the cells just built are the next candidates; after a clean check the next
  // bucket ring is the outermost used ring plus one
  save_adjacency(true);
  // as the library: cells that do not tile the target leave the evaluation unverified, and the
  // next evaluation searches bucket rings
  double area = 0.0;
  for (int r = 0; r < n; ++r) area += S.area[r];
  const bool covered = covers_domain(S.pd, area);
  const bool multiple = fails == 0 && !covered;  // edges matched, target covered twice or more
  S.have_adj = !multiple;
  if (fails == 0)
    for (int r = 0; r < n; ++r)
      S.ring[r] = (uint8_t)std::min(std::max((int)S.used[r] + 1, 1), S.ring_max);
  if (stats) {
    stats[0] = fails + (multiple ? 1 : 0);
    stats[1] = rounds;
    stats[2] = overflow;
    stats[3] = rmax;
    stats[4] = rebuilt;
    stats[5] = covered ? 1 : 0;
  }
  for (int r = 0; r < n; ++r) {
    const int o = S.order[r];
    if (mass) mass[o] = S.mass[r];
    if (fcell) fcell[o] = S.fcell[r];
    if (cen) {
      cen[2 * o] = S.cen[r].x;
      cen[2 * o + 1] = S.cen[r].y;
    }
    if (geo) {
      geo[2 * o] = S.geo[r].x;
      geo[2 * o + 1] = S.geo[r].y;
    }
    if (exist) exist[o] = S.exist[r];
  }
  return 0;
}

// The host inverse of the Newton preconditioner's coarse matrix (cot_layout.h), in place.
void emu_coarse_inverse(double* a, int n) { coarse_inverse(a, n); }

}  // extern "C"
