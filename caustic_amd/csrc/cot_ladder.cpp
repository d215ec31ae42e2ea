// Coarse-to-fine solve (cot_solve_ladder): a host driver over the library's own C ABI.
//
// Level l builds, for every problem of the batch, a jittered sides[l] x sides[l] grid of the
// problem's frame, the target box-downsampled (plus a density floor), and source masses binned
// from the fine sites; solves the batch in one context; and carries the solution to the next
// level through the weight field U = -w (the Brenier potential 2 psi minus |x|^2, smooth):
// U(node_k) = -w_k + 2 (s_k - c_k).(s_k - node_k) to first order in the jitter, c_k the cell's
// rho-weighted centroid (the gradient of U at s is -2 (s - c(s))), then bilinear at the next
// level's sites.  The full-resolution problem runs the floor stages and then the exact target
// in the caller's context.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <vector>

#include "../../include/caustic_ot.h"

namespace cot {
const cot_options* context_options(const cot_context* c);
int context_problems(const cot_context* c);
int set_error(int code, const std::string& msg);
}  // namespace cot

namespace {

using cot::set_error;

constexpr double kDefaultCoarseTol = 0.1;
constexpr double kDefaultHandoff = 2.0;

// splitmix64: the coarse grids' jitter (any fixed generator serves; the jitter only has to keep
// sites off a lattice, where cells would tie)
struct Rng {
  unsigned long long s;
  double uniform() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

struct Level {
  std::vector<double> sites, nu, rho;
  int pm = 0, pn = 0, side = 0;
  double dom[4] = {0, 0, 0, 0};
};

const double* frame_of(const cot_ladder_opts& lo, const cot_problem& pr, int p) {
  return lo.frames ? lo.frames + 4 * p : pr.proj_domain;
}

// The box-downsampling factor of a level: the largest f dividing both image sides with at least
// 32 pixels left along the shorter side and at most `want` fine pixels per coarse grid step.
int target_factor(const cot_problem& pr, int want) {
  const int shorter = std::min(pr.pm, pr.pn);
  int f = std::max(1, std::min(want, shorter / 32));
  while (f > 1 && (pr.pm % f || pr.pn % f)) --f;
  return f;
}

int level_want(const cot_problem& pr, int side) {
  const int fine_side = (int)std::lround(std::sqrt((double)pr.n_sites));
  return std::max(1, fine_side / std::max(side, 1));
}

// Sums of an image (pm x pn) over g x g blocks (g divides both sides).
void block_sums(const double* src, int pm, int pn, int g, std::vector<double>& out) {
  const int qm = pm / g, qn = pn / g;
  out.assign((size_t)qm * qn, 0.0);
  for (int iy = 0; iy < pn; ++iy) {
    const double* row = src + (size_t)iy * pm;
    double* o = out.data() + (size_t)(iy / g) * qm;
    for (int qx = 0; qx < qm; ++qx) {
      double v = 0.0;
      for (int k = 0; k < g; ++k) v += row[qx * g + k];
      o[qx] += v;
    }
  }
}

// A problem's target summed over f x f blocks for the factors of all levels, each computed from
// the largest factor already done that divides it (the full image for the finest), so the
// full-resolution image is read once for all levels.
struct Pyramid {
  std::vector<int> f;                    // ascending
  std::vector<std::vector<double>> img;  // img[k]: f[k] x f[k] block sums
  const std::vector<double>& at(int fk) const {
    for (size_t k = 0; k < f.size(); ++k)
      if (f[k] == fk) return img[k];
    return img[0];  // not reached: every level's factor is built
  }
};

void build_pyramid(const cot_problem& pr, std::vector<int> factors, Pyramid& P) {
  std::sort(factors.begin(), factors.end());
  factors.erase(std::unique(factors.begin(), factors.end()), factors.end());
  P.f.clear();
  P.img.clear();
  for (int fk : factors) {
    int from = -1;  // index of the largest built factor dividing fk
    for (int k = (int)P.f.size() - 1; k >= 0 && from < 0; --k)
      if (fk % P.f[k] == 0) from = k;
    std::vector<double> out;
    if (fk == 1) {
      out.assign(pr.rho, pr.rho + (size_t)pr.pm * pr.pn);
    } else if (from < 0) {
      block_sums(pr.rho, pr.pm, pr.pn, fk, out);
    } else {
      const int g = P.f[from];
      block_sums(P.img[from].data(), pr.pm / g, pr.pn / g, fk / g, out);
    }
    P.f.push_back(fk);
    P.img.push_back(std::move(out));
  }
}

// The target of a level: the block sums for its factor plus floor x their mean pixel, scaled to
// `total`.
void coarse_target(const cot_problem& pr, const Pyramid& P, int f, double floor_rel,
                   double total, Level& L) {
  L.pm = pr.pm / f;
  L.pn = pr.pn / f;
  L.rho = P.at(f);
  double s = 0.0;
  for (double v : L.rho) s += v;
  const double add = floor_rel * s / (double)L.rho.size();
  double s2 = 0.0;
  for (double& v : L.rho) s2 += (v += add);
  if (s2 > 0.0)
    for (double& v : L.rho) v *= total / s2;
}

// Level `side` of problem pr: jittered grid of the frame, site k = j * side + i near node (i, j)
// (the layout the field transfer assumes), masses of the fine sites binned by grid cell.
void coarse_level(const cot_problem& pr, const Pyramid& P, const double* fr, int side,
                  double floor_rel, unsigned long long seed, Level& L) {
  Rng rng{seed};
  const int n = side * side;
  L.side = side;
  L.sites.resize(2 * (size_t)n);
  const double fx = fr[1] - fr[0], fy = fr[3] - fr[2];
  for (int j = 0; j < side; ++j)
    for (int i = 0; i < side; ++i) {
      const double u = rng.uniform() * 0.5 - 0.25, v = rng.uniform() * 0.5 - 0.25;
      L.sites[2 * ((size_t)j * side + i)] = fr[0] + fx * (i + 0.5 + u) / side;
      L.sites[2 * ((size_t)j * side + i) + 1] = fr[2] + fy * (j + 0.5 + v) / side;
    }
  L.nu.assign(n, 0.0);
  double total = 0.0;
  for (int k = 0; k < pr.n_sites; ++k) {
    const int i = std::min(std::max((int)std::floor((pr.sites[2 * k] - fr[0]) / fx * side), 0),
                           side - 1);
    const int j = std::min(
        std::max((int)std::floor((pr.sites[2 * k + 1] - fr[2]) / fy * side), 0), side - 1);
    L.nu[(size_t)j * side + i] += pr.nu[k];
    total += pr.nu[k];
  }
  // a grid cell no fine site falls in keeps a sliver of mass (every site needs some)
  const double least = 1e-3 * total / n;
  double t2 = 0.0;
  for (double& v : L.nu) t2 += (v = std::max(v, least));
  for (double& v : L.nu) v *= total / t2;
  coarse_target(pr, P, target_factor(pr, level_want(pr, side)), floor_rel, total, L);
  for (int q = 0; q < 4; ++q) L.dom[q] = pr.proj_domain[q];
}

cot_problem as_problem(const Level& L, int n) {
  cot_problem p{};
  p.n_sites = n;
  p.sites = L.sites.data();
  p.nu = L.nu.data();
  p.rho = L.rho.data();
  p.pm = L.pm;
  p.pn = L.pn;
  for (int q = 0; q < 4; ++q) p.proj_domain[q] = L.dom[q];
  return p;
}

// Field U on the side x side nodes of the frame from a level's solution (mean removed).
void field_from(const Level& L, const double* fr, const double* w, const double* cen,
                std::vector<double>& U) {
  const int side = L.side, n = side * side;
  U.resize(n);
  const double fx = fr[1] - fr[0], fy = fr[3] - fr[2];
  double mean = 0.0;
  for (int k = 0; k < n; ++k) {
    const double nx = fr[0] + fx * (k % side + 0.5) / side, ny = fr[2] + fy * (k / side + 0.5) / side;
    const double sx = L.sites[2 * k], sy = L.sites[2 * k + 1];
    U[k] = -w[k] + 2.0 * ((sx - cen[2 * k]) * (sx - nx) + (sy - cen[2 * k + 1]) * (sy - ny));
    mean += U[k];
  }
  mean /= n;
  for (double& u : U) u -= mean;
}

// w_i = -U(s_i), bilinear in the field on the side x side nodes (clamped at the border), mean
// removed.
void weights_from(const std::vector<double>& U, int side, const double* fr, const double* xy,
                  int n, double* w) {
  const double fx = fr[1] - fr[0], fy = fr[3] - fr[2];
  double mean = 0.0;
  for (int k = 0; k < n; ++k) {
    const double u = std::min(std::max((xy[2 * k] - fr[0]) / fx * side - 0.5, 0.0), side - 1.0);
    const double v =
        std::min(std::max((xy[2 * k + 1] - fr[2]) / fy * side - 0.5, 0.0), side - 1.0);
    const int i0 = std::min((int)std::floor(u), side - 2), j0 = std::min((int)std::floor(v), side - 2);
    const double a = u - i0, b = v - j0;
    auto at = [&](int j, int i) { return U[(size_t)j * side + i]; };
    w[k] = -((1 - a) * (1 - b) * at(j0, i0) + a * (1 - b) * at(j0, i0 + 1) +
             (1 - a) * b * at(j0 + 1, i0) + a * b * at(j0 + 1, i0 + 1));
    mean += w[k];
  }
  mean /= std::max(n, 1);
  for (int k = 0; k < n; ++k) w[k] -= mean;
}

// One stage: L-BFGS until max-rel < handoff with no empty cell, Newton steps (at most
// kNewtonSteps) to the stopping rule of `so`.  When a problem's Newton steps stall (a direction
// no damping makes acceptable, far from the solution), the batch goes back to L-BFGS until a
// hand-off kHandoffCut times tighter -- and below the stalled problems' max-rel over kHandoffCut,
// so each of them moves before Newton restarts from a new point -- up to kNewtonRetries times;
// then L-BFGS to the stopping rule.  Problems already converged stop at once in every later
// call.  w in-out.
constexpr int kNewtonSteps = 100;
constexpr int kNewtonRetries = 2;
constexpr double kHandoffCut = 4.0;
// Verbose output: a stage's evaluation and repair counts (the difference of two snapshots).
void print_counts(const cot_stats& a, const cot_stats& b) {
  const double ev = std::max(1.0, (double)(b.evaluations - a.evaluations));
  std::printf("[ladder]   %lld evaluations; per evaluation: hand-offs %.1f degenerate + %.1f "
              "overflow, reappeared %.1f, check failures %.1f full + %.1f repair, cells "
              "rebuilt %.1f; Newton steps %lld, CG iterations %lld\n",
              b.evaluations - a.evaluations, (b.handoff_degenerate - a.handoff_degenerate) / ev,
              (b.handoff_overflow - a.handoff_overflow) / ev,
              (b.reappeared_cells - a.reappeared_cells) / ev,
              (b.check_failures - a.check_failures) / ev,
              (b.repair_check_failures - a.repair_check_failures) / ev,
              (b.repaired_cells - a.repaired_cells) / ev, b.newton_steps - a.newton_steps,
              b.cg_iterations - a.cg_iterations);
}

int stage(cot_context* ctx, double* w, const cot_solve_opts& so, double handoff, int np,
          std::vector<cot_report>& rep) {
  rep.assign(np, cot_report{});
  std::vector<cot_report> r(np);
  auto add = [&](bool last) {
    for (int p = 0; p < np; ++p) {
      const int it = rep[p].iterations + r[p].iterations;
      const int ev = rep[p].evaluations + r[p].evaluations;
      if (last) rep[p] = r[p];
      rep[p].iterations = it;
      rep[p].evaluations = ev;
    }
  };
  cot_solve_opts nt = so;
  nt.max_iter = std::min(so.max_iter, kNewtonSteps);
  bool all = false;
  double tol = handoff;
  for (int attempt = 0; attempt <= kNewtonRetries && !all && tol > so.tol; ++attempt) {
    cot_solve_opts a = so;
    a.stop_rule = COT_STOP_HANDOFF;
    a.tol = tol;
    int rc = cot_solve(ctx, w, &a, r.data());
    if (rc != COT_OK && rc != COT_ENOTCONV) return rc;
    add(false);
    rc = cot_solve_newton(ctx, w, &nt, r.data());
    if (rc != COT_OK && rc != COT_ENOTCONV) return rc;
    add(true);
    all = true;
    double stalled = INFINITY;
    for (int p = 0; p < np; ++p) {
      all = all && r[p].converged;
      if (!r[p].converged) stalled = std::min(stalled, r[p].max_rel_err);
    }
    tol = std::min(tol, stalled) / kHandoffCut;
  }
  if (!all) {
    const int rc = cot_solve(ctx, w, &so, r.data());
    if (rc != COT_OK && rc != COT_ENOTCONV) return rc;
    add(true);
  }
  return COT_OK;
}

}  // namespace

extern "C" {

void cot_default_ladder_opts(const cot_problem* problems, int n_problems, cot_ladder_opts* lo) {
  if (!lo) return;
  std::memset(lo, 0, sizeof(*lo));
  // measured on MI355X (N = 65,536 shapes, 64 uneven Gaussian-mixture problems, N = 262,144
  // tilted curves): coarse levels to 0.1 rather than 1e-2 and the Newton hand-off at max-rel 2
  // rather than 1 took 9%, 44% and 7% off the solves (the next level's transfer error dwarfs a
  // coarse level's last decade; the last L-BFGS iterations before max-rel 1 toggle empty cells
  // while Newton steps converge); a hand-off at 3 made the batch 2.6x slower
  lo->coarse_tol = kDefaultCoarseTol;
  lo->handoff = kDefaultHandoff;
  lo->seed = 0x5eedull;
  if (!problems || n_problems <= 0) return;
  int nmin = problems[0].n_sites;
  bool sparse = false;
  for (int p = 0; p < n_problems; ++p) {
    const cot_problem& pr = problems[p];
    nmin = std::min(nmin, pr.n_sites);
    if (!pr.rho) continue;
    const size_t P = (size_t)pr.pm * pr.pn;
    double s = 0.0, mn = INFINITY;
    for (size_t k = 0; k < P; ++k) {
      s += pr.rho[k];
      mn = std::min(mn, pr.rho[k]);
    }
    if (P && mn < 1e-3 * s / (double)P) sparse = true;
  }
  // each coarse level holds at most half the smallest problem's sites
  for (int side = 16; 2.0 * side * side <= (double)nmin && lo->n_levels < COT_LADDER_MAX;
       side *= 2) {
    lo->sides[lo->n_levels] = side;
    lo->floors[lo->n_levels] = sparse ? 0.1 * std::pow(10.0, -0.5 * lo->n_levels) : 0.0;
    ++lo->n_levels;
  }
  if (sparse) {
    lo->n_final_floors = 1;
    lo->final_floors[0] = 1e-3;
  }
}

int cot_solve_ladder(cot_context* ctx, const cot_problem* problems, double* w,
                     const cot_solve_opts* opt, const cot_ladder_opts* lopt,
                     cot_report* reports) {
  if (!ctx || !problems || !w) return set_error(COT_EINVAL, "null argument");
  const auto t0 = std::chrono::steady_clock::now();
  const int np = cot::context_problems(ctx);
  {
    long long tot = 0;
    for (int p = 0; p < np; ++p) {
      if (problems[p].n_sites <= 0 || !problems[p].sites || !problems[p].nu || !problems[p].rho)
        return set_error(COT_EINVAL, "ladder: problem " + std::to_string(p) + " is empty");
      tot += problems[p].n_sites;
    }
    if (tot != cot_total_sites(ctx))
      return set_error(COT_EINVAL, "ladder: the problems are not the ones of the context");
  }
  cot_ladder_opts lo;
  if (lopt) {
    lo = *lopt;
  } else {
    cot_default_ladder_opts(problems, np, &lo);
  }
  if (lo.n_levels < 0 || lo.n_levels > COT_LADDER_MAX || lo.n_final_floors < 0 ||
      lo.n_final_floors > 4)
    return set_error(COT_EINVAL, "ladder: bad level counts");
  for (int l = 0; l < lo.n_levels; ++l)
    if (lo.sides[l] < 2 || (l > 0 && lo.sides[l] <= lo.sides[l - 1]))
      return set_error(COT_EINVAL, "ladder: level sides must be >= 2 and increasing");
  cot_solve_opts so;
  if (opt) {
    so = *opt;
  } else {
    cot_default_solve_opts(&so);
  }
  const cot_options base = *cot::context_options(ctx);
  cot_solve_opts coarse = so;
  coarse.stop_rule = COT_STOP_MAXREL;
  coarse.tol = lo.coarse_tol > 0 ? lo.coarse_tol : kDefaultCoarseTol;
  const double handoff = lo.handoff > 0 ? lo.handoff : kDefaultHandoff;

  // the targets' block sums for every level
  std::vector<Pyramid> pyr(np);
  for (int p = 0; p < np; ++p) {
    std::vector<int> fs;
    for (int l = 0; l < lo.n_levels; ++l)
      fs.push_back(target_factor(problems[p], level_want(problems[p], lo.sides[l])));
    build_pyramid(problems[p], fs, pyr[p]);
  }
  // coarse levels, all problems batched per level
  std::vector<std::vector<double>> field(np);
  std::vector<cot_report> rep;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  // every level's sites and target depend only on the problems: each level's are built on a
  // worker thread from the start, while the levels below it solve
  std::vector<std::vector<Level>> lvs(lo.n_levels, std::vector<Level>(np));
  std::vector<std::future<void>> built;
  for (int l = 0; l < lo.n_levels; ++l)
    built.push_back(std::async(std::launch::async, [&, l] {
      for (int p = 0; p < np; ++p)
        coarse_level(problems[p], pyr[p], frame_of(lo, problems[p], p), lo.sides[l],
                     lo.floors[l], lo.seed + 7919ull * (l + 1) + p, lvs[l][p]);
    }));
  for (int l = 0; l < lo.n_levels; ++l) {
    const auto tl0 = now();
    const int side = lo.sides[l], n = side * side;
    built[l].get();
    std::vector<Level>& lv = lvs[l];
    std::vector<cot_problem> cp(np);
    std::vector<double> wl((size_t)np * n, 0.0);
    for (int p = 0; p < np; ++p) {
      const double* fr = frame_of(lo, problems[p], p);
      cp[p] = as_problem(lv[p], n);
      if (l > 0) weights_from(field[p], lo.sides[l - 1], fr, lv[p].sites.data(), n, &wl[(size_t)p * n]);
    }
    cot_options oo = base;
    if (lo.level_options) oo = lo.level_options[l];
    const auto tl1 = now();
    cot_context* cc = nullptr;
    int rc = cot_create(cp.data(), np, &oo, &cc);
    if (rc != COT_OK) return rc;
    const auto tl2 = now();
    rc = stage(cc, wl.data(), coarse, handoff, np, rep);
    std::vector<double> cen(2 * (size_t)np * n);
    if (rc == COT_OK) rc = cot_centroids(cc, wl.data(), 1, cen.data(), nullptr);
    const auto tl3 = now();
    cot_stats cs{};
    if (so.verbose > 0 && rc == COT_OK) rc = cot_get_stats(cc, &cs);
    cot_destroy(cc);
    if (rc != COT_OK) return rc;
    for (int p = 0; p < np; ++p)
      field_from(lv[p], frame_of(lo, problems[p], p), &wl[(size_t)p * n], &cen[2 * (size_t)p * n],
                 field[p]);
    if (so.verbose > 0) {
      int it = 0;
      for (const cot_report& r : rep) it = std::max(it, r.iterations);
      std::printf("[ladder] level %d: %d problems, iterations max %d; ms: set-up %.2f create %.2f "
                  "solve %.2f destroy+field %.2f\n",
                  side, np, it, ms(tl0, tl1), ms(tl1, tl2), ms(tl2, tl3), ms(tl3, now()));
      print_counts(cot_stats{}, cs);
      std::fflush(stdout);
    }
  }
  const auto tf0 = now();
  // the full-resolution start
  size_t off = 0;
  for (int p = 0; p < np; ++p) {
    const cot_problem& pr = problems[p];
    if (lo.n_levels > 0) {
      weights_from(field[p], lo.sides[lo.n_levels - 1], frame_of(lo, pr, p), pr.sites, pr.n_sites,
                   w + off);
    } else {
      std::fill(w + off, w + off + pr.n_sites, 0.0);
    }
    off += pr.n_sites;
  }
  // floor stages at full resolution
  for (int q = 0; q < lo.n_final_floors; ++q) {
    std::vector<Level> lv(np);
    std::vector<cot_problem> cp(np);
    for (int p = 0; p < np; ++p) {
      const cot_problem& pr = problems[p];
      const size_t P = (size_t)pr.pm * pr.pn;
      double s = 0.0;
      for (size_t k = 0; k < P; ++k) s += pr.rho[k];
      const double add = lo.final_floors[q] * s / (double)P, sc = s / (s + add * (double)P);
      lv[p].rho.resize(P);
      for (size_t k = 0; k < P; ++k) lv[p].rho[k] = (pr.rho[k] + add) * sc;
      cp[p] = pr;
      cp[p].rho = lv[p].rho.data();
    }
    cot_context* cc = nullptr;
    int rc = cot_create(cp.data(), np, &base, &cc);
    if (rc != COT_OK) return rc;
    rc = stage(cc, w, coarse, handoff, np, rep);
    cot_destroy(cc);
    if (rc != COT_OK) return rc;
    if (so.verbose > 0) {
      std::printf("[ladder] floor %g done\n", lo.final_floors[q]);
      std::fflush(stdout);
    }
  }
  // the exact target, to the caller's stopping rule
  const auto tf1 = now();
  cot_stats s0{}, s1{};
  int rc = so.verbose > 0 ? cot_get_stats(ctx, &s0) : COT_OK;
  if (rc == COT_OK) rc = stage(ctx, w, so, handoff, np, rep);
  if (rc != COT_OK) return rc;
  if (so.verbose > 0) {
    const auto tf2 = now();
    if ((rc = cot_get_stats(ctx, &s1)) != COT_OK) return rc;
    std::printf("[ladder] final: ms: start and floor stages %.2f, solve %.2f\n", ms(tf0, tf1),
                ms(tf1, tf2));
    print_counts(s0, s1);
    std::fflush(stdout);
  }
  // the reference's gauge: its L-BFGS from w = 0 keeps sum(w) = 0 (the gradient sums to zero);
  // Newton steps, whose preconditioned directions need not, may shift the constant, which
  // changes no cell
  off = 0;
  for (int p = 0; p < np; ++p) {
    const int n = problems[p].n_sites;
    double mean = 0.0;
    for (int k = 0; k < n; ++k) mean += w[off + k];
    mean /= n;
    for (int k = 0; k < n; ++k) w[off + k] -= mean;
    off += n;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  bool all = true;
  for (int p = 0; p < np; ++p) {
    all = all && rep[p].converged;
    rep[p].seconds = secs;
    if (reports) reports[p] = rep[p];
  }
  if (!all) return set_error(COT_ENOTCONV, "ladder: the final stage did not reach the tolerance");
  return COT_OK;
}

}  // extern "C"
