// caustic_gpu -- drop-in replacement of the reference's `caustic_GPU <ot_time> <xiushi>` process
// (Optimal_Transport_rl/main.cu:51-342), run with cwd = Optimal_Transport_rl/build.
//
// Reads   ../set_json/<xiushi>/_py_sp_settings.json
//         ../../data/pic_data/<picture_para.name>.txt
//         ../seed_point/<xiushi>/{points,proj_xy,left_idx,right_idx,bottom_idx,top_idx,
//                                 inside_light_direct}.txt, ../seed_point/<xiushi>/<t>/height.txt
//         ../../data/weights/<xiushi>/<t-1>/final_proj_point_weights.txt   (t >= 1)
// Writes  ../../data/weights/<xiushi>/<t>/final_proj_point_weights.txt   (%g, 6 digits)
//         ../../data/point_data/<xiushi>/<t>/{out.txt,_py_sp_settings.json}
//         ../../data/point_data/<xiushi>/in.txt                            (t = 0)
// and, when the settings hold "mi355x": {"full_precision_weights": true}, also
//         ../../data/weights/<xiushi>/<t>/final_proj_point_weights_full.txt (%.17g).
//
// "mi355x": {"solver": "ladder"} solves t = 0 coarse to fine (cot_solve_ladder: coarse grids,
// density-floor stages for targets with zero-density regions, then the exact target to the same
// stopping rule, |mass - nu|_2 < convergence_tolerance with no empty cell) instead of L-BFGS from
// w = 0; "lbfgs" (the default) is the reference's algorithm.
//
// and the diagnostics ../render_result/<xiushi>/<t>/{render_initial,render_<iter>}.txt and
// ../mesh_output/<xiushi>/<t>/{mesh_initial,mesh_<iter>}.obj (t >= 1: pre_render_initial.txt,
// pre_initial_mesh.obj, <t-1>_render_<iter>.txt, <t-1>_mesh_<iter>.obj), unless the settings
// hold "mi355x": {"diagnostics": false}.
//
// Differences from the reference executable: errors end the process with a message and a
// non-zero status (never abort or an unspecified status); the periodic render, mesh and weight
// snapshots of the iterations are not written; the meshes hold the cells inside proj_domain.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "caustic_ot.h"

namespace fs = std::filesystem;

namespace {

// ------------------------------------------------------------------ minimal JSON reader
struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<JVal> arr;
  std::map<std::string, JVal> obj;

  const JVal& at(const std::string& key) const {
    if (kind != Obj) throw std::runtime_error("JSON: not an object when reading '" + key + "'");
    auto it = obj.find(key);
    if (it == obj.end()) throw std::runtime_error("JSON: missing key '" + key + "'");
    return it->second;
  }
  const JVal* find(const std::string& key) const {
    if (kind != Obj) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
  }
  double as_num() const {
    if (kind == Num) return num;
    if (kind == Bool) return b ? 1.0 : 0.0;
    throw std::runtime_error("JSON: value is not a number");
  }
  bool as_bool() const {
    if (kind == Bool) return b;
    if (kind == Num) return num != 0.0;
    throw std::runtime_error("JSON: value is not a boolean");
  }
  const std::string& as_str() const {
    if (kind != Str) throw std::runtime_error("JSON: value is not a string");
    return str;
  }
  std::vector<double> as_nums() const {
    if (kind != Arr) throw std::runtime_error("JSON: value is not an array");
    std::vector<double> v;
    for (auto& e : arr) v.push_back(e.as_num());
    return v;
  }
};

class JParser {
 public:
  explicit JParser(const std::string& s) : s_(s) {}
  JVal parse() {
    JVal v = value();
    ws();
    if (p_ != s_.size()) throw std::runtime_error("JSON: trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;
  void ws() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
  }
  char peek() {
    ws();
    if (p_ >= s_.size()) throw std::runtime_error("JSON: unexpected end");
    return s_[p_];
  }
  void expect(char c) {
    if (peek() != c) throw std::runtime_error(std::string("JSON: expected '") + c + "'");
    ++p_;
  }
  std::string string_lit() {
    expect('"');
    std::string out;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c == '\\' && p_ < s_.size()) {
        char e = s_[p_++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {  // keep the escape verbatim; keys used here are ASCII
            out += "\\u";
            for (int k = 0; k < 4 && p_ < s_.size(); ++k) out += s_[p_++];
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (p_ >= s_.size()) throw std::runtime_error("JSON: unterminated string");
    ++p_;
    return out;
  }
  JVal value() {
    char c = peek();
    JVal v;
    if (c == '{') {
      ++p_;
      v.kind = JVal::Obj;
      if (peek() == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        std::string k = string_lit();
        expect(':');
        v.obj[k] = value();
        char d = peek();
        ++p_;
        if (d == '}') break;
        if (d != ',') throw std::runtime_error("JSON: expected ',' or '}'");
      }
    } else if (c == '[') {
      ++p_;
      v.kind = JVal::Arr;
      if (peek() == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        char d = peek();
        ++p_;
        if (d == ']') break;
        if (d != ',') throw std::runtime_error("JSON: expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = JVal::Str;
      v.str = string_lit();
    } else if (s_.compare(p_, 4, "true") == 0) {
      p_ += 4;
      v.kind = JVal::Bool;
      v.b = true;
    } else if (s_.compare(p_, 5, "false") == 0) {
      p_ += 5;
      v.kind = JVal::Bool;
      v.b = false;
    } else if (s_.compare(p_, 4, "null") == 0) {
      p_ += 4;
    } else {
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) throw std::runtime_error("JSON: bad value");
      p_ += (size_t)(e - b);
      v.kind = JVal::Num;
    }
    return v;
  }
};

// ------------------------------------------------------------------ text formats
bool read_pairs(const std::string& path, std::vector<double>& out) {
  std::ifstream f(path);
  if (!f) return false;
  double x, y;
  while (f >> x >> y) {
    out.push_back(x);
    out.push_back(y);
  }
  return true;
}

bool read_ints(const std::string& path, std::vector<int>& out) {
  std::ifstream f(path);
  if (!f) return false;
  int v;
  while (f >> v) out.push_back(v);
  return true;
}

bool read_doubles(const std::string& path, std::vector<double>& out) {
  std::ifstream f(path);
  if (!f) return false;
  double v;
  while (f >> v) out.push_back(v);
  return true;
}

// Picture-shaped text in the reader's orientation: m rows of n values, fixed with 6 decimals,
// row r column c holding data[c * m + (m - 1 - r)] * scale (save_array_to_txt,
// basic_function.cuh:205-248).
void write_picture(const std::string& path, const std::vector<double>& data, int m, int n,
                   double scale) {
  fs::create_directories(fs::path(path).parent_path());
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return;
  for (int r = 0; r < m; ++r) {
    for (int c = 0; c < n; ++c)
      std::fprintf(f, c + 1 < n ? "%.6f " : "%.6f", data[(size_t)c * m + (m - 1 - r)] * scale);
    std::fputc('\n', f);
  }
  std::fclose(f);
}

// Wavefront OBJ of the cells: each non-empty cell's vertices as "v x y 0.0", then its face
// (writeOBJ, basic_function.cuh:15-59).
void write_obj(const std::string& path, const std::vector<int>& counts,
               const std::vector<double>& verts) {
  fs::create_directories(fs::path(path).parent_path());
  std::ofstream f(path);
  int idx = 1;
  for (size_t i = 0; i < counts.size(); ++i) {
    const int nv = counts[i];
    if (nv < 3) continue;
    for (int k = 0; k < nv; ++k)
      f << "v " << verts[(i * COT_MAX_VERTS + k) * 2] << " "
        << verts[(i * COT_MAX_VERTS + k) * 2 + 1] << " 0.0\n";
    f << "f ";
    for (int k = 0; k < nv; ++k) f << idx + k << " ";
    f << "\n";
    idx += nv;
  }
}

// m text rows x n columns; element (ix, iy) = row m-1-ix, column iy, stored at iy*m + ix
// (basic_function.cuh:145-204).
bool read_picture(const std::string& path, std::vector<double>& data, int& m, int& n) {
  std::ifstream f(path);
  if (!f) return false;
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::vector<double> r;
    double v;
    while (ss >> v) r.push_back(v);
    if (r.empty()) continue;
    if (!rows.empty() && r.size() != rows[0].size()) return false;
    rows.push_back(std::move(r));
  }
  if (rows.empty()) return false;
  m = (int)rows.size();
  n = (int)rows[0].size();
  data.assign((size_t)m * n, 0.0);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < m; ++x) data[(size_t)y * m + x] = rows[m - 1 - x][y];
  return true;
}

struct CtxGuard {
  cot_context* c = nullptr;
  ~CtxGuard() {
    if (c) cot_destroy(c);
  }
};

// Solid angle under which the point light sees the sample domain [x0, x1, y0, y1] in the plane
// z: the rectangle (x0,y0) (x0,y1) (x1,y1) (x1,y0) fanned from its first corner, each spherical
// triangle 2 atan(a.(b x c) / (1 + a.b + b.c + c.a)) of the unit vectors to its corners, a
// triangle with |denominator| < 1e-6 skipped, magnitude of the sum (len_solid_angle and
// compute_solid_angle, basic_function.cuh:653-658, 558-588).  One polygon: host arithmetic.
double domain_solid_angle(const std::vector<double>& d, const std::vector<double>& light,
                          double z) {
  const double cx[4] = {d[0], d[0], d[1], d[1]}, cy[4] = {d[2], d[3], d[3], d[2]};
  double u[4][3];
  for (int k = 0; k < 4; ++k) {
    const double e[3] = {cx[k] - light[0], cy[k] - light[1], z - light[2]};
    const double r = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (int j = 0; j < 3; ++j) u[k][j] = e[j] / r;
  }
  auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  double sum = 0.0;
  for (int k = 1; k + 1 < 4; ++k) {
    const double *a = u[0], *b = u[k], *c = u[k + 1];
    const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2],
                          b[0] * c[1] - b[1] * c[0]};
    const double den = 1.0 + dot(a, b) + dot(b, c) + dot(a, c);
    if (std::fabs(den) >= 1e-6) sum += 2.0 * std::atan(dot(a, bc) / den);
  }
  return std::fabs(sum);
}

int die(const std::string& msg, int code = 1) {
  std::cerr << "caustic_gpu: " << msg << std::endl;
  return code;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return die("usage: caustic_gpu <ot_time> <xiushi>");
  const std::string ot_time = argv[1], xiushi = argv[2];
  const std::string json_path = "../set_json/" + xiushi + "/_py_sp_settings.json";
  std::ifstream jf(json_path);
  if (!jf) return die("cannot read " + json_path, 255);
  std::stringstream jss;
  jss << jf.rdbuf();

  std::string picture_name, light_type;
  int total_iter = 0, history_size = 0;
  double tol = 0, initial_step = 1.0;
  std::vector<double> sample_domain, proj_domain, point_light;
  bool full_precision = false;
  bool diagnostics = true;
  std::string solver = "lbfgs";
  int aa_history_size = 7;
  double aa_fixed_para = -2.0, z_of_len = 0.0, z_of_receiver = 0.0, refraction_index = 1.5;
  bool if_aa_op_initial = true;
  try {
    JVal J = JParser(jss.str()).parse();
    (void)J.at("test_num").as_str();
    const JVal& sp = J.at("sample_point_para");
    (void)sp.at("data_name").as_str();
    total_iter = (int)sp.at("total_iter").as_num();
    tol = sp.at("convergence_tolerance").as_num();
    history_size = (int)sp.at("history_size").as_num();
    aa_history_size = (int)sp.at("aa_history_size").as_num();
    aa_fixed_para = sp.at("aa_fixed_para").as_num();
    if_aa_op_initial = sp.at("if_aa_op_initial").as_bool();
    initial_step = sp.at("initial_step").as_num();
    sample_domain = sp.at("sample_domain").as_nums();
    const JVal& lp = J.at("len_para");
    z_of_len = lp.at("z_of_len").as_num();
    z_of_receiver = lp.at("z_of_receiver").as_num();
    refraction_index = lp.at("refraction_index").as_num();
    proj_domain = lp.at("proj_domain").as_nums();
    const JVal& li = J.at("light_para");
    light_type = li.at("light_type").as_str();
    point_light = li.at("point_light_coordinate").as_nums();
    picture_name = J.at("picture_para").at("name").as_str();
    if (const JVal* ext = J.find("mi355x"))
      if (const JVal* fp = ext->find("full_precision_weights")) full_precision = fp->as_bool();
    if (const JVal* ext = J.find("mi355x"))
      if (const JVal* dg = ext->find("diagnostics")) diagnostics = dg->as_bool();
    if (const JVal* ext = J.find("mi355x"))
      if (const JVal* sv = ext->find("solver")) solver = sv->as_str();
  } catch (const std::exception& e) {
    return die(std::string("settings: ") + e.what(), 2);
  }
  if (sample_domain.size() != 4 || proj_domain.size() != 4)
    return die("settings: sample_domain and proj_domain need 4 numbers", 2);
  if (solver != "lbfgs" && solver != "ladder")
    return die("settings: mi355x.solver must be \"lbfgs\" or \"ladder\"", 2);

  // inputs
  const std::string seed = "../seed_point/" + xiushi + "/";
  std::vector<double> picture, points, proj_xy, inside, height;
  std::vector<int> idx[4];
  int pm = 0, pn = 0;
  const std::string pic_path = "../../data/pic_data/" + picture_name + ".txt";
  if (!read_picture(pic_path, picture, pm, pn)) return die("cannot read picture " + pic_path);
  if (!read_pairs(seed + "points.txt", points)) return die("cannot read " + seed + "points.txt");
  if (!read_pairs(seed + "proj_xy.txt", proj_xy))
    return die("cannot read " + seed + "proj_xy.txt");
  const char* idx_names[4] = {"left_idx.txt", "right_idx.txt", "bottom_idx.txt", "top_idx.txt"};
  for (int k = 0; k < 4; ++k)
    if (!read_ints(seed + idx_names[k], idx[k])) return die("cannot read " + seed + idx_names[k]);
  if (!read_doubles(seed + "inside_light_direct.txt", inside))
    return die("cannot read " + seed + "inside_light_direct.txt");
  if (!read_doubles(seed + ot_time + "/height.txt", height))
    return die("cannot read " + seed + ot_time + "/height.txt");
  const int N = (int)(points.size() / 2);
  if (N == 0 || proj_xy.size() != points.size()) return die("points/proj_xy size mismatch");
  const bool is_point = light_type == "point_light";
  if (!is_point && light_type != "parallel_light")
    return die("unknown light_type '" + light_type + "'", 3);
  if (is_point && (point_light.size() != 3 || point_light[2] == z_of_len))
    return die("settings: point_light_coordinate needs 3 numbers off the lens plane", 2);
  if (inside.size() < 3 * (size_t)(points.size() / 2) || height.size() < points.size() / 2)
    return die("inside_light_direct.txt / height.txt hold fewer rows than points.txt");
  std::vector<double> w_start(points.size() / 2, 0.0);
  if (ot_time != "0") {
    // warm start from the previous run (main.cu:193-198)
    int prev = 0;
    try {
      prev = std::stoi(ot_time) - 1;
    } catch (const std::exception&) {
      return die("ot_time must be an integer", 2);
    }
    const std::string wpath =
        "../../data/weights/" + xiushi + "/" + std::to_string(prev) + "/final_proj_point_weights.txt";
    std::vector<double> wprev;
    if (!read_doubles(wpath, wprev) || wprev.size() != points.size() / 2)
      return die("cannot read " + wpath);
    w_start = wprev;
  }

  const double xs = sample_domain[1] - sample_domain[0], ys = sample_domain[3] - sample_domain[2];
  const double area_scalar = xs * ys;

  // source masses: Voronoi cells of the sample points inside the sample domain, weighted by
  // their area (parallel light) or by the solid angle the light sees them under (point light,
  // each divided by the solid angle of the whole sample domain) (ot_op.cuh:90-92,154-206)
  std::vector<double> nu(N, 1.0), area(N, 0.0), w0(N, 0.0), incident;
  {
    std::vector<double> one(1, 1.0);
    cot_problem pr{};
    pr.n_sites = N;
    pr.sites = points.data();
    pr.nu = nu.data();
    pr.rho = one.data();
    pr.pm = 1;
    pr.pn = 1;
    for (int k = 0; k < 4; ++k) pr.proj_domain[k] = sample_domain[k];
    CtxGuard g;
    if (cot_create(&pr, 1, nullptr, &g.c) != COT_OK)
      return die(std::string("source cells: ") + cot_last_error(), 4);
    if (!is_point) {
      if (cot_cell_areas(g.c, w0.data(), area.data()) != COT_OK)
        return die(std::string("source cells: ") + cot_last_error(), 4);
    } else {
      if (cot_cell_solid_angles(g.c, w0.data(), point_light.data(), z_of_len, area.data()) !=
          COT_OK)
        return die(std::string("source cells: ") + cot_last_error(), 4);
      const double whole = domain_solid_angle(sample_domain, point_light, z_of_len);
      if (!(whole > 0.0)) return die("the light sees the sample domain under no solid angle", 4);
      for (double& a : area) a = a / whole * area_scalar;
      // direction of the light at each point of the lens plane (main.cu:218-231)
      incident.resize(3 * (size_t)N);
      for (int i = 0; i < N; ++i) {
        const double d[3] = {points[2 * i] - point_light[0], points[2 * i + 1] - point_light[1],
                             z_of_len - point_light[2]};
        const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) incident[3 * (size_t)i + k] = d[k] / r;
      }
    }
    double tot = 0.0;
    for (double a : area) tot += a;
    for (int i = 0; i < N; ++i) nu[i] = area[i] / tot * area_scalar;
  }

  // target: rho = picture / (sum / area_scalar)  (ot_op.cuh:81-88,154)
  double gray = 0.0;
  for (double v : picture) gray += v;
  const double gray_multiple = gray / area_scalar;
  for (double& v : picture) v /= gray_multiple;

  std::vector<double> w = w_start;
  cot_problem pr{};
  pr.n_sites = N;
  pr.sites = proj_xy.data();
  pr.nu = nu.data();
  pr.rho = picture.data();
  pr.pm = pm;
  pr.pn = pn;
  for (int k = 0; k < 4; ++k) pr.proj_domain[k] = proj_domain[k];
  CtxGuard g;
  if (cot_create(&pr, 1, nullptr, &g.c) != COT_OK)
    return die(std::string("context: ") + cot_last_error(), 4);
  cot_solve_opts so;
  cot_default_solve_opts(&so);
  so.history = history_size;
  so.max_iter = total_iter;
  so.tol = tol;
  so.initial_step = initial_step;
  so.stop_rule = COT_STOP_REFERENCE;
  so.h0_threshold = 1e-5;
  so.verbose = 20;
  cot_report rep{};
  int rc;
  // diagnostic render and mesh of the cells (ot_op.cuh:214-219, 281-286, 499-533, 798-820,
  // 1270-1296): before the solve and at its end
  const std::string rdir = "../render_result/" + xiushi + "/" + ot_time + "/";
  const std::string mdir = "../mesh_output/" + xiushi + "/" + ot_time + "/";
  const std::string pre = ot_time == "0" ? "" : "pre_";
  auto diagnose = [&](const std::string& render_name, const std::string& mesh_name) -> bool {
    std::vector<double> img((size_t)pm * pn);
    std::vector<int> counts(N);
    std::vector<double> verts((size_t)N * COT_MAX_VERTS * 2);
    if (cot_render(g.c, w.data(), img.data()) != COT_OK ||
        cot_cell_polygons(g.c, w.data(), counts.data(), verts.data()) != COT_OK)
      return false;
    write_picture(rdir + render_name, img, pm, pn, gray_multiple);
    write_obj(mdir + mesh_name, counts, verts);
    return true;
  };
  if (diagnostics && !diagnose(pre + "render_initial.txt",
                               ot_time == "0" ? "mesh_initial.obj" : "pre_initial_mesh.obj"))
    return die(std::string("render: ") + cot_last_error(), 5);
  if (ot_time == "0" && solver == "ladder") {
    cot_ladder_opts lo;
    cot_default_ladder_opts(&pr, 1, &lo);
    rc = cot_solve_ladder(g.c, &pr, w.data(), &so, &lo, &rep);
  } else if (ot_time == "0") {
    rc = cot_solve(g.c, w.data(), &so, &rep);
  } else {
    // reflection-aware path: Fresnel-weighted source masses, L-BFGS then Anderson
    cot_reflection rf{};
    rf.points = points.data();
    rf.inside_dir = inside.data();
    rf.height = height.data();
    rf.clip_area = area.data();
    rf.z_of_len = z_of_len;
    rf.z_of_receiver = z_of_receiver;
    rf.refraction_index = refraction_index;
    rf.area_scalar = area_scalar;
    rf.aa_history_size = aa_history_size;
    rf.aa_fixed_para = aa_fixed_para;
    rf.if_aa_op_initial = if_aa_op_initial ? 1 : 0;
    rf.incident_dir = is_point ? incident.data() : nullptr;
    rc = cot_solve_reflection(g.c, w.data(), &so, &rf, &rep);
  }
  if (rc != COT_OK && rc != COT_ENOTCONV) return die(std::string("solve: ") + cot_last_error(), 5);
  if (diagnostics) {
    const std::string it = std::to_string(rep.iterations);
    const std::string tag = ot_time == "0" ? "" : std::to_string(std::stoi(ot_time) - 1) + "_";
    if (!diagnose(tag + "render_" + it + ".txt", tag + "mesh_" + it + ".obj"))
      return die(std::string("render: ") + cot_last_error(), 5);
  }
  std::printf("caustic_gpu: %s after %d iterations (%d evaluations), |g| %.6e, %.2f s\n",
              rep.converged ? "converged" : "stopped", rep.iterations, rep.evaluations,
              rep.grad_norm, rep.seconds);
  std::vector<double> cen(2 * (size_t)N);
  std::vector<unsigned char> exist(N);
  if (cot_centroids(g.c, w.data(), 1, cen.data(), exist.data()) != COT_OK)
    return die(std::string("centroids: ") + cot_last_error(), 5);

  // outputs (main.cu:291-334)
  const std::string wdir = "../../data/weights/" + xiushi + "/" + ot_time;
  fs::create_directories(wdir);
  {
    std::ofstream f(wdir + "/final_proj_point_weights.txt");
    for (int i = 0; i < N; ++i) f << w[i] << "\n";
  }
  if (full_precision) {
    FILE* f = std::fopen((wdir + "/final_proj_point_weights_full.txt").c_str(), "w");
    if (f) {
      for (int i = 0; i < N; ++i) std::fprintf(f, "%.17g\n", w[i]);
      std::fclose(f);
    }
  }
  const std::string pdir = "../../data/point_data/" + xiushi + "/" + ot_time;
  fs::create_directories(pdir);
  fs::copy_file(json_path, pdir + "/_py_sp_settings.json", fs::copy_options::overwrite_existing);
  if (ot_time == "0") {
    std::ofstream f("../../data/point_data/" + xiushi + "/in.txt");
    for (int i = 0; i < N; ++i)
      if (exist[i]) f << points[2 * i] << " " << points[2 * i + 1] << "\n";
  }
  {
    std::ofstream f(pdir + "/out.txt");
    for (int i = 0; i < N; ++i)
      if (exist[i]) f << cen[2 * i] << " " << cen[2 * i + 1] << "\n";
  }
  return rc == COT_OK ? 0 : 6;
}
