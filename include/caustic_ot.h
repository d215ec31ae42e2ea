/* caustic_ot.h -- C ABI of the MI355X semi-discrete optimal-transport dual solver.
 *
 * One context holds a batch of OT problems (one or many, uneven sizes) resident on one GPU:
 * sites, source masses nu, a piecewise-constant target image rho per problem, and the solver's
 * device state.  Pointer arguments are caller-owned HOST buffers in the caller's site order
 * (the library keeps its own bucket-sorted order internally), except where an entry point says
 * device memory (the *_dev variants and the renderers).  Calls return COT_OK or a status
 * code and never abort; cot_last_error() describes the last failure on the calling thread.
 *
 * Reference interfaces replaced (rmk523/caustic, Optimal_Transport_rl/):
 *   cot_create / cot_destroy   the device buffers direct_ot allocates and frees per call
 *                              (ot_op.cuh:102-152, 548-588)
 *   cot_evaluate               polygons_in_VD + compute_inte + compute_subtract + compute_f
 *                              (power_diagram.cpp:8-134; basic_function.cuh:743-952, 988-995)
 *   cot_solve                  the L-BFGS loop of direct_ot (ot_op.cuh:66-537)
 *   cot_centroids              compute_center, weighted or geometric, with if_exist
 *                              (basic_function.cuh:1056-1191; ot_op.cuh:539-546)
 *   cot_cell_areas             compute_clip_area, the source-mass step (basic_function.cuh:610-633)
 *   cot_irra_render            render() of the irra_render extension, the irradiance image of a
 *                              triangle mesh (render_exe_program/render.cu:93-171, 553-591)
 *   cot_irra_render_grad       render_grad() of the same extension
 *                              (render_exe_program/render.cu:173-420, 422-542, 593-728)
 *   cot_splat                  _gaussian_splat_raw_chunk and the point gradient of
 *                              render_gaussian_splat_two_pass_backward (utils/basic_function.py:674-772)
 *   cot_cell_solid_angles      solid_angle + compute_solid_angle, the point-light source-mass
 *                              step (basic_function.cuh:558-588, 635-651; ot_op.cuh:160-184)
 */
#ifndef CAUSTIC_OT_H_
#define CAUSTIC_OT_H_

#ifdef __cplusplus
extern "C" {
#endif

enum {
  COT_OK = 0,
  COT_EINVAL = 1,      /* bad argument */
  COT_ENOMEM = 2,      /* device allocation failed */
  COT_EHIP = 3,        /* HIP runtime error */
  COT_ENOTCONV = 4,    /* iteration limit reached before the stopping rule held */
  COT_ECOMPLEX = 5,    /* a cell exceeded the vertex limit at every trial step */
  COT_EUNVERIFIED = 6  /* cells still failed the conformity check at the largest ring */
};

enum { COT_PREC_FP64 = 0, COT_PREC_MIXED = 1 };
enum { COT_STOP_REFERENCE = 0, COT_STOP_MAXREL = 1, COT_STOP_HANDOFF = 2 };

typedef struct cot_context cot_context;

/* One problem.  rho is mass per pixel in the reference layout rho[ix + iy * pm], pm pixels
 * along x and pn along y over proj_domain = {x0, x1, y0, y1} (basic_function.cuh:145-204,
 * 758-760).  sites are n (x, y) pairs, nu n source masses. */
typedef struct {
  int n_sites;
  const double* sites;
  const double* nu;
  const double* rho;
  int pm, pn;
  double proj_domain[4];
} cot_problem;

typedef struct {
  int precision;            /* COT_PREC_FP64: FP64 image; COT_PREC_MIXED: FP32 image */
  double sites_per_bucket;  /* bucket grid density (default 1.0) */
  int ring_init;            /* first candidate ring radius (default 2) */
  int ring_max;             /* largest ring radius (default 32) */
  double verify_tol;        /* edge-match tolerance relative to the site spacing (default 1e-4) */
  int max_repairs;          /* conformity repair rounds per evaluation (default 12) */
  int int_lanes;            /* integration lanes per site: 4, 8 or 16 (default 8) */
  int repair_blocks;        /* blocks of the repair passes over the work list (default 256) */
  int full_pass_ring;       /* candidates of a FP32-geometry full pass: the previous cells'
                               one-ring (0: default, or 1) or two-ring (2); FP64 always uses
                               the two-ring */
  int repair_mode;          /* construction of the repair passes: 0 one wave per listed cell
                               (default), 1 one lane per cell; both build the same cells */
  int pass_mode;            /* construction of the full passes: 0 one lane per cell (default),
                               1 one wave per cell (more parallelism for small problems); both
                               build the same cells */
} cot_options;

typedef struct {
  int history;              /* L-BFGS memory m (history_size) */
  int max_iter;             /* total_iter */
  double tol;               /* convergence_tolerance */
  double initial_step;      /* initial_step */
  int stop_rule;            /* COT_STOP_REFERENCE: |g|_2 < tol with no empty cell (ot_op.cuh:510);
                               COT_STOP_MAXREL: max_i |mass_i - nu_i| / nu_i < tol;
                               COT_STOP_HANDOFF: that and no empty cell (the starting condition
                               of cot_solve_newton; the same as MAXREL for tol <= 1) */
  double h0_threshold;      /* H0 = y.s / y.y when y.y > h0_threshold, else 1 (1e-5 reference) */
  int verbose;              /* host loop: a line every `verbose` iterations; device loop: a
                               progress line every 2 s (0: silent) */
  int device_loop;          /* 1: device-resident rounds replayed from a hipGraph (default);
                               2: the same device-resident rounds launched without a graph;
                               0: host-driven rounds (one host sync per phase) */
  int accept_rule;          /* 0: accept a trial on f decrease (ot_op.cuh:397);
                               1: also accept when the directional derivative at the trial is
                                  <= 0 (the convex objective decreased) -- for FP32 geometry,
                                  whose f is too coarse to resolve the last decades */
  int repair_rounds;        /* conformity repair rounds of a line-search trial (default 3) */
  int newton_sweeps;        /* cot_solve_newton: block-local Jacobi sweeps of the conjugate
                               gradients' preconditioner (default 2; 0: plain Jacobi).  With
                               the coarse term, 2 sweeps cut the iterations ~20% at
                               N = 65,536 for ~2 us more per iteration */
  int newton_coarse;        /* cot_solve_newton: at most this many patches per problem in the
                               coarse term of the preconditioner, 0 or 2..1024 (default 256;
                               0: none).  Patches are unions of 128-site blocks; the term adds
                               P (P^T H P)^-1 P^T (inverted on the host each step, ~n^3/2
                               multiply-adds: 0.3-1.7 ms at 256) and cuts the iterations ~4x
                               at N = 65,536 */
  int small_cg;             /* cot_solve_newton: run the whole CG of a system of at most 8
                               blocks of 128 sites in one workgroup, one launch per Newton
                               step (default 1; 0: the multi-kernel iteration at every size) */
} cot_solve_opts;

typedef struct {
  int iterations;
  int evaluations;
  int converged;
  int status;
  double f;                 /* objective at the returned weights */
  double grad_norm;         /* |mass - nu|_2 */
  double max_rel_err;       /* max_i |mass_i - nu_i| / nu_i */
  double seconds;           /* wall time of the solve */
} cot_report;

void cot_default_options(cot_options* opt);
void cot_default_solve_opts(cot_solve_opts* opt);

int cot_create(const cot_problem* problems, int n_problems, const cot_options* opt,
               cot_context** out);
int cot_destroy(cot_context* ctx);
const char* cot_last_error(void);

/* Total sites over the batch. */
int cot_total_sites(const cot_context* ctx);

/* Evaluate the batch at weights w (all problems' sites concatenated in caller order).  Outputs
 * may be NULL: mass (per site), f (per problem objective), centroid (2 per site, rho-weighted,
 * geometric when the cell holds no mass), exist (1 byte per site). */
int cot_evaluate(cot_context* ctx, const double* w, double* mass, double* f, double* centroid,
                 unsigned char* exist);

/* Solve every problem from the weights in w (in-out).  reports: one per problem (may be NULL).
 * Returns COT_OK when every problem met its stopping rule, COT_ENOTCONV otherwise. */
int cot_solve(cot_context* ctx, double* w, const cot_solve_opts* opt, cot_report* reports);

/* Reflection-aware solve for ot_time >= 1 (aarl_direct_ot, ot_op.cuh:591-1385), one-problem
 * contexts only.  Source masses follow the lens's Fresnel transmittance at the current cells'
 * geometric centroids: nu_i = area_i T1_i T2_i / sum * area_scalar (ot_op.cuh:779-795).
 * Stage A (if_aa_op_initial): L-BFGS with those masses frozen (the rules of cot_solve, reference
 * stopping test).  Then the masses are re-estimated and stage B runs Anderson acceleration of
 * G(w) = w + aa_fixed_para * (mass - nu) with aa_history_size pairs, recomputing nu at every
 * trial, until |mass - nu|_2 < tol (ot_op.cuh:1112-1323).  w is in-out (caller order). */
typedef struct {
  const double* points;      /* n source points (x, y) in the lens plane (points.txt) */
  const double* inside_dir;  /* 3n directions inside the lens (inside_light_direct.txt) */
  const double* height;      /* n back-surface heights (<t>/height.txt) */
  const double* clip_area;   /* n source-cell areas inside the sample domain */
  double z_of_len, z_of_receiver, refraction_index;
  double area_scalar;        /* area of the sample domain */
  int aa_history_size;
  double aa_fixed_para;
  int if_aa_op_initial;
  /* 3n unit directions of the light arriving at the points: normalize((x, y, z_of_len) -
   * light) for a point light (main.cu:218-231); NULL for parallel light along +z (main.cu:211-217) */
  const double* incident_dir;
} cot_reflection;

int cot_solve_reflection(cot_context* ctx, double* w, const cot_solve_opts* opt,
                         const cot_reflection* refl, cot_report* report);

/* Damped Newton steps from w until the stopping rule holds -- opt->stop_rule COT_STOP_REFERENCE:
 * |mass - nu|_2 < opt->tol with no empty cell; otherwise max_i |mass_i - nu_i| / nu_i < opt->tol
 * -- (opt->max_iter Newton steps at most); w in-out, one report per problem.  No counterpart in the reference, whose
 * solver is L-BFGS only (ot_op.cuh:289-537): it serves the tail of a max-rel solve, after L-BFGS
 * has brought every cell to non-empty.  Each step builds the Hessian -- the cell-adjacency
 * Laplacian h_ij = int over the common edge of rho ds / (2 |s_i - s_j|) -- from the current
 * cells, solves H d = -g by preconditioned conjugate gradients (opt->newton_sweeps Jacobi sweeps
 * on each 128-site block's submatrix of H, plus with opt->newton_coarse the two-level patch
 * term), and halves the step until every cell keeps mass and |g|_2 falls by (1 - t/2)
 * (Kitagawa, Merigot, Thibert 2019).  The conjugate gradients run to |r| <= eta |rhs|,
 * eta = 0.1 max-rel clamped to [1e-4, 0.1], replayed from a captured graph in chunks of 32
 * iterations (16 with the patch term; launched directly when opt->device_loop == 2).  A problem with an empty or unverified cell, or whose step falls below
 * 1/4096, stops with status COT_ENOTCONV at its last accepted weights. */
int cot_solve_newton(cot_context* ctx, double* w, const cot_solve_opts* opt, cot_report* reports);

/* Coarse-to-fine solve from nothing but the problems (no counterpart in the reference, whose
 * solver starts L-BFGS from w = 0, ot_op.cuh:66-537; the stopping rule is the caller's).
 *
 * Level l solves every problem on a jittered sides[l] x sides[l] grid of its frame (the region
 * its sites sample): the target box-downsampled, plus floors[l] times its mean pixel mass, the
 * source masses of the fine sites binned on the grid; to max-rel coarse_tol, starting from the
 * previous level's solution carried up through the weight field (w = -U, U(node) =
 * -w_k + 2 (s_k - c_k).(s_k - node), c_k the cell's centroid, bilinear at the next sites).  The
 * full-resolution problem then runs the floor continuation final_floors[] (same rule, each
 * stage starting from the last) and finally the exact target, in ctx, to opt's stopping rule.
 * Every stage runs L-BFGS until max-rel < handoff with no empty cell, then damped Newton steps
 * (cot_solve_newton), and L-BFGS again if the Newton steps stall.
 *
 * The floors make a target with zero-density regions tractable: without them, cells lying on
 * zero density hold no mass, the dual objective is flat along their weights and L-BFGS
 * crawls (the reference's own configuration -- N = 160,978, 76.6 % of the pixels zero -- is
 * still at |mass - nu|_2 = 0.93 after 3,000 iterations from w = 0); each floor stage has every
 * cell holding mass, so its Newton steps apply, and the last one leaves every cell touching the
 * support.  problems: the array ctx was created from (the host copies of the targets are read
 * again); w: out (caller order); reports: one per problem, of the final stage (seconds: the
 * whole call). */
#define COT_LADDER_MAX 8
typedef struct {
  int n_levels;                     /* coarse levels, 0 .. COT_LADDER_MAX */
  int sides[COT_LADDER_MAX];        /* sites per side of each level's grid, increasing */
  double floors[COT_LADDER_MAX];    /* each level's density floor, relative to its mean pixel */
  int n_final_floors;               /* floor stages at full resolution, 0 .. 4 */
  double final_floors[4];
  double coarse_tol;                /* max-rel of the coarse levels and floor stages */
  double handoff;                   /* L-BFGS until this max-rel, then Newton steps */
  unsigned long long seed;          /* jitter of the coarse grids */
  const double* frames;             /* 4 per problem (x0, x1, y0, y1): the region the sites
                                       sample; NULL: proj_domain */
  const cot_options* level_options; /* launch shapes of each level (n_levels entries), NULL:
                                       defaults at the context's precision */
} cot_ladder_opts;

/* Defaults for these problems: levels 16, 32, 64, ... while a level holds at most half the
 * smallest problem's sites (side^2 <= n_min / 2); floors only for a target whose smallest pixel is below 1e-3 of
 * its mean (zero-density regions): 0.1 at the first level, divided by sqrt(10) per level, and one
 * full-resolution stage at 1e-3; coarse_tol 0.1, handoff 2. */
void cot_default_ladder_opts(const cot_problem* problems, int n_problems, cot_ladder_opts* lo);
int cot_solve_ladder(cot_context* ctx, const cot_problem* problems, double* w,
                     const cot_solve_opts* opt, const cot_ladder_opts* lo, cot_report* reports);

/* Diagnostic image (render, basic_function.cuh:660-741): pixel p of each problem's target gets
 * sum_i nu_i |C_i n p| / |C_i| over the cells at weights w; image holds the problems' targets one
 * after the other in the reference layout (ix + iy * pm), sum of pm * pn doubles.  The caller
 * multiplies by its grey scale (gray_multiple, ot_op.cuh:215) if it wants picture units. */
int cot_render(cot_context* ctx, const double* w, double* image);

/* The cells at weights w (the polygons writeOBJ writes, basic_function.cuh:15-59): counts[i]
 * vertices (0 for an empty cell, at most COT_MAX_VERTS), CCW, in verts[(i * COT_MAX_VERTS + k) *
 * 2 + {0, 1}] (caller order).  The cells are the restricted cells inside proj_domain. */
#define COT_MAX_VERTS 48
int cot_cell_polygons(cot_context* ctx, const double* w, int* counts, double* verts);

/* y = H x, H the Hessian of the dual objective at w (d mass_i / d w_j, the matrix the Newton
 * steps solve with); all arrays in caller order.  No counterpart in the reference. */
int cot_hessian_product(cot_context* ctx, const double* w, const double* x, double* y);

/* Centroids of the cells at weights w: weighted != 0 -> rho-weighted (geometric fallback),
 * weighted == 0 -> geometric centroid of the domain-clipped cell; exist per site. */
int cot_centroids(cot_context* ctx, const double* w, int weighted, double* centroid,
                  unsigned char* exist);

/* The same three operations on caller-owned DEVICE arrays (caller order; f: n_problems
 * doubles), ordered on the caller's stream (a hipStream_t; NULL: the default stream): the work
 * starts after what is queued there and what is queued there afterwards runs after it.  The
 * library's own stream does the work; reports stay host structs.  Replace direct_ot's device
 * buffers and copies (ot_op.cuh:66-70, 102-152, 539-588) for a caller whose data already lives
 * on the GPU (a torch pipeline). */
int cot_evaluate_dev(cot_context* ctx, const double* w, double* mass, double* f, double* centroid,
                     unsigned char* exist, void* stream);
int cot_solve_dev(cot_context* ctx, double* w, const cot_solve_opts* opt, cot_report* reports,
                  void* stream);
int cot_centroids_dev(cot_context* ctx, const double* w, int weighted, double* centroid,
                      unsigned char* exist, void* stream);

/* Area of each cell (inside proj_domain) at weights w. */
int cot_cell_areas(cot_context* ctx, const double* w, double* area);

/* Irradiance image of a triangle mesh: the forward pass of the irra_render extension's render()
 * (render_exe_program/render.cu:553-591, with compute_render, :93-171).  Face f (vertex indices
 * faces[3f..3f+2] into the n_points points xy, (x, y) pairs) spreads tri_energy[f] * total_gray
 * uniformly over its triangle; pixel (ix, iy) of the pm x pn grid over proj_domain
 * {x0, x1, y0, y1} receives e_f |T_f n pixel| / |T_f| (degenerate triangles add nothing).
 * image gets pm x pn floats in render()'s output layout: row pm - 1 - ix, column iy.  All
 * pointers are device memory; the work is ordered on stream (a hipStream_t; NULL: the default
 * stream) and the call waits for it only to check that every face index names a point
 * (COT_EINVAL otherwise).  No context is needed. */
int cot_irra_render(const float* tri_energy, const int* faces, int n_faces, const float* xy,
                    int n_points, float total_gray, const double* proj_domain, int pm, int pn,
                    float* image, void* stream);

/* The gradient pass of the irra_render extension, render_grad (render_exe_program/
 * render.cu:593-728 with compute_render_diff, computer_grad_inte, compute_grad_render and,
 * for use_sobel, compute_sobel_weight and compute_grad_render_sobel, :53-91, 422-542, 173-420).
 * With the mesh as in cot_irra_render and
 * real_picture the target in the grid layout the extension's kernels read (pixel (ix, iy) at
 * iy * pm + ix):
 *   render, render_diff   pm x pn images in render()'s output layout: the mesh's image and
 *                         image - target;
 *   xy_render_grad (2 n_points), tri_render_grad (n_faces): d/dxy and d/dtri_energy of
 *                         sum_p (render_p - real_p)^2;
 *   inte_result, inte_diff (n_faces): inte_f = sum_p orient_f |T_f n p| real_p / |p| and
 *                         inte_f - tri_energy_f * total_gray (orient_f = -1 for a clockwise
 *                         face: the reference's signed clipped areas, compute_grad_area,
 *                         render_exe_program/clip.h:277-292);
 *   xy_inte_grad (2 n_points): d/dxy of sum_f inte_diff_f^2;
 *   with use_sobel, xy_sobel_grad (2 n_points) and tri_sobel_grad (n_faces): d/dxy and
 *                         d/dtri_energy of the sum over 4-neighbouring pixel pairs of
 *                         (diff_p - diff_q)^2, diff = render - real (unused otherwise).
 * The vertex derivatives are exact where the clipped areas are differentiable (each is the
 * boundary transport of the triangle's edges through the pixel; pixels are half-open, so an
 * edge lying on a grid line counts once, with the derivative for moving it towards +x / +y).  Device pointers; work
 * ordered on stream; COT_EINVAL for a face index that names no point.  A Sobel call holds a
 * pm x pn float image of device memory for its duration.  The reference's Sobel weights read
 * past the image when pm > pn (row bound pm - 1, :77); here a neighbour outside the image is
 * absent, which is the reference's result whenever pm == pn. */
int cot_irra_render_grad(const float* tri_energy, const int* faces, int n_faces, const float* xy,
                         int n_points, float total_gray, const float* real_picture,
                         const double* proj_domain, int pm, int pn, float* render,
                         float* render_diff, float* xy_render_grad, float* tri_render_grad,
                         float* inte_diff, float* xy_inte_grad, float* inte_result,
                         int use_sobel, float* xy_sobel_grad, float* tri_sobel_grad,
                         void* stream);

/* cot_irra_render_grad with no device allocation: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, outputs and errors as cot_irra_render_grad;
 * the Sobel weight image lives in a fixed device array of the library (each GPU holds
 * COT_NOALLOC_MAX_PIXELS floats plus 256 bytes from the library's load to the end of the
 * process), so with use_sobel pm * pn may not exceed COT_NOALLOC_MAX_PIXELS (COT_EINVAL
 * otherwise).  The face indices are checked on the device and the count read back (the call
 * waits for the stream once, as cot_irra_render_grad does).  Calls on one GPU share that
 * array: calls that overlap on one GPU (other host threads, other streams) race. */
#define COT_NOALLOC_MAX_PIXELS (1 << 20)
int cot_irra_render_grad_noalloc(const float* tri_energy, const int* faces, int n_faces,
                                 const float* xy, int n_points, float total_gray,
                                 const float* real_picture, const double* proj_domain, int pm,
                                 int pn, float* render, float* render_diff, float* xy_render_grad,
                                 float* tri_render_grad, float* inte_diff, float* xy_inte_grad,
                                 float* inte_result, int use_sobel, float* xy_sobel_grad,
                                 float* tri_sobel_grad, void* stream);

/* Gaussian splatting of n ray hits (points, (x, y) pairs inside proj_domain, finite) onto a
 * res_y x res_x image (pixel (x, y) at y * res_x + x): _gaussian_splat_raw_chunk,
 * utils/basic_function.py:674-709.  Each point spreads energy over the (2 radius + 1)^2 pixels
 * around the one it falls in, weights exp(-d^2 / (2 sigma^2)) (d in pixels from the pixel
 * centre) normalised over the whole window, pixels outside the image dropped.
 * image_grad NULL: out (the image) += the splat.  image_grad given: out (2n) = d/d(points) of
 * sum_p scale * splat_p * image_grad_p with the normalisation held constant (the second pass of
 * render_gaussian_splat_two_pass_backward, :711-772).  Device pointers; work ordered on stream;
 * radius at most 16. */
int cot_splat(const float* points, const float* energy, int n, int res_y, int res_x,
              const double* proj_domain, float sigma, int radius, const float* image_grad,
              float scale, float* out, void* stream);

/* Solid angle under which a point light at light[0..2] sees each cell (inside proj_domain, at
 * weights w) placed in the plane z (solid_angle, basic_function.cuh:635-651, with
 * compute_solid_angle, :558-588): the cell fanned from its first vertex into spherical
 * triangles, each 2 atan(a.(b x c) / (1 + a.b + b.c + c.a)) for the unit vectors towards its
 * corners, triangles with |denominator| < 1e-6 skipped, the magnitude of the sum returned; 0
 * for an empty cell.  For a point light the reference's source masses are
 * nu_i = omega_i / sum_j omega_j * area_scalar (ot_op.cuh:160-206). */
int cot_cell_solid_angles(cot_context* ctx, const double* w, const double* light, double z,
                          double* omega);

/* Kernel time accumulated since the last reset.  Each launch of the cell-construction kernel,
 * of the integration kernel and of their repair passes (cells that failed the conformity check)
 * is timed inside the kernel with the device's constant-rate wall clock, from the earliest
 * start to the latest end of a block that did work; launches in which no block had work (all
 * problems finished) are not counted.  The conformity check is not timed (verify_* stay 0). */
typedef struct {
  double cell_seconds;        /* full passes of the construction kernel */
  long long cell_launches;
  double repair_seconds;      /* repair passes (construction + integration) */
  long long repair_launches;
  double verify_seconds;      /* conformity-check launches */
  long long verify_launches;
  long long evaluations;      /* evaluations (line-search trials included) */
  long long repair_rounds;    /* repair rounds run by host-driven evaluations */
  double integ_seconds;       /* full passes of the integration kernel */
  long long integ_launches;
  long long newton_steps;     /* cot_solve_newton: Hessians built (one per Newton step) */
  long long cg_iterations;    /* cot_solve_newton: conjugate-gradient iterations */
  double coarse_seconds;      /* cot_solve_newton: host time of the coarse-matrix inverses */
  long long cg_restarts;      /* cot_solve_newton: steps whose CG diverged with the full
                                 preconditioner and was redone with Jacobi alone */
  long long handoff_degenerate;    /* FP32 full passes: cells handed to the FP64 repair for a
                                      degeneracy FP32 cannot resolve (sliver, marginal
                                      emptying, a vertex the refinement could not place) */
  long long handoff_overflow;      /* full passes: cells handed to the repair for a polygon
                                      above the full pass's vertex limit */
  long long check_failures;        /* sites failing the conformity check after a full pass */
  long long repair_check_failures; /* sites failing it after a repair pass */
  long long repaired_cells;        /* cells rebuilt by repair passes */
  long long reappeared_cells;      /* cells left empty by an evaluation that the next one's
                                      pre-pass found non-empty again (queued as candidates of
                                      their new neighbours before the full pass) */
} cot_stats;

int cot_get_stats(cot_context* ctx, cot_stats* out);
int cot_reset_stats(cot_context* ctx);

#ifdef __cplusplus
}
#endif
#endif /* CAUSTIC_OT_H_ */
