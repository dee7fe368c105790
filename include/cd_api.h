/*
 * cd_api.h -- C ABI of the MI355X (gfx950) caustic-design hot path.
 *
 * Plain pointers and sizes only.  Host-pointer entry points copy in/out;
 * *_device entry points take device pointers (HIP global memory) and leave
 * everything resident.  All work is ordered on the context's HIP stream.
 *
 * Each entry point names the reference interface of
 * dylanmsu/poisson_caustic_design it replaces (file:line).
 *
 * Layouts (same as the reference, flattened):
 *   grids     row-major [y][x]: height rows of width doubles
 *             (std::vector<std::vector<double>> grid[y][x], solver.h:8)
 *   vertices  index = y*mesh_res_x + x (src/mesh.cpp:48-54)
 *   points    3 doubles {x, y, z} per vertex (point_t, src/mesh.h:17)
 *
 * Status codes replace the reference's exit(0) on a raster miss
 * (src/mesh.cpp:277-279, 334-336) and its silent iteration cap
 * (src/solver.cpp:92).
 */
#ifndef CD_API_H
#define CD_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CD_API_VERSION 2

typedef enum {
    CD_OK = 0,
    CD_ERR_RASTER_MISS = 1,     /* a raster sample / source vertex lies in no triangle */
    CD_ERR_FOLDED = 2,          /* a transport step inverted mesh triangles
                                   (only with cd_design_set_fold_check(d, 1))          */
    CD_ERR_NOT_CONVERGED = 3,   /* sweep cap reached before the stop criterion          */
    CD_ERR_NAN = 4,             /* NaN/Inf in a result                                  */
    CD_ERR_BAD_ARG = 5,
    CD_ERR_HIP = 6,             /* HIP runtime error (message via cd_context_last_error) */
    CD_ERR_NO_DEVICE = 7,
    CD_ERR_OOM = 8,
    CD_ERR_STATE = 9            /* call out of order (e.g. step before init)            */
} cd_status;

typedef enum {
    CD_PREC_MIXED = 0,  /* FP32 red-black sweeps on a correction, FP64 base, FP64
                           residual / zero-mean projection / defect correction */
    CD_PREC_FP64 = 1    /* FP64 red-black sweeps                                  */
} cd_precision;

typedef enum {
    CD_STOP_RESIDUAL = 0,   /* ||P b - A phi||_2 / ||P b||_2 <= tol (FP64)             */
    CD_STOP_MAX_UPDATE = 1  /* max |delta| over one sweep < tol  (solver.cpp:142)       */
} cd_stop_rule;

typedef struct {
    int max_sweeps;       /* sweep cap (reference max_iterations, solver.cpp:92)        */
    int stop;             /* cd_stop_rule                                               */
    double tol;           /* tolerance of the stop rule                                 */
    int precision;        /* cd_precision                                               */
    int check_every;      /* sweeps between FP64 residual checks / flushes; 0 = auto    */
    int learned_init;     /* 1: learned initial correction (needs cd_context_set_net)   */
    double omega;         /* > 0: as given; < 0: the reference's 2/(1+3.14159265/width)
                             (solver.cpp:71); 0: the optimum for this operator,
                             2/(1+sqrt(1-rho^2)), rho = 1-sin^2(pi/(2 max(w,h)))     */
} cd_solve_opts;

typedef struct {
    int status;
    int sweeps;                 /* device-counted red+black sweeps                   */
    int checks;                 /* FP64 residual evaluations                         */
    int learned_applied;        /* learned corrections accepted                      */
    int learned_rejected;       /* learned corrections rejected (residual rose)      */
    double rel_residual;        /* final ||P b - A phi||_2/||P b||_2 (FP64)          */
    double initial_rel_residual;
    double last_max_update;     /* max |delta| of the last sweep                     */
    double elapsed_ms;
} cd_solve_stats;

typedef struct cd_context cd_context;
typedef struct cd_design cd_design;

/* ---- context -------------------------------------------------------- */
void cd_solve_opts_default(cd_solve_opts* o);  /* residual 1e-6, mixed, cap 100000, omega 0 */
int cd_context_create(int device, cd_context** out);
void cd_context_destroy(cd_context* ctx);
int cd_context_synchronize(cd_context* ctx);
const char* cd_context_last_error(const cd_context* ctx);
const char* cd_status_string(int status);
int cd_get_api_version(void);
/* Kernel timing over HIP events on the context stream (off by default).
 * kernel ids: cd_kernel_id */
typedef enum {
    CD_KERNEL_SOR_SWEEP = 0,    /* the red-black SOR sweep kernel(s)                 */
    CD_KERNEL_RESIDUAL = 1,     /* FP64 residual + correction flush                  */
    CD_KERNEL_RASTER = 2,       /* mesh rasterization                                */
    CD_KERNEL_VERTEX = 3,       /* per-vertex kernels                                */
    CD_KERNEL_LEARNED = 4,      /* learned correction (restriction, network, prolongation) */
    CD_KERNEL_COUNT = 5
} cd_kernel_id;
int cd_context_set_profiling(cd_context* ctx, int enable);
/* total elapsed ms, number of launches, algorithmic bytes moved by those launches */
int cd_context_kernel_stats(cd_context* ctx, int kernel, double* total_ms, long long* launches,
                            double* bytes);
int cd_context_reset_kernel_stats(cd_context* ctx);

/* ---- Poisson solve ---------------------------------------------------
 * Replaces  void poisson_solver(std::vector<std::vector<double>>& D,
 *           std::vector<std::vector<double>>& phi, int width, int height,
 *           int max_iterations, double convergence_threshold, int max_threads)
 * (src/solver.h:8, src/solver.cpp:70-147).  rhs is D ([height][width]); NaN
 * cells are holes excluded from their neighbours' stencils (solver.cpp:29-44)
 * and come back NaN in phi.  phi is the warm start and the result.  The
 * right-hand side is projected to zero mean over non-hole cells (the Neumann
 * compatibility condition) before sweeping. */
int cd_poisson_solve(cd_context* ctx, const double* rhs, double* phi, int width, int height,
                     const cd_solve_opts* opts, cd_solve_stats* stats);
int cd_poisson_solve_device(cd_context* ctx, const double* d_rhs, double* d_phi, int width,
                            int height, const cd_solve_opts* opts, cd_solve_stats* stats);

/* Allocation-free variant of cd_poisson_solve_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * solve (tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels, results and statuses as cd_poisson_solve_device; its
 * scratch (reduction partials, solve state, the FP32 corrections, the
 * right-hand side of the sweeps, the hole mask) is one static device array
 * per GPU of CD_NOALLOC_SCRATCH_MIB MiB, reserved when the library's code
 * object is loaded on that GPU and held until the process ends, which covers
 * every grid up to CD_NOALLOC_MAX_SIDE x CD_NOALLOC_MAX_SIDE; a larger grid
 * fails with CD_ERR_BAD_ARG before any work.  learned_init must be 0.  All
 * calls on one GPU share that array (it keeps the last call's contents; each
 * call initialises what it reads): calls that overlap on one GPU, from any
 * context or thread, race and must be serialised by the caller.  Kernel
 * timers (cd_context_set_profiling) do not cover it. */
#define CD_NOALLOC_MAX_SIDE 2048
#define CD_NOALLOC_SCRATCH_MIB 53
int cd_poisson_solve_device_noalloc(cd_context* ctx, const double* d_rhs, double* d_phi,
                                    int width, int height, const cd_solve_opts* opts,
                                    cd_solve_stats* stats);

/* ---- design (class Caustic_design, src/caustic_design.h:7-66) ----------- */
typedef struct {
    int mesh_res_x, mesh_res_y;   /* set_mesh_resolution   (caustic_design.cpp:28-31) */
    int res_x, res_y;             /* set_domain_resolution (caustic_design.cpp:33-36) */
    double width, height;         /* set_mesh_size         (caustic_design.cpp:38-41) */
    double focal_l;               /* set_lens_focal_length (caustic_design.cpp:43-45) */
    double thickness;             /* set_lens_thickness    (caustic_design.cpp:47-49) */
    double ior;                   /* refractive index, 1.49 (caustic_design.cpp:271)  */
    double step_size;             /* OT step, (double)0.05f (caustic_design.cpp:250)  */
    cd_solve_opts phi_solve;      /* transport potential solve (caustic_design.cpp:222) */
    cd_solve_opts h_solve;        /* height solve               (caustic_design.cpp:311) */
} cd_design_params;

/* Defaults reproduce the reference: ior 1.49, step (double)0.05f, phi solve
 * max-update 1e-7, h solve max-update 1e-8, 100000-sweep caps, the
 * reference's omega (solve opts omega = -1). */
void cd_design_params_default(cd_design_params* p);

int cd_design_create(cd_context* ctx, const cd_design_params* p, cd_design** out);
void cd_design_destroy(cd_design* d);
int cd_design_set_solve_opts(cd_design* d, const cd_solve_opts* phi, const cd_solve_opts* h);
/* initialize_solvers(image) (caustic_design.cpp:334-364): image is res_y x res_x */
int cd_design_init(cd_design* d, const double* image);
int cd_design_init_device(cd_design* d, const double* d_image);
/* perform_transport_iteration() (caustic_design.cpp:190-266); *step_out is
 * max vertex displacement / width */
int cd_design_transport_step(cd_design* d, double* step_out);
/* perform_height_map_iteration(itr) (caustic_design.cpp:269-332); *update_out
 * is the sum of squared vertex-height updates (mesh.cpp:724-742) */
int cd_design_height_step(cd_design* d, double* update_out);
/* Mesh::laplacian_smoothing(target_points, factor) (mesh.cpp:534-606). */
int cd_design_laplacian_smoothing(cd_design* d, double factor);
/* Folded meshes.  The reference keeps stepping a mesh whose triangles have
 * inverted (mesh.cpp:485-532 does not clamp interior vertices) and reports
 * nothing.  Every cd_design_transport_step counts the triangles of the
 * stepped target mesh whose orientation (b-a)x(c-a) is <= 0 (all are > 0 in
 * the initial mesh, mesh.cpp:57-63); cd_run_stats carries the counts.
 * fail != 0: a step that leaves inverted triangles returns CD_ERR_FOLDED
 * (the step is applied); default 0 keeps the reference's behaviour. */
int cd_design_set_fold_check(cd_design* d, int fail);
/* Inverted triangles of the current target mesh (e.g. after cd_design_set). */
int cd_design_count_folds(cd_design* d, long long* folded_triangles);
/* on != 0: a raster miss in cd_design_transport_step smooths the target mesh
 * with factor 0.1f and returns CD_OK with a NaN step, the recovery of
 * caustic_design.cpp:215-218 (in the reference an exit(0) at mesh.cpp:279
 * precedes it).  Default off: the step fails with CD_ERR_RASTER_MISS. */
int cd_design_set_miss_recovery(cd_design* d, int on);
/* Which triangle a raster sample or grid vertex takes where several triangles
 * hold it (overlaps of a folded mesh; shared edges).  mode 0 (default): the
 * lowest triangle index.  mode 1: the triangle the reference's BVH query
 * returns (Bvh::query, bvh.cpp:196-266, over the tree of Bvh::build,
 * bvh.cpp:20-155, built with (5, 30) at mesh.cpp:235,292,349), i.e. the
 * holding triangle of lowest position in cd_bvh_first_hit_order's array.
 * Mode 1 builds that order on the host at every rasterization of the target
 * mesh and when the source owners are built (at cd_design_init, or here when
 * the design is initialised): a copy of the mesh to the host each time.
 * Returns CD_ERR_BAD_ARG for other modes. */
int cd_design_set_first_hit(cd_design* d, int mode);
/* order[k] = the triangle at position k of the array that the reference's
 * Bvh::build (bvh.cpp:20-155; leaves of at most 5 triangles, depth 30)
 * partitions, for the mx x my mesh with vertex coordinates xy (x,y pairs,
 * triangulation of mesh.cpp:57-63, 2(mx-1)(my-1) triangles).  Host code. */
int cd_bvh_first_hit_order(const double* xy, int mx, int my, int* order);

typedef struct {
    int status;
    int transport_steps;          /* OT steps performed                              */
    int height_steps;
    double last_step;             /* last transport step size                        */
    long long phi_sweeps, h_sweeps;
    long long checks;
    int learned_applied, learned_rejected;
    double max_phi_residual, max_h_residual;
    double elapsed_ms;
    int recoveries;               /* raster misses recovered by smoothing            */
    long long folded_triangles;   /* inverted triangles after the last OT step       */
    int folded_steps;             /* OT steps that ended with a folded mesh          */
    int first_folded_step;        /* index of the first of them (-1: none)           */
    int unconverged_solves;       /* Poisson solves that reached their sweep cap; the
                                     design goes on with that solution, as the
                                     reference does after max_iterations
                                     (solver.cpp:92), and the run still returns CD_OK */
} cd_run_stats;

/* The reference driver loop (main.cpp:243-262): up to max_transport_steps
 * OT steps, stopping after the first step whose size is < conv_tres, then
 * height_steps height steps.  Everything stays on the device. */
int cd_design_run(cd_design* d, int max_transport_steps, double conv_tres, int height_steps,
                  cd_run_stats* stats);

/* Batched designs on concurrent streams.  For each i < n: re-initialise
 * designs[i] from the device image d_images[i] (cd_design_init_device;
 * skipped when d_images or d_images[i] is NULL), then cd_design_run(designs[i],
 * max_transport_steps, conv_tres, height_steps, &stats[i]).  The n designs run
 * concurrently, one host thread each, every one on its own context's stream;
 * each design must belong to a different context (a context's stream and
 * workspace serve one design at a time), else CD_ERR_BAD_ARG before any work.
 * stats[i].status is each design's status; the return value is CD_OK or the
 * status of the lowest-index design that failed.  The reference runs one
 * design per process (main.cpp:243-262); this is that loop over a batch of
 * independent targets.  n <= CD_BATCH_MAX. */
#define CD_BATCH_MAX 64
int cd_designs_run_batch(cd_design* const* designs, const double* const* d_images, int n,
                         int max_transport_steps, double conv_tres, int height_steps,
                         cd_run_stats* stats);

typedef enum {
    CD_FIELD_TARGET_POINTS = 0,   /* V x 3  mesh->target_points            */
    CD_FIELD_SOURCE_POINTS = 1,   /* V x 3  mesh->source_points            */
    CD_FIELD_PHI = 2,             /* N     phi                             */
    CD_FIELD_H = 3,               /* N     h                               */
    CD_FIELD_RASTER = 4,          /* N     raster (after zero-mean)        */
    CD_FIELD_ERRORS = 5,          /* V     errors                          */
    CD_FIELD_TARGET_AREAS = 6,    /* V     target_areas                    */
    CD_FIELD_DIVERGENCE = 7,      /* N     divergence (after zero-mean)    */
    CD_FIELD_NORM_X = 8,          /* N     norm_x                          */
    CD_FIELD_NORM_Y = 9,          /* N     norm_y                          */
    CD_FIELD_NORMALS = 10,        /* V x 2 normals[0], normals[1]          */
    CD_FIELD_VERTEX_GRADIENT = 11,/* V x 2 vertex_gradient                 */
    CD_FIELD_PIXELS = 12,         /* N     pixels (min-max scaled image)   */
    CD_FIELD_INVERSE_MAP = 13,    /* V x 2 inverted transport map          */
    CD_FIELD_IMAGE = 14,          /* N     input image as last initialised */
    CD_FIELD_GRADIENT = 15,       /* 2 x N gradient[0], gradient[1] of phi
                                     (calculate_gradient, utils.cpp:3-20;
                                     caustic_design.cpp:225), read only    */
    CD_FIELD_COUNT = 16
} cd_field;

/* number of doubles of a field (N = res_x*res_y grid, V = vertices) */
size_t cd_design_field_size(const cd_design* d, int field);
int cd_design_get(cd_design* d, int field, double* host_out, size_t count);
int cd_design_set(cd_design* d, int field, const double* host_in, size_t count);
/* device pointer of a field's storage (NULL for composite fields) */
const void* cd_design_device_ptr(cd_design* d, int field);
int cd_design_last_solve_stats(const cd_design* d, cd_solve_stats* phi, cd_solve_stats* h);
/* Partitioned dual cells of the target mesh, the reference's target_cells
 * (initial == 0; rebuilt each transport step, caustic_design.cpp:195-196) or
 * first_target_cells (initial != 0; the initial mesh, caustic_design.cpp:
 * 343-346), built by Mesh::build_target_partitioned_dual_cells
 * (mesh.cpp:121-231).  For vertex v and slot k < 6: triangles[6v+k] is its
 * k-th adjacent triangle in ascending index (-1 past the last) and
 * quads[(6v+k)*12 ...] that triangle's quad as 4 points {x, y, z}.  (The
 * reference lists a vertex's quads in the iteration order of a
 * std::unordered_set of these indices; the C++ drop-in reproduces it.)
 * vertices >= V; triangles holds 6V ints, quads 72V doubles. */
int cd_design_get_dual_cells(cd_design* d, int initial, int* triangles, double* quads,
                             size_t vertices);

/* ---- input and output files (main.cpp, src/utils.cpp) -------------------- */
/* Parameters a design was created with. */
int cd_design_get_params(const cd_design* d, cd_design_params* p);

/* PNG -> RGBA8 rows, top row first, with the conversions main.cpp:56-83 asks
 * libpng for (16-bit stripped, palette and low-bit gray expanded, tRNS to
 * alpha, filler 0xFF).  Gray and gray-alpha images give (g, g, g, a).  With
 * rgba == NULL only *width, *height are set; otherwise capacity must be at
 * least 4*width*height.  Replaces the decoding half of image_to_grid
 * (main.cpp:29-109). */
int cd_png_read_rgba8(const char* path, int* width, int* height, unsigned char* rgba,
                      size_t capacity);
/* Gray conversion 0.299 r + 0.587 g + 0.114 b of bytes/255 (main.cpp:95-98)
 * at the pixels a nearest resize to out_w x out_h samples (resize_image,
 * main.cpp:13-27), row-major FP64 out.  Host and device-pointer variants. */
int cd_image_gray_resize(cd_context* ctx, const unsigned char* rgba, int width, int height,
                         int out_w, int out_h, double* out);
int cd_image_gray_resize_device(cd_context* ctx, const unsigned char* d_rgba, int width,
                                int height, int out_w, int out_h, double* d_out);

/* Solidified lens OBJ (save_solid_obj, src/utils.cpp:198-260): front points
 * V x 3 (x, y, height), back points V x 3 (x, y used), res_x x res_y mesh. */
int cd_obj_write_solid(const char* path, const double* front, const double* back, int res_x,
                       int res_y, double thickness, double width, double height);
/* heightmap.json of a row-major res_y x res_x grid (main.cpp:111-135). */
int cd_json_write_heightmap(const char* path, const double* h, int res_x, int res_y);
/* SVG of a point grid (export_grid_to_svg, src/utils.cpp:262-307) or of the
 * structured triangles (export_triangles_to_svg, src/utils.cpp:309-344);
 * points have `stride` doubles each, x and y first. */
int cd_svg_write_grid(const char* path, const double* points, int stride, int res_x, int res_y,
                      double width, double height, double stroke_width);
int cd_svg_write_triangles(const char* path, const double* points, int stride, int res_x,
                           int res_y, double width, double height, double stroke_width);

enum { CD_OBJ_SOURCE = 0, CD_OBJ_TARGET = 1 };
/* save_solid_obj_source / save_solid_obj_target (src/caustic_design.cpp:55-61). */
int cd_design_save_solid_obj(cd_design* d, const char* path, int front);
/* save_heightmap_as_json(caustic_design.h, ...) (main.cpp:269). */
int cd_design_save_heightmap_json(cd_design* d, const char* path);
enum { CD_SVG_PARAMETERIZATION = 0, CD_SVG_INVERTED_MAP = 1, CD_SVG_TRIANGLES = 2 };
/* export_paramererization_to_svg, export_inverted_transport_map, export_to_svg
 * (src/caustic_design.cpp:20-26, src/mesh.cpp:101-103,411-414,608-610). */
int cd_design_export_svg(cd_design* d, const char* path, int what, double stroke_width);

/* ---- learned coarse solve (no reference counterpart) ----------------------
 * A bias-free convolutional encoder-decoder (U-Net: `levels` resolution
 * levels, two 3x3 convolutions per level and per decoder stage, 2x2 average
 * pooling, nearest upsampling, concatenated skips; linear, or ReLU on the
 * hidden convolutions) approximates the inverse of the Neumann 5-point
 * operator on a grid `coarsen` times coarser than the solve: its input is a
 * coarse residual normalised to unit RMS, its output the coarse solution
 * divided by (n/pi)^2 (n = larger coarse side).  BF16 weights and
 * activations, FP32 accumulation (gfx950 MFMA), FP32 master weights and Adam.
 *
 * Solves with learned_init = 1 (FP32-mixed precision, no holes) start with
 * flexible conjugate-gradient iterations preconditioned by one cycle:
 * restrict the residual (block mean), apply the network, prolong (bilinear),
 * with mid_sweeps > 0 through an intermediate grid smoothed by that many
 * red-black sweeps, and run `smooth_sweeps` red-black SOR sweeps.  The first cycle
 * is the learned initial correction.  Every iteration recomputes the FP64
 * residual; when eight iterations bring no halving of it the solve continues
 * with plain SOR from the current iterate, so the stop rule and tolerance are
 * the same as without the network. */
typedef struct cd_net cd_net;
typedef struct {
    int levels;        /* resolution levels (>= 2)                          */
    int channels[6];   /* channels per level, multiples of 8 (e.g. 16 32 64 64) */
    int coarsen;       /* fine cells per coarse cell along each axis (4)     */
    int smooth_sweeps; /* red-black sweeps between learned corrections (4)   */
    double smooth_omega; /* over-relaxation of those sweeps (1.2)            */
    int linear;        /* 1: no activation (a linear operator, default)      */
    int out_smooth;    /* smoothing steps y += Lc y / 8 on the output (0)    */
    int loss_smooth;   /* the same steps on the operator-loss residual (0):  */
                       /* weights it towards the low frequencies             */
    int loss_energy;   /* 1: train on the energy J(y) = y'(-cs Lc)y/2 + x'y  */
                       /* (same minimiser cs Lc y = x; errors weighted by Lc  */
                       /* once instead of twice, so smooth error counts more; */
                       /* loss_smooth is then unused) (0)                    */
    int mid_sweeps;    /* > 0 with coarsen a multiple of 4, >= 8: the coarse  */
                       /* correction is prolonged to an intermediate grid     */
                       /* (coarsen/4 fine cells per cell), smoothed there by  */
                       /* this many red-black sweeps on the restricted        */
                       /* residual, then prolonged to the fine grid (0)       */
} cd_net_arch;

void cd_net_arch_default(cd_net_arch* a);
int cd_net_create(cd_context* ctx, const cd_net_arch* arch, cd_net** out);
void cd_net_destroy(cd_net* net);
size_t cd_net_num_params(const cd_net* net);
int cd_net_init_params(cd_net* net, unsigned long long seed);      /* He-normal     */
int cd_net_set_params(cd_net* net, const float* host, size_t n);  /* FP32 master   */
int cd_net_get_params(cd_net* net, float* host, size_t n);
/* Forward pass on a device batch: x, y are [batch][h][w] FP32 (normalised
 * coarse residual in, normalised coarse correction out). */
int cd_net_forward(cd_net* net, const float* d_x, float* d_y, int batch, int h, int w);
/* Inference forwards (cd_net_forward and the solver's coarse solve) read the
 * upsampled U-Net decoder inputs through the convolutions that consume them
 * (on != 0, the default) or materialise them first, as training does (0).
 * Both give bit-identical outputs; the switch exists to check that.  No
 * counterpart in the reference (the network is new). */
int cd_net_set_fused_forward(cd_net* net, int on);
/* Forward convolutions (training and inference) on 8x8 output tiles whose
 * input and weights are staged in LDS (on != 0, the default) or with
 * per-chunk global gathers (0).  Bit-identical outputs; the switch exists to
 * check that.  No counterpart in the reference. */
int cd_net_set_tiled_forward(cd_net* net, int on);
/* Mean squared error of the forward pass against d_t and its gradient with
 * respect to every parameter (FP32, host order of cd_net_get_params). */
int cd_net_loss_grad(cd_net* net, const float* d_x, const float* d_t, int batch, int h, int w,
                     double* loss, float* host_grad, size_t n);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on a device batch; *loss is the
 * mean squared error before the step. */
int cd_net_train_step(cd_net* net, const float* d_x, const float* d_t, int batch, int h, int w,
                      float lr, double* loss);
/* Operator loss of the coarse solve: with y the forward output, Lc the coarse
 * Neumann 5-point operator and cs = (n/pi)^2, the mean of (cs*Lc y - x)^2 and
 * its parameter gradient.  Needs no targets; it measures how far Lc times the
 * network is from the identity, the quantity that sets the convergence of the
 * preconditioned iterations. */
int cd_net_loss_grad_operator(cd_net* net, const float* d_x, int batch, int h, int w,
                              double* loss, float* host_grad, size_t n);
/* One Adam step on the operator loss; *loss is the loss before the step. */
int cd_net_train_step_operator(cd_net* net, const float* d_x, int batch, int h, int w, float lr,
                               double* loss);
/* Learned FCG iterations replay their kernel sequence (restriction, network,
 * prolongation, smoothing sweeps, FCG update) as one captured HIP graph
 * (default on; off: the same kernels launched one by one, identical results). */
int cd_context_set_graphs(cd_context* ctx, int on);
/* Attach (or detach with NULL) the network used by solves with learned_init. */
int cd_context_set_net(cd_context* ctx, cd_net* net);

/* Training pairs captured during solves: at every convergence check of a
 * solve of a (coarsen*h) x (coarsen*w) grid the normalised coarse residual is
 * appended to d_x and, when the solve ends, the normalised coarse difference
 * between the final and the then-current solution to d_t ([capacity][h][w]
 * FP32 each).  Pairs are taken every `every`-th check; capture stops at
 * capacity.  d_x == NULL disables capture. */
int cd_context_set_capture(cd_context* ctx, float* d_x, float* d_t, int capacity, int h, int w,
                           int coarsen, int every);
int cd_context_capture_count(const cd_context* ctx, int* count);

#ifdef __cplusplus
}
#endif
#endif /* CD_API_H */
