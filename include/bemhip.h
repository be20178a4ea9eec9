/*
 * bemhip.h -- C ABI of the MI355X full-3D boundary-element library.
 *
 * This is the drop-in boundary for the reference's full-3D BEM solve
 * (brandynlucca/AcousticScattering.jl v0.1.2).  Each entry point names the
 * reference function it replaces (paths relative to the reference repo).
 * Plain C types only: pointers, sizes, doubles.  Complex arrays are
 * interleaved (re, im) doubles, which is Julia's ComplexF64 layout.  Matrices
 * passed in are column-major (Julia order); triangles are 1-based.
 *
 * Every function returns an int status (BEM_OK on success).  The message of
 * the last failure on a context is returned by bem_last_error().  No C++
 * exception crosses this boundary.  The library never keeps a caller's host
 * pointer after a call returns; device state lives behind opaque handles.
 */
#ifndef BEMHIP_H
#define BEMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped by the host shim to the reference's errors) ---- */
#define BEM_OK 0
#define BEM_ERR_ARGUMENT 1  /* ArgumentError (e.g. src/solvers/bem/full.jl:71,297-298) */
#define BEM_ERR_DIMENSION 2 /* DimensionMismatch                                       */
#define BEM_ERR_DEVICE 3    /* HIP failure or library unusable on this device            */
#define BEM_ERR_ADMISSION 4 /* request exceeds the context's HBM budget                  */
#define BEM_ERR_NUMERIC 5   /* double-precision fallback rejected (fluid_system.jl:286)  */
#define BEM_ERR_STATE 6     /* wrong handle or request not finished                      */

/* ---- boundary conditions (src/core/materials.jl:16,23,62-72) ---- */
#define BEM_RIGID 0
#define BEM_PRESSURE_RELEASE 1
#define BEM_FLUID 2

/* ---- formulations (src/solvers/bem/full.jl:52-59,273-280) ---- */
#define BEM_BURTON_MILLER 0 /* rigid / pressure-release default                   */
#define BEM_CBIE 1          /* conventional equation (rigid, soft or fluid)        */
#define BEM_MULLER 2        /* fluid default                                       */

/* ---- precision (src/solvers/bem/full.jl:349-385) ---- */
#define BEM_DOUBLE 0 /* FP64 assembly, FP64 LU                                      */
#define BEM_MIXED 1  /* FP32 regular pairs, complex-FP32 LU, FP64 refinement         */

/* ---- refinement reasons (src/solvers/bem/fluid_system.jl:186-227,282-300) ---- */
#define BEM_REASON_NONE 0
#define BEM_REASON_CONVERSION 1
#define BEM_REASON_SINGLE_SINGULAR 2
#define BEM_REASON_CONDITION 3
#define BEM_REASON_NONFINITE 4
#define BEM_REASON_ITERATION_LIMIT 5
#define BEM_REASON_STAGNATION 6

typedef struct bem_ctx_s *bem_ctx;
typedef struct bem_mesh_s *bem_mesh;
typedef struct bem_operator_s *bem_operator;
typedef struct bem_solution_s *bem_solution;

/* Solver options; mirrors the keyword arguments of solve_full_bem
 * (src/solvers/bem/full.jl:52-59 rigid/soft, :273-280 fluid).  Fill with
 * bem_problem_defaults() and override fields. */
typedef struct {
    int32_t boundary;            /* BEM_RIGID / BEM_PRESSURE_RELEASE / BEM_FLUID        */
    int32_t formulation;         /* BEM_BURTON_MILLER / BEM_CBIE / BEM_MULLER           */
    int32_t precision;           /* BEM_DOUBLE / BEM_MIXED                              */
    int32_t equilibrate;         /* row/column max scaling before LU (full.jl:388-394)  */
    double k;                    /* exterior wavenumber [rad/m], finite and > 0        */
    double density_contrast;     /* g = rho_int / rho_ext (fluid only)                 */
    double soundspeed_contrast;  /* h = c_int / c_ext (fluid only); k_int = k / h      */
    double refine_tolerance;     /* backward-error acceptance, default 1e-13           */
    int32_t refine_maxiter;      /* default 8                                          */
    int32_t force_promotion;     /* test hook: 1 = promote to FP64 before refining     */
    double condition_guard;      /* default 10 eps(Float32); <0 disables the estimate  */
} bem_problem;

/* Per right-hand side report; field names follow the reference's diagnostics
 * (src/solvers/bem/full.jl:170-181,479-491; fluid_system.jl:172-309). */
typedef struct {
    int64_t unknown_count;
    int64_t quadrature_nodes;
    double absolute_residual;
    double relative_residual;
    double scaled_absolute_residual;
    double scaled_relative_residual;
    int32_t factor_precision;    /* 0 double, 1 single                                  */
    int32_t refinement_reason;   /* BEM_REASON_*                                        */
    int32_t refinement_iterations;
    int32_t converged;           /* 1 when the refinement accepted the solution        */
    double normwise_backward_error;
    double componentwise_backward_error;
    double single_reciprocal_condition;
    double assembly_ms;
    double factor_ms;
    double solve_ms;
    int32_t iterations;          /* GMRES iterations (hierarchical path), else 0      */
    int32_t method;              /* 0 direct LU, 1 mixed refinement, 2 GMRES           */
    int32_t derivative_evaluation; /* Mueller K' and H of both media: 1 direct quadrature,
                                      2 Calderon reconstruction (full.jl:201-207,336-339);
                                      0 for other formulations                           */
} bem_diagnostics;

/* Library version string. */
const char *bem_version(void);

/* Fill `p` with the reference defaults for boundary `bc`. */
void bem_problem_defaults(bem_problem *p, int32_t bc);

/* ---- context ---------------------------------------------------------- */
/* Create a context on HIP device `device`.  `hbm_budget_bytes` bounds the
 * device memory the context holds for meshes, operators and factors; 0 means
 * 90% of the device's free memory.  Requests beyond the budget fail with
 * BEM_ERR_ADMISSION. */
int bem_ctx_create(int32_t device, uint64_t hbm_budget_bytes, bem_ctx *out);
int bem_ctx_destroy(bem_ctx ctx);
const char *bem_last_error(bem_ctx ctx);
/* Device bytes currently held by the context. */
int bem_ctx_memory(bem_ctx ctx, uint64_t *used, uint64_t *budget);
/* Kernel timing with HIP events on the library's stream: 1 enables. */
int bem_ctx_set_profiling(bem_ctx ctx, int32_t enable);
/* Accumulated event time and algorithmic work of one kernel class since the
 * last reset.  Classes: "assemble", "near", "lu_panel", "lu_gemm" (FP32
 * trailing updates), "lu_gemm_fp64" (FP64 trailing updates), "lu_trsm",
 * "lu_swap", "solve", "residual", "far_field", "h_apply" (compressed operator
 * apply; bytes = blocks + vectors; its two kernels "h_proj" and "h_rows"),
 * "h_aca", "h_ortho" (GMRES Gram-Schmidt), "lu_split" (bf16 split planes of
 * the outer trailing updates), "lu_gemm_bf16" (those updates from the split
 * planes; flops = the complex update's 8 M N K), "lu_graph" (factorizations
 * replayed as captured graphs; flops = 8/3 n^3). */
int bem_ctx_kernel_stats(bem_ctx ctx, const char *name, int64_t *launches, double *total_ms,
                         double *flops, double *bytes);
int bem_ctx_reset_stats(bem_ctx ctx);
/* Largest number of solves bem_sweep runs concurrently (1..8, default 8; the
 * HBM budget may allow fewer). */
int bem_ctx_set_lanes(bem_ctx ctx, int32_t max_lanes);
/* Largest number of heavy solves (16384 or more unknowns, whose trailing
 * updates alone fill the GPU's matrix cores) that bem_sweep runs at once
 * (1..8, default 1); the other lanes take the lighter solves meanwhile. */
int bem_ctx_set_heavy_lanes(bem_ctx ctx, int32_t max_heavy);

/* ---- LU tuning ----------------------------------------------------------
 * No counterpart in the reference (it calls LAPACK, src/solvers/bem/full.jl:404).
 * The shape of a factorization: outer block width (the K of the trailing
 * updates: 256 or 512), tile rows per L2 group of the update's tile order
 * (4, 8, 16), whether the complex-FP32 outer updates run as bf16-split 3M
 * products (FP32-accurate piece products on the bf16 matrix cores), whether
 * the next block's panels run ahead on a side stream, and whether a sweep
 * replays the factorization as a captured HIP graph. */
typedef struct {
    int64_t unknowns;
    int32_t precision;   /* BEM_DOUBLE or BEM_MIXED */
    int32_t outer, group, split, ahead, graph;
    int32_t tuned;       /* 1: chosen by the tuner; 0: the default shape */
    int32_t candidates;  /* shapes timed */
    double best_ms;      /* factorization time of the chosen shape */
    double default_ms;   /* of the default shape (outer 512, group 8, FP32 updates, look-ahead) */
} bem_lu_tune;
/* 0: the default shape everywhere; 1: the default with bf16-split updates
 * for complex FP32; 2 (default): per size class the tuner's choice where the
 * class was tuned, else the default with bf16-split updates from 4096
 * unknowns. */
int bem_ctx_set_lu_tuning(bem_ctx ctx, int32_t mode);
/* Times candidate shapes for an n-unknown factorization of the given
 * precision on a synthetic matrix (coordinate descent over the five knobs)
 * and keeps the fastest for that size class. */
int bem_ctx_autotune(bem_ctx ctx, int64_t n, int32_t precision, bem_lu_tune *out);
/* The shape a factorization of n unknowns would use now. */
int bem_ctx_lu_tune(bem_ctx ctx, int64_t n, int32_t precision, bem_lu_tune *out);

/* ---- mesh generators (host only, no device needed) --------------------- */
/* Sizes of the level-`level` icosphere: 10*4^L+2 nodes, 20*4^L triangles. */
int bem_icosphere_size(int32_t level, int64_t *n_nodes, int64_t *n_tri);
/* Icosphere mapped by (x,y,z) -> (ax x, ay y, az z): a sphere (ax=ay=az=a) or
 * a spheroid with its axis along x (src/geometry/spheroid.jl:7-28;
 * src/geometry/canonical_mesh.jl:24-27).  nodes 3 x Nv column-major,
 * tri 3 x N, 1-based, outward. */
int bem_icosphere(int32_t level, double ax, double ay, double az, double *nodes, int64_t *tri);
/* Body of revolution about x: meridian samples (x_s, r_s), s = 0..ns-1, with
 * r_0 = r_{ns-1} = 0 (poles), revolved with `n_phi` azimuthal segments
 * (src/geometry/bodies.jl:73-89).  Sizes: nodes = (ns-2)*n_phi + 2,
 * triangles = 2*(ns-2)*n_phi. */
int bem_revolution_size(int32_t ns, int32_t n_phi, int64_t *n_nodes, int64_t *n_tri);
/* Curved icosphere: the level-`level` icosphere with order-2 (6-node) or
 * order-3 (10-node) Lagrange triangles whose extra nodes are projected to the
 * unit sphere, then mapped like bem_icosphere; the curved counterpart of
 * gmsh_sphere_mesh / gmsh_spheroid_mesh with mesh_order 2 or 3
 * (src/geometry/canonical_mesh.jl:3-36).  Sizes: N = 20*4^L triangles, nodes
 * 10*4^L+2 (order 1), + 30*4^L (order 2), + 60*4^L + N (order 3).  tri is
 * npt x N, 1-based, Gmsh order (vertices, edge nodes along 1->2, 2->3, 3->1,
 * interior). */
int bem_curved_icosphere_size(int32_t level, int32_t order, int64_t *n_nodes, int64_t *n_tri);
int bem_curved_icosphere(int32_t level, double ax, double ay, double az, int32_t order, double *nodes,
                         int64_t *tri);
int bem_revolution(int32_t ns, const double *xs, const double *rs, int32_t n_phi, double *nodes,
                   int64_t *tri);

/* ---- mesh (src/geometry/surface.jl:271-298 supplied-mesh contract) ----- */
/* Validates the topology like src/geometry/surface.jl:341-405 (closed,
 * manifold, connected, consistently oriented, positive volume) and uploads
 * the panel geometry to the device. */
int bem_mesh_create(bem_ctx ctx, int64_t n_nodes, const double *nodes, int64_t n_tri,
                    const int64_t *tri, bem_mesh *out);
/* Validation only (host), with the reference's messages. */
int bem_mesh_validate(int64_t n_nodes, const double *nodes, int64_t n_tri, const int64_t *tri,
                      double *resolution, char *message, int32_t message_len);
int bem_mesh_destroy(bem_mesh mesh);
/* N, Nv, resolution (longest edge, surface.jl:363) and enclosed volume. */
int bem_mesh_info(bem_mesh mesh, int64_t *n_tri, int64_t *n_nodes, double *resolution,
                  double *volume);
/* Panel geometry: centroids 3 x N, unit normals 3 x N, areas N. */
int bem_mesh_panels(bem_mesh mesh, double *centroids, double *normals, double *areas);
/* 7-point Gauss nodes of every panel (3 x 7N coords, 3 x 7N normals, 7N
 * weights).  The Julia shim hands these to the reference's far_field
 * (src/solvers/bem/full.jl:509-519) as its `quad`. */
int bem_mesh_quadrature(bem_mesh mesh, double *coords, double *normals, double *weights);
/* Number of near (subdivided-quadrature) panel pairs. */
int bem_mesh_near_count(bem_mesh mesh, int64_t *count);

/* ---- curved-element Nystrom meshes --------------------------------------
 * The reference's own full-3D discretisation: mesh(nodes, triangles; qorder)
 * (src/geometry/surface.jl:271-298) over 3-, 6- or 10-node triangles in Gmsh
 * order (npt x N, 1-based), with an order-`qorder` quadrature on every curved
 * element (Inti.Quadrature, surface.jl:331; qorder 1..8, the reference's
 * default 4 gives 6 nodes per element).  The unknowns are the quadrature
 * nodes, numbered element by element; the layer operators are Nystrom
 * matrices corrected by density interpolation (correction=(method=:dim,),
 * src/solvers/bem/full.jl:89-97, src/solvers/bem/quadrature/fluid.jl:78-163).
 * The returned mesh works with every operator, solve, sweep and server call;
 * bem_mesh_info / bem_mesh_panels / bem_mesh_quadrature then report the
 * quadrature nodes (count, coordinates, normals, weights), which are the
 * reference's `quad` (full.jl:509-519).  The topology checks add the
 * reference's curved ones (repeated nodes, conforming edge nodes,
 * surface.jl:406-419).  bem_pressure on their solutions follows the
 * reference's Nystrom field evaluation (the nearest element excluded and
 * recovered by density interpolation, src/fields/pressure.jl:441-462, with
 * regions by the winding number of flat triangles through the element
 * nodes).  bem_hmatrix_create compresses them too (the element corrections
 * in the dense leaves): rigid surfaces, and fluid bodies in the direct
 * Mueller form or, below the Calderon limit, the Calderon form. */
int bem_mesh_create_nystrom(bem_ctx ctx, int64_t n_nodes, const double *nodes, int64_t n_tri,
                            int32_t nodes_per_tri, const int64_t *tri, int32_t qorder, bem_mesh *out);
/* Host only: the same validation and quadrature without a device.  With
 * coords == NULL only *n_quad is set (N times the rule's node count). */
int bem_nystrom_quadrature(int64_t n_nodes, const double *nodes, int64_t n_tri, int32_t nodes_per_tri,
                           const int64_t *tri, int32_t qorder, int64_t *n_quad, double *coords,
                           double *normals, double *weights, char *message, int32_t message_len);
/* Elements, nodes per element, quadrature nodes per element (0 for flat
 * panels) and qorder of a mesh. */
int bem_mesh_nystrom_info(bem_mesh mesh, int64_t *n_elements, int32_t *nodes_per_tri, int32_t *nq,
                          int32_t *qorder);

/* ---- operators: assemble + factor once, solve many ---------------------
 * Replaces _assemble_full_boundary/_assemble_full_fluid followed by
 * _factor_fluid_system (src/solvers/bem/full.jl:68-128,289-347,364-409);
 * used by incidence_angle_sweep (src/postprocessing/sweeps.jl:626-701). */
int bem_operator_create(bem_ctx ctx, bem_mesh mesh, const bem_problem *problem,
                        bem_operator *out);
int bem_operator_destroy(bem_operator op);
/* Download the assembled (unscaled) system matrix, column-major n x n
 * complex, n = N or 2N.  Test and diagnostics use. */
int bem_operator_matrix(bem_operator op, double *A, int64_t ld);
int bem_operator_size(bem_operator op, int64_t *n);

/* Solve for `n_rhs` plane waves with incidence angles (beta_r, alpha_r)
 * stored as angles[2r], angles[2r+1] (polar from +x, azimuth from +y toward
 * +z, full.jl:15-18).  Replaces _solve_full_boundary / _solve_full_fluid
 * (full.jl:130-184,456-495). */
int bem_operator_solve(bem_operator op, int32_t n_rhs, const double *angles, bem_solution *out);

/* One-shot solve_full_bem (full.jl:52-66,273-287): assemble, factor, solve. */
int bem_solve(bem_ctx ctx, bem_mesh mesh, const bem_problem *problem, double incidence_angle,
              double incidence_azimuth, bem_solution *out);

/* ---- solutions --------------------------------------------------------- */
int bem_solution_destroy(bem_solution sol);
int bem_solution_count(bem_solution sol, int32_t *n_rhs, int64_t *n_panels);
/* Scattered surface traces p_scat, dpdn_scat for right-hand side `r`
 * (N complex each), as returned by solve_full_bem. */
int bem_solution_traces(bem_solution sol, int32_t r, double *p_scat, double *dpdn_scat);
/* Far-field amplitude f(xhat) for `n_dir` unit directions (3 x n_dir) of
 * right-hand side `r` (full.jl:509-519), FP64 on the device.  A null `dirs`
 * means backscatter xhat = -d (src/solvers/bem/postprocessing.jl:4-6). */
int bem_far_field(bem_solution sol, int32_t r, int64_t n_dir, const double *dirs, double *out);
int bem_solution_diagnostics(bem_solution sol, int32_t r, bem_diagnostics *out);

/* ---- pressure at points (src/fields/pressure.jl:404-472) ----------------
 * Field codes of pressure(solution, points; field) (pressure.jl:1-20). */
#define BEM_FIELD_TOTAL 0     /* incident + scattered outside; transmitted inside a fluid body */
#define BEM_FIELD_SCATTERED 1 /* on and outside the surface only                              */
#define BEM_FIELD_INCIDENT 2  /* the unit plane wave                                          */
#define BEM_FIELD_INTERIOR 3  /* total pressure inside a fluid body                           */
/* Complex pressure at `n_points` points (3 x n column-major, metres) for
 * right-hand side `r` of a solution (dense or hierarchical).  Outside (or
 * on) the surface: the scattered field D p - S dp/dn of the surface traces
 * at wavenumber k; inside a fluid body: S (g (q + q_inc)) - D (p + p_inc) at
 * k_int (pressure.jl:426-472); the total field adds the incident wave
 * outside.  Inside and outside are decided by the surface's winding number
 * about each point.  A point where the selected field has no pressure (the
 * scattered field inside the body, any interior field of a rigid or
 * pressure-release body) fails with BEM_ERR_ARGUMENT and the reference's
 * message (pressure.jl:496-501).  FP64 on the device; a panel within twice
 * its longest edge of a point takes the subdivided near rule. */
int bem_pressure(bem_solution sol, int32_t r, int64_t n_points, const double *points, int32_t field,
                 double *out);
/* bem_far_field without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).
 * Same arguments, results and kernels as bem_far_field, for at most 1024
 * directions per call (more: BEM_ERR_ARGUMENT).  Its staging and partial sums
 * live in a fixed 4.05 MiB device array per GPU, held until the process
 * ends; calls that overlap on one GPU (from two contexts on the same device)
 * race on that array, calls through one context are serialised. */
int bem_far_field_noalloc(bem_solution sol, int32_t r, int64_t n_dir, const double *dirs, double *out);

/* ---- batched frequency sweep (src/postprocessing/sweeps.jl:103-110) ----
 * Solves `n` independent problems (mesh_i, problem_i) at incidence
 * (beta_i, alpha_i) and writes the backscatter amplitude of each (complex,
 * 2 doubles) and its wall latency [ms] from submission to completion.
 * Work is ordered largest-first and overlapped across device streams. */
int bem_sweep(bem_ctx ctx, int32_t n, const bem_mesh *meshes, const bem_problem *problems,
              const double *angles, double *amplitudes, double *latency_ms,
              bem_diagnostics *diagnostics);

/* ---- reduced-basis frequency sweep (src/postprocessing/sweeps.jl:113-306) --
 * frequency_sweep(surface, material::FluidFilled, frequencies, c;
 * training_frequencies, ...): dense FP64 CBIE fluid equations at each
 * wavenumber.  Training solves (full FP64 solves) fill an orthonormal snapshot
 * space of at most min(capacity, 2N, maxbytes / (16 * 2N)) columns, which is
 * then frozen; every query assembles its own equations, solves the
 * least-squares problem on the row-scaled images A V / r by pivoted QR, and
 * accepts it when both the original and the row-scaled relative residuals are
 * <= tolerance; otherwise (empty basis, rank loss, non-finite values or a
 * larger residual) it takes the full equilibrated LU solve, which must meet
 * the same residual (else BEM_ERR_NUMERIC).  Queries never enrich the basis.
 * amplitudes: backscatter f(-d) per query (2 doubles each). */
#define BEM_REDUCED_ACCEPTED 0
#define BEM_REDUCED_EMPTY 1
#define BEM_REDUCED_RESIDUAL 2
#define BEM_REDUCED_RANK 3
#define BEM_REDUCED_NONFINITE 4

typedef struct {
    double tolerance;      /* relative residual acceptance (1e-10)               */
    double rank_tolerance; /* QR rank cutoff and snapshot remainder (1e-12)      */
    int32_t capacity;      /* basis columns (16)                                 */
    int64_t maxbytes;      /* basis storage bound in bytes (64 MiB)              */
} bem_reduced_options;

typedef struct {
    double k;
    int32_t reduced;  /* 1: the basis solution was accepted                      */
    int32_t reason;   /* BEM_REDUCED_*                                           */
    int32_t rank;     /* query: rank of the images; training: basis size after   */
    int32_t kept;     /* training: the snapshot enlarged the basis               */
    double projected_original, projected_scaled; /* residuals of the basis solution */
    double residual_original, residual_scaled;   /* residuals of the returned one   */
} bem_reduced_report;

void bem_reduced_defaults(bem_reduced_options *options);
int bem_reduced_sweep(bem_ctx ctx, bem_mesh mesh, double density_contrast, double soundspeed_contrast,
                      int32_t n_train, const double *train_k, int32_t n_query, const double *query_k,
                      double incidence_angle, double incidence_azimuth, const bem_reduced_options *options,
                      double *amplitudes, bem_reduced_report *train, bem_reduced_report *query,
                      int32_t *basis_rank);

/* ---- serving: continuous request loop -------------------------------------
 * The reference has no server; its unit of work is the blocking solve that
 * `bem(surface, bc, k; incidence_angle, incidence_azimuth)` performs
 * (src/solvers/bem/surface.jl:10-22) followed by the backscatter amplitude
 * (src/postprocessing/bem/postprocessing.jl:1-13).  A server runs that unit
 * for requests submitted at any time, from any thread:
 *  - `lanes` worker streams run admitted requests concurrently on the device;
 *  - a request is admitted when a lane is free and its device block (the
 *    operator, factors and solve buffers, sized up front) fits in the
 *    context's HBM budget; the oldest waiting request has first claim on
 *    memory, later ones may pass it only with memory it cannot use;
 *  - with keep_factors, assembled operators and their LU factors stay resident
 *    after a request and serve later requests with the same mesh and problem
 *    (any incidence) without reassembly; resident operators are evicted least
 *    recently used when admission needs their memory.
 * Results are bit-identical to a solo bem_solve of the same request. */
typedef struct bem_server_s *bem_server;

typedef struct {
    int32_t lanes;        /* concurrent requests on the device, 1..16            */
    int32_t keep_factors; /* keep operators + factors resident for reuse (0/1)    */
} bem_server_options;

typedef struct {
    int64_t submitted, completed, failed, queued, running;
    int64_t resident_operators, factor_reuses;
    uint64_t resident_bytes; /* device bytes held by resident operators          */
} bem_server_stats;

int bem_server_create(bem_ctx ctx, const bem_server_options *options, bem_server *out);
/* Waits for running requests; queued requests end with BEM_ERR_STATE. */
int bem_server_destroy(bem_server srv);
/* Queues a backscatter request (incidence angles as in bem_solve); the
 * server holds a reference to the mesh until the request ends. */
int bem_server_submit(bem_server srv, bem_mesh mesh, const bem_problem *problem, double incidence_angle,
                      double incidence_azimuth, int64_t *ticket);
int bem_server_poll(bem_server srv, int64_t ticket, int32_t *done);
/* Blocks until the request ends or timeout_ms passes (< 0: no limit); returns
 * BEM_ERR_STATE on timeout, else the request's own status, and on success the
 * backscatter amplitude (2 doubles), the diagnostics and timing[3] = queue,
 * run and total latency in ms.  A finished request is forgotten after a
 * successful wait. */
int bem_server_wait(bem_server srv, int64_t ticket, double timeout_ms, double *amplitude,
                    bem_diagnostics *diagnostics, double *timing);
/* Message of a failed request (valid until the next call on the server). */
const char *bem_server_error(bem_server srv, int64_t ticket);
int bem_server_stats_get(bem_server srv, bem_server_stats *out);

/* ---- hierarchical-matrix path ---------------------------------------------
 * Reference: the compressed operators of src/solvers/bem/compression/
 * (assembly.jl:3-129, ACA blocks through HMatrices.jl; workspace.jl:3-122,
 * the batched H-matvec) solved by GMRES (full.jl:156-157; compression/
 * fluid.jl:421-449).  Rigid surfaces, Burton-Miller or CBIE, with the dense
 * path's scattered-field equation (full.jl:103-111,144-148): the system
 * operator 1/2 I - D [- (i/k) H] and the right-hand-side operator
 * S [+ (i/k)(1/2 I + K')] are both compressed over one cluster tree; the
 * scattered traces are x and -dp_inc/dn.  Fluid bodies in the Mueller
 * formulation (compression/fluid.jl:20-106): the 2N x 2N system
 * [I - De + Di, Se - g Si; -He + Hi/g, I + K'e - K'i] is compressed as one
 * operator over a cluster tree of its 2N unknowns (unknown i < N the
 * pressure, N + i the normal derivative on panel i, both at the centroid),
 * the right-hand side is [p_inc; dp_inc/dn] and the scattered traces are x
 * minus the incident traces; bodies in the low-frequency regime
 * (max(k, k_int) * radius <= pi/2, where the reference takes K' and H from
 * the Calderon identities through a hierarchical LU of S) are refused with
 * BEM_ERR_ARGUMENT.  Blocks are complex FP32 (dense leaves and ACA factors
 * U V^T); GMRES runs in FP64 with FP64 residuals of that operator,
 * right-preconditioned by block Jacobi on the diagonal leaf blocks. */
typedef struct bem_hmatrix_s *bem_hmatrix;

typedef struct {
    int32_t leaf_size;       /* cluster-tree leaf size (64)                          */
    double eta;              /* admissibility min(diam) < eta * dist (2)              */
    double aca_tolerance;    /* ACA relative Frobenius tolerance per block (1e-6)     */
    int32_t max_rank;        /* rank cap of one block, <= 256 (256)                   */
    int32_t gmres_restart;   /* Krylov dimension per cycle (50)                       */
    int32_t gmres_maxiter;   /* total iterations (1000)                               */
    double gmres_tolerance;  /* relative FP64 residual ||b - A x|| / ||b|| (1e-8)      */
    int32_t preconditioner;  /* right preconditioner: 0 none, 1 block Jacobi on the
                                diagonal leaf blocks, 2 two-level: block Jacobi plus an
                                orthonormal coarse space of the degree <= 2 harmonic
                                polynomials on each trace, P^-1 = M^-1 + (I - M^-1 A) Q
                                (Q^H A Q)^-1 Q^H (compression/fluid.jl:166-301) (2) */
    int32_t recycle_dimension; /* solutions kept to warm-start later solves of the same
                                matrix, bounded by 64 MiB of solution/image pairs
                                (compression/fluid.jl:303-419); 0 disables (8) */
} bem_hmatrix_options;

typedef struct {
    int64_t n, nodes, leaves, dense_blocks, lowrank_blocks, near_pairs;
    uint64_t dense_bytes, lowrank_bytes;
    int32_t max_rank, unconverged_blocks;
    double mean_rank;
    double plan_ms, assembly_ms, aca_ms;
} bem_hmatrix_info;

void bem_hmatrix_defaults(bem_hmatrix_options *options);
/* Cluster tree and block tree only (host, no device): `points` 3 x n
 * column-major.  perm (n, 0-based) receives the tree order; blocks, when
 * non-null with room for 5 * (n_dense + n_lowrank) values, receives rows
 * (lowrank 0/1, row_lo, m, col_lo, n), dense blocks first, each list in
 * block-tree order.  The rules are stated in the library's sources. */
int bem_hmatrix_plan(int64_t n, const double *points, int32_t leaf_size, double eta, double near_radius,
                     int64_t *perm, int64_t *n_nodes, int64_t *n_dense, int64_t *n_lowrank, int64_t *blocks,
                     int64_t blocks_cap);
int bem_hmatrix_create(bem_ctx ctx, bem_mesh mesh, const bem_problem *problem,
                       const bem_hmatrix_options *options, bem_hmatrix *out);
int bem_hmatrix_destroy(bem_hmatrix h);
int bem_hmatrix_get_info(bem_hmatrix h, bem_hmatrix_info *out);
/* y = A x for n_rhs columns (unknown order, n x n_rhs column-major complex;
 * n = bem_hmatrix_info.n: N panels, or 2N unknowns for a fluid body). */
int bem_hmatrix_matvec(bem_hmatrix h, int32_t n_rhs, const double *x, double *y);
/* Either compressed operator: which = 0 the system (as bem_hmatrix_matvec),
 * 1 the right-hand-side operator S [+ (i/k)(1/2 I + K')] of a rigid surface
 * (full.jl:144-148), so a caller can recompute b = B q_inc and the residual
 * b - A x outside GMRES.  Fluid bodies have no right-hand-side operator
 * (BEM_ERR_ARGUMENT for which = 1). */
int bem_hmatrix_apply(bem_hmatrix h, int32_t which, int32_t n_rhs, const double *x, double *y);
/* GMRES for n_rhs plane-wave incidences (angles as in bem_operator_solve);
 * the solution carries traces, far field and per-incidence diagnostics
 * (method 2, iterations, relative_residual = ||b - A x|| / ||b||). */
int bem_hmatrix_solve(bem_hmatrix h, int32_t n_rhs, const double *angles, bem_solution *out);

/* ---- axisymmetric BEM (src/solvers/bem/axisymmetric.jl) -------------------
 * A body of revolution given by its meridian: n_nodes points (rho_i, z_i),
 * z along the axis of symmetry (Cartesian x of the 3D path), from the north
 * pole (rho = 0, z max) to the south pole (rho = 0), so that (-dz, drho)/L is
 * the outward normal of each straight panel (MeridianMesh,
 * src/geometry/meridian.jl:16-162).  Piecewise-constant traces collocated at
 * the panel midpoints, azimuthal Fourier modes m = 0 .. m_max (only m = 0
 * for axial incidence), FP64 throughout. */
typedef struct bem_axi_s *bem_axi;
/* Mode operators K_m, V_m of assemble_cbie_operators_modes
 * (axisymmetric.jl:389-465,574-607) for m = m0 .. m0 + n_modes - 1 at
 * wavenumber k: K and V receive n_modes column-major n x n complex matrices
 * (n = n_nodes - 1 panels). */
int bem_axi_operators(bem_ctx ctx, int64_t n_nodes, const double *rho, const double *z, double k, int32_t m0,
                      int32_t n_modes, double *K, double *V);
/* bem(body, boundary, k; method=:axisymmetric, incidence_angle, m_max) for a
 * rigid, pressure-release or fluid-filled body (problem->boundary, k and the
 * contrasts; the other fields are ignored): solve_axial at incidence_angle
 * 0 (axisymmetric.jl:683-729,806-857), solve_oblique otherwise
 * (axisymmetric.jl:995-1057,1185-1300; fluid bodies as the reduced 2n
 * system in the exterior traces).  The incidence angle is measured from the
 * axis of symmetry, the azimuth of incidence is 0. */
int bem_axi_solve(bem_ctx ctx, int64_t n_nodes, const double *rho, const double *z, const bem_problem *problem,
                  double incidence_angle, int32_t m_max, bem_axi *out);
/* Panels, modes, largest relative residual ||A x - b|| / ||b|| over the
 * modes' systems (the reference's solve reports), and timings. */
int bem_axi_info(bem_axi sol, int32_t *n_panels, int32_t *n_modes, double *max_relative_residual,
                 double *assembly_ms, double *solve_ms);
/* Scattered traces of mode index `mode` (p_scat_modes[mode + 1],
 * dpdn_scat_modes[mode + 1] of the reference), n complex values each. */
int bem_axi_traces(bem_axi sol, int32_t mode, double *p_scat, double *dpdn_scat);
/* far_field(ps, p_scat_modes, dpdn_scat_modes, k, theta, phi)
 * (src/postprocessing/farfield.jl:14-100) at n_dir directions given as
 * (theta, phi) pairs: theta from the axis of symmetry, phi from the plane of
 * incidence; backscatter is (pi - incidence_angle, pi). */
int bem_axi_far_field(bem_axi sol, int64_t n_dir, const double *angles, double *out);
int bem_axi_destroy(bem_axi sol);

/* ---- test hook: dense LU of a caller matrix ------------------------------
 * Factors the column-major n x n complex matrix A on the device with the
 * library's LU (no equilibration; FP32 when bit 0 of flags is set) and returns
 * the factors column-major in LU (unit-lower L below the diagonal, U on and
 * above) and the 1-based pivot rows in ipiv, the LAPACK zgetrf/cgetrf
 * convention the reference relies on (src/solvers/bem/full.jl:404).  When b
 * and x are non-null it also solves A x = b with those factors, or
 * A^H x = b when bit 1 of flags is set.  Bits 2, 3 and 4 force the
 * factorization's shape (FP32 only for bit 2): bf16-split outer updates, FP32
 * updates, replay through a captured graph.  No counterpart in the reference: it
 * exists so tests can check the factorization and the solves by themselves. */
int bem_debug_lu(bem_ctx ctx, int64_t n, const double *A, int32_t flags, double *LU, int64_t *ipiv,
                 const double *b, double *x);

#ifdef __cplusplus
}
#endif
#endif /* BEMHIP_H */
