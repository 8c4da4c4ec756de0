/*
 * vortexhip: the C ABI of the MI355X (gfx950) Galerkin hot path of VortexPanel.jl.
 *
 * Plain C: pointers and sizes only.  All indices are 0-based (the Julia binding subtracts 1).  Every call returns
 * 0 on success or a negative status (VP_E_*); vp_last_error(ctx) gives the message.  One call at a time per context
 * (not reentrant); a context owns one HIP stream on one device.
 */
#ifndef VORTEXHIP_H
#define VORTEXHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VP_OK = 0,
    VP_E_ARGUMENT = -1, /* bad option or input: the reference throws ArgumentError */
    VP_E_DIMENSION = -2, /* array sizes disagree: the reference throws DimensionMismatch */
    VP_E_DEVICE = -3,   /* HIP runtime error (allocation, launch) */
    VP_E_SINGULAR = -4, /* zero pivot in the factorization: the reference's LAPACK getrf reports info > 0 */
    VP_E_CONVERGENCE = -5 /* FP64 iterative refinement did not meet its stopping test (no counterpart: the
                             reference's LAPACK solve has no refinement) */
};

typedef struct vp_ctx vp_ctx;
typedef struct vp_plan vp_plan;
typedef struct vp_factor vp_factor;

/* AssemblyOptions (src/assembly/assemble.jl:26-41) plus the precision of the path.
 * test: 0 hat, 1 onering, 2 dual.  precision: 0 plain FP64, 1 mixed (FP32 multipole far field).
 * quad_touch_levels: the grading levels of the touching rules on isoparametric quads (vp_assemble_mixed only).
 * `threaded` has no counterpart (the device is always parallel); `external` is not crossable (a Julia closure). */
typedef struct vp_assembly_options {
    int test;
    double far_ratio;
    int far_points;
    int near_points;
    double near_ratio;
    int near_depth;
    int touch_levels;
    double touch_ratio;
    int touch_order;
    double corner_angle;
    double contact_rel;
    double kernel_far;
    int precision;
    int quad_touch_levels;
} vp_assembly_options;

/* The SurfaceGeometry and DOFMapping arrays assemble reads (src/geometry/types.jl:174-199,
 * src/dofmap/types.jl:73-93): nodes x (nn x 3, row-major), triangles (nt x 3), corner slots (nt x 3), nslot, and the
 * geometry's touch pairs (ntouch x 2). */
typedef struct vp_surface {
    int nn;
    const double *x;
    int nt;
    const int *tris;
    const int *corner_slot;
    int nslot;
    int ntouch;
    const int *touch;
} vp_surface;

/* RigidWake (src/wake/rigid.jl:53-62): nodes (nw x 3), triangles (ntri x 3), filaments (nfil x 2), 0-based. */
typedef struct vp_wake {
    int nw;
    const double *x;
    int ntri;
    const int *tris;
    int nfil;
    const int *filaments;
} vp_wake;

/* ---- context ---------------------------------------------------------------------------------------------- */
int vp_create(int device, vp_ctx **out);
void vp_destroy(vp_ctx *ctx);
const char *vp_last_error(const vp_ctx *ctx);
/* Wait for the context's stream. */
int vp_sync(vp_ctx *ctx);

/* ---- assembly (src/assembly/assemble.jl:377-447) ------------------------------------------------------------ */

/* One-shot `assemble(g, dm, els, sigma, Vinf; opts, wake)`: A is written column-major, nslot x (nslot + nw)
 * (A[i + j*nslot], row = test slot, column = source slot then wake node), b has nslot entries; pairs/points per
 * tier in the reference's TIER_NAMES order (self, edge, vertex, contact, near, far).  wake may be NULL. */
int vp_assemble(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const double *sigma, const double *vinf,
                const vp_assembly_options *opts, double *A, double *b, long long *pairs, long long *points);

/* MixedMesh (src/assembly/mixed.jl:13-58), 0-based: nquad isoparametric quads (corner nodes in their outward order and
 * their DOF slots, nquad x 4), then ntri flat triangles (nodes and slots, ntri x 3); element touch pairs (ntouch x 2,
 * element ids with the quads first, either order). */
typedef struct vp_mixed_mesh {
    int nn;
    const double *x;
    int nquad;
    const int *quad_nodes;
    const int *quad_slots;
    int ntri;
    const int *tri_nodes;
    const int *tri_slots;
    int nslot;
    int ntouch;
    const int *touch;
} vp_mixed_mesh;

/* `assemble(g, dm, mesh::MixedMesh, V; opts, wake)` (src/assembly/mixed.jl:444-505): the quads' inner integrals by
 * quad_field (tensor Gauss or polar), their test integrals on the parameter square's tiers, the triangles' as in
 * vp_assemble; sources of strength -V.n (the quads' along their curved surface).  A column-major nslot x
 * (nslot + nw), b, pairs/points per tier as vp_assemble; tiers (may be NULL) receives the tier of every (element,
 * body element or wake triangle) pair, (nquad + ntri) x (nquad + ntri + wake triangles) bytes row-major.  FP64 only
 * (opts->precision is ignored).  VP_E_ARGUMENT for a degenerate element, bad options, or a test point on a source
 * quad's boundary (the reference's quad_field raises there). */
int vp_assemble_mixed(vp_ctx *ctx, const vp_mixed_mesh *m, const vp_wake *wake, const double *vinf,
                      const vp_assembly_options *opts, double *A, double *b, long long *pairs, long long *points,
                      uint8_t *tiers);

/* The same in steps, with the matrix kept in device memory: a plan holds the geometry, element records, tiles and
 * rule tables; vp_plan_assemble fills the plan's A and b (on the device) for new sigma and Vinf. */
int vp_plan_create(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const vp_assembly_options *opts,
                   vp_plan **out);
/* flags: VP_PLAN_RECORD_TIERS keeps the tier of every pair of the last assembly (see vp_plan_tiers);
 * VP_PLAN_TIMING records per-kernel HIP events (see vp_plan_timings); VP_PLAN_WAKE_ONLY assembles only the pairs of
 * body test triangles with the wake's sources (the wake columns of A, no body columns, no freestream in b), for the
 * known vortex elements of explicit_rhs. */
enum { VP_PLAN_RECORD_TIERS = 1, VP_PLAN_TIMING = 2, VP_PLAN_WAKE_ONLY = 4 };
int vp_plan_create_ex(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const vp_assembly_options *opts,
                      int flags, vp_plan **out);
int vp_plan_assemble(vp_plan *plan, const double *sigma, const double *vinf, long long *pairs, long long *points);
int vp_plan_read(vp_plan *plan, double *A, double *b);
/* Tier of every (body test triangle, source triangle) pair (nt x (nt + wake triangles), uint8, TIER_NAMES index)
 * of the last vp_plan_assemble; the plan must have been created with VP_PLAN_RECORD_TIERS. */
int vp_plan_tiers(vp_plan *plan, uint8_t *tiers);
/* Per-kernel device times of the last vp_plan_assemble, from HIP events on the context's stream (the plan must be
 * created with VP_PLAN_TIMING): ms[0] far pairs and far filaments, ms[1] routing of the non-far pairs and the
 * touching pairs' rule pieces, ms[2] plain near pairs, ms[3] touching pairs' points, ms[4] freestream; then ms[5]
 * the number of single-piece near pairs, ms[6] of touching rule pieces, ms[7] the far tier's closed-form
 * (non-multipole) point evaluations (8 values). */
int vp_plan_timings(vp_plan *plan, double *ms);
/* Backward assembly (no counterpart in the reference, which has no derivatives; it serves the adjoint dJ/dx of a
 * Galerkin solution): the gradient of lambdaᵀ (b(x) - A(x) mu) with respect to the node coordinates, for nlam sets of
 * adjoint weights lambda (nlam x nslot, row-major: one per objective) and doublet values mu (nslot + nw: slots, then
 * wake nodes), with the rules of the last vp_plan_assemble and sigma = -Vinf.n (source_strengths, assemble.jl:364).
 * grad: nlam x (nn + nw) x 3, body nodes then wake nodes.  Two sets share one pass of the pair integrals.  With a
 * direction dvinf (3 values, or null) it also returns lambdaᵀ b(sigma = -dvinf.n, dvinf) per set in lambda_b_dot
 * (nlam): the derivative of lambdaᵀ b along dvinf (b is linear in Vinf), e.g. for d/dalpha, without a second
 * assembly. */
int vp_plan_adjoint(vp_plan *plan, int nlam, const double *lambda, const double *mu, const double *vinf,
                    const double *dvinf, double *grad, double *lambda_b_dot);
/* `explicit_rhs` (src/assembly/explicit.jl:31-86) of known sheets and filaments: out = A_wake mu_wake (nslot values;
 * the right hand side is its negative), the wake columns of the last vp_plan_assemble times the doublets of the wake's
 * nodes (a wake-only plan whose wake holds the known sheets and filaments, one node per corner or filament end). */
int vp_plan_wake_product(vp_plan *plan, const double *mu_wake, double *out);
/* New coordinates (nw x 3) for the plan's wake nodes, same triangles and filaments (no counterpart in the reference:
 * a free wake's known sheets move every step, `step!`, src/freewake/step.jl:113-146, while their topology stays):
 * the wake's element records are rebuilt in place and the plan is reused; assemble again before reading. */
int vp_plan_update_wake(vp_plan *plan, const double *wake_x);
/* `explicit_rhs` of nb vortex particles (positions bx and strengths alpha nb x 3, cores delta): their right hand side
 * (nslot values, added to b) on the plan's body: the far rule beyond far_ratio max(diam, core), else near_rule. */
int vp_plan_blob_rhs(vp_plan *plan, int nb, const double *bx, const double *alpha, const double *delta, double *out);
/* Device pointers of the plan's A (column-major, leading dimension nslot) and b, for chaining device calls. */
int vp_plan_device_A(vp_plan *plan, double **A, double **b, long long *lda);
void vp_plan_destroy(vp_plan *plan);

/* `explicit_rhs(els, dm, src::ExplicitSources; opts)` (src/assembly/explicit.jl:31-86; ExplicitSources :10-18) as a
 * handle on one body.  A call takes nsheet flat triangles (sheet_x nsheet x 9: corners p, q, r) with corner doublets
 * (sheet_mu nsheet x 3), nseg filaments (seg_x nseg x 6: ends a, b) with end circulations (seg_gamma nseg x 2) and nblob
 * particles (bx, alpha nblob x 3, cores delta) and returns -∫ψ_i n·u dA for every slot (out, nslot values; added to
 * assemble's b).  The sheets and filaments take the assembly's tiers: out is minus their assembled columns (one wake
 * node per corner and filament end) times their doublets; the particles take the far rule or near_rule.  The handle
 * keeps a plan of the known elements and moves it in place while nsheet and nseg stay (the free wake's rows move every
 * step, step!, src/freewake/step.jl:113-146; new plans only when rows are added), and a plan of the body for the
 * particles.  One call at a time per handle. */
typedef struct vp_explicit vp_explicit;
int vp_explicit_create(vp_ctx *ctx, const vp_surface *g, const vp_assembly_options *opts, vp_explicit **out);
int vp_explicit_rhs(vp_explicit *h, int nsheet, const double *sheet_x, const double *sheet_mu, int nseg,
                    const double *seg_x, const double *seg_gamma, int nblob, const double *bx, const double *alpha,
                    const double *delta, double *out);
/* counts[0] plans of known elements built, [1] moved in place, [2] calls */
int vp_explicit_info(const vp_explicit *h, long long *counts);
void vp_explicit_destroy(vp_explicit *h);

/* ---- the free-wake step (FreeWakeSim / step!, src/freewake/step.jl:35-146) --------------------------------------
 * A native engine for step!: it keeps the free rows (node rows with their frozen doublets), their conversion into
 * particles with strip merging (state.jl:292-357), the loads of every step (step_loads, step.jl:149-163), forward-Euler
 * advection and delete_far! (step.jl:190-195), and does each step's device work through the library's own handles:
 * explicit_rhs of the known elements (vp_explicit), the solve with the factorization given here (vp_factor_solve) and
 * the velocity sums as one HIP graph (vp_wake_field).  The caller makes the set-up the reference's constructor makes
 * (step.jl:35-58): the trailing edge panel, its assembly with the body and the bordered factorization, and the
 * resolved options. */
typedef struct vp_freewake vp_freewake;
typedef struct vp_freewake_setup {
    const double *vinf;               /* V∞ (3) */
    double dt;                        /* FreeWakeOptions.dt */
    int advection;                    /* 0 free, 1 freestream */
    double delete_beyond;             /* spans downstream of the trailing edge; non-finite: no deletion */
    double span;                      /* the trailing edge panel's span (RigidWake.span) */
    int keep_rows;                    /* free rows kept before the oldest is converted (max_rows, resolved) */
    int core_mode;                    /* CoreRule (state.jl:68-92): 0 fixed core_value, 1 auto, 2 scaled */
    double core_value, overlap, core_factor, core_floor;
    const vp_assembly_options *opts;  /* the assembly options of explicit_rhs (kernel_far also for the velocities) */
    int field_precision;              /* the velocity sums: 0 FP64, 1 FP32 kernels with FP64 sums */
    vp_factor *factor;                /* the bordered factorization of the body with the trailing edge panel (kept) */
    const double *b0;                 /* assemble's b with the trailing edge panel (nslot) */
    int nq;                           /* DOFs (the factorization's size less its border rows) */
    const int *C_rowptr, *C_col;      /* C (nslot x nq), CSR */
    const double *C_val;
    int nte;                          /* wake trailing edge nodes */
    const int *K_rowptr, *K_col;      /* K (nte x nq), CSR: mu_w = K q */
    const double *K_val;
    const double *te_x;               /* the wake trailing edge nodes' positions (nte x 3), for kutta_joukowski_force */
    int nteseg;                       /* its segments along the lines (te_segments, forces.jl:22-23), pairs */
    const int *teseg;
    const double *edge;               /* the trailing edge panel's downstream edge (nte x 3): where rows are shed */
    int nstrip;                       /* Topology (state.jl:99-121): strips (pairs of trailing edge nodes) */
    const int *strips;
    const int *merge;                 /* rows merged per strip (nstrip), or NULL for no merging */
    int nfirst;                       /* the first and last nodes of open lines (the tips) */
    const int *first;
    int nlast;
    const int *last;
    const vp_wake *panel;             /* the trailing edge panel as a wake (nodes, triangles, filaments) */
    const int *panel_te;              /* its nodes' trailing edge node (RigidWake.te) */
    /* The wake blobs' velocities by the treecode of vp_blob_velocities_tree (no counterpart in the reference):
     * tree_theta 0 sums them directly (the reference's blob_velocities); otherwise theta in (0, 1) and tree_degree
     * 1..7.  The body's kernels stay direct sums either way. */
    double tree_theta;
    int tree_degree;
} vp_freewake_setup;
/* step!'s record (step.jl:138-140): n, t, F_pressure, M_pressure, F_kj, F_impulse (NaN on the first step), Ω, I,
 * rows, particles, λ (the first border multiplier), umax, t_rhs and t_vel (s); iters: the solve's refinements. */
typedef struct vp_freewake_record {
    int n, rows, particles, iters;
    double t;
    double F_pressure[3], M_pressure[3], F_kj[3], F_impulse[3], Omega[3], I[3];
    double lambda, umax, t_rhs, t_vel;
} vp_freewake_record;
int vp_freewake_create(vp_ctx *ctx, const vp_surface *g, const vp_freewake_setup *setup, vp_freewake **out);
/* One step; mu_w (nte, or NULL) receives the step's wake doublets */
int vp_freewake_step(vp_freewake *h, vp_freewake_record *rec, double *mu_w);
/* counts: [0] node rows, [1] particles, [2] nte, [3] strips, [4] steps taken, [5] conversions */
int vp_freewake_counts(const vp_freewake *h, long long *counts);
/* The state (any pointer may be NULL): node rows (rows x nte x 3, newest first) and their doublets (rows x nte),
 * mu_line (nte), particles (positions and strengths np x 3, cores np), their merge tags and groups (np), the strips'
 * merge batches (rows, length, width: nstrip each), the last step's mu (nslot), mu_w (nte) and element Cp (nt), the
 * impulse I of the last loads (3) and the time t. */
int vp_freewake_read(const vp_freewake *h, double *rows_x, double *rows_mu, double *mu_line, double *px, double *palpha,
                     double *pdelta, long long *tag, int *group, long long *batch_rows, double *batch_length,
                     double *batch_width, double *mu, double *mu_w, double *cp, double *I_prev, double *t);
void vp_freewake_destroy(vp_freewake *h);

/* ---- bordered solve (src/assembly/solve.jl:15-99) ---------------------------------------------------------- */

/* The DOFMapping pieces of the bordered system: C (nslot x nq) and the wake rows W (nw x nq, `wake_rows`,
 * src/dofmap/system.jl:17) as CSR, and per body the unscaled border vectors e and c (nb x nq each, row-major,
 * `bordered_rows`, src/dofmap/system.jl:41-50).  W may be NULL when nw == 0. */
typedef struct vp_constraints {
    int nslot;
    int nq;
    int nw;
    int nb;
    const int *C_rowptr;
    const int *C_col;
    const double *C_val;
    const int *W_rowptr;
    const int *W_col;
    const double *W_val;
    const double *e;
    const double *c;
} vp_constraints;

/* `factorize_galerkin(dm, A; W)` (solve.jl:81-87): forms M = [Cᵀ A [C; W]  se e; sc cᵀ  0] on the device (A is
 * column-major nslot x (nslot + nw); on the device when A_on_device, e.g. from vp_plan_device_A) and factors it.
 * mode 0: plain FP64 (randomized FP64 LU on FP64 MFMA); mode 1: mixed (FP32 LU on FP32 MFMA).  Both refine in FP64
 * against the FP64 M.  The matrix size is n = nq + nb. */
int vp_factorize(vp_ctx *ctx, const double *A, long long lda, int A_on_device, const vp_constraints *k, int mode,
                 vp_factor **out);
/* Solve M X = R for nrhs right hand sides (`solve_galerkin(F, b)`, solve.jl:95-99): R and X are n x nrhs
 * row-major host arrays (R[i*nrhs + r]; the caller forms [Cᵀ b; 0]).  iters: refinement steps taken; resid: the
 * largest relative residual |R - M X|_2 / |R|_2 over the right hand sides (the reference's `residual`, solve.jl:58).
 * Refinement stops by LAPACK dsgesv's test (|r|_inf <= |x|_inf |M|_inf eps sqrt(n)) within 30 steps.  Mixed factors
 * that do not meet it are replaced by an FP64 factorization of the same M (dsgesv's fallback; vp_factor_info reports
 * it) and the solve is repeated; FP64 factors that do not meet it return VP_E_CONVERGENCE. */
int vp_factor_solve(vp_factor *f, const double *R, int nrhs, double *X, int *iters, double *resid);
/* Solve Mᵀ X = R (the adjoint system; no counterpart in the reference): same layout as vp_factor_solve, through the
 * same factors read transposed (no second factorization), FP64 refinement against Mᵀ. */
int vp_factor_solve_transposed(vp_factor *f, const double *R, int nrhs, double *X, int *iters, double *resid);
/* 1-norm condition number estimate of M (1 / gecon's rcond, solve.jl:85). */
int vp_factor_cond(vp_factor *f, double *cond);
/* `consistency(dm, A, b; W)` (solve.jl:111-115): with ACf = A [C; W] (the k's C and W; its border vectors are not
 * read), out[0] = max |row sums of ACf| / max |ACf|, out[1] = max |column sums of ACf| / max |ACf|, out[2] =
 * |sum b| / sum |b|.  A is column-major nslot x (nslot + nw), on the device when A_on_device (vp_plan_device_A); b is
 * a host array of nslot values. */
int vp_consistency(vp_ctx *ctx, const double *A, long long lda, int A_on_device, const double *b,
                   const vp_constraints *k, double *out);
/* The border column scales se (nb values, solve.jl:33) and n. */
int vp_factor_info(vp_factor *f, int *n, double *se);
/* The factorization's current precision mode (0 FP64, 1 FP32 factors) and the number of FP64 fallbacks taken. */
int vp_factor_mode(vp_factor *f, int *mode, int *fallbacks);
/* GMRES-IR (no counterpart in the reference): the solves of a mixed factorization that plain refinement left short and
 * restarted GMRES on M, preconditioned by the FP32 factors, finished without an FP64 refactorization, and the GMRES
 * iterations taken (environment VP_GMRES=0 turns it off). */
int vp_factor_gmres(vp_factor *f, long long *solves, long long *iterations);
/* The explicit inverse of the factors (no counterpart in the reference): whether it is in use, and the narrow
 * applications counted before it was made.  A factorization forms it at the start of a narrow solve (one to four
 * right-hand sides) once it has served 24 narrow applications (a refinement step each) and its padded size is at most
 * 16,384; narrow solves then apply it by one matrix-vector product instead of the triangular sweeps.  Environment
 * VP_LU_INVERSE=0 never forms it, 1 forms it at the first narrow solve. */
int vp_factor_inverse(vp_factor *f, int *in_use, long long *narrow_applies);
void vp_factor_destroy(vp_factor *f);
/* The LU trailing update's GEMM tile shapes chosen by the autotuner (no counterpart in the reference), as JSON:
 * per precision the chosen BMxBNxBK (null before the first factorization large enough to time), the update size timed
 * and each candidate's milliseconds.  The tuner runs inside the first vp_factorize per precision; the environment
 * variable VP_GEMM_TILE=<index> forces a candidate instead.  The string lives until the next call. */
const char *vp_gemm_tuning(void);
/* The assembly's launch autotuner (no counterpart in the reference), as JSON: for each forward evaluation kernel
 * (touch_simple, near_points, near_group, group_eval) the occupancy cap chosen (waves per SIMD; null while it is still
 * being timed) and each candidate's time in ms on the plan it was timed on.  Candidates run in turn on one plan's
 * successive assemblies; the fastest is kept for the process.  VP_ASM_WAVES=<3|2> forces one.  The string lives until
 * the next call. */
const char *vp_assembly_tuning(void);

/* ---- N-body velocity sums (src/post/Post.jl:32-112, src/freewake/step.jl:176-187) -------------------------- */

/* `surface_velocity(els, dm, sigma, mu, Vinf; wake, mu_wake, kernel_far)` (Post.jl:32-64): the velocity at every
 * body triangle's centroid (u, nt x 3): Vinf + sum_s sigma_s w_s + gamma_s x w_s (the own element on its normal's
 * side), gamma from the slot values mu (nslot) through corner_slot; with a wake, its sheets from mu_wake (one value
 * per wake node) and its tip filaments.  precision 0: FP64; 1: FP32 kernels with FP64 sums (own element in FP64). */
int vp_surface_velocity(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const double *sigma, const double *mu,
                        const double *mu_wake, const double *vinf, double kernel_far, int precision, double *u);
/* `body_velocity(els, sigma, gamma, xs; kernel_far)` (Post.jl:92-112): the body's sources and sheets (gamma, nt x 3)
 * at np off-body points xs (np x 3), without the freestream. */
int vp_body_velocity(vp_ctx *ctx, const vp_surface *g, const double *sigma, const double *gamma, int np,
                     const double *xs, double kernel_far, int precision, double *u);
/* `blob_velocities(blobs, X)` (step.jl:176-187) with blob_velocity (src/kernels/blob.jl:21-25): nb Rosenhead-Moore
 * particles (positions bx and strengths alpha nb x 3, core radii delta) at np points. */
int vp_blob_velocities(vp_ctx *ctx, int nb, const double *bx, const double *alpha, const double *delta, int np,
                       const double *xs, int precision, double *u);

/* vp_blob_velocities on device arrays, allocating nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops or harnesses that forbid hipMalloc).
 * bx, alpha (nb x 3), delta (nb), xs (np x 3) and u (np x 3) are device pointers on the context's device; the work is
 * queued on the context's stream and not waited for (vp_sync waits).  The partial sums live in one fixed 32 MiB device
 * array per GPU, held until the process ends: np is at most 1,398,101 (VP_E_DIMENSION beyond).  Every call on a GPU
 * shares that array, so calls that overlap on one GPU (from other contexts or threads) race. */
int vp_blob_velocities_device(vp_ctx *ctx, int nb, const double *bx, const double *alpha, const double *delta, int np,
                              const double *xs, int precision, double *u);
/* Fast summation of blob_velocities (no counterpart in the reference, which sums directly and defers fast summation,
 * docs/claude_handoff.md:177): a cluster-particle treecode.  Target clusters well separated from a source cluster
 * (box radius <= theta * distance from the box centre to the sources' box) take that cluster's field at the
 * (degree+1)^3 Chebyshev points of their box and interpolate it (barycentric Lagrange); every other pair is summed
 * directly.  Each blob keeps its own core radius and exact kernel.  theta in (0, 1), degree 1..7; precision 0 FP64,
 * 1 FP32 kernels (relative to a nearby origin) with FP64 interpolation and sums. */
typedef struct vp_tree_options {
    double theta;
    int degree;
    int precision;
} vp_tree_options;
typedef struct vp_tree_stats {
    long long far_clusters;  /* target clusters that interpolate */
    long long far_pairs;     /* (target cluster, source cluster) pairs taken at the proxies */
    long long near_pairs;    /* (target leaf, source range) pairs summed directly */
    long long proxy_evals;   /* kernel evaluations at proxies */
    long long near_evals;    /* kernel evaluations at targets */
    long long interp_evals;  /* proxy values read by the interpolation */
    double build_ms;         /* everything before the kernels: sort_ms + lists_ms + uploads */
    double device_ms;        /* device: the evaluation kernels */
    double sort_ms;          /* inputs up, Morton codes, sort and gathers on the device, sorted codes and points down */
    double lists_ms;         /* host: trees and lists */
} vp_tree_stats;
int vp_blob_velocities_tree(vp_ctx *ctx, int nb, const double *bx, const double *alpha, const double *delta, int np,
                            const double *xs, const vp_tree_options *opts, double *u, vp_tree_stats *stats);
/* The treecode's lists on the host only (test entry): *bad = the number of target leaves whose direct ranges and far
 * ancestors' ranges do not cover every source exactly once (0 when the lists are right). */
int vp_tree_coverage(int nb, const double *bx, int np, const double *xs, const vp_tree_options *opts, long long *bad,
                     vp_tree_stats *stats);
/* Device buffers on the context's device for the device-pointer entry points (hipMalloc / hipFree), and blocking copies
 * to and from them on the context's stream. */
int vp_device_alloc(vp_ctx *ctx, long long bytes, void **out);
int vp_device_free(vp_ctx *ctx, void *p);
int vp_copy_to_device(vp_ctx *ctx, void *dst, const void *src, long long bytes);
int vp_copy_from_device(vp_ctx *ctx, void *dst, const void *src, long long bytes);

/* The free-wake step's velocity sums, `wake_velocities` (src/freewake/step.jl:166-173): base (V∞, or null for 0)
 * + the body's sources and sheets (as vp_body_velocity) + nb blobs (as vp_blob_velocities) at np points, as one HIP
 * graph: the body and blob fields are independent nodes joined by one reduction.  The body's elements are set up at
 * create; the graph is instantiated once (again only if blobs appear or vanish) and each eval rewrites its nodes' grids
 * and arguments, so the particle count may change on every call.  sigma (nt) and gamma (nt x 3) are passed per call.
 * One call at a time per handle (it shares the context's stream).  info: counts[0] graphs built, [1] node updates,
 * [2] graph launches. */
typedef struct vp_wake_field vp_wake_field;
int vp_wake_field_create(vp_ctx *ctx, const vp_surface *g, double kernel_far, int precision, vp_wake_field **out);
int vp_wake_field_eval(vp_wake_field *h, const double *sigma, const double *gamma, int np, const double *xs, int nb,
                       const double *bx, const double *alpha, const double *delta, const double *base, double *u);
int vp_wake_field_info(const vp_wake_field *h, long long *counts);
void vp_wake_field_destroy(vp_wake_field *h);

/* Test entry points (host arrays, row-major): C = op(A) B or C -= op(A) B with the library's MFMA GEMM (prec 0 FP64,
 * 1 FP32; B K x N, C M x N, all doubles on the host).  flags: bit 0 subtract; bit 1 transposed A (A stored K x M,
 * op(A) = Aᵀ), else A is M x K; bits 4-7 the autotuner's candidate: 0-7 the tile shapes (0 is the default), 8 the
 * library GEMM (rocBLAS) for subtract without transposed A, the default tile otherwise.  And x = M^-1 r for a dense n x n matrix through the randomized LU (mode 0 FP64, 1 FP32 factors) with
 * FP64 refinement. */
int vp_gemm(vp_ctx *ctx, int prec, int flags, int M, int N, int K, const double *A, const double *B, double *C);
int vp_dense_solve(vp_ctx *ctx, int n, const double *M, int mode, const double *r, double *x, int *iters);
/* Test entry point: the LU's diagonal-block kernel on one 64 x 64 block (A row-major, factored in its top-left nb x nb
 * part, the rest the identity; prec 0 FP64, 1 FP32): the packed factors (L unit lower below the diagonal, U on and
 * above it) and the inverses of L and U, row-major 64 x 64. */
int vp_diag_lu_probe(vp_ctx *ctx, int prec, int nb, const double *A, double *LU, double *Linv, double *Uinv);
/* Test entry point: the device's FP64 natural log (fn 0: out = log a), two-argument arctangent (fn 1:
 * out = atan2(a, b)), reciprocal square root (fn 2: out = 1/sqrt(a), a normal and positive) and square root of a
 * squared distance (fn 3: out = sqrt(a), a zero or normal and positive) as the kernels evaluate them, and the mixed
 * path's FP32 atan2 (fn 4: out = atan2(a, b), both rounded to FP32) and log(1 + u) (fn 5: u = a >= 0 rounded to FP32),
 * elementwise over n host values. */
int vp_math_probe(vp_ctx *ctx, int fn, int n, const double *a, const double *b, double *out);

#ifdef __cplusplus
}
#endif
#endif
