// Internal types behind the C ABI: the context (device + stream + last error) and the assembly plan.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vortexhip.h"

namespace vp {

struct Elem;
struct TestFar;
struct AsmParams;
struct AdjArgs;

// Iterative refinement that does not meet its stopping test (VP_E_CONVERGENCE at the C ABI)
struct ConvergenceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
};

struct RuleTab {  // all rules' points (3 barycentric coordinates each) and weights, concatenated
    const double *l;
    const double *w;
    int off[9];
    int n[9];
};

// The forward evaluation kernels whose occupancy cap the launch autotuner picks, and its candidates (waves per SIMD;
// index 0 is the default)
enum TuneKernel { TK_TOUCH = 0, TK_NEARPTS = 1, TK_NEARGRP = 2, TK_GROUP = 3, TK_COUNT = 4 };
constexpr int kTuneWaves[2] = {3, 2};
long long next_plan_serial();

// HIP graph replay of the narrow solves and of the wake field (default on); VP_GRAPHS=0 launches the same kernels one
// by one instead (for tools that cannot follow graph launches: rocprofv3's kernel trace has crashed inside
// hipGraphLaunch here)
// GMRES-IR before the FP64 refactorization of a mixed factorization whose refinement stops short (default on);
// VP_GMRES=0 goes straight to the refactorization
inline bool gmres_enabled()
{
    const char *e = std::getenv("VP_GMRES");
    return !(e && e[0] == '0');
}
// VP_GMRES=2 (a check of the method, not a production setting): GMRES-IR right after the first application of the
// mixed factors, with no plain refinement before it
inline bool gmres_only()
{
    const char *e = std::getenv("VP_GMRES");
    return e && e[0] == '2';
}
constexpr int GMRES_M = 30, GMRES_RESTARTS = 8;
// VP_LU_INVERSE: 0 never forms the explicit inverse of the factors, 1 forms it at the first narrow apply (tests), unset
// after LU_INVERSE_AFTER narrow applies
inline int lu_inverse_policy()
{
    const char *e = std::getenv("VP_LU_INVERSE");
    return e ? (e[0] == '0' ? 0 : 1) : -1;
}
constexpr int LU_INVERSE_AFTER = 24, LU_INVERSE_MAX = 16384;

inline bool graphs_enabled()
{
    const char *e = std::getenv("VP_GRAPHS");  // read per call (cheap), so a process may switch it
    return !(e && e[0] == '0');
}
std::string asm_tuning_report();

struct Plan {
    Ctx *ctx = nullptr;
    vp_assembly_options opts{};
    int nt = 0, nn = 0, nslot = 0, nw = 0, nwt = 0, nseg = 0, nsrc = 0, ncol = 0, ntouch = 0;
    long long lda = 0;
    int ntt = 0, nts = 0;
    unsigned int near_cap = 0, nnear = 0;
    std::vector<void *> owned;
    // device buffers
    Elem *d_el = nullptr;
    TestFar *d_tf = nullptr;
    int *d_nodes = nullptr, *d_cols = nullptr, *d_cs = nullptr;
    double *d_seg = nullptr, *d_x = nullptr;  // filament ends; every node's coordinates (body, then wake)
    int *d_segcol = nullptr;
    std::vector<int> h_fil;                    // the wake's filaments (node pairs), kept for update_wake
    int2 *d_touch = nullptr;
    int2 *d_near = nullptr;
    int *d_tte = nullptr, *d_ttr = nullptr, *d_ttn = nullptr, *d_tse = nullptr, *d_tsc = nullptr, *d_tsn = nullptr;
    uint8_t *d_ttl = nullptr, *d_tsl = nullptr;
    uint8_t *d_tiers = nullptr;
    double *A = nullptr, *d_b = nullptr, *d_sigma = nullptr;
    bool ownsA = false;
    unsigned long long *d_stats = nullptr;
    unsigned int *d_count = nullptr;
    RuleTab rt{};
    bool record_tiers = false;
    bool timing = false;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t tev[TK_COUNT][2] = {};       // around the launches the autotuner times
    bool tune_timed[TK_COUNT] = {};
    long long serial = 0;
    unsigned int nonfar_cap = 0, leaf_cap = 0, nnonfar = 0, nleaves = 0, group_cap = 0, ngroups = 0;
    int4 *d_groups = nullptr;
    int2 *d_nearm = nullptr, *d_touchp = nullptr;
    unsigned int nnearm = 0, ntouchp = 0, nsimple = 0;
    int4 *d_simple = nullptr;
    bool assembled = false;
    bool lists_sized = false;        // the non-far lists sized by an upper bound (the first assembly)
    bool nonfar_cap_forced = false;  // a first budget set by VP_NONFAR_CAP (no sizing pass)
    bool wake_only = false;  // assemble only the body test triangles' pairs with the wake's sources
    long long far_queue_wgs = 0;  // > 0: the far tier as a persistent work queue, this many CUs' workgroups (0: a grid)
    uint16_t *d_fq = nullptr;     // the far queue's deferred closed-form points, 3 TT_MAX TS per workgroup
    int nts_body = 0;        // source tiles of the body (the wake's follow)
    int2 *d_bnear = nullptr; // near (test triangle, particle) pairs of blob_rhs
    unsigned int bnear_cap = 0;
    // backward assembly (adjoint) buffers, made on first use
    double *d_lam = nullptr, *d_mu = nullptr, *d_G = nullptr;
    int *d_slotnode = nullptr;
    double *d_bdot = nullptr;
    int2 *d_nonfar = nullptr;
    struct Leaf *d_leaves = nullptr;

    // grow-only device buffers of the per-call entries (wake_product, blob_rhs), kept with the plan: a free-wake
    // step's calls allocate nothing once their sizes have been seen
    std::vector<void *> sbuf;
    std::vector<size_t> scap;
    double *scratch_d(int slot, size_t n);

    ~Plan();
    void build(Ctx *ctx, const vp_surface &g, const vp_wake *wake, const vp_assembly_options &o);
    void assemble(const double *sigma, const double *vinf, bool host_sigma);
    void stats(long long *pairs, long long *points);
    void kernel_ms(double *ms);
    AsmParams params() const;
    void wake_product(const double *mu_nodes, double *out);
    void update_wake(const double *wake_x);
    int tune_pick(int k);
    void tune_collect();
    void tune_release();
    void blob_rhs(int nb, const double *bx, const double *ba, const double *bd, double *out);
    void read_stats(unsigned long long *h);
    void adjoint(int nlam, const double *lam, const double *mu, const double *vinf, const double *dvinf, double *G,
                 double *bdot);
    template <int NL> void adjoint_pass(const AdjArgs &aa);
};

// The factored bordered system (solve.jl:65-87): the original FP64 bordered matrix M (row-major, npad x npad, the
// identity on the padding), its randomized LU in FP64 (mode 0) or FP32 (mode 1), the butterflies, and workspaces.
struct Factor {
    Ctx *ctx = nullptr;
    int mode = 0, nslot = 0, nq = 0, nw = 0, nb = 0, n = 0, npad = 0;
    long long ld = 0;
    double *M = nullptr, *d_ru = nullptr, *d_rv = nullptr;
    double *G64 = nullptr, *linv64 = nullptr, *uinv64 = nullptr;
    float *G32 = nullptr, *linv32 = nullptr, *uinv32 = nullptr;
    double *X = nullptr, *R = nullptr, *Y = nullptr, *B = nullptr, *nrm = nullptr, *part = nullptr;
    float *Xf = nullptr;
    int ws_nrhs = 0, ws_epoch = 0, last_iters = 0;
    // apply_inverse replayed as a HIP graph per direction (captured on its second call with the same nrhs and
    // workspace): the triangular solves are a chain of two kernels per 64-row panel, bound by launch cost
    struct InvGraph {
        hipGraphExec_t exec = nullptr;
        int nrhs = -1, epoch = -1, calls = 0;
    } inv_graph[2];
    long long graph_launches = 0;
    int fallbacks = 0;  // FP32 factorizations replaced by FP64 ones (refinement that did not converge)
    // GMRES-IR (refinement by preconditioned GMRES when plain refinement with the FP32 factors stops short)
    double *gV = nullptr, *gx = nullptr, *gw = nullptr, *gb = nullptr, *gh = nullptr;
    long long gmres_solves = 0, gmres_iters = 0;
    bool gmres_column(int c, int nrhs, bool trans, double tolf, int max_restarts);
    // The explicit inverse of the LU factors, G^-1 (npad x npad row-major, the factors' precision), made at the start of
    // a narrow solve once the factorization has served LU_INVERSE_AFTER narrow applications (and npad <=
    // LU_INVERSE_MAX): a narrow apply is then one GEMV between the butterflies instead of the panel sweeps, whose
    // launches bound it.  Zbuf: its output.
    void *ginv = nullptr;
    double *Zbuf = nullptr;
    bool inv_ready = false;
    long long narrow_applies = 0, inverse_builds = 0;
    void build_inverse();
    double anorm1 = 0.0, anorm_inf = -1.0, cond = -1.0;
    std::vector<double> se;
    std::vector<void *> owned;

    ~Factor();
    void build(Ctx *c, const double *A, long long lda, const vp_constraints &k, int mode);
    void init_from_dense(const double *Md, long long ldd, int n, int npad, int mode, double anorm);
    void ensure_ws(int nrhs);
    void apply_inverse(int nrhs, bool trans);
    void apply_inverse_replay(int nrhs, bool trans);
    int solve(const double *B, int nrhs, double *X, bool device_io, int max_iter, double *resid, bool trans = false,
              bool strict = true);
    void refactor_fp64();
    double cond_estimate();
};

std::string tuning_report();
void gemm_host(hipStream_t s, int prec, int flags, int M, int N, int K, const double *A, const double *B, double *C);
void diag_lu_host(hipStream_t s, int prec, int nb, const double *A, double *LU, double *Linv, double *Uinv);
void surface_velocity_host(Ctx *c, const vp_surface &g, const vp_wake *w, const double *sigma, const double *mu,
                           const double *mu_wake, const double *vinf, double kernel_far, int precision, double *u);
void body_velocity_host(Ctx *c, const vp_surface &g, const double *sigma, const double *gam, int np, const double *xs,
                        double kernel_far, int precision, double *u);
// The no-allocation blob sums (vp_blob_velocities_device): the size of the per-GPU scratch array in doubles (32 MiB),
// its device address per device (taken at vp_create), and the sums on device arrays.
constexpr size_t NOALLOC_DOUBLES = (size_t)1 << 22;
hipError_t noalloc_init(int device);
void blob_velocities_noalloc(Ctx *c, int nb, const double *bx, const double *ba, const double *bd, int np,
                             const double *xs, int precision, double *u);
void blob_velocities_host(Ctx *c, int nb, const double *bx, const double *ba, const double *bd, int np,
                          const double *xs, int precision, double *u);
// The blob sums by the cluster-particle treecode (vp_tree.hip), and its lists' coverage check (host only)
void blob_velocities_tree_host(Ctx *c, int nb, const double *bx, const double *ba, const double *bd, int np,
                               const double *xs, const vp_tree_options &o, double *u, vp_tree_stats *stats);
long long tree_coverage_host(int nb, const double *bx, int np, const double *xs, const vp_tree_options &o,
                             vp_tree_stats *stats);
// The free-wake step's velocity sums (the body's kernels and the wake's blobs at np targets, plus a base velocity)
// replayed as one HIP graph per node structure, its nodes' grids and arguments rewritten on every call.
struct WakeField {
    Ctx *c = nullptr;
    int device = 0;                 // the context's device and stream, kept for the destructor
    hipStream_t stream = nullptr;
    int nt = 0, precision = 0;
    double kernel_far = 10.0;
    Elem *el = nullptr;
    double *sig = nullptr, *gam = nullptr, *X = nullptr, *u = nullptr, *part = nullptr;
    double *bx = nullptr, *ba = nullptr, *bd = nullptr;
    size_t capX = 0, capU = 0, capP = 0, capB3 = 0, capA3 = 0, capB1 = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t node[3] = {nullptr, nullptr, nullptr};
    int shape = -1;  // the instantiated graph's nodes: 0 body field and reduction, 1 with the blob field too
    long long builds = 0, updates = 0, launches = 0;

    ~WakeField();
    void init(Ctx *ctx, const vp_surface &g, double kernel_far, int precision);
    void eval(const double *sigma, const double *gamma, int np, const double *xs, int nb, const double *bx,
              const double *alpha, const double *delta, const double *base, double *u);
};
int dense_solve_host(Ctx *c, int n, const double *M, int mode, const double *r, double *x);
// consistency of A (on the device, column-major lda) and b (host), out: rows, columns, rhs
void consistency_host(Ctx *c, const double *A, long long lda, const double *b, const vp_constraints &k, double *out);
void math_probe_host(Ctx *c, int fn, int n, const double *a, const double *b, double *out);
// assemble over a mixed mesh of isoparametric quads and flat triangles (vp_mixed.hip)
void mixed_assemble_host(Ctx *c, const vp_mixed_mesh &m, const vp_wake *w, const double *vinf,
                         const vp_assembly_options &o, double *A, double *b, long long *pairs, long long *points,
                         uint8_t *tiers);

}  // namespace vp

struct vp_factor {
    vp::Factor f;
};

struct vp_ctx {
    vp::Ctx c;
};
struct vp_plan {
    vp::Plan p;
};
struct vp_wake_field {
    vp::WakeField w;
};
