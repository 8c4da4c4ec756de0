// The C ABI (include/vortexhip.h): argument checks, error strings and the context's stream.
#include <hip/hip_runtime.h>

#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "vp_capi_internal.hpp"

namespace {

int fail(vp::Ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <typename F> int guard(vp::Ctx *c, F &&f)
{
    try {
        f();
        if (c) c->err.clear();
        return VP_OK;
    } catch (const std::invalid_argument &e) {
        return fail(c, VP_E_ARGUMENT, e.what());
    } catch (const std::out_of_range &e) {
        return fail(c, VP_E_ARGUMENT, e.what());
    } catch (const std::length_error &e) {
        return fail(c, VP_E_DIMENSION, e.what());
    } catch (const std::domain_error &e) {
        return fail(c, VP_E_SINGULAR, e.what());
    } catch (const vp::ConvergenceError &e) {
        return fail(c, VP_E_CONVERGENCE, e.what());
    } catch (const std::bad_alloc &) {
        return fail(c, VP_E_DEVICE, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(c, VP_E_DEVICE, e.what());
    }
}

}  // namespace

extern "C" {

int vp_create(int device, vp_ctx **out)
{
    if (!out) return VP_E_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return VP_E_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return VP_E_DEVICE;
    vp_ctx *c = new (std::nothrow) vp_ctx;
    if (!c) return VP_E_DEVICE;
    c->c.device = device;
    if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VP_E_DEVICE;
    }
    // the device address of the no-allocation entry point's scratch array, once per device (a failure is reported
    // by that entry point, which alone needs it)
    (void)vp::noalloc_init(device);
    *out = c;
    return VP_OK;
}

void vp_destroy(vp_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

const char *vp_last_error(const vp_ctx *ctx) { return ctx ? ctx->c.err.c_str() : "no context"; }

int vp_sync(vp_ctx *ctx)
{
    if (!ctx) return VP_E_ARGUMENT;
    return guard(&ctx->c, [&] {
        hipError_t e = hipStreamSynchronize(ctx->c.stream);
        if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    });
}

int vp_plan_create(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const vp_assembly_options *opts, vp_plan **out)
{
    return vp_plan_create_ex(ctx, g, wake, opts, 0, out);
}

int vp_plan_create_ex(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const vp_assembly_options *opts, int flags,
                      vp_plan **out)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!g || !opts || !out) return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    *out = nullptr;
    vp_plan *p = new (std::nothrow) vp_plan;
    if (!p) return fail(&ctx->c, VP_E_DEVICE, "host allocation failed");
    p->p.record_tiers = (flags & VP_PLAN_RECORD_TIERS) != 0;
    p->p.timing = (flags & VP_PLAN_TIMING) != 0;
    p->p.wake_only = (flags & VP_PLAN_WAKE_ONLY) != 0;
    int rc = guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        p->p.build(&ctx->c, *g, wake && (wake->nw > 0) ? wake : nullptr, *opts);
    });
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return VP_OK;
}

int vp_plan_assemble(vp_plan *plan, const double *sigma, const double *vinf, long long *pairs, long long *points)
{
    if (!plan || !plan->p.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    if (!sigma || !vinf) return fail(c, VP_E_ARGUMENT, "null argument");
    return guard(c, [&] {
        plan->p.assemble(sigma, vinf, true);
        if (pairs && points) plan->p.stats(pairs, points);
    });
}

int vp_plan_read(vp_plan *plan, double *A, double *b)
{
    if (!plan || !plan->p.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    return guard(c, [&] {
        vp::Plan &P = plan->p;
        if (A && hipMemcpyAsync(A, P.A, sizeof(double) * (size_t)P.lda * P.ncol, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            throw std::runtime_error("copy of A failed");
        if (b && hipMemcpyAsync(b, P.d_b, sizeof(double) * P.nslot, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            throw std::runtime_error("copy of b failed");
        if (hipStreamSynchronize(c->stream) != hipSuccess) throw std::runtime_error("stream synchronization failed");
    });
}

int vp_plan_tiers(vp_plan *plan, uint8_t *tiers)
{
    if (!plan || !plan->p.ctx || !tiers) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    if (!plan->p.d_tiers) return fail(c, VP_E_ARGUMENT, "the plan does not record tiers (VP_PLAN_RECORD_TIERS)");
    return guard(c, [&] {
        vp::Plan &P = plan->p;
        if (hipMemcpyAsync(tiers, P.d_tiers, (size_t)P.nt * P.nsrc, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            throw std::runtime_error("copy of the tiers failed");
    });
}

int vp_plan_timings(vp_plan *plan, double *ms)
{
    if (!plan || !plan->p.ctx || !ms) return VP_E_ARGUMENT;
    if (!plan->p.timing) return fail(plan->p.ctx, VP_E_ARGUMENT, "the plan was not created with VP_PLAN_TIMING");
    return guard(plan->p.ctx, [&] { plan->p.kernel_ms(ms); });
}

int vp_plan_device_A(vp_plan *plan, double **A, double **b, long long *lda)
{
    if (!plan || !plan->p.A) return VP_E_ARGUMENT;
    if (A) *A = plan->p.A;
    if (b) *b = plan->p.d_b;
    if (lda) *lda = plan->p.lda;
    return VP_OK;
}

int vp_plan_adjoint(vp_plan *plan, int nlam, const double *lambda, const double *mu, const double *vinf,
                    const double *dvinf, double *grad, double *lambda_b_dot)
{
    if (!plan || !plan->p.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    if (nlam < 1) return fail(c, VP_E_ARGUMENT, "nlam must be at least 1");
    if (!lambda || !mu || !vinf || !grad || (dvinf && !lambda_b_dot)) return fail(c, VP_E_ARGUMENT, "null argument");
    return guard(c, [&] {
        (void)hipSetDevice(c->device);
        plan->p.adjoint(nlam, lambda, mu, vinf, dvinf, grad, dvinf ? lambda_b_dot : nullptr);
    });
}

int vp_plan_wake_product(vp_plan *plan, const double *mu_wake, double *out)
{
    if (!plan || !plan->p.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    if (!out || (plan->p.nw > 0 && !mu_wake)) return fail(c, VP_E_ARGUMENT, "null argument");
    return guard(c, [&] { plan->p.wake_product(mu_wake, out); });
}

int vp_plan_update_wake(vp_plan *plan, const double *wake_x)
{
    if (!plan || !plan->p.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    if (!wake_x) return fail(c, VP_E_ARGUMENT, "null argument");
    if (plan->p.nw <= 0) return fail(c, VP_E_DIMENSION, "the plan has no wake");
    return guard(c, [&] {
        (void)hipSetDevice(c->device);
        plan->p.update_wake(wake_x);
    });
}

int vp_plan_blob_rhs(vp_plan *plan, int nb, const double *bx, const double *alpha, const double *delta, double *out)
{
    if (!plan || !plan->p.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = plan->p.ctx;
    if (!out || nb < 0 || (nb > 0 && (!bx || !alpha || !delta))) return fail(c, VP_E_ARGUMENT, "null argument");
    return guard(c, [&] {
        (void)hipSetDevice(c->device);
        plan->p.blob_rhs(nb, bx, alpha, delta, out);
    });
}

void vp_plan_destroy(vp_plan *plan)
{
    if (!plan) return;
    if (plan->p.ctx) (void)hipStreamSynchronize(plan->p.ctx->stream);
    delete plan;
}

}  // extern "C"

// explicit_rhs as a handle: copies of the body's arrays (a new plan of the known elements needs them again), the
// two wake-only plans and the known elements as a wake (one node per corner and filament end)
struct vp_explicit {
    vp_ctx *ctx = nullptr;
    vp_assembly_options opts{};
    std::vector<double> x;
    std::vector<int> tris, cs, touch;
    vp_surface g{};
    vp_plan *body = nullptr, *known = nullptr;
    int nsheet = -1, nseg = -1;
    std::vector<double> wx, wmu, tmp, zero;
    std::vector<int> wtris, wfil;
    long long counts[3] = {0, 0, 0};
    ~vp_explicit()
    {
        vp_plan_destroy(known);
        vp_plan_destroy(body);
    }
};

extern "C" {

int vp_explicit_create(vp_ctx *ctx, const vp_surface *g, const vp_assembly_options *opts, vp_explicit **out)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!g || !opts || !out || !g->x || !g->tris || !g->corner_slot || (g->ntouch > 0 && !g->touch))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (g->nn <= 0 || g->nt <= 0 || g->nslot <= 0 || g->ntouch < 0) return fail(&ctx->c, VP_E_DIMENSION, "empty body");
    *out = nullptr;
    vp_explicit *h = new (std::nothrow) vp_explicit;
    if (!h) return fail(&ctx->c, VP_E_DEVICE, "host allocation failed");
    const int rc = guard(&ctx->c, [&] {
        h->ctx = ctx;
        h->opts = *opts;
        h->x.assign(g->x, g->x + 3 * (size_t)g->nn);
        h->tris.assign(g->tris, g->tris + 3 * (size_t)g->nt);
        h->cs.assign(g->corner_slot, g->corner_slot + 3 * (size_t)g->nt);
        if (g->ntouch > 0) h->touch.assign(g->touch, g->touch + 2 * (size_t)g->ntouch);
        h->g = *g;
        h->g.x = h->x.data();
        h->g.tris = h->tris.data();
        h->g.corner_slot = h->cs.data();
        h->g.touch = h->touch.empty() ? nullptr : h->touch.data();
        h->tmp.assign(g->nslot, 0.0);
        h->zero.assign(g->nt, 0.0);
    });
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return VP_OK;
}

int vp_explicit_rhs(vp_explicit *h, int nsheet, const double *sheet_x, const double *sheet_mu, int nseg,
                    const double *seg_x, const double *seg_gamma, int nblob, const double *bx, const double *alpha,
                    const double *delta, double *out)
{
    if (!h || !h->ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = &h->ctx->c;
    if (!out || nsheet < 0 || nseg < 0 || nblob < 0 || (nsheet > 0 && (!sheet_x || !sheet_mu)) ||
        (nseg > 0 && (!seg_x || !seg_gamma)) || (nblob > 0 && (!bx || !alpha || !delta)))
        return fail(c, VP_E_ARGUMENT, "null argument");
    const int nslot = h->g.nslot;
    for (int i = 0; i < nslot; i++) out[i] = 0.0;
    h->counts[2]++;
    static const double vzero[3] = {0.0, 0.0, 0.0};
    if (nsheet + nseg > 0) {
        const int nw = 3 * nsheet + 2 * nseg;
        h->wx.resize(3 * (size_t)nw);
        h->wmu.resize(nw);
        std::memcpy(h->wx.data(), sheet_x, sizeof(double) * 9 * (size_t)nsheet);
        std::memcpy(h->wmu.data(), sheet_mu, sizeof(double) * 3 * (size_t)nsheet);
        std::memcpy(h->wx.data() + 9 * (size_t)nsheet, seg_x, sizeof(double) * 6 * (size_t)nseg);
        std::memcpy(h->wmu.data() + 3 * (size_t)nsheet, seg_gamma, sizeof(double) * 2 * (size_t)nseg);
        int rc;
        if (h->known && h->nsheet == nsheet && h->nseg == nseg) {
            if ((rc = vp_plan_update_wake(h->known, h->wx.data()))) return rc;
            h->counts[1]++;
        } else {
            vp_plan_destroy(h->known);
            h->known = nullptr;
            h->wtris.resize(3 * (size_t)nsheet);
            for (int i = 0; i < 3 * nsheet; i++) h->wtris[i] = i;
            h->wfil.resize(2 * (size_t)nseg);
            for (int i = 0; i < 2 * nseg; i++) h->wfil[i] = 3 * nsheet + i;
            const vp_wake w{nw, h->wx.data(), nsheet, h->wtris.data(), nseg, h->wfil.data()};
            if ((rc = vp_plan_create_ex(h->ctx, &h->g, &w, &h->opts, VP_PLAN_WAKE_ONLY, &h->known))) return rc;
            h->nsheet = nsheet;
            h->nseg = nseg;
            h->counts[0]++;
        }
        if ((rc = vp_plan_assemble(h->known, h->zero.data(), vzero, nullptr, nullptr))) return rc;
        if ((rc = vp_plan_wake_product(h->known, h->wmu.data(), h->tmp.data()))) return rc;
        for (int i = 0; i < nslot; i++) out[i] -= h->tmp[i];
    }
    if (nblob > 0) {
        int rc;
        if (!h->body) {
            if ((rc = vp_plan_create_ex(h->ctx, &h->g, nullptr, &h->opts, VP_PLAN_WAKE_ONLY, &h->body))) return rc;
            if ((rc = vp_plan_assemble(h->body, h->zero.data(), vzero, nullptr, nullptr))) {
                vp_plan_destroy(h->body);
                h->body = nullptr;
                return rc;
            }
        }
        if ((rc = vp_plan_blob_rhs(h->body, nblob, bx, alpha, delta, h->tmp.data()))) return rc;
        for (int i = 0; i < nslot; i++) out[i] += h->tmp[i];
    }
    return VP_OK;
}

int vp_explicit_info(const vp_explicit *h, long long *counts)
{
    if (!h || !counts) return VP_E_ARGUMENT;
    for (int k = 0; k < 3; k++) counts[k] = h->counts[k];
    return VP_OK;
}

void vp_explicit_destroy(vp_explicit *h) { delete h; }

int vp_factorize(vp_ctx *ctx, const double *A, long long lda, int A_on_device, const vp_constraints *k, int mode,
                 vp_factor **out)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!A || !k || !out || !k->C_rowptr || !k->C_col || !k->C_val || (k->nw > 0 && (!k->W_rowptr || !k->W_col || !k->W_val)) ||
        (k->nb > 0 && (!k->e || !k->c)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (lda < k->nslot) return fail(&ctx->c, VP_E_DIMENSION, "lda is smaller than nslot");
    *out = nullptr;
    vp_factor *f = new (std::nothrow) vp_factor;
    if (!f) return fail(&ctx->c, VP_E_DEVICE, "host allocation failed");
    double *tmp = nullptr;
    int rc = guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        const double *Ad = A;
        if (!A_on_device) {
            size_t bytes = sizeof(double) * (size_t)lda * (size_t)(k->nslot + k->nw);
            if (hipMalloc((void **)&tmp, bytes) != hipSuccess) throw std::runtime_error("device allocation of A failed");
            if (hipMemcpyAsync(tmp, A, bytes, hipMemcpyHostToDevice, ctx->c.stream) != hipSuccess)
                throw std::runtime_error("copy of A failed");
            Ad = tmp;
        }
        f->f.build(&ctx->c, Ad, lda, *k, mode);
    });
    if (tmp) {
        (void)hipStreamSynchronize(ctx->c.stream);
        (void)hipFree(tmp);
    }
    if (rc) {
        delete f;
        return rc;
    }
    *out = f;
    return VP_OK;
}

int vp_consistency(vp_ctx *ctx, const double *A, long long lda, int A_on_device, const double *b,
                   const vp_constraints *k, double *out)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!A || !b || !k || !out || !k->C_rowptr || !k->C_col || !k->C_val ||
        (k->nw > 0 && (!k->W_rowptr || !k->W_col || !k->W_val)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (k->nslot <= 0 || k->nq <= 0 || k->nw < 0) return fail(&ctx->c, VP_E_ARGUMENT, "empty system");
    if (lda < k->nslot) return fail(&ctx->c, VP_E_DIMENSION, "lda is smaller than nslot");
    double *tmp = nullptr;
    int rc = guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        const double *Ad = A;
        if (!A_on_device) {
            size_t bytes = sizeof(double) * (size_t)lda * (size_t)(k->nslot + k->nw);
            if (hipMalloc((void **)&tmp, bytes) != hipSuccess) throw std::runtime_error("device allocation of A failed");
            if (hipMemcpyAsync(tmp, A, bytes, hipMemcpyHostToDevice, ctx->c.stream) != hipSuccess)
                throw std::runtime_error("copy of A failed");
            Ad = tmp;
        }
        vp::consistency_host(&ctx->c, Ad, lda, b, *k, out);
    });
    if (tmp) {
        (void)hipStreamSynchronize(ctx->c.stream);
        (void)hipFree(tmp);
    }
    return rc;
}

int vp_factor_solve(vp_factor *f, const double *R, int nrhs, double *X, int *iters, double *resid)
{
    if (!f || !f->f.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = f->f.ctx;
    if (!R || !X || nrhs <= 0) return fail(c, VP_E_ARGUMENT, "null argument or nrhs < 1");
    return guard(c, [&] {
        int it = f->f.solve(R, nrhs, X, false, 30, resid);
        if (iters) *iters = it;
    });
}

int vp_factor_solve_transposed(vp_factor *f, const double *R, int nrhs, double *X, int *iters, double *resid)
{
    if (!f || !f->f.ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = f->f.ctx;
    if (!R || !X || nrhs <= 0) return fail(c, VP_E_ARGUMENT, "null argument or nrhs < 1");
    return guard(c, [&] {
        int it = f->f.solve(R, nrhs, X, false, 30, resid, true);
        if (iters) *iters = it;
    });
}

int vp_factor_cond(vp_factor *f, double *cond)
{
    if (!f || !f->f.ctx || !cond) return VP_E_ARGUMENT;
    return guard(f->f.ctx, [&] {
        if (f->f.cond < 0) f->f.cond = f->f.anorm1 * f->f.cond_estimate();
        *cond = f->f.cond;
    });
}

const char *vp_gemm_tuning(void)
{
    static thread_local std::string report;
    report = vp::tuning_report();
    return report.c_str();
}

const char *vp_assembly_tuning(void)
{
    static thread_local std::string report;
    report = vp::asm_tuning_report();
    return report.c_str();
}

int vp_factor_info(vp_factor *f, int *n, double *se)
{
    if (!f) return VP_E_ARGUMENT;
    if (n) *n = f->f.n;
    if (se)
        for (size_t i = 0; i < f->f.se.size(); i++) se[i] = f->f.se[i];
    return VP_OK;
}

int vp_factor_mode(vp_factor *f, int *mode, int *fallbacks)
{
    if (!f) return VP_E_ARGUMENT;
    if (mode) *mode = f->f.mode;
    if (fallbacks) *fallbacks = f->f.fallbacks;
    return VP_OK;
}

int vp_factor_gmres(vp_factor *f, long long *solves, long long *iterations)
{
    if (!f) return VP_E_ARGUMENT;
    if (solves) *solves = f->f.gmres_solves;
    if (iterations) *iterations = f->f.gmres_iters;
    return VP_OK;
}

int vp_factor_inverse(vp_factor *f, int *in_use, long long *narrow_applies)
{
    if (!f) return VP_E_ARGUMENT;
    if (in_use) *in_use = f->f.inv_ready ? 1 : 0;
    if (narrow_applies) *narrow_applies = f->f.narrow_applies;
    return VP_OK;
}

void vp_factor_destroy(vp_factor *f)
{
    if (!f) return;
    if (f->f.ctx) (void)hipStreamSynchronize(f->f.ctx->stream);
    delete f;
}

int vp_gemm(vp_ctx *ctx, int prec, int flags, int M, int N, int K, const double *A, const double *B, double *C)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!A || !B || !C || (prec != 0 && prec != 1) || M <= 0 || N <= 0 || K <= 0 || (flags & ~0xF3) ||
        ((flags >> 4) & 15) > 8)
        return fail(&ctx->c, VP_E_ARGUMENT, "bad argument");
    return guard(&ctx->c, [&] { vp::gemm_host(ctx->c.stream, prec, flags, M, N, K, A, B, C); });
}

int vp_math_probe(vp_ctx *ctx, int fn, int n, const double *a, const double *b, double *out)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!a || !out || n <= 0 || fn < 0 || fn > 5 || ((fn == 1 || fn == 4) && !b)) return fail(&ctx->c, VP_E_ARGUMENT, "bad argument");
    return guard(&ctx->c, [&] { vp::math_probe_host(&ctx->c, fn, n, a, b, out); });
}

int vp_diag_lu_probe(vp_ctx *ctx, int prec, int nb, const double *A, double *LU, double *Linv, double *Uinv)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!A || !LU || !Linv || !Uinv || (prec != 0 && prec != 1) || nb < 1 || nb > 64)
        return fail(&ctx->c, VP_E_ARGUMENT, "bad argument");
    return guard(&ctx->c, [&] { vp::diag_lu_host(ctx->c.stream, prec, nb, A, LU, Linv, Uinv); });
}

int vp_dense_solve(vp_ctx *ctx, int n, const double *M, int mode, const double *r, double *x, int *iters)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!M || !r || !x || n <= 0 || (mode != 0 && mode != 1)) return fail(&ctx->c, VP_E_ARGUMENT, "bad argument");
    return guard(&ctx->c, [&] {
        int it = vp::dense_solve_host(&ctx->c, n, M, mode, r, x);
        if (iters) *iters = it;
    });
}

int vp_surface_velocity(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const double *sigma, const double *mu,
                        const double *mu_wake, const double *vinf, double kernel_far, int precision, double *u)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!g || !g->x || !g->tris || !g->corner_slot || !sigma || !mu || !vinf || !u)
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (precision != 0 && precision != 1) return fail(&ctx->c, VP_E_ARGUMENT, "precision must be 0 (fp64) or 1 (mixed)");
    if (g->nt <= 0 || g->nn <= 0 || g->nslot <= 0) return fail(&ctx->c, VP_E_DIMENSION, "empty surface");
    const vp_wake *w = wake && wake->nw > 0 ? wake : nullptr;
    if (w && (!w->x || !mu_wake || (w->ntri > 0 && !w->tris) || (w->nfil > 0 && !w->filaments)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null wake argument");
    return guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        vp::surface_velocity_host(&ctx->c, *g, w, sigma, mu, mu_wake, vinf, kernel_far, precision, u);
    });
}

int vp_body_velocity(vp_ctx *ctx, const vp_surface *g, const double *sigma, const double *gamma, int np,
                     const double *xs, double kernel_far, int precision, double *u)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!g || !g->x || !g->tris || !sigma || !gamma || (np > 0 && (!xs || !u)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (precision != 0 && precision != 1) return fail(&ctx->c, VP_E_ARGUMENT, "precision must be 0 (fp64) or 1 (mixed)");
    if (np < 0 || g->nt <= 0 || g->nn <= 0) return fail(&ctx->c, VP_E_DIMENSION, "bad sizes");
    if (np == 0) return VP_OK;
    return guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        vp::body_velocity_host(&ctx->c, *g, sigma, gamma, np, xs, kernel_far, precision, u);
    });
}

int vp_blob_velocities(vp_ctx *ctx, int nb, const double *bx, const double *alpha, const double *delta, int np,
                       const double *xs, int precision, double *u)
{
    if (!ctx) return VP_E_ARGUMENT;
    if ((nb > 0 && (!bx || !alpha || !delta)) || (np > 0 && (!xs || !u)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (precision != 0 && precision != 1) return fail(&ctx->c, VP_E_ARGUMENT, "precision must be 0 (fp64) or 1 (mixed)");
    if (nb < 0 || np < 0) return fail(&ctx->c, VP_E_DIMENSION, "negative size");
    if (np == 0) return VP_OK;
    return guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        vp::blob_velocities_host(&ctx->c, nb, bx, alpha, delta, np, xs, precision, u);
    });
}

int vp_blob_velocities_tree(vp_ctx *ctx, int nb, const double *bx, const double *alpha, const double *delta, int np,
                            const double *xs, const vp_tree_options *opts, double *u, vp_tree_stats *stats)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!opts || (nb > 0 && (!bx || !alpha || !delta)) || (np > 0 && (!xs || !u)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (nb < 0 || np < 0) return fail(&ctx->c, VP_E_DIMENSION, "negative size");
    if (stats) *stats = vp_tree_stats{};
    if (np == 0) return VP_OK;
    return guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        vp::blob_velocities_tree_host(&ctx->c, nb, bx, alpha, delta, np, xs, *opts, u, stats);
    });
}

int vp_tree_coverage(int nb, const double *bx, int np, const double *xs, const vp_tree_options *opts, long long *bad,
                     vp_tree_stats *stats)
{
    if (!opts || !bad || (nb > 0 && !bx) || (np > 0 && !xs)) return VP_E_ARGUMENT;
    if (nb < 0 || np < 0) return VP_E_DIMENSION;
    if (stats) *stats = vp_tree_stats{};
    return guard(nullptr, [&] { *bad = vp::tree_coverage_host(nb, bx, np, xs, *opts, stats); });
}

int vp_blob_velocities_device(vp_ctx *ctx, int nb, const double *bx, const double *alpha, const double *delta, int np,
                              const double *xs, int precision, double *u)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (nb < 0 || np < 0) return fail(&ctx->c, VP_E_DIMENSION, "negative size");
    if ((size_t)np * 3 > vp::NOALLOC_DOUBLES)
        return fail(&ctx->c, VP_E_DIMENSION,
                    "np = " + std::to_string(np) + " targets exceed the no-allocation scratch array (at most " +
                        std::to_string(vp::NOALLOC_DOUBLES / 3) + ")");
    if ((nb > 0 && (!bx || !alpha || !delta)) || (np > 0 && (!xs || !u)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (precision != 0 && precision != 1) return fail(&ctx->c, VP_E_ARGUMENT, "precision must be 0 (fp64) or 1 (mixed)");
    if (np == 0) return VP_OK;
    return guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        vp::blob_velocities_noalloc(&ctx->c, nb, bx, alpha, delta, np, xs, precision, u);
    });
}

int vp_device_alloc(vp_ctx *ctx, long long bytes, void **out)
{
    if (!ctx || !out) return VP_E_ARGUMENT;
    *out = nullptr;
    if (bytes <= 0) return fail(&ctx->c, VP_E_ARGUMENT, "bytes must be positive");
    (void)hipSetDevice(ctx->c.device);
    if (hipMalloc(out, (size_t)bytes) != hipSuccess) return fail(&ctx->c, VP_E_DEVICE, "device allocation failed");
    return VP_OK;
}

int vp_device_free(vp_ctx *ctx, void *p)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!p) return VP_OK;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    if (hipFree(p) != hipSuccess) return fail(&ctx->c, VP_E_DEVICE, "device free failed");
    return VP_OK;
}

int vp_copy_to_device(vp_ctx *ctx, void *dst, const void *src, long long bytes)
{
    if (!ctx || !dst || !src || bytes < 0) return VP_E_ARGUMENT;
    (void)hipSetDevice(ctx->c.device);
    if (hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->c.stream) != hipSuccess ||
        hipStreamSynchronize(ctx->c.stream) != hipSuccess)
        return fail(&ctx->c, VP_E_DEVICE, "copy to the device failed");
    return VP_OK;
}

int vp_copy_from_device(vp_ctx *ctx, void *dst, const void *src, long long bytes)
{
    if (!ctx || !dst || !src || bytes < 0) return VP_E_ARGUMENT;
    (void)hipSetDevice(ctx->c.device);
    if (hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->c.stream) != hipSuccess ||
        hipStreamSynchronize(ctx->c.stream) != hipSuccess)
        return fail(&ctx->c, VP_E_DEVICE, "copy from the device failed");
    return VP_OK;
}

int vp_wake_field_create(vp_ctx *ctx, const vp_surface *g, double kernel_far, int precision, vp_wake_field **out)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!out || !g || !g->x || !g->tris) return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    *out = nullptr;
    if (precision != 0 && precision != 1) return fail(&ctx->c, VP_E_ARGUMENT, "precision must be 0 (fp64) or 1 (mixed)");
    if (g->nt <= 0 || g->nn <= 0) return fail(&ctx->c, VP_E_DIMENSION, "empty surface");
    auto *h = new vp_wake_field();
    int rc = guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        h->w.init(&ctx->c, *g, kernel_far, precision);
    });
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return VP_OK;
}

int vp_wake_field_eval(vp_wake_field *h, const double *sigma, const double *gamma, int np, const double *xs, int nb,
                       const double *bx, const double *alpha, const double *delta, const double *base, double *u)
{
    if (!h || !h->w.c) return VP_E_ARGUMENT;
    vp::Ctx *c = h->w.c;
    if (!sigma || !gamma || (np > 0 && (!xs || !u)) || (nb > 0 && (!bx || !alpha || !delta)))
        return fail(c, VP_E_ARGUMENT, "null argument");
    if (np < 0 || nb < 0) return fail(c, VP_E_DIMENSION, "negative size");
    if (np == 0) return VP_OK;
    return guard(c, [&] {
        (void)hipSetDevice(c->device);
        h->w.eval(sigma, gamma, np, xs, nb, bx, alpha, delta, base, u);
    });
}

int vp_wake_field_info(const vp_wake_field *h, long long *counts)
{
    if (!h || !counts) return VP_E_ARGUMENT;
    counts[0] = h->w.builds;
    counts[1] = h->w.updates;
    counts[2] = h->w.launches;
    return VP_OK;
}

void vp_wake_field_destroy(vp_wake_field *h)
{
    if (!h) return;
    if (h->w.c) (void)hipSetDevice(h->w.device);
    delete h;
}

int vp_assemble(vp_ctx *ctx, const vp_surface *g, const vp_wake *wake, const double *sigma, const double *vinf,
                const vp_assembly_options *opts, double *A, double *b, long long *pairs, long long *points)
{
    vp_plan *p = nullptr;
    int rc = vp_plan_create(ctx, g, wake, opts, &p);
    if (rc) return rc;
    rc = vp_plan_assemble(p, sigma, vinf, pairs, points);
    if (!rc) rc = vp_plan_read(p, A, b);
    vp_plan_destroy(p);
    return rc;
}

int vp_assemble_mixed(vp_ctx *ctx, const vp_mixed_mesh *m, const vp_wake *wake, const double *vinf,
                      const vp_assembly_options *opts, double *A, double *b, long long *pairs, long long *points,
                      uint8_t *tiers)
{
    if (!ctx) return VP_E_ARGUMENT;
    if (!m || !opts || !vinf || !A || !b || !m->x) return fail(&ctx->c, VP_E_ARGUMENT, "null argument");
    if (m->nquad < 0 || m->ntri < 0 || m->ntouch < 0) return fail(&ctx->c, VP_E_DIMENSION, "negative size");
    if ((m->nquad > 0 && (!m->quad_nodes || !m->quad_slots)) || (m->ntri > 0 && (!m->tri_nodes || !m->tri_slots)) ||
        (m->ntouch > 0 && !m->touch))
        return fail(&ctx->c, VP_E_ARGUMENT, "null mesh array");
    const vp_wake *w = wake && wake->nw > 0 ? wake : nullptr;
    if (w && (!w->x || (w->ntri > 0 && !w->tris) || (w->nfil > 0 && !w->filaments)))
        return fail(&ctx->c, VP_E_ARGUMENT, "null wake argument");
    return guard(&ctx->c, [&] {
        (void)hipSetDevice(ctx->c.device);
        vp::mixed_assemble_host(&ctx->c, *m, w, vinf, *opts, A, b, pairs, points, tiers);
    });
}

}  // extern "C"
