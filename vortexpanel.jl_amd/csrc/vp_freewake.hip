// The free-wake step (step!, src/freewake/step.jl:113-146, with the bookkeeping of src/freewake/state.jl) as a native
// engine behind the C ABI: the free rows, their conversion into particles (with strip merging), the loads, the
// forward-Euler advection and the deletion of far particles run here on the host, in step order; the step's device
// work goes through the library's own handles: the known elements' right-hand side (vp_explicit), the solve with the
// factorization of the body and its trailing edge panel (vp_factor), and the velocity sums as one HIP graph
// (vp_wake_field).  The caller builds the trailing edge panel, assembles and factors once (FreeWakeSim's set-up,
// step.jl:35-58) and passes the pieces in vp_freewake_setup.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "vp_capi_internal.hpp"

namespace {

struct P3 {
    double x = 0, y = 0, z = 0;
};
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator*(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline P3 operator/(P3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(P3 a) { return std::sqrt(dot(a, a)); }
inline P3 ld(const double *p) { return {p[0], p[1], p[2]}; }

// CoreRule, state.jl:68-92: the core radius of particles and of the blobs standing for panels and filaments
struct Core {
    int mode = 0;  // 0 fixed, 1 auto, 2 scaled
    double value = 0, overlap = 1.5, factor = 1.5, floor = 0;
    double tri(P3 p, P3 q, P3 r) const
    {
        if (mode == 0) return value;
        if (mode == 1) return overlap * std::max(std::max(norm(q - p), norm(r - q)), norm(p - r)) / 2;
        return std::max(factor * std::sqrt(norm(cross(q - p, r - p))), floor);
    }
    double seg(P3 a, P3 b) const
    {
        if (mode == 0) return value;
        if (mode == 1) return overlap * norm(b - a) / 2;
        return std::max(factor * norm(b - a) / 2, floor);
    }
};

struct Blobs {
    std::vector<P3> x, a;
    std::vector<double> d;
    size_t size() const { return d.size(); }
    void push(P3 xx, P3 aa, double dd)
    {
        x.push_back(xx);
        a.push_back(aa);
        d.push_back(dd);
    }
    void append(const Blobs &o)
    {
        x.insert(x.end(), o.x.begin(), o.x.end());
        a.insert(a.end(), o.a.begin(), o.a.end());
        d.insert(d.end(), o.d.begin(), o.d.end());
    }
};

// Known vortex elements (state.jl:157-165): every triangle with its corner doublets (those with an area also as
// sheets) and filament segments with their end circulations
struct Elements {
    std::vector<P3> corners;  // 3 per triangle
    std::vector<double> cmu;  // 3 per triangle
    std::vector<double> sheet_x, sheet_mu;  // 9 and 3 per sheet (the non-degenerate triangles)
    std::vector<P3> seg;                    // 2 per segment
    std::vector<double> seg_g;              // 2 per segment
    int nsheet() const { return (int)sheet_mu.size() / 3; }
    int nseg() const { return (int)seg_g.size() / 2; }
};

// state.jl:171: |(q - p) x (r - p)| > 1e-12 times the largest squared edge
bool nondegenerate(P3 p, P3 q, P3 r)
{
    const P3 u = q - p, v = r - p, w = r - q;
    return norm(cross(u, v)) > 1e-12 * std::max(std::max(dot(u, u), dot(v, v)), dot(w, w));
}

void add_sheet(Elements &e, P3 p, P3 q, P3 r, double mp, double mq, double mr)
{
    e.corners.insert(e.corners.end(), {p, q, r});
    e.cmu.insert(e.cmu.end(), {mp, mq, mr});
    if (nondegenerate(p, q, r)) {
        e.sheet_x.insert(e.sheet_x.end(), {p.x, p.y, p.z, q.x, q.y, q.z, r.x, r.y, r.z});
        e.sheet_mu.insert(e.sheet_mu.end(), {mp, mq, mr});
    }
}

void add_segment(Elements &e, P3 a, P3 b, double ga, double gb)
{
    e.seg.insert(e.seg.end(), {a, b});
    e.seg_g.insert(e.seg_g.end(), {ga, gb});
}

// state.jl:258 and src/kernels/blob.jl:33-38: a particle at a triangle's centroid with its sheet's vorticity, two
// 2-point Gauss blobs per filament
void free_wake_blobs(const Elements &e, const Core &core, Blobs &out)
{
    const size_t nt = e.corners.size() / 3;
    for (size_t i = 0; i < nt; i++) {
        const P3 p = e.corners[3 * i], q = e.corners[3 * i + 1], r = e.corners[3 * i + 2];
        const double *m = &e.cmu[3 * i];
        const P3 w = ((m[0] + m[1]) * (q - p) + (m[1] + m[2]) * (r - q) + (m[2] + m[0]) * (p - r)) / 2.0;
        out.push((p + q + r) / 3.0, w, core.tri(p, q, r));
    }
    const double s = 1 / (2 * std::sqrt(3.0));
    for (int j = 0; j < e.nseg(); j++) {
        const P3 a = e.seg[2 * j], b = e.seg[2 * j + 1];
        const double ga = e.seg_g[2 * j], gb = e.seg_g[2 * j + 1], dl = core.seg(a, b);
        for (double xi : {0.5 - s, 0.5 + s}) out.push(a + xi * (b - a), (((1 - xi) * ga + xi * gb) / 2) * (b - a), dl);
    }
}

// CSR matrix (rows of a DOFMapping operator)
struct Csr {
    int nrow = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
    void mul(const double *x, double *y) const  // y = M x
    {
        for (int i = 0; i < nrow; i++) {
            double s = 0;
            for (int k = ptr[i]; k < ptr[i + 1]; k++) s += val[k] * x[col[k]];
            y[i] = s;
        }
    }
    void mul_t(const double *x, double *y, int ncol) const  // y = Mᵀ x
    {
        for (int j = 0; j < ncol; j++) y[j] = 0;
        for (int i = 0; i < nrow; i++)
            for (int k = ptr[i]; k < ptr[i + 1]; k++) y[col[k]] += val[k] * x[i];
    }
};

Csr copy_csr(int nrow, const int *ptr, const int *col, const double *val)
{
    Csr m;
    m.nrow = nrow;
    m.ptr.assign(ptr, ptr + nrow + 1);
    m.col.assign(col, col + ptr[nrow]);
    m.val.assign(val, val + ptr[nrow]);
    return m;
}

}  // namespace

struct vp_freewake {
    vp_ctx *ctx = nullptr;
    vp_factor *F = nullptr;  // not owned
    vp_explicit *ex = nullptr;
    vp_wake_field *field = nullptr;
    // the body
    int nt = 0, nslot = 0, nq = 0, nb = 0;
    std::vector<P3> nrm, c, gradl, gam;  // per triangle; gradl and gam 3 per triangle
    std::vector<double> area, sigma;
    std::vector<int> cs;
    Csr C, K;
    std::vector<double> b0, se;
    // the trailing edge and the panel
    int nte = 0;
    std::vector<P3> edge, te_x;
    std::vector<int> strips, first, last, teseg, merge;
    std::vector<P3> panel_x;
    std::vector<int> panel_te, panel_tris, panel_fil;
    // options
    P3 V;
    double dt = 0, delete_beyond = 0, span = 0;
    int advection = 0, keep_rows = 0;
    vp_tree_options tree{0.0, 0, 0};  // the wake blobs by the treecode when tree.theta > 0
    Core core;
    // state (state.jl:123-143)
    std::vector<std::vector<P3>> rows;  // node rows, newest (R0) first
    std::vector<std::vector<double>> rmu;
    std::vector<double> mu_line;
    Blobs particles;
    std::vector<long long> tag;
    std::vector<signed char> group;
    std::vector<long long> batch_rows;
    std::vector<double> batch_length, batch_width;
    long long conversions = 0;
    long long nstep = 0;
    double t = 0;
    std::vector<double> mu, mu_w, cp;
    P3 I_prev;
    // scratch
    std::vector<double> rhs, R, X, U, Xt, bx, ba, bd, g3, Ub;

    ~vp_freewake()
    {
        vp_explicit_destroy(ex);
        vp_wake_field_destroy(field);
    }

    int nstrip() const { return (int)strips.size() / 2; }

    // panel_row!, state.jl:188-201 (i: the row between node rows i and i + 1)
    void panel_row(Elements &e, size_t i) const
    {
        const auto &x0 = rows[i], &x1 = rows[i + 1];
        const auto &m0 = rmu[i], &m1 = rmu[i + 1];
        for (int s = 0; s < nstrip(); s++) {
            const int A = strips[2 * s], B = strips[2 * s + 1];
            add_sheet(e, x0[A], x1[A], x0[B], m0[A], m1[A], m0[B]);
            add_sheet(e, x1[A], x1[B], x0[B], m1[A], m1[B], m0[B]);
        }
        for (int k : first) add_segment(e, x1[k], x0[k], m1[k], m0[k]);
        for (int k : last) add_segment(e, x0[k], x1[k], m0[k], m1[k]);
    }

    // far_end!, state.jl:204-212
    void far_end(Elements &e, size_t i) const
    {
        const auto &x = rows[i];
        const auto &m = rmu[i];
        for (int s = 0; s < nstrip(); s++) {
            const int A = strips[2 * s], B = strips[2 * s + 1];
            const double gA = m[A] - mu_line[A], gB = m[B] - mu_line[B];
            if (gA == 0 && gB == 0) continue;
            add_segment(e, x[B], x[A], gB, gA);
        }
    }

    // vortex_elements(st, topo), state.jl:222-237
    Elements vortex_elements() const
    {
        Elements e;
        const size_t nr = rows.size();
        for (size_t i = 0; i + 1 < nr; i++) panel_row(e, i);
        if (nr >= 1) {
            for (int s = 0; s < nstrip(); s++) {
                const int A = strips[2 * s], B = strips[2 * s + 1];
                add_segment(e, rows[0][A], rows[0][B], rmu[0][A], rmu[0][B]);
            }
            far_end(e, nr - 1);
        }
        return e;
    }

    // vortex_elements(w::RigidWake, mu_w), state.jl:240-250: the trailing edge panel
    Elements panel_elements(const std::vector<double> &mw) const
    {
        Elements e;
        for (size_t k = 0; k < panel_tris.size() / 3; k++) {
            const int a = panel_tris[3 * k], b = panel_tris[3 * k + 1], c3 = panel_tris[3 * k + 2];
            add_sheet(e, panel_x[a], panel_x[b], panel_x[c3], mw[panel_te[a]], mw[panel_te[b]], mw[panel_te[c3]]);
        }
        for (size_t k = 0; k < panel_fil.size() / 2; k++) {
            const int a = panel_fil[2 * k], b = panel_fil[2 * k + 1];
            add_segment(e, panel_x[a], panel_x[b], mw[panel_te[a]], mw[panel_te[b]]);
        }
        return e;
    }

    void keep_particles(const std::vector<char> &keep)
    {
        Blobs p;
        std::vector<long long> tg;
        std::vector<signed char> gr;
        for (size_t i = 0; i < particles.size(); i++)
            if (keep[i]) {
                p.push(particles.x[i], particles.a[i], particles.d[i]);
                tg.push_back(tag[i]);
                gr.push_back(group[i]);
            }
        particles = std::move(p);
        tag = std::move(tg);
        group = std::move(gr);
    }

    // merge_strip!, state.jl:337-357 (s 0-based; tags hold s + 1, 0 once final)
    void merge_strip(int s)
    {
        const P3 o;
        const double d = core.tri(o, o + P3{batch_width[s], 0, 0}, o + P3{batch_width[s], batch_length[s], 0});
        Blobs merged;
        for (int gi = 1; gi <= 2; gi++) {
            P3 al, xs;
            double wsum = 0;
            int cnt = 0;
            for (size_t i = 0; i < particles.size(); i++)
                if (tag[i] == s + 1 && group[i] == gi) {
                    al = al + particles.a[i];
                    cnt++;
                }
            if (!cnt) continue;
            for (size_t i = 0; i < particles.size(); i++)
                if (tag[i] == s + 1 && group[i] == gi) wsum += norm(particles.a[i]);
            for (size_t i = 0; i < particles.size(); i++)
                if (tag[i] == s + 1 && group[i] == gi)
                    xs = xs + (wsum > 0 ? norm(particles.a[i]) : 1.0) * particles.x[i];
            merged.push(xs / (wsum > 0 ? wsum : (double)cnt), al, d);
        }
        std::vector<char> keep(particles.size());
        for (size_t i = 0; i < keep.size(); i++) keep[i] = tag[i] != s + 1;
        keep_particles(keep);
        for (size_t i = 0; i < merged.size(); i++) {
            particles.push(merged.x[i], merged.a[i], merged.d[i]);
            tag.push_back(0);
            group.push_back(0);
        }
        batch_rows[s] = 0;
        batch_length[s] = 0.0;
    }

    // convert_oldest!, state.jl:292-328: the oldest panel row into particles
    void convert_oldest()
    {
        const size_t nr = rows.size();
        if (nr < 2) return;
        Elements e;
        panel_row(e, nr - 2);
        far_end(e, nr - 1);
        Blobs b;
        free_wake_blobs(e, core, b);
        std::vector<long long> tg(b.size(), 0);
        std::vector<signed char> gr(b.size(), 0);
        if (!merge.empty()) {
            const auto &up = rows[nr - 2], &down = rows[nr - 1];
            for (int s = 0; s < nstrip(); s++) {
                const int k = merge[s];
                if (k <= 1) continue;
                const int A = strips[2 * s], B = strips[2 * s + 1];
                const long long m = k - batch_rows[s];
                const double uv[2] = {1.0 / 3, 2.0 / 3};
                for (int h = 0; h < 2; h++) {
                    tg[2 * s + h] = s + 1;
                    gr[2 * s + h] = uv[h] + ((double)(m - 1) + uv[h]) / k < 1 ? 1 : 2;
                }
                batch_rows[s] += 1;
                batch_length[s] += (norm(down[A] - up[A]) + norm(down[B] - up[B])) / 2;
                batch_width[s] = norm(up[B] - up[A]);
            }
        }
        particles.append(b);
        tag.insert(tag.end(), tg.begin(), tg.end());
        group.insert(group.end(), gr.begin(), gr.end());
        mu_line = rmu[nr - 2];
        rows.pop_back();
        rmu.pop_back();
        conversions++;
        if (!merge.empty())
            for (int s = 0; s < nstrip(); s++)
                if (merge[s] > 1 && batch_rows[s] >= merge[s]) merge_strip(s);
    }

    // sheet strengths n x grad mu (Post.jl:22-23)
    void sheet_strengths(const std::vector<double> &m, std::vector<double> &out) const
    {
        out.assign(3 * (size_t)nt, 0.0);
        for (int i = 0; i < nt; i++) {
            P3 g;
            for (int k = 0; k < 3; k++) g = g + m[cs[3 * i + k]] * gam[3 * i + k];
            out[3 * i] = g.x;
            out[3 * i + 1] = g.y;
            out[3 * i + 2] = g.z;
        }
    }

    void step(vp_freewake_record *rec, double *mu_w_out);
};

namespace {

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void check(vp_ctx *ctx, int rc)
{
    if (rc == VP_OK) return;
    const std::string msg = vp_last_error(ctx);
    switch (rc) {
    case VP_E_ARGUMENT: throw std::invalid_argument(msg);
    case VP_E_DIMENSION: throw std::length_error(msg);
    case VP_E_SINGULAR: throw std::domain_error(msg);
    case VP_E_CONVERGENCE: throw vp::ConvergenceError(msg);
    default: throw std::runtime_error(msg);
    }
}

void flat(const Blobs &b, std::vector<double> &x, std::vector<double> &a, std::vector<double> &d)
{
    x.resize(3 * b.size());
    a.resize(3 * b.size());
    d.assign(b.d.begin(), b.d.end());
    for (size_t i = 0; i < b.size(); i++) {
        x[3 * i] = b.x[i].x, x[3 * i + 1] = b.x[i].y, x[3 * i + 2] = b.x[i].z;
        a[3 * i] = b.a[i].x, a[3 * i + 1] = b.a[i].y, a[3 * i + 2] = b.a[i].z;
    }
}

}  // namespace

// step!, step.jl:113-146
void vp_freewake::step(vp_freewake_record *rec, double *mu_w_out)
{
    const Elements we = vortex_elements();
    // the known elements' right hand side (explicit_rhs), then the solve with the factors kept
    double t0 = now();
    const int nseg_e = we.nseg();
    std::vector<double> sx(6 * (size_t)nseg_e);
    for (int j = 0; j < 2 * nseg_e; j++) {
        sx[3 * j] = we.seg[j].x;
        sx[3 * j + 1] = we.seg[j].y;
        sx[3 * j + 2] = we.seg[j].z;
    }
    flat(particles, bx, ba, bd);
    rhs.assign(nslot, 0.0);
    check(ctx, vp_explicit_rhs(ex, we.nsheet(), we.sheet_x.data(), we.sheet_mu.data(), nseg_e, sx.data(),
                               we.seg_g.data(), (int)particles.size(), bx.data(), ba.data(), bd.data(), rhs.data()));
    const double t_rhs = now() - t0;
    for (int i = 0; i < nslot; i++) rhs[i] += b0[i];
    const int n = nq + nb;
    R.assign(n, 0.0);
    C.mul_t(rhs.data(), R.data(), nq);
    X.assign(n, 0.0);
    int iters = 0;
    double resid = 0;
    check(ctx, vp_factor_solve(F, R.data(), 1, X.data(), &iters, &resid));
    std::vector<double> m(nslot), mw(nte);
    C.mul(X.data(), m.data());
    K.mul(X.data(), mw.data());
    const double lam0 = nb > 0 ? se[0] * X[nq] : 0.0;

    // step_loads, step.jl:149-163
    const double VV = dot(V, V);
    P3 Fp, Mp;
    cp.assign(nt, 0.0);
    for (int i = 0; i < nt; i++) {
        P3 g;
        double dm = 0;
        for (int k = 0; k < 3; k++) {
            const int sl = cs[3 * i + k];
            g = g + m[sl] * gradl[3 * i + k];
            if (nstep > 0) dm += (m[sl] - mu[sl]) / dt;
        }
        const P3 u = V - dot(nrm[i], V) * nrm[i] + g;
        cp[i] = 1 - (dot(u, u) + 2 * (dm / 3)) / VV;
        const P3 f = (-(cp[i] * area[i])) * nrm[i];
        Fp = Fp + f;
        Mp = Mp + cross(c[i], f);
    }
    P3 G;
    for (size_t k = 0; k < teseg.size() / 2; k++) {
        const int a = teseg[2 * k], b = teseg[2 * k + 1];
        G = G + ((mw[a] + mw[b]) / 2) * (te_x[b] - te_x[a]);
    }
    const P3 Fkj = (2 / VV) * cross(V, G);
    sheet_strengths(m, g3);
    const Elements pe = panel_elements(mw);
    Blobs wb;  // the wake's blobs: the trailing edge panel, the free rows, the particles
    free_wake_blobs(pe, core, wb);
    free_wake_blobs(we, core, wb);
    P3 Om, I;
    for (int i = 0; i < nt; i++) {  // the body's sheets as particles at the centroids
        const P3 al = area[i] * P3{g3[3 * i], g3[3 * i + 1], g3[3 * i + 2]};
        Om = Om + al;
        I = I + cross(c[i], al);
    }
    for (size_t i = 0; i < wb.size(); i++) {
        Om = Om + wb.a[i];
        I = I + cross(wb.x[i], wb.a[i]);
    }
    for (size_t i = 0; i < particles.size(); i++) {
        Om = Om + particles.a[i];
        I = I + cross(particles.x[i], particles.a[i]);
    }
    I = I / 2.0;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const P3 Fimp = nstep == 0 ? P3{nan, nan, nan} : (-(2 / VV)) * ((I - I_prev) / dt);
    I_prev = I;

    if (rows.empty()) {
        rows.push_back(edge);
        rmu.push_back(mw);
    }
    // wake_velocities, step.jl:166-173, at every free node and particle
    size_t nnode = 0;
    for (const auto &r : rows) nnode += r.size();
    const size_t np = nnode + particles.size();
    Xt.resize(3 * np);
    {
        size_t k = 0;
        for (const auto &r : rows)
            for (const P3 &p : r) {
                Xt[3 * k] = p.x, Xt[3 * k + 1] = p.y, Xt[3 * k + 2] = p.z;
                k++;
            }
        for (const P3 &p : particles.x) {
            Xt[3 * k] = p.x, Xt[3 * k + 1] = p.y, Xt[3 * k + 2] = p.z;
            k++;
        }
    }
    U.assign(3 * np, 0.0);
    t0 = now();
    if (advection == 1) {
        for (size_t k = 0; k < np; k++) U[3 * k] = V.x, U[3 * k + 1] = V.y, U[3 * k + 2] = V.z;
    } else if (np > 0) {
        Blobs all = wb;
        all.append(particles);
        std::vector<double> ax, aa, ad;
        flat(all, ax, aa, ad);
        const double base[3] = {V.x, V.y, V.z};
        if (tree.theta > 0.0) {  // the body in the graph, the wake's blobs by the treecode
            check(ctx, vp_wake_field_eval(field, sigma.data(), g3.data(), (int)np, Xt.data(), 0, nullptr, nullptr,
                                          nullptr, base, U.data()));
            Ub.assign(3 * np, 0.0);
            check(ctx, vp_blob_velocities_tree(ctx, (int)all.size(), ax.data(), aa.data(), ad.data(), (int)np,
                                               Xt.data(), &tree, Ub.data(), nullptr));
            for (size_t k = 0; k < 3 * np; k++) U[k] += Ub[k];
        } else {
            check(ctx, vp_wake_field_eval(field, sigma.data(), g3.data(), (int)np, Xt.data(), (int)all.size(),
                                          ax.data(), aa.data(), ad.data(), base, U.data()));
        }
    }
    const double t_vel = now() - t0;
    double umax = 0;
    for (size_t k = 0; k < np; k++) umax = std::max(umax, norm(ld(&U[3 * k])));
    // forward Euler
    {
        size_t k = 0;
        for (auto &r : rows)
            for (P3 &p : r) {
                p = p + dt * ld(&U[3 * k]);
                k++;
            }
        for (P3 &p : particles.x) {
            p = p + dt * ld(&U[3 * k]);
            k++;
        }
    }
    // shed the new row at the trailing edge panel's downstream edge, convert beyond keep_rows
    rows.insert(rows.begin(), edge);
    rmu.insert(rmu.begin(), mw);
    while (rows.size() - 1 > (size_t)keep_rows) convert_oldest();
    // delete_far!, step.jl:190-195
    if (std::isfinite(delete_beyond)) {
        const P3 d = V / norm(V);
        double lim = -std::numeric_limits<double>::infinity();
        for (const P3 &p : edge) lim = std::max(lim, dot(p, d));
        lim += delete_beyond * span;
        std::vector<char> keep(particles.size());
        for (size_t i = 0; i < keep.size(); i++) keep[i] = dot(particles.x[i], d) <= lim;
        keep_particles(keep);
    }
    rec->n = (int)nstep;
    rec->t = t;
    rec->rows = (int)rows.size() - 1;
    rec->particles = (int)particles.size();
    const P3 *vs[6] = {&Fp, &Mp, &Fkj, &Fimp, &Om, &I};
    double *outs[6] = {rec->F_pressure, rec->M_pressure, rec->F_kj, rec->F_impulse, rec->Omega, rec->I};
    for (int k = 0; k < 6; k++) outs[k][0] = vs[k]->x, outs[k][1] = vs[k]->y, outs[k][2] = vs[k]->z;
    rec->lambda = lam0;
    rec->umax = umax;
    rec->t_rhs = t_rhs;
    rec->t_vel = t_vel;
    rec->iters = iters;
    if (mu_w_out) std::memcpy(mu_w_out, mw.data(), sizeof(double) * nte);
    mu = std::move(m);
    mu_w = std::move(mw);
    nstep++;
    t += dt;
}

extern "C" {

int vp_freewake_create(vp_ctx *ctx, const vp_surface *g, const vp_freewake_setup *su, vp_freewake **out)
{
    if (!ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = &ctx->c;
    auto bad = [&](const char *m) {
        c->err = m;
        return VP_E_ARGUMENT;
    };
    if (!g || !su || !out || !g->x || !g->tris || !g->corner_slot) return bad("null argument");
    if (!su->vinf || !su->opts || !su->factor || !su->b0 || !su->C_rowptr || !su->C_col || !su->C_val ||
        !su->K_rowptr || !su->K_col || !su->K_val || !su->te_x || !su->edge || (su->nteseg > 0 && !su->teseg) ||
        (su->nstrip > 0 && !su->strips) || (su->nfirst > 0 && !su->first) || (su->nlast > 0 && !su->last) ||
        !su->panel || !su->panel_te)
        return bad("null argument in the set-up");
    if (!(su->dt > 0) || su->keep_rows < 0 || su->core_mode < 0 || su->core_mode > 2 || su->advection < 0 ||
        su->advection > 1 || su->nte <= 0 || su->nq <= 0)
        return bad("bad option in the set-up");
    if (su->tree_theta != 0.0 && !(su->tree_theta > 0.0 && su->tree_theta < 1.0 && su->tree_degree >= 1 &&
                                   su->tree_degree <= 7))
        return bad("tree_theta must be 0 (direct sums) or lie in (0, 1), with tree_degree in 1..7");
    *out = nullptr;
    vp_freewake *h = new (std::nothrow) vp_freewake;
    if (!h) return VP_E_DEVICE;
    int rc = VP_OK;
    try {
        h->ctx = ctx;
        h->F = su->factor;
        h->nt = g->nt;
        h->nslot = g->nslot;
        h->nq = su->nq;
        int nfac = 0;
        check(ctx, vp_factor_info(su->factor, &nfac, nullptr));
        h->nb = nfac - su->nq;
        if (h->nb < 0) throw std::length_error("the factorization is smaller than nq");
        h->se.assign(std::max(h->nb, 1), 0.0);
        check(ctx, vp_factor_info(su->factor, nullptr, h->se.data()));
        h->V = ld(su->vinf);
        h->dt = su->dt;
        h->advection = su->advection;
        h->tree = vp_tree_options{su->tree_theta, su->tree_degree, su->field_precision};
        h->delete_beyond = su->delete_beyond;
        h->span = su->span;
        h->keep_rows = su->keep_rows;
        h->core.mode = su->core_mode;
        h->core.value = su->core_value;
        h->core.overlap = su->overlap;
        h->core.factor = su->core_factor;
        h->core.floor = su->core_floor;
        // the body's elements (build_elements, src/elements/flat_triangle.jl:53) and sigma = -V.n
        const int nt = g->nt;
        h->nrm.resize(nt);
        h->c.resize(nt);
        h->area.resize(nt);
        h->sigma.resize(nt);
        h->gradl.resize(3 * (size_t)nt);
        h->gam.resize(3 * (size_t)nt);
        h->cs.assign(g->corner_slot, g->corner_slot + 3 * (size_t)nt);
        for (int i = 0; i < nt; i++) {
            const P3 a = ld(g->x + 3 * g->tris[3 * i]), b = ld(g->x + 3 * g->tris[3 * i + 1]),
                     cc = ld(g->x + 3 * g->tris[3 * i + 2]);
            const P3 nA = cross(b - a, cc - a);
            const double A2 = norm(nA);
            if (!(A2 > 0)) throw std::invalid_argument("degenerate triangle");
            const P3 nn = nA / A2, e[3] = {cc - b, a - cc, b - a};
            h->nrm[i] = nn;
            h->area[i] = A2 / 2;
            h->c[i] = (a + b + cc) / 3.0;
            for (int k = 0; k < 3; k++) {
                h->gradl[3 * i + k] = cross(nn, e[k]) / A2;
                h->gam[3 * i + k] = (-1.0 * e[k]) / A2;
            }
            h->sigma[i] = -dot(nn, h->V);
        }
        h->C = copy_csr(g->nslot, su->C_rowptr, su->C_col, su->C_val);
        h->nte = su->nte;
        h->K = copy_csr(su->nte, su->K_rowptr, su->K_col, su->K_val);
        h->b0.assign(su->b0, su->b0 + g->nslot);
        for (int k = 0; k < su->nte; k++) {
            h->edge.push_back(ld(su->edge + 3 * k));
            h->te_x.push_back(ld(su->te_x + 3 * k));
        }
        h->teseg.assign(su->teseg, su->teseg + 2 * (size_t)std::max(su->nteseg, 0));
        h->strips.assign(su->strips, su->strips + 2 * (size_t)std::max(su->nstrip, 0));
        if (su->merge) h->merge.assign(su->merge, su->merge + std::max(su->nstrip, 0));
        if (su->nfirst > 0) h->first.assign(su->first, su->first + su->nfirst);
        if (su->nlast > 0) h->last.assign(su->last, su->last + su->nlast);
        for (int v : h->strips)
            if (v < 0 || v >= su->nte) throw std::out_of_range("strip node out of range");
        const vp_wake &pw = *su->panel;
        for (int k = 0; k < pw.nw; k++) h->panel_x.push_back(ld(pw.x + 3 * k));
        h->panel_te.assign(su->panel_te, su->panel_te + pw.nw);
        h->panel_tris.assign(pw.tris, pw.tris + 3 * (size_t)pw.ntri);
        if (pw.nfil > 0) h->panel_fil.assign(pw.filaments, pw.filaments + 2 * (size_t)pw.nfil);
        h->batch_rows.assign(std::max(su->nstrip, 0), 0);
        h->batch_length.assign(std::max(su->nstrip, 0), 0.0);
        h->batch_width.assign(std::max(su->nstrip, 0), 0.0);
        h->mu_line.assign(su->nte, 0.0);
        h->mu.assign(g->nslot, 0.0);
        h->mu_w.assign(su->nte, 0.0);
        h->cp.assign(nt, 0.0);
        check(ctx, vp_explicit_create(ctx, g, su->opts, &h->ex));
        vp_surface body = *g;
        body.corner_slot = nullptr;
        check(ctx, vp_wake_field_create(ctx, &body, su->opts->kernel_far, su->field_precision, &h->field));
    } catch (const std::invalid_argument &e) {
        c->err = e.what();
        rc = VP_E_ARGUMENT;
    } catch (const std::out_of_range &e) {
        c->err = e.what();
        rc = VP_E_ARGUMENT;
    } catch (const std::length_error &e) {
        c->err = e.what();
        rc = VP_E_DIMENSION;
    } catch (const std::exception &e) {
        if (c->err.empty()) c->err = e.what();
        rc = VP_E_DEVICE;
    }
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    c->err.clear();
    return VP_OK;
}

int vp_freewake_step(vp_freewake *h, vp_freewake_record *rec, double *mu_w)
{
    if (!h || !h->ctx) return VP_E_ARGUMENT;
    vp::Ctx *c = &h->ctx->c;
    if (!rec) {
        c->err = "null argument";
        return VP_E_ARGUMENT;
    }
    try {
        h->step(rec, mu_w);
    } catch (const std::invalid_argument &e) {
        c->err = e.what();
        return VP_E_ARGUMENT;
    } catch (const std::length_error &e) {
        c->err = e.what();
        return VP_E_DIMENSION;
    } catch (const vp::ConvergenceError &e) {
        c->err = e.what();
        return VP_E_CONVERGENCE;
    } catch (const std::exception &e) {
        c->err = e.what();
        return VP_E_DEVICE;
    }
    c->err.clear();
    return VP_OK;
}

int vp_freewake_counts(const vp_freewake *h, long long *counts)
{
    if (!h || !counts) return VP_E_ARGUMENT;
    counts[0] = (long long)h->rows.size();
    counts[1] = (long long)h->particles.size();
    counts[2] = h->nte;
    counts[3] = h->nstrip();
    counts[4] = h->nstep;
    counts[5] = h->conversions;
    return VP_OK;
}

int vp_freewake_read(const vp_freewake *h, double *rows_x, double *rows_mu, double *mu_line, double *px, double *palpha,
                     double *pdelta, long long *tag, int *group, long long *batch_rows, double *batch_length,
                     double *batch_width, double *mu, double *mu_w, double *cp, double *I_prev, double *t)
{
    if (!h) return VP_E_ARGUMENT;
    const int nte = h->nte;
    for (size_t r = 0; r < h->rows.size(); r++)
        for (int k = 0; k < nte; k++) {
            if (rows_x) {
                double *o = rows_x + 3 * ((size_t)r * nte + k);
                o[0] = h->rows[r][k].x, o[1] = h->rows[r][k].y, o[2] = h->rows[r][k].z;
            }
            if (rows_mu) rows_mu[(size_t)r * nte + k] = h->rmu[r][k];
        }
    if (mu_line) std::memcpy(mu_line, h->mu_line.data(), sizeof(double) * nte);
    for (size_t i = 0; i < h->particles.size(); i++) {
        if (px) px[3 * i] = h->particles.x[i].x, px[3 * i + 1] = h->particles.x[i].y, px[3 * i + 2] = h->particles.x[i].z;
        if (palpha)
            palpha[3 * i] = h->particles.a[i].x, palpha[3 * i + 1] = h->particles.a[i].y,
                   palpha[3 * i + 2] = h->particles.a[i].z;
        if (pdelta) pdelta[i] = h->particles.d[i];
        if (tag) tag[i] = h->tag[i];
        if (group) group[i] = h->group[i];
    }
    for (int s = 0; s < h->nstrip(); s++) {
        if (batch_rows) batch_rows[s] = h->batch_rows[s];
        if (batch_length) batch_length[s] = h->batch_length[s];
        if (batch_width) batch_width[s] = h->batch_width[s];
    }
    if (mu) std::memcpy(mu, h->mu.data(), sizeof(double) * h->nslot);
    if (mu_w) std::memcpy(mu_w, h->mu_w.data(), sizeof(double) * nte);
    if (cp) std::memcpy(cp, h->cp.data(), sizeof(double) * h->nt);
    if (I_prev) I_prev[0] = h->I_prev.x, I_prev[1] = h->I_prev.y, I_prev[2] = h->I_prev.z;
    if (t) *t = h->t;
    return VP_OK;
}

void vp_freewake_destroy(vp_freewake *h) { delete h; }

}  // extern "C"
