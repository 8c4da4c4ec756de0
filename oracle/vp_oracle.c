/*
 * FP64 CPU oracle of VortexPanel.jl's Galerkin hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * never as the thing measured or shipped.  It restates, function by function, the reference's algorithm in plain
 * C99 (gcc, -ffp-contract=off so that no multiply-add is fused, as Julia does not fuse them), in the reference's
 * operation order, so that tier classification and rule sizes are reproducible bit for bit and matrix entries to
 * rounding.  Each function cites the reference file:line it follows (paths relative to the reference checkout).
 *
 * Pinned by (see tests/): closed forms (Gauss flux = area, in-plane jump 1/2, edge 1/4, vertex alpha/4pi, far-field
 * continuity and 4th order), independent quadrature golden vectors (tests/golden/, mpmath), rule moments and log
 * integrals, the sphere's exact solution, assembly invariants (row sums, pair counts), and LAPACK for the solve.
 * Functions without such a check say "parity unpinned" in their comment.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const double *p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 dvs(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double sumsq(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double norm(v3 a) { return sqrt(sumsq(a)); }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ kernels ---------------------------------- */

/* SourceTriangle, src/kernels/source_triangle.jl:8-29 */
typedef struct {
    v3 v[3], n, c;
    double area, diam;
    v3 nu[3];
    double len[3];
} src_t;

static int make_src(v3 a, v3 b, v3 c, src_t *s)
{
    v3 nA = cross(sub(b, a), sub(c, a));
    double A2 = norm(nA);
    if (!(A2 > 0)) return -1; /* ArgumentError("degenerate triangle") */
    v3 n = dvs(nA, A2);
    s->len[0] = norm(sub(b, a));
    s->len[1] = norm(sub(c, b));
    s->len[2] = norm(sub(a, c));
    s->nu[0] = dvs(cross(sub(b, a), n), s->len[0]);
    s->nu[1] = dvs(cross(sub(c, b), n), s->len[1]);
    s->nu[2] = dvs(cross(sub(a, c), n), s->len[2]);
    s->v[0] = a; s->v[1] = b; s->v[2] = c;
    s->n = n;
    s->c = dvs(add(add(a, b), c), 3.0);
    s->area = A2 / 2;
    s->diam = dmax(dmax(s->len[0], s->len[1]), s->len[2]);
    return 0;
}

/* edge_log, src/kernels/source_triangle.jl:39-44 */
static inline double edge_log(v3 ri, v3 rj, double d)
{
    double a = norm(ri), b = norm(rj);
    double p = dot(ri, rj);
    double s2d2 = p < 0 ? 2 * sumsq(cross(ri, rj)) / (a * b - p) : 2 * (a * b + p);
    double t = a + b + d;
    return log(t * t / s2d2);
}

/* inplane_fraction, src/kernels/source_triangle.jl:52-64 */
static double inplane_fraction(const src_t *t, v3 x, double tol)
{
    v3 a = t->v[0], b = t->v[1], c = t->v[2];
    double A2 = 2 * t->area;
    double l[3];
    l[0] = dot(cross(sub(b, x), sub(c, x)), t->n) / A2;
    l[1] = dot(cross(sub(c, x), sub(a, x)), t->n) / A2;
    l[2] = dot(cross(sub(a, x), sub(b, x)), t->n) / A2;
    for (int k = 0; k < 3; k++) if (l[k] < -tol) return 0.0;
    int nzero = 0;
    for (int k = 0; k < 3; k++) if (fabs(l[k]) <= tol) nzero++;
    if (nzero == 0) return 1.0;
    if (nzero == 1) return 0.5;
    int k = 0;
    while (k < 3 && !(fabs(l[k]) > tol)) k++;
    if (k == 3) k = 0; /* degenerate: all three near zero; the reference's findfirst would fail here */
    v3 e1 = sub(t->v[(k + 1) % 3], t->v[k]), e2 = sub(t->v[(k + 2) % 3], t->v[k]);
    double ca = dot(e1, e2) / (norm(e1) * norm(e2));
    if (ca < -1) ca = -1;
    if (ca > 1) ca = 1;
    return acos(ca) / (2 * PI);
}

/* source_velocity_exact, src/kernels/source_triangle.jl:75-89 */
static v3 sv_exact(const src_t *t, v3 x, int side, double plane_tol)
{
    v3 ra = sub(x, t->v[0]), rb = sub(x, t->v[1]), rc = sub(x, t->v[2]);
    v3 wt = add(add(scl(edge_log(ra, rb, t->len[0]), t->nu[0]), scl(edge_log(rb, rc, t->len[1]), t->nu[1])),
                scl(edge_log(rc, ra, t->len[2]), t->nu[2]));
    double z = dot(sub(x, t->c), t->n);
    double Om;
    if (fabs(z) <= plane_tol * t->diam) {
        Om = side * 2 * PI * inplane_fraction(t, x, plane_tol);
    } else {
        double la = norm(ra), lb = norm(rb), lc = norm(rc);
        double N = dot(ra, cross(rb, rc));
        double D = la * lb * lc + dot(ra, rb) * lc + dot(ra, rc) * lb + dot(rb, rc) * la;
        Om = 2 * atan2(N, D);
    }
    return dvs(add(wt, scl(Om, t->n)), 4 * PI);
}

/* source_velocity_far, src/kernels/source_triangle.jl:100-119 */
static v3 sv_far(const src_t *t, v3 x)
{
    v3 r = sub(x, t->c);
    double R2 = dot(r, r);
    double R = sqrt(R2);
    double iR2 = 1 / R2, iR3 = iR2 / R, iR5 = iR3 * iR2, iR7 = iR5 * iR2, iR9 = iR7 * iR2;
    v3 w = scl(iR3, r);
    for (int k = 0; k < 3; k++) {
        v3 d = sub(t->v[k], t->c);
        double dr = dot(d, r), dd = dot(d, d);
        /* quad = 15 dr^2 iR7 r - 6 dr iR5 d - 3 dd iR5 r */
        v3 quad = sub(sub(scl(15 * (dr * dr) * iR7, r), scl(6 * dr * iR5, d)), scl(3 * dd * iR5, r));
        /* oct = 45 dr^2 iR7 d - 105 dr^3 iR9 r - 9 dd iR5 d + 45 dd dr iR7 r */
        v3 oct = add(sub(sub(scl(45 * (dr * dr) * iR7, d), scl(105 * (dr * dr * dr) * iR9, r)), scl(9 * dd * iR5, d)),
                     scl(45 * dd * dr * iR7, r));
        w = add(w, sub(dvs(quad, 24.0), dvs(oct, 180.0)));
    }
    return scl(t->area / (4 * PI), w);
}

/* source_velocity, src/kernels/source_triangle.jl:134-138 */
static inline v3 sv(const src_t *t, v3 x, int side, double far, double plane_tol)
{
    v3 r = sub(x, t->c);
    double fd = far * t->diam;
    if (dot(r, r) > fd * fd) return sv_far(t, x);
    return sv_exact(t, x, side, plane_tol);
}

/* segment_velocity, src/kernels/segment.jl:15-25 */
static void seg_vel(v3 a, v3 b, v3 x, v3 *va, v3 *vb)
{
    v3 ra = sub(x, a), rb = sub(x, b);
    double Ra = norm(ra), Rb = norm(rb);
    v3 c = cross(ra, rb);
    double p = dot(ra, rb);
    double den = p >= 0 ? Ra * Rb + p : sumsq(c) / (Ra * Rb - p);
    if (den == 0 || Ra == 0 || Rb == 0) { *va = mk(0, 0, 0); *vb = mk(0, 0, 0); return; }
    v3 f = dvs(c, 4 * PI * den);
    *va = dvs(f, Ra);
    *vb = dvs(f, Rb);
}

/* blob_velocity, src/kernels/blob.jl:21-25 */
static inline v3 blob_vel(v3 x, v3 xp, v3 al, double delta)
{
    v3 r = sub(x, xp);
    double d2 = dot(r, r) + delta * delta;
    return dvs(cross(al, r), 4 * PI * d2 * sqrt(d2));
}

/* exported kernel entry points (for the tests) */
int vpo_source_velocity(const double *a, const double *b, const double *c, const double *x, int side, double far,
                        double plane_tol, int mode, double *w)
{
    src_t s;
    if (make_src(ld3(a), ld3(b), ld3(c), &s)) return -1;
    v3 r = mode == 1 ? sv_exact(&s, ld3(x), side, plane_tol) : mode == 2 ? sv_far(&s, ld3(x)) : sv(&s, ld3(x), side, far, plane_tol);
    st3(w, r);
    return 0;
}

double vpo_inplane_fraction(const double *a, const double *b, const double *c, const double *x, double tol)
{
    src_t s;
    if (make_src(ld3(a), ld3(b), ld3(c), &s)) return NAN;
    return inplane_fraction(&s, ld3(x), tol);
}

void vpo_segment_velocity(const double *a, const double *b, const double *x, double *va, double *vb)
{
    v3 p, q;
    seg_vel(ld3(a), ld3(b), ld3(x), &p, &q);
    st3(va, p);
    st3(vb, q);
}

void vpo_blob_velocity(const double *x, const double *xp, const double *al, double delta, double *u)
{
    st3(u, blob_vel(ld3(x), ld3(xp), ld3(al), delta));
}

/* ------------------------------------------------------------------- rules ---------------------------------- */

typedef struct { int n, cap; double *l; /* 3 per point */ double *w; } rule_t;

static void rule_init(rule_t *r) { r->n = 0; r->cap = 0; r->l = NULL; r->w = NULL; }
static void rule_free(rule_t *r) { free(r->l); free(r->w); rule_init(r); }
/* empty a rule for reuse, keeping its buffers (the per-thread scratch rule of every pair) */
static void rule_clear(rule_t *r) { r->n = 0; }
static void rule_push(rule_t *r, double l1, double l2, double l3, double w)
{
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 64;
        r->l = (double *)realloc(r->l, sizeof(double) * 3 * r->cap);
        r->w = (double *)realloc(r->w, sizeof(double) * r->cap);
    }
    r->l[3 * r->n] = l1; r->l[3 * r->n + 1] = l2; r->l[3 * r->n + 2] = l3;
    r->w[r->n] = w;
    r->n++;
}

/* gauss_legendre, src/elements/rules.jl:12-40 (nodes and weights on [0, 1]) */
static void gauss_legendre(int n, double *x, double *w)
{
    for (int i = 1; i <= (n + 1) / 2; i++) {
        double z = cos(PI * (i - 0.25) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = ((2.0 * j - 1) * z * p1 - (j - 1.0) * p2) / j;
                p2 = p1; p1 = p3;
            }
            dp = n * (z * p1 - p2) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p1 = 1.0, p2 = 0.0;
        for (int j = 1; j <= n; j++) {
            double p3 = ((2.0 * j - 1) * z * p1 - (j - 1.0) * p2) / j;
            p2 = p1; p1 = p3;
        }
        dp = n * (z * p1 - p2) / (z * z - 1);
        x[i - 1] = (1 - z) / 2;
        x[n - i] = (1 + z) / 2;
        w[i - 1] = w[n - i] = 1 / ((1 - z * z) * dp * dp);
    }
}

/* graded_gauss, src/elements/rules.jl:48-63; returns the point count */
static int graded_gauss(int levels, double ratio, int order, double *x, double *w)
{
    double br[64];
    int nb = 0;
    br[nb++] = 0.0;
    for (int k = levels; k >= 1; k--) br[nb++] = pow(ratio, k);
    br[nb++] = 1.0;
    int m = 0;
    double xg[64], wg[64];
    for (int k = 1; k <= nb - 1; k++) {
        int ng = order + k - 1;
        gauss_legendre(ng, xg, wg);
        double a = br[k - 1], b = br[k];
        for (int i = 0; i < ng; i++) {
            if (k == 1) { x[m] = b * (xg[i] * xg[i]); w[m] = 2 * b * xg[i] * wg[i]; }
            else { x[m] = a + (b - a) * xg[i]; w[m] = (b - a) * wg[i]; }
            m++;
        }
    }
    return m;
}

/* graded_gauss_both, src/elements/rules.jl:66-69 */
static int graded_gauss_both(int levels, double ratio, int order, double *x, double *w)
{
    double xs[2048], ws[2048];
    int m = graded_gauss(levels, ratio, order, xs, ws);
    for (int i = 0; i < m; i++) { x[i] = xs[i] / 2; w[i] = ws[i] / 2; }
    for (int i = 0; i < m; i++) { x[m + i] = 1 - xs[m - 1 - i] / 2; w[m + i] = ws[m - 1 - i] / 2; }
    return 2 * m;
}

/* triangle_rule, src/elements/rules.jl:76-92 */
static int triangle_rule(int n, rule_t *r)
{
    rule_init(r);
    if (n == 1) { rule_push(r, 1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0); return 0; }
    if (n == 3) {
        double a = 2.0 / 3, b = 1.0 / 6;
        rule_push(r, a, b, b, 1.0 / 3); rule_push(r, b, a, b, 1.0 / 3); rule_push(r, b, b, a, 1.0 / 3);
        return 0;
    }
    if (n == 6) {
        double a = 0.445948490915965, b = 0.091576213509771;
        double wa = 0.223381589678011, wb = 0.109951743655322;
        rule_push(r, 1 - 2 * a, a, a, wa); rule_push(r, a, 1 - 2 * a, a, wa); rule_push(r, a, a, 1 - 2 * a, wa);
        rule_push(r, 1 - 2 * b, b, b, wb); rule_push(r, b, 1 - 2 * b, b, wb); rule_push(r, b, b, 1 - 2 * b, wb);
        return 0;
    }
    if (n == 7) {
        double s = sqrt(15.0);
        double a = (6 - s) / 21, b = (6 + s) / 21;
        double wa = (155 - s) / 1200, wb = (155 + s) / 1200;
        rule_push(r, 1.0 / 3, 1.0 / 3, 1.0 / 3, 9.0 / 40);
        rule_push(r, 1 - 2 * a, a, a, wa); rule_push(r, a, 1 - 2 * a, a, wa); rule_push(r, a, a, 1 - 2 * a, wa);
        rule_push(r, 1 - 2 * b, b, b, wb); rule_push(r, b, 1 - 2 * b, b, wb); rule_push(r, b, b, 1 - 2 * b, wb);
        return 0;
    }
    return -1; /* ArgumentError: no n-point triangle rule */
}

/* rotate_rule, src/elements/rules.jl:97-102 (base corner 1 becomes corner k, k = 1..3) */
static void rotate_rule(const rule_t *r, int k, rule_t *o)
{
    rule_init(o);
    /* mod1(2-k,3), mod1(3-k,3), mod1(4-k,3) as 0-based indices */
    int idx[3] = {((1 - k) % 3 + 3) % 3, ((2 - k) % 3 + 3) % 3, ((3 - k) % 3 + 3) % 3};
    for (int i = 0; i < r->n; i++)
        rule_push(o, r->l[3 * i + idx[0]], r->l[3 * i + idx[1]], r->l[3 * i + idx[2]], r->w[i]);
}

/* edge_graded_rule, src/elements/rules.jl:108-118 */
static void edge_graded_rule(int levels, double ratio, int order, rule_t *r)
{
    double ts[2048], wt[2048], ss[4096], ws[4096];
    int nt = graded_gauss(levels, ratio, order, ts, wt);
    int lv2 = levels / 2 > 1 ? levels / 2 : 1;
    int ns = graded_gauss_both(lv2, ratio, order, ss, ws);
    rule_init(r);
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < ns; j++)
            rule_push(r, (1 - ts[i]) * (1 - ss[j]), (1 - ts[i]) * ss[j], ts[i], 2 * (1 - ts[i]) * wt[i] * ws[j]);
}

/* vertex_graded_rule, src/elements/rules.jl:125-134 */
static void vertex_graded_rule(int levels, double ratio, int order, rule_t *r)
{
    double rs[2048], wr[2048], ss[256], ws[256];
    int nr = graded_gauss(levels, ratio, order, rs, wr);
    int ns = order + levels;
    gauss_legendre(ns, ss, ws);
    rule_init(r);
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < ns; j++)
            rule_push(r, 1 - rs[i], rs[i] * (1 - ss[j]), rs[i] * ss[j], 2 * rs[i] * wr[i] * ws[j]);
}

/* self_rule, src/elements/rules.jl:142-155 */
static void self_rule(int levels, double ratio, int order, rule_t *r)
{
    rule_t base;
    edge_graded_rule(levels, ratio, order, &base);
    rule_init(r);
    const double c = 1.0 / 3;
    for (int k = 0; k < 3; k++) {
        double ek[3] = {0, 0, 0}, ek1[3] = {0, 0, 0};
        ek[k] = 1.0;
        ek1[(k + 1) % 3] = 1.0;
        for (int i = 0; i < base.n; i++) {
            const double *m = base.l + 3 * i;
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * ek[j] + m[1] * ek1[j]) + m[2] * c;
            rule_push(r, l[0], l[1], l[2], base.w[i] / 3);
        }
    }
    rule_free(&base);
}

/* TierRules, src/elements/rules.jl:161-174 */
typedef struct { rule_t far, near, self, edge[3], vertex[3]; } tier_rules_t;

/* AssemblyOptions, src/assembly/assemble.jl:26-41 (test: 0 hat, 1 onering, 2 dual) */
typedef struct {
    int test;
    double far_ratio;
    int far_points, near_points;
    double near_ratio;
    int near_depth, touch_levels;
    double touch_ratio;
    int touch_order;
    double corner_angle, contact_rel, kernel_far;
} vpo_opts;

static int tier_rules(const vpo_opts *o, tier_rules_t *R)
{
    if (triangle_rule(o->far_points, &R->far)) return -1;
    if (triangle_rule(o->near_points, &R->near)) return -1;
    self_rule(o->touch_levels, o->touch_ratio, o->touch_order, &R->self);
    rule_t e, v;
    edge_graded_rule(o->touch_levels, o->touch_ratio, o->touch_order, &e);
    vertex_graded_rule(o->touch_levels, o->touch_ratio, o->touch_order, &v);
    for (int k = 0; k < 3; k++) { rotate_rule(&e, k + 1, &R->edge[k]); rotate_rule(&v, k + 1, &R->vertex[k]); }
    rule_free(&e);
    rule_free(&v);
    return 0;
}

static void tier_rules_free(tier_rules_t *R)
{
    rule_free(&R->far); rule_free(&R->near); rule_free(&R->self);
    for (int k = 0; k < 3; k++) { rule_free(&R->edge[k]); rule_free(&R->vertex[k]); }
}

/* rule export for tests: kind 0 far,1 near,2 self,3..5 edge k,6..8 vertex k; returns count, fills if buffers given */
int vpo_rule(const vpo_opts *o, int kind, double *lam, double *w)
{
    tier_rules_t R;
    if (tier_rules(o, &R)) return -1;
    rule_t *r = kind == 0 ? &R.far : kind == 1 ? &R.near : kind == 2 ? &R.self : kind <= 5 ? &R.edge[kind - 3] : &R.vertex[kind - 6];
    int n = r->n;
    if (lam) memcpy(lam, r->l, sizeof(double) * 3 * n);
    if (w) memcpy(w, r->w, sizeof(double) * n);
    tier_rules_free(&R);
    return n;
}

int vpo_gauss_legendre(int n, double *x, double *w) { gauss_legendre(n, x, w); return n; }
int vpo_graded_gauss(int levels, double ratio, int order, double *x, double *w) { return graded_gauss(levels, ratio, order, x, w); }

/* ---------------------------------------------------------------- elements ---------------------------------- */

/* FlatTriangle, src/elements/flat_triangle.jl:12-31 */
typedef struct {
    v3 x[3], n, c;
    double area, diam;
    v3 gl[3], g[3];
    src_t src;
} el_t;

static int make_el(v3 a, v3 b, v3 c, el_t *e)
{
    if (make_src(a, b, c, &e->src)) return -1;
    e->x[0] = a; e->x[1] = b; e->x[2] = c;
    e->n = e->src.n;
    e->area = e->src.area;
    e->c = e->src.c;
    e->diam = e->src.diam;
    double A2 = 2 * e->area;
    v3 ed[3] = {sub(c, b), sub(a, c), sub(b, a)};
    for (int k = 0; k < 3; k++) {
        e->gl[k] = dvs(cross(e->n, ed[k]), A2);
        e->g[k] = dvs(scl(-1.0, ed[k]), A2);
    }
    return 0;
}

/* point, src/elements/flat_triangle.jl:36 */
static inline v3 el_point(const el_t *e, const double *l)
{
    return add(add(scl(l[0], e->x[0]), scl(l[1], e->x[1])), scl(l[2], e->x[2]));
}

/* reach, src/elements/flat_triangle.jl:56 */
static inline double reach(const el_t *e)
{
    return dmax(dmax(norm(sub(e->x[0], e->c)), norm(sub(e->x[1], e->c))), norm(sub(e->x[2], e->c)));
}

/* distance_to (Ericson 5.1.5), src/elements/flat_triangle.jl:59-78 */
static double tri_distance(const v3 *x, v3 p)
{
    v3 a = x[0], b = x[1], c = x[2];
    v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return norm(ap);
    v3 bp = sub(p, b);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return norm(bp);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return norm(sub(p, add(a, scl(d1 / (d1 - d3), ab))));
    v3 cp = sub(p, c);
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return norm(cp);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return norm(sub(p, add(a, scl(d2 / (d2 - d6), ac))));
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0)
        return norm(sub(p, add(b, scl((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b)))));
    double denom = 1 / (va + vb + vc);
    return norm(sub(p, add(add(a, scl(vb * denom, ab)), scl(vc * denom, ac))));
}

/* segment_projection, src/assembly/assemble.jl:145-149 */
static inline double seg_projection(v3 p, v3 a, v3 c, double *tout)
{
    v3 e = sub(c, a);
    double t = dot(sub(p, a), e) / dot(e, e);
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    if (tout) *tout = t;
    return norm(sub(p, add(a, scl(t, e))));
}

/* a source of the pair integrals: a triangle, a filament segment (distance_to at assemble.jl:157) or a blob
 * (distance_to at src/assembly/explicit.jl:20) */
typedef struct { int kind; const el_t *tri; v3 a, b; } srcref_t;

static inline double src_distance(const srcref_t *s, v3 p)
{
    if (s->kind == 0) return tri_distance(s->tri->x, p);
    if (s->kind == 1) return seg_projection(p, s->a, s->b, NULL);
    return norm(sub(p, s->a));
}

/* ------------------------------------------------------------ rule builders ---------------------------------- */

static const double CORNERS[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

/* append_piece!, src/assembly/assemble.jl:85-91 */
static void append_piece(rule_t *out, const rule_t *r, const double b[3][3], double frac)
{
    for (int i = 0; i < r->n; i++) {
        const double *m = r->l + 3 * i;
        double l[3];
        for (int j = 0; j < 3; j++) l[j] = (m[0] * b[0][j] + m[1] * b[1][j]) + m[2] * b[2][j];
        rule_push(out, l[0], l[1], l[2], r->w[i] * frac);
    }
}

/* quarters, src/assembly/assemble.jl:79-82 */
static void quarters(const double b[3][3], double q[4][3][3])
{
    double m12[3], m23[3], m31[3];
    for (int j = 0; j < 3; j++) {
        m12[j] = (b[0][j] + b[1][j]) / 2;
        m23[j] = (b[1][j] + b[2][j]) / 2;
        m31[j] = (b[2][j] + b[0][j]) / 2;
    }
    for (int j = 0; j < 3; j++) {
        q[0][0][j] = b[0][j]; q[0][1][j] = m12[j]; q[0][2][j] = m31[j];
        q[1][0][j] = m12[j];  q[1][1][j] = b[1][j]; q[1][2][j] = m23[j];
        q[2][0][j] = m31[j];  q[2][1][j] = m23[j]; q[2][2][j] = b[2][j];
        q[3][0][j] = m12[j];  q[3][1][j] = m23[j]; q[3][2][j] = m31[j];
    }
}

/* near_pieces!, src/assembly/assemble.jl:131-142 */
static void near_pieces(rule_t *out, const el_t *T, const srcref_t *S, const rule_t *base, double ratio, int depth,
                        const double b[3][3], double frac, int level)
{
    v3 p1 = el_point(T, b[0]), p2 = el_point(T, b[1]), p3 = el_point(T, b[2]);
    double size = dmax(dmax(norm(sub(p2, p1)), norm(sub(p3, p2))), norm(sub(p1, p3)));
    if (level < depth && src_distance(S, dvs(add(add(p1, p2), p3), 3.0)) < ratio * size) {
        double q[4][3][3];
        quarters(b, q);
        for (int i = 0; i < 4; i++) near_pieces(out, T, S, base, ratio, depth, (const double(*)[3])q[i], frac / 4, level + 1);
    } else {
        append_piece(out, base, b, frac);
    }
}

/* near_rule, src/assembly/assemble.jl:69-73 */
static void near_rule(rule_t *out, const el_t *T, const srcref_t *S, const rule_t *base, double ratio, int depth)
{
    rule_clear(out);
    near_pieces(out, T, S, base, ratio, depth, CORNERS, 1.0, 0);
}

/* vertex_pieces!, src/assembly/assemble.jl:102-116 (k 0-based; p physical corners of the piece) */
static void vertex_pieces(rule_t *out, const v3 *p, int k, const rule_t *base, const double b[3][3], double frac,
                          double corner_angle)
{
    int i = (k + 1) % 3, j = (k + 2) % 3;
    v3 u = sub(p[i], p[k]), v = sub(p[j], p[k]);
    double cs = dot(u, v) / (norm(u) * norm(v));
    if (cs < -1.0) cs = -1.0;
    if (cs > 1.0) cs = 1.0;
    double th = acos(cs);
    if (th <= PI / 2) {
        double bb[3][3];
        memcpy(bb[0], b[k], sizeof(double) * 3);
        memcpy(bb[1], b[i], sizeof(double) * 3);
        memcpy(bb[2], b[j], sizeof(double) * 3);
        append_piece(out, base, (const double(*)[3])bb, frac);
        return;
    }
    int m = (int)ceil(th * (180 / PI) / corner_angle);
    double ts[64];
    ts[0] = 0.0;
    for (int q = 1; q <= m - 1; q++) {
        double ph = th * q / m;
        ts[q] = norm(u) * sin(ph) / (norm(u) * sin(ph) + norm(v) * sin(th - ph));
    }
    ts[m] = 1.0;
    for (int q = 0; q < m; q++) {
        double bb[3][3];
        for (int c = 0; c < 3; c++) {
            bb[0][c] = b[k][c];
            bb[1][c] = (1 - ts[q]) * b[i][c] + ts[q] * b[j][c];
            bb[2][c] = (1 - ts[q + 1]) * b[i][c] + ts[q + 1] * b[j][c];
        }
        append_piece(out, base, (const double(*)[3])bb, frac * (ts[q + 1] - ts[q]));
    }
}

/* vertex_rule, src/assembly/assemble.jl:119-125 (k 0-based); returns the stored rule or builds into out */
static const rule_t *vertex_rule(const el_t *T, int k, const tier_rules_t *R, const vpo_opts *o, rule_t *out)
{
    v3 u = sub(T->x[(k + 1) % 3], T->x[k]), v = sub(T->x[(k + 2) % 3], T->x[k]);
    if (dot(u, v) >= 0) return &R->vertex[k];
    rule_clear(out);
    vertex_pieces(out, T->x, k, &R->vertex[0], CORNERS, 1.0, o->corner_angle);
    return out;
}

/* edge_splits / splits_an_edge, src/assembly/assemble.jl:160-184 (S is a triangle) */
static int edge_splits_k(const v3 *p, int k, const el_t *S, double tol, double *ts)
{
    v3 a = p[k], c = p[(k + 1) % 3];
    double L = norm(sub(c, a));
    int n = 0;
    for (int q = 0; q < 3; q++) {
        double t;
        double d = seg_projection(S->x[q], a, c, &t);
        if (d <= tol && t * L > tol && (1 - t) * L > tol) ts[n++] = t;
    }
    for (int i = 1; i < n; i++) /* sort! */
        for (int j = i; j > 0 && ts[j] < ts[j - 1]; j--) { double x = ts[j]; ts[j] = ts[j - 1]; ts[j - 1] = x; }
    return n;
}

static int splits_an_edge(const v3 *p, const el_t *S, double tol)
{
    for (int k = 0; k < 3; k++) {
        v3 a = p[k], c = p[(k + 1) % 3];
        double L = norm(sub(c, a));
        for (int q = 0; q < 3; q++) {
            double t;
            double d = seg_projection(S->x[q], a, c, &t);
            if (d <= tol && t * L > tol && (1 - t) * L > tol) return 1;
        }
    }
    return 0;
}

/* contact_pieces!, src/assembly/assemble.jl:209-238 */
static void contact_pieces(rule_t *out, const el_t *T, const el_t *S, const tier_rules_t *R, const vpo_opts *o,
                           double tol, const double b[3][3], double frac, int level)
{
    v3 p[3] = {el_point(T, b[0]), el_point(T, b[1]), el_point(T, b[2])};
    if (level < 4) {
        for (int k = 0; k < 3; k++) {
            double sp[3];
            int ns = edge_splits_k(p, k, S, tol, sp);
            if (ns == 0) continue;
            double ts[5];
            ts[0] = 0.0;
            for (int i = 0; i < ns; i++) ts[i + 1] = sp[i];
            ts[ns + 1] = 1.0;
            const double *bk = b[k], *bk1 = b[(k + 1) % 3], *opp = b[(k + 2) % 3];
            for (int i = 0; i < ns + 1; i++) {
                double piece[3][3];
                for (int c = 0; c < 3; c++) {
                    piece[0][c] = (1 - ts[i]) * bk[c] + ts[i] * bk1[c];
                    piece[1][c] = (1 - ts[i + 1]) * bk[c] + ts[i + 1] * bk1[c];
                    piece[2][c] = opp[c];
                }
                contact_pieces(out, T, S, R, o, tol, (const double(*)[3])piece, frac * (ts[i + 1] - ts[i]), level + 1);
            }
            return;
        }
    }
    int on[3], nc = 0;
    for (int k = 0; k < 3; k++) { on[k] = tri_distance(S->x, p[k]) <= tol; nc += on[k]; }
    if (nc == 2) {
        int k = !on[2] ? 0 : !on[0] ? 1 : 2;
        if (tri_distance(S->x, dvs(add(p[k], p[(k + 1) % 3]), 2.0)) <= tol) { append_piece(out, &R->edge[k], b, frac); return; }
    }
    if (nc == 1) {
        int k = on[0] ? 0 : on[1] ? 1 : 2;
        vertex_pieces(out, p, k, &R->vertex[0], b, frac, o->corner_angle);
    } else if (nc == 0) {
        srcref_t sr = {0, S, {0, 0, 0}, {0, 0, 0}};
        near_pieces(out, T, &sr, &R->near, o->near_ratio, o->near_depth, b, frac, 0);
    } else {
        append_piece(out, &R->self, b, frac);
    }
}

/* contact_rule, src/assembly/assemble.jl:199-207; returns 1 and fills out when S touches T beyond shared corners */
static int contact_rule(const el_t *T, const el_t *S, const int shared[3], const tier_rules_t *R, const vpo_opts *o,
                        rule_t *out)
{
    double tol = o->contact_rel * dmin(T->diam, S->diam);
    if (norm(sub(S->c, T->c)) > reach(T) + reach(S) + tol) return 0;
    int extra = 0;
    for (int k = 0; k < 3 && !extra; k++) extra = !shared[k] && tri_distance(S->x, T->x[k]) <= tol;
    if (!extra) extra = splits_an_edge(T->x, S, tol);
    if (!extra) return 0;
    rule_clear(out);
    contact_pieces(out, T, S, R, o, tol, CORNERS, 1.0, 0);
    return 1;
}

/* tiers: TIER_NAMES, src/assembly/assemble.jl:260 (0 self, 1 edge, 2 vertex, 3 contact, 4 near, 5 far) */
enum { T_SELF = 0, T_EDGE, T_VERTEX, T_CONTACT, T_NEAR, T_FAR };

/* pair_rule, src/assembly/assemble.jl:266-282 (tri/gs node ids; wake sources have ids no body node shares) */
static const rule_t *pair_rule(const el_t *T, const el_t *S, int t, int s, const int *tri, const int *gs,
                               const tier_rules_t *R, const vpo_opts *o, int touch, rule_t *tmp, int *tier)
{
    if (s == t) { *tier = T_SELF; return &R->self; }
    if (norm(sub(S->c, T->c)) > o->far_ratio * dmax(T->diam, S->diam)) { *tier = T_FAR; return &R->far; }
    int shared[3];
    for (int k = 0; k < 3; k++) shared[k] = tri[k] == gs[0] || tri[k] == gs[1] || tri[k] == gs[2];
    if (contact_rule(T, S, shared, R, o, tmp)) { *tier = T_CONTACT; return tmp; }
    int ns = shared[0] + shared[1] + shared[2];
    if (ns == 2) {
        *tier = T_EDGE;
        return &R->edge[shared[0] && shared[1] ? 0 : shared[1] && shared[2] ? 1 : 2];
    } else if (ns == 1) {
        *tier = T_VERTEX;
        return vertex_rule(T, shared[0] ? 0 : shared[1] ? 1 : 2, R, o, tmp);
    }
    int depth = touch ? o->near_depth + 4 : o->near_depth;
    srcref_t sr = {0, S, {0, 0, 0}, {0, 0, 0}};
    *tier = T_NEAR;
    near_rule(tmp, T, &sr, &R->near, o->near_ratio, depth);
    return tmp;
}

/* segment_rule, src/assembly/assemble.jl:299-306 */
static const rule_t *segment_rule(const el_t *T, v3 fa, v3 fb, const tier_rules_t *R, const vpo_opts *o, rule_t *tmp,
                                  int *tier)
{
    double L = norm(sub(fb, fa));
    if (norm(sub(dvs(add(fa, fb), 2.0), T->c)) > o->far_ratio * dmax(T->diam, L)) { *tier = T_FAR; return &R->far; }
    for (int k = 0; k < 3; k++)
        if (dmin(norm(sub(T->x[k], fa)), norm(sub(T->x[k], fb))) <= 1e-12 * T->diam) {
            *tier = T_VERTEX;
            return vertex_rule(T, k, R, o, tmp);
        }
    srcref_t sr = {1, NULL, fa, fb};
    *tier = T_NEAR;
    near_rule(tmp, T, &sr, &R->near, o->near_ratio, o->near_depth);
    return tmp;
}

/* test_weights, src/assembly/assemble.jl:52-58 */
static inline void test_weights(int test, const double *l, double *psi)
{
    if (test == 0) { psi[0] = l[0]; psi[1] = l[1]; psi[2] = l[2]; return; }
    if (test == 1) { psi[0] = psi[1] = psi[2] = 1.0; return; }
    double m = dmax(dmax(l[0], l[1]), l[2]);
    int is[3], c = 0;
    for (int k = 0; k < 3; k++) { is[k] = fabs(l[k] - m) < 1e-12; c += is[k]; }
    for (int k = 0; k < 3; k++) psi[k] = (double)is[k] / c;
}

/* pair_block, src/assembly/assemble.jl:245-258: blk[a][b] row-major (a test corner, b source corner) */
static void pair_block(const el_t *T, const el_t *S, const rule_t *r, double sig, int test, int side, double kfar,
                       double blk[3][3], double rb[3])
{
    memset(blk, 0, sizeof(double) * 9);
    rb[0] = rb[1] = rb[2] = 0;
    for (int q = 0; q < r->n; q++) {
        const double *l = r->l + 3 * q;
        v3 w = sv(&S->src, el_point(T, l), side, kfar, 1e-12);
        double W = r->w[q] * T->area;
        v3 v = cross(w, T->n);
        double c[3] = {dot(v, S->g[0]), dot(v, S->g[1]), dot(v, S->g[2])};
        double tw[3], psi[3];
        test_weights(test, l, tw);
        for (int a = 0; a < 3; a++) psi[a] = W * tw[a];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) blk[a][b] += psi[a] * c[b];
        double f = sig * dot(w, T->n);
        for (int a = 0; a < 3; a++) rb[a] -= f * psi[a];
    }
}

/* segment_block, src/assembly/assemble.jl:312-321 */
static void segment_block(const el_t *T, v3 fa, v3 fb, const rule_t *r, int test, double ca[3], double cb[3])
{
    ca[0] = ca[1] = ca[2] = cb[0] = cb[1] = cb[2] = 0;
    for (int q = 0; q < r->n; q++) {
        const double *l = r->l + 3 * q;
        v3 va, vb;
        seg_vel(fa, fb, el_point(T, l), &va, &vb);
        double tw[3];
        test_weights(test, l, tw);
        double W = r->w[q] * T->area;
        double da = dot(va, T->n), db = dot(vb, T->n);
        for (int a = 0; a < 3; a++) { ca[a] += da * (W * tw[a]); cb[a] += db * (W * tw[a]); }
    }
}

/* ---------------------------------------------------------------- assemble ---------------------------------- */

static int build_els(int nn, const double *x, int nt, const int *tris, el_t *els)
{
    for (int t = 0; t < nt; t++) {
        const int *tr = tris + 3 * t;
        if (tr[0] < 0 || tr[1] < 0 || tr[2] < 0 || tr[0] >= nn || tr[1] >= nn || tr[2] >= nn) return -2;
        if (make_el(ld3(x + 3 * tr[0]), ld3(x + 3 * tr[1]), ld3(x + 3 * tr[2]), els + t)) return -1;
    }
    return 0;
}

static int cmp_pair(const void *a, const void *b)
{
    const int *p = (const int *)a, *q = (const int *)b;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    return 0;
}

static int is_touch(const int *tp, int ntp, int t, int s)
{
    if (!ntp) return 0;
    int key[2] = {t, s};
    return bsearch(key, tp, ntp, sizeof(int) * 2, cmp_pair) != NULL;
}

/*
 * assemble, src/assembly/assemble.jl:377-447, with add_wake! (:333-356) and the freestream RHS (:422).
 * Indices are 0-based.  A is column-major nslot x (nslot + nw): A[i + j*nslot] = row (test slot) i, column j.
 * tiers (optional, nt x (nt + nwt) bytes, row-major by test triangle) receives each body/wake pair's tier.
 * Returns 0, -1 degenerate triangle, -2 bad index, -3 bad option.
 */
int vpo_assemble_rows(int nn, const double *x, int nt, const int *tris, const int *corner_slot, int nslot,
                      const double *sigma, const double *vinf, int ntouch, const int *touch,
                      int nwx, const double *wx, int nwt, const int *wtris, int nwf, const int *wfil,
                      const vpo_opts *o, int nthreads, double *A, double *bvec, long long *pairs, long long *points,
                      unsigned char *tiers, int t_begin, int t_end)
{
    if (o->test < 0 || o->test > 2) return -3;
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    el_t *els = (el_t *)malloc(sizeof(el_t) * (nt > 0 ? nt : 1));
    el_t *wel = (el_t *)malloc(sizeof(el_t) * (nwt > 0 ? nwt : 1));
    int rc = build_els(nn, x, nt, tris, els);
    if (!rc && nwt) rc = build_els(nwx, wx, nwt, wtris, wel);
    int *tp = NULL;
    if (!rc && ntouch) {
        tp = (int *)malloc(sizeof(int) * 4 * ntouch);
        for (int i = 0; i < ntouch; i++) {
            tp[4 * i] = touch[2 * i]; tp[4 * i + 1] = touch[2 * i + 1];
            tp[4 * i + 2] = touch[2 * i + 1]; tp[4 * i + 3] = touch[2 * i];
        }
        qsort(tp, 2 * ntouch, sizeof(int) * 2, cmp_pair);
    }
    if (rc) { free(els); free(wel); free(tp); tier_rules_free(&R); return rc; }
    long long ncol = (long long)nslot + nwx;
    memset(A, 0, sizeof(double) * (size_t)nslot * (size_t)ncol);
    memset(bvec, 0, sizeof(double) * nslot);
    for (int k = 0; k < 6; k++) pairs[k] = points[k] = 0;
    v3 V = ld3(vinf);
    double ti[3] = {o->test == 1 ? 1.0 : 1.0 / 3, o->test == 1 ? 1.0 : 1.0 / 3, o->test == 1 ? 1.0 : 1.0 / 3};
    int nodeoff = nn; /* wake node ids offset past the body's so that no corner is shared */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)ncol);
        long long mp[6] = {0}, mq[6] = {0};
        rule_t tmp;
        rule_init(&tmp);
#pragma omp for schedule(dynamic, 4)
        for (int t = t_begin; t < t_end; t++) {
            const el_t *T = els + t;
            const int *tri = tris + 3 * t;
            memset(buf, 0, sizeof(double) * 3 * (size_t)ncol);
            double rows[3] = {0, 0, 0};
            for (int s = 0; s < nt; s++) {
                int tier;
                const rule_t *r = pair_rule(T, els + s, t, s, tri, tris + 3 * s, &R, o, is_touch(tp, 2 * ntouch, t, s), &tmp, &tier);
                double blk[3][3], rb[3];
                pair_block(T, els + s, r, sigma[s], o->test, s == t ? 1 : 0, o->kernel_far, blk, rb);
                const int *cs = corner_slot + 3 * s;
                for (int bb = 0; bb < 3; bb++)
                    for (int a = 0; a < 3; a++) buf[(size_t)a * ncol + cs[bb]] += blk[a][bb];
                for (int a = 0; a < 3; a++) rows[a] += rb[a];
                mp[tier]++;
                mq[tier] += r->n;
                if (tiers) tiers[(size_t)t * (nt + nwt) + s] = (unsigned char)tier;
            }
            /* add_wake!, src/assembly/assemble.jl:333-356 (wake_pair_rule :288-293) */
            for (int s = 0; s < nwt; s++) {
                const int *wt = wtris + 3 * s;
                int gs[3] = {wt[0] + nodeoff, wt[1] + nodeoff, wt[2] + nodeoff};
                int tier;
                const rule_t *r = pair_rule(T, wel + s, t, -1 - s, tri, gs, &R, o, 0, &tmp, &tier);
                double blk[3][3], rb[3];
                pair_block(T, wel + s, r, 0.0, o->test, 0, o->kernel_far, blk, rb);
                for (int bb = 0; bb < 3; bb++)
                    for (int a = 0; a < 3; a++) buf[(size_t)a * ncol + nslot + wt[bb]] += blk[a][bb];
                mp[tier]++;
                mq[tier] += r->n;
                if (tiers) tiers[(size_t)t * (nt + nwt) + nt + s] = (unsigned char)tier;
            }
            for (int f = 0; f < nwf; f++) {
                int ia = wfil[2 * f], ib = wfil[2 * f + 1];
                v3 fa = ld3(wx + 3 * ia), fb = ld3(wx + 3 * ib);
                int tier;
                const rule_t *r = segment_rule(T, fa, fb, &R, o, &tmp, &tier);
                double ca[3], cb[3];
                segment_block(T, fa, fb, r, o->test, ca, cb);
                for (int a = 0; a < 3; a++) {
                    buf[(size_t)a * ncol + nslot + ia] += ca[a];
                    buf[(size_t)a * ncol + nslot + ib] += cb[a];
                }
                mp[tier]++;
                mq[tier] += r->n;
            }
            double vn = dot(V, T->n);
            for (int a = 0; a < 3; a++) rows[a] -= vn * T->area * ti[a];
            const int *ct = corner_slot + 3 * t;
#pragma omp critical(vpo_scatter)
            {
                for (int a = 0; a < 3; a++) {
                    const double *src = buf + (size_t)a * ncol;
                    for (long long j = 0; j < ncol; j++) A[(size_t)j * nslot + ct[a]] += src[j];
                    bvec[ct[a]] += rows[a];
                }
            }
        }
#pragma omp critical(vpo_stats)
        for (int k = 0; k < 6; k++) { pairs[k] += mp[k]; points[k] += mq[k]; }
        rule_free(&tmp);
        free(buf);
    }
    free(els);
    free(wel);
    free(tp);
    tier_rules_free(&R);
    return 0;
}

/* The whole assembly (all test triangles); vpo_assemble_rows restricts the rows to test triangles
 * [t_begin, t_end) -- the bounded sample the CPU baseline times. */
int vpo_assemble(int nn, const double *x, int nt, const int *tris, const int *corner_slot, int nslot,
                 const double *sigma, const double *vinf, int ntouch, const int *touch,
                 int nwx, const double *wx, int nwt, const int *wtris, int nwf, const int *wfil,
                 const vpo_opts *o, int nthreads, double *A, double *bvec, long long *pairs, long long *points,
                 unsigned char *tiers)
{
    return vpo_assemble_rows(nn, x, nt, tris, corner_slot, nslot, sigma, vinf, ntouch, touch, nwx, wx, nwt, wtris,
                             nwf, wfil, o, nthreads, A, bvec, pairs, points, tiers, 0, nt);
}

/* The classification of assemble alone (pair_rule, wake_pair_rule and segment_rule per pair, no pair blocks): the
 * tier of every (body test triangle, source triangle) pair into tiers (nt x (nt + nwt), when not null) and the pair and
 * point counts per tier, with the same rules and arithmetic as vpo_assemble; for full-size tier checks. */
int vpo_classify(int nn, const double *x, int nt, const int *tris, int ntouch, const int *touch, int nwx,
                 const double *wx, int nwt, const int *wtris, int nwf, const int *wfil, const vpo_opts *o,
                 int nthreads, long long *pairs, long long *points, unsigned char *tiers)
{
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    el_t *els = (el_t *)malloc(sizeof(el_t) * (nt > 0 ? nt : 1));
    el_t *wel = (el_t *)malloc(sizeof(el_t) * (nwt > 0 ? nwt : 1));
    int rc = build_els(nn, x, nt, tris, els);
    if (!rc && nwt) rc = build_els(nwx, wx, nwt, wtris, wel);
    int *tp = NULL;
    if (!rc && ntouch) {
        tp = (int *)malloc(sizeof(int) * 4 * ntouch);
        for (int i = 0; i < ntouch; i++) {
            tp[4 * i] = touch[2 * i]; tp[4 * i + 1] = touch[2 * i + 1];
            tp[4 * i + 2] = touch[2 * i + 1]; tp[4 * i + 3] = touch[2 * i];
        }
        qsort(tp, 2 * ntouch, sizeof(int) * 2, cmp_pair);
    }
    if (rc) { free(els); free(wel); free(tp); tier_rules_free(&R); return rc; }
    for (int k = 0; k < 6; k++) pairs[k] = points[k] = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        long long mp[6] = {0}, mq[6] = {0};
        rule_t tmp;
        rule_init(&tmp);
#pragma omp for schedule(dynamic, 16)
        for (int t = 0; t < nt; t++) {
            const el_t *T = els + t;
            const int *tri = tris + 3 * t;
            for (int s = 0; s < nt; s++) {
                int tier;
                const rule_t *r = pair_rule(T, els + s, t, s, tri, tris + 3 * s, &R, o, is_touch(tp, 2 * ntouch, t, s), &tmp, &tier);
                mp[tier]++;
                mq[tier] += r->n;
                if (tiers) tiers[(size_t)t * (nt + nwt) + s] = (unsigned char)tier;
            }
            for (int s = 0; s < nwt; s++) {
                const int *wt = wtris + 3 * s;
                int gs[3] = {wt[0] + nn, wt[1] + nn, wt[2] + nn};
                int tier;
                const rule_t *r = pair_rule(T, wel + s, t, -1 - s, tri, gs, &R, o, 0, &tmp, &tier);
                mp[tier]++;
                mq[tier] += r->n;
                if (tiers) tiers[(size_t)t * (nt + nwt) + nt + s] = (unsigned char)tier;
            }
            for (int f = 0; f < nwf; f++) {
                int tier;
                const rule_t *r = segment_rule(T, ld3(wx + 3 * wfil[2 * f]), ld3(wx + 3 * wfil[2 * f + 1]), &R, o, &tmp, &tier);
                mp[tier]++;
                mq[tier] += r->n;
            }
        }
#pragma omp critical(vpo_cstats)
        for (int k = 0; k < 6; k++) { pairs[k] += mp[k]; points[k] += mq[k]; }
        rule_free(&tmp);
    }
    free(els);
    free(wel);
    free(tp);
    tier_rules_free(&R);
    return 0;
}

/* One pair block, for the per-pair parity tests: tier, block and RHS of test triangle t and source triangle s
 * (s >= nt: wake triangle s - nt).  Same rules and classification as vpo_assemble. */
int vpo_pair(int nn, const double *x, int nt, const int *tris, const double *sigma, int nwx, const double *wx,
             int nwt, const int *wtris, int t, int s, int touch, const vpo_opts *o, double *blk, double *rb,
             int *npoints)
{
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    el_t T, S;
    int gs[3];
    const int *tri = tris + 3 * t;
    if (make_el(ld3(x + 3 * tri[0]), ld3(x + 3 * tri[1]), ld3(x + 3 * tri[2]), &T)) return -1;
    double sig = 0.0;
    int sid;
    if (s < nt) {
        const int *st = tris + 3 * s;
        if (make_el(ld3(x + 3 * st[0]), ld3(x + 3 * st[1]), ld3(x + 3 * st[2]), &S)) return -1;
        gs[0] = st[0]; gs[1] = st[1]; gs[2] = st[2];
        sig = sigma[s];
        sid = s;
    } else {
        const int *wt = wtris + 3 * (s - nt);
        if (make_el(ld3(wx + 3 * wt[0]), ld3(wx + 3 * wt[1]), ld3(wx + 3 * wt[2]), &S)) return -1;
        gs[0] = wt[0] + nn; gs[1] = wt[1] + nn; gs[2] = wt[2] + nn;
        sid = -1;
    }
    rule_t tmp;
    rule_init(&tmp);
    int tier;
    const rule_t *r = pair_rule(&T, &S, t, sid, tri, gs, &R, o, touch, &tmp, &tier);
    double b3[3][3];
    pair_block(&T, &S, r, sig, o->test, sid == t ? 1 : 0, o->kernel_far, b3, rb);
    memcpy(blk, b3, sizeof(double) * 9);
    *npoints = r->n;
    rule_free(&tmp);
    tier_rules_free(&R);
    return tier;
}

/* ----------------------------------------------------------- test exports ---------------------------------- */

/* source_velocity at n points (xs, n x 3 -> w, n x 3), for cubature references in the tests */
int vpo_source_velocity_many(const double *a, const double *b, const double *c, int n, const double *xs, int side,
                             double far, double plane_tol, int mode, double *w)
{
    src_t s;
    if (make_src(ld3(a), ld3(b), ld3(c), &s)) return -1;
    for (int i = 0; i < n; i++) {
        v3 x = ld3(xs + 3 * i);
        v3 r = mode == 1 ? sv_exact(&s, x, side, plane_tol) : mode == 2 ? sv_far(&s, x) : sv(&s, x, side, far, plane_tol);
        st3(w + 3 * i, r);
    }
    return 0;
}

/* segment_velocity at n points: va, vb (n x 3 each) */
void vpo_segment_velocity_many(const double *a, const double *b, int n, const double *xs, double *va, double *vb)
{
    for (int i = 0; i < n; i++) {
        v3 p, q;
        seg_vel(ld3(a), ld3(b), ld3(xs + 3 * i), &p, &q);
        st3(va + 3 * i, p);
        st3(vb + 3 * i, q);
    }
}

/* The rule and pair_block of a test triangle T (tx: 3 corners) and a source triangle S (sx), given directly, with node
 * ids tid and sid (equal ids are shared corners; a wake source's ids match no body node, which makes pair_rule the
 * reference's wake_pair_rule).  self_pair: S is T (side +1).  kind -1: the classification (pair_rule, tier set);
 * otherwise a fixed rule: 0 far, 1 near (the stored 7-point base), 2 self, 3..5 edge k, 6..8 vertex k (stored), 9
 * near_rule (the recursive split), 10..12 vertex_rule k (wide corners in wedges), 13 contact_rule (-2 when it gives
 * none).  Returns the rule's point count; fills up to cap points into lam (3 each) and w, and blk (3x3 row-major, test
 * corner by source corner) and rb of pair_block with source strength sig. */
int vpo_pair_geom(const double *tx, const double *sx, const int *tid, const int *sid, int self_pair, int kind,
                  const vpo_opts *o, double sig, int *tier, int cap, double *lam, double *w, double *blk, double *rb)
{
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    el_t T, S;
    if (make_el(ld3(tx), ld3(tx + 3), ld3(tx + 6), &T) || make_el(ld3(sx), ld3(sx + 3), ld3(sx + 6), &S)) {
        tier_rules_free(&R);
        return -1;
    }
    rule_t tmp;
    rule_init(&tmp);
    const rule_t *r = NULL;
    int shared[3];
    for (int k = 0; k < 3; k++) shared[k] = tid[k] == sid[0] || tid[k] == sid[1] || tid[k] == sid[2];
    *tier = -1;
    if (kind == -1) {
        r = pair_rule(&T, &S, 0, self_pair ? 0 : 1, tid, sid, &R, o, 0, &tmp, tier);
    } else if (kind == 0) {
        r = &R.far;
    } else if (kind == 1) {
        r = &R.near;
    } else if (kind == 2) {
        r = &R.self;
    } else if (kind <= 5) {
        r = &R.edge[kind - 3];
    } else if (kind <= 8) {
        r = &R.vertex[kind - 6];
    } else if (kind == 9) {
        srcref_t sr = {0, &S, {0, 0, 0}, {0, 0, 0}};
        near_rule(&tmp, &T, &sr, &R.near, o->near_ratio, o->near_depth);
        r = &tmp;
    } else if (kind <= 12) {
        r = vertex_rule(&T, kind - 10, &R, o, &tmp);
    } else if (kind == 13) {
        if (contact_rule(&T, &S, shared, &R, o, &tmp)) r = &tmp;
    }
    int n = -2;
    if (r) {
        n = r->n;
        for (int i = 0; i < n && i < cap; i++) {
            if (lam) memcpy(lam + 3 * i, r->l + 3 * i, sizeof(double) * 3);
            if (w) w[i] = r->w[i];
        }
        double b3[3][3], r3[3];
        pair_block(&T, &S, r, sig, o->test, self_pair ? 1 : 0, o->kernel_far, b3, r3);
        if (blk) memcpy(blk, b3, sizeof(double) * 9);
        if (rb) memcpy(rb, r3, sizeof(double) * 3);
    }
    rule_free(&tmp);
    tier_rules_free(&R);
    return n;
}

/* The filament rows of a test triangle T (tx) for a filament from fa to fb: segment_rule (tier set) and
 * segment_block's ca, cb (src/assembly/assemble.jl:299-321); returns the rule's point count */
int vpo_segment_geom(const double *tx, const double *fa, const double *fb, const vpo_opts *o, int *tier, double *ca,
                     double *cb)
{
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    el_t T;
    if (make_el(ld3(tx), ld3(tx + 3), ld3(tx + 6), &T)) {
        tier_rules_free(&R);
        return -1;
    }
    rule_t tmp;
    rule_init(&tmp);
    const rule_t *r = segment_rule(&T, ld3(fa), ld3(fb), &R, o, &tmp, tier);
    const int n = r->n;
    segment_block(&T, ld3(fa), ld3(fb), r, o->test, ca, cb);
    rule_free(&tmp);
    tier_rules_free(&R);
    return n;
}

/* ------------------------------------------------------------- velocities ---------------------------------- */

/* body_velocity, src/post/Post.jl:92-112: u(x) = sum_s sigma_s w_s(x) + gamma_s x w_s(x) (no freestream).
 * gam is nt x 3 (the sheet of every element). */
int vpo_body_velocity(int nn, const double *x, int nt, const int *tris, const double *sigma, const double *gam,
                      int np, const double *xs, double kernel_far, int nthreads, double *u)
{
    el_t *els = (el_t *)malloc(sizeof(el_t) * (nt > 0 ? nt : 1));
    int rc = build_els(nn, x, nt, tris, els);
    if (rc) { free(els); return rc; }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int i = 0; i < np; i++) {
        v3 p = ld3(xs + 3 * i), acc = mk(0, 0, 0);
        for (int s = 0; s < nt; s++) {
            v3 w = sv(&els[s].src, p, 0, kernel_far, 1e-12);
            acc = add(acc, add(scl(sigma[s], w), cross(ld3(gam + 3 * s), w)));
        }
        st3(u + 3 * i, acc);
    }
    free(els);
    return 0;
}

/* surface_velocity, src/post/Post.jl:32-64: velocity on the fluid side at every centroid; the element's own
 * source and sheet on side +1; wake sheets (wgam per wake triangle) and filaments (wfil with wmu per wake node). */
int vpo_surface_velocity(int nn, const double *x, int nt, const int *tris, const double *sigma, const double *gam,
                         const double *vinf, int nwx, const double *wx, int nwt, const int *wtris, const double *wgam,
                         int nwf, const int *wfil, const double *wmu, double kernel_far, int nthreads, double *u)
{
    el_t *els = (el_t *)malloc(sizeof(el_t) * (nt > 0 ? nt : 1));
    el_t *wel = (el_t *)malloc(sizeof(el_t) * (nwt > 0 ? nwt : 1));
    int rc = build_els(nn, x, nt, tris, els);
    if (!rc && nwt) rc = build_els(nwx, wx, nwt, wtris, wel);
    if (rc) { free(els); free(wel); return rc; }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; t++) {
        v3 p = els[t].c;
        v3 acc = add(ld3(vinf), mk(0, 0, 0));
        for (int s = 0; s < nt; s++) {
            v3 w = sv(&els[s].src, p, s == t ? 1 : 0, kernel_far, 1e-12);
            acc = add(acc, add(scl(sigma[s], w), cross(ld3(gam + 3 * s), w)));
        }
        for (int s = 0; s < nwt; s++) acc = add(acc, cross(ld3(wgam + 3 * s), sv(&wel[s].src, p, 0, kernel_far, 1e-12)));
        for (int f = 0; f < nwf; f++) {
            int a = wfil[2 * f], b = wfil[2 * f + 1];
            v3 va, vb;
            seg_vel(ld3(wx + 3 * a), ld3(wx + 3 * b), p, &va, &vb);
            acc = add(acc, add(scl(wmu[a], va), scl(wmu[b], vb)));
        }
        st3(u + 3 * t, acc);
    }
    free(els);
    free(wel);
    return 0;
}

/* blob_velocities, src/freewake/step.jl:176-187 */
void vpo_blob_velocities(int nb, const double *bx, const double *ba, const double *bd, int np, const double *xs,
                         int nthreads, double *u)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int i = 0; i < np; i++) {
        v3 p = ld3(xs + 3 * i), acc = mk(0, 0, 0);
        for (int k = 0; k < nb; k++) acc = add(acc, blob_vel(p, ld3(bx + 3 * k), ld3(ba + 3 * k), bd[k]));
        st3(u + 3 * i, acc);
    }
}

/*
 * explicit_rhs, src/assembly/explicit.jl:31-86: -int psi_i n.u_src dA for every slot, from known sheets (sheet
 * triangles sx/stris with corner doublets smu), filaments (segments with end circulations gam2) and blobs.
 */
int vpo_explicit_rhs(int nn, const double *x, int nt, const int *tris, const int *corner_slot, int nslot,
                     int nsx, const double *sx, int nst, const int *stris, const double *smu,
                     int nseg, const double *seg, const double *gam2,
                     int nb, const double *bx, const double *ba, const double *bd,
                     const vpo_opts *o, int nthreads, double *bvec)
{
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    el_t *els = (el_t *)malloc(sizeof(el_t) * (nt > 0 ? nt : 1));
    el_t *sel = (el_t *)malloc(sizeof(el_t) * (nst > 0 ? nst : 1));
    int rc = build_els(nn, x, nt, tris, els);
    if (!rc && nst) rc = build_els(nsx, sx, nst, stris, sel);
    if (rc) { free(els); free(sel); tier_rules_free(&R); return rc; }
    memset(bvec, 0, sizeof(double) * nslot);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        rule_t tmp;
        rule_init(&tmp);
        const rule_t *fr = &R.far;
        v3 *ufar = (v3 *)malloc(sizeof(v3) * fr->n);
#pragma omp for schedule(dynamic, 4)
        for (int t = 0; t < nt; t++) {
            const el_t *T = els + t;
            double rows[3] = {0, 0, 0};
            static const int notri[3] = {-1, -2, -3};
            for (int s = 0; s < nst; s++) {
                int gs[3] = {-10, -11, -12}, tier;
                const rule_t *r = pair_rule(T, sel + s, t, -1 - s, notri, gs, &R, o, 0, &tmp, &tier);
                double blk[3][3], rb[3];
                pair_block(T, sel + s, r, 0.0, o->test, 0, o->kernel_far, blk, rb);
                const double *m = smu + 3 * s;
                for (int a = 0; a < 3; a++) rows[a] -= (blk[a][0] * m[0] + blk[a][1] * m[1]) + blk[a][2] * m[2];
            }
            for (int f = 0; f < nseg; f++) {
                v3 fa = ld3(seg + 6 * f), fb = ld3(seg + 6 * f + 3);
                int tier;
                const rule_t *r = segment_rule(T, fa, fb, &R, o, &tmp, &tier);
                double ca[3], cb[3];
                segment_block(T, fa, fb, r, o->test, ca, cb);
                for (int a = 0; a < 3; a++) rows[a] -= ca[a] * gam2[2 * f] + cb[a] * gam2[2 * f + 1];
            }
            for (int i = 0; i < fr->n; i++) ufar[i] = mk(0, 0, 0);
            for (int k = 0; k < nb; k++) {
                v3 xp = ld3(bx + 3 * k), al = ld3(ba + 3 * k);
                if (norm(sub(xp, T->c)) > o->far_ratio * dmax(T->diam, bd[k])) {
                    for (int i = 0; i < fr->n; i++) ufar[i] = add(ufar[i], blob_vel(el_point(T, fr->l + 3 * i), xp, al, bd[k]));
                } else {
                    srcref_t sr = {2, NULL, xp, xp};
                    near_rule(&tmp, T, &sr, &R.near, o->near_ratio, o->near_depth);
                    for (int q = 0; q < tmp.n; q++) {
                        const double *l = tmp.l + 3 * q;
                        double f = tmp.w[q] * T->area * dot(blob_vel(el_point(T, l), xp, al, bd[k]), T->n);
                        double tw[3];
                        test_weights(o->test, l, tw);
                        for (int a = 0; a < 3; a++) rows[a] -= f * tw[a];
                    }
                }
            }
            if (nb) {
                for (int i = 0; i < fr->n; i++) {
                    const double *l = fr->l + 3 * i;
                    double f = fr->w[i] * T->area * dot(ufar[i], T->n);
                    double tw[3];
                    test_weights(o->test, l, tw);
                    for (int a = 0; a < 3; a++) rows[a] -= f * tw[a];
                }
            }
            const int *ct = corner_slot + 3 * t;
#pragma omp critical(vpo_rhs)
            for (int a = 0; a < 3; a++) bvec[ct[a]] += rows[a];
        }
        free(ufar);
        rule_free(&tmp);
    }
    free(els);
    free(sel);
    tier_rules_free(&R);
    return 0;
}

/* the mixed-mesh (isoparametric quad) oracle shares this translation unit */
#include "vp_oracle_quad.c"
