/*
 * FP64 CPU oracle of the mixed-mesh assembly (isoparametric quads and flat triangles) -- TEST INFRASTRUCTURE ONLY.
 *
 * Included at the end of vp_oracle.c (one translation unit: it reuses the triangle kernel, rules and rule builders
 * there).  Like that file it is the checker: only tests/ load it.  It restates, in the reference's operation order
 * and with -ffp-contract=off:
 *   src/elements/iso_quad.jl     IsoQuad, shape functions, tangents, sheet densities, project (closest point);
 *   src/elements/quad_field.jl   kernel_terms, gauss_field, polar_field (sinh-graded polar rays, Guiggiani principal
 *                                values, one-sided limits), ray_pv, ray_near, quad_field;
 *   src/elements/quad_rules.jl   SquareRule, square_gauss, turn, the graded edge / vertex / self square rules,
 *                                SquareTiers, square_near_rule;
 *   src/assembly/mixed.jl        MixedMesh element order, square_test_weights, near_rects!, vertex_rect!,
 *                                vertex_square, side_splits, cuts, square_contact_rule, quad_pair_rule, tri_pair_rule,
 *                                quad_wake_rule, quad_segment_rule, the four block kinds, segment_block, assemble.
 * Pinned by tests/test_oracle_quads.py: adaptive cubature of the inner integrals (polar and tensor Gauss) and of
 * touching pair blocks, brute-force closest points, principal values against two-sided limits, the rules' monomial
 * and log integrals, the triangle-only mixed mesh against vpo_assemble, and the cube-sphere's exact flow.
 */

/* ------------------------------------------------------------------ IsoQuad ---------------------------------- */

/* IsoQuad, src/elements/iso_quad.jl:13-37 */
typedef struct {
    v3 x[4], warp, c, n;
    double area, diam, reach;
} quad_t;

static int make_quad(v3 a, v3 b, v3 c, v3 d, quad_t *q)
{
    v3 nc = cross(sub(c, a), sub(d, b));
    if (!(norm(nc) > 0)) return -1;
    double xs[4], ws[4];
    gauss_legendre(4, xs, ws);
    double area = 0.0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double xi = xs[i], eta = xs[j];
            v3 a1 = add(scl(1 - eta, sub(b, a)), scl(eta, sub(c, d)));
            v3 a2 = add(scl(1 - xi, sub(d, a)), scl(xi, sub(c, b)));
            area += ws[i] * ws[j] * norm(cross(a1, a2));
        }
    q->x[0] = a; q->x[1] = b; q->x[2] = c; q->x[3] = d;
    double dm = norm(sub(b, a));
    dm = dmax(dm, norm(sub(c, b)));
    dm = dmax(dm, norm(sub(d, c)));
    dm = dmax(dm, norm(sub(a, d)));
    dm = dmax(dm, norm(sub(c, a)));
    dm = dmax(dm, norm(sub(d, b)));
    q->diam = dm;
    q->c = dvs(add(add(add(a, b), c), d), 4.0);
    q->warp = sub(add(sub(a, b), c), d);
    q->n = dvs(nc, norm(nc));
    q->area = area;
    double r = 0.0;
    for (int k = 0; k < 4; k++) r = dmax(r, norm(sub(q->x[k], q->c)));
    q->reach = r;
    return 0;
}

/* shape, src/elements/iso_quad.jl:40 */
static inline void q_shape(double xi, double eta, double N[4])
{
    N[0] = (1 - xi) * (1 - eta); N[1] = xi * (1 - eta); N[2] = xi * eta; N[3] = (1 - xi) * eta;
}

/* shape_derivatives, src/elements/iso_quad.jl:43 */
static inline void q_dshape(double xi, double eta, double Nx[4], double Ne[4])
{
    Nx[0] = -(1 - eta); Nx[1] = 1 - eta; Nx[2] = eta; Nx[3] = -eta;
    Ne[0] = -(1 - xi); Ne[1] = -xi; Ne[2] = xi; Ne[3] = 1 - xi;
}

/* point, src/elements/iso_quad.jl:49-52 */
static inline v3 q_point(const quad_t *Q, double xi, double eta)
{
    double N[4];
    q_shape(xi, eta, N);
    return add(add(add(scl(N[0], Q->x[0]), scl(N[1], Q->x[1])), scl(N[2], Q->x[2])), scl(N[3], Q->x[3]));
}

/* tangents, src/elements/iso_quad.jl:55-60 */
static inline void q_tangents(const quad_t *Q, double xi, double eta, v3 *a1, v3 *a2)
{
    *a1 = add(scl(1 - eta, sub(Q->x[1], Q->x[0])), scl(eta, sub(Q->x[2], Q->x[3])));
    *a2 = add(scl(1 - xi, sub(Q->x[3], Q->x[0])), scl(xi, sub(Q->x[2], Q->x[1])));
}

/* area_normal, src/elements/iso_quad.jl:63 */
static inline v3 q_area_normal(const quad_t *Q, double xi, double eta)
{
    v3 a1, a2;
    q_tangents(Q, xi, eta, &a1, &a2);
    return cross(a1, a2);
}

/* sheet_densities, src/elements/iso_quad.jl:69-73 */
static inline void q_sheet_densities(const quad_t *Q, double xi, double eta, v3 g[4])
{
    v3 a1, a2;
    q_tangents(Q, xi, eta, &a1, &a2);
    double Nx[4], Ne[4];
    q_dshape(xi, eta, Nx, Ne);
    for (int b = 0; b < 4; b++) g[b] = sub(scl(Nx[b], a2), scl(Ne[b], a1));
}

/* segment_closest, src/elements/iso_quad.jl:94-98 */
static inline double seg_closest(v3 a, v3 b, v3 p, double *tout)
{
    v3 e = sub(b, a);
    double t = dot(sub(p, a), e) / dot(e, e);
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    *tout = t;
    return norm(sub(p, add(a, scl(t, e))));
}

/* project, src/elements/iso_quad.jl:108-146: closest point (P) and distance */
static double q_project(const quad_t *Q, v3 p, double P[2])
{
    double best_d = INFINITY, best[2] = {0.5, 0.5};
    double q[2] = {0.5, 0.5};
    double f0 = INFINITY;
    for (int a = 0; a <= 4; a++)
        for (int b = 0; b <= 4; b++) {
            double g0 = a / 4.0, g1 = b / 4.0;
            v3 r = sub(q_point(Q, g0, g1), p);
            if (dot(r, r) < f0) { f0 = dot(r, r); q[0] = g0; q[1] = g1; }
        }
    for (int it = 0; it < 60; it++) {
        v3 r = sub(q_point(Q, q[0], q[1]), p);
        v3 a1, a2;
        q_tangents(Q, q[0], q[1], &a1, &a2);
        double g1 = dot(r, a1), g2 = dot(r, a2);
        double H11 = dot(a1, a1), H22 = dot(a2, a2), H12 = dot(a1, a2) + dot(r, Q->warp);
        double det = H11 * H22 - H12 * H12;
        double st[2];
        if (det > 0) { st[0] = (H22 * g1 - H12 * g2) / det; st[1] = (H11 * g2 - H12 * g1) / det; }
        else { double hs = H11 + H22; st[0] = g1 / hs; st[1] = g2 / hs; }
        double t = 1.0;
        for (int i = 0; i < 2; i++) {
            if (st[i] > 0) t = dmin(t, (q[i] + 0.5) / st[i]);
            if (st[i] < 0) t = dmin(t, (q[i] - 1.5) / st[i]);
        }
        double f = dot(r, r);
        while (t > 1e-10) {
            v3 rn = sub(q_point(Q, q[0] - t * st[0], q[1] - t * st[1]), p);
            if (dot(rn, rn) <= f) break;
            t /= 2;
        }
        q[0] = q[0] - t * st[0];
        q[1] = q[1] - t * st[1];
        if (t * (fabs(st[0]) + fabs(st[1])) < 1e-14) break;
    }
    if (0 <= q[0] && q[0] <= 1 && 0 <= q[1] && q[1] <= 1) {
        best_d = norm(sub(q_point(Q, q[0], q[1]), p));
        best[0] = q[0]; best[1] = q[1];
    }
    for (int k = 0; k < 4; k++) {
        double t;
        double d = seg_closest(Q->x[k], Q->x[(k + 1) % 4], p, &t);
        if (d < best_d) {
            best_d = d;
            if (k == 0) { best[0] = t; best[1] = 0.0; }
            else if (k == 1) { best[0] = 1.0; best[1] = t; }
            else if (k == 2) { best[0] = 1 - t; best[1] = 1.0; }
            else { best[0] = 0.0; best[1] = 1 - t; }
        }
    }
    P[0] = best[0]; P[1] = best[1];
    return best_d;
}

/* -------------------------------------------------------------- quad_field ---------------------------------- */

static const double INV4PI = 1.0 / (4.0 * 3.14159265358979323846);

/* GAUSS = [gauss_legendre(n) for n in 1:48], src/elements/quad_field.jl:7 */
static double GX[49][48], GW[49][48];
static int g_gauss_ready = 0;
static void gauss_table(void)
{
    if (g_gauss_ready) return;
#pragma omp critical(vpo_gauss_table)
    if (!g_gauss_ready) {
        for (int n = 1; n <= 48; n++) gauss_legendre(n, GX[n], GW[n]);
        g_gauss_ready = 1;
    }
}

/* kernel_terms, src/elements/quad_field.jl:17-27 (without 1/4pi): c[4] sheet terms, *s source term */
static inline void kernel_terms(const quad_t *Q, v3 x, v3 nx, v3 V, double xi, double eta, double c[4], double *s)
{
    v3 y = q_point(Q, xi, eta);
    v3 a1, a2;
    q_tangents(Q, xi, eta, &a1, &a2);
    v3 r = sub(x, y);
    double r2 = dot(r, r);
    double ir3 = 1 / (r2 * sqrt(r2));
    v3 t = scl(ir3, cross(r, nx));
    double Nx[4], Ne[4];
    q_dshape(xi, eta, Nx, Ne);
    double ta2 = dot(t, a2), ta1 = dot(t, a1);
    for (int b = 0; b < 4; b++) c[b] = Nx[b] * ta2 - Ne[b] * ta1;
    *s = -dot(V, cross(a1, a2)) * dot(nx, r) * ir3;
}

/* gauss_order, src/elements/quad_field.jl:30 */
static inline int gauss_order(double r) { return r >= 8 ? 3 : r >= 4 ? 4 : r >= 2 ? 6 : 8; }

/* gauss_field, src/elements/quad_field.jl:48-59 */
static void gauss_field(const quad_t *Q, v3 x, v3 nx, v3 V, int m, double c[4], double *s)
{
    gauss_table();
    double cc[4] = {0, 0, 0, 0}, ss = 0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            double dc[4], ds;
            kernel_terms(Q, x, nx, V, GX[m][i], GX[m][j], dc, &ds);
            double w = GW[m][i] * GW[m][j];
            for (int b = 0; b < 4; b++) cc[b] += w * dc[b];
            ss += w * ds;
        }
    for (int b = 0; b < 4; b++) c[b] = INV4PI * cc[b];
    *s = INV4PI * ss;
}

/* pieces, src/elements/quad_field.jl:66-69: up to two (lo, hi, n) pieces; returns their count */
static int qf_pieces(double a, double b, double brk, int base, double slope, double lo[2], double hi[2], int n[2])
{
    int np;
    if (a < brk && brk < b) { lo[0] = a; hi[0] = brk; lo[1] = brk; hi[1] = b; np = 2; }
    else { lo[0] = a; hi[0] = b; np = 1; }
    for (int i = 0; i < np; i++) {
        int k = base + (int)ceil(slope * (hi[i] - lo[i]));
        n[i] = k < 48 ? k : 48;
    }
    return np;
}

/* ray_pv, src/elements/quad_field.jl:160-171 */
static void ray_pv(const quad_t *Q, v3 x, v3 nx, v3 V, const double P[2], const double e[2], double R,
                   const double f[4], double fs, int n, double c[4], double *s)
{
    double cc[4] = {0, 0, 0, 0}, ss = 0;
    for (int i = 0; i < n; i++) {
        double rho = R * GX[n][i], wq = GW[n][i];
        double dc[4], ds;
        kernel_terms(Q, x, nx, V, P[0] + rho * e[0], P[1] + rho * e[1], dc, &ds);
        for (int b = 0; b < 4; b++) cc[b] += (wq * R) * (rho * dc[b] - f[b] / rho);
        ss += (wq * R) * (rho * ds - fs / rho);
    }
    for (int b = 0; b < 4; b++) c[b] = cc[b];
    *s = ss;
}

/* ray_near, src/elements/quad_field.jl:174-192 */
static void ray_near(const quad_t *Q, v3 x, v3 nx, v3 V, const double P[2], const double e[2], double R, double rho0,
                     int npts, double slope, int nplain, double c[4], double *s)
{
    double cc[4] = {0, 0, 0, 0}, ss = 0;
    if (rho0 < R) {
        double lo[2], hi[2];
        int nn[2];
        int np = qf_pieces(0.0, asinh(R / rho0), -1.0, npts, slope, lo, hi, nn);
        for (int p = 0; p < np; p++) {
            double sa = lo[p], sb = hi[p];
            int n = nn[p];
            for (int i = 0; i < n; i++) {
                double sv = sa + (sb - sa) * GX[n][i];
                double rho = rho0 * sinh(sv);
                double jac = GW[n][i] * (sb - sa) * rho0 * cosh(sv) * rho;
                double dc[4], ds;
                kernel_terms(Q, x, nx, V, P[0] + rho * e[0], P[1] + rho * e[1], dc, &ds);
                for (int b = 0; b < 4; b++) cc[b] += jac * dc[b];
                ss += jac * ds;
            }
        }
    } else {
        for (int i = 0; i < nplain; i++) {
            double rho = R * GX[nplain][i], wq = GW[nplain][i];
            double dc[4], ds;
            kernel_terms(Q, x, nx, V, P[0] + rho * e[0], P[1] + rho * e[1], dc, &ds);
            double wr = wq * R * rho;
            for (int b = 0; b < 4; b++) cc[b] += wr * dc[b];
            ss += wr * ds;
        }
    }
    for (int b = 0; b < 4; b++) c[b] = cc[b];
    *s = ss;
}

static const double SQUARE[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {1.0, 1.0}, {0.0, 1.0}};

/* polar_field, src/elements/quad_field.jl:103-154; returns -1 for x on the quad's boundary (ArgumentError) */
static int polar_field(const quad_t *Q, v3 x, v3 nx, v3 V, const double P[2], double d, int side, double c[4],
                       double *s)
{
    const int angle_points = 5, ray_points = 6, pv_points = 12;
    const double angle_slope = 1.5, ray_slope = 1.0;
    gauss_table();
    int onq = d <= 1e-12 * Q->diam;
    if (onq) {
        double mn = dmin(dmin(P[0], P[1]), dmin(1 - P[0], 1 - P[1]));
        if (mn <= 1e-12) return -1;
    }
    v3 a1s, a2s;
    q_tangents(Q, P[0], P[1], &a1s, &a2s);
    double g11 = dot(a1s, a1s), g12 = dot(a1s, a2s), g22 = dot(a2s, a2s);
    double l11 = sqrt(g11);
    double l12 = g12 / l11, l22 = sqrt(g22 - g12 * g12 / g11);
    double J = l11 * l22;
    v3 g0[4];
    q_sheet_densities(Q, P[0], P[1], g0);
    double sJ0 = -dot(V, cross(a1s, a2s));
    double cc[4] = {0, 0, 0, 0}, ss = 0;
    for (int k = 0; k < 4; k++) {
        const double *S0 = SQUARE[k], *S1 = SQUARE[(k + 1) % 4];
        double pa0 = S0[0] - P[0], pa1 = S0[1] - P[1], pb0 = S1[0] - P[0], pb1 = S1[1] - P[1];
        double A[2] = {l11 * pa0 + l12 * pa1, l22 * pa1}, B[2] = {l11 * pb0 + l12 * pb1, l22 * pb1};
        double BA0 = B[0] - A[0], BA1 = B[1] - A[1];
        double L = sqrt(BA0 * BA0 + BA1 * BA1);
        double et[2] = {BA0 / L, BA1 / L};
        double nk[2] = {et[1], -et[0]};
        double h = A[0] * nk[0] + A[1] * nk[1];
        if (h <= 1e-14 * L) continue;
        double lo[2], hi[2];
        int nv[2];
        int np = qf_pieces(asinh((A[0] * et[0] + A[1] * et[1]) / h), asinh((B[0] * et[0] + B[1] * et[1]) / h), 0.0,
                           angle_points, angle_slope, lo, hi, nv);
        for (int p = 0; p < np; p++) {
            double va = lo[p], vb = hi[p];
            int n = nv[p];
            for (int i = 0; i < n; i++) {
                double v = va + (vb - va) * GX[n][i], wq = GW[n][i];
                double ch = cosh(v);
                double R = h * ch;
                double sh = sinh(v);
                double z0 = (nk[0] + sh * et[0]) / ch, z1 = (nk[1] + sh * et[1]) / ch;
                double e[2] = {(z0 - l12 * z1 / l22) / l11, z1 / l22};
                v3 At = add(scl(e[0], a1s), scl(e[1], a2s));
                double nA = norm(At);
                double W = wq * (vb - va) / (ch * J);
                double dc[4], ds;
                if (onq) {
                    double nA3 = nA * nA * nA;
                    double f[4];
                    for (int b = 0; b < 4; b++) f[b] = -dot(nx, cross(g0[b], At)) / nA3;
                    double fs = -sJ0 * dot(nx, At) / nA3;
                    ray_pv(Q, x, nx, V, P, e, R, f, fs, pv_points, dc, &ds);
                    double lg = log(R * nA);
                    for (int b = 0; b < 4; b++) dc[b] += f[b] * lg;
                    ds += fs * lg;
                } else {
                    ray_near(Q, x, nx, V, P, e, R, d / nA, ray_points, ray_slope, pv_points, dc, &ds);
                }
                for (int b = 0; b < 4; b++) cc[b] += W * dc[b];
                ss += W * ds;
            }
        }
    }
    for (int b = 0; b < 4; b++) c[b] = INV4PI * cc[b];
    *s = INV4PI * ss;
    if (onq && side != 0) {
        v3 nJ = cross(a1s, a2s);
        v3 nP = dvs(nJ, norm(nJ));
        double hs = side / 2.0;
        for (int b = 0; b < 4; b++) c[b] += hs * (dot(nx, cross(g0[b], nP)) / J);
        *s += side * (-dot(V, nP)) / 2 * dot(nx, nP);
    }
    return 0;
}

/* quad_field, src/elements/quad_field.jl:40-46 (polar = 1) */
static int quad_field(const quad_t *Q, v3 x, v3 nx, v3 V, int side, double polar, double c[4], double *s)
{
    double dc = norm(sub(x, Q->c)) - Q->reach;
    if (dc >= polar * Q->diam) { gauss_field(Q, x, nx, V, gauss_order(dc / Q->diam), c, s); return 0; }
    double P[2];
    double d = q_project(Q, x, P);
    if (d >= polar * Q->diam) { gauss_field(Q, x, nx, V, gauss_order(d / Q->diam), c, s); return 0; }
    return polar_field(Q, x, nx, V, P, d, side, c, s);
}

/* --------------------------------------------------------- square rules ------------------------------------- */
/* A SquareRule is held in a rule_t with the third coordinate zero (src/elements/quad_rules.jl:5-8). */

/* square_gauss, src/elements/quad_rules.jl:13-16 */
static void square_gauss(int m, rule_t *r)
{
    double x[64], w[64];
    gauss_legendre(m, x, w);
    rule_init(r);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) rule_push(r, x[i], x[j], 0.0, w[i] * w[j]);
}

/* turn, src/elements/quad_rules.jl:19-25 (k = 1..4) */
static void square_turn(const rule_t *r, int k, rule_t *o)
{
    rule_init(o);
    for (int i = 0; i < r->n; i++) {
        double p0 = r->l[3 * i], p1 = r->l[3 * i + 1];
        for (int t = 2; t <= k; t++) { double n0 = 1 - p1, n1 = p0; p0 = n0; p1 = n1; }
        rule_push(o, p0, p1, 0.0, r->w[i]);
    }
}

/* append_triangle!, src/elements/quad_rules.jl:28-34: bary rule r onto the parameter triangle (a, b, c) */
static void append_triangle(rule_t *o, const rule_t *r, const double a[2], const double b[2], const double c[2],
                            double frac)
{
    for (int i = 0; i < r->n; i++) {
        const double *l = r->l + 3 * i;
        double p0 = (l[0] * a[0] + l[1] * b[0]) + l[2] * c[0];
        double p1 = (l[0] * a[1] + l[1] * b[1]) + l[2] * c[1];
        rule_push(o, p0, p1, 0.0, r->w[i] * frac);
    }
}

/* square_edge_rule, src/elements/quad_rules.jl:42-46 */
static void square_edge_rule(int levels, double ratio, int order, rule_t *r)
{
    double es[2048], we[2048], xs[4096], wx[4096];
    int ne = graded_gauss(levels, ratio, order, es, we);
    int lv2 = levels / 2 > 1 ? levels / 2 : 1;
    int nx = graded_gauss_both(lv2, ratio, order, xs, wx);
    rule_init(r);
    for (int i = 0; i < ne; i++)
        for (int j = 0; j < nx; j++) rule_push(r, xs[j], es[i], 0.0, we[i] * wx[j]);
}

/* square_vertex_rule, src/elements/quad_rules.jl:49-55 */
static void square_vertex_rule(int levels, double ratio, int order, rule_t *r)
{
    rule_t base;
    vertex_graded_rule(levels, ratio, order, &base);
    rule_init(r);
    append_triangle(r, &base, SQUARE[0], SQUARE[1], SQUARE[2], 0.5);
    append_triangle(r, &base, SQUARE[0], SQUARE[2], SQUARE[3], 0.5);
    rule_free(&base);
}

/* square_self_rule, src/elements/quad_rules.jl:58-66 */
static void square_self_rule(int levels, double ratio, int order, rule_t *r)
{
    rule_t base;
    edge_graded_rule(levels, ratio, order, &base);
    rule_init(r);
    const double centre[2] = {0.5, 0.5};
    for (int k = 0; k < 4; k++) append_triangle(r, &base, SQUARE[k], SQUARE[(k + 1) % 4], centre, 0.25);
    rule_free(&base);
}

/* SquareTiers, src/elements/quad_rules.jl:73-87; square_tiers, src/assembly/mixed.jl:70-71 */
typedef struct { rule_t far, near, self, edge[4], vertex[4], vertex_base; } square_tiers_t;

static void square_tiers(const vpo_opts *o, int quad_levels, square_tiers_t *S)
{
    square_gauss(o->far_points <= 3 ? 2 : 3, &S->far);
    square_gauss(o->near_points <= 7 ? 3 : 4, &S->near);
    square_self_rule(quad_levels, o->touch_ratio, o->touch_order, &S->self);
    rule_t e, v;
    square_edge_rule(quad_levels, o->touch_ratio, o->touch_order, &e);
    square_vertex_rule(quad_levels, o->touch_ratio, o->touch_order, &v);
    for (int k = 0; k < 4; k++) { square_turn(&e, k + 1, &S->edge[k]); square_turn(&v, k + 1, &S->vertex[k]); }
    rule_free(&e);
    rule_free(&v);
    vertex_graded_rule(quad_levels, o->touch_ratio, o->touch_order, &S->vertex_base);
}

static void square_tiers_free(square_tiers_t *S)
{
    rule_free(&S->far); rule_free(&S->near); rule_free(&S->self); rule_free(&S->vertex_base);
    for (int k = 0; k < 4; k++) { rule_free(&S->edge[k]); rule_free(&S->vertex[k]); }
}

/* ------------------------------------------------------- generic sources ------------------------------------ */

/* A source of the mixed assembly: a flat triangle (kind 0), an isoparametric quad (kind 1) or a filament segment
 * (kind 2), with what the rule builders read of it: corners, centroid, diameter, reach and distance_to. */
typedef struct {
    int kind, nc;
    const el_t *tri;
    const quad_t *quad;
    v3 a, b;
} msrc_t;

static inline double msrc_distance(const msrc_t *S, v3 p)
{
    if (S->kind == 0) return tri_distance(S->tri->x, p);
    if (S->kind == 1) { double P[2]; return q_project(S->quad, p, P); }
    return seg_projection(p, S->a, S->b, NULL);
}
static inline v3 msrc_corner(const msrc_t *S, int k) { return S->kind == 0 ? S->tri->x[k] : S->quad->x[k]; }
static inline v3 msrc_centroid(const msrc_t *S) { return S->kind == 0 ? S->tri->c : S->quad->c; }
static inline double msrc_diam(const msrc_t *S) { return S->kind == 0 ? S->tri->diam : S->quad->diam; }
static inline double msrc_reach(const msrc_t *S) { return S->kind == 0 ? reach(S->tri) : S->quad->reach; }

/* ---------------------------------------------- rules on a test triangle (generic source) ------------------- */
/* near_pieces!, edge_splits, splits_an_edge, contact_pieces!, contact_rule and pair_rule of
 * src/assembly/assemble.jl:131-282 for a source that may be a quad (tri_pair_rule, src/assembly/mixed.jl:336-351) */

static void near_pieces_m(rule_t *out, const el_t *T, const msrc_t *S, const rule_t *base, double ratio, int depth,
                          const double b[3][3], double frac, int level)
{
    v3 p1 = el_point(T, b[0]), p2 = el_point(T, b[1]), p3 = el_point(T, b[2]);
    double size = dmax(dmax(norm(sub(p2, p1)), norm(sub(p3, p2))), norm(sub(p1, p3)));
    if (level < depth && msrc_distance(S, dvs(add(add(p1, p2), p3), 3.0)) < ratio * size) {
        double q[4][3][3];
        quarters(b, q);
        for (int i = 0; i < 4; i++) near_pieces_m(out, T, S, base, ratio, depth, (const double(*)[3])q[i], frac / 4, level + 1);
    } else {
        append_piece(out, base, b, frac);
    }
}

static int edge_splits_m(const v3 *p, int k, const msrc_t *S, double tol, double *ts)
{
    v3 a = p[k], c = p[(k + 1) % 3];
    double L = norm(sub(c, a));
    int n = 0;
    for (int q = 0; q < S->nc; q++) {
        double t;
        double d = seg_projection(msrc_corner(S, q), a, c, &t);
        if (d <= tol && t * L > tol && (1 - t) * L > tol) ts[n++] = t;
    }
    for (int i = 1; i < n; i++)
        for (int j = i; j > 0 && ts[j] < ts[j - 1]; j--) { double x = ts[j]; ts[j] = ts[j - 1]; ts[j - 1] = x; }
    return n;
}

static void contact_pieces_m(rule_t *out, const el_t *T, const msrc_t *S, const tier_rules_t *R, const vpo_opts *o,
                             double tol, const double b[3][3], double frac, int level)
{
    v3 p[3] = {el_point(T, b[0]), el_point(T, b[1]), el_point(T, b[2])};
    if (level < 4) {
        for (int k = 0; k < 3; k++) {
            double sp[4];
            int ns = edge_splits_m(p, k, S, tol, sp);
            if (ns == 0) continue;
            double ts[6];
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
                contact_pieces_m(out, T, S, R, o, tol, (const double(*)[3])piece, frac * (ts[i + 1] - ts[i]), level + 1);
            }
            return;
        }
    }
    int on[3], nc = 0;
    for (int k = 0; k < 3; k++) { on[k] = msrc_distance(S, p[k]) <= tol; nc += on[k]; }
    if (nc == 2) {
        int k = !on[2] ? 0 : !on[0] ? 1 : 2;
        if (msrc_distance(S, dvs(add(p[k], p[(k + 1) % 3]), 2.0)) <= tol) { append_piece(out, &R->edge[k], b, frac); return; }
    }
    if (nc == 1) {
        int k = on[0] ? 0 : on[1] ? 1 : 2;
        vertex_pieces(out, p, k, &R->vertex[0], b, frac, o->corner_angle);
    } else if (nc == 0) {
        near_pieces_m(out, T, S, &R->near, o->near_ratio, o->near_depth, b, frac, 0);
    } else {
        append_piece(out, &R->self, b, frac);
    }
}

static int contact_rule_m(const el_t *T, const msrc_t *S, const int shared[3], const tier_rules_t *R,
                          const vpo_opts *o, rule_t *out)
{
    double tol = o->contact_rel * dmin(T->diam, msrc_diam(S));
    if (norm(sub(msrc_centroid(S), T->c)) > reach(T) + msrc_reach(S) + tol) return 0;
    int extra = 0;
    for (int k = 0; k < 3 && !extra; k++) extra = !shared[k] && msrc_distance(S, T->x[k]) <= tol;
    for (int k = 0; k < 3 && !extra; k++) {
        double sp[4];
        extra = edge_splits_m(T->x, k, S, tol, sp) > 0;
    }
    if (!extra) return 0;
    rule_clear(out);
    contact_pieces_m(out, T, S, R, o, tol, CORNERS, 1.0, 0);
    return 1;
}

/* tri_pair_rule, src/assembly/mixed.jl:336-351 (tn: T's 3 nodes; sn: S's nn nodes) */
static const rule_t *tri_pair_rule(const el_t *T, const msrc_t *S, const int *tn, const int *sn, int nsn, int self,
                                   const tier_rules_t *R, const vpo_opts *o, int touch, rule_t *tmp, int *tier)
{
    if (self) { *tier = T_SELF; return &R->self; }
    if (norm(sub(msrc_centroid(S), T->c)) > o->far_ratio * dmax(T->diam, msrc_diam(S))) { *tier = T_FAR; return &R->far; }
    int shared[3];
    for (int k = 0; k < 3; k++) {
        shared[k] = 0;
        for (int j = 0; j < nsn; j++) shared[k] |= tn[k] == sn[j];
    }
    if (contact_rule_m(T, S, shared, R, o, tmp)) { *tier = T_CONTACT; return tmp; }
    int ns = shared[0] + shared[1] + shared[2];
    if (ns == 2) {
        *tier = T_EDGE;
        return &R->edge[shared[0] && shared[1] ? 0 : shared[1] && shared[2] ? 1 : 2];
    } else if (ns == 1) {
        *tier = T_VERTEX;
        return vertex_rule(T, shared[0] ? 0 : shared[1] ? 1 : 2, R, o, tmp);
    }
    *tier = T_NEAR;
    rule_clear(tmp);
    near_pieces_m(tmp, T, S, &R->near, o->near_ratio, touch ? o->near_depth + 4 : o->near_depth, CORNERS, 1.0, 0);
    return tmp;
}

/* ------------------------------------------------- rules on a test quad ------------------------------------- */

/* append_rect!, src/assembly/mixed.jl:74-81 */
static void append_rect(rule_t *o, const rule_t *r, const double lo[2], const double hi[2])
{
    double d0 = hi[0] - lo[0], d1 = hi[1] - lo[1];
    for (int i = 0; i < r->n; i++) rule_push(o, lo[0] + d0 * r->l[3 * i], lo[1] + d1 * r->l[3 * i + 1], 0.0, r->w[i] * d0 * d1);
}

/* near_squares!, src/elements/quad_rules.jl:103-117 */
static void near_squares(rule_t *o, const quad_t *Q, const msrc_t *S, const rule_t *base, double ratio, int depth,
                         const double lo[2], double size, int level)
{
    v3 c1 = q_point(Q, lo[0], lo[1]), c2 = q_point(Q, lo[0] + size, lo[1] + 0.0);
    v3 c3 = q_point(Q, lo[0] + size, lo[1] + size), c4 = q_point(Q, lo[0] + 0.0, lo[1] + size);
    double extent = dmax(norm(sub(c3, c1)), norm(sub(c4, c2)));
    if (level < depth && msrc_distance(S, q_point(Q, lo[0] + size / 2, lo[1] + size / 2)) < ratio * extent) {
        double h = size / 2;
        const double off[4][2] = {{0.0, 0.0}, {h, 0.0}, {0.0, h}, {h, h}};
        for (int i = 0; i < 4; i++) {
            double l2[2] = {lo[0] + off[i][0], lo[1] + off[i][1]};
            near_squares(o, Q, S, base, ratio, depth, l2, h, level + 1);
        }
    } else {
        for (int i = 0; i < base->n; i++)
            rule_push(o, lo[0] + size * base->l[3 * i], lo[1] + size * base->l[3 * i + 1], 0.0, base->w[i] * (size * size));
    }
}

/* square_near_rule, src/elements/quad_rules.jl:96-100 */
static void square_near_rule(rule_t *o, const quad_t *Q, const msrc_t *S, const rule_t *base, double ratio, int depth)
{
    rule_clear(o);
    const double lo[2] = {0.0, 0.0};
    near_squares(o, Q, S, base, ratio, depth, lo, 1.0, 0);
}

/* near_rects!, src/assembly/mixed.jl:84-96 */
static void near_rects(rule_t *o, const quad_t *T, const msrc_t *S, const rule_t *base, double ratio, int depth,
                       const double lo[2], const double hi[2], int level)
{
    v3 c1 = q_point(T, lo[0], lo[1]), c2 = q_point(T, hi[0], lo[1]);
    v3 c3 = q_point(T, hi[0], hi[1]), c4 = q_point(T, lo[0], hi[1]);
    double extent = dmax(norm(sub(c3, c1)), norm(sub(c4, c2)));
    double m[2] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2};
    if (level < depth && msrc_distance(S, q_point(T, m[0], m[1])) < ratio * extent) {
        const double r[4][4] = {{lo[0], lo[1], m[0], m[1]}, {m[0], lo[1], hi[0], m[1]},
                                {lo[0], m[1], m[0], hi[1]}, {m[0], m[1], hi[0], hi[1]}};
        for (int i = 0; i < 4; i++) near_rects(o, T, S, base, ratio, depth, r[i], r[i] + 2, level + 1);
    } else {
        append_rect(o, base, lo, hi);
    }
}

/* ray_exit, src/assembly/mixed.jl:99-106 */
static void ray_exit(const double c[2], const double d[2], const double lo[2], const double hi[2], double out[2])
{
    double t = INFINITY;
    for (int i = 0; i < 2; i++) {
        if (d[i] > 1e-14) t = dmin(t, (hi[i] - c[i]) / d[i]);
        if (d[i] < -1e-14) t = dmin(t, (lo[i] - c[i]) / d[i]);
    }
    out[0] = c[0] + t * d[0];
    out[1] = c[1] + t * d[1];
}

static inline double cross2(double a0, double a1, double b0, double b1) { return a0 * b1 - a1 * b0; }

static double along(const double P[2], const double X[2], const double Y[2])
{
    double yx0 = Y[0] - X[0], yx1 = Y[1] - X[1];
    double t = ((P[0] - X[0]) * yx0 + (P[1] - X[1]) * yx1) / (yx0 * yx0 + yx1 * yx1);
    return t < 0.0 ? 0.0 : t > 1.0 ? 1.0 : t;
}

/* vertex_rect!, src/assembly/mixed.jl:118-151 (k 0-based) */
static void vertex_rect(rule_t *o, const quad_t *T, int k, const double lo[2], const double hi[2], const rule_t *base,
                        double corner_angle)
{
    double rc[4][2] = {{lo[0], lo[1]}, {hi[0], lo[1]}, {hi[0], hi[1]}, {lo[0], hi[1]}};
    const double *C = rc[k], *A = rc[(k + 1) % 4], *O = rc[(k + 2) % 4], *B = rc[(k + 3) % 4];
    v3 a1, a2;
    q_tangents(T, C[0], C[1], &a1, &a2);
    v3 u = add(scl(A[0] - C[0], a1), scl(A[1] - C[1], a2));
    v3 v = add(scl(B[0] - C[0], a1), scl(B[1] - C[1], a2));
    double cs = dot(u, v) / (norm(u) * norm(v));
    if (cs < -1.0) cs = -1.0;
    if (cs > 1.0) cs = 1.0;
    double th = acos(cs);
    int m = th <= PI / 2 ? 1 : (int)ceil(th * (180 / PI) / corner_angle);
    v3 uh = dvs(u, norm(u));
    v3 wv = sub(v, scl(dot(v, uh), uh));
    v3 wh = dvs(wv, norm(wv));
    double G11 = dot(a1, a1), G12 = dot(a1, a2), G22 = dot(a2, a2);
    double det = G11 * G22 - G12 * G12;
    double pts[66][2];
    for (int i = 0; i <= m; i++) {
        if (i == 0) { pts[i][0] = A[0]; pts[i][1] = A[1]; continue; }
        if (i == m) { pts[i][0] = B[0]; pts[i][1] = B[1]; continue; }
        v3 d = add(scl(cos(i * th / m), uh), scl(sin(i * th / m), wh));
        double b0 = dot(a1, d), b1 = dot(a2, d);
        double dir[2] = {(G22 * b0 - G12 * b1) / det, (G11 * b1 - G12 * b0) / det};
        ray_exit(C, dir, lo, hi, pts[i]);
    }
    double tol = 1e-12 * sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]));
    double OA0 = O[0] - A[0], OA1 = O[1] - A[1];
    double nOA = sqrt(OA0 * OA0 + OA1 * OA1);
    for (int i = 0; i < m; i++) {
        const double *P = pts[i], *Qp = pts[i + 1];
        double sP = fabs(cross2(P[0] - A[0], P[1] - A[1], OA0, OA1)) / nOA <= tol ? along(P, A, O) : 1 + along(P, O, B);
        double sQ = fabs(cross2(Qp[0] - A[0], Qp[1] - A[1], OA0, OA1)) / nOA <= tol ? along(Qp, A, O) : 1 + along(Qp, O, B);
        if (sP < 1 - 1e-9 && sQ > 1 + 1e-9) {
            append_triangle(o, base, C, P, O, fabs(cross2(P[0] - C[0], P[1] - C[1], O[0] - C[0], O[1] - C[1])) / 2);
            append_triangle(o, base, C, O, Qp, fabs(cross2(O[0] - C[0], O[1] - C[1], Qp[0] - C[0], Qp[1] - C[1])) / 2);
        } else {
            append_triangle(o, base, C, P, Qp, fabs(cross2(P[0] - C[0], P[1] - C[1], Qp[0] - C[0], Qp[1] - C[1])) / 2);
        }
    }
}

/* vertex_square, src/assembly/mixed.jl:154-160 (k 0-based) */
static const rule_t *vertex_square(const quad_t *T, int k, const square_tiers_t *ST, const vpo_opts *o, rule_t *tmp)
{
    v3 u = sub(T->x[(k + 1) % 4], T->x[k]), v = sub(T->x[(k + 3) % 4], T->x[k]);
    if (dot(u, v) >= 0) return &ST->vertex[k];
    rule_clear(tmp);
    const double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0};
    vertex_rect(tmp, T, k, lo, hi, &ST->vertex_base, o->corner_angle);
    return tmp;
}

/* side_splits, src/assembly/mixed.jl:166-179: the splits of side k into ts (unsorted); returns their count */
static int side_splits_k(const quad_t *T, int k, const msrc_t *S, double tol, double *ts)
{
    v3 a = T->x[k], c = T->x[(k + 1) % 4];
    double L = norm(sub(c, a));
    int n = 0;
    for (int q = 0; q < S->nc; q++) {
        double t;
        double d = seg_projection(msrc_corner(S, q), a, c, &t);
        if (d <= tol && t * L > tol && (1 - t) * L > tol) ts[n++] = t;
    }
    return n;
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* cuts, src/assembly/mixed.jl:182-190 */
static int cuts(double *v, int n)
{
    qsort(v, n, sizeof(double), cmp_double);
    int m = 1;
    for (int i = 1; i < n; i++)
        if (v[i] - v[m - 1] > 1e-9) v[m++] = v[i];
    return m;
}

/* square_contact_rule, src/assembly/mixed.jl:203-232 */
static int square_contact_rule(const quad_t *T, const msrc_t *S, const int shared[4], const square_tiers_t *ST,
                               const vpo_opts *o, rule_t *out)
{
    double tol = o->contact_rel * dmin(T->diam, msrc_diam(S));
    if (norm(sub(msrc_centroid(S), T->c)) > T->reach + msrc_reach(S) + tol) return 0;
    double sp[4][4];
    int ns[4];
    for (int k = 0; k < 4; k++) ns[k] = side_splits_k(T, k, S, tol, sp[k]);
    int extra = 0;
    for (int k = 0; k < 4 && !extra; k++) extra = !shared[k] && msrc_distance(S, T->x[k]) <= tol;
    for (int k = 0; k < 4 && !extra; k++) extra = ns[k] > 0;
    if (!extra) return 0;
    double xc[12], ec[12];
    int nx = 0, ne = 0;
    xc[nx++] = 0.0;
    for (int i = 0; i < ns[0]; i++) xc[nx++] = sp[0][i];
    for (int i = 0; i < ns[2]; i++) xc[nx++] = 1 - sp[2][i];
    xc[nx++] = 1.0;
    ec[ne++] = 0.0;
    for (int i = 0; i < ns[1]; i++) ec[ne++] = sp[1][i];
    for (int i = 0; i < ns[3]; i++) ec[ne++] = 1 - sp[3][i];
    ec[ne++] = 1.0;
    nx = cuts(xc, nx);
    ne = cuts(ec, ne);
    rule_clear(out);
    for (int i = 0; i < nx - 1; i++)
        for (int j = 0; j < ne - 1; j++) {
            double lo[2] = {xc[i], ec[j]}, hi[2] = {xc[i + 1], ec[j + 1]};
            double c[4][2] = {{lo[0], lo[1]}, {hi[0], lo[1]}, {hi[0], hi[1]}, {lo[0], hi[1]}};
            int on[4], nc = 0;
            for (int q = 0; q < 4; q++) { on[q] = msrc_distance(S, q_point(T, c[q][0], c[q][1])) <= tol; nc += on[q]; }
            int k = -1;
            for (int q = 0; q < 4 && k < 0; q++) if (on[q] && on[(q + 1) % 4]) k = q;
            if (nc == 2 && k >= 0 &&
                msrc_distance(S, q_point(T, (c[k][0] + c[(k + 1) % 4][0]) / 2, (c[k][1] + c[(k + 1) % 4][1]) / 2)) <= tol) {
                append_rect(out, &ST->edge[k], lo, hi);
            } else if (nc == 1) {
                int f = on[0] ? 0 : on[1] ? 1 : on[2] ? 2 : 3;
                vertex_rect(out, T, f, lo, hi, &ST->vertex_base, o->corner_angle);
            } else if (nc == 0) {
                near_rects(out, T, S, &ST->near, o->near_ratio, o->near_depth + 2, lo, hi, 0);
            } else {
                append_rect(out, &ST->self, lo, hi);
            }
        }
    return 1;
}

/* quad_pair_rule, src/assembly/mixed.jl:238-255 (tn: T's 4 nodes; sn: S's nsn nodes) */
static const rule_t *quad_pair_rule(const quad_t *T, const msrc_t *S, const int *tn, const int *sn, int nsn, int self,
                                    const square_tiers_t *ST, const vpo_opts *o, int touch, rule_t *tmp, int *tier)
{
    if (self) { *tier = T_SELF; return &ST->self; }
    if (norm(sub(msrc_centroid(S), T->c)) > o->far_ratio * dmax(T->diam, msrc_diam(S))) { *tier = T_FAR; return &ST->far; }
    int shared[4];
    for (int k = 0; k < 4; k++) {
        shared[k] = 0;
        for (int j = 0; j < nsn; j++) shared[k] |= tn[k] == sn[j];
    }
    if (square_contact_rule(T, S, shared, ST, o, tmp)) { *tier = T_CONTACT; return tmp; }
    int ns = shared[0] + shared[1] + shared[2] + shared[3];
    if (ns == 2) {
        *tier = T_EDGE;
        for (int k = 0; k < 4; k++) if (shared[k] && shared[(k + 1) % 4]) return &ST->edge[k];
        return &ST->self;
    } else if (ns == 1) {
        *tier = T_VERTEX;
        return vertex_square(T, shared[0] ? 0 : shared[1] ? 1 : shared[2] ? 2 : 3, ST, o, tmp);
    } else if (ns > 2) {
        *tier = T_EDGE;
        return &ST->self;
    }
    *tier = T_NEAR;
    square_near_rule(tmp, T, S, &ST->near, o->near_ratio, touch ? o->near_depth + 4 : o->near_depth);
    return tmp;
}

/* quad_wake_rule, src/assembly/mixed.jl:354-359 */
static const rule_t *quad_wake_rule(const quad_t *T, const msrc_t *S, const square_tiers_t *ST, const vpo_opts *o,
                                    rule_t *tmp, int *tier)
{
    if (norm(sub(msrc_centroid(S), T->c)) > o->far_ratio * dmax(T->diam, msrc_diam(S))) { *tier = T_FAR; return &ST->far; }
    const int shared[4] = {0, 0, 0, 0};
    if (square_contact_rule(T, S, shared, ST, o, tmp)) { *tier = T_CONTACT; return tmp; }
    *tier = T_NEAR;
    square_near_rule(tmp, T, S, &ST->near, o->near_ratio, o->near_depth);
    return tmp;
}

/* quad_segment_rule, src/assembly/mixed.jl:362-369 */
static const rule_t *quad_segment_rule(const quad_t *T, v3 fa, v3 fb, const square_tiers_t *ST, const vpo_opts *o,
                                       rule_t *tmp, int *tier)
{
    double L = norm(sub(fb, fa));
    if (norm(sub(dvs(add(fa, fb), 2.0), T->c)) > o->far_ratio * dmax(T->diam, L)) { *tier = T_FAR; return &ST->far; }
    for (int k = 0; k < 4; k++)
        if (dmin(norm(sub(T->x[k], fa)), norm(sub(T->x[k], fb))) <= 1e-12 * T->diam) {
            *tier = T_VERTEX;
            return vertex_square(T, k, ST, o, tmp);
        }
    msrc_t f = {2, 0, NULL, NULL, fa, fb};
    *tier = T_NEAR;
    square_near_rule(tmp, T, &f, &ST->near, o->near_ratio, o->near_depth);
    return tmp;
}

/* ---------------------------------------------------------------- blocks ------------------------------------ */

/* square_test_weights, src/assembly/mixed.jl:62-67 */
static inline void square_test_weights(int test, double xi, double eta, double psi[4])
{
    if (test == 0) { q_shape(xi, eta, psi); return; }
    if (test == 1) { psi[0] = psi[1] = psi[2] = psi[3] = 1.0; return; }
    int k = xi < 0.5 ? (eta < 0.5 ? 0 : 3) : (eta < 0.5 ? 1 : 2);
    for (int i = 0; i < 4; i++) psi[i] = i == k ? 1.0 : 0.0;
}

/* block(T::IsoQuad, S::IsoQuad), src/assembly/mixed.jl:376-388: blk[a][b] (a test corner, b source corner) */
static int block_qq(const quad_t *T, const quad_t *S, const rule_t *r, v3 V, int test, int side, double blk[4][4],
                    double rb[4])
{
    memset(blk, 0, sizeof(double) * 16);
    for (int a = 0; a < 4; a++) rb[a] = 0;
    for (int q = 0; q < r->n; q++) {
        double xi = r->l[3 * q], eta = r->l[3 * q + 1], wq = r->w[q];
        v3 nJ = q_area_normal(T, xi, eta);
        double J = norm(nJ);
        v3 nx = dvs(nJ, J);
        double c[4], s;
        int rc;
        if (side != 0) { double P[2] = {xi, eta}; rc = polar_field(S, q_point(T, xi, eta), nx, V, P, 0.0, side, c, &s); }
        else rc = quad_field(S, q_point(T, xi, eta), nx, V, 0, 1.0, c, &s);
        if (rc) return rc;
        double tw[4], psi[4];
        square_test_weights(test, xi, eta, tw);
        for (int a = 0; a < 4; a++) psi[a] = (wq * J) * tw[a];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) blk[a][b] += psi[a] * c[b];
        for (int a = 0; a < 4; a++) rb[a] -= s * psi[a];
    }
    return 0;
}

/* block(T::IsoQuad, S::FlatTriangle), src/assembly/mixed.jl:391-405 */
static void block_qt(const quad_t *T, const el_t *S, const rule_t *r, double sig, int test, double kfar,
                     double blk[4][3], double rb[4])
{
    memset(blk, 0, sizeof(double) * 12);
    for (int a = 0; a < 4; a++) rb[a] = 0;
    for (int q = 0; q < r->n; q++) {
        double xi = r->l[3 * q], eta = r->l[3 * q + 1], wq = r->w[q];
        v3 nJ = q_area_normal(T, xi, eta);
        double J = norm(nJ);
        v3 nx = dvs(nJ, J);
        v3 w = sv(&S->src, q_point(T, xi, eta), 0, kfar, 1e-12);
        v3 v = cross(w, nx);
        double tw[4], psi[4];
        square_test_weights(test, xi, eta, tw);
        for (int a = 0; a < 4; a++) psi[a] = (wq * J) * tw[a];
        double cb[3] = {dot(v, S->g[0]), dot(v, S->g[1]), dot(v, S->g[2])};
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 3; b++) blk[a][b] += psi[a] * cb[b];
        double f = sig * dot(w, nx);
        for (int a = 0; a < 4; a++) rb[a] -= f * psi[a];
    }
}

/* block(T::FlatTriangle, S::IsoQuad), src/assembly/mixed.jl:408-418 */
static int block_tq(const el_t *T, const quad_t *S, const rule_t *r, v3 V, int test, double blk[3][4], double rb[3])
{
    memset(blk, 0, sizeof(double) * 12);
    for (int a = 0; a < 3; a++) rb[a] = 0;
    for (int q = 0; q < r->n; q++) {
        const double *l = r->l + 3 * q;
        double c[4], s;
        int rc = quad_field(S, el_point(T, l), T->n, V, 0, 1.0, c, &s);
        if (rc) return rc;
        double tw[3], psi[3];
        test_weights(test, l, tw);
        for (int a = 0; a < 3; a++) psi[a] = (r->w[q] * T->area) * tw[a];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 4; b++) blk[a][b] += psi[a] * c[b];
        for (int a = 0; a < 3; a++) rb[a] -= s * psi[a];
    }
    return 0;
}

/* segment_block(T::IsoQuad), src/assembly/mixed.jl:421-432 */
static void segment_block_q(const quad_t *T, v3 fa, v3 fb, const rule_t *r, int test, double ca[4], double cb[4])
{
    for (int a = 0; a < 4; a++) ca[a] = cb[a] = 0;
    for (int q = 0; q < r->n; q++) {
        double xi = r->l[3 * q], eta = r->l[3 * q + 1], wq = r->w[q];
        v3 nJ = q_area_normal(T, xi, eta);
        double J = norm(nJ);
        v3 va, vb;
        seg_vel(fa, fb, q_point(T, xi, eta), &va, &vb);
        double tw[4], psi[4];
        square_test_weights(test, xi, eta, tw);
        for (int a = 0; a < 4; a++) psi[a] = (wq * J) * tw[a];
        v3 nx = dvs(nJ, J);
        double da = dot(va, nx), db = dot(vb, nx);
        for (int a = 0; a < 4; a++) { ca[a] += da * psi[a]; cb[a] += db * psi[a]; }
    }
}

/* ------------------------------------------------------------- public ----------------------------------------- */

/* The inner integrals of one quad (tests): quad_field(Q, x, nx, V; side, polar) of src/elements/quad_field.jl:40-46,
 * or, with P given (P != NULL), polar_field about P at distance d.  qx: 4 x 3 corners.  Returns 0, -1 (x on the
 * quad's boundary) or -2 (degenerate quad). */
int vpo_quad_field(const double *qx, const double *x, const double *nx, const double *V, int side, double polar,
                   const double *P, double d, double *c, double *s)
{
    quad_t Q;
    if (make_quad(ld3(qx), ld3(qx + 3), ld3(qx + 6), ld3(qx + 9), &Q)) return -2;
    if (P) return polar_field(&Q, ld3(x), ld3(nx), ld3(V), P, d, side, c, s);
    return quad_field(&Q, ld3(x), ld3(nx), ld3(V), side, polar, c, s);
}

/* project (closest point) and the IsoQuad's derived data: out = [area, diam, reach, cx, cy, cz, nx, ny, nz] */
int vpo_quad_project(const double *qx, const double *p, double *P, double *d, double *out)
{
    quad_t Q;
    if (make_quad(ld3(qx), ld3(qx + 3), ld3(qx + 6), ld3(qx + 9), &Q)) return -2;
    if (p) *d = q_project(&Q, ld3(p), P);
    if (out) {
        out[0] = Q.area; out[1] = Q.diam; out[2] = Q.reach;
        st3(out + 3, Q.c);
        st3(out + 6, Q.n);
    }
    return 0;
}

/* Square tier rules for tests: kind 0 far, 1 near, 2 self, 3..6 edge k, 7..10 vertex k (stored), 11 vertex_base
 * (barycentric, 3 coordinates per point); returns the count, fills when the buffers are given (2 or 3 per point). */
int vpo_square_rule(const vpo_opts *o, int quad_levels, int kind, double *p, double *w)
{
    square_tiers_t ST;
    square_tiers(o, quad_levels, &ST);
    const rule_t *r = kind == 0 ? &ST.far : kind == 1 ? &ST.near : kind == 2 ? &ST.self : kind <= 6 ? &ST.edge[kind - 3]
                    : kind <= 10 ? &ST.vertex[kind - 7] : &ST.vertex_base;
    int n = r->n, stride = kind == 11 ? 3 : 2;
    for (int i = 0; i < n; i++) {
        if (p) for (int j = 0; j < stride; j++) p[stride * i + j] = r->l[3 * i + j];
        if (w) w[i] = r->w[i];
    }
    square_tiers_free(&ST);
    return n;
}

/* One pair's tier and rule on a test quad (tests): kind 0 source quad (sx 4 x 3, sn 4 nodes), 1 source triangle
 * (sx 3 x 3, sn 3), 2 wake triangle (sx 3 x 3), 3 filament (sx 2 x 3); 4 vertex_square of corner `self` alone, 5
 * square_contact_rule alone (shared given by tn/sn).  The rule's points go to p (2 per point) and w when given;
 * returns the point count (0 for no contact rule), -2 for a degenerate element. */
int vpo_quad_rule(const double *tx, const int *tn, int kind, const double *sx, const int *sn, int self, int touch,
                  const vpo_opts *o, int quad_levels, int *tier, double *p, double *w, int cap)
{
    quad_t T, SQ;
    el_t ST;
    if (make_quad(ld3(tx), ld3(tx + 3), ld3(tx + 6), ld3(tx + 9), &T)) return -2;
    square_tiers_t Ts;
    square_tiers(o, quad_levels, &Ts);
    rule_t tmp;
    rule_init(&tmp);
    const rule_t *r = NULL;
    msrc_t S = {0, 3, NULL, NULL, {0, 0, 0}, {0, 0, 0}};
    int rc = 0;
    if (kind == 0 || kind == 5) {
        if (make_quad(ld3(sx), ld3(sx + 3), ld3(sx + 6), ld3(sx + 9), &SQ)) rc = -2;
        S.kind = 1; S.nc = 4; S.quad = &SQ;
    } else if (kind == 1 || kind == 2) {
        if (make_el(ld3(sx), ld3(sx + 3), ld3(sx + 6), &ST)) rc = -2;
        S.tri = &ST;
    }
    if (!rc) {
        if (kind == 0) r = quad_pair_rule(&T, &S, tn, sn, 4, self, &Ts, o, touch, &tmp, tier);
        else if (kind == 1) r = quad_pair_rule(&T, &S, tn, sn, 3, self, &Ts, o, touch, &tmp, tier);
        else if (kind == 2) r = quad_wake_rule(&T, &S, &Ts, o, &tmp, tier);
        else if (kind == 3) r = quad_segment_rule(&T, ld3(sx), ld3(sx + 3), &Ts, o, &tmp, tier);
        else if (kind == 4) { r = vertex_square(&T, self, &Ts, o, &tmp); *tier = T_VERTEX; }
        else {
            int shared[4];
            for (int k = 0; k < 4; k++) { shared[k] = 0; for (int j = 0; j < 4; j++) shared[k] |= tn[k] == sn[j]; }
            r = square_contact_rule(&T, &S, shared, &Ts, o, &tmp) ? &tmp : NULL;
            *tier = T_CONTACT;
        }
    }
    int n = rc ? rc : r ? r->n : 0;
    if (r && n > 0)
        for (int i = 0; i < n && i < cap; i++) {
            if (p) { p[2 * i] = r->l[3 * i]; p[2 * i + 1] = r->l[3 * i + 1]; }
            if (w) w[i] = r->w[i];
        }
    rule_free(&tmp);
    square_tiers_free(&Ts);
    return n;
}

/* The pair block of source quad S on test quad T with a given square rule (tests): blk 4 x 4 row-major (a test
 * corner, b source corner) and the source part of the right hand side, rb[4] (src/assembly/mixed.jl:376-388). */
int vpo_quad_block(const double *tx, const double *sx, int n, const double *p, const double *w, const double *V,
                   int test, int side, double *blk, double *rb)
{
    quad_t T, S;
    if (make_quad(ld3(tx), ld3(tx + 3), ld3(tx + 6), ld3(tx + 9), &T)) return -2;
    if (make_quad(ld3(sx), ld3(sx + 3), ld3(sx + 6), ld3(sx + 9), &S)) return -2;
    rule_t r;
    rule_init(&r);
    for (int i = 0; i < n; i++) rule_push(&r, p[2 * i], p[2 * i + 1], 0.0, w[i]);
    int rc = block_qq(&T, &S, &r, ld3(V), test, side, (double(*)[4])blk, rb);
    rule_free(&r);
    return rc;
}

/*
 * assemble(g, dm, mesh::MixedMesh, V; opts, wake), src/assembly/mixed.jl:444-505, with test_quad! (:508-560) and
 * test_tri! (:563-598).  Elements: nq quads (quad_nodes, quad_slots: nq x 4) first, then nt triangles (tri_nodes,
 * tri_slots: nt x 3).  touch: ntouch element pairs (0-based element ids, quads first), made symmetric here.  A is
 * column-major nslot x (nslot + nwx) as in vpo_assemble; tiers (optional) ne x (ne + nwt) bytes.  Only the rows of
 * test elements [e_begin, e_end) are assembled (the bounded sample the CPU baseline times; 0, ne for all).
 * Returns 0, -1 degenerate element, -2 bad index, -3 bad option, -4 a test point on a source quad's boundary.
 */
int vpo_assemble_mixed(int nn, const double *x, int nq, const int *quad_nodes, const int *quad_slots, int nt,
                       const int *tri_nodes, const int *tri_slots, int nslot, const double *vinf, int ntouch,
                       const int *touch, int nwx, const double *wx, int nwt, const int *wtris, int nwf,
                       const int *wfil, const vpo_opts *o, int quad_levels, int nthreads, double *A, double *bvec,
                       long long *pairs, long long *points, unsigned char *tiers, int e_begin, int e_end)
{
    if (o->test < 0 || o->test > 2) return -3;
    tier_rules_t R;
    if (tier_rules(o, &R)) return -3;
    square_tiers_t ST;
    square_tiers(o, quad_levels, &ST);
    gauss_table();
    int ne = nq + nt;
    quad_t *qs = (quad_t *)malloc(sizeof(quad_t) * (nq > 0 ? nq : 1));
    el_t *ts = (el_t *)malloc(sizeof(el_t) * (nt > 0 ? nt : 1));
    el_t *wel = (el_t *)malloc(sizeof(el_t) * (nwt > 0 ? nwt : 1));
    int rc = 0;
    for (int e = 0; e < nq && !rc; e++) {
        const int *qn = quad_nodes + 4 * e;
        for (int k = 0; k < 4; k++) if (qn[k] < 0 || qn[k] >= nn) rc = -2;
        if (!rc && make_quad(ld3(x + 3 * qn[0]), ld3(x + 3 * qn[1]), ld3(x + 3 * qn[2]), ld3(x + 3 * qn[3]), qs + e)) rc = -1;
    }
    if (!rc) rc = build_els(nn, x, nt, tri_nodes, ts);
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
    if (rc) { free(qs); free(ts); free(wel); free(tp); tier_rules_free(&R); square_tiers_free(&ST); return rc; }
    long long ncol = (long long)nslot + nwx;
    memset(A, 0, sizeof(double) * (size_t)nslot * (size_t)ncol);
    memset(bvec, 0, sizeof(double) * nslot);
    for (int k = 0; k < 6; k++) pairs[k] = points[k] = 0;
    v3 V = ld3(vinf);
    double *sig = (double *)malloc(sizeof(double) * (nt > 0 ? nt : 1));
    for (int s = 0; s < nt; s++) sig[s] = -dot(V, ts[s].n);
    double ti[3] = {o->test == 1 ? 1.0 : 1.0 / 3, o->test == 1 ? 1.0 : 1.0 / 3, o->test == 1 ? 1.0 : 1.0 / 3};
    int nodeoff = nn;
    rule_t g3;
    square_gauss(3, &g3);
    int fail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 4 * (size_t)ncol);
        long long mp[6] = {0}, mq[6] = {0};
        rule_t tmp;
        rule_init(&tmp);
#pragma omp for schedule(dynamic, 1)
        for (int e = e_begin < 0 ? 0 : e_begin; e < (e_end > ne ? ne : e_end); e++) {
            if (fail) continue;
            memset(buf, 0, sizeof(double) * 4 * (size_t)ncol);
            double rows[4] = {0, 0, 0, 0};
            int na;
            const int *slots;
            int erc = 0;
            if (e < nq) {
                const quad_t *T = qs + e;
                const int *tn = quad_nodes + 4 * e;
                na = 4;
                slots = quad_slots + 4 * e;
                for (int s = 0; s < nq && !erc; s++) {
                    msrc_t S = {1, 4, NULL, qs + s, {0, 0, 0}, {0, 0, 0}};
                    int tier;
                    const rule_t *r = quad_pair_rule(T, &S, tn, quad_nodes + 4 * s, 4, s == e, &ST, o,
                                                     is_touch(tp, 2 * ntouch, e, s), &tmp, &tier);
                    double blk[4][4], rb[4];
                    erc = block_qq(T, qs + s, r, V, o->test, s == e ? 1 : 0, blk, rb);
                    const int *cs = quad_slots + 4 * s;
                    for (int bb = 0; bb < 4; bb++)
                        for (int a = 0; a < 4; a++) buf[(size_t)a * ncol + cs[bb]] += blk[a][bb];
                    for (int a = 0; a < 4; a++) rows[a] += rb[a];
                    mp[tier]++;
                    mq[tier] += r->n;
                    if (tiers) tiers[(size_t)e * (ne + nwt) + s] = (unsigned char)tier;
                }
                for (int s = 0; s < nt && !erc; s++) {
                    msrc_t S = {0, 3, ts + s, NULL, {0, 0, 0}, {0, 0, 0}};
                    int tier;
                    const rule_t *r = quad_pair_rule(T, &S, tn, tri_nodes + 3 * s, 3, 0, &ST, o,
                                                     is_touch(tp, 2 * ntouch, e, nq + s), &tmp, &tier);
                    double blk[4][3], rb[4];
                    block_qt(T, ts + s, r, sig[s], o->test, o->kernel_far, blk, rb);
                    const int *cs = tri_slots + 3 * s;
                    for (int bb = 0; bb < 3; bb++)
                        for (int a = 0; a < 4; a++) buf[(size_t)a * ncol + cs[bb]] += blk[a][bb];
                    for (int a = 0; a < 4; a++) rows[a] += rb[a];
                    mp[tier]++;
                    mq[tier] += r->n;
                    if (tiers) tiers[(size_t)e * (ne + nwt) + nq + s] = (unsigned char)tier;
                }
                for (int s = 0; s < nwt; s++) {
                    msrc_t S = {0, 3, wel + s, NULL, {0, 0, 0}, {0, 0, 0}};
                    int tier;
                    const rule_t *r = quad_wake_rule(T, &S, &ST, o, &tmp, &tier);
                    double blk[4][3], rb[4];
                    block_qt(T, wel + s, r, 0.0, o->test, o->kernel_far, blk, rb);
                    const int *c = wtris + 3 * s;
                    for (int bb = 0; bb < 3; bb++)
                        for (int a = 0; a < 4; a++) buf[(size_t)a * ncol + nslot + c[bb]] += blk[a][bb];
                    mp[tier]++;
                    mq[tier] += r->n;
                    if (tiers) tiers[(size_t)e * (ne + nwt) + ne + s] = (unsigned char)tier;
                }
                for (int f = 0; f < nwf; f++) {
                    int ia = wfil[2 * f], ib = wfil[2 * f + 1];
                    v3 fa = ld3(wx + 3 * ia), fb = ld3(wx + 3 * ib);
                    int tier;
                    const rule_t *r = quad_segment_rule(T, fa, fb, &ST, o, &tmp, &tier);
                    double ca[4], cb[4];
                    segment_block_q(T, fa, fb, r, o->test, ca, cb);
                    for (int a = 0; a < 4; a++) {
                        buf[(size_t)a * ncol + nslot + ia] += ca[a];
                        buf[(size_t)a * ncol + nslot + ib] += cb[a];
                    }
                    mp[tier]++;
                    mq[tier] += r->n;
                }
                for (int q = 0; q < g3.n; q++) { /* freestream, exact on the 3 x 3 Gauss rule */
                    double xi = g3.l[3 * q], eta = g3.l[3 * q + 1];
                    v3 nJ = q_area_normal(T, xi, eta);
                    double tw[4];
                    square_test_weights(o->test, xi, eta, tw);
                    double f = g3.w[q] * dot(V, nJ);
                    for (int a = 0; a < 4; a++) rows[a] -= f * tw[a];
                }
            } else {
                int t = e - nq;
                const el_t *T = ts + t;
                const int *tn = tri_nodes + 3 * t;
                na = 3;
                slots = tri_slots + 3 * t;
                for (int s = 0; s < nq && !erc; s++) {
                    msrc_t S = {1, 4, NULL, qs + s, {0, 0, 0}, {0, 0, 0}};
                    int tier;
                    const rule_t *r = tri_pair_rule(T, &S, tn, quad_nodes + 4 * s, 4, 0, &R, o,
                                                    is_touch(tp, 2 * ntouch, e, s), &tmp, &tier);
                    double blk[3][4], rb[3];
                    erc = block_tq(T, qs + s, r, V, o->test, blk, rb);
                    const int *cs = quad_slots + 4 * s;
                    for (int bb = 0; bb < 4; bb++)
                        for (int a = 0; a < 3; a++) buf[(size_t)a * ncol + cs[bb]] += blk[a][bb];
                    for (int a = 0; a < 3; a++) rows[a] += rb[a];
                    mp[tier]++;
                    mq[tier] += r->n;
                    if (tiers) tiers[(size_t)e * (ne + nwt) + s] = (unsigned char)tier;
                }
                for (int s = 0; s < nt; s++) {
                    msrc_t S = {0, 3, ts + s, NULL, {0, 0, 0}, {0, 0, 0}};
                    int tier;
                    const rule_t *r = tri_pair_rule(T, &S, tn, tri_nodes + 3 * s, 3, s == t, &R, o,
                                                    is_touch(tp, 2 * ntouch, e, nq + s), &tmp, &tier);
                    double blk[3][3], rb[3];
                    pair_block(T, ts + s, r, sig[s], o->test, s == t ? 1 : 0, o->kernel_far, blk, rb);
                    const int *cs = tri_slots + 3 * s;
                    for (int bb = 0; bb < 3; bb++)
                        for (int a = 0; a < 3; a++) buf[(size_t)a * ncol + cs[bb]] += blk[a][bb];
                    for (int a = 0; a < 3; a++) rows[a] += rb[a];
                    mp[tier]++;
                    mq[tier] += r->n;
                    if (tiers) tiers[(size_t)e * (ne + nwt) + nq + s] = (unsigned char)tier;
                }
                /* add_wake!, src/assembly/assemble.jl:333-356 */
                for (int s = 0; s < nwt; s++) {
                    const int *wt = wtris + 3 * s;
                    int gs[3] = {wt[0] + nodeoff, wt[1] + nodeoff, wt[2] + nodeoff};
                    int tier;
                    const rule_t *r = pair_rule(T, wel + s, t, -1 - s, tn, gs, &R, o, 0, &tmp, &tier);
                    double blk[3][3], rb[3];
                    pair_block(T, wel + s, r, 0.0, o->test, 0, o->kernel_far, blk, rb);
                    for (int bb = 0; bb < 3; bb++)
                        for (int a = 0; a < 3; a++) buf[(size_t)a * ncol + nslot + wt[bb]] += blk[a][bb];
                    mp[tier]++;
                    mq[tier] += r->n;
                    if (tiers) tiers[(size_t)e * (ne + nwt) + ne + s] = (unsigned char)tier;
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
            }
            if (erc) {
#pragma omp atomic write
                fail = 1;
                continue;
            }
#pragma omp critical(vpo_scatter)
            {
                for (int a = 0; a < na; a++) {
                    const double *src = buf + (size_t)a * ncol;
                    for (long long j = 0; j < ncol; j++) A[(size_t)j * nslot + slots[a]] += src[j];
                    bvec[slots[a]] += rows[a];
                }
            }
        }
#pragma omp critical(vpo_stats)
        for (int k = 0; k < 6; k++) { pairs[k] += mp[k]; points[k] += mq[k]; }
        rule_free(&tmp);
        free(buf);
    }
    rule_free(&g3);
    free(sig);
    free(qs);
    free(ts);
    free(wel);
    free(tp);
    tier_rules_free(&R);
    square_tiers_free(&ST);
    return fail ? -4 : 0;
}

/* quad_field at n points (xs, nxs: n x 3) for the cubature tests; with P (n x 2), polar_field about P[i] at distance
 * 0 (the self pair's principal values and one-sided limits). c: n x 4, s: n.  Returns 0, -1 or -2 as vpo_quad_field. */
int vpo_quad_field_many(const double *qx, int n, const double *xs, const double *nxs, const double *V, int side,
                        const double *P, double *c, double *s)
{
    quad_t Q;
    if (make_quad(ld3(qx), ld3(qx + 3), ld3(qx + 6), ld3(qx + 9), &Q)) return -2;
    gauss_table();
    int fail = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        int rc = P ? polar_field(&Q, ld3(xs + 3 * i), ld3(nxs + 3 * i), ld3(V), P + 2 * i, 0.0, side, c + 4 * i, s + i)
                   : quad_field(&Q, ld3(xs + 3 * i), ld3(nxs + 3 * i), ld3(V), side, 1.0, c + 4 * i, s + i);
        if (rc) {
#pragma omp atomic write
            fail = rc;
        }
    }
    return fail;
}
