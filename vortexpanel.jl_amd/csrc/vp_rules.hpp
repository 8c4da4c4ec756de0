// Quadrature tier rules, built on the host once per assembly and copied to device memory.
// Reference: src/elements/rules.jl:12-174 (gauss_legendre, graded_gauss, triangle_rule, rotate_rule,
// edge_graded_rule, vertex_graded_rule, self_rule, TierRules).
#pragma once
#include <cmath>
#include <stdexcept>
#include <vector>

namespace vp {

struct HostRule {
    std::vector<double> l;  // 3 per point
    std::vector<double> w;
    int n() const { return (int)w.size(); }
    void push(double a, double b, double c, double wt)
    {
        l.push_back(a);
        l.push_back(b);
        l.push_back(c);
        w.push_back(wt);
    }
};

// rules.jl:12-40 (nodes and weights on [0, 1], Newton on the Legendre recurrence)
inline void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 1; i <= (n + 1) / 2; i++) {
        double z = std::cos(M_PI * (i - 0.25) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = ((2.0 * j - 1) * z * p1 - (j - 1.0) * p2) / j;
                p2 = p1;
                p1 = p3;
            }
            dp = n * (z * p1 - p2) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p1 = 1.0, p2 = 0.0;
        for (int j = 1; j <= n; j++) {
            double p3 = ((2.0 * j - 1) * z * p1 - (j - 1.0) * p2) / j;
            p2 = p1;
            p1 = p3;
        }
        dp = n * (z * p1 - p2) / (z * z - 1);
        x[i - 1] = (1 - z) / 2;
        x[n - i] = (1 + z) / 2;
        w[i - 1] = w[n - i] = 1 / ((1 - z * z) * dp * dp);
    }
}

// rules.jl:48-63 (hp-graded toward 0; u^2 substitution on the innermost interval)
inline void graded_gauss(int levels, double ratio, int order, std::vector<double> &x, std::vector<double> &w)
{
    std::vector<double> br{0.0};
    for (int k = levels; k >= 1; k--) br.push_back(std::pow(ratio, k));
    br.push_back(1.0);
    x.clear();
    w.clear();
    std::vector<double> xg, wg;
    for (size_t k = 1; k < br.size(); k++) {
        gauss_legendre(order + (int)k - 1, xg, wg);
        double a = br[k - 1], b = br[k];
        for (size_t i = 0; i < xg.size(); i++) {
            if (k == 1) {
                x.push_back(b * (xg[i] * xg[i]));
                w.push_back(2 * b * xg[i] * wg[i]);
            } else {
                x.push_back(a + (b - a) * xg[i]);
                w.push_back((b - a) * wg[i]);
            }
        }
    }
}

// rules.jl:66-69
inline void graded_gauss_both(int levels, double ratio, int order, std::vector<double> &x, std::vector<double> &w)
{
    std::vector<double> xs, ws;
    graded_gauss(levels, ratio, order, xs, ws);
    size_t m = xs.size();
    x.resize(2 * m);
    w.resize(2 * m);
    for (size_t i = 0; i < m; i++) {
        x[i] = xs[i] / 2;
        w[i] = ws[i] / 2;
        x[m + i] = 1 - xs[m - 1 - i] / 2;
        w[m + i] = ws[m - 1 - i] / 2;
    }
}

// rules.jl:76-92
inline HostRule triangle_rule(int n)
{
    HostRule r;
    if (n == 1) {
        r.push(1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0);
    } else if (n == 3) {
        double a = 2.0 / 3, b = 1.0 / 6;
        r.push(a, b, b, 1.0 / 3);
        r.push(b, a, b, 1.0 / 3);
        r.push(b, b, a, 1.0 / 3);
    } else if (n == 6) {
        double a = 0.445948490915965, b = 0.091576213509771, wa = 0.223381589678011, wb = 0.109951743655322;
        r.push(1 - 2 * a, a, a, wa);
        r.push(a, 1 - 2 * a, a, wa);
        r.push(a, a, 1 - 2 * a, wa);
        r.push(1 - 2 * b, b, b, wb);
        r.push(b, 1 - 2 * b, b, wb);
        r.push(b, b, 1 - 2 * b, wb);
    } else if (n == 7) {
        double s = std::sqrt(15.0), a = (6 - s) / 21, b = (6 + s) / 21, wa = (155 - s) / 1200, wb = (155 + s) / 1200;
        r.push(1.0 / 3, 1.0 / 3, 1.0 / 3, 9.0 / 40);
        r.push(1 - 2 * a, a, a, wa);
        r.push(a, 1 - 2 * a, a, wa);
        r.push(a, a, 1 - 2 * a, wa);
        r.push(1 - 2 * b, b, b, wb);
        r.push(b, 1 - 2 * b, b, wb);
        r.push(b, b, 1 - 2 * b, wb);
    } else {
        throw std::invalid_argument("no such triangle rule (1, 3, 6 or 7 points)");
    }
    return r;
}

// rules.jl:97-102: base corner 1 becomes corner k (k = 1..3); lambda_T[j] = lambda_base[mod1(j + 1 - k, 3)]
inline HostRule rotate_rule(const HostRule &r, int k)
{
    HostRule o;
    int idx[3] = {((1 - k) % 3 + 3) % 3, ((2 - k) % 3 + 3) % 3, ((3 - k) % 3 + 3) % 3};
    for (int i = 0; i < r.n(); i++) o.push(r.l[3 * i + idx[0]], r.l[3 * i + idx[1]], r.l[3 * i + idx[2]], r.w[i]);
    return o;
}

// rules.jl:108-118
inline HostRule edge_graded_rule(int levels, double ratio, int order)
{
    std::vector<double> ts, wt, ss, ws;
    graded_gauss(levels, ratio, order, ts, wt);
    graded_gauss_both(levels / 2 > 1 ? levels / 2 : 1, ratio, order, ss, ws);
    HostRule r;
    for (size_t i = 0; i < ts.size(); i++)
        for (size_t j = 0; j < ss.size(); j++)
            r.push((1 - ts[i]) * (1 - ss[j]), (1 - ts[i]) * ss[j], ts[i], 2 * (1 - ts[i]) * wt[i] * ws[j]);
    return r;
}

// rules.jl:125-134
inline HostRule vertex_graded_rule(int levels, double ratio, int order)
{
    std::vector<double> rs, wr, ss, ws;
    graded_gauss(levels, ratio, order, rs, wr);
    gauss_legendre(order + levels, ss, ws);
    HostRule r;
    for (size_t i = 0; i < rs.size(); i++)
        for (size_t j = 0; j < ss.size(); j++)
            r.push(1 - rs[i], rs[i] * (1 - ss[j]), rs[i] * ss[j], 2 * rs[i] * wr[i] * ws[j]);
    return r;
}

// rules.jl:142-155
inline HostRule self_rule(int levels, double ratio, int order)
{
    HostRule base = edge_graded_rule(levels, ratio, order), r;
    const double c = 1.0 / 3;
    for (int k = 0; k < 3; k++) {
        double ek[3] = {0, 0, 0}, ek1[3] = {0, 0, 0};
        ek[k] = 1.0;
        ek1[(k + 1) % 3] = 1.0;
        for (int i = 0; i < base.n(); i++) {
            const double *m = &base.l[3 * i];
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * ek[j] + m[1] * ek1[j]) + m[2] * c;
            r.push(l[0], l[1], l[2], base.w[i] / 3);
        }
    }
    return r;
}

// Rule table ids on the device
enum RuleId { R_FAR = 0, R_NEAR = 1, R_SELF = 2, R_EDGE0 = 3, R_VERT0 = 6, R_COUNT = 9 };

}  // namespace vp
