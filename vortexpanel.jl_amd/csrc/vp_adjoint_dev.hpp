// Device helpers of the backward assembly (vp_adjoint.hpp has its kernels and the host side): dual numbers (for the
// filament), the reverse derivatives of the octupole multipole and of the closed form, the closed-form adjoints of a
// triangle's normal, area and sheet, and the per-pair sums.  Included by vp_assemble.hip ahead of the
// evaluation kernels, which take an adjoint mode.
#pragma once

namespace vp {

// ------------------------------------------------------------------------------------------ dual numbers -----------

struct D3 {  // a value and its derivatives along three directions
    double v, d[3];
    D3() = default;
    __device__ constexpr D3(double x) : v(x), d{0.0, 0.0, 0.0} {}
    __device__ constexpr D3(double x, double a, double b, double c) : v(x), d{a, b, c} {}
};

__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3(a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3(a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]); }
__device__ __forceinline__ D3 operator-(D3 a) { return D3(-a.v, -a.d[0], -a.d[1], -a.d[2]); }
__device__ __forceinline__ D3 operator*(D3 a, D3 b)
{
    return D3(a.v * b.v, a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]);
}
__device__ __forceinline__ D3 operator*(D3 a, double s) { return D3(a.v * s, a.d[0] * s, a.d[1] * s, a.d[2] * s); }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return a * s; }
__device__ __forceinline__ D3 operator/(D3 a, D3 b)
{
    const double q = a.v / b.v, ib = 1.0 / b.v;
    return D3(q, (a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib, (a.d[2] - q * b.d[2]) * ib);
}
__device__ __forceinline__ D3 &operator+=(D3 &a, D3 b) { return a = a + b; }
__device__ __forceinline__ D3 &operator-=(D3 &a, D3 b) { return a = a - b; }
__device__ __forceinline__ bool operator<(D3 a, D3 b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(D3 a, D3 b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(D3 a, D3 b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(D3 a, D3 b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(D3 a, D3 b) { return a.v == b.v; }
__device__ __forceinline__ D3 dsqrt(D3 a)
{
    const double s = __builtin_sqrt(a.v), h = 0.5 / s;
    return D3(s, a.d[0] * h, a.d[1] * h, a.d[2] * h);
}
__device__ __forceinline__ D3 dlog(D3 a)
{
    const double i = 1.0 / a.v;
    return D3(log(a.v), a.d[0] * i, a.d[1] * i, a.d[2] * i);
}
__device__ __forceinline__ D3 datan2(D3 y, D3 x)
{
    const double i = 1.0 / (x.v * x.v + y.v * y.v);
    return D3(atan2(y.v, x.v), (x.v * y.d[0] - y.v * x.d[0]) * i, (x.v * y.d[1] - y.v * x.d[1]) * i,
              (x.v * y.d[2] - y.v * x.d[2]) * i);
}

__device__ __forceinline__ V3<D3> dconst(V3<double> x) { return mk3(D3(x.x), D3(x.y), D3(x.z)); }
__device__ __forceinline__ V3<D3> dseed(V3<double> x, double on)
{
    return mk3(D3(x.x, on, 0.0, 0.0), D3(x.y, 0.0, on, 0.0), D3(x.z, 0.0, 0.0, on));
}

// ------------------------------------------------------------------------------------------ triangle adjoints ------

// g[k] += gradient of p.N (N = (X1 - X0) x (X2 - X0), twice the area times the normal) with respect to vertex k
__device__ __forceinline__ void cross_adj(const V3<double> *X, V3<double> p, V3<double> *g)
{
    g[0] = g[0] + cross(p, X[2] - X[1]);
    g[1] = g[1] + cross(X[2] - X[0], p);
    g[2] = g[2] + cross(p, X[1] - X[0]);
}

// g[k] += gradient of c.n (n the unit normal) with respect to vertex k
__device__ __forceinline__ void normal_adj(const V3<double> *X, V3<double> c, V3<double> *g)
{
    const V3<double> N = cross(X[1] - X[0], X[2] - X[0]);
    const double A2 = dsqrt(dot(N, N));
    const V3<double> n = (1.0 / A2) * N;
    cross_adj(X, (1.0 / A2) * (c - dot(c, n) * n), g);
}

// g[k] += s times the gradient of |N| (twice the area) with respect to vertex k
__device__ __forceinline__ void area2_adj(const V3<double> *X, double s, V3<double> *g)
{
    const V3<double> N = cross(X[1] - X[0], X[2] - X[0]);
    cross_adj(X, (s / dsqrt(dot(N, N))) * N, g);
}

// The sheet Gamma = sum_b mu_b gamma_b of a triangle (gamma_b = -e_b/(2A), src/elements/flat_triangle.jl:23-31)
__device__ __forceinline__ V3<double> sheet_of(const V3<double> *X, const double *mu)
{
    const V3<double> N = cross(X[1] - X[0], X[2] - X[0]);
    const V3<double> E = mu[0] * (X[2] - X[1]) + mu[1] * (X[0] - X[2]) + mu[2] * (X[1] - X[0]);
    return (-1.0 / dsqrt(dot(N, N))) * E;
}

// g[k] += gradient of d.Gamma with respect to vertex k
__device__ __forceinline__ void sheet_adj(const V3<double> *X, const double *mu, V3<double> d, V3<double> *g)
{
    const V3<double> N = cross(X[1] - X[0], X[2] - X[0]);
    const double A2 = dsqrt(dot(N, N));
    const V3<double> E = mu[0] * (X[2] - X[1]) + mu[1] * (X[0] - X[2]) + mu[2] * (X[1] - X[0]);
    g[0] = g[0] - ((mu[1] - mu[2]) / A2) * d;
    g[1] = g[1] - ((mu[2] - mu[0]) / A2) * d;
    g[2] = g[2] - ((mu[0] - mu[1]) / A2) * d;
    area2_adj(X, dot(d, E) / (A2 * A2), g);
}

template <typename T> __device__ __forceinline__ V3<T> cvt3(V3<double> v) { return mk3(T(v.x), T(v.y), T(v.z)); }
template <typename T> __device__ __forceinline__ V3<double> dbl3(V3<T> v)
{
    return mk3((double)v.x, (double)v.y, (double)v.z);
}

// The octupole multipole (source_velocity_far, source_triangle.jl:100-119) and the reverse derivative of
// phi = Q.w with respect to the source's vertices, by hand: phi depends on r = x - c, the vertex offsets d_k = v_k - c
// and the area A; with dr_k = d_k.r, dd_k = d_k.d_k and iRn = |r|^-n,
//   d phi/d v_j = d phi/d d_j - (1/3) sum_k d phi/d d_k - (1/3) d phi/d r + (phi/A) dA/d v_j.
// About 200 flops per point against ~3000 for dual numbers.  T: the compute type (float on the mixed path: the
// offsets r and d_k are formed in FP64 first, so FP32 carries only their rounding).
template <typename T = double>
__device__ __forceinline__ void far_point_grad(const V3<double> *XS, const Elem &S, V3<double> x, V3<double> Qd,
                                               V3<double> &wv, V3<double> *g)
{
    const V3<double> cd = ld3(S.c);
    const V3<T> n = cvt3<T>(ld3(S.n)), Q = cvt3<T>(Qd);
    const T A = T(S.area), s = A * T(1.0 / (4 * kPi));
    const V3<T> r = cvt3<T>(x - cd);
    const T R2 = dot(r, r), iR = vrsqrt(R2);
    const T iR2 = iR * iR, iR3 = iR2 * iR, iR5 = iR3 * iR2, iR7 = iR5 * iR2, iR9 = iR7 * iR2, iR11 = iR9 * iR2;
    const T qr = dot(Q, r);
    T cr = T(0), rr = T(-3) * qr * iR5;
    V3<T> w = mk3(T(0), T(0), T(0)), gr = w, gd[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const V3<T> d = cvt3<T>(XS[k] - cd);
        const T dr = dot(d, r), dd = dot(d, d), qd = dot(Q, d);
        cr += (T(15) * dr * dr * iR7 - T(3) * dd * iR5) * T(1.0 / 24) -
              (T(45) * dd * dr * iR7 - T(105) * dr * dr * dr * iR9) * T(1.0 / 180);
        const T cd_ = T(-6) * dr * iR5 * T(1.0 / 24) - (T(45) * dr * dr * iR7 - T(9) * dd * iR5) * T(1.0 / 180);
        w = w + cd_ * d;
        const T pdr = qr * (T(30) * dr * iR7 * T(1.0 / 24) - (T(45) * dd * iR7 - T(315) * dr * dr * iR9) * T(1.0 / 180)) +
                      qd * (T(-6) * iR5 * T(1.0 / 24) - T(90) * dr * iR7 * T(1.0 / 180));
        const T pdd = qr * (T(-3) * iR5 * T(1.0 / 24) - T(45) * dr * iR7 * T(1.0 / 180)) + qd * (T(9) * iR5 * T(1.0 / 180));
        gr = gr + pdr * d;
        rr += qr * (T(-105) * dr * dr * iR9 * T(1.0 / 24) + T(15) * dd * iR7 * T(1.0 / 24) +
                    T(315) * dd * dr * iR9 * T(1.0 / 180) - T(945) * dr * dr * dr * iR11 * T(1.0 / 180)) +
              qd * (T(30) * dr * iR7 * T(1.0 / 24) + T(315) * dr * dr * iR9 * T(1.0 / 180) - T(45) * dd * iR7 * T(1.0 / 180));
        gd[k] = s * (cd_ * Q + pdr * r + (T(2) * pdd) * d);
    }
    w = w + (iR3 + cr) * r;
    const V3<T> wt = s * w;
    const T phi = dot(Q, wt);
    const V3<T> g_r = s * ((iR3 + cr) * Q + gr + rr * r);
    const V3<T> base = T(-1.0 / 3) * ((gd[0] + gd[1]) + gd[2] + g_r);
    const T pa = T(0.5) * phi / A;
    wv = dbl3(wt);
    g[0] = dbl3(gd[0] + base + pa * cross(n, cvt3<T>(XS[2] - XS[1])));
    g[1] = dbl3(gd[1] + base + pa * cross(cvt3<T>(XS[2] - XS[0]), n));
    g[2] = dbl3(gd[2] + base + pa * cross(n, cvt3<T>(XS[1] - XS[0])));
}

// The closed form (source_velocity_exact, source_triangle.jl:75-89) and the reverse derivative of phi = Q.w with
// respect to the source's vertices (x fixed), by hand: w = (1/4pi)[sum_e nu_e L_e + Omega n] with
// L_e = log((a + b + d)/(a + b - d)) (dL/da = dL/db = -2d/s2d2, dL/dd = 2(a + b)/s2d2, s2d2 the cancellation-free
// (a + b)^2 - d^2), nu_e = e x n / d and Omega = 2 atan2(N, D); the normal's dependence is gathered into one vector
// and passed through its closed-form adjoint.  In the own element's plane (side != 0) Omega is the constant one-sided
// limit; off it the atan2 form is differentiated (also where the forward takes the in-plane zero of side = 0).
__device__ __forceinline__ void exact_point_grad(const V3<double> *XS, const Elem &S, V3<double> x, V3<double> Q,
                                                 int side, V3<double> &wv, V3<double> *g)
{
    const V3<double> n = ld3(S.n);
    const V3<double> r[3] = {x - XS[0], x - XS[1], x - XS[2]};
    const double l[3] = {psqrt(dot(r[0], r[0])), psqrt(dot(r[1], r[1])), psqrt(dot(r[2], r[2]))};
    const V3<double> nQ = cross(n, Q);
    V3<double> cn = mk3(0.0, 0.0, 0.0), wt = cn;
    g[0] = g[1] = g[2] = cn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = k, j = (k + 1) % 3;
        const V3<double> e = XS[j] - XS[i];
        const double d = S.len[k], a = l[i], b = l[j];
        const double p = dot(r[i], r[j]);
        const V3<double> c = cross(r[i], r[j]);
        const double s2d2 = p < 0.0 ? 2.0 * dot(c, c) / (a * b - p) : 2.0 * (a * b + p);
        const double t = a + b + d;
        const double Le = dlog(t * t / s2d2);
        const V3<double> nu = ld3(S.nu[k]);
        wt = wt + Le * nu;
        const double f = dot(Q, nu);
        const double dLa = -2.0 * d / s2d2, dLd = 2.0 * (a + b) / s2d2;
        const V3<double> ge = Le * ((1.0 / d) * nQ - (f / (d * d)) * e);
        g[i] = g[i] + (f * (-dLa / a)) * r[i] - (f * dLd / d) * e - ge;
        g[j] = g[j] + (f * (-dLa / b)) * r[j] + (f * dLd / d) * e + ge;
        cn = cn + (Le / d) * cross(Q, e);
    }
    const double qn = dot(Q, n);
    double Om;
    const double z = cls_dot(cls_sub(x, ld3(S.c)), n);
    if (side != 0 && fabs(z) <= 1e-12 * S.diam) {
        Om = side * 2 * kPi * inplane_fraction(S, x, 1e-12);
    } else {
        const V3<double> &ra = r[0], &rb = r[1], &rc = r[2];
        const double la = l[0], lb = l[1], lc = l[2];
        const double N = dot(ra, cross(rb, rc));
        const double D = la * lb * lc + dot(ra, rb) * lc + dot(ra, rc) * lb + dot(rb, rc) * la;
        Om = 2.0 * datan2(N, D);
        const double h = 2.0 / (N * N + D * D), dN = h * D * qn, dD = -h * N * qn;
        g[0] = g[0] - dN * cross(rb, rc) - dD * ((lb * lc + dot(rb, rc)) / la * ra + lc * rb + lb * rc);
        g[1] = g[1] - dN * cross(rc, ra) - dD * ((la * lc + dot(ra, rc)) / lb * rb + lc * ra + la * rc);
        g[2] = g[2] - dN * cross(ra, rb) - dD * ((la * lb + dot(ra, rb)) / lc * rc + lb * ra + la * rb);
    }
    cn = cn + Om * Q;
    normal_adj(XS, cn, g);
    const double s = 1.0 / (4 * kPi);
    wv = s * (wt + Om * n);
    g[0] = s * g[0];
    g[1] = s * g[1];
    g[2] = s * g[2];
}

// exact_point_grad in FP32 for points off the source (the mixed path's far and plain near tiers; no in-plane branch),
// with the forms of sv_exact_f32 that keep FP32's digits: each edge log as log1p(2 d t / s2d2), the solid angle's
// numerator as 2 A z, and the cross products of its derivative as rb x (b - c), rc x (c - a), ra x (a - b) (equal to
// rb x rc, rc x ra, ra x rb, without their cancellation for nearly parallel offsets).  The normal's adjoint is taken
// in FP64.  (exact_point_grad with side 0.)
__device__ __forceinline__ void exact_point_grad_f32(const V3<double> *XS, const Elem &S, V3<double> x, V3<double> Qd,
                                                     V3<double> &wv, V3<double> *g)
{
    const V3<float> n = cvt3<float>(ld3(S.n)), Q = cvt3<float>(Qd);
    V3<float> r[3], e[3], gf[3];
    float l[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r[k] = cvt3<float>(x - XS[k]);
        e[k] = cvt3<float>(XS[k == 2 ? 0 : k + 1] - XS[k]);  // edge k: vertex k -> k + 1
        l[k] = __builtin_amdgcn_sqrtf(dot(r[k], r[k]));
        gf[k] = mk3(0.f, 0.f, 0.f);
    }
    const V3<float> nQ = cross(n, Q);
    V3<float> cn = mk3(0.f, 0.f, 0.f), wt = cn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = k, j = k == 2 ? 0 : k + 1;
        const float d = (float)S.len[k], a = l[i], b = l[j], ab = a * b, p = dot(r[i], r[j]);
        const V3<float> c = cross(r[i], r[j]);
        const float s2d2 = p < 0.f ? 2.f * dot(c, c) * __builtin_amdgcn_rcpf(ab - p) : 2.f * (ab + p);
        const float is2 = __builtin_amdgcn_rcpf(s2d2), t = a + b + d, Le = flog1p(2.f * d * t * is2);
        const V3<float> nu = cvt3<float>(ld3(S.nu[k]));
        wt = wt + Le * nu;
        const float f = dot(Q, nu), id = __builtin_amdgcn_rcpf(d);
        const float dLa = -2.f * d * is2, dLd = 2.f * (a + b) * is2;
        const V3<float> ge = Le * (id * nQ - (f * id * id) * e[k]);
        gf[i] = gf[i] + (f * (-dLa * __builtin_amdgcn_rcpf(a))) * r[i] - (f * dLd * id) * e[k] - ge;
        gf[j] = gf[j] + (f * (-dLa * __builtin_amdgcn_rcpf(b))) * r[j] + (f * dLd * id) * e[k] + ge;
        cn = cn + (Le * id) * cross(Q, e[k]);
    }
    // the atan2 form differentiated everywhere, in the plane too (where the forward takes 0 and N = 0), as the FP64
    // adjoint does for points off the own element
    const float qn = dot(Q, n);
    float Om;
    const double z = dot(x - ld3(S.c), ld3(S.n));
    {
        const V3<float> &ra = r[0], &rb = r[1], &rc = r[2];
        const float la = l[0], lb = l[1], lc = l[2];
        const float N = (float)(2.0 * S.area * z);
        const float D = la * lb * lc + dot(ra, rb) * lc + dot(ra, rc) * lb + dot(rb, rc) * la;
        Om = 2.f * fatan2(N, D);
        const float h = 2.f * __builtin_amdgcn_rcpf(N * N + D * D), dN = h * D * qn, dD = -h * N * qn;
        // rb x rc = rb x (b - c) = -rb x e[1]; rc x ra = -rc x e[2]; ra x rb = -ra x e[0]
        gf[0] = gf[0] + dN * cross(rb, e[1]) - dD * (((lb * lc + dot(rb, rc)) * __builtin_amdgcn_rcpf(la)) * ra + lc * rb + lb * rc);
        gf[1] = gf[1] + dN * cross(rc, e[2]) - dD * (((la * lc + dot(ra, rc)) * __builtin_amdgcn_rcpf(lb)) * rb + lc * ra + la * rc);
        gf[2] = gf[2] + dN * cross(ra, e[0]) - dD * (((la * lb + dot(ra, rb)) * __builtin_amdgcn_rcpf(lc)) * rc + lb * ra + la * rb);
    }
    cn = cn + Om * Q;
    const double s4 = 1.0 / (4 * kPi);
    for (int k = 0; k < 3; k++) g[k] = dbl3(gf[k]);
    normal_adj(XS, dbl3(cn), g);
    wv = s4 * dbl3(wt + Om * n);
    g[0] = s4 * g[0];
    g[1] = s4 * g[1];
    g[2] = s4 * g[2];
}

// phi = Q.w(x; S) (w = source_velocity, source_triangle.jl:134-138, with the forward's switch) and d phi/d(vertex k of
// S) for k = 0, 1, 2; the value w(x) in wv.  Both branches by their reverse derivatives; T float: the mixed path's
// FP32 forms (points off the source only: side 0).
template <typename T = double>
__device__ __forceinline__ void tri_point_grad(const V3<double> *XS, const Elem &S, V3<double> x, V3<double> Q, int side,
                                               double kernel_far, V3<double> &wv, V3<double> *g)
{
    const V3<double> r0 = cls_sub(x, ld3(S.c));
    const double fd = kernel_far * S.diam;
    if (cls_dot(r0, r0) > fd * fd)
        far_point_grad<T>(XS, S, x, Q, wv, g);
    else if constexpr (std::is_same<T, float>::value)
        exact_point_grad_f32(XS, S, x, Q, wv, g);
    else
        exact_point_grad(XS, S, x, Q, side, wv, g);
}

// phi = n.(mu_a va + mu_b vb) of the filament from fa to fb (segment_velocity, src/kernels/segment.jl:15-25) and its
// derivatives with respect to fa and fb; the value mu_a va + mu_b vb in wv.
__device__ void seg_point_grad(V3<double> fa, V3<double> fb, V3<double> x, V3<double> n, double mua, double mub,
                               V3<double> &wv, V3<double> *g)
{
#pragma unroll 1
    for (int k = 0; k < 2; k++) {
        const V3<D3> a = dseed(fa, k == 0 ? 1.0 : 0.0), b = dseed(fb, k == 1 ? 1.0 : 0.0), xd = dconst(x);
        V3<D3> va, vb;
        seg_vel(xd - a, xd - b, va, vb);
        const V3<D3> w = D3(mua) * va + D3(mub) * vb;
        const D3 phi = w.x * n.x + w.y * n.y + w.z * n.z;
        g[k] = mk3(phi.d[0], phi.d[1], phi.d[2]);
        if (k == 0) wv = mk3(w.x.v, w.y.v, w.z.v);
    }
}

// A pair's accumulated sums -> the gradient of P_ts with respect to T's and S's vertices (see the file head).
struct PairSums {
    V3<double> Sv;     // sum_q a_q w_q
    V3<double> GS[3];  // sum_q a_q d(Q.w)/d(vertex k of S)
    V3<double> GX[3];  // sum_q a_q l_qk d(Q.w)/dx
};

__device__ void finish_tri_pair(const V3<double> *XT, V3<double> nT, double AT, const V3<double> *XS, const double *muS,
                                double sigS, V3<double> Gam, bool body, V3<double> Vinf, V3<double> Q,
                                const PairSums &p, V3<double> *gT, V3<double> *gS)
{
    for (int k = 0; k < 3; k++) {
        gT[k] = (-AT) * p.GX[k];
        gS[k] = (-AT) * p.GS[k];
    }
    area2_adj(XT, -0.5 * dot(Q, p.Sv), gT);
    normal_adj(XT, (-AT) * (sigS * p.Sv + cross(Gam, p.Sv)), gT);
    if (body) normal_adj(XS, (AT * dot(nT, p.Sv)) * Vinf, gS);  // sigma_S = -V∞.n_S
    sheet_adj(XS, muS, (-AT) * cross(p.Sv, nT), gS);
}

__device__ __forceinline__ void gadd(double *G, int node, V3<double> v)
{
    unsafeAtomicAdd(&G[3 * node], v.x);
    unsafeAtomicAdd(&G[3 * node + 1], v.y);
    unsafeAtomicAdd(&G[3 * node + 2], v.z);
}

__device__ __forceinline__ int col_node(int col, int nslot, int nn, const int *slot_node)
{
    return col < nslot ? slot_node[col] : nn + (col - nslot);
}

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// x of the lane that the DPP control CTRL names within the lane's row of 16
template <int CTRL> __device__ __forceinline__ double dpp_get(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double lane_value(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The sum of v over the wave, in every lane, without LDS: DPP adds within each row of 16 lanes (lane pairs by
// quad_perm [1,0,3,2], quads by [2,3,0,1], half rows by row_half_mirror, rows by row_mirror), then the four row sums
// read from lanes 0, 16, 32 and 48.  Every lane of the wave must be active.
__device__ __forceinline__ double wave_total(double v)
{
    v += dpp_get<0xB1>(v);
    v += dpp_get<0x4E>(v);
    v += dpp_get<0x141>(v);
    v += dpp_get<0x140>(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// What the evaluation kernels need in adjoint mode (null lam in the forward assembly).  The kernels take up to
// ADJ_MAX sets of adjoint weights at once (a template count NL): the point derivatives of the kernel do not depend on
// lambda, which only weights their sums, so several objectives share one pass.
constexpr int ADJ_MAX = 2;
struct AdjArgs {
    const double *lam;               // adjoint weights per slot, NL sets lstride apart
    const double *mu;                // doublet values per column (slots, then wake nodes)
    double *G;                       // gradient per node (body nodes, then wake nodes), 3 each; NL sets gstride apart
    const int *nodes;                // gradient index of every element corner (body nodes, wake nodes after them)
    double vx, vy, vz;               // V∞ (sigma = -V∞.n)
    double dx = 0, dy = 0, dz = 0;   // a direction dV∞ for ...
    double *bdot = nullptr;          // ... lambdaᵀ b(sigma = -dV∞.n, dV∞) per set, summed here when not null
    int lstride = 0;
    long long gstride = 0;
};

// A body pair's part of lambdaᵀ b for the source strength sigma' = -d.n_S: -sigma' A_T (Sv . n_T), since
// Sv = sum_q a_q w_q with A_T a_q = sum_a lambda_a W_q psi_a (pair_block's b entries, assemble.jl:253-254)
__device__ __forceinline__ double pair_bdot(const AdjArgs &aa, bool body, V3<double> nS, double AT, V3<double> nT,
                                            V3<double> Sv)
{
    return body ? dot(mk3(aa.dx, aa.dy, aa.dz), nS) * AT * dot(Sv, nT) : 0.0;
}

// Adds every thread's v to *dst (one atomic per workgroup); every thread of the workgroup calls it
__device__ __forceinline__ void block_add(double v, double *dst)
{
    __shared__ double part[4];
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0 && dst) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        if (t != 0.0) atomicAdd(dst, t);
    }
    __syncthreads();
}

template <int NL> __device__ __forceinline__ void block_add_n(const double *v, double *dst)
{
    for (int l = 0; l < NL; l++) block_add(v[l], dst ? dst + l : nullptr);
}

__device__ __forceinline__ void zero_sums(PairSums &p)
{
    p.Sv = mk3(0.0, 0.0, 0.0);
    for (int k = 0; k < 3; k++) p.GS[k] = p.GX[k] = p.Sv;
}

// One point's contribution, weight a, to a pair's sums: value wv, vertex derivatives g, l the point's barycentrics
__device__ __forceinline__ void add_point(PairSums &p, double a, V3<double> wv, const V3<double> *g, V3<double> gx,
                                          const double *l)
{
    p.Sv = p.Sv + a * wv;
    for (int k = 0; k < 3; k++) {
        p.GS[k] = p.GS[k] + a * g[k];
        p.GX[k] = p.GX[k] + (a * l[k]) * gx;
    }
}

// A (test triangle, source triangle) pair in adjoint mode: the contraction Q, the weights and the running sums
template <int NL> struct AdjPair {
    V3<double> XS[3], Gam, Q;
    double la[NL][3], muS[3], sigS;
    PairSums p[NL];
};

template <int NL>
__device__ __forceinline__ void adj_pair_begin(const AdjArgs &aa, int nt, const Elem &S, int s, const int *ct,
                                               const int *cs, V3<double> nT, AdjPair<NL> &ap)
{
    for (int k = 0; k < 3; k++) {
        ap.XS[k] = ld3(S.v[k]);
        ap.muS[k] = aa.mu[cs[k]];
        for (int l = 0; l < NL; l++) ap.la[l][k] = aa.lam[l * aa.lstride + ct[k]];
    }
    ap.Gam = sheet_of(ap.XS, ap.muS);
    ap.sigS = s < nt ? -dot(mk3(aa.vx, aa.vy, aa.vz), ld3(S.n)) : 0.0;
    ap.Q = ap.sigS * nT + cross(nT, ap.Gam);
    for (int l = 0; l < NL; l++) zero_sums(ap.p[l]);
}

// One rule point (barycentric l on the test triangle, weight W0 = rule weight x piece area fraction)
template <int NL, typename T = double>
__device__ __forceinline__ void adj_point(AdjPair<NL> &ap, const V3<double> *XT, const Elem &S, int side,
                                          double kernel_far, double W0, const double *l, int test)
{
    double psi[3], a[NL];
    test_weights(test, l, psi);
    bool any = false;
    for (int j = 0; j < NL; j++) {
        a[j] = W0 * ((ap.la[j][0] * psi[0] + ap.la[j][1] * psi[1]) + ap.la[j][2] * psi[2]);
        any = any || a[j] != 0.0;
    }
    if (!any) return;
    V3<double> wv, g[3];
    tri_point_grad<T>(ap.XS, S, cls_point(XT, l), ap.Q, side, kernel_far, wv, g);
    const V3<double> gx = -1.0 * ((g[0] + g[1]) + g[2]);
    for (int j = 0; j < NL; j++) add_point(ap.p[j], a[j], wv, g, gx, l);
}

__device__ __forceinline__ void adj_wave_reduce(PairSums &p)
{
    p.Sv = mk3(wsum(p.Sv.x), wsum(p.Sv.y), wsum(p.Sv.z));
    for (int k = 0; k < 3; k++) {
        p.GS[k] = mk3(wsum(p.GS[k].x), wsum(p.GS[k].y), wsum(p.GS[k].z));
        p.GX[k] = mk3(wsum(p.GX[k].x), wsum(p.GX[k].y), wsum(p.GX[k].z));
    }
}

// Adds the pair's parts of lambdaᵀ b along dV∞ (pair_bdot) to bd
template <int NL>
__device__ __forceinline__ void adj_pair_end(const AdjArgs &aa, int nt, const Elem &T, const V3<double> *XT, int t,
                                             int s, const int *nodes, const AdjPair<NL> &ap, const Elem &S, double *bd)
{
    for (int l = 0; l < NL; l++) {
        V3<double> gT[3], gS[3];
        finish_tri_pair(XT, ld3(T.n), T.area, ap.XS, ap.muS, ap.sigS, ap.Gam, s < nt, mk3(aa.vx, aa.vy, aa.vz), ap.Q,
                        ap.p[l], gT, gS);
        double *G = aa.G + l * aa.gstride;
        for (int k = 0; k < 3; k++) {
            gadd(G, nodes[3 * t + k], gT[k]);
            gadd(G, nodes[3 * s + k], gS[k]);
        }
        bd[l] += pair_bdot(aa, s < nt, ld3(S.n), T.area, ld3(T.n), ap.p[l].Sv);
    }
}

}  // namespace vp
