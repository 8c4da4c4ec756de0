// Device-side kernels of the Galerkin hot path for gfx950: the unit source triangle, the linear-strength filament,
// the Rosenhead-Moore blob, and the geometric predicates of the quadrature tiers.
//
// Reference: src/kernels/source_triangle.jl:8-138, src/kernels/segment.jl:15-25, src/kernels/blob.jl:21-25,
// src/elements/flat_triangle.jl:59-78 (distance_to), src/assembly/assemble.jl:145-184.
//
// Precision: everything is templated on the compute type T (double for the FP64 path, float for the far field of
// the mixed path).  Functions named cls_* decide quadrature tiers and rule pieces; they are compiled with FP
// contraction off and in the reference's operation order, so that tiers and rule sizes match the FP64 oracle bit for
// bit (hipcc contracts a*b+c into FMA by default, Julia never does).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vp {

constexpr double kPi = 3.14159265358979323846;

template <typename T> struct V3 {
    T x, y, z;
};

template <typename T> __device__ __forceinline__ V3<T> mk3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator*(V3<T> a, T s) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Orders LDS accesses between the lanes of one wavefront: what any lane wrote before the call is visible to every
// lane after it, and no access is moved across it by the compiler (a release fence at wavefront scope, the
// scheduling barrier, an acquire fence).  Kernels that hand values from some lanes to others through LDS, with no
// workgroup barrier in between, call it after the writes and again before the same locations are written anew.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP64 natural logarithm: x = m 2^e with m scaled into [sqrt(1/2), sqrt(2)), then log m = 2 atanh(s),
// s = (m - 1)/(m + 1), |s| <= 3 - 2 sqrt(2) < 0.1716, by eleven odd terms of the series (the next is below 2^-53
// of the sum), and e ln 2 with ln 2 split into a 32-bit head (exact times any exponent) and its tail.  About 40
// instructions, a third of the library routine's; within a few ulp of the C library's log (tested on the device).
__device__ __forceinline__ double vlog(double x)
{
    constexpr double kLn2Hi = 0.6931471803691238, kLn2Lo = 1.9082149292705877e-10;
    int e;
    double m = __builtin_frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.7071067811865476) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 21;
    p = fma(p, z, 1.0 / 19);
    p = fma(p, z, 1.0 / 17);
    p = fma(p, z, 1.0 / 15);
    p = fma(p, z, 1.0 / 13);
    p = fma(p, z, 1.0 / 11);
    p = fma(p, z, 1.0 / 9);
    p = fma(p, z, 1.0 / 7);
    p = fma(p, z, 1.0 / 5);
    p = fma(p, z, 1.0 / 3);
    const double lm = 2.0 * s + (2.0 * s * z) * p;
    const double ed = (double)e;
    const double r = fma(ed, kLn2Hi, fma(ed, kLn2Lo, lm));
    const double inf = __builtin_inf();
    return (x > 0.0 && x < inf) ? r : (x == 0.0 ? -inf : (x == inf ? inf : __builtin_nan("")));
}

// log(num/den) for num, den > 0 with a single division: with num = mn 2^en and den = md 2^ed (frexp), the ratio
// m = mn/md is brought into [sqrt(1/2), sqrt(2)) by doubling mn or md, and s = (m - 1)/(m + 1) = (mn - md)/(mn + md)
// feeds the series of vlog.  A zero den or infinite num gives +inf, a zero num or infinite den -inf, as log of the
// quotient would.
__device__ __forceinline__ double vlog_ratio(double num, double den)
{
    constexpr double kLn2Hi = 0.6931471803691238, kLn2Lo = 1.9082149292705877e-10;
    int en, ed;
    double mn = __builtin_frexp(num, &en), md = __builtin_frexp(den, &ed);
    int e = en - ed;
    if (mn < 0.7071067811865476 * md) {
        mn *= 2.0;
        e -= 1;
    } else if (mn >= 1.4142135623730951 * md) {
        md *= 2.0;
        e += 1;
    }
    const double s = (mn - md) / (mn + md);
    const double z = s * s;
    double p = 1.0 / 21;
    p = fma(p, z, 1.0 / 19);
    p = fma(p, z, 1.0 / 17);
    p = fma(p, z, 1.0 / 15);
    p = fma(p, z, 1.0 / 13);
    p = fma(p, z, 1.0 / 11);
    p = fma(p, z, 1.0 / 9);
    p = fma(p, z, 1.0 / 7);
    p = fma(p, z, 1.0 / 5);
    p = fma(p, z, 1.0 / 3);
    const double ed2 = (double)e;
    const double r = fma(ed2, kLn2Hi, fma(ed2, kLn2Lo, 2.0 * s + (2.0 * s * z) * p));
    const double inf = __builtin_inf();
    if (num > 0.0 && num < inf && den > 0.0 && den < inf) return r;
    return (den == 0.0 || num == inf) ? inf : (num == 0.0 || den == inf) ? -inf : __builtin_nan("");
}

// FP64 two-argument arctangent: the octant by the larger of |x| and |y| (t = min/max in [0, 1]), above
// tan(pi/12) the identity atan t = pi/6 + atan((sqrt3 t - 1)/(t + sqrt3)), so one quotient u with
// |u| <= tan(pi/12), u^2 < 0.0718, and fourteen terms of the arctangent series; quadrants and signed zeros as the C
// library's atan2.  About 50 instructions, half the library routine's (tested on the device).
__device__ __forceinline__ double vatan2(double y, double x)
{
    constexpr double kSqrt3 = 1.7320508075688772, kTan15 = 0.2679491924311227;
    const double ax = fabs(x), ay = fabs(y);
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    const bool big = num > kTan15 * den;
    const double n1 = big ? fma(num, kSqrt3, -den) : num;
    const double d1 = big ? fma(den, kSqrt3, num) : den;
    const double u = den == 0.0 ? 0.0 : n1 / d1;
    const double z = u * u;
    double p = -1.0 / 27;
    p = fma(p, z, 1.0 / 25);
    p = fma(p, z, -1.0 / 23);
    p = fma(p, z, 1.0 / 21);
    p = fma(p, z, -1.0 / 19);
    p = fma(p, z, 1.0 / 17);
    p = fma(p, z, -1.0 / 15);
    p = fma(p, z, 1.0 / 13);
    p = fma(p, z, -1.0 / 11);
    p = fma(p, z, 1.0 / 9);
    p = fma(p, z, -1.0 / 7);
    p = fma(p, z, 1.0 / 5);
    p = fma(p, z, -1.0 / 3);
    // the octant's offset and sign in one step: result = off + sgn atan(u), off among 0, pi/6, pi/3, pi/2, 2pi/3,
    // 5pi/6, pi as a double and its tail
    const double at = fma(u * z, p, u);
    const bool xneg = __builtin_signbit(x);
    double hi, lo, sgn;
    if (!swap) {
        sgn = xneg ? -1.0 : 1.0;
        hi = xneg ? (big ? 2.6179938779914944 : 3.141592653589793) : (big ? 0.5235987755982989 : 0.0);
        lo = xneg ? (big ? -4.597583668774144e-17 : 1.2246467991473532e-16) : (big ? -5.360408832255455e-17 : 0.0);
    } else {
        sgn = xneg ? 1.0 : -1.0;
        hi = xneg ? (big ? 2.0943951023931957 : 1.5707963267948966) : (big ? 1.0471975511965979 : 1.5707963267948966);
        lo = xneg ? (big ? -2.144163532902182e-16 : 6.123233995736766e-17) : (big ? -1.072081766451091e-16 : 6.123233995736766e-17);
    }
    const double a = hi + fma(sgn, at, lo);
    return __builtin_signbit(y) ? -a : a;
}

// 1/sqrt(x) for normal x > 0 (the multipole's distances): the hardware estimate refined by two Newton steps,
// y <- y + (y/2)(1 - x y^2), each squaring the estimate's relative error; about 10 instructions against ~30 for a
// correctly rounded sqrt and a division, within 2 ulp of 1/sqrt (tested on the device)
__device__ __forceinline__ double vrsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double r = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, r, y);
    r = fma(-x * y, y, 1.0);
    return fma(0.5 * y, r, y);
}
// FP32 (the mixed path's multipole and blobs): the hardware estimate itself, about 1 ulp (the FP32 kernels' stated
// errors, 2e-5 and 5e-5 of the largest velocity, are far above it)
__device__ __forceinline__ float vrsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
// The square root of a squared distance (x zero or normal positive) as x / sqrt(x) through vrsqrt: within 3 ulp of
// the correctly rounded root (tested on the device), about half the instructions of the rounded one; the closed
// form's vertex distances
__device__ __forceinline__ double psqrt(double x) { return x > 0.0 ? x * vrsqrt(x) : 0.0; }
__device__ __forceinline__ float psqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dlog(double x) { return vlog(x); }
__device__ __forceinline__ float dlog(float x) { return __logf(x); }
__device__ __forceinline__ double datan2(double y, double x) { return vatan2(y, x); }
__device__ __forceinline__ float datan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double drcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float drcp(float x) { return __frcp_rn(x); }

// Element data on the device: one record per flat triangle (body, then wake), 40 doubles (320 bytes).
// v: vertices, n: unit normal, c: centroid, nu: in-plane outward edge normals, len: edge lengths, g: gamma_b
// (the sheet of a unit doublet at corner b, -e_b/(2A)), reach: largest centroid-to-corner distance.
struct Elem {
    double v[3][3];
    double n[3];
    double c[3];
    double area, diam, reach, pad0;
    double nu[3][3];
    double len[3];
    double g[3][3];
};
static_assert(sizeof(Elem) == 40 * 8, "Elem layout");

// Source triangle in registers, in the compute type, with coordinates relative to `origin` (the element's centroid
// for the far field, so FP32 keeps its digits where the cancellation happens).
template <typename T> struct SrcT {
    V3<T> v[3], n, nu[3];
    T len[3], diam, area;
};

// edge_log, src/kernels/source_triangle.jl:39-44, with the distances a = |ri|, b = |rj| given:
// log(t^2 / s2d2), t = a + b + d, s2d2 = 2 |ri x rj|^2 / (a b - p) when p = ri.rj < 0 else 2 (a b + p)
template <typename T> __device__ __forceinline__ T edge_log(V3<T> ri, V3<T> rj, T a, T b, T d)
{
    T p = dot(ri, rj);
    V3<T> c = cross(ri, rj);
    T s2d2 = p < T(0) ? T(2) * dot(c, c) / (a * b - p) : T(2) * (a * b + p);
    T t = a + b + d;
    return dlog(t * t / s2d2);
}

// The FP64 edge_log with its two quotients folded into vlog_ratio's one division
template <> __device__ __forceinline__ double edge_log(V3<double> ri, V3<double> rj, double a, double b, double d)
{
    const double p = dot(ri, rj);
    const V3<double> c = cross(ri, rj);
    const double t = a + b + d, ab = a * b;
    const bool neg = p < 0.0;
    return vlog_ratio(neg ? t * t * (ab - p) : t * t, neg ? 2.0 * dot(c, c) : 2.0 * (ab + p));
}

// inplane_fraction, src/kernels/source_triangle.jl:52-64 (FP64 only: the self pair's in-plane points)
__device__ inline double inplane_fraction(const Elem &e, V3<double> x, double tol)
{
#pragma clang fp contract(off)
    V3<double> a = mk3(e.v[0][0], e.v[0][1], e.v[0][2]), b = mk3(e.v[1][0], e.v[1][1], e.v[1][2]),
               c = mk3(e.v[2][0], e.v[2][1], e.v[2][2]), n = mk3(e.n[0], e.n[1], e.n[2]);
    double A2 = 2 * e.area;
    double l[3] = {dot(cross(b - x, c - x), n) / A2, dot(cross(c - x, a - x), n) / A2, dot(cross(a - x, b - x), n) / A2};
    for (int k = 0; k < 3; k++)
        if (l[k] < -tol) return 0.0;
    int nz = 0;
    for (int k = 0; k < 3; k++) nz += fabs(l[k]) <= tol;
    if (nz == 0) return 1.0;
    if (nz == 1) return 0.5;
    int k = 0;
    while (k < 2 && !(fabs(l[k]) > tol)) k++;
    V3<double> vk = mk3(e.v[k][0], e.v[k][1], e.v[k][2]);
    V3<double> e1 = mk3(e.v[(k + 1) % 3][0], e.v[(k + 1) % 3][1], e.v[(k + 1) % 3][2]) - vk;
    V3<double> e2 = mk3(e.v[(k + 2) % 3][0], e.v[(k + 2) % 3][1], e.v[(k + 2) % 3][2]) - vk;
    double ca = dot(e1, e2) / (dsqrt(dot(e1, e1)) * dsqrt(dot(e2, e2)));
    ca = ca < -1 ? -1 : ca > 1 ? 1 : ca;
    return acos(ca) / (2 * kPi);
}

// source_velocity_exact, src/kernels/source_triangle.jl:75-89, from the offsets ra, rb, rc of the point from the
// vertices and its height z = (x - centroid).n over the plane.  The in-plane branch (side) needs the element in
// FP64 and is only reached by the touching tiers (ein non-null).
template <typename T>
__device__ __forceinline__ V3<T> sv_exact_abs(const SrcT<T> &s, V3<T> ra, V3<T> rb, V3<T> rc, T z, int side,
                                              const Elem *ein, V3<double> xabs)
{
    // the three vertex distances serve the three edge logs and the solid angle
    const T la = psqrt(dot(ra, ra)), lb = psqrt(dot(rb, rb)), lc = psqrt(dot(rc, rc));
    V3<T> wt = edge_log(ra, rb, la, lb, s.len[0]) * s.nu[0] + edge_log(rb, rc, lb, lc, s.len[1]) * s.nu[1] +
               edge_log(rc, ra, lc, la, s.len[2]) * s.nu[2];
    T Om;
    if (fabs(z) <= T(1e-12) * s.diam) {
        Om = ein ? T(side * 2 * kPi * inplane_fraction(*ein, xabs, 1e-12)) : T(0);
    } else {
        T N = dot(ra, cross(rb, rc));
        T D = la * lb * lc + dot(ra, rb) * lc + dot(ra, rc) * lb + dot(rb, rc) * la;
        Om = T(2) * datan2(N, D);
    }
    return T(1.0 / (4 * kPi)) * (wt + Om * s.n);
}

// source_velocity_far, src/kernels/source_triangle.jl:100-119 (vertex offsets from the centroid are s.v)
template <typename T> __device__ __forceinline__ V3<T> sv_far_rel(const SrcT<T> &s, V3<T> r)
{
    T R2 = dot(r, r);
    T iR = vrsqrt(R2);
    T iR2 = iR * iR, iR3 = iR2 * iR, iR5 = iR3 * iR2, iR7 = iR5 * iR2, iR9 = iR7 * iR2;
    // sum over the vertices of quad/24 - oct/180, collected as coefficients of r and of each d
    T cr = T(0);
    V3<T> w = mk3(T(0), T(0), T(0));
    for (int k = 0; k < 3; k++) {
        V3<T> d = s.v[k];
        T dr = dot(d, r), dd = dot(d, d);
        // r terms: (15 dr^2 iR7 - 3 dd iR5)/24 - (-105 dr^3 iR9 + 45 dd dr iR7)/180
        cr += (T(15) * dr * dr * iR7 - T(3) * dd * iR5) * T(1.0 / 24) - (T(45) * dd * dr * iR7 - T(105) * dr * dr * dr * iR9) * T(1.0 / 180);
        // d terms: (-6 dr iR5)/24 - (45 dr^2 iR7 - 9 dd iR5)/180
        T cd = T(-6) * dr * iR5 * T(1.0 / 24) - (T(45) * dr * dr * iR7 - T(9) * dd * iR5) * T(1.0 / 180);
        w = w + cd * d;
    }
    w = w + (iR3 + cr) * r;
    return (s.area * T(1.0 / (4 * kPi))) * w;
}

// The octupole multipole of source_velocity_far (src/kernels/source_triangle.jl:100-119) from the triangle's vertex
// moments about its centroid, made once per source: Q = sum_k d_k d_k^T (6), the third moment sum_k d_k d_k d_k (10),
// P = sum_k |d_k|^2 d_k (3) and D = sum_k |d_k|^2.  The reference's sums over the three vertices then regroup as
//   w = A/(4pi) [ (R^-3 + c) r - (1/4) R^-5 Q r - (1/4) R^-7 u + (1/20) R^-5 P ],   u_j = sum_ab T_jab r_a r_b,
//   c = (15/24) R^-7 r.Q r - (1/8) R^-5 D - (1/4) R^-7 P.r + (7/12) R^-9 u.r,
// the same terms summed in another order (within rounding of the per-vertex form).
template <typename T> struct FarMom {
    T q[6];   // xx xy xz yy yz zz
    T t[10];  // xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    T p[3], dd, area;
};

template <typename T> __device__ __forceinline__ FarMom<T> far_moments(const V3<T> *d, T area)
{
    FarMom<T> m;
    for (int i = 0; i < 6; i++) m.q[i] = T(0);
    for (int i = 0; i < 10; i++) m.t[i] = T(0);
    m.p[0] = m.p[1] = m.p[2] = m.dd = T(0);
    for (int k = 0; k < 3; k++) {
        const T x = d[k].x, y = d[k].y, z = d[k].z, s2 = dot(d[k], d[k]);
        m.q[0] += x * x, m.q[1] += x * y, m.q[2] += x * z, m.q[3] += y * y, m.q[4] += y * z, m.q[5] += z * z;
        m.t[0] += x * x * x, m.t[1] += x * x * y, m.t[2] += x * x * z, m.t[3] += x * y * y, m.t[4] += x * y * z;
        m.t[5] += x * z * z, m.t[6] += y * y * y, m.t[7] += y * y * z, m.t[8] += y * z * z, m.t[9] += z * z * z;
        m.p[0] += s2 * x, m.p[1] += s2 * y, m.p[2] += s2 * z, m.dd += s2;
    }
    m.area = area;
    return m;
}

template <typename T> __device__ __forceinline__ V3<T> sv_far_mom(const FarMom<T> &m, V3<T> r)
{
    const T iR = vrsqrt(dot(r, r)), iR2 = iR * iR, iR3 = iR2 * iR, iR5 = iR3 * iR2, iR7 = iR5 * iR2, iR9 = iR7 * iR2;
    const T x = r.x, y = r.y, z = r.z, xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    const V3<T> Qr = mk3(m.q[0] * x + m.q[1] * y + m.q[2] * z, m.q[1] * x + m.q[3] * y + m.q[4] * z,
                         m.q[2] * x + m.q[4] * y + m.q[5] * z);
    const V3<T> u = mk3(m.t[0] * xx + m.t[3] * yy + m.t[5] * zz + T(2) * (m.t[1] * xy + m.t[2] * xz + m.t[4] * yz),
                        m.t[1] * xx + m.t[6] * yy + m.t[8] * zz + T(2) * (m.t[3] * xy + m.t[4] * xz + m.t[7] * yz),
                        m.t[2] * xx + m.t[7] * yy + m.t[9] * zz + T(2) * (m.t[4] * xy + m.t[5] * xz + m.t[8] * yz));
    const T Pr = m.p[0] * x + m.p[1] * y + m.p[2] * z;
    const T c = T(15.0 / 24) * dot(r, Qr) * iR7 - T(1.0 / 8) * m.dd * iR5 - T(1.0 / 4) * Pr * iR7 +
                T(7.0 / 12) * dot(u, r) * iR9;
    const V3<T> w = (iR3 + c) * r + (T(-1.0 / 4) * iR5) * Qr + (T(-1.0 / 4) * iR7) * u +
                    (T(1.0 / 20) * iR5) * mk3(m.p[0], m.p[1], m.p[2]);
    return (m.area * T(1.0 / (4 * kPi))) * w;
}

// ---- FP32 closed form (the mixed path's far-tier closed-form points and plain near pairs) ----

// atan2 in FP32: atan t on [0, 1] as t P(t^2), P of degree 8 fitted here to atan(t)/t in relative error on 4000
// Chebyshev nodes of [0, 1], by Lawson's iteratively reweighted least squares towards the minimax fit (1.5e-8 of atan
// in exact arithmetic, 1.6e-7 evaluated in FP32), with the octant
// reduction atan2 = pi/2 - atan(1/t) above the diagonal and pi - atan for x < 0.
__device__ __forceinline__ float fatan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y), mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f, z = t * t;
    float p = 2.849886427e-03f;
    p = fmaf(p, z, -1.606861502e-02f);
    p = fmaf(p, z, 4.269149899e-02f);
    p = fmaf(p, z, -7.504292578e-02f);
    p = fmaf(p, z, 1.064093336e-01f);
    p = fmaf(p, z, -1.420364380e-01f);
    p = fmaf(p, z, 1.999261975e-01f);
    p = fmaf(p, z, -3.333307207e-01f);
    p = fmaf(p, z, 1.000000000e+00f);
    float r = t * p;
    if (ay > ax) r = 1.57079633f - r;
    if (x < 0.f) r = 3.14159265f - r;
    return copysignf(r, y);
}

// log(1 + u) for u >= 0 in FP32: below u = 0.5 as 2 atanh(s), s = u / (2 + u), by five terms of the odd series
// (within 1e-8 relative), above it the hardware logarithm of 1 + u (no cancellation there)
__device__ __forceinline__ float flog1p(float u)
{
    if (u < 0.5f) {
        const float s = u * __builtin_amdgcn_rcpf(2.f + u), z = s * s;
        float p = 1.f / 9;
        p = fmaf(p, z, 1.f / 7);
        p = fmaf(p, z, 1.f / 5);
        p = fmaf(p, z, 1.f / 3);
        p = fmaf(p, z, 1.f);
        return 2.f * s * p;
    }
    return __logf(1.f + u);
}

// source_velocity_exact (src/kernels/source_triangle.jl:75-89) in FP32 at a point off the source, from its offsets
// from the vertices SX and its height z over the plane, both taken in FP64:
//  - each edge_log as log1p(2 d t / s2d2): with t = a + b + d and s2d2 = (a + b)^2 - d^2 (its cancellation-free
//    form), t^2 - s2d2 = 2 d t exactly because |ri - rj| = d, so the ratio's excess over one needs no subtraction;
//  - the solid angle's numerator ra.(rb x rc) as 2 A z, which it equals exactly ((x - a).((b - a) x (c - a)), the
//    other terms of the triple product vanish), instead of a triple product of nearly parallel offsets;
// so that only the tangential sum's own cancellation (its terms ~ diam/r, their sum ~ (diam/r)^2) costs digits:
// about 1e-6 of the FP64 closed form out to the multipole switch.  In the plane (|z| <= 1e-12 diam) the solid angle
// is 0, as the FP64 form's off-element branch.
__device__ __forceinline__ V3<float> sv_exact_f32(const Elem &S, const V3<double> *SX, V3<double> x, double z)
{
    V3<float> r[3];
    float l[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const V3<double> d = x - SX[k];
        r[k] = mk3((float)d.x, (float)d.y, (float)d.z);
        l[k] = __builtin_amdgcn_sqrtf(dot(r[k], r[k]));
    }
    V3<float> wt = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = k, j = k == 2 ? 0 : k + 1;
        const float d = (float)S.len[k], ab = l[i] * l[j], p = dot(r[i], r[j]);
        const V3<float> c = cross(r[i], r[j]);
        const float s2d2 = p < 0.f ? 2.f * dot(c, c) * __builtin_amdgcn_rcpf(ab - p) : 2.f * (ab + p);
        const float L = flog1p(2.f * d * (l[i] + l[j] + d) * __builtin_amdgcn_rcpf(s2d2));
        wt = wt + L * mk3((float)S.nu[k][0], (float)S.nu[k][1], (float)S.nu[k][2]);
    }
    float Om = 0.f;
    if (!(fabs(z) <= 1e-12 * S.diam)) {
        const float N = (float)(2.0 * S.area * z);
        const float D = l[0] * l[1] * l[2] + dot(r[0], r[1]) * l[2] + dot(r[0], r[2]) * l[1] + dot(r[1], r[2]) * l[0];
        Om = 2.f * fatan2(N, D);
    }
    return (float)(1.0 / (4 * kPi)) * (wt + Om * mk3((float)S.n[0], (float)S.n[1], (float)S.n[2]));
}

// segment_velocity, src/kernels/segment.jl:15-25
template <typename T> __device__ __forceinline__ void seg_vel(V3<T> ra, V3<T> rb, V3<T> &va, V3<T> &vb)
{
    T Ra = dsqrt(dot(ra, ra)), Rb = dsqrt(dot(rb, rb));
    V3<T> c = cross(ra, rb);
    T p = dot(ra, rb);
    T den = p >= T(0) ? Ra * Rb + p : dot(c, c) / (Ra * Rb - p);
    if (den == T(0) || Ra == T(0) || Rb == T(0)) {
        va = vb = mk3(T(0), T(0), T(0));
        return;
    }
    V3<T> f = (T(1) / (T(4 * kPi) * den)) * c;
    va = (T(1) / Ra) * f;
    vb = (T(1) / Rb) * f;
}

// blob_velocity, src/kernels/blob.jl:21-25, r = x - xp
// (1/(4 pi)) (d2)^(-3/2) alpha x r with one reciprocal square root (the refined one in FP64, the hardware's in FP32)
// instead of a square root and a division
template <typename T> __device__ __forceinline__ V3<T> blob_vel(V3<T> r, V3<T> al, T d2core)
{
    const T d2 = dot(r, r) + d2core;
    const T i = vrsqrt(d2);
    return (T(1.0 / (4 * kPi)) * (i * i * i)) * cross(al, r);
}

// ---------------------------------------------------------------- classification (contraction off, FP64) ----------

__device__ inline double cls_norm(V3<double> a)
{
#pragma clang fp contract(off)
    return __builtin_sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

// The reference's far test norm(d) > T (assemble.jl:269, explicit.jl:58) without the square root wherever the squared
// comparison already decides it: s2 = d.d is formed exactly as cls_norm forms it, and with q = T*T,
// s2 > q(1 + 1e-12) puts the correctly rounded sqrt(s2) above T(1 + 4e-13), s2 < q(1 - 1e-12) puts it below
// T(1 - 4e-13) (a few units of rounding in s2, q and the bound are 1e-16 each).  In the band between, and when q is
// not a normal number, the test is the reference's own.  So the result is cls_norm(d) > T, bit for bit.
__device__ inline bool cls_far(V3<double> d, double T)
{
#pragma clang fp contract(off)
    const double s2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const double q = T * T;
    if (q >= 1e-290 && q <= 1e290) {
        if (s2 > q * (1.0 + 1e-12)) return true;
        if (s2 < q * (1.0 - 1e-12)) return false;
    }
    return __builtin_sqrt(s2) > T;
}

__device__ inline double cls_dot(V3<double> a, V3<double> b)
{
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

__device__ inline V3<double> cls_axpy(double s, V3<double> a, V3<double> b)  // s*a + b
{
#pragma clang fp contract(off)
    return mk3(s * a.x + b.x, s * a.y + b.y, s * a.z + b.z);
}

__device__ inline V3<double> cls_point(const V3<double> *x, const double *l)  // point, flat_triangle.jl:36
{
#pragma clang fp contract(off)
    return mk3((l[0] * x[0].x + l[1] * x[1].x) + l[2] * x[2].x, (l[0] * x[0].y + l[1] * x[1].y) + l[2] * x[2].y,
               (l[0] * x[0].z + l[1] * x[1].z) + l[2] * x[2].z);
}

// Vector helpers for the classification code: every floating-point operation of the tier decisions is written in a
// contraction-off scope (clang only fuses a multiply and an add that are both allowed to contract).
__device__ inline V3<double> cls_sub(V3<double> a, V3<double> b)
{
#pragma clang fp contract(off)
    return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}

__device__ inline V3<double> cls_add_scaled(V3<double> a, double s, V3<double> b)  // a + s*b
{
#pragma clang fp contract(off)
    return mk3(a.x + s * b.x, a.y + s * b.y, a.z + s * b.z);
}

// distance_to (Ericson 5.1.5), src/elements/flat_triangle.jl:59-78
__device__ inline double cls_tri_distance(const V3<double> *x, V3<double> p)
{
#pragma clang fp contract(off)
    V3<double> a = x[0], b = x[1], c = x[2];
    V3<double> ab = cls_sub(b, a), ac = cls_sub(c, a), ap = cls_sub(p, a);
    double d1 = cls_dot(ab, ap), d2 = cls_dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return cls_norm(ap);
    V3<double> bp = cls_sub(p, b);
    double d3 = cls_dot(ab, bp), d4 = cls_dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return cls_norm(bp);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return cls_norm(cls_sub(p, cls_add_scaled(a, d1 / (d1 - d3), ab)));
    V3<double> cp = cls_sub(p, c);
    double d5 = cls_dot(ab, cp), d6 = cls_dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return cls_norm(cp);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return cls_norm(cls_sub(p, cls_add_scaled(a, d2 / (d2 - d6), ac)));
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0)
        return cls_norm(cls_sub(p, cls_add_scaled(b, (d4 - d3) / ((d4 - d3) + (d5 - d6)), cls_sub(c, b))));
    double denom = 1 / (va + vb + vc);
    return cls_norm(cls_sub(p, cls_add_scaled(cls_add_scaled(a, vb * denom, ab), vc * denom, ac)));
}

// segment_projection, src/assembly/assemble.jl:145-149
__device__ inline double cls_seg_projection(V3<double> p, V3<double> a, V3<double> c, double *tout)
{
#pragma clang fp contract(off)
    V3<double> e = cls_sub(c, a);
    double t = cls_dot(cls_sub(p, a), e) / cls_dot(e, e);
    t = t < 0.0 ? 0.0 : t > 1.0 ? 1.0 : t;
    if (tout) *tout = t;
    return cls_norm(cls_sub(p, cls_add_scaled(a, t, e)));
}

__device__ __forceinline__ V3<double> ld3(const double *p) { return mk3(p[0], p[1], p[2]); }

}  // namespace vp
