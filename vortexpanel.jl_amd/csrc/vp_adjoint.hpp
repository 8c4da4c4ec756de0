// Backward assembly on gfx950: the gradient, with respect to every node coordinate, of
//
//     L(x) = lambdaᵀ (b(x) - A(x) mu)
//
// for fixed adjoint weights lambda (one per slot) and doublet values mu (one per column: slots, then wake nodes),
// with A and b assembled exactly as assemble does (src/assembly/assemble.jl:377-447) and sigma = -V∞.n (source_strengths,
// assemble.jl:364).  Together with the adjoint solve Mᵀ psi = dJ/dq this gives dJ/dx of any objective J(mu, x) of a
// Galerkin solution (the discrete adjoint); there is no counterpart in the reference, which has no derivatives.
//
// Nothing of dA/dx is stored.  Every pair's integrals are recomputed with the rule pieces the forward assembly chose
// (its far test and its non-far pair list; pieces at fixed barycentric positions), and per pair
//
//     P_ts = -A_T sum_q w_q psi_lambda(q) Q.w(x_q; S),   Q = sigma_S n_T + n_T x Gamma_S,
//
// (pair_block, assemble.jl:245-258, contracted with lambda and mu) is differentiated: the kernel w by forward-mode dual
// numbers with respect to the three vertices of S (three passes of three tangents; the derivative with respect to the
// point is minus their sum, by translation invariance), and the test triangle's area and normal, the source's normal
// (sigma) and sheet (Gamma) by their closed-form adjoints.  Filaments (segment_block, :312-321) and the freestream
// term (:422) likewise.  Gradients are added per node: body nodes, then wake nodes (offset by the body's node count).
//
// This file is included by vp_assemble.hip (it uses the plan's tiles, lists and rule generator).

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

// source_velocity_exact (src/kernels/source_triangle.jl:75-89) with the triangle's data formed from its vertices
// (the constructor, :18-27), in dual numbers.  Off the own element the solid angle is always the atan2 form (its
// derivative across the plane is the one the closed form has; the in-plane branch's zero has none).
__device__ __forceinline__ V3<D3> sv_exact_vertices(const V3<D3> *v, V3<double> x, int side, const Elem *self,
                                                    double diam)
{
    const V3<D3> e0 = v[1] - v[0], e1 = v[2] - v[1], e2 = v[0] - v[2];
    const V3<D3> nA = cross(e0, v[2] - v[0]);
    const D3 A2 = dsqrt(dot(nA, nA));
    const V3<D3> n = (D3(1.0) / A2) * nA;
    const D3 l0 = dsqrt(dot(e0, e0)), l1 = dsqrt(dot(e1, e1)), l2 = dsqrt(dot(e2, e2));
    const V3<D3> xd = dconst(x);
    const V3<D3> ra = xd - v[0], rb = xd - v[1], rc = xd - v[2];
    const V3<D3> nu0 = (D3(1.0) / l0) * cross(e0, n), nu1 = (D3(1.0) / l1) * cross(e1, n),
                 nu2 = (D3(1.0) / l2) * cross(e2, n);
    const V3<D3> wt = edge_log(ra, rb, l0) * nu0 + edge_log(rb, rc, l1) * nu1 + edge_log(rc, ra, l2) * nu2;
    D3 Om;
    const V3<D3> c = D3(1.0 / 3) * ((v[0] + v[1]) + v[2]);
    const D3 z = dot(xd - c, n);
    if (side != 0 && fabs(z.v) <= 1e-12 * diam) {
        Om = D3(side * 2 * kPi * inplane_fraction(*self, x, 1e-12));
    } else {
        const D3 la = dsqrt(dot(ra, ra)), lb = dsqrt(dot(rb, rb)), lc = dsqrt(dot(rc, rc));
        const D3 N = dot(ra, cross(rb, rc));
        const D3 D = la * lb * lc + dot(ra, rb) * lc + dot(ra, rc) * lb + dot(rb, rc) * la;
        Om = D3(2.0) * datan2(N, D);
    }
    return D3(1.0 / (4 * kPi)) * (wt + Om * n);
}

// source_velocity_far (source_triangle.jl:100-119) from the vertices, in dual numbers
__device__ __forceinline__ V3<D3> sv_far_vertices(const V3<D3> *v, V3<double> x)
{
    const V3<D3> c = D3(1.0 / 3) * ((v[0] + v[1]) + v[2]);
    const V3<D3> nA = cross(v[1] - v[0], v[2] - v[0]);
    SrcT<D3> s;
    for (int k = 0; k < 3; k++) s.v[k] = v[k] - c;
    s.area = D3(0.5) * dsqrt(dot(nA, nA));
    return sv_far_rel(s, dconst(x) - c);
}

// phi = Q.w(x; S) (w = source_velocity, source_triangle.jl:134-138, with the forward's switch) and d phi/d(vertex k of
// S) for k = 0, 1, 2; the value w(x) in wv.
__device__ void tri_point_grad(const V3<double> *XS, const Elem &S, V3<double> x, V3<double> Q, int side,
                               double kernel_far, V3<double> &wv, V3<double> *g)
{
    const V3<double> r0 = cls_sub(x, ld3(S.c));
    const double fd = kernel_far * S.diam;
    const bool far = cls_dot(r0, r0) > fd * fd;
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        V3<D3> v[3];
        for (int j = 0; j < 3; j++) v[j] = dseed(XS[j], j == k ? 1.0 : 0.0);
        const V3<D3> w = far ? sv_far_vertices(v, x) : sv_exact_vertices(v, x, side, &S, S.diam);
        const D3 phi = w.x * Q.x + w.y * Q.y + w.z * Q.z;
        g[k] = mk3(phi.d[0], phi.d[1], phi.d[2]);
        if (k == 0) wv = mk3(w.x.v, w.y.v, w.z.v);
    }
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

// ------------------------------------------------------------------------------------------ kernels ---------------

__global__ void slot_nodes(const int *__restrict__ corner_slot, const int *__restrict__ nodes, int nt, int *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt)
        for (int k = 0; k < 3; k++) out[corner_slot[3 * t + k]] = nodes[3 * t + k];
}

// The far tier (the forward's tiles and far test, far_pairs): lane = source element, wave = test triangle; the
// gradients are summed per local test slot and local source column in LDS and flushed per tile.
__global__ __launch_bounds__(256) void adj_far_pairs(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf,
                                                     const Elem *__restrict__ el, const int *__restrict__ tt_elems,
                                                     const uint8_t *__restrict__ tt_loc, const int *__restrict__ tt_rows,
                                                     const int *__restrict__ tt_nrow, const int *__restrict__ ts_elems,
                                                     const uint8_t *__restrict__ ts_loc, const int *__restrict__ ts_cols,
                                                     const int *__restrict__ ts_ncol, const int *__restrict__ src_cols,
                                                     const int *__restrict__ slot_node, const double *__restrict__ lam,
                                                     const double *__restrict__ mu, double vx, double vy, double vz,
                                                     double *__restrict__ G)
{
    __shared__ double GT[VMAX][3];
    __shared__ double GU[UMAX][3];
    const int tt = blockIdx.x, st = blockIdx.y;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < VMAX * 3; i += blockDim.x) GT[i / 3][i % 3] = 0.0;
    for (int i = threadIdx.x; i < UMAX * 3; i += blockDim.x) GU[i / 3][i % 3] = 0.0;
    const V3<double> Vinf = mk3(vx, vy, vz);
    const int s = ts_elems[st * TS + lane];
    const bool valid = s >= 0;
    V3<double> XS[3], Gam = mk3(0.0, 0.0, 0.0), sc = mk3(0.0, 0.0, 0.0);
    double muS[3] = {0, 0, 0}, sigS = 0.0, sdiam = 0.0;
    int uloc[3] = {0, 0, 0};
    if (valid) {
        const Elem &S = el[s];
        for (int k = 0; k < 3; k++) {
            XS[k] = ld3(S.v[k]);
            muS[k] = mu[src_cols[3 * s + k]];
            uloc[k] = ts_loc[(st * TS + lane) * 3 + k];
        }
        Gam = sheet_of(XS, muS);
        sigS = s < P.nt ? -dot(Vinf, ld3(S.n)) : 0.0;
        sc = ld3(S.c);
        sdiam = S.diam;
    }
    __syncthreads();
    const int off = rt.off[R_FAR];
    for (int i = wave; i < TT_MAX; i += 4) {
        const int t = tt_elems[tt * TT_MAX + i];
        if (t < 0) break;
        const TestFar &F = tf[t];
        bool far = false;
        if (valid && s != t) far = cls_norm(sc - ld3(F.c)) > P.far_ratio * fmax(F.diam, sdiam);
        if (!far) continue;
        const Elem &T = el[t];
        const V3<double> XT[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        const V3<double> nT = ld3(F.n);
        const int vl[3] = {tt_loc[(tt * TT_MAX + i) * 3], tt_loc[(tt * TT_MAX + i) * 3 + 1], tt_loc[(tt * TT_MAX + i) * 3 + 2]};
        const double la[3] = {lam[tt_rows[tt * VMAX + vl[0]]], lam[tt_rows[tt * VMAX + vl[1]]], lam[tt_rows[tt * VMAX + vl[2]]]};
        const V3<double> Q = sigS * nT + cross(nT, Gam);
        PairSums p;
        p.Sv = mk3(0.0, 0.0, 0.0);
        for (int k = 0; k < 3; k++) p.GS[k] = p.GX[k] = mk3(0.0, 0.0, 0.0);
#pragma unroll 1
        for (int q = 0; q < 3; q++) {
            const double a = (la[0] * F.wpsi[q][0] + la[1] * F.wpsi[q][1] + la[2] * F.wpsi[q][2]) / F.area;
            V3<double> wv, g[3];
            tri_point_grad(XS, el[s], ld3(F.x[q]), Q, 0, P.kernel_far, wv, g);
            const V3<double> gx = -1.0 * ((g[0] + g[1]) + g[2]);
            const double *l = rt.l + 3 * (off + q);
            p.Sv = p.Sv + a * wv;
            for (int k = 0; k < 3; k++) {
                p.GS[k] = p.GS[k] + a * g[k];
                p.GX[k] = p.GX[k] + (a * l[k]) * gx;
            }
        }
        V3<double> gT[3], gS[3];
        finish_tri_pair(XT, nT, F.area, XS, muS, sigS, Gam, s < P.nt, Vinf, Q, p, gT, gS);
        for (int k = 0; k < 3; k++) {
            atomicAdd(&GT[vl[k]][0], gT[k].x);
            atomicAdd(&GT[vl[k]][1], gT[k].y);
            atomicAdd(&GT[vl[k]][2], gT[k].z);
            atomicAdd(&GU[uloc[k]][0], gS[k].x);
            atomicAdd(&GU[uloc[k]][1], gS[k].y);
            atomicAdd(&GU[uloc[k]][2], gS[k].z);
        }
    }
    __syncthreads();
    const int nrow = tt_nrow[tt], ncol = ts_ncol[st];
    for (int i = threadIdx.x; i < nrow * 3; i += blockDim.x) {
        const double v = GT[i / 3][i % 3];
        if (v != 0.0) unsafeAtomicAdd(&G[3 * slot_node[tt_rows[tt * VMAX + i / 3]] + i % 3], v);
    }
    for (int i = threadIdx.x; i < ncol * 3; i += blockDim.x) {
        const double v = GU[i / 3][i % 3];
        if (v != 0.0) unsafeAtomicAdd(&G[3 * col_node(ts_cols[st * UMAX + i / 3], P.nslot, P.nn, slot_node) + i % 3], v);
    }
}

// Far filament pairs (seg_pairs' far test and far segment_block): one thread per (test triangle, filament).
__global__ void adj_seg_far(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf, const Elem *__restrict__ el,
                            const int *__restrict__ corner_slot, const int *__restrict__ nodes, const double *__restrict__ seg,
                            const int *__restrict__ segcol, const double *__restrict__ lam, const double *__restrict__ mu,
                            double *__restrict__ G)
{
#pragma clang fp contract(off)
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)P.nt * P.nseg) return;
    const int t = (int)(idx / P.nseg), f = (int)(idx % P.nseg);
    const TestFar &F = tf[t];
    const V3<double> fa = ld3(seg + 6 * f), fb = ld3(seg + 6 * f + 3);
    const double L = cls_norm(fb - fa);
    V3<double> mid = fa + fb;
    mid = mk3(mid.x / 2, mid.y / 2, mid.z / 2);
    if (!(cls_norm(mid - ld3(F.c)) > P.far_ratio * fmax(F.diam, L))) return;
    const Elem &T = el[t];
    const V3<double> XT[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
    const V3<double> nT = ld3(F.n);
    const double la[3] = {lam[corner_slot[3 * t]], lam[corner_slot[3 * t + 1]], lam[corner_slot[3 * t + 2]]};
    const double mua = mu[segcol[2 * f]], mub = mu[segcol[2 * f + 1]];
    V3<double> Sv = mk3(0.0, 0.0, 0.0), GF[2] = {Sv, Sv}, GX[3] = {Sv, Sv, Sv};
    for (int q = 0; q < 3; q++) {
        const double a = (la[0] * F.wpsi[q][0] + la[1] * F.wpsi[q][1] + la[2] * F.wpsi[q][2]) / F.area;
        V3<double> wv, g[2];
        seg_point_grad(fa, fb, ld3(F.x[q]), nT, mua, mub, wv, g);
        const V3<double> gx = -1.0 * (g[0] + g[1]);
        const double *l = rt.l + 3 * (rt.off[R_FAR] + q);
        Sv = Sv + a * wv;
        GF[0] = GF[0] + a * g[0];
        GF[1] = GF[1] + a * g[1];
        for (int k = 0; k < 3; k++) GX[k] = GX[k] + (a * l[k]) * gx;
    }
    V3<double> gT[3];
    for (int k = 0; k < 3; k++) gT[k] = (-F.area) * GX[k];
    area2_adj(XT, -0.5 * dot(nT, Sv), gT);
    normal_adj(XT, (-F.area) * Sv, gT);
    for (int k = 0; k < 3; k++) gadd(G, nodes[3 * t + k], gT[k]);
    gadd(G, P.nn + segcol[2 * f] - P.nslot, (-F.area) * GF[0]);
    gadd(G, P.nn + segcol[2 * f + 1] - P.nslot, (-F.area) * GF[1]);
}

// Non-far pairs: one wavefront per pair, the points of all its rule pieces (from the forward's rule generator) over
// the lanes, as group_eval does.
__global__ __launch_bounds__(256) void adj_group_eval(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                      const int *__restrict__ corner_slot, const int *__restrict__ nodes,
                                                      const int *__restrict__ src_cols, const double *__restrict__ seg,
                                                      const int *__restrict__ segcol, const Leaf *__restrict__ leaves,
                                                      const int4 *__restrict__ groups, unsigned int ngroups,
                                                      const double *__restrict__ lam, const double *__restrict__ mu,
                                                      double vx, double vy, double vz, double *__restrict__ G)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const V3<double> Vinf = mk3(vx, vy, vz);
    for (unsigned int gi = blockIdx.x * 4 + wave; gi < ngroups; gi += gridDim.x * 4) {
        const int4 grp = groups[gi];
        const int t = leaves[grp.x].t, s = leaves[grp.x].s, ntot = grp.z;
        const bool isseg = s >= P.nsrc;
        const Elem &T = el[t];
        const V3<double> XT[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        const V3<double> nT = ld3(T.n);
        const double la[3] = {lam[corner_slot[3 * t]], lam[corner_slot[3 * t + 1]], lam[corner_slot[3 * t + 2]]};
        V3<double> XS[3], Gam = mk3(0.0, 0.0, 0.0), Q = mk3(0.0, 0.0, 0.0), fa = Gam, fb = Gam;
        double muS[3] = {0, 0, 0}, sigS = 0.0, mua = 0.0, mub = 0.0;
        if (isseg) {
            const int f = s - P.nsrc;
            fa = ld3(seg + 6 * f);
            fb = ld3(seg + 6 * f + 3);
            mua = mu[segcol[2 * f]];
            mub = mu[segcol[2 * f + 1]];
        } else {
            const Elem &S = el[s];
            for (int k = 0; k < 3; k++) {
                XS[k] = ld3(S.v[k]);
                muS[k] = mu[src_cols[3 * s + k]];
            }
            Gam = sheet_of(XS, muS);
            sigS = s < P.nt ? -dot(Vinf, ld3(S.n)) : 0.0;
            Q = sigS * nT + cross(nT, Gam);
        }
        PairSums p;
        p.Sv = mk3(0.0, 0.0, 0.0);
        for (int k = 0; k < 3; k++) p.GS[k] = p.GX[k] = mk3(0.0, 0.0, 0.0);
        int li = grp.x, lbase = 0;
#pragma unroll 1
        for (int pp = lane; pp < ntot; pp += WAVE) {
            while (pp >= lbase + leaves[li].npt) {
                lbase += leaves[li].npt;
                li++;
            }
            const Leaf &lf = leaves[li];
            const int pq = pp - lbase, off = rt.off[lf.rid];
            const double *m = rt.l + 3 * (off + pq);
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * lf.b[0][j] + m[1] * lf.b[1][j]) + m[2] * lf.b[2][j];
            double psi[3];
            test_weights(P.test, l, psi);
            const double a = rt.w[off + pq] * lf.frac * (la[0] * psi[0] + la[1] * psi[1] + la[2] * psi[2]);
            const V3<double> x = cls_point(XT, l);
            V3<double> wv, g[3];
            if (isseg) {
                seg_point_grad(fa, fb, x, nT, mua, mub, wv, g);
                g[2] = mk3(0.0, 0.0, 0.0);
            } else {
                tri_point_grad(XS, el[s], x, Q, s == t ? 1 : 0, P.kernel_far, wv, g);
            }
            const V3<double> gx = -1.0 * ((g[0] + g[1]) + g[2]);
            p.Sv = p.Sv + a * wv;
            for (int k = 0; k < 3; k++) {
                p.GS[k] = p.GS[k] + a * g[k];
                p.GX[k] = p.GX[k] + (a * l[k]) * gx;
            }
        }
        // reduce the 21 sums over the wave
        p.Sv = mk3(wsum(p.Sv.x), wsum(p.Sv.y), wsum(p.Sv.z));
        for (int k = 0; k < 3; k++) {
            p.GS[k] = mk3(wsum(p.GS[k].x), wsum(p.GS[k].y), wsum(p.GS[k].z));
            p.GX[k] = mk3(wsum(p.GX[k].x), wsum(p.GX[k].y), wsum(p.GX[k].z));
        }
        if (lane == 0) {
            V3<double> gT[3], gS[3];
            if (isseg) {
                for (int k = 0; k < 3; k++) gT[k] = (-T.area) * p.GX[k];
                area2_adj(XT, -0.5 * dot(nT, p.Sv), gT);
                normal_adj(XT, (-T.area) * p.Sv, gT);
                const int f = s - P.nsrc;
                gadd(G, P.nn + segcol[2 * f] - P.nslot, (-T.area) * p.GS[0]);
                gadd(G, P.nn + segcol[2 * f + 1] - P.nslot, (-T.area) * p.GS[1]);
            } else {
                finish_tri_pair(XT, nT, T.area, XS, muS, sigS, Gam, s < P.nt, Vinf, Q, p, gT, gS);
                for (int k = 0; k < 3; k++) gadd(G, nodes[3 * s + k], gS[k]);
            }
            for (int k = 0; k < 3; k++) gadd(G, nodes[3 * t + k], gT[k]);
        }
    }
}

// The freestream term -(V∞.n_T) A_T int(psi_a)/A_T (assemble.jl:422) contracted with lambda: -(ti/2) Lambda V∞.N_T.
__global__ void adj_freestream(const Elem *__restrict__ el, const int *__restrict__ corner_slot, const int *__restrict__ nodes,
                               int nt, int test, const double *__restrict__ lam, double vx, double vy, double vz,
                               double *__restrict__ G)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const Elem &E = el[t];
    const V3<double> X[3] = {ld3(E.v[0]), ld3(E.v[1]), ld3(E.v[2])};
    const double ti = test == 1 ? 1.0 : 1.0 / 3;
    const double Lam = lam[corner_slot[3 * t]] + lam[corner_slot[3 * t + 1]] + lam[corner_slot[3 * t + 2]];
    V3<double> g[3] = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0)};
    cross_adj(X, (-0.5 * ti * Lam) * mk3(vx, vy, vz), g);
    for (int k = 0; k < 3; k++) gadd(G, nodes[3 * t + k], g[k]);
}

// ------------------------------------------------------------------------------------------ host ------------------

void Plan::adjoint(const double *lam_h, const double *mu_h, const double *vinf, double *G_h)
{
    hipStream_t st = ctx->stream;
    if (!assembled) throw std::invalid_argument("the plan has not been assembled (vp_plan_assemble first)");
    if (!d_lam) {
        d_lam = dalloc<double>(nslot);
        d_mu = dalloc<double>(ncol);
        d_G = dalloc<double>((size_t)3 * (nn + nw));
        d_slotnode = dalloc<int>(nslot);
        d_astats = dalloc<unsigned long long>(16);
        d_acount = dalloc<unsigned int>(2);
        owned.insert(owned.end(), {d_lam, d_mu, d_G, d_slotnode, d_astats, d_acount});
        slot_nodes<<<(nt + 255) / 256, 256, 0, st>>>(d_cs, d_nodes, nt, d_slotnode);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(d_lam, lam_h, sizeof(double) * nslot, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_mu, mu_h, sizeof(double) * ncol, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemsetAsync(d_G, 0, sizeof(double) * 3 * (nn + nw), st));
    const AsmParams P = params();
    const double vx = vinf[0], vy = vinf[1], vz = vinf[2];
    adj_far_pairs<<<dim3(ntt, nts), 256, 0, st>>>(P, rt, d_tf, d_el, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl, d_tsc, d_tsn,
                                                  d_cols, d_slotnode, d_lam, d_mu, vx, vy, vz, d_G);
    HIPCHECK(hipGetLastError());
    if (nseg) {
        const long long np = (long long)nt * nseg;
        adj_seg_far<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(P, rt, d_tf, d_el, d_cs, d_nodes, d_seg, d_segcol, d_lam,
                                                                   d_mu, d_G);
        HIPCHECK(hipGetLastError());
    }
    // rule pieces of every non-far pair of the last assembly (the forward's generator, all tiers)
    if (nnonfar) {
        for (int attempt = 0;; attempt++) {
            if (attempt == 3) throw std::runtime_error("adjoint rule pieces did not fit after regrowing");
            if (!d_aleaves) {
                aleaf_cap = std::max<unsigned int>(aleaf_cap, nnonfar * 4 + 1024);
                agroup_cap = std::max<unsigned int>(agroup_cap, nnonfar + 1024);
                d_aleaves = dalloc<Leaf>(aleaf_cap);
                d_agroups = dalloc<int4>(agroup_cap);
            }
            HIPCHECK(hipMemsetAsync(d_acount, 0, sizeof(unsigned int) * 2, st));
            touch_gen<<<std::min<unsigned int>((nnonfar + 255) / 256, 65535u), 256, 0, st>>>(
                P, rt, d_el, d_nodes, d_seg, d_nonfar, nnonfar, d_touch, ntouch, d_aleaves, aleaf_cap, d_agroups,
                agroup_cap, d_acount, d_astats, nullptr);
            HIPCHECK(hipGetLastError());
            unsigned int c[2];
            HIPCHECK(hipMemcpyAsync(c, d_acount, sizeof(c), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (c[0] <= aleaf_cap && c[1] <= agroup_cap) {
                adj_group_eval<<<std::min<unsigned int>((c[1] + 3) / 4, 65535u), 256, 0, st>>>(
                    P, rt, d_el, d_cs, d_nodes, d_cols, d_seg, d_segcol, d_aleaves, d_agroups, c[1], d_lam, d_mu, vx, vy,
                    vz, d_G);
                HIPCHECK(hipGetLastError());
                break;
            }
            (void)hipFree(d_aleaves);
            (void)hipFree(d_agroups);
            d_aleaves = nullptr;
            d_agroups = nullptr;
            aleaf_cap = std::max(aleaf_cap, c[0] + 1024);
            agroup_cap = std::max(agroup_cap, c[1] + 1024);
        }
    }
    adj_freestream<<<(nt + 255) / 256, 256, 0, st>>>(d_el, d_cs, d_nodes, nt, opts.test, d_lam, vx, vy, vz, d_G);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(G_h, d_G, sizeof(double) * 3 * (nn + nw), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace vp
