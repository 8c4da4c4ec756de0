// Galerkin assembly over mixed meshes on gfx950: isoparametric bilinear quads kept whole and flat triangles
// (assemble(g, dm, mesh::MixedMesh, V; opts, wake), src/assembly/mixed.jl:444-598).
//
// The quads' inner integrals have no closed form (src/elements/quad_field.jl): a tensor Gauss rule whose order grows
// as the target nears the quad, and within one diameter an adaptive polar rule about the target's closest point on
// the quad (sinh-graded angles and rays, principal values for targets on the quad).  A target point costs from 9 to
// several thousand kernel evaluations, so the work is cut by point, not by pair:
//
//   mixed_quad_setup  IsoQuad records (corners, warp, centroid, normal, area, diameter, reach), FP64, contraction off
//   mixed_nonfar_count / mixed_far
//                     thread per (test element, source) pair: the far test (centroid distance against far_ratio
//                     times the larger diameter); a far pair is evaluated on the spot with the far rule and added to A
//                     (its inner integrals are always the tensor Gauss rule at that distance); a non-far pair is
//                     appended to a list
//   mixed_rules<0|1>  thread per non-far pair: the reference's tier and rule builders (quad_pair_rule,
//                     square_contact_rule, vertex_square, square_near_rule; tri_pair_rule with a quad or triangle
//                     source; the wake and filament rules), FP64 with contraction off so that tiers and point counts
//                     are the oracle's bit for bit.  Pass 0 counts each pair's rule pieces and points; the host scans
//                     the counts; pass 1 writes the pieces (a rectangle of the parameter square carrying a square
//                     rule, a triangle of the parameter square or a barycentric piece of a test triangle carrying a
//                     triangle rule) at their offsets
//   mixed_eval        thread per rule point: the test point, its normal and test weights, the source's field (quad:
//                     tensor Gauss or polar; triangle: the closed form or multipole; filament), and the pair's block
//                     entries; a wave whose lanes share one pair sums them by shuffles and adds once
//   mixed_freestream  thread per element: the freestream part of b (3 x 3 Gauss on a quad, exact)
//
// Indices: test elements e in [0, nq) quads, [nq, ne) triangles; sources s in [0, ne) body elements, then nwt wake
// triangles, then nwf filaments.  A is column-major nslot x (nslot + nw), A[row + col * lda].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "vp_capi_internal.hpp"
#include "vp_device.hpp"
#include "vp_rules.hpp"

namespace vp {

__global__ void elem_setup(const double *x, const int *tris, int ne, Elem *out);  // vp_assemble.hip

namespace {

#define MCHECK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum { MT_SELF = 0, MT_EDGE, MT_VERTEX, MT_CONTACT, MT_NEAR, MT_FAR };

// Rule table ids: square rules (two parameters per point, the third coordinate zero) and triangle rules (barycentric)
enum {
    MR_SFAR = 0, MR_SNEAR = 1, MR_SSELF = 2, MR_SEDGE0 = 3, MR_SVERT0 = 7, MR_VBASE = 11,
    MR_TFAR = 12, MR_TNEAR = 13, MR_TSELF = 14, MR_TEDGE0 = 15, MR_TVERT0 = 18, MR_COUNT = 21
};

constexpr int MAXD = 10;  // deepest quadtree split of the rule builders (near_depth + 4 at most)

struct QuadRec {  // IsoQuad, src/elements/iso_quad.jl:13-37
    double x[4][3];
    double warp[3], c[3], n[3];
    double area, diam, reach, bad;
    double pad[7];
};
static_assert(sizeof(QuadRec) == 32 * 8, "QuadRec layout");

struct MxPiece {
    double c[9];  // kind 0: lo, hi of a parameter rectangle; 1: the parameter triangle's corners a, b, c; 2: b[3][3]
    double frac;  // kinds 1 and 2: the piece's area fraction
    int pair, rule, kind, npts;
    long long pt0;  // its first point in the flattened point list
};

struct MxArgs {
    int nq, nt, ne, nwt, nwf, nsrc, nslot, test, near_depth, ntouch;
    long long lda;
    double far_ratio, near_ratio, corner_angle, contact_rel, kernel_far;
    double V[3];
    const QuadRec *quad;
    const Elem *tri, *wake;
    const int *qnodes, *qslots, *tnodes, *tslots, *wtris, *fil;
    const double *wx;
    const int2 *touch;
    const double *rl, *rw;
    int roff[MR_COUNT], rn[MR_COUNT];
    const double *gx, *gw;  // Gauss-Legendre on [0, 1], n = 1..48, rule n at n (n - 1) / 2
    double *A, *b;
    uint8_t *tiers;
    int *err;
};

// ------------------------------------------------------------------ classification geometry (contraction off) ---

__device__ inline V3<double> mq_corner(const QuadRec &Q, int k) { return mk3(Q.x[k][0], Q.x[k][1], Q.x[k][2]); }

// point, src/elements/iso_quad.jl:49-52
__device__ inline V3<double> mq_point(const V3<double> *X, double xi, double eta)
{
#pragma clang fp contract(off)
    const double N0 = (1 - xi) * (1 - eta), N1 = xi * (1 - eta), N2 = xi * eta, N3 = (1 - xi) * eta;
    return mk3(((N0 * X[0].x + N1 * X[1].x) + N2 * X[2].x) + N3 * X[3].x,
               ((N0 * X[0].y + N1 * X[1].y) + N2 * X[2].y) + N3 * X[3].y,
               ((N0 * X[0].z + N1 * X[1].z) + N2 * X[2].z) + N3 * X[3].z);
}

// tangents, src/elements/iso_quad.jl:55-60
__device__ inline void mq_tangents(const V3<double> *X, double xi, double eta, V3<double> &a1, V3<double> &a2)
{
#pragma clang fp contract(off)
    const V3<double> d10 = cls_sub(X[1], X[0]), d23 = cls_sub(X[2], X[3]), d30 = cls_sub(X[3], X[0]),
                     d21 = cls_sub(X[2], X[1]);
    a1 = mk3((1 - eta) * d10.x + eta * d23.x, (1 - eta) * d10.y + eta * d23.y, (1 - eta) * d10.z + eta * d23.z);
    a2 = mk3((1 - xi) * d30.x + xi * d21.x, (1 - xi) * d30.y + xi * d21.y, (1 - xi) * d30.z + xi * d21.z);
}

__device__ inline V3<double> cls_cross(V3<double> a, V3<double> b)
{
#pragma clang fp contract(off)
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// project, src/elements/iso_quad.jl:108-146: the closest point's parameters P and the distance
__device__ double mq_project(const V3<double> *X, V3<double> warp, V3<double> p, double *P)
{
#pragma clang fp contract(off)
    double best_d = INFINITY, b0 = 0.5, b1 = 0.5, q0 = 0.5, q1 = 0.5, f0 = INFINITY;
    for (int a = 0; a <= 4; a++)
        for (int b = 0; b <= 4; b++) {
            const double g0 = a / 4.0, g1 = b / 4.0;
            const V3<double> r = cls_sub(mq_point(X, g0, g1), p);
            const double rr = cls_dot(r, r);
            if (rr < f0) {
                f0 = rr;
                q0 = g0;
                q1 = g1;
            }
        }
    for (int it = 0; it < 60; it++) {
        const V3<double> r = cls_sub(mq_point(X, q0, q1), p);
        V3<double> a1, a2;
        mq_tangents(X, q0, q1, a1, a2);
        const double g1 = cls_dot(r, a1), g2 = cls_dot(r, a2);
        const double H11 = cls_dot(a1, a1), H22 = cls_dot(a2, a2), H12 = cls_dot(a1, a2) + cls_dot(r, warp);
        const double det = H11 * H22 - H12 * H12;
        double s0, s1;
        if (det > 0) {
            s0 = (H22 * g1 - H12 * g2) / det;
            s1 = (H11 * g2 - H12 * g1) / det;
        } else {
            const double hs = H11 + H22;
            s0 = g1 / hs;
            s1 = g2 / hs;
        }
        double t = 1.0;
        if (s0 > 0) t = fmin(t, (q0 + 0.5) / s0);
        if (s0 < 0) t = fmin(t, (q0 - 1.5) / s0);
        if (s1 > 0) t = fmin(t, (q1 + 0.5) / s1);
        if (s1 < 0) t = fmin(t, (q1 - 1.5) / s1);
        const double f = cls_dot(r, r);
        while (t > 1e-10) {
            const V3<double> rn = cls_sub(mq_point(X, q0 - t * s0, q1 - t * s1), p);
            if (cls_dot(rn, rn) <= f) break;
            t /= 2;
        }
        q0 = q0 - t * s0;
        q1 = q1 - t * s1;
        if (t * (fabs(s0) + fabs(s1)) < 1e-14) break;
    }
    if (0 <= q0 && q0 <= 1 && 0 <= q1 && q1 <= 1) {
        best_d = cls_norm(cls_sub(mq_point(X, q0, q1), p));
        b0 = q0;
        b1 = q1;
    }
    for (int k = 0; k < 4; k++) {  // segment_closest, iso_quad.jl:94-98
        const V3<double> a = X[k], e = cls_sub(X[(k + 1) % 4], a);
        double t = cls_dot(cls_sub(p, a), e) / cls_dot(e, e);
        t = t < 0.0 ? 0.0 : t > 1.0 ? 1.0 : t;
        const double d = cls_norm(cls_sub(p, cls_add_scaled(a, t, e)));
        if (d < best_d) {
            best_d = d;
            if (k == 0) { b0 = t; b1 = 0.0; }
            else if (k == 1) { b0 = 1.0; b1 = t; }
            else if (k == 2) { b0 = 1 - t; b1 = 1.0; }
            else { b0 = 0.0; b1 = 1 - t; }
        }
    }
    P[0] = b0;
    P[1] = b1;
    return best_d;
}

// A source as the rule builders see it (distance_to, corners, centroid, diameter, reach): a flat triangle (kind 0),
// an isoparametric quad (1) or a filament segment (2)
struct GSrc {
    int kind, nc;
    V3<double> x[4];
    V3<double> c, warp;
    double diam, reach;
};

__device__ inline double gsrc_distance(const GSrc &S, V3<double> p)
{
    if (S.kind == 0) return cls_tri_distance(S.x, p);
    if (S.kind == 1) {
        double P[2];
        return mq_project(S.x, S.warp, p, P);
    }
    return cls_seg_projection(p, S.x[0], S.x[1], nullptr);
}

__device__ inline GSrc gsrc_quad(const QuadRec &Q)
{
    GSrc S;
    S.kind = 1;
    S.nc = 4;
    for (int k = 0; k < 4; k++) S.x[k] = mq_corner(Q, k);
    S.c = ld3(Q.c);
    S.warp = ld3(Q.warp);
    S.diam = Q.diam;
    S.reach = Q.reach;
    return S;
}

__device__ inline GSrc gsrc_tri(const Elem &E)
{
    GSrc S;
    S.kind = 0;
    S.nc = 3;
    for (int k = 0; k < 3; k++) S.x[k] = ld3(E.v[k]);
    S.x[3] = S.x[2];
    S.c = ld3(E.c);
    S.warp = mk3(0.0, 0.0, 0.0);
    S.diam = E.diam;
    S.reach = E.reach;
    return S;
}

// --------------------------------------------------------------------------------------------- rule emission ---

struct Emitter {
    MxPiece *out;  // null: count only
    int npieces;
    long long npts, pt0;
    int pair;
    const int *rn;

    __device__ void put(int rule, int kind, const double *c, int nc, double frac)
    {
        const int n = rn[rule];
        if (out) {
            MxPiece pc;
            for (int i = 0; i < 9; i++) pc.c[i] = i < nc ? c[i] : 0.0;
            pc.frac = frac;
            pc.pair = pair;
            pc.rule = rule;
            pc.kind = kind;
            pc.npts = n;
            pc.pt0 = pt0 + npts;
            out[npieces] = pc;
        }
        npieces++;
        npts += n;
    }
    __device__ void rect(int rule, double lo0, double lo1, double hi0, double hi1)
    {
        const double c[4] = {lo0, lo1, hi0, hi1};
        put(rule, 0, c, 4, 1.0);
    }
    __device__ void ptri(int rule, const double *a, const double *b, const double *cc, double frac)
    {
        const double c[6] = {a[0], a[1], b[0], b[1], cc[0], cc[1]};
        put(rule, 1, c, 6, frac);
    }
    __device__ void btri(int rule, const double (*b)[3], double frac)
    {
        double c[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) c[3 * i + j] = b[i][j];
        put(rule, 2, c, 9, frac);
    }
};

// ---------------------------------------------------------------- rules on a test triangle (generic source) ---

// quarters, src/assembly/assemble.jl:79-82
__device__ inline void mx_quarters(const double (*b)[3], double (*q)[3][3])
{
#pragma clang fp contract(off)
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

struct BEnt {
    double b[3][3];
    double frac;
    int level, type;  // type 0: a contact piece, 1: a near piece
};
constexpr int BSTACK = 4 * 4 + 1 + 3 * MAXD + 1;

// vertex_pieces!, src/assembly/assemble.jl:102-116 (k 0-based; p the piece's physical corners)
__device__ void mx_vertex_pieces(Emitter &em, int rule, const V3<double> *p, int k, const double (*b)[3], double frac,
                                 double corner_angle)
{
#pragma clang fp contract(off)
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    const V3<double> u = cls_sub(p[i], p[k]), v = cls_sub(p[j], p[k]);
    double cs = cls_dot(u, v) / (cls_norm(u) * cls_norm(v));
    cs = cs < -1.0 ? -1.0 : cs > 1.0 ? 1.0 : cs;
    const double th = acos(cs);
    if (th <= kPi / 2) {
        double bb[3][3];
        for (int c = 0; c < 3; c++) {
            bb[0][c] = b[k][c];
            bb[1][c] = b[i][c];
            bb[2][c] = b[j][c];
        }
        em.btri(rule, bb, frac);
        return;
    }
    int m = (int)ceil(th * (180 / kPi) / corner_angle);
    m = m < 1 ? 1 : m > 60 ? 60 : m;  // corner_angle >= 3 degrees keeps m <= 60 (a guard against NaN angles)
    double tprev = 0.0;
    for (int q = 0; q < m; q++) {
        double tnext = 1.0;
        if (q + 1 < m) {
            const double ph = th * (q + 1) / m;
            tnext = cls_norm(u) * sin(ph) / (cls_norm(u) * sin(ph) + cls_norm(v) * sin(th - ph));
        }
        double bb[3][3];
        for (int c = 0; c < 3; c++) {
            bb[0][c] = b[k][c];
            bb[1][c] = (1 - tprev) * b[i][c] + tprev * b[j][c];
            bb[2][c] = (1 - tnext) * b[i][c] + tnext * b[j][c];
        }
        em.btri(rule, bb, frac * (tnext - tprev));
        tprev = tnext;
    }
}

// edge_splits of the piece with corners p along edge k (src/assembly/assemble.jl:160-172), sorted
__device__ int mx_edge_splits(const V3<double> *p, int k, const GSrc &S, double tol, double *ts)
{
#pragma clang fp contract(off)
    const V3<double> a = p[k], c = p[(k + 1) % 3];
    const double L = cls_norm(cls_sub(c, a));
    int n = 0;
    for (int q = 0; q < S.nc; q++) {
        double t;
        const double d = cls_seg_projection(S.x[q], a, c, &t);
        if (d <= tol && t * L > tol && (1 - t) * L > tol) ts[n++] = t;
    }
    for (int i = 1; i < n; i++)
        for (int j = i; j > 0 && ts[j] < ts[j - 1]; j--) {
            const double x = ts[j];
            ts[j] = ts[j - 1];
            ts[j - 1] = x;
        }
    return n;
}

// contact_pieces! (src/assembly/assemble.jl:209-238) and near_pieces! (:131-142) from the root b0 on one stack.
// root_type 0: the contact rule's pieces; 1: the near rule alone (depth given).  Returns 0, or -1 on a stack overflow.
__device__ int mx_tri_pieces(Emitter &em, const V3<double> *TX, const GSrc &S, const MxArgs &P, double tol, int root_type,
                             int near_depth)
{
#pragma clang fp contract(off)
    BEnt st[BSTACK];
    int top = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) st[0].b[i][j] = i == j ? 1.0 : 0.0;
    st[0].frac = 1.0;
    st[0].level = 0;
    st[0].type = root_type;
    top = 1;
    while (top > 0) {
        const BEnt E = st[--top];
        V3<double> p[3];
        for (int k = 0; k < 3; k++) p[k] = cls_point(TX, E.b[k]);
        if (E.type == 1) {  // near_pieces!
            const double size = fmax(fmax(cls_norm(cls_sub(p[1], p[0])), cls_norm(cls_sub(p[2], p[1]))),
                                     cls_norm(cls_sub(p[0], p[2])));
            const V3<double> s3 = cls_add_scaled(cls_add_scaled(p[0], 1.0, p[1]), 1.0, p[2]);
            const V3<double> cen = mk3(s3.x / 3.0, s3.y / 3.0, s3.z / 3.0);
            if (E.level < near_depth && gsrc_distance(S, cen) < P.near_ratio * size) {
                if (top + 4 > BSTACK) return -1;
                double q[4][3][3];
                mx_quarters(E.b, q);
                for (int c = 3; c >= 0; c--) {
                    BEnt &N = st[top++];
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) N.b[i][j] = q[c][i][j];
                    N.frac = E.frac / 4;
                    N.level = E.level + 1;
                    N.type = 1;
                }
            } else {
                em.btri(MR_TNEAR, E.b, E.frac);
            }
            continue;
        }
        // a contact piece
        if (E.level < 4) {
            bool split = false;
            for (int k = 0; k < 3 && !split; k++) {
                double sp[4];
                const int ns = mx_edge_splits(p, k, S, tol, sp);
                if (ns == 0) continue;
                split = true;
                double ts[6];
                ts[0] = 0.0;
                for (int i = 0; i < ns; i++) ts[i + 1] = sp[i];
                ts[ns + 1] = 1.0;
                if (top + ns + 1 > BSTACK) return -1;
                const double *bk = E.b[k], *bk1 = E.b[(k + 1) % 3], *opp = E.b[(k + 2) % 3];
                for (int i = ns; i >= 0; i--) {
                    BEnt &N = st[top++];
                    for (int c = 0; c < 3; c++) {
                        N.b[0][c] = (1 - ts[i]) * bk[c] + ts[i] * bk1[c];
                        N.b[1][c] = (1 - ts[i + 1]) * bk[c] + ts[i + 1] * bk1[c];
                        N.b[2][c] = opp[c];
                    }
                    N.frac = E.frac * (ts[i + 1] - ts[i]);
                    N.level = E.level + 1;
                    N.type = 0;
                }
            }
            if (split) continue;
        }
        int on[3], nc = 0;
        for (int k = 0; k < 3; k++) {
            on[k] = gsrc_distance(S, p[k]) <= tol;
            nc += on[k];
        }
        if (nc == 2) {
            const int k = !on[2] ? 0 : !on[0] ? 1 : 2;
            const V3<double> mid = cls_add_scaled(p[k], 1.0, p[(k + 1) % 3]);
            if (gsrc_distance(S, mk3(mid.x / 2.0, mid.y / 2.0, mid.z / 2.0)) <= tol) {
                em.btri(MR_TEDGE0 + k, E.b, E.frac);
                continue;
            }
        }
        if (nc == 1) {
            mx_vertex_pieces(em, MR_TVERT0, p, on[0] ? 0 : on[1] ? 1 : 2, E.b, E.frac, P.corner_angle);
        } else if (nc == 0) {
            if (top + 1 > BSTACK) return -1;
            BEnt &N = st[top++];
            N = E;
            N.level = 0;
            N.type = 1;
        } else {
            em.btri(MR_TSELF, E.b, E.frac);
        }
    }
    return 0;
}

// vertex_rule, src/assembly/assemble.jl:119-125 (k 0-based)
__device__ void mx_vertex_rule(Emitter &em, const V3<double> *TX, int k, const MxArgs &P)
{
#pragma clang fp contract(off)
    const V3<double> u = cls_sub(TX[(k + 1) % 3], TX[k]), v = cls_sub(TX[(k + 2) % 3], TX[k]);
    double id[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (cls_dot(u, v) >= 0) {
        em.btri(MR_TVERT0 + k, id, 1.0);
        return;
    }
    mx_vertex_pieces(em, MR_TVERT0, TX, k, id, 1.0, P.corner_angle);
}

// ---------------------------------------------------------------------------------- rules on a test quad ---

// vertex_rect!, src/assembly/mixed.jl:118-151 (k 0-based), onto [lo, hi]
__device__ void mx_vertex_rect(Emitter &em, const V3<double> *X, int k, const double *lo, const double *hi,
                               double corner_angle)
{
#pragma clang fp contract(off)
    const double rc[4][2] = {{lo[0], lo[1]}, {hi[0], lo[1]}, {hi[0], hi[1]}, {lo[0], hi[1]}};
    const double *C = rc[k], *A = rc[(k + 1) % 4], *O = rc[(k + 2) % 4], *B = rc[(k + 3) % 4];
    V3<double> a1, a2;
    mq_tangents(X, C[0], C[1], a1, a2);
    const V3<double> u = cls_add_scaled(mk3((A[0] - C[0]) * a1.x, (A[0] - C[0]) * a1.y, (A[0] - C[0]) * a1.z), A[1] - C[1], a2);
    const V3<double> v = cls_add_scaled(mk3((B[0] - C[0]) * a1.x, (B[0] - C[0]) * a1.y, (B[0] - C[0]) * a1.z), B[1] - C[1], a2);
    double cs = cls_dot(u, v) / (cls_norm(u) * cls_norm(v));
    cs = cs < -1.0 ? -1.0 : cs > 1.0 ? 1.0 : cs;
    const double th = acos(cs);
    int m = th <= kPi / 2 ? 1 : (int)ceil(th * (180 / kPi) / corner_angle);
    m = m < 1 ? 1 : m > 60 ? 60 : m;
    const double nu = cls_norm(u);
    const V3<double> uh = mk3(u.x / nu, u.y / nu, u.z / nu);
    const double vu = cls_dot(v, uh);
    const V3<double> wv = cls_sub(v, mk3(vu * uh.x, vu * uh.y, vu * uh.z));
    const double nw = cls_norm(wv);
    const V3<double> wh = mk3(wv.x / nw, wv.y / nw, wv.z / nw);
    const double G11 = cls_dot(a1, a1), G12 = cls_dot(a1, a2), G22 = cls_dot(a2, a2);
    const double det = G11 * G22 - G12 * G12;
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1];
    const double tol = 1e-12 * sqrt(dx * dx + dy * dy);
    const double OA0 = O[0] - A[0], OA1 = O[1] - A[1];
    const double nOA = sqrt(OA0 * OA0 + OA1 * OA1);
    auto along = [](const double *Pp, const double *X0, const double *Y0) {
        const double yx0 = Y0[0] - X0[0], yx1 = Y0[1] - X0[1];
        const double t = ((Pp[0] - X0[0]) * yx0 + (Pp[1] - X0[1]) * yx1) / (yx0 * yx0 + yx1 * yx1);
        return t < 0.0 ? 0.0 : t > 1.0 ? 1.0 : t;
    };
    auto spos = [&](const double *Pp) {
        const double cr = (Pp[0] - A[0]) * OA1 - (Pp[1] - A[1]) * OA0;
        return fabs(cr) / nOA <= tol ? along(Pp, A, O) : 1 + along(Pp, O, B);
    };
    auto area2 = [](const double *Cc, const double *Pp, const double *Qq) {
        return fabs((Pp[0] - Cc[0]) * (Qq[1] - Cc[1]) - (Pp[1] - Cc[1]) * (Qq[0] - Cc[0])) / 2;
    };
    double prev[2] = {A[0], A[1]};
    for (int i = 0; i < m; i++) {
        double next[2];
        if (i + 1 == m) {
            next[0] = B[0];
            next[1] = B[1];
        } else {  // ray_exit, mixed.jl:99-106, along G \ (a1.d, a2.d)
            const double ang = (i + 1) * th / m;
            const double ca = cos(ang), sa = sin(ang);
            const V3<double> d = mk3(ca * uh.x + sa * wh.x, ca * uh.y + sa * wh.y, ca * uh.z + sa * wh.z);
            const double bb0 = cls_dot(a1, d), bb1 = cls_dot(a2, d);
            const double dir[2] = {(G22 * bb0 - G12 * bb1) / det, (G11 * bb1 - G12 * bb0) / det};
            double t = INFINITY;
            for (int c = 0; c < 2; c++) {
                if (dir[c] > 1e-14) t = fmin(t, (hi[c] - C[c]) / dir[c]);
                if (dir[c] < -1e-14) t = fmin(t, (lo[c] - C[c]) / dir[c]);
            }
            next[0] = C[0] + t * dir[0];
            next[1] = C[1] + t * dir[1];
        }
        const double sP = spos(prev), sQ = spos(next);
        if (sP < 1 - 1e-9 && sQ > 1 + 1e-9) {
            em.ptri(MR_VBASE, C, prev, O, area2(C, prev, O));
            em.ptri(MR_VBASE, C, O, next, area2(C, O, next));
        } else {
            em.ptri(MR_VBASE, C, prev, next, area2(C, prev, next));
        }
        prev[0] = next[0];
        prev[1] = next[1];
    }
}

// vertex_square, src/assembly/mixed.jl:154-160
__device__ void mx_vertex_square(Emitter &em, const V3<double> *X, int k, const MxArgs &P)
{
#pragma clang fp contract(off)
    const V3<double> u = cls_sub(X[(k + 1) % 4], X[k]), v = cls_sub(X[(k + 3) % 4], X[k]);
    if (cls_dot(u, v) >= 0) {
        em.rect(MR_SVERT0 + k, 0.0, 0.0, 1.0, 1.0);
        return;
    }
    const double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0};
    mx_vertex_rect(em, X, k, lo, hi, P.corner_angle);
}

struct REnt {
    double lo0, lo1, hi0, hi1;
    int level;
};
constexpr int RSTACK = 3 * MAXD + 4;

// near_squares! (src/elements/quad_rules.jl:103-117, square = true: the piece of side `size` at lo) and near_rects!
// (src/assembly/mixed.jl:84-96) on [lo, hi]; the leaves carry the near square rule.  Returns 0 or -1 (stack).
__device__ int mx_near_rects(Emitter &em, const V3<double> *X, const GSrc &S, double ratio, int depth, double lo0,
                             double lo1, double hi0, double hi1, bool square)
{
#pragma clang fp contract(off)
    REnt st[RSTACK];
    int top = 0;
    st[top++] = REnt{lo0, lo1, hi0, hi1, 0};
    while (top > 0) {
        const REnt E = st[--top];
        V3<double> c1, c2, c3, c4, mid;
        if (square) {  // corners lo + (0|size, 0|size), centre lo + size/2
            const double size = E.hi0 - E.lo0;
            c1 = mq_point(X, E.lo0, E.lo1);
            c2 = mq_point(X, E.lo0 + size, E.lo1 + 0.0);
            c3 = mq_point(X, E.lo0 + size, E.lo1 + size);
            c4 = mq_point(X, E.lo0 + 0.0, E.lo1 + size);
            mid = mq_point(X, E.lo0 + size / 2, E.lo1 + size / 2);
        } else {
            c1 = mq_point(X, E.lo0, E.lo1);
            c2 = mq_point(X, E.hi0, E.lo1);
            c3 = mq_point(X, E.hi0, E.hi1);
            c4 = mq_point(X, E.lo0, E.hi1);
            mid = mq_point(X, (E.lo0 + E.hi0) / 2, (E.lo1 + E.hi1) / 2);
        }
        const double extent = fmax(cls_norm(cls_sub(c3, c1)), cls_norm(cls_sub(c4, c2)));
        if (E.level < depth && gsrc_distance(S, mid) < ratio * extent) {
            if (top + 4 > RSTACK) return -1;
            REnt ch[4];
            if (square) {
                const double h = (E.hi0 - E.lo0) / 2;
                const double off[4][2] = {{0.0, 0.0}, {h, 0.0}, {0.0, h}, {h, h}};
                for (int c = 0; c < 4; c++)
                    ch[c] = REnt{E.lo0 + off[c][0], E.lo1 + off[c][1], E.lo0 + off[c][0] + h, E.lo1 + off[c][1] + h,
                                 E.level + 1};
            } else {
                const double m0 = (E.lo0 + E.hi0) / 2, m1 = (E.lo1 + E.hi1) / 2;
                ch[0] = REnt{E.lo0, E.lo1, m0, m1, E.level + 1};
                ch[1] = REnt{m0, E.lo1, E.hi0, m1, E.level + 1};
                ch[2] = REnt{E.lo0, m1, m0, E.hi1, E.level + 1};
                ch[3] = REnt{m0, m1, E.hi0, E.hi1, E.level + 1};
            }
            for (int c = 3; c >= 0; c--) st[top++] = ch[c];
        } else {
            em.rect(MR_SNEAR, E.lo0, E.lo1, E.hi0, E.hi1);
        }
    }
    return 0;
}

// square_contact_rule, src/assembly/mixed.jl:203-232: returns 1 (rule emitted), 0 (no contact) or -1 (stack)
__device__ int mx_square_contact(Emitter &em, const QuadRec &T, const V3<double> *X, const GSrc &S, const int *shared,
                                 const MxArgs &P)
{
#pragma clang fp contract(off)
    const double tol = P.contact_rel * fmin(T.diam, S.diam);
    if (cls_norm(cls_sub(S.c, ld3(T.c))) > T.reach + S.reach + tol) return 0;
    double sp[4][4];
    int ns[4];
    for (int k = 0; k < 4; k++) {  // side_splits, mixed.jl:166-179
        const V3<double> a = X[k], c = X[(k + 1) % 4];
        const double L = cls_norm(cls_sub(c, a));
        ns[k] = 0;
        for (int q = 0; q < S.nc; q++) {
            double t;
            const double d = cls_seg_projection(S.x[q], a, c, &t);
            if (d <= tol && t * L > tol && (1 - t) * L > tol) sp[k][ns[k]++] = t;
        }
    }
    bool extra = false;
    for (int k = 0; k < 4 && !extra; k++) extra = !shared[k] && gsrc_distance(S, X[k]) <= tol;
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
    // cuts, mixed.jl:182-190: sorted, near-duplicates merged
    auto cuts = [](double *v, int n) {
        for (int i = 1; i < n; i++)
            for (int j = i; j > 0 && v[j] < v[j - 1]; j--) {
                const double x = v[j];
                v[j] = v[j - 1];
                v[j - 1] = x;
            }
        int m = 1;
        for (int i = 1; i < n; i++)
            if (v[i] - v[m - 1] > 1e-9) v[m++] = v[i];
        return m;
    };
    nx = cuts(xc, nx);
    ne = cuts(ec, ne);
    for (int i = 0; i < nx - 1; i++)
        for (int j = 0; j < ne - 1; j++) {
            const double lo[2] = {xc[i], ec[j]}, hi[2] = {xc[i + 1], ec[j + 1]};
            const double c[4][2] = {{lo[0], lo[1]}, {hi[0], lo[1]}, {hi[0], hi[1]}, {lo[0], hi[1]}};
            int on[4], nc = 0;
            for (int q = 0; q < 4; q++) {
                on[q] = gsrc_distance(S, mq_point(X, c[q][0], c[q][1])) <= tol;
                nc += on[q];
            }
            int k = -1;
            for (int q = 0; q < 4 && k < 0; q++)
                if (on[q] && on[(q + 1) % 4]) k = q;
            if (nc == 2 && k >= 0 &&
                gsrc_distance(S, mq_point(X, (c[k][0] + c[(k + 1) % 4][0]) / 2, (c[k][1] + c[(k + 1) % 4][1]) / 2)) <= tol) {
                em.rect(MR_SEDGE0 + k, lo[0], lo[1], hi[0], hi[1]);
            } else if (nc == 1) {
                mx_vertex_rect(em, X, on[0] ? 0 : on[1] ? 1 : on[2] ? 2 : 3, lo, hi, P.corner_angle);
            } else if (nc == 0) {
                if (mx_near_rects(em, X, S, P.near_ratio, P.near_depth + 2, lo[0], lo[1], hi[0], hi[1], false)) return -1;
            } else {
                em.rect(MR_SSELF, lo[0], lo[1], hi[0], hi[1]);
            }
        }
    return 1;
}

__device__ inline bool mx_is_touch(const MxArgs &P, int a, int b)
{
    int lo = 0, hi = P.ntouch - 1;
    while (lo <= hi) {
        const int m = (lo + hi) >> 1;
        const int2 t = P.touch[m];
        if (t.x == a && t.y == b) return true;
        if (t.x < a || (t.x == a && t.y < b)) lo = m + 1;
        else hi = m - 1;
    }
    return false;
}

// The tier and rule of a non-far pair (e, s) (quad_pair_rule, tri_pair_rule, quad_wake_rule, quad_segment_rule,
// src/assembly/mixed.jl:238-369; pair_rule / wake_pair_rule / segment_rule, assemble.jl:266-306, for triangle tests
// with wake sources).  Emits the rule's pieces; returns the tier, or -1 on a rule-builder stack overflow.
__device__ int mx_pair_rule(Emitter &em, const MxArgs &P, int e, int s)
{
#pragma clang fp contract(off)
    const bool tq = e < P.nq;
    GSrc S;
    int sn[4] = {-1, -1, -1, -1}, nsn = 0;
    bool is_seg = false;
    const int swake = s - P.ne, sfil = s - P.ne - P.nwt;
    if (s < P.nq) {
        S = gsrc_quad(P.quad[s]);
        for (int k = 0; k < 4; k++) sn[k] = P.qnodes[4 * s + k];
        nsn = 4;
    } else if (s < P.ne) {
        S = gsrc_tri(P.tri[s - P.nq]);
        for (int k = 0; k < 3; k++) sn[k] = P.tnodes[3 * (s - P.nq) + k];
        nsn = 3;
    } else if (swake < P.nwt) {
        S = gsrc_tri(P.wake[swake]);  // shares no node with the body
    } else {
        is_seg = true;
        S.kind = 2;
        S.nc = 2;
        S.x[0] = ld3(P.wx + 3 * P.fil[2 * sfil]);
        S.x[1] = ld3(P.wx + 3 * P.fil[2 * sfil + 1]);
    }
    const bool body = s < P.ne;
    if (tq) {
        const QuadRec &T = P.quad[e];
        V3<double> X[4];
        for (int k = 0; k < 4; k++) X[k] = mq_corner(T, k);
        if (s == e) {
            em.rect(MR_SSELF, 0.0, 0.0, 1.0, 1.0);
            return MT_SELF;
        }
        if (is_seg) {  // quad_segment_rule
            for (int k = 0; k < 4; k++)
                if (fmin(cls_norm(cls_sub(X[k], S.x[0])), cls_norm(cls_sub(X[k], S.x[1]))) <= 1e-12 * T.diam) {
                    mx_vertex_square(em, X, k, P);
                    return MT_VERTEX;
                }
            if (mx_near_rects(em, X, S, P.near_ratio, P.near_depth, 0.0, 0.0, 1.0, 1.0, true)) return -1;
            return MT_NEAR;
        }
        int shared[4] = {0, 0, 0, 0};
        if (body)
            for (int k = 0; k < 4; k++) {
                const int v = P.qnodes[4 * e + k];
                for (int j = 0; j < nsn; j++) shared[k] |= v == sn[j];
            }
        const int c = mx_square_contact(em, T, X, S, shared, P);
        if (c < 0) return -1;
        if (c > 0) return MT_CONTACT;
        if (body) {
            const int ns = shared[0] + shared[1] + shared[2] + shared[3];
            if (ns == 2) {
                for (int k = 0; k < 4; k++)
                    if (shared[k] && shared[(k + 1) % 4]) {
                        em.rect(MR_SEDGE0 + k, 0.0, 0.0, 1.0, 1.0);
                        return MT_EDGE;
                    }
                em.rect(MR_SSELF, 0.0, 0.0, 1.0, 1.0);
                return MT_EDGE;
            } else if (ns == 1) {
                mx_vertex_square(em, X, shared[0] ? 0 : shared[1] ? 1 : shared[2] ? 2 : 3, P);
                return MT_VERTEX;
            } else if (ns > 2) {
                em.rect(MR_SSELF, 0.0, 0.0, 1.0, 1.0);
                return MT_EDGE;
            }
        }
        const int depth = body && mx_is_touch(P, e, s) ? P.near_depth + 4 : P.near_depth;
        if (mx_near_rects(em, X, S, P.near_ratio, depth, 0.0, 0.0, 1.0, 1.0, true)) return -1;
        return MT_NEAR;
    }
    // a test triangle
    const int t = e - P.nq;
    const Elem &T = P.tri[t];
    V3<double> TX[3];
    for (int k = 0; k < 3; k++) TX[k] = ld3(T.v[k]);
    if (s == e) {
        double id[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        em.btri(MR_TSELF, id, 1.0);
        return MT_SELF;
    }
    if (is_seg) {  // segment_rule, assemble.jl:299-306
        for (int k = 0; k < 3; k++)
            if (fmin(cls_norm(cls_sub(TX[k], S.x[0])), cls_norm(cls_sub(TX[k], S.x[1]))) <= 1e-12 * T.diam) {
                mx_vertex_rule(em, TX, k, P);
                return MT_VERTEX;
            }
        if (mx_tri_pieces(em, TX, S, P, 0.0, 1, P.near_depth)) return -1;
        return MT_NEAR;
    }
    int shared[3] = {0, 0, 0};
    if (body)
        for (int k = 0; k < 3; k++) {
            const int v = P.tnodes[3 * t + k];
            for (int j = 0; j < nsn; j++) shared[k] |= v == sn[j];
        }
    // contact_rule, assemble.jl:199-207
    const double tol = P.contact_rel * fmin(T.diam, S.diam);
    if (!(cls_norm(cls_sub(S.c, ld3(T.c))) > T.reach + S.reach + tol)) {
        bool extra = false;
        for (int k = 0; k < 3 && !extra; k++) extra = !shared[k] && gsrc_distance(S, TX[k]) <= tol;
        for (int k = 0; k < 3 && !extra; k++) {
            double sp[4];
            extra = mx_edge_splits(TX, k, S, tol, sp) > 0;
        }
        if (extra) {
            if (mx_tri_pieces(em, TX, S, P, tol, 0, P.near_depth)) return -1;
            return MT_CONTACT;
        }
    }
    const int ns = shared[0] + shared[1] + shared[2];
    if (ns == 2) {
        double id[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        em.btri(MR_TEDGE0 + (shared[0] && shared[1] ? 0 : shared[1] && shared[2] ? 1 : 2), id, 1.0);
        return MT_EDGE;
    } else if (ns == 1) {
        mx_vertex_rule(em, TX, shared[0] ? 0 : shared[1] ? 1 : 2, P);
        return MT_VERTEX;
    }
    const int depth = body && mx_is_touch(P, e, s) ? P.near_depth + 4 : P.near_depth;
    if (mx_tri_pieces(em, TX, S, P, 0.0, 1, depth)) return -1;
    return MT_NEAR;
}

// ------------------------------------------------------------------------------------ inner integrals ---------

__device__ __forceinline__ int gofs(int n) { return n * (n - 1) / 2; }

// kernel_terms, src/elements/quad_field.jl:17-27 (without 1/4pi), scaled by w and added to acc[0..4]
__device__ __forceinline__ void kt_terms(const V3<double> *X, V3<double> x, V3<double> nx, V3<double> V, double xi,
                                         double eta, double *c)
{
    const double oe = 1 - eta, ox = 1 - xi;
    const V3<double> y = ((ox * oe) * X[0] + (xi * oe) * X[1]) + ((xi * eta) * X[2] + (ox * eta) * X[3]);
    const V3<double> a1 = oe * (X[1] - X[0]) + eta * (X[2] - X[3]);
    const V3<double> a2 = ox * (X[3] - X[0]) + xi * (X[2] - X[1]);
    const V3<double> r = x - y;
    const double ir = vrsqrt(dot(r, r));
    const double ir3 = ir * ir * ir;
    const V3<double> t = ir3 * cross(r, nx);
    const double ta2 = dot(t, a2), ta1 = dot(t, a1);
    c[0] = -oe * ta2 + ox * ta1;
    c[1] = oe * ta2 + xi * ta1;
    c[2] = eta * ta2 - xi * ta1;
    c[3] = -eta * ta2 - ox * ta1;
    c[4] = -dot(V, cross(a1, a2)) * dot(nx, r) * ir3;
}

__device__ __forceinline__ int gauss_order(double r) { return r >= 8 ? 3 : r >= 4 ? 4 : r >= 2 ? 6 : 8; }

// gauss_field, src/elements/quad_field.jl:48-59 (acc without 1/4pi)
__device__ void gauss_field(const V3<double> *X, V3<double> x, V3<double> nx, V3<double> V, int m, const double *gx,
                            const double *gw, double *acc)
{
    const int o = gofs(m);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            double dc[5];
            kt_terms(X, x, nx, V, gx[o + i], gx[o + j], dc);
            const double w = gw[o + i] * gw[o + j];
            for (int b = 0; b < 5; b++) acc[b] += w * dc[b];
        }
}

// pieces, src/elements/quad_field.jl:66-69
__device__ __forceinline__ int qf_pieces(double a, double b, double brk, int base, double slope, double *lo, double *hi,
                                         int *n)
{
    int np;
    if (a < brk && brk < b) {
        lo[0] = a; hi[0] = brk; lo[1] = brk; hi[1] = b;
        np = 2;
    } else {
        lo[0] = a; hi[0] = b;
        np = 1;
    }
    for (int i = 0; i < np; i++) {
        const int k = base + (int)ceil(slope * (hi[i] - lo[i]));
        n[i] = k < 1 ? 1 : k < 48 ? k : 48;
    }
    return np;
}

// The integrand of polar_field along one ray p = P + rho e from the closest point P.  The bilinear surface makes
// the source point quadratic in rho and the area normal and the tangents linear in it:
//   x(p) = x(P) + rho D + rho^2 E  (D = e1 a1(P) + e2 a2(P), E = e1 e2 W, W = x1 - x2 + x3 - x4),
//   a1(p) = a1(P) + rho e2 W,  a2(p) = a2(P) + rho e1 W,  a1 x a2 (p) = a1 x a2 (P) + rho (e1 a1(P) x W + e2 W x a2(P)),
// so with r = r0 - rho (D + rho E), r0 = x - x(P), the kernel_terms of src/elements/quad_field.jl:17-27 become
//   t.a2 = |r|^-3 r.(nx x a2(p)), t.a1 = |r|^-3 r.(nx x a1(p)),  s = -V.(a1 x a2)(p) nx.r |r|^-3
// (the same integrand, about half the arithmetic of evaluating the bilinear map at every point).
struct PolarRay {
    V3<double> r0, D, E, Q1, Q2, Wn, nx;
    double P0, P1, e0, e1, vn0, vn1;
};

__device__ __forceinline__ void ray_terms(const PolarRay &R, double rho, double *c)
{
    const V3<double> r = R.r0 - rho * (R.D + rho * R.E);
    const double ir = vrsqrt(dot(r, r));
    const double ir3 = ir * ir * ir;
    const double rw = dot(r, R.Wn);
    const double ta2 = ir3 * (dot(r, R.Q2) + (rho * R.e0) * rw);
    const double ta1 = ir3 * (dot(r, R.Q1) + (rho * R.e1) * rw);
    const double xi = R.P0 + rho * R.e0, eta = R.P1 + rho * R.e1;
    const double oe = 1 - eta, ox = 1 - xi;
    c[0] = -oe * ta2 + ox * ta1;
    c[1] = oe * ta2 + xi * ta1;
    c[2] = eta * ta2 - xi * ta1;
    c[3] = -eta * ta2 - ox * ta1;
    c[4] = -(R.vn0 + rho * R.vn1) * dot(R.nx, r) * ir3;
}

// polar_field, src/elements/quad_field.jl:103-154 with ray_pv (:160-171) and ray_near (:174-192); out[0..4] with
// 1/4pi.  Returns false for a target on the quad's boundary.
__device__ bool polar_field(const V3<double> *X, V3<double> x, V3<double> nx, V3<double> V, double P0, double P1,
                            double d, double diam, int side, const double *gx, const double *gw, double *out)
{
    const bool onq = d <= 1e-12 * diam;
    if (onq && fmin(fmin(P0, P1), fmin(1 - P0, 1 - P1)) <= 1e-12) return false;
    const V3<double> W = (X[0] - X[1]) + (X[2] - X[3]);
    const V3<double> a1s = (X[1] - X[0]) + P1 * W, a2s = (X[3] - X[0]) + P0 * W;
    const double g11 = dot(a1s, a1s), g12 = dot(a1s, a2s), g22 = dot(a2s, a2s);
    const double l11 = sqrt(g11);
    const double l12 = g12 / l11, l22 = sqrt(g22 - g12 * g12 / g11);
    const double J = l11 * l22;
    const V3<double> nJs = cross(a1s, a2s);
    PolarRay ray;
    ray.P0 = P0;
    ray.P1 = P1;
    ray.nx = nx;
    ray.r0 = x - (X[0] + P0 * (X[1] - X[0]) + P1 * (X[3] - X[0]) + (P0 * P1) * W);
    ray.Q1 = cross(nx, a1s);
    ray.Q2 = cross(nx, a2s);
    ray.Wn = cross(nx, W);
    ray.vn0 = dot(V, nJs);
    const double va = dot(V, cross(a1s, W)), vb = dot(V, cross(W, a2s));
    const double sJ0 = -ray.vn0;
    // f_b = -nx.(g0_b x A)/|A|^3 with g0_b = N_b,xi a2 - N_b,eta a1 (the sheet densities at P): nx.(g0_b x A) =
    // N_b,xi nx.(a2 x A) - N_b,eta nx.(a1 x A)
    const double Nx[4] = {-(1 - P1), 1 - P1, P1, -P1}, Ne[4] = {-(1 - P0), -P0, P0, 1 - P0};
    double acc[5] = {0, 0, 0, 0, 0};
    const double SQ[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {1.0, 1.0}, {0.0, 1.0}};
    for (int k = 0; k < 4; k++) {
        const double pa0 = SQ[k][0] - P0, pa1 = SQ[k][1] - P1, pb0 = SQ[(k + 1) % 4][0] - P0, pb1 = SQ[(k + 1) % 4][1] - P1;
        const double A0 = l11 * pa0 + l12 * pa1, A1 = l22 * pa1, B0 = l11 * pb0 + l12 * pb1, B1 = l22 * pb1;
        const double BA0 = B0 - A0, BA1 = B1 - A1;
        const double L = sqrt(BA0 * BA0 + BA1 * BA1);
        const double et0 = BA0 / L, et1 = BA1 / L;
        const double nk0 = et1, nk1 = -et0;
        const double h = A0 * nk0 + A1 * nk1;
        if (h <= 1e-14 * L) continue;
        double lo[2], hi[2];
        int nv[2];
        const int np = qf_pieces(asinh((A0 * et0 + A1 * et1) / h), asinh((B0 * et0 + B1 * et1) / h), 0.0, 5, 1.5, lo, hi, nv);
        for (int pp = 0; pp < np; pp++) {
            const double va_ = lo[pp], vb_ = hi[pp];
            const int n = nv[pp], o = gofs(n);
            for (int i = 0; i < n; i++) {
                const double v = va_ + (vb_ - va_) * gx[o + i], wq = gw[o + i];
                const double ch = cosh(v), sh = sinh(v);
                const double R = h * ch;
                const double z0 = (nk0 + sh * et0) / ch, z1 = (nk1 + sh * et1) / ch;
                const double e0 = (z0 - l12 * z1 / l22) / l11, e1 = z1 / l22;
                ray.e0 = e0;
                ray.e1 = e1;
                ray.D = e0 * a1s + e1 * a2s;
                ray.E = (e0 * e1) * W;
                ray.vn1 = e0 * va + e1 * vb;
                const double nA = sqrt(dot(ray.D, ray.D));
                const double W_ = wq * (vb_ - va_) / (ch * J);
                double dc[5] = {0, 0, 0, 0, 0};
                if (onq) {  // ray_pv with the 1/rho terms added back analytically
                    const double nA3 = nA * nA * nA;
                    const double u2 = dot(nx, cross(a2s, ray.D)), u1 = dot(nx, cross(a1s, ray.D));
                    double f[5];
                    for (int b = 0; b < 4; b++) f[b] = -(Nx[b] * u2 - Ne[b] * u1) / nA3;
                    f[4] = -sJ0 * dot(nx, ray.D) / nA3;
                    const int o12 = gofs(12);
                    for (int q = 0; q < 12; q++) {
                        const double rho = R * gx[o12 + q], wr = gw[o12 + q] * R;
                        double kt[5];
                        ray_terms(ray, rho, kt);
                        for (int b = 0; b < 5; b++) dc[b] += wr * (rho * kt[b] - f[b] / rho);
                    }
                    const double lg = log(R * nA);
                    for (int b = 0; b < 5; b++) dc[b] += f[b] * lg;
                } else {
                    const double rho0 = d / nA;
                    if (rho0 < R) {
                        const double sb = asinh(R / rho0);
                        const int nr0 = 6 + (int)ceil(1.0 * (sb - 0.0));
                        const int nr = nr0 < 1 ? 1 : nr0 < 48 ? nr0 : 48, orr = gofs(nr);
                        for (int q = 0; q < nr; q++) {
                            const double sv = sb * gx[orr + q];
                            const double rho = rho0 * sinh(sv);
                            const double jac = gw[orr + q] * sb * rho0 * cosh(sv) * rho;
                            double kt[5];
                            ray_terms(ray, rho, kt);
                            for (int b = 0; b < 5; b++) dc[b] += jac * kt[b];
                        }
                    } else {
                        const int o12 = gofs(12);
                        for (int q = 0; q < 12; q++) {
                            const double rho = R * gx[o12 + q];
                            const double wr = gw[o12 + q] * R * rho;
                            double kt[5];
                            ray_terms(ray, rho, kt);
                            for (int b = 0; b < 5; b++) dc[b] += wr * kt[b];
                        }
                    }
                }
                for (int b = 0; b < 5; b++) acc[b] += W_ * dc[b];
            }
        }
    }
    const double i4 = 1.0 / (4 * kPi);
    for (int b = 0; b < 5; b++) out[b] = i4 * acc[b];
    if (onq && side != 0) {  // half the jumps: (g0_b x nP) for the sheets, sigma nP for the source
        const V3<double> nP = (1.0 / sqrt(dot(nJs, nJs))) * nJs;
        const double hs = side / 2.0;
        const double w2 = dot(nx, cross(a2s, nP)), w1 = dot(nx, cross(a1s, nP));
        for (int b = 0; b < 4; b++) out[b] += hs * ((Nx[b] * w2 - Ne[b] * w1) / J);
        out[4] += side * (-dot(V, nP)) / 2 * dot(nx, nP);
    }
    return true;
}

// quad_field, src/elements/quad_field.jl:40-46 (polar = 1); out[0..4] = c[4], s.  Returns false on a boundary target.
__device__ bool quad_field(const QuadRec &Q, const V3<double> *X, V3<double> x, V3<double> nx, V3<double> V,
                           const double *gx, const double *gw, double *out)
{
    double dc;
    {
#pragma clang fp contract(off)
        dc = cls_norm(cls_sub(x, ld3(Q.c))) - Q.reach;
    }
    if (dc >= Q.diam) {
        double acc[5] = {0, 0, 0, 0, 0};
        gauss_field(X, x, nx, V, gauss_order(dc / Q.diam), gx, gw, acc);
        const double i4 = 1.0 / (4 * kPi);
        for (int b = 0; b < 5; b++) out[b] = i4 * acc[b];
        return true;
    }
    double P[2];
    const double d = mq_project(X, ld3(Q.warp), x, P);
    if (d >= Q.diam) {
        double acc[5] = {0, 0, 0, 0, 0};
        gauss_field(X, x, nx, V, gauss_order(d / Q.diam), gx, gw, acc);
        const double i4 = 1.0 / (4 * kPi);
        for (int b = 0; b < 5; b++) out[b] = i4 * acc[b];
        return true;
    }
    return polar_field(X, x, nx, V, P[0], P[1], d, Q.diam, 0, gx, gw, out);
}

// source_velocity of a flat triangle (src/kernels/source_triangle.jl:134-138), FP64: the multipole beyond
// kernel_far diameters, else the closed form with the in-plane limit on side `side`
__device__ V3<double> tri_velocity(const Elem &S, V3<double> x, int side, double kernel_far)
{
    const V3<double> SX[3] = {ld3(S.v[0]), ld3(S.v[1]), ld3(S.v[2])};
    const V3<double> sc = ld3(S.c);
    const V3<double> r = x - sc;
    const double fd = kernel_far * S.diam;
    SrcT<double> sr;
    if (dot(r, r) > fd * fd) {
        for (int k = 0; k < 3; k++) sr.v[k] = SX[k] - sc;
        sr.area = S.area;
        return sv_far_rel(sr, r);
    }
    for (int k = 0; k < 3; k++) {
        sr.nu[k] = ld3(S.nu[k]);
        sr.len[k] = S.len[k];
    }
    sr.n = ld3(S.n);
    sr.diam = S.diam;
    return sv_exact_abs(sr, x - SX[0], x - SX[1], x - SX[2], dot(r, sr.n), side, side ? &S : nullptr, x);
}

// square_test_weights, src/assembly/mixed.jl:62-67 and test_weights, assemble.jl:52-58
__device__ __forceinline__ void mx_square_weights(int test, double xi, double eta, double *psi)
{
    if (test == 0) {
        psi[0] = (1 - xi) * (1 - eta);
        psi[1] = xi * (1 - eta);
        psi[2] = xi * eta;
        psi[3] = (1 - xi) * eta;
    } else if (test == 1) {
        psi[0] = psi[1] = psi[2] = psi[3] = 1.0;
    } else {
        const int k = xi < 0.5 ? (eta < 0.5 ? 0 : 3) : (eta < 0.5 ? 1 : 2);
        for (int i = 0; i < 4; i++) psi[i] = i == k ? 1.0 : 0.0;
    }
}

__device__ __forceinline__ void mx_tri_weights(int test, const double *l, double *psi)
{
    if (test == 0) {
        psi[0] = l[0]; psi[1] = l[1]; psi[2] = l[2];
    } else if (test == 1) {
        psi[0] = psi[1] = psi[2] = 1.0;
    } else {
        const double m = fmax(fmax(l[0], l[1]), l[2]);
        int is[3], c = 0;
        for (int k = 0; k < 3; k++) {
            is[k] = fabs(l[k] - m) < 1e-12;
            c += is[k];
        }
        for (int k = 0; k < 3; k++) psi[k] = (double)is[k] / c;
    }
    psi[3] = 0.0;
}

// One test point's block entries for pair (e, s): ent[4 a + b] += psi_a c_b, ent[16 + a] -= s psi_a (the source part
// of the right-hand side, body sources only).  Test point: a quad at parameters (xi, eta) with rule weight w (times
// the area element J), or a triangle at barycentric l with weight w (times its area).  Returns false on a boundary
// target (the pair block's quad_field raises there).
// KIND 1: the source is a quad; 0: a triangle or filament; -1: either (the far tier)
template <int KIND>
__device__ bool mx_point(const MxArgs &P, int e, int s, bool tquad, double xi, double eta, const double *l, double w,
                         double *ent)
{
    V3<double> x, nx;
    double psi[4];
    if (tquad) {
        const QuadRec &T = P.quad[e];
        V3<double> TX[4];
        for (int k = 0; k < 4; k++) TX[k] = mq_corner(T, k);
        const double oe = 1 - eta, ox = 1 - xi;
        x = ((ox * oe) * TX[0] + (xi * oe) * TX[1]) + ((xi * eta) * TX[2] + (ox * eta) * TX[3]);
        const V3<double> a1 = oe * (TX[1] - TX[0]) + eta * (TX[2] - TX[3]);
        const V3<double> a2 = ox * (TX[3] - TX[0]) + xi * (TX[2] - TX[1]);
        const V3<double> nJ = cross(a1, a2);
        const double J = sqrt(dot(nJ, nJ));
        nx = (1.0 / J) * nJ;
        mx_square_weights(P.test, xi, eta, psi);
        for (int a = 0; a < 4; a++) psi[a] *= w * J;
    } else {
        const Elem &T = P.tri[e - P.nq];
        x = (l[0] * ld3(T.v[0]) + l[1] * ld3(T.v[1])) + l[2] * ld3(T.v[2]);
        nx = ld3(T.n);
        mx_tri_weights(P.test, l, psi);
        for (int a = 0; a < 3; a++) psi[a] *= w * T.area;
    }
    const int na = tquad ? 4 : 3;
    const V3<double> V = mk3(P.V[0], P.V[1], P.V[2]);
    double c[4] = {0, 0, 0, 0}, src = 0.0;
    int nb;
    if (KIND == 1 || (KIND < 0 && s < P.nq)) {  // a quad source
        const QuadRec &S = P.quad[s];
        V3<double> SX[4];
        for (int k = 0; k < 4; k++) SX[k] = mq_corner(S, k);
        double out[5];
        bool ok;
        if (s == e) ok = polar_field(SX, x, nx, V, xi, eta, 0.0, S.diam, 1, P.gx, P.gw, out);  // the self pair: P known
        else ok = quad_field(S, SX, x, nx, V, P.gx, P.gw, out);
        if (!ok) return false;
        for (int b = 0; b < 4; b++) c[b] = out[b];
        src = out[4];
        nb = 4;
    } else if (s < P.ne + P.nwt) {  // a triangle: body or wake
        const bool wk = s >= P.ne;
        const Elem &S = wk ? P.wake[s - P.ne] : P.tri[s - P.nq];
        const V3<double> wv = tri_velocity(S, x, s == e ? 1 : 0, P.kernel_far);
        const V3<double> v = cross(wv, nx);
        for (int b = 0; b < 3; b++) c[b] = dot(v, ld3(S.g[b]));
        src = wk ? 0.0 : -dot(V, ld3(S.n)) * dot(wv, nx);
        nb = 3;
    } else {  // a filament
        const int f = s - P.ne - P.nwt;
        const V3<double> fa = ld3(P.wx + 3 * P.fil[2 * f]), fb = ld3(P.wx + 3 * P.fil[2 * f + 1]);
        V3<double> va, vb;
        seg_vel(x - fa, x - fb, va, vb);
        c[0] = dot(va, nx);
        c[1] = dot(vb, nx);
        nb = 2;
    }
    (void)na;
    (void)nb;
    // all sixteen products with constant indices (the unused rows and columns are zero: psi[3] for a triangle test,
    // c[b] past the source's corners), so that ent stays in registers
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int b = 0; b < 4; b++) ent[4 * a + b] += psi[a] * c[b];
        ent[16 + a] -= src * psi[a];
    }
    return true;
}

// Adds a pair's entries to A and b: rows are the test element's slots, columns the source's slots (body), wake node
// columns (nslot + id) or a filament's two end columns
__device__ void mx_scatter(const MxArgs &P, int e, int s, const double *ent)
{
    const bool tquad = e < P.nq;
    const int na = tquad ? 4 : 3;
    const int *rows = tquad ? P.qslots + 4 * e : P.tslots + 3 * (e - P.nq);
    int cols[4], nb;
    bool rhs = true;
    if (s < P.nq) {
        for (int b = 0; b < 4; b++) cols[b] = P.qslots[4 * s + b];
        nb = 4;
    } else if (s < P.ne) {
        for (int b = 0; b < 3; b++) cols[b] = P.tslots[3 * (s - P.nq) + b];
        nb = 3;
    } else if (s < P.ne + P.nwt) {
        for (int b = 0; b < 3; b++) cols[b] = P.nslot + P.wtris[3 * (s - P.ne) + b];
        nb = 3;
        rhs = false;
    } else {
        const int f = s - P.ne - P.nwt;
        cols[0] = P.nslot + P.fil[2 * f];
        cols[1] = P.nslot + P.fil[2 * f + 1];
        nb = 2;
        rhs = false;
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
        if (a >= na) break;
        const long long r = rows[a];
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (b < nb) unsafeAtomicAdd(&P.A[(long long)cols[b] * P.lda + r], ent[4 * a + b]);
        if (rhs) unsafeAtomicAdd(&P.b[r], ent[16 + a]);
    }
}

__device__ __forceinline__ bool mx_far_test(const MxArgs &P, int e, int s)
{
#pragma clang fp contract(off)
    V3<double> tc;
    double td;
    if (e < P.nq) {
        tc = ld3(P.quad[e].c);
        td = P.quad[e].diam;
    } else {
        tc = ld3(P.tri[e - P.nq].c);
        td = P.tri[e - P.nq].diam;
    }
    if (s == e) return false;
    V3<double> sc;
    double sd;
    if (s < P.nq) {
        sc = ld3(P.quad[s].c);
        sd = P.quad[s].diam;
    } else if (s < P.ne) {
        sc = ld3(P.tri[s - P.nq].c);
        sd = P.tri[s - P.nq].diam;
    } else if (s < P.ne + P.nwt) {
        sc = ld3(P.wake[s - P.ne].c);
        sd = P.wake[s - P.ne].diam;
    } else {  // quad_segment_rule / segment_rule: the segment's midpoint and length
        const int f = s - P.ne - P.nwt;
        const V3<double> fa = ld3(P.wx + 3 * P.fil[2 * f]), fb = ld3(P.wx + 3 * P.fil[2 * f + 1]);
        const V3<double> m = cls_add_scaled(fa, 1.0, fb);
        sc = mk3(m.x / 2.0, m.y / 2.0, m.z / 2.0);
        sd = cls_norm(cls_sub(fb, fa));
    }
    return cls_far(cls_sub(sc, tc), P.far_ratio * fmax(td, sd));
}

// ------------------------------------------------------------------------------------------------ kernels ---

// IsoQuad, src/elements/iso_quad.jl:24-37 (contraction off)
__global__ void mixed_quad_setup(const double *__restrict__ x, const int *__restrict__ qn, int nq, const double *gx,
                                 const double *gw, QuadRec *__restrict__ out)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nq) return;
    V3<double> X[4];
    for (int k = 0; k < 4; k++) X[k] = ld3(x + 3 * qn[4 * e + k]);
    const V3<double> a = X[0], b = X[1], c = X[2], d = X[3];
    const V3<double> nc = cls_cross(cls_sub(c, a), cls_sub(d, b));
    const double nn = cls_norm(nc);
    QuadRec Q;
    Q.bad = nn > 0 ? 0.0 : 1.0;
    const int o = gofs(4);
    double area = 0.0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const double xi = gx[o + i], eta = gx[o + j];
            const V3<double> ba = cls_sub(b, a), cd = cls_sub(c, d), da = cls_sub(d, a), cb = cls_sub(c, b);
            const V3<double> a1 = mk3((1 - eta) * ba.x + eta * cd.x, (1 - eta) * ba.y + eta * cd.y, (1 - eta) * ba.z + eta * cd.z);
            const V3<double> a2 = mk3((1 - xi) * da.x + xi * cb.x, (1 - xi) * da.y + xi * cb.y, (1 - xi) * da.z + xi * cb.z);
            area += gw[o + i] * gw[o + j] * cls_norm(cls_cross(a1, a2));
        }
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 3; j++) Q.x[k][j] = (&X[k].x)[j];
    double dm = cls_norm(cls_sub(b, a));
    dm = fmax(dm, cls_norm(cls_sub(c, b)));
    dm = fmax(dm, cls_norm(cls_sub(d, c)));
    dm = fmax(dm, cls_norm(cls_sub(a, d)));
    dm = fmax(dm, cls_norm(cls_sub(c, a)));
    dm = fmax(dm, cls_norm(cls_sub(d, b)));
    const V3<double> s4 = cls_add_scaled(cls_add_scaled(cls_add_scaled(a, 1.0, b), 1.0, c), 1.0, d);
    const V3<double> cc = mk3(s4.x / 4.0, s4.y / 4.0, s4.z / 4.0);
    const V3<double> wp = cls_sub(cls_add_scaled(cls_sub(a, b), 1.0, c), d);
    for (int j = 0; j < 3; j++) {
        Q.c[j] = (&cc.x)[j];
        Q.warp[j] = (&wp.x)[j];
        Q.n[j] = nn > 0 ? (&nc.x)[j] / nn : 0.0;
    }
    Q.area = area;
    Q.diam = dm;
    double r = 0.0;
    for (int k = 0; k < 4; k++) r = fmax(r, cls_norm(cls_sub(X[k], cc)));
    Q.reach = r;
    for (int k = 0; k < 7; k++) Q.pad[k] = 0.0;
    out[e] = Q;
}

// Non-far pairs per test element row (the list's size before the far pass)
__global__ __launch_bounds__(256) void mixed_nonfar_count(MxArgs P, unsigned long long *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    for (int e = blockIdx.y; e < P.ne; e += gridDim.y) {
        const bool nf = s < P.nsrc && !mx_far_test(P, e, s);
        const unsigned long long m = __ballot(nf);
        if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
    }
}

// One test element's row of the far tier: thread per pair; far pairs evaluated with the far rule and added, the
// others appended to `list`.  stats[0], stats[1]: far pairs and far points.
__device__ __forceinline__ void mixed_far_row(const MxArgs &P, int e, int2 *list, unsigned long long *count,
                                              unsigned long long cap, unsigned long long *stats)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < P.nsrc;
    const bool far = valid && mx_far_test(P, e, s);
    const bool nf = valid && !far;
    // append the non-far pairs, one atomic per wave
    const unsigned long long m = __ballot(nf);
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == 0 && m) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl(base, 0);
    if (nf) {
        const unsigned long long k = base + __popcll(m & ((1ull << lane) - 1));
        if (k < cap) list[k] = make_int2(e, s);
    }
    const bool tquad = e < P.nq;
    const int rule = tquad ? MR_SFAR : MR_TFAR;
    const unsigned long long fm = __ballot(far);
    if (lane == 0 && fm) {
        atomicAdd(&stats[0], (unsigned long long)__popcll(fm));
        atomicAdd(&stats[1], (unsigned long long)__popcll(fm) * P.rn[rule]);
    }
    if (P.tiers && far && s < P.ne + P.nwt) P.tiers[(long long)e * (P.ne + P.nwt) + s] = MT_FAR;
    if (!far) return;
    double ent[20];
    for (int i = 0; i < 20; i++) ent[i] = 0.0;
    const int o = P.roff[rule];
    bool ok = true;
    for (int q = 0; q < P.rn[rule] && ok; q++) {
        const double *l = P.rl + 3 * (o + q);
        ok = mx_point<-1>(P, e, s, tquad, l[0], l[1], l, P.rw[o + q], ent);
    }
    if (!ok) {
        atomicExch(P.err, 1);
        return;
    }
    mx_scatter(P, e, s, ent);
}

__global__ __launch_bounds__(256) void mixed_far(MxArgs P, int2 *list, unsigned long long *count, unsigned long long cap,
                                                unsigned long long *stats)
{
    for (int e = blockIdx.y; e < P.ne; e += gridDim.y) mixed_far_row(P, e, list, count, cap, stats);
}

// The non-far pairs' rules.  PASS 0: per pair its tier, piece count and point count; PASS 1: the pieces at
// piece_off[pair] (points from point_off[pair]).
template <int PASS>
__global__ __launch_bounds__(64) void mixed_rules(MxArgs P, const int2 *__restrict__ list, int npairs,
                                                  int *__restrict__ tier, int *__restrict__ npieces,
                                                  long long *__restrict__ npts, const long long *__restrict__ piece_off,
                                                  const long long *__restrict__ point_off, MxPiece *__restrict__ pieces)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const int2 pr = list[i];
    Emitter em;
    em.out = PASS ? pieces + piece_off[i] : nullptr;
    em.npieces = 0;
    em.npts = 0;
    em.pt0 = PASS ? point_off[i] : 0;
    em.pair = i;
    em.rn = P.rn;
    const int t = mx_pair_rule(em, P, pr.x, pr.y);
    if (PASS == 0) {
        tier[i] = t < 0 ? t : t | (pr.y < P.nq ? 8 : 0);  // bit 3: a quad source (its points go to mixed_eval<1>)
        npieces[i] = em.npieces;
        npts[i] = em.npts;
        if (t < 0) atomicExch(P.err, 2);
        if (P.tiers && t >= 0 && pr.y < P.ne + P.nwt) P.tiers[(long long)pr.x * (P.ne + P.nwt) + pr.y] = (uint8_t)t;
    }
}

// Thread per rule point of the non-far pairs whose sources are quads (KIND 1: the numerical inner integrals) or
// triangles and filaments (KIND 0: closed forms), points pt_base .. pt_base + ntotal - 1 and their pieces; a wave
// whose lanes share a pair sums its entries by shuffles
template <int KIND, int WPS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPS, WPS == 1 ? 8 : WPS))) void mixed_eval(MxArgs P, const int2 *__restrict__ list, const MxPiece *__restrict__ pieces,
                                                  long long npieces, long long pt_base, long long ntotal)
{
    const long long g = pt_base + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = g < pt_base + ntotal;
    int pair = -1, e = 0, s = 0;
    double ent[20];
    for (int i = 0; i < 20; i++) ent[i] = 0.0;
    if (valid) {
        long long lo = 0, hi = npieces - 1;  // the last piece whose first point is <= g
        while (lo < hi) {
            const long long m = (lo + hi + 1) >> 1;
            if (pieces[m].pt0 <= g) lo = m;
            else hi = m - 1;
        }
        const MxPiece &pc = pieces[lo];
        const int q = (int)(g - pc.pt0);
        pair = pc.pair;
        const int2 pr = list[pair];
        e = pr.x;
        s = pr.y;
        const int o = P.roff[pc.rule] + q;
        const double *bl = P.rl + 3 * o;
        const double bw = P.rw[o];
        const bool tquad = e < P.nq;
        bool ok;
        if (pc.kind == 0) {  // append_rect!: lo + d .* q, w d1 d2
            const double d0 = pc.c[2] - pc.c[0], d1 = pc.c[3] - pc.c[1];
            ok = mx_point<KIND>(P, e, s, true, pc.c[0] + d0 * bl[0], pc.c[1] + d1 * bl[1], nullptr, bw * d0 * d1, ent);
        } else if (pc.kind == 1) {  // append_triangle!: l1 a + l2 b + l3 c, w frac
            const double xi = (bl[0] * pc.c[0] + bl[1] * pc.c[2]) + bl[2] * pc.c[4];
            const double eta = (bl[0] * pc.c[1] + bl[1] * pc.c[3]) + bl[2] * pc.c[5];
            ok = mx_point<KIND>(P, e, s, true, xi, eta, nullptr, bw * pc.frac, ent);
        } else {  // append_piece!: sum_i m_i b_i, w frac
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (bl[0] * pc.c[j] + bl[1] * pc.c[3 + j]) + bl[2] * pc.c[6 + j];
            ok = mx_point<KIND>(P, e, s, tquad, 0.0, 0.0, l, bw * pc.frac, ent);
        }
        if (!ok) atomicExch(P.err, 1);
    }
    // one add per wave when every active lane holds the same pair
    const int p0 = __shfl(pair, 0);
    const bool same = !valid || pair == p0;
    const unsigned long long allsame = __ballot(same);
    const unsigned long long act = __ballot(true);
    if (allsame == act && p0 >= 0) {
        const int lane = threadIdx.x & 63;
        for (int i = 0; i < 20; i++) {
            double v = ent[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            ent[i] = v;
        }
        if (lane == 0) {
            const int2 pr = list[p0];
            mx_scatter(P, pr.x, pr.y, ent);
        }
    } else if (valid) {
        mx_scatter(P, e, s, ent);
    }
}

// The freestream part of b: -(V.n) A int psi / A on a triangle (assemble.jl:422), -sum w (V.nJ) psi on the 3 x 3
// Gauss points of a quad (mixed.jl:553-557)
__global__ void mixed_freestream(MxArgs P)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.ne) return;
    const V3<double> V = mk3(P.V[0], P.V[1], P.V[2]);
    if (e < P.nq) {
        const QuadRec &T = P.quad[e];
        V3<double> X[4];
        for (int k = 0; k < 4; k++) X[k] = mq_corner(T, k);
        double rows[4] = {0, 0, 0, 0};
        const int o = gofs(3);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double xi = P.gx[o + i], eta = P.gx[o + j], w = P.gw[o + i] * P.gw[o + j];
                const V3<double> a1 = (1 - eta) * (X[1] - X[0]) + eta * (X[2] - X[3]);
                const V3<double> a2 = (1 - xi) * (X[3] - X[0]) + xi * (X[2] - X[1]);
                double psi[4];
                mx_square_weights(P.test, xi, eta, psi);
                const double f = w * dot(V, cross(a1, a2));
                for (int a = 0; a < 4; a++) rows[a] -= f * psi[a];
            }
        for (int a = 0; a < 4; a++) unsafeAtomicAdd(&P.b[P.qslots[4 * e + a]], rows[a]);
    } else {
        const Elem &T = P.tri[e - P.nq];
        const double ti = P.test == 1 ? 1.0 : 1.0 / 3;
        const double vn = dot(V, ld3(T.n));
        for (int a = 0; a < 3; a++) unsafeAtomicAdd(&P.b[P.tslots[3 * (e - P.nq) + a]], -vn * T.area * ti);
    }
}

// ------------------------------------------------------------------------------------------------- host ---

// SquareTiers (src/elements/quad_rules.jl:41-87, square_tiers of src/assembly/mixed.jl:70-71) and TierRules
// (src/elements/rules.jl:161-174) in one table
struct MixedRules {
    std::vector<double> l, w;
    int off[MR_COUNT], n[MR_COUNT];
};

HostRule square_gauss_h(int m)
{
    std::vector<double> x, w;
    gauss_legendre(m, x, w);
    HostRule r;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) r.push(x[i], x[j], 0.0, w[i] * w[j]);
    return r;
}

HostRule turn_h(const HostRule &r, int k)
{
    HostRule o;
    for (int i = 0; i < r.n(); i++) {
        double p0 = r.l[3 * i], p1 = r.l[3 * i + 1];
        for (int t = 2; t <= k; t++) {
            const double n0 = 1 - p1, n1 = p0;
            p0 = n0;
            p1 = n1;
        }
        o.push(p0, p1, 0.0, r.w[i]);
    }
    return o;
}

void append_triangle_h(HostRule &o, const HostRule &r, const double *a, const double *b, const double *c, double frac)
{
    for (int i = 0; i < r.n(); i++) {
        const double *l = &r.l[3 * i];
        o.push((l[0] * a[0] + l[1] * b[0]) + l[2] * c[0], (l[0] * a[1] + l[1] * b[1]) + l[2] * c[1], 0.0, r.w[i] * frac);
    }
}

MixedRules mixed_rules_h(const vp_assembly_options &o, int ql)
{
    std::vector<HostRule> R(MR_COUNT);
    R[MR_SFAR] = square_gauss_h(o.far_points <= 3 ? 2 : 3);
    R[MR_SNEAR] = square_gauss_h(o.near_points <= 7 ? 3 : 4);
    const double SQ[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {1.0, 1.0}, {0.0, 1.0}}, centre[2] = {0.5, 0.5};
    {
        HostRule base = edge_graded_rule(ql, o.touch_ratio, o.touch_order), self;
        for (int k = 0; k < 4; k++) append_triangle_h(self, base, SQ[k], SQ[(k + 1) % 4], centre, 0.25);
        R[MR_SSELF] = self;
    }
    {
        std::vector<double> es, we, xs, wx;
        graded_gauss(ql, o.touch_ratio, o.touch_order, es, we);
        graded_gauss_both(ql / 2 > 1 ? ql / 2 : 1, o.touch_ratio, o.touch_order, xs, wx);
        HostRule e;
        for (size_t i = 0; i < es.size(); i++)
            for (size_t j = 0; j < xs.size(); j++) e.push(xs[j], es[i], 0.0, we[i] * wx[j]);
        for (int k = 0; k < 4; k++) R[MR_SEDGE0 + k] = turn_h(e, k + 1);
    }
    {
        HostRule base = vertex_graded_rule(ql, o.touch_ratio, o.touch_order), v;
        append_triangle_h(v, base, SQ[0], SQ[1], SQ[2], 0.5);
        append_triangle_h(v, base, SQ[0], SQ[2], SQ[3], 0.5);
        for (int k = 0; k < 4; k++) R[MR_SVERT0 + k] = turn_h(v, k + 1);
        R[MR_VBASE] = base;
    }
    R[MR_TFAR] = triangle_rule(o.far_points);
    R[MR_TNEAR] = triangle_rule(o.near_points);
    R[MR_TSELF] = self_rule(o.touch_levels, o.touch_ratio, o.touch_order);
    {
        HostRule e = edge_graded_rule(o.touch_levels, o.touch_ratio, o.touch_order);
        HostRule v = vertex_graded_rule(o.touch_levels, o.touch_ratio, o.touch_order);
        for (int k = 0; k < 3; k++) {
            R[MR_TEDGE0 + k] = rotate_rule(e, k + 1);
            R[MR_TVERT0 + k] = rotate_rule(v, k + 1);
        }
    }
    MixedRules M;
    for (int k = 0; k < MR_COUNT; k++) {
        M.off[k] = (int)M.w.size();
        M.n[k] = R[k].n();
        M.l.insert(M.l.end(), R[k].l.begin(), R[k].l.end());
        M.w.insert(M.w.end(), R[k].w.begin(), R[k].w.end());
    }
    return M;
}

struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename X> X *alloc(size_t n)
    {
        X *q = nullptr;
        MCHECK(hipMalloc((void **)&q, (n > 0 ? n : 1) * sizeof(X)));
        p.push_back(q);
        return q;
    }
    template <typename X> X *upload(const X *h, size_t n, hipStream_t st)
    {
        X *q = alloc<X>(n);
        if (n) MCHECK(hipMemcpyAsync(q, h, n * sizeof(X), hipMemcpyHostToDevice, st));
        return q;
    }
};

}  // namespace

void mixed_assemble_host(Ctx *c, const vp_mixed_mesh &m, const vp_wake *w, const double *vinf,
                         const vp_assembly_options &o, double *A, double *b, long long *pairs, long long *points,
                         uint8_t *tiers)
{
    if (o.test < 0 || o.test > 2) throw std::invalid_argument("test must be hat (0), onering (1) or dual (2)");
    if (o.quad_touch_levels < 1 || o.touch_levels < 1 || o.touch_order < 1)
        throw std::invalid_argument("touch_levels, quad_touch_levels and touch_order must be positive");
    if (o.near_depth < 0 || o.near_depth + 4 > MAXD)
        throw std::invalid_argument("near_depth must lie in 0.." + std::to_string(MAXD - 4));
    if (!(o.corner_angle >= 3.0)) throw std::invalid_argument("corner_angle must be at least 3 degrees");
    const int nq = m.nquad, nt = m.ntri, ne = nq + nt, nn = m.nn, nslot = m.nslot;
    if (ne <= 0 || nn <= 0 || nslot <= 0) throw std::invalid_argument("empty mixed mesh");
    for (int i = 0; i < 4 * nq; i++) {
        if (m.quad_nodes[i] < 0 || m.quad_nodes[i] >= nn) throw std::out_of_range("quad node index out of range");
        if (m.quad_slots[i] < 0 || m.quad_slots[i] >= nslot) throw std::out_of_range("quad slot out of range");
    }
    for (int i = 0; i < 3 * nt; i++) {
        if (m.tri_nodes[i] < 0 || m.tri_nodes[i] >= nn) throw std::out_of_range("triangle node index out of range");
        if (m.tri_slots[i] < 0 || m.tri_slots[i] >= nslot) throw std::out_of_range("triangle slot out of range");
    }
    const int nw = w ? w->nw : 0, nwt = w ? w->ntri : 0, nwf = w ? w->nfil : 0;
    for (int i = 0; i < 3 * nwt; i++)
        if (w->tris[i] < 0 || w->tris[i] >= nw) throw std::out_of_range("wake triangle index out of range");
    for (int i = 0; i < 2 * nwf; i++)
        if (w->filaments[i] < 0 || w->filaments[i] >= nw) throw std::out_of_range("filament index out of range");
    std::vector<int2> touch;
    for (int i = 0; i < m.ntouch; i++) {
        const int a = m.touch[2 * i], bb = m.touch[2 * i + 1];
        if (a < 0 || a >= ne || bb < 0 || bb >= ne) throw std::out_of_range("touch pair element out of range");
        if (a == bb) continue;
        touch.push_back(make_int2(a, bb));
        touch.push_back(make_int2(bb, a));
    }
    std::sort(touch.begin(), touch.end(), [](int2 p, int2 q) { return p.x < q.x || (p.x == q.x && p.y < q.y); });
    touch.erase(std::unique(touch.begin(), touch.end(), [](int2 p, int2 q) { return p.x == q.x && p.y == q.y; }),
                touch.end());
    const MixedRules R = mixed_rules_h(o, o.quad_touch_levels);
    std::vector<double> gx, gw;
    for (int n = 1; n <= 48; n++) {
        std::vector<double> x, ww;
        gauss_legendre(n, x, ww);
        gx.insert(gx.end(), x.begin(), x.end());
        gw.insert(gw.end(), ww.begin(), ww.end());
    }

    hipStream_t st = c->stream;
    DevBufs D;
    MxArgs P{};
    P.nq = nq;
    P.nt = nt;
    P.ne = ne;
    P.nwt = nwt;
    P.nwf = nwf;
    P.nsrc = ne + nwt + nwf;
    P.nslot = nslot;
    P.lda = nslot;
    P.test = o.test;
    P.near_depth = o.near_depth;
    P.ntouch = (int)touch.size();
    P.far_ratio = o.far_ratio;
    P.near_ratio = o.near_ratio;
    P.corner_angle = o.corner_angle;
    P.contact_rel = o.contact_rel;
    P.kernel_far = o.kernel_far;
    for (int k = 0; k < 3; k++) P.V[k] = vinf[k];
    const double *d_x = D.upload(m.x, (size_t)3 * nn, st);
    const int *d_qn = D.upload(m.quad_nodes, (size_t)4 * nq, st);
    const int *d_tn = D.upload(m.tri_nodes, (size_t)3 * nt, st);
    P.qnodes = d_qn;
    P.qslots = D.upload(m.quad_slots, (size_t)4 * nq, st);
    P.tnodes = d_tn;
    P.tslots = D.upload(m.tri_slots, (size_t)3 * nt, st);
    P.wx = D.upload(w ? w->x : nullptr, (size_t)3 * nw, st);
    P.wtris = D.upload(w ? w->tris : nullptr, (size_t)3 * nwt, st);
    P.fil = D.upload(w ? w->filaments : nullptr, (size_t)2 * nwf, st);
    P.touch = D.upload(touch.data(), touch.size(), st);
    P.rl = D.upload(R.l.data(), R.l.size(), st);
    P.rw = D.upload(R.w.data(), R.w.size(), st);
    for (int k = 0; k < MR_COUNT; k++) {
        P.roff[k] = R.off[k];
        P.rn[k] = R.n[k];
    }
    P.gx = D.upload(gx.data(), gx.size(), st);
    P.gw = D.upload(gw.data(), gw.size(), st);
    QuadRec *d_quad = D.alloc<QuadRec>(nq);
    Elem *d_tri = D.alloc<Elem>(nt);
    Elem *d_wake = D.alloc<Elem>(nwt);
    if (nq) mixed_quad_setup<<<(nq + 127) / 128, 128, 0, st>>>(d_x, d_qn, nq, P.gx, P.gw, d_quad);
    if (nt) elem_setup<<<(nt + 255) / 256, 256, 0, st>>>(d_x, d_tn, nt, d_tri);
    if (nwt) elem_setup<<<(nwt + 255) / 256, 256, 0, st>>>(P.wx, P.wtris, nwt, d_wake);
    MCHECK(hipGetLastError());
    P.quad = d_quad;
    P.tri = d_tri;
    P.wake = d_wake;
    {  // degenerate elements: the reference throws ArgumentError in the IsoQuad / FlatTriangle constructors
        std::vector<QuadRec> hq(nq);
        std::vector<Elem> ht(nt), hw(nwt);
        if (nq) MCHECK(hipMemcpyAsync(hq.data(), d_quad, nq * sizeof(QuadRec), hipMemcpyDeviceToHost, st));
        if (nt) MCHECK(hipMemcpyAsync(ht.data(), d_tri, nt * sizeof(Elem), hipMemcpyDeviceToHost, st));
        if (nwt) MCHECK(hipMemcpyAsync(hw.data(), d_wake, nwt * sizeof(Elem), hipMemcpyDeviceToHost, st));
        MCHECK(hipStreamSynchronize(st));
        for (auto &q : hq)
            if (q.bad != 0.0) throw std::invalid_argument("degenerate quadrilateral");
        for (auto &t : ht)
            if (t.pad0 != 0.0) throw std::invalid_argument("degenerate triangle");
        for (auto &t : hw)
            if (t.pad0 != 0.0) throw std::invalid_argument("degenerate wake triangle");
    }
    const long long ncol = (long long)nslot + nw;
    P.A = D.alloc<double>((size_t)nslot * ncol);
    P.b = D.alloc<double>(nslot);
    MCHECK(hipMemsetAsync(P.A, 0, (size_t)nslot * ncol * sizeof(double), st));
    MCHECK(hipMemsetAsync(P.b, 0, nslot * sizeof(double), st));
    P.err = D.alloc<int>(1);
    MCHECK(hipMemsetAsync(P.err, 0, sizeof(int), st));
    const long long tcols = (long long)ne + nwt;
    P.tiers = tiers ? D.alloc<uint8_t>((size_t)ne * tcols) : nullptr;
    unsigned long long *d_cnt = D.alloc<unsigned long long>(4);
    MCHECK(hipMemsetAsync(d_cnt, 0, 4 * sizeof(unsigned long long), st));

    // the far tier, and the list of the other pairs
    const dim3 grid((P.nsrc + 255) / 256, ne < 32768 ? ne : 32768);
    mixed_nonfar_count<<<grid, 256, 0, st>>>(P, d_cnt);
    MCHECK(hipGetLastError());
    unsigned long long nnonfar = 0;
    MCHECK(hipMemcpyAsync(&nnonfar, d_cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    if (nnonfar > (unsigned long long)INT32_MAX) throw std::length_error("too many non-far pairs");
    int2 *d_list = D.alloc<int2>(nnonfar);
    mixed_far<<<grid, 256, 0, st>>>(P, d_list, d_cnt + 1, nnonfar, d_cnt + 2);
    MCHECK(hipGetLastError());

    // rules: counts, offsets on the host, pieces
    const int np = (int)nnonfar;
    int *d_tier = D.alloc<int>(np), *d_npc = D.alloc<int>(np);
    long long *d_npt = D.alloc<long long>(np);
    if (np)
        mixed_rules<0><<<(np + 63) / 64, 64, 0, st>>>(P, d_list, np, d_tier, d_npc, d_npt, nullptr, nullptr, nullptr);
    MCHECK(hipGetLastError());
    std::vector<int> h_tier(np), h_npc(np);
    std::vector<long long> h_npt(np);
    unsigned long long hc[4] = {0, 0, 0, 0};
    int herr = 0;
    if (np) {
        MCHECK(hipMemcpyAsync(h_tier.data(), d_tier, np * sizeof(int), hipMemcpyDeviceToHost, st));
        MCHECK(hipMemcpyAsync(h_npc.data(), d_npc, np * sizeof(int), hipMemcpyDeviceToHost, st));
        MCHECK(hipMemcpyAsync(h_npt.data(), d_npt, np * sizeof(long long), hipMemcpyDeviceToHost, st));
    }
    MCHECK(hipMemcpyAsync(hc, d_cnt, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    MCHECK(hipMemcpyAsync(&herr, P.err, sizeof(int), hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    if (hc[1] != nnonfar) throw std::runtime_error("non-far pair count changed between passes");
    if (herr == 1) throw std::invalid_argument("quad_field: a far test point lies on a source quad's boundary");
    if (herr == 2) throw std::invalid_argument("the rule builder's split stack overflowed (near_depth too large)");
    long long pr[6] = {0, 0, 0, 0, 0, 0}, pt[6] = {0, 0, 0, 0, 0, 0};
    pr[MT_FAR] = (long long)hc[2];
    pt[MT_FAR] = (long long)hc[3];
    // pieces and points of the pairs with quad sources first (mixed_eval<1>), then the rest (mixed_eval<0>)
    std::vector<long long> poff(np), qoff(np);
    long long totp = 0, totq = 0, heavy_p = 0, heavy_q = 0;
    for (int pass = 1; pass >= 0; pass--) {
        for (int i = 0; i < np; i++) {
            if (((h_tier[i] & 8) != 0) != (pass == 1)) continue;
            poff[i] = totp;
            qoff[i] = totq;
            totp += h_npc[i];
            totq += h_npt[i];
            const int t = h_tier[i] & 7;
            pr[t]++;
            pt[t] += h_npt[i];
        }
        if (pass == 1) {
            heavy_p = totp;
            heavy_q = totq;
        }
    }
    long long *d_poff = D.upload(poff.data(), np, st), *d_qoff = D.upload(qoff.data(), np, st);
    MxPiece *d_pieces = D.alloc<MxPiece>(totp);
    if (np)
        mixed_rules<1><<<(np + 63) / 64, 64, 0, st>>>(P, d_list, np, d_tier, d_npc, d_npt, d_poff, d_qoff, d_pieces);
    MCHECK(hipGetLastError());
    if ((totq + 255) / 256 > INT32_MAX) throw std::length_error("too many rule points");
    // the quad-source kernel's occupancy cap in waves per SIMD (VP_MIXED_WAVES, 1 to 4)
    const char *mw = std::getenv("VP_MIXED_WAVES");
    const int wps = mw && mw[0] >= '1' && mw[0] <= '4' ? mw[0] - '0' : 2;
    const unsigned hb = (unsigned)((heavy_q + 255) / 256);
    if (heavy_q > 0) {
        if (wps == 1) mixed_eval<1, 1><<<hb, 256, 0, st>>>(P, d_list, d_pieces, heavy_p, 0, heavy_q);
        else if (wps == 2) mixed_eval<1, 2><<<hb, 256, 0, st>>>(P, d_list, d_pieces, heavy_p, 0, heavy_q);
        else if (wps == 3) mixed_eval<1, 3><<<hb, 256, 0, st>>>(P, d_list, d_pieces, heavy_p, 0, heavy_q);
        else mixed_eval<1, 4><<<hb, 256, 0, st>>>(P, d_list, d_pieces, heavy_p, 0, heavy_q);
    }
    MCHECK(hipGetLastError());
    if (totq > heavy_q)
        mixed_eval<0, 1><<<(unsigned)((totq - heavy_q + 255) / 256), 256, 0, st>>>(
            P, d_list, d_pieces + heavy_p, totp - heavy_p, heavy_q, totq - heavy_q);
    MCHECK(hipGetLastError());
    mixed_freestream<<<(ne + 255) / 256, 256, 0, st>>>(P);
    MCHECK(hipGetLastError());
    MCHECK(hipMemcpyAsync(&herr, P.err, sizeof(int), hipMemcpyDeviceToHost, st));
    MCHECK(hipMemcpyAsync(A, P.A, (size_t)nslot * ncol * sizeof(double), hipMemcpyDeviceToHost, st));
    MCHECK(hipMemcpyAsync(b, P.b, nslot * sizeof(double), hipMemcpyDeviceToHost, st));
    if (tiers) MCHECK(hipMemcpyAsync(tiers, P.tiers, (size_t)ne * tcols, hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    if (herr) throw std::invalid_argument("quad_field: a test point lies on a source quad's boundary");
    for (int k = 0; k < 6; k++) {
        if (pairs) pairs[k] = pr[k];
        if (points) points[k] = pt[k];
    }
}

}  // namespace vp
