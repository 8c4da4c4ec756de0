// Galerkin pair assembly on gfx950: A_ij = int psi_i n.u[phi_j] dA and b_i = -int psi_i n.(Vinf + u_sigma) dA.
//
// Reference: src/assembly/assemble.jl:245-447 (pair_block, pair_rule, wake_pair_rule, segment_rule, segment_block,
// add_wake!, assemble), src/assembly/assemble.jl:69-238 (near, vertex and contact rules).
//
// Kernels (one stream per plan):
//   elem_setup     FP64 element records (centroid, diameter, reach ... in the reference's operation order)
//   test_setup     the far rule's points and test weights on every body test triangle
//   far_pairs<T>   tiled test x source pairs: lane = source element, wave = one test triangle at a time; every pair
//                  is classified in FP64 (bit-exact with the oracle); far pairs evaluate their 3-point block in the
//                  compute type T and reduce into an LDS tile R[test slot][source column], flushed to A with one
//                  atomic per nonzero tile entry; the other pairs are appended to the near list
//   seg_pairs      body test triangle x tip filament, far ones evaluated here, the others appended
//   triage         one thread per non-far pair routes it: touching (filaments, self, shared corners, contact,
//                  geometric touch pairs), plain near with one piece, plain near with several pieces
//   touch_simple   one wavefront per self / shared-edge / shared-corner pair (no contact): the stored rule or the
//                  wide corner's wedges, points flattened over the 64 lanes
//   touch_gen      one thread per remaining touching pair (contact, touch pairs, filaments): its exact tier and rule
//                  pieces (contact fan splits, vertex wedges, near splits) into a leaf list
//   group_eval     one wavefront per such pair: the points of all its pieces flattened over the 64 lanes
//   near_thread    one thread per plain near pair whose near rule is the single root piece
//   near_group     one wavefront per plain near pair with several pieces: all split tests of the quadtree in
//                  parallel over the lanes, leaves by ballot, their points flattened over the lanes
//   freestream     b_i -= (Vinf.n) A int psi_i / A
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "vp_capi_internal.hpp"
#include "vp_device.hpp"
#include "vp_rules.hpp"

namespace vp {

// ------------------------------------------------------------------------------------------ device constants -------

constexpr int WAVE = 64;
constexpr int TT_MAX = 128;  // test triangles per far tile
constexpr int TS = 64;       // source elements per far tile (one per lane)
constexpr int VMAX = 96;     // distinct test slots per far tile
constexpr int UMAX = 64;     // distinct source columns per far tile

enum Tier { T_SELF = 0, T_EDGE, T_VERTEX, T_CONTACT, T_NEAR, T_FAR };

struct TestFar {  // one per body test triangle: the far rule's points, test weights x W, normal, centroid, diam
    double x[3][3];
    double wpsi[3][3];  // [q][a]
    double n[3];
    double c[3];
    double diam, area, reach, pad;
};

struct AsmParams {
    int nt, nsrc, nseg, nslot, ncol;  // body triangles, triangle sources (body + wake), filaments, rows, columns
    int nn;                           // body nodes
    int test;                         // 0 hat, 1 onering, 2 dual
    double far_ratio, near_ratio, corner_angle, contact_rel, kernel_far;
    int near_depth;
    long long lda;
};

// ------------------------------------------------------------------------------------------ setup kernels ----------

// FlatTriangle + SourceTriangle, src/elements/flat_triangle.jl:48-55 and src/kernels/source_triangle.jl:21-28,
// with reach (flat_triangle.jl:56).  Contraction off: centroid, diameter and reach feed the bit-exact tiers.
__global__ void elem_setup(const double *__restrict__ x, const int *__restrict__ tris, int ne, Elem *__restrict__ out)
{
#pragma clang fp contract(off)
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    V3<double> a = ld3(x + 3 * tris[3 * e]), b = ld3(x + 3 * tris[3 * e + 1]), c = ld3(x + 3 * tris[3 * e + 2]);
    V3<double> nA = cross(b - a, c - a);
    double A2 = cls_norm(nA);
    V3<double> n = mk3(nA.x / A2, nA.y / A2, nA.z / A2);
    double len[3] = {cls_norm(b - a), cls_norm(c - b), cls_norm(a - c)};
    V3<double> ed[3] = {b - a, c - b, a - c};
    Elem E;
    V3<double> v[3] = {a, b, c};
    for (int k = 0; k < 3; k++) {
        E.v[k][0] = v[k].x;
        E.v[k][1] = v[k].y;
        E.v[k][2] = v[k].z;
        V3<double> nu = cross(ed[k], n);
        E.nu[k][0] = nu.x / len[k];
        E.nu[k][1] = nu.y / len[k];
        E.nu[k][2] = nu.z / len[k];
        E.len[k] = len[k];
    }
    E.n[0] = n.x;
    E.n[1] = n.y;
    E.n[2] = n.z;
    V3<double> s = (a + b) + c;
    V3<double> cc = mk3(s.x / 3.0, s.y / 3.0, s.z / 3.0);
    E.c[0] = cc.x;
    E.c[1] = cc.y;
    E.c[2] = cc.z;
    E.area = A2 / 2;
    E.diam = fmax(fmax(len[0], len[1]), len[2]);
    E.reach = fmax(fmax(cls_norm(a - cc), cls_norm(b - cc)), cls_norm(c - cc));
    V3<double> eo[3] = {c - b, a - c, b - a};  // the edge opposite each corner
    for (int k = 0; k < 3; k++) {
        E.g[k][0] = -eo[k].x / A2;
        E.g[k][1] = -eo[k].y / A2;
        E.g[k][2] = -eo[k].z / A2;
    }
    E.pad0 = A2 > 0 ? 0.0 : 1.0;  // degenerate flag, checked by the host
    out[e] = E;
}

// test_weights, src/assembly/assemble.jl:52-58
__device__ __forceinline__ void test_weights(int test, const double *l, double *psi)
{
    if (test == 0) {
        psi[0] = l[0];
        psi[1] = l[1];
        psi[2] = l[2];
    } else if (test == 1) {
        psi[0] = psi[1] = psi[2] = 1.0;
    } else {
        double m = fmax(fmax(l[0], l[1]), l[2]);
        int is[3], c = 0;
        for (int k = 0; k < 3; k++) {
            is[k] = fabs(l[k] - m) < 1e-12;
            c += is[k];
        }
        for (int k = 0; k < 3; k++) psi[k] = (double)is[k] / c;
    }
}

__global__ void test_setup(const Elem *__restrict__ el, int nt, RuleTab rt, int test, TestFar *__restrict__ out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const Elem &E = el[t];
    V3<double> X[3] = {ld3(E.v[0]), ld3(E.v[1]), ld3(E.v[2])};
    TestFar F;
    for (int q = 0; q < 3; q++) {
        const double *l = rt.l + 3 * (rt.off[R_FAR] + q);
        V3<double> p = cls_point(X, l);
        F.x[q][0] = p.x;
        F.x[q][1] = p.y;
        F.x[q][2] = p.z;
        double psi[3];
        test_weights(test, l, psi);
        double W = rt.w[rt.off[R_FAR] + q] * E.area;
        for (int a = 0; a < 3; a++) F.wpsi[q][a] = W * psi[a];
    }
    for (int k = 0; k < 3; k++) {
        F.n[k] = E.n[k];
        F.c[k] = E.c[k];
    }
    F.diam = E.diam;
    F.area = E.area;
    F.reach = E.reach;
    F.pad = 0.0;
    out[t] = F;
}

}  // namespace vp

#include "vp_adjoint_dev.hpp"

namespace vp {

// ------------------------------------------------------------------------------------------ far pairs --------------

// contact_rule's early out, src/assembly/assemble.jl:200-201: a pair can only be a contact pair when its centroids
// are within reach(T) + reach(S) + contact_rel min(diam_T, diam_S)
__device__ inline bool cls_contact_possible(V3<double> cs, V3<double> ct, double rs, double rt, double ds, double dt,
                                            double contact_rel)
{
#pragma clang fp contract(off)
    double tol = contact_rel * fmin(dt, ds);
    return !(cls_norm(cls_sub(cs, ct)) > (rt + rs) + tol);
}

// Wave-aggregated append of (t, s) to the near list.
__device__ __forceinline__ void append_pair(bool want, int t, int s, int2 *list, unsigned int *count, unsigned int cap)
{
    unsigned long long mask = __ballot(want);
    if (!mask) return;
    int lane = threadIdx.x & (WAVE - 1);
    unsigned int base = 0;
    int leader = __ffsll((long long)mask) - 1;
    if (lane == leader) base = atomicAdd(count, (unsigned int)__popcll(mask));
    base = __shfl(base, leader);
    if (want) {
        unsigned int pos = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
        if (pos < cap) list[pos] = make_int2(t, s);
    }
}

template <typename T> struct FarSrc {  // per-lane source element data for the far tier
    V3<double> c;                      // centroid (FP64: the classification and the relative offsets)
    double diam, reach;
    FarMom<T> m;                       // the multipole's vertex moments about the centroid, in the compute type
    V3<T> g[3];
    T sigma;
};

template <typename T> __device__ __forceinline__ FarSrc<T> load_far_src(const Elem &E, double sigma)
{
    FarSrc<T> f;
    f.c = ld3(E.c);
    f.diam = E.diam;
    f.reach = E.reach;
    V3<T> d[3];
    for (int k = 0; k < 3; k++) {
        d[k] = mk3(T(E.v[k][0] - E.c[0]), T(E.v[k][1] - E.c[1]), T(E.v[k][2] - E.c[2]));
        f.g[k] = mk3(T(E.g[k][0]), T(E.g[k][1]), T(E.g[k][2]));
    }
    f.m = far_moments(d, T(E.area));
    f.sigma = T(sigma);
    return f;
}

// One kernel evaluation of source_velocity (src/kernels/source_triangle.jl:134-138) at x (FP64) for the far tier:
// the multipole beyond kernel_far diameters in T, the closed form in FP64 otherwise.  (The FP32 closed form,
// sv_exact_f32, measured no faster here on the swept wing, 208 against 210 ms: the far pass is not bound by its
// arithmetic there; so the mixed path keeps these points' FP64 accuracy.)
// The closed form of source_velocity at x for the far tier's points inside kernel_far diameters (FP64), c the source's
// centroid
__device__ __forceinline__ V3<double> far_exact(const Elem *__restrict__ E, V3<double> c, V3<double> x)
{
    SrcT<double> s;
    for (int k = 0; k < 3; k++) {
        s.v[k] = ld3(E->v[k]) - c;
        s.nu[k] = ld3(E->nu[k]);
        s.len[k] = E->len[k];
    }
    s.n = ld3(E->n);
    s.diam = E->diam;
    s.area = E->area;
    V3<double> ra = x - ld3(E->v[0]), rb = x - ld3(E->v[1]), rc = x - ld3(E->v[2]);
    return sv_exact_abs(s, ra, rb, rc, dot(x - c, s.n), 0, nullptr, x);
}

template <typename T>
__device__ __forceinline__ V3<double> far_eval(const FarSrc<T> &f, const Elem *__restrict__ E, V3<double> x, double kfar2,
                                               unsigned int &nexact)
{
    V3<double> r = x - f.c;
    if (cls_dot(r, r) > kfar2) {
        V3<T> w = sv_far_mom(f.m, mk3(T(r.x), T(r.y), T(r.z)));
        return mk3((double)w.x, (double)w.y, (double)w.z);
    }
    nexact++;
    return far_exact(E, f.c, x);
}

// The pair_block of a far pair (assemble.jl:245-258) from its three far-rule points: blk[a][b] += psi_a (w x n).gamma_b,
// rb[a] -= sigma_S psi_a (w.n), with the weights W psi folded into F.wpsi; EXACT admits closed-form points, and
// without it the points in the bit mask skip are left out (they are summed elsewhere)
template <typename T, bool EXACT>
__device__ __forceinline__ void far_block(const FarSrc<T> &f, const Elem *__restrict__ E, const TestFar &F, double kfar2,
                                          double (*blk)[3], double *rb, unsigned int &nexact, unsigned int skip = 0)
{
    const V3<double> n = ld3(F.n);
    // unrolled: the three points are independent (C2 far kernel 17.1 -> 16.3 ms, no spills at two waves per SIMD)
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (!EXACT && (skip >> q & 1u)) continue;
        V3<double> w;
        if constexpr (EXACT) {
            w = far_eval<T>(f, E, ld3(F.x[q]), kfar2, nexact);
        } else {
            const V3<double> r = ld3(F.x[q]) - f.c;
            const V3<T> wt = sv_far_mom(f.m, mk3(T(r.x), T(r.y), T(r.z)));
            w = mk3((double)wt.x, (double)wt.y, (double)wt.z);
        }
        V3<double> vv = cross(w, n);  // n.(gamma x w) = gamma.(w x n), assemble.jl:251
        double c0 = dot(vv, mk3((double)f.g[0].x, (double)f.g[0].y, (double)f.g[0].z));
        double c1 = dot(vv, mk3((double)f.g[1].x, (double)f.g[1].y, (double)f.g[1].z));
        double c2 = dot(vv, mk3((double)f.g[2].x, (double)f.g[2].y, (double)f.g[2].z));
        double wn = (double)f.sigma * dot(w, n);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double psi = F.wpsi[q][a];
            blk[a][0] += psi * c0;
            blk[a][1] += psi * c1;
            blk[a][2] += psi * c2;
            rb[a] -= wn * psi;
        }
    }
}

// A far tile's LDS sums (blocks per local source column x local test slot, R[u * RST + v]; RHS per local test slot)
// into A and b.  Consecutive lanes take consecutive test slots of one source column: the tile's slots are sorted, so
// a wave's atomic adds fall on ascending addresses of one column of A (column-major) rather than one cache line per
// lane (C4: the far pass was bound by those scattered atomics).
constexpr int RST = VMAX + 1;  // the tile's column stride in LDS: odd, so that the lanes' columns spread over the banks
__device__ __forceinline__ void far_flush(const AsmParams &P, const double *R, const double *rbacc, int tt, int st,
                                          const int *tt_rows, const int *tt_nrow, const int *ts_cols, const int *ts_ncol,
                                          double *A, double *bvec)
{
    const int nrow = tt_nrow[tt], ncol = ts_ncol[st];
    for (int i = threadIdx.x; i < nrow * ncol; i += blockDim.x) {
        const int u = i / nrow, v = i - u * nrow;
        const double val = R[u * RST + v];
        if (val != 0.0) unsafeAtomicAdd(&A[(long long)ts_cols[st * UMAX + u] * P.lda + tt_rows[tt * VMAX + v]], val);
    }
    for (int v = threadIdx.x; v < nrow; v += blockDim.x)
        if (rbacc[v] != 0.0) unsafeAtomicAdd(&bvec[tt_rows[tt * VMAX + v]], rbacc[v]);
}

// The far tier over tiles of TT_MAX test triangles (Morton order, waves) x TS source elements (lanes): pair_rule's far
// test (FP64, no contraction: bit-exact tiers), non-far pairs appended to the near list, far pairs evaluated at their
// three far-rule points with source_velocity's switch and summed per local test slot and source column in LDS, one
// flush per tile.  The closed form is rare here but carries most of the kernel's registers; capping the kernel at two
// waves per SIMD (a few registers spill in that path) beats the one wave it would otherwise get.  enum_only: the near
// list alone, its second pass after it grew.
#define FAR_ARGS                                                                                                    \
    AsmParams P, const TestFar *__restrict__ tf, const Elem *__restrict__ el, const double *__restrict__ sigma,        \
        const int *__restrict__ tt_elems, const uint8_t *__restrict__ tt_loc, const int *__restrict__ tt_rows,         \
        const int *__restrict__ tt_nrow, const int *__restrict__ ts_elems, const uint8_t *__restrict__ ts_loc,         \
        const int *__restrict__ ts_cols, const int *__restrict__ ts_ncol, double *__restrict__ A,                      \
        double *__restrict__ bvec, int2 *__restrict__ near_list, unsigned int *__restrict__ near_count,                \
        unsigned int near_cap, unsigned long long *__restrict__ far_count, uint8_t *__restrict__ tiers, int st0,      \
        bool enum_only
#define FAR_PASS                                                                                                    \
    P, tf, el, sigma, tt_elems, tt_loc, tt_rows, tt_nrow, ts_elems, ts_loc, ts_cols, ts_ncol, A, bvec, near_list,      \
        near_count, near_cap, far_count, tiers, st0, enum_only

// The closed-form points of a tile's far pairs (DEFER), after the tile's multipole points, one point per thread:
// entries (i << 8) | (q << 6) | lane of the workgroup's queue dq (test triangle i of the tile, far-rule point q, source
// lane), each point's terms of far_block summed into the tile's LDS sums.  Only the source's element record is read
// here (no multipole moments), so that the closed form's registers are not live beside the multipole pass's: the FP32
// far kernel runs at three waves per SIMD instead of two.
template <typename T>
__device__ __forceinline__ void far_deferred(const TestFar *__restrict__ tf, const Elem *__restrict__ el,
                                             const double *__restrict__ sigma, int nt, const int *__restrict__ tt_elems,
                                             const uint8_t *__restrict__ tt_loc, const int *__restrict__ ts_elems,
                                             const uint8_t *__restrict__ ts_loc, int tt, int st,
                                             const uint16_t *__restrict__ dq, unsigned int nq, double *R, double *rbacc)
{
    for (unsigned int j = threadIdx.x; j < nq; j += blockDim.x) {
        const unsigned int e = dq[j];
        const int i = (int)(e >> 8), q = (int)(e >> 6 & 3u), ln = (int)(e & 63u);
        const int t = tt_elems[tt * TT_MAX + i], s = ts_elems[st * TS + ln];
        const Elem *E = el + s;
        const TestFar &F = tf[t];
        const V3<double> n = ld3(F.n);
        // the same conversions as load_far_src and far_block
        const V3<double> w = far_exact(E, ld3(E->c), ld3(F.x[q])), vv = cross(w, n);
        double c[3];
        for (int b = 0; b < 3; b++)
            c[b] = dot(vv, mk3((double)T(E->g[b][0]), (double)T(E->g[b][1]), (double)T(E->g[b][2])));
        const double wn = (double)T(s < nt ? sigma[s] : 0.0) * dot(w, n);
        int uloc[3], vl[3];
        for (int k = 0; k < 3; k++) {
            uloc[k] = ts_loc[(st * TS + ln) * 3 + k];
            vl[k] = tt_loc[(tt * TT_MAX + i) * 3 + k];
        }
        for (int a = 0; a < 3; a++) {
            const double psi = F.wpsi[q][a];
            for (int b = 0; b < 3; b++) atomicAdd(&R[uloc[b] * RST + vl[a]], psi * c[b]);
            atomicAdd(&rbacc[vl[a]], -(wn * psi));
        }
    }
}

// One tile (test tile tt, source tile st) of the far tier, by the whole workgroup, with its LDS sums R and rbacc.
// DEFER: the far pairs' closed-form points go to the workgroup's queue dq (3 TT_MAX TS entries, count nq in LDS) and
// are evaluated after the tile's multipole points; otherwise every far point is evaluated in place.
template <typename T, bool DEFER>
__device__ __forceinline__ void far_tile(FAR_ARGS, int tt, int st, double *R, double *rbacc, unsigned int &nfar,
                                         unsigned int &nexw, unsigned int &nq, uint16_t *__restrict__ dq)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    if (!enum_only) {
        for (int i = threadIdx.x; i < UMAX * RST; i += blockDim.x) R[i] = 0.0;
        if (threadIdx.x < VMAX) rbacc[threadIdx.x] = 0.0;
    }
    if (threadIdx.x == 0) nfar = nexw = nq = 0;
    const int s = ts_elems[st * TS + lane];
    const bool valid = s >= 0;
    FarSrc<T> f;
    int uloc[3] = {0, 0, 0};
    if (valid) {
        f = load_far_src<T>(el[s], s < P.nt ? sigma[s] : 0.0);
        for (int k = 0; k < 3; k++) uloc[k] = ts_loc[(st * TS + lane) * 3 + k];
    }
    const double kfar2 = (P.kernel_far * (valid ? f.diam : 0.0)) * (P.kernel_far * (valid ? f.diam : 0.0));
    __syncthreads();
    unsigned int myfar = 0, myex = 0;
    for (int i = wave; i < TT_MAX; i += 4) {
        const int t = tt_elems[tt * TT_MAX + i];
        if (t < 0) break;
        const TestFar &F = tf[t];
        bool far = false;
        if (valid && s != t) {
            // pair_rule's far test, src/assembly/assemble.jl:269 (bit-exact: FP64, no contraction)
            V3<double> d = f.c - ld3(F.c);
            far = cls_far(d, P.far_ratio * fmax(F.diam, f.diam));
        }
        append_pair(valid && !far, t, s, near_list, near_count, near_cap);
        if (enum_only) continue;  // the far blocks are in A already
        if (tiers && valid && far) tiers[(size_t)t * P.nsrc + s] = T_FAR;
        if (far) {
            myfar++;
            unsigned int skip = 0;  // DEFER: the pair's closed-form points, queued
            if constexpr (DEFER) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const V3<double> r = ld3(F.x[q]) - f.c;
                    if (!(cls_dot(r, r) > kfar2)) {
                        skip |= 1u << q;
                        dq[atomicAdd(&nq, 1u)] = (uint16_t)((i << 8) | (q << 6) | lane);
                    }
                }
                myex += __builtin_popcount(skip);
            }
            double blk[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, rb[3] = {0, 0, 0};
            far_block<T, !DEFER>(f, el + s, F, kfar2, blk, rb, myex, skip);
            const int vl[3] = {tt_loc[(tt * TT_MAX + i) * 3], tt_loc[(tt * TT_MAX + i) * 3 + 1],
                               tt_loc[(tt * TT_MAX + i) * 3 + 2]};
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) atomicAdd(&R[uloc[b] * RST + vl[a]], blk[a][b]);
#pragma unroll
            for (int a = 0; a < 3; a++) atomicAdd(&rbacc[vl[a]], rb[a]);
        }
    }
    if (enum_only) return;
    if constexpr (DEFER) {
        __syncthreads();  // the queue's entries and count
        far_deferred<T>(tf, el, sigma, P.nt, tt_elems, tt_loc, ts_elems, ts_loc, tt, st, dq, nq, R, rbacc);
    }
    if (myfar) atomicAdd(&nfar, myfar);
    if (myex) atomicAdd(&nexw, myex);
    __syncthreads();
    far_flush(P, R, rbacc, tt, st, tt_rows, tt_nrow, ts_cols, ts_ncol, A, bvec);
    if (threadIdx.x == 0 && nfar) atomicAdd(far_count, (unsigned long long)nfar);
    if (threadIdx.x == 0 && nexw) atomicAdd(far_count + 7, (unsigned long long)nexw);  // stats[12]: exact far points
}

// The far tier as a grid of tiles: test tile blockIdx.x, source tile blockIdx.y + st0
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void far_pairs(FAR_ARGS)
{
    __shared__ double R[UMAX * RST];
    __shared__ double rbacc[VMAX];
    __shared__ unsigned int nfar, nexw, nq;
    far_tile<T, false>(FAR_PASS, blockIdx.x, blockIdx.y + st0, R, rbacc, nfar, nexw, nq, nullptr);
}

// far_pairs_queue<T>: whether it defers the closed-form points, and its waves per SIMD (and workgroups per CU).  FP32
// multipoles defer and run at three (C2 mixed assembly 32.3 -> 31.5 ms); FP64 evaluates in place at two (deferred, at
// three it spills and took 22.8 against 14.3 ms, at two 15.3)
#define FAR_Q_DEFER(T) (sizeof(T) == 4)
#define FAR_Q_WAVES(T) (FAR_Q_DEFER(T) ? 3 : 2)
// The far tier as one persistent work queue: a grid sized to the chip's resident workgroups takes tiles in order
// (test tile fastest, as the grid form) from one counter until the ntt x nst tiles are done; every workgroup leaves
// when the counter passes the last tile.  Tiles of bodies of any size share the queue.  Each workgroup defers a
// FP32 tile's closed-form points to its own slice of fq (3 TT_MAX TS entries), which lets that kernel run at three waves
// per SIMD (three workgroups per CU).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FAR_Q_WAVES(T), FAR_Q_WAVES(T)))) void far_pairs_queue(
    FAR_ARGS, int ntt, int nst, unsigned int *__restrict__ queue, uint16_t *__restrict__ fq)
{
    __shared__ double R[UMAX * RST];
    __shared__ double rbacc[VMAX];
    __shared__ unsigned int nfar, nexw, nq, tile;
    uint16_t *const dq = fq + (size_t)blockIdx.x * (3 * TT_MAX * TS);
    const unsigned int ntiles = (unsigned int)ntt * (unsigned int)nst;
    for (;;) {
        if (threadIdx.x == 0) tile = atomicAdd(queue, 1u);
        __syncthreads();
        const unsigned int k = tile;
        if (k >= ntiles) break;
        far_tile<T, FAR_Q_DEFER(T)>(FAR_PASS, (int)(k % (unsigned int)ntt), (int)(k / (unsigned int)ntt) + st0, R, rbacc,
                                    nfar, nexw, nq, dq);
        __syncthreads();  // the flush has read R and the counters before the next tile clears them
    }
}

// ------------------------------------------------------------------------------------------ filaments --------------

// segment_rule's far test (src/assembly/assemble.jl:299-301) and the far segment_block (:312-321)
// An upper bound of the entries the far pass appends to the non-far list: far_pairs' test norm(c_S - c_T) >
// far_ratio max(diam_T, diam_S) over the sources [s_begin, s_end), and seg_pairs' with a filament's midpoint and length,
// taken on squared distances with a relative margin of 1e-12 (the FP64 tests' rounding is a few ulps), so that no pair
// they keep is missed.  Thread per test triangle, sources in LDS tiles over grid.y chunks; one atomic per workgroup.
constexpr int BOUND_CHUNK = 4096;
__global__ __launch_bounds__(256) void nonfar_bound(const Elem *__restrict__ el, int nt, int s_begin, int s_end,
                                                    const double *__restrict__ seg, int nseg, double far_ratio,
                                                    unsigned long long *__restrict__ count)
{
    __shared__ double S[256][4];
    __shared__ unsigned int part[4];
    const int t = blockIdx.x * 256 + threadIdx.x, ne = s_end - s_begin, ns = ne + nseg;
    const bool live = t < nt;
    const V3<double> tc = live ? ld3(el[t].c) : mk3(0.0, 0.0, 0.0);
    const double td = live ? el[t].diam : 0.0;
    const int j0 = blockIdx.y * BOUND_CHUNK, j1 = min(ns, j0 + BOUND_CHUNK);
    unsigned int c = 0;
    for (int base = j0; base < j1; base += 256) {
        __syncthreads();
        const int j = base + threadIdx.x;
        if (j < j1) {
            if (j < ne) {
                const Elem &E = el[s_begin + j];
                S[threadIdx.x][0] = E.c[0];
                S[threadIdx.x][1] = E.c[1];
                S[threadIdx.x][2] = E.c[2];
                S[threadIdx.x][3] = E.diam;
            } else {
                const V3<double> fa = ld3(seg + 6 * (j - ne)), fb = ld3(seg + 6 * (j - ne) + 3);
                S[threadIdx.x][0] = (fa.x + fb.x) / 2;
                S[threadIdx.x][1] = (fa.y + fb.y) / 2;
                S[threadIdx.x][2] = (fa.z + fb.z) / 2;
                S[threadIdx.x][3] = cls_norm(fb - fa);
            }
        }
        __syncthreads();
        const int nk = min(256, j1 - base);
        if (live)
            for (int k = 0; k < nk; k++) {
                const double dx = S[k][0] - tc.x, dy = S[k][1] - tc.y, dz = S[k][2] - tc.z;
                const double r = far_ratio * fmax(td, S[k][3]);
                c += dx * dx + dy * dy + dz * dz <= r * r * (1.0 + 1e-12);
            }
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tot = (unsigned long long)part[0] + part[1] + part[2] + part[3];
        if (tot) atomicAdd(count, tot);
    }
}

__global__ void seg_pairs(AsmParams P, const TestFar *__restrict__ tf, const int *__restrict__ corner_slot,
                          const double *__restrict__ seg, const int *__restrict__ segcol, double *__restrict__ A,
                          int2 *__restrict__ near_list, unsigned int *__restrict__ near_count, unsigned int near_cap,
                          unsigned long long *__restrict__ far_count, bool enum_only)
{
#pragma clang fp contract(off)
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = idx < (long long)P.nt * P.nseg;
    int t = in ? (int)(idx / P.nseg) : 0, f = in ? (int)(idx % P.nseg) : 0;
    bool far = false;
    V3<double> fa, fb;
    if (in) {
        const TestFar &F = tf[t];
        fa = ld3(seg + 6 * f);
        fb = ld3(seg + 6 * f + 3);
        double L = cls_norm(fb - fa);
        V3<double> mid = fa + fb;
        mid = mk3(mid.x / 2, mid.y / 2, mid.z / 2);
        far = cls_far(mid - ld3(F.c), P.far_ratio * fmax(F.diam, L));
    }
    append_pair(in && !far, t, P.nsrc + f, near_list, near_count, near_cap);
    if (enum_only) return;
    unsigned long long m = __ballot(far);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(far_count, (unsigned long long)__popcll(m));
    if (!far) return;
    const TestFar &F = tf[t];
    V3<double> n = ld3(F.n);
    double ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
    for (int q = 0; q < 3; q++) {
        V3<double> x = ld3(F.x[q]), va, vb;
        seg_vel(x - fa, x - fb, va, vb);
        double da = dot(va, n), db = dot(vb, n);
        for (int a = 0; a < 3; a++) {
            ca[a] += da * F.wpsi[q][a];
            cb[a] += db * F.wpsi[q][a];
        }
    }
    const int *ct = corner_slot + 3 * t;
    int ca_col = segcol[2 * f], cb_col = segcol[2 * f + 1];
    for (int a = 0; a < 3; a++) {
        unsafeAtomicAdd(&A[(long long)ca_col * P.lda + ct[a]], ca[a]);
        unsafeAtomicAdd(&A[(long long)cb_col * P.lda + ct[a]], cb[a]);
    }
}

// ------------------------------------------------------------------------------------------ near pairs -------------

constexpr int PSTACK = 32;  // rule-piece stack entries per thread in triage

struct Piece {  // a piece of the test triangle in barycentric corners, with its area fraction
    double b[3][3];
    double frac;
    int kind;  // 0 near (recursive split), 1 contact (fan split)
    int level, depth, pad;
};

struct Leaf {  // a rule piece to evaluate: the pair, the base rule, the piece's barycentric corners, its area fraction
    double b[3][3];
    double frac;
    int t, s, rid, npt;
};

// Leaf generation for one pair on one thread (triage): the rule pieces of pair_rule's touching tiers and of the near
// rule of touch pairs and filaments.  Run twice: once counting (out null), once writing at the reserved base.
struct Gen {
    V3<double> X[3], SX[3], fa, fb;
    bool seg;
    double tol;
    double near_ratio, corner_angle;
    int near_depth;
    int t, s;
    const int *rn;  // rule sizes
    Leaf *out;
    unsigned int base, cap, n;
    long long npts;
    bool overflow;
};

__device__ void gen_emit(Gen &g, int rid, const double (*b)[3], double frac)
{
    if (g.out && g.base + g.n < g.cap) {
        Leaf &L = g.out[g.base + g.n];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) L.b[i][j] = b[i][j];
        L.frac = frac;
        L.t = g.t;
        L.s = g.s;
        L.rid = rid;
        L.npt = g.rn[rid];
    }
    g.n++;
    g.npts += g.rn[rid];
}

__device__ __forceinline__ double src_distance(const Gen &g, V3<double> p)
{
    return g.seg ? cls_seg_projection(p, g.fa, g.fb, nullptr) : cls_tri_distance(g.SX, p);
}

__device__ __forceinline__ void set_corners(double (*b)[3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) b[i][j] = i == j ? 1.0 : 0.0;
}

// vertex_pieces!, src/assembly/assemble.jl:102-116 (k 0-based, p the piece's physical corners)
__device__ void vertex_pieces(Gen &g, const V3<double> *p, int k, const double (*b)[3], double frac)
{
#pragma clang fp contract(off)
    int i = (k + 1) % 3, j = (k + 2) % 3;
    V3<double> u = cls_sub(p[i], p[k]), v = cls_sub(p[j], p[k]);
    double nu = cls_norm(u), nv = cls_norm(v);
    double cs = cls_dot(u, v) / (nu * nv);
    cs = cs < -1.0 ? -1.0 : cs > 1.0 ? 1.0 : cs;
    double th = acos(cs);
    double bb[3][3];
    if (th <= kPi / 2) {
        for (int q = 0; q < 3; q++) {
            bb[0][q] = b[k][q];
            bb[1][q] = b[i][q];
            bb[2][q] = b[j][q];
        }
        gen_emit(g, R_VERT0, bb, frac);
        return;
    }
    int m = (int)ceil(th * (180 / kPi) / g.corner_angle);
    double t0 = 0.0;
    for (int q = 1; q <= m; q++) {
        double t1 = 1.0;
        if (q < m) {
            double ph = th * q / m;
            t1 = nu * sin(ph) / (nu * sin(ph) + nv * sin(th - ph));
        }
        for (int z = 0; z < 3; z++) {
            bb[0][z] = b[k][z];
            bb[1][z] = (1 - t0) * b[i][z] + t0 * b[j][z];
            bb[2][z] = (1 - t1) * b[i][z] + t1 * b[j][z];
        }
        gen_emit(g, R_VERT0, bb, frac * (t1 - t0));
        t0 = t1;
    }
}

// vertex_rule, src/assembly/assemble.jl:119-125
__device__ void vertex_rule(Gen &g, int k)
{
#pragma clang fp contract(off)
    double b0[3][3];
    set_corners(b0);
    V3<double> u = cls_sub(g.X[(k + 1) % 3], g.X[k]), v = cls_sub(g.X[(k + 2) % 3], g.X[k]);
    if (cls_dot(u, v) >= 0)
        gen_emit(g, R_VERT0 + k, b0, 1.0);
    else
        vertex_pieces(g, g.X, k, b0, 1.0);
}

// The splits of the piece's edge k at the source triangle's corners (edge_splits, assemble.jl:160-171)
__device__ int edge_splits_k(const Gen &g, const V3<double> *p, int k, double *ts)
{
#pragma clang fp contract(off)
    V3<double> a = p[k], e = p[(k + 1) % 3];
    double L = cls_norm(cls_sub(e, a));
    int n = 0;
    for (int q = 0; q < 3; q++) {
        double t;
        double d = cls_seg_projection(g.SX[q], a, e, &t);
        if (d <= g.tol && t * L > g.tol && (1 - t) * L > g.tol) ts[n++] = t;
    }
    for (int i = 1; i < n; i++)
        for (int j = i; j > 0 && ts[j] < ts[j - 1]; j--) {
            double x = ts[j];
            ts[j] = ts[j - 1];
            ts[j - 1] = x;
        }
    return n;
}

// The rule pieces of near_pieces! (assemble.jl:131-142) and contact_pieces! (:209-238), depth first on an explicit
// stack, children pushed in reverse so they come out in the reference's order.
__device__ void walk(Gen &g, Piece root)
{
#pragma clang fp contract(off)
    Piece stk[PSTACK];
    int sp = 0;
    stk[sp++] = root;
    while (sp > 0) {
        Piece pc = stk[--sp];
        V3<double> p[3] = {cls_point(g.X, pc.b[0]), cls_point(g.X, pc.b[1]), cls_point(g.X, pc.b[2])};
        if (pc.kind == 0) {
            double size = fmax(fmax(cls_norm(cls_sub(p[1], p[0])), cls_norm(cls_sub(p[2], p[1]))), cls_norm(cls_sub(p[0], p[2])));
            V3<double> cen = (p[0] + p[1]) + p[2];
            cen = mk3(cen.x / 3.0, cen.y / 3.0, cen.z / 3.0);
            if (pc.level < pc.depth && src_distance(g, cen) < g.near_ratio * size) {
                if (sp + 4 > PSTACK) {
                    g.overflow = true;
                    continue;
                }
                double m12[3], m23[3], m31[3];
                for (int j = 0; j < 3; j++) {
                    m12[j] = (pc.b[0][j] + pc.b[1][j]) / 2;
                    m23[j] = (pc.b[1][j] + pc.b[2][j]) / 2;
                    m31[j] = (pc.b[2][j] + pc.b[0][j]) / 2;
                }
                const double *kids[4][3] = {{pc.b[0], m12, m31}, {m12, pc.b[1], m23}, {m31, m23, pc.b[2]}, {m12, m23, m31}};
                for (int q = 3; q >= 0; q--) {
                    Piece ch;
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) ch.b[i][j] = kids[q][i][j];
                    ch.frac = pc.frac / 4;
                    ch.kind = 0;
                    ch.level = pc.level + 1;
                    ch.depth = pc.depth;
                    stk[sp++] = ch;
                }
            } else {
                gen_emit(g, R_NEAR, pc.b, pc.frac);
            }
            continue;
        }
        if (pc.level < 4) {
            int k = 0, ns = 0;
            double sp3[3];
            for (k = 0; k < 3; k++) {
                ns = edge_splits_k(g, p, k, sp3);
                if (ns) break;
            }
            if (ns) {
                if (sp + ns + 1 > PSTACK) {
                    g.overflow = true;
                    continue;
                }
                double ts[5];
                ts[0] = 0.0;
                for (int i = 0; i < ns; i++) ts[i + 1] = sp3[i];
                ts[ns + 1] = 1.0;
                const double *bk = pc.b[k], *bk1 = pc.b[(k + 1) % 3], *opp = pc.b[(k + 2) % 3];
                for (int i = ns; i >= 0; i--) {
                    Piece ch;
                    for (int z = 0; z < 3; z++) {
                        ch.b[0][z] = (1 - ts[i]) * bk[z] + ts[i] * bk1[z];
                        ch.b[1][z] = (1 - ts[i + 1]) * bk[z] + ts[i + 1] * bk1[z];
                        ch.b[2][z] = opp[z];
                    }
                    ch.frac = pc.frac * (ts[i + 1] - ts[i]);
                    ch.kind = 1;
                    ch.level = pc.level + 1;
                    ch.depth = 0;
                    stk[sp++] = ch;
                }
                continue;
            }
        }
        int on[3], nc = 0;
        for (int k = 0; k < 3; k++) {
            on[k] = cls_tri_distance(g.SX, p[k]) <= g.tol;
            nc += on[k];
        }
        if (nc == 2) {
            int k = !on[2] ? 0 : !on[0] ? 1 : 2;
            V3<double> mid = p[k] + p[(k + 1) % 3];
            mid = mk3(mid.x / 2, mid.y / 2, mid.z / 2);
            if (cls_tri_distance(g.SX, mid) <= g.tol) {
                gen_emit(g, R_EDGE0 + k, pc.b, pc.frac);
                continue;
            }
        }
        if (nc == 1) {
            vertex_pieces(g, p, on[0] ? 0 : on[1] ? 1 : 2, pc.b, pc.frac);
        } else if (nc == 0) {
            Piece ch = pc;
            ch.kind = 0;
            ch.level = 0;
            ch.depth = g.near_depth;
            stk[sp++] = ch;
        } else {
            gen_emit(g, R_SELF, pc.b, pc.frac);
        }
    }
}

// Classify one non-far pair exactly as pair_rule / segment_rule do (src/assembly/assemble.jl:266-306, after the far
// test) and produce its rule pieces; returns the tier, or -1 for a plain near pair (near tier, no touch pair).
__device__ int classify_and_generate(Gen &g, const AsmParams &P, const Elem *el, const int *elem_nodes,
                                     const int2 *touch, int ntouch, bool force_walk)
{
#pragma clang fp contract(off)
    double b0[3][3];
    set_corners(b0);
    const Elem &T = el[g.t];
    if (g.seg) {
        // segment_rule: the vertex rule where the filament starts at a corner of T, else the near rule
        int kv = -1;
        for (int k = 0; k < 3 && kv < 0; k++)
            if (fmin(cls_norm(cls_sub(g.X[k], g.fa)), cls_norm(cls_sub(g.X[k], g.fb))) <= 1e-12 * T.diam) kv = k;
        if (kv >= 0) {
            vertex_rule(g, kv);
            return T_VERTEX;
        }
        Piece r0;
        set_corners(r0.b);
        r0.frac = 1.0;
        r0.kind = 0;
        r0.level = 0;
        r0.depth = g.near_depth;
        walk(g, r0);
        return T_NEAR;
    }
    if (g.s == g.t) {
        gen_emit(g, R_SELF, b0, 1.0);
        return T_SELF;
    }
    const Elem &S = el[g.s];
    int shared[3];
    for (int k = 0; k < 3; k++) {
        int v = elem_nodes[3 * g.t + k];
        shared[k] = v == elem_nodes[3 * g.s] || v == elem_nodes[3 * g.s + 1] || v == elem_nodes[3 * g.s + 2];
    }
    // contact_rule, src/assembly/assemble.jl:199-207
    g.tol = P.contact_rel * fmin(T.diam, S.diam);
    if (cls_contact_possible(ld3(S.c), ld3(T.c), S.reach, T.reach, S.diam, T.diam, P.contact_rel)) {
        bool extra = false;
        for (int k = 0; k < 3 && !extra; k++) extra = !shared[k] && cls_tri_distance(g.SX, g.X[k]) <= g.tol;
        for (int k = 0; k < 3 && !extra; k++) {
            double ts3[3];
            extra = edge_splits_k(g, g.X, k, ts3) > 0;
        }
        if (extra) {
            Piece r0;
            set_corners(r0.b);
            r0.frac = 1.0;
            r0.kind = 1;
            r0.level = 0;
            r0.depth = 0;
            walk(g, r0);
            return T_CONTACT;
        }
    }
    int ns = shared[0] + shared[1] + shared[2];
    if (ns == 2) {
        gen_emit(g, R_EDGE0 + (shared[0] && shared[1] ? 0 : shared[1] && shared[2] ? 1 : 2), b0, 1.0);
        return T_EDGE;
    }
    if (ns == 1) {
        vertex_rule(g, shared[0] ? 0 : shared[1] ? 1 : 2);
        return T_VERTEX;
    }
    bool tp = false;
    if (g.s < P.nt && ntouch) {  // geometric touch pairs get a deeper near rule (assemble.jl:280)
        int lo = 0, hi = ntouch;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            int2 q = touch[mid];
            if (q.x < g.t || (q.x == g.t && q.y < g.s)) lo = mid + 1; else hi = mid;
        }
        tp = lo < ntouch && touch[lo].x == g.t && touch[lo].y == g.s;
    }
    if (!tp && !force_walk) return -1;  // a plain near pair (near_thread / near_group take it)
    Piece r0;
    set_corners(r0.b);
    r0.frac = 1.0;
    r0.kind = 0;
    r0.level = 0;
    r0.depth = g.near_depth + (tp ? 4 : 0);
    walk(g, r0);
    return T_NEAR;
}

constexpr int NG_MAXDEPTH = 5;  // deepest near rule the parallel-tree kernel (near_group) takes

// One thread per non-far pair, a router that decides exactly as pair_rule / segment_rule do (assemble.jl:266-306,
// after the far test) but builds no rule: pairs whose rule needs the piece recursion (contact per contact_rule's
// test, geometric touch pairs, filaments, near rules deeper than near_group takes) go to the general touching list
// (touch_gen); the self pair and shared-edge / shared-corner pairs to the simple touching list as (t, s, rule code,
// tier); plain near pairs to near_thread when the root piece is not split (near_pieces!' first test, assemble.jl:134,
// on the whole triangle, whose corners, size and centroid are the element's own) and to near_group otherwise.
__global__ __launch_bounds__(256) void triage(AsmParams P, const Elem *__restrict__ el, const int *__restrict__ elem_nodes,
                                              const int2 *__restrict__ list, unsigned int npairs,
                                              const int2 *__restrict__ touch, int ntouch, int2 *__restrict__ touch_out,
                                              int2 *__restrict__ single_out, int2 *__restrict__ multi_out,
                                              int4 *__restrict__ simple_out, unsigned int *__restrict__ counts,
                                              unsigned int cap)
{
#pragma clang fp contract(off)
    for (unsigned int pi = blockIdx.x * blockDim.x + threadIdx.x; pi < npairs; pi += gridDim.x * blockDim.x) {
        const int2 pr = list[pi];
        const int t = pr.x, s = pr.y;
        int route = 0;  // 0 general touching, 1 single near piece, 2 several near pieces, 3 simple touching
        int code = 0, tier = T_SELF;
        if (s == t) {
            route = 3;
            code = R_SELF;
            tier = T_SELF;
        } else if (s < P.nsrc) {
            const Elem &T = el[t], &S = el[s];
            int sh[3];
            for (int a = 0; a < 3; a++) {
                int va = elem_nodes[3 * t + a];
                sh[a] = va == elem_nodes[3 * s] || va == elem_nodes[3 * s + 1] || va == elem_nodes[3 * s + 2];
            }
            const int ns = sh[0] + sh[1] + sh[2];
            bool contact = false;
            const V3<double> SX[3] = {ld3(S.v[0]), ld3(S.v[1]), ld3(S.v[2])};
            if (cls_contact_possible(ld3(S.c), ld3(T.c), S.reach, T.reach, S.diam, T.diam, P.contact_rel)) {
                const double tol = P.contact_rel * fmin(T.diam, S.diam);
                const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
                for (int k = 0; k < 3 && !contact; k++) contact = !sh[k] && cls_tri_distance(SX, X[k]) <= tol;
                for (int k = 0; k < 3 && !contact; k++) {  // splits_an_edge, assemble.jl:174-184
                    const V3<double> a = X[k], e = X[(k + 1) % 3];
                    const double L = cls_norm(cls_sub(e, a));
                    for (int q = 0; q < 3 && !contact; q++) {
                        double tq;
                        const double d = cls_seg_projection(SX[q], a, e, &tq);
                        contact = d <= tol && tq * L > tol && (1 - tq) * L > tol;
                    }
                }
            }
            if (contact) {
                route = 0;
            } else if (ns == 2) {
                route = 3;
                code = R_EDGE0 + (sh[0] && sh[1] ? 0 : sh[1] && sh[2] ? 1 : 2);
                tier = T_EDGE;
            } else if (ns == 1) {
                route = 3;
                code = R_VERT0 + (sh[0] ? 0 : sh[1] ? 1 : 2);
                tier = T_VERTEX;
            } else {
                bool tp = false;
                if (ntouch && s < P.nt) {
                    int lo = 0, hi = ntouch;
                    while (lo < hi) {
                        int mid = (lo + hi) / 2;
                        int2 q = touch[mid];
                        if (q.x < t || (q.x == t && q.y < s)) lo = mid + 1; else hi = mid;
                    }
                    tp = lo < ntouch && touch[lo].x == t && touch[lo].y == s;
                }
                if (!tp && P.near_depth <= NG_MAXDEPTH) {
                    const bool split = P.near_depth > 0 && cls_tri_distance(SX, ld3(T.c)) < P.near_ratio * T.diam;
                    route = split ? 2 : 1;
                }
            }
        }
        // one counter update per wavefront and route (the routes' counters are shared by the whole grid)
        const int lane = threadIdx.x & (WAVE - 1);
        unsigned int pos = 0;
        for (int r = 0; r < 4; r++) {
            const unsigned long long mask = __ballot(route == r);
            if (!mask) continue;
            const int leader = __ffsll((long long)mask) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(&counts[r], (unsigned int)__popcll(mask));
            base = __shfl(base, leader);
            if (route == r) pos = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
        }
        if (pos < cap) {
            if (route == 3)
                simple_out[pos] = make_int4(t, s, code, tier);
            else
                (route == 0 ? touch_out : route == 1 ? single_out : multi_out)[pos] = pr;
        }
    }
}

// One thread per touching pair: its exact tier (pair_rule / segment_rule after the far test) and its rule pieces,
// counted first, then written contiguously into the leaf list at a base reserved with one atomic, with one group entry
// (first piece, pieces, points, tier).
__global__ __launch_bounds__(256) void touch_gen(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                 const int *__restrict__ elem_nodes, const double *__restrict__ seg,
                                                 const int2 *__restrict__ list, unsigned int npairs,
                                                 const int2 *__restrict__ touch, int ntouch, Leaf *__restrict__ leaves,
                                                 unsigned int leaf_cap, int4 *__restrict__ groups, unsigned int group_cap,
                                                 unsigned int *__restrict__ counts, unsigned long long *__restrict__ stats,
                                                 uint8_t *__restrict__ tiers)
{
    __shared__ unsigned long long bst[13];
    if (threadIdx.x < 13) bst[threadIdx.x] = 0;
    __syncthreads();
    for (unsigned int pi = blockIdx.x * blockDim.x + threadIdx.x; pi < npairs; pi += gridDim.x * blockDim.x) {
        const int2 pr = list[pi];
        Gen g;
        g.t = pr.x;
        g.s = pr.y;
        g.seg = pr.y >= P.nsrc;
        g.near_ratio = P.near_ratio;
        g.corner_angle = P.corner_angle;
        g.near_depth = P.near_depth;
        g.rn = rt.n;
        g.tol = 0.0;
        const Elem &T = el[g.t];
        for (int k = 0; k < 3; k++) g.X[k] = ld3(T.v[k]);
        if (g.seg) {
            const int f = g.s - P.nsrc;
            g.fa = ld3(seg + 6 * f);
            g.fb = ld3(seg + 6 * f + 3);
        } else {
            for (int k = 0; k < 3; k++) g.SX[k] = ld3(el[g.s].v[k]);
        }
        g.out = nullptr;
        g.n = 0;
        g.npts = 0;
        g.overflow = false;
        const int tier = classify_and_generate(g, P, el, elem_nodes, touch, ntouch, true);
        const unsigned int n = g.n;
        const long long npts = g.npts;
        g.base = atomicAdd(&counts[0], n);
        const unsigned int gi = atomicAdd(&counts[1], 1u);
        g.cap = leaf_cap;
        g.out = leaves;
        g.n = 0;
        g.npts = 0;
        (void)classify_and_generate(g, P, el, elem_nodes, touch, ntouch, true);
        if (gi < group_cap) groups[gi] = make_int4((int)g.base, (int)n, (int)npts, tier);
        atomicAdd(&bst[tier], 1ull);
        atomicAdd(&bst[6 + tier], (unsigned long long)npts);
        if (g.overflow) atomicAdd(&bst[12], 1ull);
        if (tiers && !g.seg) tiers[(size_t)g.t * P.nsrc + g.s] = (uint8_t)tier;
    }
    __syncthreads();
    if (threadIdx.x < 12 && bst[threadIdx.x]) atomicAdd(&stats[threadIdx.x], bst[threadIdx.x]);
    if (threadIdx.x == 12 && bst[12]) atomicAdd(&stats[13], bst[12]);
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One wavefront per pair group: the points of all the pair's rule pieces flattened over the 64 lanes (a lane walks
// the pieces with a cursor), in FP64 (pair_block, assemble.jl:245-258, or segment_block, :312-321), reduced over the
// wave and added to the pair's block of A and its rows of b.  WPS: the occupancy cap (waves per SIMD), chosen by the
// launch autotuner (three measured best on MI355X: C2 2.12 -> 1.73 ms against two; a few registers spill).
template <int WPS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPS, WPS))) void group_eval(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                  const double *__restrict__ sigma, const int *__restrict__ corner_slot,
                                                  const int *__restrict__ src_cols, const double *__restrict__ seg,
                                                  const int *__restrict__ segcol, const Leaf *__restrict__ leaves,
                                                  const int4 *__restrict__ groups, unsigned int ngroups,
                                                  double *__restrict__ A, double *__restrict__ bvec)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (unsigned int gi = blockIdx.x * 4 + wave; gi < ngroups; gi += gridDim.x * 4) {
        const int4 grp = groups[gi];
        const int t = leaves[grp.x].t, s = leaves[grp.x].s, ntot = grp.z;
        const bool isseg = s >= P.nsrc;
        const Elem &T = el[t];
        const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        const V3<double> tn = ld3(T.n);
        double acc[12];
        for (int i = 0; i < 12; i++) acc[i] = 0.0;
        int li = grp.x, lbase = 0;
#pragma unroll 1
        for (int pp = lane; pp < ntot; pp += WAVE) {
            while (pp >= lbase + leaves[li].npt) {
                lbase += leaves[li].npt;
                li++;
            }
            const Leaf &lf = leaves[li];
            const int p = pp - lbase, off = rt.off[lf.rid];
            const double *m = rt.l + 3 * (off + p);
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * lf.b[0][j] + m[1] * lf.b[1][j]) + m[2] * lf.b[2][j];
            const double W = rt.w[off + p] * lf.frac * T.area;
            double psi[3];
            test_weights(P.test, l, psi);
            const V3<double> x = cls_point(X, l);
            if (isseg) {
                const int f = s - P.nsrc;
                V3<double> va, vb;
                seg_vel(x - ld3(seg + 6 * f), x - ld3(seg + 6 * f + 3), va, vb);
                const double da = dot(va, tn), db = dot(vb, tn);
                for (int a = 0; a < 3; a++) {
                    acc[a] += da * (W * psi[a]);
                    acc[3 + a] += db * (W * psi[a]);
                }
                continue;
            }
            const Elem &S = el[s];
            const V3<double> sc = ld3(S.c);
            const V3<double> r = x - sc;
            const double fd = P.kernel_far * S.diam;
            V3<double> w;
            if (cls_dot(r, r) > fd * fd) {
                SrcT<double> sr;
                for (int k = 0; k < 3; k++) sr.v[k] = ld3(S.v[k]) - sc;
                sr.area = S.area;
                w = sv_far_rel(sr, r);
            } else {
                SrcT<double> sr;
                for (int k = 0; k < 3; k++) {
                    sr.nu[k] = ld3(S.nu[k]);
                    sr.len[k] = S.len[k];
                }
                sr.n = ld3(S.n);
                sr.diam = S.diam;
                w = sv_exact_abs(sr, x - ld3(S.v[0]), x - ld3(S.v[1]), x - ld3(S.v[2]), cls_dot(r, sr.n), s == t ? 1 : 0,
                                 &S, x);
            }
            const V3<double> v = cross(w, tn);
            const double c0 = dot(v, ld3(S.g[0])), c1 = dot(v, ld3(S.g[1])), c2 = dot(v, ld3(S.g[2]));
            const double wn = (s < P.nt ? sigma[s] : 0.0) * dot(w, tn);
            for (int a = 0; a < 3; a++) {
                const double ps = W * psi[a];
                acc[3 * a] += ps * c0;
                acc[3 * a + 1] += ps * c1;
                acc[3 * a + 2] += ps * c2;
                acc[9 + a] -= wn * ps;
            }
        }
        double tot[12];
        for (int i = 0; i < 12; i++) tot[i] = wave_sum(acc[i]);
        if (lane == 0) {
            const int *ct = corner_slot + 3 * t;
            if (isseg) {
                const int f = s - P.nsrc;
                for (int a = 0; a < 3; a++) {
                    unsafeAtomicAdd(&A[(long long)segcol[2 * f] * P.lda + ct[a]], tot[a]);
                    unsafeAtomicAdd(&A[(long long)segcol[2 * f + 1] * P.lda + ct[a]], tot[3 + a]);
                }
            } else {
                const int *cs = src_cols + 3 * s;
                for (int a = 0; a < 3; a++)
                    for (int bb = 0; bb < 3; bb++) unsafeAtomicAdd(&A[(long long)cs[bb] * P.lda + ct[a]], tot[3 * a + bb]);
                if (s < P.nt)
                    for (int a = 0; a < 3; a++) unsafeAtomicAdd(&bvec[ct[a]], tot[9 + a]);
            }
        }
    }
}

// The self pair and shared-edge / shared-corner pairs (no contact): one wavefront per pair.  Its pieces are the
// stored rule on the whole triangle (self, edge k, vertex k when the corner is at most 90 degrees) or vertex_pieces!'
// equal-angle wedges of a wide corner (assemble.jl:102-125), kept in LDS; their points are flattened over the lanes.
// The forward kernel's occupancy cap WPS is chosen by the launch autotuner (three waves per SIMD measured best on
// MI355X: C2 10.2 -> 7.9 ms against the two its 249 registers give); the backward runs at one.
constexpr int MAXWEDGE = 16;

__device__ __forceinline__ double wave_sum2(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int NADJ, int WPS = 3>  // NADJ 0: the forward assembly; 1..ADJ_MAX: backward with that many weight sets
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NADJ == 0 ? WPS : 1, NADJ == 0 ? WPS : 1))) void touch_simple(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                    const double *__restrict__ sigma, const int *__restrict__ corner_slot,
                                                    const int *__restrict__ src_cols, const int4 *__restrict__ list,
                                                    unsigned int npairs, double *__restrict__ A, double *__restrict__ bvec,
                                                    unsigned long long *__restrict__ stats, uint8_t *__restrict__ tiers,
                                                    AdjArgs aa)
{
    __shared__ double wedge[4][MAXWEDGE][10];
    __shared__ unsigned long long wst[12];
    const int lane = threadIdx.x & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (threadIdx.x < 12) wst[threadIdx.x] = 0;
    double bdsum[NADJ > 0 ? NADJ : 1] = {};
    __syncthreads();
    for (unsigned int pi = blockIdx.x * 4 + wave; pi < npairs; pi += gridDim.x * 4) {
        const int4 e = list[pi];
        const int t = e.x, s = e.y, tier = e.w;
        int rid = e.z, npc = 1;
        const Elem &T = el[t], &S = el[s];
        const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        wave_lds_sync();  // the previous pair's reads of wedge are done before it is rewritten
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) wedge[wave][0][3 * i + j] = i == j ? 1.0 : 0.0;
        wedge[wave][0][9] = 1.0;
        if (tier == T_VERTEX) {
#pragma clang fp contract(off)
            const int k = rid - R_VERT0, i = (k + 1) % 3, j = (k + 2) % 3;
            const V3<double> u = cls_sub(X[i], X[k]), v = cls_sub(X[j], X[k]);
            if (cls_dot(u, v) < 0) {  // vertex_rule's wide corner: vertex_pieces! with the base rule
                rid = R_VERT0;
                const double nu = cls_norm(u), nv = cls_norm(v);
                double cs = cls_dot(u, v) / (nu * nv);
                cs = cs < -1.0 ? -1.0 : cs > 1.0 ? 1.0 : cs;
                const double th = acos(cs);
                double b[3][3];
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) b[a][c] = a == c ? 1.0 : 0.0;
                if (th <= kPi / 2) {
                    for (int c = 0; c < 3; c++) {
                        wedge[wave][0][c] = b[k][c];
                        wedge[wave][0][3 + c] = b[i][c];
                        wedge[wave][0][6 + c] = b[j][c];
                    }
                } else {
                    int m = (int)ceil(th * (180 / kPi) / P.corner_angle);
                    m = m < MAXWEDGE ? m : MAXWEDGE;
                    double t0 = 0.0;
                    for (int q = 1; q <= m; q++) {
                        double t1 = 1.0;
                        if (q < m) {
                            const double ph = th * q / m;
                            t1 = nu * sin(ph) / (nu * sin(ph) + nv * sin(th - ph));
                        }
                        for (int c = 0; c < 3; c++) {
                            wedge[wave][q - 1][c] = b[k][c];
                            wedge[wave][q - 1][3 + c] = (1 - t0) * b[i][c] + t0 * b[j][c];
                            wedge[wave][q - 1][6 + c] = (1 - t1) * b[i][c] + t1 * b[j][c];
                        }
                        wedge[wave][q - 1][9] = t1 - t0;
                        t0 = t1;
                    }
                    npc = m;
                }
            }
        }
        wave_lds_sync();  // the pieces' corners, written by every lane, read at other lanes' points below
        const int nr = rt.n[rid], off = rt.off[rid], ntot = npc * nr;
        const int side = s == t ? 1 : 0;
        const V3<double> tn = ld3(T.n), sc = ld3(S.c);
        const double fd = P.kernel_far * S.diam;
        if constexpr (NADJ > 0) {  // backward assembly: the same pieces, the pair's gradient sums
            AdjPair<NADJ> ap;
            adj_pair_begin(aa, P.nt, S, s, corner_slot + 3 * t, src_cols + 3 * s, tn, ap);
#pragma unroll 1
            for (int pp = lane; pp < ntot; pp += WAVE) {
                const int pc = pp / nr, q = pp - pc * nr;
                const double *bl = wedge[wave][pc];
                const double *m = rt.l + 3 * (off + q);
                double l[3];
                for (int j = 0; j < 3; j++) l[j] = (m[0] * bl[j] + m[1] * bl[3 + j]) + m[2] * bl[6 + j];
                adj_point(ap, X, S, side, P.kernel_far, rt.w[off + q] * bl[9], l, P.test);
            }
            for (int l = 0; l < NADJ; l++) adj_wave_reduce(ap.p[l]);
            if (lane == 0) adj_pair_end(aa, P.nt, T, X, t, s, aa.nodes, ap, S, bdsum);
            continue;
        }
        double acc[12];
        for (int q = 0; q < 12; q++) acc[q] = 0.0;
#pragma unroll 1
        for (int pp = lane; pp < ntot; pp += WAVE) {
            const int pc = pp / nr, q = pp - pc * nr;
            const double *bl = wedge[wave][pc];
            const double *m = rt.l + 3 * (off + q);
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * bl[j] + m[1] * bl[3 + j]) + m[2] * bl[6 + j];
            const double W = rt.w[off + q] * bl[9] * T.area;
            double psi[3];
            test_weights(P.test, l, psi);
            const V3<double> x = cls_point(X, l);
            const V3<double> r = x - sc;
            V3<double> w;
            if (cls_dot(r, r) > fd * fd) {
                SrcT<double> sr;
                for (int k = 0; k < 3; k++) sr.v[k] = ld3(S.v[k]) - sc;
                sr.area = S.area;
                w = sv_far_rel(sr, r);
            } else {
                SrcT<double> sr;
                for (int k = 0; k < 3; k++) {
                    sr.nu[k] = ld3(S.nu[k]);
                    sr.len[k] = S.len[k];
                }
                sr.n = ld3(S.n);
                sr.diam = S.diam;
                w = sv_exact_abs(sr, x - ld3(S.v[0]), x - ld3(S.v[1]), x - ld3(S.v[2]), cls_dot(r, sr.n), side,
                                 side ? &S : nullptr, x);
            }
            const V3<double> v = cross(w, tn);
            const double wn = dot(w, tn);
            for (int a = 0; a < 3; a++) {
                const double ps = W * psi[a];
                acc[3 * a] += ps * v.x;
                acc[3 * a + 1] += ps * v.y;
                acc[3 * a + 2] += ps * v.z;
                acc[9 + a] += ps * wn;
            }
        }
        double tot[12];
        for (int q = 0; q < 12; q++) tot[q] = wave_sum2(acc[q]);
        if (lane == 0) {
            const int *ct = corner_slot + 3 * t;
            const int *cs = src_cols + 3 * s;
            const double sig = s < P.nt ? sigma[s] : 0.0;
            for (int a = 0; a < 3; a++) {
                const V3<double> va = mk3(tot[3 * a], tot[3 * a + 1], tot[3 * a + 2]);
                for (int bb = 0; bb < 3; bb++) unsafeAtomicAdd(&A[(long long)cs[bb] * P.lda + ct[a]], dot(va, ld3(S.g[bb])));
                if (s < P.nt) unsafeAtomicAdd(&bvec[ct[a]], -sig * tot[9 + a]);
            }
            if (tiers) tiers[(size_t)t * P.nsrc + s] = (uint8_t)tier;
            atomicAdd(&wst[tier], 1ull);
            atomicAdd(&wst[6 + tier], (unsigned long long)ntot);
        }
    }
    __syncthreads();
    if constexpr (NADJ > 0) block_add_n<NADJ>(bdsum, aa.bdot);
    if (NADJ == 0 && threadIdx.x < 12 && wst[threadIdx.x]) atomicAdd(&stats[threadIdx.x], wst[threadIdx.x]);
}

// One point of a plain near pair's rule: accumulates W psi_a (w x n_T) (acc[3a..3a+2]) and W psi_a (w . n_T)
// (acc[9+a]); the sheets gamma_b and the source strength are applied once per pair (pair_block, assemble.jl:245-258).
// T float: the mixed path (multipole and closed form in FP32, the sums in FP64); the rule's points stay FP64.
template <typename T = double>
__device__ __forceinline__ void near_point(const V3<double> *X, const V3<double> *SX, const Elem &S, V3<double> tn,
                                           V3<double> sc, double fd2, double W, const double *l, int test, double *acc)
{
    double psi[3];
    test_weights(test, l, psi);
    const V3<double> x = cls_point(X, l);
    const V3<double> r = x - sc;
    V3<double> w;
    if constexpr (std::is_same<T, float>::value) {
        V3<float> wf;
        if (cls_dot(r, r) > fd2) {
            SrcT<float> sr;
            for (int k = 0; k < 3; k++) {
                const V3<double> d = SX[k] - sc;
                sr.v[k] = mk3((float)d.x, (float)d.y, (float)d.z);
            }
            sr.area = (float)S.area;
            wf = sv_far_rel(sr, mk3((float)r.x, (float)r.y, (float)r.z));
        } else {
            wf = sv_exact_f32(S, SX, x, dot(r, ld3(S.n)));
        }
        w = mk3((double)wf.x, (double)wf.y, (double)wf.z);
    } else if (cls_dot(r, r) > fd2) {
        SrcT<double> sr;
        for (int k = 0; k < 3; k++) sr.v[k] = SX[k] - sc;
        sr.area = S.area;
        w = sv_far_rel(sr, r);
    } else {
        SrcT<double> sr;
        for (int k = 0; k < 3; k++) {
            sr.nu[k] = ld3(S.nu[k]);
            sr.len[k] = S.len[k];
        }
        sr.n = ld3(S.n);
        sr.diam = S.diam;
        w = sv_exact_abs(sr, x - SX[0], x - SX[1], x - SX[2], cls_dot(r, sr.n), 0, nullptr, x);
    }
    const V3<double> v = cross(w, tn);
    const double wn = dot(w, tn);
    for (int a = 0; a < 3; a++) {
        const double ps = W * psi[a];
        acc[3 * a] += ps * v.x;
        acc[3 * a + 1] += ps * v.y;
        acc[3 * a + 2] += ps * v.z;
        acc[9 + a] += ps * wn;
    }
}

// One of the twelve adds of near_scatter (entries 0-8: block row a = e / 3 and source column b = e % 3; 9-11: right
// hand side row e - 9), so that a pair's adds can be spread over lanes that all hold its sums: e differs from lane to
// lane, so the sums are picked with constant indices (a register array indexed by e would go to scratch memory).
__device__ __forceinline__ void near_scatter_entry(const AsmParams &P, int e, int t, int s, const Elem &S,
                                                   const double *tot, const double *sigma, const int *corner_slot,
                                                   const int *src_cols, double *A, double *bvec)
{
    const int *ct = corner_slot + 3 * t;
    if (e < 9) {
        const int a = e / 3, bb = e - 3 * a;
        V3<double> va = mk3(tot[0], tot[1], tot[2]);
        if (a == 1) va = mk3(tot[3], tot[4], tot[5]);
        if (a == 2) va = mk3(tot[6], tot[7], tot[8]);
        unsafeAtomicAdd(&A[(long long)src_cols[3 * s + bb] * P.lda + ct[a]], dot(va, ld3(S.g[bb])));
    } else if (s < P.nt) {
        const double w = e == 9 ? tot[9] : e == 10 ? tot[10] : tot[11];
        unsafeAtomicAdd(&bvec[ct[e - 9]], -sigma[s] * w);
    }
}

// The pair's block from the accumulated sums: A[ct_a, cs_b] += (sum W psi_a w x n) . gamma_b, b[ct_a] -= sigma sum ...
__device__ __forceinline__ void near_scatter(const AsmParams &P, int t, int s, const Elem &S, const double *tot,
                                             const double *sigma, const int *corner_slot, const int *src_cols, double *A,
                                             double *bvec)
{
    const int *ct = corner_slot + 3 * t;
    const int *cs = src_cols + 3 * s;
    const double sig = s < P.nt ? sigma[s] : 0.0;
    for (int a = 0; a < 3; a++) {
        const V3<double> va = mk3(tot[3 * a], tot[3 * a + 1], tot[3 * a + 2]);
        for (int bb = 0; bb < 3; bb++) unsafeAtomicAdd(&A[(long long)cs[bb] * P.lda + ct[a]], dot(va, ld3(S.g[bb])));
        if (s < P.nt) unsafeAtomicAdd(&bvec[ct[a]], -sig * tot[9 + a]);
    }
}

// Plain near pairs whose near rule is the single root piece (near_rule, assemble.jl:69-73), forward: one lane per
// rule point, WAVE / n pairs of n points per wavefront (nine of the default seven).  The lanes of a pair are summed in
// point order (the reference's accumulation order) by shuffles, and the pair's first lane scatters its block.  WPS:
// the occupancy cap, chosen by the launch autotuner.  CT: the points' compute type (float on the mixed path).
template <int WPS, typename CT = double>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPS, WPS))) void near_points(
    AsmParams P, RuleTab rt, const Elem *__restrict__ el, const double *__restrict__ sigma,
    const int *__restrict__ corner_slot, const int *__restrict__ src_cols, const int2 *__restrict__ list,
    unsigned int npairs, double *__restrict__ A, double *__restrict__ bvec, unsigned long long *__restrict__ stats,
    uint8_t *__restrict__ tiers)
{
    __shared__ unsigned long long wst;
    if (threadIdx.x == 0) wst = 0;
    __syncthreads();
    const int nr = rt.n[R_NEAR], off = rt.off[R_NEAR];
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    const int per = WAVE / nr, slot = lane / nr, q = lane - slot * nr;
    unsigned long long mypairs = 0;
    const unsigned long long stride = (unsigned long long)gridDim.x * 4 * per;
    for (unsigned long long base = ((unsigned long long)blockIdx.x * 4 + wave) * per; base < npairs; base += stride) {
        const unsigned long long pi = base + slot;
        const bool live = slot < per && pi < npairs;
        double acc[12];
        for (int i = 0; i < 12; i++) acc[i] = 0.0;
        int t = 0, s = 0;
        if (live) {
            const int2 pr = list[pi];
            t = pr.x;
            s = pr.y;
            const Elem &T = el[t], &S = el[s];
            const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
            const V3<double> SX[3] = {ld3(S.v[0]), ld3(S.v[1]), ld3(S.v[2])};
            const double fd = P.kernel_far * S.diam;
            near_point<CT>(X, SX, S, ld3(T.n), ld3(S.c), fd * fd, rt.w[off + q] * T.area, rt.l + 3 * (off + q), P.test, acc);
        }
        double tot[12];
        const int first = (slot < per ? slot : 0) * nr;
        for (int i = 0; i < 12; i++) {
            double sum = 0.0;
            for (int k = 0; k < nr; k++) sum += __shfl(acc[i], first + k);
            tot[i] = sum;
        }
        if (live) {
            // the pair's twelve adds spread over its lanes (entries q and q + nr): two atomic instructions per wave
            // instead of twelve from one lane per pair (the adds took 1.8 of the kernel's 2.8 ms on the abused
            // ellipsoid; spread, it runs 2.54 ms, with the lane-indexed sums staged in LDS by the compiler)
            for (int e = q; e < 12; e += nr) near_scatter_entry(P, e, t, s, el[s], tot, sigma, corner_slot, src_cols, A, bvec);
            if (q == 0) {
                if (tiers) tiers[(size_t)t * P.nsrc + s] = (uint8_t)T_NEAR;
                mypairs++;
            }
        }
    }
    if (mypairs) atomicAdd(&wst, mypairs);
    __syncthreads();
    if (threadIdx.x == 0 && wst) {
        atomicAdd(&stats[T_NEAR], wst);
        atomicAdd(&stats[6 + T_NEAR], wst * (unsigned long long)nr);
    }
}

// Plain near pairs whose near rule is the single root piece, backward: one thread per pair, the rule's points in
// turn.
template <int NADJ, typename CT = double>  // NADJ 0: the forward assembly; 1..ADJ_MAX: backward with that many weight sets
__global__ __launch_bounds__(256) void near_thread(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                   const double *__restrict__ sigma, const int *__restrict__ corner_slot,
                                                   const int *__restrict__ src_cols, const int2 *__restrict__ list,
                                                   unsigned int npairs, double *__restrict__ A, double *__restrict__ bvec,
                                                   unsigned long long *__restrict__ stats, uint8_t *__restrict__ tiers,
                                                   AdjArgs aa)
{
    __shared__ unsigned long long wst;
    if (threadIdx.x == 0) wst = 0;
    __syncthreads();
    const int nr = rt.n[R_NEAR], off = rt.off[R_NEAR];
    unsigned long long mypairs = 0;
    double bdsum[NADJ > 0 ? NADJ : 1] = {};
    for (unsigned int pi = blockIdx.x * blockDim.x + threadIdx.x; pi < npairs; pi += gridDim.x * blockDim.x) {
        const int2 pr = list[pi];
        const int t = pr.x, s = pr.y;
        const Elem &T = el[t], &S = el[s];
        const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        const V3<double> SX[3] = {ld3(S.v[0]), ld3(S.v[1]), ld3(S.v[2])};
        const double fd = P.kernel_far * S.diam;
        if constexpr (NADJ > 0) {
            AdjPair<NADJ> ap;
            adj_pair_begin(aa, P.nt, S, s, corner_slot + 3 * t, src_cols + 3 * s, ld3(T.n), ap);
#pragma unroll 1
            for (int q = 0; q < nr; q++)
                adj_point<NADJ, CT>(ap, X, S, 0, P.kernel_far, rt.w[off + q], rt.l + 3 * (off + q), P.test);
            adj_pair_end(aa, P.nt, T, X, t, s, aa.nodes, ap, S, bdsum);
            continue;
        }
        double acc[12];
        for (int i = 0; i < 12; i++) acc[i] = 0.0;
#pragma unroll 1
        for (int q = 0; q < nr; q++)
            near_point(X, SX, S, ld3(T.n), ld3(S.c), fd * fd, rt.w[off + q] * T.area, rt.l + 3 * (off + q), P.test, acc);
        near_scatter(P, t, s, S, acc, sigma, corner_slot, src_cols, A, bvec);
        if (tiers) tiers[(size_t)t * P.nsrc + s] = (uint8_t)T_NEAR;
        mypairs++;
    }
    if (mypairs) atomicAdd(&wst, mypairs);
    __syncthreads();
    if constexpr (NADJ > 0) block_add_n<NADJ>(bdsum, aa.bdot);
    if (NADJ == 0 && threadIdx.x == 0 && wst) {
        atomicAdd(&stats[T_NEAR], wst);
        atomicAdd(&stats[6 + T_NEAR], wst * (unsigned long long)nr);
    }
}

// Corners of quadtree node (level L, index idx: its base-4 digits, most significant first, are the children taken
// from the root), by the same quarters in the same order as near_pieces! (assemble.jl:79-82, 131-142).
__device__ __forceinline__ void node_corners(int L, int idx, double (*b)[3])
{
#pragma clang fp contract(off)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) b[i][j] = i == j ? 1.0 : 0.0;
    for (int i = 0; i < L; i++) {
        const int c = (idx >> (2 * (L - 1 - i))) & 3;
        double m12[3], m23[3], m31[3];
        for (int j = 0; j < 3; j++) {
            m12[j] = (b[0][j] + b[1][j]) / 2;
            m23[j] = (b[1][j] + b[2][j]) / 2;
            m31[j] = (b[2][j] + b[0][j]) / 2;
        }
        for (int j = 0; j < 3; j++) {
            const double b0 = b[0][j], b1 = b[1][j], b2 = b[2][j];
            b[0][j] = c == 0 ? b0 : c == 1 ? m12[j] : c == 2 ? m31[j] : m12[j];
            b[1][j] = c == 0 ? m12[j] : c == 1 ? b1 : c == 2 ? m23[j] : m23[j];
            b[2][j] = c == 0 ? m31[j] : c == 1 ? m23[j] : c == 2 ? b2 : m31[j];
        }
    }
}

__device__ __forceinline__ void node_of(int nid, int &L, int &idx)  // level-order id -> (level, index)
{
    L = 0;
    int off = 0, cnt = 1;
    while (nid >= off + cnt) {
        off += cnt;
        cnt *= 4;
        L++;
    }
    idx = nid - off;
}

// Plain near pairs whose root piece splits: one wavefront per pair evaluates the split test of every node of the
// full quadtree at once across its lanes (each test depends on that node's own corners only, so this is the
// recursive near_pieces! exactly), collects the leaves -- nodes whose ancestors all split and that do not split or
// reach the depth -- by ballot, and evaluates their 7-point rules flattened over the lanes.  The forward occupancy cap
// WPS is chosen by the launch autotuner (three waves per SIMD measured best: C2 7.77 -> 6.50 ms against two); the
// backward is capped at two (faster than the register allocation the compiler picks by itself).
// The backward builds are capped at two waves per SIMD, or at one with WPS = 1 (all weight sets in one launch).
// CT: the forward points' compute type (float on the mixed path).
template <int NADJ, int WPS = 3, typename CT = double>  // NADJ 0: the forward assembly; 1..ADJ_MAX: backward with that many weight sets
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NADJ == 0 ? WPS : (WPS == 1 ? 1 : 2), NADJ == 0 ? WPS : (WPS == 1 ? 1 : 2)))) void near_group(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                  const double *__restrict__ sigma, const int *__restrict__ corner_slot,
                                                  const int *__restrict__ src_cols, const int2 *__restrict__ list,
                                                  unsigned int npairs, double *__restrict__ A, double *__restrict__ bvec,
                                                  unsigned long long *__restrict__ stats, uint8_t *__restrict__ tiers,
                                                  AdjArgs aa)
{
    constexpr int MAXINT = (1 << (2 * NG_MAXDEPTH)) / 3 + 1;  // internal nodes of a depth-5 tree (341)
    constexpr int MAXLEAF = 1 << (2 * NG_MAXDEPTH);           // 1024
    constexpr int LB = WAVE;  // leaves whose corners and area fraction are kept in LDS for the point loop
    constexpr int PBMAX = 4;  // pairs whose split tests share the lanes
    __shared__ unsigned long long splitw[4][(MAXINT + 63) / 64];
    __shared__ unsigned short leaf[4][MAXLEAF];
    __shared__ double leafb[4][LB][10];
    __shared__ unsigned long long wst[2];
    const int lane = threadIdx.x & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (threadIdx.x < 2) wst[threadIdx.x] = 0;
    __syncthreads();
    const int depth = P.near_depth;
    int nint = 0, ncand = 1;
    for (int Lv = 0, c = 1; Lv <= depth; Lv++, c *= 4) {
        if (Lv < depth) nint += c;
        if (Lv > 0) ncand += c;
    }
    // a batch of PB pairs per wavefront: their split tests fill the lanes together when PB * nint <= 64
    const int PB = nint > 0 && nint <= WAVE / 2 ? min(PBMAX, WAVE / nint) : 1;
    const int nr = rt.n[R_NEAR], off = rt.off[R_NEAR];
    unsigned long long mypairs = 0, mypts = 0;
    double bdsum[NADJ > 0 ? NADJ : 1] = {};
    // split test of internal node nid of the near tree on test triangle X against source SX (near_pieces!,
    // assemble.jl:131-142; FP64 without contraction, as the reference's rule builder)
    auto split_test = [&](const V3<double> *X, const V3<double> *SX, int nid) {
#pragma clang fp contract(off)
        int L, idx;
        node_of(nid, L, idx);
        double b[3][3];
        node_corners(L, idx, b);
        const V3<double> p1 = cls_point(X, b[0]), p2 = cls_point(X, b[1]), p3 = cls_point(X, b[2]);
        const double size = fmax(fmax(cls_norm(cls_sub(p2, p1)), cls_norm(cls_sub(p3, p2))), cls_norm(cls_sub(p1, p3)));
        V3<double> cen = (p1 + p2) + p3;
        cen = mk3(cen.x / 3.0, cen.y / 3.0, cen.z / 3.0);
        return cls_tri_distance(SX, cen) < P.near_ratio * size;
    };
    for (unsigned int b0 = (blockIdx.x * 4 + wave) * PB; b0 < npairs; b0 += gridDim.x * 4 * PB) {
        // 1a. split tests of the batch's pairs at once (one bit word per pair)
        unsigned long long sw[PBMAX] = {0, 0, 0, 0};
        if (PB > 1) {
            const int bb = lane / nint, nid = lane - bb * nint;
            bool split = false;
            if (bb < PB && b0 + bb < npairs) {
                const int2 pb = list[b0 + bb];
                const Elem &Tb = el[pb.x], &Sb = el[pb.y];
                const V3<double> Xb[3] = {ld3(Tb.v[0]), ld3(Tb.v[1]), ld3(Tb.v[2])};
                const V3<double> SXb[3] = {ld3(Sb.v[0]), ld3(Sb.v[1]), ld3(Sb.v[2])};
                split = split_test(Xb, SXb, nid);
            }
            const unsigned long long m = __ballot(split);
            for (int k = 0; k < PB; k++) sw[k] = (m >> (k * nint)) & ((1ull << nint) - 1ull);
        }
        for (int bb = 0; bb < PB; bb++) {
        const unsigned int pi = b0 + bb;
        if (pi >= npairs) break;
        const int2 pr = list[pi];
        const int t = pr.x, s = pr.y;
        const Elem &T = el[t], &S = el[s];
        const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        const V3<double> SX[3] = {ld3(S.v[0]), ld3(S.v[1]), ld3(S.v[2])};
        wave_lds_sync();  // the previous pair's reads of splitw, leaf and leafb are done before they are rewritten
        // 1b. one pair's split tests over the lanes (deeper trees)
        if (PB > 1) {
            if (lane == 0) splitw[wave][0] = sw[bb];
        } else {
            for (int base = 0; base < nint; base += WAVE) {
                const int nid = base + lane;
                const bool split = nid < nint && split_test(X, SX, nid);
                const unsigned long long m = __ballot(split);
                if (lane == 0) splitw[wave][base / WAVE] = m;
            }
        }
        wave_lds_sync();  // lane 0's split words, read by every lane below
        // 2. leaves: every ancestor splits, the node itself does not (or is at the depth)
        int nl = 0;
        for (int base = 0; base < ncand; base += WAVE) {
            const int nid = base + lane;
            bool isleaf = false;
            if (nid < ncand) {
                int L, idx;
                node_of(nid, L, idx);
                bool reach = true;
                for (int Lu = L - 1, iu = idx >> 2, offu = 0; Lu >= 0 && reach; Lu--, iu >>= 2) {
                    offu = 0;
                    for (int k = 0, c = 1; k < Lu; k++, c *= 4) offu += c;
                    const int aid = offu + iu;
                    reach = (splitw[wave][aid >> 6] >> (aid & 63)) & 1ull;
                }
                const bool own = nid < nint ? ((splitw[wave][nid >> 6] >> (nid & 63)) & 1ull) != 0 : false;
                isleaf = reach && !own;
            }
            const unsigned long long m = __ballot(isleaf);
            if (isleaf) leaf[wave][nl + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned short)nid;
            nl += __popcll(m);
        }
        wave_lds_sync();  // the leaf list, written by the leaves' lanes and read by others
        // the first LB leaves' corners and area fractions, once per leaf for the point loop
        if (lane < nl && lane < LB) {
            int L, idx;
            node_of(leaf[wave][lane], L, idx);
            double b[3][3];
            node_corners(L, idx, b);
            double frac = 1.0;
            for (int i = 0; i < L; i++) frac /= 4;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) leafb[wave][lane][3 * i + j] = b[i][j];
            leafb[wave][lane][9] = frac;
        }
        wave_lds_sync();  // the leaves' corners, written by lanes below LB and read by every lane
        // leaf li's corners b and area fraction (LDS for the first LB leaves)
        auto leaf_geom = [&](int li, double (*b)[3], double &frac) {
            if (li < LB) {
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) b[i][j] = leafb[wave][li][3 * i + j];
                frac = leafb[wave][li][9];
            } else {
                int L, idx;
                node_of(leaf[wave][li], L, idx);
                node_corners(L, idx, b);
                frac = 1.0;
                for (int i = 0; i < L; i++) frac /= 4;
            }
        };
        // 3. the leaves' points
        const double fd = P.kernel_far * S.diam;
        if constexpr (NADJ > 0) {
            AdjPair<NADJ> ap;
            adj_pair_begin(aa, P.nt, S, s, corner_slot + 3 * t, src_cols + 3 * s, ld3(T.n), ap);
#pragma unroll 1
            for (int pp = lane; pp < nl * nr; pp += WAVE) {
                const int li = pp / nr, q = pp - li * nr;
                double b[3][3], frac;
                leaf_geom(li, b, frac);
                const double *m = rt.l + 3 * (off + q);
                double l[3];
                for (int j = 0; j < 3; j++) l[j] = (m[0] * b[0][j] + m[1] * b[1][j]) + m[2] * b[2][j];
                adj_point<NADJ, CT>(ap, X, S, 0, P.kernel_far, rt.w[off + q] * frac, l, P.test);
            }
            for (int l = 0; l < NADJ; l++) adj_wave_reduce(ap.p[l]);
            if (lane == 0) adj_pair_end(aa, P.nt, T, X, t, s, aa.nodes, ap, S, bdsum);
            continue;
        }
        double acc[12];
        for (int i = 0; i < 12; i++) acc[i] = 0.0;
#pragma unroll 1
        for (int pp = lane; pp < nl * nr; pp += WAVE) {
            const int li = pp / nr, q = pp - li * nr;
            double b[3][3], frac;
            leaf_geom(li, b, frac);
            const double *m = rt.l + 3 * (off + q);
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * b[0][j] + m[1] * b[1][j]) + m[2] * b[2][j];
            near_point<CT>(X, SX, S, ld3(T.n), ld3(S.c), fd * fd, rt.w[off + q] * frac * T.area, l, P.test, acc);
        }
        double tot[12];
        for (int i = 0; i < 12; i++) tot[i] = wave_sum(acc[i]);
        // one lane adds the pair's twelve entries: spread over twelve lanes, the lane-indexed sums went to LDS and
        // scratch and the kernel ran 5.50 -> 6.05 ms on the abused ellipsoid
        if (lane == 0) {
            near_scatter(P, t, s, S, tot, sigma, corner_slot, src_cols, A, bvec);
            if (tiers) tiers[(size_t)t * P.nsrc + s] = (uint8_t)T_NEAR;
        }
        mypairs++;
        mypts += (unsigned long long)nl * nr;
        }
    }
    if (lane == 0 && mypairs) {
        atomicAdd(&wst[0], mypairs);
        atomicAdd(&wst[1], mypts);
    }
    __syncthreads();
    if constexpr (NADJ > 0) block_add_n<NADJ>(bdsum, aa.bdot);
    if (NADJ == 0 && threadIdx.x == 0 && wst[0]) {
        atomicAdd(&stats[T_NEAR], wst[0]);
        atomicAdd(&stats[6 + T_NEAR], wst[1]);
    }
}

// freestream part of b (exact), src/assembly/assemble.jl:422
__global__ void freestream(const Elem *__restrict__ el, const int *__restrict__ corner_slot, int nt, int test,
                           double vx, double vy, double vz, double *__restrict__ bvec)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const Elem &E = el[t];
    double vn;
    {
#pragma clang fp contract(off)
        vn = vx * E.n[0] + vy * E.n[1] + vz * E.n[2];
    }
    double ti = test == 1 ? 1.0 : 1.0 / 3;
    for (int a = 0; a < 3; a++) unsafeAtomicAdd(&bvec[corner_slot[3 * t + a]], -(vn * E.area * ti));
}

// ------------------------------------------------------------------------------------------ host side --------------

namespace {

// Morton code of a point in the unit cube (10 bits per axis)
uint32_t morton3(double x, double y, double z)
{
    auto spread = [](uint32_t v) {
        v &= 0x3ff;
        v = (v | (v << 16)) & 0x030000FF;
        v = (v | (v << 8)) & 0x0300F00F;
        v = (v | (v << 4)) & 0x030C30C3;
        v = (v | (v << 2)) & 0x09249249;
        return v;
    };
    auto q = [](double u) { return (uint32_t)std::min(1023.0, std::max(0.0, u * 1023.0)); };
    return spread(q(x)) | (spread(q(y)) << 1) | (spread(q(z)) << 2);
}

// Greedy tiles along a space-filling order: up to `cap` elements and `lmax` distinct columns per tile.
void build_tiles(const std::vector<int> &order, const std::vector<int> &cols3, int cap, int lmax, std::vector<int> &elems,
                 std::vector<uint8_t> &loc, std::vector<int> &ucols, std::vector<int> &nu, int ucap)
{
    size_t i = 0;
    while (i < order.size()) {
        std::vector<int> uc;
        std::vector<int> tile;
        std::vector<uint8_t> tl;
        while (i < order.size() && (int)tile.size() < cap) {
            int e = order[i];
            int add = 0;
            int idx[3];
            for (int k = 0; k < 3; k++) {
                int col = cols3[3 * e + k];
                auto it = std::find(uc.begin(), uc.end(), col);
                if (it == uc.end()) {
                    bool dup = false;
                    for (int j = 0; j < k; j++) dup |= cols3[3 * e + j] == col;
                    if (!dup) add++;
                }
            }
            if ((int)uc.size() + add > lmax) break;
            for (int k = 0; k < 3; k++) {
                int col = cols3[3 * e + k];
                auto it = std::find(uc.begin(), uc.end(), col);
                if (it == uc.end()) {
                    uc.push_back(col);
                    idx[k] = (int)uc.size() - 1;
                } else {
                    idx[k] = (int)(it - uc.begin());
                }
            }
            tile.push_back(e);
            for (int k = 0; k < 3; k++) tl.push_back((uint8_t)idx[k]);
            i++;
        }
        // the tile's slots in ascending order, so that a far tile's flush walks each column of A in address order
        std::vector<int> perm(uc.size()), rank(uc.size());
        std::iota(perm.begin(), perm.end(), 0);
        std::sort(perm.begin(), perm.end(), [&](int a, int b) { return uc[a] < uc[b]; });
        for (size_t k = 0; k < perm.size(); k++) rank[perm[k]] = (int)k;
        std::vector<int> us(uc.size());
        for (size_t k = 0; k < perm.size(); k++) us[k] = uc[perm[k]];
        for (auto &x : tl) x = (uint8_t)rank[x];
        uc.swap(us);
        for (size_t k = 0; k < (size_t)cap; k++) {
            elems.push_back(k < tile.size() ? tile[k] : -1);
            for (int j = 0; j < 3; j++) loc.push_back(k < tile.size() ? tl[3 * k + j] : 0);
        }
        for (int k = 0; k < ucap; k++) ucols.push_back(k < (int)uc.size() ? uc[k] : -1);
        nu.push_back((int)uc.size());
    }
}

std::vector<int> space_order(const std::vector<double> &cent, int ne)
{
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int e = 0; e < ne; e++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], cent[3 * e + k]);
            hi[k] = std::max(hi[k], cent[3 * e + k]);
        }
    double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<std::pair<uint32_t, int>> key(ne);
    for (int e = 0; e < ne; e++)
        key[e] = {morton3((cent[3 * e] - lo[0]) / span, (cent[3 * e + 1] - lo[1]) / span, (cent[3 * e + 2] - lo[2]) / span), e};
    std::stable_sort(key.begin(), key.end());
    std::vector<int> order(ne);
    for (int e = 0; e < ne; e++) order[e] = key[e].second;
    return order;
}

}  // namespace

#define HIPCHECK(x)                                                                                                    \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));                \
    } while (0)

template <typename X> X *dalloc(size_t n)
{
    X *p = nullptr;
    if (n == 0) n = 1;
    HIPCHECK(hipMalloc((void **)&p, n * sizeof(X)));
    return p;
}

double *Plan::scratch_d(int slot, size_t n)
{
    if ((int)sbuf.size() <= slot) {
        sbuf.resize(slot + 1, nullptr);
        scap.resize(slot + 1, 0);
    }
    if (scap[slot] < n) {
        if (sbuf[slot]) {
            HIPCHECK(hipStreamSynchronize(ctx->stream));  // the last call's work on it is done
            (void)hipFree(sbuf[slot]);
            sbuf[slot] = nullptr;
        }
        n = std::max(n, scap[slot] + scap[slot] / 2);
        sbuf[slot] = dalloc<double>(n);
        scap[slot] = n;
    }
    return (double *)sbuf[slot];
}

template <typename X> X *dupload(const std::vector<X> &v, hipStream_t s)
{
    X *p = dalloc<X>(v.size());
    if (!v.empty()) HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, s));
    return p;
}

// ------------------------------------------------------------------------------------------ launch autotuner ------

// The forward evaluation kernels' occupancy cap (waves per SIMD: the register budget the compiler had) is chosen by
// timing.  Each candidate runs in turn on one plan's successive assemblies (the same lists, so the same work), timed
// with HIP events around its launch, and the fastest is kept for the process: the rest of that plan's assemblies and
// every later plan use it.  A launch shorter than kTuneMinMs is too short to tell the candidates apart: it runs the
// default and leaves the timing to a larger plan.  VP_ASM_WAVES=<3|2> forces a candidate.
namespace {
struct AsmTuneState {
    int choice[TK_COUNT] = {-1, -1, -1, -1};
    float ms[TK_COUNT][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};
    long long plan[TK_COUNT] = {0, 0, 0, 0};  // serial of the plan whose assemblies are compared
    int next[TK_COUNT] = {0, 0, 0, 0};        // the candidate its next assembly runs
};
AsmTuneState g_asm;
std::mutex g_asm_mutex;
long long g_plan_serial = 0;
constexpr float kTuneMinMs = 0.2f;
const char *kTuneNames[TK_COUNT] = {"touch_simple", "near_points", "near_group", "group_eval"};

int forced_waves()
{
    const char *env = std::getenv("VP_ASM_WAVES");
    if (!env) return -1;
    const int w = std::atoi(env);
    for (int v = 0; v < 2; v++)
        if (kTuneWaves[v] == w) return v;
    return -1;
}
}  // namespace

long long next_plan_serial()
{
    std::lock_guard<std::mutex> lock(g_asm_mutex);
    return ++g_plan_serial;
}

int Plan::tune_pick(int k)
{
    tune_timed[k] = false;
    const int f = forced_waves();
    if (f >= 0) return f;
    std::lock_guard<std::mutex> lock(g_asm_mutex);
    if (g_asm.choice[k] >= 0) return g_asm.choice[k];
    if (g_asm.plan[k] != 0 && g_asm.plan[k] != serial) return 0;  // another plan is being timed
    g_asm.plan[k] = serial;
    tune_timed[k] = true;
    return g_asm.next[k];
}

void Plan::tune_collect()
{
    for (int k = 0; k < TK_COUNT; k++) {
        if (!tune_timed[k]) continue;
        tune_timed[k] = false;
        float t = 0;
        // an assembly that stopped on an error may not have recorded the pair: no measurement then
        if (hipEventSynchronize(tev[k][1]) != hipSuccess || hipEventElapsedTime(&t, tev[k][0], tev[k][1]) != hipSuccess) {
            (void)hipGetLastError();
            tune_release();
            continue;
        }
        std::lock_guard<std::mutex> lock(g_asm_mutex);
        if (g_asm.choice[k] >= 0 || g_asm.plan[k] != serial) continue;
        const int v = g_asm.next[k];
        if (t < kTuneMinMs) {  // too short to tell the candidates apart: left to a larger plan
            g_asm.plan[k] = 0;
            g_asm.next[k] = 0;
            g_asm.ms[k][0] = g_asm.ms[k][1] = -1;
            continue;
        }
        g_asm.ms[k][v] = t;
        if (v + 1 < 2) {
            g_asm.next[k] = v + 1;
        } else {
            g_asm.choice[k] = g_asm.ms[k][1] < g_asm.ms[k][0] ? 1 : 0;
        }
    }
}

void Plan::tune_release()
{
    std::lock_guard<std::mutex> lock(g_asm_mutex);
    for (int k = 0; k < TK_COUNT; k++)
        if (g_asm.choice[k] < 0 && g_asm.plan[k] == serial) {  // unfinished: the next plan starts over
            g_asm.plan[k] = 0;
            g_asm.next[k] = 0;
            g_asm.ms[k][0] = g_asm.ms[k][1] = -1;
        }
}

std::string asm_tuning_report()
{
    std::lock_guard<std::mutex> lock(g_asm_mutex);
    const int f = forced_waves();
    std::string out = "{";
    for (int k = 0; k < TK_COUNT; k++) {
        const int c = f >= 0 ? f : g_asm.choice[k];
        out += std::string(k ? ", " : "") + "\"" + kTuneNames[k] + "\": {\"chosen_waves\": " +
               (c < 0 ? std::string("null") : std::to_string(kTuneWaves[c])) + ", \"ms\": {";
        for (int v = 0; v < 2; v++)
            out += std::string(v ? ", " : "") + "\"" + std::to_string(kTuneWaves[v]) + "\": " +
                   (g_asm.ms[k][v] < 0 ? std::string("null") : std::to_string(g_asm.ms[k][v]));
        out += "}}";
    }
    return out + (f >= 0 ? ", \"forced\": true}" : "}");
}

Plan::~Plan()
{
    tune_release();
    for (int k = 0; k < 6; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    for (int k = 0; k < TK_COUNT; k++)
        for (int j = 0; j < 2; j++)
            if (tev[k][j]) (void)hipEventDestroy(tev[k][j]);
    for (void *p : owned) (void)hipFree(p);
    for (void *p : sbuf)
        if (p) (void)hipFree(p);
    if (d_bnear) (void)hipFree(d_bnear);
    if (A && ownsA) (void)hipFree(A);
}

void Plan::build(Ctx *ctx_, const vp_surface &g, const vp_wake *wk, const vp_assembly_options &o)
{
    ctx = ctx_;
    hipStream_t st = ctx->stream;
    opts = o;
    {  // the far tier's form: one persistent work queue over every body's tiles (two workgroups per CU, the
       // residency at its two-wave cap; C2 far tier 14.56 -> 14.27 ms, three uneven bodies 22.1 -> 21.3 ms against
       // the grid), or with VP_FAR_QUEUE=0 a grid of tiles
        const char *q = std::getenv("VP_FAR_QUEUE");
        int ncu = 0;
        if (!(q && q[0] == '0') &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess)
            far_queue_wgs = std::max(ncu, 1);  // x FAR_Q_WAVES(T) at launch
    }
    if (o.test < 0 || o.test > 2) throw std::invalid_argument("test must be hat (0), onering (1) or dual (2)");
    if (o.precision != 0 && o.precision != 1) throw std::invalid_argument("precision must be 0 (fp64) or 1 (mixed)");
    if (g.nt <= 0 || g.nn <= 0 || g.nslot <= 0) throw std::invalid_argument("empty surface");
    nt = g.nt;
    nn = g.nn;
    nslot = g.nslot;
    nw = wk ? wk->nw : 0;
    nwt = wk ? wk->ntri : 0;
    nseg = wk ? wk->nfil : 0;
    nsrc = nt + nwt;
    ncol = nslot + nw;
    lda = nslot;
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            int v = g.tris[3 * t + k], cs = g.corner_slot[3 * t + k];
            if (v < 0 || v >= nn) throw std::out_of_range("triangle node index out of range");
            if (cs < 0 || cs >= nslot) throw std::out_of_range("corner slot out of range");
        }
    for (int t = 0; t < nwt; t++)
        for (int k = 0; k < 3; k++)
            if (wk->tris[3 * t + k] < 0 || wk->tris[3 * t + k] >= nw) throw std::out_of_range("wake triangle index out of range");
    for (int f = 0; f < nseg; f++)
        for (int k = 0; k < 2; k++)
            if (wk->filaments[2 * f + k] < 0 || wk->filaments[2 * f + k] >= nw) throw std::out_of_range("filament index out of range");

    // rules (TierRules, src/elements/rules.jl:161-174)
    std::vector<HostRule> rules(R_COUNT);
    rules[R_FAR] = triangle_rule(o.far_points);
    rules[R_NEAR] = triangle_rule(o.near_points);
    if (rules[R_FAR].n() != 3) throw std::invalid_argument("the far tier kernel takes the 3-point rule (far_points = 3)");
    rules[R_SELF] = self_rule(o.touch_levels, o.touch_ratio, o.touch_order);
    HostRule e = edge_graded_rule(o.touch_levels, o.touch_ratio, o.touch_order);
    HostRule v = vertex_graded_rule(o.touch_levels, o.touch_ratio, o.touch_order);
    for (int k = 0; k < 3; k++) {
        rules[R_EDGE0 + k] = rotate_rule(e, k + 1);
        rules[R_VERT0 + k] = rotate_rule(v, k + 1);
    }
    std::vector<double> rl, rw;
    for (int r = 0; r < R_COUNT; r++) {
        rt.off[r] = (int)rw.size();
        rt.n[r] = rules[r].n();
        rl.insert(rl.end(), rules[r].l.begin(), rules[r].l.end());
        rw.insert(rw.end(), rules[r].w.begin(), rules[r].w.end());
    }
    double *dl = dupload(rl, st), *dw = dupload(rw, st);
    owned.push_back(dl);
    owned.push_back(dw);
    rt.l = dl;
    rt.w = dw;

    // elements: body then wake; node ids (wake offset by nn so no wake corner is shared with the body)
    std::vector<double> xall((size_t)3 * (nn + nw));
    std::memcpy(xall.data(), g.x, sizeof(double) * 3 * nn);
    if (nw) std::memcpy(xall.data() + 3 * nn, wk->x, sizeof(double) * 3 * nw);
    std::vector<int> en((size_t)3 * nsrc), cols((size_t)3 * nsrc);
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            en[3 * t + k] = g.tris[3 * t + k];
            cols[3 * t + k] = g.corner_slot[3 * t + k];
        }
    for (int t = 0; t < nwt; t++)
        for (int k = 0; k < 3; k++) {
            en[3 * (nt + t) + k] = nn + wk->tris[3 * t + k];
            cols[3 * (nt + t) + k] = nslot + wk->tris[3 * t + k];
        }
    double *dx = dupload(xall, st);
    d_x = dx;
    d_nodes = dupload(en, st);
    d_cols = dupload(cols, st);
    d_cs = dupload(std::vector<int>(g.corner_slot, g.corner_slot + 3 * nt), st);
    owned.insert(owned.end(), {dx, d_nodes, d_cols, d_cs});
    d_el = dalloc<Elem>(nsrc);
    owned.push_back(d_el);
    elem_setup<<<(nsrc + 255) / 256, 256, 0, st>>>(dx, d_nodes, nsrc, d_el);
    HIPCHECK(hipGetLastError());
    std::vector<Elem> hel(nsrc);
    HIPCHECK(hipMemcpyAsync(hel.data(), d_el, sizeof(Elem) * nsrc, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int e2 = 0; e2 < nsrc; e2++)
        if (hel[e2].pad0 != 0.0) throw std::invalid_argument("degenerate triangle " + std::to_string(e2));
    d_tf = dalloc<TestFar>(nt);
    owned.push_back(d_tf);
    test_setup<<<(nt + 255) / 256, 256, 0, st>>>(d_el, nt, rt, o.test, d_tf);
    HIPCHECK(hipGetLastError());

    // filaments
    std::vector<double> segx((size_t)6 * std::max(nseg, 1));
    std::vector<int> segc((size_t)2 * std::max(nseg, 1));
    for (int f = 0; f < nseg; f++) {
        int a = wk->filaments[2 * f], b = wk->filaments[2 * f + 1];
        for (int k = 0; k < 3; k++) {
            segx[6 * f + k] = wk->x[3 * a + k];
            segx[6 * f + 3 + k] = wk->x[3 * b + k];
        }
        segc[2 * f] = nslot + a;
        segc[2 * f + 1] = nslot + b;
    }
    d_seg = dupload(segx, st);
    d_segcol = dupload(segc, st);
    if (nseg) h_fil.assign(wk->filaments, wk->filaments + 2 * nseg);
    owned.insert(owned.end(), {d_seg, d_segcol});

    // touch pairs, sorted with both orders (assemble.jl:383-386)
    std::vector<int2> tp;
    for (int i = 0; i < g.ntouch; i++) {
        tp.push_back(make_int2(g.touch[2 * i], g.touch[2 * i + 1]));
        tp.push_back(make_int2(g.touch[2 * i + 1], g.touch[2 * i]));
    }
    std::sort(tp.begin(), tp.end(), [](int2 a, int2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
    ntouch = (int)tp.size();
    d_touch = dupload(tp, st);
    owned.push_back(d_touch);

    // tiles over a space-filling order of the centroids
    std::vector<double> cent((size_t)3 * nsrc);
    for (int e2 = 0; e2 < nsrc; e2++)
        for (int k = 0; k < 3; k++) cent[3 * e2 + k] = hel[e2].c[k];
    std::vector<double> tcent(cent.begin(), cent.begin() + 3 * nt);
    std::vector<int> torder = space_order(tcent, nt);
    std::vector<int> sorder_b = torder, sorder;
    sorder = sorder_b;
    if (nwt) {
        std::vector<double> wc(cent.begin() + 3 * nt, cent.end());
        std::vector<int> wo = space_order(wc, nwt);
        for (int x : wo) sorder.push_back(nt + x);
    }
    std::vector<int> tcols(g.corner_slot, g.corner_slot + 3 * nt);
    std::vector<int> tte, ttr, ttn, tse, tsc, tsn;
    std::vector<uint8_t> ttl, tsl;
    build_tiles(torder, tcols, TT_MAX, VMAX, tte, ttl, ttr, ttn, VMAX);
    // source tiles: the body's, then the wake's (a wake-only plan runs the far tier over the wake's tiles alone)
    build_tiles(sorder_b, cols, TS, UMAX, tse, tsl, tsc, tsn, UMAX);
    nts_body = (int)tsn.size();
    if (nwt) {
        std::vector<int> wo(sorder.begin() + nt, sorder.end());
        build_tiles(wo, cols, TS, UMAX, tse, tsl, tsc, tsn, UMAX);
    }
    ntt = (int)ttn.size();
    nts = (int)tsn.size();
    d_tte = dupload(tte, st);
    d_ttl = dupload(ttl, st);
    d_ttr = dupload(ttr, st);
    d_ttn = dupload(ttn, st);
    d_tse = dupload(tse, st);
    d_tsl = dupload(tsl, st);
    d_tsc = dupload(tsc, st);
    d_tsn = dupload(tsn, st);
    owned.insert(owned.end(), {d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl, d_tsc, d_tsn});

    // outputs and work lists
    A = dalloc<double>((size_t)lda * ncol);
    ownsA = true;
    d_b = dalloc<double>(nslot);
    d_sigma = dalloc<double>(nt);
    d_stats = dalloc<unsigned long long>(32);
    d_count = dalloc<unsigned int>(8);  // [7]: the far tier's work-queue counter
    owned.insert(owned.end(), {d_b, d_sigma, d_stats, d_count});
    // The non-far list's first budget, 384 pairs per test triangle (48 bytes each over the five lists; the abused
    // ellipsoid needs 244).  The first assembly grows it to an upper bound of the count (nonfar_bound) before its far
    // pass (the swept wing's long panels need ~1,130); later assemblies grow it on demand, refilling the lists by an
    // enumerate-only far pass.  VP_NONFAR_CAP sets it and skips the bound, for tests of the regrowth.
    nonfar_cap = (unsigned int)std::min<long long>((long long)nt * 384 + 1024, 1ll << 31);
    if (const char *cap = std::getenv("VP_NONFAR_CAP")) {
        nonfar_cap = (unsigned int)std::max(1ll, std::atoll(cap));
        nonfar_cap_forced = true;  // the tests of list regrowth: no sizing pass
    }
    leaf_cap = (unsigned int)std::min<long long>((long long)nt * 24 + 1024, 1ll << 31);
    group_cap = (unsigned int)std::min<long long>((long long)nt * 24 + 1024, 1ll << 31);
    if (const char *cap = std::getenv("VP_LEAF_CAP")) leaf_cap = group_cap = (unsigned int)std::max(1ll, std::atoll(cap));
    d_nonfar = dalloc<int2>(nonfar_cap);
    d_touchp = dalloc<int2>(nonfar_cap);
    d_near = dalloc<int2>(nonfar_cap);
    d_nearm = dalloc<int2>(nonfar_cap);
    d_simple = dalloc<int4>(nonfar_cap);
    d_leaves = dalloc<Leaf>(leaf_cap);
    d_groups = dalloc<int4>(group_cap);
    owned.insert(owned.end(), {d_nonfar, d_touchp, d_near, d_nearm, d_simple, d_leaves, d_groups});
    if (far_queue_wgs > 0 && opts.precision == 1) {  // the FP32 far queue's deferred closed-form points, a slice per workgroup
        static_assert(FAR_Q_DEFER(float) && !FAR_Q_DEFER(double), "d_fq is made for the FP32 far queue only");
        d_fq = dalloc<uint16_t>((size_t)far_queue_wgs * FAR_Q_WAVES(float) * 3 * TT_MAX * TS);
        owned.push_back(d_fq);
    }
    for (int k = 0; k < 6; k++) HIPCHECK(hipEventCreate(&ev[k]));
    for (int k = 0; k < TK_COUNT; k++)
        for (int j = 0; j < 2; j++) HIPCHECK(hipEventCreate(&tev[k][j]));
    serial = next_plan_serial();
    if (record_tiers) {
        d_tiers = dalloc<uint8_t>((size_t)nt * nsrc);
        owned.push_back(d_tiers);
    }
    HIPCHECK(hipStreamSynchronize(st));
}

AsmParams Plan::params() const
{
    AsmParams P;
    P.nt = nt;
    P.nsrc = nsrc;
    P.nseg = nseg;
    P.nslot = nslot;
    P.ncol = ncol;
    P.nn = nn;
    P.test = opts.test;
    P.far_ratio = opts.far_ratio;
    P.near_ratio = opts.near_ratio;
    P.corner_angle = opts.corner_angle;
    P.contact_rel = opts.contact_rel;
    P.kernel_far = opts.kernel_far;
    P.near_depth = opts.near_depth;
    P.lda = lda;
    return P;
}

void Plan::assemble(const double *sigma, const double *vinf, bool host_sigma)
{
    hipStream_t st = ctx->stream;
    tune_collect();  // the previous assembly's timed launches
    // a forward evaluation kernel launched with the occupancy cap the launch autotuner picks, timed while it tunes
    auto tuned_launch = [&](int k, auto &&launch) {
        const int v = tune_pick(k);
        if (tune_timed[k]) HIPCHECK(hipEventRecord(tev[k][0], st));
        launch(kTuneWaves[v]);
        HIPCHECK(hipGetLastError());
        if (tune_timed[k]) HIPCHECK(hipEventRecord(tev[k][1], st));
    };
    HIPCHECK(hipMemcpyAsync(d_sigma, sigma, sizeof(double) * nt, host_sigma ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, st));
    // The work lists grow when a pass finds more entries than budgeted: a non-far list that overflowed is refilled by
    // the far kernels' enumerate-only pass (their far blocks are in A already), a rule-piece list by running the
    // touching pairs' rule generator again (its tier counts are kept apart, in stats[16..31], and reset for it).
    HIPCHECK(hipMemsetAsync(A, 0, sizeof(double) * (size_t)lda * ncol, st));
    HIPCHECK(hipMemsetAsync(d_b, 0, sizeof(double) * nslot, st));
    HIPCHECK(hipMemsetAsync(d_stats, 0, sizeof(unsigned long long) * 32, st));
    HIPCHECK(hipMemsetAsync(d_count, 0, sizeof(unsigned int) * 7, st));
    if (d_tiers) HIPCHECK(hipMemsetAsync(d_tiers, 255, (size_t)nt * nsrc, st));
    const AsmParams P = params();
    const int st0 = wake_only ? nts_body : 0;
    dim3 grid(ntt, nts - st0);
    if (timing) HIPCHECK(hipEventRecord(ev[0], st));
    auto far_pass = [&](bool enum_only) {
        if (nts - st0 > 0 && far_queue_wgs > 0) {  // one persistent work queue over every tile
            HIPCHECK(hipMemsetAsync(d_count + 7, 0, sizeof(unsigned int), st));
            const long long per_cu = opts.precision == 1 ? FAR_Q_WAVES(float) : FAR_Q_WAVES(double);
            const unsigned int nwg = (unsigned int)std::min<long long>((long long)ntt * (nts - st0), per_cu * far_queue_wgs);
            if (opts.precision == 1)
                far_pairs_queue<float><<<nwg, 256, 0, st>>>(P, d_tf, d_el, d_sigma, d_tte, d_ttl, d_ttr, d_ttn, d_tse,
                                                            d_tsl, d_tsc, d_tsn, A, d_b, d_nonfar, d_count, nonfar_cap,
                                                            d_stats + 5, d_tiers, st0, enum_only, ntt, nts - st0,
                                                            d_count + 7, d_fq);
            else
                far_pairs_queue<double><<<nwg, 256, 0, st>>>(P, d_tf, d_el, d_sigma, d_tte, d_ttl, d_ttr, d_ttn, d_tse,
                                                             d_tsl, d_tsc, d_tsn, A, d_b, d_nonfar, d_count, nonfar_cap,
                                                             d_stats + 5, d_tiers, st0, enum_only, ntt, nts - st0,
                                                             d_count + 7, d_fq);
            HIPCHECK(hipGetLastError());
        } else if (nts - st0 > 0) {
            if (opts.precision == 1)
                far_pairs<float><<<grid, 256, 0, st>>>(P, d_tf, d_el, d_sigma, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl,
                                                       d_tsc, d_tsn, A, d_b, d_nonfar, d_count, nonfar_cap, d_stats + 5,
                                                       d_tiers, st0, enum_only);
            else
                far_pairs<double><<<grid, 256, 0, st>>>(P, d_tf, d_el, d_sigma, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl,
                                                        d_tsc, d_tsn, A, d_b, d_nonfar, d_count, nonfar_cap, d_stats + 5,
                                                        d_tiers, st0, enum_only);
            HIPCHECK(hipGetLastError());
        }
        if (nseg) {
            long long np = (long long)nt * nseg;
            seg_pairs<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(P, d_tf, d_cs, d_seg, d_segcol, A, d_nonfar, d_count,
                                                                    nonfar_cap, d_stats + 5, enum_only);
            HIPCHECK(hipGetLastError());
        }
    };
    // every pair list grown to hold cap entries (contents dropped)
    auto grow_pair_lists = [&](unsigned int cap) {
        HIPCHECK(hipStreamSynchronize(st));
        for (int2 **pp : {&d_nonfar, &d_touchp, &d_near, &d_nearm}) {
            owned.erase(std::find(owned.begin(), owned.end(), (void *)*pp));
            (void)hipFree(*pp);
        }
        owned.erase(std::find(owned.begin(), owned.end(), (void *)d_simple));
        (void)hipFree(d_simple);
        nonfar_cap = cap;
        d_nonfar = dalloc<int2>(nonfar_cap);
        d_touchp = dalloc<int2>(nonfar_cap);
        d_near = dalloc<int2>(nonfar_cap);
        d_nearm = dalloc<int2>(nonfar_cap);
        d_simple = dalloc<int4>(nonfar_cap);
        owned.insert(owned.end(), {d_nonfar, d_touchp, d_near, d_nearm, d_simple});
    };
    if (!lists_sized && !nonfar_cap_forced) {
        // the plan's first assembly: the lists sized from an upper bound of the non-far count, so that the far pass
        // never has to run again in its enumerate-only mode (a wing's slender panels have ~1,100 non-far pairs each)
        const int s_begin = wake_only ? nt : 0;
        const int nsb = (nsrc - s_begin) + nseg;
        if (nt > 0 && nsb > 0) {
            HIPCHECK(hipMemsetAsync(d_stats + 15, 0, sizeof(unsigned long long), st));
            nonfar_bound<<<dim3((nt + 255) / 256, (nsb + BOUND_CHUNK - 1) / BOUND_CHUNK), 256, 0, st>>>(
                d_el, nt, s_begin, nsrc, d_seg, nseg, opts.far_ratio, d_stats + 15);
            HIPCHECK(hipGetLastError());
            unsigned long long bound = 0;
            HIPCHECK(hipMemcpyAsync(&bound, d_stats + 15, sizeof(bound), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (bound > 0xFFFF0000ull) throw std::length_error("more than 2^32 non-far pairs");
            if (bound > nonfar_cap) grow_pair_lists((unsigned int)bound + 1024);
        }
        lists_sized = true;
    }
    far_pass(false);
    if (timing) HIPCHECK(hipEventRecord(ev[1], st));
    unsigned int cnt[7] = {0, 0, 0, 0, 0, 0, 0};
    HIPCHECK(hipMemcpyAsync(cnt, d_count, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (cnt[0] > nonfar_cap) {  // more non-far pairs than budgeted: grow every pair list and enumerate again
        grow_pair_lists(cnt[0] + 1024);
        HIPCHECK(hipMemsetAsync(d_count, 0, sizeof(unsigned int), st));
        far_pass(true);
        HIPCHECK(hipMemcpyAsync(cnt, d_count, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (cnt[0] > nonfar_cap) throw std::runtime_error("the non-far pair count changed between passes");
    }
    nnonfar = cnt[0];
    if (cnt[0]) {
        triage<<<std::min<unsigned int>((cnt[0] + 255) / 256, 65535u), 256, 0, st>>>(
            P, d_el, d_nodes, d_nonfar, cnt[0], d_touch, ntouch, d_touchp, d_near, d_nearm, d_simple, d_count + 1,
            nonfar_cap);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(cnt, d_count, 5 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    nsimple = cnt[4];
    ntouchp = cnt[1];
    nnear = cnt[2];
    nnearm = cnt[3];
    auto gen_pass = [&]() {
        if (cnt[1]) {
            touch_gen<<<std::min<unsigned int>((cnt[1] + 255) / 256, 65535u), 256, 0, st>>>(
                P, rt, d_el, d_nodes, d_seg, d_touchp, cnt[1], d_touch, ntouch, d_leaves, leaf_cap, d_groups, group_cap,
                d_count + 5, d_stats + 16, d_tiers);
            HIPCHECK(hipGetLastError());
        }
    };
    gen_pass();
    if (timing) HIPCHECK(hipEventRecord(ev[2], st));
    if (nsimple) {
        const unsigned int g = std::min<unsigned int>((nsimple + 3) / 4, 65535u);
        tuned_launch(TK_TOUCH, [&](int w) {
            if (w == 3)
                touch_simple<0, 3><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_simple, nsimple, A, d_b,
                                                       d_stats, d_tiers, AdjArgs{});
            else
                touch_simple<0, 2><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_simple, nsimple, A, d_b,
                                                       d_stats, d_tiers, AdjArgs{});
        });
    }
    if (cnt[2]) {
        const unsigned int per = WAVE / std::max(1, rt.n[R_NEAR]), waves = (cnt[2] + per - 1) / per;
        const unsigned int g = std::min<unsigned int>((waves + 3) / 4, 65535u), n = cnt[2];
        tuned_launch(TK_NEARPTS, [&](int w) {
            // FP64 in both paths: its FP32 build measured slower on the swept wing (24.1 against 23.1 ms)
            if (w == 3)
                near_points<3><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_near, n, A, d_b, d_stats, d_tiers);
            else
                near_points<2><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_near, n, A, d_b, d_stats, d_tiers);
        });
    }
    if (cnt[3]) {
        const unsigned int g = std::min<unsigned int>((cnt[3] + 3) / 4, 65535u), n = cnt[3];
        tuned_launch(TK_NEARGRP, [&](int w) {
            if (opts.precision == 1 && w == 3)
                near_group<0, 3, float><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, n, A, d_b, d_stats,
                                                            d_tiers, AdjArgs{});
            else if (opts.precision == 1)
                near_group<0, 2, float><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, n, A, d_b, d_stats,
                                                            d_tiers, AdjArgs{});
            else if (w == 3)
                near_group<0, 3><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, n, A, d_b, d_stats,
                                                     d_tiers, AdjArgs{});
            else
                near_group<0, 2><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, n, A, d_b, d_stats,
                                                     d_tiers, AdjArgs{});
        });
    }
    if (timing) HIPCHECK(hipEventRecord(ev[3], st));
    HIPCHECK(hipMemcpyAsync(cnt + 5, d_count + 5, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (cnt[5] > leaf_cap || cnt[6] > group_cap) {  // more rule pieces than budgeted: grow and generate again
        owned.erase(std::find(owned.begin(), owned.end(), (void *)d_leaves));
        owned.erase(std::find(owned.begin(), owned.end(), (void *)d_groups));
        (void)hipFree(d_leaves);
        (void)hipFree(d_groups);
        leaf_cap = std::max(leaf_cap, cnt[5] + 1024);
        group_cap = std::max(group_cap, cnt[6] + 1024);
        d_leaves = dalloc<Leaf>(leaf_cap);
        d_groups = dalloc<int4>(group_cap);
        owned.insert(owned.end(), {d_leaves, d_groups});
        HIPCHECK(hipMemsetAsync(d_count + 5, 0, 2 * sizeof(unsigned int), st));
        HIPCHECK(hipMemsetAsync(d_stats + 16, 0, sizeof(unsigned long long) * 16, st));
        gen_pass();
        HIPCHECK(hipMemcpyAsync(cnt + 5, d_count + 5, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (cnt[5] > leaf_cap || cnt[6] > group_cap) throw std::runtime_error("the rule-piece count changed between passes");
    }
    nleaves = cnt[5];
    ngroups = cnt[6];
    if (cnt[6]) {
        const unsigned int g = std::min<unsigned int>((cnt[6] + 3) / 4, 65535u), n = cnt[6];
        tuned_launch(TK_GROUP, [&](int w) {
            if (w == 3)
                group_eval<3><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_seg, d_segcol, d_leaves, d_groups,
                                                  n, A, d_b);
            else
                group_eval<2><<<g, 256, 0, st>>>(P, rt, d_el, d_sigma, d_cs, d_cols, d_seg, d_segcol, d_leaves, d_groups,
                                                  n, A, d_b);
        });
    }
    if (timing) HIPCHECK(hipEventRecord(ev[4], st));
    if (!wake_only) {
        freestream<<<(nt + 255) / 256, 256, 0, st>>>(d_el, d_cs, nt, opts.test, vinf[0], vinf[1], vinf[2], d_b);
        HIPCHECK(hipGetLastError());
    }
    if (timing) HIPCHECK(hipEventRecord(ev[5], st));
    assembled = true;
}

void Plan::kernel_ms(double *ms)
{
    HIPCHECK(hipEventSynchronize(ev[5]));
    for (int k = 0; k < 5; k++) {
        float t = 0;
        HIPCHECK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
        ms[k] = t;
    }
    ms[5] = (double)nnear;
    ms[6] = (double)nleaves;

    unsigned long long h[16];
    read_stats(h);
    ms[7] = (double)h[12];
}

// The assembly's counters: the evaluation kernels' (stats[0..15]) plus the rule generator's (stats[16..31])
void Plan::read_stats(unsigned long long *h)
{
    unsigned long long both[32];
    HIPCHECK(hipMemcpyAsync(both, d_stats, sizeof(both), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 16; k++) h[k] = both[k] + both[16 + k];
}

void Plan::stats(long long *pairs, long long *points)
{
    unsigned long long h[16];
    read_stats(h);
    for (int k = 0; k < 6; k++) {
        pairs[k] = (long long)h[k];
        points[k] = (long long)h[6 + k];
    }
    // far pairs were counted in slot 5 by far_pairs and seg_pairs (their points are 3 each)
    points[5] = 3 * pairs[5] + points[5];
}

}  // namespace vp

#include "vp_adjoint.hpp"
#include "vp_explicit.hpp"
