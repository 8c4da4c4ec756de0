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
// (its far tiles and its routed non-far lists, evaluated by the forward's kernels in adjoint mode; pieces at fixed
// barycentric positions), and per pair
//
//     P_ts = -A_T sum_q w_q psi_lambda(q) Q.w(x_q; S),   Q = sigma_S n_T + n_T x Gamma_S,
//
// (pair_block, assemble.jl:245-258, contracted with lambda and mu) is differentiated: the kernel w by forward-mode dual
// numbers with respect to the three vertices of S (three passes of three tangents; the derivative with respect to the
// point is minus their sum, by translation invariance) or, for the multipole, its reverse derivative by hand; the test
// triangle's area and normal, the source's normal
// (sigma) and sheet (Gamma) by their closed-form adjoints.  Filaments (segment_block, :312-321) and the freestream
// term (:422) likewise.  Gradients are added per node: body nodes, then wake nodes (offset by the body's node count).
//
// This file is included by vp_assemble.hip (it uses the plan's tiles, lists and rule generator).

namespace vp {

// ------------------------------------------------------------------------------------------ kernels ---------------

__global__ void slot_nodes(const int *__restrict__ corner_slot, const int *__restrict__ nodes, int nt, int *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt)
        for (int k = 0; k < 3; k++) out[corner_slot[3 * t + k]] = nodes[3 * t + k];
}

// A far tile's LDS gradient sums (per local test slot and local source column) added to the node gradients
__device__ __forceinline__ void far_tile_flush(const AsmParams &P, double (*GT)[3], double (*GU)[3], int tt, int st,
                                               const int *tt_rows, const int *tt_nrow, const int *ts_cols,
                                               const int *ts_ncol, const int *slot_node, double *G)
{
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

// The far tier (the forward's tiles and far test, far_pairs): lane = source element, wave = test triangle; the
// gradients are summed per local test slot and local source column in LDS and flushed per tile.  Pairs with a
// closed-form point are queued in LDS as (test index in the tile) x TS + lane and written out as one contiguous segment
// per tile (exact_segs: tile, first entry, count), for adj_far_exact, which keeps the tile's LDS sums.
__global__ __launch_bounds__(256) void adj_far_pairs(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf,
                                                     const Elem *__restrict__ el, const int *__restrict__ tt_elems,
                                                     const uint8_t *__restrict__ tt_loc, const int *__restrict__ tt_rows,
                                                     const int *__restrict__ tt_nrow, const int *__restrict__ ts_elems,
                                                     const uint8_t *__restrict__ ts_loc, const int *__restrict__ ts_cols,
                                                     const int *__restrict__ ts_ncol, const int *__restrict__ src_cols,
                                                     const int *__restrict__ slot_node, const double *__restrict__ lam,
                                                     const double *__restrict__ mu, double vx, double vy, double vz,
                                                     V3<double> dV, double *__restrict__ bdot,
                                                     double *__restrict__ G, unsigned short *__restrict__ exact_list,
                                                     int4 *__restrict__ exact_segs, unsigned int *__restrict__ exact_count,
                                                     unsigned int exact_cap)
{
    __shared__ double GT[VMAX][3];
    __shared__ double GU[UMAX][3];
    __shared__ unsigned short exq[TT_MAX * TS];
    __shared__ unsigned int nexq, exbase;
    const int tt = blockIdx.x, st = blockIdx.y;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < VMAX * 3; i += blockDim.x) GT[i / 3][i % 3] = 0.0;
    for (int i = threadIdx.x; i < UMAX * 3; i += blockDim.x) GU[i / 3][i % 3] = 0.0;
    if (threadIdx.x == 0) nexq = 0;
    const V3<double> Vinf = mk3(vx, vy, vz);
    const int s = ts_elems[st * TS + lane];
    const bool valid = s >= 0;
    V3<double> XS[3], Gam = mk3(0.0, 0.0, 0.0), sc = mk3(0.0, 0.0, 0.0);
    double muS[3] = {0, 0, 0}, sigS = 0.0, sdiam = 0.0, dVn = 0.0;
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
        dVn = s < P.nt ? dot(dV, ld3(S.n)) : 0.0;
        sc = ld3(S.c);
        sdiam = S.diam;
    }
    double bdsum = 0.0;
    __syncthreads();
    const int off = rt.off[R_FAR];
    for (int i = wave; i < TT_MAX; i += 4) {
        const int t = tt_elems[tt * TT_MAX + i];
        if (t < 0) break;
        const TestFar &F = tf[t];
        bool far = false;
        if (valid && s != t) far = cls_norm(sc - ld3(F.c)) > P.far_ratio * fmax(F.diam, sdiam);
        // pairs with a closed-form point go to adj_far_exact (dual numbers); here only all-multipole pairs
        bool multi = far;
        if (far) {
            const double fd = P.kernel_far * sdiam;
            for (int q = 0; q < 3; q++) {
                const V3<double> r0 = cls_sub(ld3(F.x[q]), sc);
                multi = multi && cls_dot(r0, r0) > fd * fd;
            }
        }
        const bool ex = far && !multi;
        const unsigned long long em = __ballot(ex);
        if (em) {
            unsigned int b0 = 0;
            if (lane == 0) b0 = atomicAdd(&nexq, (unsigned int)__popcll(em));
            b0 = __shfl(b0, 0);
            if (ex) exq[b0 + __popcll(em & ((1ull << lane) - 1ull))] = (unsigned short)(i * TS + lane);
        }
        if (!multi) continue;
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
            far_point_grad(XS, el[s], ld3(F.x[q]), Q, wv, g);
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
        bdsum += dVn * F.area * dot(p.Sv, nT);
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
    if (threadIdx.x == 0 && nexq) {
        exbase = atomicAdd(exact_count, nexq);
        const unsigned int sg = atomicAdd(exact_count + 1, 1u);
        exact_segs[sg] = make_int4(tt, st, (int)exbase, (int)nexq);  // at most one segment per tile: in bounds
    }
    __syncthreads();
    for (unsigned int j = threadIdx.x; j < nexq; j += blockDim.x)
        if (exbase + j < exact_cap) exact_list[exbase + j] = exq[j];
    block_add(bdsum, bdot);
    far_tile_flush(P, GT, GU, tt, st, tt_rows, tt_nrow, ts_cols, ts_ncol, slot_node, G);
}

// Far pairs with a closed-form point (within kernel_far diameters of the source), one workgroup per tile segment
// queued by adj_far_pairs: thread per pair, the three far points by tri_point_grad, the gradients summed in LDS per
// local test slot and source column as in adj_far_pairs and flushed once.
__global__ __launch_bounds__(256) void adj_far_exact(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf,
                                                     const Elem *__restrict__ el, const int *__restrict__ tt_elems,
                                                     const uint8_t *__restrict__ tt_loc, const int *__restrict__ tt_rows,
                                                     const int *__restrict__ tt_nrow, const int *__restrict__ ts_elems,
                                                     const uint8_t *__restrict__ ts_loc, const int *__restrict__ ts_cols,
                                                     const int *__restrict__ ts_ncol, const int *__restrict__ src_cols,
                                                     const int *__restrict__ slot_node, const double *__restrict__ lam,
                                                     const double *__restrict__ mu, double vx, double vy, double vz,
                                                     V3<double> dV, double *__restrict__ bdot,
                                                     double *__restrict__ G, const unsigned short *__restrict__ list,
                                                     const int4 *__restrict__ segs)
{
    __shared__ double GT[VMAX][3];
    __shared__ double GU[UMAX][3];
    const int4 sg = segs[blockIdx.x];
    const int tt = sg.x, st = sg.y;
    for (int i = threadIdx.x; i < VMAX * 3; i += blockDim.x) GT[i / 3][i % 3] = 0.0;
    for (int i = threadIdx.x; i < UMAX * 3; i += blockDim.x) GU[i / 3][i % 3] = 0.0;
    __syncthreads();
    const V3<double> Vinf = mk3(vx, vy, vz);
    const int off = rt.off[R_FAR];
    double bdsum = 0.0;
#pragma unroll 1
    for (int j = threadIdx.x; j < sg.w; j += blockDim.x) {
        const int code = list[sg.z + j], i = code / TS, ls = code - i * TS;
        const int t = tt_elems[tt * TT_MAX + i], s = ts_elems[st * TS + ls];
        const TestFar &F = tf[t];
        const Elem &S = el[s], &T = el[t];
        V3<double> XS[3];
        double muS[3];
        int uloc[3], vl[3];
        double la[3];
        for (int k = 0; k < 3; k++) {
            XS[k] = ld3(S.v[k]);
            muS[k] = mu[src_cols[3 * s + k]];
            uloc[k] = ts_loc[(st * TS + ls) * 3 + k];
            vl[k] = tt_loc[(tt * TT_MAX + i) * 3 + k];
            la[k] = lam[tt_rows[tt * VMAX + vl[k]]];
        }
        const V3<double> Gam = sheet_of(XS, muS);
        const double sigS = s < P.nt ? -dot(Vinf, ld3(S.n)) : 0.0;
        const V3<double> nT = ld3(F.n);
        const V3<double> Q = sigS * nT + cross(nT, Gam);
        PairSums p;
        p.Sv = mk3(0.0, 0.0, 0.0);
        for (int k = 0; k < 3; k++) p.GS[k] = p.GX[k] = mk3(0.0, 0.0, 0.0);
#pragma unroll 1
        for (int q = 0; q < 3; q++) {
            const double a = (la[0] * F.wpsi[q][0] + la[1] * F.wpsi[q][1] + la[2] * F.wpsi[q][2]) / F.area;
            V3<double> wv, g[3];
            tri_point_grad(XS, S, ld3(F.x[q]), Q, 0, P.kernel_far, wv, g);
            const V3<double> gx = -1.0 * ((g[0] + g[1]) + g[2]);
            const double *l = rt.l + 3 * (off + q);
            p.Sv = p.Sv + a * wv;
            for (int k = 0; k < 3; k++) {
                p.GS[k] = p.GS[k] + a * g[k];
                p.GX[k] = p.GX[k] + (a * l[k]) * gx;
            }
        }
        const V3<double> XT[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        V3<double> gT[3], gS[3];
        finish_tri_pair(XT, nT, F.area, XS, muS, sigS, Gam, s < P.nt, Vinf, Q, p, gT, gS);
        if (s < P.nt) bdsum += dot(dV, ld3(S.n)) * F.area * dot(p.Sv, nT);
        for (int k = 0; k < 3; k++) {
            atomicAdd(&GT[vl[k]][0], gT[k].x);
            atomicAdd(&GT[vl[k]][1], gT[k].y);
            atomicAdd(&GT[vl[k]][2], gT[k].z);
            atomicAdd(&GU[uloc[k]][0], gS[k].x);
            atomicAdd(&GU[uloc[k]][1], gS[k].y);
            atomicAdd(&GU[uloc[k]][2], gS[k].z);
        }
    }
    block_add(bdsum, bdot);
    far_tile_flush(P, GT, GU, tt, st, tt_rows, tt_nrow, ts_cols, ts_ncol, slot_node, G);
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

// The pairs of the general touching list (contact, geometric touch pairs, filaments): one wavefront per pair, the
// points of all its rule pieces (the forward's leaf list) over the lanes, as group_eval does.
__global__ __launch_bounds__(256) void adj_group_eval(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                      const int *__restrict__ corner_slot, const int *__restrict__ nodes,
                                                      const int *__restrict__ src_cols, const double *__restrict__ seg,
                                                      const int *__restrict__ segcol, const Leaf *__restrict__ leaves,
                                                      const int4 *__restrict__ groups, unsigned int ngroups,
                                                      const double *__restrict__ lam, const double *__restrict__ mu,
                                                      double vx, double vy, double vz, V3<double> dV,
                                                      double *__restrict__ bdot, double *__restrict__ G)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const V3<double> Vinf = mk3(vx, vy, vz);
    double bdsum = 0.0;
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
                if (s < P.nt) bdsum += dot(dV, ld3(el[s].n)) * T.area * dot(p.Sv, nT);
            }
            for (int k = 0; k < 3; k++) gadd(G, nodes[3 * t + k], gT[k]);
        }
    }
    block_add(bdsum, bdot);
}

// The freestream term -(V∞.n_T) A_T int(psi_a)/A_T (assemble.jl:422) contracted with lambda: -(ti/2) Lambda V∞.N_T;
// its value along dV∞ goes to bdot.
__global__ __launch_bounds__(256) void adj_freestream(const Elem *__restrict__ el, const int *__restrict__ corner_slot,
                                                      const int *__restrict__ nodes, int nt, int test,
                                                      const double *__restrict__ lam, double vx, double vy, double vz,
                                                      V3<double> dV, double *__restrict__ bdot, double *__restrict__ G)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double bd = 0.0;
    if (t < nt) {
        const Elem &E = el[t];
        const V3<double> X[3] = {ld3(E.v[0]), ld3(E.v[1]), ld3(E.v[2])};
        const double ti = test == 1 ? 1.0 : 1.0 / 3;
        const double Lam = lam[corner_slot[3 * t]] + lam[corner_slot[3 * t + 1]] + lam[corner_slot[3 * t + 2]];
        V3<double> g[3] = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0)};
        cross_adj(X, (-0.5 * ti * Lam) * mk3(vx, vy, vz), g);
        for (int k = 0; k < 3; k++) gadd(G, nodes[3 * t + k], g[k]);
        bd = -ti * Lam * E.area * dot(dV, ld3(E.n));
    }
    block_add(bd, bdot);
}

// ------------------------------------------------------------------------------------------ host ------------------

void Plan::adjoint(const double *lam_h, const double *mu_h, const double *vinf, const double *dvinf, double *G_h,
                   double *bdot_h)
{
    hipStream_t st = ctx->stream;
    if (!assembled) throw std::invalid_argument("the plan has not been assembled (vp_plan_assemble first)");
    if (!d_lam) {
        d_lam = dalloc<double>(nslot);
        d_mu = dalloc<double>(ncol);
        d_G = dalloc<double>((size_t)3 * (nn + nw));
        d_slotnode = dalloc<int>(nslot);
        d_bdot = dalloc<double>(1);
        owned.insert(owned.end(), {d_lam, d_mu, d_G, d_slotnode, d_bdot});
        slot_nodes<<<(nt + 255) / 256, 256, 0, st>>>(d_cs, d_nodes, nt, d_slotnode);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(d_lam, lam_h, sizeof(double) * nslot, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_mu, mu_h, sizeof(double) * ncol, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemsetAsync(d_G, 0, sizeof(double) * 3 * (nn + nw), st));
    HIPCHECK(hipMemsetAsync(d_bdot, 0, sizeof(double), st));
    const AsmParams P = params();
    const double vx = vinf[0], vy = vinf[1], vz = vinf[2];
    V3<double> dV;
    dV.x = dvinf ? dvinf[0] : 0.0;
    dV.y = dvinf ? dvinf[1] : 0.0;
    dV.z = dvinf ? dvinf[2] : 0.0;
    double *bd = dvinf ? d_bdot : nullptr;
    // far pairs with a closed-form point: at most the forward's count of closed-form far points
    unsigned long long hst[16];
    read_stats(hst);
    const unsigned int ecap = (unsigned int)std::min<unsigned long long>(hst[12] + 1024, 1ull << 31);
    unsigned short *d_exact = dalloc<unsigned short>(ecap);
    int4 *d_esegs = dalloc<int4>((size_t)ntt * nts);
    unsigned int *d_ecount = dalloc<unsigned int>(2);
    HIPCHECK(hipMemsetAsync(d_ecount, 0, 2 * sizeof(unsigned int), st));
    adj_far_pairs<<<dim3(ntt, nts), 256, 0, st>>>(P, rt, d_tf, d_el, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl, d_tsc, d_tsn,
                                                  d_cols, d_slotnode, d_lam, d_mu, vx, vy, vz, dV, bd, d_G, d_exact,
                                                  d_esegs, d_ecount, ecap);
    HIPCHECK(hipGetLastError());
    unsigned int ecnt[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(ecnt, d_ecount, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    auto free_exact = [&] {
        (void)hipFree(d_exact);
        (void)hipFree(d_esegs);
        (void)hipFree(d_ecount);
    };
    if (ecnt[0] > ecap) {
        free_exact();
        throw std::runtime_error("adjoint: more closed-form far pairs than the forward's closed-form far points");
    }
    if (ecnt[1]) {
        adj_far_exact<<<ecnt[1], 256, 0, st>>>(P, rt, d_tf, d_el, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl, d_tsc, d_tsn,
                                               d_cols, d_slotnode, d_lam, d_mu, vx, vy, vz, dV, bd, d_G, d_exact,
                                               d_esegs);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(st));
    free_exact();
    if (nseg) {
        const long long np = (long long)nt * nseg;
        adj_seg_far<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(P, rt, d_tf, d_el, d_cs, d_nodes, d_seg, d_segcol, d_lam,
                                                                   d_mu, d_G);
        HIPCHECK(hipGetLastError());
    }
    // the non-far pairs of the last assembly, through the forward's own routing and kernels in adjoint mode
    AdjArgs aa{d_lam, d_mu, d_G, d_nodes, vx, vy, vz, dV.x, dV.y, dV.z, bd};
    if (nsimple) {
        touch_simple<true><<<std::min<unsigned int>((nsimple + 3) / 4, 65535u), 256, 0, st>>>(
            P, rt, d_el, d_sigma, d_cs, d_cols, d_simple, nsimple, nullptr, nullptr, nullptr, nullptr, aa);
        HIPCHECK(hipGetLastError());
    }
    if (nnear) {
        near_thread<true><<<std::min<unsigned int>((nnear + 255) / 256, 65535u), 256, 0, st>>>(
            P, rt, d_el, d_sigma, d_cs, d_cols, d_near, nnear, nullptr, nullptr, nullptr, nullptr, aa);
        HIPCHECK(hipGetLastError());
    }
    if (nnearm) {
        near_group<true><<<std::min<unsigned int>((nnearm + 3) / 4, 65535u), 256, 0, st>>>(
            P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, nnearm, nullptr, nullptr, nullptr, nullptr, aa);
        HIPCHECK(hipGetLastError());
    }
    if (ngroups) {
        adj_group_eval<<<std::min<unsigned int>((ngroups + 3) / 4, 65535u), 256, 0, st>>>(
            P, rt, d_el, d_cs, d_nodes, d_cols, d_seg, d_segcol, d_leaves, d_groups, ngroups, d_lam, d_mu, vx, vy, vz, dV, bd,
            d_G);
        HIPCHECK(hipGetLastError());
    }
    adj_freestream<<<(nt + 255) / 256, 256, 0, st>>>(d_el, d_cs, d_nodes, nt, opts.test, d_lam, vx, vy, vz, dV, bd, d_G);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(G_h, d_G, sizeof(double) * 3 * (nn + nw), hipMemcpyDeviceToHost, st));
    if (bdot_h) HIPCHECK(hipMemcpyAsync(bdot_h, d_bdot, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace vp
