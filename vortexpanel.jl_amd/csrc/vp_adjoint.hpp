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
// (pair_block, assemble.jl:245-258, contracted with lambda and mu) is differentiated: the kernel w (closed form or
// multipole) by its reverse derivative by hand with respect to the three vertices of S (the derivative with respect
// to the point is minus their sum, by translation invariance); the test triangle's area and normal, the source's
// normal (sigma) and sheet (Gamma) by their closed-form adjoints; filaments (segment_block, :312-321) by dual numbers;
// the freestream term (:422) by hand.  Gradients are added per node: body nodes, then wake nodes (offset by the body's node count).
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

// A far tile's LDS gradient sums (per weight set, local test slot and local source column) added to the node
// gradients
template <int NL>
__device__ __forceinline__ void far_tile_flush(const AsmParams &P, double (*GT)[VMAX][3], double (*GU)[UMAX][3], int tt,
                                               int st, const int *tt_rows, const int *tt_nrow, const int *ts_cols,
                                               const int *ts_ncol, const int *slot_node, const AdjArgs &aa)
{
    const int nrow = tt_nrow[tt], ncol = ts_ncol[st];
    for (int l = 0; l < NL; l++) {
        double *G = aa.G + l * aa.gstride;
        for (int i = threadIdx.x; i < nrow * 3; i += blockDim.x) {
            const double v = GT[l][i / 3][i % 3];
            if (v != 0.0) unsafeAtomicAdd(&G[3 * slot_node[tt_rows[tt * VMAX + i / 3]] + i % 3], v);
        }
        for (int i = threadIdx.x; i < ncol * 3; i += blockDim.x) {
            const double v = GU[l][i / 3][i % 3];
            if (v != 0.0) unsafeAtomicAdd(&G[3 * col_node(ts_cols[st * UMAX + i / 3], P.nslot, P.nn, slot_node) + i % 3], v);
        }
    }
}

template <int NL> __device__ __forceinline__ void far_tile_zero(double (*GT)[VMAX][3], double (*GU)[UMAX][3])
{
    for (int i = threadIdx.x; i < NL * VMAX * 3; i += blockDim.x) (&GT[0][0][0])[i] = 0.0;
    for (int i = threadIdx.x; i < NL * UMAX * 3; i += blockDim.x) (&GU[0][0][0])[i] = 0.0;
}

// One far pair's three far-rule points (multipole only, or with the forward's kernel switch when EXACT) and its
// gradients per weight set: gT with respect to the test triangle's vertices, gS to the source's; bd gets the pair's
// lambdaᵀ b along dV∞ (dVn = dV∞.n_S).
template <int NL, bool EXACT, typename CT = double>
__device__ __forceinline__ void far_pair_grads(const AsmParams &P, const RuleTab &rt, const TestFar &F, const Elem &T,
                                               const V3<double> *XS, const Elem &S, bool body, const double *muS,
                                               V3<double> Gam, double sigS, double dVn, const double (*la)[3],
                                               V3<double> Vinf, V3<double> (*gT)[3], V3<double> (*gS)[3], double *bd)
{
    const V3<double> nT = ld3(F.n);
    const V3<double> Q = sigS * nT + cross(nT, Gam);
    PairSums p[NL];
    for (int j = 0; j < NL; j++) zero_sums(p[j]);
    const int off = rt.off[R_FAR];
    // unrolled: the three points' gradients are independent, and at one wave per SIMD their interleaving is the
    // latency hiding (C4: adj_far_exact 255 -> 196 ms, adj_far_pairs 652 -> 628 ms)
#pragma unroll
    for (int q = 0; q < 3; q++) {
        V3<double> wv, g[3];
        if constexpr (EXACT)
            tri_point_grad<CT>(XS, S, ld3(F.x[q]), Q, 0, P.kernel_far, wv, g);
        else
            far_point_grad<CT>(XS, S, ld3(F.x[q]), Q, wv, g);
        const V3<double> gx = -1.0 * ((g[0] + g[1]) + g[2]);
        const double *l = rt.l + 3 * (off + q);
        for (int j = 0; j < NL; j++) {
            const double a = (la[j][0] * F.wpsi[q][0] + la[j][1] * F.wpsi[q][1] + la[j][2] * F.wpsi[q][2]) / F.area;
            add_point(p[j], a, wv, g, gx, l);
        }
    }
    const V3<double> XT[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
    for (int j = 0; j < NL; j++) {
        finish_tri_pair(XT, nT, F.area, XS, muS, sigS, Gam, body, Vinf, Q, p[j], gT[j], gS[j]);
        bd[j] += dVn * F.area * dot(p[j].Sv, nT);
    }
}

__device__ __forceinline__ void lds_add3(double *G, V3<double> v)
{
    atomicAdd(G, v.x);
    atomicAdd(G + 1, v.y);
    atomicAdd(G + 2, v.z);
}

// far_pair_grads with the gradients added to the tile's LDS sums (vl, uloc: the local test slots and source columns)
template <int NL, bool EXACT, typename CT = double>
__device__ __forceinline__ void far_pair_adj(const AsmParams &P, const RuleTab &rt, const TestFar &F, const Elem &T,
                                             const V3<double> *XS, const Elem &S, bool body, const double *muS,
                                             V3<double> Gam, double sigS, double dVn, const double (*la)[3],
                                             const int *vl, const int *uloc, V3<double> Vinf, double (*GT)[VMAX][3],
                                             double (*GU)[UMAX][3], double *bd)
{
    V3<double> gT[NL][3], gS[NL][3];
    far_pair_grads<NL, EXACT, CT>(P, rt, F, T, XS, S, body, muS, Gam, sigS, dVn, la, Vinf, gT, gS, bd);
    for (int j = 0; j < NL; j++)
        for (int k = 0; k < 3; k++) {
            lds_add3(GT[j][vl[k]], gT[j][k]);
            lds_add3(GU[j][uloc[k]], gS[j][k]);
        }
}

// The far tier (the forward's tiles and far test, far_pairs): lane = source element, wave = test triangle; the
// gradients are summed per local test slot and local source column in LDS and flushed per tile.  Pairs with a
// closed-form point are queued in LDS as (test index in the tile) x TS + lane and written out as one contiguous segment
// per tile (exact_segs: tile, first entry, count), for adj_far_exact, which keeps the tile's LDS sums.
// Both builds keep the compiler's own register allocation (FP64 208 VGPRs, FP32 176; capping the FP32 build at three
// waves per SIMD left it 84 VGPRs and 840 bytes of spills, 1.8x slower).
template <int NL, typename CT = double>
__global__ __launch_bounds__(256) void adj_far_pairs(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf,
                                                     const Elem *__restrict__ el, const int *__restrict__ tt_elems,
                                                     const uint8_t *__restrict__ tt_loc, const int *__restrict__ tt_rows,
                                                     const int *__restrict__ tt_nrow, const int *__restrict__ ts_elems,
                                                     const uint8_t *__restrict__ ts_loc, const int *__restrict__ ts_cols,
                                                     const int *__restrict__ ts_ncol, const int *__restrict__ src_cols,
                                                     const int *__restrict__ slot_node, AdjArgs aa,
                                                     unsigned short *__restrict__ exact_list,
                                                     int4 *__restrict__ exact_segs, unsigned int *__restrict__ exact_count,
                                                     unsigned int exact_cap, int st0)
{
    __shared__ double GT[NL][VMAX][3];
    __shared__ double GU[NL][UMAX][3];
    __shared__ unsigned short exq[TT_MAX * TS];
    __shared__ unsigned int nexq, exbase;
    const int tt = blockIdx.x, st = blockIdx.y + st0;  // st0: the first source tile of this launch's slice
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    far_tile_zero<NL>(GT, GU);
    if (threadIdx.x == 0) nexq = 0;
    const V3<double> Vinf = mk3(aa.vx, aa.vy, aa.vz);
    const int s = ts_elems[st * TS + lane];
    const bool valid = s >= 0;
    V3<double> XS[3], Gam = mk3(0.0, 0.0, 0.0), sc = mk3(0.0, 0.0, 0.0);
    double muS[3] = {0, 0, 0}, sigS = 0.0, sdiam = 0.0, dVn = 0.0;
    int uloc[3] = {0, 0, 0};
    if (valid) {
        const Elem &S = el[s];
        for (int k = 0; k < 3; k++) {
            XS[k] = ld3(S.v[k]);
            muS[k] = aa.mu[src_cols[3 * s + k]];
            uloc[k] = ts_loc[(st * TS + lane) * 3 + k];
        }
        Gam = sheet_of(XS, muS);
        sigS = s < P.nt ? -dot(Vinf, ld3(S.n)) : 0.0;
        dVn = s < P.nt ? dot(mk3(aa.dx, aa.dy, aa.dz), ld3(S.n)) : 0.0;
        sc = ld3(S.c);
        sdiam = S.diam;
    }
    double bd[NL] = {};
    V3<double> accS[NL][3];
    for (int j = 0; j < NL; j++)
        for (int k = 0; k < 3; k++) accS[j][k] = mk3(0.0, 0.0, 0.0);
    __syncthreads();
    for (int i = wave; i < TT_MAX; i += 4) {
        const int t = tt_elems[tt * TT_MAX + i];
        if (t < 0) break;
        const TestFar &F = tf[t];
        bool far = false;
        if (valid && s != t) far = cls_far(sc - ld3(F.c), P.far_ratio * fmax(F.diam, sdiam));
        // pairs with a closed-form point go to adj_far_exact; here only all-multipole pairs
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
        if (!__ballot(multi)) continue;
        const int vl[3] = {tt_loc[(tt * TT_MAX + i) * 3], tt_loc[(tt * TT_MAX + i) * 3 + 1], tt_loc[(tt * TT_MAX + i) * 3 + 2]};
        V3<double> gT[NL][3];
        if (multi) {
            double la[NL][3];
            for (int j = 0; j < NL; j++)
                for (int k = 0; k < 3; k++) la[j][k] = aa.lam[j * aa.lstride + tt_rows[tt * VMAX + vl[k]]];
            V3<double> gS[NL][3];
            far_pair_grads<NL, false, CT>(P, rt, F, el[t], XS, el[s], s < P.nt, muS, Gam, sigS, dVn, la, Vinf, gT, gS, bd);
            for (int j = 0; j < NL; j++)
                for (int k = 0; k < 3; k++) accS[j][k] = accS[j][k] + gS[j][k];
        } else {
            for (int j = 0; j < NL; j++)
                for (int k = 0; k < 3; k++) gT[j][k] = mk3(0.0, 0.0, 0.0);
        }
        // the wave's lanes share the test triangle: its gradients are summed over the wave and added once
        for (int j = 0; j < NL; j++)
            for (int k = 0; k < 3; k++) {
                const V3<double> v = mk3(wave_total(gT[j][k].x), wave_total(gT[j][k].y), wave_total(gT[j][k].z));
                if (lane == 0) lds_add3(GT[j][vl[k]], v);
            }
    }
    // the lane's source gradients, summed over the tile's test triangles, into the tile's column sums
    if (valid)
        for (int j = 0; j < NL; j++)
            for (int k = 0; k < 3; k++) lds_add3(GU[j][uloc[k]], accS[j][k]);
    __syncthreads();
    if (threadIdx.x == 0 && nexq) {
        exbase = atomicAdd(exact_count, nexq);
        const unsigned int sg = atomicAdd(exact_count + 1, 1u);
        exact_segs[sg] = make_int4(tt, st, (int)exbase, (int)nexq);  // at most one segment per tile: in bounds
    }
    __syncthreads();
    for (unsigned int j = threadIdx.x; j < nexq; j += blockDim.x)
        if (exbase + j < exact_cap) exact_list[exbase + j] = exq[j];
    block_add_n<NL>(bd, aa.bdot);
    far_tile_flush<NL>(P, GT, GU, tt, st, tt_rows, tt_nrow, ts_cols, ts_ncol, slot_node, aa);
}

// Far pairs with a closed-form point (within kernel_far diameters of the source), one workgroup per tile segment
// queued by adj_far_pairs: thread per pair, the three far points by tri_point_grad, the gradients summed in LDS per
// local test slot and source column as in adj_far_pairs and flushed once.
template <int NL, typename CT = double>
__global__ __launch_bounds__(256) void adj_far_exact(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf,
                                                     const Elem *__restrict__ el, const int *__restrict__ tt_elems,
                                                     const uint8_t *__restrict__ tt_loc, const int *__restrict__ tt_rows,
                                                     const int *__restrict__ tt_nrow, const int *__restrict__ ts_elems,
                                                     const uint8_t *__restrict__ ts_loc, const int *__restrict__ ts_cols,
                                                     const int *__restrict__ ts_ncol, const int *__restrict__ src_cols,
                                                     const int *__restrict__ slot_node, AdjArgs aa,
                                                     const unsigned short *__restrict__ list,
                                                     const int4 *__restrict__ segs)
{
    __shared__ double GT[NL][VMAX][3];
    __shared__ double GU[NL][UMAX][3];
    const int4 sg = segs[blockIdx.x];
    const int tt = sg.x, st = sg.y;
    far_tile_zero<NL>(GT, GU);
    __syncthreads();
    const V3<double> Vinf = mk3(aa.vx, aa.vy, aa.vz), dV = mk3(aa.dx, aa.dy, aa.dz);
    double bd[NL] = {};
#pragma unroll 1
    for (int j = threadIdx.x; j < sg.w; j += blockDim.x) {
        const int code = list[sg.z + j], i = code / TS, ls = code - i * TS;
        const int t = tt_elems[tt * TT_MAX + i], s = ts_elems[st * TS + ls];
        const Elem &S = el[s];
        V3<double> XS[3];
        double muS[3], la[NL][3];
        int uloc[3], vl[3];
        for (int k = 0; k < 3; k++) {
            XS[k] = ld3(S.v[k]);
            muS[k] = aa.mu[src_cols[3 * s + k]];
            uloc[k] = ts_loc[(st * TS + ls) * 3 + k];
            vl[k] = tt_loc[(tt * TT_MAX + i) * 3 + k];
            for (int l = 0; l < NL; l++) la[l][k] = aa.lam[l * aa.lstride + tt_rows[tt * VMAX + vl[k]]];
        }
        const bool body = s < P.nt;
        const double sigS = body ? -dot(Vinf, ld3(S.n)) : 0.0, dVn = body ? dot(dV, ld3(S.n)) : 0.0;
        far_pair_adj<NL, true, CT>(P, rt, tf[t], el[t], XS, S, body, muS, sheet_of(XS, muS), sigS, dVn, la, vl, uloc, Vinf,
                               GT, GU, bd);
    }
    __syncthreads();
    block_add_n<NL>(bd, aa.bdot);
    far_tile_flush<NL>(P, GT, GU, tt, st, tt_rows, tt_nrow, ts_cols, ts_ncol, slot_node, aa);
}

// Far filament pairs (seg_pairs' far test and far segment_block): one thread per (test triangle, filament).
template <int NL>
__global__ void adj_seg_far(AsmParams P, RuleTab rt, const TestFar *__restrict__ tf, const Elem *__restrict__ el,
                            const int *__restrict__ corner_slot, const double *__restrict__ seg,
                            const int *__restrict__ segcol, AdjArgs aa)
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
    const double mua = aa.mu[segcol[2 * f]], mub = aa.mu[segcol[2 * f + 1]];
    PairSums p[NL];
    for (int j = 0; j < NL; j++) zero_sums(p[j]);
    for (int q = 0; q < 3; q++) {
        V3<double> wv, g[3];
        seg_point_grad(fa, fb, ld3(F.x[q]), nT, mua, mub, wv, g);
        g[2] = mk3(0.0, 0.0, 0.0);
        const V3<double> gx = -1.0 * (g[0] + g[1]);
        const double *l = rt.l + 3 * (rt.off[R_FAR] + q);
        for (int j = 0; j < NL; j++) {
            const double *lam = aa.lam + j * aa.lstride;
            const double a = (lam[corner_slot[3 * t]] * F.wpsi[q][0] + lam[corner_slot[3 * t + 1]] * F.wpsi[q][1] +
                              lam[corner_slot[3 * t + 2]] * F.wpsi[q][2]) / F.area;
            add_point(p[j], a, wv, g, gx, l);
        }
    }
    for (int j = 0; j < NL; j++) {
        V3<double> gT[3];
        for (int k = 0; k < 3; k++) gT[k] = (-F.area) * p[j].GX[k];
        area2_adj(XT, -0.5 * dot(nT, p[j].Sv), gT);
        normal_adj(XT, (-F.area) * p[j].Sv, gT);
        double *G = aa.G + j * aa.gstride;
        for (int k = 0; k < 3; k++) gadd(G, aa.nodes[3 * t + k], gT[k]);
        gadd(G, P.nn + segcol[2 * f] - P.nslot, (-F.area) * p[j].GS[0]);
        gadd(G, P.nn + segcol[2 * f + 1] - P.nslot, (-F.area) * p[j].GS[1]);
    }
}

// The pairs of the general touching list (contact, geometric touch pairs, filaments): one wavefront per pair, the
// points of all its rule pieces (the forward's leaf list) over the lanes, as group_eval does.
template <int NL>
__global__ __launch_bounds__(256) void adj_group_eval(AsmParams P, RuleTab rt, const Elem *__restrict__ el,
                                                      const int *__restrict__ corner_slot, const int *__restrict__ src_cols,
                                                      const double *__restrict__ seg, const int *__restrict__ segcol,
                                                      const Leaf *__restrict__ leaves, const int4 *__restrict__ groups,
                                                      unsigned int ngroups, AdjArgs aa)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const V3<double> Vinf = mk3(aa.vx, aa.vy, aa.vz), dV = mk3(aa.dx, aa.dy, aa.dz);
    double bd[NL] = {};
    for (unsigned int gi = blockIdx.x * 4 + wave; gi < ngroups; gi += gridDim.x * 4) {
        const int4 grp = groups[gi];
        const int t = leaves[grp.x].t, s = leaves[grp.x].s, ntot = grp.z;
        const bool isseg = s >= P.nsrc;
        const Elem &T = el[t];
        const V3<double> XT[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
        const V3<double> nT = ld3(T.n);
        double la[NL][3];
        for (int j = 0; j < NL; j++)
            for (int k = 0; k < 3; k++) la[j][k] = aa.lam[j * aa.lstride + corner_slot[3 * t + k]];
        V3<double> XS[3], Gam = mk3(0.0, 0.0, 0.0), Q = mk3(0.0, 0.0, 0.0), fa = Gam, fb = Gam;
        double muS[3] = {0, 0, 0}, sigS = 0.0, mua = 0.0, mub = 0.0;
        if (isseg) {
            const int f = s - P.nsrc;
            fa = ld3(seg + 6 * f);
            fb = ld3(seg + 6 * f + 3);
            mua = aa.mu[segcol[2 * f]];
            mub = aa.mu[segcol[2 * f + 1]];
        } else {
            const Elem &S = el[s];
            for (int k = 0; k < 3; k++) {
                XS[k] = ld3(S.v[k]);
                muS[k] = aa.mu[src_cols[3 * s + k]];
            }
            Gam = sheet_of(XS, muS);
            sigS = s < P.nt ? -dot(Vinf, ld3(S.n)) : 0.0;
            Q = sigS * nT + cross(nT, Gam);
        }
        PairSums p[NL];
        for (int j = 0; j < NL; j++) zero_sums(p[j]);
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
            const V3<double> x = cls_point(XT, l);
            V3<double> wv, g[3];
            if (isseg) {
                seg_point_grad(fa, fb, x, nT, mua, mub, wv, g);
                g[2] = mk3(0.0, 0.0, 0.0);
            } else {
                tri_point_grad(XS, el[s], x, Q, s == t ? 1 : 0, P.kernel_far, wv, g);
            }
            const V3<double> gx = -1.0 * ((g[0] + g[1]) + g[2]);
            const double W = rt.w[off + pq] * lf.frac;
            for (int j = 0; j < NL; j++)
                add_point(p[j], W * (la[j][0] * psi[0] + la[j][1] * psi[1] + la[j][2] * psi[2]), wv, g, gx, l);
        }
        for (int j = 0; j < NL; j++) adj_wave_reduce(p[j]);
        if (lane == 0) {
            for (int j = 0; j < NL; j++) {
                double *G = aa.G + j * aa.gstride;
                V3<double> gT[3], gS[3];
                if (isseg) {
                    for (int k = 0; k < 3; k++) gT[k] = (-T.area) * p[j].GX[k];
                    area2_adj(XT, -0.5 * dot(nT, p[j].Sv), gT);
                    normal_adj(XT, (-T.area) * p[j].Sv, gT);
                    const int f = s - P.nsrc;
                    gadd(G, P.nn + segcol[2 * f] - P.nslot, (-T.area) * p[j].GS[0]);
                    gadd(G, P.nn + segcol[2 * f + 1] - P.nslot, (-T.area) * p[j].GS[1]);
                } else {
                    finish_tri_pair(XT, nT, T.area, XS, muS, sigS, Gam, s < P.nt, Vinf, Q, p[j], gT, gS);
                    for (int k = 0; k < 3; k++) gadd(G, aa.nodes[3 * s + k], gS[k]);
                    if (s < P.nt) bd[j] += dot(dV, ld3(el[s].n)) * T.area * dot(p[j].Sv, nT);
                }
                for (int k = 0; k < 3; k++) gadd(G, aa.nodes[3 * t + k], gT[k]);
            }
        }
    }
    block_add_n<NL>(bd, aa.bdot);
}

// The freestream term -(V∞.n_T) A_T int(psi_a)/A_T (assemble.jl:422) contracted with lambda: -(ti/2) Lambda V∞.N_T;
// its value along dV∞ goes to bdot.
template <int NL>
__global__ __launch_bounds__(256) void adj_freestream(const Elem *__restrict__ el, const int *__restrict__ corner_slot,
                                                      int nt, int test, AdjArgs aa)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double bd[NL] = {};
    if (t < nt) {
        const Elem &E = el[t];
        const V3<double> X[3] = {ld3(E.v[0]), ld3(E.v[1]), ld3(E.v[2])};
        const double ti = test == 1 ? 1.0 : 1.0 / 3;
        for (int j = 0; j < NL; j++) {
            const double *lam = aa.lam + j * aa.lstride;
            const double Lam = lam[corner_slot[3 * t]] + lam[corner_slot[3 * t + 1]] + lam[corner_slot[3 * t + 2]];
            V3<double> g[3] = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0)};
            cross_adj(X, (-0.5 * ti * Lam) * mk3(aa.vx, aa.vy, aa.vz), g);
            for (int k = 0; k < 3; k++) gadd(aa.G + j * aa.gstride, aa.nodes[3 * t + k], g[k]);
            bd[j] = -ti * Lam * E.area * dot(mk3(aa.dx, aa.dy, aa.dz), ld3(E.n));
        }
    }
    block_add_n<NL>(bd, aa.bdot);
}

// ------------------------------------------------------------------------------------------ host ------------------

template <int NL> void Plan::adjoint_pass(const AdjArgs &aa)
{
    hipStream_t st = ctx->stream;
    const AsmParams P = params();
    // far pairs with a closed-form point: at most the forward's count of closed-form far points.  When that exceeds
    // the queue's budget (2^28 entries; a kernel_far of 1e300 makes every far point closed-form: 1.3e10 on the swept
    // wing), the far tier goes in slices of source tiles, each queueing at most what its pairs can add.
    unsigned long long hst[16];
    read_stats(hst);
    const unsigned long long budget = 1ull << 28, per_col = (unsigned long long)ntt * TT_MAX * TS;
    int slice = nts;
    unsigned long long ecap64 = hst[12] + 1024;
    if (ecap64 > budget) {
        slice = (int)std::max<unsigned long long>(1, budget / per_col);
        ecap64 = std::min(budget, per_col * (unsigned long long)slice);
    }
    const unsigned int ecap = (unsigned int)ecap64;
    unsigned short *d_exact = dalloc<unsigned short>(ecap);
    int4 *d_esegs = dalloc<int4>((size_t)ntt * slice);
    unsigned int *d_ecount = dalloc<unsigned int>(2);
    auto free_exact = [&] {
        (void)hipFree(d_exact);
        (void)hipFree(d_esegs);
        (void)hipFree(d_ecount);
    };
    // the far and plain near tiers' points in the plan's precision (FP32 on the mixed path); the touching tiers, the
    // filaments and the freestream in FP64
    auto run = [&](auto tag) {
        using CT = decltype(tag);
        for (int sb = 0; sb < nts; sb += slice) {
            HIPCHECK(hipMemsetAsync(d_ecount, 0, 2 * sizeof(unsigned int), st));
            adj_far_pairs<NL, CT><<<dim3(ntt, std::min(slice, nts - sb)), 256, 0, st>>>(
                P, rt, d_tf, d_el, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl, d_tsc, d_tsn, d_cols, d_slotnode, aa, d_exact,
                d_esegs, d_ecount, ecap, sb);
            HIPCHECK(hipGetLastError());
            unsigned int ecnt[2] = {0, 0};
            HIPCHECK(hipMemcpyAsync(ecnt, d_ecount, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (ecnt[0] > ecap) {
                free_exact();
                throw std::runtime_error("adjoint: more closed-form far pairs than the forward's closed-form far points");
            }
            if (ecnt[1]) {
                adj_far_exact<NL, CT><<<ecnt[1], 256, 0, st>>>(P, rt, d_tf, d_el, d_tte, d_ttl, d_ttr, d_ttn, d_tse, d_tsl,
                                                               d_tsc, d_tsn, d_cols, d_slotnode, aa, d_exact, d_esegs);
                HIPCHECK(hipGetLastError());
            }
            HIPCHECK(hipStreamSynchronize(st));
        }
        free_exact();
        if (nseg) {
            const long long np = (long long)nt * nseg;
            adj_seg_far<NL><<<(unsigned)((np + 255) / 256), 256, 0, st>>>(P, rt, d_tf, d_el, d_cs, d_seg, d_segcol, aa);
            HIPCHECK(hipGetLastError());
        }
        // the non-far pairs of the last assembly, through the forward's own routing and kernels in adjoint mode
        if (nsimple) {
            touch_simple<NL><<<std::min<unsigned int>((nsimple + 3) / 4, 65535u), 256, 0, st>>>(
                P, rt, d_el, d_sigma, d_cs, d_cols, d_simple, nsimple, nullptr, nullptr, nullptr, nullptr, aa);
            HIPCHECK(hipGetLastError());
        }
        if (nnear) {
            near_thread<NL, CT><<<std::min<unsigned int>((nnear + 255) / 256, 65535u), 256, 0, st>>>(
                P, rt, d_el, d_sigma, d_cs, d_cols, d_near, nnear, nullptr, nullptr, nullptr, nullptr, aa);
            HIPCHECK(hipGetLastError());
        }
        // near_group takes one set per launch: with two it spills (560 bytes per lane under its two-wave cap) and runs
        // slower than two launches.  VP_ADJ_NEAR_ONE_LAUNCH=1 takes every set in one launch of the build at one wave per
        // SIMD (no spills), which the tests compare with the per-set launches.
        const char *one_env = getenv("VP_ADJ_NEAR_ONE_LAUNCH");
        const bool one_launch = NL > 1 && one_env && one_env[0] == '1';
        if (nnearm && one_launch) {
            near_group<NL, 1, CT><<<std::min<unsigned int>((nnearm + 3) / 4, 65535u), 256, 0, st>>>(
                P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, nnearm, nullptr, nullptr, nullptr, nullptr, aa);
            HIPCHECK(hipGetLastError());
        }
        for (int l = 0; nnearm && !one_launch && l < NL; l++) {
            AdjArgs a1 = aa;
            a1.lam += (size_t)l * aa.lstride;
            a1.G += l * aa.gstride;
            if (a1.bdot) a1.bdot += l;
            near_group<1, 3, CT><<<std::min<unsigned int>((nnearm + 3) / 4, 65535u), 256, 0, st>>>(
                P, rt, d_el, d_sigma, d_cs, d_cols, d_nearm, nnearm, nullptr, nullptr, nullptr, nullptr, a1);
            HIPCHECK(hipGetLastError());
        }
    };
    if (opts.precision == 1)
        run(float{});
    else
        run(double{});
    if (ngroups) {
        adj_group_eval<NL><<<std::min<unsigned int>((ngroups + 3) / 4, 65535u), 256, 0, st>>>(
            P, rt, d_el, d_cs, d_cols, d_seg, d_segcol, d_leaves, d_groups, ngroups, aa);
        HIPCHECK(hipGetLastError());
    }
    adj_freestream<NL><<<(nt + 255) / 256, 256, 0, st>>>(d_el, d_cs, nt, opts.test, aa);
    HIPCHECK(hipGetLastError());
}

// nlam sets of adjoint weights (nlam x nslot, row-major) -> nlam gradients ((nn + nw) x 3 each) and, with a direction
// dvinf, nlam values of lambdaᵀ b along it; ADJ_MAX sets per pass of the kernels.
void Plan::adjoint(int nlam, const double *lam_h, const double *mu_h, const double *vinf, const double *dvinf,
                   double *G_h, double *bdot_h)
{
    hipStream_t st = ctx->stream;
    if (!assembled) throw std::invalid_argument("the plan has not been assembled (vp_plan_assemble first)");
    if (nlam < 1) throw std::invalid_argument("nlam must be at least 1");
    const long long gstride = 3ll * (nn + nw);
    if (!d_lam) {
        d_lam = dalloc<double>((size_t)ADJ_MAX * nslot);
        d_mu = dalloc<double>(ncol);
        d_G = dalloc<double>((size_t)ADJ_MAX * gstride);
        d_slotnode = dalloc<int>(nslot);
        d_bdot = dalloc<double>(ADJ_MAX);
        owned.insert(owned.end(), {d_lam, d_mu, d_G, d_slotnode, d_bdot});
        slot_nodes<<<(nt + 255) / 256, 256, 0, st>>>(d_cs, d_nodes, nt, d_slotnode);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(d_mu, mu_h, sizeof(double) * ncol, hipMemcpyHostToDevice, st));
    AdjArgs aa{d_lam, d_mu, d_G, d_nodes, vinf[0], vinf[1], vinf[2]};
    if (dvinf) {
        aa.dx = dvinf[0];
        aa.dy = dvinf[1];
        aa.dz = dvinf[2];
        aa.bdot = d_bdot;
    }
    aa.lstride = nslot;
    aa.gstride = gstride;
    for (int l0 = 0; l0 < nlam; l0 += ADJ_MAX) {
        const int nl = std::min(ADJ_MAX, nlam - l0);
        HIPCHECK(hipMemcpyAsync(d_lam, lam_h + (size_t)l0 * nslot, sizeof(double) * nslot * nl, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemsetAsync(d_G, 0, sizeof(double) * gstride * nl, st));
        HIPCHECK(hipMemsetAsync(d_bdot, 0, sizeof(double) * ADJ_MAX, st));
        if (nl == 2)
            adjoint_pass<2>(aa);
        else
            adjoint_pass<1>(aa);
        HIPCHECK(hipMemcpyAsync(G_h + (size_t)l0 * gstride, d_G, sizeof(double) * gstride * nl, hipMemcpyDeviceToHost, st));
        if (dvinf && bdot_h) HIPCHECK(hipMemcpyAsync(bdot_h + l0, d_bdot, sizeof(double) * nl, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
}

}  // namespace vp
