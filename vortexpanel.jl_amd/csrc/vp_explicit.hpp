// The right hand side of known vortex elements, explicit_rhs (src/assembly/explicit.jl:31-86), on the device:
//
//  * sheets and filaments take the assembly's own tiers: a wake-only plan (its wake = the known sheets and filaments,
//    one node per corner or filament end) assembles their columns, and b_e = -A_wake mu (wake_product);
//  * particles (Rosenhead-Moore blobs): far from a test triangle (centroid distance > far_ratio max(diam, core)) their
//    velocity is summed at the far rule's 3 points, one thread per test triangle over LDS tiles of blobs
//    (blob_rhs_far); otherwise the pair goes to a list and gets the near rule of the blob (near_rule with
//    distance_to(blob, p) = |p - blob|, assemble.jl:69-72, 131-142), one thread per pair (blob_rhs_near).
//
// Included by vp_assemble.hip (it uses the plan's elements, far points and rules).

namespace vp {

constexpr int BLOB_TILE = 256;

// out[i] = sum_j A[i, nslot + j] mu[j] (the wake columns of a column-major A)
__global__ void wake_matvec(const double *__restrict__ A, long long lda, int nslot, int nw, const double *__restrict__ mu,
                            double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslot) return;
    double acc = 0.0;
    for (int j = 0; j < nw; j++) acc += A[(long long)(nslot + j) * lda + i] * mu[j];
    out[i] = acc;
}

// Far blobs: rows[t][a] -= sum_q W_q psi_a(q) n_T . u_q, u_q the far blobs' velocity at far point q (explicit.jl:60-79).
// Near pairs are appended to near_list.
__global__ __launch_bounds__(256) void blob_rhs_far(AsmParams P, const TestFar *__restrict__ tf, const double *__restrict__ bx,
                                                    const double *__restrict__ ba, const double *__restrict__ bd, int nb,
                                                    int chunk, int2 *__restrict__ near_list,
                                                    unsigned int *__restrict__ near_count, unsigned int near_cap,
                                                    double *__restrict__ partial)
{
    __shared__ double Bt[BLOB_TILE][8];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < P.nt;
    TestFar F;
    if (live) F = tf[t];
    V3<double> u[3] = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0)};
    const int b0 = blockIdx.y * chunk, b1 = min(nb, b0 + chunk);
    for (int base = b0; base < b1; base += BLOB_TILE) {
        const int nk = min(BLOB_TILE, b1 - base);
        __syncthreads();
        if (threadIdx.x < nk) {
            const int b = base + threadIdx.x;
            for (int k = 0; k < 3; k++) {
                Bt[threadIdx.x][k] = bx[3 * b + k];
                Bt[threadIdx.x][3 + k] = ba[3 * b + k];
            }
            Bt[threadIdx.x][6] = bd[b];
        }
        __syncthreads();
        if (!live) continue;
        for (int k = 0; k < nk; k++) {
            const V3<double> xb = mk3(Bt[k][0], Bt[k][1], Bt[k][2]);
            const double del = Bt[k][6];
            if (cls_far(cls_sub(xb, ld3(F.c)), P.far_ratio * fmax(F.diam, del))) {
                const V3<double> al = mk3(Bt[k][3], Bt[k][4], Bt[k][5]);
                for (int q = 0; q < 3; q++) u[q] = u[q] + blob_vel(ld3(F.x[q]) - xb, al, del * del);
            } else {
                const unsigned int pos = atomicAdd(near_count, 1u);
                if (pos < near_cap) near_list[pos] = make_int2(t, base + k);
            }
        }
    }
    if (!live) return;
    const V3<double> n = ld3(F.n);
    double *p = partial + 3 * ((size_t)blockIdx.y * P.nt + t);
    for (int a = 0; a < 3; a++) {
        double r = 0.0;
        for (int q = 0; q < 3; q++) r -= F.wpsi[q][a] * dot(u[q], n);
        p[a] = r;
    }
}

// rows[t][a] = sum over chunks (fixed order) of the far partial sums
__global__ void blob_rhs_reduce(const double *__restrict__ partial, int nchunk, int nt, double *__restrict__ rows)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nt) return;
    double acc = 0.0;
    for (int c = 0; c < nchunk; c++) acc += partial[(size_t)c * 3 * nt + i];
    rows[i] = acc;
}

// Near blob pairs: near_rule(T, blob) (recursive 4-way split while |centroid - blob| < near_ratio size, depth
// near_depth, 7-point leaves), one thread per pair; added to rows with atomics.
__global__ void blob_rhs_near(AsmParams P, RuleTab rt, const Elem *__restrict__ el, const double *__restrict__ bx,
                              const double *__restrict__ ba, const double *__restrict__ bd, const int2 *__restrict__ list,
                              unsigned int npairs, double *__restrict__ rows)
{
    const unsigned int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= npairs) return;
    const int t = list[pi].x, b = list[pi].y;
    const Elem &T = el[t];
    const V3<double> X[3] = {ld3(T.v[0]), ld3(T.v[1]), ld3(T.v[2])};
    const V3<double> n = ld3(T.n), xb = ld3(bx + 3 * b), al = ld3(ba + 3 * b);
    const double d2 = bd[b] * bd[b];
    const int nr = rt.n[R_NEAR], off = rt.off[R_NEAR];
    double acc[3] = {0.0, 0.0, 0.0};
    // explicit stack of pieces (at most 3 per level + 1)
    struct Pc {
        double b[3][3];
        double frac;
        int level;
    };
    Pc stk[3 * 8 + 1];
    int sp = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) stk[0].b[i][j] = i == j ? 1.0 : 0.0;
    stk[0].frac = 1.0;
    stk[0].level = 0;
    sp = 1;
    while (sp > 0) {
        const Pc pc = stk[--sp];
        bool split = false;
        {
#pragma clang fp contract(off)
            const V3<double> p1 = cls_point(X, pc.b[0]), p2 = cls_point(X, pc.b[1]), p3 = cls_point(X, pc.b[2]);
            const double size = fmax(fmax(cls_norm(cls_sub(p2, p1)), cls_norm(cls_sub(p3, p2))), cls_norm(cls_sub(p1, p3)));
            V3<double> c = (p1 + p2) + p3;
            c = mk3(c.x / 3.0, c.y / 3.0, c.z / 3.0);
            split = pc.level < P.near_depth && sp + 4 <= 3 * 8 + 1 && cls_norm(cls_sub(xb, c)) < P.near_ratio * size;
        }
        if (split) {
            double m12[3], m23[3], m31[3];
            for (int j = 0; j < 3; j++) {
                m12[j] = (pc.b[0][j] + pc.b[1][j]) / 2;
                m23[j] = (pc.b[1][j] + pc.b[2][j]) / 2;
                m31[j] = (pc.b[2][j] + pc.b[0][j]) / 2;
            }
            // quarters (assemble.jl:79-82), pushed in reverse so they are evaluated in order
            const double *ch[4][3] = {{m12, m23, m31}, {m31, m23, pc.b[2]}, {m12, pc.b[1], m23}, {pc.b[0], m12, m31}};
            for (int c = 0; c < 4; c++) {
                Pc &q = stk[sp++];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) q.b[i][j] = ch[c][i][j];
                q.frac = pc.frac / 4;
                q.level = pc.level + 1;
            }
            continue;
        }
        for (int q = 0; q < nr; q++) {
            const double *m = rt.l + 3 * (off + q);
            double l[3];
            for (int j = 0; j < 3; j++) l[j] = (m[0] * pc.b[0][j] + m[1] * pc.b[1][j]) + m[2] * pc.b[2][j];
            double psi[3];
            test_weights(P.test, l, psi);
            const double W = rt.w[off + q] * pc.frac * T.area;
            const double un = dot(blob_vel(cls_point(X, l) - xb, al, d2), n);
            for (int a = 0; a < 3; a++) acc[a] -= W * un * psi[a];
        }
    }
    for (int a = 0; a < 3; a++) unsafeAtomicAdd(&rows[3 * t + a], acc[a]);
}

// rows (nt x 3) -> b[corner_slot] (atomics: slots are shared)
__global__ void rows_to_slots(const double *__restrict__ rows, const int *__restrict__ corner_slot, int nt,
                              double *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * nt) unsafeAtomicAdd(&b[corner_slot[i]], rows[i]);
}

// New coordinates for the wake's nodes, same triangles and filaments (a free wake between steps): the wake's element
// records and filament ends are rebuilt in place.  The tiles and the work lists depend on the topology alone (every
// pair is classified again at each assembly), so the plan is reused without a rebuild.
void Plan::update_wake(const double *wake_x)
{
    if (!nw) throw std::invalid_argument("the plan has no wake");
    hipStream_t st = ctx->stream;
    HIPCHECK(hipMemcpyAsync(d_x + 3 * (size_t)nn, wake_x, sizeof(double) * 3 * nw, hipMemcpyHostToDevice, st));
    if (nwt) {
        elem_setup<<<(nwt + 255) / 256, 256, 0, st>>>(d_x, d_nodes + 3 * (size_t)nt, nwt, d_el + nt);
        HIPCHECK(hipGetLastError());
        std::vector<Elem> hel(nwt);
        HIPCHECK(hipMemcpyAsync(hel.data(), d_el + nt, sizeof(Elem) * nwt, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        for (int e = 0; e < nwt; e++)
            if (hel[e].pad0 != 0.0) throw std::invalid_argument("degenerate triangle " + std::to_string(nt + e));
    }
    if (nseg) {
        std::vector<double> segx((size_t)6 * nseg);
        for (int f = 0; f < nseg; f++)
            for (int k = 0; k < 3; k++) {
                segx[6 * f + k] = wake_x[3 * h_fil[2 * f] + k];
                segx[6 * f + 3 + k] = wake_x[3 * h_fil[2 * f + 1] + k];
            }
        HIPCHECK(hipMemcpyAsync(d_seg, segx.data(), sizeof(double) * segx.size(), hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    assembled = false;
}

void Plan::wake_product(const double *mu_nodes, double *out)
{
    hipStream_t st = ctx->stream;
    if (!assembled) throw std::invalid_argument("the plan has not been assembled (vp_plan_assemble first)");
    double *dmu = scratch_d(7, std::max(nw, 1)), *dout = scratch_d(8, nslot);
    if (nw) HIPCHECK(hipMemcpyAsync(dmu, mu_nodes, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    wake_matvec<<<(nslot + 255) / 256, 256, 0, st>>>(A, lda, nslot, nw, dmu, dout);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, dout, sizeof(double) * nslot, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

void Plan::blob_rhs(int nb, const double *bx_h, const double *ba_h, const double *bd_h, double *out)
{
    hipStream_t st = ctx->stream;
    const AsmParams P = params();
    // the plan's grow-only buffers (the call ends with a synchronization, so the next call may reuse them); on an
    // error the stream is drained before the buffers can be reused
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain_{st};
    double *bx = scratch_d(0, (size_t)3 * std::max(nb, 1)), *ba = scratch_d(1, (size_t)3 * std::max(nb, 1)),
           *bd = scratch_d(2, std::max(nb, 1));
    if (nb) {
        HIPCHECK(hipMemcpyAsync(bx, bx_h, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(ba, ba_h, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(bd, bd_h, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    }
    double *rows = scratch_d(3, (size_t)3 * nt), *b = scratch_d(4, nslot);
    HIPCHECK(hipMemsetAsync(rows, 0, sizeof(double) * 3 * nt, st));
    HIPCHECK(hipMemsetAsync(b, 0, sizeof(double) * nslot, st));
    if (nb) {
        const int tblocks = (nt + 255) / 256, ntiles = (nb + BLOB_TILE - 1) / BLOB_TILE;
        const int nchunk = std::max(1, std::min(ntiles, (2048 + tblocks - 1) / tblocks));
        const int chunk = ((ntiles + nchunk - 1) / nchunk) * BLOB_TILE;
        const int nch = (nb + chunk - 1) / chunk;
        double *partial = scratch_d(5, (size_t)3 * nt * nch);
        unsigned int *cnt = (unsigned int *)scratch_d(6, 1);
        for (int attempt = 0;; attempt++) {
            if (attempt == 3) throw std::runtime_error("near blob list did not fit after regrowing");
            if (!d_bnear) {
                bnear_cap = std::max(bnear_cap, (unsigned int)std::min<long long>((long long)nt * 8 + 1024, 1ll << 30));
                d_bnear = dalloc<int2>(bnear_cap);
            }
            HIPCHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned int), st));
            blob_rhs_far<<<dim3(tblocks, nch), 256, 0, st>>>(P, d_tf, bx, ba, bd, nb, chunk, d_bnear, cnt, bnear_cap, partial);
            HIPCHECK(hipGetLastError());
            unsigned int nn_ = 0;
            HIPCHECK(hipMemcpyAsync(&nn_, cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (nn_ > bnear_cap) {
                (void)hipFree(d_bnear);
                d_bnear = nullptr;
                bnear_cap = nn_ + 1024;
                continue;
            }
            blob_rhs_reduce<<<(3 * nt + 255) / 256, 256, 0, st>>>(partial, nch, nt, rows);
            HIPCHECK(hipGetLastError());
            if (nn_) {
                blob_rhs_near<<<(nn_ + 255) / 256, 256, 0, st>>>(P, rt, d_el, bx, ba, bd, d_bnear, nn_, rows);
                HIPCHECK(hipGetLastError());
            }
            break;
        }
    }
    rows_to_slots<<<(3 * nt + 255) / 256, 256, 0, st>>>(rows, d_cs, nt, b);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, b, sizeof(double) * nslot, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace vp
