// The bordered dense solve on gfx950: build the bordered matrix on the device, randomize it with a depth-2 recursive
// butterfly transform, factor it by a blocked LU without pivoting whose trailing updates run on MFMA (FP64 MFMA for
// the plain FP64 path, FP32 MFMA for the mixed path), and solve with FP64 iterative refinement on the original
// FP64 matrix.
//
// Reference: src/assembly/solve.jl:15-99 (bordered_system, bordered_matrix, solve_galerkin, factorize_galerkin) and
// src/dofmap/system.jl:26-50 (reduce_system, bordered_rows).  The reference factors with LAPACK getrf (partial
// pivoting).  Partial pivoting needs n sequential pivot searches across the whole column; here the random butterfly
// transform (Parker 1995; Baboulin, Dongarra et al. 2013, MAGMA's gesv_rbt) makes pivoting unnecessary with high
// probability, and refinement in FP64 (as LAPACK's dsgesv) brings the solution to FP64 accuracy.  The condition
// estimate is Hager-Higham's 1-norm estimator (LAPACK dlacn2, the algorithm behind the reference's gecon call).
//
// Layouts: all device matrices are row-major with a leading dimension; right hand sides are n x nrhs row-major.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdlib>
#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "vp_capi_internal.hpp"

namespace vp {

#define HIPCHECK2(x)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));                \
    } while (0)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Both 16x16x4 MFMAs take A[m][k] from lane m + 16k and B[k][n] from lane n + 16k.  Their 16x16 results differ:
// the FP32 one holds D[4(l/16) + v][l%16] in lane l, register v; the FP64 one D[(l/16) + 4v][l%16] (read on gfx950).
template <> struct Mfma<double> {
    typedef d4 acc;
    static __device__ __forceinline__ acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
    typedef f4 acc;
    static __device__ __forceinline__ acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int NB = 64;    // diagonal block width of the LU (the blocks whose inverses the triangular solves use)
constexpr int NB2 = 256;  // outer panel width of the LU: the K of its trailing updates

// ------------------------------------------------------------------------------------------ GEMM (MFMA) ------------

// C = A B (SET) or C -= A B (SUB), row-major, arbitrary M, N, K; with TA, A is read transposed (element (m, k) at
// A[k lda + m]).  Workgroup tile BM x BN, four waves in a 2 x 2 grid each owning a (BM/2) x (BN/2) sub-tile of 16 x 16
// MFMA blocks.  A is staged in LDS as As[k][m] and B as Bs[k][n], so each lane of an MFMA reads A[m = l%16][k = l/16]
// and B[k = l/16][n = l%16].  The next k-tile is loaded into registers while the current one is multiplied.  In place
// (C aliasing B) is allowed when one workgroup row covers all of M (the whole of B's tile is read before C is
// written), and (C aliasing A) when one workgroup column covers all of N.
template <typename T, int BM, int BN, int BK, bool SUB, bool TA>
__global__ __launch_bounds__(256) void gemm(int M, int N, int K, const T *A, long long lda, const T *B, long long ldb,
                                            T *C, long long ldc)
{
    constexpr int WM = BM / 2, WN = BN / 2, TI = WM / 16, TJ = WN / 16;
    constexpr int LA = BM * BK / 256, LB = BK * BN / 256;  // elements each thread stages per k-tile
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    typename Mfma<T>::acc acc[TI][TJ];
    for (int i = 0; i < TI; i++)
        for (int j = 0; j < TJ; j++)
            for (int v = 0; v < 4; v++) acc[i][j][v] = T(0);
    T ra[LA], rb[LB];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < LA; u++) {
            const int idx = tid + 256 * u;
            int r, k;
            if (TA) {
                k = idx / BM;
                r = idx - k * BM;
            } else {
                r = idx / BK;
                k = idx - r * BK;
            }
            const long long gr = m0 + r;
            const int gk = k0 + k;
            ra[u] = (gr < M && gk < K) ? (TA ? A[(long long)gk * lda + gr] : A[gr * lda + gk]) : T(0);
        }
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int idx = tid + 256 * u;
            const int k = idx / BN, c = idx - k * BN;
            const int gk = k0 + k;
            const long long gc = n0 + c;
            rb[u] = (gk < K && gc < N) ? B[(long long)gk * ldb + gc] : T(0);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < LA; u++) {
            const int idx = tid + 256 * u;
            if (TA) {
                const int k = idx / BM;
                As[k][idx - k * BM] = ra[u];
            } else {
                const int r = idx / BK;
                As[idx - r * BK][r] = ra[u];
            }
        }
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int idx = tid + 256 * u;
            const int k = idx / BN;
            Bs[k][idx - k * BN] = rb[u];
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        store();
        __syncthreads();
        if (k0 + BK < K) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[TI], b[TJ];
#pragma unroll
            for (int i = 0; i < TI; i++) a[i] = As[kk + (lane >> 4)][wr * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TJ; j++) b[j] = Bs[kk + (lane >> 4)][wc * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TI; i++)
#pragma unroll
                for (int j = 0; j < TJ; j++) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                long long r = m0 + wr * WM + i * 16 + Mfma<T>::row(lane, v);
                long long c = n0 + wc * WN + j * 16 + (lane & 15);
                if (r < M && c < N) {
                    if (SUB)
                        C[r * ldc + c] -= acc[i][j][v];
                    else
                        C[r * ldc + c] = acc[i][j][v];
                }
            }
}

// Tile shapes (BM, BN, BK) the autotuner chooses among, per precision; every one covers the 64-wide panels, which the
// in-place panel solves need.
struct TileShape {
    int bm, bn, bk;
};
constexpr int kNumTiles = 8;
constexpr TileShape kTiles64[kNumTiles] = {{64, 64, 16}, {128, 64, 16}, {64, 128, 16}, {128, 128, 16},
                                           {128, 128, 8}, {64, 64, 32}, {64, 64, 8},   {128, 64, 8}};
constexpr TileShape kTiles32[kNumTiles] = {{128, 128, 16}, {64, 64, 16},  {128, 64, 16}, {128, 128, 32},
                                           {256, 128, 16}, {128, 64, 32}, {64, 128, 16}, {128, 64, 8}};

template <typename T, int BM, int BN, int BK, bool SUB, bool TA>
void launch_tile(hipStream_t s, int M, int N, int K, const T *A, long long lda, const T *B, long long ldb, T *C,
                 long long ldc)
{
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
    gemm<T, BM, BN, BK, SUB, TA><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc);
    HIPCHECK2(hipGetLastError());
}

// The library candidate of the trailing update, index kNumTiles after the tile shapes: rocBLAS's GEMM for the plain
// C -= A B (row-major here, so rocBLAS's column-major Cᵀ -= Bᵀ Aᵀ), one handle per device made on first use and set
// to the caller's stream.  Only the factorization's trailing update uses it (never inside a captured graph).
constexpr int kBlasCfg = kNumTiles;
static std::mutex g_blas_mutex;
static rocblas_handle g_blas[64] = {};

static rocblas_handle blas_handle(hipStream_t s)
{
    int dev = 0;
    HIPCHECK2(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw std::runtime_error("rocBLAS: device index out of range");
    std::lock_guard<std::mutex> lock(g_blas_mutex);
    if (!g_blas[dev] && rocblas_create_handle(&g_blas[dev]) != rocblas_status_success)
        throw std::runtime_error("rocblas_create_handle failed");
    if (rocblas_set_stream(g_blas[dev], s) != rocblas_status_success) throw std::runtime_error("rocblas_set_stream failed");
    return g_blas[dev];
}

template <typename T>
void blas_gemm_sub(hipStream_t s, int M, int N, int K, const T *A, long long lda, const T *B, long long ldb, T *C,
                   long long ldc)
{
    rocblas_handle h = blas_handle(s);
    const T alpha = T(-1), beta = T(1);
    rocblas_status st;
    if constexpr (sizeof(T) == 8)
        st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &alpha, B, (int)ldb, A, (int)lda,
                           &beta, C, (int)ldc);
    else
        st = rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &alpha, B, (int)ldb, A, (int)lda,
                           &beta, C, (int)ldc);
    if (st != rocblas_status_success) throw std::runtime_error("rocBLAS GEMM failed");
}

template <typename T, bool SUB, bool TA>
void launch_gemm_cfg(int cfg, hipStream_t s, int M, int N, int K, const T *A, long long lda, const T *B, long long ldb,
                     T *C, long long ldc)
{
    if (M <= 0 || N <= 0 || K <= 0) return;
    if (cfg == kBlasCfg) {
        if constexpr (SUB && !TA) return blas_gemm_sub<T>(s, M, N, K, A, lda, B, ldb, C, ldc);
        cfg = 0;  // the library candidate covers the trailing update's form only
    }
    if constexpr (sizeof(T) == 8) {
        switch (cfg) {
        case 1: return launch_tile<T, 128, 64, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 2: return launch_tile<T, 64, 128, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 3: return launch_tile<T, 128, 128, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 4: return launch_tile<T, 128, 128, 8, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 5: return launch_tile<T, 64, 64, 32, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 6: return launch_tile<T, 64, 64, 8, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 7: return launch_tile<T, 128, 64, 8, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        default: return launch_tile<T, 64, 64, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        }
    } else {
        switch (cfg) {
        case 1: return launch_tile<T, 64, 64, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 2: return launch_tile<T, 128, 64, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 3: return launch_tile<T, 128, 128, 32, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 4: return launch_tile<T, 256, 128, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 5: return launch_tile<T, 128, 64, 32, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 6: return launch_tile<T, 64, 128, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        case 7: return launch_tile<T, 128, 64, 8, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        default: return launch_tile<T, 128, 128, 16, SUB, TA>(s, M, N, K, A, lda, B, ldb, C, ldc);
        }
    }
}

// The autotuner: the trailing update's GEMM per precision (one of the tile shapes, or rocBLAS), chosen by timing every
// candidate (HIP events, three runs after a warm-up) on the first factorization's own update shape (C -= A B with
// K = NB2, M = N = the matrix size less one panel, at most 8192) in its factor buffer before the matrix is written
// there; kept for the process.  VP_GEMM_TILE=<index> forces a candidate (8: rocBLAS).
struct GemmTuning {
    int cfg = -1;
    double ms[kNumTiles + 1] = {};
    int size = 0;
};
static GemmTuning g_tune[2];  // [0] FP64, [1] FP32
static std::mutex g_tune_mutex;

template <typename T> int tuned_cfg(hipStream_t s, T *scratch, long long ld, int npad)
{
    const int p = sizeof(T) == 8 ? 0 : 1;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    GemmTuning &tu = g_tune[p];
    if (tu.cfg >= 0) return tu.cfg;
    if (const char *env = std::getenv("VP_GEMM_TILE")) {
        const int c = std::atoi(env);
        tu.cfg = (c >= 0 && c <= kBlasCfg) ? c : 0;
        return tu.cfg;
    }
    const int m = std::min(npad - NB2, 8192);
    if (m < 1024) return 0;  // too small to time well: the default shape, tuned later on a larger matrix
    HIPCHECK2(hipMemsetAsync(scratch, 0, sizeof(T) * (size_t)ld * (m + NB2), s));
    hipEvent_t e0, e1;
    HIPCHECK2(hipEventCreate(&e0));
    HIPCHECK2(hipEventCreate(&e1));
    int best = 0;
    for (int c = 0; c <= kBlasCfg; c++) {
        T *A = scratch + NB2 * ld, *B = scratch + NB2, *C = scratch + NB2 * ld + NB2;
        launch_gemm_cfg<T, true, false>(c, s, m, m, NB2, A, ld, B, ld, C, ld);
        HIPCHECK2(hipEventRecord(e0, s));
        for (int r = 0; r < 3; r++) launch_gemm_cfg<T, true, false>(c, s, m, m, NB2, A, ld, B, ld, C, ld);
        HIPCHECK2(hipEventRecord(e1, s));
        HIPCHECK2(hipEventSynchronize(e1));
        float t = 0;
        HIPCHECK2(hipEventElapsedTime(&t, e0, e1));
        tu.ms[c] = t / 3;
        if (tu.ms[c] < tu.ms[best]) best = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tu.cfg = best;
    tu.size = m;
    return best;
}

std::string tuning_report()
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    std::string out = "{";
    for (int p = 0; p < 2; p++) {
        const GemmTuning &tu = g_tune[p];
        const TileShape *sh = p == 0 ? kTiles64 : kTiles32;
        auto name = [&](int c) {
            if (c == kBlasCfg) return std::string("rocblas");
            return std::to_string(sh[c].bm) + "x" + std::to_string(sh[c].bn) + "x" + std::to_string(sh[c].bk);
        };
        out += std::string(p ? ", " : "") + (p == 0 ? "\"fp64\": {" : "\"fp32\": {");
        out += "\"chosen\": " + (tu.cfg < 0 ? std::string("null") : "\"" + name(tu.cfg) + "\"");
        out += ", \"tuned_on\": " + std::to_string(tu.size) + ", \"ms\": {";
        for (int c = 0; c <= kBlasCfg; c++) out += std::string(c ? ", " : "") + "\"" + name(c) + "\": " + std::to_string(tu.ms[c]);
        out += "}}";
    }
    return out + "}";
}

template <typename T, bool SUB>
void launch_gemm(hipStream_t s, int M, int N, int K, const T *A, long long lda, const T *B, long long ldb, T *C, long long ldc)
{
    launch_gemm_cfg<T, SUB, false>(0, s, M, N, K, A, lda, B, ldb, C, ldc);
}

// ------------------------------------------------------------------------------------------ LU pieces --------------


// Factor the NB x NB diagonal block at (k0, k0) in place without pivoting (D = L U, L unit lower), and write the
// inverses of L and U (NB x NB, row-major) for the panel's triangular solves.  A zero or non-finite pivot sets *bad.
// Every stage is right-looking over the whole workgroup: the elimination, then L^-1 and U^-1 from the identity by
// rank-1 updates (row k of the inverse is final once the updates of the rows before it, or after it for U, are in).
// (A one-wave version, lane per column with no workgroup barriers, ran 2x slower: its serial LDS loops leave the
// latency unhidden; C4's FP64 factorization 0.40 -> 0.47 s.)
template <typename T>
__global__ __launch_bounds__(256) void diag_lu(T *G, long long ld, int k0, int nb, T *Linv, T *Uinv, int *bad)
{
    __shared__ T D[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < NB * NB; idx += 256) {
        int i = idx / NB, j = idx % NB;
        D[i][j] = (i < nb && j < nb) ? G[(long long)(k0 + i) * ld + k0 + j] : T(i == j ? 1 : 0);
        X[i][j] = T(i == j ? 1 : 0);
    }
    __syncthreads();
    for (int k = 0; k < nb; k++) {
        const T p = D[k][k];
        if (tid == 0 && (!(p != T(0)) || !isfinite((double)p))) atomicExch(bad, 1);
        const T ip = T(1) / p;
        for (int i = k + 1 + tid; i < nb; i += 256) D[i][k] *= ip;
        __syncthreads();
        const int m = nb - k - 1;
        for (int idx = tid; idx < m * m; idx += 256) {
            int i = k + 1 + idx / m, j = k + 1 + idx % m;
            D[i][j] -= D[i][k] * D[k][j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nb * nb; idx += 256) {
        int i = idx / nb, j = idx % nb;
        G[(long long)(k0 + i) * ld + k0 + j] = D[i][j];
    }
    // L^-1 (unit lower): after step k, rows i > k have lost L[i][k] x (row k)
    for (int k = 0; k < NB - 1; k++) {
        const int w = k + 1, cnt = (NB - 1 - k) * w;
        for (int idx = tid; idx < cnt; idx += 256) {
            const int i = k + 1 + idx / w, j = idx % w;
            X[i][j] -= D[i][k] * X[k][j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < NB * NB; idx += 256) Linv[idx] = X[idx / NB][idx % NB];
    __syncthreads();
    for (int idx = tid; idx < NB * NB; idx += 256) X[idx / NB][idx % NB] = T(idx / NB == idx % NB ? 1 : 0);
    __syncthreads();
    // U^-1: from the last row up, row k is scaled by 1/U[k][k], then rows i < k lose U[i][k] x (row k)
    for (int k = NB - 1; k >= 0; k--) {
        const T ip = T(1) / D[k][k];
        for (int j = k + tid; j < NB; j += 256) X[k][j] *= ip;
        __syncthreads();
        const int w = NB - k, cnt = k * w;
        for (int idx = tid; idx < cnt; idx += 256) {
            const int i = idx / w, j = k + idx % w;
            X[i][j] -= D[i][k] * X[k][j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < NB * NB; idx += 256) Uinv[idx] = X[idx / NB][idx % NB];
}

// ------------------------------------------------------------------------------------------ bordered matrix --------

// M (row-major, n x n, ld) rows/cols < nq: M[p][q] = sum_i sum_j C[i,p] Cf[j,q] A[i + j lda]  (Cᵀ A [C; W]),
// with C given by column (CSC: cp/ci/cv over nq columns) and Cf = [C; W] by column (fp/fi/fv).
// Threads run along p (coalesced reads of A's columns); a 64 x 64 LDS tile transposes to row-major writes.
__global__ __launch_bounds__(256) void reduce_tile(const double *__restrict__ A, long long lda, int nq,
                                                   const int *__restrict__ cp, const int *__restrict__ ci,
                                                   const double *__restrict__ cv, const int *__restrict__ fp,
                                                   const int *__restrict__ fi, const double *__restrict__ fv,
                                                   double *__restrict__ M, long long ld)
{
    __shared__ double t[64][65];
    const int p0 = blockIdx.x * 64, q0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int p = p0 + tx;
    for (int qq = ty; qq < 64; qq += 4) {
        int q = q0 + qq;
        double s = 0.0;
        if (p < nq && q < nq) {
            for (int a = fp[q]; a < fp[q + 1]; a++) {
                const double *col = A + (long long)fi[a] * lda;
                double cs = 0.0;
                for (int b = cp[p]; b < cp[p + 1]; b++) cs += cv[b] * col[ci[b]];
                s += fv[a] * cs;
            }
        }
        t[qq][tx] = s;
    }
    __syncthreads();
    for (int pp = ty; pp < 64; pp += 4) {
        int pr = p0 + pp, qc = q0 + tx;
        if (pr < nq && qc < nq) M[(long long)pr * ld + qc] = t[tx][pp];
    }
}

__global__ void diag_abs_sum(const double *M, long long ld, int nq, double *out)
{
    __shared__ double s[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < nq; i += 256) v += fabs(M[(long long)i * ld + i]);
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// border rows/columns (solve.jl:30-35) and the identity on the padding
__global__ void border_fill(double *M, long long ld, int nq, int nb, int npad, const double *e, const double *c,
                            const double *scale_e, const double *scale_c)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // part 1: columns nq..nq+nb-1 for rows < nq, and rows nq..nq+nb-1 for columns < nq
    if (idx < (long long)nq * nb) {
        int k = (int)(idx / nq), p = (int)(idx % nq);
        M[(long long)p * ld + nq + k] = scale_e[k] * e[(long long)k * nq + p];
        M[(long long)(nq + k) * ld + p] = scale_c[k] * c[(long long)k * nq + p];
    }
    // part 2: the (nb x nb) corner is zero; padding rows/columns: zero with 1 on the diagonal
    long long j = idx;
    int n = nq + nb;
    if (j < (long long)nb * nb) M[(long long)(nq + j / nb) * ld + nq + j % nb] = 0.0;
    if (j < (long long)(npad - n) * npad) {
        int r = n + (int)(j / npad), cc = (int)(j % npad);
        M[(long long)r * ld + cc] = r == cc ? 1.0 : 0.0;
        M[(long long)cc * ld + r] = r == cc ? 1.0 : 0.0;
    }
}

// X (npad x nrhs, row-major, in place) <- L^-T U^-T X, the transposed solve with the same factors: forward through
// Uᵀ (its diagonal blocks' inverses are uinvᵀ, its block (j, k) below the diagonal is U(k, j)ᵀ), then backward through
// Lᵀ; the transposes are the GEMM's transposed-A reads, so nothing is copied.
template <typename T> void lu_solve_t(hipStream_t s, const T *G, long long ld, int npad, const T *linv, const T *uinv, T *X, int nrhs)
{
    for (int k0 = 0; k0 < npad; k0 += NB) {
        int nb = std::min(NB, npad - k0);
        T *xk = X + (long long)k0 * nrhs;
        launch_gemm_cfg<T, false, true>(0, s, nb, nrhs, nb, uinv + (size_t)(k0 / NB) * NB * NB, NB, xk, nrhs, xk, nrhs);
        int m = npad - k0 - nb;
        if (m > 0)
            launch_gemm_cfg<T, true, true>(0, s, m, nrhs, nb, G + (long long)k0 * ld + k0 + nb, ld, xk, nrhs,
                                           xk + (long long)nb * nrhs, nrhs);
    }
    int last = ((npad - 1) / NB) * NB;
    for (int k0 = last; k0 >= 0; k0 -= NB) {
        int nb = std::min(NB, npad - k0);
        T *xk = X + (long long)k0 * nrhs;
        launch_gemm_cfg<T, false, true>(0, s, nb, nrhs, nb, linv + (size_t)(k0 / NB) * NB * NB, NB, xk, nrhs, xk, nrhs);
        if (k0 > 0) launch_gemm_cfg<T, true, true>(0, s, k0, nrhs, nb, G + (long long)k0 * ld, ld, xk, nrhs, X, nrhs);
    }
}

// ------------------------------------------------------------------------------------------ butterflies -----------

// Depth-2 recursive butterfly on both sides, G' = Uᵀ M V (one pass: every thread owns a 4 x 4 group of entries at
// rows {i, i+h, i+2h, i+3h} and columns {j, j+h, j+2h, j+3h}, h = npad/4).  ru/rv hold, per side, the level-1
// diagonals (R, S of size npad/2) then the level-2 ones (Ra, Sa, Rb, Sb of size npad/4).
__device__ __forceinline__ void bfly4(double *v, const double *r, int i, int h)
{
    // level 2 within halves, then level 1:  (B^T x)_top = R (x_top + x_bot)/sqrt2, (B^T x)_bot = S (x_top - x_bot)/sqrt2
    const double s2 = 0.70710678118654752440;
    const double *R1 = r, *S1 = r + 2 * h, *Ra = r + 4 * h, *Sa = r + 5 * h, *Rb = r + 6 * h, *Sb = r + 7 * h;
    double u0 = Ra[i] * (v[0] + v[1]) * s2, u1 = Sa[i] * (v[0] - v[1]) * s2;
    double u2 = Rb[i] * (v[2] + v[3]) * s2, u3 = Sb[i] * (v[2] - v[3]) * s2;
    v[0] = R1[i] * (u0 + u2) * s2;
    v[1] = R1[i + h] * (u1 + u3) * s2;
    v[2] = S1[i] * (u0 - u2) * s2;
    v[3] = S1[i + h] * (u1 - u3) * s2;
}

template <typename T>
__global__ void rbt_both(const double *__restrict__ M, long long ld, int npad, const double *__restrict__ ru,
                         const double *__restrict__ rv, T *__restrict__ G, long long ldg)
{
    const int h = npad / 4;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)h * h) return;
    int i = (int)(idx / h), j = (int)(idx % h);
    double m[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) m[a][b] = M[(long long)(i + a * h) * ld + j + b * h];
    for (int b = 0; b < 4; b++) {  // left: columns of the group
        double v[4] = {m[0][b], m[1][b], m[2][b], m[3][b]};
        bfly4(v, ru, i, h);
        for (int a = 0; a < 4; a++) m[a][b] = v[a];
    }
    for (int a = 0; a < 4; a++) bfly4(m[a], rv, j, h);  // right: rows of the group (M V acts like B^T on rows)
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) G[(long long)(i + a * h) * ldg + j + b * h] = (T)m[a][b];
}

// Left butterfly on a block of right hand sides (npad x nrhs, row-major): out = Uᵀ in
__global__ void rbt_rows(const double *__restrict__ in, int npad, int nrhs, const double *__restrict__ ru, double *out)
{
    const int h = npad / 4;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)h * nrhs) return;
    int i = (int)(idx / nrhs), c = (int)(idx % nrhs);
    double v[4];
    for (int a = 0; a < 4; a++) v[a] = in[(long long)(i + a * h) * nrhs + c];
    bfly4(v, ru, i, h);
    for (int a = 0; a < 4; a++) out[(long long)(i + a * h) * nrhs + c] = v[a];
}

// x = V y: V = diag(C2a, C2b) C1; (C1 y)_top = (R y_top + S y_bot)/sqrt2, (C1 y)_bot = (R y_top - S y_bot)/sqrt2
__global__ void rbt_apply_v(const double *__restrict__ y, int npad, int nrhs, const double *__restrict__ rv, double *x)
{
    const int h = npad / 4;
    const double s2 = 0.70710678118654752440;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)h * nrhs) return;
    int i = (int)(idx / nrhs), c = (int)(idx % nrhs);
    const double *R1 = rv, *S1 = rv + 2 * h, *Ra = rv + 4 * h, *Sa = rv + 5 * h, *Rb = rv + 6 * h, *Sb = rv + 7 * h;
    double v[4];
    for (int a = 0; a < 4; a++) v[a] = y[(long long)(i + a * h) * nrhs + c];
    // level 1 on index pairs (i, i+2h) and (i+h, i+3h)
    double w0 = (R1[i] * v[0] + S1[i] * v[2]) * s2, w2 = (R1[i] * v[0] - S1[i] * v[2]) * s2;
    double w1 = (R1[i + h] * v[1] + S1[i + h] * v[3]) * s2, w3 = (R1[i + h] * v[1] - S1[i + h] * v[3]) * s2;
    // level 2 within halves: pairs (i, i+h) and (i+2h, i+3h)
    x[(long long)i * nrhs + c] = (Ra[i] * w0 + Sa[i] * w1) * s2;
    x[(long long)(i + h) * nrhs + c] = (Ra[i] * w0 - Sa[i] * w1) * s2;
    x[(long long)(i + 2 * h) * nrhs + c] = (Rb[i] * w2 + Sb[i] * w3) * s2;
    x[(long long)(i + 3 * h) * nrhs + c] = (Rb[i] * w2 - Sb[i] * w3) * s2;
}

template <typename S, typename D> __global__ void convert(const S *in, D *out, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (D)in[i];
}

__global__ void axpy_inplace(double *x, const double *dx, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += dx[i];
}

// ---- GMRES-IR's vector kernels (one right hand side at a time, vectors of npad entries) ----

// column c of an n x nrhs row-major block into a vector, and back
__global__ void gm_gather(const double *__restrict__ src, int nrhs, int c, int n, double *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(long long)i * nrhs + c];
}
__global__ void gm_scatter(const double *__restrict__ src, int nrhs, int c, int n, double *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[(long long)i * nrhs + c] = src[i];
}

// out[j] = V[:, j] . w for the ncol basis columns (V column-major, leading dimension ldv): a block per column, summed
// in a fixed order
__global__ void gm_dots(const double *__restrict__ V, long long ldv, const double *__restrict__ w, int n,
                        double *__restrict__ out)
{
    __shared__ double sh[256];
    const double *v = V + (long long)blockIdx.x * ldv;
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a += v[i] * w[i];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// w = alpha w + beta V[:, 0..ncol) h (alpha 0 overwrites w without reading it)
__global__ void gm_update(const double *__restrict__ V, long long ldv, int ncol, const double *__restrict__ h,
                          double alpha, double beta, double *__restrict__ w, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < ncol; j++) acc += V[(long long)j * ldv + i] * h[j];
    w[i] = alpha == 0.0 ? beta * acc : alpha * w[i] + beta * acc;
}

// dst = s src
__global__ void gm_scale(const double *__restrict__ src, double s, double *__restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = s * src[i];
}

// max |r| and max |x| (a non-finite entry makes the first infinite), for the stopping test
__global__ void gm_maxabs(const double *__restrict__ r, const double *__restrict__ x, int n, double *__restrict__ out)
{
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        a = fmax(a, (isfinite(r[i]) && isfinite(x[i])) ? fabs(r[i]) : INFINITY);
        b = fmax(b, fabs(x[i]));
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + o]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}

// The refinement residual for a few right hand sides (nrhs <= GEMV_MAX), HBM-bound (M is read once, 8 bytes per
// entry): Y -= M X with one wavefront per row of M (row-major, ld), its lanes striding over the columns, the sums
// reduced over the wave; and Y -= Mᵀ X in two passes, a thread per column of M summing a slice of KSPLIT rows into
// part[slice][column] (coalesced along the rows), then the slices added in slice order (deterministic).
constexpr int GEMV_MAX = 4, KSPLIT = 32;

template <int NR>
__global__ __launch_bounds__(256) void gemv_sub(const double *__restrict__ M, long long ld, int nrow, int ncol,
                                                const double *__restrict__ X, double *__restrict__ Y)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= nrow) return;
    const double *m = M + (long long)row * ld;
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
#pragma unroll 4
    for (int k = lane; k < ncol; k += 64) {
        const double a = m[k];
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = fma(a, X[(long long)k * NR + c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NR; c++)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NR; c++) Y[(long long)row * NR + c] -= acc[c];
}

template <int NR>
__global__ __launch_bounds__(256) void gemv_t_part(const double *__restrict__ M, long long ld, int nrow, int ncol,
                                                   const double *__restrict__ X, int kchunk, double *__restrict__ part)
{
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= ncol) return;
    const int k0 = blockIdx.y * kchunk, k1 = min(nrow, k0 + kchunk);
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const double a = M[(long long)k * ld + col];
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = fma(a, X[(long long)k * NR + c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NR; c++) part[((long long)blockIdx.y * ncol + col) * NR + c] = acc[c];
}

__global__ void gemv_t_reduce(const double *__restrict__ part, int nsplit, long long len, double *__restrict__ Y)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; k++) s += part[(long long)k * len + i];
    Y[i] -= s;
}

// Products with the explicit inverse of the factors (TM = the factors' precision, FP64 sums): Y = G^-1 X with one
// wavefront per row, and Y = G^-T X in two passes (slices of KSPLIT rows, then the slices added in order)
template <typename TM, int NR>
__global__ __launch_bounds__(256) void gemv_set(const TM *__restrict__ M, int n, const double *__restrict__ X,
                                                double *__restrict__ Y)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const TM *m = M + (long long)row * n;
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
#pragma unroll 4
    for (int k = lane; k < n; k += 64) {
        const double a = (double)m[k];
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = fma(a, X[(long long)k * NR + c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NR; c++)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NR; c++) Y[(long long)row * NR + c] = acc[c];
}

template <typename TM, int NR>
__global__ __launch_bounds__(256) void gemv_t_part_inv(const TM *__restrict__ M, int n, const double *__restrict__ X,
                                                       int kchunk, double *__restrict__ part)
{
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= n) return;
    const int k0 = blockIdx.y * kchunk, k1 = min(n, k0 + kchunk);
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const double a = (double)M[(long long)k * n + col];
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = fma(a, X[(long long)k * NR + c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NR; c++) part[((long long)blockIdx.y * n + col) * NR + c] = acc[c];
}

__global__ void gemv_t_set(const double *__restrict__ part, int nsplit, long long len, double *__restrict__ Y)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; k++) s += part[(long long)k * len + i];
    Y[i] = s;
}

template <typename TM, int NR>
void launch_inverse_product(hipStream_t s, bool trans, const TM *Gi, int npad, const double *X, double *Y, double *part)
{
    if (!trans) {
        gemv_set<TM, NR><<<(npad + 3) / 4, 256, 0, s>>>(Gi, npad, X, Y);
    } else {
        const int kchunk = (npad + KSPLIT - 1) / KSPLIT;
        gemv_t_part_inv<TM, NR><<<dim3((npad + 255) / 256, KSPLIT), 256, 0, s>>>(Gi, npad, X, kchunk, part);
        HIPCHECK2(hipGetLastError());
        const long long len = (long long)npad * NR;
        gemv_t_set<<<(unsigned)((len + 255) / 256), 256, 0, s>>>(part, KSPLIT, len, Y);
    }
    HIPCHECK2(hipGetLastError());
}

template <typename TM>
void inverse_product(hipStream_t s, bool trans, const TM *Gi, int npad, int nrhs, const double *X, double *Y,
                     double *part)
{
    switch (nrhs) {
    case 1: launch_inverse_product<TM, 1>(s, trans, Gi, npad, X, Y, part); break;
    case 2: launch_inverse_product<TM, 2>(s, trans, Gi, npad, X, Y, part); break;
    case 3: launch_inverse_product<TM, 3>(s, trans, Gi, npad, X, Y, part); break;
    default: launch_inverse_product<TM, 4>(s, trans, Gi, npad, X, Y, part); break;
    }
}

template <typename T> __global__ void set_identity(T *__restrict__ X, int n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * n) return;
    X[i] = (i / n == i % n) ? T(1) : T(0);
}

template <int NR>
void launch_residual(hipStream_t s, bool trans, const double *M, long long ld, int npad, const double *X, double *Y,
                     double *part)
{
    if (!trans) {
        gemv_sub<NR><<<(npad + 3) / 4, 256, 0, s>>>(M, ld, npad, npad, X, Y);
    } else {
        const int kchunk = (npad + KSPLIT - 1) / KSPLIT;
        gemv_t_part<NR><<<dim3((npad + 255) / 256, KSPLIT), 256, 0, s>>>(M, ld, npad, npad, X, kchunk, part);
        HIPCHECK2(hipGetLastError());
        const long long len = (long long)npad * NR;
        gemv_t_reduce<<<(unsigned)((len + 255) / 256), 256, 0, s>>>(part, KSPLIT, len, Y);
    }
    HIPCHECK2(hipGetLastError());
}

// per right hand side (columns of n x nrhs row-major): max |r|, max |x|, sum r^2 and sum rhs^2 into out[4c..4c+3];
// a non-finite r or x makes max |r| infinite (fmax alone would drop a NaN)
__global__ void colmax2(const double *r, const double *x, const double *rhs, int n, int nrhs, double *out)
{
    __shared__ double s1[256], s2[256], s3[256], s4[256];
    int c = blockIdx.x;
    double a = 0.0, b = 0.0, r2 = 0.0, b2 = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double ri = r[(long long)i * nrhs + c], bi = rhs[(long long)i * nrhs + c];
        const double xi = x[(long long)i * nrhs + c];
        a = fmax(a, (isfinite(ri) && isfinite(xi)) ? fabs(ri) : INFINITY);
        b = fmax(b, fabs(xi));
        r2 += ri * ri;
        b2 += bi * bi;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    s3[threadIdx.x] = r2;
    s4[threadIdx.x] = b2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + o]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + o]);
            s3[threadIdx.x] += s3[threadIdx.x + o];
            s4[threadIdx.x] += s4[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[4 * c] = s1[0];
        out[4 * c + 1] = s2[0];
        out[4 * c + 2] = s3[0];
        out[4 * c + 3] = s4[0];
    }
}

// infinity norm of rows (max over rows of the sum of |M[i][:]|): partial per block
__global__ void row_abs_max(const double *M, long long ld, int n, double *out)
{
    __shared__ double s[256];
    double best = 0.0;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        double v = 0.0;
        for (int j = threadIdx.x; j < n; j += 256) v += fabs(M[(long long)i * ld + j]);
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) best = fmax(best, s[0]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = best;
}

// column sums of |M| (the 1-norm's columns), one thread per column
__global__ void col_abs_sum(const double *M, long long ld, int n, double *out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double v = 0.0;
    for (int i = 0; i < n; i++) v += fabs(M[(long long)i * ld + j]);
    out[j] = v;
}

// ------------------------------------------------------------------------------------------ host side --------------

template <typename T> T *dmalloc(size_t n)
{
    T *p = nullptr;
    HIPCHECK2(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return p;
}

Factor::~Factor()
{
    for (auto &g : inv_graph)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (void *p : owned) (void)hipFree(p);
}

static unsigned blocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Right-looking blocked LU without pivoting, two levels: panels of NB2 columns factored in steps of NB (updates kept
// inside the panel), then the panel's row block U12 = L11^-1 A12 by block forward substitution, then the trailing
// update A22 -= L21 U12 with K = NB2 (the autotuned tile shape), which keeps the update's HBM traffic per flop low.
template <typename T> void lu_factor(hipStream_t s, T *G, long long ld, int npad, T *linv, T *uinv, int *bad, int cfg)
{
    auto Li = [&](int k0) { return linv + (size_t)(k0 / NB) * NB * NB; };
    auto Ui = [&](int k0) { return uinv + (size_t)(k0 / NB) * NB * NB; };
    for (int K0 = 0; K0 < npad; K0 += NB2) {
        const int w = std::min(NB2, npad - K0);
        for (int k0 = K0; k0 < K0 + w; k0 += NB) {
            const int nb = std::min(NB, npad - k0);
            diag_lu<T><<<1, 256, 0, s>>>(G, ld, k0, nb, Li(k0), Ui(k0), bad);
            HIPCHECK2(hipGetLastError());
            const int m = npad - k0 - nb, wr = K0 + w - k0 - nb;
            if (m <= 0) break;
            T *D = G + (long long)k0 * ld + k0;
            // L21 = A21 U11^-1 (in place: one workgroup column covers the block's NB columns)
            launch_gemm<T, false>(s, m, nb, nb, D + (long long)nb * ld, ld, Ui(k0), NB, D + (long long)nb * ld, ld);
            if (wr > 0) {
                // U12 = L11^-1 A12 and A22 -= L21 U12, both within the panel's columns
                launch_gemm<T, false>(s, nb, wr, nb, Li(k0), NB, D + nb, ld, D + nb, ld);
                launch_gemm<T, true>(s, m, wr, nb, D + (long long)nb * ld, ld, D + nb, ld, D + (long long)nb * ld + nb, ld);
            }
        }
        const int nr = npad - K0 - w;
        if (nr <= 0) break;
        T *R = G + (long long)K0 * ld + K0 + w;  // the panel's row block right of it
        for (int k0 = K0; k0 < K0 + w; k0 += NB) {
            T *Xk = R + (long long)(k0 - K0) * ld;
            launch_gemm<T, false>(s, NB, nr, NB, Li(k0), NB, Xk, ld, Xk, ld);
            const int below = K0 + w - k0 - NB;
            if (below > 0)
                launch_gemm<T, true>(s, below, nr, NB, G + (long long)(k0 + NB) * ld + k0, ld, Xk, ld, Xk + (long long)NB * ld, ld);
        }
        launch_gemm_cfg<T, true, false>(cfg, s, nr, nr, w, G + (long long)(K0 + w) * ld + K0, ld, R, ld,
                                        G + (long long)(K0 + w) * ld + K0 + w, ld);
    }
}

// X (npad x nrhs, row-major, in place) <- U^-1 L^-1 X
template <typename T> void lu_solve(hipStream_t s, const T *G, long long ld, int npad, const T *linv, const T *uinv, T *X, int nrhs)
{
    for (int k0 = 0; k0 < npad; k0 += NB) {
        int nb = std::min(NB, npad - k0);
        T *xk = X + (long long)k0 * nrhs;
        launch_gemm<T, false>(s, nb, nrhs, nb, linv + (size_t)(k0 / NB) * NB * NB, NB, xk, nrhs, xk, nrhs);
        int m = npad - k0 - nb;
        if (m > 0) launch_gemm<T, true>(s, m, nrhs, nb, G + (long long)(k0 + nb) * ld + k0, ld, xk, nrhs, xk + (long long)nb * nrhs, nrhs);
    }
    int last = ((npad - 1) / NB) * NB;
    for (int k0 = last; k0 >= 0; k0 -= NB) {
        int nb = std::min(NB, npad - k0);
        T *xk = X + (long long)k0 * nrhs;
        launch_gemm<T, false>(s, nb, nrhs, nb, uinv + (size_t)(k0 / NB) * NB * NB, NB, xk, nrhs, xk, nrhs);
        if (k0 > 0) launch_gemm<T, true>(s, k0, nrhs, nb, G + k0, ld, xk, nrhs, X, nrhs);
    }
}

// ------------------------------------------------------------------------------------------ narrow solves ---------

// Triangular solves for a few right hand sides (nrhs <= GEMV_MAX) in panels of NS = 256 rows: per panel one workgroup
// solves the panel's diagonal block (four 64-row steps through the stored 64 x 64 inverses, with the updates inside
// the panel between them), then one kernel subtracts the panel's contribution from every row still to be solved.
// That is two launches per 256 rows instead of two per 64: a narrow solve is bound by the latency of each launch.
constexpr int NS = 256;

// T_kk x_k = x_k for the panel of nb rows (a multiple of NB) at k0.  LOWER: forward over the 64-row blocks, else
// backward.  TRANS: the triangle is Gᵀ's (its entry (i, j) is G[j][i]) and the block inverses are applied transposed.
template <typename T, int NR, bool LOWER, bool TRANS>
__global__ __launch_bounds__(256) void trsv_panel(const T *__restrict__ G, long long ld, const T *__restrict__ dinv,
                                                  int k0, int nb, T *__restrict__ X)
{
    __shared__ T xs[NS][NR];
    __shared__ T ys[NB][NR];
    const int tid = threadIdx.x;
    if (tid < nb)
        for (int c = 0; c < NR; c++) xs[tid][c] = X[(long long)(k0 + tid) * NR + c];
    __syncthreads();
    const int nsub = nb / NB;
    for (int t = 0; t < nsub; t++) {
        const int sb = LOWER ? t : nsub - 1 - t;
        const T *D = dinv + (size_t)(k0 / NB + sb) * NB * NB;
        if (tid < NB) {
            T acc[NR];
            for (int c = 0; c < NR; c++) acc[c] = T(0);
            for (int j = 0; j < NB; j++) {
                const T d = TRANS ? D[j * NB + tid] : D[tid * NB + j];
                for (int c = 0; c < NR; c++) acc[c] += d * xs[sb * NB + j][c];
            }
            for (int c = 0; c < NR; c++) ys[tid][c] = acc[c];
        }
        __syncthreads();
        if (tid < NB)
            for (int c = 0; c < NR; c++) xs[sb * NB + tid][c] = ys[tid][c];
        __syncthreads();
        // the rows of the panel still to be solved take this block's contribution
        const int r = (LOWER ? (sb + 1) * NB : 0) + tid, r1 = LOWER ? nb : sb * NB;
        if (r < r1) {
            T acc[NR];
            for (int c = 0; c < NR; c++) acc[c] = T(0);
            const long long gi = k0 + r;
            for (int j = 0; j < NB; j++) {
                const long long gj = k0 + sb * NB + j;
                const T g = TRANS ? G[gj * ld + gi] : G[gi * ld + gj];
                for (int c = 0; c < NR; c++) acc[c] += g * xs[sb * NB + j][c];
            }
            for (int c = 0; c < NR; c++) xs[r][c] -= acc[c];
        }
        __syncthreads();
    }
    if (tid < nb)
        for (int c = 0; c < NR; c++) X[(long long)(k0 + tid) * NR + c] = xs[tid][c];
}

// x_i -= sum over the panel's columns j in [j0, j0 + nj) of G[i][j] x_j for the rows i in [i0, i1): a wavefront per
// row, coalesced along it.
template <typename T, int NR>
__global__ __launch_bounds__(256) void trsv_update(const T *__restrict__ G, long long ld, int j0, int nj, int i0, int i1,
                                                   T *__restrict__ X)
{
    __shared__ T xs[NS][NR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < nj)
        for (int c = 0; c < NR; c++) xs[tid][c] = X[(long long)(j0 + tid) * NR + c];
    __syncthreads();
    const long long i = i0 + (long long)blockIdx.x * 4 + wave;
    if (i >= i1) return;
    const T *g = G + i * ld + j0;
    T acc[NR];
    for (int c = 0; c < NR; c++) acc[c] = T(0);
    for (int j = lane; j < nj; j += 64) {
        const T a = g[j];
        for (int c = 0; c < NR; c++) acc[c] += a * xs[j][c];
    }
    for (int c = 0; c < NR; c++)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o);
    if (lane == 0)
        for (int c = 0; c < NR; c++) X[i * NR + c] -= acc[c];
}

// The same with G[j][i] for G[i][j]: a thread per row i (coalesced along G's rows), the panel's columns split over
// the four wavefronts and their sums added in wavefront order.
template <typename T, int NR>
__global__ __launch_bounds__(256) void trsv_update_t(const T *__restrict__ G, long long ld, int j0, int nj, int i0,
                                                     int i1, T *__restrict__ X)
{
    __shared__ T xs[NS][NR];
    __shared__ T part[4][64][NR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < nj)
        for (int c = 0; c < NR; c++) xs[tid][c] = X[(long long)(j0 + tid) * NR + c];
    __syncthreads();
    const long long i = i0 + (long long)blockIdx.x * 64 + lane;
    T acc[NR];
    for (int c = 0; c < NR; c++) acc[c] = T(0);
    if (i < i1) {
        const int jb = wave * (NS / 4), je = min(nj, jb + NS / 4);
        for (int j = jb; j < je; j++) {
            const T a = G[(long long)(j0 + j) * ld + i];
            for (int c = 0; c < NR; c++) acc[c] += a * xs[j][c];
        }
    }
    for (int c = 0; c < NR; c++) part[wave][lane][c] = acc[c];
    __syncthreads();
    if (wave == 0 && i < i1)
        for (int c = 0; c < NR; c++)
            X[i * NR + c] -= ((part[0][lane][c] + part[1][lane][c]) + part[2][lane][c]) + part[3][lane][c];
}

// G^-1 X (trans: G^-T X) for G = L U in place, npad a multiple of NS: forward over L (trans: Uᵀ), backward over U
// (trans: Lᵀ), as lu_solve and lu_solve_t do with the GEMM.
template <typename T, int NR>
void lu_solve_narrow(hipStream_t s, const T *G, long long ld, int npad, const T *linv, const T *uinv, T *X, bool trans)
{
    const int np = npad / NS;
    for (int p = 0; p < np; p++) {
        const int k0 = p * NS, i0 = k0 + NS;
        if (trans) {
            trsv_panel<T, NR, true, true><<<1, 256, 0, s>>>(G, ld, uinv, k0, NS, X);
            if (i0 < npad) trsv_update_t<T, NR><<<(npad - i0 + 63) / 64, 256, 0, s>>>(G, ld, k0, NS, i0, npad, X);
        } else {
            trsv_panel<T, NR, true, false><<<1, 256, 0, s>>>(G, ld, linv, k0, NS, X);
            if (i0 < npad) trsv_update<T, NR><<<(npad - i0 + 3) / 4, 256, 0, s>>>(G, ld, k0, NS, i0, npad, X);
        }
        HIPCHECK2(hipGetLastError());
    }
    for (int p = np - 1; p >= 0; p--) {
        const int k0 = p * NS;
        if (trans) {
            trsv_panel<T, NR, false, true><<<1, 256, 0, s>>>(G, ld, linv, k0, NS, X);
            if (k0 > 0) trsv_update_t<T, NR><<<(k0 + 63) / 64, 256, 0, s>>>(G, ld, k0, NS, 0, k0, X);
        } else {
            trsv_panel<T, NR, false, false><<<1, 256, 0, s>>>(G, ld, uinv, k0, NS, X);
            if (k0 > 0) trsv_update<T, NR><<<(k0 + 3) / 4, 256, 0, s>>>(G, ld, k0, NS, 0, k0, X);
        }
        HIPCHECK2(hipGetLastError());
    }
}

template <typename T>
void lu_solve_any(hipStream_t s, const T *G, long long ld, int npad, const T *linv, const T *uinv, T *X, int nrhs,
                  bool trans)
{
    if (npad % NS != 0) throw std::logic_error("the padded size must be a multiple of 256");
    switch (nrhs <= GEMV_MAX ? nrhs : 0) {
    case 1: return lu_solve_narrow<T, 1>(s, G, ld, npad, linv, uinv, X, trans);
    case 2: return lu_solve_narrow<T, 2>(s, G, ld, npad, linv, uinv, X, trans);
    case 3: return lu_solve_narrow<T, 3>(s, G, ld, npad, linv, uinv, X, trans);
    case 4: return lu_solve_narrow<T, 4>(s, G, ld, npad, linv, uinv, X, trans);
    default:
        if (trans)
            lu_solve_t<T>(s, G, ld, npad, linv, uinv, X, nrhs);
        else
            lu_solve<T>(s, G, ld, npad, linv, uinv, X, nrhs);
    }
}

void Factor::build(Ctx *c, const double *A, long long lda, const vp_constraints &k, int mode_)
{
    ctx = c;
    mode = mode_;
    hipStream_t s = c->stream;
    if (mode != 0 && mode != 1) throw std::invalid_argument("mode must be 0 (fp64) or 1 (mixed)");
    nslot = k.nslot;
    nq = k.nq;
    nw = k.nw;
    nb = k.nb;
    if (nq <= 0 || nslot <= 0 || nb < 0) throw std::invalid_argument("empty system");
    n = nq + nb;
    npad = ((n + 255) / 256) * 256;
    ld = npad;
    // CSC of C (nslot x nq) and of Cf = [C; W] ((nslot + nw) x nq) from the CSR inputs
    std::vector<std::vector<std::pair<int, double>>> ccol(nq), fcol(nq);
    for (int i = 0; i < nslot; i++)
        for (int a = k.C_rowptr[i]; a < k.C_rowptr[i + 1]; a++) {
            int q = k.C_col[a];
            if (q < 0 || q >= nq) throw std::out_of_range("C column index out of range");
            ccol[q].push_back({i, k.C_val[a]});
            fcol[q].push_back({i, k.C_val[a]});
        }
    for (int i = 0; i < nw; i++)
        for (int a = k.W_rowptr[i]; a < k.W_rowptr[i + 1]; a++) {
            int q = k.W_col[a];
            if (q < 0 || q >= nq) throw std::out_of_range("W column index out of range");
            fcol[q].push_back({nslot + i, k.W_val[a]});
        }
    auto csc = [](const std::vector<std::vector<std::pair<int, double>>> &cols, std::vector<int> &p, std::vector<int> &ix,
                  std::vector<double> &v) {
        p.assign(1, 0);
        for (auto &col : cols) {
            for (auto &e : col) {
                ix.push_back(e.first);
                v.push_back(e.second);
            }
            p.push_back((int)ix.size());
        }
    };
    std::vector<int> cp, ci, fp, fi;
    std::vector<double> cv, fv;
    csc(ccol, cp, ci, cv);
    csc(fcol, fp, fi, fv);
    auto up = [&](auto &vec) {
        using E = typename std::decay<decltype(vec)>::type::value_type;
        E *d = dmalloc<E>(vec.size());
        if (!vec.empty()) HIPCHECK2(hipMemcpyAsync(d, vec.data(), vec.size() * sizeof(E), hipMemcpyHostToDevice, s));
        owned.push_back(d);
        return d;
    };
    int *dcp = up(cp), *dci = up(ci), *dfp = up(fp), *dfi = up(fi);
    double *dcv = up(cv), *dfv = up(fv);
    M = dmalloc<double>((size_t)ld * npad);
    owned.push_back(M);
    HIPCHECK2(hipMemsetAsync(M, 0, sizeof(double) * (size_t)ld * npad, s));
    dim3 g((nq + 63) / 64, (nq + 63) / 64);
    reduce_tile<<<g, 256, 0, s>>>(A, lda, nq, dcp, dci, dcv, dfp, dfi, dfv, M, ld);
    HIPCHECK2(hipGetLastError());
    // border scaling: scale = sum |diag(Ar)| / nq; e and c scaled to 2-norm = scale (solve.jl:30-35)
    double *dsum = dmalloc<double>(1);
    owned.push_back(dsum);
    diag_abs_sum<<<1, 256, 0, s>>>(M, ld, nq, dsum);
    double hsum = 0.0;
    HIPCHECK2(hipMemcpyAsync(&hsum, dsum, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    double scale = hsum / nq;
    se.assign(nb, 0.0);
    std::vector<double> sc(nb);
    for (int b = 0; b < nb; b++) {
        double ne = 0.0, nc = 0.0;
        for (int p = 0; p < nq; p++) {
            ne += k.e[(size_t)b * nq + p] * k.e[(size_t)b * nq + p];
            nc += k.c[(size_t)b * nq + p] * k.c[(size_t)b * nq + p];
        }
        se[b] = scale / std::sqrt(ne);
        sc[b] = scale / std::sqrt(nc);
    }
    std::vector<double> ev(k.e, k.e + (size_t)nb * nq), cvv(k.c, k.c + (size_t)nb * nq);
    double *de = up(ev), *dc = up(cvv), *dse = up(se), *dsc = up(sc);
    long long work = std::max<long long>((long long)nq * nb, (long long)(npad - n) * npad);
    work = std::max<long long>(work, (long long)nb * nb);
    if (work > 0) border_fill<<<blocks(work), 256, 0, s>>>(M, ld, nq, nb, npad, de, dc, dse, dsc);
    HIPCHECK2(hipGetLastError());
    // the random butterflies (fixed seed: the factorization is deterministic)
    std::mt19937_64 rng(20261203ull);
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    std::vector<double> ru(2 * npad), rv(2 * npad);
    for (auto &x : ru) x = std::exp(U(rng) / 10);
    for (auto &x : rv) x = std::exp(U(rng) / 10);
    d_ru = up(ru);
    d_rv = up(rv);
    // 1-norm of M (for the condition estimate): column sums
    double *dcol = dmalloc<double>(npad);
    owned.push_back(dcol);
    col_abs_sum<<<blocks(n), 256, 0, s>>>(M, ld, n, dcol);
    std::vector<double> hcol(n);
    HIPCHECK2(hipMemcpyAsync(hcol.data(), dcol, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    anorm1 = *std::max_element(hcol.begin(), hcol.end());
    // randomized copy in the factor type, then LU
    int *dbad = dmalloc<int>(1);
    owned.push_back(dbad);
    HIPCHECK2(hipMemsetAsync(dbad, 0, sizeof(int), s));
    size_t nblk = (size_t)(npad / NB) * NB * NB;
    long long h = npad / 4;
    if (mode == 0) {
        G64 = dmalloc<double>((size_t)ld * npad);
        linv64 = dmalloc<double>(nblk);
        uinv64 = dmalloc<double>(nblk);
        owned.insert(owned.end(), {G64, linv64, uinv64});
        const int cfg = tuned_cfg<double>(s, G64, ld, npad);
        rbt_both<double><<<blocks(h * h), 256, 0, s>>>(M, ld, npad, d_ru, d_rv, G64, ld);
        HIPCHECK2(hipGetLastError());
        lu_factor<double>(s, G64, ld, npad, linv64, uinv64, dbad, cfg);
    } else {
        G32 = dmalloc<float>((size_t)ld * npad);
        linv32 = dmalloc<float>(nblk);
        uinv32 = dmalloc<float>(nblk);
        owned.insert(owned.end(), {G32, linv32, uinv32});
        const int cfg = tuned_cfg<float>(s, G32, ld, npad);
        rbt_both<float><<<blocks(h * h), 256, 0, s>>>(M, ld, npad, d_ru, d_rv, G32, ld);
        HIPCHECK2(hipGetLastError());
        lu_factor<float>(s, G32, ld, npad, linv32, uinv32, dbad, cfg);
    }
    int hbad = 0;
    HIPCHECK2(hipMemcpyAsync(&hbad, dbad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    if (hbad) throw std::domain_error("zero or non-finite pivot in the randomized LU (singular bordered matrix)");
    ws_nrhs = 0;
}

void Factor::ensure_ws(int nrhs)
{
    if (nrhs <= ws_nrhs) return;
    for (void *p : {(void *)X, (void *)R, (void *)Y, (void *)B, (void *)Xf, (void *)nrm, (void *)part}) {
        if (!p) continue;
        owned.erase(std::find(owned.begin(), owned.end(), p));
        (void)hipFree(p);
    }
    size_t sz = (size_t)npad * nrhs;
    X = dmalloc<double>(sz);
    R = dmalloc<double>(sz);
    Y = dmalloc<double>(sz);
    B = dmalloc<double>(sz);
    Xf = dmalloc<float>(sz);
    nrm = dmalloc<double>(4 * (size_t)nrhs);
    part = dmalloc<double>((size_t)KSPLIT * npad * std::min(nrhs, GEMV_MAX));
    owned.insert(owned.end(), {X, R, Y, B, Xf, nrm, part});
    ws_nrhs = nrhs;
    ws_epoch++;
}

// y = M^-1 r, or M^-T r with trans, for the current workspace block (in: Y holds r, out: Y holds the product),
// npad x nrhs.  M^-1 = B_v G^-1 B_uᵀ for G = B_uᵀ M B_v = L U, so M^-T = B_u G^-T B_vᵀ: the same kernels with the
// two butterflies' roles swapped and the transposed triangular solves.
void Factor::apply_inverse(int nrhs, bool trans)
{
    hipStream_t s = ctx->stream;
    long long h = npad / 4;
    rbt_rows<<<blocks(h * nrhs), 256, 0, s>>>(Y, npad, nrhs, trans ? d_rv : d_ru, R);
    HIPCHECK2(hipGetLastError());
    if (inv_ready && nrhs <= GEMV_MAX) {  // G^-1 R (G^-T R) by the explicit inverse
        if (mode == 0)
            inverse_product<double>(s, trans, (const double *)ginv, npad, nrhs, R, Zbuf, part);
        else
            inverse_product<float>(s, trans, (const float *)ginv, npad, nrhs, R, Zbuf, part);
        rbt_apply_v<<<blocks(h * nrhs), 256, 0, s>>>(Zbuf, npad, nrhs, trans ? d_ru : d_rv, Y);
        HIPCHECK2(hipGetLastError());
        return;
    }
    if (mode == 0) {
        lu_solve_any<double>(s, G64, ld, npad, linv64, uinv64, R, nrhs, trans);
    } else {
        convert<double, float><<<blocks((long long)npad * nrhs), 256, 0, s>>>(R, Xf, (long long)npad * nrhs);
        lu_solve_any<float>(s, G32, ld, npad, linv32, uinv32, Xf, nrhs, trans);
        convert<float, double><<<blocks((long long)npad * nrhs), 256, 0, s>>>(Xf, R, (long long)npad * nrhs);
    }
    rbt_apply_v<<<blocks(h * nrhs), 256, 0, s>>>(R, npad, nrhs, trans ? d_ru : d_rv, Y);
    HIPCHECK2(hipGetLastError());
}

// apply_inverse through a HIP graph: the first call for a direction, nrhs and workspace runs eagerly, the second
// captures the same launches from the stream and instantiates them, later ones replay the graph.  The captured work is
// kernels only (no memset or copy nodes); it reads the factors and reads and writes only the workspace.
void Factor::apply_inverse_replay(int nrhs, bool trans)
{
    hipStream_t s = ctx->stream;
    if (nrhs <= GEMV_MAX && !inv_ready) narrow_applies++;
    if (!graphs_enabled()) {
        apply_inverse(nrhs, trans);
        return;
    }
    InvGraph &g = inv_graph[trans ? 1 : 0];
    if (g.nrhs != nrhs || g.epoch != ws_epoch) {
        if (g.exec) HIPCHECK2(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
        g.nrhs = nrhs;
        g.epoch = ws_epoch;
        g.calls = 0;
    }
    if (!g.exec && ++g.calls < 2) {
        apply_inverse(nrhs, trans);
        return;
    }
    if (!g.exec) {
        hipGraph_t graph = nullptr;
        HIPCHECK2(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            apply_inverse(nrhs, trans);
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIPCHECK2(hipStreamEndCapture(s, &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHECK2(e);
    }
    HIPCHECK2(hipGraphLaunch(g.exec, s));
    graph_launches++;
}

// Solve M X = Bh, or Mᵀ X = Bh with trans (n x nrhs row-major host or device), with FP64 iterative refinement (LAPACK
// dsgesv's stopping test: |r|_inf <= |x|_inf |op(M)|_inf eps sqrt(n), at most max_iter refinements).  resid_out: the
// largest relative residual |R - op(M) X|_2 / |R|_2 over the right hand sides (the reference's residual, solve.jl:58).
// When the test is not met within max_iter, the FP32 factors are replaced by an FP64 factorization of the same M and
// the solve is repeated (dsgesv's fallback); FP64 factors that still do not meet it raise ConvergenceError.
// max_iter 0 applies the factors once, with no refinement and no test; strict false refines up to max_iter and keeps
// what it has (the condition estimate's products).
int Factor::solve(const double *Bin, int nrhs, double *Xout, bool device_io, int max_iter, double *resid_out, bool trans,
                  bool strict)
{
    hipStream_t s = ctx->stream;
    ensure_ws(nrhs);
    if (nrhs <= GEMV_MAX && !inv_ready) {  // at a solve's start only: a GMRES cycle keeps one preconditioner
        const int pol = lu_inverse_policy();
        if (pol != 0 && npad <= LU_INVERSE_MAX && (pol == 1 || narrow_applies >= LU_INVERSE_AFTER)) build_inverse();
    }
    size_t nn = (size_t)n * nrhs, sz = (size_t)npad * nrhs;
    HIPCHECK2(hipMemsetAsync(B, 0, sizeof(double) * sz, s));
    HIPCHECK2(hipMemcpyAsync(B, Bin, sizeof(double) * nn, device_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    HIPCHECK2(hipMemcpyAsync(Y, B, sizeof(double) * sz, hipMemcpyDeviceToDevice, s));
    apply_inverse_replay(nrhs, trans);
    HIPCHECK2(hipMemcpyAsync(X, Y, sizeof(double) * sz, hipMemcpyDeviceToDevice, s));
    if (max_iter <= 0) {
        if (resid_out) *resid_out = -1.0;
        HIPCHECK2(hipMemcpyAsync(Xout, X, sizeof(double) * nn, device_io ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
        last_iters = 0;
        return 0;
    }
    if (trans && anorm1 <= 0) {
        double *dcol = dmalloc<double>(npad);
        col_abs_sum<<<blocks(n), 256, 0, s>>>(M, ld, n, dcol);
        std::vector<double> hcol(n);
        HIPCHECK2(hipMemcpyAsync(hcol.data(), dcol, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
        (void)hipFree(dcol);
        anorm1 = *std::max_element(hcol.begin(), hcol.end());
    }
    if (!trans && anorm_inf < 0) {
        double *dpart = dmalloc<double>(1024);
        row_abs_max<<<1024, 256, 0, s>>>(M, ld, n, dpart);
        std::vector<double> part(1024);
        HIPCHECK2(hipMemcpyAsync(part.data(), dpart, sizeof(double) * 1024, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
        (void)hipFree(dpart);
        anorm_inf = *std::max_element(part.begin(), part.end());
    }
    const double opnorm = trans ? anorm1 : anorm_inf;  // |Mᵀ|_inf = |M|_1
    const double eps = 1.1102230246251565e-16, tolf = opnorm * eps * std::sqrt((double)n);
    std::vector<double> hn(4 * (size_t)nrhs);
    int it = 0;
    bool done = false, diverged = false;
    double worst = 0.0;
    std::vector<double> prev(nrhs, INFINITY);
    for (;; it++) {
        // R = B - M X in FP64: a GEMV pass over M for a few right hand sides, else the MFMA GEMM
        HIPCHECK2(hipMemcpyAsync(Y, B, sizeof(double) * sz, hipMemcpyDeviceToDevice, s));
        switch (nrhs <= GEMV_MAX ? nrhs : 0) {
        case 1: launch_residual<1>(s, trans, M, ld, npad, X, Y, part); break;
        case 2: launch_residual<2>(s, trans, M, ld, npad, X, Y, part); break;
        case 3: launch_residual<3>(s, trans, M, ld, npad, X, Y, part); break;
        case 4: launch_residual<4>(s, trans, M, ld, npad, X, Y, part); break;
        default:
            if (trans)
                launch_gemm_cfg<double, true, true>(0, s, npad, nrhs, npad, M, ld, X, nrhs, Y, nrhs);
            else
                launch_gemm<double, true>(s, npad, nrhs, npad, M, ld, X, nrhs, Y, nrhs);
        }
        colmax2<<<nrhs, 256, 0, s>>>(Y, X, B, n, nrhs, nrm);
        HIPCHECK2(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * 4 * nrhs, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
        done = true;
        worst = 0.0;
        for (int c = 0; c < nrhs; c++) {
            if (!(hn[4 * c] <= hn[4 * c + 1] * tolf)) done = false;  // a non-finite residual fails the test too
            const double rb = std::sqrt(hn[4 * c + 3]), rr = std::sqrt(hn[4 * c + 2]);
            const double rel = rb > 0 ? rr / rb : rr;
            worst = std::isnan(rel) || std::isnan(worst) ? NAN : std::max(worst, rel);
            // dsgesv stops the FP32 factors early when a refinement step does not reduce the residual
            if (it > 0 && strict && mode == 1 && !(hn[4 * c] < prev[c])) diverged = true;
            prev[c] = hn[4 * c];
        }
        if (done || it >= max_iter || (diverged && !done) || (strict && mode == 1 && gmres_only())) break;
        apply_inverse_replay(nrhs, trans);
        axpy_inplace<<<blocks((long long)sz), 256, 0, s>>>(X, Y, (long long)sz);
        HIPCHECK2(hipGetLastError());
    }
    if (!done && strict && mode == 1 && gmres_enabled()) {
        // GMRES-IR on every right hand side from the refinement's iterate, before a second factorization
        bool all = true;
        for (int c = 0; c < nrhs && all; c++) all = gmres_column(c, nrhs, trans, tolf, GMRES_RESTARTS);
        if (all) {
            gmres_solves++;
            HIPCHECK2(hipMemcpyAsync(Y, B, sizeof(double) * sz, hipMemcpyDeviceToDevice, s));
            switch (nrhs <= GEMV_MAX ? nrhs : 0) {
            case 1: launch_residual<1>(s, trans, M, ld, npad, X, Y, part); break;
            case 2: launch_residual<2>(s, trans, M, ld, npad, X, Y, part); break;
            case 3: launch_residual<3>(s, trans, M, ld, npad, X, Y, part); break;
            case 4: launch_residual<4>(s, trans, M, ld, npad, X, Y, part); break;
            default:
                if (trans)
                    launch_gemm_cfg<double, true, true>(0, s, npad, nrhs, npad, M, ld, X, nrhs, Y, nrhs);
                else
                    launch_gemm<double, true>(s, npad, nrhs, npad, M, ld, X, nrhs, Y, nrhs);
            }
            colmax2<<<nrhs, 256, 0, s>>>(Y, X, B, n, nrhs, nrm);
            HIPCHECK2(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * 4 * nrhs, hipMemcpyDeviceToHost, s));
            HIPCHECK2(hipStreamSynchronize(s));
            worst = 0.0;
            for (int c = 0; c < nrhs; c++) {
                const double rb = std::sqrt(hn[4 * c + 3]), rr = std::sqrt(hn[4 * c + 2]);
                worst = std::max(worst, rb > 0 ? rr / rb : rr);
            }
            done = true;
        }
    }
    if (!done && strict) {
        if (mode == 1) {
            refactor_fp64();
            const int it2 = solve(Bin, nrhs, Xout, device_io, max_iter, resid_out, trans);
            return it + 1 + it2;
        }
        char msg[200];
        snprintf(msg, sizeof msg,
                 "iterative refinement did not converge: relative residual %.3e after %d refinements of the FP64 factors%s",
                 worst, it, diverged ? " (the residual stopped decreasing)" : "");
        throw ConvergenceError(msg);
    }
    if (resid_out) *resid_out = worst;
    HIPCHECK2(hipMemcpyAsync(Xout, X, sizeof(double) * nn, device_io ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    last_iters = it;
    return it;
}

// GMRES-IR (Carson & Higham, SIAM J. Sci. Comput. 40, 2018): when refinement with the FP32 factors stops short, the
// correction equation is solved by restarted GMRES(m) on op(M) in FP64, right-preconditioned by the factors' inverse
// (the butterflies and the FP32 LU), from the refinement's iterate: x_{k+1} = x_k + M~^-1 V y, with V the Arnoldi basis
// of op(M) M~^-1 (classical Gram-Schmidt twice) and y the Hessenberg least-squares solution (Givens rotations on the
// host).  Column c of the workspace X (row-major npad x nrhs) is refined against column c of B; the stopping test is
// the refinement's (|r|_inf <= |x|_inf |op(M)|_inf eps sqrt(n) on the true residual).  Returns true when it is met.
bool Factor::gmres_column(int c, int nrhs, bool trans, double tolf, int max_restarts)
{
    hipStream_t s = ctx->stream;
    const int m = GMRES_M;
    if (!gV) {
        gV = dmalloc<double>((size_t)npad * (m + 1));
        gx = dmalloc<double>(npad);
        gw = dmalloc<double>(npad);
        gb = dmalloc<double>(npad);
        gh = dmalloc<double>(m + 2);
        owned.insert(owned.end(), {gV, gx, gw, gb, gh});
    }
    const unsigned nb = (unsigned)blocks(npad);
    gm_gather<<<nb, 256, 0, s>>>(X, nrhs, c, npad, gx);
    gm_gather<<<nb, 256, 0, s>>>(B, nrhs, c, npad, gb);
    HIPCHECK2(hipGetLastError());
    auto residual = [&]() {  // gw = b - op(M) x
        HIPCHECK2(hipMemcpyAsync(gw, gb, sizeof(double) * npad, hipMemcpyDeviceToDevice, s));
        launch_residual<1>(s, trans, M, ld, npad, gx, gw, part);
    };
    auto dots = [&](const double *V, int ncol, const double *w, double *h) {
        gm_dots<<<ncol, 256, 0, s>>>(V, npad, w, npad, gh);
        HIPCHECK2(hipGetLastError());
        HIPCHECK2(hipMemcpyAsync(h, gh, sizeof(double) * ncol, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
    };
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 1), h2(m + 1), y(m);
    double mx[2];
    bool converged = false;
    for (int restart = 0; restart <= max_restarts; restart++) {
        residual();
        gm_maxabs<<<1, 256, 0, s>>>(gw, gx, npad, gh);
        HIPCHECK2(hipMemcpyAsync(mx, gh, sizeof(mx), hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
        if (mx[0] <= mx[1] * tolf) {
            converged = true;
            break;
        }
        if (!std::isfinite(mx[0]) || restart == max_restarts) break;
        double beta;
        dots(gw, 1, gw, &beta);
        beta = std::sqrt(beta);
        if (!(beta > 0)) break;
        gm_scale<<<nb, 256, 0, s>>>(gw, 1.0 / beta, gV, npad);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        for (int j = 0; j < m; j++) {
            // z = M~^-1 v_j (in Y), w = op(M) z
            HIPCHECK2(hipMemcpyAsync(Y, gV + (size_t)j * npad, sizeof(double) * npad, hipMemcpyDeviceToDevice, s));
            apply_inverse_replay(1, trans);
            HIPCHECK2(hipMemsetAsync(gw, 0, sizeof(double) * npad, s));
            launch_residual<1>(s, trans, M, ld, npad, Y, gw, part);  // gw = -op(M) z
            gm_scale<<<nb, 256, 0, s>>>(gw, -1.0, gw, npad);
            // classical Gram-Schmidt, twice
            dots(gV, j + 1, gw, h.data());
            HIPCHECK2(hipMemcpyAsync(gh, h.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));
            gm_update<<<nb, 256, 0, s>>>(gV, npad, j + 1, gh, 1.0, -1.0, gw, npad);
            dots(gV, j + 1, gw, h2.data());
            HIPCHECK2(hipMemcpyAsync(gh, h2.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));
            gm_update<<<nb, 256, 0, s>>>(gV, npad, j + 1, gh, 1.0, -1.0, gw, npad);
            for (int i = 0; i <= j; i++) h[i] += h2[i];
            double hn;
            dots(gw, 1, gw, &hn);
            hn = std::sqrt(hn);
            h[j + 1] = hn;
            if (hn > 0) gm_scale<<<nb, 256, 0, s>>>(gw, 1.0 / hn, gV + (size_t)(j + 1) * npad, npad);
            HIPCHECK2(hipGetLastError());
            gmres_iters++;
            // the previous rotations on the new column, then its own
            for (int i = 0; i < j; i++) {
                const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t;
            }
            const double r = std::hypot(h[j], h[j + 1]);
            cs[j] = r > 0 ? h[j] / r : 1.0;
            sn[j] = r > 0 ? h[j + 1] / r : 0.0;
            h[j] = r;
            h[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = h[i];
            k = j + 1;
            if (!(hn > 0) || std::fabs(g[j + 1]) <= 1e-14 * beta) break;  // the Krylov space holds the correction
        }
        // y = R^-1 g, u = V y, x += M~^-1 u
        for (int i = k - 1; i >= 0; i--) {
            double t = g[i];
            for (int l = i + 1; l < k; l++) t -= H[(size_t)i * m + l] * y[l];
            y[i] = H[(size_t)i * m + i] != 0.0 ? t / H[(size_t)i * m + i] : 0.0;
        }
        HIPCHECK2(hipMemcpyAsync(gh, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        gm_update<<<nb, 256, 0, s>>>(gV, npad, k, gh, 0.0, 1.0, Y, npad);
        apply_inverse_replay(1, trans);
        axpy_inplace<<<nb, 256, 0, s>>>(gx, Y, npad);
        HIPCHECK2(hipGetLastError());
    }
    gm_scatter<<<nb, 256, 0, s>>>(gx, nrhs, c, npad, X);
    HIPCHECK2(hipGetLastError());
    return converged;
}

// G^-1 from the factors: the identity solved as npad right hand sides (the GEMM-based sweeps, 2 npad^3 flops), kept in
// the factors' precision.  The graphs captured the panel sweeps, so they are captured again.
void Factor::build_inverse()
{
    hipStream_t s = ctx->stream;
    const long long nn = (long long)npad * npad;
    if (mode == 0) {
        double *g = dmalloc<double>((size_t)nn);
        owned.push_back(g);
        set_identity<double><<<blocks(nn), 256, 0, s>>>(g, npad);
        HIPCHECK2(hipGetLastError());
        lu_solve_any<double>(s, G64, ld, npad, linv64, uinv64, g, npad, false);
        ginv = g;
    } else {
        float *g = dmalloc<float>((size_t)nn);
        owned.push_back(g);
        set_identity<float><<<blocks(nn), 256, 0, s>>>(g, npad);
        HIPCHECK2(hipGetLastError());
        lu_solve_any<float>(s, G32, ld, npad, linv32, uinv32, g, npad, false);
        ginv = g;
    }
    if (!Zbuf) {
        Zbuf = dmalloc<double>((size_t)npad * GEMV_MAX);
        owned.push_back(Zbuf);
    }
    HIPCHECK2(hipStreamSynchronize(s));
    for (InvGraph &g : inv_graph) {
        if (g.exec) HIPCHECK2(hipGraphExecDestroy(g.exec));
        g = InvGraph();
    }
    inv_ready = true;
    inverse_builds++;
}

// dsgesv's fallback: the FP32 factors did not give a converged solve, so M (kept in FP64) is factored again in FP64
// through the same butterflies, and the factorization stays FP64 from then on.
void Factor::refactor_fp64()
{
    hipStream_t s = ctx->stream;
    HIPCHECK2(hipStreamSynchronize(s));
    for (InvGraph &g : inv_graph) {
        if (g.exec) HIPCHECK2(hipGraphExecDestroy(g.exec));
        g = InvGraph();
    }
    int *dbad = dmalloc<int>(1);
    owned.push_back(dbad);
    HIPCHECK2(hipMemsetAsync(dbad, 0, sizeof(int), s));
    const size_t nblk = (size_t)(npad / NB) * NB * NB;
    const long long h = npad / 4;
    G64 = dmalloc<double>((size_t)ld * npad);
    linv64 = dmalloc<double>(nblk);
    uinv64 = dmalloc<double>(nblk);
    owned.insert(owned.end(), {G64, linv64, uinv64});
    const int cfg = tuned_cfg<double>(s, G64, ld, npad);
    rbt_both<double><<<blocks(h * h), 256, 0, s>>>(M, ld, npad, d_ru, d_rv, G64, ld);
    HIPCHECK2(hipGetLastError());
    lu_factor<double>(s, G64, ld, npad, linv64, uinv64, dbad, cfg);
    int hbad = 0;
    HIPCHECK2(hipMemcpyAsync(&hbad, dbad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    if (hbad) throw std::domain_error("zero or non-finite pivot in the randomized LU (singular bordered matrix)");
    for (void *p : {(void *)G32, (void *)linv32, (void *)uinv32}) {
        if (!p) continue;
        owned.erase(std::find(owned.begin(), owned.end(), p));
        (void)hipFree(p);
    }
    G32 = nullptr;
    linv32 = uinv32 = nullptr;
    if (ginv) {  // the FP32 factors' inverse goes with them
        owned.erase(std::find(owned.begin(), owned.end(), ginv));
        (void)hipFree(ginv);
        ginv = nullptr;
    }
    inv_ready = false;
    narrow_applies = 0;
    mode = 0;
    fallbacks++;
}

// Hager-Higham estimate of |M^-1|_1 (the iteration of LAPACK dlacn2, which gecon runs on the reference's factors).
// It needs products with M^-1 and M^-T, both from the one factorization.
double Factor::cond_estimate()
{
    // products with the factors alone, as gecon takes them (FP32 factors give the estimate to about 1e-4)
    const int kit = 0;
    std::vector<double> x(n, 1.0 / n), y(n), z(n);
    double est = 0.0;
    int jlast = -1;
    for (int iter = 0; iter < 5; iter++) {
        solve(x.data(), 1, y.data(), false, kit, nullptr, false, false);  // y = M^-1 x
        double ny = 0.0;
        for (double v : y) ny += std::fabs(v);
        if (iter > 0 && ny <= est) {
            est = std::max(est, ny);
            break;
        }
        est = ny;
        for (int i = 0; i < n; i++) z[i] = y[i] >= 0 ? 1.0 : -1.0;
        solve(z.data(), 1, x.data(), false, kit, nullptr, true, false);  // x = M^-T z
        int j = 0;
        for (int i = 1; i < n; i++)
            if (std::fabs(x[i]) > std::fabs(x[j])) j = i;
        if (j == jlast) break;
        jlast = j;
        std::fill(x.begin(), x.end(), 0.0);
        x[j] = 1.0;
    }
    // the alternative estimate of dlacn2 (altsgn)
    std::vector<double> alt(n);
    for (int i = 0; i < n; i++) alt[i] = (i % 2 ? -1.0 : 1.0) * (1.0 + (double)i / (n > 1 ? n - 1 : 1));
    solve(alt.data(), 1, y.data(), false, kit, nullptr, false, false);
    double t = 0.0;
    for (double v : y) t += std::fabs(v);
    t = 2 * t / (3.0 * n);
    est = std::max(est, t);
    return est;
}

// Factor a dense bordered matrix already on the device (n x n in an npad x npad row-major buffer).
void Factor::init_from_dense(const double *Md, long long ldd, int n_, int npad_, int mode_, double anorm)
{
    hipStream_t s = ctx->stream;
    mode = mode_;
    n = n_;
    npad = npad_;
    ld = ldd;
    nq = n;
    nb = 0;
    anorm1 = anorm;
    M = const_cast<double *>(Md);
    std::mt19937_64 rng(20261204ull);
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    std::vector<double> ru(2 * npad), rv(2 * npad);
    for (auto &v : ru) v = std::exp(U(rng) / 10);
    for (auto &v : rv) v = std::exp(U(rng) / 10);
    d_ru = dmalloc<double>(ru.size());
    d_rv = dmalloc<double>(rv.size());
    owned.insert(owned.end(), {d_ru, d_rv});
    HIPCHECK2(hipMemcpyAsync(d_ru, ru.data(), sizeof(double) * ru.size(), hipMemcpyHostToDevice, s));
    HIPCHECK2(hipMemcpyAsync(d_rv, rv.data(), sizeof(double) * rv.size(), hipMemcpyHostToDevice, s));
    int *dbad = dmalloc<int>(1);
    owned.push_back(dbad);
    HIPCHECK2(hipMemsetAsync(dbad, 0, sizeof(int), s));
    size_t nblk = (size_t)(npad / NB) * NB * NB;
    long long h = npad / 4;
    if (mode == 0) {
        G64 = dmalloc<double>((size_t)ld * npad);
        linv64 = dmalloc<double>(nblk);
        uinv64 = dmalloc<double>(nblk);
        owned.insert(owned.end(), {G64, linv64, uinv64});
        const int cfg = tuned_cfg<double>(s, G64, ld, npad);
        rbt_both<double><<<blocks(h * h), 256, 0, s>>>(M, ld, npad, d_ru, d_rv, G64, ld);
        lu_factor<double>(s, G64, ld, npad, linv64, uinv64, dbad, cfg);
    } else {
        G32 = dmalloc<float>((size_t)ld * npad);
        linv32 = dmalloc<float>(nblk);
        uinv32 = dmalloc<float>(nblk);
        owned.insert(owned.end(), {G32, linv32, uinv32});
        const int cfg = tuned_cfg<float>(s, G32, ld, npad);
        rbt_both<float><<<blocks(h * h), 256, 0, s>>>(M, ld, npad, d_ru, d_rv, G32, ld);
        lu_factor<float>(s, G32, ld, npad, linv32, uinv32, dbad, cfg);
    }
    HIPCHECK2(hipGetLastError());
}

// ------------------------------------------------------------------------------------------ test hooks ------------

template <typename T> void gemm_flags(hipStream_t s, int flags, int M_, int N_, int K_, const T *A, const T *B, T *C)
{
    const bool sub = flags & 1, ta = flags & 2;
    const int tile = (flags >> 4) & 15;
    const long long lda = ta ? M_ : K_;
    if (sub && ta)
        launch_gemm_cfg<T, true, true>(tile, s, M_, N_, K_, A, lda, B, N_, C, N_);
    else if (sub)
        launch_gemm_cfg<T, true, false>(tile, s, M_, N_, K_, A, lda, B, N_, C, N_);
    else if (ta)
        launch_gemm_cfg<T, false, true>(tile, s, M_, N_, K_, A, lda, B, N_, C, N_);
    else
        launch_gemm_cfg<T, false, false>(tile, s, M_, N_, K_, A, lda, B, N_, C, N_);
}

void gemm_host(hipStream_t s, int prec, int flags, int M_, int N_, int K_, const double *A, const double *B, double *C)
{
    size_t na = (size_t)M_ * K_, nbb = (size_t)K_ * N_, nc = (size_t)M_ * N_;
    double *dA = dmalloc<double>(na), *dB = dmalloc<double>(nbb), *dC = dmalloc<double>(nc);
    HIPCHECK2(hipMemcpyAsync(dA, A, na * 8, hipMemcpyHostToDevice, s));
    HIPCHECK2(hipMemcpyAsync(dB, B, nbb * 8, hipMemcpyHostToDevice, s));
    HIPCHECK2(hipMemcpyAsync(dC, C, nc * 8, hipMemcpyHostToDevice, s));
    if (prec == 0) {
        gemm_flags<double>(s, flags, M_, N_, K_, dA, dB, dC);
    } else {
        float *fA = dmalloc<float>(na), *fB = dmalloc<float>(nbb), *fC = dmalloc<float>(nc);
        convert<double, float><<<blocks(na), 256, 0, s>>>(dA, fA, na);
        convert<double, float><<<blocks(nbb), 256, 0, s>>>(dB, fB, nbb);
        convert<double, float><<<blocks(nc), 256, 0, s>>>(dC, fC, nc);
        gemm_flags<float>(s, flags, M_, N_, K_, fA, fB, fC);
        convert<float, double><<<blocks(nc), 256, 0, s>>>(fC, dC, nc);
        HIPCHECK2(hipStreamSynchronize(s));
        (void)hipFree(fA);
        (void)hipFree(fB);
        (void)hipFree(fC);
    }
    HIPCHECK2(hipMemcpyAsync(C, dC, nc * 8, hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
}

// Test hook: diag_lu on one block (NB x NB row-major A, nb <= NB), its packed factors and both inverses back
void diag_lu_host(hipStream_t s, int prec, int nb, const double *A, double *LU, double *Linv, double *Uinv)
{
    const size_t nn = (size_t)NB * NB;
    double *dA = dmalloc<double>(nn), *dL = dmalloc<double>(nn), *dU = dmalloc<double>(nn);
    int *bad = dmalloc<int>(1);
    HIPCHECK2(hipMemcpyAsync(dA, A, nn * 8, hipMemcpyHostToDevice, s));
    HIPCHECK2(hipMemsetAsync(bad, 0, sizeof(int), s));
    if (prec == 0) {
        diag_lu<double><<<1, 256, 0, s>>>(dA, NB, 0, nb, dL, dU, bad);
    } else {
        float *fA = dmalloc<float>(nn), *fL = dmalloc<float>(nn), *fU = dmalloc<float>(nn);
        convert<double, float><<<blocks(nn), 256, 0, s>>>(dA, fA, nn);
        diag_lu<float><<<1, 256, 0, s>>>(fA, NB, 0, nb, fL, fU, bad);
        convert<float, double><<<blocks(nn), 256, 0, s>>>(fA, dA, nn);
        convert<float, double><<<blocks(nn), 256, 0, s>>>(fL, dL, nn);
        convert<float, double><<<blocks(nn), 256, 0, s>>>(fU, dU, nn);
        HIPCHECK2(hipStreamSynchronize(s));
        (void)hipFree(fA);
        (void)hipFree(fL);
        (void)hipFree(fU);
    }
    HIPCHECK2(hipGetLastError());
    HIPCHECK2(hipMemcpyAsync(LU, dA, nn * 8, hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipMemcpyAsync(Linv, dL, nn * 8, hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipMemcpyAsync(Uinv, dU, nn * 8, hipMemcpyDeviceToHost, s));
    HIPCHECK2(hipStreamSynchronize(s));
    for (void *p : {(void *)dA, (void *)dL, (void *)dU, (void *)bad}) (void)hipFree(p);
}

int dense_solve_host(Ctx *c, int n, const double *Mh, int mode, const double *r, double *x)
{
    hipStream_t s = c->stream;
    int npad = ((n + 255) / 256) * 256;
    double *dM = dmalloc<double>((size_t)npad * npad);
    std::vector<double> hM((size_t)npad * npad, 0.0);
    for (int i = 0; i < npad; i++)
        for (int j = 0; j < npad; j++) hM[(size_t)i * npad + j] = (i < n && j < n) ? Mh[(size_t)i * n + j] : (i == j ? 1.0 : 0.0);
    HIPCHECK2(hipMemcpyAsync(dM, hM.data(), hM.size() * 8, hipMemcpyHostToDevice, s));
    Factor f;
    f.ctx = c;
    f.owned.push_back(dM);
    f.init_from_dense(dM, npad, n, npad, mode, 0.0);
    int it = f.solve(r, 1, x, false, 30, nullptr);
    return it;
}

// ------------------------------------------------------------------------------------------ consistency ------------
// `consistency(dm, A, b; W)` (src/assembly/solve.jl:111-115): with ACf = A [C; W] (nslot x nq), the largest absolute
// row sum and column sum of ACf over its largest absolute entry, and |sum b| / sum |b|.

// One workgroup per column q of Cf = [C; W] (CSC): ACf[:, q] = sum over the column's entries of v A[:, j]; its
// largest absolute entry and its sum (threads over the rows, coalesced down A's columns; fixed reduction order).
__global__ __launch_bounds__(256) void acf_columns(const double *__restrict__ A, long long lda, int nrow,
                                                   const int *__restrict__ fp, const int *__restrict__ fi,
                                                   const double *__restrict__ fv, double *__restrict__ colmax,
                                                   double *__restrict__ colsum)
{
    __shared__ double smax[256], ssum[256];
    const int q = blockIdx.x, a0 = fp[q], a1 = fp[q + 1];
    double mx = 0.0, sm = 0.0;
    for (int i = threadIdx.x; i < nrow; i += 256) {
        double v = 0.0;
        for (int a = a0; a < a1; a++) v += fv[a] * A[i + (long long)fi[a] * lda];
        mx = fmax(mx, fabs(v));
        sm += v;
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
            ssum[threadIdx.x] += ssum[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        colmax[q] = smax[0];
        colsum[q] = ssum[0];
    }
}

// Row sums of ACf as A (Cf 1): partial sums over column chunks (grid.y), thread per row, then added in chunk order
__global__ __launch_bounds__(256) void av_rows(const double *__restrict__ A, long long lda, int nrow, int ncol,
                                               const double *__restrict__ v, int chunk, double *__restrict__ part)
{
    const int i = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (i >= nrow) return;
    const int j1 = min(ncol, (c + 1) * chunk);
    double s = 0.0;
    for (int j = c * chunk; j < j1; j++) s += A[i + (long long)j * lda] * v[j];
    part[(long long)c * nrow + i] = s;
}

__global__ void av_rows_finish(const double *__restrict__ part, int nrow, int nchunk, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(long long)c * nrow + i];
    out[i] = s;
}

void consistency_host(Ctx *c, const double *A, long long lda, const double *b, const vp_constraints &k, double *out)
{
    hipStream_t s = c->stream;
    const int nslot = k.nslot, nq = k.nq, nw = k.nw, ncol = nslot + nw;
    std::vector<std::vector<std::pair<int, double>>> fcol(nq);
    std::vector<double> v1(ncol, 0.0);  // Cf 1
    for (int i = 0; i < nslot; i++)
        for (int a = k.C_rowptr[i]; a < k.C_rowptr[i + 1]; a++) {
            const int q = k.C_col[a];
            if (q < 0 || q >= nq) throw std::out_of_range("C column index out of range");
            fcol[q].push_back({i, k.C_val[a]});
            v1[i] += k.C_val[a];
        }
    for (int i = 0; i < nw; i++)
        for (int a = k.W_rowptr[i]; a < k.W_rowptr[i + 1]; a++) {
            const int q = k.W_col[a];
            if (q < 0 || q >= nq) throw std::out_of_range("W column index out of range");
            fcol[q].push_back({nslot + i, k.W_val[a]});
            v1[nslot + i] += k.W_val[a];
        }
    std::vector<int> fp(1, 0), fi;
    std::vector<double> fv;
    for (auto &col : fcol) {
        for (auto &e : col) {
            fi.push_back(e.first);
            fv.push_back(e.second);
        }
        fp.push_back((int)fi.size());
    }
    const int chunk = 512, nchunk = (ncol + chunk - 1) / chunk;
    std::vector<void *> mine;
    auto dev = [&](size_t bytes) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) throw std::runtime_error("device allocation failed");
        mine.push_back(p);
        return p;
    };
    try {
        int *dfp = (int *)dev(fp.size() * 4), *dfi = (int *)dev(fi.size() * 4);
        double *dfv = (double *)dev(fv.size() * 8), *dv = (double *)dev((size_t)ncol * 8);
        double *cmax = (double *)dev((size_t)nq * 8), *csum = (double *)dev((size_t)nq * 8);
        double *part = (double *)dev((size_t)nchunk * nslot * 8), *rsum = (double *)dev((size_t)nslot * 8);
        HIPCHECK2(hipMemcpyAsync(dfp, fp.data(), fp.size() * 4, hipMemcpyHostToDevice, s));
        if (!fi.empty()) {
            HIPCHECK2(hipMemcpyAsync(dfi, fi.data(), fi.size() * 4, hipMemcpyHostToDevice, s));
            HIPCHECK2(hipMemcpyAsync(dfv, fv.data(), fv.size() * 8, hipMemcpyHostToDevice, s));
        }
        HIPCHECK2(hipMemcpyAsync(dv, v1.data(), (size_t)ncol * 8, hipMemcpyHostToDevice, s));
        acf_columns<<<nq, 256, 0, s>>>(A, lda, nslot, dfp, dfi, dfv, cmax, csum);
        HIPCHECK2(hipGetLastError());
        av_rows<<<dim3((nslot + 255) / 256, nchunk), 256, 0, s>>>(A, lda, nslot, ncol, dv, chunk, part);
        HIPCHECK2(hipGetLastError());
        av_rows_finish<<<(nslot + 255) / 256, 256, 0, s>>>(part, nslot, nchunk, rsum);
        HIPCHECK2(hipGetLastError());
        std::vector<double> hmax(nq), hsum(nq), hrow(nslot);
        HIPCHECK2(hipMemcpyAsync(hmax.data(), cmax, (size_t)nq * 8, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipMemcpyAsync(hsum.data(), csum, (size_t)nq * 8, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipMemcpyAsync(hrow.data(), rsum, (size_t)nslot * 8, hipMemcpyDeviceToHost, s));
        HIPCHECK2(hipStreamSynchronize(s));
        double nA = 0.0, cs = 0.0, rs = 0.0;
        for (int q = 0; q < nq; q++) {
            nA = std::max(nA, hmax[q]);
            cs = std::max(cs, std::fabs(hsum[q]));
        }
        for (int i = 0; i < nslot; i++) rs = std::max(rs, std::fabs(hrow[i]));
        double bs = 0.0, ba = 0.0;
        for (int i = 0; i < nslot; i++) {
            bs += b[i];
            ba += std::fabs(b[i]);
        }
        out[0] = rs / nA;
        out[1] = cs / nA;
        out[2] = std::fabs(bs) / ba;
    } catch (...) {
        (void)hipStreamSynchronize(s);
        for (void *p : mine) (void)hipFree(p);
        throw;
    }
    for (void *p : mine) (void)hipFree(p);
}

}  // namespace vp
