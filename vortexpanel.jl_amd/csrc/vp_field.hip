// N-body velocity sums on gfx950: the surface velocity of Post.surface_velocity (src/post/Post.jl:32-64), the body's
// velocity at off-body points of Post.body_velocity (Post.jl:92-112), and the free wake's direct blob sums
// (blob_velocities, src/freewake/step.jl:176-187).
//
// Layout: one target point per thread, 256 targets per block (blockIdx.x); the sources are cut into grid.y chunks of
// whole 64-element tiles.  A tile is staged in LDS and every lane reads the same record (a broadcast, no bank
// conflicts).  Each chunk writes its partial sums to partial[chunk][target]; field_reduce adds the chunks in a fixed
// order (so results do not depend on scheduling), then the freestream and the wake's tip filaments.
//
// Precision: T = double evaluates every kernel in FP64 exactly as the reference; T = float evaluates the source
// triangles and blobs in FP32 relative to the element's centroid (offsets formed in FP64), sums a tile in FP32 and
// carries the running sum in FP64.  Sources within FP64_NEAR diameters of the target (where the closed form's logs
// and solid angle cancel) and the self term of the surface velocity (the in-plane limit) are FP64 in both.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "vp_capi_internal.hpp"
#include "vp_device.hpp"

namespace vp {

__global__ void elem_setup(const double *x, const int *tris, int ne, Elem *out);  // vp_assemble.hip

namespace {

constexpr int FT = 256;  // targets per block
constexpr int FS = 64;   // source elements per LDS tile
constexpr int BT = 256;  // blobs per LDS tile
constexpr double FP64_NEAR = 3.0;  // FP32 path: sources closer than this many diameters are evaluated in FP64

#define FCHECK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T> struct FieldSrc {  // a source element as the field kernels read it
    double c[3];                         // centroid (FP64: the far test and the offsets of the target)
    double diam;
    T v[3][3];  // vertex offsets from the centroid
    T n[3];
    T nu[3][3];
    T len[3];
    T area, sigma;
    T gam[3];
};

// Vortex sheet of every element from its corners' doublet values (sheet_strength, src/elements/flat_triangle.jl:39;
// element_values, src/post/Post.jl:19): gam = sum_b mu[corner[b]] g_b.  Rows beyond ne_mu get gam = 0.
__global__ void sheet_gamma(const Elem *__restrict__ el, const int *__restrict__ corner, const double *__restrict__ mu,
                            int ne, double *__restrict__ gam)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const Elem &E = el[e];
    double g[3] = {0, 0, 0};
    for (int b = 0; b < 3; b++) {
        const double m = mu[corner[3 * e + b]];
        for (int k = 0; k < 3; k++) g[k] += m * E.g[b][k];
    }
    for (int k = 0; k < 3; k++) gam[3 * e + k] = g[k];
}

// Sum over sources s in this block's chunk of sigma_s w_s(x) + gam_s x w_s(x) (Post.jl:43-46, :103-106).
// self_mode: target i is the centroid of element i, whose own term is the one-sided limit on the normal's side.
template <typename T>
__global__ __launch_bounds__(FT) void source_field(const Elem *__restrict__ el, const double *__restrict__ sigma,
                                                   const double *__restrict__ gam, int ns, int chunk,
                                                   const double *__restrict__ X, int np, int self_mode,
                                                   double kernel_far, double *__restrict__ partial)
{
    __shared__ FieldSrc<T> S[FS];
    const int i = blockIdx.x * FT + threadIdx.x;
    const bool live = i < np;
    const V3<double> x = live ? ld3(X + 3 * i) : mk3(0.0, 0.0, 0.0);
    const int s0 = blockIdx.y * chunk, s1 = min(ns, s0 + chunk);
    double ux = 0.0, uy = 0.0, uz = 0.0;
    for (int base = s0; base < s1; base += FS) {
        const int nk = min(FS, s1 - base);
        __syncthreads();
        if (threadIdx.x < nk) {
            const Elem &E = el[base + threadIdx.x];
            FieldSrc<T> &q = S[threadIdx.x];
            for (int k = 0; k < 3; k++) {
                q.c[k] = E.c[k];
                q.n[k] = (T)E.n[k];
                q.len[k] = (T)E.len[k];
                q.gam[k] = (T)gam[3 * (base + threadIdx.x) + k];
                for (int j = 0; j < 3; j++) {
                    q.v[k][j] = (T)(E.v[k][j] - E.c[j]);
                    q.nu[k][j] = (T)E.nu[k][j];
                }
            }
            q.diam = E.diam;
            q.area = (T)E.area;
            q.sigma = sigma ? (T)sigma[base + threadIdx.x] : T(0);
        }
        __syncthreads();
        V3<T> acc = mk3(T(0), T(0), T(0));
#pragma unroll 1
        for (int k = 0; k < nk; k++) {
            const FieldSrc<T> &q = S[k];
            const V3<double> rd = cls_sub(x, mk3(q.c[0], q.c[1], q.c[2]));
            const double fd = kernel_far * q.diam;
            V3<T> w;
            if (cls_dot(rd, rd) > fd * fd) {  // source_velocity's switch, source_triangle.jl:136
                SrcT<T> s;
                for (int a = 0; a < 3; a++) s.v[a] = mk3(q.v[a][0], q.v[a][1], q.v[a][2]);
                s.area = q.area;
                w = sv_far_rel(s, mk3((T)rd.x, (T)rd.y, (T)rd.z));
            } else if ((self_mode && base + k == i) ||
                       (sizeof(T) == 4 && cls_dot(rd, rd) <= FP64_NEAR * FP64_NEAR * q.diam * q.diam)) {
                // the own element (side +1, source_triangle.jl:80-81) and, in the FP32 path, every source within
                // FP64_NEAR diameters: FP64 from the element record, where the closed form cancels
                const Elem &E = el[base + k];
                SrcT<double> s;
                for (int a = 0; a < 3; a++) {
                    s.nu[a] = ld3(E.nu[a]);
                    s.len[a] = E.len[a];
                }
                s.n = ld3(E.n);
                s.diam = E.diam;
                const int side = self_mode && base + k == i ? 1 : 0;
                const V3<double> w64 = sv_exact_abs(s, x - ld3(E.v[0]), x - ld3(E.v[1]), x - ld3(E.v[2]),
                                                    cls_dot(rd, s.n), side, side ? &E : nullptr, x);
                w = mk3((T)w64.x, (T)w64.y, (T)w64.z);
            } else {
                SrcT<T> s;
                for (int a = 0; a < 3; a++) {
                    s.nu[a] = mk3(q.nu[a][0], q.nu[a][1], q.nu[a][2]);
                    s.len[a] = q.len[a];
                }
                s.n = mk3(q.n[0], q.n[1], q.n[2]);
                s.diam = (T)q.diam;
                const V3<T> r = mk3((T)rd.x, (T)rd.y, (T)rd.z);
                const T z = (T)cls_dot(rd, mk3((double)q.n[0], (double)q.n[1], (double)q.n[2]));
                w = sv_exact_abs(s, r - mk3(q.v[0][0], q.v[0][1], q.v[0][2]), r - mk3(q.v[1][0], q.v[1][1], q.v[1][2]),
                                 r - mk3(q.v[2][0], q.v[2][1], q.v[2][2]), z, 0, nullptr, x);
            }
            acc = acc + q.sigma * w + cross(mk3(q.gam[0], q.gam[1], q.gam[2]), w);
        }
        ux += (double)acc.x;
        uy += (double)acc.y;
        uz += (double)acc.z;
    }
    if (live) {
        double *p = partial + 3 * ((size_t)blockIdx.y * np + i);
        p[0] = ux;
        p[1] = uy;
        p[2] = uz;
    }
}

// Blob sums, step.jl:176-187 with blob_velocity (src/kernels/blob.jl:21-25): tiles of 256 blobs in LDS.  In FP32
// the positions are taken relative to the block's first target (x0, FP64), so a blob next to the block's targets
// keeps the digits of x - xp where the kernel is largest.
template <typename T>
__global__ __launch_bounds__(FT) void blob_field(const double *__restrict__ bx, const double *__restrict__ ba,
                                                 const double *__restrict__ bd, int nb, int chunk,
                                                 const double *__restrict__ X, int np, double *__restrict__ partial)
{
    __shared__ T P[BT][7];  // position - x0, strength, core radius squared
    const int i = blockIdx.x * FT + threadIdx.x;
    const bool live = i < np;
    const V3<double> x0 = ld3(X + 3 * (size_t)(blockIdx.x * FT));
    const V3<double> x = live ? ld3(X + 3 * i) : x0;
    const V3<T> xt = mk3((T)(x.x - x0.x), (T)(x.y - x0.y), (T)(x.z - x0.z));
    const int s0 = blockIdx.y * chunk, s1 = min(nb, s0 + chunk);
    double ux = 0.0, uy = 0.0, uz = 0.0;
    for (int base = s0; base < s1; base += BT) {
        const int nk = min(BT, s1 - base);
        __syncthreads();
        if (threadIdx.x < nk) {
            const int b = base + threadIdx.x;
            const double o[3] = {x0.x, x0.y, x0.z};
            for (int k = 0; k < 3; k++) {
                P[threadIdx.x][k] = (T)(bx[3 * b + k] - o[k]);
                P[threadIdx.x][3 + k] = (T)ba[3 * b + k];
            }
            P[threadIdx.x][6] = (T)(bd[b] * bd[b]);
        }
        __syncthreads();
        V3<T> acc = mk3(T(0), T(0), T(0));
        for (int k = 0; k < nk; k++) {
            const V3<T> r = xt - mk3(P[k][0], P[k][1], P[k][2]);
            acc = acc + blob_vel(r, mk3(P[k][3], P[k][4], P[k][5]), P[k][6]);
        }
        ux += (double)acc.x;
        uy += (double)acc.y;
        uz += (double)acc.z;
    }
    if (live) {
        double *p = partial + 3 * ((size_t)blockIdx.y * np + i);
        p[0] = ux;
        p[1] = uy;
        p[2] = uz;
    }
}

// blob_field in FP64 evaluates r = x - xp in FP64 (the reference's order); the FP32 variant above forms it in FP32.
template <>
__global__ __launch_bounds__(FT) void blob_field<double>(const double *__restrict__ bx, const double *__restrict__ ba,
                                                         const double *__restrict__ bd, int nb, int chunk,
                                                         const double *__restrict__ X, int np,
                                                         double *__restrict__ partial)
{
    __shared__ double P[BT][7];
    const int i = blockIdx.x * FT + threadIdx.x;
    const bool live = i < np;
    const V3<double> x = live ? ld3(X + 3 * i) : mk3(0.0, 0.0, 0.0);
    const int s0 = blockIdx.y * chunk, s1 = min(nb, s0 + chunk);
    V3<double> acc = mk3(0.0, 0.0, 0.0);
    for (int base = s0; base < s1; base += BT) {
        const int nk = min(BT, s1 - base);
        __syncthreads();
        if (threadIdx.x < nk) {
            const int b = base + threadIdx.x;
            for (int k = 0; k < 3; k++) {
                P[threadIdx.x][k] = bx[3 * b + k];
                P[threadIdx.x][3 + k] = ba[3 * b + k];
            }
            P[threadIdx.x][6] = bd[b] * bd[b];
        }
        __syncthreads();
#pragma unroll 4  // four interactions in flight (84k x 84k: 9.3 -> 8.2 ms)
        for (int k = 0; k < nk; k++)
            acc = acc + blob_vel(x - mk3(P[k][0], P[k][1], P[k][2]), mk3(P[k][3], P[k][4], P[k][5]), P[k][6]);
    }
    if (live) {
        double *p = partial + 3 * ((size_t)blockIdx.y * np + i);
        p[0] = acc.x;
        p[1] = acc.y;
        p[2] = acc.z;
    }
}

// u[i] = base + sum over chunks (in chunk order) + the wake filaments (Post.jl:50-53); base is V∞ or 0.
__global__ void field_reduce(const double *__restrict__ partial, int nchunk, int np, double bx, double by, double bz,
                             const double *__restrict__ wx, const int *__restrict__ fil, const double *__restrict__ wmu,
                             int nfil, const double *__restrict__ X, int accumulate, double *__restrict__ u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    V3<double> acc = mk3(bx, by, bz);
    if (accumulate) acc = acc + ld3(u + 3 * i);
    for (int c = 0; c < nchunk; c++) acc = acc + ld3(partial + 3 * ((size_t)c * np + i));
    if (nfil) {
        const V3<double> x = ld3(X + 3 * i);
        for (int f = 0; f < nfil; f++) {
            const int a = fil[2 * f], b = fil[2 * f + 1];
            V3<double> va, vb;
            seg_vel(x - ld3(wx + 3 * a), x - ld3(wx + 3 * b), va, vb);
            acc = acc + wmu[a] * va + wmu[b] * vb;
        }
    }
    u[3 * i] = acc.x;
    u[3 * i + 1] = acc.y;
    u[3 * i + 2] = acc.z;
}

__global__ void centroids(const Elem *__restrict__ el, int ne, double *__restrict__ X)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne)
        for (int k = 0; k < 3; k++) X[3 * e + k] = el[e].c[k];
}

template <typename T> T *dnew(size_t n, std::vector<void *> &owned)
{
    T *p = nullptr;
    if (n == 0) n = 1;
    FCHECK(hipMalloc((void **)&p, n * sizeof(T)));
    owned.push_back(p);
    return p;
}

template <typename T> T *dput(const T *h, size_t n, hipStream_t st, std::vector<void *> &owned)
{
    T *p = dnew<T>(n, owned);
    if (n && h) FCHECK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    return p;
}

struct Owned {
    std::vector<void *> v;
    hipStream_t st;
    ~Owned()
    {
        (void)hipStreamSynchronize(st);
        for (void *p : v) (void)hipFree(p);
    }
};

// Chunks of whole tiles so that the grid has at least ~2048 blocks (8 per CU) when the sources allow it.
int chunk_of(int ns, int np, int tile)
{
    const int tblocks = (np + FT - 1) / FT;
    const int ntiles = (ns + tile - 1) / tile;
    int nchunk = std::max(1, std::min(ntiles, (2048 + tblocks - 1) / tblocks));
    const int per = (ntiles + nchunk - 1) / nchunk;
    return per * tile;
}

}  // namespace

void source_field_device(hipStream_t st, int precision, const Elem *el, const double *sigma, const double *gam, int ns,
                         const double *X, int np, int self_mode, double kernel_far, double *partial, int *nchunk_out)
{
    const int chunk = chunk_of(ns, np, FS);
    const int nchunk = ns > 0 ? (ns + chunk - 1) / chunk : 0;
    *nchunk_out = nchunk;
    if (!nchunk || !np) return;
    dim3 grid((np + FT - 1) / FT, nchunk);
    if (precision == 1)
        source_field<float><<<grid, FT, 0, st>>>(el, sigma, gam, ns, chunk, X, np, self_mode, kernel_far, partial);
    else
        source_field<double><<<grid, FT, 0, st>>>(el, sigma, gam, ns, chunk, X, np, self_mode, kernel_far, partial);
    FCHECK(hipGetLastError());
}

void blob_field_device(hipStream_t st, int precision, const double *bx, const double *ba, const double *bd, int nb,
                       const double *X, int np, double *partial, int *nchunk_out)
{
    const int chunk = chunk_of(nb, np, BT);
    const int nchunk = nb > 0 ? (nb + chunk - 1) / chunk : 0;
    *nchunk_out = nchunk;
    if (!nchunk || !np) return;
    dim3 grid((np + FT - 1) / FT, nchunk);
    if (precision == 1)
        blob_field<float><<<grid, FT, 0, st>>>(bx, ba, bd, nb, chunk, X, np, partial);
    else
        blob_field<double><<<grid, FT, 0, st>>>(bx, ba, bd, nb, chunk, X, np, partial);
    FCHECK(hipGetLastError());
}

size_t field_partial_size(int ns, int np, int tile_blobs)
{
    const int tile = tile_blobs ? BT : FS;
    const int chunk = chunk_of(ns, np, tile);
    const int nchunk = ns > 0 ? (ns + chunk - 1) / chunk : 0;
    return (size_t)std::max(nchunk, 1) * np * 3;
}

void field_reduce_device(hipStream_t st, const double *partial, int nchunk, int np, const double *base,
                         const double *wx, const int *fil, const double *wmu, int nfil, const double *X, int accumulate,
                         double *u)
{
    if (!np) return;
    field_reduce<<<(np + 255) / 256, 256, 0, st>>>(partial, nchunk, np, base ? base[0] : 0.0, base ? base[1] : 0.0,
                                                   base ? base[2] : 0.0, wx, fil, wmu, nfil, X, accumulate, u);
    FCHECK(hipGetLastError());
}

void surface_velocity_host(Ctx *c, const vp_surface &g, const vp_wake *w, const double *sigma, const double *mu,
                           const double *mu_wake, const double *vinf, double kernel_far, int precision, double *u)
{
    hipStream_t st = c->stream;
    Owned o{{}, st};
    const int nt = g.nt, nwt = w ? w->ntri : 0, ne = nt + nwt;
    double *d_x = dput(g.x, (size_t)g.nn * 3, st, o.v);
    int *d_tris = dput(g.tris, (size_t)nt * 3, st, o.v);
    int *d_cs = dput(g.corner_slot, (size_t)nt * 3, st, o.v);
    double *d_mu = dput(mu, (size_t)g.nslot, st, o.v);
    Elem *d_el = dnew<Elem>(ne, o.v);
    elem_setup<<<(nt + 255) / 256, 256, 0, st>>>(d_x, d_tris, nt, d_el);
    FCHECK(hipGetLastError());
    double *d_gam = dnew<double>((size_t)ne * 3, o.v);
    sheet_gamma<<<(nt + 255) / 256, 256, 0, st>>>(d_el, d_cs, d_mu, nt, d_gam);
    FCHECK(hipGetLastError());
    double *d_sig = dnew<double>(ne, o.v);
    FCHECK(hipMemsetAsync(d_sig, 0, sizeof(double) * ne, st));
    FCHECK(hipMemcpyAsync(d_sig, sigma, sizeof(double) * nt, hipMemcpyHostToDevice, st));
    double *d_wx = nullptr, *d_wmu = nullptr;
    int *d_fil = nullptr;
    if (w) {
        d_wx = dput(w->x, (size_t)w->nw * 3, st, o.v);
        d_wmu = dput(mu_wake, (size_t)w->nw, st, o.v);
        if (nwt) {
            int *d_wt = dput(w->tris, (size_t)nwt * 3, st, o.v);
            elem_setup<<<(nwt + 255) / 256, 256, 0, st>>>(d_wx, d_wt, nwt, d_el + nt);
            FCHECK(hipGetLastError());
            sheet_gamma<<<(nwt + 255) / 256, 256, 0, st>>>(d_el + nt, d_wt, d_wmu, nwt, d_gam + 3 * (size_t)nt);
            FCHECK(hipGetLastError());
        }
        if (w->nfil) d_fil = dput(w->filaments, (size_t)w->nfil * 2, st, o.v);
    }
    double *d_X = dnew<double>((size_t)nt * 3, o.v);
    centroids<<<(nt + 255) / 256, 256, 0, st>>>(d_el, nt, d_X);
    FCHECK(hipGetLastError());
    double *d_part = dnew<double>(field_partial_size(ne, nt, 0), o.v);
    double *d_u = dnew<double>((size_t)nt * 3, o.v);
    int nchunk = 0;
    source_field_device(st, precision, d_el, d_sig, d_gam, ne, d_X, nt, 1, kernel_far, d_part, &nchunk);
    field_reduce_device(st, d_part, nchunk, nt, vinf, d_wx, d_fil, d_wmu, w ? w->nfil : 0, d_X, 0, d_u);
    FCHECK(hipMemcpyAsync(u, d_u, sizeof(double) * nt * 3, hipMemcpyDeviceToHost, st));
    FCHECK(hipStreamSynchronize(st));
}

void body_velocity_host(Ctx *c, const vp_surface &g, const double *sigma, const double *gam, int np, const double *xs,
                        double kernel_far, int precision, double *u)
{
    hipStream_t st = c->stream;
    Owned o{{}, st};
    const int nt = g.nt;
    double *d_x = dput(g.x, (size_t)g.nn * 3, st, o.v);
    int *d_tris = dput(g.tris, (size_t)nt * 3, st, o.v);
    Elem *d_el = dnew<Elem>(nt, o.v);
    elem_setup<<<(nt + 255) / 256, 256, 0, st>>>(d_x, d_tris, nt, d_el);
    FCHECK(hipGetLastError());
    double *d_sig = dput(sigma, (size_t)nt, st, o.v);
    double *d_gam = dput(gam, (size_t)nt * 3, st, o.v);
    double *d_X = dput(xs, (size_t)np * 3, st, o.v);
    double *d_part = dnew<double>(field_partial_size(nt, np, 0), o.v);
    double *d_u = dnew<double>((size_t)np * 3, o.v);
    int nchunk = 0;
    source_field_device(st, precision, d_el, d_sig, d_gam, nt, d_X, np, 0, kernel_far, d_part, &nchunk);
    field_reduce_device(st, d_part, nchunk, np, nullptr, nullptr, nullptr, nullptr, 0, d_X, 0, d_u);
    FCHECK(hipMemcpyAsync(u, d_u, sizeof(double) * np * 3, hipMemcpyDeviceToHost, st));
    FCHECK(hipStreamSynchronize(st));
}

void blob_velocities_host(Ctx *c, int nb, const double *bx, const double *ba, const double *bd, int np,
                          const double *xs, int precision, double *u)
{
    hipStream_t st = c->stream;
    Owned o{{}, st};
    double *d_bx = dput(bx, (size_t)nb * 3, st, o.v);
    double *d_ba = dput(ba, (size_t)nb * 3, st, o.v);
    double *d_bd = dput(bd, (size_t)nb, st, o.v);
    double *d_X = dput(xs, (size_t)np * 3, st, o.v);
    double *d_part = dnew<double>(field_partial_size(nb, np, 1), o.v);
    double *d_u = dnew<double>((size_t)np * 3, o.v);
    int nchunk = 0;
    blob_field_device(st, precision, d_bx, d_ba, d_bd, nb, d_X, np, d_part, &nchunk);
    field_reduce_device(st, d_part, nchunk, np, nullptr, nullptr, nullptr, nullptr, 0, d_X, 0, d_u);
    FCHECK(hipMemcpyAsync(u, d_u, sizeof(double) * np * 3, hipMemcpyDeviceToHost, st));
    FCHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------------------------------ wake field graph ------

// The free-wake step's velocity sums (wake_velocities, src/freewake/step.jl:166-173: V∞ + the body's sources and
// sheets + every wake blob) at np targets as one HIP graph of three kernel nodes: the body field and the blob field
// are independent nodes (they run side by side on the device) writing disjoint chunks of one partial buffer, and the
// reduction node adds V∞ and the chunks in chunk order.  The graph is built and instantiated once per node structure
// (with or without blobs); every later call rewrites the three nodes' grids and arguments in the executable graph, so
// a particle count that changes every step needs no new graph.  The body's elements are set up once; the device
// buffers grow by half again when a call needs more and are kept between calls.
WakeField::~WakeField()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *p : {(void *)el, (void *)sig, (void *)gam, (void *)X, (void *)bx, (void *)ba, (void *)bd, (void *)part,
                    (void *)u})
        if (p) (void)hipFree(p);
}

namespace {

// (re)allocate *p for at least need elements, keeping none of its contents
template <typename T> void grow(Ctx *c, T *&p, size_t &cap, size_t need)
{
    if (need <= cap && p) return;
    const size_t n = std::max<size_t>(need + need / 2, 1024);
    if (p) {
        FCHECK(hipStreamSynchronize(c->stream));
        FCHECK(hipFree(p));
        p = nullptr;
    }
    FCHECK(hipMalloc((void **)&p, n * sizeof(T)));
    cap = n;
}

}  // namespace

void WakeField::init(Ctx *ctx, const vp_surface &g, double kfar, int prec)
{
    c = ctx;
    device = ctx->device;
    stream = ctx->stream;
    nt = g.nt;
    kernel_far = kfar;
    precision = prec;
    hipStream_t st = c->stream;
    std::vector<void *> tmp;
    double *d_x = dput(g.x, (size_t)g.nn * 3, st, tmp);
    int *d_tris = dput(g.tris, (size_t)nt * 3, st, tmp);
    FCHECK(hipMalloc((void **)&el, sizeof(Elem) * nt));
    FCHECK(hipMalloc((void **)&sig, sizeof(double) * nt));
    FCHECK(hipMalloc((void **)&gam, sizeof(double) * 3 * nt));
    elem_setup<<<(nt + 255) / 256, 256, 0, st>>>(d_x, d_tris, nt, el);
    FCHECK(hipGetLastError());
    FCHECK(hipStreamSynchronize(st));
    for (void *p : tmp) FCHECK(hipFree(p));
}

void WakeField::eval(const double *sigma, const double *gamma, int np, const double *xs, int nb, const double *hbx,
                     const double *hba, const double *hbd, const double *base, double *hu)
{
    hipStream_t st = c->stream;
    const int chunk_s = chunk_of(nt, np, FS), nchunk_s = (nt + chunk_s - 1) / chunk_s;
    const int chunk_b = nb > 0 ? chunk_of(nb, np, BT) : 0, nchunk_b = nb > 0 ? (nb + chunk_b - 1) / chunk_b : 0;
    grow(c, X, capX, (size_t)np * 3);
    grow(c, u, capU, (size_t)np * 3);
    grow(c, part, capP, (size_t)(nchunk_s + nchunk_b) * np * 3);
    if (nb > 0) {
        grow(c, bx, capB3, (size_t)nb * 3);
        grow(c, ba, capA3, (size_t)nb * 3);
        grow(c, bd, capB1, (size_t)nb);
        FCHECK(hipMemcpyAsync(bx, hbx, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        FCHECK(hipMemcpyAsync(ba, hba, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        FCHECK(hipMemcpyAsync(bd, hbd, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    }
    FCHECK(hipMemcpyAsync(sig, sigma, sizeof(double) * nt, hipMemcpyHostToDevice, st));
    FCHECK(hipMemcpyAsync(gam, gamma, sizeof(double) * 3 * nt, hipMemcpyHostToDevice, st));
    FCHECK(hipMemcpyAsync(X, xs, sizeof(double) * 3 * np, hipMemcpyHostToDevice, st));

    // the nodes' arguments, in each kernel's parameter order
    const Elem *a_el = el;
    const double *a_sig = sig, *a_gam = gam, *a_X = X, *a_bx = bx, *a_ba = ba, *a_bd = bd, *a_part = part;
    double *a_pb = part + (size_t)nchunk_s * np * 3, *a_u = u, *a_ps = part;
    int a_nt = nt, a_cs = chunk_s, a_np = np, a_self = 0, a_nb = nb, a_cb = chunk_b, a_nch = nchunk_s + nchunk_b;
    int a_nfil = 0, a_acc = 0;
    double a_kf = kernel_far, b0 = base ? base[0] : 0.0, b1 = base ? base[1] : 0.0, b2 = base ? base[2] : 0.0;
    const double *a_null = nullptr;
    const int *a_inull = nullptr;
    void *src_args[] = {&a_el, &a_sig, &a_gam, &a_nt, &a_cs, &a_X, &a_np, &a_self, &a_kf, &a_ps};
    void *blob_args[] = {&a_bx, &a_ba, &a_bd, &a_nb, &a_cb, &a_X, &a_np, &a_pb};
    void *red_args[] = {&a_part, &a_nch, &a_np, &b0, &b1, &b2, &a_null, &a_inull, &a_null, &a_nfil, &a_X, &a_acc, &a_u};
    hipKernelNodeParams kp[3] = {};
    kp[0].func = precision == 1 ? reinterpret_cast<void *>(&source_field<float>)
                                : reinterpret_cast<void *>(&source_field<double>);
    kp[0].gridDim = dim3((np + FT - 1) / FT, nchunk_s);
    kp[0].blockDim = dim3(FT);
    kp[0].kernelParams = src_args;
    kp[1].func = precision == 1 ? reinterpret_cast<void *>(&blob_field<float>)
                                : reinterpret_cast<void *>(&blob_field<double>);
    kp[1].gridDim = dim3((np + FT - 1) / FT, std::max(nchunk_b, 1));
    kp[1].blockDim = dim3(FT);
    kp[1].kernelParams = blob_args;
    kp[2].func = reinterpret_cast<void *>(&field_reduce);
    kp[2].gridDim = dim3((np + 255) / 256);
    kp[2].blockDim = dim3(256);
    kp[2].kernelParams = red_args;

    const int want = nb > 0 ? 1 : 0;
    if (!graphs_enabled()) {  // the same three kernels, launched one by one
        for (int k = 0; k < 3; k++)
            if (k != 1 || want)
                FCHECK(hipLaunchKernel(kp[k].func, kp[k].gridDim, kp[k].blockDim, kp[k].kernelParams, 0, st));
        FCHECK(hipMemcpyAsync(hu, u, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, st));
        FCHECK(hipStreamSynchronize(st));
        return;
    }
    if (!exec || shape != want) {
        if (exec) FCHECK(hipGraphExecDestroy(exec));
        if (graph) FCHECK(hipGraphDestroy(graph));
        exec = nullptr;
        graph = nullptr;
        FCHECK(hipGraphCreate(&graph, 0));
        FCHECK(hipGraphAddKernelNode(&node[0], graph, nullptr, 0, &kp[0]));
        hipGraphNode_t deps[2] = {node[0], nullptr};
        int ndep = 1;
        if (want) {
            FCHECK(hipGraphAddKernelNode(&node[1], graph, nullptr, 0, &kp[1]));
            deps[ndep++] = node[1];
        }
        FCHECK(hipGraphAddKernelNode(&node[2], graph, deps, ndep, &kp[2]));
        FCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        shape = want;
        builds++;
    } else {
        FCHECK(hipGraphExecKernelNodeSetParams(exec, node[0], &kp[0]));
        if (want) FCHECK(hipGraphExecKernelNodeSetParams(exec, node[1], &kp[1]));
        FCHECK(hipGraphExecKernelNodeSetParams(exec, node[2], &kp[2]));
        updates++;
    }
    FCHECK(hipGraphLaunch(exec, st));
    launches++;
    FCHECK(hipMemcpyAsync(hu, u, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, st));
    FCHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------------------------------ test hook -----------

// The device's FP64 log (fn 0, of a), atan2 (fn 1, of y = a[i], x = b[i]) and 1/sqrt (fn 2, of a) as the kernels
// use them
__global__ void math_probe(int fn, int n, const double *__restrict__ a, const double *__restrict__ b, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (fn) {
    case 0: out[i] = dlog(a[i]); break;
    case 1: out[i] = datan2(a[i], b[i]); break;
    case 2: out[i] = vrsqrt(a[i]); break;
    case 3: out[i] = psqrt(a[i]); break;
    case 4: out[i] = fatan2((float)a[i], (float)b[i]); break;
    default: out[i] = flog1p((float)a[i]); break;
    }
}

void math_probe_host(Ctx *c, int fn, int n, const double *a, const double *b, double *out)
{
    hipStream_t st = c->stream;
    Owned o;
    o.st = st;
    double *da = dput(a, (size_t)n, st, o.v), *db = dput(fn == 1 || fn == 4 ? b : a, (size_t)n, st, o.v);
    double *dout = dnew<double>((size_t)n, o.v);
    math_probe<<<(n + 255) / 256, 256, 0, st>>>(fn, n, da, db, dout);
    FCHECK(hipGetLastError());
    FCHECK(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    FCHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------------------------------ no-allocation variant --
// vp_blob_velocities_device keeps its per-chunk partial sums in `scratch`, a fixed array that each GPU holds for the
// process.  In host code the array's name is not a device address: the address comes from hipGetSymbolAddress, taken
// once per device with that device current (at vp_create) and kept in scratch_dev; only that pointer is passed to
// kernels.  Calls on one GPU share the array, so calls that overlap on a GPU race.
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
std::mutex scratch_mu;
constexpr int MAX_DEVICES = 64;
double *scratch_dev[MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_init(int device)
{
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    if (scratch_dev[device]) return hipSuccess;
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) scratch_dev[device] = p;
    return e;
}

// The blob sums of blob_velocities_host on device arrays, with no allocation: the chunking of the allocating path
// (the same partial sums, added in the same order) unless it would not fit the scratch array, when the blobs are split
// into fewer, longer chunks.  Queued on the context's stream, not waited for.
void blob_velocities_noalloc(Ctx *c, int nb, const double *bx, const double *ba, const double *bd, int np,
                             const double *xs, int precision, double *u)
{
    double *scr = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (c->device >= 0 && c->device < MAX_DEVICES) scr = scratch_dev[c->device];
    }
    if (!scr) throw std::runtime_error("the scratch array's device address is not set (hipGetSymbolAddress failed at vp_create)");
    if ((size_t)np * 3 > NOALLOC_DOUBLES)
        throw std::length_error("np = " + std::to_string(np) + " targets exceed the no-allocation scratch array (at most " +
                                std::to_string(NOALLOC_DOUBLES / 3) + ")");
    if (np == 0) return;
    hipStream_t st = c->stream;
    int chunk = chunk_of(nb, np, BT);
    int nchunk = nb > 0 ? (nb + chunk - 1) / chunk : 0;
    const size_t cap = NOALLOC_DOUBLES / ((size_t)np * 3);
    if ((size_t)nchunk > cap) {
        const int ntiles = (nb + BT - 1) / BT, per = (int)((ntiles + cap - 1) / cap);
        chunk = per * BT;
        nchunk = (nb + chunk - 1) / chunk;
    }
    if (nchunk) {
        const dim3 grid((np + FT - 1) / FT, nchunk);
        if (precision == 1)
            blob_field<float><<<grid, FT, 0, st>>>(bx, ba, bd, nb, chunk, xs, np, scr);
        else
            blob_field<double><<<grid, FT, 0, st>>>(bx, ba, bd, nb, chunk, xs, np, scr);
        FCHECK(hipGetLastError());
    }
    field_reduce_device(st, scr, nchunk, np, nullptr, nullptr, nullptr, nullptr, 0, xs, 0, u);
}

}  // namespace vp
