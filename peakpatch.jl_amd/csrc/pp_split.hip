// pp_split.hip -- run_multitile_split's fields on the device: the coarse
// grid of the global white noise and, per tile, the Hoffman-Ribak residual,
// its isolated (zero-padded) convolution, the Catmull-Rom interpolation of
// the coarse fields and the tile-periodic 2LPT and k^2 fields.
//
// Reference: src/MultiResolution.jl:41-314 (building blocks), 596-667
// (phase 1a, the coarse fields), 719-850 (process_tile!'s phase 1b),
// 1239-1255 (_apply_kernel_inplace!); the reference's GPU path
// (ext/CUDAExt.jl:2480-2947) for the Hermitian planes of the isolated
// transform.
//
// Design.  Everything a tile needs is generated where it is used:
// * the coarse noise is a block sum over the global grid, one workgroup per
//   coarse cell, each lane walking counter pairs along rows (one Threefry
//   block gives both normals of a pair), Float64 partial sums reduced in a
//   fixed tree;
// * the residual is written straight into the zero-padded real buffer of an
//   in-place (2n)^3 rocFFT transform (rows padded to 2 (n+1) floats), two
//   cells per lane.  (Transforming only the planes that hold data, with the
//   z axis as a separate strided 1-D pass, measured slower: rocFFT's strided
//   1-D z pass costs more than its whole 3-D plan saves.);
// * each output's spectrum is one pass over the base spectrum that forms the
//   Float64 transfer factor, rounds the product into ComplexF32 once and,
//   on the kx = 0 and kx = Nyquist planes, averages the mode with its
//   conjugate partner (computed by the same lane);
// * the inverse transform runs in place; one pass then cuts the tile out of
//   the padded result, applies the Float32 1/(2n)^3 of the normalised
//   inverse and adds the coarse field's 64-tap Catmull-Rom interpolation in
//   the reference's Float32 order;
// * 2LPT: the source is accumulated in real space in the reference's order,
//   one potential at a time (two spectra and two real fields resident).
#include <cmath>
#include <vector>

#include "pp_common.h"
#include "pp_kspace.h"
#include "pp_noise.h"

struct pp_coarse {
    int device = 0;
    pp_split_geom g{};
    int N = 0, nsub = 0, block = 0, npk = 0;
    double boxsize_full = 0;
    double lk0 = 0, inv_dlk = 0;  // the P(k) table's first log k and 1 / spacing when even
    pp::DevBuf noise, delta, psi, pk;  // M^3 floats; 3 M^3 floats; log k, log P
};

namespace pp {

// ------------------------------------------------------------ coarse noise
// _downsample_noise: coarse cell (I, J, K) sums the normals of fine cells
// I b .. I b + b - 1 (and likewise in y, z), 0-based
__global__ void __launch_bounds__(256) k_coarse_noise(float *__restrict__ out, int N, int M, int b, uint64_t seed,
                                                      double norm)
{
    __shared__ double red[256];
    const long long ncoarse = (long long)M * M * M;
    const int per_row = b / 2 + 1;  // counter pairs a row of b cells can touch
    const int items = b * b * per_row;
    for (long long c = blockIdx.x; c < ncoarse; c += gridDim.x) {
        const int I = (int)(c % M), J = (int)((c / M) % M), K = (int)(c / ((long long)M * M));
        double acc = 0.0;
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int q = it % per_row, row = it / per_row;
            const int dj = row % b, dk = row / b;
            // idx - 1 of the row's first cell, and its last
            const long long s = (long long)I * b + ((long long)(J * b + dj) + (long long)(K * b + dk) * N) * N;
            const long long e = s + b - 1;
            const long long p = (s >> 1) + q;
            if (p > (e >> 1)) continue;
            double g1, g2;
            threefry_normals((uint64_t)p, seed, g1, g2);
            if (2 * p >= s) acc += g1;
            if (2 * p + 1 <= e) acc += g2;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = (float)(red[0] / norm);
        __syncthreads();
    }
}

// ------------------------------------------------------ spectrum transfer
enum XferMode { XF_AMP = 0, XF_AMP_IK = 1, XF_PHI = 2, XF_MIK = 3, XF_K2 = 4 };

struct Xfer {
    KGrid g;
    long long total;
    int mode, a, b;
    int nyq_zero;  // zero k = 0's partners on the Nyquist planes (_apply_kernel_inplace!)
    int herm;      // average the kx = 0 / kx = Nyquist planes with their conjugate partners
    double dk3, n3;  // dk^3 and n^3 of the amplitude sqrt((P(k) dk^3) n^3)
    const double *lk, *lp;
    int np;
    double lk0, inv_dlk;  // evenly spaced table: first log k and 1 / spacing (0: bisection)
};

// the mode at (ix, iy, iz) times its transfer factor, rounded to ComplexF32
__device__ __forceinline__ float2 xfer_mode(const Xfer &X, const float2 *__restrict__ src, int ix, int iy, int iz)
{
#pragma clang fp contract(off)
    const KGrid &g = X.g;
    const int fy = iy < (g.n + 1) / 2 ? iy : iy - g.n, fz = iz < (g.n + 1) / 2 ? iz : iz - g.n;
    double k[3] = {(double)ix * g.mult, (double)fy * g.mult, (double)fz * g.mult};
    const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    if (k2 == 0.0) return make_float2(0.f, 0.f);
    if (X.nyq_zero && (ix == g.nk - 1 || iy == g.n / 2 || iz == g.n / 2)) return make_float2(0.f, 0.f);
    const float2 v = src[ix + (long long)g.nk * (iy + (long long)g.n * iz)];
    switch (X.mode) {
    case XF_AMP: {
        const double amp = sqrt(pk_eval(X.lk, X.lp, X.np, sqrt(k2), X.lk0, X.inv_dlk) * X.dk3 * X.n3);
        return cmul(v, amp, 0.0);
    }
    case XF_AMP_IK: {
        const double amp = sqrt(pk_eval(X.lk, X.lp, X.np, sqrt(k2), X.lk0, X.inv_dlk) * X.dk3 * X.n3);
        return cmul(v, 0.0, amp * (k[X.a] / k2));
    }
    case XF_PHI:
        return cmul(v, -k[X.a] * k[X.b] / k2, 0.0);
    case XF_MIK:
        return cmul(v, 0.0, -k[X.a] / k2);
    default:
        return cmul(v, k2, 0.0);
    }
}

__global__ void __launch_bounds__(256) k_xfer(const float2 *__restrict__ src, float2 *__restrict__ dst, Xfer X)
{
#pragma clang fp contract(off)
    const KGrid &g = X.g;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < X.total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(e % g.nk);
        const long long r = e / g.nk;
        const int iy = (int)(r % g.n), iz = (int)(r / g.n);
        float2 v = xfer_mode(X, src, ix, iy, iz);
        if (X.herm && (ix == 0 || ((g.n & 1) == 0 && ix == g.nk - 1))) {
            const int iyc = iy == 0 ? 0 : g.n - iy, izc = iz == 0 ? 0 : g.n - iz;
            if (iyc == iy && izc == iz) {
                v.y = 0.f;
            } else {
                const float2 p = xfer_mode(X, src, ix, iyc, izc);
                v = make_float2((v.x + p.x) * 0.5f, (v.y - p.y) * 0.5f);
            }
        }
        dst[e] = v;
    }
}

// out[i] = in[i] * s (the Float32 normalisation of an inverse transform;
// in may be out)
__global__ void __launch_bounds__(256) k_scale(const float *in, float *out, long long n, float s)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        out[e] = in[e] * s;
}

// ------------------------------------------------------------- tile stages
struct TileGeo {
    int n, N, M, b;
    int i0, j0, k0;  // 0-based global index of the tile's first cell (before the wrap), >= -nbuff
    float inv_b3;    // Float32(1 / sqrt(block^3))
    uint64_t seed;
};

// v in [-m, 2m) folded into [0, m) (mod1 of the reference, 0-based): tile
// cells lie within nbuff <= N of the global grid, Catmull-Rom taps within
// two cells of the coarse one (M >= 2)
__device__ __forceinline__ int wrap1(int v, int m) { return v < 0 ? v + m : (v >= m ? v - m : v); }

// residual into the zero-padded (2n)^3 real buffer of pitch 2 (n + 1), every
// float of every row written
__global__ void __launch_bounds__(256) k_residual_pad(float *__restrict__ R, TileGeo T,
                                                      const float *__restrict__ coarse)
{
    const int n2 = 2 * T.n, pitch = n2 + 2, half = pitch / 2;
    const long long total = (long long)half * n2 * n2;  // two floats per item, padding floats included
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int x = 2 * (int)(e % half);
        const long long r = e / half;
        const int y = (int)(r % n2), z = (int)(r / n2);
        float2 v = make_float2(0.f, 0.f);
        if (x < T.n && y < T.n && z < T.n) {
            const int gy = wrap1(T.j0 + y, T.N), gz = wrap1(T.k0 + z, T.N);
            const long long rowb = ((long long)gy + (long long)gz * T.N) * T.N;
            const long long crow = (long long)T.M * (gy / T.b + (long long)T.M * (gz / T.b));
            float out2[2];
            long long last_pair = -1;
            double g1 = 0, g2 = 0;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int gx = wrap1(T.i0 + x + h, T.N);
                const long long idx0 = gx + rowb;  // idx - 1
                const long long pr = idx0 >> 1;
                if (pr != last_pair) {
                    threefry_normals((uint64_t)pr, T.seed, g1, g2);
                    last_pair = pr;
                }
                const float fine = (float)((idx0 & 1) ? g2 : g1);
                const float cm = coarse[gx / T.b + crow];
                out2[h] = fine - cm * T.inv_b3;
            }
            v = make_float2(out2[0], out2[1]);
        }
        *reinterpret_cast<float2 *>(R + x + (long long)pitch * (y + (long long)n2 * z)) = v;
    }
}

// Catmull-Rom weights of _catmull_rom, Float32
__device__ __forceinline__ void cr_weights(float t, float w[4])
{
#pragma clang fp contract(off)
    const float t2 = t * t, t3 = t2 * t;
    w[0] = (-0.5f * t3 + t2) - 0.5f * t;
    w[1] = (1.5f * t3 - 2.5f * t2) + 1.0f;
    w[2] = (-1.5f * t3 + 2.0f * t2) + 0.5f * t;
    w[3] = 0.5f * t3 - 0.5f * t2;
}

// 1-based coarse base index and fraction of 0-based global fine index g:
// f = (g + 1 - 0.5) / block + 0.5
__device__ __forceinline__ void cr_coord(int g, int b, int &i0, float &t)
{
    const double f = ((double)(g + 1) - 0.5) / (double)b + 0.5;
    const double fl = floor(f);
    i0 = (int)fl;
    t = (float)(f - fl);
}

// out = (padded inverse transform at the tile cell) * inv + Catmull-Rom of
// the coarse field.  LDS: the whole coarse field is first copied to LDS (its
// M^3 floats fit when M^3 <= kCoarseLds), so the 64 taps per cell are LDS reads.
constexpr int kCoarseLds = 16384;

template <bool LDS>
__global__ void __launch_bounds__(256) k_extract_add(const float *__restrict__ S, float inv, TileGeo T,
                                                     const float *__restrict__ coarse_g, float *__restrict__ out)
{
#pragma clang fp contract(off)
    const int n = T.n, n2 = 2 * n, pitch = n2 + 2, M = T.M;
    const long long total = (long long)n * n * n;
    extern __shared__ float coarse_s[];
    if (LDS) {
        for (int q = threadIdx.x; q < M * M * M; q += blockDim.x) coarse_s[q] = coarse_g[q];
        __syncthreads();
    }
    const float *coarse = LDS ? coarse_s : coarse_g;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(e % n);
        const long long r = e / n;
        const int y = (int)(r % n), z = (int)(r / n);
        const float self = S[x + (long long)pitch * (y + (long long)n2 * z)] * inv;
        int bx, by, bz;
        float tx, ty, tz, wx[4], wy[4], wz[4];
        cr_coord(wrap1(T.i0 + x, T.N), T.b, bx, tx);
        cr_coord(wrap1(T.j0 + y, T.N), T.b, by, ty);
        cr_coord(wrap1(T.k0 + z, T.N), T.b, bz, tz);
        cr_weights(tx, wx);
        cr_weights(ty, wy);
        cr_weights(tz, wz);
        // taps mod1(b + d, M) for d = -1..2, 0-based (b in 0..M)
        int ii[4], jj[4], kk[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            ii[a] = wrap1(bx + a - 2, M);
            jj[a] = wrap1(by + a - 2, M);
            kk[a] = wrap1(bz + a - 2, M);
        }
        float val = 0.0f;
        for (int c = 0; c < 4; c++) {
            for (int q = 0; q < 4; q++) {
                const float *row = coarse + (long long)M * (jj[q] + (long long)M * kk[c]);
#pragma unroll
                for (int a = 0; a < 4; a++) val = val + ((wx[a] * wy[q]) * wz[c]) * row[ii[a]];
            }
        }
        out[e] = self + val;
    }
}

// src = (delta * delta) * 0.5
__global__ void __launch_bounds__(256) k_src_init(const float *__restrict__ d, float *__restrict__ src, long long n)
{
#pragma clang fp contract(off)
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        src[e] = (d[e] * d[e]) * 0.5f;
}

// phi = t * inv; src = src - (phi * phi) * coef
__global__ void __launch_bounds__(256) k_src_sub(float *__restrict__ src, const float *__restrict__ t, long long n,
                                                 float inv, float coef)
{
#pragma clang fp contract(off)
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const float phi = t[e] * inv;
        src[e] = src[e] - (phi * phi) * coef;
    }
}

static int grid_for(pp_ctx *ctx, long long items)
{
    return (int)std::max<long long>(1, std::min<long long>((items + 255) / 256, (long long)ctx->num_cu * 16));
}

static int launch_xfer(pp_ctx *ctx, const void *src, void *dst, const Xfer &X)
{
    hipLaunchKernelGGL(k_xfer, dim3(grid_for(ctx, X.total)), dim3(256), 0, ctx->stream, (const float2 *)src,
                       (float2 *)dst, X);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

static Xfer make_xfer(int n, double dk, int mode, int a, int b, int nyq_zero, const pp_coarse *c)
{
    Xfer X{};
    X.g = KGrid{n, n / 2 + 1, ((double)n * dk) / (double)n};
    X.total = (long long)X.g.nk * n * n;
    X.mode = mode;
    X.a = a;
    X.b = b;
    X.nyq_zero = nyq_zero;
    X.herm = 1;
    X.dk3 = dk * dk * dk;
    X.n3 = (double)n * n * n;
    if (c) {
        X.lk = c->pk.as<double>();
        X.lp = X.lk + c->npk;
        X.np = c->npk;
        X.lk0 = c->lk0;
        X.inv_dlk = c->inv_dlk;
    }
    return X;
}

// Float32(1) / Float32(N): the scale of FFTW's normalised inverse on Float32 data
static float inv_count(long long count) { return 1.0f / (float)count; }

}  // namespace pp

using namespace pp;

extern "C" {

int pp_coarse_create(pp_ctx *ctx, const pp_split_geom *geom, const pp_pk_table *pk, pp_coarse **out)
{
    PP_ARG(ctx && geom && pk && out, "NULL argument");
    *out = nullptr;
    const pp_split_geom &G = *geom;
    PP_ARG(G.nmesh >= 4 && G.nbuff >= 0 && G.ntile >= 1 && G.M >= 2 && G.alatt > 0, "bad split geometry");
    const int nsub = G.nmesh - 2 * G.nbuff;
    PP_ARG(nsub >= 1, "nmesh %d leaves no core cells with nbuff %d", G.nmesh, G.nbuff);
    const long long N = (long long)nsub * G.ntile + 2ll * G.nbuff;
    PP_ARG(N % G.M == 0, "N=%lld must be divisible by M=%d", N, G.M);
    PP_ARG(N <= (1ll << 20) && (long long)G.M * G.M * G.M <= (1ll << 30), "global grid N=%lld too large", N);
    PP_ARG(pk->npk >= 2 && pk->k && pk->P, "P(k) table needs >= 2 rows");
    for (int q = 0; q < pk->npk; q++) {
        PP_ARG(pk->k[q] > 0 && pk->P[q] > 0, "P(k) table row %d not positive", q);
        PP_ARG(q == 0 || pk->k[q] > pk->k[q - 1], "P(k) table k not increasing at row %d", q);
    }
    PP_CHECK(ensure_device(ctx));
    pp_coarse *c = new pp_coarse();
    c->device = ctx->device;
    c->g = G;
    c->N = (int)N;
    c->nsub = nsub;
    c->block = (int)(N / G.M);
    c->boxsize_full = (double)N * G.alatt;
    c->npk = pk->npk;
    auto fail = [&](int rc) {
        pp_coarse_destroy(c);
        return rc;
    };
    const int M = G.M;
    const long long M3 = (long long)M * M * M;
    int rc = c->noise.grow(M3 * 4);
    if (rc == PP_OK) rc = c->delta.grow(M3 * 4);
    if (rc == PP_OK) rc = c->psi.grow(3 * M3 * 4);
    if (rc == PP_OK) rc = c->pk.grow(2 * pk->npk * sizeof(double));
    if (rc != PP_OK) return fail(rc);
    std::vector<double> lt(2 * pk->npk);
    for (int q = 0; q < pk->npk; q++) {
        lt[q] = log(pk->k[q]);
        lt[pk->npk + q] = log(pk->P[q]);
    }
    c->lk0 = lt[0];
    c->inv_dlk = pk_uniform_inv_spacing(lt.data(), pk->npk);
    hipError_t e = hipMemcpyAsync(c->pk.ptr, lt.data(), lt.size() * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e != hipSuccess) {
        set_error("hipMemcpyAsync: %s", hipGetErrorString(e));
        return fail(PP_ERR_HIP);
    }
    // 1. the coarse noise
    const int b = c->block;
    const double norm = sqrt((double)b * (double)b * (double)b);
    hipLaunchKernelGGL(k_coarse_noise, dim3((unsigned)std::min<long long>(M3, 1 << 20)), dim3(256), 0, ctx->stream,
                       c->noise.as<float>(), c->N, M, b, G.seed, norm);
    e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("k_coarse_noise: %s", hipGetErrorString(e));
        return fail(PP_ERR_HIP);
    }
    // 2. its periodic convolutions on the coarse grid
    FftPlans *F = nullptr;
    rc = fft_plans(ctx, M, 0, &F);
    if (rc != PP_OK) return fail(rc);
    const long long ck = (long long)(M / 2 + 1) * M * M;
    rc = ctx->d_k0.grow(ck * 8);
    if (rc == PP_OK) rc = ctx->d_k1.grow(ck * 8);
    if (rc == PP_OK) rc = ctx->d_field_a.grow(M3 * 4);
    if (rc != PP_OK) return fail(rc);
    const double dk = 2.0 * M_PI / c->boxsize_full;
    const float inv = inv_count(M3);
    auto run = [&]() -> int {
        void *ib[1] = {c->noise.ptr}, *ob[1] = {ctx->d_k0.ptr};
        PP_FFT(rocfft_execute(F->r2c, ib, ob, F->info));
        for (int q = 0; q < 4; q++) {
            const Xfer X = make_xfer(M, dk, q == 0 ? XF_AMP : XF_AMP_IK, q == 0 ? 0 : q - 1, 0, 0, c);
            PP_CHECK(launch_xfer(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, X));
            void *ib2[1] = {ctx->d_k1.ptr}, *ob2[1] = {ctx->d_field_a.ptr};
            PP_FFT(rocfft_execute(F->c2r, ib2, ob2, F->info));
            float *dst = q == 0 ? c->delta.as<float>() : c->psi.as<float>() + (q - 1) * M3;
            hipLaunchKernelGGL(k_scale, dim3(grid_for(ctx, M3)), dim3(256), 0, ctx->stream,
                               (const float *)ctx->d_field_a.ptr, dst, M3, inv);
            PP_HIP(hipGetLastError());
        }
        PP_HIP(hipStreamSynchronize(ctx->stream));
        return PP_OK;
    };
    rc = run();
    if (rc != PP_OK) return fail(rc);
    *out = c;
    return PP_OK;
}

int pp_coarse_download(const pp_coarse *c, float *noise, float *delta, float *psix, float *psiy, float *psiz)
{
    PP_ARG(c, "NULL coarse field");
    PP_HIP(hipSetDevice(c->device));
    const size_t M3 = (size_t)c->g.M * c->g.M * c->g.M, b = M3 * 4;
    if (noise) PP_HIP(hipMemcpy(noise, c->noise.ptr, b, hipMemcpyDeviceToHost));
    if (delta) PP_HIP(hipMemcpy(delta, c->delta.ptr, b, hipMemcpyDeviceToHost));
    float *p[3] = {psix, psiy, psiz};
    for (int q = 0; q < 3; q++)
        if (p[q]) PP_HIP(hipMemcpy(p[q], c->psi.as<float>() + q * M3, b, hipMemcpyDeviceToHost));
    return PP_OK;
}

int pp_coarse_destroy(pp_coarse *c)
{
    if (!c) return PP_OK;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    c->noise.release();
    c->delta.release();
    c->psi.release();
    c->pk.release();
    delete c;
    return PP_OK;
}

int pp_split_tile_fields(pp_ctx *ctx, const pp_coarse *c, int it, int jt, int kt, pp_tile_fields *f)
{
    PP_ARG(ctx && c && f && f->delta && f->etax && f->etay && f->etaz, "NULL argument");
    PP_ARG(c->device == ctx->device, "coarse field on device %d, context on device %d", c->device, ctx->device);
    const pp_split_geom &G = c->g;
    PP_ARG(it >= 1 && jt >= 1 && kt >= 1 && it <= G.ntile && jt <= G.ntile && kt <= G.ntile,
           "tile (%d, %d, %d) outside 1..%d", it, jt, kt, G.ntile);
    const bool lpt2 = f->eta2x != nullptr;
    PP_ARG(!lpt2 || (f->eta2y && f->eta2z), "eta2 outputs come as a set of three");
    PP_CHECK(ensure_device(ctx));
    const int n = G.nmesh, n2 = 2 * n;
    const long long n3 = (long long)n * n * n;
    // 1. residual -> padded buffer R; base spectrum in place
    FftPlans *Fi = nullptr;
    PP_CHECK(fft_plans(ctx, n2, 1, &Fi));
    // the two padded (2n)^3 buffers (the largest of the tile) are stream
    // ordered and returned when the tile's fields are written, so a context
    // does not keep them between tiles
    const size_t pad_floats = (size_t)(n2 + 2) * n2 * n2;
    float *R = nullptr, *S = nullptr;
    PP_HIP(hipMallocAsync((void **)&R, pad_floats * 4, ctx->stream));
    hipError_t ea = hipMallocAsync((void **)&S, pad_floats * 4, ctx->stream);
    if (ea != hipSuccess) {
        (void)hipFreeAsync(R, ctx->stream);
        set_error("hipMallocAsync of %zu B: %s", pad_floats * 4, hipGetErrorString(ea));
        return PP_ERR_HIP;
    }
    void *Wk = nullptr;
    struct Release {
        hipStream_t s;
        float *a, *b;
        void **w;
        ~Release()
        {
            (void)hipFreeAsync(a, s);
            (void)hipFreeAsync(b, s);
            if (*w) (void)hipFreeAsync(*w, s);
        }
    } release{ctx->stream, R, S, &Wk};
    if (Fi->work_need > 0) {
        PP_HIP(hipMallocAsync(&Wk, Fi->work_need, ctx->stream));
        PP_FFT(rocfft_execution_info_set_work_buffer(Fi->info, Wk, Fi->work_need));
    }
    TileGeo T;
    T.n = n;
    T.N = c->N;
    T.M = G.M;
    T.b = c->block;
    T.i0 = (it - 1) * c->nsub - G.nbuff;
    T.j0 = (jt - 1) * c->nsub - G.nbuff;
    T.k0 = (kt - 1) * c->nsub - G.nbuff;
    T.inv_b3 = (float)(1.0 / sqrt(pow((double)c->block, 3)));
    T.seed = G.seed;
    hipLaunchKernelGGL(k_residual_pad, dim3(grid_for(ctx, (long long)(n + 1) * n2 * n2)), dim3(256), 0, ctx->stream, R, T,
                       (const float *)c->noise.ptr);
    PP_HIP(hipGetLastError());
    {
        void *io[1] = {R};
        PP_FFT(rocfft_execute(Fi->r2c, io, nullptr, Fi->info));
    }
    // 2. delta and psi: transfer, inverse in place, cut + interpolate
    const double dx = (double)G.nmesh * G.alatt / n;  // boxsize_local / n
    const double dki = 2.0 * M_PI / (n2 * dx);
    const float inv_i = inv_count((long long)n2 * n2 * n2);
    float *outs[4] = {f->delta, f->etax, f->etay, f->etaz};
    for (int q = 0; q < 4; q++) {
        const Xfer X = make_xfer(n2, dki, q == 0 ? XF_AMP : XF_AMP_IK, q == 0 ? 0 : q - 1, 0, 0, c);
        PP_CHECK(launch_xfer(ctx, R, S, X));
        void *io[1] = {S};
        PP_FFT(rocfft_execute(Fi->c2r, io, nullptr, Fi->info));
        const float *cf = q == 0 ? c->delta.as<float>() : c->psi.as<float>() + (long long)(q - 1) * G.M * G.M * G.M;
        const long long M3 = (long long)G.M * G.M * G.M;
        if (M3 <= kCoarseLds)
            hipLaunchKernelGGL(k_extract_add<true>, dim3(grid_for(ctx, n3)), dim3(256), M3 * sizeof(float),
                               ctx->stream, (const float *)S, inv_i, T, cf, outs[q]);
        else
            hipLaunchKernelGGL(k_extract_add<false>, dim3(grid_for(ctx, n3)), dim3(256), 0, ctx->stream,
                               (const float *)S, inv_i, T, cf, outs[q]);
        PP_HIP(hipGetLastError());
    }
    // 3. the tile-periodic 2LPT and k^2 fields from the combined delta
    if (lpt2 || f->lapd) {
        PP_CHECK(fft_prepare(ctx, n));
        const long long ckn = (long long)(n / 2 + 1) * n * n;
        PP_CHECK(ctx->d_k0.grow(ckn * 8));
        PP_CHECK(ctx->d_k1.grow(ckn * 8));
        PP_CHECK(ctx->d_field_a.grow(2 * n3 * 4));
        float *src = ctx->d_field_a.as<float>(), *tmp = src + n3;
        const double dkt = 2.0 * M_PI / ((double)G.nmesh * G.alatt);
        const float inv_t = inv_count(n3);
        FftPlans *F = ctx->fftp;
        auto c2r = [&](void *in, float *o) -> int {
            void *ib[1] = {in}, *ob[1] = {(void *)o};
            PP_FFT(rocfft_execute(F->c2r, ib, ob, F->info));
            return PP_OK;
        };
        auto r2c = [&](const float *in, void *o) -> int {
            void *ib[1] = {(void *)in}, *ob[1] = {o};
            PP_FFT(rocfft_execute(F->r2c, ib, ob, F->info));
            return PP_OK;
        };
        auto scale = [&](float *a) -> int {
            hipLaunchKernelGGL(k_scale, dim3(grid_for(ctx, n3)), dim3(256), 0, ctx->stream, (const float *)a, a, n3,
                               inv_t);
            PP_HIP(hipGetLastError());
            return PP_OK;
        };
        PP_CHECK(r2c(f->delta, ctx->d_k0.ptr));
        if (f->lapd) {
            PP_CHECK(launch_xfer(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, make_xfer(n, dkt, XF_K2, 0, 0, 1, nullptr)));
            PP_CHECK(c2r(ctx->d_k1.ptr, f->lapd));
            PP_CHECK(scale(f->lapd));
        }
        if (lpt2) {
            hipLaunchKernelGGL(k_src_init, dim3(grid_for(ctx, n3)), dim3(256), 0, ctx->stream, (const float *)f->delta,
                               src, n3);
            PP_HIP(hipGetLastError());
            const int A[6] = {0, 1, 2, 0, 0, 1}, B[6] = {0, 1, 2, 1, 2, 2};
            for (int q = 0; q < 6; q++) {
                PP_CHECK(launch_xfer(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, make_xfer(n, dkt, XF_PHI, A[q], B[q], 1, nullptr)));
                PP_CHECK(c2r(ctx->d_k1.ptr, tmp));
                hipLaunchKernelGGL(k_src_sub, dim3(grid_for(ctx, n3)), dim3(256), 0, ctx->stream, src,
                                   (const float *)tmp, n3, inv_t, q < 3 ? 0.5f : 1.0f);
                PP_HIP(hipGetLastError());
            }
            PP_CHECK(r2c(src, ctx->d_k0.ptr));
            float *e2[3] = {f->eta2x, f->eta2y, f->eta2z};
            for (int d = 0; d < 3; d++) {
                PP_CHECK(launch_xfer(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, make_xfer(n, dkt, XF_MIK, d, 0, 1, nullptr)));
                PP_CHECK(c2r(ctx->d_k1.ptr, e2[d]));
                PP_CHECK(scale(e2[d]));
            }
        }
    }
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

}  // extern "C"
