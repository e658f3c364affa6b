// pp_filter.hip -- k-space filter-bank multiply, rocFFT transforms and the
// 27-point local-maximum peak stencil.
//
// Reference: smooth_field (src/HaloFinder/Filters.jl:79-106), find_peaks
// (src/HaloFinder/PeakFind.jl:15-78), driver loop (src/Pipeline.jl:160-195,
// src/MultiResolution.jl:885-907), GPU counterpart peak_find_tile_gpu
// (ext/CUDAExt.jl:2959-3197).
//
// Design:
//  * window multiply: one pass over the complex half-spectrum, |k| and W in
//    Float64 (the top-hat cancels badly at small kR in Float32), product in
//    Float32 as the reference.  The 1/n^3 of the normalised irfft is folded
//    into the same pass (a power of two for power-of-two n, so bit-identical
//    to scaling after the transform).
//  * stencil: the inner cube is linearised k-outer / j / i-inner, exactly the
//    CPU loop order; pass 1 counts candidates per 256-cell block (wave ballot),
//    an exclusive scan turns counts into offsets, pass 2 recomputes the
//    predicate and writes each peak at offset + rank.  The output order is
//    therefore the CPU order, deterministic, with no atomics.  The cap of
//    max_peaks per filter (PeakFind.jl:70-72) drops the tail of that order and
//    masks only the kept peaks, as the CPU early return does.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "pp_common.h"

namespace pp {

constexpr int kStencilBlock = 256;

// ------------------------------------------------------------- window
__global__ void __launch_bounds__(256) k_window(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                int n, int nk, long long total, double dk, double Rf,
                                                int wsmooth, float scale)
{
#pragma clang fp contract(off)
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        int ix = (int)(idx % nk);
        long long rem = idx / nk;
        int iy = (int)(rem % n);
        int iz = (int)(rem / n);
        int fy = iy < (n + 1) / 2 ? iy : iy - n;
        int fz = iz < (n + 1) / 2 ? iz : iz - n;
        double kx = dk * (double)ix, ky = dk * (double)fy, kz = dk * (double)fz;
        double k = sqrt(kx * kx + ky * ky + kz * kz);
        double fkR = k * Rf;
        float2 v = in[idx];
        if (fkR != 0.0) {
            double w;
            if (wsmooth == PP_WINDOW_GAUSSIAN) {
                double t = fkR / 2.0;
                w = exp(-(t * t) / 2.0);
            } else if (wsmooth == PP_WINDOW_K2GAUSS) {
                double t = fkR / 2.0;
                w = exp(-(t * t) / 2.0) * (k * k);
            } else {
                w = 3.0 * (sin(fkR) - fkR * cos(fkR)) / (fkR * fkR * fkR);
            }
            float wf = (float)w;
            v.x = v.x * wf;
            v.y = v.y * wf;
        }
        v.x = v.x * scale;
        v.y = v.y * scale;
        out[idx] = v;
    }
}

static int launch_window(pp_ctx *ctx, const void *in, void *out, int n, double boxsize, double Rf,
                         int wsmooth, float scale)
{
    int nk = n / 2 + 1;
    long long total = (long long)nk * n * n;
    double dk = 2.0 * M_PI / boxsize;
    long long blocks = (total + 255) / 256;
    int grid = (int)std::min<long long>(blocks, (long long)ctx->num_cu * 32);
    hipLaunchKernelGGL(k_window, dim3(grid), dim3(256), 0, ctx->stream, (const float2 *)in,
                       (float2 *)out, n, nk, total, dk, Rf, wsmooth, scale);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// ------------------------------------------------------------- stencil
struct StencilGeom {
    int n1, n2, n3, nbuff;
    int in1, in2, in3;
    long long total;
};

__device__ __forceinline__ bool stencil_pred(const float *__restrict__ d, const int8_t *__restrict__ mask,
                                             const StencilGeom &g, long long li, float fcrit,
                                             int &i, int &j, int &k, float &ff)
{
    i = (int)(li % g.in1) + g.nbuff;
    long long rem = li / g.in1;
    j = (int)(rem % g.in2) + g.nbuff;
    k = (int)(rem / g.in2) + g.nbuff;
    const long long s1 = g.n1, s12 = (long long)g.n1 * g.n2;
    const long long c = i + j * s1 + k * s12;
    if (mask[c] == 1) return false;
    ff = d[c];
    if (ff < fcrit) return false;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const float *row = d + c + dy * s1 + dz * s12;
#pragma unroll
            for (int dx = -1; dx <= 1; dx++)
                if (ff < row[dx]) return false;
        }
    return true;
}

__global__ void __launch_bounds__(kStencilBlock) k_stencil_count(const float *__restrict__ d,
                                                                 const int8_t *__restrict__ mask,
                                                                 StencilGeom g, float fcrit,
                                                                 int *__restrict__ block_count)
{
    long long li = (long long)blockIdx.x * kStencilBlock + threadIdx.x;
    bool f = false;
    if (li < g.total) {
        int i, j, k;
        float ff;
        f = stencil_pred(d, mask, g, li, fcrit, i, j, k, ff);
    }
    unsigned long long b = __ballot(f);
    __shared__ int wc[kStencilBlock / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wc[wid] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
#pragma unroll
        for (int w = 0; w < kStencilBlock / 64; w++) s += wc[w];
        block_count[blockIdx.x] = s;
    }
}

struct PeakOut {
    int32_t *i, *j, *k, *filt;
    float *delta_s;
    long long capacity;
};

// counters[0] = peaks written so far (all filters), counters[1] = peaks found
// (capped per filter), counters[2] = first index of the current filter.
__global__ void __launch_bounds__(kStencilBlock) k_stencil_write(const float *__restrict__ d, int8_t *__restrict__ mask,
                                                                 StencilGeom g, float fcrit,
                                                                 const int *__restrict__ block_off,
                                                                 const long long *__restrict__ counters,
                                                                 long long max_peaks, int filt, PeakOut o)
{
    long long li = (long long)blockIdx.x * kStencilBlock + threadIdx.x;
    bool f = false;
    int i = 0, j = 0, k = 0;
    float ff = 0.f;
    if (li < g.total) f = stencil_pred(d, mask, g, li, fcrit, i, j, k, ff);
    unsigned long long b = __ballot(f);
    __shared__ int wc[kStencilBlock / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wc[wid] = __popcll(b);
    __syncthreads();
    if (!f) return;
    int before = 0;
    for (int w = 0; w < wid; w++) before += wc[w];
    unsigned long long lt = lane == 0 ? 0ull : (b & (~0ull >> (64 - lane)));
    long long rank = (long long)block_off[blockIdx.x] + before + __popcll(lt);
    if (rank >= max_peaks) return;
    const long long c = i + (long long)j * g.n1 + (long long)k * g.n1 * g.n2;
    mask[c] = 1;
    long long slot = counters[0] + rank;
    if (slot >= o.capacity) return;
    o.i[slot] = i;
    o.j[slot] = j;
    o.k[slot] = k;
    o.filt[slot] = filt;
    o.delta_s[slot] = ff;
}

__global__ void k_stencil_finish(const int *__restrict__ block_off, const int *__restrict__ block_count,
                                 int nblocks, long long max_peaks, long long *__restrict__ counters)
{
    long long found = nblocks > 0 ? (long long)block_off[nblocks - 1] + block_count[nblocks - 1] : 0;
    long long kept = found < max_peaks ? found : max_peaks;
    counters[2] = counters[0];
    counters[0] += kept;
    counters[1] += kept;
}

__global__ void k_sample_field(const float *__restrict__ f, int n1, int n2, const int32_t *__restrict__ pi,
                               const int32_t *__restrict__ pj, const int32_t *__restrict__ pk,
                               float *__restrict__ out, const long long *__restrict__ counters,
                               long long capacity)
{
    long long lo = counters[2], hi = counters[0];
    if (hi > capacity) hi = capacity;
    for (long long s = lo + (long long)blockIdx.x * blockDim.x + threadIdx.x; s < hi;
         s += (long long)gridDim.x * blockDim.x)
        out[s] = f[pi[s] + (long long)pj[s] * n1 + (long long)pk[s] * n1 * n2];
}

// Stencil over one field; peaks appended after counters[0].
static int run_stencil(pp_ctx *ctx, const float *d_field, int8_t *d_mask, const StencilGeom &g,
                       float fcrit, long long max_peaks, int filt, const PeakOut &o,
                       long long *d_counters)
{
    long long nblocks_ll = (g.total + kStencilBlock - 1) / kStencilBlock;
    PP_ARG(nblocks_ll < (1ll << 31), "inner cube too large for the stencil grid");
    int nblocks = (int)nblocks_ll;
    if (nblocks == 0) {
        hipLaunchKernelGGL(k_stencil_finish, dim3(1), dim3(1), 0, ctx->stream, (const int *)nullptr,
                           (const int *)nullptr, 0, max_peaks, d_counters);
        PP_HIP(hipGetLastError());
        return PP_OK;
    }
    PP_CHECK(ctx->d_scan.grow(2 * (size_t)nblocks * sizeof(int)));
    int *d_cnt = ctx->d_scan.as<int>();
    int *d_off = d_cnt + nblocks;
    hipLaunchKernelGGL(k_stencil_count, dim3(nblocks), dim3(kStencilBlock), 0, ctx->stream, d_field,
                       (const int8_t *)d_mask, g, fcrit, d_cnt);
    PP_HIP(hipGetLastError());
    size_t tmp = 0;
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_cnt, d_off, nblocks, ctx->stream));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_cub.ptr, tmp, d_cnt, d_off, nblocks, ctx->stream));
    hipLaunchKernelGGL(k_stencil_write, dim3(nblocks), dim3(kStencilBlock), 0, ctx->stream, d_field,
                       d_mask, g, fcrit, (const int *)d_off, (const long long *)d_counters, max_peaks,
                       filt, o);
    PP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_stencil_finish, dim3(1), dim3(1), 0, ctx->stream, (const int *)d_off,
                       (const int *)d_cnt, nblocks, max_peaks, d_counters);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

static int alloc_peak_out(pp_ctx *ctx, long long cap, PeakOut &o, float **lapd, long long **counters)
{
    size_t per = cap > 0 ? (size_t)cap : 1;
    size_t bytes = per * (4 * sizeof(int32_t) + 2 * sizeof(float)) + 4 * sizeof(long long) + 256;
    PP_CHECK(ctx->d_peaks.grow(bytes));
    char *p = ctx->d_peaks.as<char>();
    *counters = (long long *)p;
    p += 256;
    o.i = (int32_t *)p; p += per * sizeof(int32_t);
    o.j = (int32_t *)p; p += per * sizeof(int32_t);
    o.k = (int32_t *)p; p += per * sizeof(int32_t);
    o.filt = (int32_t *)p; p += per * sizeof(int32_t);
    o.delta_s = (float *)p; p += per * sizeof(float);
    *lapd = (float *)p;
    o.capacity = cap;
    PP_HIP(hipMemsetAsync(*counters, 0, 4 * sizeof(long long), ctx->stream));
    return PP_OK;
}

static int download_peaks(pp_ctx *ctx, const PeakOut &o, const float *d_lapd, const long long *d_counters,
                          pp_peaks *out)
{
    long long cnt[4];
    PP_HIP(hipMemcpyAsync(cnt, d_counters, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    long long found = cnt[1];
    long long written = found < out->capacity ? found : out->capacity;
    out->needed = found;
    out->count = written;
    if (written > 0) {
        size_t b4 = (size_t)written * 4;
        if (out->i) PP_HIP(hipMemcpyAsync(out->i, o.i, b4, hipMemcpyDeviceToHost, ctx->stream));
        if (out->j) PP_HIP(hipMemcpyAsync(out->j, o.j, b4, hipMemcpyDeviceToHost, ctx->stream));
        if (out->k) PP_HIP(hipMemcpyAsync(out->k, o.k, b4, hipMemcpyDeviceToHost, ctx->stream));
        if (out->filt) PP_HIP(hipMemcpyAsync(out->filt, o.filt, b4, hipMemcpyDeviceToHost, ctx->stream));
        if (out->delta_s)
            PP_HIP(hipMemcpyAsync(out->delta_s, o.delta_s, b4, hipMemcpyDeviceToHost, ctx->stream));
        if (out->lapd_s && d_lapd)
            PP_HIP(hipMemcpyAsync(out->lapd_s, d_lapd, b4, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (found > out->capacity) {
        set_error("peak output capacity %lld too small: %lld peaks found", (long long)out->capacity, found);
        return PP_ERR_CAPACITY;
    }
    return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_window_multiply(pp_ctx *ctx, const void *delta_k, void *out, int on_device, int n, double boxsize,
                       double Rf, int wsmooth, float scale)
{
    PP_ARG(ctx && delta_k && out, "NULL argument");
    PP_ARG(n >= 2 && boxsize > 0, "bad n=%d / boxsize", n);
    PP_ARG(wsmooth == 0 || wsmooth == 1 || wsmooth == 3, "wsmooth %d not in {0,1,3}", wsmooth);
    PP_CHECK(ensure_device(ctx));
    size_t bytes = (size_t)(n / 2 + 1) * n * n * sizeof(float2);
    if (on_device) return launch_window(ctx, delta_k, out, n, boxsize, Rf, wsmooth, scale);
    PP_CHECK(ctx->d_k1.grow(bytes));
    PP_HIP(hipMemcpyAsync(ctx->d_k1.ptr, delta_k, bytes, hipMemcpyHostToDevice, ctx->stream));
    PP_CHECK(launch_window(ctx, ctx->d_k1.ptr, ctx->d_k1.ptr, n, boxsize, Rf, wsmooth, scale));
    PP_HIP(hipMemcpyAsync(out, ctx->d_k1.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

int pp_rfft(pp_ctx *ctx, const float *in, void *out, int on_device, int n)
{
    PP_ARG(ctx && in && out, "NULL argument");
    PP_ARG(n >= 2, "bad n=%d", n);
    PP_CHECK(ensure_device(ctx));
    PP_CHECK(fft_prepare(ctx, n));
    size_t rb = (size_t)n * n * n * sizeof(float), cb = (size_t)(n / 2 + 1) * n * n * sizeof(float2);
    void *din = (void *)in, *dout = out;
    if (!on_device) {
        PP_CHECK(ctx->d_field_a.grow(rb));
        PP_CHECK(ctx->d_k0.grow(cb));
        PP_HIP(hipMemcpyAsync(ctx->d_field_a.ptr, in, rb, hipMemcpyHostToDevice, ctx->stream));
        din = ctx->d_field_a.ptr;
        dout = ctx->d_k0.ptr;
    }
    void *ib[1] = {din}, *ob[1] = {dout};
    PP_FFT(rocfft_execute(ctx->fft.r2c, ib, ob, ctx->fft.info));
    if (!on_device) PP_HIP(hipMemcpyAsync(out, dout, cb, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

int pp_smooth_field(pp_ctx *ctx, const void *delta_k, float *out, int on_device, int n, double boxsize,
                    double Rf, int wsmooth)
{
    PP_ARG(ctx && delta_k && out, "NULL argument");
    PP_ARG(n >= 2 && boxsize > 0, "bad n=%d / boxsize", n);
    PP_ARG(wsmooth == 0 || wsmooth == 1 || wsmooth == 3, "wsmooth %d not in {0,1,3}", wsmooth);
    PP_CHECK(ensure_device(ctx));
    PP_CHECK(fft_prepare(ctx, n));
    size_t rb = (size_t)n * n * n * sizeof(float), cb = (size_t)(n / 2 + 1) * n * n * sizeof(float2);
    PP_CHECK(ctx->d_k1.grow(cb));
    const void *src = delta_k;
    if (!on_device) {
        PP_CHECK(ctx->d_k0.grow(cb));
        PP_HIP(hipMemcpyAsync(ctx->d_k0.ptr, delta_k, cb, hipMemcpyHostToDevice, ctx->stream));
        src = ctx->d_k0.ptr;
    }
    float inv = (float)(1.0 / ((double)n * n * n));
    PP_CHECK(launch_window(ctx, src, ctx->d_k1.ptr, n, boxsize, Rf, wsmooth, inv));
    float *dst = out;
    if (!on_device) {
        PP_CHECK(ctx->d_field_a.grow(rb));
        dst = ctx->d_field_a.as<float>();
    }
    void *ib[1] = {ctx->d_k1.ptr}, *ob[1] = {dst};
    PP_FFT(rocfft_execute(ctx->fft.c2r, ib, ob, ctx->fft.info));
    if (!on_device) PP_HIP(hipMemcpyAsync(out, dst, rb, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

int pp_find_peaks_field(pp_ctx *ctx, const float *delta_s, int8_t *mask, int on_device, int n1, int n2,
                        int n3, int nbuff, float fcrit, int64_t max_peaks, pp_peaks *out)
{
    PP_ARG(ctx && delta_s && mask && out, "NULL argument");
    PP_ARG(n1 >= 3 && n2 >= 3 && n3 >= 3, "grid %dx%dx%d too small", n1, n2, n3);
    PP_ARG(nbuff >= 1, "nbuff must be >= 1 (the 27-point stencil reads one cell past the inner cube)");
    PP_ARG(out->capacity >= 0, "negative capacity");
    PP_CHECK(ensure_device(ctx));
    if (max_peaks <= 0) max_peaks = 2000000;
    StencilGeom g;
    g.n1 = n1; g.n2 = n2; g.n3 = n3; g.nbuff = nbuff;
    g.in1 = std::max(0, n1 - 2 * nbuff);
    g.in2 = std::max(0, n2 - 2 * nbuff);
    g.in3 = std::max(0, n3 - 2 * nbuff);
    g.total = (long long)g.in1 * g.in2 * g.in3;
    size_t N = (size_t)n1 * n2 * n3;
    const float *d_f = delta_s;
    int8_t *d_m = mask;
    if (!on_device) {
        PP_CHECK(ctx->d_field_a.grow(N * sizeof(float)));
        PP_CHECK(ctx->d_mask.grow(N));
        PP_HIP(hipMemcpyAsync(ctx->d_field_a.ptr, delta_s, N * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(ctx->d_mask.ptr, mask, N, hipMemcpyHostToDevice, ctx->stream));
        d_f = ctx->d_field_a.as<float>();
        d_m = ctx->d_mask.as<int8_t>();
    }
    PeakOut o;
    float *d_lapd;
    long long *d_counters;
    PP_CHECK(alloc_peak_out(ctx, out->capacity, o, &d_lapd, &d_counters));
    PP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    PP_CHECK(run_stencil(ctx, d_f, d_m, g, fcrit, max_peaks, 0, o, d_counters));
    PP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    if (!on_device) PP_HIP(hipMemcpyAsync(mask, d_m, N, hipMemcpyDeviceToHost, ctx->stream));
    int rc = download_peaks(ctx, o, nullptr, d_counters, out);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->timing.peakfind_ms = ms;
    return rc;
}

int pp_find_peaks(pp_ctx *ctx, const float *delta, int on_device, int n, double boxsize, const double *Rf,
                  const float *fcrit, int nfilt, int wsmooth, int want_lapd, int nbuff,
                  int64_t max_peaks_per_filter, int8_t *mask_inout, pp_peaks *out)
{
    PP_ARG(ctx && delta && Rf && fcrit && out, "NULL argument");
    PP_ARG(n >= 3 && boxsize > 0 && nfilt >= 0, "bad n=%d / boxsize / nfilt", n);
    PP_ARG(nbuff >= 1, "nbuff must be >= 1");
    PP_ARG(wsmooth == 0 || wsmooth == 1 || wsmooth == 3, "wsmooth %d not in {0,1,3}", wsmooth);
    PP_ARG(out->capacity >= 0, "negative capacity");
    PP_CHECK(ensure_device(ctx));
    PP_CHECK(fft_prepare(ctx, n));
    if (max_peaks_per_filter <= 0) max_peaks_per_filter = 2000000;
    size_t N = (size_t)n * n * n;
    size_t rb = N * sizeof(float), cb = (size_t)(n / 2 + 1) * n * n * sizeof(float2);
    PP_CHECK(ctx->d_k0.grow(cb));
    PP_CHECK(ctx->d_k1.grow(cb));
    PP_CHECK(ctx->d_field_a.grow(rb));
    if (want_lapd) PP_CHECK(ctx->d_field_b.grow(rb));
    PP_CHECK(ctx->d_mask.grow(N));
    PP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    const float *d_delta = delta;
    if (!on_device) {
        PP_HIP(hipMemcpyAsync(ctx->d_field_a.ptr, delta, rb, hipMemcpyHostToDevice, ctx->stream));
        d_delta = ctx->d_field_a.as<float>();
    }
    {
        void *ib[1] = {(void *)d_delta}, *ob[1] = {ctx->d_k0.ptr};
        PP_FFT(rocfft_execute(ctx->fft.r2c, ib, ob, ctx->fft.info));
    }
    int8_t *d_mask = ctx->d_mask.as<int8_t>();
    if (mask_inout)
        PP_HIP(hipMemcpyAsync(d_mask, mask_inout, N, hipMemcpyHostToDevice, ctx->stream));
    else
        PP_HIP(hipMemsetAsync(d_mask, 0, N, ctx->stream));
    PeakOut o;
    float *d_lapd;
    long long *d_counters;
    PP_CHECK(alloc_peak_out(ctx, out->capacity, o, &d_lapd, &d_counters));
    StencilGeom g;
    g.n1 = g.n2 = g.n3 = n;
    g.nbuff = nbuff;
    g.in1 = g.in2 = g.in3 = std::max(0, n - 2 * nbuff);
    g.total = (long long)g.in1 * g.in2 * g.in3;
    const float inv = (float)(1.0 / ((double)n * n * n));
    for (int f = 0; f < nfilt; f++) {
        PP_CHECK(launch_window(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, n, boxsize, Rf[f], wsmooth, inv));
        void *ib[1] = {ctx->d_k1.ptr}, *ob[1] = {ctx->d_field_a.ptr};
        PP_FFT(rocfft_execute(ctx->fft.c2r, ib, ob, ctx->fft.info));
        PP_CHECK(run_stencil(ctx, ctx->d_field_a.as<float>(), d_mask, g, fcrit[f], max_peaks_per_filter, f,
                             o, d_counters));
        if (want_lapd) {
            PP_CHECK(launch_window(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, n, boxsize, Rf[f], PP_WINDOW_K2GAUSS, inv));
            void *ib2[1] = {ctx->d_k1.ptr}, *ob2[1] = {ctx->d_field_b.ptr};
            PP_FFT(rocfft_execute(ctx->fft.c2r, ib2, ob2, ctx->fft.info));
            hipLaunchKernelGGL(k_sample_field, dim3(ctx->num_cu * 4), dim3(256), 0, ctx->stream,
                               ctx->d_field_b.as<float>(), n, n, (const int32_t *)o.i, (const int32_t *)o.j,
                               (const int32_t *)o.k, d_lapd, (const long long *)d_counters, o.capacity);
            PP_HIP(hipGetLastError());
        }
    }
    PP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    if (mask_inout) PP_HIP(hipMemcpyAsync(mask_inout, d_mask, N, hipMemcpyDeviceToHost, ctx->stream));
    int rc = download_peaks(ctx, o, want_lapd ? d_lapd : nullptr, d_counters, out);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->timing.peakfind_ms = ms;
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------ input noise
// fill_noise_threefry! (src/InitialConditions/RandomField.jl:269-306): cell
// idx (1-based, x fastest) takes the first normal of counter pair
// (idx-1) >> 1 when idx-1 is even, the second otherwise; each pair is
// Threefry-2x64-20 of counter (pair, 0) under key (seed, 0) (Salmon et al.
// 2011), two 53-bit uniforms in (0, 1] and Box-Muller in Float64.
namespace pp {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

__global__ void k_noise_threefry(float *__restrict__ out, int64_t n3, uint64_t seed)
{
    constexpr int ROT[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    const uint64_t ks[3] = {seed, 0ull, 0x1BD11BDAA9FC1A22ull ^ seed};
    const int64_t npairs = (n3 + 1) / 2;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += (int64_t)gridDim.x * blockDim.x) {
        uint64_t x0 = (uint64_t)p + ks[0], x1 = ks[1];
#pragma unroll
        for (int blk = 0; blk < 5; blk++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                x0 += x1;
                x1 = rotl64(x1, ROT[(4 * blk + r) & 7]);
                x1 ^= x0;
            }
            const int inj = blk + 1;
            x0 += ks[inj % 3];
            x1 += ks[(inj + 1) % 3] + (uint64_t)inj;
        }
        const double u1 = ((double)(x0 >> 11) + 1.0) * 1.1102230246251565e-16;
        const double u2 = ((double)(x1 >> 11) + 1.0) * 1.1102230246251565e-16;
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586 * u2;
        out[2 * p] = (float)(r * cos(th));
        if (2 * p + 1 < n3) out[2 * p + 1] = (float)(r * sin(th));
    }
}

}  // namespace pp

extern "C" int pp_fill_noise_threefry(pp_ctx *ctx, float *out, int64_t ncells, uint64_t seed)
{
    PP_ARG(ctx && out, "NULL argument");
    PP_ARG(ncells >= 0, "negative cell count");
    PP_CHECK(pp::ensure_device(ctx));
    if (ncells == 0) return PP_OK;
    const int64_t npairs = (ncells + 1) / 2;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((npairs + 255) / 256, (int64_t)ctx->num_cu * 16));
    hipLaunchKernelGGL(pp::k_noise_threefry, dim3(grid), dim3(256), 0, ctx->stream, out, ncells, seed);
    PP_HIP(hipGetLastError());
    return PP_OK;
}
