// pp_fields.hip -- the input fields of one tile on the device: Gaussian
// random field, 1LPT and 2LPT displacements and the k^2 ("Laplacian") field.
//
// Reference: run_tile's field generation (src/Pipeline.jl:110-206):
// generate_grf / _convolve_noise! (src/InitialConditions/RandomField.jl:
// 123-175) with fill_noise_threefry! (:46-81) and load_pk (src/
// InitialConditions/PowerSpectrum.jl:11-23); displacements_1lpt /
// displacements_2lpt (src/InitialConditions/LPT.jl:21-191);
// _compute_laplacian (src/Pipeline.jl:35-51).
//
// Design: rocFFT for the transforms (the context's cached plans); every
// k-space step is one pass over the half spectrum that evaluates its Float64
// factor per mode and rounds the product into Float32 once, as the
// reference's Float64 factor times its ComplexF32 array does.  The 1/n^3 of
// the normalised inverse transforms is folded into that factor (an exact
// power of two for power-of-two n).  The 2LPT source is accumulated in real
// space in the reference's order, so only the diagonal potentials and the
// running sum are held at once.
#include <cmath>

#include "pp_common.h"
#include "pp_kspace.h"

namespace pp {

// noise_k * sqrt(P(k) dk^3 n^3) / n^3, k = 0 zeroed (RandomField.jl:144-167)
__global__ void __launch_bounds__(256) k_grf_amp(float2 *__restrict__ a, KGrid g, long long total, double dk3,
                                                 double n3, double inv_n3, const double *__restrict__ lk,
                                                 const double *__restrict__ lp, int np, double lk0, double inv_dlk)
{
#pragma clang fp contract(off)
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        double kx, ky, kz;
        kvec(g, e, ix, iy, iz, kx, ky, kz);
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == 0.0) {
            a[e] = make_float2(0.f, 0.f);
            continue;
        }
        const double amp = sqrt(pk_eval(lk, lp, np, sqrt(k2), lk0, inv_dlk) * dk3 * n3);
        a[e] = cmul(a[e], amp * inv_n3, 0.0);
    }
}

// one spectrum from delta_k: which = 0..2: psi_i = i k_i / k^2 (1LPT,
// Nyquist planes zeroed); 3..8: phi_ab = -k_a k_b / k^2 (ab = 11, 22, 33,
// 12, 13, 23); 9..11: psi2_i = -i k_i / k^2 (2LPT, Nyquist zeroed); all
// times scale
__global__ void __launch_bounds__(256) k_lpt(const float2 *__restrict__ in, float2 *__restrict__ out, KGrid g,
                                             long long total, int which, double scale)
{
#pragma clang fp contract(off)
    const int A[6] = {0, 1, 2, 0, 0, 1}, B[6] = {0, 1, 2, 1, 2, 2};
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        double k[3];
        kvec(g, e, ix, iy, iz, k[0], k[1], k[2]);
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const bool nyq = ix == g.nk - 1 || iy == g.n / 2 || iz == g.n / 2;
        float2 r = make_float2(0.f, 0.f);
        if (which < 3) {
            if (k2 != 0.0 && !nyq) r = cmul(in[e], 0.0, k[which] / k2 * scale);
        } else if (which < 9) {
            const int q = which - 3;
            if (k2 != 0.0) r = cmul(in[e], -k[A[q]] * k[B[q]] / k2 * scale, 0.0);
        } else {
            if (k2 != 0.0 && !nyq) r = cmul(in[e], 0.0, -k[which - 9] / k2 * scale);
        }
        out[e] = r;
    }
}

// _compute_laplacian: k^2 delta_k with k and the product in Float32,
// then the two normalisations (exact powers of two for power-of-two n)
__global__ void __launch_bounds__(256) k_lapk(const float2 *__restrict__ in, float2 *__restrict__ out, int n, int nk,
                                              long long total, float dkf, float s1, float s2)
{
#pragma clang fp contract(off)
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(e % nk);
        const long long r = e / nk;
        const int iy = (int)(r % n), iz = (int)(r / n);
        const int nyq = n / 2;
        const float kx = (float)ix * dkf;
        const float ky = (float)(iy <= nyq ? iy : iy - n) * dkf;
        const float kz = (float)(iz <= nyq ? iz : iz - n) * dkf;
        const float k2 = kx * kx + ky * ky + kz * kz;
        const float2 v = in[e];
        out[e] = make_float2(v.x * k2 * s1 * s2, v.y * k2 * s1 * s2);
    }
}

// the 2LPT source, left to right in Float32: step 0: src = p11 p22 - p12^2;
// 1: src = src + p11 p33 - p13^2 (b = p33, c = p13); 2: src = src + p22 p33 -
// p23^2 (a = p22, b = p33, c = p23)
__global__ void __launch_bounds__(256) k_src2(float *__restrict__ src, const float *__restrict__ a,
                                              const float *__restrict__ b, const float *__restrict__ c, long long total,
                                              int step)
{
#pragma clang fp contract(off)
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const float cc = c[e] * c[e];
        if (step == 0) src[e] = a[e] * b[e] - cc;
        else src[e] = (src[e] + a[e] * b[e]) - cc;
    }
}

int launch_noise(pp_ctx *ctx, float *out, int64_t ncells, uint64_t seed);  // pp_filter.hip

}  // namespace pp

using namespace pp;

extern "C" int pp_make_tile_fields(pp_ctx *ctx, int n, double boxsize, uint64_t seed, const pp_pk_table *pk,
                                   pp_tile_fields *f)
{
    PP_ARG(ctx && pk && f && f->delta && f->etax && f->etay && f->etaz, "NULL argument");
    PP_ARG(n >= 4 && boxsize > 0, "bad n=%d / boxsize", n);
    PP_ARG(pk->npk >= 2 && pk->k && pk->P, "P(k) table needs >= 2 rows");
    const bool lpt2 = f->eta2x != nullptr;
    PP_ARG(!lpt2 || (f->eta2y && f->eta2z), "eta2 outputs come as a set of three");
    for (int q = 0; q < pk->npk; q++) {
        PP_ARG(pk->k[q] > 0 && pk->P[q] > 0, "P(k) table row %d not positive", q);
        PP_ARG(q == 0 || pk->k[q] > pk->k[q - 1], "P(k) table k not increasing at row %d", q);
    }
    PP_CHECK(ensure_device(ctx));
    PP_CHECK(fft_prepare(ctx, n));
    const int nk = n / 2 + 1;
    const long long total = (long long)nk * n * n;
    const size_t N = (size_t)n * n * n, cb = (size_t)total * sizeof(float2), rb = N * sizeof(float);
    // device buffers: the log P(k) table, delta_k, one work spectrum, and
    // with 2LPT four real fields (the diagonal potentials and the source)
    std::vector<double> lk(pk->npk), lp(pk->npk);
    for (int q = 0; q < pk->npk; q++) {
        lk[q] = log(pk->k[q]);
        lp[q] = log(pk->P[q]);
    }
    PP_CHECK(ctx->d_misc.grow(2 * pk->npk * sizeof(double)));
    double *d_lk = ctx->d_misc.as<double>(), *d_lp = d_lk + pk->npk;
    PP_HIP(hipMemcpyAsync(d_lk, lk.data(), pk->npk * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipMemcpyAsync(d_lp, lp.data(), pk->npk * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PP_CHECK(ctx->d_k0.grow(cb));
    PP_CHECK(ctx->d_k1.grow(cb));
    float2 *dk0 = ctx->d_k0.as<float2>(), *dk1 = ctx->d_k1.as<float2>();
    const int grid = (int)std::min<long long>((total + 255) / 256, (long long)ctx->num_cu * 16);
    const double dk = 2.0 * M_PI / boxsize;
    KGrid g{n, nk, ((double)n * dk) / (double)n};
    const double inv_n3 = 1.0 / ((double)n * n * n);
    auto c2r = [&](void *in, float *out) -> int {
        void *ib[1] = {in}, *ob[1] = {(void *)out};
        PP_FFT(rocfft_execute(ctx->fftp->c2r, ib, ob, ctx->fftp->info));
        return PP_OK;
    };
    auto r2c = [&](const float *in, void *out) -> int {
        void *ib[1] = {(void *)in}, *ob[1] = {out};
        PP_FFT(rocfft_execute(ctx->fftp->r2c, ib, ob, ctx->fftp->info));
        return PP_OK;
    };
    // 1. delta: noise -> rfft -> sqrt(P dk^3 n^3) -> normalised irfft
    PP_CHECK(launch_noise(ctx, f->delta, (int64_t)N, seed));
    PP_CHECK(r2c(f->delta, dk0));
    hipLaunchKernelGGL(k_grf_amp, dim3(grid), dim3(256), 0, ctx->stream, dk0, g, total, dk * dk * dk,
                       (double)n * n * n, inv_n3, (const double *)d_lk, (const double *)d_lp, pk->npk, lk[0],
                       pk_uniform_inv_spacing(lk.data(), pk->npk));
    PP_HIP(hipGetLastError());
    PP_CHECK(c2r(dk0, f->delta));
    // 2. delta_k = rfft(delta), kept in dk0
    PP_CHECK(r2c(f->delta, dk0));
    auto lpt = [&](int which, float *out) -> int {
        hipLaunchKernelGGL(k_lpt, dim3(grid), dim3(256), 0, ctx->stream, (const float2 *)dk0, dk1, g, total, which,
                           inv_n3);
        PP_HIP(hipGetLastError());
        return c2r(dk1, out);
    };
    // 3. 1LPT
    PP_CHECK(lpt(0, f->etax));
    PP_CHECK(lpt(1, f->etay));
    PP_CHECK(lpt(2, f->etaz));
    // 4. Laplacian (before 2LPT reuses delta_k's buffer)
    if (f->lapd) {
        const float s = (float)inv_n3;
        hipLaunchKernelGGL(k_lapk, dim3(grid), dim3(256), 0, ctx->stream, (const float2 *)dk0, dk1, n, nk, total,
                           (float)dk, s, s);
        PP_HIP(hipGetLastError());
        PP_CHECK(c2r(dk1, f->lapd));
    }
    // 5. 2LPT: the source from the six potentials, then its gradient
    if (lpt2) {
        PP_CHECK(ctx->d_field_a.grow(4 * rb));
        float *p11 = ctx->d_field_a.as<float>(), *p22 = p11 + N, *p33 = p22 + N, *src = p33 + N;
        float *tmp = f->eta2x;  // the off-diagonal potentials pass through an output buffer
        PP_CHECK(lpt(3, p11));
        PP_CHECK(lpt(4, p22));
        PP_CHECK(lpt(6, tmp));  // p12
        hipLaunchKernelGGL(k_src2, dim3(grid), dim3(256), 0, ctx->stream, src, (const float *)p11, (const float *)p22,
                           (const float *)tmp, (long long)N, 0);
        PP_HIP(hipGetLastError());
        PP_CHECK(lpt(5, p33));
        PP_CHECK(lpt(7, tmp));  // p13
        hipLaunchKernelGGL(k_src2, dim3(grid), dim3(256), 0, ctx->stream, src, (const float *)p11, (const float *)p33,
                           (const float *)tmp, (long long)N, 1);
        PP_HIP(hipGetLastError());
        PP_CHECK(lpt(8, tmp));  // p23
        hipLaunchKernelGGL(k_src2, dim3(grid), dim3(256), 0, ctx->stream, src, (const float *)p22, (const float *)p33,
                           (const float *)tmp, (long long)N, 2);
        PP_HIP(hipGetLastError());
        PP_CHECK(r2c(src, dk0));
        PP_CHECK(lpt(9, f->eta2x));
        PP_CHECK(lpt(10, f->eta2y));
        PP_CHECK(lpt(11, f->eta2z));
    }
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}
