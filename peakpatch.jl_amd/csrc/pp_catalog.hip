// pp_catalog.hip -- halo record packing from the shell-analysis results
// (src/Pipeline.jl:247-308; src/MultiResolution.jl:1025-1094): the collapsed
// peaks, in peak order, become the 11-float HaloRecord or the 33-float
// ExtHaloRecord of the .pksc catalogue (src/IO/Catalog.jl:48-94).
//
// Every field is one Float64 operation on the results and the host's
// per-peak growth factors, then a cast to Float32, exactly as the Julia
// record constructors do; the compaction keeps the sequential loop's order
// (flag, exclusive scan, scatter).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "pp_common.h"

namespace pp {

struct HaloIn {
    const double *x, *y, *z, *Rf, *ZZon, *D, *v2fac;
    const float *FcRf, *d2Rf;
    const double *RTHL, *Srb, *Sbar, *Sbar2, *strain, *gradpk, *gradpkrf, *Fbarx, *e_v, *p_v, *zvir_half,
        *gradpkf, *d2F;
};

__global__ void k_fill_f32_zero(float *p)
{
    if (threadIdx.x == 0) *p = 0.0f;
}

// keep(p): RTHL > 0 (Pipeline.jl:249) and not z > z_max (Pipeline.jl:252-255;
// z_max = +inf or NaN: no cut)
__global__ void k_halo_flags(HaloIn H, int64_t np, double z_max, int64_t *__restrict__ flag)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const bool keep = H.RTHL[p] > 0.0 && !(H.ZZon[p] - 1.0 > z_max);
        flag[p] = keep ? 1 : 0;
    }
}

__global__ void k_halo_pack(HaloIn H, int64_t np, double z_max, double alatt, int extended,
                            const int64_t *__restrict__ pos, float *__restrict__ rec, float *__restrict__ rthl_max)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const bool keep = H.RTHL[p] > 0.0 && !(H.ZZon[p] - 1.0 > z_max);
        if (!keep) continue;
        const int nf = extended ? 33 : 11;
        float *r = rec + pos[p] * nf;
        const double D = H.D[p];
        const double v2 = H.v2fac ? H.v2fac[p] : 0.0;
        const float RTHL_phys = (float)(H.RTHL[p] * alatt);
        r[0] = (float)H.x[p]; r[1] = (float)H.y[p]; r[2] = (float)H.z[p];
        for (int L = 0; L < 3; L++) r[3 + L] = (float)(H.Sbar[p * 3 + L] * D);
        r[6] = RTHL_phys;
        for (int L = 0; L < 3; L++) r[7 + L] = H.v2fac ? (float)(H.Sbar2[p * 3 + L] * v2) : 0.0f;
        r[10] = (float)H.Fbarx[p];
        if (extended) {
            const double *sm = H.strain + p * 9;  // column-major: sm[L + 3K] = strain_mat[L+1, K+1]
            r[11] = (float)H.e_v[p]; r[12] = (float)H.p_v[p];
            r[13] = (float)sm[0]; r[14] = (float)sm[4]; r[15] = (float)sm[8];
            r[16] = (float)sm[7]; r[17] = (float)sm[6]; r[18] = (float)sm[3];
            r[19] = (float)H.d2F[p];
            r[20] = (float)H.zvir_half[p];
            for (int L = 0; L < 3; L++) r[21 + L] = (float)H.gradpk[p * 3 + L];
            for (int L = 0; L < 3; L++) r[24 + L] = (float)H.gradpkf[p * 3 + L];
            r[27] = (float)H.Rf[p];
            r[28] = H.FcRf[p];
            r[29] = H.d2Rf ? H.d2Rf[p] : 0.0f;
            for (int L = 0; L < 3; L++) r[30 + L] = (float)H.gradpkrf[p * 3 + L];
        }
        // RTHL_phys > 0, so the float's bits order like its value
        atomicMax((unsigned int *)rthl_max, __float_as_uint(RTHL_phys));
    }
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_pack_halos(pp_ctx *ctx, const pp_results *res, const pp_halo_inputs *in, double alatt, double z_max,
                  int extended, float *records, int64_t capacity, int records_on_device, int64_t *nhalo,
                  float *RTHLmax)
{
    PP_ARG(ctx && res && in && nhalo && RTHLmax, "NULL argument");
    PP_ARG(res->on_device && in->on_device, "pp_pack_halos reads device-resident results and peak arrays");
    PP_ARG(alatt > 0.0, "alatt must be > 0");
    const int64_t np = in->npeaks;
    *nhalo = 0;
    *RTHLmax = 0.0f;
    PP_CHECK(ensure_device(ctx));
    if (np == 0) return PP_OK;
    PP_ARG(in->x && in->y && in->z && in->Rf && in->ZZon && in->D && in->FcRf, "peak arrays must be given");
    PP_ARG(res->RTHL && res->Sbar && res->Sbar2 && res->Fbarx, "results RTHL, Sbar, Sbar2, Fbarx must be given");
    PP_ARG(!extended || (res->strain && res->gradpk && res->gradpkrf && res->e_v && res->p_v && res->zvir_half &&
                         res->gradpkf && res->d2F),
           "extended records need every result field");
    HaloIn H{in->x, in->y, in->z, in->Rf, in->ZZon, in->D, in->v2fac, in->FcRf, in->d2Rf,
             res->RTHL, res->Srb, res->Sbar, res->Sbar2, res->strain, res->gradpk, res->gradpkrf, res->Fbarx,
             res->e_v, res->p_v, res->zvir_half, res->gradpkf, res->d2F};
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    PP_CHECK(ctx->d_scan.grow(al((size_t)np * 8) * 2 + 4096));
    int64_t *flag = ctx->d_scan.as<int64_t>();
    int64_t *pos = (int64_t *)(ctx->d_scan.as<char>() + al((size_t)np * 8));
    PP_CHECK(ctx->d_misc.grow(4096));
    float *d_max = (float *)(ctx->d_misc.as<char>() + 2048);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((np + 255) / 256, (int64_t)ctx->num_cu * 8));
    hipLaunchKernelGGL(k_halo_flags, dim3(grid), dim3(256), 0, ctx->stream, H, np, z_max, flag);
    PP_HIP(hipGetLastError());
    size_t tmp = 0;
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag, pos, (int)np, ctx->stream));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_cub.ptr, tmp, flag, pos, (int)np, ctx->stream));
    int64_t last[2];
    PP_HIP(hipMemcpyAsync(&last[0], pos + np - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipMemcpyAsync(&last[1], flag + np - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t nh = last[0] + last[1];
    *nhalo = nh;
    if (nh > capacity) {
        set_error("record buffer holds %lld halos, %lld needed", (long long)capacity, (long long)nh);
        return PP_ERR_CAPACITY;
    }
    if (nh == 0) return PP_OK;
    const int nf = extended ? 33 : 11;
    float *d_rec = records;
    if (!records_on_device) {
        PP_CHECK(ctx->d_k1.grow((size_t)nh * nf * sizeof(float)));
        d_rec = ctx->d_k1.as<float>();
    }
    hipLaunchKernelGGL(k_fill_f32_zero, dim3(1), dim3(64), 0, ctx->stream, d_max);
    hipLaunchKernelGGL(k_halo_pack, dim3(grid), dim3(256), 0, ctx->stream, H, np, z_max, alatt, extended,
                       (const int64_t *)pos, d_rec, d_max);
    PP_HIP(hipGetLastError());
    PP_HIP(hipMemcpyAsync(RTHLmax, d_max, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (!records_on_device)
        PP_HIP(hipMemcpyAsync(records, d_rec, (size_t)nh * nf * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

}  // extern "C"
