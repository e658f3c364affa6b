// pp_bookkeep.hip -- the drivers' per-peak bookkeeping between the peak
// finder and the shell analysis, on the device.
//
// Reference: run_tile (src/Pipeline.jl:217-231, 257-271) and process_tile!
// (src/MultiResolution.jl:959-988): for every peak its filter radius, its
// position x = xbx + alatt (i - (n+1)/2) (1-based i), its redshift (z_out,
// or in light-cone mode peak_redshift: the observer distance through the
// chi -> z table by bisection, src/Cosmology/Cosmology.jl:344-372), ZZon =
// 1 + z, fcrit = fsc_of_z(z) = 1.686 D(1/(1+z)) (RadialShell.jl:191-196),
// ir2min = min(floor((1.75 Rf / alatt)^2), floor((40 / alatt - 1)^2)), the
// growth factor D_a of Dlinear_ab and the 2LPT velocity factor
// 3/7 Om_a^(-1/143) D_a^2; in light-cone mode peaks with z > z_max are
// dropped, the others kept in order.  run_multitile_split rounds ZZon and
// fcrit through Float32 (MultiResolution.jl:963-967), run_tile keeps
// Float64 (round_f32 selects).
//
// Design: one thread per peak, all Float64, the growth and chi tables read
// from global memory (L2-resident); the z_max cut is a flag, a scan and an
// ordered scatter.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "pp_common.h"
#include "pp_growth.h"

namespace pp {

struct BookArgs {
    const int32_t *i, *j, *k, *filt;
    const float *dels, *laps;
    int64_t np;
    const double *Rf_filt;
    int nfilt;
    int n, ievol, lpt2, round_f32;
    double alatt, xbx, ybx, zbx, z_out, ox, oy, oz, z_max, Om, OL, ir2cap;
    const double *at, *Dt, *Ht;
    GrowthDev g;
    const double *cz, *cc;  // chi -> z table: z and chi columns
    int nchi;
    double z_top;           // the table's z_max (chi beyond the last row)
};

struct BookOut {
    int32_t *i, *j, *k;
    double *Rf;
    int64_t *ir2min;
    double *ZZon, *fcrit, *x, *y, *z, *D, *v2fac;
    float *FcRf, *d2Rf;
};

// chi_to_z: 0 below the first row, the table's z_max beyond the last, else
// linear between the bracketing rows found by bisection
__device__ double chi_to_z(const double *zs, const double *cs, int n, double z_top, double chi)
{
    if (chi <= cs[0]) return 0.0;
    if (chi >= cs[n - 1]) return z_top;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cs[mid] <= chi) lo = mid;
        else hi = mid;
    }
    const double frac = (chi - cs[lo]) / (cs[hi] - cs[lo]);
    return zs[lo] + frac * (zs[hi] - zs[lo]);
}

// per peak: the derived values at its own index, and keep[p] (1 = kept)
__global__ void __launch_bounds__(256) k_bookkeep(BookArgs A, BookOut T, int32_t *__restrict__ keep)
{
#pragma clang fp contract(off)
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.np; p += (int64_t)gridDim.x * blockDim.x) {
        const int f = A.filt[p];
        const double Rf = A.Rf_filt[f < 0 ? 0 : (f >= A.nfilt ? A.nfilt - 1 : f)];
        const double cen = 0.5 * (A.n + 1);
        const double x = A.xbx + A.alatt * ((double)(A.i[p] + 1) - cen);
        const double y = A.ybx + A.alatt * ((double)(A.j[p] + 1) - cen);
        const double z = A.zbx + A.alatt * ((double)(A.k[p] + 1) - cen);
        double zpk = A.z_out;
        if (A.ievol == 1) {
            const double dx = x - A.ox, dy = y - A.oy, dz = z - A.oz;
            zpk = chi_to_z(A.cz, A.cc, A.nchi, A.z_top, sqrt(dx * dx + dy * dy + dz * dz));
        }
        const double a = 1.0 / (1.0 + zpk);
        double Dl, HD, Da;
        dlinear(A.at, A.Dt, A.Ht, A.g, a, Dl, HD, Da);
        double ZZon = 1.0 + zpk, fcrit = 1.686 * Dl;
        if (A.round_f32) {
            ZZon = (double)(float)ZZon;
            fcrit = (double)(float)fcrit;
        }
        const double t = 1.75 * Rf / A.alatt;
        const double ir2 = fmin(floor(t * t), A.ir2cap);
        double v2 = 0.0;
        if (A.lpt2) {
            const double a3 = a * a * a;
            const double Om_a = A.Om * a3 / (A.Om * a3 + A.OL);
            v2 = -(-3.0 / 7.0 * pow(Om_a, -1.0 / 143) * (Da * Da));
        }
        T.Rf[p] = Rf;
        T.ir2min[p] = (int64_t)ir2;
        T.ZZon[p] = ZZon;
        T.fcrit[p] = fcrit;
        T.x[p] = x;
        T.y[p] = y;
        T.z[p] = z;
        T.D[p] = Da;
        T.v2fac[p] = v2;
        keep[p] = (A.ievol != 1 || zpk <= A.z_max) ? 1 : 0;
    }
}

// ordered compaction of the kept peaks from src (computed) to dst
__global__ void __launch_bounds__(256) k_book_compact(BookArgs A, BookOut src, BookOut dst,
                                                      const int32_t *__restrict__ keep,
                                                      const int64_t *__restrict__ pos)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.np; p += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[p]) continue;
        const int64_t q = pos[p];
        dst.i[q] = A.i[p];
        dst.j[q] = A.j[p];
        dst.k[q] = A.k[p];
        dst.Rf[q] = src.Rf[p];
        dst.ir2min[q] = src.ir2min[p];
        dst.ZZon[q] = src.ZZon[p];
        dst.fcrit[q] = src.fcrit[p];
        dst.x[q] = src.x[p];
        dst.y[q] = src.y[p];
        dst.z[q] = src.z[p];
        dst.D[q] = src.D[p];
        dst.v2fac[q] = src.v2fac[p];
        dst.FcRf[q] = A.dels[p];
        dst.d2Rf[q] = A.laps ? A.laps[p] : 0.0f;
    }
}

struct KeepToI64 {
    __device__ __forceinline__ int64_t operator()(const int32_t &v) const { return (int64_t)v; }
};

}  // namespace pp

using namespace pp;

extern "C" int pp_bookkeep_peaks(pp_ctx *ctx, const pp_peaks *peaks, const double *Rf_filters, int nfilt,
                             const pp_tile_geom *geo, pp_batch_out *out, int64_t *nkept)
{
    PP_ARG(ctx && peaks && Rf_filters && geo && out && nkept, "NULL argument");
    PP_ARG(nfilt >= 1, "no filters");
    PP_ARG(peaks->i && peaks->j && peaks->k && peaks->filt && peaks->delta_s, "peak arrays missing");
    PP_ARG(out->i && out->j && out->k && out->Rf && out->ir2min && out->ZZon && out->fcrit && out->x && out->y &&
               out->z && out->D && out->v2fac && out->FcRf && out->d2Rf,
           "every output array is required");
    PP_ARG(geo->n >= 1 && geo->alatt > 0, "bad tile geometry");
    PP_ARG(geo->growth.ntab >= 2 && geo->growth.atab && geo->growth.D_a && geo->growth.dlnD_dlna,
           "growth tables missing");
    PP_ARG(geo->ievol != 1 || (geo->nchi >= 2 && geo->chi_z && geo->chi_chi), "chi -> z table missing");
    PP_CHECK(ensure_device(ctx));
    const int64_t np = peaks->count;
    *nkept = 0;
    if (np <= 0) return PP_OK;
    const int nt = geo->growth.ntab, nc = geo->ievol == 1 ? geo->nchi : 0;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // tables, per-peak staging (computed values before the cut) and the scan
    const size_t tab_b = al((3 * (size_t)nt + 2 * (size_t)nc + (size_t)nfilt) * sizeof(double));
    const size_t per = al(np * 8);
    PP_CHECK(ctx->d_misc.grow(tab_b + 11 * per + 4096));
    char *p = ctx->d_misc.as<char>();
    double *d_tab = (double *)p;
    p += tab_b;
    PP_HIP(hipMemcpyAsync(d_tab, geo->growth.atab, nt * 8, hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipMemcpyAsync(d_tab + nt, geo->growth.D_a, nt * 8, hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipMemcpyAsync(d_tab + 2 * nt, geo->growth.dlnD_dlna, nt * 8, hipMemcpyHostToDevice, ctx->stream));
    if (nc) {
        PP_HIP(hipMemcpyAsync(d_tab + 3 * nt, geo->chi_z, nc * 8, hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(d_tab + 3 * nt + nc, geo->chi_chi, nc * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    double *d_rf = d_tab + 3 * nt + 2 * nc;
    PP_HIP(hipMemcpyAsync(d_rf, Rf_filters, nfilt * 8, hipMemcpyHostToDevice, ctx->stream));
    BookOut S;
    S.Rf = (double *)p; p += per;
    S.ir2min = (int64_t *)p; p += per;
    S.ZZon = (double *)p; p += per;
    S.fcrit = (double *)p; p += per;
    S.x = (double *)p; p += per;
    S.y = (double *)p; p += per;
    S.z = (double *)p; p += per;
    S.D = (double *)p; p += per;
    S.v2fac = (double *)p; p += per;
    int32_t *keep = (int32_t *)p; p += per;
    int64_t *pos = (int64_t *)p; p += per;
    BookArgs A;
    A.i = peaks->i; A.j = peaks->j; A.k = peaks->k; A.filt = peaks->filt;
    A.dels = peaks->delta_s; A.laps = peaks->lapd_s;
    A.np = np; A.Rf_filt = d_rf; A.nfilt = nfilt;
    A.n = geo->n; A.ievol = geo->ievol; A.lpt2 = geo->ilpt >= 2; A.round_f32 = geo->round_f32;
    A.alatt = geo->alatt; A.xbx = geo->xbx; A.ybx = geo->ybx; A.zbx = geo->zbx; A.z_out = geo->z_out;
    A.ox = geo->obs[0]; A.oy = geo->obs[1]; A.oz = geo->obs[2]; A.z_max = geo->z_max;
    A.Om = geo->Om; A.OL = geo->OL;
    {
        const double c = 40.0 / geo->alatt - 1.0;
        A.ir2cap = floor(c * c);
    }
    A.at = d_tab; A.Dt = d_tab + nt; A.Ht = d_tab + 2 * nt;
    A.g.ntab = nt;
    A.g.amin = geo->growth.atab[0]; A.g.amax = geo->growth.atab[nt - 1];
    A.g.D_inf = geo->growth.D_a[0]; A.g.D_0 = geo->growth.D_a[nt - 1];
    A.g.HD_inf = geo->growth.dlnD_dlna[0]; A.g.HD_0 = geo->growth.dlnD_dlna[nt - 1];
    A.cz = nc ? d_tab + 3 * nt : nullptr;
    A.cc = nc ? d_tab + 3 * nt + nc : nullptr;
    A.nchi = nc;
    A.z_top = geo->chi_zmax;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((np + 255) / 256, (int64_t)ctx->num_cu * 8));
    BookOut T;
    T.i = out->i; T.j = out->j; T.k = out->k; T.Rf = out->Rf; T.ir2min = out->ir2min; T.ZZon = out->ZZon;
    T.fcrit = out->fcrit; T.x = out->x; T.y = out->y; T.z = out->z; T.D = out->D; T.v2fac = out->v2fac;
    T.FcRf = out->FcRf; T.d2Rf = out->d2Rf;
    hipLaunchKernelGGL(k_bookkeep, dim3(grid), dim3(256), 0, ctx->stream, A, S, keep);
    PP_HIP(hipGetLastError());
    hipcub::TransformInputIterator<int64_t, KeepToI64, const int32_t *> it(keep, KeepToI64());
    size_t tmp = 0;
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, pos, (int)np, ctx->stream));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_cub.ptr, tmp, it, pos, (int)np, ctx->stream));
    hipLaunchKernelGGL(k_book_compact, dim3(grid), dim3(256), 0, ctx->stream, A, S, T, (const int32_t *)keep,
                       (const int64_t *)pos);
    PP_HIP(hipGetLastError());
    int64_t last_pos = 0;
    int32_t last_keep = 0;
    PP_HIP(hipMemcpyAsync(&last_pos, pos + np - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipMemcpyAsync(&last_keep, keep + np - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    *nkept = last_pos + last_keep;
    return PP_OK;
}
