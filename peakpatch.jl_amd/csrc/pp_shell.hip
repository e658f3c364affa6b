// pp_shell.hip -- per-peak radial-shell analysis (analyse_peak,
// src/HaloFinder/RadialShell.jl:282-891) on the GPU.
//
// Two kernel forms share the post-process in pp_post.h:
//  * PP_SHELL_PLAIN: one 64-lane workgroup per peak, one launch per batch.
//    The wave walks the r^2-sorted cell table shell by shell, reduces each
//    shell with a wave butterfly and keeps the profile in global memory
//    (the structure of the reference's _shell_gather_full_kernel! +
//    _post_process_kernel!, ext/CUDAExt.jl:859-2051, re-done for wave64).
//  * PP_SHELL_PERSISTENT: see pp_persist.hip.
//
// Semantics follow the production CPU code, not the reference's GPU
// prototype: npart cell cut for rmax2rs > 0, d2F from cell 2 with the stop
// at the first out-of-bounds cell, zvir_half computed.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "pp_common.h"
#include "pp_post.h"
#include "pp_shell.h"

namespace pp {

// --------------------------------------------------------- profile views
// Plain form: profile in global memory, SoA [field][shell] per peak.
struct GlobalProfile {
    const float *f;   // PF_NFIELD * m floats
    const double *fb; // Fbar, m doubles
    const double *radt;
    int m;
    bool lapd;
    __device__ double rad(int s) const { return radt[s - 1]; }
    __device__ double Fbar(int s) const { return fb[s - 1]; }
    __device__ double count(int s) const { return (double)f[PF_N * m + s - 1]; }
    __device__ double S(int s, int L) const { return (double)f[(PF_S + L) * m + s - 1]; }
    __device__ double S2(int s, int L) const { return (double)f[(PF_S2 + L) * m + s - 1]; }
    __device__ double lap(int s) const { return (double)f[PF_L * m + s - 1]; }
    __device__ bool has_lapd() const { return lapd; }
    __device__ void strain_inputs(int s, double G[3], double Gf[3], double SR[9]) const
    {
        if (s == 1) {
            for (int i = 0; i < 3; i++) G[i] = Gf[i] = 0.0;
            for (int i = 0; i < 9; i++) SR[i] = 0.0;
            return;
        }
        const double r = radt[s - 1];
        for (int i = 0; i < 3; i++) {
            G[i] = (double)f[(PF_G + i) * m + s - 1] / r;
            Gf[i] = (double)f[(PF_GF + i) * m + s - 1] / r;
        }
        for (int i = 0; i < 9; i++) SR[i] = (double)f[(PF_SR + i) * m + s - 1] / r;
    }
};

__device__ __forceinline__ void gather_cell(const GatherFields &F, long c, long cT, float dx, float dy, float dz,
                                            float a[PF_NFIELD])
{
    const float d = F.d[c];
    const float dT = F.d[cT];
    const float ex = F.ex[c], ey = F.ey[c], ez = F.ez[c];
    a[PF_N] += 1.0f;
    a[PF_F] += d;
    a[PF_S + 0] += ex;
    a[PF_S + 1] += ey;
    a[PF_S + 2] += ez;
    if (F.e2x) {
        a[PF_S2 + 0] += F.e2x[c];
        a[PF_S2 + 1] += F.e2y[c];
        a[PF_S2 + 2] += F.e2z[c];
    }
    a[PF_G + 0] += d * dx;
    a[PF_G + 1] += d * dy;
    a[PF_G + 2] += d * dz;
    a[PF_GF + 0] += dT * dx;
    a[PF_GF + 1] += dT * dy;
    a[PF_GF + 2] += dT * dz;
    a[PF_SR + 0] += ex * dx; a[PF_SR + 1] += ex * dy; a[PF_SR + 2] += ex * dz;
    a[PF_SR + 3] += ey * dx; a[PF_SR + 4] += ey * dy; a[PF_SR + 5] += ey * dz;
    a[PF_SR + 6] += ez * dx; a[PF_SR + 7] += ez * dy; a[PF_SR + 8] += ez * dz;
    if (F.lap) a[PF_L] += F.lap[c];
}

// Fbar recursion of the gather loop (RadialShell.jl:383-390), one lane.
__device__ void fbar_recursion(const float *f, int m, const double *radt, double *fb)
{
    double F1 = (double)f[PF_F * m + 0];
    fb[0] = F1;
    double dFbarp = F1;
    for (int s = 2; s <= m; s++) {
        double n = (double)f[PF_N * m + s - 1];
        double dF = n > 0.0 ? (double)f[PF_F * m + s - 1] / n : 0.0;
        double rp = radt[s - 2], r = radt[s - 1];
        double r3p = rp * rp * rp, r3 = r * r * r;
        fb[s - 1] = (r3p * fb[s - 2] + 0.5 * (dFbarp + dF) * (r3 - r3p)) / r3;
        dFbarp = dF;
    }
}

__global__ void __launch_bounds__(64) k_shell_plain(GatherFields F, const PeakParam *__restrict__ pks,
                                                    const int64_t *__restrict__ prof_off, float *__restrict__ prof,
                                                    double *__restrict__ fbar, PostConst C, double *__restrict__ res,
                                                    int32_t *__restrict__ status, int64_t p0)
{
    const int64_t p = p0 + blockIdx.x;
    const int lane = threadIdx.x;
    const PeakParam pk = pks[p];
    const int m = pk.m;
    float *pf = prof + prof_off[p] * PF_NFIELD;
    double *fb = fbar + prof_off[p];
    const long n1 = F.n, n12 = (long)F.n * F.n;
    const bool inside = pk.flags & 1;
    if (lane == 0) {
        const long c = pk.ci + pk.cj * n1 + pk.ck * n12;
        pf[PF_N * m] = 1.0f;
        pf[PF_F * m] = F.d[c];
        pf[(PF_S + 0) * m] = F.ex[c];
        pf[(PF_S + 1) * m] = F.ey[c];
        pf[(PF_S + 2) * m] = F.ez[c];
        for (int L = 0; L < 3; L++) pf[(PF_S2 + L) * m] = F.e2x ? (L == 0 ? F.e2x[c] : L == 1 ? F.e2y[c] : F.e2z[c]) : 0.0f;
        for (int q = PF_G; q < PF_NFIELD; q++) pf[q * m] = 0.0f;
    }
    for (int s = 2; s <= m; s++) {
        float a[PF_NFIELD];
#pragma unroll
        for (int q = 0; q < PF_NFIELD; q++) a[q] = 0.0f;
        const int c0 = C.shell_start[s - 1], c1 = C.shell_start[s];
        for (int c = c0 + lane; c < c1; c += 64) {
            const int pc = C.cell[c];
            const int di = (pc & 255) - 128, dj = ((pc >> 8) & 255) - 128, dk = ((pc >> 16) & 255) - 128;
            const int x = pk.ci + di, y = pk.cj + dj, z = pk.ck + dk;
            if (!inside && (x < 0 || x >= F.n || y < 0 || y >= F.n || z < 0 || z >= F.n)) continue;
            gather_cell(F, x + y * n1 + z * n12, z + x * n1 + y * n12, (float)di, (float)dj, (float)dk, a);
        }
#pragma unroll
        for (int q = 0; q < PF_NFIELD; q++) a[q] = wave_sumf(a[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < PF_NFIELD; q++) pf[q * m + s - 1] = a[q];
        }
    }
    __threadfence_block();
    __syncthreads();
    if (lane == 0) fbar_recursion(pf, m, C.rad, fb);
    __threadfence_block();
    __syncthreads();
    GlobalProfile P{pf, fb, C.rad, m, F.lap != nullptr};
    int st = analyse_post(P, pk, C, res + p * R_NREC);
    if (lane == 0) status[p] = st;
}

// ------------------------------------------------------------- host side
static double akk_host[1101];
static void build_akk()
{
    // atab4 (RadialShell.jl:133-155)
    const double ca = 3.0 / (2.0 * M_PI), cb = 1.0 / (4.0 * M_PI);
    for (int i = 1; i <= 1100; i++) {
        double u = (i - 1) * 0.01, ru = std::sqrt(u);
        double v;
        if (ru < 1.0) v = ca * (2.0 / 3.0 - u + 0.5 * u * ru);
        else if (ru <= 2.0) { double t = 2.0 - ru; v = cb * (t * t * t); }
        else v = 0.0;
        akk_host[i - 1] = v;
    }
    akk_host[1100] = 0.0;
}

int prepare_post_const(pp_ctx *ctx, int n, int nbuff, double alatt, const float *lapd, int8_t *mask, PostConst &C)
{
    ShellTable &T = ctx->shells;
    if (T.d_rad.bytes < (size_t)T.nshells * sizeof(double) || T.rad_rmax != T.rmax) {
        std::vector<double> rad(T.nshells);
        for (int s = 0; s < T.nshells; s++) rad[s] = std::sqrt((double)T.shell_r2[s]);
        PP_CHECK(T.d_rad.grow(rad.size() * sizeof(double)));
        PP_HIP(hipMemcpy(T.d_rad.ptr, rad.data(), rad.size() * sizeof(double), hipMemcpyHostToDevice));
        build_akk();
        PP_CHECK(T.d_akk.grow(sizeof akk_host));
        PP_HIP(hipMemcpy(T.d_akk.ptr, akk_host, sizeof akk_host, hipMemcpyHostToDevice));
        T.rad_rmax = T.rmax;
    }
    C.rad = T.d_rad.as<double>();
    C.akk = T.d_akk.as<double>();
    C.cell = T.d_cell.as<int32_t>();
    C.shell_start = T.d_shell_start.as<int32_t>();
    C.tab.data = ctx->d_table.as<float>();
    C.tab.nx = ctx->tnx; C.tab.ny = ctx->tny; C.tab.nz = ctx->tnz;
    C.tab.X1 = ctx->tX1; C.tab.X2 = ctx->tX2; C.tab.Y1 = ctx->tY1;
    C.tab.Y2 = ctx->tY2; C.tab.Z1 = ctx->tZ1; C.tab.Z2 = ctx->tZ2;
    // normalisation constants (RadialShell.jl:298-302) with hlatt = 1
    const double anor = 1.0;
    C.aRnor = 3.0 * anor / alatt;
    const double wnor = anor / (4.0 * M_PI);
    C.wRnor = 3.0 * wnor / alatt;
    C.n1 = C.n2 = C.n3 = n;
    C.nbuff = nbuff;
    C.lapd = lapd;
    C.mask = mask;
    return PP_OK;
}

// Per-peak parameters (host, Float64): npart, finalised shells m, the
// initial m0 and the Phase-3 shell mrf.
int build_peak_params(pp_ctx *ctx, const pp_peak_batch *pb, int n, double alatt, double rmax2rs,
                      const int8_t *mask_host, std::vector<PeakParam> &out)
{
    const ShellTable &T = ctx->shells;
    const int64_t np = pb->npeaks;
    out.resize(np);
    // shell (1-based) of a cell index (0-based)
    auto shell_of_cell = [&](int64_t c) {
        return (int)(std::upper_bound(T.shell_start.begin(), T.shell_start.end(), (int32_t)c) -
                     T.shell_start.begin());
    };
    struct RfCache { double Rf; int npart, m, mrf; };
    std::vector<RfCache> cache;
    for (int64_t p = 0; p < np; p++) {
        PeakParam &q = out[p];
        const int i = pb->i[p], j = pb->j[p], k = pb->k[p];
        PP_ARG(i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n, "peak %lld at (%d,%d,%d) outside the %d^3 grid",
               (long long)p, i, j, k, n);
        const double Rf = pb->Rf[p];
        const RfCache *rc = nullptr;
        for (const auto &e : cache) if (e.Rf == Rf) { rc = &e; break; }
        if (!rc) {
            RfCache e;
            e.Rf = Rf;
            int64_t npart = T.ncells;
            if (rmax2rs > 0.0) {
                const double f3p = 4.0 / 3.0 * M_PI;
                const double t = rmax2rs * Rf / alatt;
                double v = std::floor(f3p * (t * t * t));
                if (v < (double)npart) npart = v < 0 ? 0 : (int64_t)v;
            }
            e.npart = (int)npart;
            e.m = npart >= 2 ? shell_of_cell(npart - 1) - 1 : 1;
            // Phase 3 (RadialShell.jl:457-471): only the first cell of each
            // r^2 can improve the distance; it sees the shell counter before
            // its increment (slot - 2; slot 2 sees 1).
            const double tgt = (Rf / alatt) * (Rf / alatt);
            double rmrf = 10.0;
            int mrf = 1;
            if (npart >= 2) {
                const int slast = shell_of_cell(npart - 1);
                for (int s = 2; s <= slast; s++) {
                    double dd = std::fabs((double)T.shell_r2[s - 1] - tgt);
                    if (dd < rmrf) { mrf = s == 2 ? 1 : s - 2; rmrf = dd; }
                }
            }
            e.mrf = mrf;
            cache.push_back(e);
            rc = &cache.back();
        }
        q.ci = i; q.cj = j; q.ck = k;
        q.m = rc->m;
        q.mrf = rc->mrf;
        q.npart = rc->npart;
        // m0 (RadialShell.jl:392-399): first finalised shell with r^2 > ir2min
        int m0 = q.m >= 2 ? q.m : 1;
        for (int s = 2; s <= q.m; s++)
            if ((int64_t)T.shell_r2[s - 1] > pb->ir2min[p]) { m0 = s; break; }
        q.m0 = m0;
        const int R = (int)std::floor(std::sqrt((double)T.shell_r2[q.m - 1]));
        const bool inside = i - R >= 0 && i + R < n && j - R >= 0 && j + R < n && k - R >= 0 && k + R < n;
        q.flags = inside ? 1 : 0;
        // rmax2rs pre-mask (RadialShell.jl:311-316), read from the mask as it
        // is when the batch starts.
        if (rmax2rs > 0.0 && mask_host) {
            double lim = std::floor(Rf / alatt * rmax2rs);
            if ((double)mask_host[i + (int64_t)j * n + (int64_t)k * n * n] > lim) q.flags |= 2;
        }
        q.ZZon = pb->ZZon[p];
        q.fcrit = pb->fcrit[p];
        q.order = p;
    }
    return PP_OK;
}

int upload_fields(pp_ctx *ctx, const pp_fields *f, GatherFields &G)
{
    const int n = f->n1;
    const size_t N = (size_t)n * n * n, b = N * sizeof(float);
    const float *src[8] = {f->delta, f->etax, f->etay, f->etaz, f->eta2x, f->eta2y, f->eta2z, f->lapd};
    const float *dst[8];
    for (int q = 0; q < 8; q++) {
        if (!src[q]) { dst[q] = nullptr; continue; }
        if (f->on_device) { dst[q] = src[q]; continue; }
        PP_CHECK(ctx->d_in[q].grow(b));
        PP_HIP(hipMemcpyAsync(ctx->d_in[q].ptr, src[q], b, hipMemcpyHostToDevice, ctx->stream));
        dst[q] = ctx->d_in[q].as<float>();
    }
    G.d = dst[0]; G.ex = dst[1]; G.ey = dst[2]; G.ez = dst[3];
    G.e2x = dst[4]; G.e2y = dst[5]; G.e2z = dst[6]; G.lap = dst[7];
    G.n = n;
    return PP_OK;
}

int check_fields(pp_ctx *ctx, const pp_fields *f)
{
    PP_ARG(f && f->delta && f->etax && f->etay && f->etaz, "fields: delta and eta x/y/z are required");
    PP_ARG(f->n1 == f->n2 && f->n2 == f->n3 && f->n1 >= 3,
           "fields must be a cubic grid (the transposed delta read of RadialShell.jl:448 assumes n1 = n2 = n3)");
    PP_ARG((f->eta2x == nullptr) == (f->eta2y == nullptr) && (f->eta2y == nullptr) == (f->eta2z == nullptr),
           "eta2 x/y/z must be all given or all NULL");
    if (!ctx->have_table) { set_error("collapse table not uploaded"); return PP_ERR_STATE; }
    if (ctx->shells.rmax < 0) { set_error("shell tables not built"); return PP_ERR_STATE; }
    return PP_OK;
}

static void unpack_results(const std::vector<double> &R, const std::vector<int32_t> &st,
                           const std::vector<PeakParam> &pp, pp_results *out)
{
    const int64_t np = (int64_t)pp.size();
    for (int q = 0; q < 6; q++) out->dump_counts[q] = 0;
    for (int64_t p = 0; p < np; p++) {
        const double *r = &R[(size_t)p * R_NREC];
        const int64_t o = pp[p].order;
        if (out->RTHL) out->RTHL[o] = r[R_RTHL];
        if (out->Srb) out->Srb[o] = r[R_SRB];
        for (int L = 0; L < 3; L++) {
            if (out->Sbar) out->Sbar[o * 3 + L] = r[R_SBAR + L];
            if (out->Sbar2) out->Sbar2[o * 3 + L] = r[R_SBAR2 + L];
            if (out->gradpk) out->gradpk[o * 3 + L] = r[R_GRADPK + L];
            if (out->gradpkrf) out->gradpkrf[o * 3 + L] = r[R_GRADPKRF + L];
            if (out->gradpkf) out->gradpkf[o * 3 + L] = r[R_GRADPKF + L];
        }
        if (out->strain) for (int q = 0; q < 9; q++) out->strain[o * 9 + q] = r[R_STRAIN + q];
        if (out->Fbarx) out->Fbarx[o] = r[R_FBARX];
        if (out->e_v) out->e_v[o] = r[R_EV];
        if (out->p_v) out->p_v[o] = r[R_PV];
        if (out->zvir_half) out->zvir_half[o] = r[R_ZVIRH];
        if (out->d2F) out->d2F[o] = r[R_D2F];
        if (out->status) out->status[o] = st[p];
        if (out->nshells) out->nshells[o] = pp[p].m;
        int s = st[p];
        if (s >= 1 && s <= 5) out->dump_counts[s]++;
        else if (s == 6) out->dump_counts[0]++;  // non-finite strain: counted apart
    }
}

// Runs the plain kernel over params already on the device.
static int run_plain(pp_ctx *ctx, const GatherFields &G, const std::vector<PeakParam> &pp,
                     const PeakParam *d_pp, PostConst &C, double *d_res, int32_t *d_st)
{
    const int64_t np = (int64_t)pp.size();
    // profile scratch, processed in chunks of bounded size
    const size_t budget = (size_t)1 << 31;
    std::vector<int64_t> off(np);
    int64_t p = 0;
    while (p < np) {
        int64_t q = p, tot = 0;
        while (q < np && (tot + pp[q].m) * (int64_t)(PF_NFIELD * 4 + 8) <= (int64_t)budget) {
            off[q] = tot;
            tot += pp[q].m;
            q++;
        }
        if (q == p) { set_error("peak profile exceeds scratch budget"); return PP_ERR_OOM; }
        PP_CHECK(ctx->d_prof.grow((size_t)tot * (PF_NFIELD * 4 + 8) + (size_t)(q - p) * 8 + 256));
        char *base = ctx->d_prof.as<char>();
        int64_t *d_off = (int64_t *)base;
        float *d_f = (float *)(base + (((size_t)(q - p) * 8 + 255) & ~(size_t)255));
        double *d_fb = (double *)((char *)d_f + (((size_t)tot * PF_NFIELD * 4 + 255) & ~(size_t)255));
        PP_HIP(hipMemcpyAsync(d_off, off.data() + p, (q - p) * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_shell_plain, dim3((unsigned)(q - p)), dim3(64), 0, ctx->stream, G, d_pp + p,
                           (const int64_t *)d_off, d_f, d_fb, C, d_res + p * R_NREC, d_st + p, (int64_t)0);
        PP_HIP(hipGetLastError());
        // the offsets buffer is reused by the next chunk: wait for this one
        PP_HIP(hipStreamSynchronize(ctx->stream));
        p = q;
    }
    return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_analyse_peaks(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks, double alatt,
                     double rmax2rs, int nbuff, int8_t *mask, int mask_on_device, int kernel_form,
                     pp_results *out)
{
    PP_ARG(ctx && fields && peaks && out, "NULL argument");
    PP_ARG(peaks->npeaks >= 0, "negative npeaks");
    PP_ARG(alatt > 0.0, "alatt must be > 0");
    PP_ARG(kernel_form == PP_SHELL_PLAIN || kernel_form == PP_SHELL_PERSISTENT, "unknown kernel form %d", kernel_form);
    PP_CHECK(check_fields(ctx, fields));
    PP_CHECK(ensure_device(ctx));
    for (int q = 0; q < 6; q++) out->dump_counts[q] = 0;
    const int64_t np = peaks->npeaks;
    if (np == 0) return PP_OK;
    PP_ARG(peaks->i && peaks->j && peaks->k && peaks->Rf && peaks->ir2min && peaks->ZZon && peaks->fcrit,
           "peak batch arrays must all be given");
    const int n = fields->n1;
    std::vector<int8_t> mask_host_copy;
    const int8_t *mask_for_params = nullptr;
    // the pre-mask test can only fire where floor(Rf / alatt * rmax2rs) == 0
    bool premask_possible = false;
    if (mask && rmax2rs > 0.0)
        for (int64_t p = 0; p < np && !premask_possible; p++)
            premask_possible = std::floor(peaks->Rf[p] / alatt * rmax2rs) < 1.0;
    if (premask_possible) {
        if (mask_on_device) {
            mask_host_copy.resize((size_t)n * n * n);
            PP_HIP(hipMemcpy(mask_host_copy.data(), mask, mask_host_copy.size(), hipMemcpyDeviceToHost));
            mask_for_params = mask_host_copy.data();
        } else {
            mask_for_params = mask;
        }
    }
    std::vector<PeakParam> pp;
    PP_CHECK(build_peak_params(ctx, peaks, n, alatt, rmax2rs, mask_for_params, pp));
    GatherFields G;
    PP_CHECK(upload_fields(ctx, fields, G));
    int8_t *d_mask = nullptr;
    const size_t N = (size_t)n * n * n;
    if (mask) {
        if (mask_on_device) d_mask = mask;
        else {
            PP_CHECK(ctx->d_mask.grow(N));
            PP_HIP(hipMemcpyAsync(ctx->d_mask.ptr, mask, N, hipMemcpyHostToDevice, ctx->stream));
            d_mask = ctx->d_mask.as<int8_t>();
        }
    }
    PostConst C;
    PP_CHECK(prepare_post_const(ctx, n, nbuff, alatt, G.lap, d_mask, C));
    const size_t resb = (size_t)np * R_NREC * sizeof(double), stb = (size_t)np * sizeof(int32_t);
    const size_t ppb = (size_t)np * sizeof(PeakParam);
    PP_CHECK(ctx->d_res.grow(resb + stb + ppb + 1024));
    double *d_res = ctx->d_res.as<double>();
    int32_t *d_st = (int32_t *)((char *)d_res + resb);
    PeakParam *d_pp = (PeakParam *)((char *)d_st + ((stb + 255) & ~(size_t)255));
    int rc = PP_OK;
    if (kernel_form == PP_SHELL_PERSISTENT) {
        PP_CHECK(run_persistent(ctx, G, pp, C, d_res, d_st, d_pp));
    } else {
        PP_HIP(hipMemcpyAsync(d_pp, pp.data(), ppb, hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
        rc = run_plain(ctx, G, pp, d_pp, C, d_res, d_st);
        PP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
        PP_CHECK(rc);
        PP_HIP(hipStreamSynchronize(ctx->stream));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
        ctx->timing.shell_ms = ms;
        ctx->timing.gather_ms = ms;
    }
    std::vector<double> R((size_t)np * R_NREC);
    std::vector<int32_t> st(np);
    PP_HIP(hipMemcpyAsync(R.data(), d_res, resb, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipMemcpyAsync(st.data(), d_st, stb, hipMemcpyDeviceToHost, ctx->stream));
    if (mask && !mask_on_device) PP_HIP(hipMemcpyAsync(mask, d_mask, N, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    unpack_results(R, st, pp, out);
    return rc;
}

int pp_shell_profiles(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks, double alatt,
                      double rmax2rs, int kernel_form, int max_shells, int32_t *m_out, int32_t *nshell_out,
                      double *Fbar_out)
{
    PP_ARG(ctx && fields && peaks && m_out, "NULL argument");
    PP_ARG(max_shells >= 1, "max_shells must be >= 1");
    PP_ARG(kernel_form == PP_SHELL_PLAIN || kernel_form == PP_SHELL_PERSISTENT, "unknown kernel form %d", kernel_form);
    PP_CHECK(check_fields(ctx, fields));
    PP_CHECK(ensure_device(ctx));
    const int64_t np = peaks->npeaks;
    if (np == 0) return PP_OK;
    const int n = fields->n1;
    std::vector<PeakParam> pp;
    PP_CHECK(build_peak_params(ctx, peaks, n, alatt, rmax2rs, nullptr, pp));
    GatherFields G;
    PP_CHECK(upload_fields(ctx, fields, G));
    PostConst C;
    PP_CHECK(prepare_post_const(ctx, n, 0, alatt, G.lap, nullptr, C));
    const size_t resb = (size_t)np * R_NREC * sizeof(double), stb = (size_t)np * sizeof(int32_t);
    const size_t ppb = (size_t)np * sizeof(PeakParam);
    PP_CHECK(ctx->d_res.grow(resb + stb + ppb + 1024));
    double *d_res = ctx->d_res.as<double>();
    int32_t *d_st = (int32_t *)((char *)d_res + resb);
    PeakParam *d_pp = (PeakParam *)((char *)d_st + ((stb + 255) & ~(size_t)255));
    PP_HIP(hipMemcpyAsync(d_pp, pp.data(), ppb, hipMemcpyHostToDevice, ctx->stream));
    int64_t tot = 0;
    std::vector<int64_t> off(np);
    for (int64_t p = 0; p < np; p++) { off[p] = tot; tot += pp[p].m; }
    PP_ARG(tot * (int64_t)(PF_NFIELD * 4 + 8) <= ((int64_t)1 << 31), "profile export batch too large");
    // the plain kernel keeps every profile of a single chunk in d_prof
    PP_CHECK(run_plain(ctx, G, pp, d_pp, C, d_res, d_st));
    char *base = ctx->d_prof.as<char>();
    float *d_f = (float *)(base + (((size_t)np * 8 + 255) & ~(size_t)255));
    double *d_fb = (double *)((char *)d_f + (((size_t)tot * PF_NFIELD * 4 + 255) & ~(size_t)255));
    std::vector<float> hf((size_t)tot * PF_NFIELD);
    std::vector<double> hfb((size_t)tot);
    PP_HIP(hipMemcpy(hf.data(), d_f, hf.size() * 4, hipMemcpyDeviceToHost));
    PP_HIP(hipMemcpy(hfb.data(), d_fb, hfb.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < np; p++) {
        const int m = pp[p].m;
        m_out[p] = m;
        const float *f = &hf[(size_t)off[p] * PF_NFIELD];
        for (int s = 0; s < max_shells; s++) {
            if (nshell_out) nshell_out[p * max_shells + s] = s < m ? (int32_t)f[PF_N * m + s] : 0;
            if (Fbar_out) Fbar_out[p * max_shells + s] = s < m ? hfb[(size_t)off[p] + s] : 0.0;
        }
    }
    (void)kernel_form;
    return PP_OK;
}

int pp_autotune_shell(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks, double alatt,
                      double rmax2rs, int *choice)
{
    return autotune_persistent(ctx, fields, peaks, alatt, rmax2rs, choice);
}

}  // extern "C"
