// pp_shell.hip -- per-peak radial-shell analysis (analyse_peak,
// src/HaloFinder/RadialShell.jl:282-891) on the GPU: host orchestration,
// device-side per-peak parameters, and the plain kernel form.
//
// Kernel forms (both end in the post-process of pp_post.h):
//  * PP_SHELL_PLAIN: one 64-lane workgroup per peak, one launch per batch.
//    The wave walks the r^2-sorted cell table shell by shell, reduces each
//    shell with a wave butterfly and keeps the profile in global memory (the
//    structure of the reference's _shell_gather_full_kernel! +
//    _post_process_kernel!, ext/CUDAExt.jl:859-2051, done for wave64).
//  * PP_SHELL_PERSISTENT: pp_persist.hip.
//
// Semantics follow the production CPU code, not the reference's GPU
// prototype: npart cell cut for rmax2rs > 0, d2F from cell 2 with the stop
// at the first out-of-bounds cell, zvir_half computed.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "pp_common.h"
#include "pp_post.h"
#include "pp_shell.h"

namespace pp {

// ------------------------------------------------------ per-peak params
__device__ __forceinline__ int d_shell_of_cell(const int32_t *start, int nshells, int c)
{
    // 1-based shell s with start[s-1] <= c < start[s] (upper_bound over start[0..nshells])
    int lo = 0, hi = nshells + 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (start[mid] <= c) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// npart (RadialShell.jl:289-295), finalised shells m (the last shell touched
// by cells 2..npart is never finalised, :379-453), m0 (first finalised shell
// with r^2 > ir2min, :392-399), the Phase-3 shell mrf (:457-471: only the first
// cell of an r^2 can improve the distance and it sees the counter before its
// increment, so slot s maps to s-2, slot 2 to 1), the in-grid flag and the
// rmax2rs pre-mask (:311-316).
__global__ void k_peak_params(BatchDev b, int64_t np, const int32_t *__restrict__ start,
                              const int32_t *__restrict__ r2, int nshells, int64_t ncells, int n, double alatt,
                              double rmax2rs, double f3p, const int8_t *__restrict__ mask, PeakParam *__restrict__ out,
                              int *__restrict__ nbad)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        PeakParam q;
        const int i = b.i[p], j = b.j[p], k = b.k[p];
        q.ci = i; q.cj = j; q.ck = k;
        q.order = p;
        q.ZZon = b.ZZon[p];
        q.fcrit = b.fcrit[p];
        const bool valid = i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n;
        if (!valid) {
            atomicAdd(nbad, 1);
            q.ci = q.cj = q.ck = 0;
        }
        const double Rf = b.Rf[p];
        int64_t npart = ncells;
        if (rmax2rs > 0.0) {
            const double t = rmax2rs * Rf / alatt;
            const double v = floor(f3p * (t * t * t));
            if (v < (double)npart) npart = v < 0.0 ? 0 : (int64_t)v;
        }
        q.npart = (int32_t)npart;
        int slast = npart >= 2 ? d_shell_of_cell(start, nshells, (int)(npart - 1)) : 1;
        q.m = npart >= 2 ? slast - 1 : 1;
        // mrf: argmin over slots 2..slast of |r2 - tgt| (first on ties), if < 10
        const double tgt = (Rf / alatt) * (Rf / alatt);
        int mrf = 1;
        if (slast >= 2) {
            int lo = 2, hi = slast + 1;  // first s in [2, slast] with r2 >= tgt
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if ((double)r2[mid - 1] >= tgt) hi = mid; else lo = mid + 1;
            }
            int best = -1;
            double bd = 10.0;
            if (lo - 1 >= 2) { double d = fabs((double)r2[lo - 2] - tgt); if (d < bd) { bd = d; best = lo - 1; } }
            if (lo <= slast) { double d = fabs((double)r2[lo - 1] - tgt); if (d < bd) { bd = d; best = lo; } }
            if (best >= 2) mrf = best == 2 ? 1 : best - 2;
        }
        q.mrf = mrf;
        int m0 = q.m >= 2 ? q.m : 1;
        if (q.m >= 2) {
            const int64_t thr = b.ir2min[p];
            int lo = 2, hi = q.m + 1;  // first s in [2, m] with r2 > ir2min
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if ((int64_t)r2[mid - 1] > thr) hi = mid; else lo = mid + 1;
            }
            if (lo <= q.m) m0 = lo;
        }
        q.m0 = m0;
        const int R = (int)floor(sqrt((double)r2[q.m - 1]));
        const bool inside = valid && i - R >= 0 && i + R < n && j - R >= 0 && j + R < n && k - R >= 0 && k + R < n;
        q.flags = (inside ? 1 : 0) | (valid ? 0 : 8);
        if (valid && rmax2rs > 0.0 && mask) {
            const double lim = floor(Rf / alatt * rmax2rs);
            if ((double)mask[i + (int64_t)j * n + (int64_t)k * n * n] > lim) q.flags |= 2;
        }
        out[p] = q;
    }
}

// --------------------------------------------------------- profile views
// Plain form: profile in global memory, SoA [field][shell] per peak.
struct GlobalProfile {
    const float *f;   // PF_NFIELD * m floats
    const double *fb; // Fbar, m doubles
    const double *radt;
    int m;
    bool lapd;
    const WalkRec *wt = nullptr;  // no walk table in this form
    int wmc = 0;
    __device__ double rad(int s) const { return radt[s - 1]; }
    __device__ double Fbar(int s) const { return fb[s - 1]; }
    __device__ double count(int s) const { return (double)f[PF_N * m + s - 1]; }
    __device__ double S(int s, int L) const { return (double)f[(PF_S + L) * m + s - 1]; }
    __device__ double S2(int s, int L) const { return (double)f[(PF_S2 + L) * m + s - 1]; }
    __device__ double lap(int s) const { return (double)f[PF_L * m + s - 1]; }
    __device__ bool has_lapd() const { return lapd; }
    __device__ void sums(int s, double &cnt, double Sv[3], double S2v[3]) const
    {
        cnt = count(s);
        for (int L = 0; L < 3; L++) { Sv[L] = S(s, L); S2v[L] = S2(s, L); }
    }
    // centre shell sums are zero (RadialShell.jl:357-364)
    __device__ void strain_G(int s, float G[3], float Gf[3]) const
    {
        for (int i = 0; i < 3; i++) {
            G[i] = f[(PF_G + i) * m + s - 1];
            Gf[i] = f[(PF_GF + i) * m + s - 1];
        }
    }
    __device__ void strain_E(int s, float Es[6]) const
    {
        float SR[9];
        for (int i = 0; i < 9; i++) SR[i] = f[(PF_SR + i) * m + s - 1];
        Es[0] = SR[0] + SR[0]; Es[1] = SR[4] + SR[4]; Es[2] = SR[8] + SR[8];
        Es[3] = SR[1] + SR[3]; Es[4] = SR[2] + SR[6]; Es[5] = SR[5] + SR[7];
    }
    __device__ void strain_E2(int s, float Es[12]) const { strain_E(s, Es); strain_E(s + 1, Es + 6); }
    __device__ void strain_G2(int s, float G[6], float Gf[6]) const
    {
        strain_G(s, G, Gf);
        strain_G(s + 1, G + 3, Gf + 3);
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
                                                    double *__restrict__ fbar, PostConst C, ResultView R)
{
    const int64_t p = blockIdx.x;
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
    analyse_post<64, true>(P, pk, C, R, pk.order);
    if (lane == 0) R.nshells[pk.order] = m;
}

__global__ void k_status_hist(const int32_t *__restrict__ st, int64_t np, unsigned long long *__restrict__ hist)
{
    __shared__ unsigned int h[8];
    if (threadIdx.x < 8) h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        int s = st[p];
        if (s >= 0 && s < 8) atomicAdd(&h[s], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 8 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

__global__ void k_fill_i32(int *p, int n, int v)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// kernel_strain weight rows (one block per row)
__global__ void k_wtab(const double *__restrict__ rad, const double *__restrict__ inv_rad,
                       const double *__restrict__ akk, const int4 *__restrict__ rows, int nrows,
                       double *__restrict__ w)
{
    for (int mp = blockIdx.x + 1; mp <= nrows; mp += gridDim.x) {
        const int4 r = rows[mp - 1];
        for (int m1 = r.y + (int)threadIdx.x; m1 <= r.z; m1 += blockDim.x)
            w[r.x + m1 - r.y] = strain_weight(rad, inv_rad, akk, mp, m1);
    }
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
    if (T.d_rad.bytes < 2 * (size_t)T.nshells * sizeof(double) || T.rad_rmax != T.rmax) {
        std::vector<double> rad(2 * (size_t)T.nshells);
        for (int s = 0; s < T.nshells; s++) {
            rad[s] = std::sqrt((double)T.shell_r2[s]);
            rad[T.nshells + s] = 1.0 / rad[s];
        }
        PP_CHECK(T.d_rad.grow(rad.size() * sizeof(double)));
        PP_HIP(hipMemcpy(T.d_rad.ptr, rad.data(), rad.size() * sizeof(double), hipMemcpyHostToDevice));
        build_akk();
        PP_CHECK(T.d_akk.grow(sizeof akk_host));
        PP_HIP(hipMemcpy(T.d_akk.ptr, akk_host, sizeof akk_host, hipMemcpyHostToDevice));
        // weight rows: mp > 1 keeps the m1 with (rad(mp) - rad(m1))^2 < 4
        // (a contiguous run, rad being increasing); the centre row (mp = 1)
        // has no such cut in the reference, and its callers' windows end at
        // the last shell with rad <= 2 (r2 <= 4), which the row covers.
        const int ns = T.nshells;
        std::vector<int4> rows(ns);
        int64_t tot = 0;
        int lo = 1, hi = 1;
        for (int mp = 1; mp <= ns; mp++) {
            int a, b;
            if (mp == 1) {
                a = 1;
                b = 1;
                while (b + 1 <= ns && T.shell_r2[b] <= 4) b++;
            } else {
                const double rmp = rad[mp - 1];
                auto in = [&](int m1) { const double u0 = rmp - rad[m1 - 1]; return u0 * u0 < 4.0; };
                while (lo < mp && !in(lo)) lo++;
                if (hi < mp) hi = mp;
                while (hi + 1 <= ns && in(hi + 1)) hi++;
                a = lo;
                b = hi;
            }
            if (tot + (b - a + 1) > INT32_MAX) { set_error("kernel weight table too large"); return PP_ERR_CAPACITY; }
            rows[mp - 1] = make_int4((int)tot, a, b, 0);
            tot += b - a + 1;
        }
        const int r2max = T.shell_r2[ns - 1];
        std::vector<int32_t> ge((size_t)r2max + 2);
        for (int k = 0, sh = 1; k <= r2max + 1; k++) {
            while (sh <= ns && T.shell_r2[sh - 1] < k) sh++;
            ge[k] = sh;
        }
        PP_CHECK(T.d_shell_ge.grow(ge.size() * sizeof(int32_t)));
        PP_HIP(hipMemcpy(T.d_shell_ge.ptr, ge.data(), ge.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        PP_CHECK(T.d_wrow.grow(rows.size() * sizeof(int4)));
        PP_CHECK(T.d_wtab.grow((size_t)tot * sizeof(double)));
        PP_HIP(hipMemcpy(T.d_wrow.ptr, rows.data(), rows.size() * sizeof(int4), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_wtab, dim3(std::min(ns, 4096)), dim3(256), 0, ctx->stream, T.d_rad.as<double>(),
                           T.d_rad.as<double>() + ns, T.d_akk.as<double>(), T.d_wrow.as<int4>(), ns,
                           T.d_wtab.as<double>());
        PP_HIP(hipGetLastError());
        T.rad_rmax = T.rmax;
    }
    C.rad = T.d_rad.as<double>();
    C.inv_rad = T.d_rad.as<double>() + T.nshells;
    C.akk = T.d_akk.as<double>();
    C.wrow = T.d_wrow.as<int4>();
    C.shell_ge = T.d_shell_ge.as<int32_t>();
    C.shell_r2 = T.d_shell_r2.as<int32_t>();
    C.r2max = T.shell_r2[T.nshells - 1];
    C.nshells = T.nshells;
    C.wtab = T.d_wtab.as<double>();
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
    G.dT = nullptr;
    G.p4 = nullptr;
    G.q4 = nullptr;
    G.n = n;
    return PP_OK;
}

int check_fields(pp_ctx *ctx, const pp_fields *f)
{
    PP_ARG(f && f->delta && f->etax && f->etay && f->etaz, "fields: delta and eta x/y/z are required");
    PP_ARG(f->n1 == f->n2 && f->n2 == f->n3 && f->n1 >= 3,
           "fields must be a cubic grid (the transposed delta read of RadialShell.jl:448 assumes n1 = n2 = n3)");
    PP_ARG(f->n1 <= 1290, "grid %d too large for 32-bit cell offsets", f->n1);
    PP_ARG((f->eta2x == nullptr) == (f->eta2y == nullptr) && (f->eta2y == nullptr) == (f->eta2z == nullptr),
           "eta2 x/y/z must be all given or all NULL");
    if (!ctx->have_table) { set_error("collapse table not uploaded"); return PP_ERR_STATE; }
    if (ctx->shells.rmax < 0) { set_error("shell tables not built"); return PP_ERR_STATE; }
    return PP_OK;
}

// Device copy of the batch + per-peak parameters.  Fills `S` (pointers into
// ctx scratch) and returns PP_ERR_ARG if any peak lies outside the grid.
// Results go to ctx scratch, or straight into the caller's arrays where `out`
// holds device arrays.
int prepare_batch(pp_ctx *ctx, const pp_peak_batch *pb, int n, double alatt, double rmax2rs,
                  const int8_t *d_mask, ShellRun &S, const pp_results *out)
{
    const int64_t np = pb->npeaks;
    const ShellTable &T = ctx->shells;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bb = al(np * 4) * 3 + al(np * 8) * 4;          // i,j,k + Rf, ir2min, ZZon, fcrit
    const size_t pb_ = al(np * sizeof(PeakParam));
    const size_t ob = al(np * 4) * 2;                           // order + sort keys
    const size_t rb = al(np * 8) * 31 + al(np * 4) * 4;         // 31 doubles + status, nshells, m0, q
    PP_CHECK(ctx->d_pk.grow(bb + pb_ + ob + rb + 1024));
    char *p = ctx->d_pk.as<char>();
    BatchDev B;
    if (pb->on_device) {
        B.i = pb->i; B.j = pb->j; B.k = pb->k; B.Rf = pb->Rf; B.ir2min = pb->ir2min; B.ZZon = pb->ZZon;
        B.fcrit = pb->fcrit;
    } else {
        const void *src[7] = {pb->i, pb->j, pb->k, pb->Rf, pb->ir2min, pb->ZZon, pb->fcrit};
        const size_t sz[7] = {4, 4, 4, 8, 8, 8, 8};
        void *dst[7];
        char *q = p;
        for (int t = 0; t < 7; t++) {
            dst[t] = q;
            PP_HIP(hipMemcpyAsync(q, src[t], np * sz[t], hipMemcpyHostToDevice, ctx->stream));
            q += al(np * sz[t]);
        }
        B.i = (const int32_t *)dst[0]; B.j = (const int32_t *)dst[1]; B.k = (const int32_t *)dst[2];
        B.Rf = (const double *)dst[3]; B.ir2min = (const int64_t *)dst[4]; B.ZZon = (const double *)dst[5];
        B.fcrit = (const double *)dst[6];
    }
    p += bb;
    S.batch = B;
    S.pp = (PeakParam *)p; p += pb_;
    S.order = (int32_t *)p; p += al(np * 4);
    S.keys = (int32_t *)p; p += al(np * 4);
    ResultView &R = S.R;
    double **dptr[13] = {&R.RTHL, &R.Srb, &R.Sbar, &R.Sbar2, &R.strain, &R.gradpk, &R.gradpkrf, &R.Fbarx,
                         &R.e_v, &R.p_v, &R.zvir_half, &R.gradpkf, &R.d2F};
    const int width[13] = {1, 1, 3, 3, 9, 3, 3, 1, 1, 1, 1, 3, 1};
    for (int t = 0; t < 13; t++) { *dptr[t] = (double *)p; p += al(np * 8 * width[t]); }
    R.status = (int32_t *)p; p += al(np * 4);
    R.nshells = (int32_t *)p; p += al(np * 4);
    R.m0 = (int32_t *)p; p += al(np * 4);
    R.q = (int32_t *)p; p += al(np * 4);
    if (out && out->on_device) {
        double *const dst[13] = {out->RTHL, out->Srb, out->Sbar, out->Sbar2, out->strain, out->gradpk,
                                 out->gradpkrf, out->Fbarx, out->e_v, out->p_v, out->zvir_half, out->gradpkf,
                                 out->d2F};
        for (int t = 0; t < 13; t++)
            if (dst[t]) *dptr[t] = dst[t];
        if (out->status) R.status = out->status;
        if (out->nshells) R.nshells = out->nshells;
    }
    PP_CHECK(ctx->d_misc.grow(4096));
    int *d_bad = ctx->d_misc.as<int>();
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, d_bad, 1, 0);
    const int grid = (int)std::min<int64_t>((np + 255) / 256, (int64_t)ctx->num_cu * 8);
    hipLaunchKernelGGL(k_peak_params, dim3(std::max(grid, 1)), dim3(256), 0, ctx->stream, B, np,
                       T.d_shell_start.as<int32_t>(), T.d_shell_r2.as<int32_t>(), T.nshells, T.ncells, n, alatt,
                       rmax2rs, 4.0 / 3.0 * M_PI, d_mask, S.pp, d_bad);
    PP_HIP(hipGetLastError());
    int bad = 0;
    PP_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    PP_ARG(bad == 0, "%d peak(s) outside the %d^3 grid", bad, n);
    S.np = np;
    return PP_OK;
}

// Runs the plain kernel: one launch per chunk of peaks whose profiles fit the
// scratch budget.
int run_plain(pp_ctx *ctx, const GatherFields &G, ShellRun &S, PostConst &C, std::vector<int64_t> *offs_out)
{
    const int64_t np = S.np;
    std::vector<PeakParam> hp(np);
    PP_HIP(hipMemcpyAsync(hp.data(), S.pp, np * sizeof(PeakParam), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    const size_t rec = PF_NFIELD * 4 + 8;
    const int64_t budget = (int64_t)1 << 32;
    std::vector<int64_t> off(np);
    int64_t p = 0;
    while (p < np) {
        int64_t q = p, tot = 0;
        while (q < np && (tot + hp[q].m) * (int64_t)rec <= budget) {
            off[q] = tot;
            tot += hp[q].m;
            q++;
        }
        if (q == p) { set_error("peak profile exceeds scratch budget"); return PP_ERR_OOM; }
        if (offs_out && (p != 0 || q < np)) { set_error("profile export needs a single chunk"); return PP_ERR_ARG; }
        const size_t offb = (((size_t)(q - p) * 8) + 255) & ~(size_t)255;
        const size_t fb = (((size_t)tot * PF_NFIELD * 4) + 255) & ~(size_t)255;
        PP_CHECK(ctx->d_prof.grow(offb + fb + (size_t)tot * 8 + 256));
        char *base = ctx->d_prof.as<char>();
        int64_t *d_off = (int64_t *)base;
        float *d_f = (float *)(base + offb);
        double *d_fb = (double *)(base + offb + fb);
        PP_HIP(hipMemcpyAsync(d_off, off.data() + p, (q - p) * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_shell_plain, dim3((unsigned)(q - p)), dim3(64), 0, ctx->stream, G, S.pp + p,
                           (const int64_t *)d_off, d_f, d_fb, C, S.R);
        PP_HIP(hipGetLastError());
        // the offsets buffer is reused by the next chunk
        PP_HIP(hipStreamSynchronize(ctx->stream));
        p = q;
    }
    if (offs_out) *offs_out = off;
    return PP_OK;
}

static int copy_results(pp_ctx *ctx, const ShellRun &S, pp_results *out)
{
    const int64_t np = S.np;
    const hipMemcpyKind kind = out->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const ResultView &R = S.R;
    struct Cp { void *dst; const void *src; size_t b; };
    const Cp cp[] = {
        {out->RTHL, R.RTHL, 8}, {out->Srb, R.Srb, 8}, {out->Sbar, R.Sbar, 24}, {out->Sbar2, R.Sbar2, 24},
        {out->strain, R.strain, 72}, {out->gradpk, R.gradpk, 24}, {out->gradpkrf, R.gradpkrf, 24},
        {out->Fbarx, R.Fbarx, 8}, {out->e_v, R.e_v, 8}, {out->p_v, R.p_v, 8}, {out->zvir_half, R.zvir_half, 8},
        {out->gradpkf, R.gradpkf, 24}, {out->d2F, R.d2F, 8}, {out->status, R.status, 4},
        {out->nshells, R.nshells, 4}};
    for (const Cp &c : cp)
        if (c.dst && c.dst != c.src) PP_HIP(hipMemcpyAsync(c.dst, c.src, c.b * np, kind, ctx->stream));
    unsigned long long *d_h = (unsigned long long *)(ctx->d_misc.as<char>() + 512);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, d_h, 8, 0ull);
    hipLaunchKernelGGL(k_status_hist, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((np + 255) / 256, 1024))),
                       dim3(256), 0, ctx->stream, (const int32_t *)R.status, np, d_h);
    PP_HIP(hipGetLastError());
    unsigned long long h[8];
    PP_HIP(hipMemcpyAsync(h, d_h, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < 7; q++) out->dump_counts[q] = (int64_t)h[q];
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
    PP_ARG(peaks->npeaks < (1ll << 31), "batch larger than 2^31 peaks");
    PP_ARG(alatt > 0.0, "alatt must be > 0");
    PP_ARG(kernel_form == PP_SHELL_PLAIN || kernel_form == PP_SHELL_PERSISTENT, "unknown kernel form %d", kernel_form);
    PP_CHECK(check_fields(ctx, fields));
    PP_CHECK(ensure_device(ctx));
    for (int q = 0; q < 7; q++) out->dump_counts[q] = 0;
    for (double &t : ctx->timing.stage_ms) t = 0.0;
    const int64_t np = peaks->npeaks;
    if (np == 0) return PP_OK;
    PP_ARG(peaks->i && peaks->j && peaks->k && peaks->Rf && peaks->ir2min && peaks->ZZon && peaks->fcrit,
           "peak batch arrays must all be given");
    const int n = fields->n1;
    const size_t N = (size_t)n * n * n;
    GatherFields G;
    PP_CHECK(upload_fields(ctx, fields, G));
    PP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));  // shell_ms: from here to the end of the analysis
    // the persistent form's field packing starts now, beside the set-up below;
    // whatever happens next, the side stream has finished with the caller's
    // fields before this call returns
    struct SideJoin {
        hipStream_t s = nullptr;
        ~SideJoin() { if (s) (void)hipStreamSynchronize(s); }
    } side_join;
    if (kernel_form == PP_SHELL_PERSISTENT) {
        side_join.s = ctx->side;
        PP_CHECK(start_pack(ctx, G));
    }
    int8_t *d_mask = nullptr;
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
    ShellRun S;
    PP_CHECK(prepare_batch(ctx, peaks, n, alatt, rmax2rs, d_mask, S, out));
    if (kernel_form == PP_SHELL_PERSISTENT) {
        PP_CHECK(run_persistent(ctx, G, S, C));
    } else {
        PP_CHECK(run_plain(ctx, G, S, C, nullptr));
    }
    PP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->timing.shell_ms = ms;
    PP_CHECK(copy_results(ctx, S, out));
    if (mask && !mask_on_device) {
        PP_HIP(hipMemcpyAsync(mask, d_mask, N, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
    }
    return PP_OK;
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
    GatherFields G;
    PP_CHECK(upload_fields(ctx, fields, G));
    PostConst C;
    PP_CHECK(prepare_post_const(ctx, n, 0, alatt, G.lap, nullptr, C));
    ShellRun S;
    PP_CHECK(prepare_batch(ctx, peaks, n, alatt, rmax2rs, nullptr, S));
    std::vector<PeakParam> hp(np);
    PP_HIP(hipMemcpy(hp.data(), S.pp, np * sizeof(PeakParam), hipMemcpyDeviceToHost));
    std::vector<int32_t> cnt;
    std::vector<double> fb;
    std::vector<int64_t> off(np);
    if (kernel_form == PP_SHELL_PERSISTENT) {
        PP_CHECK(persistent_profiles(ctx, G, S, C, cnt, fb, off));
    } else {
        std::vector<int64_t> o;
        PP_CHECK(run_plain(ctx, G, S, C, &o));
        int64_t tot = 0;
        for (int64_t p = 0; p < np; p++) tot += hp[p].m;
        char *base = ctx->d_prof.as<char>();
        const size_t offb = (((size_t)np * 8) + 255) & ~(size_t)255;
        const size_t fbb = (((size_t)tot * PF_NFIELD * 4) + 255) & ~(size_t)255;
        std::vector<float> hf((size_t)tot * PF_NFIELD);
        fb.resize(tot);
        PP_HIP(hipMemcpy(hf.data(), base + offb, hf.size() * 4, hipMemcpyDeviceToHost));
        PP_HIP(hipMemcpy(fb.data(), base + offb + fbb, fb.size() * 8, hipMemcpyDeviceToHost));
        cnt.resize(tot);
        for (int64_t p = 0; p < np; p++) {
            const int m = hp[p].m;
            for (int s = 0; s < m; s++) cnt[o[p] + s] = (int32_t)hf[(size_t)o[p] * PF_NFIELD + PF_N * m + s];
        }
        off = o;
    }
    for (int64_t p = 0; p < np; p++) {
        const int m = hp[p].m;
        m_out[p] = m;
        for (int s = 0; s < max_shells; s++) {
            if (nshell_out) nshell_out[p * max_shells + s] = s < m ? cnt[off[p] + s] : 0;
            if (Fbar_out) Fbar_out[p * max_shells + s] = s < m ? fb[off[p] + s] : 0.0;
        }
    }
    return PP_OK;
}

// Coordinate descent over the persistent form's launch parameters, each
// candidate timed (HIP events around the whole shell call, best of two runs)
// on the caller's batch; the fastest setting stays in the context.
int pp_autotune_shell(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks, double alatt,
                      double rmax2rs, int *choice)
{
    PP_ARG(ctx && fields && peaks, "NULL argument");
    pp_results res;
    memset(&res, 0, sizeof res);
    auto timed = [&](double &best) -> int {
        best = 1e300;
        for (int rep = 0; rep < 2; rep++) {
            PP_CHECK(pp_analyse_peaks(ctx, fields, peaks, alatt, rmax2rs, 0, nullptr, 0, PP_SHELL_PERSISTENT, &res));
            best = std::min(best, ctx->timing.shell_ms);
        }
        return PP_OK;
    };
    double t = 0;
    PP_CHECK(timed(t));  // warm-up: tables, buffers
    static const int cand_small[] = {0, 8, 12, 16, 24, 32, 48, 64};
    static const int cand_chunk[] = {32, 64, 128, 256};
    static const int cand_coop[] = {32, 64, 128, 256, 1 << 20};
    static const int cand_waves[] = {8, 12, 16, 24, 32};
    static const int cand_width[] = {1, 2, 4, 8, 16};
    static const int cand_walk[] = {0, 8, 16, 32, 64};
    struct Knob { int *v; const int *c; int nc; };
    const Knob knobs[] = {{&ctx->chunk_shells, cand_chunk, 4}, {&ctx->coop_m, cand_coop, 5},
                          {&ctx->post_width, cand_width, 5}, {&ctx->lane_walk, cand_walk, 5},
                          {&ctx->gather_waves_cu, cand_waves, 5},
                          {&ctx->tiny_m, cand_small, 8}};
    for (const Knob &k : knobs) {
        int bestv = *k.v;
        double bt = 1e300;
        for (int q = 0; q < k.nc; q++) {
            *k.v = k.c[q];
            PP_CHECK(timed(t));
            if (t < bt) { bt = t; bestv = k.c[q]; }
        }
        *k.v = bestv;
    }
    if (choice) *choice = ctx->gather_waves_cu;
    return PP_OK;
}

}  // extern "C"
