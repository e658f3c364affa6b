// pp_persist.hip -- batched persistent form of the shell analysis.
//
// Pipeline for one pp_analyse_peaks call (all on the context stream):
//   1. k_transpose: dT(x,y,z) = delta(z,x,y), so the reference's permuted read
//      delta[iv3, iv1, iv2] (RadialShell.jl:448) coalesces like the others.
//   2. peaks are ordered by finalised-shell count m, largest first (already
//      the case for the peak finder's output, else a radix sort).
//   3. tiny peaks (m <= tiny_m): k_tiny, one lane per peak: gather, Fbar,
//      virialisation walk and finish, all lane-serial (64 peaks per wave).
//   4. other peaks: k_gather, a persistent work-queue kernel (one wave per
//      peak, peaks pulled from an atomic counter in m order).  The peak's
//      sphere is read as octant groups (|di|, |dj|, |dk|) in row-major order,
//      so the lanes of a wave read contiguous x-runs of 5 fields (delta, dT,
//      eta x/y/z); each group's <= 8 images are summed in registers and
//      deposited into per-shell LDS bins with Neumaier-compensated FP32
//      adds.  The host packs the groups into 64-lane batches in which no two
//      groups share a shell, so the deposits need no atomics and the sums
//      are deterministic.  The finished profile (16 floats per shell) goes to
//      global memory.
//   5. k_post_lane: one lane per peak runs the Float64 post-process
//      (Fbar, fcrit crossing, walk with strain eigenvalues and collapse-table
//      lookups, R_TH, Srb, zvir_half) on that profile.
//   6. k_finish: one wave per collapsed peak sums eta, eta2 over the shells
//      inside m0 and lapd over the cells inside R_TH (the only cells those
//      outputs use) and writes the mask.
// Peaks whose profile would not fit the LDS bins use the plain kernel.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "pp_shell.h"

namespace pp {

constexpr int PR = 16;  // floats per shell record of the AoS profile
enum : int { P_F = 0, P_N = 1, P_G = 2, P_GF = 5, P_E = 8, P_NB = 14, P_FB = 14 /* Fbar: double in slots 14-15 */ };

// ------------------------------------------------------------- transpose
__global__ void __launch_bounds__(256) k_transpose(const float *__restrict__ in, float *__restrict__ out, int n)
{
    __shared__ float tile[32][33];
    const int y = blockIdx.z, x0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long nn = (long)n * n;
    for (int r = ty; r < 32; r += 8) {
        const int x = x0 + r, z = z0 + tx;
        if (x < n && z < n) tile[r][tx] = in[z + (long)x * n + (long)y * nn];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int z = z0 + r, x = x0 + tx;
        if (x < n && z < n) out[x + (long)y * n + (long)z * nn] = tile[tx][r];
    }
}

// ---------------------------------------------------------- ordering
__global__ void k_keys(const PeakParam *__restrict__ pp, int64_t np, int32_t *__restrict__ keys, int *__restrict__ unsorted)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const int m = pp[p].m;
        keys[p] = m;
        if (p + 1 < np && m < pp[p + 1].m) atomicOr(unsorted, 1);
    }
}

__global__ void k_iota(int32_t *__restrict__ o, int64_t np)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
        o[p] = (int32_t)p;
}

__global__ void k_permute(const PeakParam *__restrict__ pp, const int32_t *__restrict__ order, int64_t np,
                          PeakParam *__restrict__ out, int32_t *__restrict__ m_sorted)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const PeakParam q = pp[order[p]];
        out[p] = q;
        m_sorted[p] = q.m;
    }
}

__global__ void k_hist_m(const int32_t *__restrict__ m, int64_t np, int nbins, int *__restrict__ hist)
{
    extern __shared__ int h[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[min(m[p], nbins - 1)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ void k_m_to_i64(const int32_t *__restrict__ m, int64_t n, int64_t *__restrict__ out)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = m[t];
}

__global__ void k_fill_i32b(int *p, int64_t n, int v)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

// ---------------------------------------------------------- tiny peaks
// Lane-private profiles in an interleaved scratch: element (field f, shell s)
// of lane L lives at f*TM*NL + (s-1)*NL + L, so a wave's lanes, which walk
// the shells in step, store contiguously.
struct LaneProfile {
    const float *f;
    const double *fb;
    const double *radt;
    int TM;
    int64_t NL, L;
    bool lapd;
    __device__ float at(int fld, int s) const { return f[((int64_t)fld * TM + (s - 1)) * NL + L]; }
    __device__ double rad(int s) const { return radt[s - 1]; }
    __device__ double Fbar(int s) const { return fb[(int64_t)(s - 1) * NL + L]; }
    __device__ double count(int s) const { return (double)at(PF_N, s); }
    __device__ double S(int s, int q) const { return (double)at(PF_S + q, s); }
    __device__ double S2(int s, int q) const { return (double)at(PF_S2 + q, s); }
    __device__ double lap(int s) const { return (double)at(PF_L, s); }
    __device__ bool has_lapd() const { return lapd; }
    __device__ void strain_inputs(int s, double G[3], double Gf[3], double Es[6]) const
    {
        if (s == 1) {
            for (int i = 0; i < 3; i++) G[i] = Gf[i] = 0.0;
            for (int i = 0; i < 6; i++) Es[i] = 0.0;
            return;
        }
        const double r = radt[s - 1];
        for (int i = 0; i < 3; i++) {
            G[i] = (double)at(PF_G + i, s) / r;
            Gf[i] = (double)at(PF_GF + i, s) / r;
        }
        double SR[9];
        for (int i = 0; i < 9; i++) SR[i] = (double)at(PF_SR + i, s) / r;
        Es[0] = SR[0] + SR[0]; Es[1] = SR[4] + SR[4]; Es[2] = SR[8] + SR[8];
        Es[3] = SR[1] + SR[3]; Es[4] = SR[2] + SR[6]; Es[5] = SR[5] + SR[7];
    }
};

__global__ void __launch_bounds__(64) k_tiny(GatherFields F, const PeakParam *__restrict__ pps, int64_t lo, int64_t hi,
                                             float *__restrict__ scratch, double *__restrict__ fbs, int TM, PostConst C,
                                             ResultView R)
{
    const int64_t NL = (int64_t)gridDim.x * blockDim.x;
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const long n1 = F.n, n12 = (long)F.n * F.n;
    for (int64_t idx = lo + L; idx < hi; idx += NL) {
        const PeakParam pk = pps[idx];
        const int m = pk.m;
        const bool inside = pk.flags & 1;
        auto put = [&](int fld, int s, float v) { scratch[((int64_t)fld * TM + (s - 1)) * NL + L] = v; };
        {
            const long c = pk.ci + pk.cj * n1 + pk.ck * n12;
            put(PF_N, 1, 1.0f);
            put(PF_F, 1, F.d[c]);
            put(PF_S + 0, 1, F.ex[c]);
            put(PF_S + 1, 1, F.ey[c]);
            put(PF_S + 2, 1, F.ez[c]);
            put(PF_S2 + 0, 1, F.e2x ? F.e2x[c] : 0.0f);
            put(PF_S2 + 1, 1, F.e2x ? F.e2y[c] : 0.0f);
            put(PF_S2 + 2, 1, F.e2x ? F.e2z[c] : 0.0f);
            for (int q = PF_G; q < PF_NFIELD; q++) put(q, 1, 0.0f);
        }
        for (int s = 2; s <= m; s++) {
            float a[PF_NFIELD];
#pragma unroll
            for (int q = 0; q < PF_NFIELD; q++) a[q] = 0.0f;
            const int c0 = C.shell_start[s - 1], c1 = C.shell_start[s];
            for (int c = c0; c < c1; c++) {
                const int pc = C.cell[c];
                const int di = (pc & 255) - 128, dj = ((pc >> 8) & 255) - 128, dk = ((pc >> 16) & 255) - 128;
                const int x = pk.ci + di, y = pk.cj + dj, z = pk.ck + dk;
                if (!inside && (x < 0 || x >= F.n || y < 0 || y >= F.n || z < 0 || z >= F.n)) continue;
                const long a0 = x + y * n1 + z * n12;
                const float d = F.d[a0], dT = F.d[z + x * n1 + y * n12];
                const float ex = F.ex[a0], ey = F.ey[a0], ez = F.ez[a0];
                const float fx = (float)di, fy = (float)dj, fz = (float)dk;
                a[PF_N] += 1.0f;
                a[PF_F] += d;
                a[PF_S + 0] += ex; a[PF_S + 1] += ey; a[PF_S + 2] += ez;
                if (F.e2x) { a[PF_S2 + 0] += F.e2x[a0]; a[PF_S2 + 1] += F.e2y[a0]; a[PF_S2 + 2] += F.e2z[a0]; }
                a[PF_G + 0] += d * fx; a[PF_G + 1] += d * fy; a[PF_G + 2] += d * fz;
                a[PF_GF + 0] += dT * fx; a[PF_GF + 1] += dT * fy; a[PF_GF + 2] += dT * fz;
                a[PF_SR + 0] += ex * fx; a[PF_SR + 1] += ex * fy; a[PF_SR + 2] += ex * fz;
                a[PF_SR + 3] += ey * fx; a[PF_SR + 4] += ey * fy; a[PF_SR + 5] += ey * fz;
                a[PF_SR + 6] += ez * fx; a[PF_SR + 7] += ez * fy; a[PF_SR + 8] += ez * fz;
                if (F.lap) a[PF_L] += F.lap[a0];
            }
#pragma unroll
            for (int q = 0; q < PF_NFIELD; q++) put(q, s, a[q]);
        }
        // Fbar recursion (RadialShell.jl:383-390)
        {
            double F1 = (double)scratch[((int64_t)PF_F * TM) * NL + L];
            fbs[L] = F1;
            double fbp = F1, dFbarp = F1;
            for (int s = 2; s <= m; s++) {
                const double nn = (double)scratch[((int64_t)PF_N * TM + (s - 1)) * NL + L];
                const double dF = nn > 0.0 ? (double)scratch[((int64_t)PF_F * TM + (s - 1)) * NL + L] / nn : 0.0;
                const double rp = C.rad[s - 2], r = C.rad[s - 1];
                const double r3p = rp * rp * rp, r3 = r * r * r;
                fbp = (r3p * fbp + 0.5 * (dFbarp + dF) * (r3 - r3p)) / r3;
                fbs[(int64_t)(s - 1) * NL + L] = fbp;
                dFbarp = dF;
            }
        }
        LaneProfile P{scratch, fbs, C.rad, TM, NL, L, F.lap != nullptr};
        analyse_post<false, true>(P, pk, C, R, pk.order);
        R.nshells[pk.order] = m;
    }
}

// ----------------------------------------------------------- wave gather
struct GroupDesc {
    int32_t lmn;  // l | dj << 8 | dk << 16
    int32_t s;    // shell (1-based), -1 for an empty slot
};

struct GatherArgs {
    GatherFields F;
    const PeakParam *pps;
    int64_t lo, hi;
    const GroupDesc *desc;     // all batch tables
    const int64_t *tab_off;    // per m: first descriptor of its batch table
    const int32_t *tab_nb;     // per m: number of 64-slot batches
    float *prof;
    const int64_t *prof_off;   // per sorted index: first shell record
    int *queue;
    int mcap;
};

__device__ __forceinline__ void neumaier(float *hi, float *lo, float x)
{
    const float h = *hi;
    const float t = h + x;
    const float e = fabsf(h) >= fabsf(x) ? (h - t) + x : (x - t) + h;
    *hi = t;
    *lo += e;
}

__global__ void __launch_bounds__(64) k_gather(GatherArgs A)
{
    extern __shared__ float lds[];
    const int lane = threadIdx.x;
    const int mcap = A.mcap;
    float *bh = lds, *bl = lds + P_NB * mcap;
    const GatherFields &F = A.F;
    const long n1 = F.n, n12 = (long)F.n * F.n;
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(A.queue, 1);
        it = __shfl(it, 0, 64);
        const int64_t idx = A.lo + it;
        if (idx >= A.hi) break;
        const PeakParam pk = A.pps[idx];
        const int m = pk.m;
        const bool inside = pk.flags & 1;
        for (int f = 0; f < P_NB; f++)
            for (int s2 = lane; s2 < m - 1; s2 += 64) { bh[f * mcap + s2] = 0.0f; bl[f * mcap + s2] = 0.0f; }
        __syncthreads();
        const GroupDesc *tab = A.desc + A.tab_off[m];
        const int nb = A.tab_nb[m];
        for (int b = 0; b < nb; b++) {
            const GroupDesc g = tab[b * 64 + lane];
            float acc[P_NB];
#pragma unroll
            for (int f = 0; f < P_NB; f++) acc[f] = 0.0f;
            if (g.s >= 2) {
                const int l = g.lmn & 255, dj = (g.lmn >> 8) & 255, dk = (g.lmn >> 16) & 255;
#pragma unroll
                for (int img = 0; img < 8; img++) {
                    const int sx = (img & 1) ? -1 : 1, sy = (img & 2) ? -1 : 1, sz = (img & 4) ? -1 : 1;
                    if ((sx < 0 && l == 0) || (sy < 0 && dj == 0) || (sz < 0 && dk == 0)) continue;
                    const int dx = sx * l, dy = sy * dj, dz = sz * dk;
                    const int x = pk.ci + dx, y = pk.cj + dy, z = pk.ck + dz;
                    if (!inside && (x < 0 || x >= F.n || y < 0 || y >= F.n || z < 0 || z >= F.n)) continue;
                    const long a = x + y * n1 + z * n12;
                    const float d = F.d[a], dT = F.dT[a], ex = F.ex[a], ey = F.ey[a], ez = F.ez[a];
                    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
                    acc[P_F] += d;
                    acc[P_N] += 1.0f;
                    acc[P_G + 0] += d * fx; acc[P_G + 1] += d * fy; acc[P_G + 2] += d * fz;
                    acc[P_GF + 0] += dT * fx; acc[P_GF + 1] += dT * fy; acc[P_GF + 2] += dT * fz;
                    acc[P_E + 0] += 2.0f * (ex * fx);
                    acc[P_E + 1] += 2.0f * (ey * fy);
                    acc[P_E + 2] += 2.0f * (ez * fz);
                    acc[P_E + 3] += ex * fy + ey * fx;
                    acc[P_E + 4] += ex * fz + ez * fx;
                    acc[P_E + 5] += ey * fz + ez * fy;
                }
                const int s2 = g.s - 2;
#pragma unroll
                for (int f = 0; f < P_NB; f++) neumaier(&bh[f * mcap + s2], &bl[f * mcap + s2], acc[f]);
            }
        }
        __syncthreads();
        float *rec = A.prof + A.prof_off[idx] * PR;
        for (int s = 1 + lane; s <= m; s += 64) {
            float *r = rec + (int64_t)(s - 1) * PR;
            if (s == 1) {
                r[P_F] = F.d[pk.ci + pk.cj * n1 + pk.ck * n12];
                r[P_N] = 1.0f;
                for (int f = 2; f < P_NB; f++) r[f] = 0.0f;
            } else {
                for (int f = 0; f < P_NB; f++) r[f] = bh[f * mcap + s - 2] + bl[f * mcap + s - 2];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ lane post
struct AosProfile {
    float *rec;
    const double *radt;
    __device__ double rad(int s) const { return radt[s - 1]; }
    __device__ double Fbar(int s) const { return *(const double *)(rec + (int64_t)(s - 1) * PR + P_FB); }
    __device__ double count(int s) const { return (double)rec[(int64_t)(s - 1) * PR + P_N]; }
    __device__ double S(int, int) const { return 0.0; }
    __device__ double S2(int, int) const { return 0.0; }
    __device__ double lap(int) const { return 0.0; }
    __device__ bool has_lapd() const { return false; }
    __device__ void strain_inputs(int s, double G[3], double Gf[3], double Es[6]) const
    {
        if (s == 1) {
            for (int i = 0; i < 3; i++) G[i] = Gf[i] = 0.0;
            for (int i = 0; i < 6; i++) Es[i] = 0.0;
            return;
        }
        const float *r = rec + (int64_t)(s - 1) * PR;
        const double inv = 1.0 / radt[s - 1];
        const double rr = radt[s - 1];
        (void)inv;
        for (int i = 0; i < 3; i++) {
            G[i] = (double)r[P_G + i] / rr;
            Gf[i] = (double)r[P_GF + i] / rr;
        }
        for (int i = 0; i < 6; i++) Es[i] = (double)r[P_E + i] / rr;
    }
};

__global__ void __launch_bounds__(64) k_post_lane(const PeakParam *__restrict__ pps, int64_t lo, int64_t hi,
                                                  float *__restrict__ prof, const int64_t *__restrict__ prof_off,
                                                  PostConst C, ResultView R)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < hi; idx += stride) {
        const PeakParam pk = pps[idx];
        const int m = pk.m;
        float *rec = prof + prof_off[idx] * PR;
        {   // Fbar recursion (RadialShell.jl:383-390) into slots 14-15
            double fb = (double)rec[P_F];
            *(double *)(rec + P_FB) = fb;
            double dFbarp = fb;
            for (int s = 2; s <= m; s++) {
                const float *r = rec + (int64_t)(s - 1) * PR;
                const double nn = (double)r[P_N];
                const double dF = nn > 0.0 ? (double)r[P_F] / nn : 0.0;
                const double rp = C.rad[s - 2], rs = C.rad[s - 1];
                const double r3p = rp * rp * rp, r3 = rs * rs * rs;
                fb = (r3p * fb + 0.5 * (dFbarp + dF) * (r3 - r3p)) / r3;
                *(double *)(rec + (int64_t)(s - 1) * PR + P_FB) = fb;
                dFbarp = dF;
            }
        }
        AosProfile P{rec, C.rad};
        analyse_post<false, false>(P, pk, C, R, pk.order);
        R.nshells[pk.order] = m;
    }
}

// ------------------------------------------------------------ finish
__global__ void __launch_bounds__(64) k_finish(GatherFields F, const PeakParam *__restrict__ pps, int64_t lo, int64_t hi,
                                               PostConst C, ResultView R)
{
    const float *ex[3] = {F.ex, F.ey, F.ez};
    const float *e2[3] = {F.e2x, F.e2y, F.e2z};
    for (int64_t idx = lo + blockIdx.x; idx < hi; idx += gridDim.x) {
        const PeakParam pk = pps[idx];
        const int64_t o = pk.order;
        if (R.status[o] != DUMP_NONE) continue;
        finish_cells<true>(C, pk, R.m0[o], R.q[o], ex, e2, R, o);
    }
}

// ------------------------------------------------------------- host side
// Octant groups of shells 2..m in row-major order packed into 64-slot batches
// with no shell repeated inside a batch (greedy over a window of open
// batches, which keeps row-major locality).
static void pack_groups(const ShellTable &T, int m, std::vector<GroupDesc> &out, int &nbatch)
{
    const int r2m = T.shell_r2[m - 1];
    const int R = (int)std::floor(std::sqrt((double)r2m));
    std::vector<int32_t> shell_of_r2((size_t)r2m + 1, 0);
    for (int s = 1; s <= m; s++) shell_of_r2[T.shell_r2[s - 1]] = s;
    struct B { std::vector<GroupDesc> g; std::vector<int32_t> shells; };
    std::vector<B> bs;
    const size_t K = 8;  // open batches considered
    for (int dk = 0; dk <= R; dk++)
        for (int dj = 0; dj <= R; dj++)
            for (int l = 0; l <= R; l++) {
                const int r2 = l * l + dj * dj + dk * dk;
                if (r2 == 0 || r2 > r2m) continue;
                const int s = shell_of_r2[r2];
                GroupDesc g{l | (dj << 8) | (dk << 16), s};
                bool placed = false;
                for (size_t b = bs.size() > K ? bs.size() - K : 0; b < bs.size() && !placed; b++) {
                    B &bb = bs[b];
                    if (bb.g.size() >= 64) continue;
                    if (std::find(bb.shells.begin(), bb.shells.end(), s) != bb.shells.end()) continue;
                    bb.g.push_back(g);
                    bb.shells.push_back(s);
                    placed = true;
                }
                if (!placed) { bs.emplace_back(); bs.back().g.push_back(g); bs.back().shells.push_back(s); }
            }
    nbatch = (int)bs.size();
    out.assign((size_t)nbatch * 64, GroupDesc{0, -1});
    for (int b = 0; b < nbatch; b++)
        for (size_t q = 0; q < bs[b].g.size(); q++) out[(size_t)b * 64 + q] = bs[b].g[q];
}

struct PersistState {
    int rmax = -1;
    std::map<int, std::pair<int64_t, int>> tables;  // m -> (descriptor offset, nbatch)
    std::vector<GroupDesc> desc;
    DevBuf d_desc, d_off, d_nb;
    bool dirty = true;
};

static std::map<pp_ctx *, PersistState> g_state;
static std::mutex g_state_mu;

static PersistState &state_of(pp_ctx *ctx)
{
    std::lock_guard<std::mutex> lk(g_state_mu);
    return g_state[ctx];
}

static int ensure_tables(pp_ctx *ctx, const std::vector<int> &ms)
{
    PersistState &st = state_of(ctx);
    const ShellTable &T = ctx->shells;
    if (st.rmax != T.rmax) { st = PersistState(); st.rmax = T.rmax; }
    for (int m : ms) {
        if (m < 2 || st.tables.count(m)) continue;
        std::vector<GroupDesc> d;
        int nb = 0;
        pack_groups(T, m, d, nb);
        st.tables[m] = {(int64_t)st.desc.size(), nb};
        st.desc.insert(st.desc.end(), d.begin(), d.end());
        st.dirty = true;
    }
    if (st.dirty) {
        const int ns = T.nshells + 1;
        std::vector<int64_t> off(ns, 0);
        std::vector<int32_t> nb(ns, 0);
        for (auto &kv : st.tables) { off[kv.first] = kv.second.first; nb[kv.first] = kv.second.second; }
        PP_CHECK(st.d_desc.grow(std::max<size_t>(1, st.desc.size()) * sizeof(GroupDesc)));
        PP_CHECK(st.d_off.grow(ns * sizeof(int64_t)));
        PP_CHECK(st.d_nb.grow(ns * sizeof(int32_t)));
        if (!st.desc.empty())
            PP_HIP(hipMemcpy(st.d_desc.ptr, st.desc.data(), st.desc.size() * sizeof(GroupDesc), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(st.d_off.ptr, off.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(st.d_nb.ptr, nb.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice));
        st.dirty = false;
    }
    return PP_OK;
}

static int grid1d(pp_ctx *ctx, int64_t n, int per_cu = 8)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cu * per_cu));
}

// Ordered params + histogram of m; returns host histogram.
static int order_peaks(pp_ctx *ctx, ShellRun &S, PeakParam *&pps, int32_t *&m_sorted, std::vector<int> &hist)
{
    const int64_t np = S.np;
    const int nb = ctx->shells.nshells + 1;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    PP_CHECK(ctx->d_queue.grow(al(np * sizeof(PeakParam)) + al(np * 4) * 3 + al(nb * 4) + 4096));
    char *p = ctx->d_queue.as<char>();
    pps = (PeakParam *)p; p += al(np * sizeof(PeakParam));
    m_sorted = (int32_t *)p; p += al(np * 4);
    int32_t *order2 = (int32_t *)p; p += al(np * 4);
    int32_t *keys2 = (int32_t *)p; p += al(np * 4);
    int *d_hist = (int *)p; p += al(nb * 4);
    int *d_flag = (int *)p;
    hipLaunchKernelGGL(k_fill_i32b, dim3(1), dim3(64), 0, ctx->stream, d_flag, (int64_t)1, 0);
    hipLaunchKernelGGL(k_keys, dim3(grid1d(ctx, np)), dim3(256), 0, ctx->stream, (const PeakParam *)S.pp, np, S.keys, d_flag);
    hipLaunchKernelGGL(k_iota, dim3(grid1d(ctx, np)), dim3(256), 0, ctx->stream, S.order, np);
    PP_HIP(hipGetLastError());
    int unsorted = 0;
    PP_HIP(hipMemcpyAsync(&unsorted, d_flag, 4, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    const int32_t *order = S.order;
    if (unsorted) {
        size_t tmp = 0;
        PP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, S.keys, keys2, S.order, order2, (int)np, 0, 16,
                                                            ctx->stream));
        PP_CHECK(ctx->d_cub.grow(tmp));
        PP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(ctx->d_cub.ptr, tmp, S.keys, keys2, S.order, order2, (int)np,
                                                            0, 16, ctx->stream));
        order = order2;
    }
    hipLaunchKernelGGL(k_permute, dim3(grid1d(ctx, np)), dim3(256), 0, ctx->stream, (const PeakParam *)S.pp, order, np,
                       pps, m_sorted);
    hipLaunchKernelGGL(k_fill_i32b, dim3(grid1d(ctx, nb)), dim3(256), 0, ctx->stream, d_hist, (int64_t)nb, 0);
    hipLaunchKernelGGL(k_hist_m, dim3(grid1d(ctx, np, 4)), dim3(256), nb * sizeof(int), ctx->stream,
                       (const int32_t *)m_sorted, np, nb, d_hist);
    PP_HIP(hipGetLastError());
    hist.assign(nb, 0);
    PP_HIP(hipMemcpyAsync(hist.data(), d_hist, nb * 4, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

static int lds_floats_per_shell() { return 2 * P_NB; }

// The persistent pipeline on an ordered batch.  export_prof keeps the profile
// scratch for pp_shell_profiles (all peaks then go through k_gather).
static int run_persistent_impl(pp_ctx *ctx, GatherFields G, ShellRun &S, PostConst &C, bool export_all,
                               std::vector<int32_t> *cnt, std::vector<double> *fbar, std::vector<int64_t> *offs)
{
    const int64_t np = S.np;
    const int n = G.n;
    // 1. transposed delta
    PP_CHECK(ctx->d_field_b.grow((size_t)n * n * n * sizeof(float)));
    {
        dim3 grid((n + 31) / 32, (n + 31) / 32, n), block(32, 8);
        hipLaunchKernelGGL(k_transpose, grid, block, 0, ctx->stream, G.d, ctx->d_field_b.as<float>(), n);
        PP_HIP(hipGetLastError());
        G.dT = ctx->d_field_b.as<float>();
    }
    // 2. order by m (descending)
    PeakParam *pps = nullptr;
    int32_t *m_sorted = nullptr;
    std::vector<int> hist;
    PP_CHECK(order_peaks(ctx, S, pps, m_sorted, hist));
    const int nb = (int)hist.size();
    // classes: tiny (m <= tiny_m, lane per peak); LDS classes (m in (b/2, b]
    // for power-of-two b, one persistent launch each, largest first); m above
    // the LDS capacity: plain kernel.
    const int tiny_m = export_all ? 0 : (ctx->tiny_m > 0 ? ctx->tiny_m : 16);
    const size_t lds_max = std::min<size_t>(ctx->lds_per_block, 160 * 1024);
    const int mcap_max = (int)(lds_max / (lds_floats_per_shell() * sizeof(float)));
    std::vector<int64_t> above(nb + 1, 0);  // above[x] = #peaks with m >= x
    for (int x = nb - 1; x >= 0; x--) above[x] = above[x + 1] + hist[x];
    auto cnt_gt = [&](int x) -> int64_t { return x < 0 ? np : above[std::min(x + 1, nb)]; };
    const int64_t big_end = cnt_gt(mcap_max);   // sorted [0, big_end): beyond the LDS bins
    const int64_t tiny_lo = cnt_gt(tiny_m);     // sorted [tiny_lo, np): tiny
    std::vector<int> ms;
    int mtop = 0;
    for (int x = 2; x < nb; x++)
        if (hist[x] && x > tiny_m && x <= mcap_max) { ms.push_back(x); mtop = x; }
    PP_CHECK(ensure_tables(ctx, ms));
    PersistState &st = state_of(ctx);
    // profile offsets over the gathered range [big_end, tiny_lo): exclusive
    // scan of m (int64)
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const int64_t ng = tiny_lo - big_end;
    int64_t tot = 0;
    PP_CHECK(ctx->d_scan.grow(al(std::max<int64_t>(ng, 1) * 8) * 2 + 4096));
    int64_t *d_m64 = ctx->d_scan.as<int64_t>();
    int64_t *d_off = (int64_t *)(ctx->d_scan.as<char>() + al(std::max<int64_t>(ng, 1) * 8));
    if (ng > 0) {
        hipLaunchKernelGGL(k_m_to_i64, dim3(grid1d(ctx, ng)), dim3(256), 0, ctx->stream,
                           (const int32_t *)(m_sorted + big_end), ng, d_m64);
        size_t tmp = 0;
        PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_m64, d_off, (int)ng, ctx->stream));
        PP_CHECK(ctx->d_cub.grow(tmp));
        PP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_cub.ptr, tmp, d_m64, d_off, (int)ng, ctx->stream));
        int64_t last[2];
        PP_HIP(hipMemcpyAsync(&last[0], d_off + ng - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipMemcpyAsync(&last[1], d_m64 + ng - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        tot = last[0] + last[1];
    }
    // (ctx->d_prof belongs to the plain kernel, which step 3 may run)
    PP_CHECK(ctx->d_res.grow((size_t)std::max<int64_t>(tot, 1) * PR * sizeof(float) + 4096));
    float *d_rec = ctx->d_res.as<float>();
    int *d_q = (int *)(ctx->d_misc.as<char>() + 1024);
    PP_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    // 3. peaks beyond the LDS cap: plain kernel on their ordered params
    if (big_end > 0) {
        ShellRun S2 = S;
        S2.np = big_end;
        S2.pp = pps;
        PP_CHECK(run_plain(ctx, G, S2, C, nullptr));
    }
    // 4. persistent gather per LDS class (descending m)
    if (ng > 0) {
        int bcap = 32;
        while (bcap < mtop) bcap <<= 1;
        int cls = 0;
        for (int b = std::min(bcap, mcap_max);; ) {
            const int blow = std::max(tiny_m, b / 2);
            const int64_t lo = std::max(big_end, cnt_gt(b)), hi = std::min(tiny_lo, cnt_gt(blow));
            if (hi > lo) {
                const size_t lds = (size_t)b * lds_floats_per_shell() * sizeof(float);
                const int waves_cu = std::max(1, std::min<int>(ctx->gather_waves_cu > 0 ? ctx->gather_waves_cu : 16,
                                                               (int)(lds_max / lds)));
                const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(hi - lo, (int64_t)ctx->num_cu * waves_cu));
                hipLaunchKernelGGL(k_fill_i32b, dim3(1), dim3(64), 0, ctx->stream, d_q + cls, (int64_t)1, 0);
                GatherArgs A;
                A.F = G; A.pps = pps; A.lo = lo; A.hi = hi;
                A.desc = st.d_desc.as<GroupDesc>(); A.tab_off = st.d_off.as<int64_t>(); A.tab_nb = st.d_nb.as<int32_t>();
                A.prof = d_rec; A.prof_off = d_off - big_end; A.queue = d_q + cls; A.mcap = b;
                if (lds > 64 * 1024)
                    (void)hipFuncSetAttribute((const void *)k_gather, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                hipLaunchKernelGGL(k_gather, dim3(grid), dim3(64), lds, ctx->stream, A);
                PP_HIP(hipGetLastError());
                cls++;
            }
            if (blow <= tiny_m || cls >= 200) break;
            b = blow;
        }
    }
    PP_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    // 5. tiny peaks, one lane each
    if (tiny_lo < np) {
        const int TM = std::max(tiny_m, 2);
        const int64_t lanes = std::min<int64_t>(((np - tiny_lo) + 63) / 64 * 64, (int64_t)ctx->num_cu * 64 * 16);
        PP_CHECK(ctx->d_k2.grow((size_t)lanes * TM * (PF_NFIELD * 4 + 8) + 4096));
        float *scr = ctx->d_k2.as<float>();
        double *fbs = (double *)(ctx->d_k2.as<char>() + al((size_t)lanes * TM * PF_NFIELD * 4));
        hipLaunchKernelGGL(k_tiny, dim3((unsigned)(lanes / 64)), dim3(64), 0, ctx->stream, G, (const PeakParam *)pps,
                           tiny_lo, np, scr, fbs, TM, C, S.R);
        PP_HIP(hipGetLastError());
    }
    // 6. lane post + 7. finish over the gathered range
    if (tiny_lo > big_end) {
        const int64_t cnt_g = tiny_lo - big_end;
        const int gp = (int)std::max<int64_t>(1, std::min<int64_t>((cnt_g + 63) / 64, (int64_t)ctx->num_cu * 16));
        hipLaunchKernelGGL(k_post_lane, dim3(gp), dim3(64), 0, ctx->stream, (const PeakParam *)pps, big_end, tiny_lo,
                           d_rec, (const int64_t *)(d_off - big_end), C, S.R);
        PP_HIP(hipGetLastError());
        const int gf = (int)std::max<int64_t>(1, std::min<int64_t>(cnt_g, (int64_t)ctx->num_cu * 16));
        hipLaunchKernelGGL(k_finish, dim3(gf), dim3(64), 0, ctx->stream, G, (const PeakParam *)pps, big_end, tiny_lo, C,
                           S.R);
        PP_HIP(hipGetLastError());
    }
    if (cnt) {
        // profile export (all peaks gathered): counts and Fbar per shell in the
        // caller's order
        std::vector<float> hr((size_t)tot * PR);
        PP_HIP(hipMemcpyAsync(hr.data(), d_rec, hr.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        std::vector<PeakParam> hpp(np);
        PP_HIP(hipMemcpyAsync(hpp.data(), pps, np * sizeof(PeakParam), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (big_end != 0 || tiny_lo != np) { set_error("profile export: every peak must fit the LDS bins"); return PP_ERR_ARG; }
        std::vector<int64_t> start_of(np);
        int64_t t2 = 0;
        for (int64_t i = 0; i < np; i++) { start_of[hpp[i].order] = t2; t2 += hpp[i].m; }
        cnt->resize(tot);
        fbar->resize(tot);
        offs->assign(np, 0);
        int64_t o2 = 0;
        for (int64_t i = 0; i < np; i++) {
            const int m = hpp[i].m;
            (*offs)[hpp[i].order] = start_of[hpp[i].order];
            for (int s = 0; s < m; s++) {
                const float *r = &hr[(size_t)(o2 + s) * PR];
                (*cnt)[start_of[hpp[i].order] + s] = (int32_t)r[P_N];
                double fb;
                memcpy(&fb, r + P_FB, 8);
                (*fbar)[start_of[hpp[i].order] + s] = fb;
            }
            o2 += m;
        }
    }
    float gms = 0;
    PP_HIP(hipEventSynchronize(ctx->ev[3]));
    (void)hipEventElapsedTime(&gms, ctx->ev[2], ctx->ev[3]);
    ctx->timing.gather_ms = gms;
    return PP_OK;
}

int run_persistent(pp_ctx *ctx, GatherFields G, ShellRun &S, PostConst &C)
{
    return run_persistent_impl(ctx, G, S, C, false, nullptr, nullptr, nullptr);
}

int persistent_profiles(pp_ctx *ctx, GatherFields G, ShellRun &S, PostConst &C, std::vector<int32_t> &cnt,
                        std::vector<double> &fbar, std::vector<int64_t> &off)
{
    return run_persistent_impl(ctx, G, S, C, true, &cnt, &fbar, &off);
}

int autotune_persistent(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks, double alatt,
                        double rmax2rs, int *choice)
{
    (void)ctx; (void)fields; (void)peaks; (void)alatt; (void)rmax2rs;
    if (choice) *choice = 0;
    set_error("autotuner not available in this build");
    return PP_ERR_STATE;
}

int shell_groups_build(pp_ctx *) { return PP_OK; }

}  // namespace pp
