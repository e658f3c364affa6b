// pp_persist.hip -- batched persistent form of the shell analysis.
//
// Pipeline for one pp_analyse_peaks call (on the context stream; the packing
// runs on the context's second stream beside the ordering, and after the wave
// gather the small peaks' gather, post and finish run there beside the lane
// post and the finish pass, and are joined back before the call returns):
//   1. k_pack4: (delta, dT, eta x, eta y) per cell as one float4, with
//      dT(x,y,z) = delta(z,x,y), so the reference's permuted read
//      delta[iv3, iv1, iv2] (RadialShell.jl:448) coalesces like the others;
//      k_pack_q4: (eta z, eta2 x, eta2 y, eta2 z) the same way.
//   2. peaks are ordered by finalised-shell count m, largest first (already
//      the case for the peak finder's output, else a radix sort), and every
//      peak gets a profile of m 24-float shell records (exclusive scan of m).
//   3. gather.  A peak's sphere is read as octant groups (|di|, |dj|, |dk|):
//      one lane takes a group and sums its <= 8 mirror images in registers.
//      * small peaks (m <= small_m): k_small reads floor(64 / m) peaks of
//        equal m cell by cell in row order, hands each cell to the lane of its
//        (peak, shell) position by a wave shuffle and sums runs of equal shell
//        by a segmented tree into LDS bins.
//      * other peaks: k_gather, a persistent work-queue kernel (one wave per
//        peak, peaks pulled from per-XCD atomic counters).  The shells are
//        processed in chunks of chunk_shells; within a chunk the host packed
//        the groups (row-major, so lanes read neighbouring x-runs) into dense
//        64-lane batches sorted by shell; the lanes of a run of equal shells
//        sum by a segmented tree and its first lane deposits into per-shell
//        LDS bins with Neumaier-compensated FP32 adds, without atomics and in
//        a fixed order.
//   4. Float64 post-process (fcrit crossing, inward walk with strain
//      eigenvalues and collapse-table look-ups, R_TH, Srb, zvir_half):
//      k_post_wave (one wave per peak) for m > coop_m, k_post_lane (one lane
//      per peak) below; the lane post's walk steps are evaluated ahead of it
//      by lane groups (k_walk_group, the walk table).
//   5. k_finish: one wave per collapsed non-small peak sums eta, eta2 over the
//      shells inside m0 and lapd over the cells inside R_TH and writes the
//      mask; small peaks carry those sums in their profile, from which
//      k_small_finish forms them.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>

#include "pp_shell.h"

namespace pp {

constexpr int PR = 24;  // floats per shell record of the AoS profile
// In memory the profile is three arrays of shell records: the strain inputs
// E(6), which every step of the posts' inward walk reads for each shell of its
// band, the gradient inputs G(3) Gf(3), read at the few steps whose gradients
// are kept, and the rest (F, count, Fbar, S, S2, lapd), so a walk step moves
// 24 bytes per shell and its cache lines hold nothing else.  In registers a
// record is v[PR]: the strain part in v[0..11], the rest in v[12..23].
constexpr int PSR = 12;
constexpr int PSE = 6;  // floats per record of the E and the G arrays
struct ProfBuf {
    float *e, *g, *a;  // record r: E at e + 6 r, G Gf at g + 6 r, the rest at a + 12 r
};
// record: G(3) Gf(3) E(6) | F, count | Fbar (double) | S(3) S2(3) lapd | pad;
// the 12 strain inputs load as three 16-byte vectors.  S, S2, lapd are filled
// for small peaks only.
enum : int {
    P_G = 0, P_GF = 3, P_E = 6, P_F = 12, P_N = 13, P_NB = 14 /* fields binned by k_gather */,
    P_FB = 14 /* Fbar: double in slots 14-15 */, P_S = 16, P_S2 = 19, P_L = 22
};
constexpr int SMALL_MAX = 64;  // largest m of the cell-based small-peak gather (its P m bins fit 64 lanes)
// Work placement.  The gathers read the fields around each peak; peaks of
// equal m are taken in Morton order of their centres so that the waves in
// flight share cache lines, and each XCD (its own L2) works through its own
// share of every m: workgroup b runs on XCD b % NXCD (round-robin dispatch).
constexpr int NXCD = 8;

// (delta, delta transposed, eta x, eta y) of every cell as one float4, so the
// gathers read the four with one 16-byte load (and a 4-byte one for eta z):
// a sphere row then spans about half the cache lines of five separate arrays.
// With eta2, (eta z, eta2 x, eta2 y, eta2 z) go to a second float4 array
// (q4, k_pack_q4) for the small-peak gather and the finish pass.  The block
// stages delta(z, x, y) for T x T (x, z) in LDS (the transposed read).
constexpr int kPackTile = 64;  // x and z extent of k_pack4's transpose tile (the fields' row runs are T floats)
template <int T>
__global__ void __launch_bounds__(256) k_pack4(const float *__restrict__ d, const float *__restrict__ ex,
                                               const float *__restrict__ ey, float4 *__restrict__ out, int n)
{
    constexpr int R = 256 / T;  // rows per pass of the block
    __shared__ float tile[T][T + 1];
    const int y = blockIdx.z, x0 = blockIdx.x * T, z0 = blockIdx.y * T;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long nn = (long)n * n;
    for (int r = ty; r < T; r += R) {
        const int x = x0 + r, z = z0 + tx;
        if (x < n && z < n) tile[r][tx] = d[z + (long)x * n + (long)y * nn];
    }
    __syncthreads();
    for (int r = ty; r < T; r += R) {
        const int z = z0 + r, x = x0 + tx;
        if (x < n && z < n) {
            const long a = x + (long)y * n + (long)z * nn;
            // read once and written for later kernels: streamed past the caches
            typedef float v4f __attribute__((ext_vector_type(4)));
            const v4f po = {d[a], tile[tx][r], __builtin_nontemporal_load(ex + a), __builtin_nontemporal_load(ey + a)};
            __builtin_nontemporal_store(po, (v4f *)(out + a));
        }
    }
}

// (eta z, eta2 x, eta2 y, eta2 z) of every cell as one float4, streamed, one
// cell per lane over the whole grid; it follows k_pack4 on the same stream.
// Measured at c3: k_pack4 writing both arrays 1.85 ms; this split 0.3 ms
// less (the same split with a grid-stride loop over 16 workgroups per CU: no
// change).
__global__ void __launch_bounds__(256) k_pack_q4(const float *__restrict__ ez, const float *__restrict__ e2x,
                                                 const float *__restrict__ e2y, const float *__restrict__ e2z,
                                                 float4 *__restrict__ q4, int64_t ncell)
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < ncell; a += stride) {
        const v4f q = {__builtin_nontemporal_load(ez + a), __builtin_nontemporal_load(e2x + a),
                       __builtin_nontemporal_load(e2y + a), __builtin_nontemporal_load(e2z + a)};
        __builtin_nontemporal_store(q, (v4f *)(q4 + a));
    }
}

// ---------------------------------------------------------- ordering
__global__ void k_iota(int32_t *__restrict__ o, int64_t np)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
        o[p] = (int32_t)p;
}

// keys (m), the out-of-order flag and the histogram of m in one pass
// (flag[0]: some peak has a larger m than the one before it)
__global__ void k_keys_hist(const PeakParam *__restrict__ pp, int64_t np, int32_t *__restrict__ keys, int nbins,
                            int *__restrict__ hist, int *__restrict__ flag)
{
    extern __shared__ int h[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    bool up = false;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const int m = pp[p].m;
        keys[p] = m;
        up = up || (p + 1 < np && m < pp[p + 1].m);
        atomicAdd(&h[min(m, nbins - 1)], 1);
    }
    if (up) atomicOr(flag, 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// order: the sorted positions, or null when the batch is already in order
__global__ void k_permute(const PeakParam *__restrict__ pp, const int32_t *__restrict__ order, int64_t np,
                          PeakParam *__restrict__ out, int32_t *__restrict__ m_sorted)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const PeakParam q = pp[order ? order[p] : p];
        out[p] = q;
        m_sorted[p] = q.m;
    }
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

// bits 0..20 of v moved to every third bit (Morton interleave)
__device__ __forceinline__ uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Size class of the wave gather's order: floor(log16 m).  The largest class
// still starts first; the fields around a place
// are read for all its peaks of a class together.  Measured at c3 (classes
// m < 256 and m >= 256 there): per m 6.6 ms, log2 classes 6.0, log4 5.6,
// log16 5.4, one class 6.4 (its largest peaks end last), m >= 1024 first 6.5.
__host__ __device__ __forceinline__ int gather_class(int m)
{
    int c = 0;
    while ((2 << c) <= m) c++;
    return c / 4;
}

// sort key: size class descending, then the Morton code of the centre cell
// (nbits per axis), so peaks processed together lie close in space
__global__ void k_keys_morton(const PeakParam *__restrict__ pp, int64_t np, int nbits, uint64_t *__restrict__ keys)
{
    const uint32_t msk = (1u << nbits) - 1u;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const PeakParam q = pp[p];
        const uint64_t mort = spread3((uint32_t)q.ci & msk) | spread3((uint32_t)q.cj & msk) << 1 |
                              spread3((uint32_t)q.ck & msk) << 2;
        keys[p] = (uint64_t)(0xffff - min(gather_class(q.m), 0xffff)) << (3 * nbits) | mort;
    }
}

// Per-XCD work lists of the entries [0, hi) of src (positions of the ordered
// peaks; the positions themselves if src is null), which runs of equal group
// g follow each other in (g = m, or the size class with by_class): tab row g
// holds the run's first entry, its length, then the list offset of each of
// its nx parts; the run is cut into nx parts of consecutive entries and part
// x is appended to list x.
__global__ void k_xcd_order(const int32_t *__restrict__ m_sorted, const int32_t *__restrict__ src, int64_t hi,
                            const int64_t *__restrict__ tab, int nx, int by_class, int32_t *__restrict__ xorder)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < hi; p += (int64_t)gridDim.x * blockDim.x) {
        const int m = m_sorted[src ? src[p] : p];
        const int64_t *t = tab + (int64_t)(by_class ? gather_class(m) : m) * (2 + NXCD);
        const int64_t r = p - t[0], len = t[1];
        int x = (int)min<int64_t>(nx - 1, r * nx / len);
        while (x + 1 < nx && len * (x + 1) / nx <= r) x++;
        while (x > 0 && len * x / nx > r) x--;
        xorder[t[2 + x] + r - len * x / nx] = src ? src[p] : (int32_t)p;
    }
}

// ----------------------------------------------------------- group sums
struct GroupDesc {
    int32_t lmn;  // l | dj << 8 | dk << 16
    int32_t s;    // shell (1-based), -1 for an empty slot
};

// Sum of one octant group's mirror images (those inside the grid) into acc,
// indexed by record slot.  FULL adds S, S2 and lapd.
template <bool FULL>
__device__ __forceinline__ void group_sum(const GatherFields &F, const PeakParam &pk, GroupDesc g, float acc[PR])
{
    const long n1 = F.n, n12 = (long)F.n * F.n;
    const bool inside = pk.flags & 1;
    const int l = g.lmn & 255, dj = (g.lmn >> 8) & 255, dk = (g.lmn >> 16) & 255;
#pragma unroll
    for (int img = 0; img < 8; img++) {
        const int sx = (img & 1) ? -1 : 1, sy = (img & 2) ? -1 : 1, sz = (img & 4) ? -1 : 1;
        if ((sx < 0 && l == 0) || (sy < 0 && dj == 0) || (sz < 0 && dk == 0)) continue;
        const int dx = sx * l, dy = sy * dj, dz = sz * dk;
        const int x = pk.ci + dx, y = pk.cj + dy, z = pk.ck + dz;
        if (!inside && (x < 0 || x >= F.n || y < 0 || y >= F.n || z < 0 || z >= F.n)) continue;
        const long a = x + y * n1 + z * n12;
        const float4 q = F.p4[a];
        const float d = q.x, dT = q.y, ex = q.z, ey = q.w, ez = F.ez[a];
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
        if constexpr (FULL) {
            acc[P_S + 0] += ex; acc[P_S + 1] += ey; acc[P_S + 2] += ez;
            if (F.e2x) { acc[P_S2 + 0] += F.e2x[a]; acc[P_S2 + 1] += F.e2y[a]; acc[P_S2 + 2] += F.e2z[a]; }
            if (F.lap) acc[P_L] += F.lap[a];
        }
    }
}

// Record r: the strain part and the first naux float4 of the rest
__device__ __forceinline__ void store_record(ProfBuf p, int64_t r, const float v[PR], int naux)
{
    float2 *og = (float2 *)(p.g + r * PSE), *oe = (float2 *)(p.e + r * PSE);
    float4 *oa = (float4 *)(p.a + r * PSR);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        og[q] = make_float2(v[P_G + 2 * q], v[P_G + 2 * q + 1]);
        oe[q] = make_float2(v[P_E + 2 * q], v[P_E + 2 * q + 1]);
    }
    for (int q = 0; q < naux; q++)
        oa[q] = make_float4(v[PSR + 4 * q], v[PSR + 4 * q + 1], v[PSR + 4 * q + 2], v[PSR + 4 * q + 3]);
}

__device__ __forceinline__ void put_fbar(float v[PR], double fb)
{
    float w[2];
    __builtin_memcpy(w, &fb, sizeof fb);
    v[P_FB] = w[0];
    v[P_FB + 1] = w[1];
}

// Centre record (shell 1): count 1, F = delta(centre), Fbar(1) = F; FULL
// adds S, S2 of the centre cell (lapd of shell 1 is not used by d2F, which
// starts at cell 2).  Returns F.
template <bool FULL>
__device__ __forceinline__ float center_values(const GatherFields &F, const PeakParam &pk, float v[PR])
{
    const long c = pk.ci + pk.cj * (long)F.n + pk.ck * (long)F.n * F.n;
#pragma unroll
    for (int f = 0; f < PR; f++) v[f] = 0.0f;
    v[P_F] = F.d[c];
    v[P_N] = 1.0f;
    if constexpr (FULL) {
        v[P_S + 0] = F.ex[c]; v[P_S + 1] = F.ey[c]; v[P_S + 2] = F.ez[c];
        if (F.e2x) { v[P_S2 + 0] = F.e2x[c]; v[P_S2 + 1] = F.e2y[c]; v[P_S2 + 2] = F.e2z[c]; }
    }
    put_fbar(v, (double)v[P_F]);
    return v[P_F];
}

// Inclusive prefix sum over the 64 lanes (fixed pattern, deterministic).
__device__ __forceinline__ double wave_scan(double x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    return x;
}

// ------------------------------------------------------- small peaks
// Small peaks (m <= small_m) are read cell by cell: the cells of shells 1..m
// in row-major order, so each load instruction covers whole x-runs of the
// peaks' rows.  A block handles P = floor(64 / m) peaks of equal m; the host
// cut their P x ncells cells into dense 64-lane batches and sorted every
// batch by (peak, shell).  The lanes of each (peak, shell) run sum their
// cells' 21 contributions by a segmented tree over wave shuffles (step d:
// lane += lane + d while inside the run), and the run's first lane adds the
// sum to that shell's LDS bin.  Records carry S, S2 and lapd, from which
// k_small_finish forms these peaks' Sbar, Sbar2, d2F and mask.
// A lane plays two roles per batch.  Loading, it reads the cell at its
// position in the batch's row-major order (lcell, lq), so a load instruction
// covers whole x-runs of the peaks' rows; its contributions go by a wave
// shuffle to the lane at the cell's position in (peak, shell) order, which sums runs of
// equal (peak, shell) with the tree (cell, q, s, run, rend; src = the lane
// that loaded this position's cell).
struct CellSlot {
    int32_t cell;   // (di+128) | (dj+128) << 8 | (dk+128) << 16; -1 for an empty lane
    int16_t q, s;   // peak slot in the block, shell (1-based)
    int32_t run;    // length of the (q, s) run starting here, 0 if not its first lane
    int32_t rend;   // lane one past the end of this lane's run
    int32_t lcell;  // the cell this lane loads (row-major position); -1 for none
    int16_t lq;     // its peak slot
    int16_t src;    // the lane that loads this (peak, shell)-ordered position's cell
};

struct SmallSeg {
    int32_t m, P, nbat, toff;  // shells, peaks per block, batches per block, first CellSlot
    int64_t lo, hi, blk0;      // sorted peaks [lo, hi); first block
    int32_t steps, pad;        // tree steps: 2^steps >= the longest run of the tables
};

// blocks [blk_lo, blk_lo + count) of segment seg, at position cum of an XCD's
// block list (count = the next part's cum - cum)
struct SmallPart {
    int32_t seg, pad;
    int64_t blk_lo, cum;
};

struct SmallArgs {
    GatherFields F;
    const PeakParam *pps;
    const CellSlot *tab;
    const SmallSeg *seg;
    const SmallPart *part;     // list x: parts [xp[x], xp[x + 1]), xlen[x] blocks
    int32_t xp[NXCD + 1];
    int64_t xlen[NXCD];
    int nx;                    // lists (1 or NXCD); gridDim.x is a multiple of nx
    ProfBuf prof;
    const int64_t *prof_off;
    const double *rad;
};

constexpr int NSF = 21;  // staged fields: record slots 0..13 and 16..22
__device__ __forceinline__ constexpr int sf_slot(int f) { return f < P_NB ? f : f + 2; }

__global__ void __launch_bounds__(64) k_small(SmallArgs A)
{
    __shared__ float bins[NSF][64];  // (peak q, shell s) at q m + s - 1
    __shared__ int4 spk[64];         // centre and flags of the block's peaks
    __shared__ double sdF[64], sfb[64];
    const int lane = threadIdx.x;
    const GatherFields &F = A.F;
    const long n1 = F.n, n12 = (long)F.n * F.n;
    const int x = blockIdx.x % A.nx;
    const int64_t nloc = A.xlen[x], stride = gridDim.x / A.nx;
    for (int64_t j = blockIdx.x / A.nx; j < nloc; j += stride) {
        int e = A.xp[x];
        while (e + 1 < A.xp[x + 1] && j >= A.part[e + 1].cum) e++;
        const SmallPart pt = A.part[e];
        const SmallSeg S = A.seg[pt.seg];
        const int m = S.m, P = S.P;
        const int64_t pbase = S.lo + (pt.blk_lo + (j - pt.cum)) * P;
        for (int f = 0; f < NSF; f++) bins[f][lane] = 0.0f;
        if (lane < P && pbase + lane < S.hi) {
            const PeakParam pk = A.pps[pbase + lane];
            spk[lane] = make_int4(pk.ci, pk.cj, pk.ck, pk.flags);
        }
        __syncthreads();
        for (int b = 0; b < S.nbat; b++) {
            const CellSlot t = A.tab[S.toff + b * 64 + lane];
            const bool live = t.cell >= 0 && pbase + t.q < S.hi;
            float acc[NSF];
#pragma unroll
            for (int f = 0; f < NSF; f++) acc[f] = 0.0f;
            if (t.lcell >= 0 && pbase + t.lq < S.hi) {
                const int4 pk = spk[t.lq];
                const int dx = (t.lcell & 255) - 128, dy = ((t.lcell >> 8) & 255) - 128,
                          dz = ((t.lcell >> 16) & 255) - 128;
                const int x = pk.x + dx, y = pk.y + dy, z = pk.z + dz;
                if ((pk.w & 1) || (x >= 0 && x < F.n && y >= 0 && y < F.n && z >= 0 && z < F.n)) {
                    const long a = x + y * n1 + z * n12;
                    const float4 q = F.p4[a];
                    float4 q2;
                    if (F.q4) q2 = F.q4[a];
                    else q2 = make_float4(F.ez[a], F.e2x ? F.e2x[a] : 0.0f, F.e2x ? F.e2y[a] : 0.0f,
                                          F.e2x ? F.e2z[a] : 0.0f);
                    const float d = q.x, dT = q.y, ex = q.z, ey = q.w, ez = q2.x;
                    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
                    acc[P_G + 0] = d * fx; acc[P_G + 1] = d * fy; acc[P_G + 2] = d * fz;
                    acc[P_GF + 0] = dT * fx; acc[P_GF + 1] = dT * fy; acc[P_GF + 2] = dT * fz;
                    acc[P_E + 0] = 2.0f * (ex * fx);
                    acc[P_E + 1] = 2.0f * (ey * fy);
                    acc[P_E + 2] = 2.0f * (ez * fz);
                    acc[P_E + 3] = ex * fy + ey * fx;
                    acc[P_E + 4] = ex * fz + ez * fx;
                    acc[P_E + 5] = ey * fz + ez * fy;
                    acc[P_F] = d;
                    acc[P_N] = 1.0f;
                    acc[P_NB + 0] = ex; acc[P_NB + 1] = ey; acc[P_NB + 2] = ez;
                    if (F.e2x) { acc[P_NB + 3] = q2.y; acc[P_NB + 4] = q2.z; acc[P_NB + 5] = q2.w; }
                    if (F.lap) acc[P_NB + 6] = F.lap[a];
                }
            }
            // row-major loads -> (peak, shell) order: lane takes lane src's sums
#pragma unroll
            for (int f = 0; f < NSF; f++) acc[f] = __shfl(acc[f], t.src, 64);
            for (int k = 0; k < S.steps; k++) {
                const int d = 1 << k;
                const bool in = lane + d < t.rend;
#pragma unroll
                for (int f = 0; f < NSF; f++) {
                    const float u = __shfl_down(acc[f], d, 64);
                    if (in) acc[f] += u;
                }
            }
            if (live && t.run > 0) {
                const int o = t.q * m + t.s - 1;
#pragma unroll
                for (int f = 0; f < NSF; f++) bins[f][o] += acc[f];
            }
            __syncthreads();
        }
        // record of (peak q, shell s) on lane o = q m + s - 1
        const int q = lane / m, s = 1 + (lane - q * m);
        const int64_t idx = pbase + q;
        const bool act = lane < P * m && idx < S.hi;
        float v[PR];
#pragma unroll
        for (int f = 0; f < PR; f++) v[f] = 0.0f;
        if (act) {
#pragma unroll
            for (int f = 0; f < NSF; f++) v[sf_slot(f)] = bins[f][lane];
            if (s == 1) v[P_L] = 0.0f;  // d2F starts at cell 2
        }
        sdF[lane] = v[P_N] > 0.0f ? (double)v[P_F] / (double)v[P_N] : 0.0;
        __syncthreads();
        // Fbar recursion (RadialShell.jl:383-390), one lane per peak
        if (lane < P && pbase + lane < S.hi) {
            const double *dF = sdF + lane * m;
            double fb = dF[0];  // F of the centre (count 1)
            sfb[lane * m] = fb;
            for (int t = 2; t <= m; t++) {
                const double rp = A.rad[t - 2], rs = A.rad[t - 1];
                const double r3p = rp * rp * rp, r3 = rs * rs * rs;
                fb = (r3p * fb + 0.5 * (dF[t - 2] + dF[t - 1]) * (r3 - r3p)) / r3;
                sfb[lane * m + t - 1] = fb;
            }
        }
        __syncthreads();
        if (act) {
            put_fbar(v, sfb[lane]);
            store_record(A.prof, A.prof_off[idx] + (s - 1), v, 3);
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------- wave gather
struct GatherArgs {
    GatherFields F;
    const PeakParam *pps;
    const GroupDesc *desc;     // all batch tables
    const int64_t *full_off;   // per chunk c: first descriptor of the full-chunk table
    const int32_t *full_nb;    //   and its number of 64-slot batches
    const int64_t *part_off;   // per m: the table of the last, partial chunk
    const int32_t *part_nb;
    ProfBuf prof;
    const int64_t *prof_off;   // per sorted index: first shell record
    const double *rad;
    int *queue;                // nx counters, 64 ints apart
    const int32_t *xorder;     // per-XCD lists of sorted positions
    int64_t xb[NXCD + 1];      // list x = xorder[xb[x], xb[x + 1])
    int nx;                    // lists (1 or NXCD)
    int cs;                    // shells per chunk
    int steps;                 // shuffle-tree steps: 2^steps >= the longest run of equal shells in a batch
};

__device__ __forceinline__ void neumaier(float *hi, float *lo, float x)
{
    const float h = *hi;
    const float t = h + x;
    const float e = fabsf(h) >= fabsf(x) ? (h - t) + x : (x - t) + h;
    *hi = t;
    *lo += e;
}

// One wave per peak.  Per batch: each lane sums its group's images, the
// lanes of each run of equal shells (the host sorted every batch by shell)
// sum their 14 partial sums by a segmented tree over wave shuffles, and the
// run's first lane deposits the sum into the chunk's bins.  The records of a chunk are
// written with Fbar, whose recursion is carried as the prefix sum
// Q(s) = Fbar(s) r^3(s) = sum_{t=2..s} (dF(t-1) + dF(t)) / 2 (r^3(t) - r^3(t-1)).
__global__ void __launch_bounds__(64) k_gather(GatherArgs A)
{
    extern __shared__ float lds[];
    const int lane = threadIdx.x;
    const int cs = A.cs;
    float *bh = lds, *bl = bh + P_NB * cs;  // [P_NB][cs] each
    // own XCD's list first, then the others' (the tail of the work)
    const int x0 = blockIdx.x % A.nx;
    int xq = 0;
    for (;;) {
        int idx = -1;
        if (lane == 0) {
            for (; xq < A.nx; xq++) {
                const int x = (x0 + xq) % A.nx;
                const int64_t t = A.xb[x] + atomicAdd(A.queue + 64 * x, 1);
                if (t < A.xb[x + 1]) { idx = A.xorder[t]; break; }
            }
        }
        idx = __shfl(idx, 0, 64);
        if (idx < 0) break;
        const PeakParam pk = A.pps[idx];
        const int m = pk.m;
        const int64_t rec = A.prof_off[idx];
        float F1;
        {
            float v[PR];
            F1 = center_values<false>(A.F, pk, v);
            if (lane == 0) store_record(A.prof, rec, v, 1);
        }
        double carry = 0.0, dFprev = (double)F1;
        const int nch = m >= 2 ? (m - 2) / cs + 1 : 0;
        for (int c = 0; c < nch; c++) {
            const int s0 = 2 + c * cs, s1 = min(m, s0 + cs - 1), ns = s1 - s0 + 1;
            for (int f = 0; f < P_NB; f++)
                for (int b = lane; b < ns; b += 64) { bh[f * cs + b] = 0.0f; bl[f * cs + b] = 0.0f; }
            __syncthreads();
            const bool full = s0 + cs - 1 <= m;
            const GroupDesc *tab = A.desc + (full ? A.full_off[c] : A.part_off[m]);
            const int nb = full ? A.full_nb[c] : A.part_nb[m];
            for (int b = 0; b < nb; b++) {
                const GroupDesc g = tab[b * 64 + lane];
                float acc[PR];
#pragma unroll
                for (int f = 0; f < PR; f++) acc[f] = 0.0f;
                if (g.s >= 2) group_sum<false>(A.F, pk, g, acc);
                const int rend = (g.lmn >> 24) & 127;
                for (int k = 0; k < A.steps; k++) {
                    const int d = 1 << k;
                    const bool in = lane + d < rend;
#pragma unroll
                    for (int f = 0; f < P_NB; f++) {
                        const float u = __shfl_down(acc[f], d, 64);
                        if (in) acc[f] += u;
                    }
                }
                if (g.lmn < 0) {  // run head
                    const int sb = g.s - s0;
#pragma unroll
                    for (int f = 0; f < P_NB; f++) neumaier(&bh[f * cs + sb], &bl[f * cs + sb], acc[f]);
                }
            }
            for (int b0 = 0; b0 < ns; b0 += 64) {
                const int b = b0 + lane;
                const bool valid = b < ns;
                const int s = s0 + b;
                float v[PR];
#pragma unroll
                for (int f = 0; f < PR; f++) v[f] = 0.0f;
                double term = 0.0, dF = 0.0, r3 = 1.0;
                if (valid) {
#pragma unroll
                    for (int f = 0; f < P_NB; f++) v[f] = bh[f * cs + b] + bl[f * cs + b];
                    dF = v[P_N] > 0.0f ? (double)v[P_F] / (double)v[P_N] : 0.0;
                }
                double dFm = __shfl_up(dF, 1, 64);
                if (lane == 0) dFm = dFprev;
                if (valid) {
                    const double rp = A.rad[s - 2], rs = A.rad[s - 1];
                    const double r3p = rp * rp * rp;
                    r3 = rs * rs * rs;
                    term = 0.5 * (dFm + dF) * (r3 - r3p);
                }
                const double Q = carry + wave_scan(term);
                if (valid) {
                    put_fbar(v, Q / r3);
                    store_record(A.prof, rec + (s - 1), v, 1);
                }
                const int last = min(63, ns - 1 - b0);
                carry = __shfl(Q, last, 64);
                dFprev = __shfl(dF, last, 64);
            }
            __syncthreads();
        }
    }
}

// --------------------------------------------------------------- post
template <bool FULL>
struct AosProfile {
    const float *se, *sg, *ax;  // this peak's records (E, G Gf, the rest)
    const double *radt;
    bool lapd;
    const WalkRec *wt = nullptr;  // walk table: Ebar of shells 1..wmc (lane post of the walk-table class)
    int wmc = 0;
    __device__ const float *x(int s) const { return ax + (int64_t)(s - 1) * PSR; }
    __device__ double rad(int s) const { return radt[s - 1]; }
    __device__ double Fbar(int s) const { return *(const double *)(x(s) + (P_FB - PSR)); }
    __device__ double count(int s) const { return (double)x(s)[P_N - PSR]; }
    __device__ double S(int s, int q) const { return FULL ? (double)x(s)[P_S - PSR + q] : 0.0; }
    __device__ double S2(int s, int q) const { return FULL ? (double)x(s)[P_S2 - PSR + q] : 0.0; }
    __device__ double lap(int s) const { return FULL ? (double)x(s)[P_L - PSR] : 0.0; }
    __device__ bool has_lapd() const { return FULL && lapd; }
    // count, S(3), S2(3) of shell s from one load of its 48-byte record
    __device__ void sums(int s, double &cnt, double Sv[3], double S2v[3]) const
    {
        float v[PSR];
        __builtin_memcpy(v, x(s), sizeof v);
        cnt = (double)v[P_N - PSR];
#pragma unroll
        for (int L = 0; L < 3; L++) {
            Sv[L] = FULL ? (double)v[P_S - PSR + L] : 0.0;
            S2v[L] = FULL ? (double)v[P_S2 - PSR + L] : 0.0;
        }
    }
    __device__ void strain_E(int s, float Es[6]) const
    {
        const float2 *q = (const float2 *)(se + (int64_t)(s - 1) * PSE);
        const float2 a = q[0], b = q[1], c = q[2];
        Es[0] = a.x; Es[1] = a.y; Es[2] = b.x; Es[3] = b.y; Es[4] = c.x; Es[5] = c.y;
    }
    __device__ void strain_G(int s, float G[3], float Gf[3]) const
    {
        const float2 *q = (const float2 *)(sg + (int64_t)(s - 1) * PSE);
        const float2 a = q[0], b = q[1], c = q[2];
        G[0] = a.x; G[1] = a.y; G[2] = b.x;
        Gf[0] = b.y; Gf[1] = c.x; Gf[2] = c.y;
    }
    // shells s and s + 1 (adjacent records, 48 bytes; 4-byte aligned wide loads)
    __device__ void strain_E2(int s, float Es[12]) const
    {
        __builtin_memcpy(Es, se + (int64_t)(s - 1) * PSE, 12 * sizeof(float));
    }
    __device__ void strain_G2(int s, float G[6], float Gf[6]) const
    {
        float v[12];
        __builtin_memcpy(v, sg + (int64_t)(s - 1) * PSE, sizeof v);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            G[3 * h] = v[6 * h]; G[3 * h + 1] = v[6 * h + 1]; G[3 * h + 2] = v[6 * h + 2];
            Gf[3 * h] = v[6 * h + 3]; Gf[3 * h + 1] = v[6 * h + 4]; Gf[3 * h + 2] = v[6 * h + 5];
        }
    }
};

template <bool FULL>
__device__ __forceinline__ AosProfile<FULL> profile_at(ProfBuf p, int64_t rb, const double *rad, bool lapd)
{
    return AosProfile<FULL>{p.e + rb * PSE, p.g + rb * PSE, p.a + rb * PSR, rad, lapd};
}

// launch grids (measured): workgroups per CU of k_small (grid-stride), waves
// per CU of the wave post, the finish pass and the lane posts (grid-stride:
// finer rounds balance better); the lane post asks for one wave per SIMD as
// its register budget.  The finish pass at c3 (whole call, kernel time): 32
// waves per CU 15.55 ms, 64 15.33, 128 15.14, 256 15.05, 512 15.06, 4096
// (about one peak per wave) 15.10
constexpr int kSmallGridCU = 32;
constexpr int kWaveGridCU = 16;
constexpr int kFinishGridCU = 256;
constexpr int kLaneGridCU = 64;
constexpr int kWalkGridCU = 16;  // 256-lane workgroups per CU of the walk table (grid-stride over the peaks)

// The lane post's order: m descending, then the fcrit crossing descending,
// ties in position order (a stable radix sort of these keys), so the 64
// lanes of a wave mostly start their walks at the same shell and step
// through the same shell bands together.
// keys[t - lo] = (cmax - m) << cbits | (cmax - crossing), m <= cmax < 2^cbits:
// ascending keys give m and the crossing descending in 2 * cbits bits.
// The crossing of every peak also goes to mc[t - lo]: the walk table holds
// shells 1..mc of the peak.
__global__ void k_crossing_keys(const PeakParam *__restrict__ pps, int64_t lo, int64_t hi, ProfBuf prof,
                                const int64_t *__restrict__ prof_off, const double *__restrict__ rad,
                                int cmax, int cbits, uint32_t *__restrict__ keys, int64_t *__restrict__ mc)
{
    for (int64_t t = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < hi; t += (int64_t)gridDim.x * blockDim.x) {
        const PeakParam pk = pps[t];
        const int64_t rb = prof_off[t];
        const AosProfile<false> P = profile_at<false>(prof, rb, rad, false);
        const int c = (pk.flags & 2) ? 0 : max(crossing_m0(P, pk), 0);
        keys[t - lo] = (uint32_t)(cmax - pk.m) << cbits | (uint32_t)(cmax - c);
        mc[t - lo] = c;
    }
}

__global__ void k_iota_from(int32_t *__restrict__ o, int64_t n, int32_t first)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        o[p] = first + (int32_t)p;
}

// The walk table: for a peak t in [lo, hi) with fcrit crossing mc(t) (the m0
// its walk starts from), record mp - 1 at eoff(t) (eoff: exclusive prefix
// sum of mc) holds the walk's step at shell mp: kernel_strain's Ebar over the
// band walk_band gives at mp, and the step's zvir.  The lane post's walk then
// takes one record per step instead of a band sum, an eigen-solve and a
// table look-up, whose chains of dependent loads set the length of the
// longest walks.  The records are formed as far as each walk goes: a group
// of W lanes per peak (W by the peak's size, below) evaluates W consecutive
// steps at a time from mc inward (lane l: shell top - l; the lanes' band
// loads share cache lines), and the group stops after the first block
// holding a step that ends the walk (a non-finite Ebar or zvir >= ZZon); the
// steps below it in that block are written too and never read.  Each record
// is the same arithmetic on the same inputs as the walk's own step.
// Group widths by peak size: a walk is at most mc - 1 <= m - 1 steps and a
// step's band sum grows with its radius, so larger peaks get wider groups
// (fewer rounds of W steps on the longest walks); narrower groups waste fewer
// steps below a walk's end (measured at c3: widths 64/32/16/8 15.9 ms,
// 16/8/4/4 15.74, 8/4/2/2 15.80, 4/2/2/1 16.0 for the whole call).  Class k
// holds the peaks with m >= kWalkClassM[k] (and below the previous class's
// bound).
constexpr int kWalkClasses = 4;
constexpr int kWalkClassW[kWalkClasses] = {16, 8, 4, 4};
constexpr int kWalkClassM[kWalkClasses] = {256, 96, 48, 0};
struct WalkClasses {
    int64_t lo[kWalkClasses], hi[kWalkClasses];  // positions in the lane-post order, relative to its start
    int32_t blk[kWalkClasses + 1];               // first workgroup of each class
};

template <int W>
__device__ __forceinline__ void walk_group_class(const PeakParam *__restrict__ pps, int64_t lo, int64_t c_lo,
                                                 int64_t c_hi, int64_t b0, int64_t nb,
                                                 const int32_t *__restrict__ list, ProfBuf prof,
                                                 const int64_t *__restrict__ prof_off, const PostConst &C,
                                                 const int64_t *__restrict__ eoff, const int64_t *__restrict__ mc,
                                                 WalkRec *__restrict__ out)
{
    const int lane = threadIdx.x & (W - 1);
    const unsigned long long gmask = (W == 64 ? ~0ull : (1ull << W) - 1) << ((threadIdx.x & 63) & ~(W - 1));
    const int64_t ngroups = nb * blockDim.x / W;
    for (int64_t p = c_lo + ((blockIdx.x - b0) * (int64_t)blockDim.x + threadIdx.x) / W; p < c_hi; p += ngroups) {
        const int64_t idx = list ? (int64_t)list[p] : lo + p;
        const int64_t u = idx - lo;
        const int c = (int)mc[u];
        if (c == 0) continue;  // no crossing: no walk (the group's lanes hold the same peak)
        const PeakParam pk = pps[idx];
        const AosProfile<false> P = profile_at<false>(prof, prof_off[idx], C.rad, false);
        WalkRec *rec = out + eoff[u];
        for (int top = c; top >= 1; top -= W) {
            const int mp = top - lane;
            bool stop = false;
            if (mp >= 1) {
                int mlow, mupp;
                walk_band(C, mp, pk.m, mlow, mupp);
                const Strain sv = kernel_strain<1, KS_E>(P, C, mp, mlow, mupp);
                WalkRec r;
#pragma unroll
                for (int q = 0; q < 6; q++) r.E[q] = sv.E[q];
                const bool fin = finite6(sv.E);
                double Fh, ev, pv;
                r.zvir = fin ? step_zvir(C, sv.E, P.Fbar(mp), Fh, ev, pv) : -1.0;
                const int e = mupp + 1;  // the first shell beyond the band
                const bool edge = e <= pk.m && C.rad[e - 1] == C.rad[mp - 1] + 2.0;
                r.mlow = mlow;
                r.mupp = mupp | (edge ? 1 << 30 : 0);
                rec[mp - 1] = r;
                stop = mp < c && (!fin || r.zvir >= pk.ZZon);  // shell mc itself is the walk's start, not a step
            }
            if (__ballot(stop) & gmask) break;
        }
    }
}

__global__ void __launch_bounds__(256) k_walk_group(const PeakParam *__restrict__ pps, int64_t lo,
                                                    const int32_t *__restrict__ list, ProfBuf prof,
                                                    const int64_t *__restrict__ prof_off, PostConst C,
                                                    const int64_t *__restrict__ eoff, const int64_t *__restrict__ mc,
                                                    WalkRec *__restrict__ out, WalkClasses K)
{
    int k = 0;
    while (k < kWalkClasses - 1 && (int)blockIdx.x >= K.blk[k + 1]) k++;
    const int64_t b0 = K.blk[k], nb = K.blk[k + 1] - K.blk[k];
    if (nb <= 0) return;
    switch (k) {
    case 0: walk_group_class<kWalkClassW[0]>(pps, lo, K.lo[0], K.hi[0], b0, nb, list, prof, prof_off, C, eoff, mc, out); break;
    case 1: walk_group_class<kWalkClassW[1]>(pps, lo, K.lo[1], K.hi[1], b0, nb, list, prof, prof_off, C, eoff, mc, out); break;
    case 2: walk_group_class<kWalkClassW[2]>(pps, lo, K.lo[2], K.hi[2], b0, nb, list, prof, prof_off, C, eoff, mc, out); break;
    default: walk_group_class<kWalkClassW[3]>(pps, lo, K.lo[3], K.hi[3], b0, nb, list, prof, prof_off, C, eoff, mc, out); break;
    }
}

// One lane per peak.  A walk longer than max_steps (a wave runs in step with
// its slowest lane) is handed to the wave post through the defer list.  With
// a list, lane t takes the peak at list[t - lo].
template <bool FULL>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1)))
k_post_lane(const PeakParam *__restrict__ pps, int64_t lo, int64_t hi, ProfBuf prof,
            const int64_t *__restrict__ prof_off, PostConst C, ResultView R, int has_lap, int max_steps,
            int64_t *__restrict__ defer, unsigned int *__restrict__ ndefer, const int32_t *__restrict__ list,
            const WalkRec *__restrict__ wtab, const int64_t *__restrict__ eoff, const int64_t *__restrict__ wmc)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < hi; t += stride) {
        const int64_t idx = list ? (int64_t)list[t - lo] : t;
        const PeakParam pk = pps[idx];
        const int64_t rb = prof_off[idx];
        AosProfile<FULL> P = profile_at<FULL>(prof, rb, C.rad, has_lap != 0);
        if (wtab) {  // the table covers [lo, hi)
            P.wt = wtab + eoff[idx - lo];
            P.wmc = (int)wmc[idx - lo];
        }
        if (analyse_post<1, false>(P, pk, C, R, pk.order, max_steps) == POST_DEFER) {
            defer[atomicAdd(ndefer, 1u)] = idx;
            continue;
        }
        R.nshells[pk.order] = pk.m;
    }
}

// Phase 7's record sums for the collapsed small peaks (k_small's records
// carry S, S2 and lapd per shell), one lane per peak, after their post: kept
// out of the post, whose registers it would otherwise add to.
__global__ void __launch_bounds__(256) k_small_finish(const PeakParam *__restrict__ pps, int64_t lo, int64_t hi,
                                                      ProfBuf prof, const int64_t *__restrict__ prof_off,
                                                      PostConst C, ResultView R, int has_lap)
{
    for (int64_t t = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < hi; t += (int64_t)gridDim.x * blockDim.x) {
        const PeakParam pk = pps[t];
        const int64_t o = pk.order;
        if (R.status[o] != DUMP_NONE) continue;
        const AosProfile<true> P = profile_at<true>(prof, prof_off[t], C.rad, has_lap != 0);
        finish_records<1>(P, pk, C, R, o, R.m0[o], R.q[o]);
    }
}

// One wave per deferred peak.
__global__ void __launch_bounds__(64) k_post_defer(const PeakParam *__restrict__ pps, const int64_t *__restrict__ defer,
                                                   const unsigned int *__restrict__ ndefer, ProfBuf prof,
                                                   const int64_t *__restrict__ prof_off, PostConst C, ResultView R)
{
    const unsigned int n = *ndefer;
    for (unsigned int t = blockIdx.x; t < n; t += gridDim.x) {
        const int64_t idx = defer[t];
        const PeakParam pk = pps[idx];
        const int64_t rb = prof_off[idx];
        const AosProfile<false> P = profile_at<false>(prof, rb, C.rad, false);
        analyse_post<64, false>(P, pk, C, R, pk.order);
        if (threadIdx.x == 0) R.nshells[pk.order] = pk.m;
    }
}

__global__ void __launch_bounds__(64) k_post_wave(const PeakParam *__restrict__ pps, int64_t lo, int64_t hi,
                                                  ProfBuf prof, const int64_t *__restrict__ prof_off,
                                                  PostConst C, ResultView R)
{
    for (int64_t idx = lo + blockIdx.x; idx < hi; idx += gridDim.x) {
        const PeakParam pk = pps[idx];
        const int64_t rb = prof_off[idx];
        const AosProfile<false> P = profile_at<false>(prof, rb, C.rad, false);
        analyse_post<64, false>(P, pk, C, R, pk.order);
        if (threadIdx.x == 0) R.nshells[pk.order] = pk.m;
    }
}

// W lanes per peak, 64 / W peaks per wave
template <int W>
__global__ void __launch_bounds__(64) k_post_group(const PeakParam *__restrict__ pps, int64_t lo, int64_t hi,
                                                   ProfBuf prof, const int64_t *__restrict__ prof_off,
                                                   PostConst C, ResultView R)
{
    constexpr int G = 64 / W;
    const int grp = threadIdx.x / W;
    for (int64_t b0 = lo + (int64_t)blockIdx.x * G; b0 < hi; b0 += (int64_t)gridDim.x * G) {
        const int64_t idx = b0 + grp;
        if (idx >= hi) continue;  // whole lane groups drop out
        const PeakParam pk = pps[idx];
        const int64_t rb = prof_off[idx];
        const AosProfile<false> P = profile_at<false>(prof, rb, C.rad, false);
        analyse_post<W, false>(P, pk, C, R, pk.order);
        if ((threadIdx.x & (W - 1)) == 0) R.nshells[pk.order] = pk.m;
    }
}

// ------------------------------------------------------------ finish
// One wave per collapsed wave-gathered peak: Sbar, Sbar2 over the shells
// inside m0 (shells 1..m0-1), lapd over shells 2..q (cells with r <= R_TH
// but the centre) and the mask over shells 1..q (RadialShell.jl:756-819),
// read through the same octant-group batch tables as the gather (rows of
// neighbouring cells per load instruction), summed in Float64 per lane and
// reduced by a fixed butterfly.  The reference's d2F stops at the first
// out-of-bounds cell in r^2 order; a peak whose shells 2..q leave the grid
// takes the cell-order walk of finish_cells instead.
struct FinishArgs {
    GatherFields F;
    const PeakParam *pps;
    const GroupDesc *desc;
    const int64_t *full_off;
    const int32_t *full_nb;
    const int64_t *part_off;
    const int32_t *part_nb;
    const int32_t *xorder;     // per-XCD lists of sorted positions, as for k_gather
    int64_t xb[NXCD + 1];
    int nx;                    // gridDim.x is a multiple of nx
    int cs;
    PostConst C;
    ResultView R;
};

__global__ void __launch_bounds__(64) k_finish(FinishArgs A)
{
    const int lane = threadIdx.x;
    const GatherFields &F = A.F;
    const PostConst &C = A.C;
    const ResultView &R = A.R;
    const float *ex[3] = {F.ex, F.ey, F.ez};
    const float *e2[3] = {F.e2x, F.e2y, F.e2z};
    const long n1 = F.n, n12 = (long)F.n * F.n;
    const int lo1 = C.nbuff, hi1 = C.n1 - C.nbuff;
    const int x = blockIdx.x % A.nx;
    const int64_t stride = gridDim.x / A.nx;
    for (int64_t j = A.xb[x] + blockIdx.x / A.nx; j < A.xb[x + 1]; j += stride) {
        const PeakParam pk = A.pps[A.xorder[j]];
        const int64_t o = pk.order;
        if (R.status[o] != DUMP_NONE) continue;
        const int m = pk.m, m0 = R.m0[o], q = R.q[o];
        const int xs = m0 - 1, X = max(xs, q);
        const int r2X = X >= 1 ? C.shell_r2[X - 1] : 0;
        const bool inside = pk.flags & 1;
        double Sb[3] = {0, 0, 0}, Sb2[3] = {0, 0, 0}, nS = 0.0, sL = 0.0, nL = 0.0;
        bool oob = false;
        if (lane == 0) {  // centre cell: shell 1 (in Sbar, in the mask, not in d2F)
            const long a = pk.ci + pk.cj * n1 + pk.ck * n12;
            for (int L = 0; L < 3; L++) {
                Sb[L] += (double)ex[L][a];
                if (e2[0]) Sb2[L] += (double)e2[L][a];
            }
            nS += 1.0;
            if (C.mask && pk.ci >= lo1 && pk.ci < hi1 && pk.cj >= lo1 && pk.cj < hi1 && pk.ck >= lo1 && pk.ck < hi1)
                C.mask[a] = 1;
        }
        const int nch = X >= 2 ? (X - 2) / A.cs + 1 : 0;
        for (int c = 0; c < nch; c++) {
            const int s0 = 2 + c * A.cs;
            const bool full = s0 + A.cs - 1 <= m;
            const GroupDesc *tab = A.desc + (full ? A.full_off[c] : A.part_off[m]);
            const int nb = full ? A.full_nb[c] : A.part_nb[m];
            for (int b = 0; b < nb; b++) {
                const GroupDesc g = tab[b * 64 + lane];
                const int l = g.lmn & 255, dj = (g.lmn >> 8) & 255, dk = (g.lmn >> 16) & 255;
                // rows come by increasing dj^2 + dk^2: once a batch's smallest
                // row offset is beyond shell X, so is every later batch
                int amin = g.s >= 2 ? dj * dj + dk * dk : 0x7fffffff;
                for (int o2 = 32; o2 >= 1; o2 >>= 1) amin = min(amin, __shfl_xor(amin, o2, 64));
                if (amin > r2X) break;
                if (g.s < 2 || g.s > X) continue;
                const bool inS = g.s <= xs, inL = g.s <= q;
#pragma unroll
                for (int img = 0; img < 8; img++) {
                    const int sx = (img & 1) ? -1 : 1, sy = (img & 2) ? -1 : 1, sz = (img & 4) ? -1 : 1;
                    if ((sx < 0 && l == 0) || (sy < 0 && dj == 0) || (sz < 0 && dk == 0)) continue;
                    const int x = pk.ci + sx * l, y = pk.cj + sy * dj, z = pk.ck + sz * dk;
                    if (!inside && (x < 0 || x >= F.n || y < 0 || y >= F.n || z < 0 || z >= F.n)) {
                        oob = oob || inL;
                        continue;
                    }
                    const long a = x + y * n1 + z * n12;
                    if (inS) {
                        if (F.q4) {
                            const float4 p = F.p4[a], q2 = F.q4[a];
                            const float ev[3] = {p.z, p.w, q2.x}, e2v[3] = {q2.y, q2.z, q2.w};
                            for (int L = 0; L < 3; L++) {
                                Sb[L] += (double)ev[L];
                                Sb2[L] += (double)e2v[L];
                            }
                        } else {
                            for (int L = 0; L < 3; L++) {
                                Sb[L] += (double)ex[L][a];
                                if (e2[0]) Sb2[L] += (double)e2[L][a];
                            }
                        }
                        nS += 1.0;
                    }
                    if (inL) {
                        if (C.lapd) { sL += (double)C.lapd[a]; nL += 1.0; }
                        if (C.mask && x >= lo1 && x < hi1 && y >= lo1 && y < hi1 && z >= lo1 && z < hi1) C.mask[a] = 1;
                    }
                }
            }
        }
        if (__ballot(oob) != 0ull) {
            // rare: R_TH reaches past the grid edge; the ordered walk
            // recomputes this peak's outputs (its mask writes repeat the same
            // 1s)
            finish_cells<true>(C, pk, m0, q, ex, e2, R, o);
            continue;
        }
        for (int L = 0; L < 3; L++) { Sb[L] = wave_sum(Sb[L]); Sb2[L] = wave_sum(Sb2[L]); }
        nS = wave_sum(nS);
        sL = wave_sum(sL);
        nL = wave_sum(nL);
        if (lane == 0) {
            for (int L = 0; L < 3; L++) {
                R.Sbar[o * 3 + L] = nS > 0.0 ? Sb[L] / nS : 0.0;
                R.Sbar2[o * 3 + L] = nS > 0.0 ? Sb2[L] / nS : 0.0;
            }
            R.d2F[o] = (C.lapd && nL > 0.0) ? sL / nL : 0.0;
        }
    }
}

// ------------------------------------------------------------- host side
struct SmallTab {
    int32_t toff = 0, P = 0, nbat = 0, steps = 0;
};

struct PersistState {
    int rmax = -1, cs = 0;
    std::vector<int32_t> shell_of_r2;
    std::vector<GroupDesc> desc;                      // chunk tables
    std::vector<std::pair<int64_t, int>> full;        // per chunk: (offset, batches)
    std::map<int, std::pair<int64_t, int>> part;      // per m: last partial chunk
    std::vector<CellSlot> scells;                     // small-peak batch tables
    std::vector<SmallTab> stab;                       // index m (0..small_max)
    int small_max = 0;
    int gsteps = 0;                                   // shuffle-tree steps of the gather tables
    std::vector<SmallSeg> segs_host;
    std::vector<SmallPart> parts_host;                // per-XCD block lists of k_small
    std::vector<int64_t> xtab_host;                   // per m: k_xcd_order's table row (finish lists)
    std::vector<int64_t> xtab_g_host;                 // per group: the same for the gather lists
    DevBuf d_desc, d_full_off, d_full_nb, d_part_off, d_part_nb, d_scells, d_seg, d_xtab;
    bool dirty = true;
};

static std::map<pp_ctx *, PersistState> g_state;
static std::mutex g_state_mu;

static PersistState &state_of(pp_ctx *ctx)
{
    std::lock_guard<std::mutex> lk(g_state_mu);
    return g_state[ctx];
}

void persist_release(pp_ctx *ctx)
{
    std::lock_guard<std::mutex> lk(g_state_mu);
    auto it = g_state.find(ctx);
    if (it == g_state.end()) return;
    PersistState &st = it->second;
    DevBuf *bufs[] = {&st.d_desc, &st.d_full_off, &st.d_full_nb, &st.d_part_off, &st.d_part_nb, &st.d_scells,
                      &st.d_seg, &st.d_xtab};
    for (DevBuf *b : bufs) b->release();
    g_state.erase(it);
}

// Octant groups (l, dj, dk >= 0, not all zero) of shells [slo, shi], in
// row-major order (dk, dj, l).
template <class Fn>
static void for_groups(const ShellTable &T, const std::vector<int32_t> &shell_of_r2, int slo, int shi, Fn fn)
{
    const int r2lo = T.shell_r2[slo - 1], r2hi = T.shell_r2[shi - 1];
    const int R = (int)std::floor(std::sqrt((double)r2hi));
    // rows (dj, dk) by increasing a = dj^2 + dk^2, so a batch's smallest a
    // bounds every later batch's radii (the finish pass stops there); each
    // row's l-run stays contiguous, which is what the loads coalesce on
    std::vector<std::pair<int, int>> rows;
    for (int dk = 0; dk <= R; dk++)
        for (int dj = 0; dj <= R; dj++)
            if (dk * dk + dj * dj <= r2hi) rows.push_back({dk * dk + dj * dj, dk * 1024 + dj});
    std::stable_sort(rows.begin(), rows.end());
    for (const auto &rw : rows) {
        const int a = rw.first, dk = rw.second / 1024, dj = rw.second % 1024;
        int l = 0;
        if (r2lo - a > 0) {
            l = (int)std::floor(std::sqrt((double)(r2lo - a)));
            while (l * l + a < r2lo) l++;
            while (l > 0 && (l - 1) * (l - 1) + a >= r2lo) l--;
        }
        for (; l * l + a <= r2hi; l++) {
            const int r2 = l * l + a;
            if (r2 == 0) continue;
            fn(GroupDesc{l | (dj << 8) | (dk << 16), shell_of_r2[r2]});
        }
    }
}

// Groups of shells [slo, shi] in row order cut into consecutive 64-slot
// batches (neighbouring x-runs share cache lines); inside a batch the
// groups are sorted by shell; every lane carries the end of its run of equal
// shells in lmn bits 24-30 and the first lane of a run (its deposit lane)
// bit 31.  steps grows to ceil(log2) of the longest run.
static void pack_groups(const ShellTable &T, const std::vector<int32_t> &shell_of_r2, int slo, int shi,
                        std::vector<GroupDesc> &out, int &nbatch, int &steps)
{
    std::vector<GroupDesc> g;
    for_groups(T, shell_of_r2, slo, shi, [&](GroupDesc d) { g.push_back(d); });
    nbatch = (int)((g.size() + 63) / 64);
    const size_t base = out.size();
    out.resize(base + (size_t)nbatch * 64, GroupDesc{0, -1});
    for (int b = 0; b < nbatch; b++) {
        const size_t lo = (size_t)b * 64, hi = std::min(g.size(), lo + 64);
        std::stable_sort(g.begin() + lo, g.begin() + hi,
                         [](const GroupDesc &x, const GroupDesc &y) { return x.s < y.s; });
        for (size_t q = lo; q < hi;) {
            size_t r = q;
            while (r < hi && g[r].s == g[q].s) r++;
            for (size_t u = q; u < r; u++) g[u].lmn |= (int32_t)(r - lo) << 24;  // run end (lane)
            g[q].lmn |= (int32_t)(1u << 31);                                      // run head
            while ((1 << steps) < (int)(r - q)) steps++;
            q = r;
        }
        for (size_t q = lo; q < hi; q++) out[base + q] = g[q];
    }
}

// Chunk tables for the m in ms, and the small-peak tables (once per shell
// table).
static int ensure_tables(pp_ctx *ctx, const std::vector<int> &ms, int cs)
{
    PersistState &st = state_of(ctx);
    const ShellTable &T = ctx->shells;
    if (st.rmax != T.rmax || st.cs != cs) {
        DevBuf *bufs[] = {&st.d_desc, &st.d_full_off, &st.d_full_nb, &st.d_part_off, &st.d_part_nb, &st.d_scells,
                          &st.d_seg, &st.d_xtab};
        for (DevBuf *b : bufs) b->release();
        st = PersistState();
        st.rmax = T.rmax;
        st.cs = cs;
        st.shell_of_r2.assign((size_t)T.shell_r2[T.nshells - 1] + 1, 0);
        for (int s = 1; s <= T.nshells; s++) st.shell_of_r2[T.shell_r2[s - 1]] = s;
        // small-peak tables: cells of shells 1..m in row-major order (dk, dj,
        // di), P peaks per block, cut into 64-lane batches sorted by (peak,
        // shell) with the run length on each run's first lane
        st.stab.assign(SMALL_MAX + 1, SmallTab());
        st.small_max = std::min(SMALL_MAX, (int)T.nshells);
        for (int m = 1; m <= st.small_max; m++) {
            const int r2m = T.shell_r2[m - 1];
            const int R = (int)std::floor(std::sqrt((double)r2m));
            std::vector<std::pair<int32_t, int16_t>> cells;  // (packed cell, shell)
            for (int dk = -R; dk <= R; dk++)
                for (int dj = -R; dj <= R; dj++)
                    for (int di = -R; di <= R; di++) {
                        const int r2 = di * di + dj * dj + dk * dk;
                        if (r2 > r2m) continue;
                        cells.push_back({(di + 128) | ((dj + 128) << 8) | ((dk + 128) << 16),
                                         (int16_t)st.shell_of_r2[r2]});
                    }
            const int nc = (int)cells.size();
            const int P = std::max(1, 64 / m);  // the P m records of a block fit its 64 bins
            const int nv = P * nc, nbat = (nv + 63) / 64;
            SmallTab t;
            t.toff = (int32_t)st.scells.size();
            t.P = P;
            t.nbat = nbat;
            for (int b = 0; b < nbat; b++) {
                std::vector<CellSlot> bt;
                for (int v = b * 64; v < std::min(nv, b * 64 + 64); v++) {
                    CellSlot c;
                    c.cell = cells[v % nc].first;
                    c.q = (int16_t)(v / nc);
                    c.s = cells[v % nc].second;
                    c.run = 0;
                    c.rend = 0;
                    c.src = (int16_t)bt.size();  // the loading lane: its row-major position
                    bt.push_back(c);
                }
                const std::vector<CellSlot> rowmajor = bt;
                std::stable_sort(bt.begin(), bt.end(), [](const CellSlot &x, const CellSlot &y) {
                    return x.q != y.q ? x.q < y.q : x.s < y.s;
                });
                for (size_t u = 0; u < bt.size(); u++) {
                    bt[u].lcell = rowmajor[u].cell;
                    bt[u].lq = rowmajor[u].q;
                }
                for (size_t q0 = 0; q0 < bt.size();) {
                    size_t r = q0;
                    while (r < bt.size() && bt[r].q == bt[q0].q && bt[r].s == bt[q0].s) r++;
                    bt[q0].run = (int32_t)(r - q0);
                    for (size_t u = q0; u < r; u++) bt[u].rend = (int32_t)r;
                    while ((1 << t.steps) < (int)(r - q0)) t.steps++;
                    q0 = r;
                }
                while (bt.size() < 64)
                    bt.push_back(CellSlot{-1, 0, 1, 0, (int32_t)bt.size() + 1, -1, 0, (int16_t)bt.size()});
                st.scells.insert(st.scells.end(), bt.begin(), bt.end());
            }
            st.stab[m] = t;
        }
        PP_CHECK(st.d_scells.grow(std::max<size_t>(1, st.scells.size()) * sizeof(CellSlot)));
        if (!st.scells.empty())
            PP_HIP(hipMemcpy(st.d_scells.ptr, st.scells.data(), st.scells.size() * sizeof(CellSlot),
                             hipMemcpyHostToDevice));
    }
    for (int m : ms) {
        if (m < 2) continue;
        const int nfull = (m - 1) / cs;  // chunks entirely inside [2, m]
        while ((int)st.full.size() < nfull) {
            const int c = (int)st.full.size();
            int nb = 0;
            const int64_t off = (int64_t)st.desc.size();
            pack_groups(T, st.shell_of_r2, 2 + c * cs, 1 + (c + 1) * cs, st.desc, nb, st.gsteps);
            st.full.push_back({off, nb});
            st.dirty = true;
        }
        if ((m - 1) % cs != 0 && !st.part.count(m)) {
            const int c = (m - 2) / cs;
            int nb = 0;
            const int64_t off = (int64_t)st.desc.size();
            pack_groups(T, st.shell_of_r2, 2 + c * cs, m, st.desc, nb, st.gsteps);
            st.part[m] = {off, nb};
            st.dirty = true;
        }
    }
    if (st.dirty) {
        const int ns = T.nshells + 1;
        const size_t nf = std::max<size_t>(1, st.full.size());
        std::vector<int64_t> foff(nf, 0), poff(ns, 0);
        std::vector<int32_t> fnb(nf, 0), pnb(ns, 0);
        for (size_t c = 0; c < st.full.size(); c++) { foff[c] = st.full[c].first; fnb[c] = st.full[c].second; }
        for (auto &kv : st.part) { poff[kv.first] = kv.second.first; pnb[kv.first] = kv.second.second; }
        PP_CHECK(st.d_desc.grow(std::max<size_t>(1, st.desc.size()) * sizeof(GroupDesc)));
        PP_CHECK(st.d_full_off.grow(nf * sizeof(int64_t)));
        PP_CHECK(st.d_full_nb.grow(nf * sizeof(int32_t)));
        PP_CHECK(st.d_part_off.grow(ns * sizeof(int64_t)));
        PP_CHECK(st.d_part_nb.grow(ns * sizeof(int32_t)));
        if (!st.desc.empty())
            PP_HIP(hipMemcpy(st.d_desc.ptr, st.desc.data(), st.desc.size() * sizeof(GroupDesc), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(st.d_full_off.ptr, foff.data(), nf * sizeof(int64_t), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(st.d_full_nb.ptr, fnb.data(), nf * sizeof(int32_t), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(st.d_part_off.ptr, poff.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(st.d_part_nb.ptr, pnb.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice));
        st.dirty = false;
    }
    return PP_OK;
}

static int grid1d(pp_ctx *ctx, int64_t n, int per_cu = 8)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cu * per_cu));
}

// Ordered params + histogram of m; returns host histogram.  The order is m
// descending, ties in the caller's order (so the lane posts' result writes
// stay coalesced).  W: scratch for the work lists.
struct ListScratch {
    int32_t *xg, *xf;   // per-XCD lists: gather, finish
    int32_t *iota, *mo; // Morton sort values
    uint64_t *ka, *kb;  // Morton sort keys
};

static int order_peaks(pp_ctx *ctx, ShellRun &S, PeakParam *&pps, int32_t *&m_sorted, ListScratch &W,
                       std::vector<int> &hist)
{
    const int64_t np = S.np;
    const int nb = ctx->shells.nshells + 1;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    PP_CHECK(ctx->d_queue.grow(al(np * sizeof(PeakParam)) + al(np * 4) * 7 + al(np * 8) * 2 + al(nb * 4) + 4096));
    char *p = ctx->d_queue.as<char>();
    pps = (PeakParam *)p; p += al(np * sizeof(PeakParam));
    m_sorted = (int32_t *)p; p += al(np * 4);
    int32_t *order2 = (int32_t *)p; p += al(np * 4);
    int32_t *keys2 = (int32_t *)p; p += al(np * 4);
    W.xg = (int32_t *)p; p += al(np * 4);
    W.xf = (int32_t *)p; p += al(np * 4);
    W.iota = (int32_t *)p; p += al(np * 4);
    W.mo = (int32_t *)p; p += al(np * 4);
    W.ka = (uint64_t *)p; p += al(np * 8);
    W.kb = (uint64_t *)p; p += al(np * 8);
    int *d_hist = (int *)p;  // nb bins, then the out-of-order flag
    p += al((nb + 1) * 4);
    // one pass for the keys, the flag and the histogram of m; one read-back
    hipLaunchKernelGGL(k_fill_i32b, dim3(grid1d(ctx, nb + 1)), dim3(256), 0, ctx->stream, d_hist, (int64_t)nb + 1, 0);
    hipLaunchKernelGGL(k_keys_hist, dim3(grid1d(ctx, np, 4)), dim3(256), nb * sizeof(int), ctx->stream,
                       (const PeakParam *)S.pp, np, S.keys, nb, d_hist, d_hist + nb);
    PP_HIP(hipGetLastError());
    hist.assign(nb + 1, 0);
    PP_HIP(hipMemcpyAsync(hist.data(), d_hist, (nb + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    const int unsorted = hist[nb];
    hist.resize(nb);
    const int32_t *order = nullptr;
    if (unsorted) {
        hipLaunchKernelGGL(k_iota, dim3(grid1d(ctx, np)), dim3(256), 0, ctx->stream, S.order, np);
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
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// Positions [0, hi) of the ordered params with ties of m in Morton order of
// the centres in the n^3 grid (W.mo).
static int morton_order(pp_ctx *ctx, const PeakParam *pps, int64_t hi, int n, ListScratch &W)
{
    int nbits = 1;
    while ((1 << nbits) < n) nbits++;
    if (nbits > 16) { set_error("grid side %d above 65536", n); return PP_ERR_ARG; }
    hipLaunchKernelGGL(k_keys_morton, dim3(grid1d(ctx, hi)), dim3(256), 0, ctx->stream, pps, hi, nbits, W.ka);
    hipLaunchKernelGGL(k_iota, dim3(grid1d(ctx, hi)), dim3(256), 0, ctx->stream, W.iota, hi);
    PP_HIP(hipGetLastError());
    size_t tmp = 0;
    PP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, W.ka, W.kb, W.iota, W.mo, (int)hi, 0, 3 * nbits + 16,
                                              ctx->stream));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->d_cub.ptr, tmp, W.ka, W.kb, W.iota, W.mo, (int)hi, 0, 3 * nbits + 16,
                                              ctx->stream));
    return PP_OK;
}

// 1. packed (delta, delta transposed, eta x, eta y) [and (eta z, eta2)] on the
// side stream, forked from the context stream as soon as the fields are
// resident, so that it runs beside the peak set-up and ordering; the gathers
// wait for event 12.  Event 2 marks the fork (the start of the "prepare"
// stage).
int start_pack(pp_ctx *ctx, GatherFields &G)
{
    const int n = G.n;
    const bool pack_q4 = G.e2x != nullptr;
    PP_CHECK(ctx->d_field_b.grow((size_t)n * n * n * sizeof(float4) * (pack_q4 ? 2 : 1)));
    PP_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    PP_HIP(hipStreamWaitEvent(ctx->side, ctx->ev[2], 0));
    constexpr int T = kPackTile;
    dim3 grid((n + T - 1) / T, (n + T - 1) / T, n), block(T, 256 / T);
    float4 *q4 = pack_q4 ? ctx->d_field_b.as<float4>() + (size_t)n * n * n : nullptr;
    hipLaunchKernelGGL(k_pack4<T>, grid, block, 0, ctx->side, G.d, G.ex, G.ey, ctx->d_field_b.as<float4>(), n);
    if (q4) {
        const int64_t ncell = (int64_t)n * n * n;
        const int64_t gq = std::min<int64_t>((ncell + 255) / 256, INT_MAX);
        hipLaunchKernelGGL(k_pack_q4, dim3((unsigned)gq), dim3(256), 0, ctx->side, G.ez, G.e2x, G.e2y, G.e2z, q4,
                           ncell);
    }
    PP_HIP(hipGetLastError());
    PP_HIP(hipEventRecord(ctx->ev[12], ctx->side));
    G.p4 = ctx->d_field_b.as<float4>();
    G.q4 = q4;
    G.dT = nullptr;
    return PP_OK;
}

// The persistent pipeline on an ordered batch.  export_all keeps the profile
// for pp_shell_profiles (every peak then goes through k_gather).
static int run_persistent_impl(pp_ctx *ctx, GatherFields G, ShellRun &S, PostConst &C, bool export_all,
                               std::vector<int32_t> *cnt, std::vector<double> *fbar, std::vector<int64_t> *offs)
{
    const int64_t np = S.np;
    const int n = G.n;
    if (np > INT_MAX) { set_error("persistent form: at most 2^31-1 peaks per call"); return PP_ERR_CAPACITY; }
    for (double &t : ctx->timing.stage_ms) t = 0.0;
    // 1. packed fields (start_pack), unless the caller forked the packing
    // before its own set-up
    if (!G.p4) PP_CHECK(start_pack(ctx, G));
    // 2. order by m (descending)
    PeakParam *pps = nullptr;
    int32_t *m_sorted = nullptr;
    ListScratch W;
    std::vector<int> hist;
    PP_CHECK(order_peaks(ctx, S, pps, m_sorted, W, hist));
    const int nb = (int)hist.size();
    std::vector<int64_t> above(nb + 1, 0);  // above[x] = #peaks with m >= x
    for (int x = nb - 1; x >= 0; x--) above[x] = above[x + 1] + hist[x];
    auto cnt_gt = [&](int x) -> int64_t { return x < 0 ? np : above[std::min(x + 1, nb)]; };
    const int cs = std::max(8, std::min(ctx->chunk_shells > 0 ? ctx->chunk_shells : 64, 1024));
    std::vector<int> ms;
    for (int x = 2; x < nb; x++)
        if (hist[x]) ms.push_back(x);
    PP_CHECK(ensure_tables(ctx, ms, cs));
    PersistState &st = state_of(ctx);
    // classes in sorted order: [0, small_lo) wave gather, [small_lo, np) small
    // (m <= small_m); posts: [0, coop_end) one wave per peak, the rest one
    // lane per peak
    const int small_m = export_all ? 0 : std::min(std::max(ctx->tiny_m, 0), st.small_max);
    const int64_t small_lo = cnt_gt(small_m);
    const int64_t coop_end = std::min(small_lo, cnt_gt(std::max(ctx->coop_m, 1)));
    // 3. profile offsets: exclusive scan of m over all peaks
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    int64_t tot = 0;
    PP_CHECK(ctx->d_scan.grow(al(std::max<int64_t>(np, 1) * 8) * 2 + 4096));
    int64_t *d_m64 = ctx->d_scan.as<int64_t>();
    int64_t *d_off = (int64_t *)(ctx->d_scan.as<char>() + al(std::max<int64_t>(np, 1) * 8));
    if (np > 0) {
        hipLaunchKernelGGL(k_m_to_i64, dim3(grid1d(ctx, np)), dim3(256), 0, ctx->stream, (const int32_t *)m_sorted, np,
                           d_m64);
        size_t tmp = 0;
        PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_m64, d_off, (int)np, ctx->stream));
        PP_CHECK(ctx->d_cub.grow(tmp));
        PP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_cub.ptr, tmp, d_m64, d_off, (int)np, ctx->stream));
        // the records in all: the sum of m, from the histogram (no read-back)
        for (int x = 0; x < nb; x++) tot += (int64_t)x * hist[x];
    }
    PP_CHECK(ctx->d_res.grow((size_t)std::max<int64_t>(tot, 1) * PR * sizeof(float) + 4096));
    const size_t nrec = (size_t)std::max<int64_t>(tot, 1);
    const ProfBuf d_rec{ctx->d_res.as<float>(), ctx->d_res.as<float>() + nrec * PSE,
                        ctx->d_res.as<float>() + nrec * 2 * PSE};
    PP_CHECK(ctx->d_misc.grow(8192));
    int *d_q = (int *)(ctx->d_misc.as<char>() + 4096);  // NXCD queue counters, 256 B apart
    // per-XCD lists of the wave-gathered peaks [0, small_lo): part x of every
    // run of one group goes to list x; the finish pass's groups are the m
    // (its list keeps the caller's order), the gather's the size classes in
    // Morton order
    const int nx = NXCD;
    int64_t xb[NXCD + 1], xbg[NXCD + 1];
    // tab rows for groups ng (group of m: grp(m)), offsets into out; fills xbo
    auto build_xtab = [&](int ng, auto grp, std::vector<int64_t> &tab, int64_t *xbo) -> int {
        tab.assign((size_t)ng * (2 + NXCD), 0);
        std::vector<int64_t> len(ng, 0), first(ng, 0);
        for (int m = nb - 1; m > small_m; m--) len[grp(m)] += hist[m];
        {   // groups follow each other with decreasing group index (m descending)
            int64_t pos = 0;
            for (int g = ng - 1; g >= 0; g--) { first[g] = pos; pos += len[g]; }
        }
        int64_t cur[NXCD] = {0};
        for (int g = ng - 1; g >= 0; g--)
            for (int x = 0; x < nx; x++) cur[x] += len[g] * (x + 1) / nx - len[g] * x / nx;
        xbo[0] = 0;
        for (int x = 0; x < NXCD; x++) xbo[x + 1] = xbo[x] + (x < nx ? cur[x] : 0);
        for (int x = 0; x < nx; x++) cur[x] = xbo[x];
        for (int g = ng - 1; g >= 0; g--) {
            if (!len[g]) continue;
            int64_t *t = &tab[(size_t)g * (2 + NXCD)];
            t[0] = first[g];
            t[1] = len[g];
            for (int x = 0; x < nx; x++) {
                t[2 + x] = cur[x];
                cur[x] += len[g] * (x + 1) / nx - len[g] * x / nx;
            }
        }
        if (xbo[NXCD] != small_lo) {
            set_error("per-XCD lists hold %lld of %lld peaks", (long long)xbo[NXCD], (long long)small_lo);
            return PP_ERR_STATE;
        }
        return PP_OK;
    };
    {
        PP_CHECK(build_xtab(nb, [](int m) { return m; }, st.xtab_host, xb));
        PP_CHECK(build_xtab(32, [](int m) { return gather_class(m); }, st.xtab_g_host, xbg));
        if (small_lo > 0) {
            const size_t tb = st.xtab_host.size() * sizeof(int64_t), tg = st.xtab_g_host.size() * sizeof(int64_t);
            PP_CHECK(st.d_xtab.grow(tb + tg));
            PP_HIP(hipMemcpyAsync(st.d_xtab.ptr, st.xtab_host.data(), tb, hipMemcpyHostToDevice, ctx->stream));
            PP_HIP(hipMemcpyAsync(st.d_xtab.as<char>() + tb, st.xtab_g_host.data(), tg, hipMemcpyHostToDevice,
                                  ctx->stream));
            // finish: parts of consecutive positions; gather: of the Morton order
            hipLaunchKernelGGL(k_xcd_order, dim3(grid1d(ctx, small_lo)), dim3(256), 0, ctx->stream,
                               (const int32_t *)m_sorted, (const int32_t *)nullptr, small_lo,
                               st.d_xtab.as<int64_t>(), nx, 0, W.xf);
            PP_HIP(hipGetLastError());
            PP_CHECK(morton_order(ctx, pps, small_lo, n, W));
            hipLaunchKernelGGL(k_xcd_order, dim3(grid1d(ctx, small_lo)), dim3(256), 0, ctx->stream,
                               (const int32_t *)m_sorted, (const int32_t *)W.mo, small_lo,
                               (const int64_t *)(st.d_xtab.as<char>() + tb), nx, 1, W.xg);
            PP_HIP(hipGetLastError());
        }
    }
    // small-peak segments (equal m, descending)
    st.segs_host.clear();
    int64_t nblk = 0;
    for (int m = small_m; m >= 1; m--) {
        const int64_t lo = cnt_gt(m), hi = cnt_gt(m - 1);
        if (hi <= lo) continue;
        SmallSeg sg{};
        sg.m = m;
        sg.P = st.stab[m].P;
        sg.nbat = st.stab[m].nbat;
        sg.toff = st.stab[m].toff;
        sg.steps = st.stab[m].steps;
        sg.lo = lo;
        sg.hi = hi;
        sg.blk0 = nblk;
        nblk += (hi - lo + sg.P - 1) / sg.P;
        st.segs_host.push_back(sg);
    }
    // per-XCD block lists: part x of every segment's blocks goes to list x
    int32_t sxp[NXCD + 1] = {0};
    int64_t sxlen[NXCD] = {0};
    st.parts_host.clear();
    for (int x = 0; x < nx; x++) {
        sxp[x] = (int32_t)st.parts_host.size();
        int64_t cum = 0;
        for (size_t si = 0; si < st.segs_host.size(); si++) {
            const SmallSeg &sg = st.segs_host[si];
            const int64_t nbk = (sg.hi - sg.lo + sg.P - 1) / sg.P;
            const int64_t b0 = nbk * x / nx, b1 = nbk * (x + 1) / nx;
            if (b1 <= b0) continue;
            st.parts_host.push_back(SmallPart{(int32_t)si, 0, b0, cum});
            cum += b1 - b0;
        }
        sxlen[x] = cum;
    }
    for (int x = nx; x <= NXCD; x++) sxp[x] = (int32_t)st.parts_host.size();
    const size_t seg_bytes = (st.segs_host.size() * sizeof(SmallSeg) + 255) & ~(size_t)255;
    if (!st.segs_host.empty()) {
        PP_CHECK(st.d_seg.grow(seg_bytes + st.parts_host.size() * sizeof(SmallPart)));
        PP_HIP(hipMemcpyAsync(st.d_seg.ptr, st.segs_host.data(), st.segs_host.size() * sizeof(SmallSeg),
                              hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(st.d_seg.as<char>() + seg_bytes, st.parts_host.data(),
                              st.parts_host.size() * sizeof(SmallPart), hipMemcpyHostToDevice, ctx->stream));
    }
    // defer list of the lane post: [count | pad | indices]
    PP_CHECK(ctx->d_k2.grow(256 + (size_t)std::max<int64_t>(small_lo - coop_end, 1) * sizeof(int64_t)));
    unsigned int *d_ndefer = ctx->d_k2.as<unsigned int>();
    int64_t *d_defer = (int64_t *)(ctx->d_k2.as<char>() + 256);
    // The lane post's order of [coop_end, small_lo) by (m, fcrit crossing), built
    // on the main stream before the side stream's kernels start (beside them
    // its small grids wait for free CUs)
    const int32_t *cross_list = nullptr;
    const int64_t cross_lo = coop_end, cross_hi = small_lo;
    int cmax = 1, cbits = 1;  // the class's largest m; the keys take 2 * cbits bits
    for (int x = std::min(std::max(ctx->coop_m, 1), nb - 1); x > small_m; x--)
        if (hist[x]) { cmax = x; break; }
    while ((1 << cbits) <= cmax) cbits++;
    // The walk table of the same class (k_walk_strain): per peak its crossing
    // mc and the prefix sum of mc, then one record per shell 1..mc; sized for
    // the class's shells (mc <= m), known here from the m histogram.
    int64_t *d_wmc = nullptr, *d_weoff = nullptr;
    WalkRec *d_wtab = nullptr;
    int64_t wrec_max = 0;
    for (int x = small_m + 1; x <= cmax && x < nb; x++) wrec_max += (int64_t)hist[x] * x;
    auto cross_order = [&]() -> int {
        const int64_t nl = cross_hi - cross_lo;
        if (ctx->post_width > 1 || nl <= 64) return PP_OK;
        PP_CHECK(ctx->d_walk.grow(al((size_t)nl * 8) * 2 + (size_t)std::max<int64_t>(wrec_max, 1) * sizeof(WalkRec)));
        d_wmc = ctx->d_walk.as<int64_t>();
        d_weoff = (int64_t *)(ctx->d_walk.as<char>() + al((size_t)nl * 8));
        d_wtab = (WalkRec *)(ctx->d_walk.as<char>() + 2 * al((size_t)nl * 8));
        uint32_t *ka = (uint32_t *)W.ka, *kb = (uint32_t *)W.kb;
        hipLaunchKernelGGL(k_crossing_keys, dim3(grid1d(ctx, nl)), dim3(256), 0, ctx->stream, (const PeakParam *)pps,
                           cross_lo, cross_hi, d_rec, (const int64_t *)d_off, C.rad, cmax, cbits, ka, d_wmc);
        hipLaunchKernelGGL(k_iota_from, dim3(grid1d(ctx, nl)), dim3(256), 0, ctx->stream, W.iota, nl, (int32_t)cross_lo);
        PP_HIP(hipGetLastError());
        // stable: equal keys keep the caller's order (neighbouring profiles and result rows)
        size_t tmp = 0, tmp2 = 0;
        PP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ka, kb, W.iota, W.mo, (int)nl, 0, 2 * cbits,
                                                  ctx->stream));
        PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, d_wmc, d_weoff, (int)nl, ctx->stream));
        PP_CHECK(ctx->d_cub.grow(std::max(tmp, tmp2)));
        PP_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->d_cub.ptr, tmp, ka, kb, W.iota, W.mo, (int)nl, 0, 2 * cbits,
                                                  ctx->stream));
        PP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_cub.ptr, tmp2, d_wmc, d_weoff, (int)nl, ctx->stream));
        cross_list = W.mo;
        return PP_OK;
    };
    // the walk table's records (after the gather's profiles and the crossings)
    auto walk_table = [&]() -> int {
        if (!d_wtab) return PP_OK;
        // classes: consecutive runs of the lane-post order (m descending)
        WalkClasses K{};
        int64_t want[kWalkClasses], tot = 0, pos = 0;
        for (int k = 0; k < kWalkClasses; k++) {
            const int mlo = std::max(kWalkClassM[k], small_m + 1);
            const int mhi = k == 0 ? cmax : kWalkClassM[k - 1] - 1;
            int64_t cnt = 0;
            for (int x = mlo; x <= std::min(mhi, cmax) && x < nb; x++) cnt += hist[x];
            K.lo[k] = pos;
            K.hi[k] = pos += cnt;
            want[k] = (cnt * kWalkClassW[k] + 255) / 256;
            tot += want[k];
        }
        if (pos != cross_hi - cross_lo) {
            set_error("walk classes hold %lld of %lld peaks", (long long)pos, (long long)(cross_hi - cross_lo));
            return PP_ERR_STATE;
        }
        const int64_t cap = (int64_t)ctx->num_cu * kWalkGridCU;
        K.blk[0] = 0;
        for (int k = 0; k < kWalkClasses; k++) {
            int64_t b = want[k];
            if (tot > cap && b > 0) b = std::max<int64_t>(1, b * cap / tot);
            K.blk[k + 1] = K.blk[k] + (int32_t)b;
        }
        if (K.blk[kWalkClasses] == 0) return PP_OK;
        hipLaunchKernelGGL(k_walk_group, dim3(K.blk[kWalkClasses]), dim3(256), 0, ctx->stream,
                           (const PeakParam *)pps, cross_lo, cross_list, d_rec, (const int64_t *)d_off, C,
                           (const int64_t *)d_weoff, (const int64_t *)d_wmc, d_wtab, K);
        PP_HIP(hipGetLastError());
        return PP_OK;
    };
    auto post_lane = [&](int64_t lo, int64_t hi, bool full) -> int {
        if (hi <= lo) return PP_OK;
        const int gp = (int)std::max<int64_t>(1, std::min<int64_t>((hi - lo + 63) / 64,
                                                                   (int64_t)ctx->num_cu * kLaneGridCU));
        if (full) {
            hipLaunchKernelGGL(k_post_lane<true>, dim3(gp), dim3(64), 0, ctx->side, (const PeakParam *)pps, lo, hi,
                               d_rec, (const int64_t *)d_off, C, S.R, G.lap != nullptr ? 1 : 0, 0x7fffffff, d_defer,
                               d_ndefer, (const int32_t *)nullptr, (const WalkRec *)nullptr,
                               (const int64_t *)nullptr, (const int64_t *)nullptr);
        } else {
            const int budget = ctx->lane_walk > 0 ? ctx->lane_walk : 0x7fffffff;
            hipLaunchKernelGGL(k_fill_i32b, dim3(1), dim3(64), 0, ctx->stream, (int *)d_ndefer, (int64_t)1, 0);
            // lanes of a wave start at the same crossing where possible (the
            // order is built before the fork, see cross_order)
            const bool cross = lo == cross_lo && hi == cross_hi;
            const int32_t *list = cross ? cross_list : nullptr;
            hipLaunchKernelGGL(k_post_lane<false>, dim3(gp), dim3(64), 0, ctx->stream, (const PeakParam *)pps, lo, hi,
                               d_rec, (const int64_t *)d_off, C, S.R, 0, budget, d_defer, d_ndefer, list,
                               (const WalkRec *)(cross ? d_wtab : nullptr), (const int64_t *)d_weoff,
                               (const int64_t *)d_wmc);
            if (ctx->lane_walk > 0) {
                const int gw = (int)std::max<int64_t>(1, std::min<int64_t>(hi - lo, (int64_t)ctx->num_cu * 16));
                hipLaunchKernelGGL(k_post_defer, dim3(gw), dim3(64), 0, ctx->stream, (const PeakParam *)pps,
                                   (const int64_t *)d_defer, (const unsigned int *)d_ndefer, d_rec,
                                   (const int64_t *)d_off, C, S.R);
            }
        }
        PP_HIP(hipGetLastError());
        return PP_OK;
    };
    // The small peaks (own profile records, own result rows) are gathered and
    // post-processed on the side stream, beside the main stream's work from
    // the fork on; the main stream joins them before the call returns.
    // the wave post (m > coop_m, one wave per peak) on stream s, between events 13 and 10
    auto wave_post = [&](hipStream_t s) -> int {
        PP_HIP(hipEventRecord(ctx->ev[13], s));
        if (coop_end > 0) {
            const int gw = (int)std::max<int64_t>(1, std::min<int64_t>(coop_end, (int64_t)ctx->num_cu * kWaveGridCU));
            hipLaunchKernelGGL(k_post_wave, dim3(gw), dim3(64), 0, s, (const PeakParam *)pps, (int64_t)0, coop_end,
                               d_rec, (const int64_t *)d_off, C, S.R);
            PP_HIP(hipGetLastError());
        }
        // on the side stream event 10 marks its end; on the main stream, event 5
        PP_HIP(hipEventRecord(ctx->ev[s == ctx->side ? 10 : 5], s));
        return PP_OK;
    };
    auto side_work = [&]() -> int {
        PP_HIP(hipEventRecord(ctx->ev[11], ctx->stream));
        PP_HIP(hipStreamWaitEvent(ctx->side, ctx->ev[11], 0));
        // first, so that the finish pass (which needs its statuses) can start
        // as soon as the lane post ends
        PP_CHECK(wave_post(ctx->side));
        if (nblk > 0) {
            SmallArgs A;
            A.F = G; A.pps = pps; A.tab = st.d_scells.as<CellSlot>();
            A.seg = st.d_seg.as<SmallSeg>();
            A.part = (const SmallPart *)(st.d_seg.as<char>() + seg_bytes);
            for (int x = 0; x <= NXCD; x++) A.xp[x] = sxp[x];
            for (int x = 0; x < NXCD; x++) A.xlen[x] = sxlen[x];
            A.nx = nx;
            A.prof = d_rec; A.prof_off = d_off; A.rad = C.rad;
            int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nblk, (int64_t)ctx->num_cu * kSmallGridCU));
            grid = (grid + nx - 1) / nx * nx;
            hipLaunchKernelGGL(k_small, dim3(grid), dim3(64), 0, ctx->side, A);
            PP_HIP(hipGetLastError());
        }
        PP_HIP(hipEventRecord(ctx->ev[9], ctx->side));
        PP_CHECK(post_lane(small_lo, np, true));
        if (np > small_lo) {
            hipLaunchKernelGGL(k_small_finish, dim3(grid1d(ctx, np - small_lo)), dim3(256), 0, ctx->side,
                               (const PeakParam *)pps, small_lo, np, d_rec, (const int64_t *)d_off, C, S.R,
                               G.lap != nullptr ? 1 : 0);
            PP_HIP(hipGetLastError());
        }
        PP_HIP(hipEventRecord(ctx->ev[14], ctx->side));
        return PP_OK;
    };
    PP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev[12], 0));  // the packed fields
    PP_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    // 4. gathers
    if (small_lo > 0) {
        const size_t lds = (size_t)cs * 2 * P_NB * sizeof(float);
        const size_t lds_max = std::min<size_t>(ctx->lds_per_block, 160 * 1024);
        const int waves_cu = std::max(1, std::min<int>(ctx->gather_waves_cu > 0 ? ctx->gather_waves_cu : 16,
                                                       (int)(lds_max / lds)));
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(small_lo, (int64_t)ctx->num_cu * waves_cu));
        hipLaunchKernelGGL(k_fill_i32b, dim3(1), dim3(256), 0, ctx->stream, d_q, (int64_t)(64 * NXCD), 0);
        GatherArgs A;
        A.F = G; A.pps = pps;
        A.desc = st.d_desc.as<GroupDesc>();
        A.full_off = st.d_full_off.as<int64_t>(); A.full_nb = st.d_full_nb.as<int32_t>();
        A.part_off = st.d_part_off.as<int64_t>(); A.part_nb = st.d_part_nb.as<int32_t>();
        A.prof = d_rec; A.prof_off = d_off; A.rad = C.rad; A.queue = d_q; A.cs = cs; A.steps = st.gsteps;
        A.xorder = W.xg; A.nx = nx;
        for (int x = 0; x <= NXCD; x++) A.xb[x] = xbg[x];
        if (lds > 64 * 1024)
            (void)hipFuncSetAttribute((const void *)k_gather, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_gather, dim3(grid), dim3(64), lds, ctx->stream, A);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(ctx->ev[4], ctx->stream));
    PP_CHECK(cross_order());
    PP_CHECK(side_work());  // fork after the wave gather (measured best: beside the posts, not the gather)
    // 5. posts: the wave post runs on the side stream (before the small
    // peaks' work), beside the lane post on the main stream
    PP_HIP(hipEventRecord(ctx->ev[5], ctx->stream));
    if (ctx->post_width <= 1) {
        PP_CHECK(walk_table());
        PP_CHECK(post_lane(coop_end, small_lo, false));
    } else if (small_lo > coop_end) {
        const int W = ctx->post_width;
        const int64_t nwav = (small_lo - coop_end + 64 / W - 1) / (64 / W);
        const int gp = (int)std::max<int64_t>(1, std::min<int64_t>(nwav, (int64_t)ctx->num_cu * 16));
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(gp), dim3(64), 0, ctx->stream, (const PeakParam *)pps, coop_end, small_lo,
                               d_rec, (const int64_t *)d_off, C, S.R);
        };
        if (W == 2) launch(k_post_group<2>);
        else if (W == 4) launch(k_post_group<4>);
        else if (W == 8) launch(k_post_group<8>);
        else launch(k_post_group<16>);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(ctx->ev[6], ctx->stream));
    // 6. finish for the wave-gathered peaks (after the wave post's statuses)
    PP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev[10], 0));
    if (small_lo > 0) {
        int gf = (int)std::max<int64_t>(1, std::min<int64_t>(small_lo, (int64_t)ctx->num_cu * kFinishGridCU));
        gf = (gf + nx - 1) / nx * nx;
        FinishArgs A;
        A.F = G; A.pps = pps;
        A.xorder = W.xf; A.nx = nx;
        for (int x = 0; x <= NXCD; x++) A.xb[x] = xb[x];
        A.desc = st.d_desc.as<GroupDesc>();
        A.full_off = st.d_full_off.as<int64_t>(); A.full_nb = st.d_full_nb.as<int32_t>();
        A.part_off = st.d_part_off.as<int64_t>(); A.part_nb = st.d_part_nb.as<int32_t>();
        A.cs = cs; A.C = C; A.R = S.R;
        hipLaunchKernelGGL(k_finish, dim3(gf), dim3(64), 0, ctx->stream, A);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(ctx->ev[7], ctx->stream));
    // join: everything after this call on the context stream sees the small peaks' results
    PP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev[14], 0));
    if (cnt) {
        // profile export: counts and Fbar per shell in the caller's order
        std::vector<float> hr((size_t)tot * PSR);
        PP_HIP(hipMemcpyAsync(hr.data(), d_rec.a, hr.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        std::vector<PeakParam> hpp(np);
        PP_HIP(hipMemcpyAsync(hpp.data(), pps, np * sizeof(PeakParam), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<int64_t> start_of(np);
        int64_t t2 = 0;
        for (int64_t i = 0; i < np; i++) { start_of[hpp[i].order] = t2; t2 += hpp[i].m; }
        cnt->resize(tot);
        fbar->resize(tot);
        offs->assign(np, 0);
        int64_t o2 = 0;
        for (int64_t i = 0; i < np; i++) {
            const int m = hpp[i].m;
            const int64_t so = start_of[hpp[i].order];
            (*offs)[hpp[i].order] = so;
            for (int s = 0; s < m; s++) {
                const float *r = &hr[(size_t)(o2 + s) * PSR];
                (*cnt)[so + s] = (int32_t)r[P_N - PSR];
                double fb;
                memcpy(&fb, r + (P_FB - PSR), 8);
                (*fbar)[so + s] = fb;
            }
            o2 += m;
        }
    }
    // stage times
    PP_HIP(hipEventRecord(ctx->ev[8], ctx->stream));
    PP_HIP(hipEventSynchronize(ctx->ev[8]));
    // main stream: 2 start | prepare | 3 | k_gather | 4 | 5 | lane post | 6 |
    // finish | 7 | join | 8;  side: 11 (fork) | 13 | wave post | 10 | k_small |
    // 9 | small post | 14
    const int from[Timing::NSTAGE] = {2, 3, 10, 13, 5, 6, 9};
    const int to[Timing::NSTAGE] = {3, 4, 9, 10, 6, 7, 14};
    for (int q = 0; q < Timing::NSTAGE; q++) {
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, ctx->ev[from[q]], ctx->ev[to[q]]);
        ctx->timing.stage_ms[q] = t;
    }
    ctx->timing.gather_ms = ctx->timing.stage_ms[1] + ctx->timing.stage_ms[2];
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



}  // namespace pp
