// pp_merge.hip -- the catalogue merger on the device: Lagrangian exclusion
// and volume reduction of a halo catalogue.
//
// Reference: merge_catalog (src/Merger/Merger.jl:21-79) over
// lagrangian_exclusion! / volume_reduction! / sphere_overlap and the 256^3
// spatial hash (src/Merger/Exclusion.jl:14-222).
//
// The reference walks the halos serially in descending radius order (stable
// sort); a surviving halo kills every surviving, not larger halo whose centre
// lies inside its radius.  That greedy pass is the same as: halo j survives
// iff no surviving halo earlier in the order contains j's centre (a halo
// earlier in the order is never smaller).  Here:
// * the hash is a sort of the halos by their cell (cell ranges in a dense
//   256^3 table), the order a stable radix sort of the radii;
// * every halo k lists, from its own radius' cells, the later halos j it
//   contains ("killer" edges k -> j, stored by j);
// * rounds over the undecided halos settle them: j dies as soon as one of
//   its killers is known to survive, survives once all its killers are known
//   dead; the earliest undecided halo always settles, so the rounds end;
// * the reduction lists every overlapping surviving pair once, from its
//   larger member (whose cells within twice its radius hold every smaller
//   partner), with both cap volumes of sphere_overlap; each halo's deficit
//   is its own list's caps in list order plus its caps as the smaller member
//   in pair order (a stable sort), so the result does not depend on thread
//   timing.  The deficit is the reference's (which adds each pair from both
//   sides and halves) up to Float64 rounding.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "pp_common.h"

namespace pp {

constexpr int kMergeNC = 256;

struct MergeGeo {
    double ox, oy, oz, cs;  // hash origin and cell size
};

struct MergeIn {
    const float *x, *y, *z, *r;
    int64_t n;
};

__device__ __forceinline__ int cell_axis(double p, double o, double cs)
{
    int c = (int)floor((p - o) / cs) + 1;
    c = c < 1 ? 1 : (c > kMergeNC ? kMergeNC : c);
    return c - 1;
}

__device__ __forceinline__ int cell_of(const MergeIn &H, const MergeGeo &G, int64_t i, int &cx, int &cy, int &cz)
{
    cx = cell_axis((double)H.x[i], G.ox, G.cs);
    cy = cell_axis((double)H.y[i], G.oy, G.cs);
    cz = cell_axis((double)H.z[i], G.oz, G.cs);
    return cx + kMergeNC * (cy + kMergeNC * cz);
}

__device__ __forceinline__ double dist_of(const MergeIn &H, int64_t a, int64_t b)
{
#pragma clang fp contract(off)
    const double dx = (double)H.x[b] - (double)H.x[a], dy = (double)H.y[b] - (double)H.y[a],
                 dz = (double)H.z[b] - (double)H.z[a];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// sphere_overlap (Exclusion.jl:64-96): cap volumes cut from spheres 1 and 2
__device__ __forceinline__ void sphere_caps(double d, double r1, double r2, double &v1, double &v2)
{
#pragma clang fp contract(off)
    v1 = v2 = 0.0;
    if (d >= r1 + r2 || d <= fabs(r1 - r2)) return;
    const double h1 = (r2 - r1 + d) * (r2 + r1 - d) / (2.0 * d);
    const double h2 = (r1 - r2 + d) * (r1 + r2 - d) / (2.0 * d);
    const double pi3 = 3.141592653589793 / 3.0;
    v1 = pi3 * (h1 * h1) * (3.0 * r1 - h1);
    v2 = pi3 * (h2 * h2) * (3.0 * r2 - h2);
}

// min/max of the positions and max radius (one block)
__global__ void __launch_bounds__(1024) k_merge_bounds(MergeIn H, double *__restrict__ out)
{
    __shared__ float red[7][1024];
    float v[7] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = threadIdx.x; i < H.n; i += blockDim.x) {
        v[0] = fminf(v[0], H.x[i]);
        v[1] = fmaxf(v[1], H.x[i]);
        v[2] = fminf(v[2], H.y[i]);
        v[3] = fmaxf(v[3], H.y[i]);
        v[4] = fminf(v[4], H.z[i]);
        v[5] = fmaxf(v[5], H.z[i]);
        v[6] = fmaxf(v[6], H.r[i]);
    }
    for (int q = 0; q < 7; q++) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 7; q++) {
                const float a = red[q][threadIdx.x], b = red[q][threadIdx.x + w];
                red[q][threadIdx.x] = (q & 1) || q == 6 ? fmaxf(a, b) : fminf(a, b);
            }
        __syncthreads();
    }
    if (threadIdx.x < 7) out[threadIdx.x] = (double)red[threadIdx.x][0];
}

__global__ void k_merge_keys(MergeIn H, MergeGeo G, int32_t *__restrict__ cell, int32_t *__restrict__ iota)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < H.n; i += (int64_t)gridDim.x * blockDim.x) {
        int cx, cy, cz;
        cell[i] = cell_of(H, G, i, cx, cy, cz);
        iota[i] = (int32_t)i;
    }
}

__global__ void k_fill_i32(int32_t *__restrict__ a, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

// [start, end) of every occupied cell in the cell-sorted list
__global__ void k_cell_ranges(const int32_t *__restrict__ key, int64_t n, int32_t *__restrict__ start,
                              int32_t *__restrict__ end)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = key[p];
        if (p == 0 || key[p - 1] != k) start[k] = (int32_t)p;
        if (p == n - 1 || key[p + 1] != k) end[k] = (int32_t)(p + 1);
    }
}

__global__ void k_rank(const int32_t *__restrict__ order, int64_t n, int32_t *__restrict__ rank)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        rank[order[p]] = (int32_t)p;
}

struct Hash {
    const int32_t *start, *end, *list;  // cell ranges, halos in cell order
};

// pass 1 edges: every later halo j inside halo k's radius; count (fill = 0)
// or write (fill = 1) them into j's list
__global__ void __launch_bounds__(64) k_killers(MergeIn H, MergeGeo G, Hash S, const int32_t *__restrict__ rank,
                                                int32_t *__restrict__ count, const int64_t *__restrict__ off,
                                                int32_t *__restrict__ edges, int fill)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < H.n; k += (int64_t)gridDim.x * blockDim.x) {
        const double rk = (double)H.r[k];
        const int dc = (int)ceil(rk / G.cs);
        int cx, cy, cz;
        cell_of(H, G, k, cx, cy, cz);
        const int rkk = rank[k];
        for (int jz = max(cz - dc, 0); jz <= min(cz + dc, kMergeNC - 1); jz++)
            for (int jy = max(cy - dc, 0); jy <= min(cy + dc, kMergeNC - 1); jy++)
                for (int jx = max(cx - dc, 0); jx <= min(cx + dc, kMergeNC - 1); jx++) {
                    const int c = jx + kMergeNC * (jy + kMergeNC * jz);
                    for (int32_t p = S.start[c]; p < S.end[c]; p++) {
                        const int32_t j = S.list[p];
                        if (rank[j] <= rkk || dist_of(H, k, j) >= rk) continue;
                        if (!fill) atomicAdd(&count[j], 1);
                        else edges[off[j] + atomicAdd(&count[j], 1)] = (int32_t)k;
                    }
                }
    }
}

// one round: 0 undecided, 1 survives, 2 excluded
__global__ void k_decide(int64_t n, const int64_t *__restrict__ off, const int32_t *__restrict__ edges,
                         int32_t *state, int32_t *__restrict__ undecided)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        if (state[j] != 0) continue;
        bool all_dead = true, killed = false;
        for (int64_t e = off[j]; e < off[j + 1]; e++) {
            const int32_t s = __atomic_load_n(&state[edges[e]], __ATOMIC_RELAXED);
            if (s == 1) {
                killed = true;
                break;
            }
            if (s == 0) all_dead = false;
        }
        if (killed) __atomic_store_n(&state[j], 2, __ATOMIC_RELAXED);
        else if (all_dead) __atomic_store_n(&state[j], 1, __ATOMIC_RELAXED);
        else atomicAdd(undecided, 1);
    }
}

// pass 2 pairs from the larger member i: count (fill = 0) or write
__global__ void __launch_bounds__(64) k_pairs(MergeIn H, MergeGeo G, Hash S, const int32_t *__restrict__ rank,
                                              const int32_t *__restrict__ state, int64_t *__restrict__ count,
                                              const int64_t *__restrict__ off, int32_t *__restrict__ pj,
                                              double *__restrict__ cap_i, double *__restrict__ cap_j, int fill)
{
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < H.n; i += (int64_t)gridDim.x * blockDim.x) {
        if (state[i] != 1) {
            if (!fill) count[i] = 0;
            continue;
        }
        const double ri = (double)H.r[i];
        const int dc = (int)ceil(2.0 * ri / G.cs);
        int cx, cy, cz;
        cell_of(H, G, i, cx, cy, cz);
        const int rki = rank[i];
        int64_t q = fill ? off[i] : 0;
        for (int jz = max(cz - dc, 0); jz <= min(cz + dc, kMergeNC - 1); jz++)
            for (int jy = max(cy - dc, 0); jy <= min(cy + dc, kMergeNC - 1); jy++)
                for (int jx = max(cx - dc, 0); jx <= min(cx + dc, kMergeNC - 1); jx++) {
                    const int c = jx + kMergeNC * (jy + kMergeNC * jz);
                    for (int32_t p = S.start[c]; p < S.end[c]; p++) {
                        const int32_t j = S.list[p];
                        if (rank[j] <= rki || state[j] != 1) continue;
                        const double rj = (double)H.r[j];
                        const double d = dist_of(H, i, j);
                        if (!(d < ri + rj && d > fabs(ri - rj))) continue;
                        if (fill) {
                            double v1, v2;
                            sphere_caps(d, ri, rj, v1, v2);
                            pj[q] = j;
                            cap_i[q] = v1;
                            cap_j[q] = v2;
                        }
                        q++;
                    }
                }
        if (!fill) count[i] = q;
    }
}

// reduced radii: deficit = own pair list (larger member) + smaller-member caps
__global__ void k_finalize(MergeIn H, const int32_t *__restrict__ state, const int64_t *__restrict__ off,
                           const double *__restrict__ cap_i, const int32_t *__restrict__ sstart,
                           const int32_t *__restrict__ send, const double *__restrict__ cap_sorted,
                           int8_t *__restrict__ keep, float *__restrict__ new_r, int32_t *__restrict__ nkept)
{
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < H.n; i += (int64_t)gridDim.x * blockDim.x) {
        if (state[i] != 1) {
            keep[i] = 0;
            new_r[i] = 0.0f;
            continue;
        }
        double a = 0.0, b = 0.0;
        for (int64_t q = off[i]; q < off[i + 1]; q++) a += cap_i[q];
        for (int32_t q = sstart[i]; q < send[i]; q++) b += cap_sorted[q];
        const double r = (double)H.r[i];
        const double f = 4.0 * 3.141592653589793 / 3.0;
        const double vol = f * (r * r * r) - (a + b);
        if (vol <= 0.0) {
            keep[i] = 0;
            new_r[i] = 0.0f;
        } else {
            keep[i] = 1;
            new_r[i] = (float)pow(vol / f, 1.0 / 3.0);
            atomicAdd(nkept, 1);
        }
    }
}

struct I32ToI64 {
    __device__ __forceinline__ int64_t operator()(const int32_t &v) const { return (int64_t)v; }
};

// device buffers of one call, released on return
struct Scratch {
    std::vector<DevBuf> bufs;
    template <class T> int get(size_t n, T *&p)
    {
        bufs.emplace_back();
        PP_CHECK(bufs.back().grow(std::max<size_t>(n, 1) * sizeof(T)));
        p = bufs.back().as<T>();
        return PP_OK;
    }
    ~Scratch()
    {
        for (auto &b : bufs) b.release();
    }
};

static int grid_n(pp_ctx *ctx, int64_t n, int block = 256)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, (int64_t)ctx->num_cu * 16));
}

__global__ void k_fill_i64(int64_t *__restrict__ a, int64_t n, int64_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

// off[0] = 0, off[q + 1] = counts[0] + ... + counts[q]; *total = off[n]
template <class It>
static int offsets(pp_ctx *ctx, It counts, int64_t *off, int64_t n, int64_t *total)
{
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, ctx->stream, off, (int64_t)1, (int64_t)0);
    PP_HIP(hipGetLastError());
    size_t tmp = 0;
    PP_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, counts, off + 1, (int)n, ctx->stream));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceScan::InclusiveSum(ctx->d_cub.ptr, tmp, counts, off + 1, (int)n, ctx->stream));
    PP_HIP(hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" int pp_merge_halos(pp_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                              const float *rthl, int on_device, int8_t *keep, float *new_rthl, int64_t *nkept)
{
    PP_ARG(ctx && x && y && z && rthl && keep && new_rthl && nkept, "NULL argument");
    PP_ARG(n >= 0 && n < (1ll << 31), "halo count %lld outside [0, 2^31)", (long long)n);
    *nkept = 0;
    if (n == 0) return PP_OK;
    PP_CHECK(ensure_device(ctx));
    hipStream_t s = ctx->stream;
    Scratch W;
    MergeIn H;
    H.n = n;
    const hipMemcpyKind up = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    {
        float *d;
        PP_CHECK(W.get(4 * (size_t)n, d));
        const float *src[4] = {x, y, z, rthl};
        for (int q = 0; q < 4; q++) PP_HIP(hipMemcpyAsync(d + q * n, src[q], n * 4, up, s));
        H.x = d;
        H.y = d + n;
        H.z = d + 2 * n;
        H.r = d + 3 * n;
    }
    // domain (Merger.jl:40-50) and cell size (Exclusion.jl:33-38)
    double *d_b, hb[7];
    PP_CHECK(W.get(7, d_b));
    hipLaunchKernelGGL(k_merge_bounds, dim3(1), dim3(1024), 0, s, H, d_b);
    PP_HIP(hipGetLastError());
    PP_HIP(hipMemcpyAsync(hb, d_b, sizeof(hb), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < 7; q++) PP_ARG(std::isfinite(hb[q]), "non-finite position or radius in the catalogue");
    const double pad = 2.0 * hb[6];
    const double lo[3] = {hb[0] - pad, hb[2] - pad, hb[4] - pad}, hi[3] = {hb[1] + pad, hb[3] + pad, hb[5] + pad};
    const double L = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    PP_ARG(L > 0, "degenerate catalogue extent");
    MergeGeo G{lo[0], lo[1], lo[2], L / kMergeNC};
    const int64_t ncell = (int64_t)kMergeNC * kMergeNC * kMergeNC;
    // hash: halos sorted by cell, ranges per cell
    int32_t *cell, *iota, *cell_s, *list, *cstart, *cend;
    PP_CHECK(W.get(n, cell));
    PP_CHECK(W.get(n, iota));
    PP_CHECK(W.get(n, cell_s));
    PP_CHECK(W.get(n, list));
    PP_CHECK(W.get(ncell, cstart));
    PP_CHECK(W.get(ncell, cend));
    hipLaunchKernelGGL(k_merge_keys, dim3(grid_n(ctx, n)), dim3(256), 0, s, H, G, cell, iota);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, ncell)), dim3(256), 0, s, cstart, ncell, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, ncell)), dim3(256), 0, s, cend, ncell, 0);
    PP_HIP(hipGetLastError());
    size_t tmp = 0;
    PP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, cell, cell_s, iota, list, (int)n, 0, 24, s));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->d_cub.ptr, tmp, cell, cell_s, iota, list, (int)n, 0, 24, s));
    hipLaunchKernelGGL(k_cell_ranges, dim3(grid_n(ctx, n)), dim3(256), 0, s, (const int32_t *)cell_s, n, cstart, cend);
    PP_HIP(hipGetLastError());
    const Hash S{cstart, cend, list};
    // order: stable descending radius; rank of each halo
    float *r_s;
    int32_t *order, *rank;
    PP_CHECK(W.get(n, r_s));
    PP_CHECK(W.get(n, order));
    PP_CHECK(W.get(n, rank));
    tmp = 0;
    PP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, H.r, r_s, iota, order, (int)n, 0, 32, s));
    PP_CHECK(ctx->d_cub.grow(tmp));
    PP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(ctx->d_cub.ptr, tmp, H.r, r_s, iota, order, (int)n, 0, 32, s));
    hipLaunchKernelGGL(k_rank, dim3(grid_n(ctx, n)), dim3(256), 0, s, (const int32_t *)order, n, rank);
    PP_HIP(hipGetLastError());
    // pass 1: killer edges by target, then rounds
    int32_t *cnt, *state, *d_und;
    int64_t *off;
    PP_CHECK(W.get(n, cnt));
    PP_CHECK(W.get(n, state));
    PP_CHECK(W.get(1, d_und));
    PP_CHECK(W.get(n + 1, off));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, n)), dim3(256), 0, s, cnt, n, 0);
    hipLaunchKernelGGL(k_killers, dim3(grid_n(ctx, n, 64)), dim3(64), 0, s, H, G, S, (const int32_t *)rank, cnt,
                       (const int64_t *)nullptr, (int32_t *)nullptr, 0);
    PP_HIP(hipGetLastError());
    int64_t nedge = 0;
    {
        hipcub::TransformInputIterator<int64_t, I32ToI64, const int32_t *> it(cnt, I32ToI64());
        PP_CHECK(offsets(ctx, it, off, n, &nedge));
    }
    int32_t *edges;
    PP_CHECK(W.get((size_t)nedge, edges));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, n)), dim3(256), 0, s, cnt, n, 0);
    hipLaunchKernelGGL(k_killers, dim3(grid_n(ctx, n, 64)), dim3(64), 0, s, H, G, S, (const int32_t *)rank, cnt,
                       (const int64_t *)off, edges, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, n)), dim3(256), 0, s, state, n, 0);
    PP_HIP(hipGetLastError());
    for (int64_t round = 0;; round++) {
        PP_ARG(round <= n, "exclusion rounds did not settle");
        int32_t und = 0;
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_und, (int64_t)1, 0);
        hipLaunchKernelGGL(k_decide, dim3(grid_n(ctx, n)), dim3(256), 0, s, n, (const int64_t *)off,
                           (const int32_t *)edges, state, d_und);
        PP_HIP(hipGetLastError());
        PP_HIP(hipMemcpyAsync(&und, d_und, 4, hipMemcpyDeviceToHost, s));
        PP_HIP(hipStreamSynchronize(s));
        if (und == 0) break;
    }
    // pass 2: overlapping surviving pairs from the larger member
    int64_t *pcnt, *poff;
    PP_CHECK(W.get(n, pcnt));
    PP_CHECK(W.get(n + 1, poff));
    hipLaunchKernelGGL(k_pairs, dim3(grid_n(ctx, n, 64)), dim3(64), 0, s, H, G, S, (const int32_t *)rank,
                       (const int32_t *)state, pcnt, (const int64_t *)nullptr, (int32_t *)nullptr, (double *)nullptr,
                       (double *)nullptr, 0);
    PP_HIP(hipGetLastError());
    int64_t npair = 0;
    PP_CHECK(offsets(ctx, (const int64_t *)pcnt, poff, n, &npair));
    PP_ARG(npair < (1ll << 31), "too many overlapping pairs (%lld)", (long long)npair);
    int32_t *pj, *pj_s, *sstart, *send;
    double *cap_i, *cap_j, *cap_s;
    PP_CHECK(W.get((size_t)npair, pj));
    PP_CHECK(W.get((size_t)npair, pj_s));
    PP_CHECK(W.get((size_t)npair, cap_i));
    PP_CHECK(W.get((size_t)npair, cap_j));
    PP_CHECK(W.get((size_t)npair, cap_s));
    PP_CHECK(W.get(n, sstart));
    PP_CHECK(W.get(n, send));
    hipLaunchKernelGGL(k_pairs, dim3(grid_n(ctx, n, 64)), dim3(64), 0, s, H, G, S, (const int32_t *)rank,
                       (const int32_t *)state, pcnt, (const int64_t *)poff, pj, cap_i, cap_j, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, n)), dim3(256), 0, s, sstart, n, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_n(ctx, n)), dim3(256), 0, s, send, n, 0);
    PP_HIP(hipGetLastError());
    if (npair > 0) {
        // the smaller members' caps grouped by halo, pair order kept (stable)
        tmp = 0;
        PP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, pj, pj_s, cap_j, cap_s, (int)npair, 0, 32, s));
        PP_CHECK(ctx->d_cub.grow(tmp));
        PP_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->d_cub.ptr, tmp, pj, pj_s, cap_j, cap_s, (int)npair, 0, 32, s));
        hipLaunchKernelGGL(k_cell_ranges, dim3(grid_n(ctx, npair)), dim3(256), 0, s, (const int32_t *)pj_s, npair,
                           sstart, send);
        PP_HIP(hipGetLastError());
    }
    int32_t *d_nk, hk = 0;
    PP_CHECK(W.get(1, d_nk));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_nk, (int64_t)1, 0);
    int8_t *d_keep = keep;
    float *d_nr = new_rthl;
    if (!on_device) {
        PP_CHECK(W.get(n, d_keep));
        PP_CHECK(W.get(n, d_nr));
    }
    hipLaunchKernelGGL(k_finalize, dim3(grid_n(ctx, n)), dim3(256), 0, s, H, (const int32_t *)state,
                       (const int64_t *)poff, (const double *)cap_i, (const int32_t *)sstart, (const int32_t *)send,
                       (const double *)cap_s, d_keep, d_nr, d_nk);
    PP_HIP(hipGetLastError());
    if (!on_device) {
        PP_HIP(hipMemcpyAsync(keep, d_keep, n, hipMemcpyDeviceToHost, s));
        PP_HIP(hipMemcpyAsync(new_rthl, d_nr, n * 4, hipMemcpyDeviceToHost, s));
    }
    PP_HIP(hipMemcpyAsync(&hk, d_nk, 4, hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    *nkept = hk;
    return PP_OK;
}
