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
#include <mutex>

#include "pp_common.h"
#include "pp_noise.h"

namespace pp {

// ------------------------------------------------------------- window
// W depends on |k| only through kx^2 + ky^2 + kz^2 with kx = dk ix,
// ky = dk fy, kz = dk fz; negating fy or fz is exact, so the four elements
// (ix, +-fy, +-fz) share one Float64 evaluation bit for bit.  A thread takes
// one (ix, |fy|, |fz|) and scales its (up to) four mirror elements; the grid
// is (ix, |fy|) linearised in x and |fz| in y.  With two windows the thread
// reads each element once and writes both products.
struct WinSel {
    float2 *out;
    int wsmooth;  // PP_WINDOW_*; < 0: no output
};

__device__ __forceinline__ double window_value(int wsmooth, double k, double fkR)
{
    if (wsmooth == PP_WINDOW_GAUSSIAN) {
        double t = fkR / 2.0;
        return exp(-(t * t) / 2.0);
    }
    if (wsmooth == PP_WINDOW_K2GAUSS) {
        double t = fkR / 2.0;
        return exp(-(t * t) / 2.0) * (k * k);
    }
    double s, c;
    sincos(fkR, &s, &c);
    return 3.0 * (s - fkR * c) / (fkR * fkR * fkR);
}

__global__ void __launch_bounds__(256) k_window(const float2 *__restrict__ in, WinSel a, WinSel b, int n, int nk,
                                                double dk, double Rf, float scale)
{
#pragma clang fp contract(off)
    const int half = n / 2;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int hy = idx / nk;
    const int ix = idx - hy * nk;
    const int hz = blockIdx.y;
    if (hy > half) return;
    const int iy2 = hy == 0 ? 0 : n - hy, iz2 = hz == 0 ? 0 : n - hz;
    const int ny = iy2 == hy ? 1 : 2, nz = iz2 == hz ? 1 : 2;
    // the (up to) four elements are loaded before the Float64 window is
    // evaluated, so their latency hides behind it
    size_t el[4];
    float2 vin[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int iy = (q & 1) ? iy2 : hy, iz = (q & 2) ? iz2 : hz;
        el[q] = (size_t)ix + (size_t)nk * ((size_t)iy + (size_t)n * (size_t)iz);
        if ((q & 1) < ny && (q >> 1) < nz) vin[q] = in[el[q]];
    }
    const double kx = dk * (double)ix, ky = dk * (double)hy, kz = dk * (double)hz;
    const double k = sqrt(kx * kx + ky * ky + kz * kz);
    const double fkR = k * Rf;
    float wa = 1.f, wb = 1.f;
    const bool scaled = fkR != 0.0;
    if (scaled) {
        wa = (float)window_value(a.wsmooth, k, fkR);
        if (b.wsmooth >= 0) wb = (float)window_value(b.wsmooth, k, fkR);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if ((q & 1) >= ny || (q >> 1) >= nz) continue;
        {
            const size_t e = el[q];
            const float2 v = vin[q];
            float2 r = v;
            if (scaled) { r.x = r.x * wa; r.y = r.y * wa; }
            r.x = r.x * scale;
            r.y = r.y * scale;
            a.out[e] = r;
            if (b.wsmooth >= 0) {
                float2 s = v;
                if (scaled) { s.x = s.x * wb; s.y = s.y * wb; }
                s.x = s.x * scale;
                s.y = s.y * scale;
                b.out[e] = s;
            }
        }
    }
}

static int launch_window2(pp_ctx *ctx, const void *in, void *out_a, int ws_a, void *out_b, int ws_b, int n,
                          double boxsize, double Rf, float scale)
{
    const int nk = n / 2 + 1, half = n / 2;
    const long long per = (long long)nk * (half + 1);
    PP_ARG(per < (1ll << 31) && half + 1 <= 65535, "mesh %d too large for the window grid", n);
    const double dk = 2.0 * M_PI / boxsize;
    dim3 grid((unsigned)((per + 255) / 256), (unsigned)(half + 1));
    WinSel a{(float2 *)out_a, ws_a}, b{(float2 *)out_b, out_b ? ws_b : -1};
    hipLaunchKernelGGL(k_window, grid, dim3(256), 0, ctx->stream, (const float2 *)in, a, b, n, nk, dk, Rf, scale);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

static int launch_window(pp_ctx *ctx, const void *in, void *out, int n, double boxsize, double Rf,
                         int wsmooth, float scale)
{
    return launch_window2(ctx, in, out, wsmooth, nullptr, -1, n, boxsize, Rf, scale);
}

// ------------------------------------------------------------- stencil
// One pass over the field per filter.  A workgroup takes a 64 x 16 tile of
// inner (x, y) columns and marches along z through kStencilKZ planes: each
// plane's (64+2) x (16+2) neighbourhood is staged in LDS, two planes per
// barrier (four buffers), every column keeps the 3x3 (x, y) maxima of planes
// z-1, z, z+1 in registers, and the strict
// 27-point test "no neighbour above ff" is ff >= max of the 27 values.  A
// wave is one x-run of 64 cells of one row, so its ballot of accepted cells
// is one 64-bit word of the row's peak bit mask.  The mask byte is read only
// for cells that pass the field test.
//
// A second, small pass turns the bit mask into the output: the exclusive
// scan of the words' popcounts in (k, j, x) order is the CPU loop's order
// (k outer, j, i inner), so each set bit is written at its rank with no
// atomics, the first max_peaks ranks are kept (PeakFind.jl:70-72) and only
// the kept peak cells are masked.
//
// Tiles are numbered x fastest, then y, then z, and handed out so that each
// XCD (workgroups go round-robin to the eight XCDs, each with its own L2)
// works through one contiguous run of them: neighbouring tiles share their
// halo rows and planes, and those lines are then read from HBM once per run
// instead of once per tile.
constexpr int kStencilBlock = 512;
constexpr int kStencilTY = 16;  // rows of the tile: each wave takes two, 8 apart
constexpr int kStencilKZ = 32;
constexpr int kStencilXCD = 8;

struct StencilTiles {
    int nbx, nby, nbz, per_xcd;  // tiles along x, y, z; tiles per XCD run
};

struct StencilGeom {
    int n1, n2, n3, nbuff;
    int in1, in2, in3;
    int words;            // 64-bit words per inner row, ceil(in1 / 64)
    long long nwords;     // words * in2 * in3
};

static StencilGeom stencil_geom(int n1, int n2, int n3, int nbuff)
{
    StencilGeom g;
    g.n1 = n1; g.n2 = n2; g.n3 = n3; g.nbuff = nbuff;
    g.in1 = std::max(0, n1 - 2 * nbuff);
    g.in2 = std::max(0, n2 - 2 * nbuff);
    g.in3 = std::max(0, n3 - 2 * nbuff);
    g.words = (g.in1 + 63) / 64;
    g.nwords = (long long)g.words * g.in2 * g.in3;
    return g;
}

// The staged (64+2) x (8+2) neighbourhood of a plane is 660 floats, three
// per thread at most.  Each thread's loads for plane z+3 are issued right
// after it stages plane z+1 (two planes in flight per thread), so the HBM
// latency hides behind the barrier and the comparisons of the planes
// between.
__global__ void __launch_bounds__(kStencilBlock) k_peak_bits(const float *__restrict__ d,
                                                             const int8_t *__restrict__ mask, StencilGeom g,
                                                             StencilTiles tl, float fcrit, uint64_t *__restrict__ bits)
{
    constexpr int LX = 64 + 2, LY = kStencilTY + 2, LN = LX * LY, NQ = (LN + kStencilBlock - 1) / kStencilBlock;
    __shared__ float tile[4][LN];  // two pairs of planes: one pair staged while the other is read
    const int t = (int)(blockIdx.x % kStencilXCD) * tl.per_xcd + (int)(blockIdx.x / kStencilXCD);
    if (t >= tl.nbx * tl.nby * tl.nbz) return;  // the whole workgroup leaves before any barrier
    const int tbx = t % tl.nbx, tby = (t / tl.nbx) % tl.nby, tbz = t / (tl.nbx * tl.nby);
    const int lane = threadIdx.x & 63, wy = threadIdx.x >> 6;
    const int x0 = g.nbuff + 64 * tbx, y0 = g.nbuff + kStencilTY * tby;
    const int z0 = g.nbuff + kStencilKZ * tbz;
    const int z1 = min(z0 + kStencilKZ, g.nbuff + g.in3);  // planes z0..z1-1; z0-1..z1 are read
    const long long s12 = (long long)g.n1 * g.n2;
    const int x = x0 + lane;
    const int ya = y0 + wy, yb = ya + kStencilTY / 2;
    const bool xin = x < g.nbuff + g.in1;
    const bool ina = xin && ya < g.nbuff + g.in2, inb = xin && yb < g.nbuff + g.in2;
    // the plane elements each thread stages (clamped to the grid: clamped
    // cells only neighbour columns outside the inner cube)
    int off[NQ];
    bool has[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = threadIdx.x + q * kStencilBlock;
        has[q] = e < LN;
        const int ly = e / LX, lx = e - ly * LX;
        const int gx = min(x0 - 1 + lx, g.n1 - 1), gy = min(y0 - 1 + ly, g.n2 - 1);
        off[q] = gx + gy * g.n1;
    }
    float ra[NQ], rb[NQ];
    auto load = [&](int z, float *r) {
        const float *p = d + (long long)z * s12;
#pragma unroll
        for (int q = 0; q < NQ; q++) r[q] = has[q] ? p[off[q]] : 0.0f;
    };
    auto store = [&](const float *r, int buf) {
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (has[q]) tile[buf][threadIdx.x + q * kStencilBlock] = r[q];
    };
    auto max9 = [&](int buf, int row, float &centre) {
        const float *t = tile[buf] + row * LX + lane;
        centre = t[LX + 1];
        float m = fmaxf(fmaxf(t[0], t[1]), t[2]);
        m = fmaxf(m, fmaxf(fmaxf(t[LX], t[LX + 1]), t[LX + 2]));
        m = fmaxf(m, fmaxf(fmaxf(t[2 * LX], t[2 * LX + 1]), t[2 * LX + 2]));
        return m;
    };
    const long long word_row = (long long)g.words * g.in2;
    float ca, cb, c1a, c1b, c2a, c2b;
    // planes z0-1 and z0 to buffers 2, 3; planes z0+1, z0+2 in flight
    load(z0 - 1, ra);
    load(z0, rb);
    store(ra, 2);
    store(rb, 3);
    load(z0 + 1, ra);
    if (z0 + 2 <= z1) load(z0 + 2, rb);
    __syncthreads();
    float pa = max9(2, wy, ca), pb = max9(2, wy + kStencilTY / 2, cb);
    float qa = max9(3, wy, ca), qb = max9(3, wy + kStencilTY / 2, cb);
    auto test = [&](int z, float c_a, float c_b, float mxa, float mxb) {
        bool ka = ina && c_a >= fcrit && !(c_a < mxa);
        bool kb = inb && c_b >= fcrit && !(c_b < mxb);
        if (ka) ka = mask[x + (long long)ya * g.n1 + (long long)z * s12] != 1;
        if (kb) kb = mask[x + (long long)yb * g.n1 + (long long)z * s12] != 1;
        const unsigned long long wa = __ballot(ka), wb = __ballot(kb);
        const long long base = (long long)(z - g.nbuff) * word_row + tbx;
        if (lane == 0 && ya < g.nbuff + g.in2) bits[base + (long long)(ya - g.nbuff) * g.words] = wa;
        if (lane == 0 && yb < g.nbuff + g.in2) bits[base + (long long)(yb - g.nbuff) * g.words] = wb;
    };
    // two planes per barrier: planes z+1, z+2 (in ra, rb) go to one pair of
    // buffers while z+3, z+4 are fetched; the pair written here was last
    // read two barriers ago
    for (int z = z0, k = 0; z < z1; z += 2, k ^= 1) {
        const int b1 = 2 * k, b2 = 2 * k + 1;
        store(ra, b1);
        if (z + 2 <= z1) store(rb, b2);
        if (z + 3 <= z1) load(z + 3, ra);
        if (z + 4 <= z1) load(z + 4, rb);
        __syncthreads();
        const float m1a = max9(b1, wy, c1a), m1b = max9(b1, wy + kStencilTY / 2, c1b);
        test(z, ca, cb, fmaxf(fmaxf(pa, qa), m1a), fmaxf(fmaxf(pb, qb), m1b));
        if (z + 1 < z1) {
            const float m2a = max9(b2, wy, c2a), m2b = max9(b2, wy + kStencilTY / 2, c2b);
            test(z + 1, c1a, c1b, fmaxf(fmaxf(qa, m1a), m2a), fmaxf(fmaxf(qb, m1b), m2b));
            pa = m1a; pb = m1b;
            qa = m2a; qb = m2b;
            ca = c2a; cb = c2b;
        }
    }
}

struct Popc {
    __device__ __forceinline__ int operator()(const uint64_t &w) const { return __popcll(w); }
};

struct PeakOut {
    int32_t *i, *j, *k, *filt;
    float *delta_s;
    long long capacity;
};

// counters[0] = peaks written so far (all filters), counters[1] = peaks found
// (capped per filter), counters[2] = first index of the current filter.
__global__ void __launch_bounds__(256) k_peak_emit(const float *__restrict__ d, int8_t *__restrict__ mask,
                                                   StencilGeom g, const uint64_t *__restrict__ bits,
                                                   const int *__restrict__ word_off,
                                                   const long long *__restrict__ counters, long long max_peaks,
                                                   int filt, PeakOut o)
{
    const long long w = (long long)blockIdx.x * 256 + threadIdx.x;
    if (w >= g.nwords) return;
    uint64_t b = bits[w];
    if (!b) return;
    const long long row = w / g.words;
    const int wx = (int)(w - row * g.words);
    const int j = g.nbuff + (int)(row % g.in2), k = g.nbuff + (int)(row / g.in2);
    long long rank = word_off[w];
    const long long base = counters[0];
    while (b) {
        if (rank >= max_peaks) return;
        const int bit = __ffsll((unsigned long long)b) - 1;
        b &= b - 1;
        const int i = g.nbuff + 64 * wx + bit;
        const long long c = i + (long long)j * g.n1 + (long long)k * g.n1 * g.n2;
        mask[c] = 1;
        const long long slot = base + rank;
        if (slot < o.capacity) {
            o.i[slot] = i;
            o.j[slot] = j;
            o.k[slot] = k;
            o.filt[slot] = filt;
            o.delta_s[slot] = d[c];
        }
        rank++;
    }
}

__global__ void k_stencil_finish(const uint64_t *__restrict__ bits, const int *__restrict__ word_off,
                                 long long nwords, long long max_peaks, long long *__restrict__ counters)
{
    long long found = nwords > 0 ? (long long)word_off[nwords - 1] + __popcll(bits[nwords - 1]) : 0;
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

// Device scratch of one stencil pass: the peak bit mask, the words' output
// offsets and the scan's temporary storage, carved from one buffer.
struct StencilScratch {
    uint64_t *bits;
    int *off;
    void *tmp;
    size_t tmp_bytes;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static int stencil_scan_bytes(const StencilGeom &g, size_t *tmp)
{
    hipcub::TransformInputIterator<int, Popc, const uint64_t *> it((const uint64_t *)nullptr, Popc());
    *tmp = 0;
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, *tmp, it, (int *)nullptr, (int)std::max(1ll, g.nwords)));
    return PP_OK;
}

static int stencil_scratch_bytes(const StencilGeom &g, size_t *bytes)
{
    size_t tmp = 0;
    PP_CHECK(stencil_scan_bytes(g, &tmp));
    *bytes = align256((size_t)g.nwords * 8) + align256((size_t)g.nwords * 4) + align256(tmp);
    return PP_OK;
}

static StencilScratch stencil_carve(void *base, const StencilGeom &g, size_t tmp)
{
    StencilScratch s;
    char *p = (char *)base;
    s.bits = (uint64_t *)p;
    p += align256((size_t)g.nwords * 8);
    s.off = (int *)p;
    p += align256((size_t)g.nwords * 4);
    s.tmp = p;
    s.tmp_bytes = tmp;
    return s;
}

// Stencil over one field with the scratch given; peaks appended after
// counters[0].
static int run_stencil_with(pp_ctx *ctx, const float *d_field, int8_t *d_mask, const StencilGeom &g,
                            float fcrit, long long max_peaks, int filt, const PeakOut &o, long long *d_counters,
                            const StencilScratch &s)
{
    PP_ARG(g.nwords < (1ll << 31), "inner cube too large for the stencil");
    if (g.nwords == 0) {
        hipLaunchKernelGGL(k_stencil_finish, dim3(1), dim3(1), 0, ctx->stream, (const uint64_t *)nullptr,
                           (const int *)nullptr, 0ll, max_peaks, d_counters);
        PP_HIP(hipGetLastError());
        return PP_OK;
    }
    StencilTiles tl;
    tl.nbx = g.words;
    tl.nby = (g.in2 + kStencilTY - 1) / kStencilTY;
    tl.nbz = (g.in3 + kStencilKZ - 1) / kStencilKZ;
    const long long ntiles = (long long)tl.nbx * tl.nby * tl.nbz;
    PP_ARG(ntiles < (1ll << 30), "inner cube too large for the stencil grid");
    tl.per_xcd = (int)((ntiles + kStencilXCD - 1) / kStencilXCD);
    hipLaunchKernelGGL(k_peak_bits, dim3((unsigned)tl.per_xcd * kStencilXCD), dim3(kStencilBlock), 0, ctx->stream,
                       d_field, (const int8_t *)d_mask, g, tl, fcrit, s.bits);
    PP_HIP(hipGetLastError());
    hipcub::TransformInputIterator<int, Popc, const uint64_t *> it(s.bits, Popc());
    size_t tmp = s.tmp_bytes;
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp, tmp, it, s.off, (int)g.nwords, ctx->stream));
    const unsigned eblocks = (unsigned)((g.nwords + 255) / 256);
    hipLaunchKernelGGL(k_peak_emit, dim3(eblocks), dim3(256), 0, ctx->stream, d_field, d_mask, g,
                       (const uint64_t *)s.bits, (const int *)s.off, (const long long *)d_counters, max_peaks,
                       filt, o);
    PP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_stencil_finish, dim3(1), dim3(1), 0, ctx->stream, (const uint64_t *)s.bits,
                       (const int *)s.off, g.nwords, max_peaks, d_counters);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

static int run_stencil(pp_ctx *ctx, const float *d_field, int8_t *d_mask, const StencilGeom &g,
                       float fcrit, long long max_peaks, int filt, const PeakOut &o,
                       long long *d_counters)
{
    size_t bytes = 0, tmp = 0;
    PP_CHECK(stencil_scan_bytes(g, &tmp));
    PP_CHECK(stencil_scratch_bytes(g, &bytes));
    PP_CHECK(ctx->d_scan.grow(bytes));
    return run_stencil_with(ctx, d_field, d_mask, g, fcrit, max_peaks, filt, o, d_counters,
                            stencil_carve(ctx->d_scan.ptr, g, tmp));
}

// ------------------------------------------- peak finding without allocation
// pp_find_peaks_field_noalloc keeps its counters and the stencil's bit mask,
// offsets and scan storage in `scratch`, a file-scope array every device the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel: it passes the device address hipGetSymbolAddress
// gives for the current device, taken once per device when a context is
// created there (noalloc_scratch_init).
static __device__ __align__(256) unsigned char scratch[PP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

int noalloc_scratch_init(int device)
{
    PP_ARG(device >= 0 && device < kMaxDevices, "device %d beyond the %d the library tracks", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return PP_OK;
    void *p = nullptr;
    PP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PP_ARG(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[device] = p;
    return PP_OK;
}

static void *noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// clears the counters with a kernel rather than a memset node, so that a
// caller may capture the call in a HIP graph
__global__ void k_zero_counters(long long *c)
{
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
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
        if (out->i) PP_HIP(hipMemcpyAsync(out->i, o.i, b4, hipMemcpyDefault, ctx->stream));
        if (out->j) PP_HIP(hipMemcpyAsync(out->j, o.j, b4, hipMemcpyDefault, ctx->stream));
        if (out->k) PP_HIP(hipMemcpyAsync(out->k, o.k, b4, hipMemcpyDefault, ctx->stream));
        if (out->filt) PP_HIP(hipMemcpyAsync(out->filt, o.filt, b4, hipMemcpyDefault, ctx->stream));
        if (out->delta_s)
            PP_HIP(hipMemcpyAsync(out->delta_s, o.delta_s, b4, hipMemcpyDefault, ctx->stream));
        if (out->lapd_s && d_lapd)
            PP_HIP(hipMemcpyAsync(out->lapd_s, d_lapd, b4, hipMemcpyDefault, ctx->stream));
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
    PP_FFT(rocfft_execute(ctx->fftp->r2c, ib, ob, ctx->fftp->info));
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
    PP_FFT(rocfft_execute(ctx->fftp->c2r, ib, ob, ctx->fftp->info));
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
    const StencilGeom g = stencil_geom(n1, n2, n3, nbuff);
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
    if (want_lapd) {
        PP_CHECK(ctx->d_k2.grow(cb));
        PP_CHECK(ctx->d_field_b.grow(rb));
    }
    PP_CHECK(ctx->d_mask.grow(N));
    PP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    const float *d_delta = delta;
    if (!on_device) {
        PP_HIP(hipMemcpyAsync(ctx->d_field_a.ptr, delta, rb, hipMemcpyHostToDevice, ctx->stream));
        d_delta = ctx->d_field_a.as<float>();
    }
    {
        void *ib[1] = {(void *)d_delta}, *ob[1] = {ctx->d_k0.ptr};
        PP_FFT(rocfft_execute(ctx->fftp->r2c, ib, ob, ctx->fftp->info));
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
    const StencilGeom g = stencil_geom(n, n, n, nbuff);
    const float inv = (float)(1.0 / ((double)n * n * n));
    for (int f = 0; f < nfilt; f++) {
        // one read of delta_k gives the filter's spectrum and, with want_lapd,
        // the wsmooth = 3 spectrum sampled at the filter's new peaks
        PP_CHECK(launch_window2(ctx, ctx->d_k0.ptr, ctx->d_k1.ptr, wsmooth, want_lapd ? ctx->d_k2.ptr : nullptr,
                                PP_WINDOW_K2GAUSS, n, boxsize, Rf[f], inv));
        void *ib[1] = {ctx->d_k1.ptr}, *ob[1] = {ctx->d_field_a.ptr};
        PP_FFT(rocfft_execute(ctx->fftp->c2r, ib, ob, ctx->fftp->info));
        PP_CHECK(run_stencil(ctx, ctx->d_field_a.as<float>(), d_mask, g, fcrit[f], max_peaks_per_filter, f,
                             o, d_counters));
        if (want_lapd) {
            void *ib2[1] = {ctx->d_k2.ptr}, *ob2[1] = {ctx->d_field_b.ptr};
            PP_FFT(rocfft_execute(ctx->fftp->c2r, ib2, ob2, ctx->fftp->info));
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

int pp_find_peaks_field_noalloc(pp_ctx *ctx, const float *delta_s, int8_t *mask, int n1, int n2, int n3, int nbuff,
                                float fcrit, int64_t max_peaks, pp_peaks *out)
{
    PP_ARG(ctx && delta_s && mask && out, "NULL argument");
    PP_ARG(out->i && out->j && out->k && out->filt && out->delta_s, "every output array is required");
    PP_ARG(n1 >= 3 && n2 >= 3 && n3 >= 3, "grid %dx%dx%d too small", n1, n2, n3);
    PP_ARG(nbuff >= 1, "nbuff must be >= 1 (the 27-point stencil reads one cell past the inner cube)");
    PP_ARG(out->capacity >= 0, "negative capacity");
    PP_CHECK(ensure_device(ctx));
    if (max_peaks <= 0) max_peaks = 2000000;
    const StencilGeom g = stencil_geom(n1, n2, n3, nbuff);
    size_t need = 0, tmp = 0;
    PP_CHECK(stencil_scan_bytes(g, &tmp));
    PP_CHECK(stencil_scratch_bytes(g, &need));
    need += 256;
    PP_ARG(need <= (size_t)PP_NOALLOC_SCRATCH_BYTES,
           "inner cube %dx%dx%d needs %zu B of scratch, more than the %zu B this entry point has; use "
           "pp_find_peaks_field", g.in1, g.in2, g.in3, need, (size_t)PP_NOALLOC_SCRATCH_BYTES);
    char *base = (char *)noalloc_scratch(ctx->device);
    if (!base) {
        set_error("no scratch address for device %d (context creation records it)", ctx->device);
        return PP_ERR_STATE;
    }
    long long *d_counters = (long long *)base;
    PeakOut o;
    o.i = out->i; o.j = out->j; o.k = out->k; o.filt = out->filt; o.delta_s = out->delta_s;
    o.capacity = out->capacity;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, d_counters);
    PP_HIP(hipGetLastError());
    PP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    PP_CHECK(run_stencil_with(ctx, delta_s, mask, g, fcrit, max_peaks, 0, o, d_counters,
                              stencil_carve(base + 256, g, tmp)));
    PP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    long long cnt[4];
    PP_HIP(hipMemcpyAsync(cnt, d_counters, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->timing.peakfind_ms = ms;
    out->needed = cnt[1];
    out->count = cnt[1] < out->capacity ? cnt[1] : out->capacity;
    if (cnt[1] > out->capacity) {
        set_error("peak output capacity %lld too small: %lld peaks found", (long long)out->capacity, cnt[1]);
        return PP_ERR_CAPACITY;
    }
    return PP_OK;
}

}  // extern "C"

// ------------------------------------------------------------ input noise
// fill_noise_threefry! (src/InitialConditions/RandomField.jl:46-81) over n3
// cells, one thread per counter pair (threefry_normals).
namespace pp {

__global__ void k_noise_threefry(float *__restrict__ out, int64_t n3, uint64_t seed)
{
    const int64_t npairs = (n3 + 1) / 2;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += (int64_t)gridDim.x * blockDim.x) {
        double g1, g2;
        threefry_normals((uint64_t)p, seed, g1, g2);
        out[2 * p] = (float)g1;
        if (2 * p + 1 < n3) out[2 * p + 1] = (float)g2;
    }
}

int launch_noise(pp_ctx *ctx, float *out, int64_t ncells, uint64_t seed)
{
    if (ncells == 0) return PP_OK;
    const int64_t npairs = (ncells + 1) / 2;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((npairs + 255) / 256, (int64_t)ctx->num_cu * 16));
    hipLaunchKernelGGL(k_noise_threefry, dim3(grid), dim3(256), 0, ctx->stream, out, ncells, seed);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

}  // namespace pp

extern "C" int pp_fill_noise_threefry(pp_ctx *ctx, float *out, int64_t ncells, uint64_t seed)
{
    PP_ARG(ctx && out, "NULL argument");
    PP_ARG(ncells >= 0, "negative cell count");
    PP_CHECK(pp::ensure_device(ctx));
    return pp::launch_noise(ctx, out, ncells, seed);
}
