// pp_common.h -- internal declarations shared by the libpeakpatch_hip sources.
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/peakpatch_hip.h"

namespace pp {

// ----------------------------------------------------------------- errors
void set_error(const char *fmt, ...);

struct Status {
    int code = PP_OK;
    bool ok() const { return code == PP_OK; }
};

#define PP_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::pp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                            __FILE__, __LINE__);                                          \
            return e_ == hipErrorOutOfMemory ? PP_ERR_OOM : PP_ERR_HIP;                   \
        }                                                                                 \
    } while (0)

#define PP_FFT(expr)                                                                      \
    do {                                                                                  \
        rocfft_status s_ = (expr);                                                        \
        if (s_ != rocfft_status_success) {                                                \
            ::pp::set_error("%s failed: rocfft status %d (%s:%d)", #expr, (int)s_,        \
                            __FILE__, __LINE__);                                          \
            return PP_ERR_FFT;                                                            \
        }                                                                                 \
    } while (0)

#define PP_CHECK(expr)                                                                    \
    do {                                                                                  \
        int r_ = (expr);                                                                  \
        if (r_ != PP_OK) return r_;                                                       \
    } while (0)

#define PP_ARG(cond, ...)                                                                 \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            ::pp::set_error(__VA_ARGS__);                                                 \
            return PP_ERR_ARG;                                                            \
        }                                                                                 \
    } while (0)

// ------------------------------------------------------- device buffers
// A growable device buffer owned by a context; contents are not preserved
// across grow().
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int grow(size_t need);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// ------------------------------------------------------------ shell table
// precompute_shells(rmax) cells, r^2-sorted with generation order (dk, dj,
// di) kept inside each r^2, grouped into shells.  Shell s (1-based; s = 1
// is the centre cell) spans cells [start[s-1], start[s]).
struct ShellTable {
    int rmax = -1;
    int64_t ncells = 0;
    int32_t nshells = 0;                 // distinct r^2 values incl. 0
    std::vector<int32_t> di, dj, dk, r2; // host copy, r^2-sorted
    std::vector<int32_t> shell_start;    // nshells + 1 entries (0-based cells)
    std::vector<int32_t> shell_r2;       // nshells entries
    // device copies
    DevBuf d_cell;        // int32 packed (di+128) | (dj+128)<<8 | (dk+128)<<16
    DevBuf d_shell_start; // int32 [nshells + 1]
    DevBuf d_shell_r2;    // int32 [nshells]
    DevBuf d_shell_of_r2; // int32 [rmax^2 + 1] -> shell index (1-based), 0 if none
    // octant groups (|di|, |dj|, |dk|) in row-major order for the coalesced
    // gather; see pp_shell.hip
    DevBuf d_groups;
    std::vector<int64_t> class_groups; // groups with r^2 <= class bound (prefix)
    DevBuf d_rad;         // double [nshells], sqrt(r2)
    DevBuf d_akk;         // double [1101], atab4
    // kernel_strain weights w(mp, m1) for every shell pair the radial kernel
    // touches: row mp-1 of d_wrow = {offset, first m1, last m1, 0}
    DevBuf d_wrow;        // int4 [nshells]
    DevBuf d_wtab;        // double
    DevBuf d_shell_ge;    // int32 [max r2 + 2]: first shell with r2 >= k
    int rad_rmax = -1;
};

// ------------------------------------------------------------ context
struct Timing {
    double gather_ms = 0, shell_ms = 0, peakfind_ms = 0;
    // persistent form, per stage: prepare (transpose, order, scan), k_gather,
    // k_small, wave post, lane post, finish, small-peak post (k_small and its
    // post run on the side stream, beside the wave and lane posts)
    static constexpr int NSTAGE = 7;
    double stage_ms[NSTAGE] = {0, 0, 0, 0, 0, 0, 0};
};

// A pair of single-precision real transforms of an n^3 grid (x fastest):
// out of place (kind 0, real n^3 <-> complex (n/2+1) n^2), or in place on
// the padded layout (kind 1: real rows of pitch 2 (n/2+1) floats).
struct FftPlans {
    int n = 0, kind = 0;
    unsigned long long last_use = 0;
    size_t work_need = 0;  // work area bytes; kind 1 sets its work area per use (stream ordered)
    rocfft_plan r2c = nullptr, c2r = nullptr;
    rocfft_execution_info info = nullptr;
    DevBuf work;
    void release();
};
constexpr int kFftCache = 6;  // plan sets a context keeps (least recently used replaced)

}  // namespace pp

struct pp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // second stream (default priority) for work beside the main stream's
    int num_cu = 0;
    size_t lds_per_block = 0;
    // collapse table
    pp::DevBuf d_table;
    int tnx = 0, tny = 0, tnz = 0;
    double tX1 = 0, tX2 = 0, tY1 = 0, tY2 = 0, tZ1 = 0, tZ2 = 0;
    bool have_table = false;
    pp::ShellTable shells;
    // FFT: plan sets by (n, kind); fft is the one fft_prepare made current
    pp::FftPlans fft_cache[pp::kFftCache];
    pp::FftPlans *fftp = nullptr;
    unsigned long long fft_clock = 0;
    // scratch
    pp::DevBuf d_field_a, d_field_b, d_k0, d_k1, d_k2, d_mask, d_misc, d_peaks, d_scan, d_cub;
    pp::DevBuf d_in[8], d_pk, d_res, d_prof, d_queue, d_walk;
    // persistent-form tunables (pp_autotune_shell): small-peak threshold in
    // finalised shells (several peaks per wave), shells per LDS chunk of the
    // wave gather, the shell count above which a peak's Float64 post runs on
    // a whole wave, the cap on gather waves per CU, and the lanes per peak of
    // the post below that count
    int tiny_m = 16;
    int chunk_shells = 64;
    int coop_m = 64;
    int gather_waves_cu = 16;
    int post_width = 1;  // lanes per peak of the post below coop_m (1, 4, 8, 16)
    int lane_walk = 0;   // lane post: walks longer than this go to the wave post (0: no limit)
    // events: [0,1] whole call, [2..13] persistent-form stage boundaries
    static constexpr int NEV = 15;
    hipEvent_t ev[NEV] = {};
    hipEvent_t ev_join = nullptr, ev_join2 = nullptr;  // order the context's streams against a caller's
    pp::Timing timing;
};

namespace pp {
int fft_prepare(pp_ctx *ctx, int n);                       // makes the n^3 out-of-place set current
int fft_plans(pp_ctx *ctx, int n, int kind, FftPlans **out);  // a cached set, without changing the current one
void persist_release(pp_ctx *ctx);    // pp_persist.hip: drops the context's gather tables
int noalloc_scratch_init(int device);  // pp_filter.hip: device address of the static scratch
int ensure_device(pp_ctx *ctx);
}  // namespace pp
