// pp_context.hip -- context, error reporting, static tables (collapse table,
// shell tables) and the rocFFT plan cache of libpeakpatch_hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "pp_common.h"

namespace pp {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int DevBuf::grow(size_t need)
{
    if (need <= bytes && ptr) return PP_OK;
    release();
    size_t b = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc(%zu bytes) failed: %s", b, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? PP_ERR_OOM : PP_ERR_HIP;
    }
    bytes = b;
    return PP_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int ensure_device(pp_ctx *ctx)
{
    PP_HIP(hipSetDevice(ctx->device));
    return PP_OK;
}

static std::once_flag g_fft_once;

void FftPlans::release()
{
    if (r2c) rocfft_plan_destroy(r2c);
    if (c2r) rocfft_plan_destroy(c2r);
    if (info) rocfft_execution_info_destroy(info);
    r2c = c2r = nullptr;
    info = nullptr;
    work.release();
    n = 0;
    kind = 0;
    last_use = 0;
    work_need = 0;
}

static int fft_create(pp_ctx *ctx, FftPlans &F, int n, int kind)
{
    std::call_once(g_fft_once, [] { rocfft_setup(); });
    F.release();
    // lengths are fastest-dimension first; the field is x fastest.
    size_t len[3] = {(size_t)n, (size_t)n, (size_t)n};
    if (kind == 0) {
        PP_FFT(rocfft_plan_create(&F.r2c, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                  rocfft_precision_single, 3, len, 1, nullptr));
        PP_FFT(rocfft_plan_create(&F.c2r, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                  rocfft_precision_single, 3, len, 1, nullptr));
    } else {
        // in place: real rows padded to 2 (n/2+1) floats, the complex view
        // (n/2+1) x n x n on the same memory
        const size_t nk = (size_t)n / 2 + 1, pitch = 2 * nk;
        size_t rs[3] = {1, pitch, pitch * n}, cs[3] = {1, nk, nk * n};
        const size_t rd = pitch * n * n, cd = nk * n * n;
        rocfft_plan_description fwd = nullptr, inv = nullptr;
        PP_FFT(rocfft_plan_description_create(&fwd));
        PP_FFT(rocfft_plan_description_create(&inv));
        rocfft_status st = rocfft_plan_description_set_data_layout(
            fwd, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved, nullptr, nullptr, 3, rs, rd, 3, cs,
            cd);
        if (st == rocfft_status_success)
            st = rocfft_plan_description_set_data_layout(inv, rocfft_array_type_hermitian_interleaved,
                                                         rocfft_array_type_real, nullptr, nullptr, 3, cs, cd, 3, rs,
                                                         rd);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&F.r2c, rocfft_placement_inplace, rocfft_transform_type_real_forward,
                                    rocfft_precision_single, 3, len, 1, fwd);
        if (st == rocfft_status_success)
            st = rocfft_plan_create(&F.c2r, rocfft_placement_inplace, rocfft_transform_type_real_inverse,
                                    rocfft_precision_single, 3, len, 1, inv);
        rocfft_plan_description_destroy(fwd);
        rocfft_plan_description_destroy(inv);
        PP_FFT(st);
    }
    size_t w1 = 0, w2 = 0;
    PP_FFT(rocfft_plan_get_work_buffer_size(F.r2c, &w1));
    PP_FFT(rocfft_plan_get_work_buffer_size(F.c2r, &w2));
    size_t w = std::max(w1, w2);
    PP_FFT(rocfft_execution_info_create(&F.info));
    PP_FFT(rocfft_execution_info_set_stream(F.info, ctx->stream));
    F.work_need = w;
    if (w > 0 && kind == 0) {
        PP_CHECK(F.work.grow(w));
        PP_FFT(rocfft_execution_info_set_work_buffer(F.info, F.work.ptr, F.work.bytes));
    }
    F.n = n;
    F.kind = kind;
    return PP_OK;
}

int fft_plans(pp_ctx *ctx, int n, int kind, FftPlans **out)
{
    FftPlans *hit = nullptr, *victim = &ctx->fft_cache[0];
    for (FftPlans &F : ctx->fft_cache) {
        if (F.r2c && F.n == n && F.kind == kind) hit = &F;
        if (!F.r2c) {
            if (victim->r2c) victim = &F;
        } else if (victim->r2c && F.last_use < victim->last_use) {
            victim = &F;
        }
    }
    if (!hit) {
        // the stream may still run transforms of the plan set being replaced
        if (victim->r2c) PP_HIP(hipStreamSynchronize(ctx->stream));
        int rc = fft_create(ctx, *victim, n, kind);
        if (rc != PP_OK) {
            victim->release();
            return rc;
        }
        hit = victim;
    }
    hit->last_use = ++ctx->fft_clock;
    *out = hit;
    return PP_OK;
}

int fft_prepare(pp_ctx *ctx, int n)
{
    FftPlans *F = nullptr;
    PP_CHECK(fft_plans(ctx, n, 0, &F));
    ctx->fftp = F;
    return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

const char *pp_last_error(void) { return g_err; }

int pp_version(void) { return 10000; }

int pp_device_count(int *count)
{
    PP_ARG(count != nullptr, "count is NULL");
    int c = 0;
    PP_HIP(hipGetDeviceCount(&c));
    *count = c;
    return PP_OK;
}

int pp_ctx_create(int device, pp_ctx **out)
{
    PP_ARG(out != nullptr, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    PP_HIP(hipGetDeviceCount(&ndev));
    PP_ARG(device >= 0 && device < ndev, "device %d out of range (%d devices)", device, ndev);
    pp_ctx *c = new pp_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    for (int i = 0; i < pp_ctx::NEV && e == hipSuccess; i++) e = hipEventCreate(&c->ev[i]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error("context creation on device %d failed: %s", device, hipGetErrorString(e));
        for (auto &ev : c->ev)
            if (ev) (void)hipEventDestroy(ev);
        if (c->ev_join) (void)hipEventDestroy(c->ev_join);
        if (c->ev_join2) (void)hipEventDestroy(c->ev_join2);
        if (c->side) (void)hipStreamDestroy(c->side);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return PP_ERR_HIP;
    }
    c->num_cu = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    int rc = noalloc_scratch_init(device);  // once per device, with the device current
    if (rc != PP_OK) {
        pp_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return PP_OK;
}

int pp_ctx_destroy(pp_ctx *ctx)
{
    if (!ctx) return PP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (FftPlans &F : ctx->fft_cache) F.release();
    DevBuf *bufs[] = {&ctx->d_table, &ctx->d_field_a, &ctx->d_field_b, &ctx->d_k0, &ctx->d_k1,
                      &ctx->d_k2, &ctx->d_mask, &ctx->d_misc, &ctx->d_peaks, &ctx->d_scan,
                      &ctx->d_cub, &ctx->d_pk, &ctx->d_res, &ctx->d_prof, &ctx->d_queue, &ctx->d_walk,
                      &ctx->shells.d_cell, &ctx->shells.d_shell_start, &ctx->shells.d_shell_r2,
                      &ctx->shells.d_shell_of_r2, &ctx->shells.d_groups, &ctx->shells.d_rad,
                      &ctx->shells.d_akk, &ctx->shells.d_wrow, &ctx->shells.d_wtab,
                      &ctx->shells.d_shell_ge};
    for (DevBuf *b : bufs) b->release();
    persist_release(ctx);
    for (auto &b : ctx->d_in) b.release();
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_join2) (void)hipEventDestroy(ctx->ev_join2);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PP_OK;
}

int pp_ctx_synchronize(pp_ctx *ctx)
{
    PP_ARG(ctx, "ctx is NULL");
    PP_CHECK(ensure_device(ctx));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->side));
    return PP_OK;
}

int pp_ctx_stream(pp_ctx *ctx, void **stream)
{
    PP_ARG(ctx && stream, "NULL argument");
    *stream = (void *)ctx->stream;
    return PP_OK;
}

int pp_ctx_wait_stream(pp_ctx *ctx, void *stream)
{
    PP_ARG(ctx, "ctx is NULL");
    PP_CHECK(ensure_device(ctx));
    hipStream_t s = (hipStream_t)stream;
    if (s == ctx->stream) return PP_OK;
    // the event is recorded on the caller's queue and waited on by both of
    // the context's streams, so whatever either launches next runs after
    // the work the caller has queued so far
    PP_HIP(hipEventRecord(ctx->ev_join, s));
    PP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    PP_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_join, 0));
    return PP_OK;
}

int pp_ctx_signal_stream(pp_ctx *ctx, void *stream)
{
    PP_ARG(ctx, "ctx is NULL");
    PP_CHECK(ensure_device(ctx));
    hipStream_t s = (hipStream_t)stream;
    if (s == ctx->stream) return PP_OK;
    // the side stream's work joins the main stream, whose tail the caller's
    // queue then waits for (two events: each wait keeps its own record)
    PP_HIP(hipEventRecord(ctx->ev_join2, ctx->side));
    PP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join2, 0));
    PP_HIP(hipEventRecord(ctx->ev_join, ctx->stream));
    PP_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
    return PP_OK;
}

int pp_ctx_release_scratch(pp_ctx *ctx)
{
    PP_ARG(ctx, "NULL context");
    PP_CHECK(ensure_device(ctx));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->side));
    DevBuf *bufs[] = {&ctx->d_field_a, &ctx->d_field_b, &ctx->d_k0, &ctx->d_k1, &ctx->d_k2, &ctx->d_mask,
                      &ctx->d_misc, &ctx->d_peaks, &ctx->d_scan, &ctx->d_cub, &ctx->d_pk, &ctx->d_res,
                      &ctx->d_prof, &ctx->d_queue, &ctx->d_walk};
    for (DevBuf *b : bufs) b->release();
    for (auto &b : ctx->d_in) b.release();
    return PP_OK;
}

int pp_ctx_mem_info(pp_ctx *ctx, size_t *free_bytes, size_t *total_bytes)
{
    PP_ARG(ctx && free_bytes && total_bytes, "NULL argument");
    PP_CHECK(ensure_device(ctx));
    PP_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return PP_OK;
}

int pp_upload_collapse_table(pp_ctx *ctx, const float *table, int nx, int ny, int nz, double X1,
                             double X2, double Y1, double Y2, double Z1, double Z2)
{
    PP_ARG(ctx && table, "NULL argument");
    PP_ARG(nx >= 2 && ny >= 2 && nz >= 2, "collapse table needs >= 2 points per axis (%d,%d,%d)",
           nx, ny, nz);
    PP_ARG(X2 > X1 && Y2 > Y1 && Z2 > Z1, "collapse table bounds must be increasing");
    PP_CHECK(ensure_device(ctx));
    size_t bytes = (size_t)nx * ny * nz * sizeof(float);
    PP_CHECK(ctx->d_table.grow(bytes));
    PP_HIP(hipMemcpyAsync(ctx->d_table.ptr, table, bytes, hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    ctx->tnx = nx; ctx->tny = ny; ctx->tnz = nz;
    ctx->tX1 = X1; ctx->tX2 = X2; ctx->tY1 = Y1; ctx->tY2 = Y2; ctx->tZ1 = Z1; ctx->tZ2 = Z2;
    ctx->have_table = true;
    return PP_OK;
}



int pp_build_shell_tables(pp_ctx *ctx, int rmax)
{
    PP_ARG(ctx, "ctx is NULL");
    PP_ARG(rmax >= 1 && rmax <= 120, "rmax %d outside [1, 120]", rmax);
    PP_CHECK(ensure_device(ctx));
    ShellTable &T = ctx->shells;
    if (T.rmax == rmax) return PP_OK;
    const int r2max = rmax * rmax;
    // counting sort on r^2 keeps generation order (dk outer, dj, di inner)
    // inside each r^2: a stable sort as RadialShell.jl:160-175.
    std::vector<int64_t> count((size_t)r2max + 2, 0);
    for (int k = -rmax; k <= rmax; k++)
        for (int j = -rmax; j <= rmax; j++)
            for (int i = -rmax; i <= rmax; i++) {
                int q = i * i + j * j + k * k;
                if (q <= r2max) count[(size_t)q + 1]++;
            }
    for (int q = 1; q <= r2max + 1; q++) count[q] += count[q - 1];
    int64_t ncells = count[(size_t)r2max + 1];
    T.di.assign(ncells, 0); T.dj.assign(ncells, 0); T.dk.assign(ncells, 0); T.r2.assign(ncells, 0);
    std::vector<int64_t> pos(count.begin(), count.end() - 1);
    for (int k = -rmax; k <= rmax; k++)
        for (int j = -rmax; j <= rmax; j++)
            for (int i = -rmax; i <= rmax; i++) {
                int q = i * i + j * j + k * k;
                if (q > r2max) continue;
                int64_t p = pos[q]++;
                T.di[p] = i; T.dj[p] = j; T.dk[p] = k; T.r2[p] = q;
            }
    T.shell_start.clear();
    T.shell_r2.clear();
    std::vector<int32_t> shell_of_r2((size_t)r2max + 1, 0);
    for (int64_t c = 0; c < ncells; c++) {
        if (c == 0 || T.r2[c] != T.r2[c - 1]) {
            T.shell_start.push_back((int32_t)c);
            T.shell_r2.push_back(T.r2[c]);
            shell_of_r2[T.r2[c]] = (int32_t)T.shell_r2.size();  // 1-based
        }
    }
    T.shell_start.push_back((int32_t)ncells);
    T.nshells = (int32_t)T.shell_r2.size();
    T.ncells = ncells;
    std::vector<int32_t> packed(ncells);
    for (int64_t c = 0; c < ncells; c++)
        packed[c] = (T.di[c] + 128) | ((T.dj[c] + 128) << 8) | ((T.dk[c] + 128) << 16);
    PP_CHECK(T.d_cell.grow(ncells * sizeof(int32_t)));
    PP_CHECK(T.d_shell_start.grow(T.shell_start.size() * sizeof(int32_t)));
    PP_CHECK(T.d_shell_r2.grow(T.shell_r2.size() * sizeof(int32_t)));
    PP_CHECK(T.d_shell_of_r2.grow(shell_of_r2.size() * sizeof(int32_t)));
    PP_HIP(hipMemcpy(T.d_cell.ptr, packed.data(), ncells * sizeof(int32_t), hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(T.d_shell_start.ptr, T.shell_start.data(), T.shell_start.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(T.d_shell_r2.ptr, T.shell_r2.data(), T.shell_r2.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(T.d_shell_of_r2.ptr, shell_of_r2.data(), shell_of_r2.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    T.rmax = rmax;
    return PP_OK;
}

int pp_shell_table_info(pp_ctx *ctx, int64_t *ncells, int32_t *nshells)
{
    PP_ARG(ctx && ncells && nshells, "NULL argument");
    if (ctx->shells.rmax < 0) {
        set_error("shell tables not built (call pp_build_shell_tables)");
        return PP_ERR_STATE;
    }
    *ncells = ctx->shells.ncells;
    *nshells = ctx->shells.nshells;
    return PP_OK;
}

int pp_shell_tunables(pp_ctx *ctx, const int32_t *set, int32_t *get)
{
    PP_ARG(ctx, "null context");
    if (set) {
        if (set[0] > 64) { set_error("small-peak threshold above 64 shells"); return PP_ERR_ARG; }
        if (set[1] > 0 && (set[1] < 8 || set[1] > 1024)) { set_error("chunk shells outside [8, 1024]"); return PP_ERR_ARG; }
        if (set[3] > 64) { set_error("more than 64 gather waves per CU"); return PP_ERR_ARG; }
        if (set[4] > 0 && set[4] != 1 && set[4] != 2 && set[4] != 4 && set[4] != 8 && set[4] != 16) {
            set_error("post lanes per peak must be 1, 2, 4, 8 or 16");
            return PP_ERR_ARG;
        }
        if (set[0] >= 0) ctx->tiny_m = set[0];
        if (set[1] > 0) ctx->chunk_shells = set[1];
        if (set[2] > 0) ctx->coop_m = set[2];
        if (set[3] > 0) ctx->gather_waves_cu = set[3];
        if (set[4] > 0) ctx->post_width = set[4];
        if (set[5] >= 0) ctx->lane_walk = set[5];
    }
    if (get) {
        get[0] = ctx->tiny_m;
        get[1] = ctx->chunk_shells;
        get[2] = ctx->coop_m;
        get[3] = ctx->gather_waves_cu;
        get[4] = ctx->post_width;
        get[5] = ctx->lane_walk;
    }
    return PP_OK;
}

int pp_last_shell_stages(pp_ctx *ctx, double *ms)
{
    PP_ARG(ctx && ms, "NULL argument");
    for (int q = 0; q < Timing::NSTAGE; q++) ms[q] = ctx->timing.stage_ms[q];
    return PP_OK;
}

int pp_last_timing(pp_ctx *ctx, double *g, double *s, double *p)
{
    PP_ARG(ctx, "ctx is NULL");
    if (g) *g = ctx->timing.gather_ms;
    if (s) *s = ctx->timing.shell_ms;
    if (p) *p = ctx->timing.peakfind_ms;
    return PP_OK;
}

}  // extern "C"
