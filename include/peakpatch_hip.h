/*
 * peakpatch_hip.h -- C ABI of libpeakpatch_hip.so, the MI355X (gfx950) build of
 * PeakPatch.jl's halo-finding hot path: filter-bank multiply, local-maximum
 * peak stencil, and per-peak radial-shell analysis with the ellipsoidal
 * collapse-table lookup.
 *
 * Conventions
 *  - Fields are Float32, n1 x n2 x n3, x fastest (Julia column-major layout,
 *    zero-copy from a Julia Array{Float32,3}).
 *  - Grid indices at this ABI are 0-based (the Julia shim subtracts 1).
 *  - Every entry returns PP_OK (0) or a negative PP_ERR_* code; the message
 *    of the last failure on the calling thread is pp_last_error().  Nothing
 *    aborts or throws across the ABI.
 *  - A pp_ctx is bound to one device and owns one HIP stream, the rocFFT plan
 *    cache, the shell tables, the collapse table and pooled device buffers.
 *    A context may be used by one host thread at a time.
 *  - Host buffers are owned by the caller.  Pointers flagged on_device are
 *    device memory of the context's device, also owned by the caller.
 */
#ifndef PEAKPATCH_HIP_H
#define PEAKPATCH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_OK 0
#define PP_ERR_ARG (-1)       /* bad argument / shape mismatch          */
#define PP_ERR_OOM (-2)       /* device allocation failed              */
#define PP_ERR_HIP (-3)       /* HIP runtime error                     */
#define PP_ERR_FFT (-4)       /* rocFFT error                          */
#define PP_ERR_STATE (-5)     /* missing table / wrong call order      */
#define PP_ERR_CAPACITY (-6)  /* caller's output arrays too small      */
#define PP_ERR_BUSY (-7)      /* serving queue full / not admitted     */
#define PP_ERR_TIMEOUT (-8)   /* serving wait timed out                */
#define PP_ERR_INJECTED (-9)  /* fault injected for failure tests      */

/* smooth_field window selector (Filters.jl:18-36, 79-106, wsmooth) */
#define PP_WINDOW_GAUSSIAN 0  /* exp(-k^2 R^2 / 8)          */
#define PP_WINDOW_TOPHAT 1    /* 3 (sin x - x cos x) / x^3  */
#define PP_WINDOW_K2GAUSS 3   /* k^2 exp(-k^2 R^2 / 8)      */

/* shell-analysis kernel form */
#define PP_SHELL_PLAIN 0      /* one workgroup per peak, one launch per batch */
#define PP_SHELL_PERSISTENT 1 /* persistent work-queue kernel, radius-sorted */

/* analyse_peak early-exit reasons (RadialShell.jl:14-34 dump counters) */
#define PP_DUMP_NONE 0
#define PP_DUMP_PREMASK 1
#define PP_DUMP_NO_FCRIT 2
#define PP_DUMP_M0_ONE 3
#define PP_DUMP_NO_COLLAPSE 4
#define PP_DUMP_RTHL_NEG 5
#define PP_DUMP_NONFINITE 6

typedef struct pp_ctx pp_ctx;

/* ---------------------------------------------------------------- basics */
const char *pp_last_error(void);
int pp_version(void);
int pp_device_count(int *count);

/* Replaces the CUDA.jl context set-up of ext/CUDAExt.jl (device selection,
 * set_cuda_device! CUDAExt.jl:3322-3325; plan cache CUDAExt.jl:82-106). */
int pp_ctx_create(int device, pp_ctx **out);
int pp_ctx_destroy(pp_ctx *ctx);
int pp_ctx_synchronize(pp_ctx *ctx);
/* The context's HIP stream (a hipStream_t), so a caller can order its own
 * device work (input synthesis, copies) on the same queue.  No counterpart in
 * the reference (CUDA.jl's task-local stream). */
int pp_ctx_stream(pp_ctx *ctx, void **stream);
/* Stream ordering with a caller's HIP stream (NULL = the legacy default
 * stream), for callers that produce or consume device operands on their own
 * queue (the task-local stream of CUDA.jl / AMDGPU.jl, torch's current
 * stream).  pp_ctx_wait_stream: work the context launches next runs after
 * the work queued on `stream` so far.  pp_ctx_signal_stream: work queued on
 * `stream` next runs after everything the context has launched so far.
 * Both only enqueue waits; neither blocks the host.  The reference has no
 * counterpart: CUDA.jl orders its own task-local stream implicitly. */
int pp_ctx_wait_stream(pp_ctx *ctx, void *stream);
int pp_ctx_signal_stream(pp_ctx *ctx, void *stream);
/* Free / total device memory as seen by this context (for admission). */
int pp_ctx_mem_info(pp_ctx *ctx, size_t *free_bytes, size_t *total_bytes);
/* Returns the context's pooled working memory (field, spectrum, peak and
 * profile scratch) to the device, keeping the shell, gather and collapse
 * tables and the transform plans; the next call grows the pools again.  Waits for the context's streams.  No counterpart in the
 * reference: a serving loop trims an idle slot after a large tile so the
 * memory is free for tiles on other slots. */
int pp_ctx_release_scratch(pp_ctx *ctx);

/* ------------------------------------------------------- static tables */
/* Collapse table upload: read_homeltab + CollapseTableInterp
 * (src/EllipsoidalCollapse/CollapseTable.jl:102-141).  table is Nx*Ny*Nz
 * Float32, x fastest; the bounds are the header values widened to Float64. */
int pp_upload_collapse_table(pp_ctx *ctx, const float *table, int nx, int ny, int nz,
                             double X1, double X2, double Y1, double Y2, double Z1, double Z2);

/* Linear growth tables of a cosmology (DlinearTables, src/Cosmology/
 * Cosmology.jl:56-66, 148-270): ntab rows of scale factor (ascending),
 * D(a)/a normalised to D(1) = 1, and dlnD/dlna.  Host arrays. */
typedef struct {
    int ntab;
    const double *atab, *D_a, *dlnD_dlna;
} pp_growth;

/* Collapse-table generation: make_table (src/EllipsoidalCollapse/
 * CollapseTable.jl:33-55) with evolve_ellipse_full's Fortran-compatible RK4
 * solver and default EllipsoidParams (src/EllipsoidalCollapse/
 * EllipsoidalCollapse.jl:527-624; bin/generate_table.jl --solver rk4), one
 * table point per GPU lane.  Point (i, j, k) is log10 F = X1 + i dX,
 * e = Y1 + j dY, p/e = Z1 + k dZ; its value is 1 + z of the first axis's
 * virialisation, or -1.  table_out (host, nx*ny*nz Float32, x fastest) may
 * be NULL; install = 1 also makes it the context's collapse table with
 * these bounds (as pp_upload_collapse_table). */
int pp_make_collapse_table(pp_ctx *ctx, double Om, double OL, const pp_growth *growth, int nx, int ny,
                           int nz, double X1, double X2, double Y1, double Y2, double Z1, double Z2,
                           float *table_out, int install);

/* precompute_shells(rmax) (src/HaloFinder/RadialShell.jl:160-175) grouped per
 * distinct r^2 as build_shell_tables (src/ShellAnalysisGPU.jl:60-88). */
int pp_build_shell_tables(pp_ctx *ctx, int rmax);
int pp_shell_table_info(pp_ctx *ctx, int64_t *ncells, int32_t *nshells);

/* ---------------------------------------------------------- filter bank */
/* In-place or out-of-place k-space window multiply of smooth_field
 * (src/HaloFinder/Filters.jl:79-106; ext/CUDAExt.jl:2959-3000): every
 * element of delta_k (complex64, (n/2+1) x n x n, x fastest) is multiplied by
 * Float32(W(|k| Rf)) computed in Float64; k = 0 is left unscaled; Nyquist kept.
 * scale multiplies every element afterwards (1 for the bare window). */
int pp_window_multiply(pp_ctx *ctx, const void *delta_k, void *out, int on_device, int n,
                       double boxsize, double Rf, int wsmooth, float scale);

/* smooth_field (Filters.jl:79-106): window multiply + normalised C2R.
 * delta_k complex64 (n/2+1)xnxn; out Float32 n^3. */
int pp_smooth_field(pp_ctx *ctx, const void *delta_k, float *out, int on_device, int n,
                    double boxsize, double Rf, int wsmooth);

/* rfft of a real n^3 Float32 field into complex64 (n/2+1)xnxn. */
int pp_rfft(pp_ctx *ctx, const float *in, void *out, int on_device, int n);

typedef struct {
    int64_t capacity;   /* in: length of each array below                    */
    int64_t count;      /* out: peaks written                                */
    int64_t needed;     /* out: peaks found (> capacity => PP_ERR_CAPACITY)  */
    int32_t *i, *j, *k; /* 0-based grid indices                              */
    int32_t *filt;      /* index into the Rf[] array that found the peak     */
    float *delta_s;     /* smoothed delta at the peak (FcRf)                 */
    float *lapd_s;      /* k^2-Gaussian smoothed field at the peak (d2Rf);
                           may be NULL                                       */
} pp_peaks;

/* find_peaks on one filtered field (src/HaloFinder/PeakFind.jl:15-78): strict
 * 27-point local maxima with ff >= fcrit over the inner cube, masked cells
 * skipped, each found peak cell masked.  Deterministic order k, j, i as the
 * CPU loop; at most max_peaks (the CPU early return, PeakFind.jl:70-72).
 * delta_s and mask are device or host as on_device says (mask n^3 Int8). */
int pp_find_peaks_field(pp_ctx *ctx, const float *delta_s, int8_t *mask, int on_device,
                        int n1, int n2, int n3, int nbuff, float fcrit, int64_t max_peaks,
                        pp_peaks *out);

/* pp_find_peaks_field for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same stencil, order, max_peaks cut and mask
 * update as pp_find_peaks_field; every pointer is device memory of the
 * context's device and every array of `out` (i, j, k, filt, delta_s; lapd_s
 * is ignored) is required.  The call allocates nothing: its counters, peak
 * bit mask, offsets and scan storage live in a static device array of
 * PP_NOALLOC_SCRATCH_BYTES per device (held until the process ends), which
 * limits the inner cube to about 8 M 64-cell row words (ceil(in1 / 64) x in2
 * x in3; any nbuff at n = 768); larger grids are refused with PP_ERR_ARG
 * before any work is queued.  Calls that overlap on one GPU share that array
 * and race: serialise them (one stream, or one call at a time).  Blocks until
 * the counts are back (out->count, out->needed). */
#define PP_NOALLOC_SCRATCH_BYTES (96u << 20)
int pp_find_peaks_field_noalloc(pp_ctx *ctx, const float *delta_s, int8_t *mask, int n1, int n2, int n3,
                                int nbuff, float fcrit, int64_t max_peaks, pp_peaks *out);

/* peak_find_tile_gpu (ext/CUDAExt.jl:3068-3197; src/PeakPatch.jl:341-366) and
 * the CPU filter loop (src/Pipeline.jl:160-195; MultiResolution.jl:885-907):
 * rfft(delta) once, then per filter in the given order (the driver passes
 * descending Rf): window -> C2R -> stencil (mask shared across filters), plus
 * the wsmooth=3 field sampled at the new peaks when want_lapd.  Output order:
 * filter order, then k, j, i.  mask_inout (n^3 Int8, may be NULL = zeros)
 * is read and updated.  max_peaks_per_filter <= 0 means 2,000,000. */
int pp_find_peaks(pp_ctx *ctx, const float *delta, int on_device, int n, double boxsize,
                  const double *Rf, const float *fcrit, int nfilt, int wsmooth, int want_lapd,
                  int nbuff, int64_t max_peaks_per_filter, int8_t *mask_inout, pp_peaks *out);

/* ------------------------------------------------ per-peak bookkeeping */
/* What the drivers know about a tile when they turn peaks into a batch
 * (src/Pipeline.jl:217-231, 257-271; src/MultiResolution.jl:959-988). */
typedef struct {
    int n;                  /* tile mesh                                      */
    double alatt;           /* cell size, Mpc/h                               */
    double xbx, ybx, zbx;   /* tile centre, Mpc/h                             */
    int ievol;              /* 0: every peak at z_out; 1: light cone from obs */
    double z_out, obs[3], z_max;
    int ilpt;               /* >= 2: the 2LPT velocity factor is formed       */
    double Om, OL;
    int round_f32;          /* 1: ZZon and fcrit rounded through Float32 as
                               run_multitile_split does (MultiResolution.jl:
                               963-967); 0: Float64 as run_tile              */
    pp_growth growth;       /* host tables                                    */
    int nchi;               /* chi -> z table (build_chi_to_z, Cosmology.jl:  */
    const double *chi_z, *chi_chi; /* 333-337), host, ievol = 1 only         */
    double chi_zmax;        /* the table's z_max                             */
} pp_tile_geom;

/* Device arrays, one row per kept peak (capacity: the peak count). */
typedef struct {
    int32_t *i, *j, *k;
    double *Rf;
    int64_t *ir2min;
    double *ZZon, *fcrit, *x, *y, *z, *D, *v2fac;
    float *FcRf, *d2Rf;
} pp_batch_out;

/* The drivers' per-peak loop on the device: for the peaks of pp_find_peaks
 * (device arrays; peaks->count rows) the filter radius Rf_filters[filt]
 * (host array of nfilt), position, redshift (z_out or peak_redshift through
 * the chi -> z table, Cosmology.jl:344-372), ZZon = 1 + z, fcrit =
 * fsc_of_z(z) (RadialShell.jl:191-196), ir2min, growth factor D_a
 * (Dlinear_ab, Cosmology.jl:273-293) and 2LPT velocity factor
 * 3/7 Om_a^(-1/143) D_a^2; light-cone peaks with z > z_max are dropped, the
 * others kept in order (*nkept rows written).  The output feeds
 * pp_analyse_peaks (i, j, k, Rf, ir2min, ZZon, fcrit) and pp_pack_halos
 * (x, y, z, Rf, ZZon, D, v2fac, FcRf, d2Rf) without leaving the device. */
int pp_bookkeep_peaks(pp_ctx *ctx, const pp_peaks *peaks, const double *Rf_filters, int nfilt,
                      const pp_tile_geom *geom, pp_batch_out *out, int64_t *nkept);

/* --------------------------------------------------------- shell analysis */
typedef struct {
    int n1, n2, n3;
    const float *delta, *etax, *etay, *etaz; /* required                     */
    const float *eta2x, *eta2y, *eta2z;      /* NULL when ilpt < 2          */
    const float *lapd;                       /* NULL when ioutshear == 0    */
    int on_device;                           /* 1: device pointers           */
} pp_fields;

/* One batch of peaks: the tuple (pi, pj, pk, Rf, ir2min, ZZon_pp, fcrit_pp)
 * of analyse_peaks_gpu_cuda_multirf (ext/CUDAExt.jl:2246-2258, 2326-2334), one
 * row per peak.  Host arrays. */
typedef struct {
    int64_t npeaks;
    const int32_t *i, *j, *k; /* 0-based                                   */
    const double *Rf;         /* Rfclvi                                    */
    const int64_t *ir2min;
    const double *ZZon;       /* 1 + z of the peak                         */
    const double *fcrit;      /* collapse threshold (fsc_of_z or override) */
    int on_device;            /* 1: the arrays above are device memory     */
} pp_peak_batch;

/* PeakResult (RadialShell.jl:58-72) as SoA host arrays, Float64.  Vectors
 * are 3 per peak, strain 9 per peak (column-major 3x3).  Any pointer may be
 * NULL to skip that output. */
typedef struct {
    double *RTHL, *Srb, *Sbar, *Sbar2, *strain, *gradpk, *gradpkrf;
    double *Fbarx, *e_v, *p_v, *zvir_half, *gradpkf, *d2F;
    int32_t *status;          /* PP_DUMP_* per peak (6: non-finite strain,
                                 where the reference's LAPACK call raises) */
    int32_t *nshells;         /* finalised shells per peak (diagnostic)    */
    int on_device;            /* 1: the arrays above are device memory; the
                                 kernels then write them directly, so they
                                 must not overlap each other or the inputs */
    int64_t dump_counts[7];   /* out: per status totals (get_dump_counts)  */
} pp_results;

/* analyse_peak over a batch (src/HaloFinder/RadialShell.jl:282-891 semantics;
 * replaces analyse_peaks_gpu_cuda_multirf, ext/CUDAExt.jl:2246-2462, and the
 * threaded CPU loop src/Pipeline.jl:233-241).  rmax2rs > 0 limits each peak
 * to npart = floor(4/3 pi (rmax2rs Rf / alatt)^3) cells.  mask (n^3 Int8,
 * host or device per mask_on_device, may be NULL) receives 1 for the cells
 * within R_TH inside [nbuff, n - nbuff) (RadialShell.jl:792-802).  Requires
 * pp_build_shell_tables and pp_upload_collapse_table first. */
int pp_analyse_peaks(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks,
                     double alatt, double rmax2rs, int nbuff, int8_t *mask, int mask_on_device,
                     int kernel_form, pp_results *out);

/* Per-peak radial profiles of the same gather (shell counts and cumulative
 * mean overdensity Fbar, RadialShell.jl:379-453), for profile parity tests:
 * nshell_out and Fbar_out are npeaks x max_shells (row per peak); m_out the
 * finalised shell count per peak. */
int pp_shell_profiles(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks,
                      double alatt, double rmax2rs, int kernel_form, int max_shells,
                      int32_t *m_out, int32_t *nshell_out, double *Fbar_out);

/* Autotuner: times the persistent form on the given batch over its launch
 * parameters (the pp_shell_tunables below, one at a time) and keeps the
 * fastest setting in ctx.  Returns the chosen gather waves per CU in *choice
 * (may be NULL).  Counterpart of the reference's per-radius-class launch
 * heuristics (ext/CUDAExt.jl:2246-2462, src/GPUBackend.jl). */
int pp_autotune_shell(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks,
                      double alatt, double rmax2rs, int *choice);

/* Launch parameters of the persistent form, in this order:
 *   [0] small-peak threshold in finalised shells (peaks with m at or below it
 *       are gathered cell by cell, floor(64 / m) per wave; 0 disables that
 *       path; at most 64),
 *   [1] shells per LDS chunk of the wave gather (8..1024),
 *   [2] shell count above which a peak's Float64 post runs on a whole wave,
 *   [3] cap on resident gather waves per CU,
 *   [4] lanes per peak of the Float64 post below [2] (1, 2, 4, 8 or 16),
 *   [5] with one lane per peak: virialisation walks longer than this many
 *       shells are re-run one wave per peak (0: no limit).
 * set (may be NULL): new values, entries <= 0 keep the current one except
 * [0] and [5], where 0 is a value and < 0 keeps; get (may be NULL): the
 * values in force.  They change speed only: results do not depend on them
 * beyond the last bits of the Float64 sums. */
#define PP_NTUNABLES 6
int pp_shell_tunables(pp_ctx *ctx, const int32_t *set, int32_t *get);

/* Kernel timing of the last pp_analyse_peaks / pp_find_peaks call (HIP
 * events on the context stream), milliseconds. */
int pp_last_timing(pp_ctx *ctx, double *shell_gather_ms, double *shell_total_ms,
                   double *peakfind_ms);

/* ------------------------------------------------------ input synthesis */
/* White noise of fill_noise_threefry! (src/InitialConditions/RandomField.jl:
 * 269-306) into out (device, ncells Float32, x fastest), on the context
 * stream: Threefry-2x64-20 counter pairs under key (seed, 0), Box-Muller in
 * Float64.  The tile synthesis uses it; it is not on the analysis path. */
int pp_fill_noise_threefry(pp_ctx *ctx, float *out, int64_t ncells, uint64_t seed);

/* P(k) as a table (the two columns of the reference's P(k) file, load_pk,
 * src/InitialConditions/PowerSpectrum.jl:11-23): log P is interpolated
 * linearly in log k and extrapolated linearly beyond the ends.  Host arrays,
 * k increasing, both positive. */
typedef struct {
    int npk;
    const double *k, *P;
} pp_pk_table;

/* The fields of one tile, device arrays of n^3 Float32 (x fastest).  eta2x..z
 * are all NULL (no 2LPT) or all set; lapd may be NULL. */
typedef struct {
    float *delta, *etax, *etay, *etaz, *eta2x, *eta2y, *eta2z, *lapd;
} pp_tile_fields;

/* The inputs run_tile builds for a periodic tile (src/Pipeline.jl:110-206):
 * generate_grf (src/InitialConditions/RandomField.jl:123-175: Threefry noise
 * of fill_noise_threefry!, rfft, every mode times sqrt(P(k) dk^3 n^3) with
 * k = 0 zeroed, normalised irfft), then from delta_k = rfft(delta) the 1LPT
 * and 2LPT displacements (src/InitialConditions/LPT.jl:21-191, Nyquist planes
 * zeroed) and the k^2 field of _compute_laplacian (src/Pipeline.jl:35-51).
 * Runs on the context stream and returns when the fields are written. */
int pp_make_tile_fields(pp_ctx *ctx, int n, double boxsize, uint64_t seed, const pp_pk_table *pk,
                        pp_tile_fields *fields);

/* ------------------------------------------- multi-resolution tile fields */
/* Geometry of run_multitile_split (src/MultiResolution.jl:546-553, 596-613):
 * tiles of nmesh^3 cells with a buffer of nbuff, nsub = nmesh - 2 nbuff core
 * cells, ntile^3 tiles of a global N^3 grid, N = nsub ntile + 2 nbuff; a
 * coarse M^3 grid (coarse_factor x ntile, or coarse_grid) that divides N. */
typedef struct {
    int nmesh, nbuff, ntile, M;
    double alatt;           /* cell size, Mpc/h                               */
    uint64_t seed;          /* Threefry key of the global white noise         */
} pp_split_geom;

/* The coarse fields, device-resident (any context on the same device may
 * use them once pp_coarse_create has returned). */
typedef struct pp_coarse pp_coarse;

/* Phase 1a of run_multitile_split (src/MultiResolution.jl:617-631): the
 * coarse noise of _downsample_noise (:58-83: block sums of the global
 * Threefry normals in Float64 over sqrt(block^3), Float32) and its periodic
 * convolutions (_periodic_convolve!, :288-307) delta_coarse and
 * psi_coarse[1..3] (sqrt(P(k) dk^3 M^3), times i k_d / k^2 for psi, k = 0
 * zeroed, normalised irfft).  Runs on ctx's stream and returns when done. */
int pp_coarse_create(pp_ctx *ctx, const pp_split_geom *geom, const pp_pk_table *pk, pp_coarse **out);

/* Host copies of the coarse fields (M^3 Float32 each, x fastest); any
 * pointer may be NULL. */
int pp_coarse_download(const pp_coarse *coarse, float *noise, float *delta, float *psix, float *psiy, float *psiz);

int pp_coarse_destroy(pp_coarse *coarse);

/* One tile's fields on process_tile!'s path (src/MultiResolution.jl:
 * 719-850; tile indices it, jt, kt 1-based): the Hoffman-Ribak residual
 * (_generate_extended_residual, :197-236: fine normal minus its coarse
 * block mean), its isolated convolution (_isolated_convolve, :246-283: zero
 * padded to (2 nmesh)^3, sqrt(P(k) dk^3 (2 nmesh)^3) [times i k_d / k^2],
 * k = 0 zeroed, the kx = 0 and Nyquist planes made Hermitian as the
 * reference's GPU path does, ext/CUDAExt.jl:2561-2611), plus the coarse
 * field's Catmull-Rom interpolation (_interpolate_to_tile / _tricubic,
 * :95-122, 156-187) for delta and psi; then from the combined delta the
 * tile-periodic 2LPT displacements (src2 = delta^2/2 - sum phi_dd^2/2 -
 * sum_{a<b} phi_ab^2, :806-831) and k^2 field (:843-847) with
 * _apply_kernel_inplace!'s zeroing of k = 0 and the Nyquist planes
 * (:1239-1255).  fields: device arrays of nmesh^3 (eta2x..z all NULL to
 * skip 2LPT, lapd NULL to skip it).  coarse was made on ctx's device. */
int pp_split_tile_fields(pp_ctx *ctx, const pp_coarse *coarse, int it, int jt, int kt, pp_tile_fields *fields);

/* ------------------------------------------------------- halo records */
/* Per-peak inputs of the record packing besides the pp_results: the host
 * bookkeeping of the drivers (Pipeline.jl:217-231, 257-271).  Device arrays. */
typedef struct {
    int64_t npeaks;
    const double *x, *y, *z;  /* peak position, Mpc/h (PeakCandidate x, y, z)  */
    const double *Rf;         /* filter radius                                */
    const double *ZZon;       /* 1 + z of the peak                            */
    const double *D;          /* D(a_pk), Dlinear_ab                          */
    const double *v2fac;      /* 3/7 Om_a^(-1/143) D^2 (Pipeline.jl:270-271);
                                 NULL without 2LPT fields (zero velocities)   */
    const float *FcRf;        /* smoothed delta at the peak                   */
    const float *d2Rf;        /* wsmooth=3 field at the peak, may be NULL     */
    int on_device;            /* must be 1                                    */
} pp_halo_inputs;

/* Record packing of run_tile / run_multitile_split (Pipeline.jl:247-308;
 * MultiResolution.jl:1025-1094): every peak with RTHL > 0 and not
 * z > z_max (pass +INFINITY for no cut), in peak order, becomes a HaloRecord (11 Float32) or, with
 * extended, an ExtHaloRecord (33 Float32) in the .pksc field order
 * (src/IO/Catalog.jl:48-94).  res must hold device results of
 * pp_analyse_peaks for the same peaks.  *nhalo is the number of records
 * (> capacity => PP_ERR_CAPACITY, nothing written); *RTHLmax the largest
 * RTHL * alatt (0 when empty), the .pksc header value. */
int pp_pack_halos(pp_ctx *ctx, const pp_results *res, const pp_halo_inputs *in, double alatt, double z_max,
                  int extended, float *records, int64_t capacity, int records_on_device, int64_t *nhalo,
                  float *RTHLmax);

/* merge_catalog (src/Merger/Merger.jl:21-79; src/Merger/Exclusion.jl:14-222)
 * over n halos given by their catalogue columns x, y, z and RTHL (Float32,
 * device arrays when on_device, else host): Lagrangian exclusion in
 * descending RTHL order (stable: a surviving halo removes every not larger
 * halo whose centre lies inside its RTHL), then volume reduction of the
 * survivors (sphere_overlap caps of every overlapping pair,
 * r_new = ((4/3 pi r^3 - dV) / (4/3 pi))^(1/3), halos left without volume
 * removed).  keep[i] = 1 for the halos of the merged catalogue, whose RTHL
 * becomes new_rthl[i] (0 for removed halos); *nkept counts them.  keep and
 * new_rthl are device arrays when on_device, else host. */
int pp_merge_halos(pp_ctx *ctx, int64_t n, const float *x, const float *y, const float *z, const float *rthl,
                   int on_device, int8_t *keep, float *new_rthl, int64_t *nkept);

/* ------------------------------------------------------------- serving */
/* A tile server on one GPU: the per-device task loop of run_multitile_split
 * (src/MultiResolution.jl:1183-1203, one process_tile! per tile, :719-1180)
 * with continuous batching.  `slots` worker threads each own a pp_ctx on
 * `device`; a submitted tile waits (FIFO) until a slot is idle, the declared
 * footprints of the tiles in flight plus its own fit the admission budget and
 * the device's live free memory covers its footprint, then the slot's thread
 * calls fn(user, slot_ctx, slot, ticket, request).  fn does the tile's work
 * on that context and returns PP_OK or a PP_ERR_* code; a failing tile fails
 * alone.  With device = -1 the slots get no context (ctx = NULL): the
 * scheduling runs without a GPU. */
typedef struct pp_server pp_server;
typedef int (*pp_tile_fn)(void *user, pp_ctx *ctx, int slot, int64_t ticket, void *request);
typedef struct {
    int device;           /* GPU of the slots' contexts, or -1             */
    int slots;            /* 1..64                                         */
    int queue_capacity;   /* waiting tiles beyond which submit is refused  */
    size_t budget_bytes;  /* admission budget; 0: hbm_fraction x the free
                             device memory when the server starts          */
    double hbm_fraction;
    pp_tile_fn fn;
    void *user;
} pp_serve_config;

int pp_serve_create(const pp_serve_config *cfg, pp_server **out);
/* Queue a tile.  PP_ERR_BUSY: queue_capacity tiles already wait;
 * PP_ERR_CAPACITY: footprint_bytes exceeds the budget (it could never be
 * admitted); PP_ERR_STATE: the server is closing. */
int pp_serve_submit(pp_server *srv, void *request, size_t footprint_bytes, int64_t *ticket);
/* Wait up to timeout_ms (< 0: no limit) for the tile to finish: *status is
 * fn's return (or PP_ERR_INJECTED); times (may be NULL) receives seconds
 * since the server started at submission, start and completion, and the
 * slot.  PP_ERR_TIMEOUT if it has not finished in time. */
int pp_serve_poll(pp_server *srv, int64_t ticket, int timeout_ms, int *status, double *times);
/* The next finished tile in completion order (each reported once), with
 * the request pointer it was submitted with (request may be NULL). */
int pp_serve_next(pp_server *srv, int timeout_ms, int64_t *ticket, void **request, int *status, double *times);
/* pp_last_error() of a failed tile, as its worker thread left it. */
int pp_serve_message(pp_server *srv, int64_t ticket, char *buf, size_t len);
/* Fault injection for failure tests: the tile, not yet started, fails with
 * PP_ERR_INJECTED when its turn comes instead of running fn. */
int pp_serve_inject_fault(pp_server *srv, int64_t ticket);
/* stats[0..6]: waiting, in flight, most in flight so far, finished, failed,
 * reserved bytes, budget bytes. */
int pp_serve_stats(pp_server *srv, int64_t *stats);
/* The context of a slot (NULL with device = -1), e.g. to load tables. */
int pp_serve_context(pp_server *srv, int slot, pp_ctx **ctx);
/* Finish every queued tile, stop the workers, destroy the contexts. */
int pp_serve_destroy(pp_server *srv);

/* Per-stage device time of the last persistent-form call, milliseconds
 * (HIP events), ms[0 .. PP_NSTAGES): ms[0] preparation (from the fork of
 * the field packing, beside the peak set-up, ordering and offsets, to the
 * start of the gather), ms[1] wave gather, ms[2] small-peak gather, ms[3] wave post,
 * ms[4] lane post, ms[5] finish, ms[6] small-peak post.  The small-peak
 * gather and post run on the context's second stream beside ms[3..5], so
 * the stages overlap.  No counterpart in the reference (its profile=true
 * stage timers, MultiResolution.jl, time whole stages). */
#define PP_NSTAGES 7
int pp_last_shell_stages(pp_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* PEAKPATCH_HIP_H */
