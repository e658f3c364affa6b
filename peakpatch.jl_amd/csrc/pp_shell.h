// pp_shell.h -- internal interfaces between the shell-analysis sources.
#pragma once

#include <vector>

#include "pp_common.h"
#include "pp_post.h"

namespace pp {

// Device pointers of the PeakGrid fields (RadialShell.jl:45-56); eta2 / lapd
// may be NULL.  dT is the library's transposed copy of delta (dT(x,y,z) =
// delta(z,x,y)), made by the persistent form.
struct GatherFields {
    const float *d, *ex, *ey, *ez, *e2x, *e2y, *e2z, *lap;
    const float *dT;
    const float4 *p4;  // persistent form: (delta, delta transposed, eta x, eta y) per cell, or null
    const float4 *q4;  // persistent form with eta2: (eta z, eta2 x, eta2 y, eta2 z) per cell, or null
    int n;
};

struct BatchDev {
    const int32_t *i, *j, *k;
    const double *Rf;
    const int64_t *ir2min;
    const double *ZZon, *fcrit;
};

// One pp_analyse_peaks call: device batch, per-peak params, results.
struct ShellRun {
    int64_t np = 0;
    BatchDev batch;
    PeakParam *pp = nullptr;   // params in the caller's order
    int32_t *order = nullptr;  // processing order (indices into pp)
    int32_t *keys = nullptr;   // scratch for the sort
    ResultView R;
};

int check_fields(pp_ctx *ctx, const pp_fields *f);
int upload_fields(pp_ctx *ctx, const pp_fields *f, GatherFields &G);
int prepare_post_const(pp_ctx *ctx, int n, int nbuff, double alatt, const float *lapd, int8_t *mask,
                       PostConst &C);
int prepare_batch(pp_ctx *ctx, const pp_peak_batch *pb, int n, double alatt, double rmax2rs,
                  const int8_t *d_mask, ShellRun &S, const pp_results *out = nullptr);
int run_plain(pp_ctx *ctx, const GatherFields &G, ShellRun &S, PostConst &C, std::vector<int64_t> *offs_out);
int start_pack(pp_ctx *ctx, GatherFields &G);
int run_persistent(pp_ctx *ctx, GatherFields G, ShellRun &S, PostConst &C);
int persistent_profiles(pp_ctx *ctx, GatherFields G, ShellRun &S, PostConst &C, std::vector<int32_t> &cnt,
                        std::vector<double> &fbar, std::vector<int64_t> &off);

}  // namespace pp
