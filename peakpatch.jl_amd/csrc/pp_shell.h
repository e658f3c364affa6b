// pp_shell.h -- internal interfaces between the shell-analysis sources.
#pragma once

#include <vector>

#include "pp_common.h"
#include "pp_post.h"

namespace pp {

// Device pointers of the PeakGrid fields (RadialShell.jl:45-56); eta2 / lapd
// may be NULL.
struct GatherFields {
    const float *d, *ex, *ey, *ez, *e2x, *e2y, *e2z, *lap;
    int n;
};

int check_fields(pp_ctx *ctx, const pp_fields *f);
int upload_fields(pp_ctx *ctx, const pp_fields *f, GatherFields &G);
int build_peak_params(pp_ctx *ctx, const pp_peak_batch *pb, int n, double alatt, double rmax2rs,
                      const int8_t *mask_host, std::vector<PeakParam> &out);
int prepare_post_const(pp_ctx *ctx, int n, int nbuff, double alatt, const float *lapd, int8_t *mask,
                       PostConst &C);
int run_persistent(pp_ctx *ctx, const GatherFields &G, std::vector<PeakParam> &pp, PostConst &C,
                   double *d_res, int32_t *d_st, PeakParam *d_pp);
int autotune_persistent(pp_ctx *ctx, const pp_fields *fields, const pp_peak_batch *peaks, double alatt,
                        double rmax2rs, int *choice);

}  // namespace pp
