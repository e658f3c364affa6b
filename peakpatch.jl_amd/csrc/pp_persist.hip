// pp_persist.hip -- persistent work-queue form of the shell analysis.
#include "pp_shell.h"

namespace pp {

int run_persistent(pp_ctx *ctx, const GatherFields &, std::vector<PeakParam> &, PostConst &, double *,
                   int32_t *, PeakParam *)
{
    (void)ctx;
    set_error("persistent shell kernel not available in this build");
    return PP_ERR_STATE;
}

int autotune_persistent(pp_ctx *, const pp_fields *, const pp_peak_batch *, double, double, int *)
{
    set_error("persistent shell kernel not available in this build");
    return PP_ERR_STATE;
}

int shell_groups_build(pp_ctx *) { return PP_OK; }

}  // namespace pp
