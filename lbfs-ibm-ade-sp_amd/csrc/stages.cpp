// Named stages of one time step, run one at a time.  They let a host drive
// the step the way the reference's gpu_solver does, kernel group by kernel
// group (Solver_gpu.cu:749-1246), and check each group on its own.  Running
// every stage of lbfs_stage_names in order and then lbfs_stage_end_step is
// one lbfs_step.
#include <hip/hip_runtime.h>

#include <cstring>

#include "context.hpp"

extern "C" int lbfs_stage_names(lbfs_ctx *ctx, int max_count, const char **names, int *count)
{
    if (!ctx || !count || max_count <= 0) return LBFS_E_INVALID_ARG;
    auto stages = ctx->step_stages(ctx->visc_due(ctx->t_step));
    *count = (int)stages.size() < max_count ? (int)stages.size() : max_count;
    if (names)
        for (int k = 0; k < *count; k++) names[k] = stages[k].name;
    return LBFS_OK;
}

extern "C" int lbfs_stage_run(lbfs_ctx *ctx, const char *name)
{
    if (!ctx || !name) return LBFS_E_INVALID_ARG;
    for (auto &st : ctx->step_stages(ctx->visc_due(ctx->t_step)))
        if (std::strcmp(st.name, name) == 0) return st.run();
    return ctx->fail(LBFS_E_INVALID_ARG, "no stage '%s' in this step", name);
}

extern "C" int lbfs_stage_end_step(lbfs_ctx *ctx)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    ctx->t_step++;
    return LBFS_OK;
}
