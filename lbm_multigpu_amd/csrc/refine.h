// refine.h -- 2:1 local refinement (coarse-fine interface) driver.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/lbm.h"

struct RefineState {
    bool active = false;
};

struct RefineLevels {
    void** ca; void** cb; const unsigned char* cmask; long long cqs; int cnxp;
    void** fa; void** fb; const unsigned char* fmask; long long fqs; int fnxp;
};

inline bool refine_validate(const lbm_config&, std::string& why)
{
    why = "local refinement is not available in this build";
    return false;
}
inline int refine_create(RefineState&, const lbm_config&, hipStream_t, std::string& err)
{
    err = "local refinement is not available in this build";
    return LBM_EINVAL;
}
inline void refine_destroy(RefineState&) {}
inline void refine_classify_fine(RefineState&, const lbm_config&, const std::vector<unsigned char>&,
                                 std::vector<unsigned char>&) {}
inline int refine_init(RefineState&, hipStream_t, std::string&) { return LBM_OK; }
inline int refine_step(RefineState&, const lbm_config&, const RefineLevels&, hipStream_t,
                       const std::function<int()>&, const std::function<int(float)>&, std::string& err)
{
    err = "local refinement is not available in this build";
    return LBM_EINVAL;
}
inline void refine_fine_coords(const lbm_config& c, int x, int y, int z, float& X, float& Y, float& Z)
{
    X = c.lr_x0 + 0.5f * x;
    Y = c.lr_y0 + 0.5f * y;
    Z = c.lr_z0 + 0.5f * z;
}
