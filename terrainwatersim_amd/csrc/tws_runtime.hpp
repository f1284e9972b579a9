// tws_runtime.hpp -- host runtime pieces shared by the C-ABI front ends.
#pragma once

#include <cstdint>

#include "tws.h"
#include "tws_internal.hpp"

namespace tws {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int64_t pitch_for(int width);
void fill_ghosts(GridDesc& g, const tws_params& p);
bool valid_params(const tws_params* p);

// Launch configuration (kernel variant, steps per launch) for a grid size,
// from the tuning table in tws_tune.cpp.
void choose_launch(int width, int height, int* kernel, int* k);

}  // namespace tws
