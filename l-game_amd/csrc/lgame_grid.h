// lgame_grid.h -- grid generation on the device (lgame_grid.hip), behind lgame_grid_setup.
#pragma once

#include <string>

#include "../../include/lgame_gpu.h"

namespace lgame {

// Fills out from dims and params (grid_setup and bg_setup of mo_grid_generator.F90); returns an
// LGAME_* code and, on failure, a message in err.
int grid_generate(const lgame_dims &dims, const lgame_grid_params &params, lgame_grid_out &out, std::string &err);

}  // namespace lgame
