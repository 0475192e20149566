// The reference's include path otlib/utils/rasterizer.h (ggael/otmap): the
// drop-in rasterize_image of include/otmap/rasterizer.h.
#pragma once

#include "../rasterizer.h"
