// UtcGpu.h -- umbrella header of the GPU part of the UTC task surface.
#pragma once

#include <hip/hip_runtime_api.h>

#include "GpuData.h"
#include "Utc.h"
