// Wavefront path tracer: device scene data and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "isect.h"

namespace bgpu {

struct PtConfig {
  int block = 128;
};

struct PtScene {
  std::vector<int32_t> shape_light_host;  // device light index per shape, -1 if none
};

}  // namespace bgpu
