// Test harness: the rate / trellis translation unit built for the CPU emulation.
#include "hip/hip_runtime.h"
#include "../../maroontree_amd/csrc/mt_rate.hip"
