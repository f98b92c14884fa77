// Test harness: the library's main translation unit built for the CPU emulation.
#include "hip/hip_runtime.h"
#include "../../maroontree_amd/csrc/mtgpu.hip"
