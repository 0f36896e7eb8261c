// Thread-index state and launch-order switch of the CPU stand-in (TEST INFRASTRUCTURE).
#include "hip/hip_runtime.h"

thread_local emu_uint3 blockIdx, threadIdx;
thread_local dim3 blockDim, gridDim;

static int g_order = 0;
int emu_launch_order() { return g_order; }
extern "C" void emu_set_launch_order(int order) { g_order = order; }
