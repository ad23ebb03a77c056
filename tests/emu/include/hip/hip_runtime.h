// CPU stand-in for the HIP device environment, used only to execute the
// classify kernels' device code on the host for testing (one workgroup at a
// time, 64 threads in lockstep at every wave-collective operation).
#pragma once
#include "../simt_shim.h"
