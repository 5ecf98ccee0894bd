// Stand-in for <hip/hip_runtime.h> in the host build of the kernels (tests/hostwave): the HIP
// qualifiers, launch and cross-lane builtins the kernel sources use, emulated with host threads.
#pragma once
#include "../hostwave.h"
