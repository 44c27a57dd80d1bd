// srw_multie.h - partially coherent (multi-electron) accumulation kernels.
#pragma once
#include <hip/hip_runtime.h>
