// Drop-in for dmrg/GPUPlugin/BatchedGemm.h (BatchedGemm.h:1-56): the reference's global
// template BatchedGemm<T> and IntegerType (dmrg_types.h:4), implemented by dmrgpp_amd's
// header-only class over the C ABI. A DMRG++ build with -DDMRG_BATCHED_GEMM_TYPE=Plugin that puts
// this directory before dmrg/GPUPlugin on the include path and links libdmrgpp_amd.so compiles
// BatchedGemmPluginSc.h (its "#include "BatchedGemm.h"", the constructor call at :119-120 and
// the apply_Htarget(view, view) call at :159) unchanged.
#ifndef DMRGPPPLUGINSC_H
#define DMRGPPPLUGINSC_H

#include "dmrgpp_amd_batched_gemm.hpp"

using IntegerType = int;

template <typename T> using BatchedGemm = dmrgpp_amd::BatchedGemm<T>;

#endif
