// kmc_mlp.h -- internal interface of the rate surrogate used by the KMC context.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cntkmc.h"

namespace cntkmc {

// Device of the surrogate and a status-returning preparation step (packs the inference weights
// after a parameter change); call before mlp_launch_pairs.
int mlp_device(const kmc_mlp* m);
int mlp_prepare(kmc_mlp* m);

// Fused inference over n CSR entries: entry p has class pair_cls[p] = ci * n_chir + cj of the
// context's chirality ids, mapped to surrogate classes by chir_class[n_chir] (-1 = unknown).
hipError_t mlp_launch_pairs(const kmc_mlp* m, int64_t n, const double* theta, const double* z,
                            const double* a1, const double* a2, const uint8_t* pair_cls,
                            int32_t n_chir, const int32_t* chir_class, double temperature,
                            double* rate, uint8_t* flag, hipStream_t st);

}  // namespace cntkmc
