// tune.h -- the kernel-variant table consulted by the dispatchers and filled by
// the autotuner (tune.hip).  It replaces the reference's static tuning tables
// (control/get_batched_crossover.cpp:25-228, control/batched_kernel_param.h,
// control/get_batched_gemm_decision.cpp, magmablas/gemm_config/).
#pragma once

#include "internal.h"

namespace magma_amd {

enum TuneFamily {
    kTuneGetrf = 0,    // LU, 64 < n <= 512: 0 MFMA kernel at full occupancy, 1 MFMA one workgroup
                       //   per CU, 2 VALU blocked kernel, 3 / 4 register-resident kernel at 3 / 2 waves per
                       //   SIMD (n <= 128 only)
    kTuneMixed = 1,    // mixed-precision solver: block-column width / right-hand-side group variants,
                       //   and workgroups per CU
    kTuneGemm = 2,     // batched / vbatched DGEMM: 0 32x32 tiles per wavefront, 1 16x16, 2 48x48,
                       //   3 64x64 per workgroup through LDS
    kTuneFamilies = 3
};
constexpr int kTuneClasses = 10;     // size classes of workqueue.h

// Variant for (family, size class), the default when untuned or out of range.
int tuned_variant(TuneFamily f, int size_class);
// C tile edge of a DGEMM variant
inline int gemm_tile_of(int variant) { return variant == 1 ? 16 : (variant == 2 ? 48 : (variant == 3 ? 64 : 32)); }
int size_class_of(int dim);
// number of variants of a family
int tune_variant_count(TuneFamily f);

}  // namespace magma_amd
