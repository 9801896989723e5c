// lu_reg.hip -- size dispatch for the register-resident factorizations
// (kernels in lu_reg.h, one translation unit per NMAX).
#include "lu_reg.h"

namespace magma_amd {

magma_int_t launch_getrf_small_v(const MatSet &s, int max_mn_dim, bool solve, bool var, hipStream_t st);
magma_int_t launch_potrf_small_v(const MatSet &s, int max_n, bool solve, bool var, hipStream_t st);

magma_int_t launch_getrf_small(const MatSet &s, int max_mn_dim, bool solve, hipStream_t st)
{
    return launch_getrf_small_v(s, max_mn_dim, solve, false, st);
}

magma_int_t launch_getrf_small_v(const MatSet &s, int max_mn_dim, bool solve, bool var, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    if (max_mn_dim <= 8) return launch_getrf_reg_n<8>(s, solve, var, st);
    if (max_mn_dim <= 16) return launch_getrf_reg_n<16>(s, solve, var, st);
    if (max_mn_dim <= 32) return launch_getrf_reg_n<32>(s, solve, var, st);
    return launch_getrf_reg_n<64>(s, solve, var, st);
}

magma_int_t launch_potrf_small(const MatSet &s, int max_n, bool solve, hipStream_t st)
{
    return launch_potrf_small_v(s, max_n, solve, false, st);
}

magma_int_t launch_potrf_small_v(const MatSet &s, int max_n, bool solve, bool var, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    if (max_n <= 8) return launch_potrf_reg_n<8>(s, solve, var, st);
    if (max_n <= 16) return launch_potrf_reg_n<16>(s, solve, var, st);
    if (max_n <= 32) return launch_potrf_reg_n<32>(s, solve, var, st);
    return launch_potrf_reg_n<64>(s, solve, var, st);
}

}  // namespace magma_amd
