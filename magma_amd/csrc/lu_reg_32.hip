// lu_reg_32.hip -- register-resident LU / Cholesky kernels, NMAX = 32.
#include "lu_reg.h"
MAGMA_AMD_LU_REG_INSTANTIATE(32)
