// lu_reg_16.hip -- register-resident LU / Cholesky kernels, NMAX = 16.
#include "lu_reg.h"
MAGMA_AMD_LU_REG_INSTANTIATE(16)
