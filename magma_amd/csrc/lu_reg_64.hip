// lu_reg_64.hip -- register-resident LU / Cholesky kernels, NMAX = 64.
#include "lu_reg.h"
MAGMA_AMD_LU_REG_INSTANTIATE(64)
