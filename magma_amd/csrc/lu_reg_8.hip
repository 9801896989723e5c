// lu_reg_8.hip -- register-resident LU / Cholesky kernels, NMAX = 8.
#include "lu_reg.h"
MAGMA_AMD_LU_REG_INSTANTIATE(8)
