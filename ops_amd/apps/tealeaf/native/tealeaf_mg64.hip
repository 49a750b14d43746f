// The multigrid-preconditioned CG of tealeaf_mg.hip with its V-cycle in FP64 storage and
// arithmetic (entry point tea_native_mgcg64).
#define TEA_MG_FP64
#include "tealeaf_mg.hip"
