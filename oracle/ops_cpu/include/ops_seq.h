/*
 * ORACLE — test infrastructure only. API v1 entry header of the oracle build: the loop
 * driver of ops_seq_v2.h (which also accepts pointer kernels) and the OPS_ACC<n>
 * accessor macros.
 */
#ifndef ORACLE_OPS_SEQ_H
#define ORACLE_OPS_SEQ_H
#include "ops_seq_v2.h"
#include "ops_macros.h"
#endif
