/* API v1 entry header (the reference's ops/c/include/ops_seq.h): the v2 entry header plus
 * the OPS_ACC<n> / OPS_ACC_MD<n> pointer accessors (ops_macros.h). */
#ifndef OPS_AMD_SEQ_H
#define OPS_AMD_SEQ_H
#include "ops_seq_v2.h"
#include "ops_macros.h"
#endif
