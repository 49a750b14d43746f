/* API v1 entry header (the reference's ops/c/include/ops_seq.h); see ops_seq_v2.h. */
#ifndef OPS_AMD_SEQ_H
#define OPS_AMD_SEQ_H
#include "ops_seq_v2.h"
#endif
