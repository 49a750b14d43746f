/*
 * Allocation helpers the reference's applications call (ops/c/include/ops_util.h:41-47;
 * used by apps/c/CloverLeaf/read_input.cpp:169 and apps/c/TeaLeaf/read_input.cpp:279).
 * They behave as the C library's malloc/realloc/calloc, except that a failed allocation
 * of a non-zero size prints a message and ends the process, as the reference's do.
 */
#ifndef OPS_AMD_UTIL_H
#define OPS_AMD_UTIL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

void *xmalloc(size_t size);
void *xrealloc(void *ptr, size_t size);
void *xcalloc(size_t number, size_t size);

#ifdef __cplusplus
}
#endif

#endif /* OPS_AMD_UTIL_H */
