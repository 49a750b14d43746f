/*
 * Application entry header (the reference's ops/c/include/ops_seq_v2.h, API v2).
 *
 * With this backend every ops_par_loop call is rewritten by the translator
 * (python -m ops_amd.translator -t hip_gfx950) into a call of a generated
 * ops_par_loop_<kernel>() that launches a gfx950 kernel, so no host-side loop template
 * is provided here: an untranslated ops_par_loop call fails to compile, rather than
 * silently running on the CPU.
 */
#ifndef OPS_AMD_SEQ_V2_H
#define OPS_AMD_SEQ_V2_H
#include "ops_lib_core.h"
#include "ops_tridiag.h"

#ifdef __cplusplus
template <typename... T> struct ops_untranslated_loop { static constexpr bool value = false; };
template <typename... T> void ops_par_loop(T... args) {
  static_assert(ops_untranslated_loop<T...>::value,
                "ops_par_loop must be translated with ops_amd.translator (target hip_gfx950)");
}
#endif
#endif
