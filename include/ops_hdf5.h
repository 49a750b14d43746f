/*
 * HDF5 dat I/O entry points of the reference (ops/c/include/ops_hdf5.h:61-183,
 * implemented there over libhdf5 in ops/c/src/externlib/ops_hdf5.cpp:68-1738). This image
 * has no HDF5 library, so each of them throws OPSException(OPS_HDF5_ERROR) naming itself:
 * applications that call them (apps/c/adi/adi.cpp:441-442 writes its solution) compile,
 * link and run up to that call.
 */
#ifndef OPS_AMD_HDF5_H
#define OPS_AMD_HDF5_H

#include "ops_lib_core.h"

#ifdef __cplusplus
extern "C" {
#endif

ops_dat ops_decl_dat_hdf5(ops_block block, int dat_size, char const *type, char const *dat_name,
                          char const *file_name);
ops_block ops_decl_block_hdf5(int dims, const char *block_name, char const *file_name);
ops_stencil ops_decl_stencil_hdf5(int dims, int points, const char *stencil_name,
                                  char const *file_name);
ops_stencil ops_decl_strided_stencil_hdf5(int dims, int points, const char *stencil_name,
                                          char const *file_name);
ops_halo ops_decl_halo_hdf5(ops_dat from, ops_dat to, char const *file_name);
void ops_fetch_dat_hdf5_file(ops_dat dat, char const *file_name);
void ops_fetch_block_hdf5_file(ops_block block, char const *file_name);
void ops_fetch_stencil_hdf5_file(ops_stencil stencil, char const *file_name);
void ops_fetch_halo_hdf5_file(ops_halo halo, char const *file_name);
void ops_read_dat_hdf5(ops_dat dat);
void ops_dump_to_hdf5(char const *file_name);
void ops_write_const_hdf5(char const *name, int dim, char const *type, char *const_data,
                          char const *file_name);
void ops_get_const_hdf5(char const *name, int dim, char const *type, char *const_data,
                        char const *file_name);

#ifdef __cplusplus
}
#endif

#endif /* OPS_AMD_HDF5_H */
