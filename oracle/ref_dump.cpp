// Test infrastructure (oracle side only): linked into the reference dev_seq
// builds under oracle/_ref by oracle/Makefile.ref. It wraps ops_exit
// (ops/c/src/core/ops_dummy_singlenode.cpp:983) and, before the reference tears
// its state down, writes every declared dat at full precision so that fixtures
// can be compared bit for bit. Nothing in the product links this file.
//
// Output, when the environment variable OPS_REF_DUMP_DIR names a directory:
//   <dir>/index.txt            one line per dat: name dim type_size ndims
//                              size[5] d_m[5] d_p[5] base[5] x_pad base_offset
//   <dir>/<name>.bin           the host buffer exactly as the reference holds it
//   <dir>/reductions.txt       name type size and the bytes (hex) of every reduction
//                              handle's last value
#include <cstdio>
#include <cstdlib>
#include <string>
#include <ops_lib_core.h>
#include <ops_instance.h>

extern "C" void __real__Z8ops_exitv();

extern "C" void __wrap__Z8ops_exitv() {
  const char *dir = getenv("OPS_REF_DUMP_DIR");
  if (dir != nullptr) {
    OPS_instance *inst = OPS_instance::getOPSInstance();
    std::string idx = std::string(dir) + "/index.txt";
    FILE *fi = fopen(idx.c_str(), "w");
    ops_dat_entry *item;
    TAILQ_FOREACH(item, &inst->OPS_dat_list, entries) {
      ops_dat d = item->dat;
      ops_get_data(d);
      size_t bytes = (size_t)d->elem_size;
      for (int i = 0; i < d->block->dims; i++) bytes *= (size_t)d->size[i];
      std::string name(d->name);  // some applications pad names with blanks
      while (!name.empty() && name.back() == ' ') name.pop_back();
      fprintf(fi, "%s %d %d %d", name.c_str(), d->dim, d->type_size, d->block->dims);
      for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->size[i]);
      for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->d_m[i]);
      for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->d_p[i]);
      for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->base[i]);
      fprintf(fi, " %d %zu\n", d->x_pad, d->base_offset);
      std::string fn = std::string(dir) + "/" + name + ".bin";
      FILE *fb = fopen(fn.c_str(), "wb");
      fwrite(d->data, 1, bytes, fb);
      fclose(fb);
    }
    fclose(fi);
    // reduction handles keep the last value read by ops_reduction_result
    std::string rfn = std::string(dir) + "/reductions.txt";
    FILE *fr = fopen(rfn.c_str(), "w");
    for (int i = 0; i < inst->OPS_reduction_index; i++) {
      ops_reduction r = inst->OPS_reduction_list[i];
      fprintf(fr, "%s %s %d", r->name, r->type, r->size);
      for (int b = 0; b < r->size; b++) fprintf(fr, " %02x", (unsigned char)r->data[b]);
      fprintf(fr, "\n");
    }
    fclose(fr);
  }
  __real__Z8ops_exitv();
}
