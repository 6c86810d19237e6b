// libvdhmm_core.so: the reference's base/core/core.cpp:9,176 names
// (read_haplotype / read_haplotype_batch with per-row normalised, mode-0
// semantics), forwarded to libvdhmm.so.  It exists because libwide.so
// (wide.cpp:9) exports a different read_haplotype under the same name, so a
// drop-in for each reference library needs its own shared object.
#include "vdhmm.h"

extern "C" double read_haplotype(const char* read, const unsigned char* qual, int n, const char* hap, int m,
                                 double gap, double ext) {
  return vd_core_read_haplotype(read, qual, n, hap, m, gap, ext);
}

extern "C" void read_haplotype_batch(const char* read, const unsigned char* qual, int n, const char* const* hap,
                                     const int* lengths, int count, double gap, double ext, double* out) {
  vd_core_read_haplotype_batch(read, qual, n, hap, lengths, count, gap, ext, out);
}
