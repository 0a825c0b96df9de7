// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference's own siting.cpp (taken from where it lies, never
// copied: the path comes from the SITING_REF_SRC macro set by oracle/Makefile)
// into a shared library with plain C entry points, so the CPU restatement in
// siting_oracle.c can be checked stage by stage against the reference's
// sequential build (-DSEQ -O2, makefile:10) -- and, built with -fopenmp
// instead, the reference's OpenMP stages are timed as the CPU baseline.
// main() of the reference is
// renamed so it does not clash; it is never called from here.
#define main siting_reference_main
#include SITING_REF_SRC
#undef main

extern "C" {

void ref_calc_vix(int nrows, const unsigned short *elevs, int roi, int oht,
                  int tht, int ntests, unsigned char *vix) {
  calc_vix(nrows, elevs, roi, oht, tht, ntests, vix);  // siting.cpp:192
}

void ref_find_max(int nrows, unsigned char *vix, float blocksize, int nwanted,
                  int nblocks, int nwpb, int *obs) {
  find_max(nrows, vix, blocksize, nwanted, nblocks, nwpb, obs);  // :235
}

void ref_calc_sheds(int nrows, const unsigned short *elevs, int roi, int oht,
                    int tht, int nsheds, const int *obs,
                    unsigned long long *sheds) {
  calc_sheds(nrows, elevs, roi, oht, tht, nsheds, obs, sheds);  // :349
}

// site_it prints " nusedsheds:N coverage:C" to stdout (siting.cpp:617).
void ref_site_it(int nrows, int roi, int intervis, int nsheds, const int *obs,
                 const unsigned long long *sheds, char *selected) {
  site_it(nrows, roi, intervis, nsheds, obs, sheds, selected);  // :501
  cout.flush();
}

int ref_test_one_target(int nrows, const unsigned short *elevs, int ox, int oy,
                        int oz, int tx, int ty, int tz) {
  return test_one_target(nrows, elevs, ox, oy, oz, tx, ty, tz);  // :54
}

// The OpenMP build (no -DSEQ, -fopenmp, as makefile:8) times the reference's
// parallel stages on the CPU.  main() fixes 50 threads (siting.cpp:634-635);
// called from here, the stages take the count set by this function.
int ref_set_threads(int n) {
#ifdef _OPENMP
  omp_set_dynamic(0);
  omp_set_num_threads(n);
  return n;
#else
  return 1;
#endif
}

}  // extern "C"
