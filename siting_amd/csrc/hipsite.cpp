// hipsite.cpp -- drop-in command line for cusite / ompsite / seqsite
// (README.md:27-38; main() at siting.cpp:632-780).
//
//   hipsite nrows roi height targets bw nwanted intervis inputfile outputfile
//
// Same argument contract, validation order and messages (printed to stdout
// as "ERROR: <msg>", exit status 1, siting.cpp:22-25,663-695), same raw
// uint16 input, same observers file ("row,col" per selected observer in
// tentative-observer order, siting.cpp:761-765) and the same one-line
// stdout timing report (siting.cpp:735-779 with the SITE summary of
// siting.cpp:617).  The stages run on the GPU through libsiting_amd.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "siting.h"

namespace {

[[noreturn]] void die(const char *msg) {
  std::printf("ERROR: %s\n", msg);
  std::fflush(stdout);
  std::exit(1);
}

double now() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

void check(siting_status s) {
  if (s != SITING_OK) die(siting_status_str(s));
}

}  // namespace

int main(int argc, char **argv) {
  const double t_begin = now();
  double t_mark = t_begin;
  auto lap = [&]() {
    const double t = now();
    const double e = t - t_mark;
    t_mark = t;
    return e;
  };

  if (argc != 10) {
    std::printf("argc=%d\n", argc);
    die("SITE requires 9 arguments: nrows, roi, oht/tht, ntests, blocksize, "
        "nwanted, intervis, infile, outfile");
  }
  siting_params p{};
  p.nrows = std::atoi(argv[1]);
  p.roi = std::atoi(argv[2]);
  p.oht = p.tht = std::atoi(argv[3]);
  p.ntests = std::atoi(argv[4]);
  p.bw = std::atoi(argv[5]);
  p.nwanted0 = std::atoi(argv[6]);
  p.intervis = std::atoi(argv[7]);

  int which = 0;
  const char *msg = nullptr;
  if (siting_check_cli(&p, &which, &msg) != SITING_OK) die(msg);
  siting_derived d{};
  const siting_status ds = siting_derive(&p, &d);
  if (ds == SITING_ESMALLBLOCK) die("The last block is too small for nwantedperblock.");
  check(ds);

  const size_t n2 = (size_t)p.nrows * p.nrows;
  std::vector<uint16_t> elevs(n2);
  std::vector<int32_t> obs((size_t)d.nwanted);
  std::vector<uint8_t> selected((size_t)d.nwanted);
  if (siting_dem_read(argv[8], p.nrows, elevs.data()) != SITING_OK) die("Input failed");
  std::printf("input:%g", lap());

  siting_job *job = nullptr;
  check(siting_job_create(0, &p, &job));
  check(siting_job_upload_dem(job, elevs.data()));
  check(siting_job_vix(job));
  check(siting_job_sync(job));
  std::printf(" vix:%g", lap());

  check(siting_job_findmax(job));
  check(siting_job_sync(job));
  std::printf(" findmax:%g", lap());

  check(siting_job_viewshed(job));
  check(siting_job_sync(job));
  std::printf(" viewshed:%g", lap());

  check(siting_job_site(job));
  check(siting_job_sync(job));
  siting_stats st{};
  check(siting_job_stats(job, &st));
  check(siting_job_download(job, SITING_BUF_OBS, obs.data(), obs.size() * 4));
  check(siting_job_download(job, SITING_BUF_SELECTED, selected.data(), selected.size()));
  siting_job_destroy(job);
  std::printf(" nusedsheds:%d coverage:%g", st.nused,
              100.0 * (double)st.cumarea / (double)n2);
  std::printf(" site:%g", lap());

  FILE *fp = std::fopen(argv[9], "w");
  if (fp) {
    for (int i = 0; i < d.nwanted; i++)
      if (selected[(size_t)i])
        std::fprintf(fp, "%d,%d\n", obs[(size_t)i] / p.nrows, obs[(size_t)i] % p.nrows);
    std::fclose(fp);
  }
  std::printf(" output:%g", lap());
  std::printf(" total:%g\n", now() - t_begin);
  return 0;
}
