// job.hpp -- the device-resident job (internal to the library).
#pragma once

#include "common.hpp"

struct siting_job {
  int device = 0;
  hipStream_t stream = nullptr;
  siting::Geometry g{};
  void *base = nullptr;
  size_t bytes = 0;
  uint16_t *elevs = nullptr;
  uint16_t *elevs_t = nullptr;  // the DEM transposed, written by VIEWSHED
  uint8_t *vix = nullptr;
  int32_t *obs = nullptr;
  uint64_t *sheds = nullptr;
  siting::SiteBuffers site{};
  float *recip = nullptr;
  uint32_t *jump = nullptr;
  uint8_t *vstab = nullptr;  // VIEWSHED tables (vs_tables)
  hipEvent_t ev_start[4] = {};
  hipEvent_t ev_end[4] = {};
  bool ran[4] = {};          // stage events recorded
  bool site_state = false;   // SITE state written (job or batch run)
  bool owns_mem = true;      // base came from hipMalloc here
  bool owns_stream = true;   // stream created here
  int site_groups = 0;       // cap on the SITE loop's workgroups (0: auto)
};

namespace siting {

// Device bytes a job with these parameters occupies.
size_t job_bytes(const siting_params &p, const siting_derived &d);

// Creates a job.  mem: device memory of at least job_bytes() to carve the
// buffers from (nullptr: allocate); stream: the stream to use (nullptr:
// create one).  Borrowed memory and streams are not released on destroy.
siting_status job_create_in(int device, const siting_params *p, void *mem,
                            hipStream_t stream, siting_job **out);

}  // namespace siting
