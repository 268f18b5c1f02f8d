// dev_store.h -- the structure-of-arrays photon store in HBM and the launch
// helpers shared by the stage kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "marx_gpu.h"
#include "marx_gpu_cfg.h"

namespace mxg {

// Column pointers of the resident photon store (one element per photon).
struct Soa
{
   double *energy;
   double *x0, *x1, *x2;
   double *p0, *p1, *p2;
   double *t;
   uint32_t *flags;
   float *ypix, *zpix, *upix, *vpix;
   float *dith[6];
   float *pi;
   int16_t *pha;
   uint8_t *shell;
   int8_t *ccd, *region, *order;
   int8_t *sup[4];
   uint32_t *tag;
};

// Per-stage launch arguments: the active index list in, the survivors out.
struct ListIO
{
   const uint32_t *idx_in;
   const uint32_t *n_in;
   uint32_t *idx_out;
   uint32_t *n_out;
};

constexpr int kBlock = 256;

// Stage launchers (defined in the stage .hip files).
hipError_t launch_hrma (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                        ListIO io, uint32_t n_upper, double source_distance,
                        unsigned long long seed, uint32_t source_id, hipStream_t st);
hipError_t launch_grating (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                           ListIO io, uint32_t n_upper, unsigned long long seed,
                           uint32_t source_id, hipStream_t st);
hipError_t launch_detect (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                          ListIO io, uint32_t n_upper, double tstart,
                          unsigned long long seed, uint32_t source_id, hipStream_t st);

hipError_t launch_source (const mxg_config *cfg, const double *da, Soa s, uint32_t n,
                          uint32_t tag_start, double exposure_time, double start_time,
                          unsigned long long seed, uint32_t source_id, uint64_t *ticks,
                          uint64_t *tile_tot, uint32_t *first_over, uint32_t *count,
                          uint32_t *idx, hipStream_t st);

} // namespace mxg
