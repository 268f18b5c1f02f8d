/* mxg_device.h -- extern "C" interface of the HIP layer of libmarx_gpu.
 *
 * The host C code owns configuration and the AoS photon list; the HIP layer
 * owns HBM: the configuration copy, the structure-of-arrays photon store,
 * the active-photon index lists and the kernels.  All functions return 0 on
 * success and -1 on error (message in mxgd_last_error()).
 */
#ifndef MXG_DEVICE_H
#define MXG_DEVICE_H

#include "marx_gpu.h"
#include "marx_gpu_cfg.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *mxgd_last_error (void);
int mxgd_init (void);
int mxgd_device_info (char *name, unsigned int namelen, int *num_cus);

/* Copies the configuration and arenas to HBM when generation changed. */
int mxgd_sync_config (const mxg_config *cfg,
                      const float *fa, unsigned int nf,
                      const double *da, unsigned int nd,
                      unsigned int generation);

/* AoS list <-> resident SoA store.  The active list becomes the given
 * index list (the reference's sorted_index). */
int mxgd_upload_aos (const Marx_Photon_Attr_Type *at, unsigned int n,
                     const unsigned int *active, unsigned int nactive);
int mxgd_download_aos (Marx_Photon_Attr_Type *at, unsigned int n);
unsigned int mxgd_store_size (void);
int mxgd_get_soa (mxg_soa_view *view);

/* Stages on the resident store.  Each one processes the current active
 * list and replaces it by the photons still unabsorbed. */
/* slot >= 0 records HIP events around the stage's kernels into timing slot
 * `slot` (read with mxgd_read_timings); ms, when non-NULL, synchronises and
 * returns the stage time at once. */
int mxgd_run_mirror (double source_distance, unsigned long long seed, int slot, float *ms);
/* Timing slots that mxgd_run_mirror, given a slot, also fills with its two
 * kernels alone. */
#define MXGD_SLOT_HRMA_ENTER 4
#define MXGD_SLOT_HRMA_REFLECT 5
int mxgd_run_grating (unsigned long long seed, int slot, float *ms);
int mxgd_run_detect (double tstart, unsigned long long seed, int slot, float *ms);
/* Synchronises and returns the times (ms) of the stages recorded in slots
 * 0..nslots-1 and, for stage slots, the survivor count each one left. */
int mxgd_read_timings (float *ms, unsigned int *counts, int nslots);

/* Source: creates n photons (tags tag_start..) with energies, directions,
 * arrival times and dither; when exposure_time > 0 the list ends at the
 * first photon whose arrival time reaches it (source.c:313-342).  The
 * active list becomes those photons; *n_out receives their number and
 * *last_time the arrival time of the last one. */
int mxgd_run_source (unsigned int n, unsigned int tag_start, double start_time,
                     double exposure_time, unsigned long long seed, unsigned int source_id,
                     int slot, unsigned int *n_out, double *last_time, float *ms);

/* Number of photons in the active list (device -> host copy). */
int mxgd_active_count (unsigned int *n);
/* Active list indices (host copy, n = mxgd_active_count). */
int mxgd_active_list (unsigned int *idx, unsigned int n);

int mxgd_synchronize (void);

#ifdef __cplusplus
}
#endif
#endif
