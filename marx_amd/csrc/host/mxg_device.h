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
/* relay_out (mirror, grating): the stage also hands its survivors densely
 * to the next stage, which must run next with relay_in set (with no other
 * stage, upload or list change in between) and reads them from there.  The
 * stage then leaves its survivors' x, p and flags in the store to the next
 * one, which stores them for every photon it reads. */
/* fresh: the active photons come straight from mxgd_launch_source with
 * entry_follows set (the mirror's entry writes their position, shell and
 * flags). */
int mxgd_run_mirror (double source_distance, unsigned long long seed, int slot, float *ms, int relay_out,
                     int fresh);
/* Timing slots that mxgd_run_mirror, given a slot, also fills with its two
 * kernels alone. */
#define MXGD_SLOT_HRMA_ENTER 4
#define MXGD_SLOT_HRMA_REFLECT 5
int mxgd_run_grating (unsigned long long seed, int slot, float *ms, int relay_in, int relay_out);
int mxgd_run_detect (double tstart, unsigned long long seed, int slot, float *ms, int relay_in);
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

/* As mxgd_run_source without reading anything back (no host round trip
 * before the next stage): the number created goes to the count of timing
 * slot `slot` (read with mxgd_read_timings), the stages that follow size
 * their launches for n, and mxgd_set_created must then be given that
 * number.  entry_follows: mxgd_run_mirror with fresh set runs next, and
 * the source leaves the photons' position, shell and flags to it. */
int mxgd_launch_source (unsigned int n, unsigned int tag_start, double start_time, double exposure_time,
                        unsigned long long seed, unsigned int source_id, int slot, int entry_follows);
int mxgd_set_created (unsigned int m);

/* ---- batched point sources ---- */

/* Uploads the directions (3 per source) and photon counts of a batch;
 * *n_total receives the number of photons. */
int mxgd_batch_set_sources (unsigned int nsrc, const double *dirs, const unsigned int *counts,
                            unsigned int *n_total);
/* Sizes the store for n photons and resets the list state; call before a
 * batch is launched or captured. */
int mxgd_batch_prepare (unsigned int n);
/* Tag of the first photon of the next launch (read by the kernel from
 * device memory, so a captured launch sequence picks it up). */
int mxgd_batch_set_tag_start (unsigned int tag_start);
/* Launches the batch source kernel over the n photons of the uploaded
 * sources: arrival times uniform over [0, exposure).  entry_follows: as
 * for mxgd_launch_source (mxgd_run_mirror with fresh set runs next). */
int mxgd_run_batch_source (unsigned int n, double exposure, double start_time,
                           unsigned long long seed, int slot, int entry_follows);
/* The whole batched pass (source, mirror, grating, detector) as one
 * persistent launch under the schedule's tile (not with a FEF file): the
 * same photons and final list as the staged pass; timing slot `slot` holds
 * the pass, the survivor counts land in the stage slots. */
int mxgd_run_batch_persistent (unsigned int n, double exposure, double start_time,
                               unsigned long long seed, int slot);

/* Launch schedule: blocks per CU of the stage kernels (the grating's 0: one
 * round of the blocks a CU holds at once) and the minimum waves
 * per SIMD of the reflection kernel (2, 3 or 4); persist_tile != 0: the
 * batched pass as one persistent launch of persist_blocks_per_cu blocks per
 * CU over tiles of persist_tile photons (a multiple of 256, at most 4096). */
enum { MXGD_K_SOURCE, MXGD_K_ENTER, MXGD_K_REFLECT, MXGD_K_GRATING, MXGD_K_DETECT, MXGD_NUM_K };
typedef struct
{
   unsigned int blocks_per_cu[MXGD_NUM_K];
   unsigned int reflect_waves;
   unsigned int persist_tile;
   unsigned int persist_blocks_per_cu;
}
mxgd_schedule;
int mxgd_set_schedule (const mxgd_schedule *s);
void mxgd_get_schedule (mxgd_schedule *s);

/* HIP graph of a launch sequence: mxgd_graph_begin starts capturing the
 * library's stream (stage runners then record no timing events and must be
 * given ms = NULL), mxgd_graph_end instantiates what was captured.
 * mxgd_graph_launch replays it (ms: its device time, or NULL) and restores
 * the list state the sequence left at capture. */
typedef struct mxgd_graph mxgd_graph;
/* Changes whenever a device buffer that a captured sequence may hold is
 * reallocated; a graph captured under another value must not be replayed. */
unsigned int mxgd_alloc_epoch (void);
int mxgd_graph_begin (void);
int mxgd_graph_end (mxgd_graph **g);
int mxgd_graph_launch (mxgd_graph *g, float *ms);
void mxgd_graph_destroy (mxgd_graph *g);

/* ---- differentiable mirror ---- */

/* Weighted forward pass over n photons with the shell geometry d; image is
 * a device pointer of spec->npix^2 floats; keeps the per-photon weights and
 * positions, and with tangents != 0 their derivatives.  Synchronises before
 * returning. */
int mxgd_diff_forward (const mxg_diff_cfg *d, unsigned int n, unsigned int tag_start,
                       unsigned long long seed, const mxg_image_spec *spec, float *image,
                       double *total_weight, int tangents);
/* Forward passes made so far (the id of the last one). */
unsigned long long mxgd_diff_pass (void);
/* grad[MXG_NUM_SHELLS * MXG_DIFF_NPAR] of forward pass `pass`, which must be
 * the last one, contracted with g (device, npix^2 floats, npix that pass's
 * image side).  Synchronises before returning. */
int mxgd_diff_adjoint (unsigned long long pass, unsigned int npix, const float *g, double *grad);
/* Device times (ms) of the last k_diff_forward, k_diff_adjoint, and
 * k_diff_enter + its list compaction; photons k_diff_forward ran on. */
void mxgd_diff_kernel_ms (float ms[3], unsigned int *listed);
/* Host copies of the per-photon state of the last forward pass. */
int mxgd_diff_state (unsigned int n, double *w, double *y, double *z, double *dw, double *dy,
                     double *dz, unsigned char *shell);

/* The configured dither applied to every photon of the store at its own
 * arrival time (photons replayed from a ray file).  Synchronises. */
int mxgd_dither_store (double start_time, unsigned long long seed);

/* ---- output files ---- */

/* Packs the columns cols[0..ncols) (MXG_COL_*) of the store's photons whose
 * flags & BAD_PHOTON_MASK is zero, in order, big-endian, into pinned host
 * memory: *bytes + col_off[k] holds column k's *n_rows values.  The memory
 * stays valid until the next call.  Synchronises before returning. */
int mxgd_pack_output (const int *cols, int ncols, double total_time, unsigned int *n_rows,
                      const unsigned char **bytes, size_t *col_off);

/* Event coordinates of the rows of mxgd_pack_output (same selection and
 * order) into the host arrays of *out (NULL ones skipped): the frame heads
 * of the held dither state, then k_event_coords.  subtab holds n_subtab
 * floats of subpixel tables indexed by par->sub_off / sub_n.  Fails when the
 * rows exceed capacity.  Synchronises. */
int mxgd_event_coords (const mxg_event_params *par, const float *subtab, unsigned int n_subtab,
                       double total_time, double time_del, unsigned long long seed,
                       const mxg_event_columns *out, unsigned int capacity, unsigned int *num_rows);

/* GPU time (ms) of the last mxgd_event_coords kernels (heads and coordinates). */
float mxgd_event_ms (void);

/* Number of photons in the active list (device -> host copy). */
int mxgd_active_count (unsigned int *n);
/* Active list indices (host copy, n = mxgd_active_count). */
int mxgd_active_list (unsigned int *idx, unsigned int n);

int mxgd_synchronize (void);

/* Pinned host memory for photon lists (NULL without a GPU: use malloc);
 * mxgd_host_free returns 1 when p was such an allocation (and frees it). */
void *mxgd_host_alloc (size_t bytes);
int mxgd_host_free (void *p);
/* The store's first n photons in energy order with the index as tie-break:
 * their indices and energies into host arrays.  Synchronises. */
int mxgd_sort_store_by_energy (unsigned int n, unsigned int *index, double *energies);
/* The active list (the last stage's survivors, in input order) and its
 * energies into host arrays of cap entries; *n its length.  Synchronises. */
int mxgd_download_active (unsigned int *index, double *energies, unsigned int cap, unsigned int *n);

/* The trace of mxg_trace_device into caller-owned columns v (capacity >=
 * n), with every list and counter in the library's fixed device scratch and
 * the configuration in the device buffers an earlier call allocated: no
 * allocation.  n <= MXG_NOALLOC_MAX_PHOTONS.  Synchronises. */
int mxgd_trace_noalloc (const mxg_soa_view *v, unsigned int n, unsigned int tag_start,
                        double start_time, double exposure_time, unsigned long long seed,
                        const mxg_config *cfg, const float *fa, unsigned int nf, const double *da,
                        unsigned int nd, unsigned int generation, unsigned int *n_collected,
                        unsigned int *n_detected);

#ifdef __cplusplus
}
#endif
#endif
