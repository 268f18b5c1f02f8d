/* oracle.h -- FP64 CPU oracle of the MARX photon ray-trace.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's per-photon algorithm (marx/libsrc/{hrma,reflect,wfold,
 * diffract,detector,acis-*,hrc-*,hrcblur,drake,source,s-point,s-disk,
 * spectrum,dither}.c and the jdmath/src sources), written to follow the reference's
 * own loop structure: stage-major passes over an energy-sorted index list,
 * pruning between passes, shell-major grating passes, the HESF pass before
 * HRC detection.  It is used by the tests, by the smoke check and as the
 * CPU baseline of the benchmark, as the checker -- never by the library.
 *
 * Random numbers: where the reference draws from its single serial stream
 * (jdmath/src/random.c:151-154), the oracle draws from the per-photon
 * counter streams that the GPU library uses: Philox4x32-10 with
 * key = (seed low, seed high) and counter = (photon tag, stage, block,
 * source id), 4 draws per block, uniform = u32 / (2^32 - 1) as in
 * random.c:151-154.  Each photon-stage stream keeps its own cached second
 * polar-Gaussian variate (the reference keeps one global cache,
 * gaussrnd.c:30-55).  With that stream definition the GPU and the oracle
 * must agree photon for photon.
 *
 * Tables and geometry come from the library's initialised mxg_config
 * (include/marx_gpu_cfg.h), so both sides read identical inputs.
 */
#ifndef MARX_ORACLE_H
#define MARX_ORACLE_H

#include "marx_gpu.h"
#include "marx_gpu_cfg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Work-accounting marks, active only in the op-counting build (which
 * force-includes opcount/opcount.hpp): ORC_PHASE selects the bucket that
 * the following FP64 operations are counted in, ORC_MARK counts a photon
 * entering a bucket.  Buckets of the mirror pass: 0 up to and including the
 * paraboloid hit, 1 after it, 2 the energy-only optical-constant lookups. */
#ifdef ORC_OPCOUNT_HPP
# define ORC_PHASE(k) (Orc_Opcount_Phase = (k))
# define ORC_MARK(k) (Orc_Opcount_Marks[k]++)
#else
# define ORC_PHASE(k) ((void) 0)
# define ORC_MARK(k) ((void) 0)
#endif

typedef struct
{
   const mxg_config *cfg;
   const float *fa;
   const double *da;
   unsigned long long seed;
   unsigned int source_id;
}
orc_ctx;

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), one block. */
void orc_philox4x32_10 (const unsigned int ctr[4], const unsigned int key[2],
                        unsigned int out[4]);

/* First n uniforms / Gaussians of one photon-stage stream. */
void orc_stream_uniforms (unsigned long long seed, unsigned int tag,
                          unsigned int stage, unsigned int source_id,
                          double *u, unsigned int n);
void orc_stream_gaussians (unsigned long long seed, unsigned int tag,
                           unsigned int stage, unsigned int source_id,
                           double *g, unsigned int n);

/* jdmath restatements exposed for the known-answer tests. */
unsigned int orc_binary_search_f (float x, const float *xp, unsigned int n);
float orc_interpolate_f (float x, const float *xp, const float *yp, unsigned int n);
unsigned int orc_binary_search_d (double x, const double *xp, unsigned int n);
double orc_interpolate_d (double x, const double *xp, const double *yp, unsigned int n);
int orc_quadratic_root (double a, double b, double c, double *rplus, double *rminus);
double orc_reflectivity (double cos_theta, double beta, double delta);
double orc_wfold_interp (const orc_ctx *c, const mxg_wfold *w, double energy,
                         double sin_alpha, double r);
/* JDMinterpolate_n_fvector over one column set: xp sorted, out[k][i]. */
int orc_interpolate_n_fvector (const float *xp, float **yp, unsigned int n,
                               const float *oldxp, float **oldyp,
                               unsigned int oldn, unsigned int n_yp);

/* marx2fits over the n rows of an ACIS simulation's dump columns, in row
 * order (time as read from time.dat, chip pixels from xpixel/ypixel.dat,
 * b_energy.dat, the six dither columns or NULL): tiled, focal-plane and sky
 * pixels and flight grades.  subtab / sub_off / sub_n: subpixel tables
 * indexed [256 * t + fltgrade] (t 0 front-, 1 back-illuminated), each
 * table its energies, then its chip x, then its chip y offsets. */
int orc_event_coords (const mxg_config *cfg, const mxg_event_cfg *ec, unsigned int n,
                      const float *time, const unsigned int *tag, const signed char *ccd,
                      const float *xpixel, const float *ypixel, const float *b_energy,
                      const float *const *dither, const float *subtab,
                      const unsigned int *sub_off, const unsigned int *sub_n,
                      int *tdetx, int *tdety, double *detx, double *dety,
                      double *sky_x, double *sky_y, short *fltgrade);
/* acis_subpix.c:283-322 over one table of n energies (then n chip x and n
 * chip y offsets). */
int orc_subpix (const float *e, unsigned int n, float energy, float *dx, float *dy);

/* Stage-major passes, in the order marx/src/marx.c:240-273 calls them. */
int orc_create_photons (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int num,
                        unsigned int *num_collected, double *exposure_time);
/* A USER source: ray is the plug-in's user_create_ray; tstart the time
 * origin of the exposure cut of plug-in times (_Marx_TStart_MJDsecs). */
int orc_create_user (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int num,
                     unsigned int *num_collected, double *exposure_time, double tstart,
                     int (*ray) (double *, double *, double *, double *, double *));
/* The dither applied to the photons of pt as replayed from a ray file. */
int orc_dither_photons (const orc_ctx *c, Marx_Photon_Type *pt);
/* The differentiable mirror (mxg_diff_forward / mxg_diff_adjoint): weights,
 * focal-plane positions and their derivatives per photon, an FP64 image,
 * and the gradient for d(loss)/d(image) = g. */
int orc_diff_forward (const orc_ctx *c, const mxg_diff_cfg *d, unsigned int n,
                      unsigned int tag_start, const mxg_image_spec *sp, double *image,
                      double *w, double *y, double *z, double *dw, double *dy, double *dz,
                      unsigned char *shell);
int orc_diff_adjoint (unsigned int n, const double *w, const double *y, const double *z,
                      const double *dw, const double *dy, const double *dz,
                      const unsigned char *shell, const mxg_image_spec *sp, const double *g,
                      double *grad);
/* The same on nthreads OpenMP threads (0: all); per-thread images and
 * partial gradients are summed in thread order. */
int orc_diff_forward_parallel (const orc_ctx *c, const mxg_diff_cfg *d, unsigned int n,
                               unsigned int tag_start, const mxg_image_spec *sp, double *image,
                               double *w, double *y, double *z, double *dw, double *dy,
                               double *dz, unsigned char *shell, int nthreads);
int orc_diff_adjoint_parallel (unsigned int n, const double *w, const double *y, const double *z,
                               const double *dw, const double *dy, const double *dz,
                               const unsigned char *shell, const mxg_image_spec *sp,
                               const double *g, double *grad, int nthreads);

/* The photons of a batch of point sources (mxg_trace_batch). */
int orc_create_batch (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int nsrc,
                      const double *dirs, const unsigned int *counts, unsigned int tag_start,
                      double start_time, double exposure);
int orc_mirror_reflect (const orc_ctx *c, Marx_Photon_Type *pt);
/* Test tap of orc_mirror_reflect: taps[18 i + 6 k + (0..5)] receives the ray
 * (x, then p) of the photon at attribute index i < n as it reaches strut set
 * k (0 pre-, 1 cap, 2 post-collimator), whether or not struts are on.
 * NULL switches the tap off. */
void orc_set_strut_taps (double *taps, unsigned int n);
int orc_grating_diffract (const orc_ctx *c, Marx_Photon_Type *pt);
int orc_detect (const orc_ctx *c, Marx_Photon_Type *pt);

/* Sorts the photons by energy (index tie-break) into sorted_index /
 * sorted_energies as source.c:361-371 does, allocating sorted_index. */
int orc_sort_photons (Marx_Photon_Type *pt);
void orc_prune_photons (Marx_Photon_Type *pt);

/* Whole pipeline over num photons, OpenMP-parallel over photons (chunks of
 * chunk photons, each chunk traced stage-major).  Used as the CPU baseline.
 * Returns the number of photons detected, -1 on error. */
long orc_trace_parallel (const orc_ctx *c, unsigned int num, unsigned int chunk,
                         int nthreads);
/* A batch of point sources (orc_create_batch), one OpenMP task per source. */
long orc_trace_batch_parallel (const orc_ctx *c, unsigned int nsrc, const double *dirs,
                               const unsigned int *counts, unsigned int tag_start,
                               double exposure, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
