/* marx_gpu.h -- C-ABI of libmarx_gpu, the MI355X photon ray-trace library.
 *
 * The library is a drop-in for the libmarx entry points that the marx
 * driver calls on its per-photon hot path (marx/src/marx.c:240-273 and
 * :641-656).  Every entry point below keeps the reference's name, argument
 * meaning and 0 / -1 / module-id return convention.  The photon record
 * keeps the reference's in-memory layout byte for byte, so an unchanged
 * marx.o can hand its Marx_Photon_Type to this library.
 *
 * Functions whose names start with mxg_ have no counterpart in the
 * reference: they expose the device-resident pipeline (photons created,
 * traced and binned without leaving HBM), the batched multi-source launch
 * and the survival-weighted / adjoint mode.
 *
 * All signatures use plain C types and pointers.  Device pointers are
 * passed as void * or as typed pointers documented as device memory.
 */
#ifndef MARX_GPU_H
#define MARX_GPU_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Photon list.  Layout identical to marx/libsrc/marx.h:32-151 (136 bytes per
 * attribute record, offsets energy@0 x@8 p@32 arrival_time@56 flags@64
 * y_pixel@68 dither_state@84 pi@108 pulse_height@112 mirror_shell@116
 * ccd_num@120 order@122 support_orders@123 tag@128).                      */

typedef struct
{
   double x, y, z;
}
JDMVector_Type;

typedef struct
{
   float ra, dec, roll;
   float dy, dz, dtheta;
}
Marx_Dither_Type;

typedef struct
{
   double energy;
   JDMVector_Type x;
   JDMVector_Type p;
   double arrival_time;
   unsigned int flags;
   float y_pixel;
   float z_pixel;
   float u_pixel;
   float v_pixel;
   Marx_Dither_Type dither_state;
   float pi;
   short pulse_height;
   unsigned int mirror_shell;
   signed char ccd_num;
   signed char detector_region;
   signed char order;
   signed char support_orders[4];
   unsigned int tag;
}
Marx_Photon_Attr_Type;

/* Absorption / history bits of Marx_Photon_Attr_Type.flags (marx.h:65-75) */
#define PHOTON_UNDETECTED        0x01
#define PHOTON_UNREFLECTED       0x02
#define PHOTON_UNDIFFRACTED      0x04
#define PHOTON_MISSED_DETECTOR   0x08
#define PHOTON_MIRROR_VBLOCKED   0x10
#define PHOTON_DRAKE_BLOCKED     0x20
#define PHOTON_GRATING_VBLOCKED  0x40
#define BAD_PHOTON_MASK          0xFF
#define PHOTON_DRAKE_REFLECTED   0x100
#define PHOTON_ACIS_STREAKED     0x200

typedef struct
{
   double source_distance;
   Marx_Photon_Attr_Type *attributes;
   unsigned int n_photons;
   unsigned int max_n_photons;
   double total_time;
   double start_time;
   unsigned int *sorted_index;
   double *sorted_energies;
   unsigned int num_sorted;
   unsigned long history;
   unsigned int tag_start;
}
Marx_Photon_Type;

/* Marx_Photon_Type.history bits (marx.h:126-147) */
#define MARX_ENERGY_OK        0x00000001
#define MARX_TIME_OK          0x00000002
#define MARX_X_VECTOR_OK      0x00000004
#define MARX_P_VECTOR_OK      0x00000008
#define MARX_TAG_OK           0x00000010
#define MARX_PULSEHEIGHT_OK   0x00000020
#define MARX_PI_OK            0x00000040
#define MARX_DET_PIXEL_OK     0x00000080
#define MARX_DET_NUM_OK       0x00000100
#define MARX_DET_REGION_OK    0x00000200
#define MARX_DET_UV_PIXEL_OK  0x00000400
#define MARX_MIRROR_SHELL_OK  0x00000800
#define MARX_SKY_DITHER_OK    0x00001000
#define MARX_DET_DITHER_OK    0x00002000
#define MARX_ORDER_OK         0x00100000
#define MARX_ORDER1_OK        0x00200000
#define MARX_ORDER2_OK        0x00400000
#define MARX_ORDER3_OK        0x00800000
#define MARX_ORDER4_OK        0x01000000

/* Module ids returned by the init entry points (marx.h:364-376). */
#define MARX_DETECTOR_HRC_S   1
#define MARX_DETECTOR_HRC_I   2
#define MARX_DETECTOR_ACIS_S  3
#define MARX_DETECTOR_ACIS_I  4
#define MARX_MIRROR_HRMA      2
#define MARX_GRATING_HETG     1
#define MARX_GRATING_LETG     2

/* Photon list management.  Replaces marx/libsrc/photon.c:40-108
 * (declared at marx.h:155-157). */
Marx_Photon_Type *marx_alloc_photon_type (unsigned int num);
int marx_dealloc_photon_type (Marx_Photon_Type *pt);
/* Rebuilds sorted_index / sorted_energies from the photons whose
 * flags & BAD_PHOTON_MASK is zero, keeping their order (photon.c:40-63). */
void marx_prune_photons (Marx_Photon_Type *pt);

/* ------------------------------------------------------------------------ */
/* Parameter files.  The subset of the pfile API (pfile/src/pfile.h:43-64)
 * that the path's init functions use, over IRAF-style .par files
 * (name,type,mode,value,min,max,prompt).  Getters return 0 on success and
 * -1 when the parameter is absent or malformed.                            */

typedef struct _Param_File_Type Param_File_Type;

Param_File_Type *pf_open_parameter_file (char *file, char *mode);
int pf_close_parameter_file (Param_File_Type *pf);
int pf_get_double (Param_File_Type *pf, char *name, double *v);
int pf_get_integer (Param_File_Type *pf, char *name, int *v);
int pf_get_boolean (Param_File_Type *pf, char *name, int *v);
int pf_get_string (Param_File_Type *pf, char *name, char *buf, unsigned int buflen);
int pf_get_file (Param_File_Type *pf, char *name, char *buf, unsigned int buflen);
/* Replaces (or adds) the value of a parameter; the equivalent of a
 * name=value command-line override (marx/libsrc/util.c:238). */
int pf_set_value (Param_File_Type *pf, char *name, char *value);

/* Directory that relative calibration-file names are resolved against
 * (marx/libsrc/data.c:83-126).  "$VAR" takes the value of an environment
 * variable.  When never set, the data directory bundled with the library
 * is used. */
int marx_set_data_directory (char *dir);

/* The observation start, in years and in seconds from MJDref (1998.0), as
 * marx.c's setup_tstart computes it from TStart before module set-up
 * (marx/libsrc/marx.h:578, util.c:231-236).  Used by the exposure cut of
 * photons with their own times and by the ACIS contamination's time
 * interpolation, read at detector set-up. */
int marx_set_time (double tyrs, double mjdsecs);
extern double _Marx_TStart_MJDsecs;
extern double _Marx_TStart_Yrs;

/* ------------------------------------------------------------------------ */
/* Hot-path entry points (marx/libsrc/marx.h:354-362), called from
 * marx/src/marx.c:240-273 (run) and :641-656 (init).  The init functions
 * return the module id (0 for NONE) or -1 on error.  The run functions
 * return 0 or -1.  Each run function sets its history bit and returns at
 * once when that bit is already set, as the reference stages do.           */

/* Replaces marx/libsrc/mirror.c:48-91 (HRMA: hrma.c:1097-1157). */
int marx_mirror_init (Param_File_Type *pf);
/* Replaces marx/libsrc/mirror.c:93-108 -> hrma.c:1161-1341. */
int marx_mirror_reflect (Marx_Photon_Type *pt, int verbose);
/* Replaces marx/libsrc/grating.c:44-86 (diffract.c:1687-1863). */
int marx_grating_init (Param_File_Type *pf);
/* Replaces marx/libsrc/grating.c:88-114 -> diffract.c:974-1130. */
int marx_grating_diffract (Marx_Photon_Type *pt, int verbose);
/* Replaces marx/libsrc/detector.c:298-357. */
int marx_detector_init (Param_File_Type *pf);
/* Replaces marx/libsrc/detector.c:361-379 (acis-s.c:177-248,
 * acis-i.c:96-167, hrc-s.c:236-312, hrc-i.c:118-185). */
int marx_detect (Marx_Photon_Type *pt, int verbose);

/* Geometric collecting area of the open HRMA shells in cm^2, set by
 * marx_mirror_init (hrma.c:766-770). */
extern double Marx_Mirror_Geometric_Area;

/* ------------------------------------------------------------------------ */
/* Source side (marx/libsrc/marx.h:287-293; source.c:268-417).  POINT and
 * DISK sources with FLAT or FILE spectra and the INTERNAL / NONE dither
 * models.  Photons are created on the GPU and copied into pt.              */

typedef struct _Marx_Source_Type Marx_Source_Type;

Marx_Source_Type *marx_create_source (Param_File_Type *pf);
int marx_open_source (Marx_Source_Type *st);
int marx_close_source (Marx_Source_Type *st);
int marx_create_photons (Marx_Source_Type *st, Marx_Photon_Type *pt,
                         unsigned int num, unsigned int *num_collected,
                         double *exposure_time);

/* ------------------------------------------------------------------------ */
/* Output files (marx/libsrc/marxio.c, declared at marx.h:161-186, 451-479;
 * format: marx/doc/dump.txt).  One file per column: a 32-byte header (magic
 * 0x83 0x13 0x89 0x8D, type 'A' int8 / 'I' int16 / 'J' int32 / 'E' float32
 * / 'D' float64, the column name in 15 bytes padded with 0, the number of
 * rows and of columns (0: a vector) as big-endian int32, 4 reserved bytes)
 * followed by the values, big-endian.                                      */

#define MARX_MAX_COLUMN_NAME_LEN 16
typedef struct Marx_Dump_File_Type
{
   FILE *fp;
   char type;                          /* A, I, J, E, D */
   char colname[MARX_MAX_COLUMN_NAME_LEN];
   int num_rows;
   int num_cols;                       /* 0 ==> vector */
   struct Marx_Dump_File_Type *next;
}
Marx_Dump_File_Type;

/* Replace marxio.c:466-520 and :523-533: open a column file and read its
 * header (NULL on error), close it. */
Marx_Dump_File_Type *marx_open_read_dump_file (char *file);
void marx_close_read_dump_file (Marx_Dump_File_Type *dft);
/* Replace marxio.c:143-200 and :79-115: create a column file and write the
 * header of dft (type, colname, num_rows, num_cols; dft->fp is set), and
 * write the row count n into the header of such a file and close it. */
FILE *marx_create_write_dump_file (char *filename, Marx_Dump_File_Type *dft);
int marx_close_write_dump_file (FILE *fp, unsigned long n);

/* Replaces marxio.c:402-453 (marx.c:157-170 calls it once per chunk).
 * Writes the photons of p whose flags & BAD_PHOTON_MASK is zero, in list
 * order, into the column files of dir whose history bits are in
 * write_mask & p->history (plus the sky and detector dither columns when
 * dither is on), as in marxio.c:259-289: b_energy energy time tag
 * xpos ypos zpos xcos ycos zcos pha detector xpixel ypixel hrc_u hrc_v
 * mirror hrcregion order ofine ocoarse1-3 sky_ra sky_dec sky_roll det_dy
 * det_dz det_theta (.dat).  time = arrival_time + total_time.  open_mode
 * != 0 creates the files; 0 appends to them and updates their row counts.
 * The directory must exist.  Returns 0 or -1. */
int marx_write_photons (Marx_Photon_Type *p, unsigned long write_mask, char *dir,
                        int open_mode, double total_time);

/* Replaces marxio.c:697-756: prints the column files named in argv[2..]
 * side by side as text on stdout ("marx --dump file..."), flushed before
 * returning. */
int marx_dump (int argc, char **argv);

/* Ray files (marx/libsrc/s-rayfile.c:43-145): the magic 0xF8161961 and the
 * list's history as native unsigned longs, then the photons whose flags &
 * BAD_PHOTON_MASK is zero as native Marx_Photon_Attr_Type records.  As in
 * the reference, writing adds total_time to those photons' arrival_time in
 * p.  new_file != 0 creates the file, 0 appends.  marx_dump_rayfile prints
 * a ray file as text on stdout. */
#define MARX_RAYFILE_MAGIC 0xF8161961UL
int marx_dump_to_rayfile (char *file, int new_file, Marx_Photon_Type *p, double total_time);
int marx_dump_rayfile (char *file);

/* ------------------------------------------------------------------------ */
/* mxg_ extensions (no counterpart in the reference).                       */

/* Last error message of this library (also printed on stderr). */
const char *mxg_last_error (void);

/* Selects the HIP device, creates the library's stream and checks that the
 * device is a gfx950.  Returns 0, or -1 with mxg_last_error() set when no
 * usable GPU is present.  Every entry point that launches work calls it. */
int mxg_init (void);

/* Random-number seed of the per-photon counter streams (Philox4x32-10 keyed
 * by seed; counter = (photon tag, stage, block, source id)).  Until this is
 * called, marx_mirror_init takes the seed from the RandomSeed parameter as
 * the reference's driver does (marx/src/marx.c:846-849; -1 seeds from the
 * clock); once called, RandomSeed is ignored. */
void mxg_set_seed (unsigned long long seed);
unsigned long long mxg_get_seed (void);

/* When nonzero, marx_mirror_reflect and marx_grating_diffract leave the
 * photon state resident in HBM and only marx_detect copies it back into
 * the AoS list; when zero (the default) every stage leaves pt->attributes
 * valid on return, like the reference. */
void mxg_set_lazy_sync (int on);

/* Arrival-time model of the initialised source: tmax > 0 draws times
 * uniformly over [0, tmax); tmax <= 0 restores the reference's Poisson
 * arrivals at the source's flux (source.c:313-342). */
void mxg_set_uniform_times (double tmax);

/* Device-resident pipeline: creates num photons of the initialised source
 * on the GPU and runs dither, mirror, grating and detector stages without
 * leaving HBM.  tag_start / start_time play the role of the fields of
 * Marx_Photon_Type.  *num_collected receives the number of photons created
 * (fewer than num when exposure_time truncates the chunk), *num_detected
 * the number that survive every stage.  stage_ms, when non-NULL, receives
 * 6 per-stage GPU times in ms measured with HIP events recorded on the
 * library's stream around each stage's kernels (read after the call's
 * final synchronisation, so timing adds no host round trip): source
 * (including dither), 0 (dither is fused into the source stage), mirror,
 * grating, detector, and their sum. */
int mxg_trace_device (unsigned int num, unsigned int tag_start,
                      double start_time, double exposure_time,
                      unsigned int *num_collected, unsigned int *num_detected,
                      float *stage_ms);

/* Photons alive after each stage of the last mxg_trace_device or
 * mxg_trace_batch call: created, after the mirror, after the grating, after
 * the detector. */
int mxg_last_stage_counts (unsigned int counts[4]);

/* Device time in ms of the stages of the last timed mxg_trace_device or
 * uncaptured mxg_trace_batch pass: source (with dither), mirror, grating,
 * detector (0 for a stage not run or a graph-launched pass). */
int mxg_last_stage_ms (float ms[4]);

/* ---- differentiable mirror (no counterpart in the reference) ---- */

/* A focal-plane image: npix x npix pixels of pixel_size mm centred on
 * (center_y, center_z) mm at x = 0 (the HRMA focus). */
typedef struct
{
   unsigned int npix;
   double pixel_size;
   double center_y, center_z;
}
mxg_image_spec;

/* Fills *d with the initialised mirror's shell geometry at misalignment
 * params[20] (MXG_DIFF_NPAR per shell pair, see mxg_diff_cfg) and its
 * derivatives.  Host only. */
int mxg_diff_setup (const double *params, void *d /* mxg_diff_cfg */);

/* Weighted-photon forward pass: num photons (tags tag_start...) of the
 * initialised source (its energies and direction) go through the mirror at
 * misalignment params[20], with survival weights in place of the vignetting
 * and reflectivity draws (vignetting x reflectivity x scale factor per
 * surface; a photon that misses a surface, hits a strut or scatters past
 * pi/4 gets weight 0), and are projected to the focal plane.  Each photon's
 * weight is spread bilinearly over the four pixels around it and summed in
 * FP32 into image (a device pointer of spec->npix^2 floats, overwritten).
 * The photons' weights, positions and their derivatives with respect to
 * their shell pair's parameters stay in HBM for mxg_diff_adjoint.
 * *total_weight, when non-NULL, receives the summed weight. */
int mxg_diff_forward (const double *params, unsigned int num, unsigned int tag_start,
                      const mxg_image_spec *spec, float *image, double *total_weight);

/* The same pass without the derivatives: the same image, weights and
 * positions (mxg_diff_state's w, y, z and shell) at a fraction of the cost,
 * for evaluations that need no gradient (targets, line searches, finite
 * differences).  mxg_diff_adjoint refuses to follow it. */
int mxg_diff_image (const double *params, unsigned int num, unsigned int tag_start,
                    const mxg_image_spec *spec, float *image, double *total_weight);

/* Id of the last forward pass (mxg_diff_forward or mxg_diff_image) made in
 * this process by any caller; 0 before the first.  The adjoint needs it. */
unsigned long long mxg_diff_last_pass (void);

/* Adjoint of forward pass `pass` (the value mxg_diff_last_pass gave right
 * after it): grad[20] = d(sum_pix grad_image * image)/d(params), i.e. the
 * gradient of a loss whose derivative with respect to the image is
 * grad_image (device pointer, npix^2 floats).  The device keeps the state
 * of one forward pass: -1 when another pass ran since (its state would give
 * another pass's gradient), when npix is not that pass's image side, or
 * when that pass kept no derivatives.  Sums per-block partials in a fixed
 * order (deterministic). */
int mxg_diff_adjoint (unsigned long long pass, unsigned int npix, const float *grad_image,
                      double *grad);

/* ms[0..2]: device time in ms of the last mxg_diff_forward's or
 * mxg_diff_image's main kernel (both reflections, focal plane, image), of
 * the last mxg_diff_adjoint's kernel, and of the last forward pass's first
 * kernel (every photon to its paraboloid intersection) with the list
 * compaction after it; HIP events on the library's stream around each
 * launch, 0 before the first.  *listed (may be NULL): the photons of the
 * last forward pass that reached the paraboloid, which the main kernel ran
 * on. */
int mxg_diff_kernel_ms (float *ms, unsigned int *listed);

/* Host copies of the per-photon state of the last mxg_diff_forward (n =
 * its photon count): weight w, focal-plane position y, z, and their
 * derivatives dw, dy, dz (MXG_DIFF_NPAR x n, parameter-major) by the
 * photon's shell-pair parameters; shell (0..3) of each photon.  Any
 * pointer may be NULL. */
int mxg_diff_state (unsigned int n, double *w, double *y, double *z, double *dw, double *dy,
                    double *dz, unsigned char *shell);

/* ---- batched point sources (no counterpart in the reference, which runs
 * one source per process) ---- */

/* Traces num_sources point sources in one device pass through the
 * initialised mirror, grating and detector.  Source k emits counts[k]
 * photons travelling along the unit vector dirs[3k..3k+2] (MARX frame; an
 * on-axis source is (-1, 0, 0)), with energies from the spectrum of the
 * source set up by marx_create_source (FLAT or FILE), arrival times
 * uniform over [0, exposure) after start_time, and the configured dither.
 * Photon i of the batch (sources in order) gets tag tag_start + i, and its
 * random streams depend on the seed and its tag alone: tracing source k by
 * itself with tag_start + (counts[0] + ... + counts[k-1]) gives the same
 * photons.  The photons stay resident (mxg_download_photons, mxg_get_soa);
 * a photon's source follows from its tag.  With use_graph != 0 the launch
 * sequence is captured into a HIP graph on first use and replayed while the
 * number of sources and photons, the configuration, seed, exposure,
 * start_time and schedule stay the same (counts, directions and tag_start
 * may change).  *num_detected receives the photons that survive every
 * stage; ms, when non-NULL, the device time of the whole pass.  Returns 0,
 * or -1 with mxg_last_error() set. */
int mxg_trace_batch (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                     unsigned int tag_start, double start_time, double exposure, int use_graph,
                     unsigned int *num_detected, float *ms);

/* Launch schedule of the batched pass.  Staged (persistent_tile = 0): one
 * kernel per stage, with blocks per CU of the source, paraboloid,
 * reflection, grating and detector kernels (1..32; for the grating also 0,
 * the default: one round of the blocks a CU holds at once, which its LDS
 * tables and register budget set) and the minimum waves per
 * SIMD the reflection kernel is compiled for (2, 3, 4).  Persistent
 * (persistent_tile = a multiple of 256 up to 4096): the whole pass as one
 * launch of persistent_blocks_per_cu (1..8) blocks per CU that take tiles of
 * persistent_tile photons from a work counter and run every stage on them
 * (a FEF run takes the staged pass).  Both give the same photons. */
typedef struct
{
   unsigned int blocks_per_cu[5];
   unsigned int reflect_waves;
   unsigned int persistent_tile;
   unsigned int persistent_blocks_per_cu;
}
mxg_schedule;
int mxg_set_schedule (const mxg_schedule *s);
int mxg_get_schedule (mxg_schedule *s);

/* Autotuner: times the batch of mxg_trace_batch (same arguments) under
 * candidate schedules -- for the staged pass, blocks per CU in {2, 4, 8, 16}
 * for every stage kernel and one round (0) for the grating, then 2, 3 and 4
 * waves for the reflection kernel,
 * keeping for each kernel the fastest (HIP-event time, best of `repeats`
 * runs); then the persistent pass with tiles of {512, 1024, 2048, 4096}
 * photons at {2, 3, 4} blocks per CU -- and installs the schedule whose
 * graph-launched pass is fastest, returning it in *best and that pass time
 * in *best_ms. */
int mxg_autotune_batch (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                        unsigned int tag_start, double start_time, double exposure,
                        unsigned int repeats, mxg_schedule *best, float *best_ms);

/* The two mirror kernels of the last mxg_trace_device call, each timed
 * alone with HIP events on the library's stream: [0] aperture + paraboloid
 * intersection, [1] the two reflections.  ms receives their device times in
 * milliseconds, photons_in the photons each one processed (the photons
 * entering the mirror, and those that hit the paraboloid).  Either pointer
 * may be NULL. */
int mxg_last_kernel_stats (float ms[2], unsigned int photons_in[2]);

/* Copies the device-resident photons of the last mxg_trace_device call into
 * pt (AoS), setting n_photons, history and the pruned sorted_index. */
int mxg_download_photons (Marx_Photon_Type *pt);

/* Event coordinates (marx/src/marx2fits.c:3584-3735, 3813-3910 over
 * marx/libsrc/pixlib.c, detpix.c:41-209, acis_geom.c:37-230,
 * hrc_s_geom.c:321-510, hrc_i_geom.c:130-236, acis_subpix.c:270-322,
 * dither.c:583-607) of the detected photons of an ACIS or HRC simulation:
 * tiled (TDETX/Y), focal-plane (DETX/Y) and sky (X/Y) pixels and, for ACIS,
 * the flight grade, from the values marx2fits reads (the float chip pixels
 * plus one, the float dither state, the PI energy).  A pixel is placed
 * within its chip pixel as chosen by pix_adj: EXACT (the simulated
 * position), NONE (the centre), RANDOMIZE (uniform), EDSER (ACIS: the
 * centre plus the CALDB subpixel offset of the event's energy and flight
 * grade; HRC: RANDOMIZE, as marx2fits.c:3307-3308).  The flight grade is
 * one of the four of the event's 3x3 subpixel (marx2fits.c:3822-3866).
 * For ACIS, except in EXACT mode, the dither state that DET and sky pixels
 * use is held over a frame, as marx2fits re-reads it only when the
 * exposure number (long)(time / TimeDel) changes from one row to the next
 * (marx2fits.c:3571-3578, 3741-3760), TimeDel being the ACIS exposure time
 * plus the frame transfer time; HRC rows use their own (:2850-2851).  The
 * random choices come from per-event streams (seed, tag), not from a
 * serial generator. */
#define MXG_PIX_ADJ_NONE       0
#define MXG_PIX_ADJ_RANDOMIZE  1
#define MXG_PIX_ADJ_EDSER      2
#define MXG_PIX_ADJ_EXACT      3
typedef struct
{
   double focal_length;        /* mm (marx.par FocalLength) */
   double det_offset[3];       /* DetOffsetX/Y/Z, mm */
   double roll_nom;            /* degrees (Roll_Nom), for undithered simulations */
   double exposure_time;       /* ACIS_Exposure_Time, s */
   double frame_transfer_time; /* ACIS_Frame_Transfer_Time, s */
   double total_time;          /* added to the photons' times, as by mxg_write_device */
   unsigned long long seed;    /* of the per-event streams */
   int pix_adj;                /* MXG_PIX_ADJ_* */
   int dither_used;            /* the simulation dithered: undither for the sky pixels */
}
mxg_event_cfg;

typedef struct
{
   /* host arrays of capacity rows each; any may be NULL */
   unsigned int *tag;
   int *ccd_id;
   int *tdetx, *tdety;
   double *detx, *dety;
   double *x, *y;              /* sky */
   short *fltgrade;
}
mxg_event_columns;

/* The event coordinates of the device-resident photons of the last
 * mxg_trace_device / mxg_trace_batch call whose flags & BAD_PHOTON_MASK is
 * zero (the rows of mxg_write_device, in the same order), computed on the
 * GPU.  The detector must be ACIS or HRC.  *num_rows receives the rows;
 * -1 with the error set when they exceed capacity. */
int mxg_event_coords (const mxg_event_cfg *cfg, mxg_event_columns *out, unsigned int capacity,
                      unsigned int *num_rows);

/* GPU time (ms) of the kernels of the last mxg_event_coords call (frame
 * heads and coordinates; no keep list, no copies). */
int mxg_event_kernel_ms (float *ms);

/* marx_write_photons for the device-resident photons of the last
 * mxg_trace_device or mxg_trace_batch call, without the AoS copy: the GPU
 * selects the photons whose flags & BAD_PHOTON_MASK is zero (in tag order),
 * packs each requested column big-endian in HBM, and only the file bytes
 * cross to the host.  The files are byte-identical to marx_write_photons on
 * the downloaded list with the same arguments.  *num_written (may be NULL)
 * receives the rows written.  Returns 0 or -1. */
int mxg_write_device (unsigned long write_mask, char *dir, int open_mode, double total_time,
                      unsigned int *num_written);

/* Device pointers of the resident structure-of-arrays columns.  Returned
 * pointers stay valid until the next call that grows the store. */
typedef struct
{
   unsigned int n;             /* photons in the store */
   unsigned int capacity;
   double *energy;             /* [n] */
   double *x[3];               /* [n] each */
   double *p[3];
   double *arrival_time;
   unsigned int *flags;
   float *y_pixel, *z_pixel, *u_pixel, *v_pixel;
   float *dither[6];           /* ra dec roll dy dz dtheta */
   float *pi;
   short *pulse_height;
   unsigned char *mirror_shell;
   signed char *ccd_num, *detector_region, *order;
   signed char *support_orders[4];
   unsigned int *tag;
}
mxg_soa_view;
int mxg_get_soa (mxg_soa_view *view);

/* No counterpart in the reference: mxg_trace_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The photons of the chunk go to
 * the caller's columns *store (device pointers, capacity >= num; n is
 * ignored); every list, counter and scan buffer lives in a fixed device
 * array (29 MB) that the library holds on each GPU it runs on until the
 * process ends.  The configuration is the library's device copy: at least
 * one allocating call (mxg_trace_device, marx_mirror_reflect, ...) must
 * have run since the last init, so that its device buffers exist and are
 * large enough; the entry point then copies a changed configuration into
 * them, else it fails with a message.  Limits: num <=
 * MXG_NOALLOC_MAX_PHOTONS (marx.par bounds dNumRays, the chunk size, by
 * 1000000); a larger chunk fails with a message.  mxg_init must have run.
 * The entry point calls no allocator; it synchronises the library's stream
 * before returning.  All calls on one GPU share that array: calls that
 * overlap on one GPU (other threads) race.  Same photons and records as
 * mxg_trace_device with the same arguments. */
#define MXG_NOALLOC_MAX_PHOTONS 1048576u
int mxg_trace_device_noalloc (const mxg_soa_view *store, unsigned int num, unsigned int tag_start,
                              double start_time, double exposure_time, unsigned int *num_collected,
                              unsigned int *num_detected);

/* Name of the device the library runs on, and its compute-unit count. */
int mxg_device_info (char *name, unsigned int namelen, int *num_cus);

#ifdef __cplusplus
}
#endif

#endif /* MARX_GPU_H */
