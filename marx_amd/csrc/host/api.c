/* api.c -- the drop-in entry points of libmarx_gpu.
 *
 * marx_mirror_* / marx_grating_* / marx_detect / photon-list management keep
 * the reference's signatures and conventions (marx/libsrc/marx.h:155-157,
 * 354-362; mirror.c:48-108; grating.c:44-114; detector.c:298-379;
 * photon.c:40-108).  The photon state of a chunk lives in HBM as columns;
 * the AoS list the caller owns is filled from it.  By default every stage
 * call leaves pt->attributes valid on return, as the reference does; with
 * mxg_set_lazy_sync(1) the state stays resident between the mirror, grating
 * and detector calls of one chunk and only marx_detect copies it back.
 */
#ifndef _GNU_SOURCE
# define _GNU_SOURCE   /* qsort_r */
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <time.h>

#include "mxg_host.h"

static int Mirror_Type = -1;
static int Grating_Type = -1;
static int Detector_Type = -1;
static unsigned long long Seed = 1;
static int Seed_Set = 0;             /* mxg_set_seed called: RandomSeed is ignored */
static int Lazy_Sync = 0;

/* the list whose photons are currently resident in HBM */
static const Marx_Photon_Type *Resident = NULL;
/* the device store holds the photons of the last mxg_trace_device /
 * mxg_trace_batch call (cleared when an AoS list is uploaded over them) */
static int Device_Trace = 0;
static unsigned int Resident_Tag_Start = 0;

void mxg_set_seed (unsigned long long seed) { Seed = seed; Seed_Set = 1; }
unsigned long long mxg_get_seed (void) { return Seed; }
void mxg_set_lazy_sync (int on) { Lazy_Sync = on; }

int mxg_init (void)
{
   if (-1 == mxgd_init ())
     {
        mxg_error ("GPU initialisation failed: %s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

int mxg_device_info (char *name, unsigned int namelen, int *num_cus)
{
   return mxgd_device_info (name, namelen, num_cus);
}

static int sync_config (void)
{
   if (-1 == mxg_init ())
     return -1;
   if (-1 == mxgd_sync_config (&Mxg_Cfg, mxg_farena_base (), mxg_farena_size (),
                               mxg_darena_base (), mxg_darena_size (), Mxg_Cfg_Generation))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

/* ---- photon lists (photon.c) ------------------------------------------------ */

/* Zeroed host memory for a photon list's arrays: pinned when a GPU is
 * there, so that the stages' copies run at the link's DMA rate straight
 * into the caller's records, else plain. */
static void *list_alloc (size_t n, size_t size)
{
   void *p = mxgd_host_alloc (n * size);
   return (p != NULL) ? p : calloc (n, size);
}

static void list_free (void *p)
{
   if ((p != NULL) && !mxgd_host_free (p))
     free (p);
}

Marx_Photon_Type *marx_alloc_photon_type (unsigned int num)
{
   Marx_Photon_Type *pt = (Marx_Photon_Type *) calloc (1, sizeof (Marx_Photon_Type));
   if (pt == NULL) return NULL;
   pt->attributes = (Marx_Photon_Attr_Type *) list_alloc (num ? num : 1, sizeof (Marx_Photon_Attr_Type));
   pt->sorted_energies = (double *) list_alloc (num ? num : 1, sizeof (double));
   pt->sorted_index = (unsigned int *) list_alloc (num ? num : 1, sizeof (unsigned int));
   if ((pt->attributes == NULL) || (pt->sorted_energies == NULL) || (pt->sorted_index == NULL))
     {
        marx_dealloc_photon_type (pt);
        return NULL;
     }
   pt->max_n_photons = num;
   return pt;
}

int marx_dealloc_photon_type (Marx_Photon_Type *pt)
{
   if (pt == NULL) return -1;
   if (Resident == pt) Resident = NULL;
   list_free (pt->attributes);
   list_free (pt->sorted_index);
   list_free (pt->sorted_energies);
   free (pt);
   return 0;
}

void marx_prune_photons (Marx_Photon_Type *pt)
{
   unsigned int i, n = 0;
   if ((pt == NULL) || (pt->sorted_index == NULL)) return;
   for (i = 0; i < pt->num_sorted; i++)
     {
        unsigned int indx = pt->sorted_index[i];
        Marx_Photon_Attr_Type *at = pt->attributes + indx;
        if (at->flags & BAD_PHOTON_MASK) continue;
        if (pt->sorted_energies != NULL) pt->sorted_energies[n] = at->energy;
        pt->sorted_index[n] = indx;
        n++;
     }
   pt->num_sorted = n;
}

/* ---- residency ---------------------------------------------------------------- */

static int make_resident (Marx_Photon_Type *pt)
{
   if ((Resident == pt) && (Resident_Tag_Start == pt->tag_start) && Lazy_Sync)
     return 0;
   if (pt->num_sorted > pt->n_photons)
     {
        mxg_error ("photon list has num_sorted > n_photons");
        return -1;
     }
   Device_Trace = 0;
   if (-1 == mxgd_upload_aos (pt->attributes, pt->n_photons, pt->sorted_index, pt->num_sorted))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   Resident = pt;
   Resident_Tag_Start = pt->tag_start;
   return 0;
}

/* The records back into the caller's list, and the pruned sorted_index:
 * the last stage's output list on the device is the input list (the
 * sorted_index uploaded) less the photons it absorbed, in the same order,
 * which is what marx_prune_photons would rebuild (photon.c:40-63). */
static int copy_back (Marx_Photon_Type *pt)
{
   unsigned int n = 0;
   if ((-1 == mxgd_download_aos (pt->attributes, pt->n_photons))
       || (-1 == mxgd_download_active (pt->sorted_index, pt->sorted_energies, pt->max_n_photons, &n)))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   pt->num_sorted = n;
   return 0;
}

static int finish_stage (Marx_Photon_Type *pt, int force_copy)
{
   if (Lazy_Sync && !force_copy)
     return 0;
   return copy_back (pt);
}

/* ---- mirror (mirror.c:48-108) ------------------------------------------------- */

int marx_mirror_init (Param_File_Type *pf)
{
   char buf[256];
   if (-1 == pf_get_string (pf, "MirrorType", buf, sizeof (buf)))
     return -1;
   if (strcmp (buf, "HRMA"))
     {
        mxg_error ("MirrorType %s not supported.", buf);
        Mirror_Type = -1;
        return -1;
     }
   if (-1 == mxg_hrma_init (pf))
     {
        mxg_error ("Error initializing MirrorType %s.", buf);
        Mirror_Type = -1;
        return -1;
     }
   /* The seed of the driver's RandomSeed parameter (marx/src/marx.c:846-849:
    * -1 seeds from the clock), unless the caller set one with mxg_set_seed. */
   if (Seed_Set == 0)
     {
        int rs;
        if (0 == pf_get_integer (pf, "RandomSeed", &rs))
          Seed = (rs == -1) ? (unsigned long long) time (NULL) : (unsigned long long) (unsigned int) rs;
     }
   Mirror_Type = MARX_MIRROR_HRMA;
   return Mirror_Type;
}

int marx_mirror_reflect (Marx_Photon_Type *pt, int verbose)
{
   if (Mirror_Type != MARX_MIRROR_HRMA)
     {
        mxg_error ("Mirror not initialized.");
        return -1;
     }
   if (verbose > 0) mxg_message (0, "Reflecting from HRMA\n");
   if (pt->history & MARX_MIRROR_SHELL_OK)
     return 0;
   pt->history |= MARX_MIRROR_SHELL_OK;
   if ((-1 == sync_config ()) || (-1 == make_resident (pt)))
     return -1;
   if (-1 == mxgd_run_mirror (pt->source_distance, Seed, -1, NULL, 0, 0))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return finish_stage (pt, 0);
}

/* ---- grating (grating.c:44-114) ------------------------------------------------ */

/* The history bits a grating pass sets: MARX_ORDER_OK, and for LETG the
 * support-grating order columns of the gratings that have orders
 * (diffract.c:1098-1117: the fine grating ORDER1, the coarse one ORDER2-4). */
static unsigned long grating_history (void)
{
   unsigned long hist = MARX_ORDER_OK;
   if (Grating_Type != MARX_GRATING_LETG)
     return hist;
   if (Mxg_Cfg.grating.has_fine)
     hist |= MARX_ORDER1_OK;
   if (Mxg_Cfg.grating.has_coarse)
     hist |= (MARX_ORDER2_OK | MARX_ORDER3_OK | MARX_ORDER4_OK);
   return hist;
}

int marx_grating_init (Param_File_Type *pf)
{
   char buf[256];
   int type;
   if (-1 == pf_get_string (pf, "GratingType", buf, sizeof (buf)))
     return -1;
   if (!strcmp (buf, "NONE"))
     {
        memset (&Mxg_Cfg.grating, 0, sizeof (Mxg_Cfg.grating));
        Mxg_Cfg_Generation++;
        Grating_Type = 0;
        return 0;
     }
   if (strcmp (buf, "HETG") && strcmp (buf, "LETG"))
     {
        mxg_error ("GratingType %s not supported.", buf);
        Grating_Type = -1;
        return -1;
     }
   type = mxg_grating_init (pf);
   if (type == -1)
     mxg_error ("Error initializing GratingType %s.", buf);
   Grating_Type = type;
   return type;
}

int marx_grating_diffract (Marx_Photon_Type *pt, int verbose)
{
   if (Grating_Type == 0)
     return 0;
   if ((Grating_Type != MARX_GRATING_HETG) && (Grating_Type != MARX_GRATING_LETG))
     {
        mxg_error ("Grating not initialized.");
        return -1;
     }
   if (verbose > 0)
     mxg_message (0, "Diffracting from %s.\n", (Grating_Type == MARX_GRATING_LETG) ? "LETG" : "HETG");
   if (pt->history & MARX_ORDER_OK)
     return 0;
   pt->history |= grating_history ();
   if ((-1 == sync_config ()) || (-1 == make_resident (pt)))
     return -1;
   if (-1 == mxgd_run_grating (Seed, -1, NULL, 0, 0))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return finish_stage (pt, 0);
}

/* ---- detector (detector.c:298-379) ------------------------------------------- */

int marx_detector_init (Param_File_Type *pf)
{
   int type = mxg_detector_init (pf);
   Detector_Type = type;
   return type;
}

/* the columns a detector fills (acis-s.c:189-190, acis-i.c:108-109,
 * hrc-s.c:245-248, hrc-i.c:128-129) */
static unsigned long detector_history (void)
{
   if ((Detector_Type == MARX_DETECTOR_ACIS_S) || (Detector_Type == MARX_DETECTOR_ACIS_I))
     return MARX_DET_PIXEL_OK | MARX_DET_NUM_OK | MARX_PULSEHEIGHT_OK | MARX_PI_OK;
   if (Detector_Type == MARX_DETECTOR_HRC_S)
     return MARX_DET_REGION_OK | MARX_PULSEHEIGHT_OK | MARX_DET_PIXEL_OK | MARX_DET_NUM_OK
       | MARX_DET_UV_PIXEL_OK;
   return MARX_DET_REGION_OK | MARX_PULSEHEIGHT_OK | MARX_DET_PIXEL_OK | MARX_DET_NUM_OK;
}

int marx_detect (Marx_Photon_Type *pt, int verbose)
{
   if (Detector_Type == -1)
     {
        mxg_error ("Detector not initialized.");
        return -1;
     }
   if (Detector_Type == 0)
     {
        /* NONE: nothing to detect, but a lazily synchronised chunk ends here */
        if (Lazy_Sync && (Resident == pt))
          return copy_back (pt);
        return 0;
     }
   if (verbose > 0) mxg_message (0, "Detecting\n");
   if (pt->history & MARX_DET_NUM_OK)
     return 0;
   pt->history |= detector_history ();
   if ((-1 == sync_config ()) || (-1 == make_resident (pt)))
     return -1;
   if (-1 == mxgd_run_detect (pt->start_time, Seed, -1, NULL, 0))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return finish_stage (pt, 1);
}

/* ---- source (source.c:268-384) ------------------------------------------------- */

/* the energies travel with the sort (qsort_r): no state shared between
 * concurrent sorts */
static int energy_cmp (const void *a, const void *b, void *energies)
{
   const double *en = (const double *) energies;
   unsigned int ia = *(const unsigned int *) a, ib = *(const unsigned int *) b;
   double x = en[ia], y = en[ib];
   if (x > y) return 1;
   if (x < y) return -1;
   return (ia > ib) - (ia < ib);
}

/* sorted_index / sorted_energies in energy order with index tie-break
 * (source.c:361-371, dsort.c:28-61) */
static int sort_by_energy (Marx_Photon_Type *pt)
{
   unsigned int n = pt->n_photons, i;
   double *e = (double *) malloc ((n ? n : 1) * sizeof (double));
   if (e == NULL) return -1;
   for (i = 0; i < n; i++)
     {
        e[i] = pt->attributes[i].energy;
        pt->sorted_index[i] = i;
     }
   qsort_r (pt->sorted_index, n, sizeof (unsigned int), energy_cmp, e);
   for (i = 0; i < n; i++)
     pt->sorted_energies[i] = e[pt->sorted_index[i]];
   pt->num_sorted = n;
   free (e);
   return 0;
}

/* A RAYFILE source (source.c:268-382 over s-rayfile.c:176-207): the records
 * are read on the host, cut at the exposure, given new tags, dithered again
 * on the device (the reference applies the dither to replayed photons as to
 * new ones) and sorted by energy. */
static int replay_rayfile (Marx_Source_Type *st, Marx_Photon_Type *pt, unsigned int num,
                           unsigned int *num_collected, double *exposure_time)
{
   unsigned int n = 0, i;
   unsigned long hist = 0;

   memset (pt->attributes, 0, (size_t) num * sizeof (Marx_Photon_Attr_Type));
   if (-1 == mxg_rayfile_read (st, pt->attributes, num, &n, &hist))
     {
        mxg_error ("marx_create_photons: the source is not an open ray file");
        return -1;
     }
   pt->history = hist;
   if (0 == (hist & MARX_TIME_OK))
     {
        mxg_error ("ray file without arrival times (history lacks MARX_TIME_OK)");
        return -1;
     }
   if (exposure_time != NULL)
     {
        double tmax = _Marx_TStart_MJDsecs + *exposure_time;
        for (i = 0; i < n; i++)
          if (pt->attributes[i].arrival_time >= tmax)
            {
               n = i + 1;
               break;
            }
     }
   if (n)
     {
        for (i = 0; i < n; i++) pt->attributes[i].tag = pt->tag_start + i;
        pt->history |= MARX_TAG_OK;
        pt->tag_start += n;
        if (Mxg_Cfg.src.dither_mode != MXG_DITHER_NONE)
          {
             if (-1 == sync_config ())
               return -1;
             Device_Trace = 0;
             if ((-1 == mxgd_upload_aos (pt->attributes, n, NULL, 0))
                 || (-1 == mxgd_dither_store (pt->start_time, Seed))
                 || (-1 == mxgd_download_aos (pt->attributes, n)))
               {
                  mxg_error ("%s", mxgd_last_error ());
                  return -1;
               }
          }
     }
   pt->n_photons = n;
   Resident = NULL;
   if (-1 == sort_by_energy (pt))
     return -1;
   pt->total_time = n ? pt->attributes[n - 1].arrival_time : 0.0;
   *num_collected = n;
   return 0;
}

/* A USER source (source.c:268-384 over s-user.c:140-205): the device draws
 * each photon's energy from the spectrum (and Poisson arrival times) in its
 * own stream, with the dither held back; the plug-in's user_create_ray then
 * gives the direction, and may replace the energy and give time steps,
 * photon by photon on the host until it returns -1.  The exposure cut,
 * tags, dither (on the device) and the energy sort follow as for any
 * source. */
static int create_user (Marx_Photon_Type *pt, unsigned int num, unsigned int *num_collected,
                        double *exposure_time)
{
   mxg_source_cfg *s = &Mxg_Cfg.src;
   Marx_Photon_Attr_Type *at = pt->attributes;
   unsigned int n = 0, i;
   double last_t = 0.0, t = -1.0, user_time = 0.0;
   int dither_mode = s->dither_mode, status;

   memset (at, 0, (size_t) num * sizeof (Marx_Photon_Attr_Type));
   if (-1 == mxg_user_start_iteration ())
     {
        pt->n_photons = 0;
        pt->total_time = 0.0;
        return sort_by_energy (pt);
     }
   if (num > 0)
     {
        s->dither_mode = MXG_DITHER_NONE;
        Mxg_Cfg_Generation++;
        status = sync_config ();
        Device_Trace = 0;
        if (status == 0)
          status = mxgd_run_source (num, pt->tag_start, pt->start_time, -1.0, Seed, 0, -1, &n, &last_t, NULL);
        s->dither_mode = dither_mode;
        Mxg_Cfg_Generation++;
        if ((status == -1) || ((n > 0) && (-1 == mxgd_download_aos (at, n))))
          {
             mxg_error ("%s", mxgd_last_error ());
             return -1;
          }
     }
   for (i = 0; i < n; i++)
     {
        double energy = at[i].energy, drawn = energy, px, py, pz;
        if (-1 == mxg_user_create_ray (&t, &energy, &px, &py, &pz))
          break;
        at[i].flags = 0;
        at[i].p.x = px;
        at[i].p.y = py;
        at[i].p.z = pz;
        if (t >= 0.0)
          {
             user_time += t;
             at[i].arrival_time = user_time;
          }
        if (energy < 0.0) energy = drawn;
        at[i].energy = energy;
     }
   n = i;
   if (t >= 0.0)
     {
        /* the plug-in's times: cut at the exposure from the start time */
        if (exposure_time != NULL)
          {
             double tmax = _Marx_TStart_MJDsecs + *exposure_time;
             for (i = 0; i < n; i++)
               if (at[i].arrival_time >= tmax)
                 {
                    n = i + 1;
                    break;
                 }
          }
     }
   else if (exposure_time != NULL)
     {
        /* Poisson times: the first photon at or past the exposure is the last */
        for (i = 0; i < n; i++)
          if (*exposure_time <= at[i].arrival_time)
            {
               n = i + 1;
               break;
            }
     }
   pt->history = MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK;
   if (n)
     {
        for (i = 0; i < n; i++) at[i].tag = pt->tag_start + i;
        pt->history |= MARX_TAG_OK;
        pt->tag_start += n;
        if (dither_mode != MXG_DITHER_NONE)
          {
             if ((-1 == sync_config ())
                 || (-1 == mxgd_upload_aos (at, n, NULL, 0))
                 || (-1 == mxgd_dither_store (pt->start_time, Seed))
                 || (-1 == mxgd_download_aos (at, n)))
               {
                  mxg_error ("%s", mxgd_last_error ());
                  return -1;
               }
          }
     }
   pt->n_photons = n;
   Resident = NULL;
   if (-1 == sort_by_energy (pt))
     return -1;
   pt->total_time = n ? at[n - 1].arrival_time : 0.0;
   *num_collected = n;
   return 0;
}

int marx_create_photons (Marx_Source_Type *st, Marx_Photon_Type *pt, unsigned int num,
                         unsigned int *num_collected, double *exposure_time)
{
   const mxg_source_cfg *s = &Mxg_Cfg.src;
   unsigned int n = 0;
   double last_t = 0.0;

   *num_collected = 0;
   if ((st == NULL) || (pt == NULL))
     return -1;
   if (num > pt->max_n_photons)
     {
        mxg_error ("marx_create_photons: %u photons requested, list holds %u", num, pt->max_n_photons);
        return -1;
     }
   pt->source_distance = s->distance;
   pt->history = 0;
   pt->start_time += pt->total_time;
   if (s->type == MXG_SOURCE_RAYFILE)
     return replay_rayfile (st, pt, num, num_collected, exposure_time);
   if (s->type == MXG_SOURCE_USER)
     return create_user (pt, num, num_collected, exposure_time);

   if (-1 == sync_config ())
     return -1;
   Device_Trace = 0;
   if (-1 == mxgd_run_source (num, pt->tag_start, pt->start_time,
                              (exposure_time != NULL) ? *exposure_time : -1.0,
                              Seed, 0, -1, &n, &last_t, NULL))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   /* the records, and their energy order sorted on the device */
   if ((n > 0) && ((-1 == mxgd_download_aos (pt->attributes, n))
                   || (-1 == mxgd_sort_store_by_energy (n, pt->sorted_index, pt->sorted_energies))))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   Resident = pt;
   pt->n_photons = n;
   pt->num_sorted = n;
   pt->history |= (MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK | MARX_TAG_OK);
   pt->tag_start += n;
   Resident_Tag_Start = pt->tag_start;
   if (n == 0) pt->total_time = 0.0;
   else pt->total_time = (s->time_mode == MXG_TIME_UNIFORM) ? s->uniform_tmax : last_t;
   *num_collected = n;
   /* the device list is in tag order; the next stage uploads sorted_index */
   Resident = NULL;
   return 0;
}

/* ---- device-resident pipeline ------------------------------------------------------ */

static unsigned int Last_Collected = 0;
static unsigned int Last_Tag_Start = 0;
static unsigned long Last_History = 0;

static unsigned int Stage_Counts[4];
static float Kernel_Ms[2];
static unsigned int Kernel_In[2];
/* stage times (ms) of the last pass by timing slot; zero for a
 * graph-launched pass */
static float Last_Slot_Ms[8];

int mxg_trace_device (unsigned int num, unsigned int tag_start, double start_time,
                      double exposure_time, unsigned int *num_collected,
                      unsigned int *num_detected, float *stage_ms)
{
   float ms[8];
   unsigned int counts[8];
   unsigned int n = 0, ndet;
   unsigned long hist = MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK | MARX_TAG_OK;
   int with_grating = (Grating_Type == MARX_GRATING_HETG) || (Grating_Type == MARX_GRATING_LETG);
   /* stages hand their survivors densely to the next one: the grating to
    * the detector; the mirror to the LETG grating (HETG + ACIS-S measured
    * the reflection's extra writes costing what its lighter grating pass
    * saves, LETG + HRC-S 4 percent faster per step) */
   int relay = (Grating_Type == MARX_GRATING_LETG);

   if (-1 == sync_config ())
     return -1;
   if (Mirror_Type != MARX_MIRROR_HRMA)
     {
        mxg_error ("Mirror not initialized.");
        return -1;
     }
   if ((Mxg_Cfg.src.type == MXG_SOURCE_RAYFILE) || (Mxg_Cfg.src.type == MXG_SOURCE_USER))
     {
        mxg_error ("mxg_trace_device creates its photons on the device; a RAYFILE or USER source "
                   "makes its rays on the host, through marx_create_photons");
        return -1;
     }
   /* timing slots: 0 source+dither (its count: the photons created, read
    * back with the timings, so no stage waits on the host), 1 mirror, 2
    * grating, 3 detector, 4 and 5 the two mirror kernels alone */
   if ((-1 == mxgd_launch_source (num, tag_start, start_time, exposure_time, Seed, 0, 0, 1))
       || (-1 == mxgd_run_mirror (Mxg_Cfg.src.distance, Seed, 1, NULL, relay, 1))
       || (with_grating && (-1 == mxgd_run_grating (Seed, 2, NULL, relay, Detector_Type > 0)))
       || ((Detector_Type > 0) && (-1 == mxgd_run_detect (start_time, Seed, 3, NULL, with_grating)))
       || (-1 == mxgd_read_timings (ms, counts, 6))
       || (-1 == mxgd_set_created (counts[0])))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   n = counts[0];
   hist |= MARX_MIRROR_SHELL_OK;
   if (with_grating) hist |= grating_history ();
   if (Detector_Type > 0) hist |= detector_history ();
   /* photons alive after each stage (a skipped stage passes them through) */
   Stage_Counts[0] = n;
   Stage_Counts[1] = counts[1];
   Stage_Counts[2] = with_grating ? counts[2] : counts[1];
   Stage_Counts[3] = (Detector_Type > 0) ? counts[3] : Stage_Counts[2];
   ndet = Stage_Counts[3];
   Kernel_Ms[0] = ms[MXGD_SLOT_HRMA_ENTER];
   Kernel_Ms[1] = ms[MXGD_SLOT_HRMA_REFLECT];
   Kernel_In[0] = counts[MXGD_SLOT_HRMA_ENTER];
   Kernel_In[1] = counts[MXGD_SLOT_HRMA_REFLECT];
   memcpy (Last_Slot_Ms, ms, sizeof (Last_Slot_Ms));
   Resident = NULL;
   Last_Collected = n;
   Last_Tag_Start = tag_start;
   Last_History = hist;
   Device_Trace = 1;
   if (num_collected) *num_collected = n;
   if (num_detected) *num_detected = ndet;
   if (stage_ms)
     {
        stage_ms[0] = ms[0];
        stage_ms[1] = 0.0f;           /* dither runs inside the source stage */
        stage_ms[2] = ms[1];
        stage_ms[3] = with_grating ? ms[2] : 0.0f;
        stage_ms[4] = (Detector_Type > 0) ? ms[3] : 0.0f;
        stage_ms[5] = stage_ms[0] + stage_ms[2] + stage_ms[3] + stage_ms[4];
     }
   return 0;
}

/* The device-resident trace into caller-owned columns with no allocation
 * (no counterpart in the reference). */
int mxg_trace_device_noalloc (const mxg_soa_view *store, unsigned int num, unsigned int tag_start,
                              double start_time, double exposure_time, unsigned int *num_collected,
                              unsigned int *num_detected)
{
   if (store == NULL)
     {
        mxg_error ("mxg_trace_device_noalloc: no columns given");
        return -1;
     }
   if (Mirror_Type != MARX_MIRROR_HRMA)
     {
        mxg_error ("Mirror not initialized.");
        return -1;
     }
   if ((Mxg_Cfg.src.type == MXG_SOURCE_RAYFILE) || (Mxg_Cfg.src.type == MXG_SOURCE_USER))
     {
        mxg_error ("mxg_trace_device_noalloc creates its photons on the device; a RAYFILE or USER source "
                   "makes its rays on the host, through marx_create_photons");
        return -1;
     }
   if (-1 == mxgd_trace_noalloc (store, num, tag_start, start_time, exposure_time, Seed, &Mxg_Cfg,
                                 mxg_farena_base (), mxg_farena_size (), mxg_darena_base (),
                                 mxg_darena_size (), Mxg_Cfg_Generation, num_collected, num_detected))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

/* ---- batched point sources -------------------------------------------------- */

typedef struct
{
   unsigned int nsrc, n, generation, grating, detector, alloc_epoch;
   unsigned long long seed;
   double exposure, start_time, distance;
   mxgd_schedule sched;
}
Batch_Key;

static mxgd_graph *Batch_Graph = NULL;
static Batch_Key Batch_Graph_Key;


/* one pass over the uploaded batch (n photons): source, mirror, grating,
 * detector, in timing slots 0-3, or the persistent pass in slot 0 (during
 * graph capture the stages record no events but still copy their survivor
 * counts) */
static int batch_pass (unsigned int n, double start_time, double exposure, int with_grating)
{
   /* stages hand their survivors densely to the next one: the grating to
    * the detector; the mirror to the LETG grating (HETG + ACIS-S measured
    * the reflection's extra writes costing what its lighter grating pass
    * saves, LETG + HRC-S 4 percent faster per step) */
   int relay = (Grating_Type == MARX_GRATING_LETG);
   mxgd_schedule sched;
   mxgd_get_schedule (&sched);
   /* the persistent pass traces point sources without the FEF step */
   if (sched.persist_tile && !Mxg_Cfg.det.fef_on && (Mxg_Cfg.src.type == MXG_SOURCE_POINT))
     return mxgd_run_batch_persistent (n, exposure, start_time, Seed, 0);
   if ((-1 == mxgd_run_batch_source (n, exposure, start_time, Seed, 0, 1))
       || (-1 == mxgd_run_mirror (Mxg_Cfg.src.distance, Seed, 1, NULL, relay, 1))
       || (with_grating && (-1 == mxgd_run_grating (Seed, 2, NULL, relay, Detector_Type > 0)))
       || ((Detector_Type > 0) && (-1 == mxgd_run_detect (start_time, Seed, 3, NULL, with_grating))))
     return -1;
   return 0;
}

static int batch_finish (unsigned int n, unsigned int tag_start, int with_grating,
                         unsigned int *num_detected)
{
   float ms[8];
   unsigned int counts[8];
   unsigned long hist = MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK
     | MARX_TAG_OK | MARX_MIRROR_SHELL_OK;
   if (-1 == mxgd_read_timings (ms, counts, 8))
     return -1;
   Stage_Counts[0] = n;
   Stage_Counts[1] = counts[1];
   Stage_Counts[2] = with_grating ? counts[2] : Stage_Counts[1];
   Stage_Counts[3] = (Detector_Type > 0) ? counts[3] : Stage_Counts[2];
   if (with_grating) hist |= grating_history ();
   if (Detector_Type > 0) hist |= detector_history ();
   Kernel_Ms[0] = ms[MXGD_SLOT_HRMA_ENTER];
   Kernel_Ms[1] = ms[MXGD_SLOT_HRMA_REFLECT];
   Kernel_In[0] = counts[MXGD_SLOT_HRMA_ENTER];
   Kernel_In[1] = counts[MXGD_SLOT_HRMA_REFLECT];
   Resident = NULL;
   Last_Collected = n;
   Last_Tag_Start = tag_start;
   Last_History = hist;
   Device_Trace = 1;
   memcpy (Last_Slot_Ms, ms, sizeof (ms));
   if (num_detected) *num_detected = Stage_Counts[3];
   return 0;
}

static int batch_setup (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                        unsigned int *n)
{
   if (-1 == sync_config ())
     return -1;
   if (Mirror_Type != MARX_MIRROR_HRMA)
     {
        mxg_error ("Mirror not initialized.");
        return -1;
     }
   if ((Mxg_Cfg.src.spectrum != MXG_SPECTRUM_FLAT) && (Mxg_Cfg.src.spectrum != MXG_SPECTRUM_FILE))
     {
        mxg_error ("Source not initialized (marx_create_source of a POINT or DISK source with a spectrum).");
        return -1;
     }
   if ((num_sources == 0) || (dirs == NULL) || (counts == NULL))
     {
        mxg_error ("mxg_trace_batch: no sources");
        return -1;
     }
   if ((Mxg_Cfg.src.type == MXG_SOURCE_IMAGE) || (Mxg_Cfg.src.type == MXG_SOURCE_RAYFILE)
       || (Mxg_Cfg.src.type == MXG_SOURCE_USER))
     {
        mxg_error ("mxg_trace_batch: the sources of a batch are points, disks, Gaussians, beta models "
                   "or lines about their directions, not images, ray files or plug-ins");
        return -1;
     }
   if (-1 == mxgd_batch_set_sources (num_sources, dirs, counts, n))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

int mxg_trace_batch (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                     unsigned int tag_start, double start_time, double exposure, int use_graph,
                     unsigned int *num_detected, float *ms)
{
   int with_grating = (Grating_Type == MARX_GRATING_HETG) || (Grating_Type == MARX_GRATING_LETG);
   unsigned int n;
   Batch_Key key;

   if (-1 == batch_setup (num_sources, dirs, counts, &n))
     return -1;
   /* size the store first: the key then names the buffers a capture holds */
   if (-1 == mxgd_batch_prepare (n))
     goto device_error;
   memset (&key, 0, sizeof (key));
   key.nsrc = num_sources;
   key.n = n;
   key.generation = Mxg_Cfg_Generation;
   key.grating = (unsigned int) Grating_Type;
   key.detector = (unsigned int) Detector_Type;
   key.seed = Seed;
   key.exposure = exposure;
   key.start_time = start_time;
   key.distance = Mxg_Cfg.src.distance;
   key.alloc_epoch = mxgd_alloc_epoch ();
   mxgd_get_schedule (&key.sched);

   if (use_graph)
     {
        if ((Batch_Graph == NULL) || memcmp (&key, &Batch_Graph_Key, sizeof (key)))
          {
             mxgd_graph_destroy (Batch_Graph);
             Batch_Graph = NULL;
             if (-1 == mxgd_graph_begin ())
               goto device_error;
             if (-1 == batch_pass (n, start_time, exposure, with_grating))
               {
                  mxgd_graph *junk = NULL;
                  (void) mxgd_graph_end (&junk);
                  mxgd_graph_destroy (junk);
                  goto device_error;
               }
             if (-1 == mxgd_graph_end (&Batch_Graph))
               goto device_error;
             Batch_Graph_Key = key;
          }
        if ((-1 == mxgd_batch_set_tag_start (tag_start))
            || (-1 == mxgd_graph_launch (Batch_Graph, ms)))
          goto device_error;
     }
   else
     {
        if ((-1 == mxgd_batch_set_tag_start (tag_start))
            || (-1 == batch_pass (n, start_time, exposure, with_grating))
            || (-1 == batch_finish (n, tag_start, with_grating, num_detected)))
          goto device_error;
        if (ms) *ms = Last_Slot_Ms[0] + Last_Slot_Ms[1] + Last_Slot_Ms[2] + Last_Slot_Ms[3];
        return 0;
     }
   if (-1 == batch_finish (n, tag_start, with_grating, num_detected))
     goto device_error;
   return 0;

device_error:
   mxg_error ("%s", mxgd_last_error ());
   return -1;
}

int mxg_set_schedule (const mxg_schedule *s)
{
   mxgd_schedule d;
   if (s == NULL) return -1;
   memcpy (d.blocks_per_cu, s->blocks_per_cu, sizeof (d.blocks_per_cu));
   d.reflect_waves = s->reflect_waves;
   d.persist_tile = s->persistent_tile;
   d.persist_blocks_per_cu = s->persistent_blocks_per_cu;
   if (-1 == mxgd_set_schedule (&d))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

int mxg_get_schedule (mxg_schedule *s)
{
   mxgd_schedule d;
   if (s == NULL) return -1;
   mxgd_get_schedule (&d);
   memcpy (s->blocks_per_cu, d.blocks_per_cu, sizeof (d.blocks_per_cu));
   s->reflect_waves = d.reflect_waves;
   s->persistent_tile = d.persist_tile;
   s->persistent_blocks_per_cu = d.persist_blocks_per_cu;
   return 0;
}

/* per-kernel times of one timed (uncaptured) batch pass, best of repeats:
 * source, paraboloid, reflection, grating stage, detector stage */
static int time_kernels (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                         unsigned int tag_start, double start_time, double exposure,
                         unsigned int repeats, float best[5])
{
   unsigned int r;
   for (int k = 0; k < 5; k++) best[k] = 1e30f;
   for (r = 0; r < repeats; r++)
     {
        float t[5];
        if (-1 == mxg_trace_batch (num_sources, dirs, counts, tag_start, start_time, exposure, 0,
                                   NULL, NULL))
          return -1;
        t[0] = Last_Slot_Ms[0];
        t[1] = Kernel_Ms[0];
        t[2] = Kernel_Ms[1];
        t[3] = Last_Slot_Ms[2];
        t[4] = Last_Slot_Ms[3];
        for (int k = 0; k < 5; k++)
          if (t[k] < best[k]) best[k] = t[k];
     }
   return 0;
}

/* best graph-launched pass time of the batch under the installed schedule
 * (the first launch, which captures, is not counted) */
static int graph_pass_ms (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                          unsigned int tag_start, double start_time, double exposure, unsigned int repeats,
                          float *best_ms)
{
   float ms = 0.0f;
   *best_ms = 1e30f;
   for (unsigned int r = 0; r < repeats + 1; r++)
     {
        if (-1 == mxg_trace_batch (num_sources, dirs, counts, tag_start, start_time, exposure, 1, NULL, &ms))
          return -1;
        if ((r > 0) && (ms < *best_ms)) *best_ms = ms;
     }
   return 0;
}

int mxg_autotune_batch (unsigned int num_sources, const double *dirs, const unsigned int *counts,
                        unsigned int tag_start, double start_time, double exposure,
                        unsigned int repeats, mxg_schedule *best, float *best_ms)
{
   static const unsigned int bpc[4] = {2, 4, 8, 16};
   static const unsigned int tiles[4] = {512, 1024, 2048, 4096};
   static const unsigned int tile_bpc[3] = {2, 3, 4};
   mxg_schedule cur, sel;
   float best_t[5], t[5], pass_ms, ms;
   int k, c;

   if (repeats == 0) repeats = 1;
   if (-1 == mxg_get_schedule (&cur))
     return -1;
   sel = cur;
   sel.persistent_tile = 0;
   for (k = 0; k < 5; k++) best_t[k] = 1e30f;
   /* the staged pass: blocks per CU, every kernel at once (their times are
    * independent) */
   for (c = 0; c < 4; c++)
     {
        mxg_schedule s = sel;
        for (k = 0; k < 5; k++) s.blocks_per_cu[k] = bpc[c];
        if ((-1 == mxg_set_schedule (&s))
            || (-1 == time_kernels (num_sources, dirs, counts, tag_start, start_time, exposure,
                                    repeats, t)))
          goto fail;
        for (k = 0; k < 5; k++)
          if (t[k] < best_t[k])
            {
               best_t[k] = t[k];
               sel.blocks_per_cu[k] = bpc[c];
            }
     }
   /* the grating at one round of the blocks a CU holds at once (0), which
    * none of the counts above need be (three for HETG) */
   {
      mxg_schedule s = sel;
      s.blocks_per_cu[3] = 0;
      if ((-1 == mxg_set_schedule (&s))
          || (-1 == time_kernels (num_sources, dirs, counts, tag_start, start_time, exposure,
                                  repeats, t)))
        goto fail;
      if (t[3] < best_t[3])
        {
           best_t[3] = t[3];
           sel.blocks_per_cu[3] = 0;
        }
   }
   /* the reflection kernel's register budget */
   best_t[2] = 1e30f;
   for (c = 2; c <= 4; c++)
     {
        mxg_schedule s = sel;
        s.reflect_waves = (unsigned int) c;
        if ((-1 == mxg_set_schedule (&s))
            || (-1 == time_kernels (num_sources, dirs, counts, tag_start, start_time, exposure,
                                    repeats, t)))
          goto fail;
        if (t[2] < best_t[2])
          {
             best_t[2] = t[2];
             sel.reflect_waves = (unsigned int) c;
          }
     }
   /* the staged pass against the persistent one, whole graph-launched
    * passes (the persistent pass traces point sources without the FEF
    * step) */
   if ((-1 == mxg_set_schedule (&sel))
       || (-1 == graph_pass_ms (num_sources, dirs, counts, tag_start, start_time, exposure, repeats, &pass_ms)))
     goto fail;
   if (!Mxg_Cfg.det.fef_on && (Mxg_Cfg.src.type == MXG_SOURCE_POINT))
     {
        mxg_schedule staged = sel;
        for (c = 0; c < 4; c++)
          for (k = 0; k < 3; k++)
            {
               mxg_schedule s = staged;
               s.persistent_tile = tiles[c];
               s.persistent_blocks_per_cu = tile_bpc[k];
               if ((-1 == mxg_set_schedule (&s))
                   || (-1 == graph_pass_ms (num_sources, dirs, counts, tag_start, start_time, exposure, repeats,
                                            &ms)))
                 goto fail;
               if (ms < pass_ms)
                 {
                    pass_ms = ms;
                    sel = s;
                 }
            }
     }
   if (-1 == mxg_set_schedule (&sel))
     goto fail;
   if (best) *best = sel;
   if (best_ms) *best_ms = pass_ms;
   return 0;

fail:
   (void) mxg_set_schedule (&cur);
   return -1;
}

int mxg_last_stage_ms (float ms[4])
{
   memcpy (ms, Last_Slot_Ms, 4 * sizeof (float));
   return 0;
}

int mxg_last_stage_counts (unsigned int counts[4])
{
   memcpy (counts, Stage_Counts, sizeof (Stage_Counts));
   return 0;
}

int mxg_last_kernel_stats (float ms[2], unsigned int photons_in[2])
{
   if (ms) memcpy (ms, Kernel_Ms, sizeof (Kernel_Ms));
   if (photons_in) memcpy (photons_in, Kernel_In, sizeof (Kernel_In));
   return 0;
}

int mxg_download_photons (Marx_Photon_Type *pt)
{
   unsigned int n = Last_Collected;
   if (n > pt->max_n_photons)
     {
        mxg_error ("photon list too small: %u > %u", n, pt->max_n_photons);
        return -1;
     }
   if ((n > 0) && (-1 == mxgd_download_aos (pt->attributes, n)))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   pt->n_photons = n;
   pt->tag_start = Last_Tag_Start + n;
   pt->history = Last_History;
   if ((n > 0) && (-1 == mxgd_sort_store_by_energy (n, pt->sorted_index, pt->sorted_energies)))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   pt->num_sorted = n;
   marx_prune_photons (pt);
   return 0;
}

int mxg_resident_history (unsigned long *history)
{
   if (!Device_Trace)
     {
        mxg_error ("no device-resident trace (mxg_trace_device or mxg_trace_batch) to write");
        return -1;
     }
   *history = Last_History;
   return 0;
}

int mxg_get_soa (mxg_soa_view *view)
{
   if (-1 == mxgd_get_soa (view))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}
