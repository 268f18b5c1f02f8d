/* events.c -- event coordinates (marx2fits) of the detected photons of an
 * ACIS or HRC simulation, computed on the GPU from the device-resident trace.
 *
 * The host side gathers what marx2fits reads besides the photon columns:
 * the tiling offsets of the tiled detector system (acis_geom.c:37-49,
 * 104-141, 175-196; hrc_s_geom.c:321-327, 477-510 for AXAF-HRC-2.7S;
 * hrc_i_geom.c:130-140, 216-236), the chip pixel of each facet's lower-left
 * corner (hrc_s_geom.c:329-336, 344-394; hrc_i_geom.c:139-140), the
 * detector's focal-plane system (detpix.c:41-61), the frame length TimeDel
 * (marx2fits.c:2536-2545) and the CALDB subpixel tables
 * (acis_subpix.c:121-268), and hands them to the device layer. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "mxg_host.h"
#include "mxg_device.h"

/* ACIS tiled-system offsets by chip id (acis_geom.c:38-48) */
static const double Tdet_Xoff[MXG_MAX_CHIPS] = {3061, 5131, 3061, 5131, 791, 1833, 2875, 3917, 4959, 6001};
static const double Tdet_Yoff[MXG_MAX_CHIPS] = {5131, 4107, 4085, 3061, 1702, 1702, 1702, 1702, 1702, 1702};
/* HRC-S (AXAF-HRC-2.7S) by chip id 1-3, HRC-I (AXAF-HRC-2.4I) chip 0 */
static const double Hrc_S_Tdet_Yoff[3] = {-12.0, 16444.0, 32900.0};

/* subpixel tables: for table t (0 FI, 1 BI) and flight grade g, Sub_N[t][g]
 * energies at Sub_Off[t][g] in Sub_Tab, then as many chip x and chip y
 * offsets */
static float *Sub_Tab;
static unsigned int Sub_Len, Sub_Cap;
static unsigned int Sub_Off[2][256], Sub_N[2][256];
static int Sub_Loaded;

static int sub_push (const float *v, unsigned int n)
{
   if (Sub_Len + n > Sub_Cap)
     {
        unsigned int cap = 2 * (Sub_Len + n) + 1024;
        float *p = (float *) realloc (Sub_Tab, cap * sizeof (float));
        if (p == NULL)
          {
             mxg_error ("Not enough memory for the subpixel tables");
             return -1;
          }
        Sub_Tab = p;
        Sub_Cap = cap;
     }
   memcpy (Sub_Tab + Sub_Len, v, n * sizeof (float));
   Sub_Len += n;
   return 0;
}

static int match_extname (mxg_fits_table *t, void *client)
{
   return mxg_fits_keyword_is (t, "EXTNAME", (const char *) client);
}

/* One extension of the subpixel file (acis_subpix.c:121-222): rows of
 * fltgrade, npoints and the energy, chipx_offset, chipy_offset vectors. */
static int read_subpix_ext (const char *file, const char *extname, int tab)
{
   mxg_fits_table *t = mxg_fits_find_table (file, match_extname, (void *) extname);
   unsigned int rep[3], i, j, num_rows;
   const char *vec[3] = {"ENERGY", "CHIPX_OFFSET", "CHIPY_OFFSET"};
   float *buf = NULL;

   if (t == NULL)
     {
        mxg_error ("Unable to open ACIS subpix file %s[%s]", file, extname);
        return -1;
     }
   for (j = 0; j < 3; j++)
     if (-1 == mxg_fits_column (t, vec[j], &rep[j], NULL))
       {
          mxg_error ("Error processing ACIS subpix file %s[%s]", file, extname);
          goto fail;
       }
   if ((-1 == mxg_fits_column (t, "FLTGRADE", NULL, NULL)) || (-1 == mxg_fits_column (t, "NPOINTS", NULL, NULL)))
     {
        mxg_error ("Error processing ACIS subpix file %s[%s]", file, extname);
        goto fail;
     }
   if (0 == (num_rows = mxg_fits_num_rows (t)))
     {
        mxg_error ("ACIS subpix file %s contains NO data.", file);
        goto fail;
     }
   if ((rep[1] != rep[0]) || (rep[2] != rep[0]))
     {
        mxg_error ("The chipx/y_offset_column repeat count does not match the energy repeat count");
        goto fail;
     }
   if (NULL == (buf = (float *) malloc (3 * rep[0] * sizeof (float))))
     {
        mxg_error ("Not enough memory for the subpixel tables");
        goto fail;
     }
   for (i = 0; i < 256; i++) Sub_N[tab][i] = Sub_Off[tab][i] = 0;
   for (i = 0; i < num_rows; i++)
     {
        double g, np;
        int fltgrade, npoints;
        if ((-1 == mxg_fits_get_double (t, "FLTGRADE", i, 0, &g))
            || (-1 == mxg_fits_get_double (t, "NPOINTS", i, 0, &np)))
          goto fail_row;
        fltgrade = (int) g;
        npoints = (int) np;
        if ((fltgrade < 0) || (fltgrade >= 256))
          {
             mxg_error ("Illegal flight grade in subpix column");
             goto fail_row;
          }
        if (npoints <= 0)
          {
             Sub_N[tab][fltgrade] = 0;
             continue;
          }
        if (npoints == 1)
          {
             mxg_error ("Expecting Number of points (NPOINTS) to be greater than 1");
             goto fail_row;
          }
        if (npoints > (int) rep[0])
          {
             mxg_error ("NPOINTS is greater than repeat count");
             goto fail_row;
          }
        for (j = 0; j < 3; j++)
          {
             unsigned int k;
             for (k = 0; k < (unsigned int) npoints; k++)
               if (-1 == mxg_fits_get_float (t, vec[j], i, k, buf + j * npoints + k))
                 goto fail_row;
          }
        Sub_Off[tab][fltgrade] = Sub_Len;
        Sub_N[tab][fltgrade] = (unsigned int) npoints;
        if (-1 == sub_push (buf, 3 * (unsigned int) npoints))
          goto fail;
        continue;
fail_row:
        mxg_error ("Error encountered while processing row %u of %s", i + 1, file);
        goto fail;
     }
   free (buf);
   mxg_fits_close (t);
   return 0;
fail:
   free (buf);
   mxg_fits_close (t);
   return -1;
}

/* The FI and BI tables of the CALDB ACISSUBPIX file (acis_subpix.c:224-268). */
static int load_subpix (void)
{
   char *file;
   int status;
   if (Sub_Loaded) return 0;
   if (NULL == (file = mxg_caldb_file ("ACISSUBPIX")))
     return -1;
   mxg_message (1, "Reading subpix file %s\n", file);
   Sub_Len = 0;
   status = read_subpix_ext (file, "MARX_ACIS_SUBPIX_FI", 0);
   if (status == 0) status = read_subpix_ext (file, "MARX_ACIS_SUBPIX_BI", 1);
   free (file);
   if (status == 0) Sub_Loaded = 1;
   return status;
}

int mxg_event_coords (const mxg_event_cfg *cfg, mxg_event_columns *out, unsigned int capacity,
                      unsigned int *num_rows)
{
   const mxg_detector_cfg *d = &Mxg_Cfg.det;
   mxg_event_params par;
   unsigned long history;
   double time_del, roll;
   int k, is_acis;

   if ((cfg == NULL) || (out == NULL) || (num_rows == NULL))
     {
        mxg_error ("mxg_event_coords: NULL argument");
        return -1;
     }
   *num_rows = 0;
   if (-1 == mxg_resident_history (&history))
     return -1;
   is_acis = (d->type == MARX_DETECTOR_ACIS_S) || (d->type == MARX_DETECTOR_ACIS_I);
   if (!is_acis && (d->type != MARX_DETECTOR_HRC_S) && (d->type != MARX_DETECTOR_HRC_I))
     {
        mxg_error ("event coordinates need an ACIS or HRC simulation");
        return -1;
     }
   if ((cfg->pix_adj < MXG_PIX_ADJ_NONE) || (cfg->pix_adj > MXG_PIX_ADJ_EXACT))
     {
        mxg_error ("Unsupported pixel adjustment %d", cfg->pix_adj);
        return -1;
     }

   memset (&par, 0, sizeof (par));
   /* EDSER is for ACIS; marx2fits randomizes HRC pixels instead
    * (marx2fits.c:3296-3308) */
   par.pix_adj = cfg->pix_adj;
   if (!is_acis && (par.pix_adj == MXG_PIX_ADJ_EDSER))
     par.pix_adj = MXG_PIX_ADJ_RANDOMIZE;
   if ((par.pix_adj == MXG_PIX_ADJ_EDSER) && (-1 == load_subpix ()))
     return -1;
   par.focal_length = cfg->focal_length;
   for (k = 0; k < 3; k++) par.det_offset[k] = cfg->det_offset[k];
   roll = cfg->roll_nom * PI / 180.0;
   par.cos_roll = cos (roll);
   par.sin_roll = sin (roll);
   par.tiling = (d->type == MARX_DETECTOR_ACIS_I) ? MXG_TILE_ACIS_I : MXG_TILE_SHIFT;
   par.dither_used = (cfg->dither_used != 0);
   par.has_grade = is_acis;
   par.held = is_acis;          /* marx2fits.c:2850-2851 */
   for (k = 0; k < MXG_MAX_CHIPS; k++)
     par.chip_index[k] = -1;
   if (is_acis)
     {
        par.fp_delta_s0 = 0.492 * (PI / (3600.0 * 180.0));
        par.fp_x0 = par.fp_y0 = 4096.5;
        for (k = 0; k < MXG_MAX_CHIPS; k++)
          {
             par.tdet_xoff[k] = Tdet_Xoff[k];
             par.tdet_yoff[k] = Tdet_Yoff[k];
          }
     }
   else if (d->type == MARX_DETECTOR_HRC_S)
     {
        par.fp_delta_s0 = 0.1318 * (PI / (3600.0 * 180.0));     /* AXAF-FP-2.3 */
        par.fp_x0 = par.fp_y0 = 32768.5;
        for (k = 1; k <= 3; k++)
          par.tdet_yoff[k] = Hrc_S_Tdet_Yoff[k - 1];
     }
   else
     {
        par.fp_delta_s0 = 0.13175 * (PI / (3600.0 * 180.0));    /* AXAF-FP-2.1 */
        par.fp_x0 = par.fp_y0 = 16384.5;
        par.tdet_xoff[0] = -1.0;
     }
   for (k = 0; k < d->num_chips; k++)
     {
        const mxg_chip *g = d->chip + k;
        int id = g->id;
        if ((id < 0) || (id >= MXG_MAX_CHIPS)) continue;
        /* the detector's own chips: 0-3 ACIS-I, 4-9 ACIS-S, 1-3 HRC-S, 0 HRC-I */
        if ((d->type == MARX_DETECTOR_ACIS_I) && (id > 3)) continue;
        if ((d->type == MARX_DETECTOR_ACIS_S) && (id < 4)) continue;
        par.chip_index[id] = k;
        if (d->type == MARX_DETECTOR_HRC_S)
          {
             /* the pixel of the corner: _marx_hrc_s_compute_pixel (id, 0, 0) */
             par.pixel_xoff[k] = d->pix_x0[id] + (d->pix_u[id] - d->pix_u0[id]);
             par.pixel_yoff[k] = d->pix_y0[id] + (d->pix_v[id] - d->pix_v0[id]);
          }
        else
          {
             par.pixel_xoff[k] = g->xpixel_offset;
             par.pixel_yoff[k] = g->ypixel_offset;
          }
     }
   if (par.pix_adj == MXG_PIX_ADJ_EDSER)
     {
        memcpy (par.sub_off, Sub_Off, sizeof (Sub_Off));
        memcpy (par.sub_n, Sub_N, sizeof (Sub_N));
     }

   /* TimeDel: exposure plus frame transfer (marx2fits.c:2538-2542) */
   time_del = cfg->exposure_time;
   if ((cfg->exposure_time > 0.0) && (cfg->frame_transfer_time > 0.0))
     time_del += cfg->frame_transfer_time;

   if (-1 == mxgd_event_coords (&par, (par.pix_adj == MXG_PIX_ADJ_EDSER) ? Sub_Tab : NULL,
                                (par.pix_adj == MXG_PIX_ADJ_EDSER) ? Sub_Len : 0,
                                cfg->total_time, time_del, cfg->seed, out, capacity, num_rows))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

int mxg_event_kernel_ms (float *ms)
{
   if (ms == NULL) return -1;
   *ms = mxgd_event_ms ();
   return 0;
}
