/* detector_init.c -- ACIS-S, ACIS-I, HRC-S and HRC-I initialisation.
 *
 * Follows the reference's init chain:
 *   marx_detector_init            marx/libsrc/detector.c:298-357
 *   ACIS geometry                 acis_geom.c:30-238, caldb.c:256-491
 *                                 (INSTGEOM corners + INSTRUMENTS origin +
 *                                 AIMPOINTS offset), detector.c:383-400 basis
 *   ACIS QE                       caldb.c:505-576 (AXAF_QE, CCD_ID)
 *   ACIS frame-transfer streak    acis-i.c:170-197
 *   HRC-S geometry / shield       hrc_s_geom.c:119-579, hrc-s.c:383-464
 *   HRC-I geometry                hrc_i_geom.c:51-295, hrc-i.c:186-228
 *   MCP blur parameters           hrcblur.c:115-236
 *   HESF (Drake flat)             drake.c:170-260
 *
 *   ACIS contamination            aciscontam.c:218-517 (ACIS<ccd>_CONTAM or
 *                                 AXAF_CONTAM + CCD_ID; taus at TStart)
 *   ACIS FEF                      acis_fef.c:194-298, 706-874
 *
 * The CALDB contamination and FEF files are missing from the reference's
 * data tree (marx/data/caldb/.MISSING_LARGE_BLOBS); ACISContamFile and
 * ACISFEFFile name other files.  Without a contamination file the QE draw
 * uses a transmission of 1 and a message says so (the reference refuses to
 * initialise ACIS then, unless DetIdeal is set).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <math.h>

#include "mxg_host.h"

static void set3 (double *dst, JDMVector_Type v)
{
   dst[0] = v.x; dst[1] = v.y; dst[2] = v.z;
}

static JDMVector_Type get3 (const double *s)
{
   return mxg_v (s[0], s[1], s[2]);
}

/* ---- CALDB geometry --------------------------------------------------------- */

typedef struct { const char *inst; } Inst_Match;

static int match_instgeom (mxg_fits_table *t, void *client)
{
   return mxg_fits_keyword_is (t, "EXTNAME", "INSTGEOM")
     && mxg_fits_keyword_is (t, "INSTRUME", ((Inst_Match *) client)->inst);
}

static int match_extname (mxg_fits_table *t, void *client)
{
   return mxg_fits_keyword_is (t, "EXTNAME", (const char *) client);
}

typedef struct
{
   int id;
   JDMVector_Type ll, lr, ul, ur;
}
Corner_Type;

/* chip corners of one instrument from the GEOM INSTGEOM extension */
static int read_instgeom (const char *inst, Corner_Type *c, unsigned int nchips)
{
   char *file = mxg_caldb_file ("GEOM");
   mxg_fits_table *t;
   Inst_Match m;
   unsigned int r, k;

   if (file == NULL) return -1;
   m.inst = inst;
   if (NULL == (t = mxg_fits_find_table (file, match_instgeom, &m)))
     {
        mxg_error ("Unable to find the proper INSTGEOM extension in %s", file);
        free (file);
        return -1;
     }
   free (file);
   for (r = 0; r < mxg_fits_num_rows (t); r++)
     {
        double id, v[4][3];
        static const char *cols[4] = {"LL", "LR", "UL", "UR"};
        unsigned int j;
        if (-1 == mxg_fits_get_double (t, "CHIP_ID", r, 0, &id)) goto fail;
        for (j = 0; j < 4; j++)
          for (k = 0; k < 3; k++)
            if (-1 == mxg_fits_get_double (t, cols[j], r, k, &v[j][k])) goto fail;
        for (k = 0; k < nchips; k++)
          if (c[k].id == (int) id)
            {
               c[k].ll = get3 (v[0]);
               c[k].lr = get3 (v[1]);
               c[k].ul = get3 (v[2]);
               c[k].ur = get3 (v[3]);
            }
     }
   mxg_fits_close (t);
   return 0;
fail:
   mxg_fits_close (t);
   return -1;
}

static int read_named_3d (const char *object, const char *extname, const char *namecol,
                          const char *valcol, const char *name, JDMVector_Type *v)
{
   char *file = mxg_caldb_file (object);
   mxg_fits_table *t;
   unsigned int r;
   if (file == NULL) return -1;
   if (NULL == (t = mxg_fits_find_table (file, match_extname, (void *) extname)))
     {
        mxg_error ("Unable to find %s in %s", extname, file);
        free (file);
        return -1;
     }
   for (r = 0; r < mxg_fits_num_rows (t); r++)
     {
        char buf[64];
        double d[3];
        unsigned int k;
        if (-1 == mxg_fits_get_string (t, namecol, r, buf, sizeof (buf))) break;
        if (0 != strcasecmp (buf, name)) continue;
        for (k = 0; k < 3; k++)
          if (-1 == mxg_fits_get_double (t, valcol, r, k, &d[k])) goto fail;
        *v = get3 (d);
        mxg_fits_close (t);
        free (file);
        return 0;
     }
fail:
   mxg_error ("Unable to find %s %s data in %s", name, valcol, file);
   mxg_fits_close (t);
   free (file);
   return -1;
}

/* corners += aimpoint (STF-STT) + instrument origin (STT-LSI) (caldb.c:469-491) */
static int patch_aimpoint (const char *inst, const char *aimpt, Corner_Type *c, unsigned int n)
{
   JDMVector_Type origin, aim, ofs;
   unsigned int k;
   if ((-1 == read_named_3d ("GEOM", "INSTRUMENTS", "INSTRUMENT", "ORIGIN", inst, &origin))
       || (-1 == read_named_3d ("AIMPTS", "AIMPOINTS", "AIMPOINT_NAME", "AIMPOINT", aimpt, &aim)))
     return -1;
   ofs = mxg_v_sum (aim, origin);
   for (k = 0; k < n; k++)
     {
        c[k].lr = mxg_v_sum (c[k].lr, ofs);
        c[k].ll = mxg_v_sum (c[k].ll, ofs);
        c[k].ur = mxg_v_sum (c[k].ur, ofs);
        c[k].ul = mxg_v_sum (c[k].ul, ofs);
     }
   return 0;
}

/* basis from corners (detector.c:383-400) */
static void set_chip (mxg_chip *g, const Corner_Type *c)
{
   JDMVector_Type xhat = mxg_v_diff (c->lr, c->ll), yhat = mxg_v_diff (c->ul, c->ll);
   g->id = c->id;
   g->xlen = mxg_v_length (xhat);
   g->ylen = mxg_v_length (yhat);
   mxg_v_normalize (&xhat);
   mxg_v_normalize (&yhat);
   set3 (g->x_ll, c->ll);
   set3 (g->xhat, xhat);
   set3 (g->yhat, yhat);
   set3 (g->normal, mxg_v_cross (xhat, yhat));
}

/* ---- efficiency tables -------------------------------------------------------- */

typedef struct { int ccd; } Ccd_Match;

static int match_acis_qe (mxg_fits_table *t, void *client)
{
   char buf[32];
   if (!mxg_fits_keyword_is (t, "EXTNAME", "AXAF_QE")) return 0;
   if (-1 == mxg_fits_keyword (t, "CCD_ID", buf, sizeof (buf))) return 0;
   return atoi (buf) == ((Ccd_Match *) client)->ccd;
}

static int read_acis_qe (mxg_chip *g)
{
   char *file = mxg_caldb_file ("ACISQE");
   mxg_fits_table *t;
   Ccd_Match m;
   unsigned int n, i;
   float *e, *q;

   if (file == NULL) return -1;
   m.ccd = g->id;
   if (NULL == (t = mxg_fits_find_table (file, match_acis_qe, &m)))
     {
        mxg_error ("Unable to find ACIS-%d QE extension in %s", g->id, file);
        free (file);
        return -1;
     }
   mxg_message (2, "\t%s for [CCDID = %d]\n", file, g->id);
   free (file);
   n = mxg_fits_num_rows (t);
   if (n < 2)
     {
        mxg_error ("Expecting more than 2 rows of QE for CCD_ID=%d", g->id);
        mxg_fits_close (t);
        return -1;
     }
   g->qe_n = n;
   g->qe_e_off = mxg_farena_reserve (n);
   g->qe_off = mxg_farena_reserve (n);
   if ((g->qe_e_off == (unsigned int) -1) || (g->qe_off == (unsigned int) -1))
     {
        mxg_fits_close (t);
        return -1;
     }
   e = mxg_farena_ptr (g->qe_e_off);
   q = mxg_farena_ptr (g->qe_off);
   for (i = 0; i < n; i++)
     if ((-1 == mxg_fits_get_float (t, "ENERGY", i, 0, e + i))
         || (-1 == mxg_fits_get_float (t, "QE", i, 0, q + i)))
       {
          mxg_fits_close (t);
          return -1;
       }
   mxg_fits_close (t);
   return mxg_bucket_of_farena (g->qe_e_off, n, MXG_BUCKETS_QE, &g->qe_bucket);
}

/* two-column bdat efficiency file; an empty name means unit efficiency */
static int read_qe_file (Param_File_Type *pf, const char *pname, int ideal, mxg_qe_table *q)
{
   char buf[1024], *file;
   float *cols[2];
   int status;
   memset (q, 0, sizeof (*q));
   if (-1 == pf_get_file (pf, (char *) pname, buf, sizeof (buf)))
     return -1;
   if (ideal || (buf[0] == 0))
     return 0;
   if (NULL == (file = mxg_data_file_name (buf)))
     return -1;
   mxg_message (2, "\t%s\n", file);
   if (-1 == mxg_read_bdat (file, 2, cols, &q->n))
     {
        free (file);
        return -1;
     }
   free (file);
   q->e_off = mxg_farena_push (cols[0], q->n);
   q->eff_off = mxg_farena_push (cols[1], q->n);
   free (cols[1]);
   if ((q->e_off == (unsigned int) -1) || (q->eff_off == (unsigned int) -1))
     {
        free (cols[0]);
        return -1;
     }
   /* the bucket index of the energies: the MCP QE tables hold 2e4-4e4 rows */
   status = mxg_build_bucket (cols[0], q->n, 0, MXG_BUCKETS_QE, &q->bucket);
   free (cols[0]);
   return status;
}

static int read_opt_consts (Param_File_Type *pf, const char *pname, unsigned int *n,
                            unsigned int *e_off, unsigned int *b_off, unsigned int *d_off)
{
   char buf[1024], *file;
   float *cols[3];
   *n = 0;
   if (-1 == pf_get_file (pf, (char *) pname, buf, sizeof (buf)))
     return -1;
   if (buf[0] == 0) return 0;
   if (NULL == (file = mxg_data_file_name (buf)))
     return -1;
   if (-1 == mxg_read_bdat (file, 3, cols, n))
     {
        free (file);
        return -1;
     }
   free (file);
   *e_off = mxg_farena_push (cols[0], *n);
   *b_off = mxg_farena_push (cols[1], *n);
   *d_off = mxg_farena_push (cols[2], *n);
   free (cols[0]); free (cols[1]); free (cols[2]);
   return 0;
}

/* ---- ACIS ------------------------------------------------------------------------ */

/* ---- ACIS FEF (acis_fef.c:330-899) -------------------------------------------------- */

#define FEF_REGION 32
#define FEF_NREG (1024 / FEF_REGION)
#define FEF_FWHM_TO_SIGMA 0.4246609001440095

static int match_function (mxg_fits_table *t, void *client)
{
   (void) client;
   return mxg_fits_keyword_is (t, "EXTNAME", "FUNCTION");
}

/* One region's rows [r0, r1) of the FUNCTION table into the arenas; returns
 * its index in the descriptor list or -1. */
static int fef_region (mxg_fits_table *t, unsigned int r0, unsigned int r1, unsigned int ng,
                       double *desc, unsigned int *ndesc)
{
   unsigned int ne = r1 - r0, i, k, eo, co, go;
   float *e, *c, *g;
   if ((ne < 2) || (*ndesc >= 10 * FEF_NREG * FEF_NREG))
     {
        mxg_error ("FEF region of rows %u-%u: %s", r0 + 1, r1, (ne < 2) ? "fewer than 2 energies" : "too many regions");
        return -1;
     }
   eo = mxg_farena_reserve (ne);
   co = mxg_farena_reserve (ne);
   go = mxg_farena_reserve (3 * ne * ng);
   if ((eo == (unsigned int) -1) || (co == (unsigned int) -1) || (go == (unsigned int) -1)) return -1;
   e = mxg_farena_ptr (eo);
   c = mxg_farena_ptr (co);
   g = mxg_farena_ptr (go);
   for (i = 0; i < ne; i++)
     {
        if ((-1 == mxg_fits_get_float (t, "ENERGY", r0 + i, 0, e + i))
            || (-1 == mxg_fits_get_float (t, "CHANNEL", r0 + i, 0, c + i)))
          return -1;
        for (k = 0; k < ng; k++)
          {
             char name[32];
             float fwhm, pos, ampl;
             snprintf (name, sizeof (name), "G%u_FWHM", k + 1);
             if (-1 == mxg_fits_get_float (t, name, r0 + i, 0, &fwhm)) return -1;
             snprintf (name, sizeof (name), "G%u_POS", k + 1);
             if (-1 == mxg_fits_get_float (t, name, r0 + i, 0, &pos)) return -1;
             snprintf (name, sizeof (name), "G%u_AMPL", k + 1);
             if (-1 == mxg_fits_get_float (t, name, r0 + i, 0, &ampl)) return -1;
             g[3 * (i * ng + k) + 0] = ampl;
             g[3 * (i * ng + k) + 1] = pos;
             g[3 * (i * ng + k) + 2] = (float) (FEF_FWHM_TO_SIGMA * fwhm);
          }
     }
   /* energies and channels must increase (acis_fef.c:299-312) */
   for (i = 1; i < ne; i++)
     if ((e[i] <= e[i - 1]) || (c[i] <= c[i - 1]))
       {
          mxg_error ("Fef Map does not have monotonically increasing %s", (e[i] <= e[i - 1]) ? "energies" : "PHAs");
          return -1;
       }
   desc[5 * *ndesc + 0] = ne;
   desc[5 * *ndesc + 1] = ng;
   desc[5 * *ndesc + 2] = eo;
   desc[5 * *ndesc + 3] = co;
   desc[5 * *ndesc + 4] = go;
   return (int) (*ndesc)++;
}

/* The FEF file of the ACIS chips min_id..max_id: rows grouped into regions by
 * runs of REGNUM, each region mapped onto its 32-pixel cells
 * (acis_fef.c:194-298, 706-874).  The file is ACISFEFFile when set, else the
 * CALDB ACISFEF file; when neither exists the FEF is off and PHA and PI stay 0. */
static int read_acis_fef (Param_File_Type *pf, int min_id, int max_id)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   char buf[1024], *file = NULL;
   mxg_fits_table *t;
   unsigned int nrows, ng, r0, i, ndesc = 0, id;
   double *map, *desc;
   FILE *fp;

   d->fef_on = 0;
   buf[0] = 0;
   if ((-1 == pf_get_string (pf, "ACISFEFFile", buf, sizeof (buf))) || (buf[0] == 0))
     file = mxg_caldb_file ("ACISFEF");
   else
     file = strdup (buf);
   if ((file == NULL) || (NULL == (fp = fopen (file, "rb"))))
     {
        mxg_message (0, "ACIS FEF file not available: PHA and PI are not computed\n");
        free (file);
        return 0;
     }
   fclose (fp);
   if (NULL == (t = mxg_fits_find_table (file, match_function, NULL)))
     {
        mxg_error ("Unable to open ACIS FEF file %s", file);
        free (file);
        return -1;
     }
   mxg_message (2, "\t%s\n", file);
   free (file);
   for (ng = 0; ng < 64; ng++)
     {
        char name[32];
        snprintf (name, sizeof (name), "G%u_AMPL", ng + 1);
        if (-1 == mxg_fits_column (t, name, NULL, NULL)) break;
     }
   nrows = mxg_fits_num_rows (t);
   d->fef_map_off = mxg_darena_reserve (10 * FEF_NREG * FEF_NREG);
   d->fef_desc_off = mxg_darena_reserve (5 * 10 * FEF_NREG * FEF_NREG);
   if ((ng == 0) || (ng > 16) || (d->fef_map_off == (unsigned int) -1) || (d->fef_desc_off == (unsigned int) -1))
     {
        mxg_error ("ACIS FEF file with %u Gaussians (1 to 16 supported)", ng);
        mxg_fits_close (t);
        return -1;
     }
   map = mxg_darena_ptr (d->fef_map_off);
   for (i = 0; i < 10 * FEF_NREG * FEF_NREG; i++) map[i] = -1.0;
   r0 = 0;
   while (r0 < nrows)
     {
        double v[6];
        unsigned int r1 = r0 + 1;
        int k, x, y, xmin, xmax, ymin, ymax, ccd, regnum;
        if ((-1 == mxg_fits_get_double (t, "CCD_ID", r0, 0, v + 0))
            || (-1 == mxg_fits_get_double (t, "REGNUM", r0, 0, v + 1)))
          goto fail;
        ccd = (int) v[0];
        regnum = (int) v[1];
        while (r1 < nrows)
          {
             double w[2];
             if ((-1 == mxg_fits_get_double (t, "CCD_ID", r1, 0, w + 0))
                 || (-1 == mxg_fits_get_double (t, "REGNUM", r1, 0, w + 1)))
               goto fail;
             if (((int) w[1] != regnum) || ((int) w[0] != ccd)) break;
             r1++;
          }
        if ((ccd < min_id) || (ccd > max_id))
          {
             r0 = r1;
             continue;
          }
        if ((-1 == mxg_fits_get_double (t, "CHIPX_LO", r0, 0, v + 2))
            || (-1 == mxg_fits_get_double (t, "CHIPX_HI", r0, 0, v + 3))
            || (-1 == mxg_fits_get_double (t, "CHIPY_LO", r0, 0, v + 4))
            || (-1 == mxg_fits_get_double (t, "CHIPY_HI", r0, 0, v + 5)))
          goto fail;
        xmin = (int) v[2] - 1; xmax = (int) v[3];
        ymin = (int) v[4] - 1; ymax = (int) v[5];
        if ((xmin < 0) || (ymin < 0) || (xmax > 1024) || (ymax > 1024)
            || (xmin % FEF_REGION) || (xmax % FEF_REGION) || (ymin % FEF_REGION) || (ymax % FEF_REGION))
          {
             mxg_error ("FEF region CHIPX/Y_LO/HI values must be multiples of %d within 1-1024", FEF_REGION);
             goto fail;
          }
        desc = mxg_darena_ptr (d->fef_desc_off);
        if (-1 == (k = fef_region (t, r0, r1, ng, desc, &ndesc)))
          goto fail;
        map = mxg_darena_ptr (d->fef_map_off);
        for (x = xmin / FEF_REGION; x < xmax / FEF_REGION; x++)
          for (y = ymin / FEF_REGION; y < ymax / FEF_REGION; y++)
            map[(ccd * FEF_NREG + x) * FEF_NREG + y] = k;
        r0 = r1;
     }
   mxg_fits_close (t);
   for (id = (unsigned int) min_id; id <= (unsigned int) max_id; id++)
     for (i = 0; i < FEF_NREG * FEF_NREG; i++)
       if (map[id * FEF_NREG * FEF_NREG + i] < 0)
         {
            mxg_message (0, "***Warning: FEF for CCDID=%u contains holes\n", id);
            break;
          }
   d->fef_on = 1;
   d->fef_max_gaussians = (int) ng;
   return 0;
fail:
   mxg_fits_close (t);
   return -1;
}

/* ---- ACIS contamination (aciscontam.c) ------------------------------------------ */

static int match_contam (mxg_fits_table *t, void *client)
{
   int ccd = ((Ccd_Match *) client)->ccd;
   char name[32], buf[32];
   snprintf (name, sizeof (name), "ACIS%d_CONTAM", ccd);
   if (mxg_fits_keyword_is (t, "EXTNAME", name))
     return 1;
   if (!mxg_fits_keyword_is (t, "EXTNAME", "AXAF_CONTAM"))
     return 0;
   if (-1 == mxg_fits_keyword (t, "CCD_ID", buf, sizeof (buf)))
     return 0;
   return atoi (buf) == ccd;
}

/* tau0 / tau1 of one layer at the observation start, linear in time and
 * extrapolated past the ends (aciscontam.c:236-269) */
static void contam_taus (const double *times, const float *tau0, const float *tau1, unsigned int n,
                         double *t0_out, double *t1_out)
{
   double t = _Marx_TStart_MJDsecs, w;
   unsigned int n0, n1 = 0;
   if (n == 1)
     {
        *t0_out = tau0[0];
        *t1_out = tau1[0];
        return;
     }
   /* the first node above t (JDMbinary_search_d) */
   while ((n1 < n) && (times[n1] <= t)) n1++;
   if (n1 == n) n1--;
   if (n1 == 0) n1 = 1;
   n0 = n1 - 1;
   w = (t - times[n0]) / (times[n1] - times[n0]);
   *t0_out = tau0[n0] + (tau0[n1] - tau0[n0]) * w;
   *t1_out = tau1[n0] + (tau1[n1] - tau1[n0]) * w;
}

static int contam_fail (mxg_fits_table *t, const char *file, unsigned int row, int ccd, const char *why)
{
   mxg_error ("%s: error in row %u of %s for CCD=%d", why, row + 1, file, ccd);
   mxg_fits_close (t);
   return -1;
}

/* One chip's layers from the contamination file (aciscontam.c:328-465). */
static int read_chip_contam (const char *file, int ccd, mxg_contam *c)
{
   mxg_fits_table *t;
   Ccd_Match m;
   unsigned int nrows, layer, rep_e, rep_mu, rep_t, rep_t0, rep_t1, rep_f = 0;
   int has_fxy;
   char ty, buf[32];

   memset (c, 0, sizeof (*c));
   m.ccd = ccd;
   if (NULL == (t = mxg_fits_find_table (file, match_contam, &m)))
     {
        mxg_error ("Unable to open acis contam file %s for CCD %d", file, ccd);
        return -1;
     }
   has_fxy = (0 == mxg_fits_column (t, "fxy", &rep_f, &ty));
   if ((-1 == mxg_fits_column (t, "component", NULL, &ty)) || (-1 == mxg_fits_column (t, "n_energy", NULL, &ty))
       || (-1 == mxg_fits_column (t, "energy", &rep_e, &ty)) || (-1 == mxg_fits_column (t, "mu", &rep_mu, &ty))
       || (-1 == mxg_fits_column (t, "n_time", NULL, &ty)) || (-1 == mxg_fits_column (t, "time", &rep_t, &ty))
       || (-1 == mxg_fits_column (t, "tau0", &rep_t0, &ty)) || (-1 == mxg_fits_column (t, "tau1", &rep_t1, &ty)))
     {
        mxg_error ("Error processing ACIS contam table %s for CCD=%d", file, ccd);
        mxg_fits_close (t);
        return -1;
     }
   nrows = mxg_fits_num_rows (t);
   if (nrows > MXG_MAX_CONTAM_LAYERS)
     {
        mxg_error ("The number of layers in %s for CCD=%d is larger than supported", file, ccd);
        mxg_fits_close (t);
        return -1;
     }
   if (has_fxy)
     {
        int b = 1;
        if (0 == mxg_fits_keyword (t, "FXYBLK", buf, sizeof (buf)))
          {
             b = atoi (buf);
             if ((b <= 0) || (b * (1024 / b) != 1024))
               {
                  mxg_error ("Invalid fxyblk value");
                  mxg_fits_close (t);
                  return -1;
               }
          }
        c->blocking = (unsigned int) b;
        c->fxy_model = MXG_CONTAM_TABLE;
        if (rep_f != (1024u / c->blocking) * (1024u / c->blocking))
          {
             mxg_error ("The FXY column is expected to have %u * %u values", 1024 / c->blocking,
                        1024 / c->blocking);
             mxg_fits_close (t);
             return -1;
          }
     }
   else if (ccd <= 3)
     {
        /* aciscontam.c:141-172: radial about the array's centre */
        c->fxy_model = MXG_CONTAM_RADIAL;
        c->x0 = ((ccd == 1) || (ccd == 2)) ? 0.5 : 1024.5;
        c->y0 = 1024.5;
     }
   else
     c->fxy_model = MXG_CONTAM_ROWS;

   for (layer = 0; layer < nrows; layer++)
     {
        double comp, ne, nt, *times;
        float *tau0, *tau1, *e, *mu;
        unsigned int j, n_time, n_energy;

        if ((-1 == mxg_fits_get_double (t, "component", layer, 0, &comp))
            || (-1 == mxg_fits_get_double (t, "n_energy", layer, 0, &ne))
            || (-1 == mxg_fits_get_double (t, "n_time", layer, 0, &nt)))
          return contam_fail (t, file, layer, ccd, "unreadable");
        if (comp != 0.0)
          return contam_fail (t, file, layer, ccd, "Expecting COMPONENT=0");
        n_time = (unsigned int) nt;
        n_energy = (unsigned int) ne;
        if ((n_time <= 1) || (n_energy <= 1))
          return contam_fail (t, file, layer, ccd, "Expecting n_energy or n_time to be greater than 1");
        if ((n_time > rep_t) || (n_time > rep_t0) || (n_time > rep_t1))
          return contam_fail (t, file, layer, ccd, "n_time value is larger than data arrays");
        if ((n_energy > rep_e) || (n_energy > rep_mu))
          return contam_fail (t, file, layer, ccd, "n_energy value is larger than data arrays");

        times = (double *) malloc (n_time * sizeof (double));
        tau0 = (float *) malloc (n_time * sizeof (float));
        tau1 = (float *) malloc (n_time * sizeof (float));
        c->mu_n[layer] = n_energy;
        c->mu_e_off[layer] = mxg_farena_reserve (n_energy);
        c->mu_off[layer] = mxg_farena_reserve (n_energy);
        if ((times == NULL) || (tau0 == NULL) || (tau1 == NULL)
            || (c->mu_e_off[layer] == (unsigned int) -1) || (c->mu_off[layer] == (unsigned int) -1))
          {
             free (times); free (tau0); free (tau1);
             return contam_fail (t, file, layer, ccd, "out of memory");
          }
        e = mxg_farena_ptr (c->mu_e_off[layer]);
        mu = mxg_farena_ptr (c->mu_off[layer]);
        for (j = 0; j < n_time; j++)
          if ((-1 == mxg_fits_get_double (t, "time", layer, j, times + j))
              || (-1 == mxg_fits_get_float (t, "tau0", layer, j, tau0 + j))
              || (-1 == mxg_fits_get_float (t, "tau1", layer, j, tau1 + j))
              || ((j > 0) && !(times[j] > times[j - 1])))
            {
               free (times); free (tau0); free (tau1);
               return contam_fail (t, file, layer, ccd, "Expecting the time grid to be monotonically increasing");
            }
        for (j = 0; j < n_energy; j++)
          if ((-1 == mxg_fits_get_float (t, "energy", layer, j, e + j))
              || (-1 == mxg_fits_get_float (t, "mu", layer, j, mu + j))
              || ((j > 0) && !(e[j] > e[j - 1])))
            {
               free (times); free (tau0); free (tau1);
               return contam_fail (t, file, layer, ccd, "Expecting the ENERGY grid to be monotonically increasing");
            }
        contam_taus (times, tau0, tau1, n_time, c->tau0 + layer, c->tau1 + layer);
        free (times); free (tau0); free (tau1);
        if (has_fxy)
          {
             float *f;
             c->fxy_off[layer] = mxg_farena_reserve (rep_f);
             if (c->fxy_off[layer] == (unsigned int) -1)
               return contam_fail (t, file, layer, ccd, "out of memory");
             f = mxg_farena_ptr (c->fxy_off[layer]);
             for (j = 0; j < rep_f; j++)
               if (-1 == mxg_fits_get_float (t, "fxy", layer, j, f + j))
                 return contam_fail (t, file, layer, ccd, "unreadable FXY");
          }
     }
   c->num_layers = nrows;
   mxg_fits_close (t);
   return 0;
}

/* The contamination of every chip (acis-s.c:486-487, acis-i.c:349): the
 * file is ACISContamFile when set (an error when it cannot be read, as in
 * aciscontam.c:311-316), else the CALDB ACISCONTAM file.  The CALDB file
 * is missing from the reference's data: without it no contamination is
 * applied (said once). */
static int read_acis_contam (Param_File_Type *pf)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   char buf[1024], *file;
   FILE *fp;
   int k, explicit_file;

   d->contam_on = 0;
   memset (d->contam, 0, sizeof (d->contam));
   buf[0] = 0;
   explicit_file = (0 == pf_get_string (pf, "ACISContamFile", buf, sizeof (buf))) && (buf[0] != 0);
   file = explicit_file ? strdup (buf) : mxg_caldb_file ("ACISCONTAM");
   if ((file == NULL) || (NULL == (fp = fopen (file, "rb"))))
     {
        if (explicit_file)
          {
             mxg_error ("Unable to open acis contam file %s", buf);
             free (file);
             return -1;
          }
        static int said = 0;
        if (!said)
          mxg_message (0, "ACIS contamination file not available: no contamination is applied "
                          "(the reference needs the file unless DetIdeal is set)\n");
        said = 1;
        free (file);
        return 0;
     }
   fclose (fp);
   mxg_message (2, "\t%s\n", file);
   for (k = 0; k < d->num_chips; k++)
     if (-1 == read_chip_contam (file, d->chip[k].id, d->contam + k))
       {
          free (file);
          return -1;
       }
   free (file);
   d->contam_on = 1;
   return 0;
}

static int acis_init (Param_File_Type *pf, int is_s)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   Corner_Type c[6];
   unsigned int n = is_s ? 6 : 4, k;
   double exposure, transfer;

   if ((-1 == pf_get_double (pf, "ACIS_Exposure_Time", &exposure))
       || (-1 == pf_get_double (pf, "ACIS_Frame_Transfer_Time", &transfer)))
     return -1;
   if (exposure <= 0)
     {
        mxg_error ("ACIS_Exposure_Time must be greater than 0");
        return -1;
     }
   if (transfer < 0.0) transfer = 0.0;
   d->exposure_time = exposure;
   d->frame_transfer_time = transfer;
   d->frame_time = transfer + exposure;

   memset (c, 0, sizeof (c));
   for (k = 0; k < n; k++) c[k].id = (int) k + (is_s ? 4 : 0);
   if ((-1 == read_instgeom ("ACIS", c, n))
       || (-1 == patch_aimpoint ("ACIS", is_s ? "AS1" : "AI2", c, n)))
     return -1;
   d->num_chips = (int) n;
   for (k = 0; k < n; k++)
     {
        mxg_chip *g = d->chip + k;
        set_chip (g, c + k);
        /* corners do not give exactly 1024 x 0.024 mm (acis_geom.c:52-67) */
        g->x_pixel_size = g->xlen / 1024.0;
        g->y_pixel_size = g->ylen / 1024.0;
        g->xpixel_offset = 0.0;
        g->ypixel_offset = 0.0;
        if (d->ideal == 0)
          if (-1 == read_acis_qe (g))
            return -1;
     }
   if ((d->ideal == 0) && (-1 == read_acis_contam (pf)))
     return -1;
   return read_acis_fef (pf, is_s ? 4 : 0, is_s ? 9 : 3);
}

/* ---- HRC blur (hrcblur.c:115-236) ------------------------------------------------ */

static int hrc_blur_init (Param_File_Type *pf, char det, double dx, double dy)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   const double fwhm_to_sigma = 0.42466090014400953;
   double pixelsize, norm, g2;
   char name[64];

   dx *= 0.25;
   dy *= 0.25;
   pixelsize = 0.5 * (dx + dy);

#define GETB(s, v) snprintf (name, sizeof (name), "HRC-%c-Blur%s", det, s); \
   if (-1 == pf_get_double (pf, name, &d->v)) return -1;
   GETB ("G1FWHM", g1_sigma); GETB ("G1XCTR", g1_xctr); GETB ("G1YCTR", g1_yctr); GETB ("G1AMP", g1_wgt);
   GETB ("G2FWHM", g2_sigma); GETB ("G2XCTR", g2_xctr); GETB ("G2YCTR", g2_yctr); GETB ("G2AMP", g2_wgt);
   GETB ("L1FWHM", l1_hwhm); GETB ("L1XCTR", l1_xctr); GETB ("L1YCTR", l1_yctr); GETB ("L1AMP", l1_wgt);
   GETB ("L1RMAX", l1_rmax);
#undef GETB

   if ((d->g1_wgt < 0) || (d->g2_wgt < 0) || (d->l1_wgt < 0) || (d->g1_sigma <= 0)
       || (d->g2_sigma <= 0) || (d->l1_hwhm <= 0) || (d->l1_rmax <= 0))
     {
        mxg_error ("One or more of the HRC blur parameters has an invalid/unsupported value");
        return -1;
     }
   d->g1_sigma *= pixelsize; d->g1_xctr *= pixelsize; d->g1_yctr *= pixelsize;
   d->g2_sigma *= pixelsize; d->g2_xctr *= pixelsize; d->g2_yctr *= pixelsize;
   d->l1_hwhm *= pixelsize; d->l1_xctr *= pixelsize; d->l1_yctr *= pixelsize;
   d->l1_rmax *= pixelsize;

   d->g2_xctr -= d->g1_xctr;
   d->g2_yctr -= d->g1_yctr;
   d->l1_xctr -= d->g1_xctr;
   d->l1_yctr -= d->g1_yctr;
   d->g1_xctr = 0;
   d->g1_yctr = 0;

   d->g1_sigma *= fwhm_to_sigma;
   d->g2_sigma *= fwhm_to_sigma;
   d->l1_hwhm *= 0.5;

   norm = 1.0 / (2 * PI * d->g1_sigma * d->g1_sigma);
   d->g1_wgt /= norm;
   norm = 1.0 / (2 * PI * d->g2_sigma * d->g2_sigma);
   d->g2_wgt /= norm;
   g2 = d->l1_hwhm;
   g2 = g2 * g2;
   norm = 1.0 / (g2 * PI * log (1 + d->l1_rmax * d->l1_rmax / g2));
   d->l1_wgt /= norm;
   norm = d->g1_wgt + d->g2_wgt + d->l1_wgt;
   if (norm > 0)
     {
        d->g1_wgt /= norm;
        d->g2_wgt /= norm;
        d->l1_wgt /= norm;
     }
   return 0;
}

/* ---- HRC-S --------------------------------------------------------------------------- */

static int hrc_s_init (Param_File_Type *pf)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   double v_pix, v_phys[6], v_act[6], u_pix, u_phys[2], u_act[2], aim[3];
   double l_uv[2], r_uv[2], m_uv[2], l_xyz[3], r_xyz[3], m_xyz[3], l_c[2], r_c[2], m_c[2];
   double l_th, r_th, det_th;
   double sh_t, sh_l, sh_r, sh_x, sh_sl, sh_sr, gap_l, gap_r;
   double la[3], lb[3], ma[3], mb[3], ra[3], rb[3];
   mxg_simple_item geom[] =
     {
          {"V_Pixel_Size", 1, &v_pix, 1.0, 0}, {"V_Physical", 6, v_phys, 1.0, 0},
          {"V_Active", 6, v_act, 1.0, 0}, {"U_Pixel_Size", 1, &u_pix, 1.0, 0},
          {"U_Physical", 2, u_phys, 1.0, 0}, {"U_Active", 2, u_act, 1.0, 0},
          {"Aimpoint_Offset", 3, aim, 1.0, 0},
          {"Left_UV", 2, l_uv, 1.0, 0}, {"Right_UV", 2, r_uv, 1.0, 0}, {"Middle_UV", 2, m_uv, 1.0, 0},
          {"Left_XYZ", 3, l_xyz, 1.0, 0}, {"Right_XYZ", 3, r_xyz, 1.0, 0}, {"Middle_XYZ", 3, m_xyz, 1.0, 0},
          {"Left_CXCY", 2, l_c, 1.0, 0}, {"Right_CXCY", 2, r_c, 1.0, 0}, {"Middle_CXCY", 2, m_c, 1.0, 0},
          {"Left_Angle", 1, &l_th, PI / 180.0, 0}, {"Right_Angle", 1, &r_th, PI / 180.0, 0},
          {"Detector_Angle", 1, &det_th, PI / 180.0, 0},
          {NULL, 0, NULL, 0.0, 0}
     };
   mxg_simple_item shield[] =
     {
          {"Ion_Shield_T", 1, &sh_t, 1.0, 0}, {"Ion_Shield_L", 1, &sh_l, 1.0, 0},
          {"Ion_Shield_R", 1, &sh_r, 1.0, 0}, {"Ion_Shield_X", 1, &sh_x, 1.0, 0},
          {"Ion_Shield_SL", 1, &sh_sl, 1.0, 0}, {"Ion_Shield_SR", 1, &sh_sr, 1.0, 0},
          {"Ion_Shield_GapL", 1, &gap_l, 1.0, 0}, {"Ion_Shield_GapR", 1, &gap_r, 1.0, 0},
          {NULL, 0, NULL, 0.0, 0}
     };
   char *file;
   Corner_Type c[3];
   unsigned int i, k;
   JDMVector_Type center;

   if (NULL == (file = mxg_data_file_name ("hrc/hrc_s_geom.txt")))
     return -1;
   if ((-1 == mxg_read_simple_data_file (file, geom))
       || (-1 == mxg_read_simple_data_file (file, shield)))
     {
        free (file);
        return -1;
     }
   free (file);
   for (i = 0; i < 3; i++)
     {
        l_xyz[i] += aim[i];
        r_xyz[i] += aim[i];
        m_xyz[i] += aim[i];
     }
   /* (u,v) -> xyz direction vectors (hrc_s_geom.c:160-205) */
   la[0] = 0.0; la[1] = 0.0; la[2] = u_pix;
   lb[0] = -1.0 * sin (l_th) * v_pix; lb[1] = -1.0 * cos (l_th) * v_pix; lb[2] = 0.0;
   ma[0] = 0.0; ma[1] = 0.0; ma[2] = u_pix;
   mb[0] = 0.0; mb[1] = -1.0 * v_pix; mb[2] = 0.0;
   ra[0] = 0.0; ra[1] = 0.0; ra[2] = u_pix;
   rb[0] = 1.0 * sin (r_th) * v_pix; rb[1] = -1.0 * cos (r_th) * v_pix; rb[2] = 0.0;

   if (-1 == hrc_blur_init (pf, 'S', u_pix, v_pix))
     return -1;

   memset (c, 0, sizeof (c));
   for (k = 0; k < 3; k++)
     {
        const double *uv0 = (k == 0) ? l_uv : (k == 1) ? m_uv : r_uv;
        const double *x0 = (k == 0) ? l_xyz : (k == 1) ? m_xyz : r_xyz;
        const double *al = (k == 0) ? la : (k == 1) ? ma : ra;
        const double *be = (k == 0) ? lb : (k == 1) ? mb : rb;
        double vlo = v_act[2 * k], vhi = v_act[2 * k + 1];
        double uvs[4][2];
        JDMVector_Type *dst[4];
        unsigned int j;
        JDMVector_Type axis = mxg_v (-1, 0, 0);
        double s = sin (det_th), co = cos (det_th);

        uvs[0][0] = u_act[0]; uvs[0][1] = vhi;   /* ul */
        uvs[1][0] = u_act[1]; uvs[1][1] = vhi;   /* ur */
        uvs[2][0] = u_act[0]; uvs[2][1] = vlo;   /* ll */
        uvs[3][0] = u_act[1]; uvs[3][1] = vlo;   /* lr */
        dst[0] = &c[k].ul; dst[1] = &c[k].ur; dst[2] = &c[k].ll; dst[3] = &c[k].lr;
        for (j = 0; j < 4; j++)
          {
             double du = uvs[j][0] - uv0[0], dv = uvs[j][1] - uv0[1];
             *dst[j] = mxg_v (x0[0] + du * al[0] + dv * be[0],
                              x0[1] + du * al[1] + dv * be[1],
                              x0[2] + du * al[2] + dv * be[2]);
          }
        c[k].id = (int) k + 1;
        /* rotate the MCP about the optical axis (hrc_s_geom.c:262-272) */
        c[k].ll = mxg_v_rotate_vector1 (c[k].ll, axis, co, s);
        c[k].lr = mxg_v_rotate_vector1 (c[k].lr, axis, co, s);
        c[k].ul = mxg_v_rotate_vector1 (c[k].ul, axis, co, s);
        c[k].ur = mxg_v_rotate_vector1 (c[k].ur, axis, co, s);

        /* pixel mapping of chip id k+1 (hrc_s_geom.c:344-394) */
        d->pix_u[k + 1] = u_act[0];
        d->pix_v[k + 1] = vlo;
        d->pix_u0[k + 1] = uv0[0];
        d->pix_v0[k + 1] = uv0[1];
        d->pix_x0[k + 1] = ((k == 0) ? l_c : (k == 1) ? m_c : r_c)[0];
        d->pix_y0[k + 1] = ((k == 0) ? l_c : (k == 1) ? m_c : r_c)[1];
     }
   if (-1 == patch_aimpoint ("HRC-S", "HS1", c, 3))
     return -1;
   d->num_chips = 3;
   for (k = 0; k < 3; k++)
     {
        set_chip (d->chip + k, c + k);
        d->chip[k].x_pixel_size = 6.429e-3;
        d->chip[k].y_pixel_size = 6.429e-3;
     }
   d->u_pixel_size = u_pix;
   d->v_pixel_size = v_pix;

   if ((-1 == read_qe_file (pf, "HRC-S-QEFile0", d->ideal, d->mcp_qe + 0))
       || (-1 == read_qe_file (pf, "HRC-S-QEFile1", d->ideal, d->mcp_qe + 1))
       || (-1 == read_qe_file (pf, "HRC-S-QEFile2", d->ideal, d->mcp_qe + 2))
       || (-1 == read_qe_file (pf, "HRC-S-UVISFile0", d->ideal, d->filter_qe + 0))
       || (-1 == read_qe_file (pf, "HRC-S-UVISFile1", d->ideal, d->filter_qe + 1))
       || (-1 == read_qe_file (pf, "HRC-S-UVISFile2", d->ideal, d->filter_qe + 2))
       || (-1 == read_qe_file (pf, "HRC-S-UVISFile3", d->ideal, d->filter_qe + 3)))
     return -1;
   d->mcp_map[0] = -1; d->mcp_map[1] = 2; d->mcp_map[2] = 1; d->mcp_map[3] = 0;

   /* geometric centre of the middle MCP plus the detector offset */
   center = mxg_v_sum (mxg_v (d->offset[0], d->offset[1], d->offset[2]),
                       mxg_v_ax1_bx2 (0.5, c[1].ll, 0.5, c[1].ur));
   d->shield_t = sh_t; d->shield_l = sh_l; d->shield_r = sh_r; d->shield_x = sh_x;
   d->shield_sl = sh_sl; d->shield_sr = sh_sr;
   d->shield_y_center = center.y;
   d->shield_z_center = center.z;
   d->shield_sr_gap = sh_sr + gap_r;
   d->shield_sl_gap = sh_sl + gap_l;

   if (-1 == pf_get_boolean (pf, "HRC-HESF", &d->use_drake))
     return -1;
   if (d->use_drake)
     {
        double xoff, zoff, gap1, gap2, len, crw, corner_x, corner_z;
        int nplates, i2;
        if ((-1 == pf_get_double (pf, "HESFOffsetX", &xoff))
            || (-1 == pf_get_double (pf, "HESFOffsetZ", &zoff))
            || (-1 == pf_get_double (pf, "HESFGapY1", &gap1))
            || (-1 == pf_get_double (pf, "HESFGapY2", &gap2))
            || (-1 == pf_get_integer (pf, "HESFN", &nplates))
            || (-1 == pf_get_double (pf, "HESFLength", &len))
            || (-1 == pf_get_double (pf, "HESFCrWidth", &crw)))
          return -1;
        if ((nplates <= 0) || (nplates > MXG_MAX_DRAKE / 2))
          {
             mxg_error ("HESFN must be between 1 and %d", MXG_MAX_DRAKE / 2);
             return -1;
          }
        if ((-1 == read_opt_consts (pf, "HESFOptConstC", &d->drake_c_n, &d->drake_c_e_off,
                                    &d->drake_c_beta_off, &d->drake_c_delta_off))
            || (-1 == read_opt_consts (pf, "HESFOptConstCr", &d->drake_cr_n, &d->drake_cr_e_off,
                                       &d->drake_cr_beta_off, &d->drake_cr_delta_off)))
          return -1;
        d->drake_n = 2 * nplates;
        d->drake_cr_width = crw;
        corner_x = xoff;
        corner_z = zoff;
        for (i2 = 0; i2 < nplates; i2++)
          {
             char name[32];
             double h, theta, h_tan;
             JDMVector_Type e1, e2, a;
             mxg_rect *r1 = d->drake + i2, *r2 = d->drake + i2 + nplates;
             snprintf (name, sizeof (name), "HESFHeight%d", i2 + 1);
             if (-1 == pf_get_double (pf, name, &h)) return -1;
             snprintf (name, sizeof (name), "HESFTheta%d", i2 + 1);
             if (-1 == pf_get_double (pf, name, &theta)) return -1;
             theta = theta * (PI / 180.0);
             h_tan = h * tan (theta);

             e1 = mxg_v (0.0, 1.0, 0.0);
             r1->len1 = len;
             e2 = mxg_v (h, 0.0, -h_tan);
             r1->len2 = mxg_v_length (e2);
             mxg_v_normalize (&e2);
             set3 (r1->e1, e1);
             set3 (r1->e2, e2);
             set3 (r1->normal, mxg_v_cross (e1, e2));
             a = mxg_v_sum (center, mxg_v (corner_x, gap1, corner_z));
             set3 (r1->a, a);

             e1 = mxg_v (0.0, -1.0, 0.0);
             r2->len1 = len;
             e2 = mxg_v (h, 0.0, -h_tan);
             r2->len2 = mxg_v_length (e2);
             mxg_v_normalize (&e2);
             set3 (r2->e1, e1);
             set3 (r2->e2, e2);
             set3 (r2->normal, mxg_v_cross (e2, e1));
             a = mxg_v_sum (center, mxg_v (corner_x, -gap2, corner_z));
             set3 (r2->a, a);

             corner_z -= h_tan;
             corner_x += h;
          }
     }
   return 0;
}

/* ---- HRC-I ----------------------------------------------------------------------------- */

static int hrc_i_init (Param_File_Type *pf)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   double xps, yps, ll_xyz[3], c_xyz[3], lr_vec[3], c_cxcy[2], aim_cxcy[2];
   mxg_simple_item geom[] =
     {
          {"HRC_I_X_Pixel_Size", 1, &xps, 1.0, 0}, {"HRC_I_Y_Pixel_Size", 1, &yps, 1.0, 0},
          {"HRC_I_LL_XYZ", 3, ll_xyz, 1.0, 0}, {"HRC_I_Center_XYZ", 3, c_xyz, 1.0, 0},
          {"HRC_I_LL_to_LR_Vector", 3, lr_vec, 1.0, 0}, {"HRC_I_Center_CXCY", 2, c_cxcy, 1.0, 0},
          {"HRC_I_Aimpoint_CXCY", 2, aim_cxcy, 1.0, 0},
          {NULL, 0, NULL, 0.0, 0}
     };
   JDMVector_Type a, b, center, ll, lr, LL, LR, UL, UR, dxv;
   double len, ll_cx, ll_cy;
   char *file;
   Corner_Type c;

   if (NULL == (file = mxg_data_file_name ("hrc/hrc_i_geom.txt")))
     return -1;
   if (-1 == mxg_read_simple_data_file (file, geom))
     {
        free (file);
        return -1;
     }
   free (file);
   /* (hrc_i_geom.c:51-105) */
   a = get3 (lr_vec);
   len = mxg_v_length (a);
   if (len == 0.0)
     {
        mxg_error ("Length of vector from LL to LR must be non-zero");
        return -1;
     }
   a = mxg_v (a.x * (1.0 / len), a.y * (1.0 / len), a.z * (1.0 / len));
   center = get3 (c_xyz);
   LL = get3 (ll_xyz);
   ll = mxg_v_diff (LL, center);
   if (0.0 == mxg_v_length (ll))
     {
        mxg_error ("The LL corner must not be the center");
        return -1;
     }
   lr = mxg_v_ax1_bx2 (1.0, ll, -2.0 * mxg_v_dot (a, ll), a);
   LR = mxg_v_sum (center, lr);
   UR = mxg_v_sum (center, mxg_v (-1.0 * ll.x, -1.0 * ll.y, -1.0 * ll.z));
   UL = mxg_v_sum (center, mxg_v (-1.0 * lr.x, -1.0 * lr.y, -1.0 * lr.z));
   if ((xps <= 0.0) || (yps <= 0.0))
     {
        mxg_error ("X and Y Pixel sizes must be larger than 0");
        return -1;
     }
   b = mxg_v_diff (UL, LL);
   mxg_v_normalize (&b);
   ll_cx = c_cxcy[0] + mxg_v_dot (ll, a) / xps;
   ll_cy = c_cxcy[1] + mxg_v_dot (ll, b) / yps;
   dxv = mxg_v_diff (mxg_v_ax1_bx2 ((ll_cx - aim_cxcy[0]) * xps, a, (ll_cy - aim_cxcy[1]) * yps, b), LL);
   c.ll = mxg_v_sum (LL, dxv);
   c.lr = mxg_v_sum (LR, dxv);
   c.ur = mxg_v_sum (UR, dxv);
   c.ul = mxg_v_sum (UL, dxv);
   c.id = 0;

   if (-1 == hrc_blur_init (pf, 'I', xps, yps))
     return -1;
   if (-1 == patch_aimpoint ("HRC-I", "HI1", &c, 1))
     return -1;
   d->num_chips = 1;
   set_chip (d->chip, &c);
   d->chip[0].x_pixel_size = 6.429e-3;
   d->chip[0].y_pixel_size = 6.429e-3;
   d->chip[0].xpixel_offset = ll_cx;
   d->chip[0].ypixel_offset = ll_cy;
   d->u_pixel_size = xps;
   d->v_pixel_size = yps;
   d->pix_x0[0] = ll_cx;
   d->pix_y0[0] = ll_cy;
   if ((-1 == read_qe_file (pf, "HRC-I-QEFile", d->ideal, d->mcp_qe + 0))
       || (-1 == read_qe_file (pf, "HRC-I-UVISFile", d->ideal, d->filter_qe + 0)))
     return -1;
   d->mcp_map[0] = 0;
   return 0;
}

/* ---- entry --------------------------------------------------------------------------------- */

int mxg_detector_init (Param_File_Type *pf)
{
   mxg_detector_cfg *d = &Mxg_Cfg.det;
   char buf[64];
   int type, status;

   if (-1 == pf_get_string (pf, "DetectorType", buf, sizeof (buf)))
     return -1;
   memset (d, 0, sizeof (*d));
   Mxg_Cfg_Generation++;
   if (!strcmp (buf, "NONE")) return 0;
   if (!strcmp (buf, "ACIS-S")) type = MARX_DETECTOR_ACIS_S;
   else if (!strcmp (buf, "ACIS-I")) type = MARX_DETECTOR_ACIS_I;
   else if (!strcmp (buf, "HRC-S")) type = MARX_DETECTOR_HRC_S;
   else if (!strcmp (buf, "HRC-I")) type = MARX_DETECTOR_HRC_I;
   else
     {
        mxg_error ("DetectorType \"%s\" not supported.", buf);
        return -1;
     }
   d->matrix[0] = 1; d->matrix[4] = 1; d->matrix[8] = 1;
   if ((-1 == pf_get_double (pf, "DetOffsetX", &d->offset[0]))
       || (-1 == pf_get_double (pf, "DetOffsetY", &d->offset[1]))
       || (-1 == pf_get_double (pf, "DetOffsetZ", &d->offset[2]))
       || (-1 == pf_get_boolean (pf, "DetIdeal", &d->ideal))
       || (-1 == pf_get_boolean (pf, "DetExtendFlag", &d->extend)))
     return -1;
   d->focal_length = 10065.5;

   mxg_message (0, "Initializing %s detector...\n", buf);
   switch (type)
     {
      case MARX_DETECTOR_ACIS_S: status = acis_init (pf, 1); break;
      case MARX_DETECTOR_ACIS_I: status = acis_init (pf, 0); break;
      case MARX_DETECTOR_HRC_S: status = hrc_s_init (pf); break;
      default: status = hrc_i_init (pf); break;
     }
   if (status == -1)
     {
        mxg_error ("Error initializing DetectorType %s.", buf);
        memset (d, 0, sizeof (*d));
        return -1;
     }
   d->type = type;
   Mxg_Cfg_Generation++;
   return type;
}
