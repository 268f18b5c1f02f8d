/* hrma_init.c -- HRMA initialisation: shell geometry, tilt matrices,
 * optical constants, effective-area correction factors and WFold tables.
 *
 * Follows marx/libsrc/hrma.c:1097-1157 (_marx_hrma_mirror_init) with
 * get_shell_geometry (:551-602), init_hrma_shells (:671-776),
 * read_hrma_opt_constants (:791-823), read_hrma_correction_factors
 * (:834-866) and init_yaw_pitch (:842-897).  The result is written into
 * Mxg_Cfg.hrma and the table arenas.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "mxg_host.h"

double Marx_Mirror_Geometric_Area = 1.0;

static const unsigned int Shell_Numbers[MXG_NUM_SHELLS] = {1, 3, 4, 6};

static int get_par_d (Param_File_Type *pf, const char *fmt, unsigned int id, double *v)
{
   char name[64];
   snprintf (name, sizeof (name), fmt, id);
   return pf_get_double (pf, name, v);
}

/* "0000": '0' = open, '1' = closed; bit q set when quadrant q is open
 * (mirror.c:110-141). */
static int parse_shutters (const char *s, unsigned int *bitmap, unsigned int *num_open)
{
   unsigned int count = 0;
   *bitmap = 0;
   *num_open = 0;
   for (; *s; s++, count++)
     {
        if (*s == '0')
          {
             *bitmap |= (1u << count);
             *num_open += 1;
          }
        else if (*s != '1')
          {
             mxg_error ("Illegal character in shutter bitmap");
             return -1;
          }
     }
   if (count != 4)
     {
        mxg_error ("A shutter bitmap must have 4 characters");
        return -1;
     }
   return 0;
}

static int read_geometry (const char *file, double geom[MXG_NUM_SHELLS][2][9])
{
   /* per surface: x0 y0 z0 p k rho0 az_mis el_mis l */
   static const char *cols[9] = {"x0", "y0", "z0", "p", "k", "rho0", "az_mis", "el_mis", "l"};
   mxg_rdb *rdb = mxg_rdb_open (file);
   unsigned int i, s, c;
   if (rdb == NULL) return -1;
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     for (s = 0; s < 2; s++)
       {
          char key[8];
          snprintf (key, sizeof (key), "%c%u", s ? 'h' : 'p', Shell_Numbers[i]);
          for (c = 0; c < 9; c++)
            if (-1 == mxg_rdb_get_double (rdb, "mirror", key, cols[c], &geom[i][s][c]))
              {
                 mxg_rdb_close (rdb);
                 return -1;
              }
       }
   mxg_rdb_close (rdb);
   return 0;
}

static unsigned int push_float_table (float *v, unsigned int n)
{
   unsigned int off = mxg_farena_push (v, n);
   free (v);
   return off;
}

static int read_bdat_into_arena (const char *relname, unsigned int ncols,
                                 unsigned int *n, unsigned int *offs)
{
   char *file = mxg_data_file_name (relname);
   float *cols[8];
   unsigned int j;
   if (file == NULL) return -1;
   mxg_message (2, "\t%s\n", file);
   if (-1 == mxg_read_bdat (file, ncols, cols, n))
     {
        free (file);
        return -1;
     }
   free (file);
   for (j = 0; j < ncols; j++)
     if ((unsigned int) -1 == (offs[j] = push_float_table (cols[j], *n)))
       return -1;
   return 0;
}

/* Tilt matrices of one surface from its azimuth / elevation misalignment
 * in degrees (hrma.c:842-897). */
/* bucket index of a WFold table's e_alpha values (double arena), as the
 * kernels stage them: float values compared as doubles */
static int bucket_of_e_alphas (const mxg_wfold *w, unsigned int nbmax, mxg_bucket *b)
{
   const double *da;
   float *t;
   unsigned int i;
   int status;
   memset (b, 0, sizeof (*b));
   if (w->n_arrays < 2) return 0;
   if (NULL == (t = (float *) malloc (w->n_arrays * sizeof (float))))
     {
        mxg_error ("out of memory");
        return -1;
     }
   da = mxg_darena_ptr (w->e_alphas_off);
   for (i = 0; i < w->n_arrays; i++)
     t[i] = (float) da[i];
   status = mxg_build_bucket (t, w->n_arrays, 1, nbmax, b);
   free (t);
   return status;
}

static void yaw_pitch_matrices (double az_deg, double el_deg, double fwd[9], double bwd[9])
{
   double rz[9], ry[9];
   double el = -el_deg * (PI / 180.0);
   double az = -az_deg * (PI / 180);
   JDMVector_Type new_y_axis;

   mxg_m3_rot_z (rz, -az);
   new_y_axis = mxg_m3_vmul (rz, mxg_v (0, 1, 0));
   mxg_m3_rot (ry, new_y_axis, -el);
   mxg_m3_mul (fwd, ry, rz);
   mxg_m3_rot (ry, new_y_axis, el);
   mxg_m3_rot_z (rz, az);
   mxg_m3_mul (bwd, rz, ry);
}

int mxg_hrma_init (Param_File_Type *pf)
{
   mxg_hrma_cfg *h = &Mxg_Cfg.hrma;
   double geom[MXG_NUM_SHELLS][2][9];
   char buf[1024], *file;
   double total_area;
   unsigned int i;
   int verbose;

   memset (h, 0, sizeof (*h));

   if ((-1 == pf_get_integer (pf, "Verbose", &verbose))
       || (-1 == pf_get_double (pf, "HRMAVig", &h->vignetting))
       || (-1 == pf_get_double (pf, "HRMA_Cap_X", &h->cap_position))
       || (-1 == pf_get_boolean (pf, "HRMA_Use_Blur", &h->use_blur))
       || (-1 == pf_get_boolean (pf, "HRMA_Use_Scale_Factors", &h->use_scale))
       || (-1 == pf_get_boolean (pf, "HRMA_Use_WFold", &h->use_wfold))
       || (-1 == pf_get_boolean (pf, "HRMA_Ideal", &h->ideal))
       || (-1 == pf_get_boolean (pf, "HRMA_Use_Struts", &h->use_struts)))
     return -1;
   Mxg_Verbose = verbose;

   if (h->ideal)
     {
        h->use_blur = 0;
        h->use_wfold = 0;
     }

   if (-1 == pf_get_file (pf, "HRMA_Geometry_File", buf, sizeof (buf)))
     return -1;
   if (NULL == (file = mxg_data_file_name (buf)))
     return -1;
   mxg_message (2, "\t%s\n", file);
   if (-1 == read_geometry (file, geom))
     {
        free (file);
        return -1;
     }
   free (file);

   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        mxg_shell_cfg *s = h->shell + i;
        unsigned int id = Shell_Numbers[i];
        char sh[16], name[64];
        double xoffset, cap_pos;
        unsigned int k;

        s->mirror_number = id;
        snprintf (name, sizeof (name), "Shutters%u", id);
        if ((-1 == pf_get_string (pf, name, sh, sizeof (sh)))
            || (-1 == parse_shutters (sh, &s->shutter_bitmap, &s->num_open_shutters))
            || (-1 == get_par_d (pf, "P%uBlur", id, &s->blur[0]))
            || (-1 == get_par_d (pf, "H%uBlur", id, &s->blur[1]))
            || (-1 == get_par_d (pf, "P%uScatFactor", id, &s->scat_factor[0]))
            || (-1 == get_par_d (pf, "H%uScatFactor", id, &s->scat_factor[1])))
          return -1;
        snprintf (name, sizeof (name), "HRMA_P%uH%u_XOffset", id, id);
        if (-1 == pf_get_double (pf, name, &xoffset))
          return -1;

        cap_pos = h->cap_position + xoffset;
        for (k = 0; k < 2; k++)
          {
             double *g = geom[i][k];   /* x0 y0 z0 p k rho0 az el l */
             double len = g[8];
             s->conic_c[k] = g[5] * g[5];
             s->conic_b[k] = -2.0 * g[4];
             s->conic_a[k] = 1.0 - g[3];
             s->conic_xmin[k] = -0.5 * len;
             s->conic_xmax[k] = 0.5 * len;
             /* MARX_X = -OSAC_Z, MARX_Y = +OSAC_X, MARX_Z = -OSAC_Y */
             s->to_osac[k][0] = -(-g[2]);
             s->to_osac[k][1] = (-g[0]);
             s->to_osac[k][2] = -(-g[1]);
             s->to_osac[k][0] -= cap_pos;
             yaw_pitch_matrices (g[6], g[7], s->fwd[k], s->bwd[k]);
             s->tilt_deg[k][0] = g[6];
             s->tilt_deg[k][1] = g[7];
          }
        s->front_position = cap_pos - geom[i][0][2] + s->conic_xmax[0];
     }

   if (h->ideal == 0)
     {
        if (-1 == pf_get_file (pf, "HRMAOptConst", buf, sizeof (buf)))
          return -1;
        if (buf[0] != 0)
          {
             unsigned int offs[3];
             if (-1 == read_bdat_into_arena (buf, 3, &h->opt_n, offs))
               return -1;
             h->opt_e_off = offs[0];
             h->opt_beta_off = offs[1];
             h->opt_delta_off = offs[2];
          }
        if (h->use_scale)
          {
             for (i = 0; i < MXG_NUM_SHELLS; i++)
               {
                  unsigned int offs[2];
                  snprintf (buf, sizeof (buf), "hrma/corr_%u.dat", Shell_Numbers[i]);
                  if (-1 == read_bdat_into_arena (buf, 2, &h->shell[i].corr_n, offs))
                    return -1;
                  h->shell[i].corr_e_off = offs[0];
                  h->shell[i].corr_f_off = offs[1];
               }
          }
     }

   total_area = 0.0;
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        mxg_shell_cfg *s = h->shell + i;
        double a = s->conic_a[0], b = s->conic_b[0], c = s->conic_c[0];
        double x;

        if (h->use_wfold)
          {
             unsigned int k;
             for (k = 0; k < 2; k++)
               {
                  char name[64];
                  snprintf (name, sizeof (name), k ? "WFold_H%u_File" : "WFold_P%u_File",
                            Shell_Numbers[i]);
                  if (-1 == pf_get_file (pf, name, buf, sizeof (buf)))
                    return -1;
                  if (NULL == (file = mxg_data_file_name (buf)))
                    return -1;
                  mxg_message (2, "\t%s\n", file);
                  if (-1 == mxg_read_wfold (file, &s->wfold[k]))
                    {
                       free (file);
                       return -1;
                    }
                  free (file);
               }
          }

        /* back and front of the paraboloid bound the aperture */
        x = s->conic_xmin[0];
        s->min_radius = sqrt (c + x * (b + x * (a - 1)));
        x = s->conic_xmax[0];
        s->max_radius = sqrt (c + x * (b + x * (a - 1)));
        /* widen for rays up to ~40' off axis: dr = L tan(0.67 deg) */
        s->min_radius -= geom[i][0][8] * tan (0.67 * PI / 180);
        if (s->num_open_shutters)
          total_area += (s->num_open_shutters / 4.0)
            * (s->max_radius - s->min_radius) * (s->max_radius + s->min_radius);
        s->area_fraction = total_area;
     }
   if (total_area <= 0.0)
     {
        mxg_error ("The mirror geometric area is 0.  Check shutters.");
        return -1;
     }
   h->geometric_area = (total_area * PI) / 100.0;
   Marx_Mirror_Geometric_Area = h->geometric_area;
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     h->shell[i].area_fraction = h->shell[i].area_fraction / total_area;

   /* strut shadows (hrma.c:900-930): positions relative to the cap, mm */
   h->pre_struts[0].xpos = 1492.060;  h->pre_struts[0].half_width = 0.5 * 0.5 * 25.4;
   h->pre_struts[1].xpos = 942.266;   h->pre_struts[1].half_width = 0.5 * 0.5 * 25.4;
   h->cap_struts[0].xpos = 0.5 * 1.965 * 25.4;  h->cap_struts[0].half_width = 0.5 * 0.75 * 25.4;
   h->cap_struts[1].xpos = -0.5 * 1.965 * 25.4; h->cap_struts[1].half_width = 0.5 * 0.75 * 25.4;
   h->post_struts[0].xpos = -1050.353; h->post_struts[0].half_width = 0.5 * 0.5 * 25.4;
   h->post_struts[1].xpos = -1271.333; h->post_struts[1].half_width = 0.5 * 0.5 * 25.4;
   {
      volatile double theta = 30.0 * (PI / 180.0);
      h->strut_cos = cos (theta);
      h->strut_sin = sin (theta);
   }

   /* bucket indices of the tables the reflection kernel searches per photon */
   if (h->opt_n && (-1 == mxg_bucket_of_farena (h->opt_e_off, h->opt_n, MXG_BUCKETS_OPT, &h->opt_bucket)))
     return -1;
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        mxg_shell_cfg *s = h->shell + i;
        unsigned int k;
        if (s->corr_n && (-1 == mxg_bucket_of_farena (s->corr_e_off, s->corr_n, MXG_BUCKETS_CORR, &s->corr_bucket)))
          return -1;
        for (k = 0; k < 2; k++)
          if (-1 == bucket_of_e_alphas (&s->wfold[k], MXG_BUCKETS_EALPHA, &s->ea_bucket[k]))
            return -1;
     }

   h->enabled = 1;
   Mxg_Cfg_Generation++;
   return 0;
}
