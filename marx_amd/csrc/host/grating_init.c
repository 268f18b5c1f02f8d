/* grating_init.c -- HETG / LETG initialisation: efficiency tables,
 * dispersion geometry, facet-sector tables and the LETG support gratings.
 *
 * Follows marx/libsrc/diffract.c:
 *   get_pfile_parms (:1148-1167), grating_pre_init / create_opt_const_tables
 *   (:1169-1230), read_grating_optical_consts (:146-232), grating_efficiency
 *   (:269-307), init_one_grating (:371-452), create_order_list (:1232-1249),
 *   sum_file_efficiencies (:1252-1300), read_geff_caldb_file (:1343-1418),
 *   init_file_efficiencies (:1550-1625), read_sector_info (:1900-1948),
 *   _marx_hetg_init (:1687-1726), _marx_letg_init (:1728-1863).
 *
 * Cumulative efficiency tables are stored energy-major
 * ([num_energies][num_orders]) so that one photon's order walk reads two
 * contiguous rows.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "mxg_host.h"

typedef struct
{
   double period, bar_height, bar_width, polyimide, gold, chromium, nickel;
   double dp_over_p, dispersion_angle, theta_blur;
   int num_orders;
   double vig;
}
Grating_Info;

/* Optical-constant tables on the computed-efficiency energy grid */
static unsigned int Num_Energies = 1024;
static double Min_Energy = 0.03, Max_Energy = 10.0;
static float *Energies, *Beta_F, *Delta_F, *Poly_F, *Gold_F, *Cr_F, *Ni_F;

static void free_opt_tables (void)
{
   free (Energies); free (Beta_F); free (Delta_F);
   free (Poly_F); free (Gold_F); free (Cr_F); free (Ni_F);
   Energies = Beta_F = Delta_F = Poly_F = Gold_F = Cr_F = Ni_F = NULL;
}

/* Linear interpolation of the columns of a sorted table onto sorted x,
 * walking both in step (JDMinterpolate_n_fvector, finterpo.c:134-193). */
static void interpolate_columns (const float *xp, float **yp, unsigned int n,
                                 const float *oldxp, float **oldyp, unsigned int oldn,
                                 unsigned int n_yp)
{
   unsigned int count = 1, i, j;
   for (j = 0; j < n; j++)
     {
        double x = xp[j], x_0, x_1, dx_10;
        while ((count < oldn - 1) && (x > oldxp[count])) count++;
        x_0 = oldxp[count - 1];
        x_1 = oldxp[count];
        dx_10 = x_1 - x_0;
        for (i = 0; i < n_yp; i++)
          {
             if (dx_10 == 0.0)
               yp[i][j] = oldyp[i][count - 1];
             else
               {
                  double y_0 = oldyp[i][count - 1], y_1 = oldyp[i][count];
                  yp[i][j] = (float) (y_0 + (y_1 - y_0) * (x - x_0) / dx_10);
               }
          }
     }
}

static int create_opt_const_tables (Param_File_Type *pf)
{
   char buf[1024], *file;
   float *data[7], *dst[6];
   unsigned int nread, i;
   double energy, d_energy;

   free_opt_tables ();
   if (-1 == pf_get_file (pf, "GratingOptConsts", buf, sizeof (buf)))
     return -1;
   if (NULL == (file = mxg_data_file_name (buf)))
     return -1;
   mxg_message (2, "\t%s\n", file);
   if (-1 == mxg_read_bdat (file, 7, data, &nread))
     {
        free (file);
        return -1;
     }
   free (file);
   if (nread < 2)
     {
        mxg_error ("grating optical-constant table has fewer than 2 rows");
        return -1;
     }

   Energies = (float *) malloc (Num_Energies * sizeof (float));
   Beta_F = (float *) malloc (Num_Energies * sizeof (float));
   Delta_F = (float *) malloc (Num_Energies * sizeof (float));
   Gold_F = (float *) malloc (Num_Energies * sizeof (float));
   Poly_F = (float *) malloc (Num_Energies * sizeof (float));
   Cr_F = (float *) malloc (Num_Energies * sizeof (float));
   Ni_F = (float *) malloc (Num_Energies * sizeof (float));
   if (!Energies || !Beta_F || !Delta_F || !Gold_F || !Poly_F || !Cr_F || !Ni_F)
     {
        for (i = 0; i < 7; i++) free (data[i]);
        free_opt_tables ();
        return -1;
     }

   d_energy = (Max_Energy - Min_Energy) / Num_Energies;
   energy = Min_Energy;
   for (i = 0; i < Num_Energies; i++)
     {
        Energies[i] = (float) energy;
        energy += d_energy;
     }

   /* file columns: energy beta delta t-gold t-poly t-cr t-ni */
   dst[0] = Beta_F; dst[1] = Delta_F; dst[2] = Gold_F;
   dst[3] = Poly_F; dst[4] = Cr_F; dst[5] = Ni_F;
   interpolate_columns (Energies, dst, Num_Energies, data[0], data + 1, nread, 6);
   for (i = 0; i < 7; i++) free (data[i]);

   /* transmissions are per 0.001 um of material: keep 1000 log(t) */
   for (i = 0; i < Num_Energies; i++)
     {
        float *tabs[4];
        unsigned int k;
        tabs[0] = Gold_F; tabs[1] = Poly_F; tabs[2] = Cr_F; tabs[3] = Ni_F;
        for (k = 0; k < 4; k++)
          {
             double factor = tabs[k][i];
             if (factor <= 0.0) factor = 0.0;
             else factor = log (factor);
             tabs[k][i] = (float) (1000.0 * factor);
          }
     }
   return 0;
}

/* Efficiency of one order of a bar grating at grid energy n (diffract.c:269-321). */
static double grating_efficiency (int order, unsigned int n, const Grating_Info *g)
{
   double energy = Energies[n], beta = Beta_F[n], delta = Delta_F[n];
   double a, k, kh, khb, ekhb, eff, tfactor;

   a = 1.0 - g->bar_width / g->period;
   k = energy / MXG_HBAR_C;
   kh = g->bar_height * k;
   khb = kh * beta;
   ekhb = exp (-khb);
   if (order != 0)
     {
        double mpi = order * PI;
        a = sin (mpi * a) / mpi;
        eff = a * a * (1 + ekhb * (ekhb - 2.0 * cos (kh * delta)));
     }
   else
     {
        double b = 1 - a;
        eff = a * a + b * ekhb * (b * ekhb + 2 * a * cos (kh * delta));
     }
   tfactor = exp (g->polyimide * (double) Poly_F[n] + g->gold * (double) Gold_F[n]
                  + g->chromium * (double) Cr_F[n] + g->nickel * (double) Ni_F[n]);
   return eff * tfactor;
}

static int set_order_list (mxg_grating *gt, const int *orders)
{
   unsigned int k;
   float *o;
   unsigned int off = mxg_farena_reserve (gt->num_orders);
   if (off == (unsigned int) -1) return -1;
   o = mxg_farena_ptr (off);
   for (k = 0; k < gt->num_orders; k++) o[k] = (float) orders[k];
   gt->orders_off = off;
   return 0;
}

/* Grating whose efficiencies are computed on the optical-constant grid,
 * orders listed as 0, +1, -1, +2, -2, ... (diffract.c:371-452). */
static int init_computed_grating (const Grating_Info *info, int use_this_order, int use_unit_eff,
                                  mxg_grating *gt)
{
   unsigned int num_orders = (unsigned int) info->num_orders, i, k;
   int *orders;
   float *cum;

   if (use_this_order) num_orders = 2;
   if (num_orders == 0)
     {
        memset (gt, 0, sizeof (*gt));
        return 0;
     }
   orders = (int *) malloc (num_orders * sizeof (int));
   if (orders == NULL) return -1;
   if (use_this_order)
     {
        orders[0] = use_this_order;
        orders[1] = -use_this_order;
     }
   else
     {
        int j = 1;
        orders[0] = 0;
        for (i = 1; i + 1 < num_orders; i += 2)
          {
             orders[i] = j;
             orders[i + 1] = -j;
             j++;
          }
     }
   gt->num_orders = num_orders;
   gt->num_energies = Num_Energies;
   gt->dispersion_angle = info->dispersion_angle;
   gt->theta_blur = info->theta_blur;
   gt->period = info->period;
   gt->vig = info->vig;
   gt->dp_over_p = info->dp_over_p;
   if ((-1 == set_order_list (gt, orders))
       || ((unsigned int) -1 == (gt->energies_off = mxg_farena_push (Energies, Num_Energies)))
       || ((unsigned int) -1 == (gt->cum_eff_off = mxg_farena_reserve (Num_Energies * num_orders))))
     {
        free (orders);
        return -1;
     }
   cum = mxg_farena_ptr (gt->cum_eff_off);
   for (i = 0; i < Num_Energies; i++)
     {
        double sum = 0.0;
        for (k = 0; k < num_orders; k++)
          {
             if (use_unit_eff)
               sum += (1.0 / num_orders);
             else
               sum += grating_efficiency (orders[k], i, info);
             cum[i * num_orders + k] = (float) sum;
          }
     }
   free (orders);
   return 0;
}

typedef struct
{
   int shell;
}
Greff_Match;

static int match_greff (mxg_fits_table *t, void *client)
{
   char hdu[32];
   snprintf (hdu, sizeof (hdu), "AXAF_GREFF%d", ((Greff_Match *) client)->shell + 1);
   return mxg_fits_keyword_is (t, "HDUNAME", hdu);
}

/* Grating from a CALDB GREFF extension: orders -N..N, cumulative sums,
 * normalised where they exceed 1 (diffract.c:1252-1300, 1343-1418). */
static int init_file_grating (const char *caldb_obj, int shell, double scale, int use_this_order,
                              mxg_grating *gt)
{
   char *file = mxg_caldb_file (caldb_obj);
   mxg_fits_table *t;
   Greff_Match m;
   unsigned int ne, no, rep, i, j;
   char type;
   int max_order, *orders;
   float *en, *cum;

   if (file == NULL) return -1;
   m.shell = shell;
   t = mxg_fits_find_table (file, match_greff, &m);
   if (t == NULL)
     {
        mxg_error ("Unable to find an extension with HDUNAME=AXAF_GREFF%d in %s", shell + 1, file);
        free (file);
        return -1;
     }
   mxg_message (2, "\t%s[AXAF_GREFF%d]\n", file, shell + 1);
   free (file);
   if ((-1 == mxg_fits_column (t, "ENERGY", &rep, &type)) || (rep != 1))
     {
        mxg_error ("The ENERGY column must be a scalar column");
        mxg_fits_close (t);
        return -1;
     }
   if (-1 == mxg_fits_column (t, "EFF", &no, &type))
     {
        mxg_error ("GREFF table has no EFF column");
        mxg_fits_close (t);
        return -1;
     }
   ne = mxg_fits_num_rows (t);
   if ((ne == 0) || (no < 3) || ((no % 2) != 1))
     {
        mxg_error ("Grating efficiency file does not contain enough columns or the wrong number");
        mxg_fits_close (t);
        return -1;
     }
   max_order = (int) no / 2;
   gt->num_orders = 2 * max_order + 1;
   gt->num_energies = ne;
   gt->energies_off = mxg_farena_reserve (ne);
   gt->cum_eff_off = mxg_farena_reserve (ne * gt->num_orders);
   orders = (int *) malloc (gt->num_orders * sizeof (int));
   if ((gt->energies_off == (unsigned int) -1) || (gt->cum_eff_off == (unsigned int) -1)
       || (orders == NULL))
     {
        free (orders);
        mxg_fits_close (t);
        return -1;
     }
   for (i = 0; i < gt->num_orders; i++) orders[i] = (int) i - max_order;
   if (-1 == set_order_list (gt, orders))
     {
        free (orders);
        mxg_fits_close (t);
        return -1;
     }
   /* no arena growth below this point: the pointers stay valid */
   en = mxg_farena_ptr (gt->energies_off);
   cum = mxg_farena_ptr (gt->cum_eff_off);
   for (i = 0; i < ne; i++)
     {
        if (-1 == mxg_fits_get_float (t, "ENERGY", i, 0, en + i))
          {
             free (orders);
             mxg_fits_close (t);
             return -1;
          }
        for (j = 0; j < gt->num_orders; j++)
          if (-1 == mxg_fits_get_float (t, "EFF", i, j, cum + i * gt->num_orders + j))
            {
               free (orders);
               mxg_fits_close (t);
               return -1;
            }
     }
   mxg_fits_close (t);

   for (i = 0; i < ne; i++)
     {
        float *row = cum + i * gt->num_orders;
        double sum = 0.0;
        if (use_this_order > 0)
          {
             for (j = 0; j < gt->num_orders; j++)
               {
                  if (abs (orders[j]) == use_this_order)
                    sum += scale * (double) row[j];
                  row[j] = (float) sum;
               }
             if (sum > 0)
               for (j = 0; j < gt->num_orders; j++) row[j] = (float) (row[j] / sum);
             continue;
          }
        for (j = 0; j < gt->num_orders; j++)
          {
             sum += scale * (double) row[j];
             row[j] = (float) sum;
          }
        if (sum > 1.0)
          for (j = 0; j < gt->num_orders; j++) row[j] = (float) (row[j] / sum);
     }
   free (orders);
   return 0;
}

static int read_sectors (Param_File_Type *pf, const char *name, mxg_sectors *sec)
{
   static const char shells[] = "1346";
   unsigned int i, k;
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        char pname[64], buf[1024], *file;
        double *cols[6];
        unsigned int n;
        double *dst;

        snprintf (pname, sizeof (pname), "%s_Sector%c_File", name, shells[i]);
        if (-1 == pf_get_file (pf, pname, buf, sizeof (buf)))
          return -1;
        if (NULL == (file = mxg_data_file_name (buf)))
          return -1;
        mxg_message (2, "\t%s\n", file);
        if (-1 == mxg_read_columns (file, 6, cols, &n))
          {
             free (file);
             return -1;
          }
        free (file);
        sec[i].n = n;
        sec[i].off = mxg_darena_reserve (6 * n);
        if (sec[i].off == (unsigned int) -1) return -1;
        dst = mxg_darena_ptr (sec[i].off);
        for (k = 0; k < n; k++)
          {
             dst[k] = cols[0][k] * (PI / 180.0);
             dst[n + k] = cols[1][k] * (PI / 180.0);
             dst[2 * n + k] = cols[2][k] * (PI / (180.0 * 60.0));
             dst[3 * n + k] = cols[3][k] * (PI / (180.0 * 60.0));
             dst[4 * n + k] = cols[4][k];
             dst[5 * n + k] = cols[5][k];
          }
        for (k = 0; k < 6; k++) free (cols[k]);
     }
   return 0;
}

static int get_info (Param_File_Type *pf, const char *prefix, Grating_Info *g)
{
   char name[64];
#define GET_D(suffix, field) \
   snprintf (name, sizeof (name), "%s" suffix, prefix); \
   if (-1 == pf_get_double (pf, name, &g->field)) return -1;
   GET_D ("Gold", gold);
   GET_D ("BarHeight", bar_height);
   GET_D ("BarWidth", bar_width);
   GET_D ("Period", period);
   GET_D ("dPoverP", dp_over_p);
   GET_D ("Theta", dispersion_angle);
   GET_D ("dTheta", theta_blur);
#undef GET_D
   snprintf (name, sizeof (name), "%sNumOrders", prefix);
   if (-1 == pf_get_integer (pf, name, &g->num_orders)) return -1;
   g->dispersion_angle *= PI / 180.0;
   g->theta_blur *= (PI / 180.0) / 60.0;
   return 0;
}

int mxg_grating_init (Param_File_Type *pf)
{
   mxg_grating_cfg *c = &Mxg_Cfg.grating;
   static const char shells[] = "1346";
   char type[64];
   int is_letg, use_file, use_hetg_sect, use_letg_sect, use_unit, use_this_order;
   double meg_d, heg_d, leg_d;
   Grating_Info heg, meg, leg, fine, coarse;
   unsigned int i;

   memset (c, 0, sizeof (*c));
   memset (&heg, 0, sizeof (heg)); memset (&meg, 0, sizeof (meg)); memset (&leg, 0, sizeof (leg));
   memset (&fine, 0, sizeof (fine)); memset (&coarse, 0, sizeof (coarse));

   if (-1 == pf_get_string (pf, "GratingType", type, sizeof (type)))
     return -1;
   is_letg = !strcmp (type, "LETG");

   if ((-1 == pf_get_boolean (pf, "UseGratingEffFiles", &use_file))
       || (-1 == pf_get_boolean (pf, "Use_HETG_Sector_Files", &use_hetg_sect))
       || (-1 == pf_get_boolean (pf, "Use_LETG_Sector_Files", &use_letg_sect))
       || (-1 == pf_get_boolean (pf, "Use_Unit_Efficiencies", &use_unit))
       || (-1 == pf_get_integer (pf, "Use_This_Order", &use_this_order))
       || (-1 == pf_get_double (pf, "MEGRowlandDiameter", &meg_d))
       || (-1 == pf_get_double (pf, "HEGRowlandDiameter", &heg_d))
       || (-1 == pf_get_double (pf, "LEGRowlandDiameter", &leg_d))
       || (-1 == pf_get_double (pf, "HEGVig", &heg.vig))
       || (-1 == pf_get_double (pf, "MEGVig", &meg.vig))
       || (-1 == pf_get_double (pf, "LEGVig", &leg.vig))
       || (-1 == get_info (pf, "heg", &heg))
       || (-1 == get_info (pf, "meg", &meg))
       || (-1 == get_info (pf, "leg", &leg))
       || (-1 == pf_get_double (pf, "hegNickel", &heg.nickel))
       || (-1 == pf_get_double (pf, "hegChromium", &heg.chromium))
       || (-1 == pf_get_double (pf, "hegPolyimide", &heg.polyimide))
       || (-1 == pf_get_double (pf, "megNickel", &meg.nickel))
       || (-1 == pf_get_double (pf, "megChromium", &meg.chromium))
       || (-1 == pf_get_double (pf, "megPolyimide", &meg.polyimide))
       || (-1 == pf_get_integer (pf, "legFineNumOrders", &fine.num_orders))
       || (-1 == pf_get_integer (pf, "legCoarseNumOrders", &coarse.num_orders))
       || (-1 == pf_get_double (pf, "legFineBarWidth", &fine.bar_width))
       || (-1 == pf_get_double (pf, "legFineBarHeight", &fine.bar_height))
       || (-1 == pf_get_double (pf, "legFinePeriod", &fine.period))
       || (-1 == pf_get_double (pf, "legCoarseBarWidth", &coarse.bar_width))
       || (-1 == pf_get_double (pf, "legCoarseBarHeight", &coarse.bar_height))
       || (-1 == pf_get_double (pf, "legCoarsePeriod", &coarse.period)))
     return -1;
   if (use_unit) use_file = 0;

   if ((use_file == 0) || is_letg)
     if (-1 == create_opt_const_tables (pf))
       return -1;

   if (is_letg)
     for (i = 0; i < MXG_NUM_SHELLS; i++) c->rowland[i] = leg_d;
   else
     {
        c->rowland[0] = c->rowland[1] = meg_d;
        c->rowland[2] = c->rowland[3] = heg_d;
     }

   if ((is_letg && use_letg_sect) || (!is_letg && use_hetg_sect))
     if (-1 == read_sectors (pf, is_letg ? "LETG" : "HETG", c->sectors))
       return -1;

   if (use_file)
     {
        const char *prefix = is_letg ? "LETG" : "HETG";
        for (i = 0; i < MXG_NUM_SHELLS; i++)
          {
             char pname[64];
             double vig, theta, dtheta, period, dpop;
             mxg_grating *gt = c->g + i;
#define GET_SHELL(fmt, v) snprintf (pname, sizeof (pname), fmt, prefix, shells[i]); \
             if (-1 == pf_get_double (pf, pname, &v)) return -1;
             GET_SHELL ("%s_Shell%c_Vig", vig);
             GET_SHELL ("%s_Shell%c_Theta", theta);
             GET_SHELL ("%s_Shell%c_dTheta", dtheta);
             GET_SHELL ("%s_Shell%c_Period", period);
             GET_SHELL ("%s_Shell%c_dPoverP", dpop);
#undef GET_SHELL
             if (-1 == init_file_grating (is_letg ? "LETGEFF" : "HETGEFF", (int) i, 1.0,
                                          use_this_order, gt))
               return -1;
             gt->dispersion_angle = (theta * (PI / 180.0));
             gt->theta_blur = (dtheta * (PI / 180.0) / 60.0);
             gt->period = period;
             gt->dp_over_p = dpop;
             gt->vig = vig;
          }
     }
   else if (is_letg)
     {
        mxg_grating g;
        if (-1 == init_computed_grating (&leg, use_this_order, use_unit, &g)) return -1;
        for (i = 0; i < MXG_NUM_SHELLS; i++) c->g[i] = g;
     }
   else
     {
        mxg_grating gh, gm;
        if ((-1 == init_computed_grating (&heg, use_this_order, use_unit, &gh))
            || (-1 == init_computed_grating (&meg, use_this_order, use_unit, &gm)))
          return -1;
        c->g[0] = c->g[1] = gm;
        c->g[2] = c->g[3] = gh;
     }

   if (is_letg)
     {
        char det[64];
        if (-1 == pf_get_string (pf, "DetectorType", det, sizeof (det)))
          return -1;
        if (0 == strncmp (det, "ACIS", 4))
          {
             /* aspect / focal-plane misalignment hack (diffract.c:1775-1789) */
             double dth;
             if (-1 == pf_get_double (pf, "LETG_ACIS_dTheta", &dth)) return -1;
             dth *= PI / 180.0;
             for (i = 0; i < MXG_NUM_SHELLS; i++)
               c->g[i].dispersion_angle += dth;
          }
        if ((fine.num_orders && (-1 == init_computed_grating (&fine, 0, 0, &c->fine)))
            || (coarse.num_orders && (-1 == init_computed_grating (&coarse, 0, 0, &c->coarse))))
          return -1;
        c->has_fine = (c->fine.num_orders != 0);
        c->has_coarse = (c->coarse.num_orders != 0);

        if (use_file || use_unit)
          {
             /* divide out the support-grating throughput that the tabulated
              * primary efficiencies already include (diffract.c:1790-1861) */
             for (i = 0; i < MXG_NUM_SHELLS; i++)
               {
                  mxg_grating *gt = c->g + i;
                  const float *en = mxg_farena_ptr (gt->energies_off);
                  float *cum = mxg_farena_ptr (gt->cum_eff_off);
                  unsigned int j, k;
                  if (i > 0 && (gt->cum_eff_off == c->g[i - 1].cum_eff_off))
                    continue;      /* shared table already corrected */
                  if (c->has_fine)
                    {
                       const float *se = mxg_farena_ptr (c->fine.energies_off);
                       const float *sf = mxg_farena_ptr (c->fine.cum_eff_off);
                       unsigned int no = c->fine.num_orders;
                       float *tot = (float *) malloc (c->fine.num_energies * sizeof (float));
                       if (tot == NULL) return -1;
                       for (j = 0; j < c->fine.num_energies; j++) tot[j] = sf[j * no + no - 1];
                       for (j = 0; j < gt->num_energies; j++)
                         {
                            float factor = mxg_interpolate_f (en[j], se, tot, c->fine.num_energies);
                            factor = (float) (1.0 / factor);
                            if (factor <= 0) continue;
                            for (k = 0; k < gt->num_orders; k++)
                              cum[j * gt->num_orders + k] *= factor;
                         }
                       free (tot);
                    }
                  if (c->has_coarse)
                    {
                       const float *se = mxg_farena_ptr (c->coarse.energies_off);
                       const float *sf = mxg_farena_ptr (c->coarse.cum_eff_off);
                       unsigned int no = c->coarse.num_orders;
                       float *tot = (float *) malloc (c->coarse.num_energies * sizeof (float));
                       if (tot == NULL) return -1;
                       for (j = 0; j < c->coarse.num_energies; j++) tot[j] = sf[j * no + no - 1];
                       for (j = 0; j < gt->num_energies; j++)
                         {
                            float factor = mxg_interpolate_f (en[j], se, tot, c->coarse.num_energies);
                            factor = factor * factor * factor;
                            if (factor <= 0) continue;
                            factor = (float) (1.0 / factor);
                            for (k = 0; k < gt->num_orders; k++)
                              cum[j * gt->num_orders + k] *= factor;
                         }
                       free (tot);
                    }
               }
          }
     }

   /* the rotation into and out of each grating's frame, the same for every
    * photon of a shell: computed once (diffract.c:974-980 rotates by
    * -/+ theta; sine is odd and cosine even, so these give both) */
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        c->g[i].disp_cos = cos (c->g[i].dispersion_angle);
        c->g[i].disp_sin = sin (c->g[i].dispersion_angle);
     }
   /* bucket indices of the energy grids searched per photon, and whether
    * each energy's cumulative efficiencies rise with the order index (the
    * order draw may then search the orders by bisection) */
   for (i = 0; i < MXG_NUM_SHELLS + 2; i++)
     {
        mxg_grating *gt = (i < MXG_NUM_SHELLS) ? &c->g[i] : ((i == MXG_NUM_SHELLS) ? &c->fine : &c->coarse);
        unsigned int j, k;
        const float *cum;
        if ((gt->num_energies >= 2)
            && (-1 == mxg_bucket_of_farena (gt->energies_off, gt->num_energies, MXG_BUCKETS_GRATING, &gt->e_bucket)))
          return -1;
        gt->cum_sorted = (gt->num_orders > 0);
        cum = mxg_farena_ptr (gt->cum_eff_off);
        for (j = 0; gt->cum_sorted && (j < gt->num_energies); j++)
          for (k = 1; k < gt->num_orders; k++)
            if (!(cum[j * gt->num_orders + k] >= cum[j * gt->num_orders + k - 1]))
              {
                 gt->cum_sorted = 0;
                 break;
              }
     }
   c->type = is_letg ? MARX_GRATING_LETG : MARX_GRATING_HETG;
   Mxg_Cfg_Generation++;
   return c->type;
}
