/* orc_grating.c -- oracle restatement of the HETG / LETG diffraction pass.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Follows marx/libsrc/diffract.c:
 *   diffract (:974-1130): vignetting pass, prune, rotate into the grating
 *     frame, torus intersection, shell-major order selection over the
 *     energy-sorted photons with JDMinterpolate_n_fvector, LETG support
 *     gratings, rotate back;
 *   newtons_quartic / intersect / rotate (:622-700);
 *   diffract_photon (:706-823), diffract_photon_from_grating (:828-849),
 *   diffract_from_support_grating (:917-970).
 * Efficiency tables are read from the library's energy-major arena and
 * transposed to the reference's [order][energy] layout first.
 */
#include <stdlib.h>
#include <math.h>

#include "oracle_int.h"

typedef struct
{
   const mxg_grating *g;
   float **cum;          /* [order][energy] */
   float *energies;
   int *orders;
}
Orc_Grating;

static int load_grating (const orc_ctx *c, const mxg_grating *g, Orc_Grating *og)
{
   unsigned int k, j;
   og->g = g;
   og->cum = (float **) calloc (g->num_orders ? g->num_orders : 1, sizeof (float *));
   og->orders = (int *) malloc ((g->num_orders ? g->num_orders : 1) * sizeof (int));
   og->energies = (float *) (c->fa + g->energies_off);
   if ((og->cum == NULL) || (og->orders == NULL)) return -1;
   for (k = 0; k < g->num_orders; k++)
     {
        og->cum[k] = (float *) malloc (g->num_energies * sizeof (float));
        if (og->cum[k] == NULL) return -1;
        for (j = 0; j < g->num_energies; j++)
          og->cum[k][j] = c->fa[g->cum_eff_off + j * g->num_orders + k];
        og->orders[k] = (int) c->fa[g->orders_off + k];
     }
   return 0;
}

static void free_grating (Orc_Grating *og)
{
   unsigned int k;
   if (og->cum != NULL)
     for (k = 0; k < og->g->num_orders; k++) free (og->cum[k]);
   free (og->cum);
   free (og->orders);
}

static int newtons_quartic (double a, double b, double c, double d, double t0, double *tp)
{
   unsigned int max_it = 10;
   double a2 = 2.0 * a, a3 = 3.0 * a, b2 = 2.0 * b;
   double t2, t, eps = 1.0e-4, num, den;

   while (1)
     {
        t2 = t0 * t0;
        num = t2 * (3.0 * t2 + a2 * t0 + b) - d;
        den = t2 * (4.0 * t0 + a3) + b2 * t0 + c;
        t = num / den;
        if (fabs (t - t0) < eps) break;
        max_it--;
        if (max_it == 0) return -1;
        t0 = t;
     }
   *tp = t;
   return 0;
}

static int intersect (orc_vec *x0, orc_vec p, double rowland)
{
   double a, b, c, d, t0, t, x2, pdotx, pxpz_len, r2;

   t = -x0->x / p.x;
   x0->x = 0.0;
   x0->y = x0->y + p.y * t;
   x0->z = x0->z + p.z * t;

   pxpz_len = p.x * p.x + p.z * p.z;
   x2 = x0->z * x0->z + x0->y * x0->y;
   r2 = rowland * rowland;
   pdotx = orc_v_dot (p, *x0);

   a = 4.0 * pdotx;
   b = 2.0 * x2 + a * pdotx - r2 * pxpz_len;
   c = a * x2 - 2.0 * r2 * p.z * x0->z;
   d = x2 * x2 - r2 * x0->z * x0->z;
   t0 = -rowland * sqrt (pxpz_len);

   if (-1 == newtons_quartic (a, b, c, d, t0, &t))
     return -1;
   *x0 = orc_v_ax1_bx2 (1.0, *x0, t, p);
   return 0;
}

static orc_vec rotate_x (orc_vec a, double theta)
{
   double c = cos (theta), s = sin (theta), ay = a.y, az = a.z;
   a.y = c * ay - s * az;
   a.z = s * ay + c * az;
   return a;
}

static int diffract_photon (const Orc_Grating *og, double theta, Marx_Photon_Attr_Type *at,
                            int order, const orc_ctx *c, const mxg_sectors *gs, orc_stream *st)
{
   const mxg_grating *g = og->g;
   orc_vec p, x, n, l, d, dp;
   double factor, n_lambda_over_d, dp_over_p, dtheta, p_n, p_d, p_l;

   n_lambda_over_d = order * (2.0 * PI * ORC_HBAR_C) / g->period / at->energy;
   x = at->x;
   p = at->p;

   n = orc_v_unit (x);
   n.x = -n.x; n.y = -n.y; n.z = -n.z;
   l.x = n.z;
   l.y = 0.0;
   l.z = -n.x;
   l = orc_v_unit (l);
   d = orc_v_cross (n, l);

   if ((gs != NULL) && (gs->n > 0))
     {
        const double *min_a = c->da + gs->off, *max_a = min_a + gs->n;
        const double *dth = min_a + 2 * gs->n, *dthb = min_a + 3 * gs->n;
        const double *dpp = min_a + 4 * gs->n, *dppb = min_a + 5 * gs->n;
        unsigned int sn;
        double sector = atan2 (x.y, x.z);
        if (sector < 0) sector = 2 * PI + sector;
        sn = orc_binary_search_d (sector, min_a, gs->n) - 1;
        if (sn >= gs->n) return -1;
        if ((max_a[sn] <= sector) || (min_a[sn] > sector))
          return -1;
        dtheta = dth[sn] + dthb[sn] * orc_gaussian (st);
        dp_over_p = dpp[sn] + dppb[sn] * orc_gaussian (st);
     }
   else
     {
        dtheta = g->theta_blur * orc_gaussian (st);
        dp_over_p = g->dp_over_p * orc_gaussian (st);
     }

   theta -= dtheta;
   if (theta != 0.0)
     {
        orc_vec l_tmp = l, d_tmp = d;
        double cs = cos (theta), sn = sin (theta);
        l = orc_v_ax1_bx2 (cs, l_tmp, sn, d_tmp);
        d = orc_v_ax1_bx2 (-sn, l_tmp, cs, d_tmp);
     }

   p_d = n_lambda_over_d + orc_v_dot (p, d);
   p_l = orc_v_dot (p, l);
   p_n = 1.0 - p_l * p_l - p_d * p_d;
   if (p_n < 0.0) return -1;
   p_n = sqrt (p_n);

   p = orc_v_ax1_bx2 (p_d, d, p_n, n);
   at->p = orc_v_ax1_bx2 (1.0, p, p_l, l);

   if (dp_over_p == 0) return 0;
   factor = n_lambda_over_d * dp_over_p;
   dp = orc_v_ax1_bx2 (-factor, d, factor * (p_d / p_n), n);
   at->p.x += dp.x;
   at->p.y += dp.y;
   at->p.z += dp.z;
   at->p = orc_v_unit (at->p);
   return 0;
}

static int diffract_photon_from_grating (const Orc_Grating *og, double theta,
                                         Marx_Photon_Attr_Type *at, float **cum_eff,
                                         unsigned int photon_number, signed char *order,
                                         const orc_ctx *c, const mxg_sectors *gs, orc_stream *st)
{
   double r = orc_uniform (st);
   unsigned int k;
   for (k = 0; k < og->g->num_orders; k++)
     if (r <= cum_eff[k][photon_number])
       {
          *order = (signed char) og->orders[k];
          return diffract_photon (og, theta, at, *order, c, gs, st);
       }
   return -1;
}

static void rotate_photons (const orc_ctx *c, Marx_Photon_Type *pt, int dir)
{
   unsigned int i;
   for (i = 0; i < pt->num_sorted; i++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[i];
        double theta = dir * c->cfg->grating.g[at->mirror_shell].dispersion_angle;
        at->x = rotate_x (at->x, theta);
        at->p = rotate_x (at->p, theta);
     }
}

static int support_pass (const orc_ctx *c, Marx_Photon_Type *pt, const Orc_Grating *og,
                         double theta, int support_index, orc_stream *streams)
{
   unsigned int n = pt->num_sorted, j, num = 0, k;
   float *tmp_e = (float *) malloc ((n ? n : 1) * sizeof (float));
   float **tmp_cum = (float **) calloc (og->g->num_orders, sizeof (float *));
   int ret = 0;

   if ((tmp_e == NULL) || (tmp_cum == NULL)) { free (tmp_e); free (tmp_cum); return -1; }
   for (j = 0; j < n; j++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[j];
        if (at->flags & BAD_PHOTON_MASK) continue;
        tmp_e[num++] = (float) at->energy;
     }
   if (num == 0) { free (tmp_e); free (tmp_cum); return 0; }
   for (k = 0; k < og->g->num_orders; k++)
     if (NULL == (tmp_cum[k] = (float *) malloc (num * sizeof (float)))) { ret = -1; goto done; }
   orc_interpolate_n_fvector (tmp_e, tmp_cum, num, og->energies, og->cum, og->g->num_energies,
                              og->g->num_orders);
   num = 0;
   for (j = 0; j < n; j++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[j];
        orc_stream *st;
        if (at->flags & BAD_PHOTON_MASK) continue;
        st = orc_photon_stream (c, streams, pt, at, MXG_STAGE_GRATING);
        if (-1 == diffract_photon_from_grating (og, theta, at, tmp_cum, num,
                                                at->support_orders + support_index, c, NULL, st))
          at->flags |= PHOTON_UNDIFFRACTED;
        num++;
     }
done:
   for (k = 0; k < og->g->num_orders; k++) free (tmp_cum[k]);
   free (tmp_cum);
   free (tmp_e);
   return ret;
}

int orc_grating_diffract (const orc_ctx *c, Marx_Photon_Type *pt)
{
   const mxg_grating_cfg *gc = &c->cfg->grating;
   Orc_Grating og[MXG_NUM_SHELLS], fine, coarse;
   orc_stream *streams;
   unsigned int i, n;
   float *tmp_e;
   int ret = 0;

   if (gc->type == 0) return 0;
   if (pt->history & MARX_ORDER_OK) return 0;
   pt->history |= MARX_ORDER_OK;

   if (NULL == (streams = orc_streams_alloc (pt->n_photons))) return -1;
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     if (-1 == load_grating (c, gc->g + i, og + i)) return -1;
   if (gc->has_fine && (-1 == load_grating (c, &gc->fine, &fine))) return -1;
   if (gc->has_coarse && (-1 == load_grating (c, &gc->coarse, &coarse))) return -1;

   orc_prune_photons (pt);
   n = pt->num_sorted;
   for (i = 0; i < n; i++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[i];
        orc_stream *st = orc_photon_stream (c, streams, pt, at, MXG_STAGE_GRATING);
        if (orc_uniform (st) > gc->g[at->mirror_shell].vig)
          at->flags |= PHOTON_MIRROR_VBLOCKED;
     }
   orc_prune_photons (pt);
   n = pt->num_sorted;

   rotate_photons (c, pt, -1);
   for (i = 0; i < n; i++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[i];
        if (-1 == intersect (&at->x, at->p, gc->rowland[at->mirror_shell]))
          at->flags |= PHOTON_UNDIFFRACTED;
     }

   tmp_e = (float *) malloc ((n ? n : 1) * sizeof (float));
   if (tmp_e == NULL) return -1;

   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        const Orc_Grating *g = og + i;
        unsigned int j, k, num = 0;
        float **tmp_cum;

        for (j = 0; j < n; j++)
          {
             Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[j];
             if (at->mirror_shell != i) continue;
             tmp_e[num++] = (float) at->energy;
          }
        if (num == 0) continue;
        tmp_cum = (float **) calloc (g->g->num_orders, sizeof (float *));
        for (k = 0; k < g->g->num_orders; k++)
          tmp_cum[k] = (float *) malloc (num * sizeof (float));
        orc_interpolate_n_fvector (tmp_e, tmp_cum, num, g->energies, g->cum,
                                   g->g->num_energies, g->g->num_orders);
        num = 0;
        for (j = 0; j < n; j++)
          {
             Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[j];
             orc_stream *st;
             if (at->mirror_shell != i) continue;
             st = orc_photon_stream (c, streams, pt, at, MXG_STAGE_GRATING);
             if (-1 == diffract_photon_from_grating (g, 0.0, at, tmp_cum, num, &at->order, c,
                                                     gc->sectors + i, st))
               at->flags |= PHOTON_UNDIFFRACTED;
             num++;
          }
        for (k = 0; k < g->g->num_orders; k++) free (tmp_cum[k]);
        free (tmp_cum);
     }
   free (tmp_e);

   if (gc->type == MARX_GRATING_LETG)
     {
        if (gc->has_fine)
          {
             ret |= support_pass (c, pt, &fine, PI / 2.0, 0, streams);
             pt->history |= MARX_ORDER1_OK;
          }
        if (gc->has_coarse)
          {
             ret |= support_pass (c, pt, &coarse, PI / 3.0, 1, streams);
             ret |= support_pass (c, pt, &coarse, 2.0 * PI / 3.0, 2, streams);
             ret |= support_pass (c, pt, &coarse, 0.0, 3, streams);
             pt->history |= (MARX_ORDER2_OK | MARX_ORDER3_OK | MARX_ORDER4_OK);
          }
     }

   rotate_photons (c, pt, 1);

   for (i = 0; i < MXG_NUM_SHELLS; i++) free_grating (og + i);
   if (gc->has_fine) free_grating (&fine);
   if (gc->has_coarse) free_grating (&coarse);
   free (streams);
   return ret ? -1 : 0;
}
