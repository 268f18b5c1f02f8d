/* orc_detect.c -- oracle restatement of the detection pass.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Follows
 *   marx/libsrc/detector.c:56-168     plane intersection, first-hit chip search
 *   marx/libsrc/trans.c:40-88         detector transform and its inverse
 *   marx/libsrc/acis-s.c:138-248      ACIS-S detect, QE draw
 *   marx/libsrc/acis-i.c:56-167       ACIS-I detect, frame-transfer streak
 *   marx/libsrc/aciscontam.c:93-185   ACIS contamination transmission
 *   marx/libsrc/hrc-s.c:136-312       HRC-S detect, UV/ion-shield regions
 *   marx/libsrc/hrc-i.c:66-185        HRC-I detect, HRC pulse height
 *   marx/libsrc/hrcblur.c:239-297     MCP position blur
 *   marx/libsrc/hrc_s_geom.c:344-394, hrc_i_geom.c:146-156  pixel mapping
 *   marx/libsrc/drake.c:268-369       HESF pass before HRC-S detection
 *   marx/libsrc/reflect.c:79-91       marx_interp_reflectivity
 *   marx/libsrc/acis_fef.c:380-579    ACIS FEF pulse height and PI
 * Contamination and FEF apply when the set-up read their files (the CALDB
 * ones are absent from the reference's data tree; tests write synthetic
 * files).
 */
#include <stdlib.h>
#include <math.h>

#include "oracle_int.h"

static void rotate_vector (const double *m, orc_vec *v)
{
   double x = v->x, y = v->y, z = v->z;
   v->x = m[0] * x + m[1] * y + m[2] * z;
   v->y = m[3] * x + m[4] * y + m[5] * z;
   v->z = m[6] * x + m[7] * y + m[8] * z;
}

static void rotate_vector_inv (const double *m, orc_vec *v)
{
   double x = v->x, y = v->y, z = v->z;
   v->x = m[0] * x + m[3] * y + m[6] * z;
   v->y = m[1] * x + m[4] * y + m[7] * z;
   v->z = m[2] * x + m[5] * y + m[8] * z;
}

static void transform_ray (const mxg_detector_cfg *d, orc_vec *x, orc_vec *p)
{
   x->x -= d->offset[0];
   x->y -= d->offset[1];
   x->z -= d->offset[2];
   rotate_vector (d->matrix, x);
   rotate_vector (d->matrix, p);
}

static void transform_ray_reverse (const mxg_detector_cfg *d, orc_vec *x, orc_vec *p)
{
   rotate_vector_inv (d->matrix, p);
   rotate_vector_inv (d->matrix, x);
   x->x += d->offset[0];
   x->y += d->offset[1];
   x->z += d->offset[2];
}

static orc_vec v3 (const double *a)
{
   return orc_v (a[0], a[1], a[2]);
}

static int plane_hit (const mxg_chip *g, orc_vec x0, orc_vec p, orc_vec *x,
                      double *dx, double *dy, int must_hit)
{
   orc_vec normal = v3 (g->normal), r, xll = v3 (g->x_ll);
   double pdotn, rx, ry;
   int hit = 1;

   pdotn = orc_v_dot (p, normal);
   if (pdotn == 0) return -1;
   r = orc_v_diff (x0, xll);
   r = orc_v_ax1_bx2 (1.0, r, -1.0 * orc_v_dot (r, normal) / pdotn, p);
   rx = orc_v_dot (r, v3 (g->xhat));
   if ((rx < 0.0) || (rx >= g->xlen))
     {
        if (must_hit) return 0;
        hit = 0;
     }
   ry = orc_v_dot (r, v3 (g->yhat));
   if ((ry < 0.0) || (ry >= g->ylen))
     {
        if (must_hit) return 0;
        hit = 0;
     }
   *x = orc_v_sum (r, xll);
   *dx = rx;
   *dy = ry;
   return hit;
}

static const mxg_chip *intersect_with_detector (const mxg_detector_cfg *d, orc_vec x0, orc_vec p,
                                                orc_vec *xp, double *dxp, double *dyp)
{
   const mxg_chip *best = NULL;
   double best_dx = 0, best_dy = 0, best_r2 = -1;
   orc_vec best_x = x0;
   int k;

   for (k = 0; k < d->num_chips; k++)
     if (1 == plane_hit (d->chip + k, x0, p, xp, dxp, dyp, 1))
       return d->chip + k;
   if (d->extend == 0) return NULL;
   for (k = 0; k < d->num_chips; k++)
     {
        orc_vec x;
        double dx, dy, ddx, ddy, r2;
        const mxg_chip *g = d->chip + k;
        if (-1 == plane_hit (g, x0, p, &x, &dx, &dy, 0)) continue;
        ddx = dx - 0.5 * g->xlen;
        ddy = dy - 0.5 * g->ylen;
        r2 = ddx * ddx + ddy * ddy;
        if ((r2 < best_r2) || (best == NULL))
          {
             best = g; best_x = x; best_dx = dx; best_dy = dy; best_r2 = r2;
          }
     }
   if (best == NULL) return NULL;
   *xp = best_x;
   *dxp = best_dx;
   *dyp = best_dy;
   return best;
}

/* ---- ACIS ------------------------------------------------------------------------ */

/* ---- ACIS FEF (acis_fef.c:380-503, 509-579, 910-1079) ---------------------------- */

typedef struct
{
   float amp, center, sigma, cum_area;
   int use_tail_dist;
}
fef_gauss;

static double fef_integral (double xmin, double xmax, const fef_gauss *g)
{
   double sigma, x0;
   if (0.0 == (sigma = g->sigma * 1.4142135623730951)) return 0.0;
   x0 = g->center;
   return 0.5 * g->amp * (erf ((xmax - x0) / sigma) - erf ((xmin - x0) / sigma)) * (2.5066282746310002 * g->sigma);
}

static double fef_evaluate (double x, const fef_gauss *g)
{
   double sigma = g->sigma;
   if (sigma == 0.0) return 0.0;
   x = (x - g->center) / sigma;
   return g->amp * exp (-0.5 * x * x);
}

/* compute_pha_with_pos_amps, its unbounded outer loop bounded */
static int fef_pos_pha (const fef_gauss *gs, unsigned int ng, orc_stream *st, double *phap)
{
   unsigned int it;
   for (it = 0; it < 10000; it++)
     {
        double r = orc_uniform (st);
        const fef_gauss *g = gs, *gmax = gs + ng;
        while (g < gmax)
          {
             double pha;
             if (g->cum_area <= r)
               {
                  g++;
                  continue;
               }
             if (g->use_tail_dist == 0)
               {
                  unsigned int count = 0;
                  do
                    {
                       pha = g->center + g->sigma * orc_gaussian (st);
                       count++;
                    }
                  while ((pha < 0) && (count < 100));
               }
             else
               {
                  double u, v, x;
                  double s = (0 - g->center) / g->sigma;
                  unsigned int guard = 0;
                  do
                    {
                       u = orc_uniform (st);
                       do v = orc_uniform (st);
                       while (v == 0.0);
                       x = sqrt (s * s - 2 * log (v));
                    }
                  while ((x * u > s) && (++guard < 10000));
                  pha = g->center + x * g->sigma;
               }
             if (pha < 0) break;
             *phap = pha;
             return 0;
          }
     }
   return -1;
}

/* marx_apply_acis_rmf: 0 with *phap and *pip, or -1 (the photon is undetected) */
static int apply_fef (const orc_ctx *c, const mxg_detector_cfg *d, int ccd, float x, float y, double energy,
                      orc_stream *st, short *phap, float *pip)
{
   fef_gauss gs[16];
   const double *desc;
   const float *e, *ch, *g0, *g1;
   unsigned int i, j, n, ne, ng, k;
   double t, kf, total_pos = 0.0, total_neg = 0.0, pha;

   if (x < 0) { if (d->extend == 0) return -1; x = 0; }
   else if (x >= 1024) { if (d->extend == 0) return -1; x = 1023; }
   if (y < 0) { if (d->extend == 0) return -1; y = 0; }
   else if (y >= 1024) { if (d->extend == 0) return -1; y = 1023; }
   i = (unsigned int) (x / 32);
   j = (unsigned int) (y / 32);
   if ((i >= 32) || (j >= 32) || (ccd < 0) || (ccd >= 10)) return -1;
   kf = c->da[d->fef_map_off + (ccd * 32 + i) * 32 + j];
   if (kf < 0) return -1;
   desc = c->da + d->fef_desc_off + 5 * (unsigned int) kf;
   ne = (unsigned int) desc[0];
   ng = (unsigned int) desc[1];
   e = c->fa + (unsigned int) desc[2];
   ch = c->fa + (unsigned int) desc[3];
   n = orc_binary_search_f ((float) energy, e, ne);
   if (n == 0) n++;
   if (n == ne) n--;
   t = (energy - e[n - 1]) / (e[n] - e[n - 1]);
   g0 = c->fa + (unsigned int) desc[4] + 3 * (n - 1) * ng;
   g1 = g0 + 3 * ng;
   for (k = 0; k < ng; k++)
     {
        double v, area1, area2;
        fef_gauss *g = gs + k;
        g->center = (float) (g0[3 * k + 1] + t * (g1[3 * k + 1] - g0[3 * k + 1]));
        v = g0[3 * k + 2] + t * (g1[3 * k + 2] - g0[3 * k + 2]);
        if (v <= 0.0)
          {
             g->sigma = 0.0f;
             g->amp = 0.0f;
          }
        else
          {
             g->sigma = (float) v;
             v = g0[3 * k] + t * (g1[3 * k] - g0[3 * k]);
             if ((v < 0.0) && ((g1[3 * k] > 0.0f) || (g0[3 * k] > 0.0f))) v = 0.0;
             g->amp = (float) v;
          }
        /* normalize_gaussians */
        area1 = fef_integral (0, 1e37, g);
        area2 = fef_integral (-1e37, 0, g);
        g->use_tail_dist = (area2 > area1) && (area1 / area2 < 0.1);
        if (area1 >= 0) total_pos += area1;
        else total_neg -= area1;
        g->cum_area = (float) total_pos;
     }
   if (total_pos <= total_neg) return -1;
   if (total_pos > 0)
     for (k = 0; k < ng; k++) gs[k].cum_area = (float) (gs[k].cum_area / total_pos);
   if (total_neg == 0.0)
     {
        if (-1 == fef_pos_pha (gs, ng, st, &pha)) return -1;
     }
   else
     {
        /* compute_pha_with_neg_amps */
        int count = 0, ok = 0;
        while ((count < 100) && !ok)
          {
             double pos_sum = 0.0, sum = 0.0;
             if (-1 == fef_pos_pha (gs, ng, st, &pha)) return -1;
             for (k = 0; k < ng; k++)
               {
                  double ds = fef_evaluate (pha, gs + k);
                  sum += ds;
                  if (ds > 0) pos_sum += ds;
               }
             ok = (orc_uniform (st) * pos_sum < sum);
             count++;
          }
        if (!ok) return -1;
     }
   *phap = (short) pha;
   pha = *phap - orc_uniform (st);
   *pip = orc_interpolate_f ((float) pha, ch, e, ne);
   if (*pip < 0) return -1;
   return 0;
}

/* aciscontam.c:141-185: the analytic spatial profiles */
static double fxy_acis_i (double x, double y, double x_0, double y_0)
{
   double dx = x - x_0;
   double dy = y - y_0;
   double r = (8.0 / 1024.0) * sqrt (dx * dx + dy * dy);
   r /= 8.07;
   return 1.29 * r * r;
}

static double fxy_acis_s (double x, double y)
{
   double y_0 = 512.0;
   (void) x;
   if (y <= 512.0)
     return pow (fabs ((y - y_0) / (64.0 - y_0)), 5.5);
   return pow (fabs ((y - y_0) / (964.0 - y_0)), 4.5);
}

/* aciscontam.c:93-136: the transmission of chip slot k's contamination at
 * chip pixel (cx, cy) */
static double compute_contamination (const orc_ctx *c, const mxg_contam *ct, double en, double cx,
                                     double cy)
{
   double v = 0.0, fxy = 0.0;
   unsigned int i, ofs = 0;
   if (ct->fxy_model == MXG_CONTAM_TABLE)
     {
        if ((cx < 0) || (cx >= 1024) || (cy < 0) || (cy >= 1024))
          return 0.0;
        cx /= ct->blocking;
        cy /= ct->blocking;
        ofs = (1024 / ct->blocking) * (unsigned int) cy + (unsigned int) cx;
     }
   else if (ct->fxy_model == MXG_CONTAM_RADIAL)
     fxy = fxy_acis_i (cx, cy, ct->x0, ct->y0);
   else
     fxy = fxy_acis_s (cx, cy);
   for (i = 0; i < ct->num_layers; i++)
     {
        double mu = orc_interpolate_f ((float) en, c->fa + ct->mu_e_off[i], c->fa + ct->mu_off[i],
                                       ct->mu_n[i]);
        if (ct->fxy_model == MXG_CONTAM_TABLE)
          fxy = c->fa[ct->fxy_off[i] + ofs];
        v += mu * (ct->tau0[i] + ct->tau1[i] * fxy);
     }
   return exp (-v);
}

static void acis_photon (const orc_ctx *c, Marx_Photon_Attr_Type *at, double tstart, orc_stream *st)
{
   const mxg_detector_cfg *d = &c->cfg->det;
   const mxg_chip *g;
   double dx, dy;

   transform_ray (d, &at->x, &at->p);
   g = intersect_with_detector (d, at->x, at->p, &at->x, &dx, &dy);
   if (g == NULL)
     {
        at->flags |= PHOTON_MISSED_DETECTOR;
        at->ccd_num = -1;
     }
   else
     {
        int detected = 1;
        at->ccd_num = (signed char) g->id;
        at->y_pixel = (float) (dx / g->x_pixel_size);
        at->z_pixel = (float) (dy / g->y_pixel_size);
        if (d->ideal == 0)
          {
             double r = orc_uniform (st);
             double qe = 1.0, qe_filter = 1.0, qe_contam = 1.0;
             if (g->qe_n != 0)
               qe = orc_interpolate_f ((float) at->energy, c->fa + g->qe_e_off, c->fa + g->qe_off, g->qe_n);
             if (d->contam_on && d->contam[g - d->chip].num_layers)
               qe_contam = compute_contamination (c, d->contam + (g - d->chip), at->energy, at->y_pixel,
                                                  at->z_pixel);
             if (r >= qe * qe_filter * qe_contam)
               {
                  at->flags |= PHOTON_UNDETECTED;
                  detected = 0;
               }
          }
        if (detected && d->fef_on)
          {
             if (-1 == apply_fef (c, d, g->id, at->y_pixel, at->z_pixel, at->energy, st,
                                  &at->pulse_height, &at->pi))
               {
                  at->pulse_height = -1;
                  at->pi = 0;
                  at->flags |= PHOTON_UNDETECTED;
                  detected = 0;
               }
          }
        if (detected && (d->frame_transfer_time > 0.0))
          {
             /* frame-transfer streak (acis-i.c:60-92) */
             double t = fmod (tstart + at->arrival_time, d->frame_time);
             if (t > d->exposure_time)
               {
                  double xpixel, ypixel;
                  orc_vec dxv;
                  at->z_pixel = (float) (1.0 + 1022.0 * orc_uniform (st));
                  xpixel = (at->y_pixel - g->xpixel_offset) * g->x_pixel_size;
                  ypixel = (at->z_pixel - g->ypixel_offset) * g->y_pixel_size;
                  dxv = orc_v_ax1_bx2 (xpixel, v3 (g->xhat), ypixel, v3 (g->yhat));
                  at->x = orc_v_sum (v3 (g->x_ll), dxv);
                  at->p = orc_v_diff (at->x, orc_v (d->focal_length, 0, 0));
                  orc_v_normalize (&at->p);
                  at->flags |= PHOTON_ACIS_STREAKED;
               }
          }
     }
   transform_ray_reverse (d, &at->x, &at->p);
}

/* ---- HRC ---------------------------------------------------------------------------- */

static short hrc_pha (double energy, orc_stream *st)
{
   if (energy <= 0.5) energy = 141.582 * sqrt (energy);
   else if (energy < 2.0) energy = 107.299 * pow (energy, 0.1);
   else energy = 115.0;
   energy = energy * (1.0 + 0.424661 * orc_gaussian (st));
   if (energy < 0.0) energy = 0.0;
   return (short) energy;
}

static int filter_region (const mxg_detector_cfg *d, double y, double z, signed char *r)
{
   y -= d->shield_y_center;
   z -= d->shield_z_center;
   if (y < 0)
     {
        y = -y;
        if (y < d->shield_l) *r = 0;
        else if (y < d->shield_sl) *r = (z >= d->shield_t) ? 0 : 1;
        else if (y >= d->shield_sl_gap) *r = (z >= d->shield_t) ? 2 : 3;
        else return -1;
     }
   else
     {
        if (y < d->shield_r) *r = 0;
        else if (y < d->shield_sr) *r = (z >= d->shield_t) ? 0 : 1;
        else if (y >= d->shield_sr_gap) *r = (z >= d->shield_t) ? 2 : 3;
        else return -1;
     }
   return 0;
}

static double interp_qe (const orc_ctx *c, const mxg_qe_table *q, double energy)
{
   return orc_interpolate_f ((float) energy, c->fa + q->e_off, c->fa + q->eff_off, q->n);
}

static int apply_hrc_qe (const orc_ctx *c, Marx_Photon_Attr_Type *at, int mcp_id, orc_stream *st)
{
   const mxg_detector_cfg *d = &c->cfg->det;
   const mxg_qe_table *q = d->mcp_qe + d->mcp_map[mcp_id];
   double energy = at->energy;
   signed char region = 0;

   if (q->n != 0)
     {
        double qe = interp_qe (c, q, energy);
        if (orc_uniform (st) >= qe) return -1;
     }
   if (d->type == MARX_DETECTOR_HRC_S)
     {
        double t = (d->shield_x - at->x.x) / at->p.x;
        double y = at->x.y + t * at->p.y;
        double z = at->x.z + t * at->p.z;
        if (-1 == filter_region (d, y, z, &region)) return -1;
        q = d->filter_qe + region;
     }
   else
     q = d->filter_qe + mcp_id;
   if (q->n != 0)
     {
        double qe = interp_qe (c, q, energy);
        if (orc_uniform (st) >= qe) return -1;
     }
   if (d->type == MARX_DETECTOR_HRC_S)
     at->detector_region = region;
   at->pulse_height = hrc_pha (energy, st);
   return 0;
}

static void hrc_blur (const mxg_detector_cfg *d, double *dx, double *dy, orc_stream *st)
{
   double r, theta, x_0, y_0;
   if (d->ideal) return;
   r = orc_uniform (st);
   if (r < d->g1_wgt)
     {
        double cc;
        do cc = orc_uniform (st); while (cc == 0.0);
        r = d->g1_sigma * sqrt (-2 * log (cc));
        x_0 = d->g1_xctr;
        y_0 = d->g1_yctr;
     }
   else if (r < d->g1_wgt + d->g2_wgt)
     {
        double cc;
        do cc = orc_uniform (st); while (cc == 0.0);
        r = d->g2_sigma * sqrt (-2 * log (cc));
        x_0 = d->g2_xctr;
        y_0 = d->g2_yctr;
     }
   else
     {
        double cc = orc_uniform (st);
        double rmax = d->l1_rmax / d->l1_hwhm;
        r = d->l1_hwhm * sqrt (expm1 (cc * log1p (rmax * rmax)));
        x_0 = d->l1_xctr;
        y_0 = d->l1_yctr;
     }
   theta = (2.0 * PI) * orc_uniform (st);
   *dx += x_0 + r * cos (theta);
   *dy += y_0 + r * sin (theta);
   if (d->extend == 0)
     {
        if (*dx < 0.0) *dx = 0.0;
        if (*dy < 0.0) *dy = 0.0;
     }
}

static void hrc_photon (const orc_ctx *c, Marx_Photon_Attr_Type *at, orc_stream *st)
{
   const mxg_detector_cfg *d = &c->cfg->det;
   const mxg_chip *g;
   double dx, dy;

   transform_ray (d, &at->x, &at->p);
   g = intersect_with_detector (d, at->x, at->p, &at->x, &dx, &dy);
   if (g == NULL)
     {
        at->flags |= PHOTON_MISSED_DETECTOR;
        at->ccd_num = -1;
     }
   else if (-1 == apply_hrc_qe (c, at, g->id, st))
     {
        at->flags |= PHOTON_UNDETECTED;
        at->ccd_num = -1;
     }
   else
     {
        hrc_blur (d, &dx, &dy, st);
        at->ccd_num = (signed char) g->id;
        if (d->type == MARX_DETECTOR_HRC_S)
          {
             int id = g->id;
             double u = d->pix_u[id], v = d->pix_v[id];
             u += dx / d->u_pixel_size;
             v += dy / d->v_pixel_size;
             at->u_pixel = (float) u;
             at->v_pixel = (float) v;
             at->y_pixel = (float) (d->pix_x0[id] + (u - d->pix_u0[id]));
             at->z_pixel = (float) (d->pix_y0[id] + (v - d->pix_v0[id]));
          }
        else
          {
             dx /= d->u_pixel_size;
             dy /= d->v_pixel_size;
             at->y_pixel = (float) (d->pix_x0[0] + dx);
             at->z_pixel = (float) (d->pix_y0[0] + dy);
          }
     }
   transform_ray_reverse (d, &at->x, &at->p);
}

/* HESF pass over the unabsorbed photons (drake.c:268-369) */
static void drake_pass (const orc_ctx *c, Marx_Photon_Type *pt, orc_stream *streams)
{
   const mxg_detector_cfg *d = &c->cfg->det;
   unsigned int i;

   orc_prune_photons (pt);
   for (i = 0; i < pt->num_sorted; i++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[i];
        const mxg_rect *plate = NULL;
        double p_dot_n = 0.0, rfl;
        int use_cr = 0, k;
        orc_vec old_x = at->x, old_p = at->p;

        if (at->flags & BAD_PHOTON_MASK) continue;
        for (k = 0; k < d->drake_n; k++)
          {
             const mxg_rect *rect = d->drake + k;
             orc_vec nrm = v3 (rect->normal), a = v3 (rect->a), new_x, xp;
             double t, xx, yy;
             if (0.0 == (p_dot_n = orc_v_dot (old_p, nrm))) continue;
             t = orc_v_dot (orc_v_diff (a, old_x), nrm) / p_dot_n;
             new_x = orc_v_ax1_bx2 (1.0, old_x, t, old_p);
             xp = orc_v_diff (new_x, a);
             xx = orc_v_dot (xp, v3 (rect->e1));
             if ((xx < 0.0) || (xx >= rect->len1)) continue;
             yy = orc_v_dot (xp, v3 (rect->e2));
             if ((yy < 0.0) || (yy >= rect->len2)) continue;
             use_cr = (xx < d->drake_cr_width);
             at->x = new_x;
             plate = rect;
             break;
          }
        if (plate == NULL) continue;

        if (use_cr)
          {
             if (d->drake_cr_n != 0)
               {
                  float beta = orc_interpolate_f ((float) at->energy, c->fa + d->drake_cr_e_off,
                                                  c->fa + d->drake_cr_beta_off, d->drake_cr_n);
                  float delta = orc_interpolate_f ((float) at->energy, c->fa + d->drake_cr_e_off,
                                                   c->fa + d->drake_cr_delta_off, d->drake_cr_n);
                  rfl = orc_reflectivity (fabs (p_dot_n), beta, delta);
               }
             else rfl = 1.0;
          }
        else
          {
             if (d->drake_c_n != 0)
               {
                  float beta = orc_interpolate_f ((float) at->energy, c->fa + d->drake_c_e_off,
                                                  c->fa + d->drake_c_beta_off, d->drake_c_n);
                  float delta = orc_interpolate_f ((float) at->energy, c->fa + d->drake_c_e_off,
                                                   c->fa + d->drake_c_delta_off, d->drake_c_n);
                  rfl = orc_reflectivity (fabs (p_dot_n), beta, delta);
               }
             else rfl = 1.0;
          }
        if (rfl < orc_uniform (orc_photon_stream (c, streams, pt, at, MXG_STAGE_DETECT)))
          {
             at->flags |= PHOTON_DRAKE_BLOCKED;
             continue;
          }
        at->p = orc_v_ax1_bx2 (1.0, at->p, -2.0 * p_dot_n, v3 (plate->normal));
        at->flags |= PHOTON_DRAKE_REFLECTED;
     }
}

int orc_detect (const orc_ctx *c, Marx_Photon_Type *pt)
{
   const mxg_detector_cfg *d = &c->cfg->det;
   orc_stream *streams;
   unsigned int i;

   if (d->type == 0) return 0;
   if (pt->history & MARX_DET_NUM_OK) return 0;
   if ((d->type == MARX_DETECTOR_ACIS_S) || (d->type == MARX_DETECTOR_ACIS_I))
     pt->history |= (MARX_DET_PIXEL_OK | MARX_DET_NUM_OK | MARX_PULSEHEIGHT_OK | MARX_PI_OK);
   else
     pt->history |= (MARX_DET_REGION_OK | MARX_PULSEHEIGHT_OK | MARX_DET_PIXEL_OK | MARX_DET_NUM_OK);
   if (d->type == MARX_DETECTOR_HRC_S)
     pt->history |= MARX_DET_UV_PIXEL_OK;

   if (NULL == (streams = orc_streams_alloc (pt->n_photons))) return -1;
   if ((d->type == MARX_DETECTOR_HRC_S) && d->use_drake)
     drake_pass (c, pt, streams);
   orc_prune_photons (pt);

   for (i = 0; i < pt->num_sorted; i++)
     {
        Marx_Photon_Attr_Type *at = pt->attributes + pt->sorted_index[i];
        orc_stream *st = orc_photon_stream (c, streams, pt, at, MXG_STAGE_DETECT);
        if ((d->type == MARX_DETECTOR_ACIS_S) || (d->type == MARX_DETECTOR_ACIS_I))
          acis_photon (c, at, pt->start_time, st);
        else
          hrc_photon (c, at, st);
     }
   free (streams);
   return 0;
}
