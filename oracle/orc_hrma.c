/* orc_hrma.c -- oracle restatement of the HRMA reflection pass.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Follows
 *   marx/libsrc/hrma.c:1161-1341  _marx_hrma_mirror_reflect (vignetting pass,
 *                                 prune, per-photon P then H reflection)
 *   marx/libsrc/hrma.c:984-1050   project_photon_to_hrma
 *   marx/libsrc/hrma.c:411-478    compute_conic_intersection
 *   marx/libsrc/hrma.c:488-545    reflect_from_conic
 *   marx/libsrc/hrma.c:1054-1093  blur_normal
 *   marx/libsrc/hrma.c:930-981    intersects_struts
 */
#include <stdlib.h>
#include <math.h>

#include "oracle_int.h"

int orc_strut_shadow (const mxg_hrma_cfg *h, orc_vec xv, orc_vec pv, const mxg_strut *s,
                      unsigned int num)
{
   double theta = 30.0 * (PI / 180.0);
   double cos_theta = cos (theta);
   double sin_theta = sin (theta);
   const mxg_strut *smax = s + num;
   unsigned int i;

   while (s < smax)
     {
        double half_width = s->half_width;
        double x = xv.x, y = xv.y, z = xv.z;
        double px = pv.x, py = pv.y, pz = pv.z;
        double t = (s->xpos + h->cap_position - x) / px;
        y += py * t;
        z += pz * t;
        for (i = 0; i < 3; i++)
          {
             if (i != 0)
               {
                  double tmp = cos_theta * y - sin_theta * z;
                  z = sin_theta * y + cos_theta * z;
                  y = tmp;
               }
             if (((-half_width < y) && (y < half_width))
                 || ((-half_width < z) && (z < half_width)))
               return 1;
          }
        s++;
     }
   return 0;
}

/* Test tap: when set, the ray (x, p) that reaches each strut set (0 pre-,
 * 1 cap, 2 post-collimator) is recorded per photon, whether or not struts
 * are on, at taps[(18 * attribute index) + 6 * set + (x, p)]; sets a photon
 * does not reach stay as the caller filled them. */
static double *Strut_Taps = NULL;
static unsigned int Strut_Taps_N = 0;
static const Marx_Photon_Attr_Type *Strut_Taps_Base = NULL;

void orc_set_strut_taps (double *taps, unsigned int n)
{
   Strut_Taps = taps;
   Strut_Taps_N = n;
}

static void tap_strut_ray (const Marx_Photon_Attr_Type *at, unsigned int set)
{
   size_t i;
   double *t;
   if ((Strut_Taps == NULL) || (Strut_Taps_Base == NULL)) return;
   i = (size_t) (at - Strut_Taps_Base);
   if (i >= Strut_Taps_N) return;
   t = Strut_Taps + 18 * i + 6 * set;
   t[0] = at->x.x; t[1] = at->x.y; t[2] = at->x.z;
   t[3] = at->p.x; t[4] = at->p.y; t[5] = at->p.z;
}

static int intersects_struts (const mxg_hrma_cfg *h, Marx_Photon_Attr_Type *at,
                              const mxg_strut *s, unsigned int num)
{
   if (0 == orc_strut_shadow (h, at->x, at->p, s, num)) return 0;
   at->flags |= PHOTON_MIRROR_VBLOCKED;
   return 1;
}

static int project_photon_to_hrma (const mxg_hrma_cfg *h, Marx_Photon_Attr_Type *at,
                                   double source_distance, orc_stream *st)
{
   unsigned int i;
   while (1)
     {
        double r = orc_uniform (st);
        for (i = 0; i < MXG_NUM_SHELLS; i++)
          {
             const mxg_shell_cfg *s = h->shell + i;
             double theta, radius;
             unsigned int quad;

             if (!(r < s->area_fraction)) continue;
             at->mirror_shell = i;
             radius = s->min_radius + (s->max_radius - s->min_radius) * orc_uniform (st);
             do
               {
                  theta = orc_uniform (st);
                  quad = (unsigned int) (4.0 * theta);
               }
             while (0 == (s->shutter_bitmap & (1u << quad)));
             theta = (2.0 * PI) * (theta - 1.0 / 8.0);
             at->x.z = radius * cos (theta);
             at->x.y = radius * sin (theta);
             at->x.x = s->front_position;
             at->x.z -= s->to_osac[0][2];
             at->x.y -= s->to_osac[0][1];
             if (source_distance > 0.0)
               {
                  at->p = orc_v_ax1_bx2 (1.0, at->x, source_distance, at->p);
                  orc_v_normalize (&at->p);
               }
             tap_strut_ray (at, 0);
             if (h->use_struts && (1 == intersects_struts (h, at, h->pre_struts, 2)))
               return -1;
             return 0;
          }
     }
}

static int compute_conic_intersection (double a, double b, double c, orc_vec *x0, orc_vec p,
                                       orc_vec *normal, double xmin, double xmax)
{
   double alpha, beta, gamma, x_y, x_z;
   double t_plus, t_minus, x_plus, x_minus;

   t_plus = -x0->x / p.x;
   x_y = x0->y + t_plus * p.y;
   x_z = x0->z + t_plus * p.z;

   alpha = a * p.x * p.x - 1.0;
   beta = b * p.x - 2.0 * (p.y * x_y + p.z * x_z);
   gamma = c - x_z * x_z - x_y * x_y;

   if (alpha == 0.0)
     {
        if (beta == 0.0) return -1;
        t_plus = t_minus = -gamma / beta;
     }
   else if (0 >= orc_quadratic_root (alpha, beta, gamma, &t_plus, &t_minus))
     return -1;

   x_plus = p.x * t_plus;
   x_minus = p.x * t_minus;

   if ((x_plus >= xmin) && (x_plus < xmax))
     {
        if ((x_minus >= xmin) && (x_minus < xmax) && (x_minus > x_plus))
          {
             x0->x = x_minus;
             x0->y = x_y + p.y * t_minus;
             x0->z = x_z + p.z * t_minus;
          }
        else
          {
             x0->x = x_plus;
             x0->y = x_y + p.y * t_plus;
             x0->z = x_z + p.z * t_plus;
          }
     }
   else if ((x_minus >= xmin) && (x_minus < xmax))
     {
        x0->x = x_minus;
        x0->y = x_y + p.y * t_minus;
        x0->z = x_z + p.z * t_minus;
     }
   else return -1;

   normal->x = (a - 1) * x0->x + 0.5 * b;
   normal->y = -x0->y;
   normal->z = -x0->z;
   orc_v_normalize (normal);
   return 0;
}

static void blur_normal (orc_vec *n, double blur, orc_stream *st)
{
   double phi, n_y = n->y, n_z = n->z, len;
   orc_vec perp;

   len = sqrt (n_y * n_y + n_z * n_z);
   perp.x = 0.0;
   perp.y = n_z / len;
   perp.z = -n_y / len;
   phi = (2.0 * PI) * orc_uniform (st);
   perp = orc_v_rotate_unit_vector (perp, *n, phi);
   phi = blur * (1.0 / 3600.0 * PI / 180.0);
   phi = phi * orc_gaussian (st);
   *n = orc_v_rotate_unit_vector (*n, perp, phi);
}

/* 0 reflected, -1 absorbed, -2 missed the surface */
static int reflect_from_conic (const orc_ctx *c, const mxg_shell_cfg *s, unsigned int k,
                               orc_vec *x, orc_vec *p, double energy, double beta,
                               double delta, double correction_factor, orc_stream *st)
{
   const mxg_hrma_cfg *h = &c->cfg->hrma;
   orc_vec normal;
   double p_dot_n, r, rfl, sin_grazing, delta_grazing;

   if (-1 == compute_conic_intersection (s->conic_a[k], s->conic_b[k], s->conic_c[k],
                                         x, *p, &normal, s->conic_xmin[k], s->conic_xmax[k]))
     return -2;
   if (k == 0)
     {
        ORC_PHASE (1);
        ORC_MARK (1);
     }

   if (h->use_blur)
     blur_normal (&normal, s->blur[k], st);

   p_dot_n = orc_v_dot (*p, normal);

   if (h->ideal == 0)
     {
        r = orc_uniform (st);
        rfl = orc_reflectivity (fabs (p_dot_n), beta, delta);
        if (r >= rfl * correction_factor)
          return -1;
     }

   *p = orc_v_ax1_bx2 (1.0, *p, -2.0 * p_dot_n, normal);

   if (h->use_wfold == 0)
     return 0;

   sin_grazing = -p_dot_n;
   r = orc_uniform (st);
   delta_grazing = orc_wfold_interp (c, &s->wfold[k], energy, sin_grazing, r);
   delta_grazing *= s->scat_factor[k];
   if (delta_grazing > PI / 4) return -1;
   if (orc_uniform (st) < 0.5) delta_grazing = -delta_grazing;
   *p = orc_v_rotate_unit_vector (*p, orc_v_cross (*p, normal), delta_grazing);
   return 0;
}

int orc_mirror_reflect (const orc_ctx *c, Marx_Photon_Type *pt)
{
   const mxg_hrma_cfg *h = &c->cfg->hrma;
   Marx_Photon_Attr_Type *attrs = pt->attributes;
   orc_stream *streams;
   unsigned int n, i;
   double source_distance;

   if (pt->history & MARX_MIRROR_SHELL_OK)
     return 0;
   pt->history |= MARX_MIRROR_SHELL_OK;

   if (NULL == (streams = orc_streams_alloc (pt->n_photons)))
     return -1;
   Strut_Taps_Base = attrs;

   orc_prune_photons (pt);
   n = pt->num_sorted;

   if (h->ideal == 0)
     {
        for (i = 0; i < n; i++)
          {
             Marx_Photon_Attr_Type *at = attrs + pt->sorted_index[i];
             orc_stream *st = orc_photon_stream (c, streams, pt, at, MXG_STAGE_MIRROR);
             if (orc_uniform (st) > h->vignetting)
               at->flags |= PHOTON_MIRROR_VBLOCKED;
          }
        orc_prune_photons (pt);
        n = pt->num_sorted;
     }

   source_distance = pt->source_distance;

   for (i = 0; i < n; i++)
     {
        Marx_Photon_Attr_Type *at = attrs + pt->sorted_index[i];
        orc_stream *st = orc_photon_stream (c, streams, pt, at, MXG_STAGE_MIRROR);
        const mxg_shell_cfg *s;
        double energy, beta, delta, correction_factor = 1.0;
        int status;

        ORC_PHASE (0);
        (void) project_photon_to_hrma (h, at, source_distance, st);
        s = h->shell + at->mirror_shell;

        at->x = orc_v_sum (at->x, orc_v (s->to_osac[0][0], s->to_osac[0][1], s->to_osac[0][2]));
        at->x = orc_m3_vmul (s->fwd[0], at->x);
        at->p = orc_m3_vmul (s->fwd[0], at->p);

        energy = at->energy;
        ORC_PHASE (2);
        ORC_MARK (2);
        if (h->opt_n == 0)
          {
             beta = 0.0;
             delta = 1.0;
             correction_factor = 1.0;
          }
        else
          {
             beta = orc_interpolate_f ((float) energy, c->fa + h->opt_e_off,
                                       c->fa + h->opt_beta_off, h->opt_n);
             delta = orc_interpolate_f ((float) energy, c->fa + h->opt_e_off,
                                        c->fa + h->opt_delta_off, h->opt_n);
             if (h->use_scale)
               {
                  correction_factor = orc_interpolate_f ((float) energy, c->fa + s->corr_e_off,
                                                         c->fa + s->corr_f_off, s->corr_n);
                  correction_factor = sqrt (correction_factor);
               }
          }
        ORC_PHASE (0);

        status = reflect_from_conic (c, s, 0, &at->x, &at->p, energy, beta, delta,
                                     correction_factor, st);
        if (status != 0)
          {
             at->flags |= PHOTON_UNREFLECTED;
             continue;
          }

        at->p = orc_m3_vmul (s->bwd[0], at->p);
        at->x = orc_m3_vmul (s->bwd[0], at->x);
        at->x = orc_v_diff (at->x, orc_v (s->to_osac[0][0], s->to_osac[0][1], s->to_osac[0][2]));

        tap_strut_ray (at, 1);
        if (h->use_struts && (1 == intersects_struts (h, at, h->cap_struts, 2)))
          continue;

        at->x = orc_v_sum (at->x, orc_v (s->to_osac[1][0], s->to_osac[1][1], s->to_osac[1][2]));
        at->p = orc_m3_vmul (s->fwd[1], at->p);
        at->x = orc_m3_vmul (s->fwd[1], at->x);

        if (0 != reflect_from_conic (c, s, 1, &at->x, &at->p, energy, beta, delta,
                                     correction_factor, st))
          {
             at->flags |= PHOTON_UNREFLECTED;
             continue;
          }

        at->p = orc_m3_vmul (s->bwd[1], at->p);
        at->x = orc_m3_vmul (s->bwd[1], at->x);
        at->x = orc_v_diff (at->x, orc_v (s->to_osac[1][0], s->to_osac[1][1], s->to_osac[1][2]));

        tap_strut_ray (at, 2);
        if (h->use_struts)
          (void) intersects_struts (h, at, h->post_struts, 2);
     }

   ORC_PHASE (0);
   Strut_Taps_Base = NULL;
   free (streams);
   return 0;
}
