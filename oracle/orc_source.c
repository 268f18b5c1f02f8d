/* orc_source.c -- oracle restatement of photon creation and dither.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Follows
 *   marx/libsrc/source.c:268-384  marx_create_photons (source call, arrival
 *                                 times with the exposure cut, tags, dither,
 *                                 energy sort)
 *   marx/libsrc/s-point.c:59-83, s-disk.c:63-110, s-gauss.c:78-143,
 *   s-beta.c:81-130, s-line.c:60-101             directions
 *   marx/libsrc/spectrum.c:138-189, prob.c:46-58  FLAT / FILE energies
 *   marx/libsrc/dither.c:167-182, 551-581, 609-658  INTERNAL dither
 * Deviations fixed by the per-photon stream design, shared with the GPU:
 *  - the disk, Gaussian and beta-model azimuth of each photon is drawn
 *    independently (the reference rotates one normal cumulatively from
 *    photon to photon, s-disk.c:99, s-gauss.c:124, s-beta.c:111);
 *  - the exponential gaps are summed as 64-bit integers in units of 2^-32
 *    of the mean gap, so that a parallel scan gives the same times;
 *  - MXG_TIME_UNIFORM draws times uniformly over [0, tmax).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "oracle_int.h"

static orc_vec apply_dither (double ra, double dec, double roll, orc_vec p)
{
   double cos_dec, sin_dec, cos_ra, sin_ra, cos_theta, sin_theta;
   orc_vec n;

   p = orc_v_rotate_unit_vector (p, orc_v (1, 0, 0), -roll);
   cos_ra = cos (ra);
   sin_ra = sin (ra);
   cos_dec = cos (dec);
   sin_dec = sin (dec);
   cos_theta = cos_dec * cos_ra;
   n = orc_v (0, sin_dec, -cos_dec * sin_ra);
   sin_theta = orc_v_length (n);
   if (sin_theta <= 1e-20) return p;
   n.x /= sin_theta;
   n.y /= sin_theta;
   n.z /= sin_theta;
   return orc_v_rotate_unit_vector1 (p, n, cos_theta, sin_theta);
}

static double draw_energy (const orc_ctx *c, const mxg_source_cfg *s, orc_stream *st)
{
   if (s->spectrum == MXG_SPECTRUM_FLAT)
     {
        double emin = s->emin, de = s->emax - emin;
        return emin + de * orc_uniform (st);
     }
   return orc_interpolate_d (orc_uniform (st), c->da + s->file_cum_off, c->da + s->file_e_off,
                             s->file_n);
}

/* The direction of a photon of the configured shape about the source
 * direction p (s-disk.c:95-106, s-gauss.c:121-137, s-beta.c:108-124,
 * s-line.c:89-98, s-image.c:284-324); p_normal is normal to p, line_normal
 * and line_alpha rotate -x onto p; an image's cumulative pixel
 * distribution is at fa + image_off. */
static orc_vec extended_direction (const mxg_source_cfg *s, const float *fa, orc_vec p, orc_vec p_normal,
                                   orc_vec line_normal, double line_alpha, orc_stream *st)
{
   if (s->type == MXG_SOURCE_DISK)
     {
        double x0 = s->disk_theta_min / s->disk_theta, x1, rnd;
        orc_vec normal;
        x0 = x0 * x0;
        x1 = 1.0 - x0;
        normal = orc_v_rotate_unit_vector (p_normal, p, 2.0 * PI * orc_uniform (st));
        rnd = s->disk_theta * sqrt (x0 + x1 * orc_uniform (st));
        return orc_v_rotate_unit_vector (p, normal, rnd);
     }
   if ((s->type == MXG_SOURCE_GAUSS) || (s->type == MXG_SOURCE_BETA))
     {
        orc_vec normal = orc_v_rotate_unit_vector (p_normal, p, 2.0 * PI * orc_uniform (st));
        double rnd;
        do rnd = orc_uniform (st);
        while (rnd == 0.0);
        if (s->type == MXG_SOURCE_GAUSS)
          rnd = s->gauss_sigma * sqrt (-log (rnd));
        else
          rnd = s->beta_core * sqrt (pow (rnd, s->beta_xpon) - 1.0);
        return orc_v_rotate_unit_vector (p, normal, rnd);
     }
   if (s->type == MXG_SOURCE_IMAGE)
     {
        /* s-image.c:284-324 */
        unsigned int ofs = orc_binary_search_f ((float) orc_uniform (st), fa + s->image_off, s->image_n);
        double x, y, cos_y;
        orc_vec p1;
        y = (double) (ofs / s->image_nx);
        x = (double) (ofs % s->image_nx);
        y += -0.5 * s->image_ny + (orc_uniform (st) - 0.5);
        x += -0.5 * s->image_nx + (orc_uniform (st) - 0.5);
        y = y * s->image_yscale;
        x = x * s->image_xscale;
        cos_y = cos (y);
        p1.x = -cos_y * cos (x);
        p1.y = cos_y * sin (x);
        p1.z = -sin (y);
        return orc_v_rotate_unit_vector (p1, line_normal, line_alpha);
     }
   if (s->type == MXG_SOURCE_LINE)
     {
        double theta = s->line_theta * (-1.0 + 2.0 * orc_uniform (st));
        double sin_theta = -sin (theta);
        orc_vec q = orc_v (-cos (theta), sin_theta * s->line_cos_phi, sin_theta * s->line_sin_phi);
        return orc_v_rotate_unit_vector (q, line_normal, line_alpha);
     }
   return p;
}

/* the angle and unit axis rotating a onto b (jdmath/src/vector.c:289-331) */
static double rotation_axis (orc_vec a, orc_vec b, orc_vec *c)
{
   double theta = orc_v_dot (a, b), len;
   orc_vec cc;
   if (theta > 1.0) theta = 1.0;
   else if (theta < -1.0) theta = -1.0;
   theta = acos (theta);
   cc = orc_v_cross (a, b);
   len = orc_v_length (cc);
   if (len > 0.0)
     {
        *c = orc_v (cc.x / len, cc.y / len, cc.z / len);
        return theta;
     }
   *c = ((a.x != 0.0) || (a.y != 0.0)) ? orc_v (-a.y, -a.x, 0.0) : orc_v (-a.z, 0.0, 0.0);
   orc_v_normalize (c);
   return theta;
}

/* INTERNAL dither of one photon (dither.c:167-182, 551-581, 609-658) */
static void dither_photon (const orc_ctx *c, const mxg_source_cfg *s, Marx_Photon_Attr_Type *at,
                           double start_time)
{
   orc_stream st;
   double t, ra, dec, roll, delta_ra, delta_dec;
   Marx_Dither_Type *d = &at->dither_state;
   orc_stream_init (&st, c->seed, at->tag, MXG_STAGE_DITHER, c->source_id);
   t = start_time + at->arrival_time;
   t = (2.0 * PI) * t;
   ra = d->ra = (float) (s->ra_amp * sin (t / s->ra_period + s->ra_phase));
   dec = d->dec = (float) (s->dec_amp * sin (t / s->dec_period + s->dec_phase));
   roll = d->roll = (float) (s->nominal_roll + s->roll_amp * sin (t / s->roll_period + s->roll_phase));
   d->dy = d->dz = d->dtheta = 0;
   delta_ra = s->aspect_blur * orc_gaussian (&st);
   delta_dec = s->aspect_blur * orc_gaussian (&st);
   at->p = apply_dither (ra + delta_ra, dec + delta_dec, roll, at->p);
}

/* The configured dither applied to every photon of pt at its own arrival
 * time and pt->start_time (dither.c:630-658), as marx_create_photons applies
 * it to photons replayed from a ray file (source.c:344-359). */
int orc_dither_photons (const orc_ctx *c, Marx_Photon_Type *pt)
{
   const mxg_source_cfg *s = &c->cfg->src;
   unsigned int i;
   if (s->dither_mode == MXG_DITHER_INTERNAL)
     for (i = 0; i < pt->n_photons; i++)
       dither_photon (c, s, pt->attributes + i, pt->start_time);
   return 0;
}

int orc_create_photons (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int num,
                        unsigned int *num_collected, double *exposure_time)
{
   const mxg_source_cfg *s = &c->cfg->src;
   Marx_Photon_Attr_Type *at = pt->attributes;
   orc_stream *streams;
   unsigned int n = num, i;
   unsigned long long S = 0;

   *num_collected = 0;
   if (num > pt->max_n_photons) return -1;
   pt->source_distance = s->distance;
   pt->history = 0;
   pt->start_time += pt->total_time;
   memset (at, 0, num * sizeof (Marx_Photon_Attr_Type));
   if (NULL == (streams = orc_streams_alloc (num))) return -1;

   /* source: energies and directions */
   for (i = 0; i < num; i++)
     {
        orc_stream *st = streams + i;
        at[i].tag = pt->tag_start + i;
        orc_stream_init (st, c->seed, at[i].tag, MXG_STAGE_SOURCE, c->source_id);
        at[i].energy = draw_energy (c, s, st);
        at[i].p = extended_direction (s, c->fa, orc_v (s->p[0], s->p[1], s->p[2]),
                                      orc_v (s->p_normal[0], s->p_normal[1], s->p_normal[2]),
                                      orc_v (s->line_normal[0], s->line_normal[1], s->line_normal[2]),
                                      s->line_alpha, st);
     }

   /* arrival times */
   for (i = 0; i < num; i++)
     {
        orc_stream *st = streams + i;
        double t;
        if (s->time_mode == MXG_TIME_POISSON)
          {
             double e = orc_expn (st);
             S += (unsigned long long) (e * 4294967296.0 + 0.5);
             t = s->mean_time * ((double) S * (1.0 / 4294967296.0));
          }
        else
          t = s->uniform_tmax * orc_uniform (st);
        at[i].arrival_time = t;
        at[i].flags = 0;
        if ((s->time_mode == MXG_TIME_POISSON) && (exposure_time != NULL)
            && (*exposure_time > 0.0) && (*exposure_time <= t))
          {
             n = i + 1;
             break;
          }
     }
   free (streams);

   pt->history |= (MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK | MARX_TAG_OK);
   pt->n_photons = n;
   pt->tag_start += n;

   /* dither (dither.c:609-658) */
   if ((n > 0) && (s->dither_mode == MXG_DITHER_INTERNAL))
     for (i = 0; i < n; i++)
       dither_photon (c, s, at + i, pt->start_time);

   if (-1 == orc_sort_photons (pt)) return -1;
   if (n == 0) pt->total_time = 0.0;
   else pt->total_time = (s->time_mode == MXG_TIME_UNIFORM) ? s->uniform_tmax : at[n - 1].arrival_time;
   *num_collected = n;
   return 0;
}

/* A USER source (source.c:268-384 over s-user.c:140-205): energies and
 * Poisson times from the photons' streams for all num photons, then the
 * plug-in's rays in creation order until it returns -1; its time steps
 * (when it gives them) replace the times, its energy the drawn one unless
 * negative; then the exposure cut of the time model in use (tstart +
 * exposure for the plug-in's times, exposure for Poisson ones), tags,
 * dither and the energy sort. */
int orc_create_user (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int num,
                     unsigned int *num_collected, double *exposure_time, double tstart,
                     int (*ray) (double *, double *, double *, double *, double *))
{
   const mxg_source_cfg *s = &c->cfg->src;
   Marx_Photon_Attr_Type *at = pt->attributes;
   unsigned int n, i;
   unsigned long long S = 0;
   double t = -1.0, user_time = 0.0;

   *num_collected = 0;
   if (num > pt->max_n_photons) return -1;
   pt->source_distance = s->distance;
   pt->history = 0;
   pt->start_time += pt->total_time;
   memset (at, 0, num * sizeof (Marx_Photon_Attr_Type));
   for (i = 0; i < num; i++)
     {
        orc_stream st;
        at[i].tag = pt->tag_start + i;
        orc_stream_init (&st, c->seed, at[i].tag, MXG_STAGE_SOURCE, c->source_id);
        at[i].energy = draw_energy (c, s, &st);
        if (s->time_mode == MXG_TIME_POISSON)
          {
             double e = orc_expn (&st);
             S += (unsigned long long) (e * 4294967296.0 + 0.5);
             at[i].arrival_time = s->mean_time * ((double) S * (1.0 / 4294967296.0));
          }
        else
          at[i].arrival_time = s->uniform_tmax * orc_uniform (&st);
     }
   for (i = 0; i < num; i++)
     {
        double energy = at[i].energy, drawn = energy, px, py, pz;
        if (-1 == (*ray) (&t, &energy, &px, &py, &pz))
          break;
        at[i].flags = 0;
        at[i].p = orc_v (px, py, pz);
        if (t >= 0.0)
          {
             user_time += t;
             at[i].arrival_time = user_time;
          }
        if (energy < 0.0) energy = drawn;
        at[i].energy = energy;
     }
   n = i;
   if (exposure_time != NULL)
     for (i = 0; i < n; i++)
       if ((t >= 0.0) ? (at[i].arrival_time >= tstart + *exposure_time)
           : (*exposure_time <= at[i].arrival_time))
         {
            n = i + 1;
            break;
         }
   pt->history = MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK;
   if (n) pt->history |= MARX_TAG_OK;
   pt->n_photons = n;
   pt->tag_start += n;
   if ((n > 0) && (s->dither_mode == MXG_DITHER_INTERNAL))
     for (i = 0; i < n; i++)
       dither_photon (c, s, at + i, pt->start_time);
   if (-1 == orc_sort_photons (pt)) return -1;
   pt->total_time = n ? at[n - 1].arrival_time : 0.0;
   *num_collected = n;
   return 0;
}

/* Batched point sources: the photons mxg_trace_batch creates.  Photon i of
 * the batch (sources in order) has tag tag_start + i, the direction of its
 * source, an energy from the configured spectrum, an arrival time uniform
 * over [0, exposure) and the configured dither.  The list is energy-sorted
 * for the stage passes like a created chunk. */
int orc_create_batch (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int nsrc,
                      const double *dirs, const unsigned int *counts, unsigned int tag_start,
                      double start_time, double exposure)
{
   const mxg_source_cfg *s = &c->cfg->src;
   Marx_Photon_Attr_Type *at = pt->attributes;
   unsigned long long total = 0;
   unsigned int i = 0, k, j;

   for (k = 0; k < nsrc; k++) total += counts[k];
   if (total > pt->max_n_photons) return -1;
   memset (at, 0, total * sizeof (Marx_Photon_Attr_Type));
   pt->source_distance = s->distance;
   pt->history = 0;
   pt->start_time = start_time;
   pt->tag_start = tag_start;
   for (k = 0; k < nsrc; k++)
     for (j = 0; j < counts[k]; j++, i++)
       {
          orc_stream st;
          at[i].tag = tag_start + i;
          orc_stream_init (&st, c->seed, at[i].tag, MXG_STAGE_SOURCE, c->source_id);
          at[i].energy = draw_energy (c, s, &st);
          at[i].p = orc_v (dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]);
          if (s->type != MXG_SOURCE_POINT)
            {
               orc_vec p = at[i].p, nrm = orc_v (-p.y / p.x, 1.0, 0.0), ln = orc_v (0, 0, 0);
               double la = 0.0;
               orc_v_normalize (&nrm);
               if (s->type == MXG_SOURCE_LINE) la = rotation_axis (orc_v (-1.0, 0.0, 0.0), p, &ln);
               at[i].p = extended_direction (s, c->fa, p, nrm, ln, la, &st);
            }
          at[i].arrival_time = exposure * orc_uniform (&st);
          if (s->dither_mode == MXG_DITHER_INTERNAL)
            dither_photon (c, s, at + i, start_time);
       }
   pt->n_photons = (unsigned int) total;
   pt->history |= (MARX_ENERGY_OK | MARX_TIME_OK | MARX_X_VECTOR_OK | MARX_P_VECTOR_OK | MARX_TAG_OK);
   pt->total_time = exposure;
   return orc_sort_photons (pt);
}
