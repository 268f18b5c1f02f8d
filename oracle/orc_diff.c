/* orc_diff.c -- oracle of the differentiable mirror: the weighted-photon
 * HRMA trace with forward-mode derivatives in FP64, its focal-plane image
 * (FP64) and the gradient of a loss given d(loss)/d(image).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The photon steps follow
 *   marx/libsrc/hrma.c:984-1050   aperture (project_photon_to_hrma)
 *   marx/libsrc/hrma.c:411-478    conic intersection
 *   marx/libsrc/hrma.c:488-545    reflection: normal blur, reflectivity,
 *                                 WFold scatter
 *   marx/libsrc/hrma.c:900-981    strut shadows
 *   marx/libsrc/hrma.c:1161-1341  the driver's frame changes
 * with survival weights in place of the vignetting and reflectivity draws
 * (a photon missing a surface, shadowed by a strut or scattered past pi/4
 * gets weight 0).  Each quantity carries its derivatives by the photon's
 * shell-pair parameters (tilt az, tilt el, decenter y, decenter z, despace
 * x; the geometry and its derivatives come in an mxg_diff_cfg).  The
 * reflectivity's derivative by the incidence cosine is taken here by a
 * Richardson-extrapolated central difference of orc_reflectivity.
 * Draw order per photon: source stream: energy; mirror stream: shell,
 * radius, azimuth (rejection loop), then per surface the blur azimuth and
 * Gaussian, and the WFold draw and sign.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
# include <omp.h>
#endif

#include "oracle_int.h"

#define NP MXG_DIFF_NPAR

typedef struct { double v, d[NP]; } od;
typedef struct { od x, y, z; } odv;

/* Op-count buckets (the op-counting build, oracle.h): part + kind, with
 * part 0 up to and including the paraboloid intersection and 4 after it,
 * and kind 0 the values, 1 the tilt tangents (0, 1), 2 the translation
 * tangents (2..4), 3 the finite-difference reflectivity slope.  No effect in
 * the plain build. */
#ifdef ORC_OPCOUNT_HPP
static int Diff_Part;
# define SET_PART(k) (Diff_Part = (k))
#else
# define SET_PART(k) ((void) 0)
#endif
#define PH(k) ORC_PHASE (Diff_Part + (k))
#define TAN(j) PH ((j) < 2 ? 1 : 2)

static od oc (double v) { od r; int j; r.v = v; for (j = 0; j < NP; j++) r.d[j] = 0; return r; }
static od o_add (od a, od b)
{
   int j;
   a.v += b.v;
   for (j = 0; j < NP; j++) { TAN (j); a.d[j] += b.d[j]; }
   PH (0);
   return a;
}
static od o_sub (od a, od b)
{
   int j;
   a.v -= b.v;
   for (j = 0; j < NP; j++) { TAN (j); a.d[j] -= b.d[j]; }
   PH (0);
   return a;
}
static od o_neg (od a)
{
   int j;
   a.v = -a.v;
   for (j = 0; j < NP; j++) { TAN (j); a.d[j] = -a.d[j]; }
   PH (0);
   return a;
}
static od o_mul (od a, od b)
{
   od r; int j;
   r.v = a.v * b.v;
   for (j = 0; j < NP; j++) { TAN (j); r.d[j] = a.d[j] * b.v + a.v * b.d[j]; }
   PH (0);
   return r;
}
static od o_scale (double s, od a)
{
   int j;
   a.v *= s;
   for (j = 0; j < NP; j++) { TAN (j); a.d[j] *= s; }
   PH (0);
   return a;
}
static od o_div (od a, od b)
{
   od r; int j;
   r.v = a.v / b.v;
   for (j = 0; j < NP; j++) { TAN (j); r.d[j] = (a.d[j] - r.v * b.d[j]) / b.v; }
   PH (0);
   return r;
}
static od o_addc (od a, double c) { a.v += c; return a; }
static od o_sqrt (od a)
{
   od r; int j;
   r.v = sqrt (a.v);
   for (j = 0; j < NP; j++) { TAN (j); r.d[j] = (r.v > 0) ? a.d[j] / (2 * r.v) : 0.0; }
   PH (0);
   return r;
}
static od o_fabs (od a) { return (a.v < 0) ? o_neg (a) : a; }
static void o_sincos (od a, od *s, od *c)
{
   int j;
   s->v = sin (a.v);
   c->v = cos (a.v);
   for (j = 0; j < NP; j++)
     {
        TAN (j);
        s->d[j] = c->v * a.d[j];
        c->d[j] = -s->v * a.d[j];
     }
   PH (0);
}

static odv ov (orc_vec a) { odv r; r.x = oc (a.x); r.y = oc (a.y); r.z = oc (a.z); return r; }
static orc_vec ovval (odv a) { return orc_v (a.x.v, a.y.v, a.z.v); }
static odv ov_add (odv a, odv b) { a.x = o_add (a.x, b.x); a.y = o_add (a.y, b.y); a.z = o_add (a.z, b.z); return a; }
static odv ov_sub (odv a, odv b) { a.x = o_sub (a.x, b.x); a.y = o_sub (a.y, b.y); a.z = o_sub (a.z, b.z); return a; }
static od ov_dot (odv a, odv b) { return o_add (o_add (o_mul (a.x, b.x), o_mul (a.y, b.y)), o_mul (a.z, b.z)); }
static odv ov_cross (odv a, odv b)
{
   odv c;
   c.x = o_sub (o_mul (a.y, b.z), o_mul (a.z, b.y));
   c.y = o_sub (o_mul (a.z, b.x), o_mul (a.x, b.z));
   c.z = o_sub (o_mul (a.x, b.y), o_mul (a.y, b.x));
   return c;
}
static odv ov_lin2 (od a, odv x1, od b, odv x2)
{
   odv r;
   r.x = o_add (o_mul (a, x1.x), o_mul (b, x2.x));
   r.y = o_add (o_mul (a, x1.y), o_mul (b, x2.y));
   r.z = o_add (o_mul (a, x1.z), o_mul (b, x2.z));
   return r;
}
static odv ov_unit (odv a)
{
   od len = o_sqrt (ov_dot (a, a));
   if (len.v == 0) return a;
   a.x = o_div (a.x, len);
   a.y = o_div (a.y, len);
   a.z = o_div (a.z, len);
   return a;
}
/* p rotated about the unit axis n by theta, renormalised (JDMv_rotate_unit_vector) */
static odv ov_rotate (odv p, odv n, od theta)
{
   od s, c, pn;
   odv r;
   o_sincos (theta, &s, &c);
   pn = ov_dot (p, n);
   r = ov_lin2 (c, p, o_mul (pn, o_sub (oc (1.0), c)), n);
   r = ov_add (r, ov_lin2 (s, ov_cross (n, p), oc (0.0), n));
   return ov_unit (r);
}
/* M x with M(az, el): value m, derivatives dm[0] (az), dm[1] (el) */
static odv ov_mat (const double m[9], const double dm[2][9], odv x)
{
   odv r;
   od *out[3];
   int row;
   out[0] = &r.x; out[1] = &r.y; out[2] = &r.z;
   for (row = 0; row < 3; row++)
     {
        const double *a = m + 3 * row;
        od o = o_add (o_add (o_scale (a[0], x.x), o_scale (a[1], x.y)), o_scale (a[2], x.z));
        PH (1);
        o.d[0] += dm[0][3 * row] * x.x.v + dm[0][3 * row + 1] * x.y.v + dm[0][3 * row + 2] * x.z.v;
        o.d[1] += dm[1][3 * row] * x.x.v + dm[1][3 * row + 1] * x.y.v + dm[1][3 * row + 2] * x.z.v;
        PH (0);
        *out[row] = o;
     }
   return r;
}
static odv to_osac (const mxg_diff_shell *ds, int k)
{
   odv t = ov (orc_v (ds->to_osac[k][0], ds->to_osac[k][1], ds->to_osac[k][2]));
   t.y.d[2] = -1.0;
   t.z.d[3] = -1.0;
   t.x.d[4] = -1.0;
   return t;
}

/* reflectivity derivative by the incidence cosine: Richardson's
 * extrapolation of central differences with steps h and h/2 */
static double reflectivity_slope (double c, double beta, double delta)
{
   double h = 1e-3 * c, d1, d2;
   d1 = (orc_reflectivity (c + h, beta, delta) - orc_reflectivity (c - h, beta, delta)) / (2 * h);
   h *= 0.5;
   d2 = (orc_reflectivity (c + h, beta, delta) - orc_reflectivity (c - h, beta, delta)) / (2 * h);
   return (4 * d2 - d1) / 3;
}

static int conic (double a, double b, double cc, odv *x0, odv p, double xmin, double xmax)
{
   od t_plus = o_div (o_neg (x0->x), p.x), t_minus, t, x_y, x_z, alpha, beta, gamma;
   double x_plus, x_minus;
   int plus_ok, minus_ok;

   x_y = o_add (x0->y, o_mul (t_plus, p.y));
   x_z = o_add (x0->z, o_mul (t_plus, p.z));
   alpha = o_addc (o_scale (a, o_mul (p.x, p.x)), -1.0);
   beta = o_sub (o_scale (b, p.x), o_scale (2.0, o_add (o_mul (p.y, x_y), o_mul (p.z, x_z))));
   gamma = o_sub (o_sub (oc (cc), o_mul (x_z, x_z)), o_mul (x_y, x_y));
   if (alpha.v == 0.0)
     {
        if (beta.v == 0.0) return -1;
        t_plus = t_minus = o_neg (o_div (gamma, beta));
     }
   else
     {
        /* the stable roots of jdmath/src/qroot.c */
        od bsqr = o_mul (beta, beta), ac4 = o_scale (4.0, o_mul (alpha, gamma));
        od nb2a = o_neg (o_div (beta, o_scale (2.0, alpha)));
        if (bsqr.v > ac4.v)
          {
             od factor = o_addc (o_sqrt (o_sub (oc (1.0), o_div (ac4, bsqr))), 1.0);
             t_plus = o_div (o_scale (-2.0, gamma), o_mul (beta, factor));
             t_minus = o_mul (nb2a, factor);
          }
        else if (bsqr.v == ac4.v)
          t_plus = t_minus = nb2a;
        else
          return -1;
     }
   x_plus = p.x.v * t_plus.v;
   x_minus = p.x.v * t_minus.v;
   plus_ok = (x_plus >= xmin) && (x_plus < xmax);
   minus_ok = (x_minus >= xmin) && (x_minus < xmax);
   if (plus_ok)
     t = (minus_ok && (x_minus > x_plus)) ? t_minus : t_plus;
   else if (minus_ok)
     t = t_minus;
   else
     return -1;
   x0->x = o_mul (p.x, t);
   x0->y = o_add (x_y, o_mul (p.y, t));
   x0->z = o_add (x_z, o_mul (p.z, t));
   return 0;
}

static int reflect (const orc_ctx *c, const mxg_shell_cfg *s, int k, odv x, odv *p, od *w,
                    double energy, double beta, double delta, double corr, orc_stream *st)
{
   const mxg_hrma_cfg *h = &c->cfg->hrma;
   odv n;
   od pdn;
   n.x = o_addc (o_scale (s->conic_a[k] - 1, x.x), 0.5 * s->conic_b[k]);
   n.y = o_neg (x.y);
   n.z = o_neg (x.z);
   n = ov_unit (n);
   if (h->use_blur)
     {
        od len = o_sqrt (o_add (o_mul (n.y, n.y), o_mul (n.z, n.z)));
        odv perp;
        double phi;
        perp.x = oc (0.0);
        perp.y = o_div (n.z, len);
        perp.z = o_neg (o_div (n.y, len));
        phi = 2.0 * PI * orc_uniform (st);
        perp = ov_rotate (perp, n, oc (phi));
        phi = s->blur[k] * (1.0 / 3600.0 * PI / 180.0) * orc_gaussian (st);
        n = ov_rotate (n, perp, oc (phi));
     }
   pdn = ov_dot (*p, n);
   if (h->ideal == 0)
     {
        od cabs = o_fabs (pdn), R;
        double slope;
        /* the op count leaves out the finite-difference slope (phase 3):
         * it is this restatement's way to the derivative, not the path's */
        PH (3);
        slope = reflectivity_slope (cabs.v, beta, delta);
        PH (0);
        R = o_scale (slope, cabs);
        R.v = orc_reflectivity (cabs.v, beta, delta);
        *w = o_mul (*w, o_scale (corr, R));
     }
   *p = ov_lin2 (oc (1.0), *p, o_scale (-2.0, pdn), n);
   if (h->use_wfold)
     {
        const mxg_wfold *wf = &s->wfold[k];
        double r = orc_uniform (st), g, slope = 0.0;
        od dg;
        g = orc_wfold_interp (c, wf, energy, -pdn.v, r);
        if (wf->n_arrays > 1)
          {
             /* the interpolation is linear in e_alpha between two tables */
             const double *ea = c->da + wf->e_alphas_off;
             const double *arr = c->da + wf->arrays_off;
             unsigned int i = orc_binary_search_d (energy * -pdn.v, ea, wf->n_arrays);
             if (i == wf->n_arrays) i--;
             if (i == 0) i++;
             slope = (orc_fold_theta (c, arr + i * MXG_FOLD_STRIDE, r)
                      - orc_fold_theta (c, arr + (i - 1) * MXG_FOLD_STRIDE, r)) / (ea[i] - ea[i - 1]) * energy;
          }
        dg = o_scale (-slope, pdn);
        dg.v = g;
        dg = o_scale (s->scat_factor[k], dg);
        if (dg.v > PI / 4) return -1;
        if (orc_uniform (st) < 0.5) dg = o_neg (dg);
        *p = ov_rotate (*p, ov_cross (*p, n), dg);
     }
   return 0;
}

static void pixel (const mxg_image_spec *sp, double y, double z, long *i0, long *j0,
                   double *fu, double *fv)
{
   double u = (y - sp->center_y) / sp->pixel_size + 0.5 * sp->npix - 0.5;
   double v = (z - sp->center_z) / sp->pixel_size + 0.5 * sp->npix - 0.5;
   double a = floor (u), b = floor (v);
   if (a < -2) a = -2;
   if (b < -2) b = -2;
   if (a > sp->npix + 1.0) a = sp->npix + 1.0;
   if (b > sp->npix + 1.0) b = sp->npix + 1.0;
   *i0 = (long) a;
   *j0 = (long) b;
   *fu = u - a;
   *fv = v - b;
}

/* photons lo..hi-1 of n: their state, and their splat added to image.  The
 * optical constants are looked up once a photon hits the paraboloid (the
 * values do not depend on where). */
static void diff_range (const orc_ctx *c, const mxg_diff_cfg *d, unsigned int lo, unsigned int hi,
                        unsigned int n, unsigned int tag_start, const mxg_image_spec *sp,
                        double *image, double *w_out, double *y_out, double *z_out, double *dw,
                        double *dy, double *dz, unsigned char *shell_out)
{
   const mxg_hrma_cfg *h = &c->cfg->hrma;
   const mxg_source_cfg *src = &c->cfg->src;
   unsigned int i;
   int j;

   for (i = lo; i < hi; i++)
     {
        orc_stream ss, st;
        double energy, r, radius, theta, beta = 0, delta = 1, corr = 1;
        unsigned int si = 0, quad, k;
        const mxg_shell_cfg *s;
        const mxg_diff_shell *ds;
        odv x, p, t[2];
        od w = oc (h->ideal ? 1.0 : h->vignetting), yf = oc (0), zf = oc (0);
        int alive = 1;

        SET_PART (0);
        PH (0);
        orc_stream_init (&ss, c->seed, tag_start + i, MXG_STAGE_SOURCE, 0);
        if (src->spectrum == MXG_SPECTRUM_FLAT)
          energy = src->emin + (src->emax - src->emin) * orc_uniform (&ss);
        else
          energy = orc_interpolate_d (orc_uniform (&ss), c->da + src->file_cum_off,
                                      c->da + src->file_e_off, src->file_n);
        orc_stream_init (&st, c->seed, tag_start + i, MXG_STAGE_MIRROR, 0);
        for (;;)
          {
             int found = 0;
             r = orc_uniform (&st);
             for (k = 0; k < MXG_NUM_SHELLS; k++)
               if (r < h->shell[k].area_fraction) { si = k; found = 1; break; }
             if (found) break;
          }
        s = h->shell + si;
        ds = d->shell + si;
        radius = s->min_radius + (s->max_radius - s->min_radius) * orc_uniform (&st);
        do
          {
             theta = orc_uniform (&st);
             quad = (unsigned int) (4.0 * theta);
          }
        while (0 == (s->shutter_bitmap & (1u << quad)));
        theta = (2.0 * PI) * (theta - 1.0 / 8.0);
        t[0] = to_osac (ds, 0);
        t[1] = to_osac (ds, 1);
        x.x = oc (ds->front_position);
        x.x.d[4] = 1.0;
        x.z = o_sub (oc (radius * cos (theta)), t[0].z);
        x.y = o_sub (oc (radius * sin (theta)), t[0].y);
        p = ov (orc_v (src->p[0], src->p[1], src->p[2]));
        if (src->distance > 0.0)
          p = ov_unit (ov_lin2 (oc (1.0), x, oc (src->distance), p));
        if (h->use_struts && orc_strut_shadow (h, ovval (x), ovval (p), h->pre_struts, 2))
          alive = 0;
        for (k = 0; (k < 2) && alive; k++)
          {
             x = ov_add (x, t[k]);
             x = ov_mat (ds->fwd[k], ds->dfwd[k], x);
             p = ov_mat (ds->fwd[k], ds->dfwd[k], p);
             if (0 != conic (s->conic_a[k], s->conic_b[k], s->conic_c[k], &x, p, s->conic_xmin[k],
                             s->conic_xmax[k]))
               {
                  alive = 0;
                  break;
               }
             if ((k == 0) && (h->opt_n != 0) && (h->ideal == 0))
               {
                  /* the optical constants, needed from the first reflection on */
                  float ef = (float) energy;
                  SET_PART (4);
                  PH (0);
                  beta = orc_interpolate_f (ef, c->fa + h->opt_e_off, c->fa + h->opt_beta_off, h->opt_n);
                  delta = orc_interpolate_f (ef, c->fa + h->opt_e_off, c->fa + h->opt_delta_off, h->opt_n);
                  if (h->use_scale)
                    corr = sqrt (orc_interpolate_f (ef, c->fa + s->corr_e_off, c->fa + s->corr_f_off,
                                                    s->corr_n));
               }
             SET_PART (4);
             PH (0);
             if (0 != reflect (c, s, k, x, &p, &w, energy, beta, delta, corr, &st))
               {
                  alive = 0;
                  break;
               }
             p = ov_mat (ds->bwd[k], ds->dbwd[k], p);
             x = ov_mat (ds->bwd[k], ds->dbwd[k], x);
             x = ov_sub (x, t[k]);
             if (h->use_struts
                 && orc_strut_shadow (h, ovval (x), ovval (p), k ? h->post_struts : h->cap_struts, 2))
               alive = 0;
          }
        if (alive)
          {
             od tt = o_div (o_neg (x.x), p.x);
             yf = o_add (x.y, o_mul (tt, p.y));
             zf = o_add (x.z, o_mul (tt, p.z));
          }
        else
          w = oc (0.0);
        w_out[i] = w.v; y_out[i] = yf.v; z_out[i] = zf.v;
        for (j = 0; j < NP; j++)
          {
             dw[(size_t) j * n + i] = w.d[j];
             dy[(size_t) j * n + i] = yf.d[j];
             dz[(size_t) j * n + i] = zf.d[j];
          }
        shell_out[i] = (unsigned char) si;
        if (w.v != 0.0)
          {
             long i0, j0, cy, cz;
             double fu, fv;
             pixel (sp, yf.v, zf.v, &i0, &j0, &fu, &fv);
             for (cz = 0; cz < 2; cz++)
               for (cy = 0; cy < 2; cy++)
                 {
                    long iy = i0 + cy, iz = j0 + cz;
                    double b = (cy ? fu : 1 - fu) * (cz ? fv : 1 - fv);
                    if ((iy < 0) || (iz < 0) || (iy >= (long) sp->npix) || (iz >= (long) sp->npix))
                      continue;
                    image[iz * (long) sp->npix + iy] += w.v * b;
                 }
          }
     }
}

int orc_diff_forward (const orc_ctx *c, const mxg_diff_cfg *d, unsigned int n,
                      unsigned int tag_start, const mxg_image_spec *sp, double *image,
                      double *w_out, double *y_out, double *z_out, double *dw, double *dy,
                      double *dz, unsigned char *shell_out)
{
   memset (image, 0, (size_t) sp->npix * sp->npix * sizeof (double));
   diff_range (c, d, 0, n, n, tag_start, sp, image, w_out, y_out, z_out, dw, dy, dz, shell_out);
   return 0;
}

/* orc_diff_forward on nthreads OpenMP threads (0: all), each over a
 * contiguous range of photons into an image of its own; the images are
 * summed in thread order, so the result depends on nthreads only through
 * the FP64 summation order of the image. */
int orc_diff_forward_parallel (const orc_ctx *c, const mxg_diff_cfg *d, unsigned int n,
                               unsigned int tag_start, const mxg_image_spec *sp, double *image,
                               double *w_out, double *y_out, double *z_out, double *dw,
                               double *dy, double *dz, unsigned char *shell_out, int nthreads)
{
   const size_t npix2 = (size_t) sp->npix * sp->npix;
   double *img;
   int nt = nthreads, t;
#ifdef _OPENMP
   if (nt <= 0) nt = omp_get_max_threads ();
#else
   nt = 1;
#endif
   img = (double *) calloc ((size_t) nt * npix2, sizeof (double));
   if (img == NULL) return -1;
#pragma omp parallel for num_threads (nt) schedule (static, 1)
   for (t = 0; t < nt; t++)
     {
        unsigned int lo = (unsigned int) ((unsigned long long) n * t / nt);
        unsigned int hi = (unsigned int) ((unsigned long long) n * (t + 1) / nt);
        diff_range (c, d, lo, hi, n, tag_start, sp, img + (size_t) t * npix2, w_out, y_out, z_out,
                    dw, dy, dz, shell_out);
     }
   memset (image, 0, npix2 * sizeof (double));
   for (t = 0; t < nt; t++)
     {
        size_t k;
        for (k = 0; k < npix2; k++) image[k] += img[(size_t) t * npix2 + k];
     }
   free (img);
   return 0;
}

/* the gradient terms of photons lo..hi-1 of n added to grad */
static void adjoint_range (unsigned int lo, unsigned int hi, unsigned int n, const double *w,
                           const double *y, const double *z, const double *dw, const double *dy,
                           const double *dz, const unsigned char *shell,
                           const mxg_image_spec *sp, const double *g, double *grad)
{
   unsigned int i;
   int j;
   for (i = lo; i < hi; i++)
     {
        long i0, j0, cy, cz;
        double fu, fv, gs = 0, gu = 0, gv = 0;
        if (w[i] == 0.0) continue;
        pixel (sp, y[i], z[i], &i0, &j0, &fu, &fv);
        for (cz = 0; cz < 2; cz++)
          for (cy = 0; cy < 2; cy++)
            {
               long iy = i0 + cy, iz = j0 + cz;
               double gc, bu, bv, b;
               if ((iy < 0) || (iz < 0) || (iy >= (long) sp->npix) || (iz >= (long) sp->npix))
                 continue;
               gc = g[iz * (long) sp->npix + iy];
               b = (cy ? fu : 1 - fu) * (cz ? fv : 1 - fv);
               bu = (cy ? 1 : -1) * (cz ? fv : 1 - fv);
               bv = (cy ? fu : 1 - fu) * (cz ? 1 : -1);
               gs += gc * b;
               gu += gc * bu;
               gv += gc * bv;
            }
        for (j = 0; j < NP; j++)
          grad[shell[i] * NP + j] += gs * dw[(size_t) j * n + i]
            + w[i] / sp->pixel_size * (gu * dy[(size_t) j * n + i] + gv * dz[(size_t) j * n + i]);
     }
}

int orc_diff_adjoint (unsigned int n, const double *w, const double *y, const double *z,
                      const double *dw, const double *dy, const double *dz,
                      const unsigned char *shell, const mxg_image_spec *sp, const double *g,
                      double *grad)
{
   memset (grad, 0, MXG_NUM_SHELLS * NP * sizeof (double));
   adjoint_range (0, n, n, w, y, z, dw, dy, dz, shell, sp, g, grad);
   return 0;
}

/* orc_diff_adjoint on nthreads OpenMP threads (0: all), per-thread partial
 * gradients summed in thread order */
int orc_diff_adjoint_parallel (unsigned int n, const double *w, const double *y, const double *z,
                               const double *dw, const double *dy, const double *dz,
                               const unsigned char *shell, const mxg_image_spec *sp,
                               const double *g, double *grad, int nthreads)
{
   enum { NG = MXG_NUM_SHELLS * NP };
   double *part;
   int nt = nthreads, t, k;
#ifdef _OPENMP
   if (nt <= 0) nt = omp_get_max_threads ();
#else
   nt = 1;
#endif
   part = (double *) calloc ((size_t) nt * NG, sizeof (double));
   if (part == NULL) return -1;
#pragma omp parallel for num_threads (nt) schedule (static, 1)
   for (t = 0; t < nt; t++)
     {
        unsigned int lo = (unsigned int) ((unsigned long long) n * t / nt);
        unsigned int hi = (unsigned int) ((unsigned long long) n * (t + 1) / nt);
        adjoint_range (lo, hi, n, w, y, z, dw, dy, dz, shell, sp, g, part + (size_t) t * NG);
     }
   for (k = 0; k < NG; k++)
     {
        grad[k] = 0.0;
        for (t = 0; t < nt; t++) grad[k] += part[(size_t) t * NG + k];
     }
   free (part);
   return 0;
}
