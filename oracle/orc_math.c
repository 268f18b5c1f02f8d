/* orc_math.c -- oracle restatements of the jdmath primitives on the path
 * and of the per-photon random streams.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   binary search / interpolation: jdmath/src/finterpo.c:37-85, 134-193;
 *                                  jdmath/src/dinterpo.c:30-79
 *   quadratic root:                jdmath/src/qroot.c:36-70
 *   vectors / rotations:           jdmath/src/vector.c:30-212, rotate.c:81-130
 *   complex arithmetic:            jdmath/src/complex.c:40-230
 *   Fresnel reflectivity:          marx/libsrc/reflect.c:39-77
 *   WFold interpolation:           marx/libsrc/wfold.c:304-369
 *   polar Gaussian, exponential:   jdmath/src/gaussrnd.c:30-67
 *   Philox4x32-10:                 Salmon et al., SC'11 (Random123)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "oracle_int.h"

/* ---- Philox4x32-10 -------------------------------------------------------- */

static void mulhilo32 (unsigned int a, unsigned int b, unsigned int *hi, unsigned int *lo)
{
   unsigned long long p = (unsigned long long) a * (unsigned long long) b;
   *hi = (unsigned int) (p >> 32);
   *lo = (unsigned int) p;
}

void orc_philox4x32_10 (const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4])
{
   unsigned int c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
   unsigned int k0 = key[0], k1 = key[1];
   int r;

   for (r = 0; r < 10; r++)
     {
        unsigned int hi0, lo0, hi1, lo1;
        if (r > 0)
          {
             k0 += 0x9E3779B9u;
             k1 += 0xBB67AE85u;
          }
        mulhilo32 (0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32 (0xCD9E8D57u, c2, &hi1, &lo1);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
     }
   out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_stream_init (orc_stream *s, unsigned long long seed, unsigned int tag,
                      unsigned int stage, unsigned int source_id)
{
   s->key[0] = (unsigned int) seed;
   s->key[1] = (unsigned int) (seed >> 32);
   s->ctr[0] = tag;
   s->ctr[1] = stage;
   s->ctr[2] = 0;
   s->ctr[3] = source_id;
   s->pos = 4;
   s->have_gauss = 0;
   s->g2 = 0.0;
   s->inited = 1;
}

double orc_uniform (orc_stream *s)
{
   if (s->pos == 4)
     {
        orc_philox4x32_10 (s->ctr, s->key, s->buf);
        s->ctr[2]++;
        s->pos = 0;
     }
   /* same mapping as JDMrandom: u32 / (2^32 - 1), both ends included */
   return ((double) s->buf[s->pos++] * (1.0 / ((double) 0xFFFFFFFFu)));
}

double orc_gaussian (orc_stream *s)
{
   double g1, g2, g, f;

   if (s->have_gauss)
     {
        s->have_gauss = 0;
        return s->g2;
     }
   s->have_gauss = 1;
   do
     {
        g1 = 2.0 * orc_uniform (s) - 1.0;
        g2 = 2.0 * orc_uniform (s) - 1.0;
        g = g1 * g1 + g2 * g2;
     }
   while ((g >= 1.0) || (g == 0.0));
   f = sqrt (-2.0 * log (g) / g);
   s->g2 = g2 * f;
   return g1 * f;
}

double orc_expn (orc_stream *s)
{
   double r;
   do r = orc_uniform (s); while (r == 0.0);
   return -log (r);
}

void orc_stream_uniforms (unsigned long long seed, unsigned int tag, unsigned int stage,
                          unsigned int source_id, double *u, unsigned int n)
{
   orc_stream s;
   unsigned int i;
   orc_stream_init (&s, seed, tag, stage, source_id);
   for (i = 0; i < n; i++) u[i] = orc_uniform (&s);
}

void orc_stream_gaussians (unsigned long long seed, unsigned int tag, unsigned int stage,
                           unsigned int source_id, double *g, unsigned int n)
{
   orc_stream s;
   unsigned int i;
   orc_stream_init (&s, seed, tag, stage, source_id);
   for (i = 0; i < n; i++) g[i] = orc_gaussian (&s);
}

/* ---- searches and interpolation ------------------------------------------- */

/* index of the first element >= x, n if none; an exact hit returns its index */
unsigned int orc_binary_search_f (float x, const float *xp, unsigned int n)
{
   unsigned int n0 = 0, n1 = n, n2;
   while (n1 > n0 + 1)
     {
        n2 = (n0 + n1) / 2;
        if (xp[n2] >= x)
          {
             if (xp[n2] == x) return n2;
             n1 = n2;
          }
        else n0 = n2;
     }
   if (x >= xp[n0]) return n1;
   return n0;
}

float orc_interpolate_f (float x, const float *xp, const float *yp, unsigned int n)
{
   unsigned int n0, n1;
   double x0, x1;

   if (n == 1) return yp[0];
   n1 = orc_binary_search_f (x, xp, n);
   n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1]))
     return yp[n1];
   if (n1 == n)
     {
        n1--;
        n0--;
     }
   if (n1 == 0)
     n0 = 1;
   x0 = xp[n0];
   x1 = xp[n1];
   if (x1 == x0) return yp[n1];
   /* the y difference is a float subtraction, the rest double */
   return (float) (yp[n0] + (yp[n1] - yp[n0]) / (x1 - x0) * (x - x0));
}

unsigned int orc_binary_search_d (double x, const double *xp, unsigned int n)
{
   unsigned int n0 = 0, n1 = n, n2;
   while (n1 > n0 + 1)
     {
        n2 = (n0 + n1) / 2;
        if (xp[n2] >= x)
          {
             if (xp[n2] == x) return n2;
             n1 = n2;
          }
        else n0 = n2;
     }
   if (x >= xp[n0]) return n1;
   return n0;
}

double orc_interpolate_d (double x, const double *xp, const double *yp, unsigned int n)
{
   unsigned int n0, n1;
   double x0, x1;

   if (n == 1) return yp[0];
   n1 = orc_binary_search_d (x, xp, n);
   n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1]))
     return yp[n1];
   if (n1 == n)
     {
        n1--;
        n0--;
     }
   if (n1 == 0)
     n0 = 1;
   x0 = xp[n0];
   x1 = xp[n1];
   if (x1 == x0) return yp[n1];
   return yp[n0] + (yp[n1] - yp[n0]) / (x1 - x0) * (x - x0);
}

int orc_interpolate_n_fvector (const float *xp, float **yp, unsigned int n,
                               const float *oldxp, float **oldyp, unsigned int oldn,
                               unsigned int n_yp)
{
   const float *xpmax, *oldxpmax;
   unsigned int count, yp_count, i;

   if (oldn < 2) return -1;
   xpmax = xp + n;
   oldxpmax = oldxp + (oldn - 1);
   oldxp++;
   count = 1;
   yp_count = 0;
   while (xp < xpmax)
     {
        double x = *xp++, x_0, x_1, dx_10;
        while ((oldxp < oldxpmax) && (x > *oldxp))
          {
             oldxp++;
             count++;
          }
        x_0 = *(oldxp - 1);
        x_1 = *oldxp;
        dx_10 = x_1 - x_0;
        if (dx_10 == 0.0)
          {
             for (i = 0; i < n_yp; i++)
               yp[i][yp_count] = oldyp[i][count - 1];
          }
        else for (i = 0; i < n_yp; i++)
          {
             double y_0 = oldyp[i][count - 1];
             double y_1 = oldyp[i][count];
             yp[i][yp_count] = (float) (y_0 + (y_1 - y_0) * (x - x_0) / dx_10);
          }
        yp_count++;
     }
   return 0;
}

int orc_quadratic_root (double a, double b, double c, double *rplus, double *rminus)
{
   double bsqr, ac4, factor, neg_b_over_2a;
   if (a == 0.0) return -1;
   bsqr = b * b;
   ac4 = a * c * 4;
   neg_b_over_2a = -b / (2.0 * a);
   if (bsqr > ac4)
     {
        factor = 1.0 + sqrt (1.0 - ac4 / bsqr);
        *rplus = -2.0 * c / (b * factor);
        *rminus = neg_b_over_2a * factor;
        return 1;
     }
   if (bsqr == ac4)
     {
        *rplus = *rminus = neg_b_over_2a;
        return 1;
     }
   *rplus = neg_b_over_2a;
   *rminus = sqrt (c / a) * sqrt (1.0 - bsqr / ac4);
   return 0;
}

/* ---- vectors ------------------------------------------------------------------ */

orc_vec orc_v (double x, double y, double z)
{
   orc_vec a;
   a.x = x; a.y = y; a.z = z;
   return a;
}

double orc_v_length (orc_vec a)
{
   double x = fabs (a.x), y = fabs (a.y), z = fabs (a.z), tmp;
   if (z < x) { tmp = z; z = x; x = tmp; }
   if (z < y) { tmp = z; z = y; y = tmp; }
   if (z == 0.0) return 0.0;
   x = x / z;
   y = y / z;
   z = z * sqrt (1.0 + x * x + y * y);
   return z;
}

void orc_v_normalize (orc_vec *a)
{
   double len;
   if ((len = orc_v_length (*a)) != 0.0)
     {
        a->x = a->x / len;
        a->y = a->y / len;
        a->z = a->z / len;
     }
}

orc_vec orc_v_unit (orc_vec a)
{
   orc_v_normalize (&a);
   return a;
}

orc_vec orc_v_cross (orc_vec a, orc_vec b)
{
   orc_vec c;
   c.z = a.x * b.y - a.y * b.x;
   c.x = a.y * b.z - a.z * b.y;
   c.y = a.z * b.x - a.x * b.z;
   return c;
}

double orc_v_dot (orc_vec a, orc_vec b)
{
   return a.x * b.x + a.y * b.y + a.z * b.z;
}

orc_vec orc_v_ax1_bx2 (double a, orc_vec x1, double b, orc_vec x2)
{
   orc_vec c;
   c.x = a * x1.x + b * x2.x;
   c.y = a * x1.y + b * x2.y;
   c.z = a * x1.z + b * x2.z;
   return c;
}

orc_vec orc_v_ax1_bx2_cx3 (double a, orc_vec x1, double b, orc_vec x2, double c, orc_vec x3)
{
   orc_vec d;
   d.x = a * x1.x + b * x2.x + c * x3.x;
   d.y = a * x1.y + b * x2.y + c * x3.y;
   d.z = a * x1.z + b * x2.z + c * x3.z;
   return d;
}

orc_vec orc_v_sum (orc_vec a, orc_vec b)
{
   return orc_v (a.x + b.x, a.y + b.y, a.z + b.z);
}

orc_vec orc_v_diff (orc_vec a, orc_vec b)
{
   return orc_v (a.x - b.x, a.y - b.y, a.z - b.z);
}

orc_vec orc_v_rotate_vector1 (orc_vec p, orc_vec n, double c, double s)
{
   double pn = orc_v_dot (p, n);
   return orc_v_ax1_bx2_cx3 (c, p, pn * (1.0 - c), n, s, orc_v_cross (n, p));
}

orc_vec orc_v_rotate_unit_vector1 (orc_vec p, orc_vec n, double c, double s)
{
   orc_vec u = orc_v_rotate_vector1 (p, n, c, s);
   orc_v_normalize (&u);
   return u;
}

orc_vec orc_v_rotate_unit_vector (orc_vec p, orc_vec n, double theta)
{
   return orc_v_rotate_unit_vector1 (p, n, cos (theta), sin (theta));
}

orc_vec orc_m3_vmul (const double a[9], orc_vec v)
{
   orc_vec b;
   b.x = a[0] * v.x + a[1] * v.y + a[2] * v.z;
   b.y = a[3] * v.x + a[4] * v.y + a[5] * v.z;
   b.z = a[6] * v.x + a[7] * v.y + a[8] * v.z;
   return b;
}

/* ---- complex arithmetic and Fresnel reflectivity ---------------------------- */

typedef struct { double r, i; } cplx;

static cplx c_mul (cplx a, cplx b)
{
   cplx z;
   z.r = a.r * b.r - a.i * b.i;
   z.i = a.r * b.i + a.i * b.r;
   return z;
}

static cplx c_div (cplx z1, cplx z2)
{
   cplx z;
   double r1 = z1.r, i1 = z1.i, r2 = z2.r, i2 = z2.i, ratio, denom;
   if (fabs (r2) > fabs (i2))
     {
        ratio = i2 / r2;
        denom = r2 + i2 * ratio;
        z.r = (r1 + ratio * i1) / denom;
        z.i = (i1 - r1 * ratio) / denom;
     }
   else
     {
        ratio = r2 / i2;
        denom = r2 * ratio + i2;
        z.r = (r1 * ratio + i1) / denom;
        z.i = (i1 * ratio - r1) / denom;
     }
   return z;
}

static cplx c_a_bz (double a, double b, cplx z1)
{
   cplx z;
   z.r = a + b * z1.r;
   z.i = b * z1.i;
   return z;
}

static cplx c_az1_bz2 (double a, cplx z1, double b, cplx z2)
{
   cplx z;
   z.r = a * z1.r + b * z2.r;
   z.i = a * z1.i + b * z2.i;
   return z;
}

static double c_abs (cplx z)
{
   double fr = fabs (z.r), fi = fabs (z.i), ratio;
   if (fr > fi)
     {
        ratio = z.i / z.r;
        return fr * sqrt (1.0 + ratio * ratio);
     }
   if (fi == 0.0) return 0.0;
   ratio = z.r / z.i;
   return fi * sqrt (1.0 + ratio * ratio);
}

static cplx c_sqrt (cplx a)
{
   double r = c_abs (a);
   if (r == 0.0) return a;
   if (a.r >= 0.0)
     {
        a.r = sqrt (0.5 * (r + a.r));
        a.i = 0.5 * a.i / a.r;
     }
   else
     {
        r = sqrt (0.5 * (r - a.r));
        a.r = 0.5 * a.i / r;
        a.i = r;
        if (a.r < 0.0)
          {
             a.r = -a.r;
             a.i = -a.i;
          }
     }
   return a;
}

/* Polarisation-averaged Fresnel reflectivity, Jackson (1975) eq 7.39/7.41,
 * index n = 1 - delta + i beta, cos_theta >= 0 w.r.t. the normal. */
double orc_reflectivity (double cos_theta, double beta, double delta)
{
   cplx n, root, nsqr, e_perp, e_par, num, den;
   double sin_theta;

   n.r = (1.0 - delta);
   n.i = beta;
   sin_theta = sqrt (1.0 - cos_theta * cos_theta);
   nsqr = c_mul (n, n);
   root = c_sqrt (c_a_bz (-sin_theta * sin_theta, 1.0, nsqr));
   num = c_a_bz (cos_theta, -1.0, root);
   den = c_a_bz (cos_theta, 1.0, root);
   e_perp = c_div (num, den);
   num = c_az1_bz2 (cos_theta, nsqr, -1.0, root);
   den = c_az1_bz2 (cos_theta, nsqr, 1.0, root);
   e_par = c_div (num, den);
   return 0.5 * (e_par.r * e_par.r + e_par.i * e_par.i
                 + e_perp.r * e_perp.r + e_perp.i * e_perp.i);
}

/* ---- WFold ----------------------------------------------------------------------- */

double orc_fold_theta (const orc_ctx *c, const double *fa, double p)
{
   double p_min = fa[1], delta_p = fa[2], p_max = fa[3];
   unsigned int num = (unsigned int) fa[6];
   const float *t = c->fa + (unsigned int) fa[7];
   double delta_i;
   unsigned int i;

   if (p < p_min) return 0.0;
   if (p > p_max) return pow (fa[4] * (1.0 - p), fa[5]);
   delta_i = (p - p_min) / delta_p;
   i = (unsigned int) delta_i;
   if (i + 1 >= num) return (double) t[num - 1];
   delta_i -= (double) i;
   return (1.0 - delta_i) * t[i] + delta_i * t[i + 1];
}

double orc_wfold_interp (const orc_ctx *c, const mxg_wfold *w, double energy,
                         double sin_alpha, double r)
{
   const double *arrays = c->da + w->arrays_off;
   const double *e_alphas = c->da + w->e_alphas_off;
   double e_alpha, theta_0, theta_1;
   unsigned int i;

   if (w->n_arrays == 0) return 0.0;
   if (w->n_arrays == 1) return orc_fold_theta (c, arrays, r);
   e_alpha = energy * sin_alpha;
   i = orc_binary_search_d (e_alpha, e_alphas, w->n_arrays);
   if (i == w->n_arrays) i--;
   if (i == 0) i++;
   theta_0 = orc_fold_theta (c, arrays + (i - 1) * MXG_FOLD_STRIDE, r);
   theta_1 = orc_fold_theta (c, arrays + i * MXG_FOLD_STRIDE, r);
   return theta_0 + (theta_1 - theta_0) * (e_alpha - e_alphas[i - 1])
     / (e_alphas[i] - e_alphas[i - 1]);
}
