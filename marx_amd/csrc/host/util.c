/* util.c -- errors, configuration arenas and the jdmath-style vector helpers
 * used at init time.
 *
 * The vector helpers restate the arithmetic of jdmath/src/vector.c:30-330 and
 * jdmath/src/rotate.c:32-130 operation for operation, because geometry built
 * here (shell tilt matrices, detector bases, source directions) feeds the
 * per-photon trace and must round exactly as the reference's does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <math.h>

#include "mxg_host.h"

int Mxg_Verbose = 0;
mxg_config Mxg_Cfg;
unsigned int Mxg_Cfg_Generation = 1;

static char Last_Error[1024];

void mxg_error (const char *fmt, ...)
{
   va_list ap;
   va_start (ap, fmt);
   vsnprintf (Last_Error, sizeof (Last_Error), fmt, ap);
   va_end (ap);
   fprintf (stderr, "***marx_gpu: %s\n", Last_Error);
}

const char *mxg_last_error (void)
{
   return Last_Error;
}

void mxg_message (int level, const char *fmt, ...)
{
   va_list ap;
   if (level > Mxg_Verbose)
     return;
   va_start (ap, fmt);
   vfprintf (stdout, fmt, ap);
   va_end (ap);
}

/* ---- arenas ------------------------------------------------------------ */

typedef struct
{
   void *data;
   unsigned int n, cap;
   size_t elsize;
}
Arena_Type;

static Arena_Type F_Arena = {NULL, 0, 0, sizeof (float)};
static Arena_Type D_Arena = {NULL, 0, 0, sizeof (double)};

static unsigned int arena_reserve (Arena_Type *a, unsigned int n)
{
   unsigned int off = a->n;
   if (n > 0xF0000000u - a->n)
     {
        mxg_error ("table arena overflow");
        return (unsigned int) -1;
     }
   if (a->n + n > a->cap)
     {
        unsigned int cap = a->cap ? a->cap : 4096;
        void *d;
        while (cap < a->n + n) cap *= 2;
        d = realloc (a->data, (size_t) cap * a->elsize);
        if (d == NULL)
          {
             mxg_error ("out of memory growing a table arena");
             return (unsigned int) -1;
          }
        a->data = d;
        a->cap = cap;
     }
   a->n += n;
   Mxg_Cfg_Generation++;
   return off;
}

unsigned int mxg_farena_reserve (unsigned int n)
{
   return arena_reserve (&F_Arena, n);
}

unsigned int mxg_darena_reserve (unsigned int n)
{
   return arena_reserve (&D_Arena, n);
}

unsigned int mxg_farena_push (const float *v, unsigned int n)
{
   unsigned int off = arena_reserve (&F_Arena, n);
   if ((off != (unsigned int) -1) && n)
     memcpy ((float *) F_Arena.data + off, v, n * sizeof (float));
   return off;
}

unsigned int mxg_darena_push (const double *v, unsigned int n)
{
   unsigned int off = arena_reserve (&D_Arena, n);
   if ((off != (unsigned int) -1) && n)
     memcpy ((double *) D_Arena.data + off, v, n * sizeof (double));
   return off;
}

float *mxg_farena_ptr (unsigned int off) { return (float *) F_Arena.data + off; }
double *mxg_darena_ptr (unsigned int off) { return (double *) D_Arena.data + off; }
unsigned int mxg_farena_size (void) { return F_Arena.n; }
unsigned int mxg_darena_size (void) { return D_Arena.n; }
const float *mxg_farena_base (void) { return (const float *) F_Arena.data; }
const double *mxg_darena_base (void) { return (const double *) D_Arena.data; }

int mxg_get_config (const mxg_config **cfg, const float **farena, unsigned int *nf,
                    const double **darena, unsigned int *nd)
{
   if (cfg != NULL) *cfg = &Mxg_Cfg;
   if (farena != NULL) *farena = (const float *) F_Arena.data;
   if (nf != NULL) *nf = F_Arena.n;
   if (darena != NULL) *darena = (const double *) D_Arena.data;
   if (nd != NULL) *nd = D_Arena.n;
   return 0;
}

/* ---- float-table interpolation (jdmath/src/finterpo.c:37-85) --------------- */

static unsigned int search_f (float x, const float *xp, unsigned int n)
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

/* float abscissa and table, double arithmetic, linear extrapolation */
float mxg_interpolate_f (float x, const float *xp, const float *yp, unsigned int n)
{
   unsigned int n0, n1;
   double x0, x1;
   if (n == 1) return yp[0];
   n1 = search_f (x, xp, n);
   n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1])) return yp[n1];
   if (n1 == n) { n1--; n0--; }
   if (n1 == 0) n0 = 1;
   x0 = xp[n0];
   x1 = xp[n1];
   if (x1 == x0) return yp[n1];
   return (float) (yp[n0] + (yp[n1] - yp[n0]) / (x1 - x0) * (x - x0));
}

/* ---- vectors (jdmath/src/vector.c) -------------------------------------- */

JDMVector_Type mxg_v (double x, double y, double z)
{
   JDMVector_Type a;
   a.x = x; a.y = y; a.z = z;
   return a;
}

/* Scaled length: largest component factored out before the square root. */
double mxg_v_length (JDMVector_Type a)
{
   double x = fabs (a.x), y = fabs (a.y), z = fabs (a.z), tmp;
   if (z < x) { tmp = z; z = x; x = tmp; }
   if (z < y) { tmp = z; z = y; y = tmp; }
   if (z == 0.0) return 0.0;
   x = x / z;
   y = y / z;
   return z * sqrt (1.0 + x * x + y * y);
}

void mxg_v_normalize (JDMVector_Type *a)
{
   double len = mxg_v_length (*a);
   if (len != 0.0)
     {
        a->x = a->x / len;
        a->y = a->y / len;
        a->z = a->z / len;
     }
}

JDMVector_Type mxg_v_cross (JDMVector_Type a, JDMVector_Type b)
{
   JDMVector_Type c;
   c.z = a.x * b.y - a.y * b.x;
   c.x = a.y * b.z - a.z * b.y;
   c.y = a.z * b.x - a.x * b.z;
   return c;
}

double mxg_v_dot (JDMVector_Type a, JDMVector_Type b)
{
   return a.x * b.x + a.y * b.y + a.z * b.z;
}

JDMVector_Type mxg_v_sum (JDMVector_Type a, JDMVector_Type b)
{
   return mxg_v (a.x + b.x, a.y + b.y, a.z + b.z);
}

JDMVector_Type mxg_v_diff (JDMVector_Type a, JDMVector_Type b)
{
   return mxg_v (a.x - b.x, a.y - b.y, a.z - b.z);
}

JDMVector_Type mxg_v_ax1_bx2 (double a, JDMVector_Type x1, double b, JDMVector_Type x2)
{
   return mxg_v (a * x1.x + b * x2.x, a * x1.y + b * x2.y, a * x1.z + b * x2.z);
}

/* Rotation of p about unit axis n: cos p + (p.n)(1-cos) n + sin (n x p). */
JDMVector_Type mxg_v_rotate_vector1 (JDMVector_Type p, JDMVector_Type n, double c, double s)
{
   double pn = mxg_v_dot (p, n);
   JDMVector_Type nxp = mxg_v_cross (n, p);
   double b = pn * (1.0 - c);
   return mxg_v (c * p.x + b * n.x + s * nxp.x,
                 c * p.y + b * n.y + s * nxp.y,
                 c * p.z + b * n.z + s * nxp.z);
}

JDMVector_Type mxg_v_rotate_unit_vector (JDMVector_Type p, JDMVector_Type n, double theta)
{
   JDMVector_Type u = mxg_v_rotate_vector1 (p, n, cos (theta), sin (theta));
   mxg_v_normalize (&u);
   return u;
}

JDMVector_Type mxg_v_spherical_to_vector (double r, double theta, double phi)
{
   double s = sin (theta);
   return mxg_v (r * s * cos (phi), r * s * sin (phi), r * cos (theta));
}

void mxg_v_unit_vector_to_spherical (JDMVector_Type p, double *thetap, double *phip)
{
   double theta, phi, sin_theta;

   if (fabs (p.z) >= 1.0)
     {
        *thetap = (p.z >= 1.0) ? 0.0 : PI;
        *phip = 0.0;
        return;
     }
   theta = acos (p.z);
   sin_theta = sin (theta);
   if (fabs (p.x) <= fabs (p.y))
     {
        phi = acos (p.x / sin_theta);
        if (p.y < 0.0) phi = -phi;
     }
   else
     {
        phi = asin (p.y / sin_theta);
        if (p.x < 0)
          {
             if (phi >= 0) phi = PI - phi;
             else phi = -PI - phi;
          }
     }
   *thetap = theta;
   *phip = phi;
}

/* ---- 3x3 matrices (jdmath/src/rotate.c) ---------------------------------- */

void mxg_m3_rot_z (double m[9], double theta)
{
   double c = cos (theta), s = -sin (theta);
   m[0] = c;   m[1] = s;   m[2] = 0.0;
   m[3] = -s;  m[4] = c;   m[5] = 0.0;
   m[6] = 0.0; m[7] = 0.0; m[8] = 1.0;
}

void mxg_m3_rot (double m[9], JDMVector_Type n, double theta)
{
   double c = cos (theta), s = sin (theta);
   double c1 = 1 - c;
   double n0 = n.x, n1 = n.y, n2 = n.z;
   double n01 = n0 * n1 * c1, n02 = n0 * n2 * c1, n12 = n1 * n2 * c1;

   m[0] = n0 * n0 * c1 + c;
   m[1] = n01 - s * n2;
   m[2] = n02 + s * n1;
   m[3] = n01 + s * n2;
   m[4] = n1 * n1 * c1 + c;
   m[5] = n12 - s * n0;
   m[6] = n02 - s * n1;
   m[7] = n12 + s * n0;
   m[8] = n2 * n2 * c1 + c;
}

void mxg_m3_mul (double c[9], const double a[9], const double b[9])
{
   unsigned int i, j;
   double t[9];
   for (i = 0; i < 3; i++)
     for (j = 0; j < 3; j++)
       t[3*i+j] = a[3*i] * b[j] + a[3*i+1] * b[3+j] + a[3*i+2] * b[6+j];
   memcpy (c, t, sizeof (t));
}

JDMVector_Type mxg_m3_vmul (const double a[9], JDMVector_Type v)
{
   return mxg_v (a[0] * v.x + a[1] * v.y + a[2] * v.z,
                 a[3] * v.x + a[4] * v.y + a[5] * v.z,
                 a[6] * v.x + a[7] * v.y + a[8] * v.z);
}

/* ---- bucket indices of search tables ------------------------------------------ */

/* The bucket index of a search table (mxg_bucket in marx_gpu_cfg.h) with at
 * most nbmax buckets: keys are the bits of the float values, or of the
 * values widened to double (double_keys).  A table that is not strictly
 * increasing and positive gets ok = 0 (searched whole). */
static unsigned long long table_key (float v, int double_keys)
{
   if (double_keys)
     {
        double d = (double) v;
        unsigned long long k;
        memcpy (&k, &d, sizeof (k));
        return k;
     }
   else
     {
        unsigned int k;
        memcpy (&k, &v, sizeof (k));
        return k;
     }
}

/* the linear bucket key of x >= b->x0 (see mxg_lbucket) */
static unsigned int lbucket_key (const mxg_lbucket *b, double x)
{
   double t = (x - b->x0) * b->scale;
   return (t < (double) b->nb) ? (unsigned int) t : b->nb - 1;
}

int mxg_build_lbucket (const double *xp, unsigned int n, unsigned int nbmax, mxg_lbucket *b)
{
   unsigned int i, j, nb;
   double range, *cnt;

   memset (b, 0, sizeof (*b));
   if ((n < 2) || (n > (1u << 24)))
     return 0;
   for (i = 1; i < n; i++)
     if (!(xp[i] > xp[i - 1]))
       return 0;
   range = xp[n - 1] - xp[0];
   if (!(range > 0.0) || !isfinite (range))
     return 0;
   nb = (n < nbmax) ? n : nbmax;
   b->x0 = xp[0];
   b->scale = (double) nb / range;
   b->nb = nb;
   if (!isfinite (b->scale))
     {
        memset (b, 0, sizeof (*b));
        return 0;
     }
   if (NULL == (cnt = (double *) malloc ((nb + 1) * sizeof (double))))
     {
        mxg_error ("out of memory");
        return -1;
     }
   for (j = 0, i = 0; j < nb; j++)
     {
        while ((i < n) && (lbucket_key (b, xp[i]) < j))
          i++;
        cnt[j] = (double) i;
     }
   cnt[nb] = (double) n;
   b->off = mxg_darena_push (cnt, nb + 1);
   free (cnt);
   if (b->off == (unsigned int) -1)
     return -1;
   b->ok = 1;
   return 0;
}

int mxg_build_bucket (const float *xp, unsigned int n, int double_keys, unsigned int nbmax, mxg_bucket *b)
{
   unsigned long long k0, range;
   unsigned int i, j, shift, nb;
   float *cnt;
   unsigned int ok = 1;

   memset (b, 0, sizeof (*b));
   if ((n < 2) || (n > (1u << 24)) || !(xp[0] > 0.0f))
     return 0;
   for (i = 1; i < n; i++)
     if (!(xp[i] > xp[i - 1]))
       {
          if (!(xp[i] == xp[i - 1]))
            return 0;
          ok = 2;
       }
   k0 = table_key (xp[0], double_keys);
   range = table_key (xp[n - 1], double_keys) - k0;
   shift = 0;
   while ((range >> shift) + 1 > nbmax)
     shift++;
   nb = (unsigned int) (range >> shift) + 1;
   if (NULL == (cnt = (float *) malloc ((nb + 1) * sizeof (float))))
     {
        mxg_error ("out of memory");
        return -1;
     }
   for (i = 0, j = 0; i < nb; i++)
     {
        unsigned long long boundary = k0 + ((unsigned long long) i << shift);
        while ((j < n) && (table_key (xp[j], double_keys) < boundary))
          j++;
        cnt[i] = (float) j;
     }
   cnt[nb] = (float) n;
   b->off = mxg_farena_push (cnt, nb + 1);
   free (cnt);
   if (b->off == (unsigned int) -1)
     return -1;
   b->base = k0;
   b->nb = nb;
   b->shift = shift;
   b->ok = ok;
   return 0;
}

/* bucket index of float-arena table (off, n): copied first, as pushing to
 * the arena may move it */
int mxg_bucket_of_farena (unsigned int off, unsigned int n, unsigned int nbmax, mxg_bucket *b)
{
   float *t;
   int status;
   memset (b, 0, sizeof (*b));
   if (n < 2) return 0;
   if (NULL == (t = (float *) malloc (n * sizeof (float))))
     {
        mxg_error ("out of memory");
        return -1;
     }
   memcpy (t, mxg_farena_ptr (off), n * sizeof (float));
   status = mxg_build_bucket (t, n, 0, nbmax, b);
   free (t);
   return status;
}

