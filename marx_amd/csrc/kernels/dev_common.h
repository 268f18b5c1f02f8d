// dev_common.h -- per-photon device primitives for gfx950.
//
// Everything here runs in one lane per photon.  Arithmetic that feeds
// positions, directions and threshold tests is written in the operation
// order of the reference's jdmath routines (jdmath/src/vector.c,
// finterpo.c, dinterpo.c, qroot.c, complex.c) and the file is compiled with
// -ffp-contract=off, so a lane rounds exactly like the reference's CPU code
// apart from the last-bit differences of the transcendental functions.
//
// Random numbers: one Philox4x32-10 counter stream per (photon tag, stage):
// key = seed, counter = (tag, stage, block, source id); 4 uniforms per
// block, u = u32 / (2^32 - 1) (the jdmath/src/random.c:151-154 mapping);
// polar Gaussians keep their second variate inside the stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "marx_gpu_cfg.h"

#define MXG_PI 3.14159265358979323846264338327950288
#define MXG_HBAR_C 1.973269631e-4

namespace mxg {

struct vec3 { double x, y, z; };

__device__ __forceinline__ vec3 v3 (double x, double y, double z) { vec3 a; a.x = x; a.y = y; a.z = z; return a; }

// ---- Philox4x32-10 ----------------------------------------------------------

// b0..b3: the unused words of the current block, next first (a draw takes
// b0 and shifts the rest down, so no word is selected by a run-time index:
// an indexed pick puts the whole stream in scratch memory); pos: the words
// of the block already used.
struct Stream
{
   uint32_t k0, k1;
   uint32_t c0, c1, c2, c3;
   uint32_t b0, b1, b2, b3;
   int pos;
   int have_gauss;
   double g2;
};

// Philox4x32-10 (Salmon et al., SC'11).  Each round's two 32x32->64-bit
// products are formed as one 64-bit multiply (v_mad_u64_u32), half the
// multiply issue of separate low and high products.
__device__ __forceinline__ void philox (uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1,
                                        uint32_t &o0, uint32_t &o1, uint32_t &o2, uint32_t &o3)
{
#pragma unroll
   for (int r = 0; r < 10; r++)
     {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t) 0xD2511F53u * c0;
        uint64_t p1 = (uint64_t) 0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t) (p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t) (p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t) p1; c2 = n2; c3 = (uint32_t) p0;
     }
   o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

__device__ __forceinline__ void stream_init (Stream &s, unsigned long long seed, uint32_t tag,
                                             uint32_t stage, uint32_t source_id)
{
   s.k0 = (uint32_t) seed;
   s.k1 = (uint32_t) (seed >> 32);
   s.c0 = tag; s.c1 = stage; s.c2 = 0; s.c3 = source_id;
   s.pos = 4;
   s.have_gauss = 0;
   s.g2 = 0.0;
}

__device__ __forceinline__ double uniform (Stream &s)
{
   if (s.pos == 4)
     {
        philox (s.c0, s.c1, s.c2, s.c3, s.k0, s.k1, s.b0, s.b1, s.b2, s.b3);
        s.c2++;
        s.pos = 0;
     }
   const uint32_t u = s.b0;
   s.b0 = s.b1;
   s.b1 = s.b2;
   s.b2 = s.b3;
   s.pos++;
   return (double) u * (1.0 / ((double) 0xFFFFFFFFu));
}

__device__ __forceinline__ double gaussian (Stream &s)
{
   if (s.have_gauss)
     {
        s.have_gauss = 0;
        return s.g2;
     }
   s.have_gauss = 1;
   double g1, g2, g;
   do
     {
        g1 = 2.0 * uniform (s) - 1.0;
        g2 = 2.0 * uniform (s) - 1.0;
        g = g1 * g1 + g2 * g2;
     }
   while ((g >= 1.0) || (g == 0.0));
   double f = sqrt (-2.0 * log (g) / g);
   s.g2 = g2 * f;
   return g1 * f;
}

__device__ __forceinline__ double expn (Stream &s)
{
   double r;
   do r = uniform (s); while (r == 0.0);
   return -log (r);
}

// ---- searches and interpolation ----------------------------------------------

__device__ __forceinline__ unsigned int search_f (float x, const float *__restrict__ xp, unsigned int n)
{
   unsigned int n0 = 0, n1 = n;
   while (n1 > n0 + 1)
     {
        unsigned int n2 = (n0 + n1) >> 1;
        float v = xp[n2];
        if (v >= x)
          {
             if (v == x) return n2;
             n1 = n2;
          }
        else n0 = n2;
     }
   if (x >= xp[n0]) return n1;
   return n0;
}

// float table, float abscissa, double internal arithmetic, float result
__device__ __forceinline__ float interp_f (float x, const float *__restrict__ xp,
                                           const float *__restrict__ yp, unsigned int n)
{
   if (n == 1) return yp[0];
   unsigned int n1 = search_f (x, xp, n);
   unsigned int n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1])) return yp[n1];
   if (n1 == n) { n1--; n0--; }
   if (n1 == 0) n0 = 1;
   double x0 = xp[n0], x1 = xp[n1];
   if (x1 == x0) return yp[n1];
   float y0 = yp[n0];
   float dy = yp[n1] - y0;
   return (float) ((double) y0 + (double) dy / (x1 - x0) * ((double) x - x0));
}

// interp_f of two tables on one abscissa with one search (the same values)
__device__ __forceinline__ void interp_f2 (float x, const float *__restrict__ xp, const float *__restrict__ ya,
                                           const float *__restrict__ yb, unsigned int n, float &a, float &b)
{
   if (n == 1) { a = ya[0]; b = yb[0]; return; }
   unsigned int n1 = search_f (x, xp, n);
   unsigned int n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1])) { a = ya[n1]; b = yb[n1]; return; }
   if (n1 == n) { n1--; n0--; }
   if (n1 == 0) n0 = 1;
   double x0 = xp[n0], x1 = xp[n1];
   if (x1 == x0) { a = ya[n1]; b = yb[n1]; return; }
   float y0 = ya[n0];
   float dy = ya[n1] - y0;
   a = (float) ((double) y0 + (double) dy / (x1 - x0) * ((double) x - x0));
   y0 = yb[n0];
   dy = yb[n1] - y0;
   b = (float) ((double) y0 + (double) dy / (x1 - x0) * ((double) x - x0));
}

__device__ __forceinline__ unsigned int search_d (double x, const double *__restrict__ xp, unsigned int n)
{
   unsigned int n0 = 0, n1 = n;
   while (n1 > n0 + 1)
     {
        unsigned int n2 = (n0 + n1) >> 1;
        double v = xp[n2];
        if (v >= x)
          {
             if (v == x) return n2;
             n1 = n2;
          }
        else n0 = n2;
     }
   if (x >= xp[n0]) return n1;
   return n0;
}

// search over a table whose elements widen exactly to double (a float table
// holding float32 file values, or a double table): same result as search_d
template <class T>
__device__ __forceinline__ unsigned int search_w (double x, const T *__restrict__ xp, unsigned int n)
{
   unsigned int n0 = 0, n1 = n;
   while (n1 > n0 + 1)
     {
        unsigned int n2 = (n0 + n1) >> 1;
        double v = (double) xp[n2];
        if (v >= x)
          {
             if (v == x) return n2;
             n1 = n2;
          }
        else n0 = n2;
     }
   if (x >= (double) xp[n0]) return n1;
   return n0;
}

// search_f / search_w of a table with its bucket index (mxg_bucket; cnt: its
// counts, possibly in LDS): the same index, from one bucket's entries.  For
// a non-decreasing table the bisection returns the first index whose entry
// is >= x, except for x equal to the first entry, which it never probes (1
// there), and for x equal to a repeated entry, where its pick among the
// equal entries depends on its path (the whole bisection runs there).
__device__ __forceinline__ unsigned int search_fb (float x, const float *__restrict__ xp, unsigned int n,
                                                   const mxg_bucket &b, const float *__restrict__ cnt)
{
   if (!b.ok) return search_f (x, xp, n);
   const float x0 = xp[0];
   if (!(x > x0))
     {
        if (x == x0) return 1;
        if (x < x0) return 0;
        return search_f (x, xp, n);          // NaN
     }
   const uint64_t k = ((uint64_t) __float_as_uint (x) - b.base) >> b.shift;
   const uint32_t j = (k < b.nb) ? (uint32_t) k : b.nb - 1;
   unsigned int lo = (unsigned int) cnt[j], hi = (unsigned int) cnt[j + 1];
   while (lo < hi)
     {
        unsigned int mid = (lo + hi) >> 1;
        if (xp[mid] < x) lo = mid + 1;
        else hi = mid;
     }
   if ((b.ok == 2) && (lo + 1 < n) && (xp[lo] == x) && (xp[lo + 1] == x))
     return search_f (x, xp, n);
   return lo;
}

__device__ __forceinline__ unsigned int search_wb (double x, const float *__restrict__ xp, unsigned int n,
                                                   const mxg_bucket &b, const float *__restrict__ cnt)
{
   if (!b.ok) return search_w (x, xp, n);
   const double x0 = (double) xp[0];
   if (!(x > x0))
     {
        if (x == x0) return 1;
        if (x < x0) return 0;
        return search_w (x, xp, n);
     }
   const uint64_t k = ((uint64_t) __double_as_longlong (x) - b.base) >> b.shift;
   const uint32_t j = (k < b.nb) ? (uint32_t) k : b.nb - 1;
   unsigned int lo = (unsigned int) cnt[j], hi = (unsigned int) cnt[j + 1];
   while (lo < hi)
     {
        unsigned int mid = (lo + hi) >> 1;
        if ((double) xp[mid] < x) lo = mid + 1;
        else hi = mid;
     }
   if ((b.ok == 2) && (lo + 1 < n) && ((double) xp[lo] == x) && ((double) xp[lo + 1] == x))
     return search_w (x, xp, n);
   return lo;
}

// counts of a bucket index (floats), or none
inline __host__ __device__ uint32_t bucket_floats (const mxg_bucket &b) { return b.ok ? b.nb + 1 : 0u; }

// interp_f / interp_f2 with a bucket-indexed search (the same values)
__device__ __forceinline__ float interp_fb (float x, const float *__restrict__ xp, const float *__restrict__ yp,
                                            unsigned int n, const mxg_bucket &bk, const float *__restrict__ cnt)
{
   if (n == 1) return yp[0];
   unsigned int n1 = search_fb (x, xp, n, bk, cnt);
   unsigned int n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1])) return yp[n1];
   if (n1 == n) { n1--; n0--; }
   if (n1 == 0) n0 = 1;
   double x0 = xp[n0], x1 = xp[n1];
   if (x1 == x0) return yp[n1];
   float y0 = yp[n0];
   float dy = yp[n1] - y0;
   return (float) ((double) y0 + (double) dy / (x1 - x0) * ((double) x - x0));
}

__device__ __forceinline__ void interp_f2b (float x, const float *__restrict__ xp, const float *__restrict__ ya,
                                            const float *__restrict__ yb, unsigned int n, const mxg_bucket &bk,
                                            const float *__restrict__ cnt, float &a, float &b)
{
   if (n == 1) { a = ya[0]; b = yb[0]; return; }
   unsigned int n1 = search_fb (x, xp, n, bk, cnt);
   unsigned int n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1])) { a = ya[n1]; b = yb[n1]; return; }
   if (n1 == n) { n1--; n0--; }
   if (n1 == 0) n0 = 1;
   double x0 = xp[n0], x1 = xp[n1];
   if (x1 == x0) { a = ya[n1]; b = yb[n1]; return; }
   float y0 = ya[n0];
   float dy = ya[n1] - y0;
   a = (float) ((double) y0 + (double) dy / (x1 - x0) * ((double) x - x0));
   y0 = yb[n0];
   dy = yb[n1] - y0;
   b = (float) ((double) y0 + (double) dy / (x1 - x0) * ((double) x - x0));
}

// search_d over a strictly increasing table with the linear bucket index b
// (mxg_lbucket; counts cnt[0..nb]): the same index.  For x > xp[0] both
// are the first entry >= x; the key is monotone in x, so that entry lies
// between the counts of x's bucket and of the next.
__device__ __forceinline__ unsigned int search_dl (double x, const double *__restrict__ xp, unsigned int n,
                                                   const mxg_lbucket &b, const double *__restrict__ cnt)
{
   if (!(x > xp[0])) return (x == xp[0]) ? 1u : 0u;
   const double t = (x - b.x0) * b.scale;
   const unsigned int j = (t < (double) b.nb) ? (unsigned int) t : b.nb - 1;
   unsigned int lo = (unsigned int) cnt[j], hi = (unsigned int) cnt[j + 1];
   while (lo < hi)
     {
        const unsigned int mid = (lo + hi) >> 1;
        if (xp[mid] < x) lo = mid + 1;
        else hi = mid;
     }
   return lo;
}

// interp_d's interpolation once the search gave n1
__device__ __forceinline__ double interp_d_at (double x, const double *__restrict__ xp,
                                               const double *__restrict__ yp, unsigned int n, unsigned int n1)
{
   unsigned int n0 = n1 - 1;
   if ((n1 < n) && (x == xp[n1])) return yp[n1];
   if (n1 == n) { n1--; n0--; }
   if (n1 == 0) n0 = 1;
   double x0 = xp[n0], x1 = xp[n1];
   if (x1 == x0) return yp[n1];
   return yp[n0] + (yp[n1] - yp[n0]) / (x1 - x0) * (x - x0);
}

__device__ __forceinline__ double interp_d (double x, const double *__restrict__ xp,
                                            const double *__restrict__ yp, unsigned int n)
{
   if (n == 1) return yp[0];
   return interp_d_at (x, xp, yp, n, search_d (x, xp, n));
}

// numerically stable roots; returns 1 real, 0 complex, -1 if a == 0
__device__ __forceinline__ int quadratic_root (double a, double b, double c, double &rplus, double &rminus)
{
   if (a == 0.0) return -1;
   double bsqr = b * b;
   double ac4 = a * c * 4;
   double neg_b_over_2a = -b / (2.0 * a);
   if (bsqr > ac4)
     {
        double factor = 1.0 + sqrt (1.0 - ac4 / bsqr);
        rplus = -2.0 * c / (b * factor);
        rminus = neg_b_over_2a * factor;
        return 1;
     }
   if (bsqr == ac4)
     {
        rplus = rminus = neg_b_over_2a;
        return 1;
     }
   rplus = neg_b_over_2a;
   rminus = sqrt (c / a) * sqrt (1.0 - bsqr / ac4);
   return 0;
}

// ---- vectors -------------------------------------------------------------------

__device__ __forceinline__ double length (vec3 a)
{
   double x = fabs (a.x), y = fabs (a.y), z = fabs (a.z), t;
   if (z < x) { t = z; z = x; x = t; }
   if (z < y) { t = z; z = y; y = t; }
   if (z == 0.0) return 0.0;
   x = x / z;
   y = y / z;
   return z * sqrt (1.0 + x * x + y * y);
}

// a / |a|: one square root and one division (the reference divides each
// component by its overflow-safe length, vector.c; for the unit-scale
// directions and normals normalised here the results agree to an ulp)
__device__ __forceinline__ void normalize (vec3 &a)
{
   double l2 = a.x * a.x + a.y * a.y + a.z * a.z;
   if (l2 != 0.0)
     {
        double inv = 1.0 / sqrt (l2);
        a.x = a.x * inv;
        a.y = a.y * inv;
        a.z = a.z * inv;
     }
}

__device__ __forceinline__ vec3 unit (vec3 a) { normalize (a); return a; }

__device__ __forceinline__ double dot (vec3 a, vec3 b)
{
   return a.x * b.x + a.y * b.y + a.z * b.z;
}

__device__ __forceinline__ vec3 cross (vec3 a, vec3 b)
{
   vec3 c;
   c.z = a.x * b.y - a.y * b.x;
   c.x = a.y * b.z - a.z * b.y;
   c.y = a.z * b.x - a.x * b.z;
   return c;
}

__device__ __forceinline__ vec3 ax1_bx2 (double a, vec3 x1, double b, vec3 x2)
{
   return v3 (a * x1.x + b * x2.x, a * x1.y + b * x2.y, a * x1.z + b * x2.z);
}

__device__ __forceinline__ vec3 ax1_bx2_cx3 (double a, vec3 x1, double b, vec3 x2, double c, vec3 x3)
{
   return v3 (a * x1.x + b * x2.x + c * x3.x,
              a * x1.y + b * x2.y + c * x3.y,
              a * x1.z + b * x2.z + c * x3.z);
}

__device__ __forceinline__ vec3 add (vec3 a, vec3 b) { return v3 (a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ vec3 sub (vec3 a, vec3 b) { return v3 (a.x - b.x, a.y - b.y, a.z - b.z); }

__device__ __forceinline__ vec3 rotate_vector1 (vec3 p, vec3 n, double c, double s)
{
   double pn = dot (p, n);
   return ax1_bx2_cx3 (c, p, pn * (1.0 - c), n, s, cross (n, p));
}

__device__ __forceinline__ vec3 rotate_unit_vector1 (vec3 p, vec3 n, double c, double s)
{
   vec3 u = rotate_vector1 (p, n, c, s);
   normalize (u);
   return u;
}

// sin and cos of t; below 2^-6 rad (the arcsecond-to-arcminute angles of
// the mirror blur, WFold scatter and dither offsets) from their Taylor
// polynomials through t^7 and t^8 (truncation below 1e-20 relative: the
// result differs from sincos's by rounding only), else sincos.
__device__ __forceinline__ void sincos_small (double t, double *sn, double *cs)
{
   if (fabs (t) < 0.015625)
     {
        const double t2 = t * t;
        *sn = t * (1.0 - t2 * (1.0 / 6.0 - t2 * (1.0 / 120.0 - t2 * (1.0 / 5040.0))));
        *cs = 1.0 - t2 * (0.5 - t2 * (1.0 / 24.0 - t2 * (1.0 / 720.0 - t2 * (1.0 / 40320.0))));
     }
   else
     sincos (t, sn, cs);
}

__device__ __forceinline__ vec3 rotate_unit_vector (vec3 p, vec3 n, double theta)
{
   double s, c;
   sincos (theta, &s, &c);
   return rotate_unit_vector1 (p, n, c, s);
}

__device__ __forceinline__ vec3 m3_vmul (const double *__restrict__ a, vec3 v)
{
   return v3 (a[0] * v.x + a[1] * v.y + a[2] * v.z,
              a[3] * v.x + a[4] * v.y + a[5] * v.z,
              a[6] * v.x + a[7] * v.y + a[8] * v.z);
}

__device__ __forceinline__ vec3 ld3 (const double *a) { return v3 (a[0], a[1], a[2]); }

// ---- Fresnel reflectivity (Jackson 1975, eq 7.39 / 7.41) ----------------------

struct cplx { double r, i; };

__device__ __forceinline__ cplx c_mul (cplx a, cplx b)
{
   cplx z; z.r = a.r * b.r - a.i * b.i; z.i = a.r * b.i + a.i * b.r; return z;
}

__device__ __forceinline__ cplx c_a_bz (double a, double b, cplx z1)
{
   cplx z; z.r = a + b * z1.r; z.i = b * z1.i; return z;
}

__device__ __forceinline__ cplx c_az1_bz2 (double a, cplx z1, double b, cplx z2)
{
   cplx z; z.r = a * z1.r + b * z2.r; z.i = a * z1.i + b * z2.i; return z;
}

// The principal square root (non-negative real part) of a, by the half-angle
// formulas; |a| = sqrt (re^2 + im^2) (reflectivity's arguments are of order
// one, far from overflow).
__device__ __forceinline__ cplx c_sqrt (cplx a)
{
   double r = sqrt (a.r * a.r + a.i * a.i);
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
        if (a.r < 0.0) { a.r = -a.r; a.i = -a.i; }
     }
   return a;
}

// Reflectivity of a surface of complex index n = 1 - delta + i beta for a
// ray at angle theta to the normal: the mean of |E_perp|^2 and |E_par|^2,
// E_perp = (cos - root) / (cos + root), E_par = (n^2 cos - root) /
// (n^2 cos + root), root = sqrt (n^2 - sin^2) (reflect.c:39-77).  The
// squared modulus of each ratio is taken as the ratio of the squared moduli,
// which needs no complex division (the value agrees with the divided form to
// rounding).
__device__ __forceinline__ double reflectivity (double cos_theta, double beta, double delta)
{
   cplx n; n.r = 1.0 - delta; n.i = beta;
   double sin_theta = sqrt (1.0 - cos_theta * cos_theta);
   cplx nsqr = c_mul (n, n);
   cplx root = c_sqrt (c_a_bz (-sin_theta * sin_theta, 1.0, nsqr));
   cplx a1 = c_a_bz (cos_theta, -1.0, root), b1 = c_a_bz (cos_theta, 1.0, root);
   cplx a2 = c_az1_bz2 (cos_theta, nsqr, -1.0, root), b2 = c_az1_bz2 (cos_theta, nsqr, 1.0, root);
   double perp = (a1.r * a1.r + a1.i * a1.i) / (b1.r * b1.r + b1.i * b1.i);
   double par = (a2.r * a2.r + a2.i * a2.i) / (b2.r * b2.r + b2.i * b2.i);
   return 0.5 * (par + perp);
}

// ---- WFold scattering-angle tables -------------------------------------------------

__device__ __forceinline__ double fold_theta (const double *__restrict__ fa, const float *__restrict__ farena, double p)
{
   double p_min = fa[1];
   if (p < p_min) return 0.0;
   double p_max = fa[3];
   if (p > p_max) return pow (fa[4] * (1.0 - p), fa[5]);
   unsigned int num = (unsigned int) fa[6];
   const float *t = farena + (unsigned int) fa[7];
   double delta_i = (p - p_min) / fa[2];
   unsigned int i = (unsigned int) delta_i;
   if (i + 1 >= num) return (double) t[num - 1];
   delta_i -= (double) i;
   return (1.0 - delta_i) * (double) t[i] + delta_i * (double) t[i + 1];
}

// e_alphas: the table's e_alpha values (float32 file values, exact in
// float), possibly staged in LDS; fold arrays stay in the double arena
// bk, cnt: the bucket index of a float e_alphas table (nullptr: search it
// whole)
template <class T>
__device__ __forceinline__ double wfold_interp (const mxg_wfold &w, const float *__restrict__ fa,
                                                const double *__restrict__ da,
                                                const T *__restrict__ e_alphas,
                                                double energy, double sin_alpha, double r,
                                                const mxg_bucket *bk = nullptr,
                                                const float *__restrict__ cnt = nullptr)
{
   if (w.n_arrays == 0) return 0.0;
   const double *arrays = da + w.arrays_off;
   if (w.n_arrays == 1) return fold_theta (arrays, fa, r);
   double e_alpha = energy * sin_alpha;
   unsigned int i;
   if constexpr (sizeof (T) == sizeof (float))
     i = (bk != nullptr) ? search_wb (e_alpha, e_alphas, w.n_arrays, *bk, cnt)
                         : search_w (e_alpha, e_alphas, w.n_arrays);
   else
     i = search_w (e_alpha, e_alphas, w.n_arrays);
   if (i == w.n_arrays) i--;
   if (i == 0) i++;
   double theta_0 = fold_theta (arrays + (i - 1) * MXG_FOLD_STRIDE, fa, r);
   double theta_1 = fold_theta (arrays + i * MXG_FOLD_STRIDE, fa, r);
   double e0 = (double) e_alphas[i - 1], e1 = (double) e_alphas[i];
   return theta_0 + (theta_1 - theta_0) * (e_alpha - e0) / (e1 - e0);
}

// ---- shared by the mirror and source kernels --------------------------------------

// precollimator / cap / post-collimator strut shadows (hrma.c:900-981)
__device__ __forceinline__ bool in_strut_shadow (const mxg_hrma_cfg &h, vec3 x, vec3 p,
                                                 const mxg_strut *s)
{
   const double cos_theta = h.strut_cos, sin_theta = h.strut_sin;
   for (int k = 0; k < 2; k++)
     {
        double half_width = s[k].half_width;
        double t = (s[k].xpos + h.cap_position - x.x) / p.x;
        double y = x.y + p.y * t;
        double z = x.z + p.z * t;
        for (int i = 0; i < 3; i++)
          {
             if (i != 0)
               {
                  double tmp = cos_theta * y - sin_theta * z;
                  z = sin_theta * y + cos_theta * z;
                  y = tmp;
               }
             if (((-half_width < y) && (y < half_width)) || ((-half_width < z) && (z < half_width)))
               return true;
          }
     }
   return false;
}

// energy of the FLAT or FILE spectrum (spectrum.c:138-147, 184-189); one
// uniform either way
__device__ __forceinline__ double draw_energy (const mxg_source_cfg &src, const double *__restrict__ da,
                                               Stream &st)
{
   if (src.spectrum == MXG_SPECTRUM_FLAT)
     {
        double emin = src.emin, de = src.emax - emin;
        return emin + de * uniform (st);
     }
   double r = uniform (st);
   const double *cum = da + src.file_cum_off;
   if (src.file_bucket.ok)
     return interp_d_at (r, cum, da + src.file_e_off, src.file_n,
                         search_dl (r, cum, src.file_n, src.file_bucket, da + src.file_bucket.off));
   return interp_d (r, cum, da + src.file_e_off, src.file_n);
}

} // namespace mxg
