// diff.hip -- the differentiable mirror: a weighted-photon HRMA trace with
// forward-mode derivatives, its focal-plane image, and the adjoint that turns
// d(loss)/d(image) into d(loss)/d(misalignment parameters).
//
// Per photon, the same steps as hrma.hip (marx/libsrc/hrma.c:984-1050,
// 411-545, 900-981; reflect.c:39-77; wfold.c:334-369) with two changes:
//  - survival weights replace the vignetting and reflectivity draws: the
//    weight is the product of the vignetting factor and, per surface, the
//    Fresnel reflectivity times the scale factor; a photon that misses a
//    surface, is shadowed by a strut or scatters past pi/4 gets weight 0;
//  - every quantity that depends on the photon's shell-pair parameters
//    (tilt az, tilt el, decenter y, decenter z, despace x) carries their
//    five derivatives (a dual number).  Draws are consumed in the same
//    order as the oracle's weighted trace; each photon's draws do not depend
//    on the parameters.
// At the focal plane (x = 0) the weight is spread bilinearly over the four
// pixels around the photon.  Weights, positions and their derivatives stay
// in HBM for the adjoint, which contracts them with d(loss)/d(image).
#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

constexpr int kNP = MXG_DIFF_NPAR;

// ---- dual numbers: a value and its derivatives by N parameters ---------------------
//
// N = kNP carries all five parameters of the photon's shell pair.  N = 2
// carries the two tilts only: for a source at infinity a translation of the
// shell pair moves the photon's aperture position with it (the annulus is
// sampled about the displaced shell), so in the frame of each surface the
// path does not depend on the translations, whose derivatives are then
// fixed at the focal plane (kernel below).  N = 0 is the value-only pass.

template <int N>
struct Dual
{
   double v;
   double d[N > 0 ? N : 1];
};

template <int N> __device__ __forceinline__ Dual<N> dc (double v)
{
   Dual<N> r;
   r.v = v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = 0.0;
   return r;
}

template <int N> __device__ __forceinline__ Dual<N> operator+ (Dual<N> a, Dual<N> b)
{
   Dual<N> r; r.v = a.v + b.v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = a.d[j] + b.d[j];
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator- (Dual<N> a, Dual<N> b)
{
   Dual<N> r; r.v = a.v - b.v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = a.d[j] - b.d[j];
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator- (Dual<N> a)
{
   Dual<N> r; r.v = -a.v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = -a.d[j];
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator* (Dual<N> a, Dual<N> b)
{
   Dual<N> r; r.v = a.v * b.v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator* (double a, Dual<N> b)
{
   Dual<N> r; r.v = a * b.v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = a * b.d[j];
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator* (Dual<N> b, double a) { return a * b; }
template <int N> __device__ __forceinline__ Dual<N> operator/ (Dual<N> a, Dual<N> b)
{
   Dual<N> r; r.v = a.v / b.v;
   double ib = 1.0 / b.v;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = (a.d[j] - r.v * b.d[j]) * ib;
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/ (Dual<N> a, double b)
{
   Dual<N> r; r.v = a.v / b;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = a.d[j] / b;
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator+ (Dual<N> a, double b) { a.v += b; return a; }
template <int N> __device__ __forceinline__ Dual<N> operator+ (double b, Dual<N> a) { a.v = b + a.v; return a; }
template <int N> __device__ __forceinline__ Dual<N> operator- (Dual<N> a, double b) { a.v -= b; return a; }
template <int N> __device__ __forceinline__ Dual<N> operator- (double a, Dual<N> b) { return dc<N> (a) - b; }

template <int N> __device__ __forceinline__ Dual<N> dsqrt (Dual<N> a)
{
   Dual<N> r; r.v = sqrt (a.v);
   double f = (r.v > 0.0) ? 0.5 / r.v : 0.0;
#pragma unroll
   for (int j = 0; j < N; j++) r.d[j] = a.d[j] * f;
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> dfabs (Dual<N> a) { return (a.v < 0.0) ? -a : a; }
template <int N> __device__ __forceinline__ void dsincos (Dual<N> a, Dual<N> &s, Dual<N> &c)
{
   double sv, cv;
   sincos (a.v, &sv, &cv);
   s.v = sv; c.v = cv;
#pragma unroll
   for (int j = 0; j < N; j++) { s.d[j] = cv * a.d[j]; c.d[j] = -sv * a.d[j]; }
}

template <int N> struct DVec { Dual<N> x, y, z; };

template <int N> __device__ __forceinline__ DVec<N> dv (vec3 a)
{
   DVec<N> r; r.x = dc<N> (a.x); r.y = dc<N> (a.y); r.z = dc<N> (a.z); return r;
}
template <int N> __device__ __forceinline__ vec3 val (const DVec<N> &a) { return v3 (a.x.v, a.y.v, a.z.v); }
template <int N> __device__ __forceinline__ DVec<N> dadd (const DVec<N> &a, const DVec<N> &b)
{
   DVec<N> r; r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; return r;
}
template <int N> __device__ __forceinline__ DVec<N> dsub (const DVec<N> &a, const DVec<N> &b)
{
   DVec<N> r; r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z; return r;
}
template <int N> __device__ __forceinline__ Dual<N> ddot (const DVec<N> &a, const DVec<N> &b)
{
   return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <int N> __device__ __forceinline__ DVec<N> dcross (const DVec<N> &a, const DVec<N> &b)
{
   DVec<N> c;
   c.z = a.x * b.y - a.y * b.x;
   c.x = a.y * b.z - a.z * b.y;
   c.y = a.z * b.x - a.x * b.z;
   return c;
}
template <int N>
__device__ __forceinline__ DVec<N> dax1_bx2 (Dual<N> a, const DVec<N> &x1, Dual<N> b, const DVec<N> &x2)
{
   DVec<N> r;
   r.x = a * x1.x + b * x2.x;
   r.y = a * x1.y + b * x2.y;
   r.z = a * x1.z + b * x2.z;
   return r;
}
template <int N>
__device__ __forceinline__ DVec<N> dax1_bx2_cx3 (Dual<N> a, const DVec<N> &x1, Dual<N> b, const DVec<N> &x2,
                                                 Dual<N> c, const DVec<N> &x3)
{
   DVec<N> r;
   r.x = a * x1.x + b * x2.x + c * x3.x;
   r.y = a * x1.y + b * x2.y + c * x3.y;
   r.z = a * x1.z + b * x2.z + c * x3.z;
   return r;
}

// the vector length of dev_common.h's length(), on values for the ordering
template <int N> __device__ __forceinline__ Dual<N> dlength (const DVec<N> &a)
{
   Dual<N> x = dfabs (a.x), y = dfabs (a.y), z = dfabs (a.z), t;
   if (z.v < x.v) { t = z; z = x; x = t; }
   if (z.v < y.v) { t = z; z = y; y = t; }
   if (z.v == 0.0) return dc<N> (0.0);
   x = x / z;
   y = y / z;
   return z * dsqrt (1.0 + x * x + y * y);
}

template <int N> __device__ __forceinline__ void dnormalize (DVec<N> &a)
{
   Dual<N> len = dlength (a);
   if (len.v != 0.0)
     {
        a.x = a.x / len;
        a.y = a.y / len;
        a.z = a.z / len;
     }
}

template <int N>
__device__ __forceinline__ DVec<N> drotate_unit_vector1 (const DVec<N> &p, const DVec<N> &n, Dual<N> c, Dual<N> s)
{
   Dual<N> pn = ddot (p, n);
   DVec<N> u = dax1_bx2_cx3 (c, p, pn * (1.0 - c), n, s, dcross (n, p));
   dnormalize (u);
   return u;
}

template <int N>
__device__ __forceinline__ DVec<N> drotate_unit_vector (const DVec<N> &p, const DVec<N> &n, Dual<N> theta)
{
   Dual<N> s, c;
   dsincos (theta, s, c);
   return drotate_unit_vector1 (p, n, c, s);
}

// M x for a tilt matrix M(az, el): value m, derivatives dm[0] (az, tangent
// 0), dm[1] (el, tangent 1)
template <int N>
__device__ __forceinline__ Dual<N> dm3_row (const double *m, const double (*dm)[9], int row, const DVec<N> &x)
{
   const double *a = m + 3 * row;
   Dual<N> o = a[0] * x.x + a[1] * x.y + a[2] * x.z;
   if constexpr (N >= 2)
     {
        o.d[0] += dm[0][3 * row] * x.x.v + dm[0][3 * row + 1] * x.y.v + dm[0][3 * row + 2] * x.z.v;
        o.d[1] += dm[1][3 * row] * x.x.v + dm[1][3 * row + 1] * x.y.v + dm[1][3 * row + 2] * x.z.v;
     }
   return o;
}

template <int N>
__device__ __forceinline__ DVec<N> dm3_vmul (const double *m, const double (*dm)[9], const DVec<N> &x)
{
   DVec<N> r;
   r.x = dm3_row (m, dm, 0, x);
   r.y = dm3_row (m, dm, 1, x);
   r.z = dm3_row (m, dm, 2, x);
   return r;
}

// to_osac of surface k as a dual: d/d(decenter y, z, despace x) = -1 (only
// carried when the translations are)
template <int N> __device__ __forceinline__ DVec<N> dto_osac (const mxg_diff_shell &D, int k)
{
   DVec<N> t = dv<N> (ld3 (D.to_osac[k]));
   if constexpr (N == kNP)
     {
        t.y.d[2] = -1.0;
        t.z.d[3] = -1.0;
        t.x.d[4] = -1.0;
     }
   return t;
}

// ---- the reflectivity and its derivative by cos(theta) ---------------------------

struct D1 { double v, d; };
__device__ __forceinline__ D1 d1 (double v, double d = 0.0) { D1 r; r.v = v; r.d = d; return r; }
__device__ __forceinline__ D1 operator+ (D1 a, D1 b) { return d1 (a.v + b.v, a.d + b.d); }
__device__ __forceinline__ D1 operator- (D1 a, D1 b) { return d1 (a.v - b.v, a.d - b.d); }
__device__ __forceinline__ D1 operator- (D1 a) { return d1 (-a.v, -a.d); }
__device__ __forceinline__ D1 operator* (D1 a, D1 b) { return d1 (a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ D1 operator/ (D1 a, D1 b)
{
   double q = a.v / b.v;
   return d1 (q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ D1 d1sqrt (D1 a)
{
   double s = sqrt (a.v);
   return d1 (s, (s > 0.0) ? a.d * 0.5 / s : 0.0);
}
__device__ __forceinline__ D1 d1fabs (D1 a) { return (a.v < 0.0) ? -a : a; }

struct C1 { D1 r, i; };

__device__ __forceinline__ C1 c1mul (C1 a, C1 b)
{
   C1 z; z.r = a.r * b.r - a.i * b.i; z.i = a.r * b.i + a.i * b.r; return z;
}
__device__ __forceinline__ C1 c1div (C1 z1, C1 z2)
{
   C1 z;
   D1 ratio, denom;
   if (fabs (z2.r.v) > fabs (z2.i.v))
     {
        ratio = z2.i / z2.r;
        denom = z2.r + z2.i * ratio;
        z.r = (z1.r + ratio * z1.i) / denom;
        z.i = (z1.i - z1.r * ratio) / denom;
     }
   else
     {
        ratio = z2.r / z2.i;
        denom = z2.r * ratio + z2.i;
        z.r = (z1.r * ratio + z1.i) / denom;
        z.i = (z1.i * ratio - z1.r) / denom;
     }
   return z;
}
__device__ __forceinline__ C1 c1_a_bz (D1 a, double b, C1 z1)
{
   C1 z; z.r = a + d1 (b) * z1.r; z.i = d1 (b) * z1.i; return z;
}
__device__ __forceinline__ C1 c1_az1_bz2 (D1 a, C1 z1, double b, C1 z2)
{
   C1 z; z.r = a * z1.r + d1 (b) * z2.r; z.i = a * z1.i + d1 (b) * z2.i; return z;
}
__device__ __forceinline__ D1 c1abs (C1 z)
{
   D1 fr = d1fabs (z.r), fi = d1fabs (z.i), ratio;
   if (fr.v > fi.v)
     {
        ratio = z.i / z.r;
        return fr * d1sqrt (d1 (1.0) + ratio * ratio);
     }
   if (fi.v == 0.0) return d1 (0.0);
   ratio = z.r / z.i;
   return fi * d1sqrt (d1 (1.0) + ratio * ratio);
}
__device__ __forceinline__ C1 c1sqrt (C1 a)
{
   D1 r = c1abs (a);
   if (r.v == 0.0) return a;
   if (a.r.v >= 0.0)
     {
        a.r = d1sqrt (d1 (0.5) * (r + a.r));
        a.i = d1 (0.5) * a.i / a.r;
     }
   else
     {
        r = d1sqrt (d1 (0.5) * (r - a.r));
        a.r = d1 (0.5) * a.i / r;
        a.i = r;
        if (a.r.v < 0.0) { a.r = -a.r; a.i = -a.i; }
     }
   return a;
}

// reflectivity (value and d/d cos_theta) of dev_common.h's reflectivity()
__device__ __forceinline__ D1 reflectivity_d1 (double cos_theta, double beta, double delta)
{
   D1 c = d1 (cos_theta, 1.0);
   C1 n; n.r = d1 (1.0 - delta); n.i = d1 (beta);
   D1 sin_theta = d1sqrt (d1 (1.0) - c * c);
   C1 nsqr = c1mul (n, n);
   C1 root = c1sqrt (c1_a_bz (-(sin_theta * sin_theta), 1.0, nsqr));
   C1 e_perp = c1div (c1_a_bz (c, -1.0, root), c1_a_bz (c, 1.0, root));
   C1 e_par = c1div (c1_az1_bz2 (c, nsqr, -1.0, root), c1_az1_bz2 (c, nsqr, 1.0, root));
   return d1 (0.5) * (e_par.r * e_par.r + e_par.i * e_par.i + e_perp.r * e_perp.r + e_perp.i * e_perp.i);
}

// WFold angle at (energy, sin_alpha, r) and its derivative by sin_alpha
// (linear in e_alpha = energy sin_alpha between the bracketing tables)
__device__ __forceinline__ double wfold_d1 (const mxg_wfold &w, const float *__restrict__ fa,
                                            const double *__restrict__ da, double energy,
                                            double sin_alpha, double r, double &dval)
{
   dval = 0.0;
   if (w.n_arrays == 0) return 0.0;
   const double *arrays = da + w.arrays_off;
   if (w.n_arrays == 1) return fold_theta (arrays, fa, r);
   const double *e_alphas = da + w.e_alphas_off;
   double e_alpha = energy * sin_alpha;
   unsigned int i = search_w (e_alpha, e_alphas, w.n_arrays);
   if (i == w.n_arrays) i--;
   if (i == 0) i++;
   double theta_0 = fold_theta (arrays + (i - 1) * MXG_FOLD_STRIDE, fa, r);
   double theta_1 = fold_theta (arrays + i * MXG_FOLD_STRIDE, fa, r);
   double e0 = e_alphas[i - 1], e1 = e_alphas[i];
   dval = (theta_1 - theta_0) / (e1 - e0) * energy;
   return theta_0 + (theta_1 - theta_0) * (e_alpha - e0) / (e1 - e0);
}

// ---- one photon -----------------------------------------------------------------------

template <int N>
__device__ __forceinline__ bool dconic_intersection (double a, double b, double c, DVec<N> &x0, const DVec<N> &p,
                                                     double xmin, double xmax)
{
   Dual<N> t_plus = -x0.x / p.x;
   Dual<N> x_y = x0.y + t_plus * p.y;
   Dual<N> x_z = x0.z + t_plus * p.z;
   Dual<N> alpha = a * (p.x * p.x) - 1.0;
   Dual<N> beta = b * p.x - 2.0 * (p.y * x_y + p.z * x_z);
   Dual<N> gamma = c - x_z * x_z - x_y * x_y;
   Dual<N> t_minus;
   if (alpha.v == 0.0)
     {
        if (beta.v == 0.0) return false;
        t_plus = t_minus = -gamma / beta;
     }
   else
     {
        // the roots of dev_common.h's quadratic_root(), on duals
        Dual<N> bsqr = beta * beta;
        Dual<N> ac4 = alpha * gamma * 4.0;
        Dual<N> neg_b_over_2a = -beta / (2.0 * alpha);
        if (bsqr.v > ac4.v)
          {
             Dual<N> factor = 1.0 + dsqrt (1.0 - ac4 / bsqr);
             t_plus = -2.0 * gamma / (beta * factor);
             t_minus = neg_b_over_2a * factor;
          }
        else if (bsqr.v == ac4.v)
          t_plus = t_minus = neg_b_over_2a;
        else
          return false;
     }
   Dual<N> x_plus = p.x * t_plus;
   Dual<N> x_minus = p.x * t_minus;
   bool plus_ok = (x_plus.v >= xmin) && (x_plus.v < xmax);
   bool minus_ok = (x_minus.v >= xmin) && (x_minus.v < xmax);
   Dual<N> t;
   if (plus_ok)
     t = (minus_ok && (x_minus.v > x_plus.v)) ? t_minus : t_plus;
   else if (minus_ok)
     t = t_minus;
   else
     return false;
   x0.x = p.x * t;
   x0.y = x_y + p.y * t;
   x0.z = x_z + p.z * t;
   return true;
}

// reflection at x on conic k: normal blur, survival weight, specular
// reflection, WFold scatter.  false: scattered past pi/4.
template <int N>
__device__ __forceinline__ bool dreflect_at (const mxg_hrma_cfg &h, const mxg_shell_cfg &S, int k,
                                             const DVec<N> &x, DVec<N> &p, Dual<N> &w, double energy,
                                             double beta, double delta, double corr,
                                             const float *__restrict__ fa,
                                             const double *__restrict__ da, Stream &st)
{
   DVec<N> n;
   n.x = (S.conic_a[k] - 1) * x.x + 0.5 * S.conic_b[k];
   n.y = -x.y;
   n.z = -x.z;
   dnormalize (n);
   if (h.use_blur)
     {
        Dual<N> len = dsqrt (n.y * n.y + n.z * n.z);
        DVec<N> perp;
        perp.x = dc<N> (0.0);
        perp.y = n.z / len;
        perp.z = -n.y / len;
        double phi = (2.0 * MXG_PI) * uniform (st);
        perp = drotate_unit_vector (perp, n, dc<N> (phi));
        phi = S.blur[k] * (1.0 / 3600.0 * MXG_PI / 180.0);
        phi = phi * gaussian (st);
        n = drotate_unit_vector (n, perp, dc<N> (phi));
     }
   Dual<N> p_dot_n = ddot (p, n);
   if (h.ideal == 0)
     {
        D1 r = reflectivity_d1 (fabs (p_dot_n.v), beta, delta);
        Dual<N> c = dfabs (p_dot_n);
        Dual<N> R = r.d * c;             // the derivative part; its value is replaced next
        R.v = r.v;
        w = w * (R * corr);
     }
   p = dax1_bx2 (dc<N> (1.0), p, -2.0 * p_dot_n, n);
   if (h.use_wfold == 0)
     return true;
   double r = uniform (st);
   double slope;
   double g = wfold_d1 (S.wfold[k], fa, da, energy, -p_dot_n.v, r, slope);
   Dual<N> dg = (-slope) * p_dot_n;      // d/d(params) of the angle through sin_alpha = -p.n
   dg.v = g;
   dg = dg * S.scat_factor[k];
   if (dg.v > MXG_PI / 4) return false;
   if (uniform (st) < 0.5) dg = -dg;
   p = drotate_unit_vector (p, dcross (p, n), dg);
   return true;
}

struct ImageSpec
{
   uint32_t npix;
   double pix, cy, cz;
};

constexpr int kWin = 64;         // central window of the image accumulated in LDS

__device__ __forceinline__ void pixel_coords (const ImageSpec &sp, double y, double z, int &i0, int &j0,
                                              double &fu, double &fv)
{
   double u = (y - sp.cy) / sp.pix + 0.5 * sp.npix - 0.5;
   double v = (z - sp.cz) / sp.pix + 0.5 * sp.npix - 0.5;
   double fu0 = floor (u), fv0 = floor (v);
   // clamp far-away photons to an index range that keeps the arithmetic in int
   fu0 = fmin (fmax (fu0, -2.0), (double) sp.npix + 1.0);
   fv0 = fmin (fmax (fv0, -2.0), (double) sp.npix + 1.0);
   i0 = (int) fu0;
   j0 = (int) fv0;
   fu = u - fu0;
   fv = v - fv0;
}

// A photon up to its paraboloid intersection: its draws (source stream:
// energy; mirror stream: shell, radius, azimuth), the aperture position in
// MARX coordinates, the precollimator struts, and the intersection in the
// paraboloid's frame.  false: the photon is lost before the first
// reflection.  k_diff_enter runs it on values, k_diff_forward again with
// tangents on the photons that pass; the values come out the same.
template <int N>
struct DiffPhoton
{
   double energy;
   int si;
   Stream st;
   DVec<N> x, p, t0, t1;
};

template <int N>
__device__ __forceinline__ bool diff_enter (const mxg_hrma_cfg &h, const mxg_source_cfg &src,
                                            const double *__restrict__ da,
                                            const mxg_diff_cfg *__restrict__ dcfg, uint32_t tag,
                                            unsigned long long seed, DiffPhoton<N> &ph)
{
   Stream ss;
   stream_init (ss, seed, tag, MXG_STAGE_SOURCE, 0u);
   ph.energy = draw_energy (src, da, ss);
   stream_init (ph.st, seed, tag, MXG_STAGE_MIRROR, 0u);
   int si;
   for (;;)
     {
        double r = uniform (ph.st);
        si = -1;
        for (int k = 0; k < MXG_NUM_SHELLS; k++)
          if (r < h.shell[k].area_fraction) { si = k; break; }
        if (si >= 0) break;
     }
   ph.si = si;
   const mxg_shell_cfg &S = h.shell[si];
   const mxg_diff_shell &D = dcfg->shell[si];
   double radius = S.min_radius + (S.max_radius - S.min_radius) * uniform (ph.st);
   double theta;
   unsigned int quad;
   do
     {
        theta = uniform (ph.st);
        quad = (unsigned int) (4.0 * theta);
     }
   while (0 == (S.shutter_bitmap & (1u << quad)));
   theta = (2.0 * MXG_PI) * (theta - 1.0 / 8.0);
   double sn, cs;
   sincos (theta, &sn, &cs);

   ph.t0 = dto_osac<N> (D, 0);
   ph.t1 = dto_osac<N> (D, 1);
   DVec<N> &x = ph.x, &p = ph.p;
   x.x = dc<N> (D.front_position);
   if constexpr (N == kNP) x.x.d[4] = 1.0;
   x.z = dc<N> (radius * cs) - ph.t0.z;
   x.y = dc<N> (radius * sn) - ph.t0.y;
   p = dv<N> (ld3 (src.p));
   if (src.distance > 0.0)
     {
        p = dax1_bx2 (dc<N> (1.0), x, dc<N> (src.distance), p);
        dnormalize (p);
     }
   if (h.use_struts && in_strut_shadow (h, val (x), val (p), h.pre_struts))
     return false;
   x = dadd (x, ph.t0);
   x = dm3_vmul (D.fwd[0], D.dfwd[0], x);
   p = dm3_vmul (D.fwd[0], D.dfwd[0], p);
   return dconic_intersection (S.conic_a[0], S.conic_b[0], S.conic_c[0], x, p, S.conic_xmin[0],
                               S.conic_xmax[0]);
}

// Every photon to its paraboloid intersection, values only.  A lost photon
// gets its final weight, position and shell here (weight 0, position 0; its
// derivative columns are not written: a photon of weight 0 adds nothing to
// the gradient and the adjoint does not read them); the others are listed,
// in order, for k_diff_forward.
__global__ __launch_bounds__ (kBlock)
void k_diff_enter (const mxg_config *__restrict__ cfg, const double *__restrict__ da,
                   const mxg_diff_cfg *__restrict__ dcfg, uint32_t n, uint32_t tag_start,
                   unsigned long long seed, DiffOut out, ListIO io)
{
   __shared__ uint32_t wc[2 * (kBlock / 64)];
   const mxg_hrma_cfg &h = cfg->hrma;
   const mxg_source_cfg &src = cfg->src;
   const BlockRange range = block_range (n, blockIdx.x, gridDim.x);
   SegWriter sw (io, range, wc);
   for (uint32_t base = range.lo; base < range.hi; base += kBlock)
     {
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        if (i < range.hi)
          {
             DiffPhoton<0> ph;
             keep = diff_enter<0> (h, src, da, dcfg, tag_start + i, seed, ph);
             if (!keep)
               {
                  out.w[i] = 0.0; out.y[i] = 0.0; out.z[i] = 0.0;
                  out.shell[i] = (uint8_t) ph.si;
               }
          }
        sw.add (keep, i);
     }
   sw.finish (io);
}

// waves per SIMD the forward kernel's registers are bounded for; timed on
// MI355X over 1e7 photons, 1 to 4 waves: 2 with tangents, 3 value only
template <int N> constexpr int kDiffWaves = (N == 0) ? 3 : 2;

// The photons listed by k_diff_enter (idx[0 .. *n_list)), with N tangents
// each (see Dual): kNP, 2 (source at infinity: the translation derivatives
// follow at the focal plane), or 0 (value only).  From the paraboloid
// intersection on: both reflections, the cap and post struts, the focal
// plane, the state, and the splat into the image (central window in LDS,
// summed per block in win_partial; the rest by global atomics).
template <int N>
__global__ __launch_bounds__ (kBlock, kDiffWaves<N>)
void k_diff_forward (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                     const double *__restrict__ da, const mxg_diff_cfg *__restrict__ dcfg,
                     const uint32_t *__restrict__ idx, const uint32_t *__restrict__ n_list, uint32_t n,
                     uint32_t tag_start, unsigned long long seed, ImageSpec sp,
                     float *__restrict__ image, float *__restrict__ win_partial, DiffOut out)
{
   __shared__ float win[kWin * kWin];
   __shared__ double wsum[kBlock / 64];
   const int W = (sp.npix < (uint32_t) kWin) ? (int) sp.npix : kWin;
   const int wo = ((int) sp.npix - W) / 2;
   for (int k = threadIdx.x; k < W * W; k += blockDim.x) win[k] = 0.0f;
   __syncthreads ();

   const mxg_hrma_cfg &h = cfg->hrma;
   const mxg_source_cfg &src = cfg->src;
   const BlockRange range = block_range (*n_list, blockIdx.x, gridDim.x);
   double my_w = 0.0;

   for (uint32_t q = range.lo + threadIdx.x; q < range.hi; q += blockDim.x)
     {
        const uint32_t i = idx[q];
        DiffPhoton<N> ph;
        bool alive = diff_enter<N> (h, src, da, dcfg, tag_start + i, seed, ph);
        const int si = ph.si;
        const mxg_shell_cfg &S = h.shell[si];
        const mxg_diff_shell &D = dcfg->shell[si];
        DVec<N> &x = ph.x, &p = ph.p;
        Dual<N> w = dc<N> ((h.ideal == 0) ? h.vignetting : 1.0);

        double beta = 0.0, delta = 1.0, corr = 1.0;
        if (alive && (h.opt_n != 0) && (h.ideal == 0))
          {
             float ef = (float) ph.energy;
             beta = interp_f (ef, fa + h.opt_e_off, fa + h.opt_beta_off, h.opt_n);
             delta = interp_f (ef, fa + h.opt_e_off, fa + h.opt_delta_off, h.opt_n);
             if (h.use_scale)
               corr = sqrt ((double) interp_f (ef, fa + S.corr_e_off, fa + S.corr_f_off, S.corr_n));
          }
        for (int k = 0; (k < 2) && alive; k++)
          {
             if (k == 1)
               {
                  x = dadd (x, ph.t1);
                  x = dm3_vmul (D.fwd[1], D.dfwd[1], x);
                  p = dm3_vmul (D.fwd[1], D.dfwd[1], p);
                  if (!dconic_intersection (S.conic_a[1], S.conic_b[1], S.conic_c[1], x, p,
                                            S.conic_xmin[1], S.conic_xmax[1]))
                    {
                       alive = false;
                       break;
                    }
               }
             if (!dreflect_at (h, S, k, x, p, w, ph.energy, beta, delta, corr, fa, da, ph.st))
               {
                  alive = false;
                  break;
               }
             p = dm3_vmul (D.bwd[k], D.dbwd[k], p);
             x = dm3_vmul (D.bwd[k], D.dbwd[k], x);
             x = dsub (x, (k == 0) ? ph.t0 : ph.t1);
             if (h.use_struts && in_strut_shadow (h, val (x), val (p), (k == 0) ? h.cap_struts : h.post_struts))
               alive = false;
          }

        // focal plane
        Dual<N> yf = dc<N> (0.0), zf = dc<N> (0.0);
        double ib = 0.0;
        if (alive)
          {
             Dual<N> t = -x.x / p.x;
             yf = x.y + t * p.y;
             zf = x.z + t * p.z;
             ib = 1.0 / p.x.v;
          }
        else
          w = dc<N> (0.0);

        out.w[i] = w.v; out.y[i] = yf.v; out.z[i] = zf.v;
        if constexpr (N == kNP)
          {
#pragma unroll
             for (int j = 0; j < kNP; j++)
               {
                  out.dw[(size_t) j * n + i] = w.d[j];
                  out.dy[(size_t) j * n + i] = yf.d[j];
                  out.dz[(size_t) j * n + i] = zf.d[j];
               }
          }
        else if constexpr (N == 2)
          {
             // tilts as traced; the translations (decenter y, decenter z,
             // despace x) move the exit point by themselves and leave the
             // direction and weight alone: d(y, z) = (1, 0), (0, 1),
             // -(p_y, p_z) / p_x, dw = 0
             const double one = alive ? 1.0 : 0.0;
             const double dyx = -ib * p.y.v, dzx = -ib * p.z.v;
             const double dws[kNP] = {w.d[0], w.d[1], 0.0, 0.0, 0.0};
             const double dys[kNP] = {yf.d[0], yf.d[1], one, 0.0, alive ? dyx : 0.0};
             const double dzs[kNP] = {zf.d[0], zf.d[1], 0.0, one, alive ? dzx : 0.0};
#pragma unroll
             for (int j = 0; j < kNP; j++)
               {
                  out.dw[(size_t) j * n + i] = dws[j];
                  out.dy[(size_t) j * n + i] = dys[j];
                  out.dz[(size_t) j * n + i] = dzs[j];
               }
          }
        out.shell[i] = (uint8_t) si;

        if (w.v != 0.0)
          {
             my_w += w.v;
             int i0, j0;
             double fu, fv;
             pixel_coords (sp, yf.v, zf.v, i0, j0, fu, fv);
             const double b[4] = {(1 - fu) * (1 - fv), fu * (1 - fv), (1 - fu) * fv, fu * fv};
#pragma unroll
             for (int c = 0; c < 4; c++)
               {
                  const int iy = i0 + (c & 1), iz = j0 + (c >> 1);
                  if ((iy < 0) || (iz < 0) || (iy >= (int) sp.npix) || (iz >= (int) sp.npix)) continue;
                  const float add = (float) (w.v * b[c]);
                  const int wy = iy - wo, wz = iz - wo;
                  if ((wy >= 0) && (wz >= 0) && (wy < W) && (wz < W))
                    atomicAdd (&win[wz * W + wy], add);
                  else
                    atomicAdd (&image[(size_t) iz * sp.npix + iy], add);
               }
          }
     }

   // per-block summed weight and the window
   for (int off = 32; off > 0; off >>= 1) my_w += __shfl_xor (my_w, off);
   if (__lane_id () == 0) wsum[threadIdx.x >> 6] = my_w;
   __syncthreads ();
   if (threadIdx.x == 0)
     {
        double s = 0.0;
        for (int k = 0; k < kBlock / 64; k++) s += wsum[k];
        out.block_w[blockIdx.x] = s;
     }
   for (int k = threadIdx.x; k < W * W; k += blockDim.x)
     win_partial[(size_t) blockIdx.x * (kWin * kWin) + k] = win[k];
}

// image window = sum of the blocks' windows in a fixed order: slice q
// (blockIdx.y) sums blocks q, q + kSlices, ... per pixel into slice_sum; then
// k_diff_window_final adds the slices in order
constexpr int kSlices = 64;

__global__ __launch_bounds__ (kBlock)
void k_diff_window (const float *__restrict__ win_partial, uint32_t nblocks, ImageSpec sp,
                    float *__restrict__ slice_sum)
{
   const int W = (sp.npix < (uint32_t) kWin) ? (int) sp.npix : kWin;
   const int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= W * W) return;
   float s = 0.0f;
   for (uint32_t b = blockIdx.y; b < nblocks; b += kSlices) s += win_partial[(size_t) b * (kWin * kWin) + k];
   slice_sum[(size_t) blockIdx.y * (kWin * kWin) + k] = s;
}

__global__ __launch_bounds__ (kBlock)
void k_diff_window_final (const float *__restrict__ slice_sum, ImageSpec sp, float *__restrict__ image)
{
   const int W = (sp.npix < (uint32_t) kWin) ? (int) sp.npix : kWin;
   const int wo = ((int) sp.npix - W) / 2;
   const int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= W * W) return;
   float s = 0.0f;
   for (int q = 0; q < kSlices; q++) s += slice_sum[(size_t) q * (kWin * kWin) + k];
   const int wy = k % W, wz = k / W;
   image[(size_t) (wz + wo) * sp.npix + (wy + wo)] = s;
}

__global__ void k_zero_f (float *__restrict__ a, size_t n)
{
   size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
   const size_t stride = (size_t) gridDim.x * blockDim.x;
   for (; i < n; i += stride) a[i] = 0.0f;
}

// per-block partial gradients: 4 shell pairs x kNP parameters
__global__ __launch_bounds__ (kBlock)
void k_diff_adjoint (uint32_t n, DiffOut in, ImageSpec sp, const float *__restrict__ g,
                     double *__restrict__ partial)
{
   constexpr int kG = MXG_NUM_SHELLS * kNP;
   __shared__ double red[kBlock / 64][kG];
   double acc[kG];
#pragma unroll
   for (int q = 0; q < kG; q++) acc[q] = 0.0;
   const BlockRange range = block_range (n, blockIdx.x, gridDim.x);
   const double inv_pix = 1.0 / sp.pix;

   for (uint32_t i = range.lo + threadIdx.x; i < range.hi; i += blockDim.x)
     {
        // a photon of weight 0 (lost, or vignetted to nothing) adds nothing
        const double w = in.w[i];
        if (w == 0.0) continue;
        double dw[kNP], dy[kNP], dz[kNP];
#pragma unroll
        for (int j = 0; j < kNP; j++)
          {
             dw[j] = in.dw[(size_t) j * n + i];
             dy[j] = in.dy[(size_t) j * n + i];
             dz[j] = in.dz[(size_t) j * n + i];
          }
        int i0, j0;
        double fu, fv;
        pixel_coords (sp, in.y[i], in.z[i], i0, j0, fu, fv);
        // bilinear weights and their derivatives by u and v
        const double b[4] = {(1 - fu) * (1 - fv), fu * (1 - fv), (1 - fu) * fv, fu * fv};
        const double bu[4] = {-(1 - fv), (1 - fv), -fv, fv};
        const double bv[4] = {-(1 - fu), -fu, (1 - fu), fu};
        double gsum = 0.0, gu = 0.0, gv = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++)
          {
             const int iy = i0 + (c & 1), iz = j0 + (c >> 1);
             if ((iy < 0) || (iz < 0) || (iy >= (int) sp.npix) || (iz >= (int) sp.npix)) continue;
             const double gc = (double) g[(size_t) iz * sp.npix + iy];
             gsum += gc * b[c];
             gu += gc * bu[c];
             gv += gc * bv[c];
          }
        const int s = in.shell[i];
#pragma unroll
        for (int j = 0; j < kNP; j++)
          {
             const double c = gsum * dw[j] + w * inv_pix * (gu * dy[j] + gv * dz[j]);
#pragma unroll
             for (int k = 0; k < MXG_NUM_SHELLS; k++)
               acc[k * kNP + j] += (s == k) ? c : 0.0;
          }
     }
   const int wave = threadIdx.x >> 6;
#pragma unroll
   for (int q = 0; q < kG; q++)
     {
        double v = acc[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off);
        if (__lane_id () == 0) red[wave][q] = v;
     }
   __syncthreads ();
   if (threadIdx.x < kG)
     {
        double s = 0.0;
        for (int k = 0; k < kBlock / 64; k++) s += red[k][threadIdx.x];
        partial[(size_t) blockIdx.x * kG + threadIdx.x] = s;
     }
}

// out[q] = sum over b of partial[b * width + q], one block per q, in a fixed
// order: thread t sums b = t, t + kBlock, ...; then a tree over the threads
__global__ __launch_bounds__ (kBlock)
void k_diff_sum (const double *__restrict__ partial, uint32_t nblocks, double *__restrict__ out,
                 uint32_t width)
{
   __shared__ double red[kBlock];
   const uint32_t q = blockIdx.x;
   double s = 0.0;
   for (uint32_t b = threadIdx.x; b < nblocks; b += kBlock) s += partial[(size_t) b * width + q];
   red[threadIdx.x] = s;
   __syncthreads ();
   for (int half = kBlock / 2; half > 0; half >>= 1)
     {
        if ((int) threadIdx.x < half) red[threadIdx.x] += red[threadIdx.x + half];
        __syncthreads ();
     }
   if (threadIdx.x == 0) out[q] = red[0];
}

hipError_t launch_diff_forward (const mxg_config *cfg, const float *fa, const double *da,
                                const mxg_diff_cfg *dcfg, uint32_t n, uint32_t tag_start,
                                unsigned long long seed, const mxg_image_spec *spec, float *image,
                                float *win_partial, float *slice_sum, DiffOut out, const ListIO &enter,
                                uint32_t grid, double *wsum_dev, int tangents, hipStream_t st,
                                hipEvent_t *ev)
{
   ImageSpec sp = {spec->npix, spec->pixel_size, spec->center_y, spec->center_z};
   const size_t npix2 = (size_t) spec->npix * spec->npix;
   hipLaunchKernelGGL (k_zero_f, dim3 (256), dim3 (kBlock), 0, st, image, npix2);
   if (ev) (void) hipEventRecord (ev[0], st);
   hipLaunchKernelGGL (k_diff_enter, dim3 (enter.grid), dim3 (kBlock), 0, st, cfg, da, dcfg, n, tag_start,
                       seed, out, enter);
   hipError_t e = launch_compact (enter, st);
   if (e != hipSuccess) return e;
   if (ev) (void) hipEventRecord (ev[1], st);
   if (ev) (void) hipEventRecord (ev[2], st);
   const uint32_t *idx = enter.idx_out, *n_list = enter.n_out;
   if (tangents == kNP)
     hipLaunchKernelGGL (k_diff_forward<kNP>, dim3 (grid), dim3 (kBlock), 0, st, cfg, fa, da, dcfg, idx,
                         n_list, n, tag_start, seed, sp, image, win_partial, out);
   else if (tangents == 2)
     hipLaunchKernelGGL (k_diff_forward<2>, dim3 (grid), dim3 (kBlock), 0, st, cfg, fa, da, dcfg, idx,
                         n_list, n, tag_start, seed, sp, image, win_partial, out);
   else
     hipLaunchKernelGGL (k_diff_forward<0>, dim3 (grid), dim3 (kBlock), 0, st, cfg, fa, da, dcfg, idx,
                         n_list, n, tag_start, seed, sp, image, win_partial, out);
   if (ev) (void) hipEventRecord (ev[3], st);
   const int W = (spec->npix < (uint32_t) kWin) ? (int) spec->npix : kWin;
   const uint32_t wblocks = (W * W + kBlock - 1) / kBlock;
   hipLaunchKernelGGL (k_diff_window, dim3 (wblocks, kSlices), dim3 (kBlock), 0, st, win_partial, grid, sp,
                       slice_sum);
   hipLaunchKernelGGL (k_diff_window_final, dim3 (wblocks), dim3 (kBlock), 0, st, slice_sum, sp, image);
   hipLaunchKernelGGL (k_diff_sum, dim3 (1), dim3 (kBlock), 0, st, out.block_w, grid, wsum_dev, 1u);
   return hipGetLastError ();
}

hipError_t launch_diff_adjoint (uint32_t n, DiffOut in, const mxg_image_spec *spec, const float *g,
                                double *partial, uint32_t grid, double *grad_dev, hipStream_t st,
                                hipEvent_t *ev)
{
   ImageSpec sp = {spec->npix, spec->pixel_size, spec->center_y, spec->center_z};
   if (ev) (void) hipEventRecord (ev[0], st);
   hipLaunchKernelGGL (k_diff_adjoint, dim3 (grid), dim3 (kBlock), 0, st, n, in, sp, g, partial);
   if (ev) (void) hipEventRecord (ev[1], st);
   hipLaunchKernelGGL (k_diff_sum, dim3 (MXG_NUM_SHELLS * kNP), dim3 (kBlock), 0, st, partial, grid,
                       grad_dev, (uint32_t) (MXG_NUM_SHELLS * kNP));
   return hipGetLastError ();
}

} // namespace mxg
