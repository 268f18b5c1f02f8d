// grating_stage.h -- the per-photon work of the HETG / LETG stage, shared by
// the stage kernel (grating.hip) and the one-launch batch pass (persist.hip).
//
// Per photon (the operations of marx/libsrc/diffract.c:974-1130):
//   grating vignetting draw -> rotate about X into the grating frame ->
//   Rowland-torus intersection by Newton on the quartic (:622-687) ->
//   order draw against the cumulative efficiency table interpolated at the
//   photon energy (the per-photon form of the shell-major
//   JDMinterpolate_n_fvector pass, finterpo.c:134-193) -> facet basis,
//   sector lookup, dtheta and dP/P blur, grating equation (:706-823) ->
//   LETG: fine support grating once, coarse three times (:917-970, 1081-1122)
//   -> rotate back.  Efficiency tables are energy-major so the order walk of
//   one lane touches two contiguous rows.
#pragma once

#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

// rotation about x by the angle of cosine c and sine s
__device__ __forceinline__ vec3 rotate_x (vec3 a, double c, double s)
{
   double ay = a.y, az = a.z;
   a.y = c * ay - s * az;
   a.z = s * ay + c * az;
   return a;
}

__device__ __forceinline__ int newtons_quartic (double a, double b, double c, double d, double t0,
                                                double &tp)
{
   unsigned int max_it = 10;
   double a2 = 2.0 * a, a3 = 3.0 * a, b2 = 2.0 * b, t;
   while (true)
     {
        double t2 = t0 * t0;
        double num = t2 * (3.0 * t2 + a2 * t0 + b) - d;
        double den = t2 * (4.0 * t0 + a3) + b2 * t0 + c;
        t = num / den;
        if (fabs (t - t0) < 1.0e-4) break;
        max_it--;
        if (max_it == 0) return -1;
        t0 = t;
     }
   tp = t;
   return 0;
}

__device__ __forceinline__ int torus_intersect (vec3 &x0, vec3 p, double rowland)
{
   double t = -x0.x / p.x;
   x0.x = 0.0;
   x0.y = x0.y + p.y * t;
   x0.z = x0.z + p.z * t;
   double pxpz_len = p.x * p.x + p.z * p.z;
   double x2 = x0.z * x0.z + x0.y * x0.y;
   double r2 = rowland * rowland;
   double pdotx = dot (p, x0);
   double a = 4.0 * pdotx;
   double b = 2.0 * x2 + a * pdotx - r2 * pxpz_len;
   double c = a * x2 - 2.0 * r2 * p.z * x0.z;
   double d = x2 * x2 - r2 * x0.z * x0.z;
   double t0 = -rowland * sqrt (pxpz_len);
   if (-1 == newtons_quartic (a, b, c, d, t0, t))
     return -1;
   x0 = ax1_bx2 (1.0, x0, t, p);
   return 0;
}

// Order draw: first k with r <= cumulative efficiency(E); -1 when none.
// The bracket is the one the reference's merge walk over sorted energies
// reaches: the first j >= 1 with E[j] >= e, else the last interval.  E:
// the grating's energies, bcnt: their bucket counts (LDS or the arena).
__device__ __forceinline__ int draw_order (const mxg_grating &g, const float *__restrict__ fa,
                                           const float *E, const float *bcnt, double e, double r)
{
   const unsigned int ne = g.num_energies, no = g.num_orders;
   unsigned int lo = 1, hi = ne - 1;
   const mxg_bucket &b = g.e_bucket;
   const float ef = (float) e;         // e holds a float value
   if (b.ok && (ef > E[0]))
     {
        // the first entry >= e of a strictly increasing table, from the
        // bucket that holds e (mxg_bucket); the bisection below then starts
        // on the answer clamped to [1, ne - 1]
        const uint64_t k = ((uint64_t) __float_as_uint (ef) - b.base) >> b.shift;
        const uint32_t j = (k < b.nb) ? (uint32_t) k : b.nb - 1;
        unsigned int l = (unsigned int) bcnt[j], h = (unsigned int) bcnt[j + 1];
        while (l < h)
          {
             unsigned int mid = (l + h) >> 1;
             if (E[mid] < ef) l = mid + 1;
             else h = mid;
          }
        lo = hi = (l < 1) ? 1u : ((l > ne - 1) ? ne - 1 : l);
     }
   while (lo < hi)
     {
        unsigned int mid = (lo + hi) >> 1;
        if (e > (double) E[mid]) lo = mid + 1;
        else hi = mid;
     }
   const unsigned int count = lo;
   const double x_0 = E[count - 1], x_1 = E[count];
   const double dx_10 = x_1 - x_0;
   const float *y0 = fa + g.cum_eff_off + (count - 1) * no;
   const float *y1 = y0 + no;
   // Inside its interval the interpolated value of order k lies between
   // y0[k] and y1[k] (every step of it rounds within those bounds), so the
   // division is needed only for an order whose bounds straddle r; outside
   // (extrapolation) every order is interpolated.
   const bool inside = (e >= x_0) && (e <= x_1);
   unsigned int k0 = 0;
   if (inside && g.cum_sorted)
     {
        // rows that rise with the order: the bounds do too, so every order
        // before the first whose upper bound reaches r fails, found by
        // bisection
        unsigned int l = 0, h = no;
        while (l < h)
          {
             const unsigned int mid = (l + h) >> 1;
             if (r > (double) fmaxf (y0[mid], y1[mid])) l = mid + 1;
             else h = mid;
          }
        k0 = l;
     }
   for (unsigned int k = k0; k < no; k++)
     {
        const float ya = y0[k], yb = y1[k];
        if (inside)
          {
             if (r > (double) fmaxf (ya, yb)) continue;
             if (r <= (double) fminf (ya, yb)) return (int) k;
          }
        float v;
        if (dx_10 == 0.0)
          v = ya;
        else
          {
             double a = ya, b = yb;
             v = (float) (a + (b - a) * (e - x_0) / dx_10);
          }
        if (r <= (double) v)
          return (int) k;
     }
   return -1;
}

// nsec, sec: the shell's sector count and its six columns (LDS or the
// arena), nsec = 0 without a sector table
__device__ __forceinline__ int diffract_photon (const mxg_grating &g, unsigned int nsec, const double *sec,
                                                double theta, double energy, vec3 x, vec3 &p, int order,
                                                Stream &st)
{
   double n_lambda_over_d = (double) order * (2.0 * MXG_PI * MXG_HBAR_C) / g.period / energy;
   vec3 n = unit (x);
   n.x = -n.x; n.y = -n.y; n.z = -n.z;
   vec3 l = v3 (n.z, 0.0, -n.x);
   l = unit (l);
   vec3 d = cross (n, l);
   double dtheta, dp_over_p;

   if (nsec > 0)
     {
        const double *min_a = sec;
        const double *max_a = min_a + nsec;
        double sector = atan2 (x.y, x.z);
        if (sector < 0) sector = 2 * MXG_PI + sector;
        unsigned int sn = search_d (sector, min_a, nsec) - 1;
        if (sn >= nsec) return -1;
        if ((max_a[sn] <= sector) || (min_a[sn] > sector))
          return -1;
        dtheta = min_a[2 * nsec + sn] + min_a[3 * nsec + sn] * gaussian (st);
        dp_over_p = min_a[4 * nsec + sn] + min_a[5 * nsec + sn] * gaussian (st);
     }
   else
     {
        dtheta = g.theta_blur * gaussian (st);
        dp_over_p = g.dp_over_p * gaussian (st);
     }

   theta -= dtheta;
   if (theta != 0.0)
     {
        double s, c;
        sincos (theta, &s, &c);
        vec3 l_tmp = l, d_tmp = d;
        l = ax1_bx2 (c, l_tmp, s, d_tmp);
        d = ax1_bx2 (-s, l_tmp, c, d_tmp);
     }

   double p_d = n_lambda_over_d + dot (p, d);
   double p_l = dot (p, l);
   double p_n = 1.0 - p_l * p_l - p_d * p_d;
   if (p_n < 0.0) return -1;
   p_n = sqrt (p_n);
   vec3 q = ax1_bx2 (p_d, d, p_n, n);
   p = ax1_bx2 (1.0, q, p_l, l);
   if (dp_over_p == 0) return 0;
   double factor = n_lambda_over_d * dp_over_p;
   vec3 dp = ax1_bx2 (-factor, d, factor * (p_d / p_n), n);
   p.x += dp.x;
   p.y += dp.y;
   p.z += dp.z;
   p = unit (p);
   return 0;
}

// The small tables of grating g (energies, their bucket counts, orders)
// and its shell's sector columns, where the order draw and the diffraction
// search them.
struct GratingTabs
{
   const float *E, *bcnt, *orders;
   const double *sec;
   unsigned int nsec;
};

// one order draw plus diffraction off grating g; returns 0 or -1
__device__ __forceinline__ int diffract_from (const mxg_grating &g, const GratingTabs &t,
                                              const float *__restrict__ fa, double theta,
                                              double energy, double ef, vec3 x, vec3 &p,
                                              int8_t &order_out, Stream &st)
{
   double r = uniform (st);
   int k = draw_order (g, fa, t.E, t.bcnt, ef, r);
   if (k < 0) return -1;
   int order = (int) t.orders[k];
   order_out = (int8_t) order;
   return diffract_photon (g, t.nsec, t.sec, theta, energy, x, p, order, st);
}

// The search tables of the grating stage staged in LDS by k_grating:
// offsets in floats from the table area of gratings 0-3 (by shell), 4 (the
// LETG fine support grating) and 5 (coarse), and of each shell's sector
// columns (doubles, at even offsets).
struct GratingLds
{
   uint32_t e[6], bk[6], ord[6];
   uint32_t sec[MXG_NUM_SHELLS];
   uint32_t floats;
};

// Dynamic LDS of k_grating: the layout at 0, the tables from kGratingTabOff.
constexpr size_t kGratingTabOff = (sizeof (GratingLds) + 15) & ~(size_t) 15;

__host__ __device__ inline const mxg_grating &grating_of (const mxg_grating_cfg &gc, int k)
{
   return (k < MXG_NUM_SHELLS) ? gc.g[k] : ((k == MXG_NUM_SHELLS) ? gc.fine : gc.coarse);
}

// The LDS layout of configuration gc's search tables (host side); letg:
// with the support gratings.
inline GratingLds grating_lds_layout (const mxg_grating_cfg &gc, bool letg)
{
   GratingLds lay = {};
   uint32_t f = 0;
   for (int k = 0; k < 6; k++)
     {
        if ((k >= MXG_NUM_SHELLS) && !(letg && ((k == MXG_NUM_SHELLS) ? gc.has_fine : gc.has_coarse)))
          continue;
        const mxg_grating &g = grating_of (gc, k);
        lay.e[k] = f; f += g.num_energies;
        lay.bk[k] = f; f += bucket_floats (g.e_bucket);
        lay.ord[k] = f; f += g.num_orders;
     }
   for (int sh = 0; sh < MXG_NUM_SHELLS; sh++)
     {
        f = (f + 1) & ~1u;
        lay.sec[sh] = f;
        f += 2 * 6 * gc.sectors[sh].n;
     }
   lay.floats = f;
   return lay;
}

// Stages the tables of lay in LDS (the layout at lds, the tables at ltab =
// lds + kGratingTabOff); letg as for the layout.  The caller synchronises.
__device__ __forceinline__ void stage_grating_tables (const mxg_grating_cfg &gc, const float *__restrict__ fa,
                                                      const double *__restrict__ da, const GratingLds &lay,
                                                      bool letg, unsigned char *lds, float *ltab)
{
   if (threadIdx.x == 0) *(GratingLds *) lds = lay;
   for (int k = 0; k < 6; k++)
     {
        if ((k >= MXG_NUM_SHELLS) && !(letg && ((k == MXG_NUM_SHELLS) ? gc.has_fine : gc.has_coarse)))
          continue;
        const mxg_grating &g = grating_of (gc, k);
        for (uint32_t w = threadIdx.x; w < g.num_energies; w += blockDim.x)
          ltab[lay.e[k] + w] = fa[g.energies_off + w];
        const uint32_t nb = bucket_floats (g.e_bucket);
        for (uint32_t w = threadIdx.x; w < nb; w += blockDim.x)
          ltab[lay.bk[k] + w] = fa[g.e_bucket.off + w];
        for (uint32_t w = threadIdx.x; w < g.num_orders; w += blockDim.x)
          ltab[lay.ord[k] + w] = fa[g.orders_off + w];
     }
   for (int sh = 0; sh < MXG_NUM_SHELLS; sh++)
     {
        double *dst = (double *) (ltab + lay.sec[sh]);
        const uint32_t n = 6 * gc.sectors[sh].n;
        for (uint32_t w = threadIdx.x; w < n; w += blockDim.x)
          dst[w] = da[gc.sectors[sh].off + w];
     }
}

// The tables of grating k (0-3 by shell, 4 fine, 5 coarse) and, for k < 4,
// of shell k's sectors: from LDS (STAGED: layout L, tables ltab) or from the
// arenas.
template <bool STAGED>
__device__ __forceinline__ GratingTabs grating_tabs (const mxg_grating_cfg &gc, int k, const float *__restrict__ fa,
                                                     const double *__restrict__ da, const GratingLds *L,
                                                     const float *ltab)
{
   GratingTabs t;
   const mxg_grating &g = grating_of (gc, k);
   const bool shell = k < MXG_NUM_SHELLS;
   t.nsec = shell ? gc.sectors[k].n : 0u;
   if constexpr (STAGED)
     {
        t.E = ltab + L->e[k];
        t.bcnt = ltab + L->bk[k];
        t.orders = ltab + L->ord[k];
        t.sec = shell ? (const double *) (ltab + L->sec[k]) : nullptr;
     }
   else
     {
        t.E = fa + g.energies_off;
        t.bcnt = fa + g.e_bucket.off;
        t.orders = fa + g.orders_off;
        t.sec = shell ? da + gc.sectors[k].off : nullptr;
     }
   return t;
}

// k_grating's work for photon id: vignetting, the Rowland torus, the order
// draw and diffraction, and for LETG the support gratings.  True when the
// photon leaves the grating stage alive, its columns then also in o.
// RELAY: its incoming x, p, energy, tag, flags and shell are read from row
// `slot` of *relay instead of the store, where the mirror left x, p and
// flags for this stage to write (a vignetted photon's as they came).
// HANDOFF: the photons it keeps go on to the detector through o, and the
// detector stores their x, p and flags; only the photons it drops have
// theirs stored here.  STAGED: the search tables are read from LDS (layout
// L, tables ltab, as k_grating stages them), else from the arenas.
template <bool LETG, bool RELAY = false, bool HANDOFF = false, bool STAGED = false>
__device__ __forceinline__ bool grating_one (const mxg_grating_cfg &gc, const float *__restrict__ fa,
                                             const double *__restrict__ da, Soa s, uint32_t id,
                                             unsigned long long seed, uint32_t source_id, PhotonOut &o,
                                             const Relay *relay = nullptr, uint32_t slot = 0,
                                             const GratingLds *L = nullptr, const float *ltab = nullptr)
{
   bool keep = false;
   uint32_t flags = RELAY ? relay->flags[slot] : s.flags[id];
   if ((flags & BAD_PHOTON_MASK) == 0)
     {
        Stream st;
        const uint32_t tag = RELAY ? relay->tag[slot] : s.tag[id];
        stream_init (st, seed, tag, MXG_STAGE_GRATING, source_id);
        const int shell = RELAY ? (int) relay->shell[slot] : (int) s.shell[id];
        const mxg_grating &g = gc.g[shell];
        if (uniform (st) > g.vig)
          {
             s.flags[id] = flags | PHOTON_MIRROR_VBLOCKED;
             if (RELAY)
               {
                  s.x0[id] = relay->x0[slot]; s.x1[id] = relay->x1[slot]; s.x2[id] = relay->x2[slot];
                  s.p0[id] = relay->p0[slot]; s.p1[id] = relay->p1[slot]; s.p2[id] = relay->p2[slot];
               }
          }
        else
          {
             vec3 x = RELAY ? v3 (relay->x0[slot], relay->x1[slot], relay->x2[slot])
                            : v3 (s.x0[id], s.x1[id], s.x2[id]);
             vec3 p = RELAY ? v3 (relay->p0[slot], relay->p1[slot], relay->p2[slot])
                            : v3 (s.p0[id], s.p1[id], s.p2[id]);
             const double energy = RELAY ? relay->energy[slot] : s.energy[id];
             const double ef = (double) (float) energy;
             int8_t order = s.order[id];

             x = rotate_x (x, g.disp_cos, -g.disp_sin);
             p = rotate_x (p, g.disp_cos, -g.disp_sin);
             if (-1 == torus_intersect (x, p, gc.rowland[shell]))
               flags |= PHOTON_UNDIFFRACTED;
             if (-1 == diffract_from (g, grating_tabs<STAGED> (gc, shell, fa, da, L, ltab), fa, 0.0, energy,
                                      ef, x, p, order, st))
               flags |= PHOTON_UNDIFFRACTED;
             s.order[id] = order;

             if constexpr (LETG)
               {
                  for (int pass = 0; pass < 4; pass++)
                    {
                       const bool fine = (pass == 0);
                       if (fine ? !gc.has_fine : !gc.has_coarse) continue;
                       if (flags & BAD_PHOTON_MASK) continue;
                       const double theta = (pass == 0) ? MXG_PI / 2.0
                         : (pass == 1) ? MXG_PI / 3.0
                         : (pass == 2) ? 2.0 * MXG_PI / 3.0 : 0.0;
                       int8_t so = s.sup[pass][id];
                       if (-1 == diffract_from (fine ? gc.fine : gc.coarse,
                                                grating_tabs<STAGED> (gc, fine ? MXG_NUM_SHELLS : MXG_NUM_SHELLS + 1,
                                                                      fa, da, L, ltab),
                                                fa, theta, energy, ef, x, p, so, st))
                         flags |= PHOTON_UNDIFFRACTED;
                       s.sup[pass][id] = so;
                    }
               }

             x = rotate_x (x, g.disp_cos, g.disp_sin);
             p = rotate_x (p, g.disp_cos, g.disp_sin);
             keep = (flags & BAD_PHOTON_MASK) == 0;
             if (!(HANDOFF && keep))
               {
                  s.x0[id] = x.x; s.x1[id] = x.y; s.x2[id] = x.z;
                  s.p0[id] = p.x; s.p1[id] = p.y; s.p2[id] = p.z;
                  s.flags[id] = flags;
               }
             o.x0 = x.x; o.x1 = x.y; o.x2 = x.z;
             o.p0 = p.x; o.p1 = p.y; o.p2 = p.z;
             o.energy = energy;
             o.tag = tag;
             o.flags = flags;
             o.shell = shell;
          }
     }
   return keep;
}

} // namespace mxg
