// detect_stage.h -- the per-photon work of focal-plane detection, shared by
// the stage kernel (detect.hip) and the one-launch batch pass (persist.hip).
//
// ACIS-S / ACIS-I (marx/libsrc/acis-s.c:138-248, acis-i.c:56-167): detector
//   transform, first chip whose plane hit lies inside it (detector.c:56-168),
//   pixel = offset / pixel size, QE draw against the CALDB QE curve times
//   the contamination's transmission (aciscontam.c:93-185), FEF PHA and PI
//   (acis_fef.c), frame-transfer streak.
// HRC-S / HRC-I (hrc-s.c:136-312, hrc-i.c:66-185): HESF plates first
//   (drake.c:268-369), MCP QE draw, UV/ion-shield region and filter draw,
//   pulse height, MCP position blur (hrcblur.c:239-297), u/v and chip pixel
//   (hrc_s_geom.c:344-394, hrc_i_geom.c:146-156).
#pragma once

#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

__device__ __forceinline__ vec3 rot (const double *m, vec3 v)
{
   return v3 (m[0] * v.x + m[1] * v.y + m[2] * v.z,
              m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

__device__ __forceinline__ vec3 rot_inv (const double *m, vec3 v)
{
   return v3 (m[0] * v.x + m[3] * v.y + m[6] * v.z,
              m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

// 1 inside, 0 outside (must_hit), -1 parallel; writes x/dx/dy unless 0
__device__ __forceinline__ int plane_hit (const mxg_chip &g, vec3 x0, vec3 p, vec3 &x,
                                          double &dx, double &dy, bool must_hit)
{
   vec3 normal = ld3 (g.normal);
   double pdotn = dot (p, normal);
   if (pdotn == 0) return -1;
   vec3 xll = ld3 (g.x_ll);
   vec3 r = sub (x0, xll);
   r = ax1_bx2 (1.0, r, -1.0 * dot (r, normal) / pdotn, p);
   int hit = 1;
   double rx = dot (r, ld3 (g.xhat));
   if ((rx < 0.0) || (rx >= g.xlen))
     {
        if (must_hit) return 0;
        hit = 0;
     }
   double ry = dot (r, ld3 (g.yhat));
   if ((ry < 0.0) || (ry >= g.ylen))
     {
        if (must_hit) return 0;
        hit = 0;
     }
   x = add (r, xll);
   dx = rx;
   dy = ry;
   return hit;
}

__device__ __forceinline__ int find_chip (const mxg_detector_cfg &d, vec3 &x, vec3 p, double &dx, double &dy)
{
   for (int k = 0; k < d.num_chips; k++)
     if (1 == plane_hit (d.chip[k], x, p, x, dx, dy, true))
       return k;
   if (d.extend == 0) return -1;
   int best = -1;
   double best_r2 = -1, bdx = 0, bdy = 0;
   vec3 bx = x;
   for (int k = 0; k < d.num_chips; k++)
     {
        vec3 xk;
        double ddx, ddy;
        if (-1 == plane_hit (d.chip[k], x, p, xk, ddx, ddy, false)) continue;
        double ex = ddx - 0.5 * d.chip[k].xlen, ey = ddy - 0.5 * d.chip[k].ylen;
        double r2 = ex * ex + ey * ey;
        if ((r2 < best_r2) || (best < 0))
          {
             best = k; best_r2 = r2; bx = xk; bdx = ddx; bdy = ddy;
          }
     }
   if (best >= 0)
     {
        x = bx; dx = bdx; dy = bdy;
     }
   return best;
}

__device__ __forceinline__ double qe_at (const mxg_qe_table &q, const float *__restrict__ fa, double energy)
{
   return interp_fb ((float) energy, fa + q.e_off, fa + q.eff_off, q.n, q.bucket, fa + q.bucket.off);
}

__device__ __forceinline__ int filter_region (const mxg_detector_cfg &d, double y, double z)
{
   y -= d.shield_y_center;
   z -= d.shield_z_center;
   if (y < 0)
     {
        y = -y;
        if (y < d.shield_l) return 0;
        if (y < d.shield_sl) return (z >= d.shield_t) ? 0 : 1;
        if (y >= d.shield_sl_gap) return (z >= d.shield_t) ? 2 : 3;
        return -1;
     }
   if (y < d.shield_r) return 0;
   if (y < d.shield_sr) return (z >= d.shield_t) ? 0 : 1;
   if (y >= d.shield_sr_gap) return (z >= d.shield_t) ? 2 : 3;
   return -1;
}

__device__ __forceinline__ int16_t hrc_pha (double energy, Stream &st)
{
   if (energy <= 0.5) energy = 141.582 * sqrt (energy);
   else if (energy < 2.0) energy = 107.299 * pow (energy, 0.1);
   else energy = 115.0;
   energy = energy * (1.0 + 0.424661 * gaussian (st));
   if (energy < 0.0) energy = 0.0;
   return (int16_t) energy;
}

// HESF: returns the plate index hit (and moves x onto it), -1 when none
__device__ __forceinline__ int drake_hit (const mxg_detector_cfg &d, vec3 &x, vec3 p,
                                          double &p_dot_n, bool &use_cr)
{
   const vec3 old_x = x;
   for (int k = 0; k < d.drake_n; k++)
     {
        const mxg_rect &r = d.drake[k];
        vec3 nrm = ld3 (r.normal);
        p_dot_n = dot (p, nrm);
        if (p_dot_n == 0.0) continue;
        vec3 a = ld3 (r.a);
        double t = dot (sub (a, old_x), nrm) / p_dot_n;
        vec3 new_x = ax1_bx2 (1.0, old_x, t, p);
        vec3 xp = sub (new_x, a);
        double xx = dot (xp, ld3 (r.e1));
        if ((xx < 0.0) || (xx >= r.len1)) continue;
        double yy = dot (xp, ld3 (r.e2));
        if ((yy < 0.0) || (yy >= r.len2)) continue;
        use_cr = (xx < d.drake_cr_width);
        x = new_x;
        return k;
     }
   return -1;
}

// ---- ACIS FEF: PHA and PI (acis_fef.c:380-503, 509-579, 910-1079) ---------------

constexpr int kFefMaxGauss = 16;
constexpr int kFefRetries = 10000;   // the reference's outer sampling loop has no bound

struct FefGauss
{
   float amp[kFefMaxGauss], center[kFefMaxGauss], sigma[kFefMaxGauss], cum[kFefMaxGauss];
   bool tail[kFefMaxGauss];
};

// the reference's amplitudes carry the 1 / sqrt(2 pi sigma^2) (acis_fef.c:372-392)
__device__ __forceinline__ double fef_integral (double xmin, double xmax, float amp, float center, float sigma)
{
   const double s = sigma * 1.4142135623730951;
   if (s == 0.0) return 0.0;
   const double x0 = center;
   return 0.5 * amp * (erf ((xmax - x0) / s) - erf ((xmin - x0) / s)) * (2.5066282746310002 * sigma);
}

// one PHA from the positive-amplitude Gaussians; false when none is found
static __device__ bool fef_pos_pha (const FefGauss &G, int ng, Stream &st, double &pha_out)
{
   for (int it = 0; it < kFefRetries; it++)
     {
        const double r = uniform (st);
        for (int k = 0; k < ng; k++)
          {
             if (G.cum[k] <= r) continue;
             double pha;
             if (!G.tail[k])
               {
                  int count = 0;
                  do
                    {
                       pha = G.center[k] + G.sigma[k] * gaussian (st);
                       count++;
                    }
                  while ((pha < 0) && (count < 100));
               }
             else
               {
                  // Gaussian tail beyond 0 (the GSL algorithm, as acis_fef.c:444-458)
                  const double s = (0 - G.center[k]) / G.sigma[k];
                  double u, x;
                  int guard = 0;
                  do
                    {
                       double v;
                       u = uniform (st);
                       do v = uniform (st);
                       while (v == 0.0);
                       x = sqrt (s * s - 2 * log (v));
                    }
                  while ((x * u > s) && (++guard < kFefRetries));
                  pha = G.center[k] + x * G.sigma[k];
               }
             if (pha < 0) break;
             pha_out = pha;
             return true;
          }
     }
   return false;
}

// PHA and PI of a detected photon at chip pixel (x, y); false marks it undetected
static __device__ __noinline__ bool apply_fef (const mxg_detector_cfg &d, const float *__restrict__ fa,
                           const double *__restrict__ da, int ccd, float x, float y, double energy,
                           Stream &st, short &pha_out, float &pi_out)
{
   if (x < 0) { if (d.extend == 0) return false; x = 0; }
   else if (x >= 1024) { if (d.extend == 0) return false; x = 1023; }
   if (y < 0) { if (d.extend == 0) return false; y = 0; }
   else if (y >= 1024) { if (d.extend == 0) return false; y = 1023; }
   const unsigned int i = (unsigned int) (x / 32), j = (unsigned int) (y / 32);
   if ((i >= 32) || (j >= 32) || (ccd < 0) || (ccd >= 10)) return false;
   const double kf = da[d.fef_map_off + (ccd * 32 + i) * 32 + j];
   if (kf < 0) return false;
   const double *desc = da + d.fef_desc_off + 5 * (unsigned int) kf;
   const unsigned int ne = (unsigned int) desc[0];
   const int ng = (int) desc[1];
   const float *e = fa + (unsigned int) desc[2];
   const float *ch = fa + (unsigned int) desc[3];
   unsigned int n = search_f ((float) energy, e, ne);
   if (n == 0) n++;
   if (n == ne) n--;
   const double t = (energy - e[n - 1]) / (e[n] - e[n - 1]);
   const float *g0 = fa + (unsigned int) desc[4] + 3 * (n - 1) * ng, *g1 = g0 + 3 * ng;
   FefGauss G;
   double total_pos = 0.0, total_neg = 0.0;
   for (int k = 0; k < ng; k++)
     {
        const float a0 = g0[3 * k], c0 = g0[3 * k + 1], s0 = g0[3 * k + 2];
        const float a1 = g1[3 * k], c1 = g1[3 * k + 1], s1 = g1[3 * k + 2];
        G.center[k] = (float) (c0 + t * (c1 - c0));
        double v = s0 + t * (s1 - s0);
        if (v <= 0.0)
          {
             G.sigma[k] = 0.0f;
             G.amp[k] = 0.0f;
          }
        else
          {
             G.sigma[k] = (float) v;
             v = a0 + t * (a1 - a0);
             if ((v < 0.0) && ((a1 > 0.0f) || (a0 > 0.0f))) v = 0.0;
             G.amp[k] = (float) v;
          }
        // normalize_gaussians
        const double area1 = fef_integral (0, 1e37, G.amp[k], G.center[k], G.sigma[k]);
        const double area2 = fef_integral (-1e37, 0, G.amp[k], G.center[k], G.sigma[k]);
        G.tail[k] = (area2 > area1) && (area1 / area2 < 0.1);
        if (area1 >= 0) total_pos += area1;
        else total_neg -= area1;
        G.cum[k] = (float) total_pos;
     }
   if (total_pos <= total_neg) return false;
   if (total_pos > 0)
     for (int k = 0; k < ng; k++) G.cum[k] = (float) (G.cum[k] / total_pos);
   double pha;
   if (total_neg == 0.0)
     {
        if (!fef_pos_pha (G, ng, st, pha)) return false;
     }
   else
     {
        // rejection against the full sum, positive parts as the envelope
        bool ok = false;
        for (int count = 0; (count < 100) && !ok; count++)
          {
             if (!fef_pos_pha (G, ng, st, pha)) return false;
             double pos_sum = 0.0, sum = 0.0;
             for (int k = 0; k < ng; k++)
               {
                  double ds = 0.0;
                  if (G.sigma[k] != 0.0f)
                    {
                       const double z = (pha - G.center[k]) / G.sigma[k];
                       ds = G.amp[k] * exp (-0.5 * z * z);
                    }
                  sum += ds;
                  if (ds > 0) pos_sum += ds;
               }
             ok = (uniform (st) * pos_sum < sum);
          }
        if (!ok) return false;
     }
   const short ipha = (short) pha;
   const double frac = ipha - uniform (st);
   const float pi = interp_f ((float) frac, ch, e, ne);
   if (pi < 0) return false;
   pha_out = ipha;
   pi_out = pi;
   return true;
}

// FEF: compiled with the ACIS FEF step (launched only when a FEF file was
// read, so that the common case keeps the registers of the kernel without
// it); W: waves per SIMD to leave room for (0 unbounded: ACIS and HRC both
// measured slower bounded)
// ACIS contamination: the transmission exp(-sum_i mu_i(E) (tau0_i + tau1_i
// f(x, y))) at the chip pixel (cx, cy) (aciscontam.c:93-136), f the chip's
// table or one of the two analytic models (aciscontam.c:141-185).  Out of
// line: only runs with a contamination file.
static __device__ __noinline__ double contam_transmission (const mxg_contam &c, const float *__restrict__ fa,
                                                    double energy, double cx, double cy)
{
   const float ef = (float) energy;
   double v = 0.0;
   if (c.fxy_model == MXG_CONTAM_TABLE)
     {
        if ((cx < 0) || (cx >= 1024) || (cy < 0) || (cy >= 1024))
          return 0.0;
        cx /= c.blocking;
        cy /= c.blocking;
        const unsigned int ofs = (1024u / c.blocking) * (unsigned int) cy + (unsigned int) cx;
        for (unsigned int i = 0; i < c.num_layers; i++)
          {
             double mu = interp_f (ef, fa + c.mu_e_off[i], fa + c.mu_off[i], c.mu_n[i]);
             double fxy = fa[c.fxy_off[i] + ofs];
             v += mu * (c.tau0[i] + c.tau1[i] * fxy);
          }
        return exp (-v);
     }
   double fxy;
   if (c.fxy_model == MXG_CONTAM_RADIAL)
     {
        double dx = cx - c.x0, dy = cy - c.y0;
        double r = (8.0 / 1024.0) * sqrt (dx * dx + dy * dy);
        r /= 8.07;
        fxy = 1.29 * r * r;
     }
   else if (cy <= 512.0)
     fxy = pow (fabs ((cy - 512.0) / (64.0 - 512.0)), 5.5);
   else
     fxy = pow (fabs ((cy - 512.0) / (964.0 - 512.0)), 4.5);
   for (unsigned int i = 0; i < c.num_layers; i++)
     {
        double mu = interp_f (ef, fa + c.mu_e_off[i], fa + c.mu_off[i], c.mu_n[i]);
        v += mu * (c.tau0[i] + c.tau1[i] * fxy);
     }
   return exp (-v);
}

// k_detect's work for photon id (KIND: MXG_DET_ACIS, MXG_DET_HRC_S,
// MXG_DET_HRC_I; FEF: with the ACIS FEF step).  True when the photon is
// detected.  RELAY: the photon's incoming x, p, energy, tag and flags are
// read from row `slot` of *relay instead of the store.
template <bool FEF, int KIND, bool RELAY = false>
__device__ __forceinline__ bool detect_one (const mxg_detector_cfg &d, const float *__restrict__ fa,
                                            const double *__restrict__ da, Soa s, uint32_t id, double tstart,
                                            unsigned long long seed, uint32_t source_id,
                                            const Relay *relay = nullptr, uint32_t slot = 0)
{
   constexpr bool is_acis = (KIND == MXG_DET_ACIS);
   constexpr bool is_hrc_s = (KIND == MXG_DET_HRC_S);
   bool keep = false;
   uint32_t flags = RELAY ? relay->flags[slot] : s.flags[id];
   if ((flags & BAD_PHOTON_MASK) == 0)
     {
        Stream st;
        stream_init (st, seed, RELAY ? relay->tag[slot] : s.tag[id], MXG_STAGE_DETECT, source_id);
        vec3 x = RELAY ? v3 (relay->x0[slot], relay->x1[slot], relay->x2[slot]) : v3 (s.x0[id], s.x1[id], s.x2[id]);
        vec3 p = RELAY ? v3 (relay->p0[slot], relay->p1[slot], relay->p2[slot]) : v3 (s.p0[id], s.p1[id], s.p2[id]);
        const double energy = RELAY ? relay->energy[slot] : s.energy[id];
        bool detect = true;

        if (is_hrc_s && d.use_drake)
          {
             double p_dot_n = 0.0;
             bool use_cr = false;
             int k = drake_hit (d, x, p, p_dot_n, use_cr);
             if (k >= 0)
               {
                  unsigned int nt = use_cr ? d.drake_cr_n : d.drake_c_n;
                  double rfl = 1.0;
                  if (nt != 0)
                    {
                       unsigned int eo = use_cr ? d.drake_cr_e_off : d.drake_c_e_off;
                       unsigned int bo = use_cr ? d.drake_cr_beta_off : d.drake_c_beta_off;
                       unsigned int dof = use_cr ? d.drake_cr_delta_off : d.drake_c_delta_off;
                       float ef = (float) energy;
                       float beta = interp_f (ef, fa + eo, fa + bo, nt);
                       float delta = interp_f (ef, fa + eo, fa + dof, nt);
                       rfl = reflectivity (fabs (p_dot_n), (double) beta, (double) delta);
                    }
                  if (rfl < uniform (st))
                    {
                       flags |= PHOTON_DRAKE_BLOCKED;
                       detect = false;
                    }
                  else
                    {
                       p = ax1_bx2 (1.0, p, -2.0 * p_dot_n, ld3 (d.drake[k].normal));
                       flags |= PHOTON_DRAKE_REFLECTED;
                    }
               }
          }

        if (detect)
          {
             x.x -= d.offset[0];
             x.y -= d.offset[1];
             x.z -= d.offset[2];
             x = rot (d.matrix, x);
             p = rot (d.matrix, p);
             double dx = 0.0, dy = 0.0;
             int k = find_chip (d, x, p, dx, dy);
             if (k < 0)
               {
                  flags |= PHOTON_MISSED_DETECTOR;
                  s.ccd[id] = -1;
               }
             else if constexpr (is_acis)
               {
                  const mxg_chip &g = d.chip[k];
                  s.ccd[id] = (int8_t) g.id;
                  float ypix = (float) (dx / g.x_pixel_size);
                  float zpix = (float) (dy / g.y_pixel_size);
                  bool detected = true;
                  if (d.ideal == 0)
                    {
                       double r = uniform (st);
                       double qe = 1.0, qe_contam = 1.0;
                       if (g.qe_n != 0)
                         qe = interp_fb ((float) energy, fa + g.qe_e_off, fa + g.qe_off, g.qe_n, g.qe_bucket,
                                         fa + g.qe_bucket.off);
                       if (d.contam_on && d.contam[k].num_layers)
                         qe_contam = contam_transmission (d.contam[k], fa, energy, ypix, zpix);
                       // the filter table is empty on the CALDB QE path
                       // (acis-s.c:303-304): filter transmission 1
                       if (r >= qe * 1.0 * qe_contam)
                         {
                            flags |= PHOTON_UNDETECTED;
                            detected = false;
                         }
                    }
                  if (FEF && detected && d.fef_on)
                    {
                       short pha = -1;
                       float pi = 0.0f;
                       if (apply_fef (d, fa, da, g.id, ypix, zpix, energy, st, pha, pi))
                         {
                            s.pha[id] = pha;
                            s.pi[id] = pi;
                         }
                       else
                         {
                            s.pha[id] = -1;
                            s.pi[id] = 0.0f;
                            flags |= PHOTON_UNDETECTED;
                            detected = false;
                         }
                    }
                  if (detected && (d.frame_transfer_time > 0.0))
                    {
                       double t = fmod (tstart + s.t[id], d.frame_time);
                       if (t > d.exposure_time)
                         {
                            zpix = (float) (1.0 + 1022.0 * uniform (st));
                            double xpixel = ((double) ypix - g.xpixel_offset) * g.x_pixel_size;
                            double ypixel = ((double) zpix - g.ypixel_offset) * g.y_pixel_size;
                            vec3 dxv = ax1_bx2 (xpixel, ld3 (g.xhat), ypixel, ld3 (g.yhat));
                            x = add (ld3 (g.x_ll), dxv);
                            p = sub (x, v3 (d.focal_length, 0, 0));
                            normalize (p);
                            flags |= PHOTON_ACIS_STREAKED;
                         }
                    }
                  s.ypix[id] = ypix;
                  s.zpix[id] = zpix;
               }
             else
               {
                  const mxg_chip &g = d.chip[k];
                  // MCP QE, shield region, filter QE, pulse height
                  bool ok = true;
                  int region = 0;
                  const mxg_qe_table &mq = d.mcp_qe[d.mcp_map[g.id]];
                  if (mq.n != 0)
                    {
                       double qe = qe_at (mq, fa, energy);
                       if (uniform (st) >= qe) ok = false;
                    }
                  if (ok)
                    {
                       const mxg_qe_table *fq;
                       if (is_hrc_s)
                         {
                            double t = (d.shield_x - x.x) / p.x;
                            region = filter_region (d, x.y + t * p.y, x.z + t * p.z);
                            if (region < 0) ok = false;
                            fq = d.filter_qe + (region < 0 ? 0 : region);
                         }
                       else
                         fq = d.filter_qe + g.id;
                       if (ok && (fq->n != 0))
                         {
                            double qe = qe_at (*fq, fa, energy);
                            if (uniform (st) >= qe) ok = false;
                         }
                    }
                  if (!ok)
                    {
                       flags |= PHOTON_UNDETECTED;
                       s.ccd[id] = -1;
                    }
                  else
                    {
                       if (is_hrc_s) s.region[id] = (int8_t) region;
                       s.pha[id] = hrc_pha (energy, st);
                       if (d.ideal == 0)
                         {
                            double r = uniform (st), rr, x_0, y_0;
                            if (r < d.g1_wgt)
                              {
                                 double c;
                                 do c = uniform (st); while (c == 0.0);
                                 rr = d.g1_sigma * sqrt (-2 * log (c));
                                 x_0 = d.g1_xctr; y_0 = d.g1_yctr;
                              }
                            else if (r < d.g1_wgt + d.g2_wgt)
                              {
                                 double c;
                                 do c = uniform (st); while (c == 0.0);
                                 rr = d.g2_sigma * sqrt (-2 * log (c));
                                 x_0 = d.g2_xctr; y_0 = d.g2_yctr;
                              }
                            else
                              {
                                 double c = uniform (st);
                                 double rmax = d.l1_rmax / d.l1_hwhm;
                                 rr = d.l1_hwhm * sqrt (expm1 (c * log1p (rmax * rmax)));
                                 x_0 = d.l1_xctr; y_0 = d.l1_yctr;
                              }
                            double theta = (2.0 * MXG_PI) * uniform (st);
                            double sn, cs;
                            sincos (theta, &sn, &cs);
                            dx += x_0 + rr * cs;
                            dy += y_0 + rr * sn;
                            if (d.extend == 0)
                              {
                                 if (dx < 0.0) dx = 0.0;
                                 if (dy < 0.0) dy = 0.0;
                              }
                         }
                       s.ccd[id] = (int8_t) g.id;
                       if (is_hrc_s)
                         {
                            const int cid = g.id;
                            double u = d.pix_u[cid], v = d.pix_v[cid];
                            u += dx / d.u_pixel_size;
                            v += dy / d.v_pixel_size;
                            s.upix[id] = (float) u;
                            s.vpix[id] = (float) v;
                            s.ypix[id] = (float) (d.pix_x0[cid] + (u - d.pix_u0[cid]));
                            s.zpix[id] = (float) (d.pix_y0[cid] + (v - d.pix_v0[cid]));
                         }
                       else
                         {
                            s.ypix[id] = (float) (d.pix_x0[0] + dx / d.u_pixel_size);
                            s.zpix[id] = (float) (d.pix_y0[0] + dy / d.v_pixel_size);
                         }
                    }
               }
             p = rot_inv (d.matrix, p);
             x = rot_inv (d.matrix, x);
             x.x += d.offset[0];
             x.y += d.offset[1];
             x.z += d.offset[2];
          }
        s.x0[id] = x.x; s.x1[id] = x.y; s.x2[id] = x.z;
        s.p0[id] = p.x; s.p1[id] = p.y; s.p2[id] = p.z;
        s.flags[id] = flags;
        keep = (flags & BAD_PHOTON_MASK) == 0;
     }
   return keep;
}

} // namespace mxg
