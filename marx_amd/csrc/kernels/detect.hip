// detect.hip -- focal-plane detection, one lane per photon.
//
// ACIS-S / ACIS-I (marx/libsrc/acis-s.c:138-248, acis-i.c:56-167): detector
//   transform, first chip whose plane hit lies inside it (detector.c:56-168),
//   pixel = offset / pixel size, QE draw against the CALDB QE curve
//   (contamination 1), frame-transfer streak.
// HRC-S / HRC-I (hrc-s.c:136-312, hrc-i.c:66-185): HESF plates first
//   (drake.c:268-369), MCP QE draw, UV/ion-shield region and filter draw,
//   pulse height, MCP position blur (hrcblur.c:239-297), u/v and chip pixel
//   (hrc_s_geom.c:344-394, hrc_i_geom.c:146-156).
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
   return interp_f ((float) energy, fa + q.e_off, fa + q.eff_off, q.n);
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

__global__ __launch_bounds__ (kBlock)
void k_detect (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
               const double *__restrict__ da, Soa s, ListIO io, double tstart,
               unsigned long long seed, uint32_t source_id)
{
   const mxg_detector_cfg &d = cfg->det;
   const bool is_acis = (d.type == MARX_DETECTOR_ACIS_S) || (d.type == MARX_DETECTOR_ACIS_I);
   const bool is_hrc_s = (d.type == MARX_DETECTOR_HRC_S);
   __shared__ uint32_t seg_wc[2 * (kBlock / 64)];
   const BlockRange range = block_range (*io.n_in, blockIdx.x, gridDim.x);
   SegWriter survivors (io, range, seg_wc);
   (void) da;

   for (uint32_t base = range.lo; base < range.hi; base += kBlock)
     {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t id = 0;
        if (i < range.hi)
          {
             id = io.idx_in[i];
             uint32_t flags = s.flags[id];
             if ((flags & BAD_PHOTON_MASK) == 0)
               {
                  Stream st;
                  stream_init (st, seed, s.tag[id], MXG_STAGE_DETECT, source_id);
                  vec3 x = v3 (s.x0[id], s.x1[id], s.x2[id]);
                  vec3 p = v3 (s.p0[id], s.p1[id], s.p2[id]);
                  const double energy = s.energy[id];
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
                       else if (is_acis)
                         {
                            const mxg_chip &g = d.chip[k];
                            s.ccd[id] = (int8_t) g.id;
                            float ypix = (float) (dx / g.x_pixel_size);
                            float zpix = (float) (dy / g.y_pixel_size);
                            bool detected = true;
                            if (d.ideal == 0)
                              {
                                 double r = uniform (st);
                                 double qe = 1.0;
                                 if (g.qe_n != 0)
                                   qe = interp_f ((float) energy, fa + g.qe_e_off, fa + g.qe_off, g.qe_n);
                                 if (r >= qe * 1.0 * 1.0)
                                   {
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
          }
        survivors.add (keep, id);
     }
   survivors.finish (io);
}

hipError_t launch_detect (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                          ListIO io, uint32_t n_upper, double tstart, unsigned long long seed,
                          uint32_t source_id, hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   hipLaunchKernelGGL (k_detect, dim3 (io.grid), dim3 (kBlock), 0, st, cfg, fa, da, s, io, tstart,
                       seed, source_id);
   hipError_t e = hipGetLastError ();
   if (e == hipSuccess) e = launch_compact (io, st);
   return e;
}

} // namespace mxg
