// hrma.hip -- HRMA Wolter-I reflection, one lane per photon.
//
// Per photon (the operations of marx/libsrc/hrma.c:1161-1341):
//   vignetting draw -> aperture sampling by cumulative shell area, radius and
//   shutter-checked azimuth (hrma.c:984-1050) -> paraboloid: conic
//   intersection (hrma.c:411-478), normal blur (:1054-1093), Fresnel
//   reflectivity draw (reflect.c:39-77), specular reflection, WFold scatter
//   (wfold.c:334-369) -> cap struts -> hyperboloid: the same -> post-collimator
//   struts.  Geometry is FP64.  Survivors are appended to the next active list
//   with one atomic per wavefront.
#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

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

__device__ __forceinline__ int conic_intersection (double a, double b, double c, vec3 &x0, vec3 p,
                                                   vec3 &normal, double xmin, double xmax)
{
   double t_plus = -x0.x / p.x;
   double x_y = x0.y + t_plus * p.y;
   double x_z = x0.z + t_plus * p.z;
   double alpha = a * p.x * p.x - 1.0;
   double beta = b * p.x - 2.0 * (p.y * x_y + p.z * x_z);
   double gamma = c - x_z * x_z - x_y * x_y;
   double t_minus;

   if (alpha == 0.0)
     {
        if (beta == 0.0) return -1;
        t_plus = t_minus = -gamma / beta;
     }
   else if (0 >= quadratic_root (alpha, beta, gamma, t_plus, t_minus))
     return -1;

   double x_plus = p.x * t_plus;
   double x_minus = p.x * t_minus;
   bool plus_ok = (x_plus >= xmin) && (x_plus < xmax);
   bool minus_ok = (x_minus >= xmin) && (x_minus < xmax);
   double t;
   if (plus_ok)
     t = (minus_ok && (x_minus > x_plus)) ? t_minus : t_plus;
   else if (minus_ok)
     t = t_minus;
   else
     return -1;

   x0.x = p.x * t;
   x0.y = x_y + p.y * t;
   x0.z = x_z + p.z * t;
   normal.x = (a - 1) * x0.x + 0.5 * b;
   normal.y = -x0.y;
   normal.z = -x0.z;
   normalize (normal);
   return 0;
}

__device__ __forceinline__ void blur_normal (vec3 &n, double blur, Stream &st)
{
   double n_y = n.y, n_z = n.z;
   double len = sqrt (n_y * n_y + n_z * n_z);
   vec3 perp = v3 (0.0, n_z / len, -n_y / len);
   double phi = (2.0 * MXG_PI) * uniform (st);
   perp = rotate_unit_vector (perp, n, phi);
   phi = blur * (1.0 / 3600.0 * MXG_PI / 180.0);
   phi = phi * gaussian (st);
   n = rotate_unit_vector (n, perp, phi);
}

// 0 reflected, -1 absorbed or scattered out, -2 missed the surface
__device__ __forceinline__ int reflect_surface (const mxg_hrma_cfg &h, const mxg_shell_cfg &S, int k,
                                                vec3 &x, vec3 &p, double energy, double beta,
                                                double delta, double corr,
                                                const float *__restrict__ fa,
                                                const double *__restrict__ da, Stream &st)
{
   vec3 normal;
   if (-1 == conic_intersection (S.conic_a[k], S.conic_b[k], S.conic_c[k], x, p, normal,
                                 S.conic_xmin[k], S.conic_xmax[k]))
     return -2;
   if (h.use_blur)
     blur_normal (normal, S.blur[k], st);
   double p_dot_n = dot (p, normal);
   if (h.ideal == 0)
     {
        double r = uniform (st);
        double rfl = reflectivity (fabs (p_dot_n), beta, delta);
        if (r >= rfl * corr)
          return -1;
     }
   p = ax1_bx2 (1.0, p, -2.0 * p_dot_n, normal);
   if (h.use_wfold == 0)
     return 0;
   double r = uniform (st);
   double dg = wfold_interp (S.wfold[k], fa, da, energy, -p_dot_n, r);
   dg *= S.scat_factor[k];
   if (dg > MXG_PI / 4) return -1;
   if (uniform (st) < 0.5) dg = -dg;
   p = rotate_unit_vector (p, cross (p, normal), dg);
   return 0;
}

__global__ __launch_bounds__ (kBlock)
void k_hrma (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
             const double *__restrict__ da, Soa s, ListIO io, double source_distance,
             unsigned long long seed, uint32_t source_id)
{
   const mxg_hrma_cfg &h = cfg->hrma;
   const uint32_t n = *io.n_in;
   const uint32_t stride = gridDim.x * blockDim.x;

   for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
     {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t id = 0;
        if (i < n)
          {
             id = io.idx_in[i];
             uint32_t flags = s.flags[id];
             if ((flags & BAD_PHOTON_MASK) == 0)
               {
                  Stream st;
                  stream_init (st, seed, s.tag[id], MXG_STAGE_MIRROR, source_id);
                  if ((h.ideal == 0) && (uniform (st) > h.vignetting))
                    {
                       s.flags[id] = flags | PHOTON_MIRROR_VBLOCKED;
                    }
                  else
                    {
                       vec3 x, p = v3 (s.p0[id], s.p1[id], s.p2[id]);
                       double energy = s.energy[id];

                       // aperture: shell by cumulative area, radius, open quadrant
                       int si;
                       for (;;)
                         {
                            double r = uniform (st);
                            si = -1;
                            for (int k = 0; k < MXG_NUM_SHELLS; k++)
                              if (r < h.shell[k].area_fraction) { si = k; break; }
                            if (si >= 0) break;
                         }
                       const mxg_shell_cfg &S = h.shell[si];
                       double radius = S.min_radius + (S.max_radius - S.min_radius) * uniform (st);
                       double theta;
                       unsigned int quad;
                       do
                         {
                            theta = uniform (st);
                            quad = (unsigned int) (4.0 * theta);
                         }
                       while (0 == (S.shutter_bitmap & (1u << quad)));
                       theta = (2.0 * MXG_PI) * (theta - 1.0 / 8.0);
                       double sn, cs;
                       sincos (theta, &sn, &cs);
                       x.z = radius * cs;
                       x.y = radius * sn;
                       x.x = S.front_position;
                       x.z -= S.to_osac[0][2];
                       x.y -= S.to_osac[0][1];
                       if (source_distance > 0.0)
                         {
                            p = ax1_bx2 (1.0, x, source_distance, p);
                            normalize (p);
                         }
                       if (h.use_struts && in_strut_shadow (h, x, p, h.pre_struts))
                         flags |= PHOTON_MIRROR_VBLOCKED;

                       x = add (x, ld3 (S.to_osac[0]));
                       x = m3_vmul (S.fwd[0], x);
                       p = m3_vmul (S.fwd[0], p);

                       double beta = 0.0, delta = 1.0, corr = 1.0;
                       if (h.opt_n != 0)
                         {
                            float ef = (float) energy;
                            beta = interp_f (ef, fa + h.opt_e_off, fa + h.opt_beta_off, h.opt_n);
                            delta = interp_f (ef, fa + h.opt_e_off, fa + h.opt_delta_off, h.opt_n);
                            if (h.use_scale)
                              {
                                 corr = interp_f (ef, fa + S.corr_e_off, fa + S.corr_f_off, S.corr_n);
                                 corr = sqrt (corr);
                              }
                         }

                       if (0 != reflect_surface (h, S, 0, x, p, energy, beta, delta, corr, fa, da, st))
                         flags |= PHOTON_UNREFLECTED;
                       else
                         {
                            p = m3_vmul (S.bwd[0], p);
                            x = m3_vmul (S.bwd[0], x);
                            x = sub (x, ld3 (S.to_osac[0]));
                            if (h.use_struts && in_strut_shadow (h, x, p, h.cap_struts))
                              flags |= PHOTON_MIRROR_VBLOCKED;
                            else
                              {
                                 x = add (x, ld3 (S.to_osac[1]));
                                 p = m3_vmul (S.fwd[1], p);
                                 x = m3_vmul (S.fwd[1], x);
                                 if (0 != reflect_surface (h, S, 1, x, p, energy, beta, delta, corr,
                                                           fa, da, st))
                                   flags |= PHOTON_UNREFLECTED;
                                 else
                                   {
                                      p = m3_vmul (S.bwd[1], p);
                                      x = m3_vmul (S.bwd[1], x);
                                      x = sub (x, ld3 (S.to_osac[1]));
                                      if (h.use_struts && in_strut_shadow (h, x, p, h.post_struts))
                                        flags |= PHOTON_MIRROR_VBLOCKED;
                                   }
                              }
                         }
                       s.x0[id] = x.x; s.x1[id] = x.y; s.x2[id] = x.z;
                       s.p0[id] = p.x; s.p1[id] = p.y; s.p2[id] = p.z;
                       s.shell[id] = (uint8_t) si;
                       s.flags[id] = flags;
                       keep = (flags & BAD_PHOTON_MASK) == 0;
                    }
               }
          }
        append_survivor (keep, id, io.idx_out, io.n_out);
     }
}

hipError_t launch_hrma (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                        ListIO io, uint32_t n_upper, double source_distance,
                        unsigned long long seed, uint32_t source_id, hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   uint32_t blocks = (n_upper + kBlock - 1) / kBlock;
   hipLaunchKernelGGL (k_hrma, dim3 (blocks), dim3 (kBlock), 0, st, cfg, fa, da, s, io,
                       source_distance, seed, source_id);
   return hipGetLastError ();
}

} // namespace mxg
