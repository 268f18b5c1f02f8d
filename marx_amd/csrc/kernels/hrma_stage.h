// hrma_stage.h -- the per-photon work of the HRMA stage (marx/libsrc/hrma.c),
// shared by the stage kernels (hrma.hip) and the one-launch batch pass
// (persist.hip).
//
// Per photon (the operations of marx/libsrc/hrma.c:1161-1341):
//   vignetting draw -> aperture sampling by cumulative shell area, radius and
//   shutter-checked azimuth (hrma.c:984-1050) -> paraboloid: conic
//   intersection (hrma.c:411-478), normal blur (:1054-1093), Fresnel
//   reflectivity draw (reflect.c:39-77), specular reflection, WFold scatter
//   (wfold.c:334-369) -> cap struts -> hyperboloid: the same -> post-collimator
//   struts.  Geometry is FP64.
//
// The mirror's work is split in two at the paraboloid hit (hrma_enter_one,
// hrma_reflect_one).  The stream position is the Philox block counter c2
// and the index of the next unused word of the current block, packed as
// (c2 << 3) | pos.  No Gaussian variate is pending at the split (none is
// drawn before it), and the buffered block is recomputed from c2 - 1 on
// resumption.  (A second split at the hyperboloid hit, each part on the
// photons still alive, measured 37 percent slower: the second kernel
// repeats the stream resumption, the table lookups and the LDS staging.)
#pragma once

#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

// The hit of a ray on the quadric x^2+y^2+z^2 = a x^2 + b x + c between
// xmin <= x < xmax (the larger valid root in x).  On success x0 becomes the
// hit; on a miss x0 is unchanged and -1 is returned.
__device__ __forceinline__ int conic_intersection (double a, double b, double c, vec3 &x0, vec3 p,
                                                   double xmin, double xmax)
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
   return 0;
}

// inward unit normal of the quadric at a point on it (hrma.c:470-476)
__device__ __forceinline__ vec3 conic_normal (double a, double b, vec3 x)
{
   vec3 n;
   n.x = (a - 1) * x.x + 0.5 * b;
   n.y = -x.y;
   n.z = -x.z;
   normalize (n);
   return n;
}

// The rotation of rotate_unit_vector, for the small angles of the normal
// blur and the WFold scatter (see sincos_small).
__device__ __forceinline__ vec3 rotate_small (vec3 p, vec3 n, double t)
{
   double sn, cs;
   sincos_small (t, &sn, &cs);
   return rotate_unit_vector1 (p, n, cs, sn);
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
   n = rotate_small (n, perp, phi);
}

// Reflection at a surface point x already on conic k of shell S: normal
// blur, Fresnel draw, specular reflection, WFold scatter (hrma.c:488-545).
// 0 reflected, -1 absorbed or scattered out.
__device__ __forceinline__ int reflect_at (const mxg_hrma_cfg &h, const mxg_shell_cfg &S, int k,
                                           vec3 x, vec3 &p, double energy, double beta,
                                           double delta, double corr,
                                           const float *__restrict__ fa,
                                           const double *__restrict__ da,
                                           const float *__restrict__ e_alphas,
                                           const float *__restrict__ ea_cnt, Stream &st)
{
   vec3 normal = conic_normal (S.conic_a[k], S.conic_b[k], x);
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
   double dg = wfold_interp (S.wfold[k], fa, da, e_alphas, energy, -p_dot_n, r, &S.ea_bucket[k], ea_cnt);
   dg *= S.scat_factor[k];
   if (dg > MXG_PI / 4) return -1;
   if (uniform (st) < 0.5) dg = -dg;
   p = rotate_small (p, cross (p, normal), dg);
   return 0;
}

// Copies the mirror configuration (a multiple of 4 bytes) into LDS.
__device__ __forceinline__ void stage_hrma_cfg (const mxg_config *__restrict__ cfg, unsigned char *lds)
{
   constexpr uint32_t cfg_words = sizeof (mxg_hrma_cfg) / 4;
   const uint32_t *src = (const uint32_t *) &cfg->hrma;
   uint32_t *dst = (uint32_t *) lds;
   for (uint32_t w = threadIdx.x; w < cfg_words; w += blockDim.x) dst[w] = src[w];
}

// Offsets (in floats) of the tables staged in LDS by k_hrma_reflect.
struct HrmaLds
{
   uint32_t opt;                  // energies, betas, deltas: 3 * opt_n
   uint32_t corr_e[MXG_NUM_SHELLS], corr_f[MXG_NUM_SHELLS];
   uint32_t ea[MXG_NUM_SHELLS][2];  // e_alpha of each WFold table (float32 on disk: exact)
   // the bucket counts of the searched tables (mxg_bucket)
   uint32_t bk_opt, bk_corr[MXG_NUM_SHELLS], bk_ea[MXG_NUM_SHELLS][2];
   uint32_t floats;               // staged floats after the configuration copy
};

// Dynamic LDS of the reflection: the configuration copy, a copy of the
// layout at kHrmaLayOff (the reflection reads a table's offset by the
// photon's shell from LDS, instead of holding every shell's offsets in
// scalar registers and selecting per lane), the tables from kHrmaTabOff.
constexpr size_t kHrmaLayOff = (sizeof (mxg_hrma_cfg) + 15) & ~(size_t) 15;
constexpr size_t kHrmaTabOff = (kHrmaLayOff + sizeof (HrmaLds) + 15) & ~(size_t) 15;

// The LDS layout of the reflection tables for configuration h (host side).
inline HrmaLds hrma_lds_layout (const mxg_hrma_cfg &h)
{
   HrmaLds lay;
   uint32_t f = 0;
   lay.opt = f; f += 3 * h.opt_n;
   lay.bk_opt = f; f += bucket_floats (h.opt_bucket);
   for (int sh = 0; sh < MXG_NUM_SHELLS; sh++)
     {
        const mxg_shell_cfg &S = h.shell[sh];
        lay.corr_e[sh] = f; f += S.corr_n;
        lay.corr_f[sh] = f; f += S.corr_n;
        lay.bk_corr[sh] = f; f += bucket_floats (S.corr_bucket);
        for (int k = 0; k < 2; k++)
          {
             lay.ea[sh][k] = f;
             f += S.wfold[k].n_arrays;
             lay.bk_ea[sh][k] = f;
             f += bucket_floats (S.ea_bucket[k]);
          }
     }
   lay.floats = f;
   return lay;
}

// bucket counts of table b from the float arena into LDS at dst
__device__ __forceinline__ void stage_bucket (const mxg_bucket &b, const float *__restrict__ fa, float *dst)
{
   if (!b.ok) return;
   for (uint32_t w = threadIdx.x; w <= b.nb; w += blockDim.x) dst[w] = fa[b.off + w];
}

// Stages the mirror configuration and the reflection tables (the
// binary-search targets of every photon) in LDS: the configuration copy at
// lds, lay at lds + kHrmaLayOff, the tables at ltab (lds + kHrmaTabOff) +
// the offsets of lay.  The caller synchronises.
__device__ __forceinline__ void stage_hrma_tables (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                                                   const double *__restrict__ da, const HrmaLds &lay,
                                                   unsigned char *lds, float *ltab)
{
   stage_hrma_cfg (cfg, lds);
   if (threadIdx.x == 0) *(HrmaLds *) (lds + kHrmaLayOff) = lay;
   {
      const mxg_hrma_cfg &gh = cfg->hrma;
      for (uint32_t w = threadIdx.x; w < gh.opt_n; w += blockDim.x)
        {
           ltab[lay.opt + w] = fa[gh.opt_e_off + w];
           ltab[lay.opt + gh.opt_n + w] = fa[gh.opt_beta_off + w];
           ltab[lay.opt + 2 * gh.opt_n + w] = fa[gh.opt_delta_off + w];
        }
      stage_bucket (gh.opt_bucket, fa, ltab + lay.bk_opt);
      for (int sh = 0; sh < MXG_NUM_SHELLS; sh++)
        {
           const mxg_shell_cfg &S = gh.shell[sh];
           for (uint32_t w = threadIdx.x; w < S.corr_n; w += blockDim.x)
             {
                ltab[lay.corr_e[sh] + w] = fa[S.corr_e_off + w];
                ltab[lay.corr_f[sh] + w] = fa[S.corr_f_off + w];
             }
           stage_bucket (S.corr_bucket, fa, ltab + lay.bk_corr[sh]);
           for (int k = 0; k < 2; k++)
             {
                for (uint32_t w = threadIdx.x; w < S.wfold[k].n_arrays; w += blockDim.x)
                  ltab[lay.ea[sh][k] + w] = (float) da[S.wfold[k].e_alphas_off + w];
                stage_bucket (S.ea_bucket[k], fa, ltab + lay.bk_ea[sh][k]);
             }
        }
   }
}

// The source-to-paraboloid part of the mirror for an unabsorbed photon with
// tag `tag` and direction p: vignetting, aperture, precollimator struts,
// frame change and the paraboloid intersection.  False when vignetting
// blocks the photon (nothing else is drawn); else true, with its position
// x, direction p, shell si and the mirror's marks added to flags; *hit: it
// hit the paraboloid, its random-stream position in *rng.
__device__ __forceinline__ bool hrma_enter_core (const mxg_hrma_cfg &h, uint32_t tag, double source_distance,
                                                 unsigned long long seed, uint32_t source_id, vec3 &x,
                                                 vec3 &p, int &si, uint32_t &flags, bool &hit, uint32_t &rng)
{
   Stream st;
   stream_init (st, seed, tag, MXG_STAGE_MIRROR, source_id);
   hit = false;
   if ((h.ideal == 0) && (uniform (st) > h.vignetting))
     return false;

   // aperture: shell by cumulative area, radius, open quadrant
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

   if (-1 == conic_intersection (S.conic_a[0], S.conic_b[0], S.conic_c[0], x, p,
                                 S.conic_xmin[0], S.conic_xmax[0]))
     flags |= PHOTON_UNREFLECTED;
   else
     {
        hit = true;
        rng = (st.c2 << 3) | (uint32_t) st.pos;
     }
   return true;
}

// k_hrma_enter's work for photon id of the store (hrma_enter_core).  True
// when the photon hit the paraboloid (its stream position is saved in
// rng_pos[id] for hrma_reflect_one).  FRESH: the photon is as the source
// made it without its position, shell and flags (flags zero, not read);
// they are written here, the position and shell zero when vignetting
// blocks it.
template <bool FRESH = false>
__device__ __forceinline__ bool hrma_enter_one (const mxg_hrma_cfg &h, Soa s, uint32_t id,
                                                uint32_t *__restrict__ rng_pos, double source_distance,
                                                unsigned long long seed, uint32_t source_id)
{
   bool keep = false;
   uint32_t flags = FRESH ? 0u : s.flags[id];
   if ((flags & BAD_PHOTON_MASK) == 0)
     {
        vec3 x, p = v3 (s.p0[id], s.p1[id], s.p2[id]);
        int si;
        uint32_t rng = 0;
        if (!hrma_enter_core (h, s.tag[id], source_distance, seed, source_id, x, p, si, flags, keep, rng))
          {
             s.flags[id] = flags | PHOTON_MIRROR_VBLOCKED;
             if (FRESH)
               {
                  s.x0[id] = 0.0; s.x1[id] = 0.0; s.x2[id] = 0.0;
                  s.shell[id] = 0;
               }
          }
        else
          {
             if (keep) rng_pos[id] = rng;
             s.x0[id] = x.x; s.x1[id] = x.y; s.x2[id] = x.z;
             s.p0[id] = p.x; s.p1[id] = p.y; s.p2[id] = p.z;
             s.shell[id] = (uint8_t) si;
             s.flags[id] = flags;
          }
     }
   return keep;
}

// The stream of photon id at the mirror's split, saved by hrma_enter_one.
__device__ __forceinline__ void resume_stream (Stream &st, const uint32_t *__restrict__ rng_pos, uint32_t id)
{
   const uint32_t w = rng_pos[id];
   st.c2 = w >> 3;
   st.pos = (int) (w & 7u);
   if (st.pos < 4)
     {
        // the block again, less the words already used
        philox (st.c0, st.c1, st.c2 - 1u, st.c3, st.k0, st.k1, st.b0, st.b1, st.b2, st.b3);
        for (int k = 0; k < st.pos; k++)
          {
             st.b0 = st.b1;
             st.b1 = st.b2;
             st.b2 = st.b3;
          }
     }
}

// The photon's optical constants (beta, delta) and the square root of its
// shell's area correction factor, interpolated at its energy from the
// tables in LDS (the same values at both surfaces).
__device__ __forceinline__ void mirror_constants (const mxg_hrma_cfg &h, const mxg_shell_cfg &S, int si,
                                                  const float *ltab, const HrmaLds &lay, double energy,
                                                  double &beta, double &delta, double &corr)
{
   beta = 0.0;
   delta = 1.0;
   corr = 1.0;
   if (h.opt_n != 0)
     {
        const float *opt_e = ltab + lay.opt;
        float ef = (float) energy, fb, fd;
        interp_f2b (ef, opt_e, opt_e + h.opt_n, opt_e + 2 * h.opt_n, h.opt_n, h.opt_bucket, ltab + lay.bk_opt,
                    fb, fd);
        beta = fb;
        delta = fd;
        if (h.use_scale)
          {
             corr = interp_fb (ef, ltab + lay.corr_e[si], ltab + lay.corr_f[si], S.corr_n, S.corr_bucket,
                               ltab + lay.bk_corr[si]);
             corr = sqrt (corr);
          }
     }
}

// k_hrma_reflect's work for photon id of the middle list: paraboloid
// reflection, cap struts, hyperboloid, post-collimator struts.  ltab: the
// tables staged in LDS at the offsets lay.  True when the photon leaves the
// mirror, its columns then also in o.  HANDOFF: such photons go on to the
// grating through o, and the grating stores their x, p and flags; only the
// photons the mirror drops have theirs stored here.
template <bool HANDOFF = false>
__device__ __forceinline__ bool hrma_reflect_one (const mxg_hrma_cfg &h, const float *ltab, const HrmaLds &lay,
                                                  const float *__restrict__ fa, const double *__restrict__ da,
                                                  Soa s, uint32_t id, const uint32_t *__restrict__ rng_pos,
                                                  unsigned long long seed, uint32_t source_id, PhotonOut &o)
{
   uint32_t flags = s.flags[id];
   const int si = s.shell[id];
   const mxg_shell_cfg &S = h.shell[si];
   Stream st;
   const uint32_t tag = s.tag[id];
   stream_init (st, seed, tag, MXG_STAGE_MIRROR, source_id);
   resume_stream (st, rng_pos, id);
   vec3 x = v3 (s.x0[id], s.x1[id], s.x2[id]);
   vec3 p = v3 (s.p0[id], s.p1[id], s.p2[id]);
   const double energy = s.energy[id];
   double beta, delta, corr;
   mirror_constants (h, S, si, ltab, lay, energy, beta, delta, corr);

   if (0 != reflect_at (h, S, 0, x, p, energy, beta, delta, corr, fa, da,
                        ltab + lay.ea[si][0], ltab + lay.bk_ea[si][0], st))
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
             if ((-1 == conic_intersection (S.conic_a[1], S.conic_b[1], S.conic_c[1], x, p,
                                            S.conic_xmin[1], S.conic_xmax[1]))
                 || (0 != reflect_at (h, S, 1, x, p, energy, beta, delta, corr, fa, da,
                                      ltab + lay.ea[si][1], ltab + lay.bk_ea[si][1], st)))
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
   const bool keep = (flags & BAD_PHOTON_MASK) == 0;
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
   o.shell = si;
   return keep;
}

} // namespace mxg
