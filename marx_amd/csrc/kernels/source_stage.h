// source_stage.h -- the per-photon work of photon creation shared by the
// source kernels (source.hip) and the one-launch batch pass (persist.hip).
#pragma once

#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

// every column of photon i set as a fresh photon of the given energy,
// direction, time and tag, with no dither offsets; a column that a later
// step writes is left out (with_t, with_p, with_dith false; with_entry
// false: the position, shell and flags, which the mirror's entry writes
// for every photon it is given as fresh)
__device__ __forceinline__ void init_photon (Soa s, uint32_t i, double energy, vec3 p, double t,
                                             uint32_t tag, bool with_t = true, bool with_p = true,
                                             bool with_dith = true, bool with_entry = true)
{
   s.energy[i] = energy;
   if (with_p)
     {
        s.p0[i] = p.x; s.p1[i] = p.y; s.p2[i] = p.z;
     }
   if (with_entry)
     {
        s.x0[i] = 0.0; s.x1[i] = 0.0; s.x2[i] = 0.0;
        s.flags[i] = 0u;
        s.shell[i] = 0;
     }
   if (with_t) s.t[i] = t;
   s.ypix[i] = 0.0f; s.zpix[i] = 0.0f; s.upix[i] = 0.0f; s.vpix[i] = 0.0f;
   if (with_dith)
     for (int q = 0; q < 6; q++) s.dith[q][i] = 0.0f;
   s.pi[i] = 0.0f;
   s.pha[i] = 0;
   s.ccd[i] = 0; s.region[i] = 0; s.order[i] = 0;
   for (int q = 0; q < 4; q++) s.sup[q][i] = 0;
   s.tag[i] = tag;
}

// the dither state of photon i: the three pointing offsets d (ra, dec,
// roll), the other three columns zero
__device__ __forceinline__ void store_dith (Soa s, uint32_t i, const float d[3])
{
   s.dith[0][i] = d[0]; s.dith[1][i] = d[1]; s.dith[2][i] = d[2];
   s.dith[3][i] = 0.0f; s.dith[4][i] = 0.0f; s.dith[5][i] = 0.0f;
}

// The direction of a photon of the configured source shape about the source
// direction p (p_normal: a unit vector normal to p; line_normal and
// line_alpha: the rotation of -x onto p): a point source's p; uniform over
// a disk (s-disk.c:95-106); a Gaussian or beta-model profile
// (s-gauss.c:121-137, s-beta.c:108-124); uniform along a line
// (s-line.c:89-98); a pixel of an image drawn from its cumulative
// distribution (fa + image_off), uniform within the pixel, the image centre
// at -x rotated onto the source direction (s-image.c:284-324).  The draws
// follow the energy in the source stream.
__device__ __forceinline__ vec3 extended_direction (const mxg_source_cfg &src, const float *__restrict__ fa,
                                                    vec3 p, vec3 p_normal, vec3 line_normal, double line_alpha,
                                                    Stream &st)
{
   if (src.type == MXG_SOURCE_DISK)
     {
        vec3 normal = rotate_unit_vector (p_normal, p, 2.0 * MXG_PI * uniform (st));
        double x0 = src.disk_theta_min / src.disk_theta;
        x0 = x0 * x0;
        double x1 = 1.0 - x0;
        double rnd = src.disk_theta * sqrt (x0 + x1 * uniform (st));
        return rotate_unit_vector (p, normal, rnd);
     }
   if ((src.type == MXG_SOURCE_GAUSS) || (src.type == MXG_SOURCE_BETA))
     {
        vec3 normal = rotate_unit_vector (p_normal, p, 2.0 * MXG_PI * uniform (st));
        double rnd;
        do rnd = uniform (st);
        while (rnd == 0.0);
        if (src.type == MXG_SOURCE_GAUSS)
          rnd = src.gauss_sigma * sqrt (-log (rnd));
        else
          rnd = src.beta_core * sqrt (pow (rnd, src.beta_xpon) - 1.0);
        return rotate_unit_vector (p, normal, rnd);
     }
   if (src.type == MXG_SOURCE_IMAGE)
     {
        const unsigned int ofs = search_f ((float) uniform (st), fa + src.image_off, src.image_n);
        double y = (double) (ofs / src.image_nx);
        double x = (double) (ofs % src.image_nx);
        y += -0.5 * src.image_ny + (uniform (st) - 0.5);
        x += -0.5 * src.image_nx + (uniform (st) - 0.5);
        y = y * src.image_yscale;
        x = x * src.image_xscale;
        double sx, cx, sy, cy;
        sincos (x, &sx, &cx);
        sincos (y, &sy, &cy);
        return rotate_unit_vector (v3 (-cy * cx, cy * sx, -sy), line_normal, line_alpha);
     }
   if (src.type == MXG_SOURCE_LINE)
     {
        double theta = src.line_theta * (-1.0 + 2.0 * uniform (st));
        double sn, cs;
        sincos (theta, &sn, &cs);
        const double sin_theta = -sn;
        vec3 q = v3 (-cs, sin_theta * src.line_cos_phi, sin_theta * src.line_sin_phi);
        return rotate_unit_vector (q, line_normal, line_alpha);
     }
   return p;
}

// the angle and unit axis that rotate a onto b (jdmath/src/vector.c:289-331)
__device__ __forceinline__ double rotation_axis (vec3 a, vec3 b, vec3 &c)
{
   double theta = dot (a, b);
   if (theta > 1.0) theta = 1.0;
   else if (theta < -1.0) theta = -1.0;
   theta = acos (theta);
   vec3 cc = cross (a, b);
   double len = length (cc);
   if (len > 0.0)
     {
        c = v3 (cc.x / len, cc.y / len, cc.z / len);
        return theta;
     }
   c = ((a.x != 0.0) || (a.y != 0.0)) ? v3 (-a.y, -a.x, 0.0) : v3 (-a.z, 0.0, 0.0);
   normalize (c);
   return theta;
}

// INTERNAL dither (dither.c:167-182, 551-581, 609-658) of the photon with
// tag `tag`, direction p and arrival time t: the Lissajous pointing offset
// at that time, returned in d as float (as the reference stores the dither
// state), Gaussian aspect blur, and the rotation of the direction, which is
// returned.
__device__ __forceinline__ vec3 internal_dither (const mxg_source_cfg &src, vec3 p, uint32_t tag, double t,
                                                 double start_time, unsigned long long seed,
                                                 uint32_t source_id, float d[3])
{
   Stream st;
   stream_init (st, seed, tag, MXG_STAGE_DITHER, source_id);
   double tt = (2.0 * MXG_PI) * (start_time + t);
   float fra = (float) (src.ra_amp * sin (tt / src.ra_period + src.ra_phase));
   float fdec = (float) (src.dec_amp * sin (tt / src.dec_period + src.dec_phase));
   float froll = (float) (src.nominal_roll + src.roll_amp * sin (tt / src.roll_period + src.roll_phase));
   d[0] = fra; d[1] = fdec; d[2] = froll;
   double delta_ra = src.aspect_blur * gaussian (st);
   double delta_dec = src.aspect_blur * gaussian (st);
   double ra = (double) fra + delta_ra, dec = (double) fdec + delta_dec, roll = froll;

   {
      double sr, cr;
      sincos_small (-roll, &sr, &cr);
      p = rotate_unit_vector1 (p, v3 (1, 0, 0), cr, sr);
   }
   double sin_ra, cos_ra, sin_dec, cos_dec;
   sincos_small (ra, &sin_ra, &cos_ra);
   sincos_small (dec, &sin_dec, &cos_dec);
   double cos_theta = cos_dec * cos_ra;
   vec3 nv = v3 (0, sin_dec, -cos_dec * sin_ra);
   double sin_theta = length (nv);
   if (sin_theta > 1e-20)
     {
        nv.x /= sin_theta;
        nv.y /= sin_theta;
        nv.z /= sin_theta;
        p = rotate_unit_vector1 (p, nv, cos_theta, sin_theta);
     }
   return p;
}

// internal_dither applied in place to photon i of the store at time t
__device__ __forceinline__ void apply_internal_dither (const mxg_source_cfg &src, Soa s, uint32_t i,
                                                       double t, double start_time,
                                                       unsigned long long seed, uint32_t source_id)
{
   float d[3];
   vec3 p = internal_dither (src, v3 (s.p0[i], s.p1[i], s.p2[i]), s.tag[i], t, start_time, seed, source_id, d);
   store_dith (s, i, d);
   s.p0[i] = p.x; s.p1[i] = p.y; s.p2[i] = p.z;
}

// ---- batched point sources ------------------------------------------------------
//
// Photon i (0 <= i < n) of a batch belongs to the source k with
// off[k] <= i < off[k + 1]: direction dir[3k..3k+2] (unit vector of travel),
// tag *tag_start + i, arrival time uniform over [0, exposure), energy from
// the configured spectrum, INTERNAL dither when configured.  The tag alone
// keys the photon's random streams, so tracing a source alone with
// tag_start = off[k] gives the same photons.  tag_start is read from device
// memory so that a captured launch sequence can be replayed for new tags.
// The kernel also writes the identity active list and its length.

__device__ __forceinline__ uint32_t source_of (const uint32_t *__restrict__ off, uint32_t nsrc, uint32_t i)
{
   uint32_t lo = 0, hi = nsrc;         // off[lo] <= i < off[hi]
   while (hi > lo + 1)
     {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
     }
   return lo;
}

// Photon i of the batch (tag t0 + i): k_batch_source's work for one photon.
// POINT: compiled for point sources only (the extended shapes' code, and
// its registers, left out); ENTRY: the mirror's entry follows with the
// photon as fresh (see init_photon), so its position, shell and flags are
// left to it.
template <bool POINT, bool ENTRY = false>
__device__ __forceinline__ void batch_source_one (const mxg_source_cfg &src, const double *__restrict__ da, Soa s,
                                                  uint32_t i, const uint32_t *__restrict__ off, uint32_t nsrc,
                                                  const double *__restrict__ dir, uint32_t t0, double exposure,
                                                  double start_time, unsigned long long seed)
{
   const uint32_t k = source_of (off, nsrc, i);
   const uint32_t tag = t0 + i;
   Stream st;
   stream_init (st, seed, tag, MXG_STAGE_SOURCE, 0u);
   double energy = draw_energy (src, da, st);
   vec3 p = v3 (dir[3 * k], dir[3 * k + 1], dir[3 * k + 2]);
   if (!POINT && (src.type != MXG_SOURCE_POINT))
     {
        // the shape about this source's direction: its normal as the
        // source set-up makes it, and the rotation of -x onto it
        vec3 nrm = v3 (-p.y / p.x, 1.0, 0.0);
        normalize (nrm);
        vec3 ln = v3 (0.0, 0.0, 0.0);
        double la = 0.0;
        if (src.type == MXG_SOURCE_LINE) la = rotation_axis (v3 (-1.0, 0.0, 0.0), p, ln);
        p = extended_direction (src, nullptr, p, nrm, ln, la, st);
     }
   double t = exposure * uniform (st);
   if (src.dither_mode == MXG_DITHER_INTERNAL)
     {
        float d[3];
        p = internal_dither (src, p, tag, t, start_time, seed, 0u, d);
        init_photon (s, i, energy, p, t, tag, true, true, false, !ENTRY);
        store_dith (s, i, d);
     }
   else
     init_photon (s, i, energy, p, t, tag, true, true, true, !ENTRY);
}

} // namespace mxg
