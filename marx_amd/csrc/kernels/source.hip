// source.hip -- photon creation on the GPU: energies, directions, arrival
// times, exposure cut and the INTERNAL dither model.
//
// Per photon (marx/libsrc/source.c:268-384, s-point.c:59-83, s-disk.c:63-110,
// s-gauss.c:78-143, s-beta.c:81-130, s-line.c:60-101,
// spectrum.c:138-189, prob.c:46-58, dither.c:167-182, 551-581, 609-658):
//   energy from a FLAT or tabulated (inverse-CDF) spectrum; direction of a
//   point source, uniform over a disk, Gaussian or beta-model profile about
//   the source, or uniform along a line; exponential gap to the previous
//   photon (Poisson arrivals) or a uniform time; then Lissajous dither with
//   Gaussian aspect blur applied to the direction.
//
// The arrival time is a running sum of the gaps, which the reference adds up
// serially in FP64.  Here each gap is quantised to 2^-32 of the mean gap and
// summed as 64-bit integers -- exact and associative -- in a tiled scan, so
// every photon's time is the same whatever the summation order.
#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

constexpr int kTile = 2048;          // photons per scan tile
constexpr int kPer = kTile / kBlock; // photons per thread

// every column of photon i set as a fresh photon of the given energy,
// direction, time and tag; the time and the dither state are left to a
// later pass that writes them (with_t, with_dith false)
__device__ __forceinline__ void init_photon (Soa s, uint32_t i, double energy, vec3 p, double t,
                                             uint32_t tag, bool with_t = true, bool with_dith = true)
{
   s.energy[i] = energy;
   s.p0[i] = p.x; s.p1[i] = p.y; s.p2[i] = p.z;
   s.x0[i] = 0.0; s.x1[i] = 0.0; s.x2[i] = 0.0;
   if (with_t) s.t[i] = t;
   s.flags[i] = 0u;
   s.ypix[i] = 0.0f; s.zpix[i] = 0.0f; s.upix[i] = 0.0f; s.vpix[i] = 0.0f;
   if (with_dith)
     for (int q = 0; q < 6; q++) s.dith[q][i] = 0.0f;
   s.pi[i] = 0.0f;
   s.pha[i] = 0;
   s.shell[i] = 0;
   s.ccd[i] = 0; s.region[i] = 0; s.order[i] = 0;
   for (int q = 0; q < 4; q++) s.sup[q][i] = 0;
   s.tag[i] = tag;
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

__device__ __forceinline__ uint64_t gap_ticks (double e)
{
   return (uint64_t) (e * 4294967296.0 + 0.5);
}

#if MXG_SOURCE_WAVES > 0
__global__ __launch_bounds__ (kBlock, MXG_SOURCE_WAVES)
#else
__global__ __launch_bounds__ (kBlock)
#endif
void k_source (const mxg_config *__restrict__ cfg, const float *__restrict__ fa, const double *__restrict__ da, Soa s,
               uint32_t n, uint32_t tag_start, unsigned long long seed, uint32_t source_id,
               uint64_t *__restrict__ ticks, uint64_t *__restrict__ tile_tot)
{
   const mxg_source_cfg &src = cfg->src;
   __shared__ uint64_t sk[kTile];
   __shared__ uint64_t warp_tot[kBlock / 64];
   const uint32_t base = blockIdx.x * kTile;
   const bool poisson = (src.time_mode == MXG_TIME_POISSON);

   for (int j = 0; j < kPer; j++)
     {
        const uint32_t li = j * kBlock + threadIdx.x;
        const uint32_t i = base + li;
        uint64_t k = 0;
        if (i < n)
          {
             const uint32_t tag = tag_start + i;
             Stream st;
             stream_init (st, seed, tag, MXG_STAGE_SOURCE, source_id);
             double energy = draw_energy (src, da, st);
             vec3 p = extended_direction (src, fa, ld3 (src.p), ld3 (src.p_normal), ld3 (src.line_normal),
                                          src.line_alpha, st);
             double t = 0.0;
             if (poisson)
               k = gap_ticks (expn (st));
             else
               t = src.uniform_tmax * uniform (st);

             // k_source_finish writes Poisson times and the INTERNAL dither
             init_photon (s, i, energy, p, t, tag, !poisson, src.dither_mode != MXG_DITHER_INTERNAL);
          }
        sk[li] = k;
     }
   if (!poisson) return;
   __syncthreads ();

   // inclusive scan of the tile: 8 serial per thread, then across threads
   uint64_t run = 0;
   const int t0 = threadIdx.x * kPer;
   for (int j = 0; j < kPer; j++)
     {
        run += sk[t0 + j];
        sk[t0 + j] = run;
     }
   // exclusive scan of the per-thread totals: within the wave, then waves
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   uint64_t incl = run;
   for (int off = 1; off < 64; off <<= 1)
     {
        uint64_t v = __shfl_up (incl, off);
        if (lane >= off) incl += v;
     }
   if (lane == 63) warp_tot[wave] = incl;
   __syncthreads ();
   uint64_t wave_off = 0;
   for (int w = 0; w < wave; w++) wave_off += warp_tot[w];
   const uint64_t excl = wave_off + incl - run;
   for (int j = 0; j < kPer; j++) sk[t0 + j] += excl;
   __syncthreads ();
   for (int j = 0; j < kPer; j++)
     {
        const uint32_t li = j * kBlock + threadIdx.x;
        const uint32_t i = base + li;
        if (i < n) ticks[i] = sk[li];
     }
   if (threadIdx.x == kBlock - 1)
     tile_tot[blockIdx.x] = sk[kTile - 1];
}

// exclusive scan of the tile totals (one block, sequential chunks of 256)
__global__ __launch_bounds__ (kBlock)
void k_scan_tiles (uint64_t *__restrict__ tile_tot, uint32_t ntiles)
{
   __shared__ uint64_t buf[kBlock];
   uint64_t carry = 0;
   for (uint32_t c0 = 0; c0 < ntiles; c0 += kBlock)
     {
        const uint32_t i = c0 + threadIdx.x;
        uint64_t v = (i < ntiles) ? tile_tot[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads ();
        for (int off = 1; off < kBlock; off <<= 1)
          {
             uint64_t add = (threadIdx.x >= (unsigned) off) ? buf[threadIdx.x - off] : 0;
             __syncthreads ();
             buf[threadIdx.x] += add;
             __syncthreads ();
          }
        const uint64_t incl = buf[threadIdx.x];
        if (i < ntiles) tile_tot[i] = carry + incl - v;
        __syncthreads ();
        carry += buf[kBlock - 1];
        __syncthreads ();
     }
}

// INTERNAL dither of photon i (dither.c:167-182, 551-581, 609-658): the
// Lissajous pointing offset at its arrival time (stored as float, as the
// reference stores the dither state), Gaussian aspect blur, and the rotation
// of its direction.
__device__ __forceinline__ void apply_internal_dither (const mxg_source_cfg &src, Soa s, uint32_t i,
                                                       double t, double start_time,
                                                       unsigned long long seed, uint32_t source_id)
{
   Stream st;
   stream_init (st, seed, s.tag[i], MXG_STAGE_DITHER, source_id);
   double tt = (2.0 * MXG_PI) * (start_time + t);
   float fra = (float) (src.ra_amp * sin (tt / src.ra_period + src.ra_phase));
   float fdec = (float) (src.dec_amp * sin (tt / src.dec_period + src.dec_phase));
   float froll = (float) (src.nominal_roll + src.roll_amp * sin (tt / src.roll_period + src.roll_phase));
   s.dith[0][i] = fra; s.dith[1][i] = fdec; s.dith[2][i] = froll;
   s.dith[3][i] = 0.0f; s.dith[4][i] = 0.0f; s.dith[5][i] = 0.0f;
   double delta_ra = src.aspect_blur * gaussian (st);
   double delta_dec = src.aspect_blur * gaussian (st);
   double ra = (double) fra + delta_ra, dec = (double) fdec + delta_dec, roll = froll;

   vec3 p = v3 (s.p0[i], s.p1[i], s.p2[i]);
   p = rotate_unit_vector (p, v3 (1, 0, 0), -roll);
   double sin_ra, cos_ra, sin_dec, cos_dec;
   sincos (ra, &sin_ra, &cos_ra);
   sincos (dec, &sin_dec, &cos_dec);
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
   s.p0[i] = p.x; s.p1[i] = p.y; s.p2[i] = p.z;
}

// arrival times, first photon at or past the exposure, dither
__global__ __launch_bounds__ (kBlock)
void k_source_finish (const mxg_config *__restrict__ cfg, Soa s, uint32_t n,
                      const uint64_t *__restrict__ ticks, const uint64_t *__restrict__ tile_off,
                      double exposure_time, double start_time, unsigned long long seed,
                      uint32_t source_id, uint32_t *__restrict__ first_over)
{
   const mxg_source_cfg &src = cfg->src;
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   double t;
   if (src.time_mode == MXG_TIME_POISSON)
     {
        const uint32_t tile = i / kTile;
        const uint64_t S = ticks[i] + tile_off[tile];
        const double mt = src.mean_time;
        t = mt * ((double) S * (1.0 / 4294967296.0));
        s.t[i] = t;
        if ((exposure_time > 0.0) && (exposure_time <= t))
          {
             bool first = (i == 0);
             if (!first)
               {
                  const uint64_t Sp = ((i % kTile) == 0) ? tile_off[tile] : ticks[i - 1] + tile_off[tile];
                  const double tp = mt * ((double) Sp * (1.0 / 4294967296.0));
                  first = !(exposure_time <= tp);
               }
             if (first) *first_over = i;
          }
     }
   else
     t = s.t[i];

   if (src.dither_mode == MXG_DITHER_INTERNAL)
     apply_internal_dither (src, s, i, t, start_time, seed, source_id);
}

// the configured dither applied to photons 0..n-1 of the store, each at its
// own arrival time (dither.c:630-658 on photons read from a ray file)
__global__ __launch_bounds__ (kBlock)
void k_dither_store (const mxg_config *__restrict__ cfg, Soa s, uint32_t n, double start_time,
                     unsigned long long seed)
{
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   if (cfg->src.dither_mode == MXG_DITHER_INTERNAL)
     apply_internal_dither (cfg->src, s, i, s.t[i], start_time, seed, 0u);
}

hipError_t launch_dither_store (const mxg_config *cfg, Soa s, uint32_t n, double start_time,
                                unsigned long long seed, hipStream_t st)
{
   if (n == 0) return hipSuccess;
   hipLaunchKernelGGL (k_dither_store, dim3 ((n + kBlock - 1) / kBlock), dim3 (kBlock), 0, st, cfg, s, n,
                       start_time, seed);
   return hipGetLastError ();
}

__global__ void k_source_count (uint32_t n, const uint32_t *__restrict__ first_over,
                                uint32_t *__restrict__ count, uint32_t *__restrict__ idx)
{
   // the active list of a fresh chunk is the identity
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   uint32_t f = *first_over;
   uint32_t m = (f < n) ? f + 1 : n;
   if (i < m) idx[i] = i;
   if (i == 0) *count = m;
}

__global__ void k_reset (uint32_t *p, uint32_t v) { *p = v; }

hipError_t launch_source (const mxg_config *cfg, const float *fa, const double *da, Soa s, uint32_t n,
                          uint32_t tag_start, double exposure_time, double start_time,
                          unsigned long long seed, uint32_t source_id, uint64_t *ticks,
                          uint64_t *tile_tot, uint32_t *first_over, uint32_t *count,
                          uint32_t *idx, hipStream_t st)
{
   if (n == 0)
     {
        hipLaunchKernelGGL (k_reset, dim3 (1), dim3 (1), 0, st, count, 0u);
        return hipGetLastError ();
     }
   const uint32_t ntiles = (n + kTile - 1) / kTile;
   hipLaunchKernelGGL (k_reset, dim3 (1), dim3 (1), 0, st, first_over, 0xFFFFFFFFu);
   hipLaunchKernelGGL (k_source, dim3 (ntiles), dim3 (kBlock), 0, st, cfg, fa, da, s, n, tag_start,
                       seed, source_id, ticks, tile_tot);
   hipLaunchKernelGGL (k_scan_tiles, dim3 (1), dim3 (kBlock), 0, st, tile_tot, ntiles);
   hipLaunchKernelGGL (k_source_finish, dim3 ((n + kBlock - 1) / kBlock), dim3 (kBlock), 0, st,
                       cfg, s, n, ticks, tile_tot, exposure_time, start_time, seed, source_id,
                       first_over);
   hipLaunchKernelGGL (k_source_count, dim3 ((n + kBlock - 1) / kBlock), dim3 (kBlock), 0, st, n,
                       first_over, count, idx);
   return hipGetLastError ();
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

#if MXG_BATCH_SOURCE_WAVES > 0
__global__ __launch_bounds__ (kBlock, MXG_BATCH_SOURCE_WAVES)
#else
__global__ __launch_bounds__ (kBlock)
#endif
void k_batch_source (const mxg_config *__restrict__ cfg, const double *__restrict__ da, Soa s,
                     uint32_t n, const uint32_t *__restrict__ off, uint32_t nsrc,
                     const double *__restrict__ dir, const uint32_t *__restrict__ tag_start,
                     double exposure, double start_time, unsigned long long seed,
                     uint32_t *__restrict__ idx, uint32_t *__restrict__ count)
{
   const mxg_source_cfg &src = cfg->src;
   const uint32_t t0 = *tag_start;
   const uint32_t stride = gridDim.x * blockDim.x;
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
     {
        const uint32_t k = source_of (off, nsrc, i);
        const uint32_t tag = t0 + i;
        Stream st;
        stream_init (st, seed, tag, MXG_STAGE_SOURCE, 0u);
        double energy = draw_energy (src, da, st);
        vec3 p = v3 (dir[3 * k], dir[3 * k + 1], dir[3 * k + 2]);
        if (src.type != MXG_SOURCE_POINT)
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
        init_photon (s, i, energy, p, t, tag);
        if (src.dither_mode == MXG_DITHER_INTERNAL)
          apply_internal_dither (src, s, i, t, start_time, seed, 0u);
        idx[i] = i;
     }
   if ((blockIdx.x == 0) && (threadIdx.x == 0)) *count = n;
}

hipError_t launch_batch_source (const mxg_config *cfg, const double *da, Soa s, uint32_t n,
                                const uint32_t *off, uint32_t nsrc, const double *dir,
                                const uint32_t *tag_start, double exposure, double start_time,
                                unsigned long long seed, uint32_t *idx, uint32_t *count,
                                uint32_t grid, hipStream_t st)
{
   if (grid == 0) grid = 1;
   hipLaunchKernelGGL (k_batch_source, dim3 (grid), dim3 (kBlock), 0, st, cfg, da, s, n, off, nsrc,
                       dir, tag_start, exposure, start_time, seed, idx, count);
   return hipGetLastError ();
}

} // namespace mxg
