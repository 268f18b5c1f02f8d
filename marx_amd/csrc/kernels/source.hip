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
#include "source_stage.h"

namespace mxg {

constexpr int kTile = 2048;          // photons per scan tile
constexpr int kPer = kTile / kBlock; // photons per thread

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

// POINT: the instance for point sources (see batch_source_one)
template <bool POINT>
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
        batch_source_one<POINT> (src, da, s, i, off, nsrc, dir, t0, exposure, start_time, seed);
        idx[i] = i;
     }
   if ((blockIdx.x == 0) && (threadIdx.x == 0)) *count = n;
}

hipError_t launch_batch_source (const mxg_config *cfg, const double *da, Soa s, uint32_t n,
                                const uint32_t *off, uint32_t nsrc, const double *dir,
                                const uint32_t *tag_start, double exposure, double start_time,
                                unsigned long long seed, uint32_t *idx, uint32_t *count,
                                uint32_t grid, bool point, hipStream_t st)
{
   if (grid == 0) grid = 1;
   if (point)
     hipLaunchKernelGGL (k_batch_source<true>, dim3 (grid), dim3 (kBlock), 0, st, cfg, da, s, n, off, nsrc,
                         dir, tag_start, exposure, start_time, seed, idx, count);
   else
     hipLaunchKernelGGL (k_batch_source<false>, dim3 (grid), dim3 (kBlock), 0, st, cfg, da, s, n, off, nsrc,
                         dir, tag_start, exposure, start_time, seed, idx, count);
   return hipGetLastError ();
}

} // namespace mxg
