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
// every photon's time is the same whatever the summation order.  A first
// pass (k_source_scan) draws the gaps only and scans them per tile; one
// block scans the tile sums; the second pass (k_source, one photon per
// thread) draws every photon again and writes each column once, its time
// and dithered direction included.
#include "source_stage.h"

namespace mxg {

constexpr int kTile = 2048;          // photons per scan tile
constexpr int kPer = kTile / kBlock; // photons per thread

__device__ __forceinline__ uint64_t gap_ticks (double e)
{
   return (uint64_t) (e * 4294967296.0 + 0.5);
}

// The source stream of photon `tag` (source.c:268-384): its energy, its
// direction and, for Poisson arrivals, the gap to the photon before in
// 2^-32 units of the mean gap (else its uniform arrival time t).  Both
// passes over the photons draw through this, so they see the same values.
// POINT: compiled for a point source only (its direction is the source's,
// with no draw); GAP: only the gap is wanted (the energy draw, one uniform
// for either spectrum, is skipped over without its table search).
template <bool POINT, bool GAP = false>
__device__ __forceinline__ void source_draws (const mxg_source_cfg &src, const float *__restrict__ fa,
                                              const double *__restrict__ da, uint32_t tag,
                                              unsigned long long seed, uint32_t source_id, bool poisson,
                                              double &energy, vec3 &p, uint64_t &gap, double &t)
{
   Stream st;
   stream_init (st, seed, tag, MXG_STAGE_SOURCE, source_id);
   if (GAP)
     {
        (void) uniform (st);
        energy = 0.0;
     }
   else
     energy = draw_energy (src, da, st);
   if (POINT)
     p = ld3 (src.p);
   else
     p = extended_direction (src, fa, ld3 (src.p), ld3 (src.p_normal), ld3 (src.line_normal), src.line_alpha, st);
   gap = 0;
   t = 0.0;
   if (poisson)
     gap = gap_ticks (expn (st));
   else
     t = src.uniform_tmax * uniform (st);
}

// Poisson arrivals, first pass: each photon's gap (the source stream's
// draws up to it, no photon stored), the inclusive scan of the gaps within
// its tile of kTile photons in ticks[i], the tile's total in tile_tot.
template <bool POINT>
__global__ __launch_bounds__ (kBlock)
void k_source_scan (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                    const double *__restrict__ da, uint32_t n, uint32_t tag_start,
                    unsigned long long seed, uint32_t source_id, uint64_t *__restrict__ ticks,
                    uint64_t *__restrict__ tile_tot)
{
   const mxg_source_cfg &src = cfg->src;
   __shared__ uint64_t sk[kTile];
   __shared__ uint64_t warp_tot[kBlock / 64];
   if (src.time_mode != MXG_TIME_POISSON)
     {
        if (threadIdx.x == 0) tile_tot[blockIdx.x] = 0;
        return;
     }
   const uint32_t base = blockIdx.x * kTile;
   for (int j = 0; j < kPer; j++)
     {
        const uint32_t li = j * kBlock + threadIdx.x;
        const uint32_t i = base + li;
        uint64_t k = 0;
        if (i < n)
          {
             double energy, t;
             vec3 p;
             source_draws<POINT, true> (src, fa, da, tag_start + i, seed, source_id, true, energy, p, k, t);
          }
        sk[li] = k;
     }
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

// exclusive scan of the tile totals in place: one block, each thread
// summing a run of consecutive totals
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__ (kScanThreads)
void k_scan_tiles (uint64_t *__restrict__ tile_tot, uint32_t ntiles)
{
   __shared__ uint64_t wave_tot[kScanThreads / 64];
   const uint32_t per = (ntiles + kScanThreads - 1) / kScanThreads;
   const uint32_t lo = threadIdx.x * per;
   const uint32_t hi = (lo + per < ntiles) ? lo + per : ntiles;
   uint64_t run = 0;
   for (uint32_t i = lo; i < hi; i++) run += tile_tot[i];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   uint64_t incl = run;
   for (int off = 1; off < 64; off <<= 1)
     {
        uint64_t v = __shfl_up (incl, off);
        if (lane >= off) incl += v;
     }
   if (lane == 63) wave_tot[wave] = incl;
   __syncthreads ();
   uint64_t acc = incl - run;
   for (int w = 0; w < wave; w++) acc += wave_tot[w];
   for (uint32_t i = lo; i < hi; i++)
     {
        const uint64_t v = tile_tot[i];
        tile_tot[i] = acc;
        acc += v;
     }
}

// A photon as the source makes it: energy, time, direction (dithered when
// configured, with the offsets d) and whether it is active.
struct NewPhoton
{
   double energy, t;
   vec3 p;
   float d[3];
   bool dithered, active;
};

// Photon i: its draws; with Poisson arrivals its time from the first pass's
// scan and its tile's offset tile_off, whether it is active (the first
// photon, or its predecessor arrives before the exposure) and, for the first
// photon at or past the exposure, *first_over (unique: the times do not
// decrease); the INTERNAL dither.  POINT: as in source_draws.
template <bool POINT>
__device__ __forceinline__ NewPhoton source_photon (const mxg_source_cfg &src, const float *__restrict__ fa,
                                                    const double *__restrict__ da, uint32_t i, uint32_t tag,
                                                    double exposure_time, double start_time,
                                                    unsigned long long seed, uint32_t source_id,
                                                    const uint64_t *__restrict__ ticks,
                                                    const uint64_t *__restrict__ tile_off,
                                                    uint32_t *__restrict__ first_over)
{
   NewPhoton ph;
   const bool poisson = (src.time_mode == MXG_TIME_POISSON);
   uint64_t k;
   source_draws<POINT> (src, fa, da, tag, seed, source_id, poisson, ph.energy, ph.p, k, ph.t);
   ph.active = true;
   if (poisson)
     {
        const uint64_t toff = tile_off[i / kTile];
        const double mt = src.mean_time;
        ph.t = mt * ((double) (ticks[i] + toff) * (1.0 / 4294967296.0));
        if ((exposure_time > 0.0) && (i > 0))
          {
             const uint64_t Sp = ((i % kTile) == 0) ? toff : ticks[i - 1] + toff;
             const double tp = mt * ((double) Sp * (1.0 / 4294967296.0));
             ph.active = !(exposure_time <= tp);
          }
        if (ph.active && (exposure_time > 0.0) && (exposure_time <= ph.t)) *first_over = i;
     }
   ph.dithered = (src.dither_mode == MXG_DITHER_INTERNAL);
   if (ph.dithered)
     ph.p = internal_dither (src, ph.p, tag, ph.t, start_time, seed, source_id, ph.d);
   return ph;
}

// Photon i, one per thread (source_photon): every column written once, and
// idx[i] = i for an active photon.  ENTRY: the mirror's entry follows with
// the photons as fresh and writes their position, shell and flags, so
// those are left out (a photon past the exposure cut keeps stale ones: it
// is not among the photons created).
template <bool POINT, bool ENTRY>
#if MXG_SOURCE_WAVES > 0
__global__ __launch_bounds__ (kBlock, MXG_SOURCE_WAVES)
#else
__global__ __launch_bounds__ (kBlock)
#endif
void k_source (const mxg_config *__restrict__ cfg, const float *__restrict__ fa, const double *__restrict__ da, Soa s,
               uint32_t n, uint32_t tag_start, double exposure_time, double start_time,
               unsigned long long seed, uint32_t source_id, const uint64_t *__restrict__ ticks,
               const uint64_t *__restrict__ tile_off, uint32_t *__restrict__ first_over,
               uint32_t *__restrict__ idx)
{
   const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
   if (i >= n) return;
   const uint32_t tag = tag_start + i;
   const NewPhoton ph = source_photon<POINT> (cfg->src, fa, da, i, tag, exposure_time, start_time, seed,
                                              source_id, ticks, tile_off, first_over);
   init_photon (s, i, ph.energy, ph.p, ph.t, tag, true, true, !ph.dithered, !ENTRY);
   if (ph.dithered) store_dith (s, i, ph.d);
   if (ph.active) idx[i] = i;
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

// the length of the active list: up to and including the first photon at
// or past the exposure, else all (also to *count_copy when non-null)
__global__ void k_source_count (uint32_t n, const uint32_t *__restrict__ first_over, uint32_t *__restrict__ count,
                                uint32_t *__restrict__ count_copy)
{
   const uint32_t f = *first_over;
   const uint32_t m = (f < n) ? f + 1 : n;
   *count = m;
   if (count_copy) *count_copy = m;
}

__global__ void k_reset (uint32_t *p, uint32_t v) { *p = v; }

hipError_t launch_source (const mxg_config *cfg, const float *fa, const double *da, Soa s, uint32_t n,
                          uint32_t tag_start, double exposure_time, double start_time,
                          unsigned long long seed, uint32_t source_id, uint64_t *ticks, uint64_t *tile_tot,
                          uint32_t *first_over, uint32_t *count, uint32_t *count_copy, uint32_t *idx, bool point,
                          bool entry_follows, hipStream_t st)
{
   if (n == 0)
     {
        hipLaunchKernelGGL (k_reset, dim3 (1), dim3 (1), 0, st, count, 0u);
        if (count_copy) hipLaunchKernelGGL (k_reset, dim3 (1), dim3 (1), 0, st, count_copy, 0u);
        return hipGetLastError ();
     }
   const uint32_t ntiles = (n + kTile - 1) / kTile;
   const uint32_t blocks = (n + kBlock - 1) / kBlock;
   hipLaunchKernelGGL (k_reset, dim3 (1), dim3 (1), 0, st, first_over, 0xFFFFFFFFu);
   if (point)
     hipLaunchKernelGGL (k_source_scan<true>, dim3 (ntiles), dim3 (kBlock), 0, st, cfg, fa, da, n, tag_start,
                         seed, source_id, ticks, tile_tot);
   else
     hipLaunchKernelGGL (k_source_scan<false>, dim3 (ntiles), dim3 (kBlock), 0, st, cfg, fa, da, n, tag_start,
                         seed, source_id, ticks, tile_tot);
   hipLaunchKernelGGL (k_scan_tiles, dim3 (1), dim3 (kScanThreads), 0, st, tile_tot, ntiles);
   void (*kern) (const mxg_config *, const float *, const double *, Soa, uint32_t, uint32_t, double, double,
                 unsigned long long, uint32_t, const uint64_t *, const uint64_t *, uint32_t *, uint32_t *)
     = point ? (entry_follows ? k_source<true, true> : k_source<true, false>)
             : (entry_follows ? k_source<false, true> : k_source<false, false>);
   hipLaunchKernelGGL (kern, dim3 (blocks), dim3 (kBlock), 0, st, cfg, fa, da, s, n, tag_start, exposure_time,
                       start_time, seed, source_id, ticks, tile_tot, first_over, idx);
   hipLaunchKernelGGL (k_source_count, dim3 (1), dim3 (1), 0, st, n, first_over, count, count_copy);
   return hipGetLastError ();
}

// POINT: the instance for point sources (see batch_source_one)
template <bool POINT, bool ENTRY>
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
        batch_source_one<POINT, ENTRY> (src, da, s, i, off, nsrc, dir, t0, exposure, start_time, seed);
        idx[i] = i;
     }
   if ((blockIdx.x == 0) && (threadIdx.x == 0)) *count = n;
}

hipError_t launch_batch_source (const mxg_config *cfg, const double *da, Soa s, uint32_t n,
                                const uint32_t *off, uint32_t nsrc, const double *dir,
                                const uint32_t *tag_start, double exposure, double start_time,
                                unsigned long long seed, uint32_t *idx, uint32_t *count,
                                uint32_t grid, bool point, bool entry_follows, hipStream_t st)
{
   if (grid == 0) grid = 1;
   void (*kern) (const mxg_config *, const double *, Soa, uint32_t, const uint32_t *, uint32_t, const double *,
                 const uint32_t *, double, double, unsigned long long, uint32_t *, uint32_t *)
     = point ? (entry_follows ? k_batch_source<true, true> : k_batch_source<true, false>)
             : (entry_follows ? k_batch_source<false, true> : k_batch_source<false, false>);
   hipLaunchKernelGGL (kern, dim3 (grid), dim3 (kBlock), 0, st, cfg, da, s, n, off, nsrc, dir, tag_start,
                       exposure, start_time, seed, idx, count);
   return hipGetLastError ();
}

} // namespace mxg
