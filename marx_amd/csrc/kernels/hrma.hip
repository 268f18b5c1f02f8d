// hrma.hip -- the HRMA stage kernels, one lane per photon (the per-photon
// work is in hrma_stage.h).
//
// The mirror runs as two kernels so that the reflection work runs on full
// wavefronts: about half of the photons sampled over the aperture annulus
// miss the paraboloid, and in a single kernel their lanes would idle through
// both reflections.
//
//   k_hrma_enter:   vignetting, aperture, precollimator struts, frame change
//                   and the paraboloid intersection, for every active photon.
//                   Photons that hit go to the middle list, and their
//                   random-stream position is saved.
//   k_hrma_reflect: paraboloid reflection, cap struts, hyperboloid
//                   intersection and reflection, post-collimator struts, for
//                   the middle list.
//
// Survivors go to the next active list in input order (SegWriter +
// k_compact).
#include "hrma_stage.h"

namespace mxg {

// FRESH: the photons come straight from the source (hrma_enter_one<true>)
template <bool FRESH>
#if MXG_ENTER_WAVES > 0
__global__ __launch_bounds__ (kBlock, MXG_ENTER_WAVES)
#else
__global__ __launch_bounds__ (kBlock)
#endif
void k_hrma_enter (const mxg_config *__restrict__ cfg, Soa s, ListIO io,
                   uint32_t *__restrict__ rng_pos, double source_distance,
                   unsigned long long seed, uint32_t source_id)
{
   __shared__ __align__ (16) unsigned char lds[sizeof (mxg_hrma_cfg)];
   stage_hrma_cfg (cfg, lds);
   __syncthreads ();
   const mxg_hrma_cfg &h = *(const mxg_hrma_cfg *) lds;
   __shared__ uint32_t seg_wc[2 * (kBlock / 64)];
   const BlockRange range = block_range (*io.n_in, blockIdx.x, gridDim.x);
   SegWriter survivors (io, range, seg_wc);

   for (uint32_t base = range.lo; base < range.hi; base += kBlock)
     {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t id = 0;
        if (i < range.hi)
          {
             id = io.idx_in[i];
             keep = hrma_enter_one<FRESH> (h, s, id, rng_pos, source_distance, seed, source_id);
          }
        survivors.add (keep, id);
     }
   survivors.finish (io);
}

// The block stages the mirror configuration and its small, hot tables (the
// binary-search targets of every photon) in LDS once, then loops over its
// range of the middle list.
// WAVES: minimum waves per SIMD the register allocation must allow; RELAY:
// the survivors' columns go to relay at their segment slots, for the
// grating, which stores their x, p and flags.
template <int WAVES, bool RELAY>
__global__ __launch_bounds__ (kBlock, WAVES)
void k_hrma_reflect (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                     const double *__restrict__ da, Soa s, ListIO io,
                     const uint32_t *__restrict__ rng_pos, unsigned long long seed,
                     uint32_t source_id, HrmaLds lay, Relay relay)
{
   extern __shared__ __align__ (16) unsigned char lds[];
   float *ltab = (float *) (lds + kHrmaTabOff);
   stage_hrma_tables (cfg, fa, da, lay, lds, ltab);
   __syncthreads ();

   const mxg_hrma_cfg &h = *(const mxg_hrma_cfg *) lds;
   const HrmaLds &L = *(const HrmaLds *) (lds + kHrmaLayOff);
   __shared__ uint32_t seg_wc[2 * (kBlock / 64)];
   const BlockRange range = block_range (*io.n_in, blockIdx.x, gridDim.x);
   SegWriter survivors (io, range, seg_wc);

   for (uint32_t base = range.lo; base < range.hi; base += kBlock)
     {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t id = 0;
        PhotonOut o;
        if (i < range.hi)
          {
             id = io.idx_in[i];
             keep = hrma_reflect_one<RELAY> (h, ltab, L, fa, da, s, id, rng_pos, seed, source_id, o);
          }
        const uint32_t q = survivors.add (keep, id);
        if (RELAY && keep) relay_put (relay, q, o);
     }
   survivors.finish (io);
}

hipError_t launch_hrma_reflect (const mxg_config *cfg, const mxg_config *hcfg, const float *fa,
                                const double *da, Soa s, ListIO io, const uint32_t *rng_pos, int waves,
                                unsigned long long seed, uint32_t source_id, const Relay *relay_out,
                                hipStream_t st)
{
   const HrmaLds lay = hrma_lds_layout (hcfg->hrma);
   size_t lds_bytes = kHrmaTabOff + (size_t) lay.floats * sizeof (float);
   if (lds_bytes > 64 * 1024)
     return hipErrorInvalidConfiguration;
   typedef void (*ReflectFn) (const mxg_config *, const float *, const double *, Soa, ListIO,
                              const uint32_t *, unsigned long long, uint32_t, HrmaLds, Relay);
   const bool r = relay_out != nullptr;
   const ReflectFn reflect
     = (waves <= 2) ? (r ? k_hrma_reflect<2, true> : k_hrma_reflect<2, false>)
       : (waves == 3) ? (r ? k_hrma_reflect<3, true> : k_hrma_reflect<3, false>)
       : (r ? k_hrma_reflect<MXG_REFLECT_WAVES_TOP, true> : k_hrma_reflect<MXG_REFLECT_WAVES_TOP, false>);
   const Relay rel = r ? *relay_out : Relay {};
   hipLaunchKernelGGL (reflect, dim3 (io.grid), dim3 (kBlock), lds_bytes, st, cfg, fa, da, s, io, rng_pos,
                       seed, source_id, lay, rel);
   return hipGetLastError ();
}

hipError_t launch_hrma (const mxg_config *cfg, const mxg_config *hcfg, const float *fa,
                        const double *da, Soa s, ListIO io, HrmaScratch hs, uint32_t n_upper,
                        double source_distance, unsigned long long seed, uint32_t source_id,
                        const Relay *relay_out, bool fresh, hipStream_t st, hipEvent_t *ev)
{
   if (n_upper == 0) return hipSuccess;
   const HrmaLds lay = hrma_lds_layout (hcfg->hrma);
   if (kHrmaTabOff + (size_t) lay.floats * sizeof (float) > 64 * 1024)
     return hipErrorInvalidConfiguration;

   // io.idx_in -> k_hrma_enter -> middle list -> k_hrma_reflect -> io.idx_out
   ListIO a = io, b = io;
   a.idx_out = hs.idx_mid;
   a.n_out = hs.n_mid;
   a.n_copy = hs.n_mid_copy;
   b.idx_in = hs.idx_mid;
   b.n_in = hs.n_mid;
   b.grid = hs.reflect_grid;
   hipError_t e = hipSuccess;
   if (ev) e = hipEventRecord (ev[0], st);
   if (e != hipSuccess) return e;
   if (fresh)
     hipLaunchKernelGGL (k_hrma_enter<true>, dim3 (a.grid), dim3 (kBlock), 0, st, cfg, s, a, hs.rng_pos,
                         source_distance, seed, source_id);
   else
     hipLaunchKernelGGL (k_hrma_enter<false>, dim3 (a.grid), dim3 (kBlock), 0, st, cfg, s, a, hs.rng_pos,
                         source_distance, seed, source_id);
   e = hipGetLastError ();
   if ((e == hipSuccess) && ev) e = hipEventRecord (ev[1], st);
   if (e == hipSuccess) e = launch_compact (a, st);
   if ((e == hipSuccess) && ev) e = hipEventRecord (ev[2], st);
   if (e == hipSuccess)
     e = launch_hrma_reflect (cfg, hcfg, fa, da, s, b, hs.rng_pos, hs.reflect_waves, seed, source_id, relay_out,
                              st);
   if ((e == hipSuccess) && ev) e = hipEventRecord (ev[3], st);
   if (e == hipSuccess) e = launch_compact (b, st, relay_out ? relay_out->pos : nullptr);
   return e;
}

} // namespace mxg
