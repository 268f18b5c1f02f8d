// grating.hip -- the HETG / LETG stage kernel, one lane per photon (the
// per-photon work is in grating_stage.h).
#include "grating_stage.h"

namespace mxg {

// One kernel per grating: the LETG instance adds the support-grating draws,
// which the HETG instance leaves out (so it does not hold their registers).
// W: the occupancy each measured fastest at (waves per SIMD, 0 unbounded).
// RELAY: the survivors' columns also go to relay at their segment slots.
template <bool LETG, int W, bool RELAY>
__global__ __launch_bounds__ (kBlock, W > 0 ? W : 1)
void k_grating (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                const double *__restrict__ da, Soa s, ListIO io,
                unsigned long long seed, uint32_t source_id, Relay relay)
{
   const mxg_grating_cfg &gc = cfg->grating;
   __shared__ uint32_t seg_wc[2 * (kBlock / 64)];
   const BlockRange range = block_range (*io.n_in, blockIdx.x, gridDim.x);
   SegWriter survivors (io, range, seg_wc);

   for (uint32_t base = range.lo; base < range.hi; base += kBlock)
     {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t id = 0;
        GratingOut o;
        if (i < range.hi)
          {
             id = io.idx_in[i];
             keep = grating_one<LETG> (gc, fa, da, s, id, seed, source_id, o);
          }
        const uint32_t q = survivors.add (keep, id);
        if (RELAY && keep)
          {
             relay.x0[q] = o.x.x; relay.x1[q] = o.x.y; relay.x2[q] = o.x.z;
             relay.p0[q] = o.p.x; relay.p1[q] = o.p.y; relay.p2[q] = o.p.z;
             relay.energy[q] = o.energy;
             relay.tag[q] = o.tag;
             relay.flags[q] = o.flags;
          }
     }
   survivors.finish (io);
}

hipError_t launch_grating (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                           ListIO io, uint32_t n_upper, unsigned long long seed,
                           uint32_t source_id, bool letg, const Relay *relay, hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   const dim3 g (io.grid), b (kBlock);
   const Relay r = relay ? *relay : Relay {};
   if (letg && relay)
     hipLaunchKernelGGL ((k_grating<true, MXG_LETG_GRATING_WAVES, true>), g, b, 0, st, cfg, fa, da, s, io, seed,
                         source_id, r);
   else if (letg)
     hipLaunchKernelGGL ((k_grating<true, MXG_LETG_GRATING_WAVES, false>), g, b, 0, st, cfg, fa, da, s, io, seed,
                         source_id, r);
   else if (relay)
     hipLaunchKernelGGL ((k_grating<false, MXG_GRATING_WAVES, true>), g, b, 0, st, cfg, fa, da, s, io, seed,
                         source_id, r);
   else
     hipLaunchKernelGGL ((k_grating<false, MXG_GRATING_WAVES, false>), g, b, 0, st, cfg, fa, da, s, io, seed,
                         source_id, r);
   hipError_t e = hipGetLastError ();
   if (e == hipSuccess) e = launch_compact (io, st, relay ? relay->pos : nullptr);
   return e;
}

} // namespace mxg
