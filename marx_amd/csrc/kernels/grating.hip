// grating.hip -- the HETG / LETG stage kernel, one lane per photon (the
// per-photon work is in grating_stage.h).
#include "grating_stage.h"

namespace mxg {

// One kernel per grating: the LETG instance adds the support-grating draws,
// which the HETG instance leaves out (so it does not hold their registers).
// W: the occupancy each measured fastest at (waves per SIMD, 0 unbounded).
// RIN: the input list is the mirror's with relay rin (list entry i at row
// rin.pos[i]); ROUT: the survivors' columns also go to rout at their
// segment slots.
template <bool LETG, int W, bool RIN, bool ROUT>
__global__ __launch_bounds__ (kBlock, W > 0 ? W : 1)
void k_grating (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                const double *__restrict__ da, Soa s, ListIO io,
                unsigned long long seed, uint32_t source_id, Relay rin, Relay rout)
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
        PhotonOut o;
        if (i < range.hi)
          {
             id = io.idx_in[i];
             keep = grating_one<LETG, RIN, ROUT> (gc, fa, da, s, id, seed, source_id, o, &rin,
                                                  RIN ? rin.pos[i] : 0u);
          }
        const uint32_t q = survivors.add (keep, id);
        if (ROUT && keep) relay_put (rout, q, o);
     }
   survivors.finish (io);
}

template <bool LETG, int W>
static void launch_io (dim3 g, hipStream_t st, const mxg_config *cfg, const float *fa, const double *da, Soa s,
                       ListIO io, unsigned long long seed, uint32_t source_id, const Relay *rin,
                       const Relay *rout)
{
   const Relay a = rin ? *rin : Relay {}, b = rout ? *rout : Relay {};
   const dim3 blk (kBlock);
   if (rin && rout)
     hipLaunchKernelGGL ((k_grating<LETG, W, true, true>), g, blk, 0, st, cfg, fa, da, s, io, seed, source_id, a, b);
   else if (rin)
     hipLaunchKernelGGL ((k_grating<LETG, W, true, false>), g, blk, 0, st, cfg, fa, da, s, io, seed, source_id, a,
                         b);
   else if (rout)
     hipLaunchKernelGGL ((k_grating<LETG, W, false, true>), g, blk, 0, st, cfg, fa, da, s, io, seed, source_id, a,
                         b);
   else
     hipLaunchKernelGGL ((k_grating<LETG, W, false, false>), g, blk, 0, st, cfg, fa, da, s, io, seed, source_id, a,
                         b);
}

hipError_t launch_grating (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                           ListIO io, uint32_t n_upper, unsigned long long seed,
                           uint32_t source_id, bool letg, const Relay *relay_in, const Relay *relay_out,
                           hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   const dim3 g (io.grid);
   if (letg)
     launch_io<true, MXG_LETG_GRATING_WAVES> (g, st, cfg, fa, da, s, io, seed, source_id, relay_in, relay_out);
   else
     launch_io<false, MXG_GRATING_WAVES> (g, st, cfg, fa, da, s, io, seed, source_id, relay_in, relay_out);
   hipError_t e = hipGetLastError ();
   if (e == hipSuccess) e = launch_compact (io, st, relay_out ? relay_out->pos : nullptr);
   return e;
}

} // namespace mxg
