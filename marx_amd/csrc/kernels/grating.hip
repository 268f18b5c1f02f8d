// grating.hip -- the HETG / LETG stage kernel, one lane per photon (the
// per-photon work is in grating_stage.h).
#include "grating_stage.h"

namespace mxg {

// One kernel per grating: the LETG instance adds the support-grating draws,
// which the HETG instance leaves out (so it does not hold their registers).
// W: the occupancy each measured fastest at (waves per SIMD, 0 unbounded).
// RIN: the input list is the mirror's with relay rin (list entry i at row
// rin.pos[i]); ROUT: the survivors' columns also go to rout at their
// segment slots.  The block first stages the search tables (energies,
// bucket counts, orders, sectors) in LDS at the layout lay: they are the
// targets of each photon's chains of dependent loads.
template <bool LETG, int W, bool RIN, bool ROUT>
__global__ __launch_bounds__ (kBlock, W > 0 ? W : 1)
void k_grating (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                const double *__restrict__ da, Soa s, ListIO io,
                unsigned long long seed, uint32_t source_id, GratingLds lay, Relay rin, Relay rout)
{
   const mxg_grating_cfg &gc = cfg->grating;
   extern __shared__ __align__ (16) unsigned char lds[];
   float *ltab = (float *) (lds + kGratingTabOff);
   stage_grating_tables (gc, fa, da, lay, LETG, lds, ltab);
   __syncthreads ();
   const GratingLds *L = (const GratingLds *) lds;
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
             keep = grating_one<LETG, RIN, ROUT, true> (gc, fa, da, s, id, seed, source_id, o, &rin,
                                                        RIN ? rin.pos[i] : 0u, L, ltab);
          }
        const uint32_t q = survivors.add (keep, id);
        if (ROUT && keep) relay_put (rout, q, o);
     }
   survivors.finish (io);
}

template <bool LETG, int W>
static void launch_io (dim3 g, size_t lds_bytes, hipStream_t st, const mxg_config *cfg, const float *fa,
                       const double *da, Soa s, ListIO io, unsigned long long seed, uint32_t source_id,
                       const GratingLds &lay, const Relay *rin, const Relay *rout)
{
   const Relay a = rin ? *rin : Relay {}, b = rout ? *rout : Relay {};
   const dim3 blk (kBlock);
   if (rin && rout)
     hipLaunchKernelGGL ((k_grating<LETG, W, true, true>), g, blk, lds_bytes, st, cfg, fa, da, s, io, seed,
                         source_id, lay, a, b);
   else if (rin)
     hipLaunchKernelGGL ((k_grating<LETG, W, true, false>), g, blk, lds_bytes, st, cfg, fa, da, s, io, seed,
                         source_id, lay, a, b);
   else if (rout)
     hipLaunchKernelGGL ((k_grating<LETG, W, false, true>), g, blk, lds_bytes, st, cfg, fa, da, s, io, seed,
                         source_id, lay, a, b);
   else
     hipLaunchKernelGGL ((k_grating<LETG, W, false, false>), g, blk, lds_bytes, st, cfg, fa, da, s, io, seed,
                         source_id, lay, a, b);
}

uint32_t grating_blocks_per_cu (const mxg_config *hcfg)
{
   const bool letg = hcfg->grating.type == MARX_GRATING_LETG;
   const GratingLds lay = grating_lds_layout (hcfg->grating, letg);
   const size_t lds_bytes = kGratingTabOff + (size_t) lay.floats * sizeof (float);
   int n = 0;
   const hipError_t e
     = letg ? hipOccupancyMaxActiveBlocksPerMultiprocessor (&n, k_grating<true, MXG_LETG_GRATING_WAVES, false, false>,
                                                             kBlock, lds_bytes)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor (&n, k_grating<false, MXG_GRATING_WAVES, false, false>,
                                                             kBlock, lds_bytes);
   return ((e == hipSuccess) && (n > 0)) ? (uint32_t) n : 8u;
}

hipError_t launch_grating (const mxg_config *cfg, const mxg_config *hcfg, const float *fa, const double *da,
                           Soa s, ListIO io, uint32_t n_upper, unsigned long long seed,
                           uint32_t source_id, bool letg, const Relay *relay_in, const Relay *relay_out,
                           hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   const dim3 g (io.grid);
   const GratingLds lay = grating_lds_layout (hcfg->grating, letg);
   const size_t lds_bytes = kGratingTabOff + (size_t) lay.floats * sizeof (float);
   if (lds_bytes > 64 * 1024)
     return hipErrorInvalidConfiguration;
   if (letg)
     launch_io<true, MXG_LETG_GRATING_WAVES> (g, lds_bytes, st, cfg, fa, da, s, io, seed, source_id, lay,
                                             relay_in, relay_out);
   else
     launch_io<false, MXG_GRATING_WAVES> (g, lds_bytes, st, cfg, fa, da, s, io, seed, source_id, lay, relay_in,
                                         relay_out);
   hipError_t e = hipGetLastError ();
   if (e == hipSuccess) e = launch_compact (io, st, relay_out ? relay_out->pos : nullptr);
   return e;
}

} // namespace mxg
