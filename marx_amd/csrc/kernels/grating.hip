// grating.hip -- the HETG / LETG stage kernel, one lane per photon (the
// per-photon work is in grating_stage.h).
#include "grating_stage.h"

namespace mxg {

// One kernel per grating: the LETG instance adds the support-grating draws,
// which the HETG instance leaves out (so it does not hold their registers).
// W: the occupancy each measured fastest at (waves per SIMD, 0 unbounded).
template <bool LETG, int W>
__global__ __launch_bounds__ (kBlock, W > 0 ? W : 1)
void k_grating (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
                const double *__restrict__ da, Soa s, ListIO io,
                unsigned long long seed, uint32_t source_id)
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
        if (i < range.hi)
          {
             id = io.idx_in[i];
             keep = grating_one<LETG> (gc, fa, da, s, id, seed, source_id);
          }
        survivors.add (keep, id);
     }
   survivors.finish (io);
}

hipError_t launch_grating (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                           ListIO io, uint32_t n_upper, unsigned long long seed,
                           uint32_t source_id, bool letg, hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   if (letg)
     hipLaunchKernelGGL ((k_grating<true, MXG_LETG_GRATING_WAVES>), dim3 (io.grid), dim3 (kBlock), 0, st, cfg,
                         fa, da, s, io, seed, source_id);
   else
     hipLaunchKernelGGL ((k_grating<false, MXG_GRATING_WAVES>), dim3 (io.grid), dim3 (kBlock), 0, st, cfg, fa,
                         da, s, io, seed, source_id);
   hipError_t e = hipGetLastError ();
   if (e == hipSuccess) e = launch_compact (io, st);
   return e;
}

} // namespace mxg
