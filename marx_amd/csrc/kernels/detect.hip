// detect.hip -- the focal-plane detection kernel, one lane per photon (the
// per-photon work is in detect_stage.h).
#include "detect_stage.h"

namespace mxg {

// One instance per detector family (KIND: MXG_DET_ACIS, MXG_DET_HRC_S,
// MXG_DET_HRC_I), so that the ACIS pass does not hold the registers of the
// HRC and HESF code; FEF: the ACIS instance with the FEF step; RELAY: the
// input list is the grating's with its relay (list entry i at relay row
// relay.pos[i]).
template <bool FEF, int KIND, int W, bool RELAY>
__global__ __launch_bounds__ (kBlock, W > 0 ? W : 1)
void k_detect (const mxg_config *__restrict__ cfg, const float *__restrict__ fa,
               const double *__restrict__ da, Soa s, ListIO io, double tstart,
               unsigned long long seed, uint32_t source_id, Relay relay)
{
   const mxg_detector_cfg &d = cfg->det;
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
             keep = detect_one<FEF, KIND, RELAY> (d, fa, da, s, id, tstart, seed, source_id, &relay,
                                                  RELAY ? relay.pos[i] : 0u);
          }
        survivors.add (keep, id);
     }
   survivors.finish (io);
}

template <bool FEF, int KIND, int W>
static void launch_kind (dim3 g, hipStream_t st, const mxg_config *cfg, const float *fa, const double *da, Soa s,
                         ListIO io, double tstart, unsigned long long seed, uint32_t source_id,
                         const Relay *relay)
{
   const Relay r = relay ? *relay : Relay {};
   if (relay)
     hipLaunchKernelGGL ((k_detect<FEF, KIND, W, true>), g, dim3 (kBlock), 0, st, cfg, fa, da, s, io, tstart, seed,
                         source_id, r);
   else
     hipLaunchKernelGGL ((k_detect<FEF, KIND, W, false>), g, dim3 (kBlock), 0, st, cfg, fa, da, s, io, tstart, seed,
                         source_id, r);
}

hipError_t launch_detect (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                          ListIO io, uint32_t n_upper, double tstart, unsigned long long seed,
                          uint32_t source_id, bool fef, int kind, const Relay *relay, hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   const dim3 g (io.grid);
   if (kind == MXG_DET_HRC_S)
     launch_kind<false, MXG_DET_HRC_S, MXG_DETECT_WAVES> (g, st, cfg, fa, da, s, io, tstart, seed, source_id, relay);
   else if (kind == MXG_DET_HRC_I)
     launch_kind<false, MXG_DET_HRC_I, MXG_DETECT_WAVES> (g, st, cfg, fa, da, s, io, tstart, seed, source_id, relay);
   else if (fef)
     launch_kind<true, MXG_DET_ACIS, MXG_DETECT_WAVES> (g, st, cfg, fa, da, s, io, tstart, seed, source_id, relay);
   else
     launch_kind<false, MXG_DET_ACIS, MXG_ACIS_DETECT_WAVES> (g, st, cfg, fa, da, s, io, tstart, seed, source_id,
                                                             relay);
   hipError_t e = hipGetLastError ();
   if (e == hipSuccess) e = launch_compact (io, st);
   return e;
}

} // namespace mxg
