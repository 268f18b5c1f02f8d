// persist.hip -- the batched pass as one persistent launch.
//
// The staged batch pass runs one kernel per stage over the whole batch, with
// a compaction between stages.  Here one launch does all of it: each block
// stages the mirror configuration and tables in LDS once, then takes tiles
// of `tile` consecutive batch photons from a work counter until the batch is
// done.  Within a tile the stages run back to back with the surviving
// photon ids kept in an LDS list, compacted in place and in order after each
// stage (one block-wide prefix per kBlock photons), so every stage runs on
// full wavefronts without a global round trip:
//
//   batch source + HRMA up to the paraboloid  (every photon of the tile)
//   -> paraboloid, struts, hyperboloid         (the paraboloid hits)
//   -> HETG / LETG                             (when configured)
//   -> detector                                (when configured)
//
// The detected photons of tile t go to seg[t * tile ...] in id order and
// their number to tile_count[t]; k_tile_offsets and k_tile_gather then
// concatenate the tiles into the output list, so the list is the one the
// staged pass makes.  The per-photon work is the stage kernels' own
// (source_stage.h, hrma_stage.h, grating_stage.h, detect_stage.h) and every
// draw is keyed by the photon's tag, so the photons are the same whichever
// block and tile trace them.
#include "source_stage.h"
#include "hrma_stage.h"
#include "grating_stage.h"
#include "detect_stage.h"

namespace mxg {

__global__ void k_zero_u32 (uint32_t *p, uint32_t n)
{
   if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

// In-place order-preserving compaction of an LDS id list: like SegWriter,
// but the output is the list being read.  add() is called by every lane
// once per pass of kBlock inputs and holds one barrier; the outputs of a
// pass land at or before its inputs, which every lane has read before the
// barrier.
struct LdsCompactor
{
   uint32_t *list;
   uint32_t *wc;
   uint32_t count;
   uint32_t parity;

   __device__ __forceinline__ LdsCompactor (uint32_t *l, uint32_t *lds_wc)
      : list (l), wc (lds_wc), count (0), parity (0) {}

   __device__ __forceinline__ void add (bool keep, uint32_t id, uint32_t *out)
   {
      constexpr int kWaves = kBlock / 64;
      unsigned long long mask = __ballot (keep);
      const int wave = threadIdx.x >> 6, lane = __lane_id ();
      uint32_t *w = wc + parity * kWaves;
      if (lane == 0) w[wave] = (uint32_t) __popcll (mask);
      __syncthreads ();
      uint32_t before = 0, total = 0;
#pragma unroll
      for (int k = 0; k < kWaves; k++)
        {
           uint32_t c = w[k];
           before += (k < wave) ? c : 0u;
           total += c;
        }
      if (keep)
        out[count + before + (uint32_t) __popcll (mask & ((1ull << lane) - 1ull))] = id;
      count += total;
      parity ^= 1u;
   }
};

struct PersistArgs
{
   const mxg_config *cfg;
   const float *fa;
   const double *da;
   Soa s;
   uint32_t n;                      // photons of the batch
   const uint32_t *off;             // nsrc + 1 prefix offsets of the sources
   uint32_t nsrc;
   const double *dir;               // 3 nsrc directions
   const uint32_t *tag_start;       // tag of photon 0 (device memory)
   double exposure, start_time, tstart;
   unsigned long long seed;
   uint32_t *rng_pos;               // per photon: stream position at the mirror split
   uint32_t *work;                  // next tile (0 at the launch)
   uint32_t *stage_count;           // photons into the reflection, out of the mirror,
                                    // grating and detector (0 at the launch)
   uint32_t tile;                   // photons per tile, a multiple of kBlock
   uint32_t num_tiles;
   uint32_t *seg;                   // n entries: the survivors of tile t at t * tile
   uint32_t *tile_count;            // num_tiles entries
   HrmaLds lay;
   uint32_t list_off;               // byte offset of the tile's id list in dynamic LDS
};

// G: 0 no grating, 1 HETG, 2 LETG; KIND: -1 no detector, else the detector
// family of detect_one.  Point sources, without the FEF step: a batch of
// extended sources or a FEF run takes the staged pass.
template <int G, int KIND>
__global__ __launch_bounds__ (kBlock, MXG_PERSIST_WAVES)
void k_batch_persistent (PersistArgs a)
{
   extern __shared__ __align__ (16) unsigned char lds[];
   float *ltab = (float *) (lds + kHrmaTabOff);
   uint32_t *list = (uint32_t *) (lds + a.list_off);
   __shared__ uint32_t wc[2 * (kBlock / 64)];
   __shared__ uint32_t sh_tile;
   stage_hrma_tables (a.cfg, a.fa, a.da, a.lay, lds, ltab);
   __syncthreads ();

   const mxg_hrma_cfg &h = *(const mxg_hrma_cfg *) lds;
   const HrmaLds &L = *(const HrmaLds *) (lds + kHrmaLayOff);
   const mxg_source_cfg &src = a.cfg->src;
   const uint32_t t0 = *a.tag_start;
   const double distance = src.distance;

   for (;;)
     {
        if (threadIdx.x == 0) sh_tile = atomicAdd (a.work, 1u);
        __syncthreads ();
        const uint32_t t = sh_tile;
        if (t >= a.num_tiles) break;            // the same for the whole block
        const uint32_t lo = t * a.tile;
        const uint32_t hi = (a.n - lo < a.tile) ? a.n : lo + a.tile;

        // the tile's photons, up to the paraboloid; the hits are listed
        LdsCompactor c (list, wc);
        for (uint32_t base = lo; base < hi; base += kBlock)
          {
             const uint32_t i = base + threadIdx.x;
             bool keep = false;
             if (i < hi)
               {
                  batch_source_one<true, true> (src, a.da, a.s, i, a.off, a.nsrc, a.dir, t0, a.exposure,
                                                a.start_time, a.seed);
                  keep = hrma_enter_one<true> (h, a.s, i, a.rng_pos, distance, a.seed, 0u);
               }
             c.add (keep, i, list);
          }
        __syncthreads ();
        uint32_t m = c.count;
        const uint32_t m_mid = m;

        // the rest of the mirror
        {
           LdsCompactor r (list, wc);
           for (uint32_t base = 0; base < m; base += kBlock)
             {
                const uint32_t j = base + threadIdx.x;
                bool keep = false;
                uint32_t id = 0;
                if (j < m)
                  {
                     id = list[j];
                     PhotonOut o;
                     keep = hrma_reflect_one (h, ltab, L, a.fa, a.da, a.s, id, a.rng_pos, a.seed, 0u, o);
                  }
                r.add (keep, id, list);
             }
           __syncthreads ();
           m = r.count;
        }
        const uint32_t m_mirror = m;

        if constexpr (G != 0)
          {
             LdsCompactor r (list, wc);
             for (uint32_t base = 0; base < m; base += kBlock)
               {
                  const uint32_t j = base + threadIdx.x;
                  bool keep = false;
                  uint32_t id = 0;
                  if (j < m)
                    {
                       id = list[j];
                       PhotonOut o;
                       keep = grating_one<G == 2> (a.cfg->grating, a.fa, a.da, a.s, id, a.seed, 0u, o);
                    }
                  r.add (keep, id, list);
               }
             __syncthreads ();
             m = r.count;
          }
        const uint32_t m_grating = m;

        // the detector, its detected photons straight to the tile's segment
        uint32_t *out = a.seg + lo;
        if constexpr (KIND >= 0)
          {
             LdsCompactor r (list, wc);
             for (uint32_t base = 0; base < m; base += kBlock)
               {
                  const uint32_t j = base + threadIdx.x;
                  bool keep = false;
                  uint32_t id = 0;
                  if (j < m)
                    {
                       id = list[j];
                       keep = detect_one<false, KIND> (a.cfg->det, a.fa, a.da, a.s, id, a.tstart, a.seed, 0u);
                    }
                  r.add (keep, id, out);
               }
             m = r.count;
          }
        else
          {
             for (uint32_t j = threadIdx.x; j < m; j += kBlock) out[j] = list[j];
          }
        if (threadIdx.x == 0)
          {
             a.tile_count[t] = m;
             atomicAdd (a.stage_count + 0, m_mid);
             atomicAdd (a.stage_count + 1, m_mirror);
             atomicAdd (a.stage_count + 2, m_grating);
             atomicAdd (a.stage_count + 3, m);
          }
        // every lane is past its reads of sh_tile and of the list
        __syncthreads ();
     }
}

// Exclusive prefix of the tile counts (one block): tile_off[t], and the
// total in *n_out.
__global__ __launch_bounds__ (kBlock)
void k_tile_offsets (const uint32_t *__restrict__ tile_count, uint32_t num_tiles, uint32_t *__restrict__ tile_off,
                     uint32_t *__restrict__ n_out)
{
   __shared__ uint32_t part[kBlock];
   const uint32_t per = (num_tiles + kBlock - 1) / kBlock;
   const uint32_t lo = threadIdx.x * per;
   const uint32_t hi = (lo + per < num_tiles) ? lo + per : num_tiles;
   uint32_t sum = 0;
   for (uint32_t k = lo; k < hi; k++) sum += tile_count[k];
   part[threadIdx.x] = sum;
   __syncthreads ();
   // Hillis-Steele inclusive scan of the kBlock partial sums
   for (uint32_t d = 1; d < kBlock; d <<= 1)
     {
        uint32_t v = (threadIdx.x >= d) ? part[threadIdx.x - d] : 0u;
        __syncthreads ();
        part[threadIdx.x] += v;
        __syncthreads ();
     }
   uint32_t run = part[threadIdx.x] - sum;
   for (uint32_t k = lo; k < hi; k++)
     {
        tile_off[k] = run;
        run += tile_count[k];
     }
   if (threadIdx.x == kBlock - 1) *n_out = part[kBlock - 1];
}

// Tile t's segment copied to out[tile_off[t] ...] (grid-stride over tiles).
__global__ __launch_bounds__ (kBlock)
void k_tile_gather (const uint32_t *__restrict__ seg, const uint32_t *__restrict__ tile_count,
                    const uint32_t *__restrict__ tile_off, uint32_t num_tiles, uint32_t tile,
                    uint32_t *__restrict__ out)
{
   for (uint32_t t = blockIdx.x; t < num_tiles; t += gridDim.x)
     {
        const uint32_t cnt = tile_count[t], o = tile_off[t];
        const uint32_t *src = seg + (size_t) t * tile;
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) out[o + k] = src[k];
     }
}

// Dynamic LDS of the persistent kernel for a tile of `tile` photons.
size_t persist_lds_bytes (const HrmaLds &lay, uint32_t tile, uint32_t *list_off)
{
   size_t off = kHrmaTabOff + (size_t) lay.floats * sizeof (float);
   off = (off + 15) & ~(size_t) 15;
   if (list_off) *list_off = (uint32_t) off;
   return off + (size_t) tile * sizeof (uint32_t);
}

template <int G>
static hipError_t launch_g (int kind, dim3 grid, size_t lds, hipStream_t st, const PersistArgs &a)
{
   switch (kind)
     {
      case MXG_DET_ACIS:
        hipLaunchKernelGGL ((k_batch_persistent<G, MXG_DET_ACIS>), grid, dim3 (kBlock), lds, st, a);
        break;
      case MXG_DET_HRC_S:
        hipLaunchKernelGGL ((k_batch_persistent<G, MXG_DET_HRC_S>), grid, dim3 (kBlock), lds, st, a);
        break;
      case MXG_DET_HRC_I:
        hipLaunchKernelGGL ((k_batch_persistent<G, MXG_DET_HRC_I>), grid, dim3 (kBlock), lds, st, a);
        break;
      default:
        hipLaunchKernelGGL ((k_batch_persistent<G, -1>), grid, dim3 (kBlock), lds, st, a);
        break;
     }
   return hipGetLastError ();
}

hipError_t launch_batch_persistent (const PersistLaunch &p, hipStream_t st)
{
   if (p.n == 0) return hipSuccess;
   if ((p.tile == 0) || (p.tile % kBlock) || (p.tile > kMaxPersistTile))
     return hipErrorInvalidValue;
   const mxg_hrma_cfg &h = p.hcfg->hrma;
   PersistArgs a;
   a.lay = hrma_lds_layout (h);
   const size_t lds = persist_lds_bytes (a.lay, p.tile, &a.list_off);
   if (lds > 64 * 1024)
     return hipErrorInvalidConfiguration;
   a.cfg = p.cfg;
   a.fa = p.fa;
   a.da = p.da;
   a.s = p.s;
   a.n = p.n;
   a.off = p.off;
   a.nsrc = p.nsrc;
   a.dir = p.dir;
   a.tag_start = p.tag_start;
   a.exposure = p.exposure;
   a.start_time = p.start_time;
   a.tstart = p.start_time;
   a.seed = p.seed;
   a.rng_pos = p.rng_pos;
   a.work = p.work;
   a.stage_count = p.stage_count;
   a.tile = p.tile;
   a.num_tiles = (p.n + p.tile - 1) / p.tile;
   a.seg = p.seg;
   a.tile_count = p.tile_count;

   // counters start at zero (kernels, not memsets: the sequence may be
   // captured into a graph)
   hipLaunchKernelGGL (k_zero_u32, dim3 (1), dim3 (64), 0, st, p.work, 1u);
   hipLaunchKernelGGL (k_zero_u32, dim3 (1), dim3 (64), 0, st, p.stage_count, 4u);
   hipError_t e = hipGetLastError ();
   if (e != hipSuccess) return e;
   if (p.ev) e = hipEventRecord (p.ev[0], st);
   if (e != hipSuccess) return e;
   uint32_t grid = p.blocks;
   if (grid > a.num_tiles) grid = a.num_tiles;
   if (grid == 0) grid = 1;
   if (p.grating == 1) e = launch_g<1> (p.kind, dim3 (grid), lds, st, a);
   else if (p.grating == 2) e = launch_g<2> (p.kind, dim3 (grid), lds, st, a);
   else e = launch_g<0> (p.kind, dim3 (grid), lds, st, a);
   if (e != hipSuccess) return e;
   if (p.ev) e = hipEventRecord (p.ev[1], st);
   if (e != hipSuccess) return e;
   hipLaunchKernelGGL (k_tile_offsets, dim3 (1), dim3 (kBlock), 0, st, p.tile_count, a.num_tiles, p.tile_off,
                       p.n_out);
   uint32_t g2 = a.num_tiles < 4096u ? a.num_tiles : 4096u;
   hipLaunchKernelGGL (k_tile_gather, dim3 (g2), dim3 (kBlock), 0, st, p.seg, p.tile_count, p.tile_off,
                       a.num_tiles, p.tile, p.idx_out);
   return hipGetLastError ();
}

} // namespace mxg
