// dev_store.h -- the structure-of-arrays photon store in HBM and the launch
// helpers shared by the stage kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "marx_gpu.h"
#include "marx_gpu_cfg.h"

namespace mxg {

// Column pointers of the resident photon store (one element per photon).
struct Soa
{
   double *energy;
   double *x0, *x1, *x2;
   double *p0, *p1, *p2;
   double *t;
   uint32_t *flags;
   float *ypix, *zpix, *upix, *vpix;
   float *dith[6];
   float *pi;
   int16_t *pha;
   uint8_t *shell;
   int8_t *ccd, *region, *order;
   int8_t *sup[4];
   uint32_t *tag;
};

// Per-stage launch arguments: the active index list in, the survivors out.
//
// Survivors are collected without global atomics: block b of a stage kernel
// (grid = `grid` blocks of kBlock lanes) owns the input positions
// [lo_b, hi_b) of block_range() and writes the ids of its survivors, in input
// order, to seg[lo_b ...] and their number to seg_count[b].  k_compact then
// concatenates the segments into idx_out / n_out.  An input list sorted by id
// therefore yields an output list sorted by id.
struct ListIO
{
   const uint32_t *idx_in;
   const uint32_t *n_in;
   uint32_t *idx_out;      // compacted survivors (written by k_compact)
   uint32_t *n_out;
   uint32_t *seg;          // per-block survivor segments, capacity entries
   uint32_t *seg_count;    // grid entries
   uint32_t grid;          // blocks of the stage kernel (<= kMaxSegments)
   uint32_t *n_copy = nullptr;  // when non-null, k_compact also writes the length here
};

constexpr int kBlock = 256;
constexpr uint32_t kMaxSegments = 8192;

// A stage's survivors handed densely to the next stage: the columns the
// grating and the detector read of a photon (x, p, energy, tag, flags,
// shell), written by the producing kernel (the reflection, the grating) at
// the survivor's segment slot (its position in seg), and pos, written by
// k_compact: the slot of each entry of the compacted list.  After the
// mirror about one photon id in three is alive, after the grating one in
// sixteen, so reading the double columns at the ids costs a cache line per
// column for a few photons; the relay's rows are contiguous.
struct Relay
{
   double *x0, *x1, *x2, *p0, *p1, *p2, *energy;
   uint32_t *tag, *flags;
   uint8_t *shell;
   uint32_t *pos;
};

// A photon leaving a stage: the columns of its relay row.
struct PhotonOut
{
   double x0, x1, x2, p0, p1, p2, energy;
   uint32_t tag, flags;
   int shell;
};

// a survivor's row of a relay
__device__ __forceinline__ void relay_put (const Relay &r, uint32_t q, const PhotonOut &o)
{
   r.x0[q] = o.x0; r.x1[q] = o.x1; r.x2[q] = o.x2;
   r.p0[q] = o.p0; r.p1[q] = o.p1; r.p2[q] = o.p2;
   r.energy[q] = o.energy;
   r.tag[q] = o.tag;
   r.flags[q] = o.flags;
   r.shell[q] = (uint8_t) o.shell;
}


struct BlockRange { uint32_t lo, hi; };

// The input positions owned by block b of nblocks: equal chunks rounded up
// to whole blocks of lanes.
__device__ __forceinline__ BlockRange block_range (uint32_t n, uint32_t b, uint32_t nblocks)
{
   uint64_t per = ((uint64_t) n + nblocks - 1) / nblocks;
   uint64_t chunk = (per + kBlock - 1) / kBlock * kBlock;
   uint64_t lo = (uint64_t) b * chunk;
   uint64_t hi = lo + chunk;
   BlockRange r;
   r.lo = (uint32_t) (lo < n ? lo : n);
   r.hi = (uint32_t) (hi < n ? hi : n);
   return r;
}

// Order-preserving survivor output of one block (see ListIO).  add() is
// called by every lane of the block once per pass of kBlock inputs; it
// holds one barrier.  wc is a 2 x (kBlock / 64) LDS array.
struct SegWriter
{
   uint32_t *out;
   uint32_t *wc;
   uint32_t lo;
   uint32_t count;
   uint32_t parity;

   __device__ __forceinline__ SegWriter (const ListIO &io, const BlockRange &r, uint32_t *lds_wc)
      : out (io.seg + r.lo), wc (lds_wc), lo (r.lo), count (0), parity (0) {}

   // returns the slot (position in seg) of a kept id
   __device__ __forceinline__ uint32_t add (bool keep, uint32_t id)
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
      const uint32_t q = count + before + (uint32_t) __popcll (mask & ((1ull << lane) - 1ull));
      if (keep) out[q] = id;
      count += total;
      parity ^= 1u;
      return lo + q;
   }

   __device__ __forceinline__ void finish (const ListIO &io)
   {
      if (threadIdx.x == 0) io.seg_count[blockIdx.x] = count;
   }
};

// Launches k_compact over the segments a stage kernel of io.grid blocks
// wrote; pos_out, when non-null, receives the slot of each list entry.
hipError_t launch_compact (const ListIO &io, hipStream_t st, uint32_t *pos_out = nullptr);

// Output files (output.hip): the photons not flagged bad, in order, then the
// requested columns of them packed big-endian, column k at out + offset[k].
constexpr int kMaxPackCols = 32;
struct PackArgs
{
   int32_t ncols;
   int32_t col[kMaxPackCols];
   uint64_t offset[kMaxPackCols];
   double total_time;
};
hipError_t launch_keep_list (const uint32_t *flags, uint32_t n, const ListIO &io, hipStream_t st);

// Event coordinates (events.hip): row q of each column for photon idx[q].
struct EventOut
{
   uint32_t *tag;
   int32_t *ccd, *tdetx, *tdety;
   double *detx, *dety, *x, *y;
   int16_t *fltgrade;
};
// The frame heads (first row of each row's frame, for the held dither
// state) take head[m], chunk_max[ceil(m / kBlock)] and chunk_pre[same].
hipError_t launch_event_coords (const mxg_config *cfg, const mxg_event_params *params, const float *subtab,
                                const Soa &s, const uint32_t *idx, uint32_t m, double total_time, double time_del,
                                unsigned long long seed, uint32_t *head, uint32_t *chunk_max, uint32_t *chunk_pre,
                                const EventOut &out, hipStream_t st);
hipError_t launch_pack_columns (const Soa &s, const uint32_t *idx, const uint32_t *n_kept, uint32_t m,
                                const PackArgs &a, uint8_t *out, hipStream_t st);

// Minimum waves per SIMD of the mirror's reflection kernel unless the
// launch schedule says otherwise (256 lanes per block = one wave per SIMD
// per block).  Unbounded, the compiler gives it ~200 VGPRs, i.e. two waves
// per SIMD, too few to hide its LDS and table-load latency; four waves
// (128 VGPRs, a few spills) measured fastest on the HETG workload.
#define MXG_REFLECT_WAVES_DEFAULT 4
// waves per SIMD that the compiler must leave room for (0: unbounded).
// k_source needs fewer VGPRs than 4 waves allow (56 for a point source),
// k_batch_source at 4 (from 175) measured 29 percent faster, the LETG
// grating pass at 4 (from 218) 25 percent, the HETG grating pass at 3 (from
// 208) 8 percent; the ACIS detector kernel needs 112 VGPRs (4 waves)
// unbounded
#ifndef MXG_SOURCE_WAVES
#define MXG_SOURCE_WAVES 4
#endif
#ifndef MXG_BATCH_SOURCE_WAVES
#define MXG_BATCH_SOURCE_WAVES 4
#endif
// waves of the reflection kernel of the schedule's largest setting (4; 5
// measured 21 percent slower), and of the aperture kernel (unbounded, 95
// VGPRs; 6 and 8 measured slower)
#ifndef MXG_REFLECT_WAVES_TOP
#define MXG_REFLECT_WAVES_TOP 4
#endif
#ifndef MXG_ENTER_WAVES
#define MXG_ENTER_WAVES 0
#endif
#ifndef MXG_LETG_GRATING_WAVES
#define MXG_LETG_GRATING_WAVES 4
#endif
#ifndef MXG_GRATING_WAVES
#define MXG_GRATING_WAVES 3
#endif
#ifndef MXG_DETECT_WAVES
#define MXG_DETECT_WAVES 0
#endif
// the ACIS instance of the detector kernel (without FEF): waves per SIMD
#ifndef MXG_ACIS_DETECT_WAVES
#define MXG_ACIS_DETECT_WAVES 0
#endif

// Device buffers and launch shape of the mirror stage's second kernel.
struct HrmaScratch
{
   uint32_t *idx_mid;      // photons that hit the paraboloid (capacity entries)
   uint32_t *n_mid;        // its length (written by k_compact)
   uint32_t *rng_pos;      // per photon: random-stream position at the split
   uint32_t reflect_grid;  // blocks of k_hrma_reflect
   int reflect_waves;      // 2, 3 or 4: the instance of k_hrma_reflect launched
   uint32_t *n_mid_copy = nullptr;  // when non-null, also receives the middle list's length
};

// Stage launchers (defined in the stage .hip files).
// cfg: device copy; hcfg: host copy of the same configuration (launch
// geometry and LDS layout are computed from it).  ev, when non-null, holds
// four events recorded around k_hrma_enter (0, 1) and k_hrma_reflect (2, 3).
// relay_out, when non-null: the mirror's survivors go to it for the grating
// or the detector, which store their x, p and flags (the reflection kernel
// stores those of the photons it drops only)
// fresh: the input photons come straight from launch_source with
// entry_follows (unread flags, all zero; the entry writes every photon's
// position, shell and flags)
hipError_t launch_hrma (const mxg_config *cfg, const mxg_config *hcfg, const float *fa,
                        const double *da, Soa s, ListIO io, HrmaScratch hs, uint32_t n_upper,
                        double source_distance, unsigned long long seed, uint32_t source_id,
                        const Relay *relay_out, bool fresh, hipStream_t st, hipEvent_t *ev);
// k_hrma_reflect alone, from io.idx_in (photons that hit the paraboloid,
// stream positions in rng_pos) to io's survivor segments (io.grid blocks;
// the caller compacts them); waves: the instance launched (2, 3 or 4)
hipError_t launch_hrma_reflect (const mxg_config *cfg, const mxg_config *hcfg, const float *fa,
                                const double *da, Soa s, ListIO io, const uint32_t *rng_pos, int waves,
                                unsigned long long seed, uint32_t source_id, const Relay *relay_out,
                                hipStream_t st);
// The blocks of k_grating that one CU holds at once for configuration hcfg
// (its waves per SIMD, fewer if its LDS tables do not fit that many), 8 if
// the runtime cannot say.  A grid of whole rounds of them measured fastest
// (HETG + ACIS-S: 3 blocks per CU 0.36 ms, 4 0.47 ms, 8 0.40 ms).
uint32_t grating_blocks_per_cu (const mxg_config *hcfg);
// hcfg: the host copy of cfg (for the layout of the tables the kernel
// stages in LDS); letg: the instance compiled for the LETG pass; relay_in,
// when non-null:
// the input list is the mirror's with this relay, whose rows the kernel
// reads instead of the store's; relay_out, when non-null: the survivors go
// to it for launch_detect, which stores their x, p and flags (the kernel
// stores those of the photons it drops only)
hipError_t launch_grating (const mxg_config *cfg, const mxg_config *hcfg, const float *fa, const double *da, Soa s,
                           ListIO io, uint32_t n_upper, unsigned long long seed,
                           uint32_t source_id, bool letg, const Relay *relay_in, const Relay *relay_out,
                           hipStream_t st);
// fef: run the variant with the ACIS FEF step (the configuration read a FEF
// file); kind: the detector family (detector_kind of the configuration)
enum { MXG_DET_ACIS = 0, MXG_DET_HRC_S = 1, MXG_DET_HRC_I = 2 };
inline int detector_kind (const mxg_config &c)
{
   return (c.det.type == MARX_DETECTOR_HRC_S) ? MXG_DET_HRC_S
     : (c.det.type == MARX_DETECTOR_HRC_I) ? MXG_DET_HRC_I : MXG_DET_ACIS;
}
// relay, when non-null: the input list is the one the previous stage left
// with this relay, whose rows the kernel reads instead of the store's
hipError_t launch_detect (const mxg_config *cfg, const float *fa, const double *da, Soa s,
                          ListIO io, uint32_t n_upper, double tstart,
                          unsigned long long seed, uint32_t source_id, bool fef, int kind,
                          const Relay *relay, hipStream_t st);

hipError_t launch_dither_store (const mxg_config *cfg, Soa s, uint32_t n, double start_time,
                                unsigned long long seed, hipStream_t st);
// Photon creation (source.hip): photons 0..n-1 of the store from tag
// tag_start, the exposure cut in *first_over, the active list idx and its
// length *count (and *count_copy when non-null); ticks (n entries) and
// tile_tot (ceil(n / 2048) entries) are scratch; point: the configured
// shape is a point (the instance without the extended shapes);
// entry_follows: launch_hrma with fresh set follows on the created photons
// (the source leaves their position, shell and flags to it).
hipError_t launch_source (const mxg_config *cfg, const float *fa, const double *da, Soa s, uint32_t n,
                          uint32_t tag_start, double exposure_time, double start_time,
                          unsigned long long seed, uint32_t source_id, uint64_t *ticks, uint64_t *tile_tot,
                          uint32_t *first_over, uint32_t *count, uint32_t *count_copy, uint32_t *idx, bool point,
                          bool entry_follows, hipStream_t st);

// Batched point sources (see source.hip): n photons over nsrc sources with
// prefix offsets off[nsrc + 1] and unit directions dir[3 nsrc]; point: the
// configured shape is a point (the instance without the extended shapes);
// entry_follows: launch_hrma with fresh set follows (as for launch_source).
hipError_t launch_batch_source (const mxg_config *cfg, const double *da, Soa s, uint32_t n,
                                const uint32_t *off, uint32_t nsrc, const double *dir,
                                const uint32_t *tag_start, double exposure, double start_time,
                                unsigned long long seed, uint32_t *idx, uint32_t *count,
                                uint32_t grid, bool point, bool entry_follows,
                                hipStream_t st);

// The batched pass as one persistent launch (persist.hip): blocks take
// tiles of `tile` photons (a multiple of kBlock, at most kMaxPersistTile)
// from *work and run every stage on them; the detected photons of the batch
// end in idx_out / *n_out in id order, as after the staged pass.
// stage_count (4 entries) receives the photons into the reflection and out
// of the mirror, grating and detector; seg (n entries), tile_count and
// tile_off (ceil(n / tile) entries each) are scratch.  ev, when non-null:
// two events around the persistent kernel.
constexpr uint32_t kMaxPersistTile = 4096;
// waves per SIMD of the persistent kernel: 3 (168 VGPRs, 250-660 B of
// spills per lane) measured 7 percent faster than 4 (128 VGPRs) and than 2
// (no spills) on the batched ACIS-I pass
#ifndef MXG_PERSIST_WAVES
#define MXG_PERSIST_WAVES 3
#endif
struct PersistLaunch
{
   const mxg_config *cfg;        // device copy
   const mxg_config *hcfg;       // host copy (LDS layout)
   const float *fa;
   const double *da;
   Soa s;
   uint32_t n;
   const uint32_t *off;
   uint32_t nsrc;
   const double *dir;
   const uint32_t *tag_start;
   double exposure, start_time;
   unsigned long long seed;
   uint32_t *rng_pos, *work, *stage_count;
   uint32_t tile, blocks;
   uint32_t *seg, *tile_count, *tile_off, *idx_out, *n_out;
   int grating;                  // 0 none, 1 HETG, 2 LETG
   int kind;                     // detector family, -1 none
   hipEvent_t *ev;
};
hipError_t launch_batch_persistent (const PersistLaunch &p, hipStream_t st);

// The differentiable mirror (diff.hip).
struct DiffOut
{
   double *w, *y, *z;           // n values each
   double *dw, *dy, *dz;        // MXG_DIFF_NPAR x n derivatives each
   uint8_t *shell;
   double *block_w;             // per-block summed weight
};
hipError_t launch_diff_forward (const mxg_config *cfg, const float *fa, const double *da,
                                const mxg_diff_cfg *dcfg, uint32_t n, uint32_t tag_start,
                                unsigned long long seed, const mxg_image_spec *spec, float *image,
                                float *win_partial, float *slice_sum, DiffOut out, const ListIO &enter,
                                uint32_t grid, double *wsum_dev, int tangents, hipStream_t st,
                                hipEvent_t *ev);
hipError_t launch_diff_adjoint (uint32_t n, DiffOut in, const mxg_image_spec *spec, const float *g,
                                double *partial, uint32_t grid, double *grad_dev, hipStream_t st,
                                hipEvent_t *ev);

// The energy order of the store's first n photons (sort.hip): stable
// radix sort of (energy, index) pairs.  With tmp == nullptr only
// *tmp_bytes (rocPRIM's temporary storage) is set.
hipError_t sort_energies (const double *energy, uint32_t n, double *keys_in, uint32_t *vals_in,
                          double *keys_out, uint32_t *vals_out, void *tmp, size_t *tmp_bytes,
                          hipStream_t st);
// out[i] = energy[idx[i]] for i < *n_dev (n_upper an upper bound of *n_dev)
hipError_t gather_energies (const double *energy, const uint32_t *idx, const uint32_t *n_dev, uint32_t n_upper,
                            double *out, hipStream_t st);

} // namespace mxg
