// device.hip -- HBM side of libmarx_gpu: device selection, the configuration
// copy, the SoA photon store, AoS <-> SoA conversion, active-list ping-pong
// and the stage runners called by the host C layer (mxg_device.h).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <stdint.h>

#include "dev_store.h"
#include "mxg_columns.h"
#include "mxg_device.h"

using namespace mxg;

namespace {

char g_err[512];
int g_inited = 0;
hipStream_t g_stream = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
// asynchronous stage timing: events recorded around each stage, read after
// the stream has been synchronised for another reason
constexpr int kSlots = 8;
hipEvent_t g_tev[kSlots][2];
int g_slot_used[kSlots];

// configuration in HBM
mxg_config *d_cfg = nullptr;
mxg_config g_hcfg;                 // host copy of the device configuration
int g_num_cus = 0;
float *d_fa = nullptr;
double *d_da = nullptr;
size_t d_fa_cap = 0, d_da_cap = 0;
unsigned int g_cfg_gen = 0;

// photon store
void *d_store = nullptr;
size_t g_capacity = 0;
Soa g_soa;
unsigned int g_n = 0;              // photons in the store
uint32_t *d_idx[2] = {nullptr, nullptr};
uint32_t *d_idx_mid = nullptr;     // the mirror's middle list
uint32_t *d_seg = nullptr;         // per-block survivor segments of the running stage
uint32_t *d_seg_count = nullptr;   // kMaxSegments entries
uint32_t *d_rng_pos = nullptr;     // the mirror's per-photon stream position
uint32_t *d_counts = nullptr;      // [0],[1] list lengths, [2] source cut, [4] middle list,
                                   // [8..15] per-stage survivor counts, [32] the persistent
                                   // pass's tile counter, [36..39] its stage counts
uint32_t *h_counts = nullptr;      // pinned host copy of d_counts[8..15], read once per call
uint32_t *d_tile_cnt = nullptr;    // the persistent pass's per-tile counts and offsets
size_t g_tile_cap = 0;             // entries of each
int g_cur = 0;                     // list holding the active photons
unsigned int g_upper = 0;          // host-side upper bound of the active count
Marx_Photon_Attr_Type *d_aos = nullptr;
size_t g_aos_cap = 0;
uint64_t *d_ticks = nullptr;       // arrival-time scan (fixed-point gaps)
void *d_relay_mem = nullptr;       // two stage-to-stage relays (capacity rows each)
Relay g_relay[2] = {};
int g_relay_have = -1;             // the relay the last stage wrote for the current list, or -1
// batched point sources
uint32_t *d_batch_off = nullptr;   // nsrc + 1 prefix offsets
double *d_batch_dir = nullptr;     // 3 nsrc directions
uint32_t *d_batch_tag = nullptr;   // tag of the first photon of the launch
unsigned int g_batch_cap = 0, g_batch_nsrc = 0;
// differentiable mirror: per-photon weights, positions and derivatives
void *d_diff = nullptr;
size_t g_diff_cap = 0;
DiffOut g_diff_out;
unsigned int g_diff_n = 0;
mxg_image_spec g_diff_spec;
uint32_t g_diff_grid = 0;
bool g_diff_tangents = false;      // the last forward pass kept derivatives
unsigned long long g_diff_pass = 0; // forward passes made (over every caller)
hipEvent_t g_diff_ev[6];           // around k_diff_enter + k_compact, k_diff_forward, k_diff_adjoint
float g_diff_ms[3] = {0.0f, 0.0f, 0.0f};
uint32_t *d_diff_list = nullptr;   // photons past the paraboloid: segments, list
uint32_t *d_diff_seg_count = nullptr;
uint32_t *d_diff_u32 = nullptr;    // [0] photons of the pass, [1] listed
uint32_t g_diff_listed = 0;
float *d_diff_win = nullptr;       // per-block image windows
float *d_diff_slices = nullptr;    // their partial sums (64 slices)
double *d_diff_partial = nullptr;  // per-block gradient partials
double *d_diff_small = nullptr;    // summed weight, gradient
mxg_diff_cfg *d_diff_cfg = nullptr;
// launch schedule and graph capture
// Default grids: whole rounds of the blocks a CU holds at once, where that
// measured fastest (HETG + ACIS-S): the entry (five blocks a CU) at three
// rounds, 0.240 ms against 0.264 at 8 blocks per CU; the reflection (two to
// four) at 12 blocks, 0.668 against 0.676; the grating one round (0).
mxgd_schedule g_sched = {{8, 15, 12, 0, 8}, MXG_REFLECT_WAVES_DEFAULT, 0, 2};
uint32_t g_grating_bpc = 8;        // one round of k_grating blocks per CU (grating_blocks_per_cu)
int g_capturing = 0;
uint64_t *d_tile_tot = nullptr;
// bumped whenever a buffer that a captured launch sequence may hold is
// reallocated (photon store, lists, batch sources, configuration arenas):
// a graph captured in an older epoch holds freed addresses
unsigned int g_alloc_epoch = 0;

int fail (const char *fmt, ...)
{
   va_list ap;
   va_start (ap, fmt);
   vsnprintf (g_err, sizeof (g_err), fmt, ap);
   va_end (ap);
   fprintf (stderr, "***marx_gpu (device): %s\n", g_err);
   return -1;
}

#define HIPCHECK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) \
   return fail ("%s failed: %s", #call, hipGetErrorString (e_)); } while (0)

size_t align256 (size_t n) { return (n + 255) & ~(size_t) 255; }

__global__ void k_set_u32 (uint32_t *p, uint32_t v) { *p = v; }

__global__ void k_iota (uint32_t *idx, uint32_t n)
{
   uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) idx[i] = i;
}

__global__ __launch_bounds__ (kBlock)
void k_aos_to_soa (const Marx_Photon_Attr_Type *__restrict__ a, Soa s, uint32_t n)
{
   uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const Marx_Photon_Attr_Type &at = a[i];
   s.energy[i] = at.energy;
   s.x0[i] = at.x.x; s.x1[i] = at.x.y; s.x2[i] = at.x.z;
   s.p0[i] = at.p.x; s.p1[i] = at.p.y; s.p2[i] = at.p.z;
   s.t[i] = at.arrival_time;
   s.flags[i] = at.flags;
   s.ypix[i] = at.y_pixel; s.zpix[i] = at.z_pixel;
   s.upix[i] = at.u_pixel; s.vpix[i] = at.v_pixel;
   s.dith[0][i] = at.dither_state.ra; s.dith[1][i] = at.dither_state.dec;
   s.dith[2][i] = at.dither_state.roll; s.dith[3][i] = at.dither_state.dy;
   s.dith[4][i] = at.dither_state.dz; s.dith[5][i] = at.dither_state.dtheta;
   s.pi[i] = at.pi;
   s.pha[i] = at.pulse_height;
   s.shell[i] = (uint8_t) at.mirror_shell;
   s.ccd[i] = at.ccd_num; s.region[i] = at.detector_region; s.order[i] = at.order;
   for (int k = 0; k < 4; k++) s.sup[k][i] = at.support_orders[k];
   s.tag[i] = at.tag;
}

__global__ __launch_bounds__ (kBlock)
void k_soa_to_aos (Marx_Photon_Attr_Type *__restrict__ a, Soa s, uint32_t n)
{
   static_assert (sizeof (Marx_Photon_Attr_Type) % 8 == 0, "record is whole 8-byte words");
   constexpr int kWords = sizeof (Marx_Photon_Attr_Type) / 8;
   uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   // built in zeroed words and stored word by word, so that the record's
   // padding bytes are zero (a struct copy may leave them indeterminate)
   union { Marx_Photon_Attr_Type at; uint64_t w[kWords]; } u;
   for (int k = 0; k < kWords; k++) u.w[k] = 0;
   Marx_Photon_Attr_Type &at = u.at;
   at.energy = s.energy[i];
   at.x.x = s.x0[i]; at.x.y = s.x1[i]; at.x.z = s.x2[i];
   at.p.x = s.p0[i]; at.p.y = s.p1[i]; at.p.z = s.p2[i];
   at.arrival_time = s.t[i];
   at.flags = s.flags[i];
   at.y_pixel = s.ypix[i]; at.z_pixel = s.zpix[i];
   at.u_pixel = s.upix[i]; at.v_pixel = s.vpix[i];
   at.dither_state.ra = s.dith[0][i]; at.dither_state.dec = s.dith[1][i];
   at.dither_state.roll = s.dith[2][i]; at.dither_state.dy = s.dith[3][i];
   at.dither_state.dz = s.dith[4][i]; at.dither_state.dtheta = s.dith[5][i];
   at.pi = s.pi[i];
   at.pulse_height = s.pha[i];
   at.mirror_shell = s.shell[i];
   at.ccd_num = s.ccd[i]; at.detector_region = s.region[i]; at.order = s.order[i];
   for (int k = 0; k < 4; k++) at.support_orders[k] = s.sup[k][i];
   at.tag = s.tag[i];
   uint64_t *dst = (uint64_t *) (a + i);
   for (int k = 0; k < kWords; k++) dst[k] = u.w[k];
}

int ensure_capacity (size_t n)
{
   if (n <= g_capacity) return 0;
   size_t cap = g_capacity ? g_capacity : 1024;
   while (cap < n) cap *= 2;
   // bytes per photon of all columns, each column 256-byte aligned
   size_t sz[] = {8, 8, 8, 8, 8, 8, 8, 8, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 2, 1, 1, 1, 1, 1, 1, 1, 1, 4};
   size_t total = 0;
   for (size_t k = 0; k < sizeof (sz) / sizeof (sz[0]); k++) total += align256 (sz[k] * cap);
   void *mem = nullptr;
   if (d_store != nullptr) { (void) hipStreamSynchronize (g_stream); (void) hipFree (d_store); d_store = nullptr; }
   for (int k = 0; k < 2; k++) if (d_idx[k]) { (void) hipFree (d_idx[k]); d_idx[k] = nullptr; }
   if (d_ticks) { (void) hipFree (d_ticks); d_ticks = nullptr; }
   if (d_tile_tot) { (void) hipFree (d_tile_tot); d_tile_tot = nullptr; }
   if (d_idx_mid) { (void) hipFree (d_idx_mid); d_idx_mid = nullptr; }
   if (d_seg) { (void) hipFree (d_seg); d_seg = nullptr; }
   if (d_rng_pos) { (void) hipFree (d_rng_pos); d_rng_pos = nullptr; }
   if (d_tile_cnt) { (void) hipFree (d_tile_cnt); d_tile_cnt = nullptr; }
   if (d_relay_mem) { (void) hipFree (d_relay_mem); d_relay_mem = nullptr; }
   g_capacity = 0;
   g_tile_cap = 0;
   HIPCHECK (hipMalloc (&mem, total));
   HIPCHECK (hipMalloc (&d_ticks, cap * sizeof (uint64_t)));
   HIPCHECK (hipMalloc (&d_tile_tot, (cap / 2048 + 2) * sizeof (uint64_t)));
   HIPCHECK (hipMalloc (&d_idx[0], cap * sizeof (uint32_t)));
   HIPCHECK (hipMalloc (&d_idx[1], cap * sizeof (uint32_t)));
   HIPCHECK (hipMalloc (&d_idx_mid, cap * sizeof (uint32_t)));
   HIPCHECK (hipMalloc (&d_seg, cap * sizeof (uint32_t)));
   HIPCHECK (hipMalloc (&d_rng_pos, cap * sizeof (uint32_t)));
   // tiles of at least kBlock photons
   HIPCHECK (hipMalloc (&d_tile_cnt, 2 * (cap / kBlock + 1) * sizeof (uint32_t)));
   g_tile_cap = cap / kBlock + 1;
   {
      // each relay: seven double, three 32-bit and one byte column
      const size_t dcol = align256 (cap * sizeof (double)), ucol = align256 (cap * sizeof (uint32_t));
      const size_t bcol = align256 (cap);
      const size_t each = 7 * dcol + 3 * ucol + bcol;
      HIPCHECK (hipMalloc (&d_relay_mem, 2 * each));
      for (int k = 0; k < 2; k++)
        {
           Relay &r = g_relay[k];
           char *q = (char *) d_relay_mem + k * each;
           double **dc[7] = {&r.x0, &r.x1, &r.x2, &r.p0, &r.p1, &r.p2, &r.energy};
           for (int j = 0; j < 7; j++) { *dc[j] = (double *) q; q += dcol; }
           r.tag = (uint32_t *) q; q += ucol;
           r.flags = (uint32_t *) q; q += ucol;
           r.pos = (uint32_t *) q; q += ucol;
           r.shell = (uint8_t *) q;
        }
   }
   g_relay_have = -1;
   d_store = mem;
   char *p = (char *) mem;
   size_t k = 0;
   auto take = [&] (void) -> void * { void *r = p; p += align256 (sz[k++] * cap); return r; };
   g_soa.energy = (double *) take ();
   g_soa.x0 = (double *) take (); g_soa.x1 = (double *) take (); g_soa.x2 = (double *) take ();
   g_soa.p0 = (double *) take (); g_soa.p1 = (double *) take (); g_soa.p2 = (double *) take ();
   g_soa.t = (double *) take ();
   g_soa.flags = (uint32_t *) take ();
   g_soa.ypix = (float *) take (); g_soa.zpix = (float *) take ();
   g_soa.upix = (float *) take (); g_soa.vpix = (float *) take ();
   for (int j = 0; j < 6; j++) g_soa.dith[j] = (float *) take ();
   g_soa.pi = (float *) take ();
   g_soa.pha = (int16_t *) take ();
   g_soa.shell = (uint8_t *) take ();
   g_soa.ccd = (int8_t *) take (); g_soa.region = (int8_t *) take (); g_soa.order = (int8_t *) take ();
   for (int j = 0; j < 4; j++) g_soa.sup[j] = (int8_t *) take ();
   g_soa.tag = (uint32_t *) take ();
   g_capacity = cap;
   g_alloc_epoch++;
   return 0;
}

int ensure_aos (size_t n)
{
   if (n <= g_aos_cap) return 0;
   if (d_aos) { (void) hipStreamSynchronize (g_stream); (void) hipFree (d_aos); d_aos = nullptr; }
   g_aos_cap = 0;
   HIPCHECK (hipMalloc (&d_aos, n * sizeof (Marx_Photon_Attr_Type)));
   g_aos_cap = n;
   return 0;
}

// The schedule's blocks per CU of stage kernel k; 0 (the grating's
// default): one round of the blocks a CU holds at once.
uint32_t sched_blocks_per_cu (int k)
{
   const uint32_t b = g_sched.blocks_per_cu[k];
   return b ? b : g_grating_bpc;
}

// Blocks of stage kernel k over at most n_upper active photons: the
// schedule's blocks per CU times the CUs (each block owning an equal range),
// fewer when there are fewer photons than that.
uint32_t stage_grid (uint32_t n_upper, int k)
{
   uint32_t blocks = (n_upper + kBlock - 1) / kBlock;
   uint32_t cap = (uint32_t) (g_num_cus > 0 ? g_num_cus : 256) * sched_blocks_per_cu (k);
   if (cap > kMaxSegments) cap = kMaxSegments;
   if (blocks > cap) blocks = cap;
   return blocks ? blocks : 1u;
}

// count slot `slot` (d_counts[8 + slot]), or null for none
uint32_t *slot_count (int slot)
{
   return ((slot >= 0) && (slot < kSlots)) ? d_counts + 8 + slot : nullptr;
}

// Stage k's lists: the active list in, the other one out; k_compact also
// writes the survivors' number to count slot `slot` (when there is one).
ListIO list_io (int k, int slot = -1)
{
   ListIO io;
   io.idx_in = d_idx[g_cur];
   io.n_in = d_counts + g_cur;
   io.idx_out = d_idx[1 - g_cur];
   io.n_out = d_counts + (1 - g_cur);
   io.seg = d_seg;
   io.seg_count = d_seg_count;
   io.grid = stage_grid (g_upper, k);
   io.n_copy = slot_count (slot);
   return io;
}

__global__ void k_copy_u32 (uint32_t *dst, const uint32_t *src) { *dst = *src; }

int begin_stage (int slot)
{
   if ((slot >= 0) && (slot < kSlots) && !g_capturing)
     {
        HIPCHECK (hipEventRecord (g_tev[slot][0], g_stream));
        g_slot_used[slot] = 1;
     }
   return 0;
}

// A stage whose kernels do not run (no photon can be active): its output
// list and count slot are empty.  (A stage that runs gets both from
// k_compact.)
int skip_stage (const ListIO &io)
{
   hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, io.n_out, 0u);
   if (io.n_copy) hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, io.n_copy, 0u);
   HIPCHECK (hipGetLastError ());
   return 0;
}

int end_stage (int slot, float *ms)
{
   if ((slot >= 0) && (slot < kSlots) && !g_capturing)
     HIPCHECK (hipEventRecord (g_tev[slot][1], g_stream));
   g_cur = 1 - g_cur;
   if (ms != nullptr)
     {
        HIPCHECK (hipEventRecord (g_ev1, g_stream));
        HIPCHECK (hipEventSynchronize (g_ev1));
        HIPCHECK (hipEventElapsedTime (ms, g_ev0, g_ev1));
     }
   return 0;
}

} // namespace

namespace mxg {

// Concatenates the survivor segments of a stage kernel of gridDim.x blocks
// (block b of this kernel copies segment b after the sum of the counts
// before it) and writes the total to n_out.
__global__ __launch_bounds__ (kBlock)
void k_compact (const uint32_t *__restrict__ seg, const uint32_t *__restrict__ seg_count,
                const uint32_t *__restrict__ n_in, uint32_t *__restrict__ out,
                uint32_t *__restrict__ n_out, uint32_t *__restrict__ pos_out,
                uint32_t *__restrict__ n_copy)
{
   __shared__ uint32_t part[kBlock / 64];
   const uint32_t b = blockIdx.x;
   uint32_t sum = 0;
   for (uint32_t k = threadIdx.x; k < b; k += kBlock) sum += seg_count[k];
   for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor (sum, off);
   if (__lane_id () == 0) part[threadIdx.x >> 6] = sum;
   __syncthreads ();
   uint32_t prefix = 0;
   for (int w = 0; w < kBlock / 64; w++) prefix += part[w];
   const uint32_t cnt = seg_count[b];
   const BlockRange r = block_range (*n_in, b, gridDim.x);
   for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) out[prefix + k] = seg[r.lo + k];
   if (pos_out)
     for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) pos_out[prefix + k] = r.lo + k;
   if ((b == gridDim.x - 1) && (threadIdx.x == 0))
     {
        *n_out = prefix + cnt;
        if (n_copy) *n_copy = prefix + cnt;
     }
}

hipError_t launch_compact (const ListIO &io, hipStream_t st, uint32_t *pos_out)
{
   hipLaunchKernelGGL (k_compact, dim3 (io.grid), dim3 (kBlock), 0, st, io.seg, io.seg_count,
                       io.n_in, io.idx_out, io.n_out, pos_out, io.n_copy);
   return hipGetLastError ();
}

} // namespace mxg

extern "C" {

const char *mxgd_last_error (void) { return g_err; }

/* ---- the allocation-free trace (mxg_trace_device_noalloc) --------------------------- */

namespace {

// Fixed-size device memory for the allocation-free trace: the per-photon
// lists, counters and scan buffers of one chunk of at most kNoAllocPhotons
// photons (29 MB).
constexpr size_t kNoAllocPhotons = MXG_NOALLOC_MAX_PHOTONS;
constexpr size_t na256 (size_t n) { return (n + 255) & ~(size_t) 255; }
constexpr size_t kNaCounts = 0;
constexpr size_t kNaSegCount = kNaCounts + na256 (64 * sizeof (uint32_t));
constexpr size_t kNaIdx0 = kNaSegCount + na256 (kMaxSegments * sizeof (uint32_t));
constexpr size_t kNaIdx1 = kNaIdx0 + na256 (kNoAllocPhotons * sizeof (uint32_t));
constexpr size_t kNaMid = kNaIdx1 + na256 (kNoAllocPhotons * sizeof (uint32_t));
constexpr size_t kNaSeg = kNaMid + na256 (kNoAllocPhotons * sizeof (uint32_t));
constexpr size_t kNaRng = kNaSeg + na256 (kNoAllocPhotons * sizeof (uint32_t));
constexpr size_t kNaTicks = kNaRng + na256 (kNoAllocPhotons * sizeof (uint32_t));
constexpr size_t kNaTiles = kNaTicks + na256 (kNoAllocPhotons * sizeof (uint64_t));
constexpr size_t kNaBytes = kNaTiles + na256 ((kNoAllocPhotons / 2048 + 2) * sizeof (uint64_t));

constexpr int kMaxDevices = 16;
unsigned char *g_scratch_ptr[kMaxDevices];      // device address of `scratch`, per device

} // namespace

// The scratch memory itself: one copy per GPU, held until the process ends.
// Host code never uses this name as an address (it is a host-side stand-in):
// g_scratch_ptr holds each device's address from hipGetSymbolAddress.
static __device__ __align__ (256) unsigned char scratch[kNaBytes];

namespace {

// At set-up, once per device with that device current.
int scratch_setup (int dev)
{
   if ((dev < 0) || (dev >= kMaxDevices))
     return fail ("device %d out of the %d supported", dev, kMaxDevices);
   if (g_scratch_ptr[dev] != nullptr) return 0;
   void *ptr = nullptr;
   hipError_t e = hipGetSymbolAddress (&ptr, HIP_SYMBOL (scratch));
   if ((e != hipSuccess) || (ptr == nullptr))
     return fail ("hipGetSymbolAddress of the trace scratch failed: %s", hipGetErrorString (e));
   g_scratch_ptr[dev] = (unsigned char *) ptr;
   return 0;
}

} // namespace

int mxgd_init (void)
{
   if (g_inited) return 0;
   int n = 0;
   hipError_t e = hipGetDeviceCount (&n);
   if ((e != hipSuccess) || (n == 0))
     return fail ("no HIP device available (%s)", hipGetErrorString (e));
   int dev = 0;
   HIPCHECK (hipGetDevice (&dev));
   hipDeviceProp_t prop;
   HIPCHECK (hipGetDeviceProperties (&prop, dev));
   if (strncmp (prop.gcnArchName, "gfx950", 6) != 0)
     return fail ("device %d is %s; this library is built for gfx950", dev, prop.gcnArchName);
   g_num_cus = prop.multiProcessorCount;
   HIPCHECK (hipStreamCreateWithFlags (&g_stream, hipStreamNonBlocking));
   HIPCHECK (hipEventCreate (&g_ev0));
   HIPCHECK (hipEventCreate (&g_ev1));
   for (int k = 0; k < kSlots; k++)
     {
        HIPCHECK (hipEventCreate (&g_tev[k][0]));
        HIPCHECK (hipEventCreate (&g_tev[k][1]));
        g_slot_used[k] = 0;
     }
   HIPCHECK (hipMalloc (&d_counts, 64 * sizeof (uint32_t)));
   HIPCHECK (hipMalloc (&d_seg_count, kMaxSegments * sizeof (uint32_t)));
   HIPCHECK (hipMemset (d_counts, 0, 64 * sizeof (uint32_t)));
   HIPCHECK (hipHostMalloc ((void **) &h_counts, kSlots * sizeof (uint32_t), hipHostMallocDefault));
   if (scratch_setup (dev)) return -1;
   g_inited = 1;
   return 0;
}

int mxgd_device_info (char *name, unsigned int namelen, int *num_cus)
{
   if (mxgd_init ()) return -1;
   int dev = 0;
   hipDeviceProp_t prop;
   HIPCHECK (hipGetDevice (&dev));
   HIPCHECK (hipGetDeviceProperties (&prop, dev));
   if (name && namelen) { strncpy (name, prop.gcnArchName, namelen - 1); name[namelen - 1] = 0; }
   if (num_cus) *num_cus = prop.multiProcessorCount;
   return 0;
}

int mxgd_sync_config (const mxg_config *cfg, const float *fa, unsigned int nf,
                      const double *da, unsigned int nd, unsigned int generation)
{
   if (mxgd_init ()) return -1;
   if ((generation == g_cfg_gen) && d_cfg) return 0;
   HIPCHECK (hipStreamSynchronize (g_stream));
   if (!d_cfg) HIPCHECK (hipMalloc (&d_cfg, sizeof (mxg_config)));
   if ((nf + 1 > d_fa_cap) || (nd + 1 > d_da_cap))
     g_alloc_epoch++;
   if (nf + 1 > d_fa_cap)
     {
        if (d_fa) HIPCHECK (hipFree (d_fa));
        d_fa = nullptr;
        d_fa_cap = 0;
        HIPCHECK (hipMalloc (&d_fa, (nf + 1) * sizeof (float)));
        d_fa_cap = nf + 1;
     }
   if (nd + 1 > d_da_cap)
     {
        if (d_da) HIPCHECK (hipFree (d_da));
        d_da = nullptr;
        d_da_cap = 0;
        HIPCHECK (hipMalloc (&d_da, (nd + 1) * sizeof (double)));
        d_da_cap = nd + 1;
     }
   HIPCHECK (hipMemcpy (d_cfg, cfg, sizeof (mxg_config), hipMemcpyHostToDevice));
   memcpy (&g_hcfg, cfg, sizeof (mxg_config));
   g_grating_bpc = grating_blocks_per_cu (&g_hcfg);
   if (nf) HIPCHECK (hipMemcpy (d_fa, fa, nf * sizeof (float), hipMemcpyHostToDevice));
   if (nd) HIPCHECK (hipMemcpy (d_da, da, nd * sizeof (double), hipMemcpyHostToDevice));
   g_cfg_gen = generation;
   return 0;
}

int mxgd_upload_aos (const Marx_Photon_Attr_Type *at, unsigned int n,
                     const unsigned int *active, unsigned int nactive)
{
   g_relay_have = -1;
   if (mxgd_init ()) return -1;
   if (nactive > n) return fail ("active list longer than the photon list");
   if (ensure_capacity (n ? n : 1) || ensure_aos (n ? n : 1)) return -1;
   if (n)
     {
        HIPCHECK (hipMemcpyAsync (d_aos, at, (size_t) n * sizeof (Marx_Photon_Attr_Type),
                                  hipMemcpyHostToDevice, g_stream));
        hipLaunchKernelGGL (k_aos_to_soa, dim3 ((n + kBlock - 1) / kBlock), dim3 (kBlock), 0,
                            g_stream, d_aos, g_soa, (uint32_t) n);
        HIPCHECK (hipGetLastError ());
     }
   g_cur = 0;
   if (nactive)
     HIPCHECK (hipMemcpyAsync (d_idx[0], active, (size_t) nactive * sizeof (uint32_t),
                               hipMemcpyHostToDevice, g_stream));
   hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, d_counts, (uint32_t) nactive);
   HIPCHECK (hipGetLastError ());
   g_n = n;
   g_upper = nactive;
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

int mxgd_download_aos (Marx_Photon_Attr_Type *at, unsigned int n)
{
   if (mxgd_init ()) return -1;
   if (n > g_n) return fail ("download of %u photons from a store of %u", n, g_n);
   if (n == 0) return 0;
   if (ensure_aos (n)) return -1;
   hipLaunchKernelGGL (k_soa_to_aos, dim3 ((n + kBlock - 1) / kBlock), dim3 (kBlock), 0,
                       g_stream, d_aos, g_soa, (uint32_t) n);
   HIPCHECK (hipGetLastError ());
   HIPCHECK (hipMemcpyAsync (at, d_aos, (size_t) n * sizeof (Marx_Photon_Attr_Type),
                             hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

int mxgd_dither_store (double start_time, unsigned long long seed)
{
   if (mxgd_init ()) return -1;
   if (!d_cfg) return fail ("configuration not on the device");
   HIPCHECK (launch_dither_store (d_cfg, g_soa, g_n, start_time, seed, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

unsigned int mxgd_store_size (void) { return g_n; }

int mxgd_get_soa (mxg_soa_view *v)
{
   if (mxgd_init ()) return -1;
   memset (v, 0, sizeof (*v));
   v->n = g_n;
   v->capacity = (unsigned int) g_capacity;
   if (g_capacity == 0) return 0;
   v->energy = g_soa.energy;
   v->x[0] = g_soa.x0; v->x[1] = g_soa.x1; v->x[2] = g_soa.x2;
   v->p[0] = g_soa.p0; v->p[1] = g_soa.p1; v->p[2] = g_soa.p2;
   v->arrival_time = g_soa.t;
   v->flags = g_soa.flags;
   v->y_pixel = g_soa.ypix; v->z_pixel = g_soa.zpix; v->u_pixel = g_soa.upix; v->v_pixel = g_soa.vpix;
   for (int k = 0; k < 6; k++) v->dither[k] = g_soa.dith[k];
   v->pi = g_soa.pi;
   v->pulse_height = g_soa.pha;
   v->mirror_shell = g_soa.shell;
   v->ccd_num = g_soa.ccd; v->detector_region = g_soa.region; v->order = g_soa.order;
   for (int k = 0; k < 4; k++) v->support_orders[k] = g_soa.sup[k];
   v->tag = g_soa.tag;
   return 0;
}

// The relays of a stage: in, the one the previous stage left (fails when
// there is none); out, the other one.  After the stage, the relay it wrote
// (or none) is the current list's.
int stage_relays (int relay_in, int relay_out, const Relay **in, const Relay **out, int *out_k)
{
   *in = nullptr;
   *out = nullptr;
   *out_k = -1;
   int k_in = -1;
   if (relay_in)
     {
        if (g_relay_have < 0) return fail ("no relay was left for this stage by the stage before");
        k_in = g_relay_have;
        *in = &g_relay[k_in];
     }
   if (relay_out)
     {
        *out_k = (k_in == 0) ? 1 : 0;
        *out = &g_relay[*out_k];
     }
   return 0;
}

int mxgd_run_mirror (double source_distance, unsigned long long seed, int slot, float *ms, int relay_out,
                     int fresh)
{
   if (!d_cfg) return fail ("configuration not on the device");
   const Relay *rin, *rout;
   int out_k;
   if (stage_relays (0, relay_out, &rin, &rout, &out_k)) return -1;
   if (ms) HIPCHECK (hipEventRecord (g_ev0, g_stream));
   if (begin_stage (slot)) return -1;
   bool timed = (slot >= 0) && (g_upper > 0);
   HrmaScratch hs = {d_idx_mid, d_counts + 4, d_rng_pos, stage_grid (g_upper, MXGD_K_REFLECT),
                     (int) g_sched.reflect_waves};
   // photons entering each kernel: the active list, the middle list (the
   // latter's length copied by k_compact)
   if (timed) hs.n_mid_copy = slot_count (MXGD_SLOT_HRMA_REFLECT);
   hipEvent_t kev[4] = {g_tev[MXGD_SLOT_HRMA_ENTER][0], g_tev[MXGD_SLOT_HRMA_ENTER][1],
                        g_tev[MXGD_SLOT_HRMA_REFLECT][0], g_tev[MXGD_SLOT_HRMA_REFLECT][1]};
   const ListIO io = list_io (MXGD_K_ENTER, slot);
   if (g_upper == 0)
     {
        if (skip_stage (io)) return -1;
     }
   else
     HIPCHECK (launch_hrma (d_cfg, &g_hcfg, d_fa, d_da, g_soa, io, hs, g_upper, source_distance, seed, 0, rout,
                            fresh != 0, g_stream, (timed && !g_capturing) ? kev : nullptr));
   g_relay_have = out_k;
   if (timed)
     {
        if (!g_capturing)
          g_slot_used[MXGD_SLOT_HRMA_ENTER] = g_slot_used[MXGD_SLOT_HRMA_REFLECT] = 1;
        hipLaunchKernelGGL (k_copy_u32, dim3 (1), dim3 (1), 0, g_stream,
                            d_counts + 8 + MXGD_SLOT_HRMA_ENTER, d_counts + g_cur);
        HIPCHECK (hipGetLastError ());
     }
   return end_stage (slot, ms);
}

int mxgd_run_grating (unsigned long long seed, int slot, float *ms, int relay_in, int relay_out)
{
   if (!d_cfg) return fail ("configuration not on the device");
   const Relay *rin, *rout;
   int out_k;
   if (stage_relays (relay_in, relay_out, &rin, &rout, &out_k)) return -1;
   if (ms) HIPCHECK (hipEventRecord (g_ev0, g_stream));
   if (begin_stage (slot)) return -1;
   const ListIO io = list_io (MXGD_K_GRATING, slot);
   if (g_upper == 0)
     {
        if (skip_stage (io)) return -1;
     }
   else
     HIPCHECK (launch_grating (d_cfg, &g_hcfg, d_fa, d_da, g_soa, io, g_upper, seed, 0,
                               g_hcfg.grating.type == MARX_GRATING_LETG, rin, rout, g_stream));
   g_relay_have = out_k;
   return end_stage (slot, ms);
}

int mxgd_run_detect (double tstart, unsigned long long seed, int slot, float *ms, int relay_in)
{
   if (!d_cfg) return fail ("configuration not on the device");
   const Relay *rin, *rout;
   int out_k;
   if (stage_relays (relay_in, 0, &rin, &rout, &out_k)) return -1;
   if (ms) HIPCHECK (hipEventRecord (g_ev0, g_stream));
   if (begin_stage (slot)) return -1;
   const ListIO io = list_io (MXGD_K_DETECT, slot);
   if (g_upper == 0)
     {
        if (skip_stage (io)) return -1;
     }
   else
     HIPCHECK (launch_detect (d_cfg, d_fa, d_da, g_soa, io, g_upper, tstart, seed, 0, g_hcfg.det.fef_on != 0,
                              detector_kind (g_hcfg), rin, g_stream));
   g_relay_have = -1;
   return end_stage (slot, ms);
}

int mxgd_active_count (unsigned int *n)
{
   uint32_t v = 0;
   HIPCHECK (hipMemcpyAsync (&v, d_counts + g_cur, sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   *n = v;
   return 0;
}

int mxgd_active_list (unsigned int *idx, unsigned int n)
{
   if (n == 0) return 0;
   HIPCHECK (hipMemcpyAsync (idx, d_idx[g_cur], (size_t) n * sizeof (uint32_t),
                             hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

int mxgd_read_timings (float *ms, unsigned int *counts, int nslots)
{
   // One wait for the stages and the counts: the copy into pinned memory is
   // queued behind the last stage, so no second round trip follows the sync.
   HIPCHECK (hipMemcpyAsync (h_counts, d_counts + 8, kSlots * sizeof (uint32_t),
                             hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   for (int k = 0; k < nslots && k < kSlots; k++)
     {
        ms[k] = 0.0f;
        if (counts) counts[k] = h_counts[k];
        if (g_slot_used[k])
          HIPCHECK (hipEventElapsedTime (ms + k, g_tev[k][0], g_tev[k][1]));
        g_slot_used[k] = 0;
     }
   return 0;
}

int mxgd_synchronize (void)
{
   if (!g_inited) return 0;
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

int mxgd_run_source (unsigned int n, unsigned int tag_start, double start_time,
                     double exposure_time, unsigned long long seed, unsigned int source_id,
                     int slot, unsigned int *n_out, double *last_time, float *ms)
{
   g_relay_have = -1;
   if (!d_cfg) return fail ("configuration not on the device");
   if (n > (1u << 26)) return fail ("chunk of %u photons exceeds the 2^26 limit of one launch", n);
   if (ensure_capacity (n ? n : 1)) return -1;
   if (ms) HIPCHECK (hipEventRecord (g_ev0, g_stream));
   g_cur = 0;
   if (slot >= 0 && slot < kSlots)
     {
        HIPCHECK (hipEventRecord (g_tev[slot][0], g_stream));
        g_slot_used[slot] = 1;
     }
   HIPCHECK (launch_source (d_cfg, d_fa, d_da, g_soa, n, tag_start, exposure_time, start_time, seed,
                            source_id, d_ticks, d_tile_tot, d_counts + 2, d_counts, nullptr, d_idx[0],
                            g_hcfg.src.type == MXG_SOURCE_POINT, false, g_stream));
   if (slot >= 0 && slot < kSlots)
     HIPCHECK (hipEventRecord (g_tev[slot][1], g_stream));
   if (ms)
     {
        HIPCHECK (hipEventRecord (g_ev1, g_stream));
        HIPCHECK (hipEventSynchronize (g_ev1));
        HIPCHECK (hipEventElapsedTime (ms, g_ev0, g_ev1));
     }
   uint32_t m = 0;
   HIPCHECK (hipMemcpyAsync (&m, d_counts, sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   double t = 0.0;
   if (m > 0)
     {
        HIPCHECK (hipMemcpyAsync (&t, g_soa.t + (m - 1), sizeof (double), hipMemcpyDeviceToHost, g_stream));
        HIPCHECK (hipStreamSynchronize (g_stream));
     }
   g_n = m;
   g_upper = m;
   if (n_out) *n_out = m;
   if (last_time) *last_time = t;
   return 0;
}

int mxgd_launch_source (unsigned int n, unsigned int tag_start, double start_time, double exposure_time,
                        unsigned long long seed, unsigned int source_id, int slot, int entry_follows)
{
   g_relay_have = -1;
   if (!d_cfg) return fail ("configuration not on the device");
   if (n > (1u << 26)) return fail ("chunk of %u photons exceeds the 2^26 limit of one launch", n);
   if (ensure_capacity (n ? n : 1)) return -1;
   g_cur = 0;
   if ((slot >= 0) && (slot < kSlots) && !g_capturing)
     {
        HIPCHECK (hipEventRecord (g_tev[slot][0], g_stream));
        g_slot_used[slot] = 1;
     }
   HIPCHECK (launch_source (d_cfg, d_fa, d_da, g_soa, n, tag_start, exposure_time, start_time, seed,
                            source_id, d_ticks, d_tile_tot, d_counts + 2, d_counts, slot_count (slot), d_idx[0],
                            g_hcfg.src.type == MXG_SOURCE_POINT, entry_follows != 0, g_stream));
   if ((slot >= 0) && (slot < kSlots) && !g_capturing)
     HIPCHECK (hipEventRecord (g_tev[slot][1], g_stream));
   // upper bounds until the created number is read back
   g_n = n;
   g_upper = n;
   return 0;
}

int mxgd_set_created (unsigned int m)
{
   if (m > g_n) return fail ("%u photons created in a store of %u", m, g_n);
   g_n = m;
   return 0;
}

/* ---- differentiable mirror ------------------------------------------------------ */

namespace {

constexpr int kDiffWin = 64 * 64;       // floats of one block's image window
constexpr int kDiffGrad = MXG_NUM_SHELLS * MXG_DIFF_NPAR;

int diff_buffers (size_t n, uint32_t grid)
{
   if ((n > g_diff_cap) || (d_diff_win == nullptr))
     {
        size_t cap = g_diff_cap ? g_diff_cap : 1024;
        while (cap < n) cap *= 2;
        if (d_diff) { HIPCHECK (hipStreamSynchronize (g_stream)); (void) hipFree (d_diff); d_diff = nullptr; }
        g_diff_cap = 0;
        // 3 + 3 * MXG_DIFF_NPAR doubles, one byte and two list words per photon
        size_t dbl = align256 (cap * sizeof (double));
        size_t u32 = align256 (cap * sizeof (uint32_t));
        size_t total = (3 + 3 * MXG_DIFF_NPAR) * dbl + align256 (cap) + 2 * u32;
        HIPCHECK (hipMalloc (&d_diff, total));
        char *q = (char *) d_diff;
        g_diff_out.w = (double *) q; q += dbl;
        g_diff_out.y = (double *) q; q += dbl;
        g_diff_out.z = (double *) q; q += dbl;
        g_diff_out.dw = (double *) q; q += MXG_DIFF_NPAR * dbl;
        g_diff_out.dy = (double *) q; q += MXG_DIFF_NPAR * dbl;
        g_diff_out.dz = (double *) q; q += MXG_DIFF_NPAR * dbl;
        g_diff_out.shell = (uint8_t *) q; q += align256 (cap);
        d_diff_list = (uint32_t *) q;      // segments, then the list
        g_diff_cap = cap;
     }
   if (d_diff_win == nullptr)
     {
        HIPCHECK (hipMalloc (&d_diff_win, (size_t) kMaxSegments * kDiffWin * sizeof (float)));
        HIPCHECK (hipMalloc (&d_diff_slices, (size_t) 64 * kDiffWin * sizeof (float)));
        HIPCHECK (hipMalloc (&d_diff_partial, (size_t) kMaxSegments * kDiffGrad * sizeof (double)));
        HIPCHECK (hipMalloc (&g_diff_out.block_w, (size_t) kMaxSegments * sizeof (double)));
        HIPCHECK (hipMalloc (&d_diff_small, 64 * sizeof (double)));
        HIPCHECK (hipMalloc (&d_diff_cfg, sizeof (mxg_diff_cfg)));
        HIPCHECK (hipMalloc (&d_diff_seg_count, kMaxSegments * sizeof (uint32_t)));
        HIPCHECK (hipMalloc (&d_diff_u32, 4 * sizeof (uint32_t)));
        for (int k = 0; k < 6; k++) HIPCHECK (hipEventCreate (&g_diff_ev[k]));
     }
   (void) grid;
   return 0;
}

} // namespace

int mxgd_diff_forward (const mxg_diff_cfg *d, unsigned int n, unsigned int tag_start,
                       unsigned long long seed, const mxg_image_spec *spec, float *image,
                       double *total_weight, int tangents)
{
   if (!d_cfg) return fail ("configuration not on the device");
   if (n > (1u << 28)) return fail ("%u photons exceed 2^28", n);
   if (spec->npix > 16384) return fail ("image of %u pixels a side exceeds 16384", spec->npix);
   uint32_t grid = stage_grid (n, MXGD_K_REFLECT);
   if (diff_buffers (n ? n : 1, grid)) return -1;
   g_diff_grid = 0;
   HIPCHECK (hipMemcpyAsync (d_diff_cfg, d, sizeof (mxg_diff_cfg), hipMemcpyHostToDevice, g_stream));
   // tangents by the tilts only for a source at infinity (the translations
   // follow at the focal plane), by all five parameters otherwise
   const int nt = (tangents == 0) ? 0 : (g_hcfg.src.distance > 0.0) ? MXG_DIFF_NPAR : 2;
   // k_diff_enter lists the photons that reach the paraboloid
   ListIO enter;
   enter.idx_in = nullptr;
   enter.n_in = d_diff_u32;
   enter.seg = d_diff_list;
   enter.idx_out = d_diff_list + align256 (g_diff_cap * sizeof (uint32_t)) / sizeof (uint32_t);
   enter.n_out = d_diff_u32 + 1;
   enter.seg_count = d_diff_seg_count;
   enter.grid = stage_grid (n, MXGD_K_ENTER);
   hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, d_diff_u32, (uint32_t) n);
   HIPCHECK (hipGetLastError ());
   HIPCHECK (launch_diff_forward (d_cfg, d_fa, d_da, d_diff_cfg, n, tag_start, seed, spec, image,
                                  d_diff_win, d_diff_slices, g_diff_out, enter, grid, d_diff_small, nt,
                                  g_stream, g_diff_ev));
   double w = 0.0;
   HIPCHECK (hipMemcpyAsync (&w, d_diff_small, sizeof (double), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipMemcpyAsync (&g_diff_listed, d_diff_u32 + 1, sizeof (uint32_t), hipMemcpyDeviceToHost,
                             g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   HIPCHECK (hipEventElapsedTime (&g_diff_ms[2], g_diff_ev[0], g_diff_ev[1]));
   HIPCHECK (hipEventElapsedTime (&g_diff_ms[0], g_diff_ev[2], g_diff_ev[3]));
   if (total_weight) *total_weight = w;
   g_diff_n = n;
   g_diff_spec = *spec;
   g_diff_grid = grid;
   g_diff_tangents = (tangents != 0);
   g_diff_pass++;
   return 0;
}

unsigned long long mxgd_diff_pass (void) { return g_diff_pass; }

int mxgd_diff_adjoint (unsigned long long pass, unsigned int npix, const float *g, double *grad)
{
   if (g_diff_grid == 0) return fail ("no forward pass to take the adjoint of");
   if (pass != g_diff_pass)
     return fail ("the adjoint of forward pass %llu was asked for, but the device holds pass %llu "
                  "(another forward pass ran since)", pass, g_diff_pass);
   if (npix != g_diff_spec.npix)
     return fail ("gradient image of %u pixels a side for a forward pass of %u", npix, g_diff_spec.npix);
   if (!g_diff_tangents)
     return fail ("the last forward pass (mxg_diff_image) kept no derivatives; call mxg_diff_forward");
   HIPCHECK (launch_diff_adjoint (g_diff_n, g_diff_out, &g_diff_spec, g, d_diff_partial, g_diff_grid,
                                  d_diff_small + 8, g_stream, g_diff_ev + 4));
   HIPCHECK (hipMemcpyAsync (grad, d_diff_small + 8, kDiffGrad * sizeof (double),
                             hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   HIPCHECK (hipEventElapsedTime (&g_diff_ms[1], g_diff_ev[4], g_diff_ev[5]));
   return 0;
}

void mxgd_diff_kernel_ms (float ms[3], unsigned int *listed)
{
   ms[0] = g_diff_ms[0];
   ms[1] = g_diff_ms[1];
   ms[2] = g_diff_ms[2];
   if (listed) *listed = g_diff_listed;
}

int mxgd_diff_state (unsigned int n, double *w, double *y, double *z, double *dw, double *dy,
                     double *dz, unsigned char *shell)
{
   if ((g_diff_grid == 0) || (n != g_diff_n))
     return fail ("no forward pass of %u photons", n);
   if (!g_diff_tangents && (dw || dy || dz))
     return fail ("the last forward pass (mxg_diff_image) kept no derivatives");
   const size_t b = (size_t) n * sizeof (double);
   HIPCHECK (hipStreamSynchronize (g_stream));
   if (w) HIPCHECK (hipMemcpy (w, g_diff_out.w, b, hipMemcpyDeviceToHost));
   if (y) HIPCHECK (hipMemcpy (y, g_diff_out.y, b, hipMemcpyDeviceToHost));
   if (z) HIPCHECK (hipMemcpy (z, g_diff_out.z, b, hipMemcpyDeviceToHost));
   if (dw) HIPCHECK (hipMemcpy (dw, g_diff_out.dw, MXG_DIFF_NPAR * b, hipMemcpyDeviceToHost));
   if (dy) HIPCHECK (hipMemcpy (dy, g_diff_out.dy, MXG_DIFF_NPAR * b, hipMemcpyDeviceToHost));
   if (dz) HIPCHECK (hipMemcpy (dz, g_diff_out.dz, MXG_DIFF_NPAR * b, hipMemcpyDeviceToHost));
   if (dw || dy || dz)
     {
        // photons lost before the paraboloid have no derivative columns on
        // the device (see k_diff_enter): their derivatives are 0
        double *wv = w;
        if (wv == nullptr)
          {
             wv = (double *) malloc (b);
             if (wv == nullptr) return fail ("out of memory");
             hipError_t e = hipMemcpy (wv, g_diff_out.w, b, hipMemcpyDeviceToHost);
             if (e != hipSuccess) { free (wv); return fail ("hipMemcpy: %s", hipGetErrorString (e)); }
          }
        for (unsigned int i = 0; i < n; i++)
          if (wv[i] == 0.0)
            for (int j = 0; j < MXG_DIFF_NPAR; j++)
              {
                 if (dw) dw[(size_t) j * n + i] = 0.0;
                 if (dy) dy[(size_t) j * n + i] = 0.0;
                 if (dz) dz[(size_t) j * n + i] = 0.0;
              }
        if (wv != w) free (wv);
     }
   if (shell) HIPCHECK (hipMemcpy (shell, g_diff_out.shell, n, hipMemcpyDeviceToHost));
   return 0;
}

/* ---- batched point sources ---------------------------------------------------- */

int mxgd_batch_set_sources (unsigned int nsrc, const double *dirs, const unsigned int *counts,
                            unsigned int *n_total)
{
   if (mxgd_init ()) return -1;
   if (g_capturing) return fail ("sources cannot change during graph capture");
   uint64_t tot = 0;
   for (unsigned int k = 0; k < nsrc; k++) tot += counts[k];
   if (tot > (1ull << 28)) return fail ("batch of %llu photons exceeds 2^28", (unsigned long long) tot);
   if (nsrc > g_batch_cap)
     {
        HIPCHECK (hipStreamSynchronize (g_stream));
        if (d_batch_off) (void) hipFree (d_batch_off);
        if (d_batch_dir) (void) hipFree (d_batch_dir);
        d_batch_off = nullptr;
        d_batch_dir = nullptr;
        g_batch_cap = 0;
        HIPCHECK (hipMalloc (&d_batch_off, (nsrc + 1) * sizeof (uint32_t)));
        HIPCHECK (hipMalloc (&d_batch_dir, 3 * (size_t) nsrc * sizeof (double)));
        g_batch_cap = nsrc;
        g_alloc_epoch++;
     }
   if (!d_batch_tag)
     {
        HIPCHECK (hipMalloc (&d_batch_tag, sizeof (uint32_t)));
        g_alloc_epoch++;
     }
   uint32_t *off = (uint32_t *) malloc ((nsrc + 1) * sizeof (uint32_t));
   if (off == nullptr) return fail ("out of memory");
   off[0] = 0;
   for (unsigned int k = 0; k < nsrc; k++) off[k + 1] = off[k] + counts[k];
   hipError_t e = hipMemcpy (d_batch_off, off, (nsrc + 1) * sizeof (uint32_t), hipMemcpyHostToDevice);
   free (off);
   HIPCHECK (e);
   if (nsrc)
     HIPCHECK (hipMemcpy (d_batch_dir, dirs, 3 * (size_t) nsrc * sizeof (double), hipMemcpyHostToDevice));
   g_batch_nsrc = nsrc;
   *n_total = (unsigned int) tot;
   return 0;
}

int mxgd_batch_prepare (unsigned int n)
{
   if (mxgd_init ()) return -1;
   if (ensure_capacity (n ? n : 1)) return -1;
   g_cur = 0;
   g_n = n;
   g_upper = n;
   return 0;
}

int mxgd_batch_set_tag_start (unsigned int tag_start)
{
   if (!d_batch_tag) return fail ("no batch sources uploaded");
   hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, d_batch_tag, (uint32_t) tag_start);
   HIPCHECK (hipGetLastError ());
   return 0;
}

int mxgd_run_batch_source (unsigned int n, double exposure, double start_time,
                           unsigned long long seed, int slot, int entry_follows)
{
   g_relay_have = -1;
   if (!d_cfg) return fail ("configuration not on the device");
   if ((n > g_capacity) || (g_batch_nsrc == 0)) return fail ("batch not prepared");
   bool timed = (slot >= 0) && (slot < kSlots) && !g_capturing;
   if (timed)
     {
        HIPCHECK (hipEventRecord (g_tev[slot][0], g_stream));
        g_slot_used[slot] = 1;
     }
   uint32_t grid = stage_grid (n, MXGD_K_SOURCE);
   HIPCHECK (launch_batch_source (d_cfg, d_da, g_soa, n, d_batch_off, g_batch_nsrc, d_batch_dir,
                                  d_batch_tag, exposure, start_time, seed, d_idx[0], d_counts, grid,
                                  g_hcfg.src.type == MXG_SOURCE_POINT, entry_follows != 0, g_stream));
   if (timed)
     HIPCHECK (hipEventRecord (g_tev[slot][1], g_stream));
   g_cur = 0;
   g_n = n;
   g_upper = n;
   return 0;
}

int mxgd_run_batch_persistent (unsigned int n, double exposure, double start_time,
                               unsigned long long seed, int slot)
{
   g_relay_have = -1;
   if (!d_cfg) return fail ("configuration not on the device");
   if ((n > g_capacity) || (g_batch_nsrc == 0)) return fail ("batch not prepared");
   if (g_sched.persist_tile == 0) return fail ("the schedule has no persistent tile");
   if (g_hcfg.det.fef_on) return fail ("the persistent pass has no FEF step");
   if (g_hcfg.src.type != MXG_SOURCE_POINT) return fail ("the persistent pass traces point sources only");
   const uint32_t tiles = (n + g_sched.persist_tile - 1) / g_sched.persist_tile;
   if (tiles > g_tile_cap) return fail ("persistent pass of %u tiles exceeds %zu", tiles, g_tile_cap);
   bool timed = (slot >= 0) && (slot < kSlots) && !g_capturing;
   if (timed)
     {
        HIPCHECK (hipEventRecord (g_tev[slot][0], g_stream));
        g_slot_used[slot] = 1;
     }
   PersistLaunch p;
   p.cfg = d_cfg;
   p.hcfg = &g_hcfg;
   p.fa = d_fa;
   p.da = d_da;
   p.s = g_soa;
   p.n = n;
   p.off = d_batch_off;
   p.nsrc = g_batch_nsrc;
   p.dir = d_batch_dir;
   p.tag_start = d_batch_tag;
   p.exposure = exposure;
   p.start_time = start_time;
   p.seed = seed;
   p.rng_pos = d_rng_pos;
   p.work = d_counts + 32;
   p.stage_count = d_counts + 36;
   p.tile = g_sched.persist_tile;
   p.blocks = (uint32_t) (g_num_cus > 0 ? g_num_cus : 256) * g_sched.persist_blocks_per_cu;
   p.seg = d_seg;
   p.tile_count = d_tile_cnt;
   p.tile_off = d_tile_cnt + g_tile_cap;
   p.idx_out = d_idx[1];
   p.n_out = d_counts + 1;
   p.grating = (g_hcfg.grating.type == MARX_GRATING_HETG) ? 1 : (g_hcfg.grating.type == MARX_GRATING_LETG) ? 2 : 0;
   p.kind = (g_hcfg.det.type > 0) ? detector_kind (g_hcfg) : -1;
   p.ev = nullptr;
   HIPCHECK (launch_batch_persistent (p, g_stream));
   if (timed)
     HIPCHECK (hipEventRecord (g_tev[slot][1], g_stream));
   // the per-stage counts where the staged pass leaves them: photons into
   // the two mirror kernels, out of the mirror, grating and detector
   hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, d_counts + 8 + MXGD_SLOT_HRMA_ENTER, (uint32_t) n);
   hipLaunchKernelGGL (k_copy_u32, dim3 (1), dim3 (1), 0, g_stream, d_counts + 8 + MXGD_SLOT_HRMA_REFLECT,
                       d_counts + 36);
   for (int k = 1; k <= 3; k++)
     hipLaunchKernelGGL (k_copy_u32, dim3 (1), dim3 (1), 0, g_stream, d_counts + 8 + k, d_counts + 36 + k);
   HIPCHECK (hipGetLastError ());
   g_cur = 1;
   g_n = n;
   g_upper = n;
   return 0;
}

int mxgd_set_schedule (const mxgd_schedule *s)
{
   for (int k = 0; k < MXGD_NUM_K; k++)
     if (((s->blocks_per_cu[k] < 1) && !((k == MXGD_K_GRATING) && (s->blocks_per_cu[k] == 0)))
         || (s->blocks_per_cu[k] > 32))
       return fail ("blocks per CU must be 1..32 (0 for the grating: one round of the blocks a CU holds)");
   if ((s->reflect_waves < 2) || (s->reflect_waves > 4))
     return fail ("reflection kernel waves per SIMD must be 2, 3 or 4");
   if ((s->persist_tile % kBlock) || (s->persist_tile > kMaxPersistTile))
     return fail ("persistent tile must be 0 or a multiple of %d up to %u photons", kBlock,
                  (unsigned int) kMaxPersistTile);
   if (s->persist_tile && ((s->persist_blocks_per_cu < 1) || (s->persist_blocks_per_cu > 8)))
     return fail ("persistent blocks per CU must be 1..8");
   g_sched = *s;
   return 0;
}

void mxgd_get_schedule (mxgd_schedule *s) { *s = g_sched; }

unsigned int mxgd_alloc_epoch (void) { return g_alloc_epoch; }

struct mxgd_graph
{
   hipGraph_t graph;
   hipGraphExec_t exec;
   int cur;                 // list state at the end of the captured sequence
   unsigned int n, upper;
};

int mxgd_graph_begin (void)
{
   if (mxgd_init ()) return -1;
   if (g_capturing) return fail ("graph capture already running");
   HIPCHECK (hipStreamSynchronize (g_stream));
   HIPCHECK (hipStreamBeginCapture (g_stream, hipStreamCaptureModeRelaxed));
   g_capturing = 1;
   return 0;
}

int mxgd_graph_end (mxgd_graph **out)
{
   *out = nullptr;
   if (!g_capturing) return fail ("no graph capture running");
   g_capturing = 0;
   hipGraph_t graph = nullptr;
   HIPCHECK (hipStreamEndCapture (g_stream, &graph));
   mxgd_graph *g = (mxgd_graph *) calloc (1, sizeof (mxgd_graph));
   if (g == nullptr) { (void) hipGraphDestroy (graph); return fail ("out of memory"); }
   g->graph = graph;
   hipError_t e = hipGraphInstantiate (&g->exec, graph, nullptr, nullptr, 0);
   if (e != hipSuccess)
     {
        (void) hipGraphDestroy (graph);
        free (g);
        return fail ("hipGraphInstantiate failed: %s", hipGetErrorString (e));
     }
   g->cur = g_cur;
   g->n = g_n;
   g->upper = g_upper;
   *out = g;
   return 0;
}

int mxgd_graph_launch (mxgd_graph *g, float *ms)
{
   if (ms) HIPCHECK (hipEventRecord (g_ev0, g_stream));
   HIPCHECK (hipGraphLaunch (g->exec, g_stream));
   if (ms)
     {
        HIPCHECK (hipEventRecord (g_ev1, g_stream));
        HIPCHECK (hipEventSynchronize (g_ev1));
        HIPCHECK (hipEventElapsedTime (ms, g_ev0, g_ev1));
     }
   g_cur = g->cur;
   g_n = g->n;
   g_upper = g->upper;
   return 0;
}

void mxgd_graph_destroy (mxgd_graph *g)
{
   if (g == nullptr) return;
   (void) hipStreamSynchronize (g_stream);
   (void) hipGraphExecDestroy (g->exec);
   (void) hipGraphDestroy (g->graph);
   free (g);
}

} // extern "C"

/* ---- host photon lists: pinned memory, energy order, active list ------------------- */

namespace {
void *d_sort = nullptr;              // keys and values in / out, then rocPRIM's storage
size_t g_sort_cap = 0, g_sort_tmp = 0;
double *d_gather = nullptr;          // energies of the active list
size_t g_gather_cap = 0;
// pinned host allocations made for photon lists
void **g_pinned = nullptr;
size_t g_pinned_n = 0, g_pinned_cap = 0;

int sort_buffers (size_t n)
{
   if (n <= g_sort_cap) return 0;
   size_t cap = g_sort_cap ? g_sort_cap : 1024;
   while (cap < n) cap *= 2;
   size_t tmp = 0;
   HIPCHECK (sort_energies (nullptr, (uint32_t) cap, nullptr, nullptr, nullptr, nullptr, nullptr, &tmp, g_stream));
   if (d_sort) { HIPCHECK (hipStreamSynchronize (g_stream)); (void) hipFree (d_sort); d_sort = nullptr; }
   g_sort_cap = 0;
   HIPCHECK (hipMalloc (&d_sort, 2 * align256 (cap * sizeof (double)) + 2 * align256 (cap * sizeof (uint32_t))
                                 + align256 (tmp)));
   g_sort_cap = cap;
   g_sort_tmp = tmp;
   return 0;
}
} // namespace

extern "C" {

void *mxgd_host_alloc (size_t bytes)
{
   int n = 0;
   if (!g_inited && ((hipGetDeviceCount (&n) != hipSuccess) || (n == 0)))
     return nullptr;                       // no GPU: the caller uses malloc
   if (mxgd_init ()) return nullptr;
   void *p = nullptr;
   if (hipHostMalloc (&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
     return nullptr;
   memset (p, 0, bytes);
   if (g_pinned_n == g_pinned_cap)
     {
        size_t cap = g_pinned_cap ? 2 * g_pinned_cap : 64;
        void **q = (void **) realloc (g_pinned, cap * sizeof (void *));
        if (q == nullptr) { (void) hipHostFree (p); return nullptr; }
        g_pinned = q;
        g_pinned_cap = cap;
     }
   g_pinned[g_pinned_n++] = p;
   return p;
}

int mxgd_host_free (void *p)
{
   for (size_t k = 0; k < g_pinned_n; k++)
     if (g_pinned[k] == p)
       {
          g_pinned[k] = g_pinned[--g_pinned_n];
          (void) hipStreamSynchronize (g_stream);
          (void) hipHostFree (p);
          return 1;
       }
   return 0;
}

int mxgd_sort_store_by_energy (unsigned int n, unsigned int *index, double *energies)
{
   if (mxgd_init ()) return -1;
   if (n > g_n) return fail ("sort of %u photons in a store of %u", n, g_n);
   if (n == 0) return 0;
   if (sort_buffers (n)) return -1;
   char *b = (char *) d_sort;
   double *k_in = (double *) b; b += align256 (g_sort_cap * sizeof (double));
   double *k_out = (double *) b; b += align256 (g_sort_cap * sizeof (double));
   uint32_t *v_in = (uint32_t *) b; b += align256 (g_sort_cap * sizeof (uint32_t));
   uint32_t *v_out = (uint32_t *) b; b += align256 (g_sort_cap * sizeof (uint32_t));
   size_t tmp = g_sort_tmp;
   HIPCHECK (sort_energies (g_soa.energy, n, k_in, v_in, k_out, v_out, b, &tmp, g_stream));
   HIPCHECK (hipMemcpyAsync (index, v_out, (size_t) n * sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipMemcpyAsync (energies, k_out, (size_t) n * sizeof (double), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

int mxgd_download_active (unsigned int *index, double *energies, unsigned int cap, unsigned int *n_out)
{
   if (mxgd_init ()) return -1;
   uint32_t m = 0;
   HIPCHECK (hipMemcpyAsync (&m, d_counts + g_cur, sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   if (m > cap) return fail ("active list of %u photons for a list of %u", m, cap);
   *n_out = m;
   if (m == 0) return 0;
   if (m > g_gather_cap)
     {
        if (d_gather) (void) hipFree (d_gather);
        d_gather = nullptr;
        g_gather_cap = 0;
        HIPCHECK (hipMalloc (&d_gather, (size_t) m * sizeof (double)));
        g_gather_cap = m;
     }
   HIPCHECK (gather_energies (g_soa.energy, d_idx[g_cur], d_counts + g_cur, m, d_gather, g_stream));
   HIPCHECK (hipMemcpyAsync (index, d_idx[g_cur], (size_t) m * sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipMemcpyAsync (energies, d_gather, (size_t) m * sizeof (double), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}

} // extern "C"

/* ---- the allocation-free trace ------------------------------------------------------- */

extern "C" {

int mxgd_trace_noalloc (const mxg_soa_view *v, unsigned int n, unsigned int tag_start,
                        double start_time, double exposure_time, unsigned long long seed,
                        const mxg_config *cfg, const float *fa, unsigned int nf, const double *da,
                        unsigned int nd, unsigned int generation, unsigned int *n_collected,
                        unsigned int *n_detected)
{
   // mxg_init sets the library up (streams, the scratch's address); this
   // entry point allocates nothing, so it does not do that itself
   if (!g_inited) return fail ("mxg_init has not run: the allocation-free trace needs the library set up");
   if (g_capturing) return fail ("the allocation-free trace cannot run during the library's graph capture");
   if (n > kNoAllocPhotons)
     return fail ("chunk of %u photons exceeds the %u photons of mxg_trace_device_noalloc's scratch", n,
                  (unsigned int) kNoAllocPhotons);
   if (v->capacity < n)
     return fail ("caller's columns hold %u photons, %u asked for", v->capacity, n);
   // the configuration: the library's device copy, brought up to date in
   // place when the buffers an earlier call allocated hold it
   if (!d_cfg || (g_cfg_gen != generation))
     {
        if (!d_cfg || (nf + 1 > d_fa_cap) || (nd + 1 > d_da_cap))
          return fail ("the configuration's device buffers are missing or too small for this set-up: make "
                       "one allocating call (mxg_trace_device, marx_mirror_reflect, ...) after the last "
                       "init first");
        HIPCHECK (hipMemcpyAsync (d_cfg, cfg, sizeof (mxg_config), hipMemcpyHostToDevice, g_stream));
        if (nf) HIPCHECK (hipMemcpyAsync (d_fa, fa, nf * sizeof (float), hipMemcpyHostToDevice, g_stream));
        if (nd) HIPCHECK (hipMemcpyAsync (d_da, da, nd * sizeof (double), hipMemcpyHostToDevice, g_stream));
        HIPCHECK (hipStreamSynchronize (g_stream));
        memcpy (&g_hcfg, cfg, sizeof (mxg_config));
        g_grating_bpc = grating_blocks_per_cu (&g_hcfg);
        g_cfg_gen = generation;
     }
   int dev = 0;
   HIPCHECK (hipGetDevice (&dev));
   if (scratch_setup (dev)) return -1;
   unsigned char *base = g_scratch_ptr[dev];
   uint32_t *counts = (uint32_t *) (base + kNaCounts);
   uint32_t *seg_count = (uint32_t *) (base + kNaSegCount);
   uint32_t *idx[2] = {(uint32_t *) (base + kNaIdx0), (uint32_t *) (base + kNaIdx1)};
   uint32_t *mid = (uint32_t *) (base + kNaMid);
   uint32_t *seg = (uint32_t *) (base + kNaSeg);
   uint32_t *rng = (uint32_t *) (base + kNaRng);
   uint64_t *ticks = (uint64_t *) (base + kNaTicks);
   uint64_t *tiles = (uint64_t *) (base + kNaTiles);
   const mxg_config *c_cfg = d_cfg;
   const float *c_fa = d_fa;
   const double *c_da = d_da;

   Soa s;
   s.energy = v->energy;
   s.x0 = v->x[0]; s.x1 = v->x[1]; s.x2 = v->x[2];
   s.p0 = v->p[0]; s.p1 = v->p[1]; s.p2 = v->p[2];
   s.t = v->arrival_time;
   s.flags = v->flags;
   s.ypix = v->y_pixel; s.zpix = v->z_pixel; s.upix = v->u_pixel; s.vpix = v->v_pixel;
   for (int k = 0; k < 6; k++) s.dith[k] = v->dither[k];
   s.pi = v->pi;
   s.pha = v->pulse_height;
   s.shell = v->mirror_shell;
   s.ccd = v->ccd_num; s.region = v->detector_region; s.order = v->order;
   for (int k = 0; k < 4; k++) s.sup[k] = v->support_orders[k];
   s.tag = v->tag;

   // counters start at zero (one-thread kernels, not memsets): [0] and [1]
   // the two lists, [2] the source's exposure cut, [4] the mirror's middle
   // list, [6] the photons collected
   for (int k = 0; k < 8; k++)
     hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, counts + k, 0u);
   HIPCHECK (hipGetLastError ());
   HIPCHECK (launch_source (c_cfg, c_fa, c_da, s, n, tag_start, exposure_time, start_time, seed, 0, ticks, tiles,
                            counts + 2, counts + 0, nullptr, idx[0], cfg->src.type == MXG_SOURCE_POINT, false,
                            g_stream));
   hipLaunchKernelGGL (k_copy_u32, dim3 (1), dim3 (1), 0, g_stream, counts + 6, counts + 0);
   HIPCHECK (hipGetLastError ());

   // stages: list idx[cur] (length counts[cur]) in, idx[1 - cur] out; grids
   // sized by n, an upper bound of every list (the kernels read the
   // lengths from device memory)
   int cur = 0;
   auto io = [&] (int k) {
      ListIO l;
      l.idx_in = idx[cur];
      l.n_in = counts + cur;
      l.idx_out = idx[1 - cur];
      l.n_out = counts + (1 - cur);
      l.seg = seg;
      l.seg_count = seg_count;
      uint32_t blocks = (n + kBlock - 1) / kBlock;
      uint32_t cap = (uint32_t) (g_num_cus > 0 ? g_num_cus : 256) * sched_blocks_per_cu (k);
      if (cap > kMaxSegments) cap = kMaxSegments;
      l.grid = blocks < cap ? (blocks ? blocks : 1u) : cap;
      return l;
   };
   auto clear_out = [&] (void) {
      hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, counts + (1 - cur), 0u);
      return hipGetLastError ();
   };
   HIPCHECK (clear_out ());
   HrmaScratch hs = {mid, counts + 4, rng, io (MXGD_K_REFLECT).grid, (int) g_sched.reflect_waves};
   HIPCHECK (launch_hrma (c_cfg, cfg, c_fa, c_da, s, io (MXGD_K_ENTER), hs, n, cfg->src.distance, seed, 0,
                          nullptr, false, g_stream, nullptr));
   cur = 1 - cur;
   if ((cfg->grating.type == MARX_GRATING_HETG) || (cfg->grating.type == MARX_GRATING_LETG))
     {
        HIPCHECK (clear_out ());
        HIPCHECK (launch_grating (c_cfg, cfg, c_fa, c_da, s, io (MXGD_K_GRATING), n, seed, 0,
                                  cfg->grating.type == MARX_GRATING_LETG, nullptr, nullptr, g_stream));
        cur = 1 - cur;
     }
   if (cfg->det.type > 0)
     {
        HIPCHECK (clear_out ());
        HIPCHECK (launch_detect (c_cfg, c_fa, c_da, s, io (MXGD_K_DETECT), n, start_time, seed, 0,
                                 cfg->det.fef_on != 0, detector_kind (*cfg), nullptr, g_stream));
        cur = 1 - cur;
     }
   uint32_t out[2] = {0, 0};
   HIPCHECK (hipMemcpyAsync (out, counts + 6, sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipMemcpyAsync (out + 1, counts + cur, sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   if (n_collected) *n_collected = out[0];
   if (n_detected) *n_detected = out[1];
   return 0;
}

} // extern "C"

/* ---- output files ------------------------------------------------------------------ */

namespace {
uint32_t *d_out_list = nullptr;      // segments, then the list of photons written
size_t g_out_list_cap = 0;
uint32_t *d_out_u32 = nullptr;       // [0] photons in the store, [1] photons written
uint8_t *d_out_bytes = nullptr;      // packed columns
size_t g_out_bytes_cap = 0;
unsigned char *h_out_bytes = nullptr;    // their pinned host copy
size_t g_out_host_cap = 0;
// event coordinates
void *d_ev = nullptr;                // frame heads, chunk maxima and prefixes, columns
size_t g_ev_cap = 0;
void *d_ev_par = nullptr;            // mxg_event_params, then the subpixel tables
size_t g_ev_par_cap = 0;
hipEvent_t g_ev_ev[2] = {nullptr, nullptr};
float g_ev_ms = 0.0f;

// The list of the store's photons whose flags & BAD_PHOTON_MASK is zero, in
// store order: io.idx_out, *m rows.  Synchronises.
int keep_rows (ListIO &io, uint32_t *m)
{
   const uint32_t n = g_n;
   if (g_out_list_cap < n)
     {
        HIPCHECK (hipStreamSynchronize (g_stream));
        if (d_out_list) (void) hipFree (d_out_list);
        d_out_list = nullptr;
        g_out_list_cap = 0;
        HIPCHECK (hipMalloc (&d_out_list, 2 * align256 ((size_t) n * sizeof (uint32_t))));
        g_out_list_cap = n;
     }
   if (!d_out_u32) HIPCHECK (hipMalloc (&d_out_u32, 4 * sizeof (uint32_t)));
   io.idx_in = nullptr;
   io.n_in = d_out_u32;
   io.seg = d_out_list;
   io.idx_out = d_out_list + align256 (g_out_list_cap * sizeof (uint32_t)) / sizeof (uint32_t);
   io.n_out = d_out_u32 + 1;
   io.seg_count = d_seg_count;
   io.grid = stage_grid (n, MXGD_K_DETECT);
   hipLaunchKernelGGL (k_set_u32, dim3 (1), dim3 (1), 0, g_stream, d_out_u32, n);
   HIPCHECK (hipGetLastError ());
   HIPCHECK (launch_keep_list (g_soa.flags, n, io, g_stream));
   *m = 0;
   HIPCHECK (hipMemcpyAsync (m, d_out_u32 + 1, sizeof (uint32_t), hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   return 0;
}
} // namespace

extern "C" {

int mxgd_pack_output (const int *cols, int ncols, double total_time, unsigned int *n_rows,
                      const unsigned char **bytes, size_t *col_off)
{
   if (mxgd_init ()) return -1;
   if ((ncols < 0) || (ncols > kMaxPackCols)) return fail ("%d output columns (at most %d)", ncols, kMaxPackCols);
   if (g_capturing) return fail ("output cannot be written during graph capture");
   *n_rows = 0;
   *bytes = nullptr;
   if ((g_n == 0) || (ncols == 0)) return 0;
   ListIO io;
   uint32_t m = 0;
   if (keep_rows (io, &m)) return -1;
   if (m == 0) return 0;

   PackArgs a;
   memset (&a, 0, sizeof (a));
   a.ncols = ncols;
   a.total_time = total_time;
   size_t off = 0;
   for (int k = 0; k < ncols; k++)
     {
        if ((cols[k] < 0) || (cols[k] >= MXG_NUM_COLS)) return fail ("unknown output column %d", cols[k]);
        a.col[k] = cols[k];
        a.offset[k] = off;
        col_off[k] = off;
        off += align256 ((size_t) m * MXG_COL_WIDTH (cols[k]));
     }
   if (g_out_bytes_cap < off)
     {
        if (d_out_bytes) (void) hipFree (d_out_bytes);
        d_out_bytes = nullptr;
        g_out_bytes_cap = 0;
        HIPCHECK (hipMalloc (&d_out_bytes, off));
        g_out_bytes_cap = off;
     }
   if (g_out_host_cap < off)
     {
        if (h_out_bytes) (void) hipHostFree (h_out_bytes);
        h_out_bytes = nullptr;
        g_out_host_cap = 0;
        HIPCHECK (hipHostMalloc ((void **) &h_out_bytes, off, hipHostMallocDefault));
        g_out_host_cap = off;
     }
   HIPCHECK (launch_pack_columns (g_soa, io.idx_out, io.n_out, m, a, d_out_bytes, g_stream));
   HIPCHECK (hipMemcpyAsync (h_out_bytes, d_out_bytes, off, hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   *n_rows = m;
   *bytes = h_out_bytes;
   return 0;
}

int mxgd_event_coords (const mxg_event_params *par, const float *subtab, unsigned int n_subtab,
                       double total_time, double time_del, unsigned long long seed,
                       const mxg_event_columns *out, unsigned int capacity, unsigned int *num_rows)
{
   if (mxgd_init ()) return -1;
   if (g_capturing) return fail ("event coordinates cannot be computed during graph capture");
   *num_rows = 0;
   g_ev_ms = 0.0f;
   if (g_n == 0) return 0;
   ListIO io;
   uint32_t m = 0;
   if (keep_rows (io, &m)) return -1;
   if (m > capacity) return fail ("%u event rows exceed the capacity of %u", m, capacity);
   if (m == 0) return 0;

   const size_t par_bytes = align256 (sizeof (mxg_event_params)) + (size_t) n_subtab * sizeof (float);
   if (g_ev_par_cap < par_bytes)
     {
        HIPCHECK (hipStreamSynchronize (g_stream));
        if (d_ev_par) (void) hipFree (d_ev_par);
        d_ev_par = nullptr;
        g_ev_par_cap = 0;
        HIPCHECK (hipMalloc (&d_ev_par, par_bytes));
        g_ev_par_cap = par_bytes;
     }
   mxg_event_params *d_par = (mxg_event_params *) d_ev_par;
   float *d_sub = (float *) ((char *) d_ev_par + align256 (sizeof (mxg_event_params)));
   HIPCHECK (hipMemcpyAsync (d_par, par, sizeof (mxg_event_params), hipMemcpyHostToDevice, g_stream));
   if (n_subtab)
     HIPCHECK (hipMemcpyAsync (d_sub, subtab, n_subtab * sizeof (float), hipMemcpyHostToDevice, g_stream));

   // frame heads, chunk maxima, chunk prefixes, then the columns
   const size_t nc = (m + kBlock - 1) / kBlock;
   const size_t w4 = align256 ((size_t) m * 4), w8 = align256 ((size_t) m * 8), w2 = align256 ((size_t) m * 2);
   const size_t wc = align256 (nc * 4);
   const size_t need = w4 + 2 * wc + 4 * w4 + 4 * w8 + w2;
   if (g_ev_cap < need)
     {
        HIPCHECK (hipStreamSynchronize (g_stream));
        if (d_ev) (void) hipFree (d_ev);
        d_ev = nullptr;
        g_ev_cap = 0;
        HIPCHECK (hipMalloc (&d_ev, need));
        g_ev_cap = need;
     }
   char *b = (char *) d_ev;
   uint32_t *head = (uint32_t *) b; b += w4;
   uint32_t *cmax = (uint32_t *) b; b += wc;
   uint32_t *cpre = (uint32_t *) b; b += wc;
   EventOut eo;
   eo.tag = (uint32_t *) b; b += w4;
   eo.ccd = (int32_t *) b; b += w4;
   eo.tdetx = (int32_t *) b; b += w4;
   eo.tdety = (int32_t *) b; b += w4;
   eo.detx = (double *) b; b += w8;
   eo.dety = (double *) b; b += w8;
   eo.x = (double *) b; b += w8;
   eo.y = (double *) b; b += w8;
   eo.fltgrade = (int16_t *) b;
   if (!g_ev_ev[0])
     for (int k = 0; k < 2; k++) HIPCHECK (hipEventCreate (&g_ev_ev[k]));
   HIPCHECK (hipEventRecord (g_ev_ev[0], g_stream));
   HIPCHECK (launch_event_coords (d_cfg, d_par, d_sub, g_soa, io.idx_out, m, total_time, time_del, seed, head,
                                  cmax, cpre, eo, g_stream));
   HIPCHECK (hipEventRecord (g_ev_ev[1], g_stream));
   struct { void *h; const void *d; size_t w; } cp[9] = {
      {out->tag, eo.tag, 4}, {out->ccd_id, eo.ccd, 4}, {out->tdetx, eo.tdetx, 4}, {out->tdety, eo.tdety, 4},
      {out->detx, eo.detx, 8}, {out->dety, eo.dety, 8}, {out->x, eo.x, 8}, {out->y, eo.y, 8},
      {out->fltgrade, eo.fltgrade, 2}};
   for (int k = 0; k < 9; k++)
     if (cp[k].h)
       HIPCHECK (hipMemcpyAsync (cp[k].h, cp[k].d, (size_t) m * cp[k].w, hipMemcpyDeviceToHost, g_stream));
   HIPCHECK (hipStreamSynchronize (g_stream));
   HIPCHECK (hipEventElapsedTime (&g_ev_ms, g_ev_ev[0], g_ev_ev[1]));
   *num_rows = m;
   return 0;
}

float mxgd_event_ms (void)
{
   return g_ev_ms;
}

} // extern "C"
