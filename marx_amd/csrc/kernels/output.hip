// output.hip -- the output files' bytes made in HBM: the photons whose
// flags & BAD_PHOTON_MASK is zero, in store (tag) order, and each requested
// column of them as big-endian values of the file's type, the conversions
// of marx/libsrc/marxio.c:230-257 (float32 of the double attributes, time +
// total_time in double before the cast, int16 mirror shell and PHA, int8
// chip, region and orders, int32 tag).
#include "dev_common.h"
#include "dev_store.h"
#include "mxg_columns.h"

namespace mxg {

// photons of the store not flagged bad, listed in order (SegWriter)
__global__ __launch_bounds__ (kBlock)
void k_keep_list (const uint32_t *__restrict__ flags, uint32_t n, ListIO io)
{
   __shared__ uint32_t wc[2 * (kBlock / 64)];
   const BlockRange range = block_range (n, blockIdx.x, gridDim.x);
   SegWriter sw (io, range, wc);
   for (uint32_t base = range.lo; base < range.hi; base += kBlock)
     {
        const uint32_t i = base + threadIdx.x;
        const bool keep = (i < range.hi) && ((flags[i] & 0xFFu) == 0u);
        sw.add (keep, i);
     }
   sw.finish (io);
}

__device__ __forceinline__ void put4 (uint8_t *col, uint32_t q, uint32_t bits)
{
   reinterpret_cast<uint32_t *> (col)[q] = __builtin_bswap32 (bits);
}
__device__ __forceinline__ void putf (uint8_t *col, uint32_t q, float v) { put4 (col, q, __float_as_uint (v)); }
__device__ __forceinline__ void put2 (uint8_t *col, uint32_t q, int16_t v)
{
   reinterpret_cast<uint16_t *> (col)[q] = __builtin_bswap16 ((uint16_t) v);
}
__device__ __forceinline__ void put1 (uint8_t *col, uint32_t q, int8_t v)
{
   reinterpret_cast<int8_t *> (col)[q] = v;
}

// Row q of every requested column: the photon idx[q].  Column k's values
// start at out + a.offset[k] (aligned for its width).
__global__ __launch_bounds__ (kBlock)
void k_pack_columns (Soa s, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ n_kept,
                     PackArgs a, uint8_t *__restrict__ out)
{
   const uint32_t m = *n_kept;
   const uint32_t stride = gridDim.x * blockDim.x;
   for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += stride)
     {
        const uint32_t i = idx[q];
        for (int k = 0; k < a.ncols; k++)
          {
             uint8_t *c = out + a.offset[k];
             switch (a.col[k])
               {
                case MXG_COL_B_ENERGY: putf (c, q, s.pi[i]); break;
                case MXG_COL_ENERGY: putf (c, q, (float) s.energy[i]); break;
                case MXG_COL_TIME: putf (c, q, (float) (s.t[i] + a.total_time)); break;
                case MXG_COL_TAG: put4 (c, q, s.tag[i]); break;
                case MXG_COL_XPOS: putf (c, q, (float) s.x0[i]); break;
                case MXG_COL_YPOS: putf (c, q, (float) s.x1[i]); break;
                case MXG_COL_ZPOS: putf (c, q, (float) s.x2[i]); break;
                case MXG_COL_COSX: putf (c, q, (float) s.p0[i]); break;
                case MXG_COL_COSY: putf (c, q, (float) s.p1[i]); break;
                case MXG_COL_COSZ: putf (c, q, (float) s.p2[i]); break;
                case MXG_COL_PHA: put2 (c, q, s.pha[i]); break;
                case MXG_COL_CCDID: put1 (c, q, s.ccd[i]); break;
                case MXG_COL_CHIPX: putf (c, q, s.ypix[i]); break;
                case MXG_COL_CHIPY: putf (c, q, s.zpix[i]); break;
                case MXG_COL_U: putf (c, q, s.upix[i]); break;
                case MXG_COL_V: putf (c, q, s.vpix[i]); break;
                case MXG_COL_MIRROR: put2 (c, q, (int16_t) s.shell[i]); break;
                case MXG_COL_HRCREGION: put1 (c, q, s.region[i]); break;
                case MXG_COL_ORDER: put1 (c, q, s.order[i]); break;
                case MXG_COL_FINE: put1 (c, q, s.sup[0][i]); break;
                case MXG_COL_COARSE1: put1 (c, q, s.sup[1][i]); break;
                case MXG_COL_COARSE2: put1 (c, q, s.sup[2][i]); break;
                case MXG_COL_COARSE3: put1 (c, q, s.sup[3][i]); break;
                case MXG_COL_RA: putf (c, q, s.dith[0][i]); break;
                case MXG_COL_DEC: putf (c, q, s.dith[1][i]); break;
                case MXG_COL_ROLL: putf (c, q, s.dith[2][i]); break;
                case MXG_COL_DET_DY: putf (c, q, s.dith[3][i]); break;
                case MXG_COL_DET_DZ: putf (c, q, s.dith[4][i]); break;
                case MXG_COL_DET_THETA: putf (c, q, s.dith[5][i]); break;
                default: break;
               }
          }
     }
}

hipError_t launch_keep_list (const uint32_t *flags, uint32_t n, const ListIO &io, hipStream_t st)
{
   hipLaunchKernelGGL (k_keep_list, dim3 (io.grid), dim3 (kBlock), 0, st, flags, n, io);
   hipError_t e = hipGetLastError ();
   if (e != hipSuccess) return e;
   return launch_compact (io, st);
}

hipError_t launch_pack_columns (const Soa &s, const uint32_t *idx, const uint32_t *n_kept, uint32_t m,
                                const PackArgs &a, uint8_t *out, hipStream_t st)
{
   uint32_t blocks = (m + kBlock - 1) / kBlock;
   if (blocks > 4096) blocks = 4096;
   if (blocks == 0) blocks = 1;
   hipLaunchKernelGGL (k_pack_columns, dim3 (blocks), dim3 (kBlock), 0, st, s, idx, n_kept, a, out);
   return hipGetLastError ();
}

} // namespace mxg
