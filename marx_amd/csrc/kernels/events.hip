// events.hip -- event coordinates of detected photons: tiled, focal-plane and
// sky pixels and (ACIS) the flight grade, as marx2fits computes them from the
// simulation's output columns (marx/src/marx2fits.c:3584-3735, 3813-3910,
// marx/libsrc/pixlib.c:134-238, 486-573, detpix.c:182-209,
// acis_geom.c:112-200, hrc_s_geom.c:477-510, hrc_i_geom.c:216-236,
// acis_subpix.c:283-322, dither.c:583-607).
#include "dev_common.h"
#include "dev_store.h"

namespace mxg {

// flight grades of the 3x3 subpixels, four each (marx2fits.c:3822-3838)
__constant__ int kFltGrades[9][4] = {
   {10, 11, 138, 139}, {2, 34, 130, 162}, {18, 22, 50, 54},
   {8, 12, 136, 140},  {0, 0, 0, 0},      {16, 17, 48, 49},
   {72, 76, 104, 108}, {64, 65, 68, 69},  {80, 81, 208, 209}};

// a / |a| as vector.c divides (the overflow-safe length, three divisions):
// the sky angles below come from acos of numbers near 1, which turns an ulp
// of a component into up to 1e-8 rad near the axis
__device__ __forceinline__ void normalize_div (vec3 &a)
{
   const double len = length (a);
   if (len != 0.0)
     {
        a.x = a.x / len;
        a.y = a.y / len;
        a.z = a.z / len;
     }
}

// rotations of vector.c with normalize_div
__device__ __forceinline__ vec3 rot_unit1_div (vec3 p, vec3 n, double c, double s)
{
   vec3 u = rotate_vector1 (p, n, c, s);
   normalize_div (u);
   return u;
}

__device__ __forceinline__ vec3 rot_unit_div (vec3 p, vec3 n, double theta)
{
   double s, c;
   sincos (theta, &s, &c);
   return rot_unit1_div (p, n, c, s);
}

// a vector of dither.c:583-607 with the pointing offsets taken out
__device__ __forceinline__ vec3 unapply_dither (double ra, double dec, double roll, vec3 p)
{
   double sin_ra, cos_ra, sin_dec, cos_dec;
   sincos (ra, &sin_ra, &cos_ra);
   sincos (dec, &sin_dec, &cos_dec);
   const double cos_theta = cos_dec * cos_ra;
   vec3 n = v3 (0.0, -sin_dec, cos_dec * sin_ra);
   const double sin_theta = length (n);
   if (sin_theta >= 1e-20)
     {
        n.x /= sin_theta;
        n.y /= sin_theta;
        n.z /= sin_theta;
        p = rot_unit1_div (p, n, cos_theta, sin_theta);
     }
   return rot_unit_div (p, v3 (1.0, 0.0, 0.0), roll);
}

// pixlib.c:486-513: the sky position of a mirror-nodal vector
__device__ __forceinline__ void mnc_to_ra_dec (vec3 m, double &ra, double &dec)
{
   double perp = sqrt (m.x * m.x + m.y * m.y);
   if (perp > 1.0) perp = 1.0;
   dec = (m.z <= 0) ? acos (perp) : -acos (perp);
   perp = -m.x / perp;
   ra = (m.y <= 0.0) ? acos (perp) : -acos (perp);
}

// pixlib.c:522-573 with (ra_0, dec_0) = (0, 0): the aspect offsets
__device__ __forceinline__ void ra_dec_offsets0 (double ra, double dec, double &d_ra_out, double &d_dec_out)
{
   const double d_ra = ra, d_dec = dec;
   const double c = cos (dec);
   const double factor = c * (1 - cos (d_ra));
   double sin_delta_dec = sin (d_dec) + factor * sin (0.0);
   if (fabs (sin_delta_dec) > 1.0) sin_delta_dec = (sin_delta_dec < 0) ? -1.0 : 1.0;
   d_dec_out = asin (sin_delta_dec);
   const double num = c * sin (d_ra);
   const double den = cos (d_dec) - factor * cos (0.0);
   if (den >= 0) d_ra_out = atan (num / den);
   else d_ra_out = (num >= 0) ? atan (num / den) + MXG_PI : atan (num / den) - MXG_PI;
}

// acis_subpix.c:283-322
__device__ __forceinline__ void subpix (const mxg_event_params &P, const float *__restrict__ tab, int ccd,
                                        float energy, int fltgrade, float &dx, float &dy)
{
   const int t = ((ccd == 5) || (ccd == 7)) ? 1 : 0;
   const unsigned int n = P.sub_n[t][fltgrade];
   dx = dy = 0.0f;
   if (n == 0) return;
   const float *e = tab + P.sub_off[t][fltgrade];
   unsigned int j = search_f (energy, e, n);
   if (j == 0) j++;
   if (j == n) j--;
   const double w1 = ((double) energy - e[j - 1]) / (e[j] - e[j - 1]);
   const double w0 = 1.0 - w1;
   dx = (float) (w0 * e[n + j - 1] + w1 * e[n + j]);
   dy = (float) (w0 * e[2 * n + j - 1] + w1 * e[2 * n + j]);
}

// The frame of row q: marx2fits reads the time column (float) into a double
// and takes (long)(time / TimeDel); with TimeDel <= 0 every row is a frame.
__device__ __forceinline__ long long frame_of (const Soa &s, uint32_t i, double total_time, double time_del)
{
   const float tf = (float) (s.t[i] + total_time);
   return (long long) ((double) tf / time_del);
}

__device__ __forceinline__ uint32_t wave_max_scan (uint32_t v)
{
   const int lane = threadIdx.x & 63;
#pragma unroll
   for (int d = 1; d < 64; d <<= 1)
     {
        const uint32_t u = __shfl_up (v, d, 64);
        if (lane >= d) v = max (v, u);
     }
   return v;
}

// First pass of the frame heads: within each chunk of kBlock rows, the
// largest row at or before q that starts a frame (0 if none), and the
// chunk's largest such row.
__global__ __launch_bounds__ (kBlock)
void k_event_heads (Soa s, const uint32_t *__restrict__ idx, uint32_t m, double total_time, double time_del,
                    uint32_t *__restrict__ head, uint32_t *__restrict__ chunk_max)
{
   __shared__ uint32_t wmax[kBlock / 64];
   const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
   uint32_t v = 0;
   if (q < m)
     {
        bool start = (q == 0) || !(time_del > 0.0);
        if (!start)
          start = frame_of (s, idx[q], total_time, time_del) != frame_of (s, idx[q - 1], total_time, time_del);
        v = start ? q : 0u;
     }
   v = wave_max_scan (v);
   const int w = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 63) wmax[w] = v;
   __syncthreads ();
   for (int k = 0; k < w; k++) v = max (v, wmax[k]);
   if (q < m) head[q] = v;
   if (threadIdx.x == kBlock - 1) chunk_max[blockIdx.x] = v;
}

// Second pass: pre[c] = the largest frame start in chunks before c (one
// block; each lane scans a run of consecutive chunks).
__global__ __launch_bounds__ (1024)
void k_event_chunk_prefix (const uint32_t *__restrict__ chunk_max, uint32_t nc, uint32_t *__restrict__ pre)
{
   __shared__ uint32_t wmax[1024 / 64];
   const uint32_t per = (nc + 1023) / 1024;
   const uint32_t a = threadIdx.x * per;
   const uint32_t b = min (a + per, nc);
   uint32_t run = 0;
   for (uint32_t c = a; c < b; c++) run = max (run, chunk_max[c]);
   // exclusive scan of the runs' maxima over the lanes
   uint32_t incl = wave_max_scan (run);
   const int w = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 63) wmax[w] = incl;
   __syncthreads ();
   uint32_t before = __shfl_up (incl, 1, 64);
   if ((threadIdx.x & 63) == 0) before = 0;
   for (int k = 0; k < w; k++) before = max (before, wmax[k]);
   for (uint32_t c = a; c < b; c++)
     {
        pre[c] = before;
        before = max (before, chunk_max[c]);
     }
}

__global__ __launch_bounds__ (kBlock)
void k_event_coords (const mxg_config *__restrict__ cfg, const mxg_event_params *__restrict__ params,
                     const float *__restrict__ subtab, Soa s, const uint32_t *__restrict__ idx,
                     const uint32_t *__restrict__ head, const uint32_t *__restrict__ chunk_pre, uint32_t m,
                     unsigned long long seed, EventOut out)
{
   const mxg_event_params &P = *params;
   const mxg_detector_cfg &d = cfg->det;
   const uint32_t stride = gridDim.x * blockDim.x;
   for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += stride)
     {
        const uint32_t i = idx[q];
        // the row whose dither state this row uses: its own in EXACT mode,
        // else the first of its frame
        const uint32_t h = ((P.pix_adj == MXG_PIX_ADJ_EXACT) || !P.held) ? i
          : idx[max (head[q], chunk_pre[q / kBlock])];
        const int ccd = s.ccd[i];
        out.tag[q] = s.tag[i];
        out.ccd[q] = ccd;
        const int k = ((ccd >= 0) && (ccd < MXG_MAX_CHIPS)) ? P.chip_index[ccd] : -1;
        if (k < 0)
          {
             out.tdetx[q] = out.tdety[q] = -1;
             out.detx[q] = out.dety[q] = out.x[q] = out.y[q] = 0.0;
             out.fltgrade[q] = -1;
             continue;
          }
        const mxg_chip &g = d.chip[k];
        // the columns as marx2fits reads them: float chip pixels plus one
        const float chipx = s.ypix[i] + 1.0f, chipy = s.zpix[i] + 1.0f;

        // tiled pixels (marx2fits.c:3584-3597, acis_geom.c:112-200)
        {
           const unsigned int ux = (unsigned int) chipx, uy = (unsigned int) chipy;
           float xf, yf;
           if (P.tiling == MXG_TILE_SHIFT)
             {
                xf = (float) (ux + P.tdet_xoff[ccd]);
                yf = (float) (uy + P.tdet_yoff[ccd]);
             }
           else if ((ccd == 0) || (ccd == 2))
             {
                xf = (float) (P.tdet_xoff[ccd] + uy);
                yf = (float) (P.tdet_yoff[ccd] - ux);
             }
           else
             {
                xf = (float) (P.tdet_xoff[ccd] - uy);
                yf = (float) (P.tdet_yoff[ccd] + ux);
             }
           if (xf < 0.0f) xf = 0.0f;
           if (yf < 0.0f) yf = 0.0f;
           out.tdetx[q] = (int) ((unsigned int) xf) + 1;
           out.tdety[q] = (int) ((unsigned int) yf) + 1;
        }

        // flight grade (ACIS): one of the four of its 3x3 subpixel
        Stream st;
        stream_init (st, seed, s.tag[i], MXG_STAGE_EVENT, 0u);
        const double cx = (double) chipx, cy = (double) chipy;
        int fltgrade = -1;
        if (P.has_grade)
          {
             const int sx = (int) (3.0 * (cx - floor (cx)));
             const int sy = (int) (3.0 * (cy - floor (cy)));
             // the uniform is in [0, 1]: a draw of exactly 1 takes the fourth
             // grade (marx2fits.c:3864 would index one past the row)
             fltgrade = kFltGrades[3 * sy + sx][min ((int) (4 * uniform (st)), 3)];
          }
        out.fltgrade[q] = (short) fltgrade;

        // the position within the chip pixel (marx2fits.c:3693-3717)
        double x = cx - 1.0, y = cy - 1.0;
        if (P.pix_adj == MXG_PIX_ADJ_NONE)
          {
             x = floor (x) + 0.5;
             y = floor (y) + 0.5;
          }
        else if (P.pix_adj == MXG_PIX_ADJ_RANDOMIZE)
          {
             x = floor (x) + uniform (st);
             y = floor (y) + uniform (st);
          }
        else if (P.pix_adj == MXG_PIX_ADJ_EDSER)
          {
             float dx, dy;
             subpix (P, subtab, ccd, s.pi[i], fltgrade, dx, dy);
             x = floor (x) + 0.5 + dx;
             y = floor (y) + 0.5 + dy;
          }

        // chip -> mirror nodal coordinates with the dither's detector
        // offsets (pixlib.c:134-238), then the focal plane (detpix.c:182-209)
        float dth[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (P.dither_used)
          for (int k = 0; k < 6; k++) dth[k] = s.dith[k][h];
        const double theta = (double) dth[5];
        double ct, stt;
        sincos (theta, &stt, &ct);
        const double yoff = P.det_offset[1] + (double) dth[3];
        const double zoff = P.det_offset[2] + (double) dth[4];
        vec3 ofs = v3 (P.det_offset[0] - P.focal_length, yoff + ct * g.x_ll[1] - stt * g.x_ll[2],
                       zoff + stt * g.x_ll[1] + ct * g.x_ll[2]);
        vec3 e1 = v3 (g.xhat[0], ct * g.xhat[1] - stt * g.xhat[2], stt * g.xhat[1] + ct * g.xhat[2]);
        vec3 e2 = v3 (g.yhat[0], ct * g.yhat[1] - stt * g.yhat[2], stt * g.yhat[1] + ct * g.yhat[2]);
        e1 = v3 (g.x_pixel_size * e1.x, g.x_pixel_size * e1.y, g.x_pixel_size * e1.z);
        e2 = v3 (g.y_pixel_size * e2.x, g.y_pixel_size * e2.y, g.y_pixel_size * e2.z);
        x -= P.pixel_xoff[k];
        y -= P.pixel_yoff[k];
        vec3 mnc = v3 (ofs.x + e1.x * x + e2.x * y, ofs.y + e1.y * x + e2.y * y, ofs.z + e1.z * x + e2.z * y);
        normalize_div (mnc);
        const double factor = 1.0 / (P.fp_delta_s0 * mnc.x);
        const double detx = P.fp_x0 - factor * mnc.y;
        const double dety = P.fp_y0 + factor * mnc.z;
        out.detx[q] = detx;
        out.dety[q] = dety;

        // sky pixels (marx2fits.c:3869-3910)
        if (!P.dither_used)
          {
             const double ux = detx - P.fp_x0, uy = dety - P.fp_y0;
             out.x[q] = P.fp_x0 + P.cos_roll * ux + P.sin_roll * uy;
             out.y[q] = P.fp_y0 - P.sin_roll * ux + P.cos_roll * uy;
          }
        else
          {
             vec3 u = unapply_dither ((double) dth[0], (double) dth[1], (double) dth[2], mnc);
             double ra, dec, dra, ddec;
             mnc_to_ra_dec (u, ra, dec);
             ra_dec_offsets0 (ra, dec, dra, ddec);
             dra = -dra;
             out.x[q] = dra / P.fp_delta_s0 + P.fp_x0;
             out.y[q] = ddec / P.fp_delta_s0 + P.fp_y0;
          }
     }
}

hipError_t launch_event_coords (const mxg_config *cfg, const mxg_event_params *params, const float *subtab,
                                const Soa &s, const uint32_t *idx, uint32_t m, double total_time, double time_del,
                                unsigned long long seed, uint32_t *head, uint32_t *chunk_max, uint32_t *chunk_pre,
                                const EventOut &out, hipStream_t st)
{
   if (m == 0) return hipSuccess;
   const uint32_t nc = (m + kBlock - 1) / kBlock;
   hipLaunchKernelGGL (k_event_heads, dim3 (nc), dim3 (kBlock), 0, st, s, idx, m, total_time, time_del, head,
                       chunk_max);
   hipLaunchKernelGGL (k_event_chunk_prefix, dim3 (1), dim3 (1024), 0, st, chunk_max, nc, chunk_pre);
   uint32_t blocks = nc;
   if (blocks > 8192) blocks = 8192;
   hipLaunchKernelGGL (k_event_coords, dim3 (blocks), dim3 (kBlock), 0, st, cfg, params, subtab, s, idx, head,
                       chunk_pre, m, seed, out);
   return hipGetLastError ();
}

} // namespace mxg
