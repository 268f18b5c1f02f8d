/* orc_event.c -- oracle of the marx2fits event coordinates (test
 * infrastructure only, see oracle.h).
 *
 * A row-by-row restatement of what marx2fits does for each row of an ACIS or
 * HRC simulation's dump files, in its order and with its serial state
 * (marx/src/marx2fits.c): the exposure number and the held dither state
 * (compute_expno :3741-3760, read_dither_value :3567-3581; HRC re-reads
 * the dither on every row, :2850-2851), the tiled pixel (compute_tdetxy
 * :3584-3597 over acis_geom.c:112-200, hrc_s_geom.c:477-510,
 * hrc_i_geom.c:216-236), the flight grade of ACIS
 * (compute_fltgrade :3855-3866), the focal-plane pixel (compute_detxy
 * :3681-3735 over pixlib.c:133-238 and detpix.c:182-209, with the EDSER
 * offsets of acis_subpix.c:283-322; HRC randomizes instead, :3307-3308)
 * and the sky pixel (compute_xy_sky :3869-3910 over dither.c:583-607,
 * 703-707 and pixlib.c:498-573), in the detector's focal-plane system
 * (detpix.c:41-61).
 * Random choices come from the per-event streams (seed, tag, stage 6). */
#include <math.h>
#include <string.h>

#include "oracle_int.h"

#define EVENT_STAGE 6u

/* acis_geom.c:38-48, by chip id */
static const float Tdet_Xoff[10] = {3061, 5131, 3061, 5131, 791, 1833, 2875, 3917, 4959, 6001};
static const float Tdet_Yoff[10] = {5131, 4107, 4085, 3061, 1702, 1702, 1702, 1702, 1702, 1702};

/* HRC-S AXAF-HRC-2.7S y offsets of chips 1-3 (hrc_s_geom.c:321-327) */
static const float Hrc_S_Tdet_Yoff[3] = {-12.0f, 16444.0f, 32900.0f};

/* marx2fits.c:3822-3833 */
static const int Flight_Grades[9][4] =
{
   {10, 11, 138, 139}, {2, 34, 130, 162}, {18, 22, 50, 54},
   {8, 12, 136, 140}, {0, 0, 0, 0}, {16, 17, 48, 49},
   {72, 76, 104, 108}, {64, 65, 68, 69}, {80, 81, 208, 209}
};

typedef struct
{
   double ra, dec, roll, dy, dz, dtheta;
}
dither_state;

/* acis_subpix.c:283-322 over one table (energies, chip x, chip y offsets) */
int orc_subpix (const float *e, unsigned int n, float energy, float *dx, float *dy)
{
   unsigned int j;
   double w0, w1;
   if (n == 0)
     {
        *dx = *dy = 0;
        return 0;
     }
   j = orc_binary_search_f (energy, e, n);
   if (j == 0) j++;
   if (j == n) j--;
   w1 = ((double) energy - e[j - 1]) / (e[j] - e[j - 1]);
   w0 = 1.0 - w1;
   *dx = (float) (w0 * e[n + j - 1] + w1 * e[n + j]);
   *dy = (float) (w0 * e[2 * n + j - 1] + w1 * e[2 * n + j]);
   return 0;
}

/* dither.c:583-607 */
static orc_vec unapply_dither (double ra, double dec, double roll, orc_vec p)
{
   double cos_ra = cos (ra), sin_ra = sin (ra), cos_dec = cos (dec), sin_dec = sin (dec);
   double cos_theta = cos_dec * cos_ra, sin_theta;
   orc_vec n = orc_v (0, -sin_dec, cos_dec * sin_ra);
   sin_theta = orc_v_length (n);
   if (sin_theta >= 1e-20)
     {
        n.x /= sin_theta;
        n.y /= sin_theta;
        n.z /= sin_theta;
        p = orc_v_rotate_unit_vector1 (p, n, cos_theta, sin_theta);
     }
   return orc_v_rotate_unit_vector (p, orc_v (1, 0, 0), roll);
}

/* pixlib.c:498-520 */
static void mnc_to_ra_dec (orc_vec m, double *ra, double *dec)
{
   double perp = sqrt (m.x * m.x + m.y * m.y);
   if (perp > 1.0) perp = 1.0;
   if (m.z <= 0) *dec = acos (perp);
   else *dec = -acos (perp);
   perp = -m.x / perp;
   if (m.y <= 0.0) *ra = acos (perp);
   else *ra = -acos (perp);
}

/* pixlib.c:527-573 */
static void ra_dec_offsets (double ra_0, double dec_0, double ra, double dec, double *d_ra_out, double *d_dec_out)
{
   double d_ra = ra - ra_0, d_dec = dec - dec_0;
   double c = cos (dec), factor = c * (1 - cos (d_ra));
   double sin_delta_dec = sin (d_dec) + factor * sin (dec_0), num, den;
   if (fabs (sin_delta_dec) > 1.0)
     sin_delta_dec = (sin_delta_dec < 0) ? -1 : 1;
   *d_dec_out = asin (sin_delta_dec);
   num = c * sin (d_ra);
   den = cos (d_dec) - factor * cos (dec_0);
   if (den >= 0) *d_ra_out = atan (num / den);
   else if (num >= 0) *d_ra_out = atan (num / den) + PI;
   else *d_ra_out = atan (num / den) - PI;
}

int orc_event_coords (const mxg_config *cfg, const mxg_event_cfg *ec, unsigned int n,
                      const float *time, const unsigned int *tag, const signed char *ccd,
                      const float *xpixel, const float *ypixel, const float *b_energy,
                      const float *const *dither, const float *subtab,
                      const unsigned int *sub_off, const unsigned int *sub_n,
                      int *tdetx, int *tdety, double *detx, double *dety,
                      double *sky_x, double *sky_y, short *fltgrade)
{
   const mxg_detector_cfg *d = &cfg->det;
   const int type = d->type;
   const int is_acis_i = (type == MARX_DETECTOR_ACIS_I);
   const int is_acis = is_acis_i || (type == MARX_DETECTOR_ACIS_S);
   double fp_delta_s0, fp_x0, fp_y0;
   double time_del = ec->exposure_time;
   long last_expno = -1;
   int pix_adj = ec->pix_adj;
   dither_state ds;
   unsigned int i;

   /* detpix.c:43-61: AXAF-FP-1.1 (ACIS), 2.1 (HRC-I), 2.3 (HRC-S) */
   if (is_acis)
     {
        fp_delta_s0 = 0.492 * (PI / (3600.0 * 180.0));
        fp_x0 = fp_y0 = 4096.5;
     }
   else if (type == MARX_DETECTOR_HRC_I)
     {
        fp_delta_s0 = 0.13175 * (PI / (3600.0 * 180.0));
        fp_x0 = fp_y0 = 16384.5;
     }
   else if (type == MARX_DETECTOR_HRC_S)
     {
        fp_delta_s0 = 0.1318 * (PI / (3600.0 * 180.0));
        fp_x0 = fp_y0 = 32768.5;
     }
   else
     return -1;
   if (!is_acis && (pix_adj == MXG_PIX_ADJ_EDSER))
     pix_adj = MXG_PIX_ADJ_RANDOMIZE;
   if ((ec->exposure_time > 0.0) && (ec->frame_transfer_time > 0.0))
     time_del += ec->frame_transfer_time;
   memset (&ds, 0, sizeof (ds));

   for (i = 0; i < n; i++)
     {
        int id = ccd[i], update, k, kc = -1;
        float chipx = xpixel[i] + 1, chipy = ypixel[i] + 1;
        double x, y;
        orc_stream st;
        const mxg_chip *g;

        /* compute_expno (ACIS; HRC updates the dither on every row) */
        if (!is_acis)
          update = 1;
        else if (time_del <= 0.0)
          {
             last_expno++;
             update = 1;
          }
        else
          {
             long expno = (long) ((double) time[i] / time_del);
             update = (expno != last_expno);
             last_expno = expno;
          }
        /* read_dither_value: zero without dither; re-read on a new frame
         * or in EXACT mode */
        if (!ec->dither_used)
          memset (&ds, 0, sizeof (ds));
        else if (update || (pix_adj == MXG_PIX_ADJ_EXACT))
          {
             ds.ra = dither[0][i];
             ds.dec = dither[1][i];
             ds.roll = dither[2][i];
             ds.dy = dither[3][i];
             ds.dz = dither[4][i];
             ds.dtheta = dither[5][i];
          }

        for (k = 0; k < d->num_chips; k++)
          if ((d->chip[k].id == id) && ((type != MARX_DETECTOR_ACIS_I) || (id <= 3))
              && ((type != MARX_DETECTOR_ACIS_S) || (id >= 4)))
            kc = k;
        if (kc < 0)
          {
             tdetx[i] = tdety[i] = -1;
             detx[i] = dety[i] = sky_x[i] = sky_y[i] = 0.0;
             fltgrade[i] = -1;
             continue;
          }
        g = &d->chip[kc];

        /* compute_tdetxy */
        {
           unsigned int ux = (unsigned int) chipx, uy = (unsigned int) chipy;
           float xf, yf;
           if (type == MARX_DETECTOR_HRC_S)
             {
                yf = uy;
                yf = yf + Hrc_S_Tdet_Yoff[id - 1];
                xf = ux + 0.0f;
             }
           else if (type == MARX_DETECTOR_HRC_I)
             {
                xf = ux + (-1.0f);
                yf = uy + 0.0f;
             }
           else if (!is_acis_i)
             {
                xf = ux + Tdet_Xoff[id];
                yf = uy + Tdet_Yoff[id];
             }
           else if ((id == 0) || (id == 2))
             {
                xf = Tdet_Xoff[id] + uy;
                yf = Tdet_Yoff[id] - ux;
             }
           else
             {
                xf = Tdet_Xoff[id] - uy;
                yf = Tdet_Yoff[id] + ux;
             }
           if (xf < 0.0) xf = 0.0;
           if (yf < 0.0) yf = 0.0;
           tdetx[i] = (int) (unsigned int) xf + 1;
           tdety[i] = (int) (unsigned int) yf + 1;
        }

        /* compute_fltgrade (ACIS) */
        orc_stream_init (&st, ec->seed, tag[i], EVENT_STAGE, 0);
        fltgrade[i] = -1;
        if (is_acis)
        {
           int dx = (int) (3.0 * (chipx - floor (chipx)));
           int dy = (int) (3.0 * (chipy - floor (chipy)));
           int pick = (int) (4 * orc_uniform (&st));
           /* a uniform of exactly 1 takes the fourth grade (marx2fits.c:3864
            * would index one past the row) */
           if (pick > 3) pick = 3;
           fltgrade[i] = (short) Flight_Grades[3 * dy + dx][pick];
        }

        /* compute_detxy */
        {
           double ct = cos (ds.dtheta), s = sin (ds.dtheta);
           double yoff = ec->det_offset[1] + ds.dy, zoff = ec->det_offset[2] + ds.dz;
           orc_vec ofs, e1, e2, mnc;
           double factor;
           ofs.x = ec->det_offset[0] - ec->focal_length;
           ofs.y = yoff + ct * g->x_ll[1] - s * g->x_ll[2];
           ofs.z = zoff + s * g->x_ll[1] + ct * g->x_ll[2];
           e1 = orc_v (g->xhat[0], ct * g->xhat[1] - s * g->xhat[2], s * g->xhat[1] + ct * g->xhat[2]);
           e2 = orc_v (g->yhat[0], ct * g->yhat[1] - s * g->yhat[2], s * g->yhat[1] + ct * g->yhat[2]);
           e1 = orc_v (g->x_pixel_size * e1.x, g->x_pixel_size * e1.y, g->x_pixel_size * e1.z);
           e2 = orc_v (g->y_pixel_size * e2.x, g->y_pixel_size * e2.y, g->y_pixel_size * e2.z);

           x = chipx - 1.0;
           y = chipy - 1.0;
           switch (pix_adj)
             {
              case MXG_PIX_ADJ_EXACT:
                break;
              case MXG_PIX_ADJ_NONE:
                x = floor (x) + 0.5;
                y = floor (y) + 0.5;
                break;
              case MXG_PIX_ADJ_RANDOMIZE:
                x = floor (x) + orc_uniform (&st);
                y = floor (y) + orc_uniform (&st);
                break;
              case MXG_PIX_ADJ_EDSER:
                {
                   int t = ((id == 5) || (id == 7)) ? 1 : 0;
                   float sdx, sdy;
                   orc_subpix (subtab + sub_off[256 * t + fltgrade[i]], sub_n[256 * t + fltgrade[i]],
                               b_energy[i], &sdx, &sdy);
                   x = floor (x) + 0.5 + sdx;
                   y = floor (y) + 0.5 + sdy;
                }
                break;
             }
           if (type == MARX_DETECTOR_HRC_S)
             {
                /* the chip pixel of the facet's corner, as
                 * _marx_hrc_s_compute_pixel (id, 0, 0) gives it
                 * (hrc_s_geom.c:329-336, 344-394) */
                x -= d->pix_x0[id] + (d->pix_u[id] - d->pix_u0[id]);
                y -= d->pix_y0[id] + (d->pix_v[id] - d->pix_v0[id]);
             }
           else
             {
                x -= g->xpixel_offset;
                y -= g->ypixel_offset;
             }
           mnc.x = ofs.x + e1.x * x + e2.x * y;
           mnc.y = ofs.y + e1.y * x + e2.y * y;
           mnc.z = ofs.z + e1.z * x + e2.z * y;
           orc_v_normalize (&mnc);
           factor = 1.0 / (fp_delta_s0 * mnc.x);
           detx[i] = fp_x0 - factor * mnc.y;
           dety[i] = fp_y0 + factor * mnc.z;

           /* compute_xy_sky */
           if (!ec->dither_used)
             {
                double theta = ec->roll_nom * PI / 180.0;
                double c = cos (theta), sn = sin (theta);
                double ux = detx[i] - fp_x0, uy = dety[i] - fp_y0;
                sky_x[i] = fp_x0 + c * ux + sn * uy;
                sky_y[i] = fp_y0 - sn * ux + c * uy;
             }
           else
             {
                double ra, dec;
                mnc = unapply_dither (ds.ra, ds.dec, ds.roll, mnc);
                mnc_to_ra_dec (mnc, &ra, &dec);
                ra_dec_offsets (0, 0, ra, dec, &ra, &dec);
                ra = -ra;
                sky_x[i] = ra / fp_delta_s0 + fp_x0;
                sky_y[i] = dec / fp_delta_s0 + fp_y0;
             }
        }
     }
   return 0;
}
