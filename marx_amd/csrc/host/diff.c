/* diff.c -- host side of the differentiable mirror: the shell geometry at
 * given misalignment parameters with its derivatives, and the
 * mxg_diff_forward / mxg_diff_adjoint entry points.
 *
 * Parameters per shell pair (MXG_DIFF_NPAR): tilt az and el in arcsec, added
 * to the geometry file's az_mis / el_mis of both surfaces (the rotations of
 * marx/libsrc/hrma.c:842-897), decenter y and z and despace x in mm, moving
 * both surfaces of the pair along MARX +y, +z, +x (the offsets of
 * hrma.c:551-602; despace moves the pair's cap position as
 * HRMA_PnHn_XOffset does).
 *
 * The tilt matrices' derivatives come from complex-step differentiation of
 * the same construction: for an analytic f, f'(a) = Im f(a + i h) / h with
 * no subtractive cancellation, exact to rounding for tiny h.
 */
#include <complex.h>
#include <math.h>
#include <string.h>

#include "mxg_host.h"

typedef double complex cd;

static void c_rot_z (cd m[9], cd theta)
{
   cd c = ccos (theta), s = -csin (theta);
   m[0] = c;   m[1] = s;   m[2] = 0.0;
   m[3] = -s;  m[4] = c;   m[5] = 0.0;
   m[6] = 0.0; m[7] = 0.0; m[8] = 1.0;
}

static void c_rot (cd m[9], const cd n[3], cd theta)
{
   cd c = ccos (theta), s = csin (theta), c1 = 1 - c;
   cd n01 = n[0] * n[1] * c1, n02 = n[0] * n[2] * c1, n12 = n[1] * n[2] * c1;
   m[0] = n[0] * n[0] * c1 + c;
   m[1] = n01 - s * n[2];
   m[2] = n02 + s * n[1];
   m[3] = n01 + s * n[2];
   m[4] = n[1] * n[1] * c1 + c;
   m[5] = n12 - s * n[0];
   m[6] = n02 - s * n[1];
   m[7] = n12 + s * n[0];
   m[8] = n[2] * n[2] * c1 + c;
}

static void c_mul (cd c[9], const cd a[9], const cd b[9])
{
   cd t[9];
   for (int i = 0; i < 3; i++)
     for (int j = 0; j < 3; j++)
       t[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
   memcpy (c, t, sizeof (t));
}

/* the tilt matrices of one surface (as hrma_init.c builds them, hrma.c:842-897)
 * at complex angles in degrees */
static void c_yaw_pitch (cd az_deg, cd el_deg, cd fwd[9], cd bwd[9])
{
   cd rz[9], ry[9], y_axis[3];
   cd el = -el_deg * (PI / 180.0);
   cd az = -az_deg * (PI / 180);
   c_rot_z (rz, -az);
   y_axis[0] = rz[1];
   y_axis[1] = rz[4];
   y_axis[2] = rz[7];
   c_rot (ry, y_axis, -el);
   c_mul (fwd, ry, rz);
   c_rot (ry, y_axis, el);
   c_rot_z (rz, az);
   c_mul (bwd, rz, ry);
}

int mxg_diff_setup (const double *params, void *out)
{
   mxg_diff_cfg *d = (mxg_diff_cfg *) out;
   const mxg_hrma_cfg *h = &Mxg_Cfg.hrma;
   const double step = 1e-20;          /* complex step, arcsec */
   unsigned int i, k, j;

   if ((params == NULL) || (d == NULL)) return -1;
   memset (d, 0, sizeof (*d));
   memcpy (d->params, params, sizeof (d->params));
   for (i = 0; i < MXG_NUM_SHELLS; i++)
     {
        const mxg_shell_cfg *s = h->shell + i;
        mxg_diff_shell *ds = d->shell + i;
        const double *q = params + MXG_DIFF_NPAR * i;
        for (k = 0; k < 2; k++)
          {
             double az = s->tilt_deg[k][0] + q[0] / 3600.0;
             double el = s->tilt_deg[k][1] + q[1] / 3600.0;
             cd f[9], b[9];
             c_yaw_pitch (az, el, f, b);
             for (j = 0; j < 9; j++)
               {
                  ds->fwd[k][j] = creal (f[j]);
                  ds->bwd[k][j] = creal (b[j]);
               }
             c_yaw_pitch (az + I * (step / 3600.0), el, f, b);
             for (j = 0; j < 9; j++)
               {
                  ds->dfwd[k][0][j] = cimag (f[j]) / step;
                  ds->dbwd[k][0][j] = cimag (b[j]) / step;
               }
             c_yaw_pitch (az, el + I * (step / 3600.0), f, b);
             for (j = 0; j < 9; j++)
               {
                  ds->dfwd[k][1][j] = cimag (f[j]) / step;
                  ds->dbwd[k][1][j] = cimag (b[j]) / step;
               }
             ds->to_osac[k][0] = s->to_osac[k][0] - q[4];
             ds->to_osac[k][1] = s->to_osac[k][1] - q[2];
             ds->to_osac[k][2] = s->to_osac[k][2] - q[3];
          }
        ds->front_position = s->front_position + q[4];
     }
   return 0;
}

static int diff_pass (const char *name, const double *params, unsigned int num,
                      unsigned int tag_start, const mxg_image_spec *spec, float *image,
                      double *total_weight, int tangents)
{
   mxg_diff_cfg d;
   if ((params == NULL) || (spec == NULL) || (image == NULL) || (spec->npix == 0)
       || !(spec->pixel_size > 0.0))
     {
        mxg_error ("%s: bad arguments", name);
        return -1;
     }
   if (Mxg_Cfg.hrma.enabled == 0)
     {
        mxg_error ("Mirror not initialized.");
        return -1;
     }
   if ((Mxg_Cfg.src.spectrum != MXG_SPECTRUM_FLAT) && (Mxg_Cfg.src.spectrum != MXG_SPECTRUM_FILE))
     {
        mxg_error ("Source not initialized (marx_create_source of a POINT or DISK source with a spectrum).");
        return -1;
     }
   if ((-1 == mxg_diff_setup (params, &d))
       || (-1 == mxg_init ())
       || (-1 == mxgd_sync_config (&Mxg_Cfg, mxg_farena_base (), mxg_farena_size (),
                                   mxg_darena_base (), mxg_darena_size (), Mxg_Cfg_Generation))
       || (-1 == mxgd_diff_forward (&d, num, tag_start, mxg_get_seed (), spec, image, total_weight,
                                    tangents)))
     {
        if (mxgd_last_error ()[0]) mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

int mxg_diff_forward (const double *params, unsigned int num, unsigned int tag_start,
                      const mxg_image_spec *spec, float *image, double *total_weight)
{
   return diff_pass ("mxg_diff_forward", params, num, tag_start, spec, image, total_weight, 1);
}

int mxg_diff_image (const double *params, unsigned int num, unsigned int tag_start,
                    const mxg_image_spec *spec, float *image, double *total_weight)
{
   return diff_pass ("mxg_diff_image", params, num, tag_start, spec, image, total_weight, 0);
}

unsigned long long mxg_diff_last_pass (void)
{
   return mxgd_diff_pass ();
}

int mxg_diff_adjoint (unsigned long long pass, unsigned int npix, const float *grad_image,
                      double *grad)
{
   if ((grad_image == NULL) || (grad == NULL))
     {
        mxg_error ("mxg_diff_adjoint: bad arguments");
        return -1;
     }
   if (-1 == mxgd_diff_adjoint (pass, npix, grad_image, grad))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}

int mxg_diff_kernel_ms (float *ms, unsigned int *listed)
{
   if (ms == NULL) return -1;
   mxgd_diff_kernel_ms (ms, listed);
   return 0;
}

int mxg_diff_state (unsigned int n, double *w, double *y, double *z, double *dw, double *dy,
                    double *dz, unsigned char *shell)
{
   if (-1 == mxgd_diff_state (n, w, y, z, dw, dy, dz, shell))
     {
        mxg_error ("%s", mxgd_last_error ());
        return -1;
     }
   return 0;
}
