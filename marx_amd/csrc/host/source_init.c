/* source_init.c -- source, spectrum and dither initialisation, and the
 * marx_create_source / marx_create_photons entry points.
 *
 * Follows marx/libsrc/source.c:44-417 (select_source, marx_compute_elaz,
 * marx_create_photons), s-point.c:59-97, s-disk.c:63-127,
 * spectrum.c:48-300 (FLAT and FILE spectra, rectangle-rule cumulative),
 * dither.c:111-155 (setup_pointing_ra_dec), :350-450 (_marx_init_dither,
 * INTERNAL model) and pixlib.c:531-577 (marx_compute_ra_dec_offsets).
 *
 * Photons are generated on the GPU from per-photon counter streams; see
 * the mxg_source_cfg fields for what the kernels read.
 */
#include <stdio.h>
#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "mxg_host.h"

struct _Marx_Source_Type
{
   int open;
   /* RAYFILE sources (s-rayfile.c:148-254) */
   FILE *ray_fp;
   unsigned long ray_history;
   double ray_last_time;
};

/* The observation start in MJD seconds and in years, as the reference's
 * driver sets them through marx_set_time (marx.c:285-373 -> util.c:231-236).
 * The start enters the exposure cut of sources whose photons carry their own
 * times (source.c:295-309) and the ACIS contamination's time interpolation
 * (aciscontam.c:236-269). */
double _Marx_TStart_MJDsecs = 0.0;
double _Marx_TStart_Yrs = 0.0;

int marx_set_time (double tyrs, double mjdsecs)
{
   _Marx_TStart_Yrs = tyrs;
   _Marx_TStart_MJDsecs = mjdsecs;
   return 0;
}

static char Ray_File[1024];

static double Nominal_Ra, Nominal_Dec, Nominal_Roll;
static double Pointing_Ra, Pointing_Dec, Pointing_Roll;

void mxg_get_nominal_pointing (double *ra, double *dec, double *roll)
{
   *ra = Nominal_Ra; *dec = Nominal_Dec; *roll = Nominal_Roll;
}

void mxg_get_pointing (double *ra, double *dec, double *roll)
{
   *ra = Pointing_Ra; *dec = Pointing_Dec; *roll = Pointing_Roll;
}

/* x to [-PI, PI) */
static void normalize_angle (double *xp)
{
   double x = fmod (*xp, 2 * PI);
   if (x >= PI) x -= (2 * PI);
   if (x < -PI) x += (2 * PI);
   *xp = x;
}

static void spherical_to_triad (double theta, double phi, JDMVector_Type *r_hat,
                                JDMVector_Type *theta_hat, JDMVector_Type *phi_hat)
{
   double ct = cos (theta), st = sin (theta), cp = cos (phi), sp = sin (phi);
   *r_hat = mxg_v (st * cp, st * sp, ct);
   *phi_hat = mxg_v (-sp, cp, 0.0);
   *theta_hat = mxg_v (ct * cp, ct * sp, -st);
}

static void setup_pointing_ra_dec (double offset_y, double offset_z)
{
   JDMVector_Type p, nom, a, d;
   double cos_delta, alpha = offset_y, delta = offset_z;

   spherical_to_triad (PI / 2.0 - Nominal_Dec, Nominal_Ra, &nom, &d, &a);
   d.x = -d.x; d.y = -d.y; d.z = -d.z;
   cos_delta = cos (delta);
   {
      double a1 = cos_delta * cos (alpha), b1 = cos_delta * sin (alpha), c1 = sin (delta);
      p = mxg_v (a1 * nom.x + b1 * a.x + c1 * d.x,
                 a1 * nom.y + b1 * a.y + c1 * d.y,
                 a1 * nom.z + b1 * a.z + c1 * d.z);
   }
   p = mxg_v_rotate_unit_vector (p, nom, Nominal_Roll);
   mxg_v_unit_vector_to_spherical (p, &Pointing_Dec, &Pointing_Ra);
   Pointing_Dec = PI * 0.5 - Pointing_Dec;
   Pointing_Roll = Nominal_Roll;
}

int mxg_dither_init (Param_File_Type *pf, double *yoff, double *zoff)
{
   mxg_source_cfg *s = &Mxg_Cfg.src;
   char model[64];
   double offset_y, offset_z;

   if ((-1 == pf_get_string (pf, "DitherModel", model, sizeof (model)))
       || (-1 == pf_get_double (pf, "DitherAmp_RA", &s->ra_amp))
       || (-1 == pf_get_double (pf, "DitherAmp_Dec", &s->dec_amp))
       || (-1 == pf_get_double (pf, "DitherAmp_Roll", &s->roll_amp))
       || (-1 == pf_get_double (pf, "DitherPeriod_RA", &s->ra_period))
       || (-1 == pf_get_double (pf, "DitherPeriod_Dec", &s->dec_period))
       || (-1 == pf_get_double (pf, "DitherPeriod_Roll", &s->roll_period))
       || (-1 == pf_get_double (pf, "DitherPhase_RA", &s->ra_phase))
       || (-1 == pf_get_double (pf, "DitherPhase_Dec", &s->dec_phase))
       || (-1 == pf_get_double (pf, "DitherPhase_Roll", &s->roll_phase))
       || (-1 == pf_get_double (pf, "AspectBlur", &s->aspect_blur))
       || (-1 == pf_get_double (pf, "RA_Nom", &Nominal_Ra))
       || (-1 == pf_get_double (pf, "Dec_Nom", &Nominal_Dec))
       || (-1 == pf_get_double (pf, "Roll_Nom", &Nominal_Roll))
       || (-1 == pf_get_double (pf, "PointingOffsetY", &offset_y))
       || (-1 == pf_get_double (pf, "PointingOffsetZ", &offset_z)))
     {
        mxg_error ("error getting dither parameters");
        return -1;
     }
   s->aspect_blur *= PI / (180.0 * 3600.0);
   if (s->aspect_blur < 0) s->aspect_blur = 0.0;
   offset_y *= PI / (180.0 * 3600.0);
   offset_z *= PI / (180.0 * 3600.0);
   *yoff = offset_y;
   *zoff = offset_z;
   Nominal_Ra *= PI / 180.0;
   Nominal_Dec *= PI / 180.0;
   Nominal_Roll *= PI / 180.0;
   normalize_angle (&Nominal_Roll);
   s->nominal_roll = Nominal_Roll;

   if (0 == strcmp (model, "NONE"))
     {
        s->dither_mode = MXG_DITHER_NONE;
        Pointing_Roll = 0.0;
        setup_pointing_ra_dec (offset_y, offset_z);
        return 0;
     }
   if (0 == strcmp (model, "INTERNAL"))
     {
        s->dither_mode = MXG_DITHER_INTERNAL;
        s->ra_amp = s->ra_amp * (PI / (180.0 * 3600));
        s->dec_amp = s->dec_amp * (PI / (180.0 * 3600));
        s->roll_amp = s->roll_amp * (PI / (180.0 * 3600));
        setup_pointing_ra_dec (offset_y, offset_z);
        return 0;
     }
   mxg_error ("DitherModel = %s is not supported (NONE and INTERNAL are)", model);
   return -1;
}

static void compute_ra_dec_offsets (double ra_0, double dec_0, double ra, double dec,
                                    double *delta_ra, double *delta_dec)
{
   double d_ra = ra - ra_0, d_dec = dec - dec_0;
   double c = cos (dec);
   double factor = c * (1 - cos (d_ra));
   double sin_delta_dec = sin (d_dec) + factor * sin (dec_0);
   double num, den;
   if (fabs (sin_delta_dec) > 1.0)
     sin_delta_dec = (sin_delta_dec < 0) ? -1 : 1;
   *delta_dec = asin (sin_delta_dec);
   num = c * sin (d_ra);
   den = cos (d_dec) - factor * cos (dec_0);
   if (den >= 0) *delta_ra = atan (num / den);
   else if (num >= 0) *delta_ra = atan (num / den) + PI;
   else *delta_ra = atan (num / den) - PI;
}

static void compute_elaz (double ra, double dec, double *azp, double *elp)
{
   JDMVector_Type src, pnt;
   double az, el;
   src = mxg_v_spherical_to_vector (1.0, PI / 2.0 - dec, ra);
   pnt = mxg_v_spherical_to_vector (1.0, PI / 2.0 - Pointing_Dec, Pointing_Ra);
   src = mxg_v_rotate_unit_vector (src, pnt, -Pointing_Roll);
   mxg_v_unit_vector_to_spherical (src, &el, &az);
   el = PI / 2.0 - el;
   compute_ra_dec_offsets (Pointing_Ra, Pointing_Dec, az, el, &az, &el);
   *elp = el;
   *azp = az;
}

/* Cumulative spectrum by the rectangle rule, normalised (spectrum.c:48-92) */
static int file_spectrum (Param_File_Type *pf, mxg_source_cfg *s)
{
   char buf[1024], *file;
   double *cols[2], tot = 0.0, prev_x, flux;
   unsigned int n, i;
   double *cum;

   if ((-1 == pf_get_file (pf, "SpectrumFile", buf, sizeof (buf)))
       || (-1 == pf_get_double (pf, "SourceFlux", &flux)))
     return -1;
   file = (buf[0] == '/') ? strdup (buf) : mxg_data_file_name (buf);
   if ((buf[0] != '/') && (NULL != file))
     {
        FILE *fp = fopen (buf, "r");   /* relative to the working directory first */
        if (fp != NULL)
          {
             fclose (fp);
             free (file);
             file = strdup (buf);
          }
     }
   if (file == NULL) return -1;
   if ((-1 == mxg_read_columns (file, 2, cols, &n)) || (n == 0))
     {
        mxg_error ("Error reading spectrum from %s", file);
        free (file);
        return -1;
     }
   free (file);
   cum = (double *) malloc (n * sizeof (double));
   if (cum == NULL) return -1;
   prev_x = cols[0][0];
   for (i = 0; i < n; i++)
     {
        double this_x = cols[0][i];
        if (this_x - prev_x < 0.0)
          {
             mxg_error ("Your file spectrum appears corrupt.  Perhaps you are using a wavelength grid");
             free (cum); free (cols[0]); free (cols[1]);
             return -1;
          }
        tot += cols[1][i] * (this_x - prev_x);
        cum[i] = tot;
        prev_x = this_x;
     }
   if (tot == 0.0)
     {
        mxg_error ("Your file spectrum seems to have zero flux in it.");
        free (cum); free (cols[0]); free (cols[1]);
        return -1;
     }
   for (i = 0; i < n; i++) cum[i] = cum[i] / tot;
   if (flux <= 0.0) flux = tot;
   if ((flux <= 0.0) || !isfinite (flux))
     {
        mxg_error ("The flux must be positive and finite.  Check your spectrum file");
        free (cum); free (cols[0]); free (cols[1]);
        return -1;
     }
   s->file_n = n;
   s->file_e_off = mxg_darena_push (cols[0], n);
   s->file_cum_off = mxg_darena_push (cum, n);
   s->total_flux = flux;
   s->spectrum = MXG_SPECTRUM_FILE;
   /* the energy draw searches the cumulative table by bucket */
   if (-1 == mxg_build_lbucket (cum, n, MXG_BUCKETS_SPECTRUM, &s->file_bucket))
     {
        free (cum); free (cols[0]); free (cols[1]);
        return -1;
     }
   free (cum); free (cols[0]); free (cols[1]);
   return 0;
}

static int spectrum_init (Param_File_Type *pf, mxg_source_cfg *s, const char *name)
{
   char buf[64];
   if (-1 == pf_get_string (pf, "SpectrumType", buf, sizeof (buf)))
     return -1;
   if (!strcmp (buf, "FLAT"))
     {
        double emin, emax, flux;
        if ((-1 == pf_get_double (pf, "MinEnergy", &emin))
            || (-1 == pf_get_double (pf, "MaxEnergy", &emax))
            || (-1 == pf_get_double (pf, "SourceFlux", &flux)))
          return -1;
        if (flux <= 0.0)
          {
             mxg_error ("SourceFlux for FLAT spectrum must be positive, not %e", flux);
             return -1;
          }
        if ((emax - emin < 0.0) || (emax <= 0.0))
          {
             mxg_error ("MinEnergy/MaxEnergy out of range");
             return -1;
          }
        s->spectrum = MXG_SPECTRUM_FLAT;
        s->emin = emin;
        s->emax = emax;
        s->total_flux = flux;
        return 0;
     }
   if (!strcmp (buf, "FILE"))
     return file_spectrum (pf, s);
   mxg_error ("SpectrumType: \"%s\" not supported by \"%s\" source", buf, name);
   return -1;
}

/* The angle and unit axis that rotate a onto b (JDMv_find_rotation_axis,
 * jdmath/src/vector.c:289-331, with its choice of axis when a and b are
 * parallel). */
static double rotation_axis (JDMVector_Type a, JDMVector_Type b, JDMVector_Type *c)
{
   double theta = mxg_v_dot (a, b), len;
   JDMVector_Type cc;
   if (theta > 1.0) theta = 1.0;
   else if (theta < -1.0) theta = -1.0;
   theta = acos (theta);
   cc = mxg_v_cross (a, b);
   len = mxg_v_length (cc);
   if (len > 0.0)
     {
        c->x = cc.x / len;
        c->y = cc.y / len;
        c->z = cc.z / len;
        return theta;
     }
   if ((a.x != 0.0) || (a.y != 0.0))
     *c = mxg_v (-a.y, -a.x, 0.0);
   else
     *c = mxg_v (-a.z, 0.0, 0.0);
   mxg_v_normalize (c);
   return theta;
}

/* ---- USER (s-user.c) -------------------------------------------------------------- */

static void *User_Handle;
static int (*User_Open_Source) (char **, int, double, double, double, double);
static void (*User_Close_Source) (void);
static int (*User_Create_Ray) (double *, double *, double *, double *, double *);
static int (*User_Start_Iteration) (void);
static char User_Args[1024];
static char *User_Argv[64];
static int User_Argc;
static int User_Opened;

static void user_source_close (void)
{
   if (User_Opened && (User_Close_Source != NULL))
     (*User_Close_Source) ();
   if (User_Handle != NULL)
     dlclose (User_Handle);
   User_Handle = NULL;
   User_Open_Source = NULL;
   User_Close_Source = NULL;
   User_Create_Ray = NULL;
   User_Start_Iteration = NULL;
   User_Argc = 0;
   User_Opened = 0;
}

/* s-user.c:229-300: link the plug-in named by UserSourceFile; its
 * arguments are UserSourceArgs split at white space. */
static int user_source_init (Param_File_Type *pf)
{
   char file[1024], *p;
   user_source_close ();
   if (-1 == pf_get_file (pf, "UserSourceFile", file, sizeof (file)))
     {
        mxg_error ("Unable to find parameter 'UserSourceFile'");
        return -1;
     }
   User_Args[0] = 0;
   if (-1 == pf_get_string (pf, "UserSourceArgs", User_Args, sizeof (User_Args)))
     return -1;
   mxg_message (1, "Dynamically linking to file %s\n", file);
   if (NULL == (User_Handle = dlopen (file, RTLD_LAZY)))
     {
        const char *err = dlerror ();
        mxg_error ("Error linking to %s\nReason: %s", file, err ? err : "UNKNOWN");
        return -1;
     }
   *(void **) (&User_Open_Source) = dlsym (User_Handle, "user_open_source");
   *(void **) (&User_Close_Source) = dlsym (User_Handle, "user_close_source");
   *(void **) (&User_Create_Ray) = dlsym (User_Handle, "user_create_ray");
   *(void **) (&User_Start_Iteration) = dlsym (User_Handle, "user_start_iteration");
   if ((User_Open_Source == NULL) || (User_Close_Source == NULL) || (User_Create_Ray == NULL))
     {
        mxg_error ("Unable to get symbol '%s'", (User_Open_Source == NULL) ? "user_open_source"
                   : (User_Close_Source == NULL) ? "user_close_source" : "user_create_ray");
        user_source_close ();
        return -1;
     }
   p = User_Args;
   while (*p && (User_Argc < 63))
     {
        while ((*p == ' ') || (*p == '\t') || (*p == '\n')) p++;
        if (*p == 0) break;
        User_Argv[User_Argc++] = p;
        while (*p && (*p != ' ') && (*p != '\t') && (*p != '\n')) p++;
        if (*p) *p++ = 0;
     }
   User_Argv[User_Argc] = NULL;
   return 0;
}

/* s-user.c:122-130: user_open_source (argv, argc, area, source direction) */
static int user_source_open (void)
{
   const mxg_source_cfg *s = &Mxg_Cfg.src;
   if (User_Open_Source == NULL) return -1;
   if (User_Opened) return 0;
   if (0 != (*User_Open_Source) (User_Argv, User_Argc, Marx_Mirror_Geometric_Area, s->p[0], s->p[1], s->p[2]))
     {
        mxg_error ("user_open_source failed");
        return -1;
     }
   User_Opened = 1;
   return 0;
}

int mxg_user_start_iteration (void)
{
   if (-1 == user_source_open ()) return -1;
   if (User_Start_Iteration == NULL) return 0;
   return (*User_Start_Iteration) ();
}

int mxg_user_create_ray (double *dt, double *energy, double *cx, double *cy, double *cz)
{
   return (*User_Create_Ray) (dt, energy, cx, cy, cz);
}

/* ---- IMAGE (s-image.c) ------------------------------------------------------------- */

/* the value of keyword key in a FITS header of ncards cards, or -1 */
static int fits_card (const char *hdr, unsigned int ncards, const char *key, char *val, unsigned int vlen)
{
   unsigned int i, n = (unsigned int) strlen (key);
   for (i = 0; i < ncards; i++)
     {
        const char *c = hdr + 80 * i;
        unsigned int j = 0, k;
        if ((0 != strncmp (c, key, n)) || ((n < 8) && (c[n] != ' ')) || (c[8] != '=')) continue;
        for (k = 10; (k < 80) && (c[k] == ' '); k++) ;
        while ((k < 80) && (c[k] != '/') && (j + 1 < vlen)) val[j++] = c[k++];
        while ((j > 0) && (val[j - 1] == ' ')) j--;
        val[j] = 0;
        return 0;
     }
   return -1;
}

static int fits_double (const char *hdr, unsigned int ncards, const char *key, double *v)
{
   char buf[80], *end;
   if (-1 == fits_card (hdr, ncards, key, buf, sizeof (buf))) return -1;
   *v = strtod (buf, &end);
   return (end == buf) ? -1 : 0;
}

/* The primary image of a FITS file as the cumulative distribution of its
 * pixels in the float arena, and its pixel scale (s-image.c:44-134,
 * 162-238): BITPIX 8, 16, 32, -32 or -64 (values as stored), CDELT1/2 or
 * the CD matrix, else 0.5 arcsec; the running sum is kept in double,
 * stored as float, then divided by the total. */
static int read_image_source (const char *file, mxg_source_cfg *s)
{
   FILE *fp = fopen (file, "rb");
   char *hdr = NULL, val[80];
   unsigned int ncards = 0, done = 0, i, n, nx, ny, bytes;
   double naxis1, naxis2, bitpix, s1, s2, sum;
   unsigned char *buf = NULL;
   float *cdf;

   if (fp == NULL)
     {
        mxg_error ("Failed to open %s", file);
        return -1;
     }
   while (!done)
     {
        char *h = (char *) realloc (hdr, 80 * (ncards + 36));
        if (h == NULL) goto nomem;
        hdr = h;
        if (2880 != fread (hdr + 80 * ncards, 1, 2880, fp))
          {
             mxg_error ("%s: truncated FITS header", file);
             goto fail;
          }
        for (i = 0; i < 36; i++)
          if (0 == strncmp (hdr + 80 * (ncards + i), "END     ", 8)) done = 1;
        ncards += 36;
     }
   if ((-1 == fits_double (hdr, ncards, "NAXIS1", &naxis1))
       || (-1 == fits_double (hdr, ncards, "NAXIS2", &naxis2))
       || (-1 == fits_double (hdr, ncards, "BITPIX", &bitpix)))
     {
        mxg_error ("%s: NAXIS1, NAXIS2 or BITPIX missing", file);
        goto fail;
     }
   if (0 == fits_double (hdr, ncards, "CDELT1", &s1))
     {
        if (-1 == fits_double (hdr, ncards, "CDELT2", &s2))
          {
             mxg_error ("Error locating keyword CDELT2");
             goto fail;
          }
        s1 = fabs (s1);
        s2 = fabs (s2);
     }
   else if (0 == fits_double (hdr, ncards, "CD1_1", &s1))
     {
        double c12, c21, c22;
        if ((-1 == fits_double (hdr, ncards, "CD1_2", &c12))
            || (-1 == fits_double (hdr, ncards, "CD2_1", &c21))
            || (-1 == fits_double (hdr, ncards, "CD2_2", &c22)))
          {
             mxg_error ("Error locating a CD keyword");
             goto fail;
          }
        s2 = sqrt (c21 * c21 + c22 * c22);
        s1 = sqrt (s1 * s1 + c12 * c12);
     }
   else s1 = s2 = 0.0;
   if ((s1 == 0.0) || (s2 == 0.0))
     {
        mxg_message (0, "Cannot determine scale of image.  Assuming 0.5 arc-sec per pixel\n");
        s1 = s2 = 0.5 / 3600.0;
     }
   (void) val;
   nx = (unsigned int) naxis1;
   ny = (unsigned int) naxis2;
   n = nx * ny;
   switch ((int) bitpix)
     {
      case 8: bytes = 1; break;
      case 16: bytes = 2; break;
      case 32: case -32: bytes = 4; break;
      case -64: bytes = 8; break;
      default:
        mxg_error ("BITPIX = %d is not supported", (int) bitpix);
        goto fail;
     }
   if ((n == 0) || (n > (1u << 28)))
     {
        mxg_error ("%s: image of %u x %u pixels not supported", file, nx, ny);
        goto fail;
     }
   if (NULL == (buf = (unsigned char *) malloc ((size_t) n * bytes))) goto nomem;
   if (n != fread (buf, bytes, n, fp))
     {
        mxg_error ("Error reading image file");
        goto fail;
     }
   s->image_off = mxg_farena_reserve (n);
   if (s->image_off == (unsigned int) -1) goto nomem;
   cdf = mxg_farena_ptr (s->image_off);
   sum = 0.0;
   for (i = 0; i < n; i++)
     {
        const unsigned char *b = buf + (size_t) i * bytes;
        double v;
        if (bytes == 1) v = (double) (float) b[0];
        else if (bytes == 2) v = (double) (float) (short) ((b[0] << 8) | b[1]);
        else
          {
             unsigned long long u = 0;
             unsigned int k;
             for (k = 0; k < bytes; k++) u = (u << 8) | b[k];
             if (bitpix == 32) v = (double) (float) (int) (unsigned int) u;
             else if (bitpix == -32)
               {
                  unsigned int w = (unsigned int) u;
                  float f;
                  memcpy (&f, &w, 4);
                  v = (double) f;
               }
             else
               {
                  double d;
                  memcpy (&d, &u, 8);
                  v = (double) (float) d;
               }
          }
        sum += v;
        cdf[i] = (float) sum;
     }
   for (i = 0; i < n; i++)
     cdf[i] = (float) (cdf[i] / sum);
   s->image_n = n;
   s->image_nx = nx;
   s->image_ny = ny;
   s->image_xscale = s1 * PI / 180.0;
   s->image_yscale = s2 * PI / 180.0;
   free (buf);
   free (hdr);
   fclose (fp);
   return 0;
nomem:
   mxg_error ("Not enough memory for the image source");
fail:
   free (buf);
   free (hdr);
   fclose (fp);
   return -1;
}

int mxg_source_init (Param_File_Type *pf)
{
   mxg_source_cfg *s = &Mxg_Cfg.src;
   char name[64];
   double ra, dec, dist, yoff, zoff, az, el;
   JDMVector_Type p;

   memset (s, 0, sizeof (*s));
   if ((-1 == pf_get_double (pf, "SourceDistance", &dist))
       || (-1 == pf_get_double (pf, "SourceRA", &ra))
       || (-1 == pf_get_double (pf, "SourceDEC", &dec))
       || (-1 == pf_get_string (pf, "SourceType", name, sizeof (name))))
     return -1;
   if (!strcmp (name, "POINT")) s->type = MXG_SOURCE_POINT;
   else if (!strcmp (name, "DISK")) s->type = MXG_SOURCE_DISK;
   else if (!strcmp (name, "RAYFILE")) s->type = MXG_SOURCE_RAYFILE;
   else if (!strcmp (name, "GAUSS")) s->type = MXG_SOURCE_GAUSS;
   else if (!strcmp (name, "BETA")) s->type = MXG_SOURCE_BETA;
   else if (!strcmp (name, "LINE")) s->type = MXG_SOURCE_LINE;
   else if (!strcmp (name, "IMAGE")) s->type = MXG_SOURCE_IMAGE;
   else if (!strcmp (name, "USER")) s->type = MXG_SOURCE_USER;
   else
     {
        mxg_error ("SourceType %s not supported (POINT, DISK, GAUSS, BETA, LINE, IMAGE, USER and RAYFILE "
                   "are)", name);
        return -1;
     }
   if (s->type == MXG_SOURCE_RAYFILE)
     {
        /* s-rayfile.c:226-254: the photons come from the file */
        int dump;
        if ((-1 != pf_get_boolean (pf, "DumpToRayFile", &dump)) && dump)
          {
             mxg_error ("You cannot specify a RAYFILE source AND dump to a ray file.");
             return -1;
          }
        if (-1 == pf_get_string (pf, "RayFile", Ray_File, sizeof (Ray_File)))
          return -1;
        if (-1 == mxg_dither_init (pf, &yoff, &zoff))
          return -1;
        s->distance = dist * 1000.0;
        Mxg_Cfg_Generation++;
        return 0;
     }
   dist *= 1000.0;
   ra *= (PI / 180.0);
   dec *= (PI / 180.0);

   if (-1 == mxg_dither_init (pf, &yoff, &zoff))
     return -1;
   compute_elaz (ra, dec, &az, &el);

   p = mxg_v_spherical_to_vector (1.0, 0.5 * PI - zoff, yoff);
   p = mxg_v_rotate_unit_vector (p, mxg_v (0, -1, 0), el);
   p = mxg_v_rotate_unit_vector (p, mxg_v (0, 0, 1), az);
   /* roll so the point is invariant under the (average) dither roll */
   p = mxg_v_rotate_unit_vector (p, mxg_v (1, 0, 0), Nominal_Roll);
   s->p[0] = -p.x; s->p[1] = -p.y; s->p[2] = -p.z;
   if (p.x <= 0.0)
     {
        mxg_error ("Source is located at RA=%f, Dec=%f: rays will not hit the telescope",
                   ra * 180.0 / PI, dec * 180.0 / PI);
        return -1;
     }
   {
      JDMVector_Type nrm = mxg_v (-p.y / p.x, 1.0, 0.0);
      mxg_v_normalize (&nrm);
      s->p_normal[0] = nrm.x; s->p_normal[1] = nrm.y; s->p_normal[2] = nrm.z;
   }
   s->distance = dist;

   if (s->type == MXG_SOURCE_DISK)
     {
        double th1, th0;
        if ((-1 == pf_get_double (pf, "S-DiskTheta1", &th1))
            || (-1 == pf_get_double (pf, "S-DiskTheta0", &th0)))
          return -1;
        s->disk_theta = th1 * (1.0 / 3600.0 * PI / 180.0);
        s->disk_theta_min = th0 * (1.0 / 3600.0 * PI / 180.0);
     }
   else if (s->type == MXG_SOURCE_GAUSS)
     {
        /* s-gauss.c:146-160, 87-88 */
        if (-1 == pf_get_double (pf, "S-GaussSigma", &s->gauss_sigma))
          return -1;
        s->gauss_sigma *= (1.0 / 3600.0 * PI / 180.0);
     }
   else if (s->type == MXG_SOURCE_BETA)
     {
        /* s-beta.c:132-162, 99 */
        double core, beta, alpha;
        if ((-1 == pf_get_double (pf, "S-BetaCoreRadius", &core))
            || (-1 == pf_get_double (pf, "S-BetaBeta", &beta)))
          return -1;
        alpha = 3.0 * beta - 0.5;
        if ((alpha <= 1.0) || (core <= 0.0))
          {
             mxg_error ("Beta parameters out of range.  Must be greater than 1/2.");
             return -1;
          }
        s->beta_core = core * (1.0 / 3600.0) * (PI / 180.0);
        s->beta_xpon = 1.0 / (1.0 - alpha);
     }
   else if (s->type == MXG_SOURCE_LINE)
     {
        /* s-line.c:104-122, 75-80: the line along +-theta in the plane of
         * angle phi, rotated from -x onto the source direction */
        double theta, phi;
        JDMVector_Type a = mxg_v (-1.0, 0.0, 0.0), b = mxg_v (s->p[0], s->p[1], s->p[2]), c;
        if ((-1 == pf_get_double (pf, "S-LinePhi", &phi))
            || (-1 == pf_get_double (pf, "S-LineTheta", &theta)))
          return -1;
        s->line_theta = theta * (1.0 / 3600.0 * PI / 180.0);
        phi = phi * (PI / 180.0);
        s->line_cos_phi = cos (phi);
        s->line_sin_phi = sin (phi);
        s->line_alpha = rotation_axis (a, b, &c);
        s->line_normal[0] = c.x; s->line_normal[1] = c.y; s->line_normal[2] = c.z;
     }
   else if (s->type == MXG_SOURCE_USER)
     {
        if (-1 == user_source_init (pf))
          return -1;
     }
   else if (s->type == MXG_SOURCE_IMAGE)
     {
        /* s-image.c:370-385, 262-282: the image centre is -x, rotated onto
         * the source direction about cross(-x, p) (left unnormalised, and
         * unused, when the angle is 0) */
        char file[1024];
        JDMVector_Type m = mxg_v (-1.0, 0.0, 0.0), p0 = mxg_v (s->p[0], s->p[1], s->p[2]), nrm;
        double theta;
        if (-1 == pf_get_file (pf, "S-ImageFile", file, sizeof (file)))
          return -1;
        if (-1 == read_image_source (file, s))
          return -1;
        theta = mxg_v_dot (p0, m);
        if (fabs (theta) > 1.0) theta = (theta < 0) ? -1.0 : 1.0;
        theta = acos (theta);
        nrm = mxg_v_cross (m, p0);
        if (theta != 0.0) mxg_v_normalize (&nrm);
        s->line_alpha = theta;
        s->line_normal[0] = nrm.x; s->line_normal[1] = nrm.y; s->line_normal[2] = nrm.z;
     }
   if (-1 == spectrum_init (pf, s, name))
     return -1;
   /* mean time between photons: 1 / (flux * area), source.c:258-264 */
   s->mean_time = (s->total_flux <= 0.0) ? 0.0
     : 1.0 / s->total_flux / Marx_Mirror_Geometric_Area;
   s->time_mode = MXG_TIME_POISSON;
   Mxg_Cfg_Generation++;
   return 0;
}

/* ---- entry points (source.c:268-417) --------------------------------------- */

Marx_Source_Type *marx_create_source (Param_File_Type *pf)
{
   Marx_Source_Type *st;
   if (-1 == mxg_source_init (pf))
     return NULL;
   st = (Marx_Source_Type *) calloc (1, sizeof (Marx_Source_Type));
   if ((st != NULL) && (Mxg_Cfg.src.type == MXG_SOURCE_RAYFILE))
     {
        st->ray_fp = mxg_open_rayfile (Ray_File, &st->ray_history);
        if (st->ray_fp == NULL)
          {
             free (st);
             return NULL;
          }
     }
   return st;
}

/* s-rayfile.c:176-207: up to num records of the ray file, their arrival
 * times made relative to the last photon of the previous call. */
int mxg_rayfile_read (Marx_Source_Type *st, Marx_Photon_Attr_Type *at, unsigned int num,
                      unsigned int *n, unsigned long *history)
{
   unsigned int i;
   double t = 0.0;
   if ((st == NULL) || (st->ray_fp == NULL)) return -1;
   *n = (unsigned int) fread (at, sizeof (Marx_Photon_Attr_Type), num, st->ray_fp);
   for (i = 0; i < *n; i++)
     {
        t = at[i].arrival_time;
        at[i].arrival_time -= st->ray_last_time;
     }
   st->ray_last_time = t;
   *history = st->ray_history;
   return 0;
}

int marx_open_source (Marx_Source_Type *st)
{
   if (st == NULL) return -1;
   if ((Mxg_Cfg.src.type == MXG_SOURCE_USER) && (-1 == user_source_open ()))
     return -1;
   st->open = 1;
   return 0;
}

int marx_close_source (Marx_Source_Type *st)
{
   if (st == NULL) return -1;
   if (st->ray_fp != NULL) fclose (st->ray_fp);
   if (Mxg_Cfg.src.type == MXG_SOURCE_USER)
     user_source_close ();
   free (st);
   return 0;
}

/* Arrival-time model of the source: Poisson (the reference's) or, with
 * tmax > 0, uniform over [0, tmax). */
void mxg_set_uniform_times (double tmax)
{
   Mxg_Cfg.src.time_mode = (tmax > 0.0) ? MXG_TIME_UNIFORM : MXG_TIME_POISSON;
   Mxg_Cfg.src.uniform_tmax = tmax;
   Mxg_Cfg_Generation++;
}
