/* marx_gpu_cfg.h -- layout of the initialised configuration of libmarx_gpu.
 *
 * The init entry points (marx_mirror_init, marx_grating_init,
 * marx_detector_init, marx_create_source) turn the parameter file and the
 * calibration tables into one plain-old-data mxg_config plus two table
 * arenas (float and double).  Tables are referenced by element offsets into
 * the arenas, never by pointers, so the same bytes are valid on the host and,
 * after one copy, in HBM.  The GPU kernels and any CPU re-implementation read
 * exactly this structure.
 *
 * Units follow the reference: mm, keV, radians, seconds.
 */
#ifndef MARX_GPU_CFG_H
#define MARX_GPU_CFG_H

#define MXG_NUM_SHELLS     4
#define MXG_MAX_CHIPS      10
#define MXG_MAX_DRAKE      8

/* Counter-stream stage ids of the per-photon random streams. */
#define MXG_STAGE_SOURCE   1u
#define MXG_STAGE_DITHER   2u
#define MXG_STAGE_MIRROR   3u
#define MXG_STAGE_GRATING  4u
#define MXG_STAGE_DETECT   5u
#define MXG_STAGE_EVENT    6u       /* event coordinates: grade choice, pixel randomisation */

/* One scattering array of a WFold table (wfold.c:42-57), stored in the
 * double arena as 8 consecutive values:
 *   e_alpha, p_min, delta_p, p_max, pow_law_norm, pow_law_expon,
 *   num_theta_values, theta float-arena offset                           */
#define MXG_FOLD_STRIDE 8

/* Bucket index of a strictly increasing, positive search table: for
 * positive values the IEEE bit pattern orders like the value (and spaces
 * like its logarithm), so the keys (bits - base) >> shift split the table
 * into nb buckets; float-arena entries off .. off + nb hold, for bucket j,
 * the number of table entries whose key is below j (entry nb: all of
 * them).  A search then looks at one bucket's entries.  ok = 1: the table
 * is strictly increasing and positive; ok = 2: it repeats some entries (the
 * HRC MCP and UV/ion shield tables do), and a value equal to a repeated
 * entry is searched whole, since the bisection's pick among equal entries
 * depends on its path; ok = 0: search it whole.  Float keys
 * for tables searched with float abscissae, double keys (the float values
 * widened) for those searched with double ones. */
typedef struct
{
   unsigned long long base;
   unsigned int off, nb, shift, ok;
}
mxg_bucket;
/* Linear bucket index of a strictly increasing double table xp[0..n-1]
 * whose entries spread evenly over their range (the cumulative spectrum,
 * searched with a uniform variate): the key of x > x0 is
 * min(nb - 1, (unsigned) ((x - x0) * scale)), monotone in x under IEEE
 * rounding and computed alike on host and device; double-arena entries
 * off .. off + nb hold, for bucket j, the number of table entries whose key
 * is below j (entry nb: all of them).  ok = 0: search the table whole. */
typedef struct
{
   double x0, scale;
   unsigned int off, nb, ok, pad;
}
mxg_lbucket;
/* largest linear bucket count (the cumulative spectrum) */
#define MXG_BUCKETS_SPECTRUM 4096
/* largest bucket counts: optical constants, correction factors, e_alpha */
#define MXG_BUCKETS_OPT 256
#define MXG_BUCKETS_CORR 64
#define MXG_BUCKETS_EALPHA 32
#define MXG_BUCKETS_GRATING 256
#define MXG_BUCKETS_QE 256

typedef struct
{
   unsigned int n_arrays;      /* arrays sorted by e_alpha, duplicates removed */
   unsigned int arrays_off;    /* double-arena offset of n_arrays*MXG_FOLD_STRIDE */
   unsigned int e_alphas_off;  /* double-arena offset of n_arrays e_alpha values */
   unsigned int pad;
}
mxg_wfold;

typedef struct
{
   /* [0] = paraboloid, [1] = hyperboloid.  x^2+y^2+z^2 = a x^2 + b x + c */
   double conic_a[2], conic_b[2], conic_c[2];
   double conic_xmin[2], conic_xmax[2];
   double to_osac[2][3];        /* OSAC = MARX + to_osac */
   double fwd[2][9], bwd[2][9]; /* tilt matrices, row major */
   double blur[2];              /* arcsec */
   double scat_factor[2];
   double tilt_deg[2][2];       /* [surface][az, el] misalignment (geometry file), degrees */
   double front_position;
   double area_fraction;        /* cumulative, normalised */
   double min_radius, max_radius;
   unsigned int shutter_bitmap; /* bit q set = quadrant q open */
   unsigned int num_open_shutters;
   unsigned int corr_n, corr_e_off, corr_f_off;   /* float arena */
   unsigned int mirror_number;
   mxg_wfold wfold[2];
   mxg_bucket corr_bucket;      /* of the correction table's energies (float keys) */
   mxg_bucket ea_bucket[2];     /* of each WFold table's e_alpha values (double keys) */
}
mxg_shell_cfg;

typedef struct
{
   double xpos;         /* offset from the cap position */
   double half_width;
}
mxg_strut;

typedef struct
{
   int enabled;
   int ideal, use_blur, use_wfold, use_scale, use_struts;
   double vignetting;
   double cap_position;
   double geometric_area;          /* cm^2 */
   unsigned int opt_n, opt_e_off, opt_beta_off, opt_delta_off;   /* float arena */
   mxg_shell_cfg shell[MXG_NUM_SHELLS];
   mxg_strut pre_struts[2], cap_struts[2], post_struts[2];
   double strut_cos, strut_sin;    /* cos, sin of the 30 degree strut clocking */
   mxg_bucket opt_bucket;          /* of the optical constants' energies (float keys) */
}
mxg_hrma_cfg;

typedef struct
{
   unsigned int num_orders, num_energies;
   unsigned int energies_off;   /* float arena: num_energies */
   unsigned int cum_eff_off;    /* float arena: [num_energies][num_orders] */
   unsigned int orders_off;     /* float arena: num_orders diffraction orders */
   unsigned int cum_sorted;     /* every energy's cumulative row is non-decreasing in order */
   double dispersion_angle;     /* radians */
   double period;               /* um */
   double dp_over_p;
   double theta_blur;           /* radians */
   double vig;
   double disp_cos, disp_sin;   /* of dispersion_angle (host libm), for the frame rotations */
   mxg_bucket e_bucket;         /* of the energies (float keys) */
}
mxg_grating;

typedef struct
{
   unsigned int n;             /* 0 = no sector table */
   unsigned int off;           /* double arena: 6 columns of n values:
                                  min, max, dtheta, dtheta_blur,
                                  dp_over_p, dp_over_p_blur (radians) */
}
mxg_sectors;

typedef struct
{
   int type;                    /* 0 NONE, MARX_GRATING_HETG, MARX_GRATING_LETG */
   double rowland[MXG_NUM_SHELLS];   /* torus diameter per shell, mm */
   mxg_grating g[MXG_NUM_SHELLS];
   mxg_sectors sectors[MXG_NUM_SHELLS];
   int has_fine, has_coarse;    /* LETG support gratings */
   mxg_grating fine, coarse;
}
mxg_grating_cfg;

typedef struct
{
   int id;
   int pad;
   double x_ll[3], xhat[3], yhat[3], normal[3];
   double xlen, ylen;
   double x_pixel_size, y_pixel_size;
   double xpixel_offset, ypixel_offset;
   unsigned int qe_n, qe_e_off, qe_off;          /* float arena */
   unsigned int filter_n, filter_e_off, filter_off;
   mxg_bucket qe_bucket;                          /* of the QE energies */
}
mxg_chip;

typedef struct
{
   unsigned int n, e_off, eff_off;   /* float arena; n = 0 -> unit efficiency */
   unsigned int pad;
   mxg_bucket bucket;                 /* of the energies */
}
mxg_qe_table;

typedef struct
{
   double a[3], e1[3], e2[3], normal[3];
   double len1, len2;
}
mxg_rect;

/* ACIS contamination of one chip (aciscontam.c:93-136, 218-465): the
 * transmission exp(-sum_i mu_i(E) (tau0_i + tau1_i f(x, y))) over up to
 * MXG_MAX_CONTAM_LAYERS layers, mu_i tabulated in energy (float arena),
 * tau interpolated in time at set-up.  f is a table of (1024 / blocking)^2
 * floats per layer (fxy_model 0, float arena at fxy_off[i]), the ACIS-I
 * radial model about (x0, y0) (1) or the ACIS-S row profile (2).
 * num_layers = 0: no contamination (transmission 1). */
#define MXG_MAX_CONTAM_LAYERS 5
#define MXG_CONTAM_TABLE    0
#define MXG_CONTAM_RADIAL   1
#define MXG_CONTAM_ROWS     2
typedef struct
{
   unsigned int num_layers, fxy_model, blocking, pad;
   double x0, y0;
   double tau0[MXG_MAX_CONTAM_LAYERS], tau1[MXG_MAX_CONTAM_LAYERS];
   unsigned int mu_n[MXG_MAX_CONTAM_LAYERS], mu_e_off[MXG_MAX_CONTAM_LAYERS];
   unsigned int mu_off[MXG_MAX_CONTAM_LAYERS], fxy_off[MXG_MAX_CONTAM_LAYERS];
}
mxg_contam;

typedef struct
{
   int type;                     /* 0 NONE or MARX_DETECTOR_* */
   int ideal, extend;
   int num_chips;
   double matrix[9];             /* detector transform (detector.c:330-337) */
   double offset[3];             /* DetOffsetX/Y/Z */
   double focal_length;          /* Marx_Focal_Length, detpix.c:41 */
   mxg_chip chip[MXG_MAX_CHIPS];
   /* ACIS frame-transfer streak (acis-i.c:56-92) */
   double frame_time, exposure_time, frame_transfer_time;
   /* HRC */
   mxg_qe_table mcp_qe[3];       /* HRC-S MCP QE, index = mcp id mapping */
   mxg_qe_table filter_qe[4];    /* HRC-S UV/ion shield regions 0..3 */
   int mcp_map[4];               /* mcp id -> mcp_qe index (hrc-s.c:49-52) */
   double shield_t, shield_l, shield_r, shield_x, shield_sl, shield_sr;
   double shield_sl_gap, shield_sr_gap, shield_y_center, shield_z_center;
   /* MCP position blur (hrcblur.c:53-192), lengths in mm */
   double g1_sigma, g1_xctr, g1_yctr, g1_wgt;
   double g2_sigma, g2_xctr, g2_yctr, g2_wgt;
   double l1_hwhm, l1_xctr, l1_yctr, l1_rmax, l1_wgt;
   /* pixel mapping (hrc_s_geom.c:344-394, hrc_i_geom.c:146-156) */
   double u_pixel_size, v_pixel_size;
   double pix_u[4], pix_v[4], pix_u0[4], pix_v0[4], pix_x0[4], pix_y0[4];
   /* HESF / Drake flat (drake.c:170-260) */
   int use_drake, drake_n;       /* drake_n = 2 * number of plates */
   double drake_cr_width;
   mxg_rect drake[MXG_MAX_DRAKE];
   unsigned int drake_c_n, drake_c_e_off, drake_c_beta_off, drake_c_delta_off;
   unsigned int drake_cr_n, drake_cr_e_off, drake_cr_beta_off, drake_cr_delta_off;
   /* ACIS FEF (acis_fef.c): fef_on when a FEF file was read.  Double arena
    * at fef_map_off: by chip id (10) and 32 x 32 regions of 32 pixels
    * ([id][x/32][y/32]), the FEF index or -1; at fef_desc_off, 5 per FEF:
    * energies, Gaussians, and the float-arena offsets of its energies, its
    * channels and its Gaussians ([energy][gaussian][amp, centre, sigma]). */
   int fef_on, fef_max_gaussians;
   unsigned int fef_map_off, fef_desc_off;
   /* ACIS contamination by chip slot (chip[k] <-> contam[k]); contam_on when
    * any chip has layers */
   int contam_on, contam_pad;
   mxg_contam contam[MXG_MAX_CHIPS];
}
mxg_detector_cfg;

#define MXG_SOURCE_POINT   0
#define MXG_SOURCE_DISK    1
#define MXG_SOURCE_RAYFILE 2       /* replayed on the host by marx_create_photons */
#define MXG_SOURCE_GAUSS   3
#define MXG_SOURCE_BETA    4
#define MXG_SOURCE_LINE    5
#define MXG_SOURCE_IMAGE   6
#define MXG_SOURCE_USER    7       /* rays from a dlopen()ed plug-in, on the host */
#define MXG_SPECTRUM_FLAT  1
#define MXG_SPECTRUM_FILE  2
#define MXG_TIME_POISSON   0
#define MXG_TIME_UNIFORM   1
#define MXG_DITHER_NONE    0
#define MXG_DITHER_INTERNAL 1

typedef struct
{
   int type;
   int spectrum;
   double p[3], p_normal[3];
   double distance;                 /* mm, <= 0 means infinity */
   double disk_theta, disk_theta_min;    /* radians */
   double emin, emax, total_flux;
   unsigned int file_n, file_e_off, file_cum_off, pad;  /* double arena */
   double mean_time;                /* s between photons */
   int time_mode;
   int dither_mode;
   double uniform_tmax;             /* MXG_TIME_UNIFORM: times ~ U[0,tmax] */
   double ra_amp, dec_amp, roll_amp;        /* radians */
   double ra_period, dec_period, roll_period;
   double ra_phase, dec_phase, roll_phase;
   double nominal_roll;
   double aspect_blur;              /* radians, sigma */
   double gauss_sigma;              /* GAUSS: radians (s-gauss.c) */
   double beta_core, beta_xpon;     /* BETA: core radius (radians), 1 / (1 - alpha) (s-beta.c) */
   double line_theta;               /* LINE: half length (radians) (s-line.c) */
   double line_cos_phi, line_sin_phi;    /* of its orientation angle */
   double line_alpha, line_normal[3];    /* LINE, IMAGE: the rotation of -x onto the source direction */
   unsigned int image_n, image_nx, image_ny, image_off;   /* IMAGE: cumulative pixel distribution */
   double image_xscale, image_yscale;    /* radians per pixel (float arena at image_off) */
   mxg_lbucket file_bucket;              /* FILE: of the cumulative distribution */
}
mxg_source_cfg;

typedef struct
{
   mxg_hrma_cfg hrma;
   mxg_grating_cfg grating;
   mxg_detector_cfg det;
   mxg_source_cfg src;
}
mxg_config;

/* Event coordinates (mxg_event_coords): the settings with the detector's
 * tiling and focal-plane system, and the CALDB subpixel (EDSER) tables of
 * ACIS: for table t (0 front-, 1 back-illuminated chips) and flight grade
 * g, sub_n[t][g] points (0: no offset) at sub_off[t][g] in a float array
 * holding the energies, then the chip x offsets, then the chip y offsets. */
#define MXG_TILE_SHIFT 0                 /* tiled = chip + offset (ACIS-S, HRC) */
#define MXG_TILE_ACIS_I 1                /* chips 0, 2 and 1, 3 turned (acis_geom.c:112-141) */
typedef struct
{
   double focal_length, det_offset[3];
   double cos_roll, sin_roll;
   double fp_delta_s0, fp_x0, fp_y0;    /* the detector's focal-plane system (detpix.c:41-61) */
   double tdet_xoff[MXG_MAX_CHIPS], tdet_yoff[MXG_MAX_CHIPS];   /* by chip id */
   double pixel_xoff[MXG_MAX_CHIPS], pixel_yoff[MXG_MAX_CHIPS]; /* chip pixel of the LL corner, by index */
   int chip_index[MXG_MAX_CHIPS];       /* chip id -> mxg_detector_cfg.chip[], -1 */
   int tiling;                          /* MXG_TILE_* */
   int pix_adj, dither_used;
   int has_grade;                       /* ACIS: flight grades (and EDSER) */
   int held;                            /* dither state held over a frame (ACIS) */
   unsigned int sub_off[2][256], sub_n[2][256];
}
mxg_event_params;

#ifdef __cplusplus
extern "C" {
#endif

/* Differentiable mirror: the shell geometry at given misalignment
 * parameters, with the derivatives of each piece with respect to them.
 * Parameters per shell pair, in this order: tilt az, tilt el (arcsec, added
 * to both surfaces' yaw / pitch misalignment), decenter y, decenter z,
 * despace x (mm: the pair moves along MARX +y, +z, +x). */
#define MXG_DIFF_NPAR 5
typedef struct
{
   double fwd[2][9], bwd[2][9];              /* tilt matrices at the parameters */
   double dfwd[2][2][9], dbwd[2][2][9];      /* [surface][d/d az, d/d el per arcsec] */
   double to_osac[2][3];                     /* d/d(decenter y, z, despace x) = -1 */
   double front_position;                    /* d/d despace x = +1 */
}
mxg_diff_shell;

typedef struct
{
   double params[MXG_NUM_SHELLS * MXG_DIFF_NPAR];
   mxg_diff_shell shell[MXG_NUM_SHELLS];
}
mxg_diff_cfg;

/* The library's current configuration and arenas (host copies). */
int mxg_get_config (const mxg_config **cfg,
                    const float **farena, unsigned int *nf,
                    const double **darena, unsigned int *nd);

#ifdef __cplusplus
}
#endif

#endif /* MARX_GPU_CFG_H */
