/* mxg_host.h -- internal interfaces of the host (C) side of libmarx_gpu. */
#ifndef MXG_HOST_H
#define MXG_HOST_H

#include <stddef.h>
#include "marx_gpu.h"
#include "marx_gpu_cfg.h"

#ifndef PI
# define PI 3.14159265358979323846
#endif
/* keV um (2006 NIST value), marx/libsrc/marx.h:471-477 */
#define MXG_HBAR_C 1.973269631e-4

/* ---- errors ---------------------------------------------------------- */
void mxg_error (const char *fmt, ...);
void mxg_message (int level, const char *fmt, ...);
extern int Mxg_Verbose;

/* ---- configuration and arenas ---------------------------------------- */
extern mxg_config Mxg_Cfg;
/* Bumped whenever Mxg_Cfg or an arena changes; the device copy follows. */
extern unsigned int Mxg_Cfg_Generation;

/* Append n values and return their element offset, or (unsigned)-1. */
unsigned int mxg_farena_push (const float *v, unsigned int n);
unsigned int mxg_darena_push (const double *v, unsigned int n);
unsigned int mxg_farena_reserve (unsigned int n);   /* uninitialised */
unsigned int mxg_darena_reserve (unsigned int n);
float *mxg_farena_ptr (unsigned int off);
double *mxg_darena_ptr (unsigned int off);
unsigned int mxg_farena_size (void);
/* The bucket index (mxg_bucket) of a search table of n floats, keys the
 * float bits (double_keys = 0) or those of the values widened to double, at
 * most nbmax buckets; its counts are pushed to the float arena.  A table
 * that is not strictly increasing and positive gets ok = 0.  0 or -1. */
int mxg_build_bucket (const float *xp, unsigned int n, int double_keys, unsigned int nbmax, mxg_bucket *b);
/* the linear bucket index of a double table (counts in the double arena);
 * b->ok = 0 when the table is not strictly increasing */
int mxg_build_lbucket (const double *xp, unsigned int n, unsigned int nbmax, mxg_lbucket *b);
/* the same for float-arena table (off, n), float keys */
int mxg_bucket_of_farena (unsigned int off, unsigned int n, unsigned int nbmax, mxg_bucket *b);
unsigned int mxg_darena_size (void);
const float *mxg_farena_base (void);
const double *mxg_darena_base (void);

/* ---- data files -------------------------------------------------------- */
/* Malloced absolute path of a calibration file (data.c:57-81). */
char *mxg_data_file_name (const char *file);
/* CALDB object (GEOM, AIMPTS, ACISQE, HETGEFF, LETGEFF, ACISSUBPIX) -> malloced path
 * (caldb.c:37-133). */
char *mxg_caldb_file (const char *object);

/* Big-endian "bdat" float tables (data/FORMAT.TXT; data.c:134-254).
 * Reads the first ncols columns into malloced float arrays cols[0..ncols). */
int mxg_read_bdat (const char *file, unsigned int ncols,
                   float **cols, unsigned int *nrows);

/* WFold scattering tables (wfold.c:112-299): sorted, de-duplicated arrays
 * appended to the arenas. */
int mxg_read_wfold (const char *file, mxg_wfold *w);

/* Tab-separated RDB table (rdb.c) */
typedef struct mxg_rdb mxg_rdb;
mxg_rdb *mxg_rdb_open (const char *file);
void mxg_rdb_close (mxg_rdb *r);
int mxg_rdb_get_double (mxg_rdb *r, const char *key_col, const char *key,
                        const char *col, double *v);

/* "keyword v1 v2 ..." text file (data.c:286-395) */
typedef struct
{
   const char *name;
   unsigned int num;
   double *data;
   double scale;
   int processed;
}
mxg_simple_item;
int mxg_read_simple_data_file (const char *file, mxg_simple_item *table);

/* Whitespace-separated numeric columns, '#' comments (JDMread_column_ddata):
 * reads the first ncols columns into malloced double arrays. */
int mxg_read_columns (const char *file, unsigned int ncols,
                      double **cols, unsigned int *nrows);

/* Minimal FITS binary-table access (enough for the CALDB subset). */
typedef struct mxg_fits_table mxg_fits_table;
typedef int (*mxg_fits_match_fn) (mxg_fits_table *t, void *client);
/* Opens the first BINTABLE HDU for which match returns 1. */
mxg_fits_table *mxg_fits_find_table (const char *file, mxg_fits_match_fn match,
                                     void *client);
void mxg_fits_close (mxg_fits_table *t);
/* Keyword value as a string (quotes and trailing blanks removed). */
int mxg_fits_keyword (mxg_fits_table *t, const char *key, char *buf, unsigned int buflen);
int mxg_fits_keyword_is (mxg_fits_table *t, const char *key, const char *value);
unsigned int mxg_fits_num_rows (mxg_fits_table *t);
/* Column repeat count and type code ('E','D','J','A',...); -1 if absent. */
int mxg_fits_column (mxg_fits_table *t, const char *name, unsigned int *repeat, char *type);
/* Element j of column 'name' in row i, converted to double. */
int mxg_fits_get_double (mxg_fits_table *t, const char *name, unsigned int row,
                         unsigned int j, double *v);
int mxg_fits_get_float (mxg_fits_table *t, const char *name, unsigned int row,
                        unsigned int j, float *v);
int mxg_fits_get_string (mxg_fits_table *t, const char *name, unsigned int row,
                         char *buf, unsigned int buflen);

float mxg_interpolate_f (float x, const float *xp, const float *yp, unsigned int n);

/* ---- vector helpers restating jdmath's arithmetic (vector.c, rotate.c) --- */
JDMVector_Type mxg_v (double x, double y, double z);
double mxg_v_length (JDMVector_Type a);
void mxg_v_normalize (JDMVector_Type *a);
JDMVector_Type mxg_v_cross (JDMVector_Type a, JDMVector_Type b);
double mxg_v_dot (JDMVector_Type a, JDMVector_Type b);
JDMVector_Type mxg_v_sum (JDMVector_Type a, JDMVector_Type b);
JDMVector_Type mxg_v_diff (JDMVector_Type a, JDMVector_Type b);
JDMVector_Type mxg_v_ax1_bx2 (double a, JDMVector_Type x1, double b, JDMVector_Type x2);
JDMVector_Type mxg_v_rotate_vector1 (JDMVector_Type p, JDMVector_Type n, double c, double s);
JDMVector_Type mxg_v_rotate_unit_vector (JDMVector_Type p, JDMVector_Type n, double theta);
JDMVector_Type mxg_v_spherical_to_vector (double r, double theta, double phi);
void mxg_v_unit_vector_to_spherical (JDMVector_Type p, double *theta, double *phi);
void mxg_m3_rot_z (double m[9], double theta);
void mxg_m3_rot (double m[9], JDMVector_Type n, double theta);
void mxg_m3_mul (double c[9], const double a[9], const double b[9]);
JDMVector_Type mxg_m3_vmul (const double a[9], JDMVector_Type v);

/* ---- module init ------------------------------------------------------- */
int mxg_hrma_init (Param_File_Type *pf);
int mxg_grating_init (Param_File_Type *pf);
int mxg_detector_init (Param_File_Type *pf);
int mxg_source_init (Param_File_Type *pf);
/* Pointing used by the dither model and source setup (dither.c:125-155). */
int mxg_dither_init (Param_File_Type *pf, double *yoff, double *zoff);
void mxg_get_nominal_pointing (double *ra, double *dec, double *roll);
void mxg_get_pointing (double *ra, double *dec, double *roll);

/* History bits of the device-resident photons of the last mxg_trace_device
 * or mxg_trace_batch call; -1 (with the error set) when the store no longer
 * holds them. */
int mxg_resident_history (unsigned long *history);
/* Opens a ray file and reads its header (s-rayfile.c:43-65); NULL with
 * the error set when it is not one. */
FILE *mxg_open_rayfile (const char *file, unsigned long *history);
/* Up to num records of a RAYFILE source, times relative to the previous
 * call's last photon (s-rayfile.c:176-207). */
/* USER sources (s-user.c): the plug-in's user_start_iteration (opening the
 * source first if needed; 0 when it has none) and user_create_ray. */
int mxg_user_start_iteration (void);
int mxg_user_create_ray (double *dt, double *energy, double *cx, double *cy, double *cz);
int mxg_rayfile_read (Marx_Source_Type *st, Marx_Photon_Attr_Type *at, unsigned int num,
                      unsigned int *n, unsigned long *history);
extern double _Marx_TStart_MJDsecs;

/* ---- photon store on the device (implemented in the HIP layer) ---------- */
#include "mxg_device.h"

#endif
