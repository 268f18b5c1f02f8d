/* tables.c -- readers for the calibration-table formats on the path.
 *
 *  - bdat: 32-byte big-endian header + comment + row-major float32 data
 *    (marx/data/FORMAT.TXT; reader semantics marx/libsrc/data.c:134-254).
 *  - WFold scattering tables (marx/libsrc/wfold.c:112-299): sequence of
 *    arrays {int32 n; float32 e_alpha, p_min, delta_p, p_max, norm, expon;
 *    float32 theta[n]}, terminated by end of file or n == 0.
 *  - RDB tab-separated tables (marx/libsrc/rdb.c).
 *  - "keyword values..." text files (marx/libsrc/data.c:286-395).
 *  - whitespace-separated numeric columns (JDMread_column_ddata).
 *  - FITS binary tables (the CALDB subset read by caldb.c and
 *    diffract.c:1343-1418), big-endian, header cards of 80 bytes.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <ctype.h>
#include <dlfcn.h>
#include <libgen.h>

#include "mxg_host.h"

/* ---- data directory ----------------------------------------------------- */

static char *Data_Directory = NULL;

static char *xstrdup (const char *s)
{
   size_t n = strlen (s) + 1;
   char *d = (char *) malloc (n);
   if (d != NULL) memcpy (d, s, n);
   return d;
}

int marx_set_data_directory (char *dir)
{
   if (dir == NULL) return -1;
   if (dir[0] == '$')
     {
        char *env = getenv (dir + 1);
        if (env == NULL)
          {
             mxg_error ("Unable to set data directory: environment variable %s does not exist", dir + 1);
             return -1;
          }
        dir = env;
     }
   if (*dir == 0)
     {
        mxg_error ("Data path is empty");
        return -1;
     }
   free (Data_Directory);
   Data_Directory = xstrdup (dir);
   return (Data_Directory == NULL) ? -1 : 0;
}

/* The bundled tables live in ../data relative to this library's file. */
static const char *default_data_directory (void)
{
   static char buf[4096];
   Dl_info info;
   char tmp[4096];

   if (buf[0]) return buf;
   if ((0 == dladdr ((void *) &default_data_directory, &info)) || (info.dli_fname == NULL))
     return ".";
   strncpy (tmp, info.dli_fname, sizeof (tmp) - 1);
   tmp[sizeof (tmp) - 1] = 0;
   snprintf (buf, sizeof (buf), "%s/../data", dirname (tmp));
   return buf;
}

char *mxg_data_file_name (const char *file)
{
   const char *dir;
   char *path;
   size_t n;

   if ((file == NULL) || (*file == 0))
     {
        mxg_error ("empty data file name");
        return NULL;
     }
   if (file[0] == '/')
     return xstrdup (file);
   dir = Data_Directory ? Data_Directory : default_data_directory ();
   n = strlen (dir) + strlen (file) + 2;
   if (NULL == (path = (char *) malloc (n)))
     return NULL;
   snprintf (path, n, "%s/%s", dir, file);
   return path;
}

/* CALDB object names and the files that hold them (marx/data/marxcaldb.par). */
char *mxg_caldb_file (const char *object)
{
   static const struct { const char *obj, *file; } Caldb[] =
     {
          {"GEOM", "caldb/telD1999-07-23geomN0006.fits"},
          {"AIMPTS", "caldb/telD1999-07-23aimptsN0002.fits"},
          {"ACISQE", "caldb/acisD1997-04-17qeN0006.fits"},
          {"HETGEFF", "caldb/hetgD1996-11-01greffpr001N0007.fits"},
          {"LETGEFF", "caldb/letgD1996-11-01greffpr001N0007.fits"},
          {"ACISSUBPIX", "caldb/acisD1999-07-22subpixN0001_marx.fits"},
          {"ACISFEF", "caldb/acisfef.fits"},
          {"ACISCONTAM", "caldb/acisD1999-08-13contamN0015_marx.fits"},
          {NULL, NULL}
     };
   unsigned int i;
   for (i = 0; Caldb[i].obj != NULL; i++)
     if (0 == strcmp (Caldb[i].obj, object))
       return mxg_data_file_name (Caldb[i].file);
   mxg_error ("Unable to find %s caldb file", object);
   return NULL;
}

/* ---- big-endian helpers -------------------------------------------------- */

static unsigned int be32 (const unsigned char *b)
{
   return ((unsigned int) b[0] << 24) | ((unsigned int) b[1] << 16)
     | ((unsigned int) b[2] << 8) | (unsigned int) b[3];
}

static float be_float (const unsigned char *b)
{
   unsigned int u = be32 (b);
   float f;
   memcpy (&f, &u, 4);
   return f;
}

static double be_double (const unsigned char *b)
{
   unsigned long long u = ((unsigned long long) be32 (b) << 32) | be32 (b + 4);
   double d;
   memcpy (&d, &u, 8);
   return d;
}

static unsigned char *read_whole_file (const char *file, size_t *len)
{
   FILE *fp = fopen (file, "rb");
   unsigned char *buf;
   long n;

   if (fp == NULL)
     {
        mxg_error ("Unable to open %s", file);
        return NULL;
     }
   fseek (fp, 0, SEEK_END);
   n = ftell (fp);
   fseek (fp, 0, SEEK_SET);
   if ((n < 0) || (NULL == (buf = (unsigned char *) malloc ((size_t) n + 1))))
     {
        fclose (fp);
        return NULL;
     }
   if ((size_t) n != fread (buf, 1, (size_t) n, fp))
     {
        mxg_error ("Read error on %s", file);
        free (buf);
        fclose (fp);
        return NULL;
     }
   buf[n] = 0;
   fclose (fp);
   *len = (size_t) n;
   return buf;
}

/* ---- bdat ---------------------------------------------------------------- */

int mxg_read_bdat (const char *file, unsigned int ncols, float **cols, unsigned int *nrows)
{
   static const unsigned char magic[4] = {0x83, 0x43, 0x59, 0x8E};
   unsigned char *buf;
   size_t len;
   unsigned int nr, nc, clen, i, j;
   const unsigned char *data;

   if (NULL == (buf = read_whole_file (file, &len)))
     return -1;
   if ((len < 32) || memcmp (buf, magic, 4))
     {
        mxg_error ("%s: bad bdat magic number", file);
        free (buf);
        return -1;
     }
   if (buf[4] != 'E')
     {
        mxg_error ("%s has unsupported data type '%c'", file, buf[4]);
        free (buf);
        return -1;
     }
   nr = be32 (buf + 5);
   nc = be32 (buf + 9);
   clen = be32 (buf + 13);
   if ((ncols > nc) || (32 + (size_t) clen + 4 * (size_t) nr * nc > len))
     {
        mxg_error ("%s does not have enough columns or rows", file);
        free (buf);
        return -1;
     }
   data = buf + 32 + clen;
   for (j = 0; j < ncols; j++)
     {
        cols[j] = (float *) malloc ((nr ? nr : 1) * sizeof (float));
        if (cols[j] == NULL)
          {
             while (j--) free (cols[j]);
             free (buf);
             return -1;
          }
     }
   for (i = 0; i < nr; i++)
     for (j = 0; j < ncols; j++)
       cols[j][i] = be_float (data + 4 * ((size_t) i * nc + j));
   *nrows = nr;
   free (buf);
   return 0;
}

/* ---- WFold ----------------------------------------------------------------- */

typedef struct
{
   double v[6];            /* e_alpha p_min delta_p p_max norm expon */
   unsigned int n;
   unsigned int order;     /* position in the file */
   const unsigned char *theta;
}
Fold_Array;

static int fold_cmp (const void *a, const void *b)
{
   const Fold_Array *fa = (const Fold_Array *) a, *fb = (const Fold_Array *) b;
   if (fa->v[0] > fb->v[0]) return 1;
   if (fa->v[0] < fb->v[0]) return -1;
   /* keep file order among equal e_alpha so the first one survives */
   return (fa->order < fb->order) ? -1 : (fa->order > fb->order);
}

int mxg_read_wfold (const char *file, mxg_wfold *w)
{
   unsigned char *buf;
   size_t len, pos = 0;
   Fold_Array *arrays = NULL;
   unsigned int num = 0, cap = 0, i, nkept;
   double *dst;
   unsigned int doff, eoff;

   if (NULL == (buf = read_whole_file (file, &len)))
     return -1;

   while (pos + 4 <= len)
     {
        Fold_Array *a;
        unsigned int n = be32 (buf + pos), k;
        if (n == 0) break;
        if (pos + 4 + 24 + 4 * (size_t) n > len)
          {
             mxg_error ("%s: truncated wfold array", file);
             goto fail;
          }
        if (num == cap)
          {
             Fold_Array *na;
             cap = cap ? 2 * cap : 256;
             na = (Fold_Array *) realloc (arrays, cap * sizeof (Fold_Array));
             if (na == NULL) goto fail;
             arrays = na;
          }
        a = arrays + num;
        for (k = 0; k < 6; k++)
          a->v[k] = (double) be_float (buf + pos + 4 + 4 * k);
        a->n = n;
        a->order = num;
        a->theta = buf + pos + 28;
        /* delta_p re-derived from the end points (wfold.c:164-173) */
        if (n == 1)
          {
             a->v[3] = a->v[1];
             a->v[2] = 0.1;
          }
        else a->v[2] = (a->v[3] - a->v[1]) / (double) (n - 1);
        if ((a->v[2] <= 0.0) || (a->v[2] > 1.0) || (a->v[0] < 0.0)
            || (a->v[3] > 1.0) || (a->v[3] < 0.0) || (a->v[1] > 1.0) || (a->v[1] < 0.0))
          {
             mxg_error ("%s: table has inconsistent values", file);
             goto fail;
          }
        num++;
        pos += 28 + 4 * (size_t) n;
     }
   if (num == 0)
     {
        mxg_error ("%s: no wfold arrays", file);
        goto fail;
     }

   /* sort by e_alpha and drop later duplicates (wfold.c:381-427) */
   qsort (arrays, num, sizeof (Fold_Array), fold_cmp);
   nkept = 0;
   for (i = 0; i < num; i++)
     {
        if ((nkept > 0) && (arrays[i].v[0] == arrays[nkept - 1].v[0]))
          continue;
        arrays[nkept++] = arrays[i];
     }

   doff = mxg_darena_reserve (nkept * MXG_FOLD_STRIDE);
   eoff = mxg_darena_reserve (nkept);
   if ((doff == (unsigned int) -1) || (eoff == (unsigned int) -1))
     goto fail;
   for (i = 0; i < nkept; i++)
     {
        unsigned int k, toff = mxg_farena_reserve (arrays[i].n);
        float *t;
        if (toff == (unsigned int) -1) goto fail;
        t = mxg_farena_ptr (toff);
        for (k = 0; k < arrays[i].n; k++)
          t[k] = be_float (arrays[i].theta + 4 * k);
        dst = mxg_darena_ptr (doff + i * MXG_FOLD_STRIDE);
        for (k = 0; k < 6; k++) dst[k] = arrays[i].v[k];
        dst[6] = (double) arrays[i].n;
        dst[7] = (double) toff;
        mxg_darena_ptr (eoff)[i] = arrays[i].v[0];
     }
   w->n_arrays = nkept;
   w->arrays_off = doff;
   w->e_alphas_off = eoff;
   w->pad = 0;
   free (arrays);
   free (buf);
   return 0;

fail:
   free (arrays);
   free (buf);
   return -1;
}

/* ---- RDB ----------------------------------------------------------------- */

struct mxg_rdb
{
   unsigned int ncols, nrows;
   char **colnames;
   char ***cells;        /* [row][col] */
};

static unsigned int split_tabs (char *line, char **out, unsigned int nmax)
{
   unsigned int n = 0;
   char *p = line;
   while (n < nmax)
     {
        char *start = p;
        while ((*p != 0) && (*p != '\t') && (*p != '\n') && (*p != '\r')) p++;
        out[n++] = start;
        if (*p != '\t')
          {
             *p = 0;
             break;
          }
        *p++ = 0;
     }
   return n;
}

void mxg_rdb_close (mxg_rdb *r)
{
   unsigned int i, j;
   if (r == NULL) return;
   for (i = 0; i < r->nrows; i++)
     {
        for (j = 0; j < r->ncols; j++) free (r->cells[i][j]);
        free (r->cells[i]);
     }
   for (j = 0; j < r->ncols; j++) free (r->colnames[j]);
   free (r->cells);
   free (r->colnames);
   free (r);
}

mxg_rdb *mxg_rdb_open (const char *file)
{
   FILE *fp;
   char line[8192];
   char *f[256];
   mxg_rdb *r;
   int stage = 0;
   unsigned int cap = 0;

   if (NULL == (fp = fopen (file, "r")))
     {
        mxg_error ("Unable to open %s", file);
        return NULL;
     }
   r = (mxg_rdb *) calloc (1, sizeof (mxg_rdb));
   if (r == NULL) { fclose (fp); return NULL; }

   while (NULL != fgets (line, sizeof (line), fp))
     {
        unsigned int n, j;
        if (line[0] == '#') continue;
        n = split_tabs (line, f, 256);
        if (stage == 0)
          {
             r->ncols = n;
             r->colnames = (char **) calloc (n, sizeof (char *));
             if (r->colnames == NULL) goto fail;
             for (j = 0; j < n; j++) r->colnames[j] = xstrdup (f[j]);
             stage = 1;
             continue;
          }
        if (stage == 1)
          {
             stage = 2;        /* column definition line */
             continue;
          }
        if ((n == 1) && (f[0][0] == 0)) continue;
        if (r->nrows == cap)
          {
             char ***nc;
             cap = cap ? 2 * cap : 16;
             nc = (char ***) realloc (r->cells, cap * sizeof (char **));
             if (nc == NULL) goto fail;
             r->cells = nc;
          }
        r->cells[r->nrows] = (char **) calloc (r->ncols, sizeof (char *));
        if (r->cells[r->nrows] == NULL) goto fail;
        for (j = 0; j < r->ncols; j++)
          r->cells[r->nrows][j] = xstrdup ((j < n) ? f[j] : "");
        r->nrows++;
     }
   fclose (fp);
   return r;

fail:
   fclose (fp);
   mxg_rdb_close (r);
   return NULL;
}

int mxg_rdb_get_double (mxg_rdb *r, const char *key_col, const char *key,
                        const char *col, double *v)
{
   unsigned int i, kc = (unsigned int) -1, vc = (unsigned int) -1;
   for (i = 0; i < r->ncols; i++)
     {
        if (0 == strcmp (r->colnames[i], key_col)) kc = i;
        if (0 == strcmp (r->colnames[i], col)) vc = i;
     }
   if ((kc == (unsigned int) -1) || (vc == (unsigned int) -1))
     {
        mxg_error ("rdb column %s or %s not found", key_col, col);
        return -1;
     }
   for (i = 0; i < r->nrows; i++)
     {
        if (0 != strcmp (r->cells[i][kc], key)) continue;
        if (1 != sscanf (r->cells[i][vc], "%lf", v))
          {
             mxg_error ("Error parsing field %s as a double", col);
             return -1;
          }
        return 0;
     }
   mxg_error ("rdb row %s=%s not found", key_col, key);
   return -1;
}

/* ---- simple keyword data files ------------------------------------------- */

int mxg_read_simple_data_file (const char *file, mxg_simple_item *table)
{
   FILE *fp;
   char line[0x3FFF];
   mxg_simple_item *t;
   int ret = 0;

   for (t = table; t->name != NULL; t++) t->processed = 0;
   if (NULL == (fp = fopen (file, "r")))
     {
        mxg_error ("Unable to open %s", file);
        return -1;
     }
   while (NULL != fgets (line, sizeof (line), fp))
     {
        char *p0 = line, *p1;
        while (isspace ((unsigned char) *p0)) p0++;
        if ((*p0 == '#') || (*p0 == 0)) continue;
        p1 = p0;
        while ((*p1 != 0) && !isspace ((unsigned char) *p1)) p1++;
        if (*p1 != 0) *p1++ = 0;
        for (t = table; t->name != NULL; t++)
          {
             unsigned int i;
             if (0 != strcmp (t->name, p0)) continue;
             for (i = 0; i < t->num; i++)
               {
                  double d;
                  char *end;
                  d = strtod (p1, &end);
                  if (end == p1)
                    {
                       mxg_error ("%s: expecting %u numeric fields for %s", file, t->num, t->name);
                       fclose (fp);
                       return -1;
                    }
                  t->data[i] = d * t->scale;
                  p1 = end;
               }
             t->processed = 1;
             break;
          }
     }
   fclose (fp);
   for (t = table; t->name != NULL; t++)
     if (t->processed == 0)
       {
          mxg_error ("%s not in %s", t->name, file);
          ret = -1;
       }
   return ret;
}

/* ---- numeric columns ------------------------------------------------------ */

int mxg_read_columns (const char *file, unsigned int ncols, double **cols, unsigned int *nrows)
{
   FILE *fp;
   char line[8192];
   unsigned int n = 0, cap = 0, j;

   if (NULL == (fp = fopen (file, "r")))
     {
        mxg_error ("Unable to open %s", file);
        return -1;
     }
   for (j = 0; j < ncols; j++) cols[j] = NULL;
   while (NULL != fgets (line, sizeof (line), fp))
     {
        char *p = line, *end;
        double v[64];
        unsigned int k;
        while (isspace ((unsigned char) *p)) p++;
        if ((*p == '#') || (*p == 0) || (*p == '%')) continue;
        for (k = 0; k < ncols; k++)
          {
             v[k] = strtod (p, &end);
             if (end == p) break;
             p = end;
          }
        if (k < ncols) continue;
        if (n == cap)
          {
             cap = cap ? 2 * cap : 1024;
             for (j = 0; j < ncols; j++)
               {
                  double *nc = (double *) realloc (cols[j], cap * sizeof (double));
                  if (nc == NULL) { fclose (fp); return -1; }
                  cols[j] = nc;
               }
          }
        for (j = 0; j < ncols; j++) cols[j][n] = v[j];
        n++;
     }
   fclose (fp);
   if (n == 0)
     {
        mxg_error ("%s contains no data rows", file);
        return -1;
     }
   *nrows = n;
   return 0;
}

/* ---- FITS binary tables ----------------------------------------------------- */

#define FITS_BLOCK 2880
#define FITS_MAX_COLS 64

struct mxg_fits_table
{
   unsigned char *file_buf;
   size_t file_len;
   const unsigned char *header;    /* start of this HDU's header */
   unsigned int header_cards;
   const unsigned char *data;
   unsigned int row_bytes, nrows, ncols;
   char colname[FITS_MAX_COLS][72];
   char coltype[FITS_MAX_COLS];
   unsigned int repeat[FITS_MAX_COLS], offset[FITS_MAX_COLS];
};

static int card_key (const unsigned char *card, const char *key)
{
   size_t n = strlen (key);
   unsigned int i;
   if (n > 8) return 0;
   if (memcmp (card, key, n)) return 0;
   for (i = (unsigned int) n; i < 8; i++)
     if (card[i] != ' ') return 0;
   return 1;
}

/* Value field of a card without quotes, comment and trailing blanks. */
static void card_value (const unsigned char *card, char *buf, unsigned int buflen)
{
   char tmp[81];
   char *p, *e;
   memcpy (tmp, card, 80);
   tmp[80] = 0;
   p = tmp + 10;
   while (*p == ' ') p++;
   if (*p == '\'')
     {
        p++;
        e = strchr (p, '\'');
        if (e != NULL) *e = 0;
     }
   else
     {
        e = strchr (p, '/');
        if (e != NULL) *e = 0;
     }
   e = p + strlen (p);
   while ((e > p) && (e[-1] == ' ')) *--e = 0;
   strncpy (buf, p, buflen - 1);
   buf[buflen - 1] = 0;
}

static int header_value (const unsigned char *hdr, unsigned int ncards, const char *key,
                         char *buf, unsigned int buflen)
{
   unsigned int i;
   for (i = 0; i < ncards; i++)
     if (card_key (hdr + 80 * i, key) && (hdr[80 * i + 8] == '='))
       {
          card_value (hdr + 80 * i, buf, buflen);
          return 0;
       }
   return -1;
}

static long header_long (const unsigned char *hdr, unsigned int ncards, const char *key, long dflt)
{
   char buf[80];
   if (-1 == header_value (hdr, ncards, key, buf, sizeof (buf)))
     return dflt;
   return strtol (buf, NULL, 10);
}

static unsigned int tform_size (char t)
{
   switch (t)
     {
      case 'L': case 'B': case 'A': return 1;
      case 'I': return 2;
      case 'J': case 'E': return 4;
      case 'K': case 'D': case 'C': return 8;
      case 'M': return 16;
     }
   return 0;
}

void mxg_fits_close (mxg_fits_table *t)
{
   if (t == NULL) return;
   free (t->file_buf);
   free (t);
}

mxg_fits_table *mxg_fits_find_table (const char *file, mxg_fits_match_fn match, void *client)
{
   size_t len, pos = 0;
   unsigned char *buf = read_whole_file (file, &len);
   mxg_fits_table *t;

   if (buf == NULL) return NULL;
   t = (mxg_fits_table *) calloc (1, sizeof (mxg_fits_table));
   if (t == NULL) { free (buf); return NULL; }
   t->file_buf = buf;
   t->file_len = len;

   while (pos + FITS_BLOCK <= len)
     {
        const unsigned char *hdr = buf + pos;
        unsigned int ncards = 0, i;
        int found_end = 0;
        long naxis, bitpix, pcount, gcount;
        size_t datasize = 0;
        char xt[80];

        while (!found_end && (pos + FITS_BLOCK <= len))
          {
             for (i = 0; i < 36; i++)
               {
                  if (card_key (buf + pos + 80 * i, "END"))
                    {
                       found_end = 1;
                       ncards += i + 1;
                       break;
                    }
               }
             if (!found_end) ncards += 36;
             pos += FITS_BLOCK;
          }
        if (!found_end) break;

        naxis = header_long (hdr, ncards, "NAXIS", 0);
        bitpix = header_long (hdr, ncards, "BITPIX", 8);
        pcount = header_long (hdr, ncards, "PCOUNT", 0);
        gcount = header_long (hdr, ncards, "GCOUNT", 1);
        if (naxis > 0)
          {
             long k;
             datasize = (size_t) ((bitpix < 0) ? -bitpix : bitpix) / 8;
             for (k = 1; k <= naxis; k++)
               {
                  char key[32];
                  snprintf (key, sizeof (key), "NAXIS%ld", k);
                  datasize *= (size_t) header_long (hdr, ncards, key, 0);
               }
             datasize = (datasize + (size_t) pcount) * (size_t) gcount;
          }

        if ((0 == header_value (hdr, ncards, "XTENSION", xt, sizeof (xt)))
            && (0 == strcmp (xt, "BINTABLE")))
          {
             long tfields = header_long (hdr, ncards, "TFIELDS", 0);
             unsigned int off = 0;
             long k;
             t->header = hdr;
             t->header_cards = ncards;
             t->data = buf + pos;
             t->row_bytes = (unsigned int) header_long (hdr, ncards, "NAXIS1", 0);
             t->nrows = (unsigned int) header_long (hdr, ncards, "NAXIS2", 0);
             t->ncols = (tfields > FITS_MAX_COLS) ? FITS_MAX_COLS : (unsigned int) tfields;
             for (k = 1; k <= (long) t->ncols; k++)
               {
                  char key[16], form[80], *p;
                  unsigned int rep;
                  snprintf (key, sizeof (key), "TTYPE%ld", k);
                  if (-1 == header_value (hdr, ncards, key, t->colname[k-1], 72))
                    t->colname[k-1][0] = 0;
                  snprintf (key, sizeof (key), "TFORM%ld", k);
                  if (-1 == header_value (hdr, ncards, key, form, sizeof (form)))
                    form[0] = 0;
                  p = form;
                  rep = 1;
                  if (isdigit ((unsigned char) *p))
                    rep = (unsigned int) strtoul (p, &p, 10);
                  t->coltype[k-1] = *p;
                  t->repeat[k-1] = rep;
                  t->offset[k-1] = off;
                  off += rep * tform_size (*p);
               }
             if ((pos + datasize <= len) && (off <= t->row_bytes)
                 && ((match == NULL) || (1 == match (t, client))))
               return t;
          }
        pos += ((datasize + FITS_BLOCK - 1) / FITS_BLOCK) * FITS_BLOCK;
     }
   mxg_fits_close (t);
   return NULL;
}

int mxg_fits_keyword (mxg_fits_table *t, const char *key, char *buf, unsigned int buflen)
{
   return header_value (t->header, t->header_cards, key, buf, buflen);
}

int mxg_fits_keyword_is (mxg_fits_table *t, const char *key, const char *value)
{
   char buf[80];
   if (-1 == mxg_fits_keyword (t, key, buf, sizeof (buf))) return 0;
   return 0 == strcasecmp (buf, value);
}

unsigned int mxg_fits_num_rows (mxg_fits_table *t)
{
   return t->nrows;
}

static int find_col (mxg_fits_table *t, const char *name)
{
   unsigned int i;
   for (i = 0; i < t->ncols; i++)
     if (0 == strcasecmp (t->colname[i], name))
       return (int) i;
   return -1;
}

int mxg_fits_column (mxg_fits_table *t, const char *name, unsigned int *repeat, char *type)
{
   int c = find_col (t, name);
   if (c < 0) return -1;
   if (repeat != NULL) *repeat = t->repeat[c];
   if (type != NULL) *type = t->coltype[c];
   return 0;
}

int mxg_fits_get_double (mxg_fits_table *t, const char *name, unsigned int row,
                         unsigned int j, double *v)
{
   int c = find_col (t, name);
   const unsigned char *b;
   if ((c < 0) || (row >= t->nrows) || (j >= t->repeat[c]))
     {
        mxg_error ("FITS column %s element (%u,%u) not available", name, row, j);
        return -1;
     }
   b = t->data + (size_t) row * t->row_bytes + t->offset[c] + j * tform_size (t->coltype[c]);
   switch (t->coltype[c])
     {
      case 'E': *v = (double) be_float (b); return 0;
      case 'D': *v = be_double (b); return 0;
      case 'J': *v = (double) (int) be32 (b); return 0;
      case 'I': *v = (double) (short) (((unsigned int) b[0] << 8) | b[1]); return 0;
      case 'B': *v = (double) b[0]; return 0;
     }
   mxg_error ("FITS column %s has unsupported type %c", name, t->coltype[c]);
   return -1;
}

int mxg_fits_get_float (mxg_fits_table *t, const char *name, unsigned int row,
                        unsigned int j, float *v)
{
   int c = find_col (t, name);
   if ((c >= 0) && (t->coltype[c] == 'E') && (row < t->nrows) && (j < t->repeat[c]))
     {
        *v = be_float (t->data + (size_t) row * t->row_bytes + t->offset[c] + 4 * j);
        return 0;
     }
   else
     {
        double d;
        if (-1 == mxg_fits_get_double (t, name, row, j, &d)) return -1;
        *v = (float) d;
        return 0;
     }
}

int mxg_fits_get_string (mxg_fits_table *t, const char *name, unsigned int row,
                         char *buf, unsigned int buflen)
{
   int c = find_col (t, name);
   unsigned int n;
   char *e;
   if ((c < 0) || (t->coltype[c] != 'A') || (row >= t->nrows))
     return -1;
   n = t->repeat[c];
   if (n >= buflen) n = buflen - 1;
   memcpy (buf, t->data + (size_t) row * t->row_bytes + t->offset[c], n);
   buf[n] = 0;
   e = buf + strlen (buf);
   while ((e > buf) && (e[-1] == ' ')) *--e = 0;
   return 0;
}
