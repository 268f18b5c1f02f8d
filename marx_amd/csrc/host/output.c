/* output.c -- marx's output files: the per-column dump files of
 * marx/libsrc/marxio.c (format: marx/doc/dump.txt) and the ray files of
 * marx/libsrc/s-rayfile.c:43-145.
 *
 * Column files are written a column at a time from a buffer of big-endian
 * values (the reference writes each value with its own fwrite).  Rows come
 * either from an AoS photon list (marx_write_photons) or, for the photons
 * resident in HBM, from the device packer (mxg_write_device), which selects
 * the rows and makes the bytes on the GPU.  Both give the same files.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mxg_host.h"
#include "mxg_columns.h"

#define HEADER_SIZE 32
#define ROWS_OFFSET 20          /* magic (4), type (1), name (15) */
#define NAME_BYTES 15

static const unsigned char Magic[4] = {0x83, 0x13, 0x89, 0x8D};

typedef struct
{
   const char *file, *name;
   char type;
   unsigned long mask;
}
column_info;

/* the table of marxio.c:259-289, in its order; indexed by MXG_COL_* */
static const column_info Columns[MXG_NUM_COLS] =
{
   {"b_energy.dat", "B_ENERGY", 'E', MARX_PI_OK},
   {"energy.dat", "ENERGY", 'E', MARX_ENERGY_OK},
   {"time.dat", "TIME", 'E', MARX_TIME_OK},
   {"tag.dat", "TAG", 'J', MARX_TAG_OK},
   {"xpos.dat", "XPOS", 'E', MARX_X_VECTOR_OK},
   {"ypos.dat", "YPOS", 'E', MARX_X_VECTOR_OK},
   {"zpos.dat", "ZPOS", 'E', MARX_X_VECTOR_OK},
   {"xcos.dat", "COSX", 'E', MARX_P_VECTOR_OK},
   {"ycos.dat", "COSY", 'E', MARX_P_VECTOR_OK},
   {"zcos.dat", "COSZ", 'E', MARX_P_VECTOR_OK},
   {"pha.dat", "PHA", 'I', MARX_PULSEHEIGHT_OK},
   {"detector.dat", "CCDID", 'A', MARX_DET_NUM_OK},
   {"xpixel.dat", "CHIPX", 'E', MARX_DET_PIXEL_OK},
   {"ypixel.dat", "CHIPY", 'E', MARX_DET_PIXEL_OK},
   {"hrc_u.dat", "U", 'E', MARX_DET_UV_PIXEL_OK},
   {"hrc_v.dat", "V", 'E', MARX_DET_UV_PIXEL_OK},
   {"mirror.dat", "MIRROR", 'I', MARX_MIRROR_SHELL_OK},
   {"hrcregion.dat", "HRCREGION", 'A', MARX_DET_REGION_OK},
   {"order.dat", "ORDER", 'A', MARX_ORDER_OK},
   {"ofine.dat", "FINE", 'A', MARX_ORDER1_OK},
   {"ocoarse1.dat", "COARSE1", 'A', MARX_ORDER2_OK},
   {"ocoarse2.dat", "COARSE2", 'A', MARX_ORDER3_OK},
   {"ocoarse3.dat", "COARSE3", 'A', MARX_ORDER4_OK},
   {"sky_ra.dat", "RA", 'E', MARX_SKY_DITHER_OK},
   {"sky_dec.dat", "DEC", 'E', MARX_SKY_DITHER_OK},
   {"sky_roll.dat", "ROLL", 'E', MARX_SKY_DITHER_OK},
   {"det_dy.dat", "DET_DY", 'E', MARX_DET_DITHER_OK},
   {"det_dz.dat", "DET_DZ", 'E', MARX_DET_DITHER_OK},
   {"det_theta.dat", "DET_THETA", 'E', MARX_DET_DITHER_OK},
};

/* ---- big-endian values -------------------------------------------------- */

static void put_be32 (unsigned char *b, uint32_t v)
{
   b[0] = (unsigned char) (v >> 24);
   b[1] = (unsigned char) (v >> 16);
   b[2] = (unsigned char) (v >> 8);
   b[3] = (unsigned char) v;
}

static uint32_t get_be32 (const unsigned char *b)
{
   return ((uint32_t) b[0] << 24) | ((uint32_t) b[1] << 16) | ((uint32_t) b[2] << 8) | b[3];
}

static void put_be16 (unsigned char *b, uint16_t v)
{
   b[0] = (unsigned char) (v >> 8);
   b[1] = (unsigned char) v;
}

static uint32_t float_bits (float f)
{
   uint32_t u;
   memcpy (&u, &f, 4);
   return u;
}

/* ---- dump-file headers (marxio.c:79-200, 466-533) -------------------------- */

static int write_i32_at (FILE *fp, long offset, int32_t v)
{
   unsigned char b[4];
   put_be32 (b, (uint32_t) v);
   if ((0 != fseek (fp, offset, SEEK_SET)) || (4 != fwrite (b, 1, 4, fp)))
     return -1;
   return 0;
}

FILE *marx_create_write_dump_file (char *filename, Marx_Dump_File_Type *dft)
{
   unsigned char h[HEADER_SIZE];
   size_t len = 0;
   FILE *fp;

   dft->fp = NULL;
   if (NULL == (fp = fopen (filename, "w+b")))
     {
        mxg_error ("Unable to open %s for writing.", filename);
        return NULL;
     }
   memset (h, 0, sizeof (h));
   memcpy (h, Magic, 4);
   h[4] = (unsigned char) dft->type;
   while ((len < NAME_BYTES) && (len < MARX_MAX_COLUMN_NAME_LEN) && dft->colname[len]) len++;
   memcpy (h + 5, dft->colname, len);
   put_be32 (h + ROWS_OFFSET, (uint32_t) dft->num_rows);
   put_be32 (h + ROWS_OFFSET + 4, (uint32_t) dft->num_cols);
   if (HEADER_SIZE != fwrite (h, 1, HEADER_SIZE, fp))
     {
        fclose (fp);
        mxg_error ("Write to %s failed.", filename);
        return NULL;
     }
   dft->fp = fp;
   return fp;
}

int marx_close_write_dump_file (FILE *fp, unsigned long n)
{
   int ret = 0;
   if (fp == NULL) return -1;
   if (-1 == write_i32_at (fp, ROWS_OFFSET, (int32_t) n))
     {
        mxg_error ("marx_close_write_dump_file: write error");
        ret = -1;
     }
   if (EOF == fclose (fp))
     {
        mxg_error ("marx_close_write_dump_file: write error");
        ret = -1;
     }
   return ret;
}

Marx_Dump_File_Type *marx_open_read_dump_file (char *file)
{
   unsigned char h[HEADER_SIZE];
   Marx_Dump_File_Type *dft;
   FILE *fp;

   if (NULL == (dft = (Marx_Dump_File_Type *) calloc (1, sizeof (*dft))))
     {
        mxg_error ("Memory allocation failure.");
        return NULL;
     }
   if (NULL == (fp = fopen (file, "rb")))
     {
        mxg_error ("Unable to open %s.", file);
        free (dft);
        return NULL;
     }
   if (HEADER_SIZE != fread (h, 1, HEADER_SIZE, fp))
     {
        mxg_error ("Error reading header information for %s.", file);
        fclose (fp);
        free (dft);
        return NULL;
     }
   if (memcmp (h, Magic, 4))
     {
        mxg_error ("File %s does not appear to be a marx output file.  Bad Magic Number.", file);
        fclose (fp);
        free (dft);
        return NULL;
     }
   dft->type = (char) h[4];
   memcpy (dft->colname, h + 5, NAME_BYTES);
   dft->colname[NAME_BYTES] = 0;
   dft->num_rows = (int) get_be32 (h + ROWS_OFFSET);
   dft->num_cols = (int) get_be32 (h + ROWS_OFFSET + 4);
   dft->fp = fp;
   return dft;
}

void marx_close_read_dump_file (Marx_Dump_File_Type *dft)
{
   if (dft == NULL) return;
   if (dft->fp != NULL) (void) fclose (dft->fp);
   free (dft);
}

/* ---- the column files of one write -------------------------------------- */

typedef struct
{
   int ncols;
   int col[MXG_NUM_COLS];
   FILE *fp[MXG_NUM_COLS];
   unsigned long rows[MXG_NUM_COLS];   /* rows in the file before this write */
}
column_set;

/* The columns of mask, in table order (marxio.c:402-412 for the mask). */
static unsigned long effective_mask (unsigned long write_mask, unsigned long history)
{
   if (Mxg_Cfg.src.dither_mode != MXG_DITHER_NONE)
     return write_mask & (history | MARX_SKY_DITHER_OK | MARX_DET_DITHER_OK);
   return write_mask & history;
}

static void close_columns (column_set *cs, int with_counts, unsigned long added, int *status)
{
   int k;
   for (k = 0; k < cs->ncols; k++)
     {
        if (cs->fp[k] == NULL) continue;
        if (with_counts)
          {
             if (-1 == marx_close_write_dump_file (cs->fp[k], cs->rows[k] + added))
               *status = -1;
          }
        else
          (void) fclose (cs->fp[k]);
        cs->fp[k] = NULL;
     }
}

/* Opens (new_file: creates with a header; else: opens for appending, taking
 * the row count from the header) every column file of mask in dir.  On an
 * error the files opened so far are closed unchanged. */
static int open_columns (column_set *cs, const char *dir, unsigned long mask, int new_file)
{
   int c, st = 0;
   memset (cs, 0, sizeof (*cs));
   for (c = 0; c < MXG_NUM_COLS; c++)
     {
        const column_info *ci = Columns + c;
        char *path;
        FILE *fp;
        size_t dl;
        if (0 == (ci->mask & mask)) continue;
        dl = strlen (dir);
        path = (char *) malloc (dl + strlen (ci->file) + 2);
        if (path == NULL)
          {
             mxg_error ("Memory allocation failure.");
             close_columns (cs, 0, 0, &st);
             return -1;
          }
        strcpy (path, dir);
        if ((dl > 0) && (dir[dl - 1] != '/')) strcat (path, "/");
        strcat (path, ci->file);
        cs->rows[cs->ncols] = 0;
        if (new_file)
          {
             Marx_Dump_File_Type dft;
             memset (&dft, 0, sizeof (dft));
             dft.type = ci->type;
             strncpy (dft.colname, ci->name, sizeof (dft.colname) - 1);
             fp = marx_create_write_dump_file (path, &dft);
          }
        else
          {
             unsigned char h[HEADER_SIZE];
             fp = fopen (path, "r+b");
             if ((fp != NULL) && ((HEADER_SIZE != fread (h, 1, HEADER_SIZE, fp)) || memcmp (h, Magic, 4)
                                  || (0 != fseek (fp, 0, SEEK_END))))
               {
                  fclose (fp);
                  fp = NULL;
               }
             else if (fp != NULL)
               cs->rows[cs->ncols] = get_be32 (h + ROWS_OFFSET);
          }
        if (fp == NULL)
          {
             mxg_error ("Unable to open %s.", path);
             free (path);
             close_columns (cs, 0, 0, &st);
             return -1;
          }
        free (path);
        cs->col[cs->ncols] = c;
        cs->fp[cs->ncols] = fp;
        cs->ncols++;
     }
   return 0;
}

/* one column's big-endian values for the rows rows[0..m) of an AoS list */
static void pack_aos_column (int c, const Marx_Photon_Attr_Type *a, const unsigned int *rows,
                             unsigned int m, double total_time, unsigned char *out)
{
   unsigned int q;
   for (q = 0; q < m; q++)
     {
        const Marx_Photon_Attr_Type *at = a + rows[q];
        unsigned char *b4 = out + 4 * (size_t) q;
        switch (c)
          {
           case MXG_COL_B_ENERGY: put_be32 (b4, float_bits (at->pi)); break;
           case MXG_COL_ENERGY: put_be32 (b4, float_bits ((float) at->energy)); break;
           case MXG_COL_TIME: put_be32 (b4, float_bits ((float) (at->arrival_time + total_time))); break;
           case MXG_COL_TAG: put_be32 (b4, (uint32_t) at->tag); break;
           case MXG_COL_XPOS: put_be32 (b4, float_bits ((float) at->x.x)); break;
           case MXG_COL_YPOS: put_be32 (b4, float_bits ((float) at->x.y)); break;
           case MXG_COL_ZPOS: put_be32 (b4, float_bits ((float) at->x.z)); break;
           case MXG_COL_COSX: put_be32 (b4, float_bits ((float) at->p.x)); break;
           case MXG_COL_COSY: put_be32 (b4, float_bits ((float) at->p.y)); break;
           case MXG_COL_COSZ: put_be32 (b4, float_bits ((float) at->p.z)); break;
           case MXG_COL_PHA: put_be16 (out + 2 * (size_t) q, (uint16_t) at->pulse_height); break;
           case MXG_COL_CCDID: out[q] = (unsigned char) at->ccd_num; break;
           case MXG_COL_CHIPX: put_be32 (b4, float_bits (at->y_pixel)); break;
           case MXG_COL_CHIPY: put_be32 (b4, float_bits (at->z_pixel)); break;
           case MXG_COL_U: put_be32 (b4, float_bits (at->u_pixel)); break;
           case MXG_COL_V: put_be32 (b4, float_bits (at->v_pixel)); break;
           case MXG_COL_MIRROR: put_be16 (out + 2 * (size_t) q, (uint16_t) (int16_t) at->mirror_shell); break;
           case MXG_COL_HRCREGION: out[q] = (unsigned char) at->detector_region; break;
           case MXG_COL_ORDER: out[q] = (unsigned char) at->order; break;
           case MXG_COL_FINE: out[q] = (unsigned char) at->support_orders[0]; break;
           case MXG_COL_COARSE1: out[q] = (unsigned char) at->support_orders[1]; break;
           case MXG_COL_COARSE2: out[q] = (unsigned char) at->support_orders[2]; break;
           case MXG_COL_COARSE3: out[q] = (unsigned char) at->support_orders[3]; break;
           case MXG_COL_RA: put_be32 (b4, float_bits (at->dither_state.ra)); break;
           case MXG_COL_DEC: put_be32 (b4, float_bits (at->dither_state.dec)); break;
           case MXG_COL_ROLL: put_be32 (b4, float_bits (at->dither_state.roll)); break;
           case MXG_COL_DET_DY: put_be32 (b4, float_bits (at->dither_state.dy)); break;
           case MXG_COL_DET_DZ: put_be32 (b4, float_bits (at->dither_state.dz)); break;
           case MXG_COL_DET_THETA: put_be32 (b4, float_bits (at->dither_state.dtheta)); break;
           default: break;
          }
     }
}

int marx_write_photons (Marx_Photon_Type *p, unsigned long write_mask, char *dir, int open_mode,
                        double total_time)
{
   column_set cs;
   unsigned int *rows = NULL, m = 0, i;
   unsigned char *buf = NULL;
   int k, status = 0;

   if ((p == NULL) || (dir == NULL)) return -1;
   write_mask = effective_mask (write_mask, p->history);
   if (-1 == open_columns (&cs, dir, write_mask, open_mode))
     return -1;
   if (p->n_photons > 0)
     {
        rows = (unsigned int *) malloc (p->n_photons * sizeof (unsigned int));
        buf = (unsigned char *) malloc ((size_t) p->n_photons * 4);
        if ((rows == NULL) || (buf == NULL))
          {
             mxg_error ("Memory allocation failure.");
             free (rows);
             free (buf);
             close_columns (&cs, 0, 0, &status);
             return -1;
          }
     }
   /* the photons written: not flagged bad, in list order (marxio.c:425-435) */
   for (i = 0; i < p->n_photons; i++)
     if (0 == (p->attributes[i].flags & BAD_PHOTON_MASK))
       rows[m++] = i;
   for (k = 0; k < cs.ncols; k++)
     {
        const int c = cs.col[k];
        pack_aos_column (c, p->attributes, rows, m, total_time, buf);
        if (m != fwrite (buf, MXG_COL_WIDTH (c), m, cs.fp[k]))
          {
             mxg_error ("write error.");
             free (rows);
             free (buf);
             close_columns (&cs, 0, 0, &status);
             return -1;
          }
     }
   free (rows);
   free (buf);
   close_columns (&cs, 1, m, &status);
   if (status == -1) mxg_error ("write error.");
   return status;
}

int mxg_write_device (unsigned long write_mask, char *dir, int open_mode, double total_time,
                      unsigned int *num_written)
{
   column_set cs;
   unsigned long history;
   const unsigned char *bytes = NULL;
   size_t off[MXG_NUM_COLS];
   unsigned int m = 0;
   int k, status = 0;

   if (num_written) *num_written = 0;
   if (dir == NULL) return -1;
   if (-1 == mxg_resident_history (&history))
     return -1;
   write_mask = effective_mask (write_mask, history);
   if (-1 == open_columns (&cs, dir, write_mask, open_mode))
     return -1;
   if (-1 == mxgd_pack_output (cs.col, cs.ncols, total_time, &m, &bytes, off))
     {
        mxg_error ("%s", mxgd_last_error ());
        close_columns (&cs, 0, 0, &status);
        return -1;
     }
   for (k = 0; k < cs.ncols; k++)
     if ((m > 0) && (m != fwrite (bytes + off[k], MXG_COL_WIDTH (cs.col[k]), m, cs.fp[k])))
       {
          mxg_error ("write error.");
          close_columns (&cs, 0, 0, &status);
          return -1;
       }
   close_columns (&cs, 1, m, &status);
   if (status == -1)
     {
        mxg_error ("write error.");
        return -1;
     }
   if (num_written) *num_written = m;
   return 0;
}

/* ---- text dump of column files (marxio.c:506-620, 697-756) ----------------- */

static int read_be (FILE *fp, int width, unsigned char *b)
{
   return (int) fread (b, 1, (size_t) width, fp) == width;
}

int marx_dump (int argc, char **argv)
{
   Marx_Dump_File_Type **files;
   int nf, i, k, ret = 0;
   long num_rows = 0;

   if (argc < 3)
     {
        fprintf (stderr, "Usage: marx --dump filename [filename...]\n");
        return -1;
     }
   nf = argc - 2;
   files = (Marx_Dump_File_Type **) calloc ((size_t) nf, sizeof (*files));
   if (files == NULL) return -1;
   for (i = 0; i < nf; i++)
     if (NULL == (files[i] = marx_open_read_dump_file (argv[i + 2])))
       {
          mxg_error ("Unable to open %s.", argv[i + 2]);
          for (k = 0; k < i; k++) marx_close_read_dump_file (files[k]);
          free (files);
          return -1;
       }

   fputc ('#', stdout);
   for (i = 0; i < nf; i++)
     {
        int nc = files[i]->num_cols ? files[i]->num_cols : 1;
        while (nc--) fprintf (stdout, "%16s", files[i]->colname);
     }
   fputc ('\n', stdout);

   for (;;)
     {
        int not_done = 1;
        for (i = 0; i < nf; i++)
          {
             Marx_Dump_File_Type *d = files[i];
             int nc = d->num_cols ? d->num_cols : 1;
             unsigned char b[8];
             while (nc && not_done)
               {
                  switch (d->type)
                    {
                     case 'A':
                       if (read_be (d->fp, 1, b)) fprintf (stdout, "%16d", (int) (signed char) b[0]);
                       else not_done = 0;
                       break;
                     case 'I':
                       if (read_be (d->fp, 2, b))
                         fprintf (stdout, "%16d", (int) (int16_t) (((uint16_t) b[0] << 8) | b[1]));
                       else not_done = 0;
                       break;
                     case 'J':
                       if (read_be (d->fp, 4, b)) fprintf (stdout, "%16ld", (long) (int32_t) get_be32 (b));
                       else not_done = 0;
                       break;
                     case 'E':
                       if (read_be (d->fp, 4, b))
                         {
                            uint32_t u = get_be32 (b);
                            float f;
                            memcpy (&f, &u, 4);
                            fprintf (stdout, "%16e", f);
                         }
                       else not_done = 0;
                       break;
                     case 'D':
                       if (read_be (d->fp, 8, b))
                         {
                            uint64_t u = ((uint64_t) get_be32 (b) << 32) | get_be32 (b + 4);
                            double x;
                            memcpy (&x, &u, 8);
                            fprintf (stdout, "%16e", x);
                         }
                       else not_done = 0;
                       break;
                     default:
                       mxg_error ("Unknown data type: '%c'", d->type);
                       ret = -1;
                       goto done;
                    }
                  nc--;
               }
             if (not_done == 0)
               {
                  if (num_rows != d->num_rows)
                    {
                       mxg_error ("Read error. (%ld/%d rows)", num_rows, d->num_rows);
                       ret = -1;
                       goto done;
                    }
                  break;
               }
          }
        num_rows++;
        putc ('\n', stdout);
        if (not_done == 0) break;
     }
done:
   for (i = 0; i < nf; i++) marx_close_read_dump_file (files[i]);
   free (files);
   fflush (stdout);
   return ret;
}

/* ---- ray files (s-rayfile.c:43-145) --------------------------------------- */

int marx_dump_to_rayfile (char *file, int new_file, Marx_Photon_Type *p, double total_time)
{
   unsigned long magic = MARX_RAYFILE_MAGIC;
   unsigned int i;
   FILE *fp;
   int ret = -1;

   if ((file == NULL) || (p == NULL)) return -1;
   fp = fopen (file, new_file ? "wb" : "ab");
   if (fp == NULL) return -1;
   if (new_file
       && ((1 != fwrite (&magic, sizeof (unsigned long), 1, fp))
           || (1 != fwrite (&p->history, sizeof (unsigned long), 1, fp))))
     goto done;
   for (i = 0; i < p->n_photons; i++)
     {
        Marx_Photon_Attr_Type *at = p->attributes + i;
        if (at->flags & BAD_PHOTON_MASK) continue;
        at->arrival_time += total_time;
        if (1 != fwrite (at, sizeof (Marx_Photon_Attr_Type), 1, fp))
          goto done;
     }
   ret = 0;
done:
   if (EOF == fclose (fp)) ret = -1;
   return ret;
}

FILE *mxg_open_rayfile (const char *file, unsigned long *history)
{
   unsigned long magic;
   FILE *fp = fopen (file, "rb");
   if (fp == NULL)
     {
        mxg_error ("Unable to open rayfile %s.", file);
        return NULL;
     }
   if ((1 != fread (&magic, sizeof (magic), 1, fp)) || (magic != MARX_RAYFILE_MAGIC)
       || (1 != fread (history, sizeof (unsigned long), 1, fp)))
     {
        fclose (fp);
        mxg_error ("Magic number mismatch in rayfile %s.", file);
        return NULL;
     }
   return fp;
}

int marx_dump_rayfile (char *file)
{
   Marx_Photon_Attr_Type at;
   unsigned long history;
   FILE *fp = mxg_open_rayfile (file, &history);

   if (fp == NULL) return -1;
   fprintf (stdout, "#Energy\t\tTime");
   fprintf (stdout, "\t\tX\t\tY\t\tZ");
   fprintf (stdout, "\t\tPX\t\tPY\t\tPZ");
   fprintf (stdout, "\tY-Pixel\tZ-Pixel");
   fprintf (stdout, "\tOrder\tMirrorShell\tCCD_Num\tHRC_Region");
   fprintf (stdout, "\tRA\tDEC\tROLL");
   fprintf (stdout, "\tDY\tDZ\tDTHETA");
   putc ('\n', stdout);
   while (1 == fread (&at, sizeof (at), 1, fp))
     {
        fprintf (stdout, "%f\t%f\t", at.energy, at.arrival_time);
        fprintf (stdout, "%f\t%f\t%f\t", at.x.x, at.x.y, at.x.z);
        fprintf (stdout, "%f\t%f\t%f\t", at.p.x, at.p.y, at.p.z);
        fprintf (stdout, "%f\t%f\t", at.y_pixel, at.z_pixel);
        fprintf (stdout, "%d\t%u\t%d\t%d", at.order, at.mirror_shell, at.ccd_num, at.detector_region);
        putc ('\n', stdout);
     }
   fclose (fp);
   fflush (stdout);
   return 0;
}
