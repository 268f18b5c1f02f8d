/* marx_driver.c -- the per-chunk sequence of the reference's driver
 * (marx/src/marx.c:285-373 setup_tstart, :641-656 module_init, :545-560
 * marx_create_photons, :240-273 process_photons) as a plain C program
 * linked against libmarx_gpu.so, with a parameter-file reader of its own.
 *
 * Test program (tests/test_cpu_driver.py, tests/test_gpu_driver.py).  Its
 * pf_* functions use a record layout different from the library's: when
 * the library's init functions read parameters through this program's
 * handle, they can only succeed if their pf_* calls bind to these
 * definitions (ELF interposition by the executable), which is what a marx
 * executable linked with the reference's libpfile.a relies on.
 *
 *    marx_driver PARFILE NUM OUTFILE
 * prints "modules M G D", "pf_calls N", "process RC"; on success writes
 * the n_photons records, then num_sorted and sorted_index, to OUTFILE.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "marx_gpu.h"

/* ---- a parameter-file reader of the driver's own --------------------------------- */

struct _Param_File_Type
{
   unsigned int magic;
   unsigned int n;
   struct { char value[512]; char name[64]; } *par;   /* not the library's layout */
};

static unsigned long Pf_Calls = 0;

static int split_par_line (char *line, char **f, int nmax)
{
   int n = 0;
   char *p = line;
   while (n < nmax)
     {
        char *start;
        if (*p == '"')
          {
             start = ++p;
             while (*p && (*p != '"')) p++;
             if (*p == '"') *p++ = 0;
             while (*p && (*p != ',')) p++;
          }
        else
          {
             start = p;
             while (*p && (*p != ',') && (*p != '\n')) p++;
          }
        f[n++] = start;
        if (*p != ',') { *p = 0; break; }
        *p++ = 0;
     }
   return n;
}

Param_File_Type *pf_open_parameter_file (char *file, char *mode)
{
   FILE *fp = fopen (file, "r");
   char line[4096];
   Param_File_Type *pf;
   (void) mode;
   Pf_Calls++;
   if (fp == NULL) return NULL;
   pf = (Param_File_Type *) calloc (1, sizeof (*pf));
   pf->magic = 0xC0FFEEu;
   pf->par = calloc (4096, sizeof (*pf->par));
   while (fgets (line, sizeof (line), fp))
     {
        char *f[8];
        if ((line[0] == '#') || (line[0] == '\n')) continue;
        if (split_par_line (line, f, 8) < 4) continue;
        strncpy (pf->par[pf->n].name, f[0], 63);
        strncpy (pf->par[pf->n].value, f[3], 511);
        pf->n++;
     }
   fclose (fp);
   return pf;
}

int pf_close_parameter_file (Param_File_Type *pf)
{
   Pf_Calls++;
   if ((pf == NULL) || (pf->magic != 0xC0FFEEu)) return -1;
   free (pf->par);
   free (pf);
   return 0;
}

static const char *value_of (Param_File_Type *pf, const char *name)
{
   unsigned int i;
   Pf_Calls++;
   if ((pf == NULL) || (pf->magic != 0xC0FFEEu))
     {
        fprintf (stderr, "marx_driver: a foreign handle reached the driver's pfile\n");
        exit (3);
     }
   for (i = pf->n; i-- > 0;)
     if (0 == strcasecmp (pf->par[i].name, name))     /* parameter names are case-blind */
       return pf->par[i].value;
   return NULL;
}

int pf_set_value (Param_File_Type *pf, char *name, char *value)
{
   if (value_of (pf, name) == NULL) return -1;
   strncpy (pf->par[pf->n].name, name, 63);
   strncpy (pf->par[pf->n].value, value, 511);
   pf->n++;
   return 0;
}

int pf_get_double (Param_File_Type *pf, char *name, double *v)
{
   const char *s = value_of (pf, name);
   if (s == NULL) return -1;
   *v = atof (s);
   return 0;
}

int pf_get_integer (Param_File_Type *pf, char *name, int *v)
{
   const char *s = value_of (pf, name);
   if (s == NULL) return -1;
   *v = atoi (s);
   return 0;
}

int pf_get_boolean (Param_File_Type *pf, char *name, int *v)
{
   const char *s = value_of (pf, name);
   if (s == NULL) return -1;
   *v = (0 == strcasecmp (s, "yes")) || (0 == strcasecmp (s, "y")) || (0 == strcmp (s, "1"));
   return 0;
}

int pf_get_string (Param_File_Type *pf, char *name, char *buf, unsigned int buflen)
{
   const char *s = value_of (pf, name);
   if ((s == NULL) || (buflen == 0)) return -1;
   strncpy (buf, s, buflen - 1);
   buf[buflen - 1] = 0;
   return 0;
}

int pf_get_file (Param_File_Type *pf, char *name, char *buf, unsigned int buflen)
{
   return pf_get_string (pf, name, buf, buflen);
}

/* ---- the driver's sequence ------------------------------------------------------ */

static int process_photons (Marx_Photon_Type *pt)
{
   if (-1 == marx_mirror_reflect (pt, 0)) return -1;
   if (-1 == marx_grating_diffract (pt, 0)) return -1;
   if (-1 == marx_detect (pt, 0)) return -1;
   return 0;
}

int main (int argc, char **argv)
{
   Param_File_Type *pf;
   Marx_Source_Type *st;
   Marx_Photon_Type *pt;
   unsigned int num, num_collected = 0;
   double tstart;
   int m, g, d, rc;
   FILE *fp;

   if (argc != 4)
     {
        fprintf (stderr, "usage: marx_driver PARFILE NUM OUTFILE\n");
        return 2;
     }
   num = (unsigned int) atoi (argv[2]);
   if (NULL == (pf = pf_open_parameter_file (argv[1], "r")))
     return 2;
   /* setup_tstart: TStart in years */
   if (-1 == pf_get_double (pf, "TStart", &tstart)) return 2;
   marx_set_time (tstart, (tstart - 1998.0) * 365.25 * 24.0 * 3600.0);

   m = marx_mirror_init (pf);
   g = (m == -1) ? -1 : marx_grating_init (pf);
   d = (g == -1) ? -1 : marx_detector_init (pf);
   printf ("modules %d %d %d\n", m, g, d);
   if ((m == -1) || (g == -1) || (d == -1)) return 1;
   if ((NULL == (st = marx_create_source (pf))) || (-1 == marx_open_source (st)))
     return 1;
   printf ("pf_calls %lu\n", Pf_Calls);
   if (NULL == (pt = marx_alloc_photon_type (num)))
     return 1;
   rc = marx_create_photons (st, pt, num, &num_collected, NULL);
   if (rc == 0) rc = process_photons (pt);
   printf ("process %d\n", rc);
   if (rc == 0)
     {
        if (NULL == (fp = fopen (argv[3], "wb"))) return 1;
        fwrite (&pt->n_photons, sizeof (unsigned int), 1, fp);
        fwrite (pt->attributes, sizeof (Marx_Photon_Attr_Type), pt->n_photons, fp);
        fwrite (&pt->num_sorted, sizeof (unsigned int), 1, fp);
        fwrite (pt->sorted_index, sizeof (unsigned int), pt->num_sorted, fp);
        fclose (fp);
     }
   else
     printf ("error %s\n", mxg_last_error ());
   marx_dealloc_photon_type (pt);
   marx_close_source (st);
   pf_close_parameter_file (pf);
   return rc == 0 ? 0 : 4;
}
