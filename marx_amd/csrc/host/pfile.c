/* pfile.c -- the subset of the pfile API (pfile/src/pfile.h:43-64) that the
 * hot-path init functions call, over IRAF-style parameter files.
 *
 * Each non-comment line is  name,type,mode,value,min,max,"prompt".  Fields
 * may be double-quoted.  Names compare case-insensitively, as in
 * pfile/src/pf.c:616.  Booleans accept yes/no/true/false/1/0.  String
 * values beginning with '$' are replaced by the named environment variable.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <ctype.h>

#include "mxg_host.h"

typedef struct
{
   char *name;
   char type;
   char *value;
}
Par_Entry;

struct _Param_File_Type
{
   Par_Entry *entries;
   unsigned int num, cap;
   char *filename;
};

static char *xstrdup (const char *s)
{
   size_t n = strlen (s) + 1;
   char *d = (char *) malloc (n);
   if (d != NULL) memcpy (d, s, n);
   return d;
}

/* Splits one line into up to nmax comma-separated fields in place. */
static unsigned int split_fields (char *line, char **fields, unsigned int nmax)
{
   unsigned int n = 0;
   char *p = line;

   while (n < nmax)
     {
        char *start, *end;
        int at_end;

        while ((*p == ' ') || (*p == '\t')) p++;
        if (*p == '"')
          {
             start = ++p;
             while ((*p != 0) && (*p != '"')) p++;
             end = p;
             if (*p == '"') p++;
             while ((*p != 0) && (*p != ',') && (*p != '\n')) p++;
          }
        else
          {
             start = p;
             while ((*p != 0) && (*p != ',') && (*p != '\n') && (*p != '\r')) p++;
             end = p;
             while ((end > start) && ((end[-1] == ' ') || (end[-1] == '\t'))) end--;
          }
        at_end = (*p != ',');
        if (*p != 0) p++;
        *end = 0;
        fields[n++] = start;
        if (at_end) break;
     }
   return n;
}

static Par_Entry *find_entry (Param_File_Type *pf, const char *name)
{
   unsigned int i;
   if (pf == NULL) return NULL;
   for (i = 0; i < pf->num; i++)
     if (0 == strcasecmp (pf->entries[i].name, name))
       return pf->entries + i;
   return NULL;
}

static int add_entry (Param_File_Type *pf, const char *name, char type, const char *value)
{
   Par_Entry *e = find_entry (pf, name);
   if (e != NULL)
     {
        char *v = xstrdup (value);
        if (v == NULL) return -1;
        free (e->value);
        e->value = v;
        if (type) e->type = type;
        return 0;
     }
   if (pf->num == pf->cap)
     {
        unsigned int cap = pf->cap ? 2 * pf->cap : 256;
        Par_Entry *ne = (Par_Entry *) realloc (pf->entries, cap * sizeof (Par_Entry));
        if (ne == NULL) return -1;
        pf->entries = ne;
        pf->cap = cap;
     }
   e = pf->entries + pf->num;
   e->name = xstrdup (name);
   e->value = xstrdup (value);
   e->type = type ? type : 's';
   if ((e->name == NULL) || (e->value == NULL)) return -1;
   pf->num++;
   return 0;
}

Param_File_Type *pf_open_parameter_file (char *file, char *mode)
{
   FILE *fp;
   char line[4096];
   Param_File_Type *pf;

   (void) mode;
   if (file == NULL) return NULL;
   if (NULL == (fp = fopen (file, "r")))
     {
        mxg_error ("Unable to open parameter file %s", file);
        return NULL;
     }
   pf = (Param_File_Type *) calloc (1, sizeof (Param_File_Type));
   if (pf == NULL)
     {
        fclose (fp);
        return NULL;
     }
   pf->filename = xstrdup (file);

   while (NULL != fgets (line, sizeof (line), fp))
     {
        char *fields[7];
        unsigned int n;
        char *p = line;
        while (isspace ((unsigned char) *p)) p++;
        if ((*p == '#') || (*p == 0)) continue;
        n = split_fields (p, fields, 7);
        if (n < 4) continue;
        if (-1 == add_entry (pf, fields[0], fields[1][0], fields[3]))
          {
             fclose (fp);
             pf_close_parameter_file (pf);
             return NULL;
          }
     }
   fclose (fp);
   return pf;
}

int pf_close_parameter_file (Param_File_Type *pf)
{
   unsigned int i;
   if (pf == NULL) return -1;
   for (i = 0; i < pf->num; i++)
     {
        free (pf->entries[i].name);
        free (pf->entries[i].value);
     }
   free (pf->entries);
   free (pf->filename);
   free (pf);
   return 0;
}

/* Like pfile's pf_set_value (pfile/src/pfcmd.c:107-128), the parameter must
 * exist in the file. */
int pf_set_value (Param_File_Type *pf, char *name, char *value)
{
   if ((pf == NULL) || (name == NULL) || (value == NULL)) return -1;
   if (find_entry (pf, name) == NULL)
     {
        mxg_error ("Unable to find a parameter named %s.", name);
        return -1;
     }
   return add_entry (pf, name, 0, value);
}

static const char *get_value (Param_File_Type *pf, const char *name)
{
   Par_Entry *e = find_entry (pf, name);
   if (e == NULL)
     {
        mxg_error ("parameter %s not found", name);
        return NULL;
     }
   return e->value;
}

int pf_get_double (Param_File_Type *pf, char *name, double *v)
{
   const char *s = get_value (pf, name);
   char *end;
   if (s == NULL) return -1;
   *v = strtod (s, &end);
   if (end == s)
     {
        mxg_error ("parameter %s: \"%s\" is not a number", name, s);
        return -1;
     }
   return 0;
}

int pf_get_integer (Param_File_Type *pf, char *name, int *v)
{
   const char *s = get_value (pf, name);
   char *end;
   if (s == NULL) return -1;
   *v = (int) strtol (s, &end, 0);
   if (end == s)
     {
        mxg_error ("parameter %s: \"%s\" is not an integer", name, s);
        return -1;
     }
   return 0;
}

int pf_get_boolean (Param_File_Type *pf, char *name, int *v)
{
   const char *s = get_value (pf, name);
   if (s == NULL) return -1;
   if (!strcasecmp (s, "yes") || !strcasecmp (s, "y") || !strcasecmp (s, "true")
       || !strcmp (s, "1"))
     *v = 1;
   else if (!strcasecmp (s, "no") || !strcasecmp (s, "n") || !strcasecmp (s, "false")
            || !strcmp (s, "0"))
     *v = 0;
   else
     {
        mxg_error ("parameter %s: \"%s\" is not a boolean", name, s);
        return -1;
     }
   return 0;
}

int pf_get_string (Param_File_Type *pf, char *name, char *buf, unsigned int buflen)
{
   const char *s = get_value (pf, name);
   if (s == NULL) return -1;
   if ((s[0] == '$') && (s[1] != 0))
     {
        const char *env = getenv (s + 1);
        if (env != NULL) s = env;
     }
   if (strlen (s) + 1 > buflen)
     {
        mxg_error ("parameter %s: value too long", name);
        return -1;
     }
   strcpy (buf, s);
   return 0;
}

int pf_get_file (Param_File_Type *pf, char *name, char *buf, unsigned int buflen)
{
   return pf_get_string (pf, name, buf, buflen);
}
