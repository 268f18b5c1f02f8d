/* orc_photon.c -- oracle photon-list bookkeeping: energy sort, prune and
 * the per-photon stream tables.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *   sort:  marx/libsrc/source.c:361-371 with jdmath/src/dsort.c:28-61
 *          (qsort of indices by energy, index tie-break)
 *   prune: marx/libsrc/photon.c:40-63
 */
#ifndef _GNU_SOURCE
# define _GNU_SOURCE   /* qsort_r */
#endif
#include <stdlib.h>
#include <string.h>

#include "oracle_int.h"

/* the energies travel with the sort (qsort_r), so that concurrent sorts on
 * OpenMP threads do not share them */
static int energy_cmp (const void *a, const void *b, void *energies)
{
   const double *en = (const double *) energies;
   unsigned int ia = *(const unsigned int *) a, ib = *(const unsigned int *) b;
   double x = en[ia], y = en[ib];
   if (x > y) return 1;
   if (x < y) return -1;
   if (ia > ib) return 1;
   if (ia == ib) return 0;
   return -1;
}

int orc_sort_photons (Marx_Photon_Type *pt)
{
   unsigned int n = pt->n_photons, i;
   double *e;
   unsigned int *idx;

   e = (double *) malloc ((n ? n : 1) * sizeof (double));
   idx = (unsigned int *) malloc ((n ? n : 1) * sizeof (unsigned int));
   if ((e == NULL) || (idx == NULL))
     {
        free (e);
        free (idx);
        return -1;
     }
   for (i = 0; i < n; i++)
     {
        e[i] = pt->attributes[i].energy;
        idx[i] = i;
     }
   qsort_r (idx, n, sizeof (unsigned int), energy_cmp, e);
   /* a list made by marx_alloc_photon_type holds its index in the library's
    * (pinned) host memory, which free() must not see: the order is copied
    * into it */
   if (pt->sorted_index == NULL)
     {
        pt->sorted_index = idx;
        idx = NULL;
     }
   else
     memcpy (pt->sorted_index, idx, n * sizeof (unsigned int));
   free (idx);
   if (pt->sorted_energies == NULL)
     pt->sorted_energies = (double *) malloc ((pt->max_n_photons ? pt->max_n_photons : 1) * sizeof (double));
   for (i = 0; i < n; i++)
     pt->sorted_energies[i] = e[pt->sorted_index[i]];
   pt->num_sorted = n;
   free (e);
   return 0;
}

void orc_prune_photons (Marx_Photon_Type *pt)
{
   unsigned int i, new_num = 0;
   for (i = 0; i < pt->num_sorted; i++)
     {
        unsigned int indx = pt->sorted_index[i];
        Marx_Photon_Attr_Type *at = pt->attributes + indx;
        if (at->flags & BAD_PHOTON_MASK) continue;
        pt->sorted_energies[new_num] = at->energy;
        pt->sorted_index[new_num] = indx;
        new_num++;
     }
   pt->num_sorted = new_num;
}

orc_stream *orc_streams_alloc (unsigned int n)
{
   return (orc_stream *) calloc (n ? n : 1, sizeof (orc_stream));
}

orc_stream *orc_photon_stream (const orc_ctx *c, orc_stream *streams,
                               const Marx_Photon_Type *pt,
                               const Marx_Photon_Attr_Type *at, unsigned int stage)
{
   orc_stream *s = streams + (at - pt->attributes);
   if (!s->inited)
     orc_stream_init (s, c->seed, at->tag, stage, c->source_id);
   return s;
}
