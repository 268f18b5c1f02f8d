/* orc_trace.c -- whole-pipeline oracle run used as the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Runs the reference's host loop
 * body (marx/src/marx.c:545-608: create photons, then process_photons
 * :240-273) over independent chunks, one OpenMP thread per chunk.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
# include <omp.h>
#endif

#include "oracle_int.h"

static Marx_Photon_Type *alloc_list (unsigned int num)
{
   Marx_Photon_Type *pt = (Marx_Photon_Type *) calloc (1, sizeof (Marx_Photon_Type));
   if (pt == NULL) return NULL;
   pt->attributes = (Marx_Photon_Attr_Type *) calloc (num, sizeof (Marx_Photon_Attr_Type));
   pt->sorted_energies = (double *) calloc (num, sizeof (double));
   pt->max_n_photons = num;
   if ((pt->attributes == NULL) || (pt->sorted_energies == NULL))
     {
        free (pt->attributes);
        free (pt->sorted_energies);
        free (pt);
        return NULL;
     }
   return pt;
}

static void free_list (Marx_Photon_Type *pt)
{
   if (pt == NULL) return;
   free (pt->attributes);
   free (pt->sorted_energies);
   free (pt->sorted_index);
   free (pt);
}

long orc_trace_parallel (const orc_ctx *c, unsigned int num, unsigned int chunk, int nthreads)
{
   unsigned int nchunks, k;
   long detected = 0;
   int failed = 0;

   if (chunk == 0) chunk = 100000;
   nchunks = (num + chunk - 1) / chunk;
#ifdef _OPENMP
   if (nthreads > 0) omp_set_num_threads (nthreads);
#else
   (void) nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+:detected)
   for (k = 0; k < nchunks; k++)
     {
        unsigned int m = ((k + 1) * chunk > num) ? num - k * chunk : chunk, n;
        Marx_Photon_Type *pt = alloc_list (m);
        if (pt == NULL) { failed = 1; continue; }
        pt->tag_start = k * chunk;
        if ((0 != orc_create_photons (c, pt, m, &n, NULL))
            || (0 != orc_mirror_reflect (c, pt))
            || (0 != orc_grating_diffract (c, pt))
            || (0 != orc_detect (c, pt)))
          failed = 1;
        orc_prune_photons (pt);
        detected += pt->num_sorted;
        free_list (pt);
     }
   return failed ? -1 : detected;
}

/* Batched point sources traced source by source, one OpenMP task per
 * source (the reference runs one source per process).  Source k's photons
 * get tags tag_start + off[k] ..., as in one batch pass.  Returns the
 * number of photons detected, -1 on error. */
long orc_trace_batch_parallel (const orc_ctx *c, unsigned int nsrc, const double *dirs,
                               const unsigned int *counts, unsigned int tag_start,
                               double exposure, int nthreads)
{
   unsigned int *off = (unsigned int *) malloc ((nsrc + 1) * sizeof (unsigned int));
   long detected = 0;
   int failed = 0;
   unsigned int k;

   if (off == NULL) return -1;
   off[0] = 0;
   for (k = 0; k < nsrc; k++) off[k + 1] = off[k] + counts[k];
#ifdef _OPENMP
   if (nthreads > 0) omp_set_num_threads (nthreads);
#else
   (void) nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+:detected)
   for (k = 0; k < nsrc; k++)
     {
        Marx_Photon_Type *pt = alloc_list (counts[k] ? counts[k] : 1);
        if (pt == NULL) { failed = 1; continue; }
        if ((0 != orc_create_batch (c, pt, 1, dirs + 3 * k, counts + k, tag_start + off[k], 0.0,
                                    exposure))
            || (0 != orc_mirror_reflect (c, pt))
            || (0 != orc_grating_diffract (c, pt))
            || (0 != orc_detect (c, pt)))
          failed = 1;
        orc_prune_photons (pt);
        detected += pt->num_sorted;
        free_list (pt);
     }
   free (off);
   return failed ? -1 : detected;
}
