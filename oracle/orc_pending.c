/* orc_pending.c -- oracle passes not restated yet (test infrastructure). */
#include "oracle_int.h"

int orc_create_photons (const orc_ctx *c, Marx_Photon_Type *pt, unsigned int num,
                        unsigned int *num_collected, double *exposure_time)
{ (void) c; (void) pt; (void) num; (void) num_collected; (void) exposure_time; return -1; }
long orc_trace_parallel (const orc_ctx *c, unsigned int num, unsigned int chunk, int nthreads)
{ (void) c; (void) num; (void) chunk; (void) nthreads; return -1; }
