/* oracle_int.h -- internal declarations of the CPU oracle (test
 * infrastructure only, see oracle.h). */
#ifndef MARX_ORACLE_INT_H
#define MARX_ORACLE_INT_H

#include "oracle.h"

#ifndef PI
# define PI 3.14159265358979323846264338327950288
#endif
#define ORC_HBAR_C 1.973269631e-4

typedef JDMVector_Type orc_vec;

typedef struct
{
   unsigned int key[2];
   unsigned int ctr[4];
   unsigned int buf[4];
   int pos;
   int inited;
   int have_gauss;
   double g2;
}
orc_stream;

void orc_stream_init (orc_stream *s, unsigned long long seed, unsigned int tag,
                      unsigned int stage, unsigned int source_id);
double orc_uniform (orc_stream *s);
double orc_gaussian (orc_stream *s);
double orc_expn (orc_stream *s);

/* Per-attribute stream table of one stage pass: the stream of the photon
 * at attribute index (at - pt->attributes), created on first use. */
orc_stream *orc_streams_alloc (unsigned int n);
orc_stream *orc_photon_stream (const orc_ctx *c, orc_stream *streams,
                               const Marx_Photon_Type *pt,
                               const Marx_Photon_Attr_Type *at, unsigned int stage);

orc_vec orc_v (double x, double y, double z);
double orc_v_length (orc_vec a);
void orc_v_normalize (orc_vec *a);
orc_vec orc_v_unit (orc_vec a);
orc_vec orc_v_cross (orc_vec a, orc_vec b);
double orc_v_dot (orc_vec a, orc_vec b);
orc_vec orc_v_ax1_bx2 (double a, orc_vec x1, double b, orc_vec x2);
orc_vec orc_v_ax1_bx2_cx3 (double a, orc_vec x1, double b, orc_vec x2, double c, orc_vec x3);
orc_vec orc_v_sum (orc_vec a, orc_vec b);
orc_vec orc_v_diff (orc_vec a, orc_vec b);
orc_vec orc_v_rotate_vector1 (orc_vec p, orc_vec n, double c, double s);
orc_vec orc_v_rotate_unit_vector1 (orc_vec p, orc_vec n, double c, double s);
orc_vec orc_v_rotate_unit_vector (orc_vec p, orc_vec n, double theta);
orc_vec orc_m3_vmul (const double a[9], orc_vec v);
/* 1 when the ray x + t p is in the shadow of one of the num struts s */
int orc_strut_shadow (const mxg_hrma_cfg *h, orc_vec x, orc_vec p, const mxg_strut *s,
                      unsigned int num);
double orc_fold_theta (const orc_ctx *c, const double *fa, double p);

#endif
