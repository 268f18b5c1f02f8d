// opcount_api.cpp -- counters of the op-counting oracle build (test
// infrastructure).
#include <string.h>
enum { OC_N_ = 17 };
extern "C" {
unsigned long long Orc_Opcount[OC_N_];
void orc_opcount_reset (void) { memset (Orc_Opcount, 0, sizeof (Orc_Opcount)); }
void orc_opcount_get (unsigned long long *out) { memcpy (out, Orc_Opcount, sizeof (Orc_Opcount)); }
}
