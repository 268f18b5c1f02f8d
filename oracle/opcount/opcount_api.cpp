// opcount_api.cpp -- counters of the op-counting oracle build (test
// infrastructure).
#include <string.h>
enum { OC_N_ = 17, OC_PHASES_ = 8 };
extern "C" {
unsigned long long Orc_Opcount[OC_PHASES_][OC_N_];
unsigned long long Orc_Opcount_Marks[OC_PHASES_];
int Orc_Opcount_Phase;

void orc_opcount_reset (void)
{
   memset (Orc_Opcount, 0, sizeof (Orc_Opcount));
   memset (Orc_Opcount_Marks, 0, sizeof (Orc_Opcount_Marks));
   Orc_Opcount_Phase = 0;
}

// counts summed over the buckets
void orc_opcount_get (unsigned long long *out)
{
   for (int k = 0; k < OC_N_; k++)
     {
        out[k] = 0;
        for (int ph = 0; ph < OC_PHASES_; ph++) out[k] += Orc_Opcount[ph][k];
     }
}

void orc_opcount_get_phase (int phase, unsigned long long *out)
{
   memcpy (out, Orc_Opcount[phase], sizeof (Orc_Opcount[phase]));
}

unsigned long long orc_opcount_marks (int phase) { return Orc_Opcount_Marks[phase]; }
}
