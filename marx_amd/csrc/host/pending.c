/* pending.c -- modules not yet built into this library report so. */
#include <string.h>
#include "mxg_host.h"


int mxg_detector_init (Param_File_Type *pf)
{
   char buf[64];
   if (-1 == pf_get_string (pf, "DetectorType", buf, sizeof (buf)))
     return -1;
   if (!strcmp (buf, "NONE"))
     {
        memset (&Mxg_Cfg.det, 0, sizeof (Mxg_Cfg.det));
        return 0;
     }
   mxg_error ("detector module not built yet");
   return -1;
}
