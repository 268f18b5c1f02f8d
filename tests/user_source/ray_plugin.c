/* A USER-source plug-in for the tests (the interface of marx's
 * marx/doc/examples/user-source): every ray is a pure function of the
 * energy it is handed, so that the library and the oracle, calling it in
 * the same order, must make the same photons.  Arguments: "times" makes it
 * give time steps; "stop=E" makes it end the source at the first photon of
 * energy >= E.  It records what user_open_source received. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

double Open_Area, Open_Dir[3];
int Open_Argc, Open_Count, Close_Count;
char Open_Argv0[64], Open_Argv1[64];
static int Times;
static double Stop = 1e30;

int user_open_source (char **argv, int argc, double area, double cx, double cy, double cz)
{
   int i;
   Open_Count++;
   Open_Area = area;
   Open_Dir[0] = cx; Open_Dir[1] = cy; Open_Dir[2] = cz;
   Open_Argc = argc;
   Open_Argv0[0] = Open_Argv1[0] = 0;
   if (argc > 0) strncpy (Open_Argv0, argv[0], 63);
   if (argc > 1) strncpy (Open_Argv1, argv[1], 63);
   Times = 0;
   for (i = 0; i < argc; i++)
     {
        if (0 == strcmp (argv[i], "times")) Times = 1;
        if (0 == strncmp (argv[i], "stop=", 5)) Stop = atof (argv[i] + 5);
     }
   return 0;
}

void user_close_source (void)
{
   Close_Count++;
}

int user_create_ray (double *dt, double *energy, double *cx, double *cy, double *cz)
{
   double e = *energy, a = 2e-4 * sin (7.0 * e), b = 2e-4 * cos (5.0 * e), len;
   double x = Open_Dir[0], y = Open_Dir[1] + a, z = Open_Dir[2] + b;
   if (e >= Stop) return -1;
   len = sqrt (x * x + y * y + z * z);
   *cx = x / len;
   *cy = y / len;
   *cz = z / len;
   *dt = Times ? 0.25 + 0.05 * e : -1.0;
   *energy = (e > 6.0) ? e - 1.0 : -1.0;
   return 0;
}
