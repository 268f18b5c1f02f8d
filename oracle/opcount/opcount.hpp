// opcount.hpp -- FP64 operation counting for the oracle (test
// infrastructure).  Force-included when the oracle's C sources are compiled
// as C++: every `double` of the oracle becomes orc_fp, a layout-compatible
// wrapper that counts each add/sub, mul, div and sqrt it executes and each
// call of a transcendental function.  The counts are the algorithmic FP64
// work of the restated reference algorithm, independent of any GPU code.
#ifndef ORC_OPCOUNT_HPP
#define ORC_OPCOUNT_HPP

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <type_traits>
#ifdef _OPENMP
# include <omp.h>
#endif

enum { OC_ADD, OC_MUL, OC_DIV, OC_SQRT, OC_SIN, OC_COS, OC_TAN, OC_ATAN, OC_ATAN2, OC_ASIN,
       OC_ACOS, OC_EXP, OC_EXPM1, OC_LOG, OC_LOG1P, OC_POW, OC_FMOD, OC_ERF, OC_N };
enum { OC_PHASES = 8 };
extern "C" unsigned long long Orc_Opcount[OC_PHASES][OC_N];
extern "C" unsigned long long Orc_Opcount_Marks[OC_PHASES];
extern "C" int Orc_Opcount_Phase;
#define OC_INC(cnt) (Orc_Opcount[Orc_Opcount_Phase][cnt]++)

struct orc_fp
{
   double v;
   orc_fp () = default;
   constexpr orc_fp (double x) : v (x) {}
   template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
   constexpr orc_fp (T x) : v ((double) x) {}
   template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
   explicit operator T () const { return (T) v; }
   orc_fp &operator+= (orc_fp b) { OC_INC (OC_ADD); v += b.v; return *this; }
   orc_fp &operator-= (orc_fp b) { OC_INC (OC_ADD); v -= b.v; return *this; }
   orc_fp &operator*= (orc_fp b) { OC_INC (OC_MUL); v *= b.v; return *this; }
   orc_fp &operator/= (orc_fp b) { OC_INC (OC_DIV); v /= b.v; return *this; }
   orc_fp operator- () const { return orc_fp (-v); }
};
static_assert (sizeof (orc_fp) == 8, "orc_fp must keep the layout of double");

#define OC_BIN(op, cnt) \
   inline orc_fp operator op (orc_fp a, orc_fp b) { OC_INC (cnt); return orc_fp (a.v op b.v); } \
   template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type> \
   inline orc_fp operator op (orc_fp a, T b) { OC_INC (cnt); return orc_fp (a.v op (double) b); } \
   template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type> \
   inline orc_fp operator op (T a, orc_fp b) { OC_INC (cnt); return orc_fp ((double) a op b.v); }
OC_BIN (+, OC_ADD)
OC_BIN (-, OC_ADD)
OC_BIN (*, OC_MUL)
OC_BIN (/, OC_DIV)
#undef OC_BIN

#define OC_CMP(op) \
   inline bool operator op (orc_fp a, orc_fp b) { return a.v op b.v; } \
   template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type> \
   inline bool operator op (orc_fp a, T b) { return a.v op (double) b; } \
   template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type> \
   inline bool operator op (T a, orc_fp b) { return (double) a op b.v; }
OC_CMP (<)
OC_CMP (>)
OC_CMP (<=)
OC_CMP (>=)
OC_CMP (==)
OC_CMP (!=)
#undef OC_CMP

#define OC_F1(name, cnt) \
   inline orc_fp name (orc_fp a) { OC_INC (cnt); return orc_fp (::name (a.v)); }
OC_F1 (sqrt, OC_SQRT)
OC_F1 (sin, OC_SIN)
OC_F1 (cos, OC_COS)
OC_F1 (tan, OC_TAN)
OC_F1 (atan, OC_ATAN)
OC_F1 (asin, OC_ASIN)
OC_F1 (acos, OC_ACOS)
OC_F1 (exp, OC_EXP)
OC_F1 (expm1, OC_EXPM1)
OC_F1 (log, OC_LOG)
OC_F1 (log1p, OC_LOG1P)
OC_F1 (erf, OC_ERF)
#undef OC_F1
inline orc_fp fabs (orc_fp a) { return orc_fp (::fabs (a.v)); }
inline orc_fp floor (orc_fp a) { return orc_fp (::floor (a.v)); }
inline orc_fp atan2 (orc_fp a, orc_fp b) { OC_INC (OC_ATAN2); return orc_fp (::atan2 (a.v, b.v)); }
inline orc_fp pow (orc_fp a, orc_fp b) { OC_INC (OC_POW); return orc_fp (::pow (a.v, b.v)); }
inline orc_fp fmod (orc_fp a, orc_fp b) { OC_INC (OC_FMOD); return orc_fp (::fmod (a.v, b.v)); }
inline bool isfinite (orc_fp a) { return std::isfinite (a.v); }

#define double orc_fp

#endif
