// Complex helpers (reference: include/tmfqs/core/math.h:14-27).
#ifndef TMFQS_CORE_MATH_H
#define TMFQS_CORE_MATH_H
#include "tmfqs/core/types.h"
namespace tmfqs {
Amplitude amplitudeAdd(Amplitude a, Amplitude b);
Amplitude amplitudeMultiply(Amplitude a, Amplitude b);
/** e^(x + i y) = e^x (cos y + i sin y). */
Amplitude complexExp(Amplitude exponent);
}  // namespace tmfqs
#endif
