// Reference: include/tmfqs/core/constants.h:6.
#ifndef TMFQS_CORE_CONSTANTS_H
#define TMFQS_CORE_CONSTANTS_H
namespace tmfqs {
inline constexpr double kPi = 3.14159265358979323846;
}
#endif
