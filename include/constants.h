/* Physical constants (same values as engine/core/constants.h:4-6; the
 * exact digits are part of FP64 parity). */
#ifndef RFSYN_FDTD_CONSTANTS_H
#define RFSYN_FDTD_CONSTANTS_H

#define C0    299792458.0
#define EPS0  8.8541878128e-12
#define MU0   1.25663706212e-6

#endif
