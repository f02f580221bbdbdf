/*
 * real_type.h -- data type of the drop-in API.
 *
 * Default build of the reference: single precision (include/real_type.h:81-85 with
 * NMFGPU_SINGLE_PREC=1), R_MIN = FLT_MIN (include/real_type.h:120-126).
 */
#ifndef NMFGPU_REAL_TYPE_H
#define NMFGPU_REAL_TYPE_H

#include <float.h>

typedef float real;

#define REAL_C(x) ((real)(x##f))
#define R_MIN (FLT_MIN)
#define R_MAX (FLT_MAX)
#define SQRTR sqrtf

#endif /* NMFGPU_REAL_TYPE_H */
