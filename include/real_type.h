/*
 * real_type.h -- data type of the drop-in API.
 *
 * Default build of the reference: single precision (include/real_type.h:81-85 with
 * NMFGPU_SINGLE_PREC=1), R_MIN = FLT_MIN (include/real_type.h:120-126).  The ISO-C helpers
 * its host sources use are defined with the same names and meanings for that build:
 * SCNgREAL (include/real_type.h:133-135), STRTOREAL (:169-174), SQRTR (:180-186), and the
 * <math.h>/<string.h>/<float.h> includes the reference header provides (:71-73).
 */
#ifndef NMFGPU_REAL_TYPE_H
#define NMFGPU_REAL_TYPE_H

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef float real;

#define REAL_C(x) ((real)(x##f))
#define REAL_MC(m) (REAL_C(m))
#define R_MIN (FLT_MIN)
#define R_MAX (FLT_MAX)
#define R_EPSILON (FLT_EPSILON)

/* scanf(3) conversion of a real. */
#define SCNgREAL "g"
/* String to real. */
#define STRTOREAL strtof
/* Square root of a real. */
#define SQRTR sqrtf

#endif /* NMFGPU_REAL_TYPE_H */
