/*
 * index_type.h -- index type of the drop-in API.
 *
 * Default build of the reference: unsigned 32-bit index (include/index_type.h:67-71 with
 * NMFGPU_UINDEX=1, Makefile:253-259).  The printf/scanf conversions (PRI_IDX, SCN_IDX,
 * include/index_type.h:105-123), the limits (IDX_MAX, IDX_MIN, GPUSIZE_MAX) and the MIN/MAX
 * helpers (:160-170) keep the reference's names and meanings for that build.
 */
#ifndef NMFGPU_INDEX_TYPE_H
#define NMFGPU_INDEX_TYPE_H

#include <inttypes.h>
#include <limits.h>
#include <stdint.h>

typedef uint32_t index_t;

#define INDEX_C(x) ((index_t)(x##U))
#define INDEX_MC(m) (INDEX_C(m))
#define PRI_IDX PRIu32
#define SCN_IDX SCNu32
#define IDX_MAX UINT32_MAX
#define IDX_MIN ((index_t)0)
#define GPUSIZE_MAX (IDX_MAX)

#undef MAX
#define MAX(a, b) (((a) >= (b)) ? (a) : (b))
#undef MIN
#define MIN(a, b) (((a) <= (b)) ? (a) : (b))
#undef MAX3
#define MAX3(a, b, c) (MAX(MAX((a), (b)), (c)))
#undef MIN3
#define MIN3(a, b, c) (MIN(MIN((a), (b)), (c)))

#endif /* NMFGPU_INDEX_TYPE_H */
