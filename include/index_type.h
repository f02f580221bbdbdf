/*
 * index_type.h -- index type of the drop-in API.
 *
 * Default build of the reference: unsigned 32-bit index (include/index_type.h:67-71 with
 * NMFGPU_UINDEX=1, Makefile:253-259).
 */
#ifndef NMFGPU_INDEX_TYPE_H
#define NMFGPU_INDEX_TYPE_H

#include <inttypes.h>
#include <stdint.h>

typedef uint32_t index_t;

#define INDEX_C(x) ((index_t)(x##U))
#define PRI_IDX PRIu32
#define IDX_MAX UINT32_MAX

#endif /* NMFGPU_INDEX_TYPE_H */
