/*
 * common.h -- drop-in for the reference's include/common.h (globals, options, helpers).
 *
 * Mirrors include/common.h:150-534 of NMF-mGPU for a single process: the matrix
 * dimensions and host buffers are process globals (include/common.h:211-245), options come
 * from the same getopt flags (src/common.c:847-1119), and get_difference keeps the
 * connectivity semantics of src/common.c:1195-1209.
 */
#ifndef NMFGPU_COMMON_H
#define NMFGPU_COMMON_H

#include <stdbool.h>
#include <stddef.h>

#include "index_type.h"
#include "real_type.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Defaults (include/common_defaults.h:54-96). */
#define FIXED_SEED (INDEX_C(3))
#define DEFAULT_K (INDEX_C(2))
#define DEFAULT_NITERS (INDEX_C(2000))
#define DEFAULT_NITER_CONV (INDEX_C(10))
#define DEFAULT_STOP_THRESHOLD (INDEX_C(40))
#define DEFAULT_GPU_DEVICE (INDEX_C(0))

/* File formats (include/common.h:160-182). */
typedef enum file_format {
	ASCII_TEXT_FMT,        /* tab-separated text, optional headers/labels */
	NON_NATIVE_BINARY_FMT, /* little endian, FP64 data, uint32 dimensions, file signature */
	NATIVE_BINARY_FMT      /* native endianness and types, no signature */
} file_fmt_t;

/* Command-line arguments (include/common.h:186-205). */
struct input_arguments {
	char const *filename;
	bool numeric_hdrs;
	bool numeric_lbls;
	file_fmt_t input_file_fmt;
	file_fmt_t output_file_fmt;
	index_t k;
	index_t kp;
	index_t nIters;
	index_t niter_test_conv;
	index_t stop_threshold;
	index_t gpu_device;
	index_t idx_other_args;
};

/* Globals (include/common.h:211-245). */
extern index_t process_id, num_processes, num_act_processes;
extern index_t memory_alignment;
extern size_t matrix_max_num_items;
extern index_t matrix_max_pitch, matrix_max_non_padded_dim;
extern index_t N, M, K;
extern index_t NpP, MpP, bN, bM;
extern index_t Mp, Kp, MpPp;
extern index_t *classification, *last_classification;
extern real *W, *H, *Vcol, *Vrow;
extern char const *const file_extension[3];

/* Messages: stdout for information, stderr for errors (src/common.c:249-566).  Return
 * EXIT_SUCCESS or EXIT_FAILURE. */
int print_message(bool all_processes, char const *fmt, ...);
int append_printed_message(bool all_processes, char const *fmt, ...);
int print_error(bool all_processes, char const *fmt, ...);
int append_printed_error(bool all_processes, char const *fmt, ...);
int print_errnum(bool all_processes, int errnum, char const *fmt, ...);
int append_printed_errnum(bool all_processes, int errnum, char const *fmt, ...);
int flush_output(bool permanently);

/* Help texts (src/common.c:567-800). */
int help_file_formats(void);
int help_nmf(void);
int print_nmf_gpu_help(char const *execname);

/* Matrix limits (src/common.c:725-834): alignment (0 = default, 16 items), largest
 * dimension (0 = none) and largest number of items (0 = none).  Return EXIT_SUCCESS. */
int set_matrix_limits(index_t data_alignment, index_t max_dimension, size_t max_nitems);
int set_default_matrix_limits(void);
index_t get_padding(index_t dim);

/* Parses argv with the reference's option string ":B:b:CcE:e:HhI:i:J:j:K:k:RrT:t:Z:z:"
 * (src/common.c:847-1119); help_requested is set for -h. */
int check_arguments(int argc, char const *restrict *restrict argv, bool *restrict help_requested,
		    struct input_arguments *restrict arguments);

size_t next_power_2(size_t x);
size_t prev_power_2(size_t x);

/* Number of pairs i<j whose "same class" relation differs between the two vectors
 * (src/common.c:1195-1209). */
size_t get_difference(index_t const *classification, index_t const *last_classification, index_t m);

/* Seed: FIXED_SEED when built with NMFGPU_FIXED_INIT, else read from /dev/urandom with the
 * clock as fallback (src/common.c:1218-1258).  A decimal NMFGPU_SEED in the environment
 * takes precedence (an addition, for reproducible runs of any driver on this library). */
index_t get_seed(void);

#ifdef __cplusplus
}
#endif

#endif /* NMFGPU_COMMON_H */
