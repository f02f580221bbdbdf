/*
 * nmfk_io.h -- matrix files of bioNMF-GPU (text and binary), for the command-line driver.
 *
 * Formats (doc/user_guide.txt.md:186-293; reference implementation
 * src/matrix_io/matrix_io.c:697-4433):
 *   text        values separated by single tabs (or single spaces when the file has no
 *               tab); optional description string and column headers on the first line,
 *               optional row labels first on every line; an empty field is 0.0 / "".
 *   binary      "non-native": uint32 signature 0xEDB10EA1, uint32 rows, uint32 columns,
 *               rows x columns little-endian doubles (row-major), then optional tags as
 *               text: row labels (tab-separated) '\n' column headers '\n' name.
 *   native      uint32 rows, uint32 columns, raw floats, then the same optional tags.
 * Every value must be >= 0 and no row or column may be all zeros (checked on load).
 */
#ifndef NMFK_IO_H
#define NMFK_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NMFK_FMT_TEXT 0
#define NMFK_FMT_BINARY 1
#define NMFK_FMT_NATIVE 2

/* Optional tag elements of a matrix; NULL when absent.  Strings are owned by the struct. */
typedef struct nmfk_tags {
	char *name;
	char **headers;      /* one per column */
	char **labels;       /* one per row */
	uint32_t nheaders, nlabels;
} nmfk_tags;

void nmfk_tags_free(nmfk_tags *t);

/* Loads a matrix (rows x cols, row-major, leading dimension cols) into *A (malloc'ed).
 * numeric_hdrs / numeric_lbls force the first line / first column of a text file to be
 * read as headers / labels even when numeric (the -c / -r options). */
int nmfk_matrix_load(const char *path, int fmt, int numeric_hdrs, int numeric_lbls, float **A, uint32_t *rows,
		     uint32_t *cols, nmfk_tags *tags);

/* Saves A (rows x cols, leading dimension lda) or, with transpose, A^T (cols x rows); tags
 * describe the matrix as written.  Native files ignore transpose (no transformation, as
 * in the reference).  Text values are written with "%g", as the reference does. */
int nmfk_matrix_save(const char *path, int fmt, const float *A, uint32_t rows, uint32_t cols, size_t lda,
		     int transpose, const nmfk_tags *tags);

#ifdef __cplusplus
}
#endif

#endif /* NMFK_IO_H */
