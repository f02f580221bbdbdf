/*
 * nmfk_matrix_io.c -- the matrix file formats of bioNMF-GPU (doc/user_guide.txt.md:186-293),
 * read and written by the command-line driver.  See include/nmfk_io.h for the formats.
 *
 * Text files are read whole, split into lines (LF or CR+LF), and each line into fields at
 * single delimiters (tab, or space when the file has no tab).  The first line holds column
 * headers when any of its fields is not a number (or -c is given); the first field of the
 * first data line is a row label when it is not a number (or -r is given).  A header line
 * with one field more than the data columns starts with the description string.
 */
#include "nmfk_io.h"
#include "nmfk_internal.h"

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SIG_REAL 0xEDB10EA1u

void nmfk_tags_free(nmfk_tags *t)
{
	if (!t)
		return;
	free(t->name);
	for (uint32_t i = 0; t->headers && i < t->nheaders; i++)
		free(t->headers[i]);
	for (uint32_t i = 0; t->labels && i < t->nlabels; i++)
		free(t->labels[i]);
	free(t->headers);
	free(t->labels);
	memset(t, 0, sizeof *t);
}

static char *dup_range(const char *s, size_t n)
{
	char *d = (char *)malloc(n + 1);
	if (d) {
		memcpy(d, s, n);
		d[n] = '\0';
	}
	return d;
}

/* Whole file into memory (NUL-terminated). */
static char *slurp(const char *path, size_t *len)
{
	FILE *f = fopen(path, "rb");
	if (!f) {
		nmfk_set_error("cannot open %s: %s", path, strerror(errno));
		return NULL;
	}
	size_t cap = 1 << 16, n = 0;
	char *buf = (char *)malloc(cap + 1);
	while (buf) {
		const size_t r = fread(buf + n, 1, cap - n, f);
		n += r;
		if (n < cap)
			break;
		cap *= 2;
		char *nb = (char *)realloc(buf, cap + 1);
		if (!nb) {
			free(buf);
			buf = NULL;
		}
		buf = nb;
	}
	fclose(f);
	if (!buf) {
		nmfk_set_error("out of host memory reading %s", path);
		return NULL;
	}
	buf[n] = '\0';
	*len = n;
	return buf;
}

/* ---------------------------------------------------------------- text */

typedef struct line_t {
	char *s;       /* start; the line is NUL-terminated in place (CR and LF removed) */
	uint32_t nf;   /* fields */
} line_t;

static uint32_t count_fields(const char *s, char delim)
{
	uint32_t n = 1;
	for (; *s; s++)
		n += *s == delim;
	return n;
}

/* i-th field of a line (pointer and length). */
static const char *field(const char *s, char delim, uint32_t i, size_t *len)
{
	for (uint32_t k = 0; k < i; k++) {
		s = strchr(s, delim);
		if (!s)
			return NULL;
		s++;
	}
	const char *e = strchr(s, delim);
	*len = e ? (size_t)(e - s) : strlen(s);
	return s;
}

/* A field is a number when strtod consumes all of it (an empty field counts as 0). */
static int parse_num(const char *s, size_t len, double *v)
{
	if (len == 0) {
		*v = 0.0;
		return 1;
	}
	char tmp[128];
	if (len >= sizeof tmp)
		return 0;
	memcpy(tmp, s, len);
	tmp[len] = '\0';
	char *end = NULL;
	errno = 0;
	const double d = strtod(tmp, &end);
	if (end == tmp || *end != '\0' || errno == ERANGE || !isfinite(d))
		return 0;
	*v = d;
	return 1;
}

static int line_is_numeric(const char *s, char delim, uint32_t first)
{
	const uint32_t nf = count_fields(s, delim);
	for (uint32_t i = first; i < nf; i++) {
		size_t len;
		const char *f = field(s, delim, i, &len);
		double v;
		if (!f || !parse_num(f, len, &v))
			return 0;
	}
	return 1;
}

static char **fields_dup(const char *s, char delim, uint32_t first, uint32_t count)
{
	char **out = (char **)calloc(count ? count : 1, sizeof(char *));
	if (!out)
		return NULL;
	for (uint32_t i = 0; i < count; i++) {
		size_t len = 0;
		const char *f = field(s, delim, first + i, &len);
		out[i] = f ? dup_range(f, len) : dup_range("", 0);
		if (!out[i]) {
			for (uint32_t j = 0; j < i; j++)
				free(out[j]);
			free(out);
			return NULL;
		}
	}
	return out;
}

static int check_values(const float *A, uint32_t rows, uint32_t cols, const char *path)
{
	if (rows < 2 || cols < 2) {
		nmfk_set_error("%s: both matrix dimensions must be >= 2 (got %u x %u)", path, rows, cols);
		return NMFK_EINVAL;
	}
	uint8_t *colpos = (uint8_t *)calloc(cols, 1);
	if (!colpos) {
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	int e = 0;
	for (uint32_t i = 0; i < rows && !e; i++) {
		int rowpos = 0;
		for (uint32_t j = 0; j < cols; j++) {
			const float v = A[(size_t)i * cols + j];
			if (!(v >= 0.0f)) {
				nmfk_set_error("%s: invalid value %g at row %u, column %u (data must be non-negative)", path,
					       (double)v, i + 1, j + 1);
				e = NMFK_EINVAL;
				break;
			}
			if (v > 0.0f)
				rowpos = 1, colpos[j] = 1;
		}
		if (!e && !rowpos) {
			nmfk_set_error("%s: row %u has no positive value", path, i + 1);
			e = NMFK_EINVAL;
		}
	}
	for (uint32_t j = 0; j < cols && !e; j++)
		if (!colpos[j]) {
			nmfk_set_error("%s: column %u has no positive value", path, j + 1);
			e = NMFK_EINVAL;
		}
	free(colpos);
	return e;
}

static int load_text(const char *path, int numeric_hdrs, int numeric_lbls, float **A, uint32_t *rows, uint32_t *cols,
		     nmfk_tags *tags)
{
	size_t len = 0;
	char *buf = slurp(path, &len);
	if (!buf)
		return NMFK_EINVAL;
	const char delim = memchr(buf, '\t', len) ? '\t' : ' ';
	/* split into lines */
	size_t nl = 1;
	for (size_t i = 0; i < len; i++)
		nl += buf[i] == '\n';
	line_t *L = (line_t *)malloc(nl * sizeof(line_t));
	if (!L) {
		free(buf);
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	size_t n = 0;
	for (char *p = buf; p < buf + len || (p == buf && len == 0);) {
		char *e = strchr(p, '\n');
		char *end = e ? e : buf + len;
		if (end > p && end[-1] == '\r')
			end[-1] = '\0';
		*end = '\0';
		L[n].s = p;
		n++;
		if (!e)
			break;
		p = e + 1;
	}
	while (n > 0 && L[n - 1].s[0] == '\0')   /* trailing blank lines */
		n--;
	int e = 0;
	if (n == 0) {
		nmfk_set_error("%s: empty file", path);
		e = NMFK_EINVAL;
	}
	uint32_t first = 0, lblcol = 0, ncol = 0;
	int has_hdr = 0;
	if (!e) {
		/* Headers when a field after the first is not a number; a first line whose only
		 * non-numeric field is the first is a labelled data row, unless its field count
		 * differs from the next line's (a description string alone). */
		size_t flen;
		const char *f0 = field(L[0].s, delim, 0, &flen);
		double v;
		const int first_num = parse_num(f0, flen, &v);
		has_hdr = numeric_hdrs || !line_is_numeric(L[0].s, delim, 1) ||
			  (!first_num && (n == 1 || count_fields(L[0].s, delim) != count_fields(L[1].s, delim)));
		first = has_hdr ? 1 : 0;
		if (first >= n) {
			nmfk_set_error("%s: no matrix data", path);
			e = NMFK_EINVAL;
		}
	}
	if (!e) {
		size_t flen;
		const char *f0 = field(L[first].s, delim, 0, &flen);
		double v;
		lblcol = (numeric_lbls || !parse_num(f0, flen, &v)) ? 1 : 0;
		const uint32_t nf = count_fields(L[first].s, delim);
		if (nf <= lblcol) {
			nmfk_set_error("%s: no matrix data at line %u", path, first + 1);
			e = NMFK_EINVAL;
		}
		ncol = nf - lblcol;
	}
	const uint32_t nrow = (uint32_t)(n - first);
	float *M = NULL;
	if (!e) {
		M = (float *)malloc((size_t)nrow * ncol * sizeof(float));
		if (!M) {
			nmfk_set_error("out of host memory (%u x %u matrix)", nrow, ncol);
			e = NMFK_ENOMEM;
		}
	}
	if (!e && lblcol) {
		tags->labels = (char **)calloc(nrow, sizeof(char *));
		tags->nlabels = nrow;
		if (!tags->labels)
			e = NMFK_ENOMEM;
	}
	for (uint32_t i = 0; i < nrow && !e; i++) {
		const char *s = L[first + i].s;
		const uint32_t nf = count_fields(s, delim);
		if (nf != ncol + lblcol) {
			nmfk_set_error("%s: line %u has %u fields, expected %u", path, first + i + 1, nf, ncol + lblcol);
			e = NMFK_EINVAL;
			break;
		}
		const char *p = s;
		for (uint32_t j = 0; j < nf; j++) {
			const char *q = strchr(p, delim);
			const size_t flen = q ? (size_t)(q - p) : strlen(p);
			if (j < lblcol) {
				tags->labels[i] = dup_range(p, flen);
				if (!tags->labels[i]) {
					e = NMFK_ENOMEM;
					break;
				}
			} else {
				double v;
				if (!parse_num(p, flen, &v)) {
					nmfk_set_error("%s: invalid number at line %u, field %u", path, first + i + 1, j + 1);
					e = NMFK_EINVAL;
					break;
				}
				M[(size_t)i * ncol + (j - lblcol)] = (float)v;
			}
			p = q ? q + 1 : p + flen;
		}
	}
	if (!e && has_hdr) {
		const char *s = L[0].s;
		const uint32_t nf = count_fields(s, delim);
		uint32_t h0 = 0;
		if (nf == ncol + 1) {
			size_t flen;
			const char *f0 = field(s, delim, 0, &flen);
			tags->name = dup_range(f0, flen);
			h0 = 1;
		} else if (nf == 1 && ncol > 1) {
			tags->name = dup_range(s, strlen(s));   /* description string only */
			h0 = 1;
		} else if (nf != ncol) {
			nmfk_set_error("%s: header line has %u fields for %u data columns", path, nf, ncol);
			e = NMFK_EINVAL;
		}
		if (!e && nf > h0) {
			tags->headers = fields_dup(s, delim, h0, ncol);
			tags->nheaders = ncol;
			if (!tags->headers)
				e = NMFK_ENOMEM;
		}
	}
	if (!e)
		e = check_values(M, nrow, ncol, path);
	free(L);
	free(buf);
	if (e) {
		free(M);
		nmfk_tags_free(tags);
		return e;
	}
	*A = M;
	*rows = nrow;
	*cols = ncol;
	return 0;
}

/* ---------------------------------------------------------------- binary */

static uint32_t bswap32(uint32_t x)
{
	return (x >> 24) | ((x >> 8) & 0xFF00u) | ((x << 8) & 0xFF0000u) | (x << 24);
}

static uint64_t bswap64(uint64_t x)
{
	return ((uint64_t)bswap32((uint32_t)x) << 32) | bswap32((uint32_t)(x >> 32));
}

/* Tags after the binary data: labels line, headers line, name (each optional). */
static int parse_binary_tags(const char *p, const char *end, uint32_t rows, uint32_t cols, nmfk_tags *t)
{
	if (p >= end)
		return 0;
	const char *nl = memchr(p, '\n', (size_t)(end - p));
	const char *le = nl ? nl : end;
	size_t llen = (size_t)(le - p);
	if (llen && p[llen - 1] == '\r')
		llen--;
	char *line = dup_range(p, llen);
	if (!line)
		return NMFK_ENOMEM;
	if (llen) {
		t->labels = fields_dup(line, '\t', 0, rows);
		t->nlabels = rows;
	}
	free(line);
	if (!nl)
		return 0;
	p = nl + 1;
	nl = memchr(p, '\n', (size_t)(end - p));
	le = nl ? nl : end;
	llen = (size_t)(le - p);
	if (llen && p[llen - 1] == '\r')
		llen--;
	line = dup_range(p, llen);
	if (!line)
		return NMFK_ENOMEM;
	if (llen) {
		t->headers = fields_dup(line, '\t', 0, cols);
		t->nheaders = cols;
	}
	free(line);
	if (!nl)
		return 0;
	p = nl + 1;
	size_t nlen = (size_t)(end - p);
	while (nlen && (p[nlen - 1] == '\n' || p[nlen - 1] == '\r'))
		nlen--;
	if (nlen)
		t->name = dup_range(p, nlen);
	return 0;
}

static int load_binary(const char *path, int native, float **A, uint32_t *rows, uint32_t *cols, nmfk_tags *tags)
{
	size_t len = 0;
	char *buf = slurp(path, &len);
	if (!buf)
		return NMFK_EINVAL;
	const char *p = buf, *end = buf + len;
	int swap = 0, e = 0;
	uint32_t hdr[3];
	const size_t nh = native ? 2 : 3;
	if (len < nh * 4) {
		nmfk_set_error("%s: file too short", path);
		free(buf);
		return NMFK_EINVAL;
	}
	memcpy(hdr, p, nh * 4);
	p += nh * 4;
	if (!native) {
		if (hdr[0] == bswap32(SIG_REAL))
			swap = 1;
		else if (hdr[0] != SIG_REAL) {
			nmfk_set_error("%s: invalid signature %08X (not a bioNMF binary real-data file)", path, hdr[0]);
			free(buf);
			return NMFK_EINVAL;
		}
	}
	uint32_t r = native ? hdr[0] : hdr[1], c = native ? hdr[1] : hdr[2];
	if (swap)
		r = bswap32(r), c = bswap32(c);
	const size_t esz = native ? sizeof(float) : sizeof(double);
	const size_t need = (size_t)r * c * esz;
	if (r == 0 || c == 0 || need / esz / c != r || (size_t)(end - p) < need) {
		nmfk_set_error("%s: dimensions %u x %u do not match the file size", path, r, c);
		free(buf);
		return NMFK_EINVAL;
	}
	float *M = (float *)malloc((size_t)r * c * sizeof(float));
	if (!M) {
		free(buf);
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	for (size_t i = 0; i < (size_t)r * c; i++) {
		if (native) {
			memcpy(&M[i], p + 4 * i, 4);
		} else {
			uint64_t u;
			memcpy(&u, p + 8 * i, 8);
			if (swap)
				u = bswap64(u);
			double d;
			memcpy(&d, &u, 8);
			M[i] = (float)d;
		}
	}
	p += need;
	e = parse_binary_tags(p, end, r, c, tags);
	if (!e && !native)
		e = check_values(M, r, c, path);
	free(buf);
	if (e) {
		free(M);
		nmfk_tags_free(tags);
		return e;
	}
	*A = M;
	*rows = r;
	*cols = c;
	return 0;
}

int nmfk_matrix_load(const char *path, int fmt, int numeric_hdrs, int numeric_lbls, float **A, uint32_t *rows,
		     uint32_t *cols, nmfk_tags *tags)
{
	memset(tags, 0, sizeof *tags);
	*A = NULL;
	switch (fmt) {
	case NMFK_FMT_TEXT:
		return load_text(path, numeric_hdrs, numeric_lbls, A, rows, cols, tags);
	case NMFK_FMT_BINARY:
		return load_binary(path, 0, A, rows, cols, tags);
	case NMFK_FMT_NATIVE:
		return load_binary(path, 1, A, rows, cols, tags);
	default:
		nmfk_set_error("unknown file format %d", fmt);
		return NMFK_EINVAL;
	}
}

/* ---------------------------------------------------------------- save */

static float elem(const float *A, size_t lda, int transpose, uint32_t i, uint32_t j)
{
	return transpose ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j];
}

static int write_joined(FILE *f, char **s, uint32_t n, char delim)
{
	for (uint32_t i = 0; i < n; i++)
		if ((i && fputc(delim, f) == EOF) || fputs(s[i] ? s[i] : "", f) == EOF)
			return -1;
	return 0;
}

int nmfk_matrix_save(const char *path, int fmt, const float *A, uint32_t rows, uint32_t cols, size_t lda,
		     int transpose, const nmfk_tags *tags)
{
	/* Native files skip every transformation (transpose included), as the reference's do. */
	if (fmt == NMFK_FMT_NATIVE)
		transpose = 0;
	/* output shape */
	const uint32_t R = transpose ? cols : rows, Cn = transpose ? rows : cols;
	const nmfk_tags none = {0};
	const nmfk_tags *t = tags ? tags : &none;
	if ((t->headers && t->nheaders != Cn) || (t->labels && t->nlabels != R)) {
		nmfk_set_error("nmfk_matrix_save: tags do not match the %u x %u output", R, Cn);
		return NMFK_EINVAL;
	}
	FILE *f = fopen(path, fmt == NMFK_FMT_TEXT ? "w" : "wb");
	if (!f) {
		nmfk_set_error("cannot create %s: %s", path, strerror(errno));
		return NMFK_EINVAL;
	}
	int bad = 0;
	if (fmt == NMFK_FMT_TEXT) {
		if (t->name) {
			bad |= fputs(t->name, f) == EOF;
			if (!t->headers)
				bad |= fputc('\n', f) == EOF;
		}
		if (t->headers) {
			if (t->name)
				bad |= fputc('\t', f) == EOF;
			bad |= write_joined(f, t->headers, Cn, '\t') != 0;
			bad |= fputc('\n', f) == EOF;
		}
		for (uint32_t i = 0; i < R && !bad; i++) {
			if (t->labels)
				bad |= fprintf(f, "%s\t", t->labels[i] ? t->labels[i] : "") < 0;
			for (uint32_t j = 0; j < Cn; j++)
				bad |= fprintf(f, j ? "\t%g" : "%g", (double)elem(A, lda, transpose, i, j)) < 0;
			bad |= fputc('\n', f) == EOF;
		}
	} else {
		if (fmt == NMFK_FMT_BINARY) {
			const uint32_t sig = SIG_REAL;
			bad |= fwrite(&sig, 4, 1, f) != 1;
		}
		bad |= fwrite(&R, 4, 1, f) != 1;
		bad |= fwrite(&Cn, 4, 1, f) != 1;
		for (uint32_t i = 0; i < R && !bad; i++)
			for (uint32_t j = 0; j < Cn && !bad; j++) {
				const float v = elem(A, lda, transpose, i, j);
				if (fmt == NMFK_FMT_BINARY) {
					const double d = (double)v;
					bad |= fwrite(&d, 8, 1, f) != 1;
				} else {
					bad |= fwrite(&v, 4, 1, f) != 1;
				}
			}
		if (t->labels || t->headers || t->name) {
			if (t->labels)
				bad |= write_joined(f, t->labels, R, '\t') != 0;
			if (t->headers || t->name) {
				bad |= fputc('\n', f) == EOF;
				if (t->headers)
					bad |= write_joined(f, t->headers, Cn, '\t') != 0;
				if (t->name) {
					bad |= fputc('\n', f) == EOF;
					bad |= fputs(t->name, f) == EOF;
					bad |= fputc('\n', f) == EOF;
				}
			}
		}
	}
	bad |= fclose(f) != 0;
	if (bad) {
		nmfk_set_error("error writing %s", path);
		return NMFK_EINVAL;
	}
	return 0;
}
