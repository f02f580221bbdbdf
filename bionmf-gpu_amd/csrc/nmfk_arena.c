/*
 * nmfk_arena.c -- device memory arena of the serving loop.
 *
 * The serving loop creates and retires thousands of jobs.  With one hipMalloc per buffer
 * and one hipFree per buffer at retirement, the host thread that keeps the device busy
 * stalls: hipFree waits for the device to go idle.  The arena reserves device memory in a
 * few large segments (up to a byte limit, the HBM admission budget) and sub-allocates job
 * buffers from them, first fit with coalescing on release; nothing is returned to the
 * runtime before nmfk_arena_destroy, except by an explicit nmfk_arena_trim.  Blocks are
 * 256-byte aligned.
 *
 * Not thread-safe: one arena per serving loop, used from its host thread.
 */
#include "nmfk_internal.h"

#include <stdlib.h>
#include <string.h>

#define ARENA_ALIGN 256

typedef struct arena_block {
	size_t off, size;
	int used;
} arena_block;

typedef struct arena_seg {
	char *base;
	size_t size;
	arena_block *blk;   /* sorted by offset, covering [0, size) */
	uint32_t nblk, cap;
} arena_seg;

struct nmfk_arena {
	arena_seg *seg;
	uint32_t nseg, segcap;
	size_t limit, reserved, in_use, seg_bytes;
};

static size_t align_up(size_t x)
{
	return (x + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
}

size_t nmfk_arena_round(size_t bytes)
{
	return align_up(bytes ? bytes : 1);
}

nmfk_arena *nmfk_arena_create(size_t limit, size_t seg_bytes)
{
	nmfk_arena *a = (nmfk_arena *)calloc(1, sizeof(nmfk_arena));
	if (!a) {
		nmfk_set_error("nmfk_arena_create: out of host memory");
		return NULL;
	}
	a->limit = limit;
	a->seg_bytes = align_up(seg_bytes ? seg_bytes : ((size_t)4 << 30));
	return a;
}

void nmfk_arena_destroy(nmfk_arena *a)
{
	if (!a)
		return;
	for (uint32_t i = 0; i < a->nseg; i++) {
		hipFree(a->seg[i].base);
		free(a->seg[i].blk);
	}
	free(a->seg);
	free(a);
}

size_t nmfk_arena_reserved(const nmfk_arena *a)
{
	return a->reserved;
}

size_t nmfk_arena_in_use(const nmfk_arena *a)
{
	return a->in_use;
}

static int seg_insert(arena_seg *s, uint32_t at, arena_block b)
{
	if (s->nblk == s->cap) {
		const uint32_t nc = s->cap ? 2 * s->cap : 16;
		arena_block *p = (arena_block *)realloc(s->blk, nc * sizeof(arena_block));
		if (!p)
			return NMFK_ENOMEM;
		s->blk = p;
		s->cap = nc;
	}
	memmove(s->blk + at + 1, s->blk + at, (s->nblk - at) * sizeof(arena_block));
	s->blk[at] = b;
	s->nblk++;
	return 0;
}

static void seg_remove(arena_seg *s, uint32_t at)
{
	memmove(s->blk + at, s->blk + at + 1, (s->nblk - at - 1) * sizeof(arena_block));
	s->nblk--;
}

/* First free block of segment s with room for <bytes>, split off; NULL if none. */
static void *seg_take(arena_seg *s, size_t bytes)
{
	for (uint32_t i = 0; i < s->nblk; i++) {
		arena_block *b = &s->blk[i];
		if (b->used || b->size < bytes)
			continue;
		if (b->size > bytes) {
			arena_block rest = {b->off + bytes, b->size - bytes, 0};
			if (seg_insert(s, i + 1, rest))
				return NULL;
			b = &s->blk[i];
			b->size = bytes;
		}
		b->used = 1;
		return s->base + b->off;
	}
	return NULL;
}

static int add_segment(nmfk_arena *a, size_t need)
{
	size_t size = need > a->seg_bytes ? need : a->seg_bytes;
	if (a->reserved + size > a->limit)
		size = a->limit - a->reserved;   /* the last segment takes what the limit leaves */
	if (size < need)
		return NMFK_EBUDGET;
	if (a->nseg == a->segcap) {
		const uint32_t nc = a->segcap ? 2 * a->segcap : 8;
		arena_seg *p = (arena_seg *)realloc(a->seg, nc * sizeof(arena_seg));
		if (!p)
			return NMFK_ENOMEM;
		a->seg = p;
		a->segcap = nc;
	}
	arena_seg *s = &a->seg[a->nseg];
	memset(s, 0, sizeof *s);
	if (hipMalloc((void **)&s->base, size) != hipSuccess) {
		(void)hipGetLastError();
		return NMFK_EBUDGET;
	}
	s->size = size;
	arena_block whole = {0, size, 0};
	if (seg_insert(s, 0, whole)) {
		hipFree(s->base);
		return NMFK_ENOMEM;
	}
	a->nseg++;
	a->reserved += size;
	return 0;
}

void *nmfk_arena_alloc(nmfk_arena *a, size_t bytes)
{
	bytes = nmfk_arena_round(bytes);
	for (uint32_t i = 0; i < a->nseg; i++) {
		void *p = seg_take(&a->seg[i], bytes);
		if (p) {
			a->in_use += bytes;
			return p;
		}
	}
	if (add_segment(a, bytes))
		return NULL;
	void *p = seg_take(&a->seg[a->nseg - 1], bytes);
	if (p)
		a->in_use += bytes;
	return p;
}

int nmfk_arena_free(nmfk_arena *a, void *ptr)
{
	if (!ptr)
		return 0;
	const char *p = (const char *)ptr;
	for (uint32_t i = 0; i < a->nseg; i++) {
		arena_seg *s = &a->seg[i];
		if (p < s->base || p >= s->base + s->size)
			continue;
		const size_t off = (size_t)(p - s->base);
		uint32_t lo = 0, hi = s->nblk;
		while (lo < hi) {   /* block whose offset is off */
			const uint32_t mid = (lo + hi) / 2;
			if (s->blk[mid].off < off)
				lo = mid + 1;
			else
				hi = mid;
		}
		if (lo == s->nblk || s->blk[lo].off != off || !s->blk[lo].used) {
			nmfk_set_error("nmfk_arena_free: %p is not a block of this arena", ptr);
			return NMFK_EINVAL;
		}
		s->blk[lo].used = 0;
		a->in_use -= s->blk[lo].size;
		if (lo + 1 < s->nblk && !s->blk[lo + 1].used) {
			s->blk[lo].size += s->blk[lo + 1].size;
			seg_remove(s, lo + 1);
		}
		if (lo > 0 && !s->blk[lo - 1].used) {
			s->blk[lo - 1].size += s->blk[lo].size;
			seg_remove(s, lo);
		}
		return 0;
	}
	nmfk_set_error("nmfk_arena_free: %p is outside the arena", ptr);
	return NMFK_EINVAL;
}

/* Returns segments with no block in use to the runtime (hipFree waits for the device). */
void nmfk_arena_trim(nmfk_arena *a)
{
	uint32_t w = 0;
	for (uint32_t i = 0; i < a->nseg; i++) {
		arena_seg *s = &a->seg[i];
		if (s->nblk == 1 && !s->blk[0].used) {
			hipFree(s->base);
			free(s->blk);
			a->reserved -= s->size;
			continue;
		}
		a->seg[w++] = *s;
	}
	a->nseg = w;
}
