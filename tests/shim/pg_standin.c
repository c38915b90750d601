/*
 * Test harness: the few Postgres services the gpuqo_run shim calls
 * (palloc, Bitmapset members, elog, CHECK_FOR_INTERRUPTS) and the GUC globals
 * it reads, with the defaults of Postgres 12 and of the reference's guc.c
 * (bools :1102-1171, ints :2139-2298, enums :4809-4826).  Linked with the
 * shim into one library that the tests drive through ctypes.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include "postgres.h"
#include "miscadmin.h"
#include "nodes/bitmapset.h"
#include "optimizer/cost.h"
#include "optimizer/gpuqo_common.h"

/* GUC globals */
int			gpuqo_algorithm = 0;	/* dpsize */
int			gpuqo_scratchpad_size_mb = 10;
int			gpuqo_max_memo_size_mb = 1024;
int			gpuqo_min_memo_size_mb = 1;
int			gpuqo_n_parallel = 10240;
bool		gpuqo_dpsub_filter_enable = true;
int			gpuqo_dpsub_filter_threshold = 16;
int			gpuqo_dpsub_filter_cpu_enum_threshold = 1024;
int			gpuqo_dpsub_filter_keys_overprovisioning = 100;
bool		gpuqo_dpsub_ccc_enable = true;
bool		gpuqo_dpsub_csg_enable = true;
int			gpuqo_dpsub_csg_threshold = 128;
int			gpuqo_dpe_n_threads = 4;
int			gpuqo_cpu_dpsub_parallel_chunk_size = 64;
int			gpuqo_dpe_pairs_per_depbuf = 128;
bool		gpuqo_dpsub_tree_enable = false;
bool		gpuqo_dpsub_bicc_enable = false;
bool		gpuqo_spanning_tree_enable = false;
int			gpuqo_idp_n_iters = 0;
int			gpuqo_idp_type = 1;		/* idp2 */

double		seq_page_cost = 1.0;
double		random_page_cost = 4.0;
double		cpu_tuple_cost = 0.01;
double		cpu_index_tuple_cost = 0.005;
double		cpu_operator_cost = 0.0025;
int			effective_cache_size = 524288;
Cost		disable_cost = 1.0e10;
int			work_mem = 4096;
bool		enable_seqscan = true;
bool		enable_indexscan = true;
bool		enable_tidscan = true;
bool		enable_sort = true;
bool		enable_hashagg = true;
bool		enable_nestloop = true;
bool		enable_mergejoin = true;
bool		enable_hashjoin = true;

/* palloc: one arena of chunks, each headed by the previous chunk */
typedef struct Chunk
{
	struct Chunk *prev;
	size_t		size;
	long double data[];			/* maximally aligned, as palloc's result */
} Chunk;

static Chunk *arena;
static long n_chunks;

void *
palloc0(size_t size)
{
	Chunk	   *c = calloc(1, sizeof(Chunk) + size);

	if (c == NULL)
		abort();
	c->prev = arena;
	c->size = size;
	arena = c;
	n_chunks++;
	return c->data;
}

void *
palloc(size_t size)
{
	void	   *p = palloc0(size);

	memset(p, 0xA5, size);		/* unwritten bytes are garbage, as in palloc */
	return p;
}

/* Free every palloc'd chunk (a memory-context reset). */
void
pg_standin_reset(void)
{
	while (arena)
	{
		Chunk	   *p = arena->prev;

		free(arena);
		arena = p;
	}
	n_chunks = 0;
}

long
pg_standin_palloc_count(void)
{
	return n_chunks;
}

/* Bitmapset */
int
bms_next_member(const Bitmapset *a, int prevbit)
{
	if (a == NULL)
		return -2;
	for (int x = prevbit + 1; x < 64 * a->nwords; x++)
		if ((a->words[x / 64] >> (x % 64)) & 1)
			return x;
	return -2;
}

int
bms_num_members(const Bitmapset *a)
{
	int			k = 0;

	if (a == NULL)
		return 0;
	for (int w = 0; w < a->nwords; w++)
		k += __builtin_popcountll(a->words[w]);
	return k;
}

Bitmapset *
bms_add_member(Bitmapset *a, int x)
{
	int			need = x / 64 + 1;

	if (a == NULL || a->nwords < need)
	{
		Bitmapset  *b = palloc0(sizeof(Bitmapset) + sizeof(bitmapword) * need);

		b->nwords = need;
		if (a != NULL)
			memcpy(b->words, a->words, sizeof(bitmapword) * a->nwords);
		a = b;
	}
	a->words[x / 64] |= (bitmapword) 1 << (x % 64);
	return a;
}

/* A palloc'd Bitmapset holding the given words (NULL when all are zero,
 * as Postgres represents the empty set). */
Bitmapset *
pg_standin_bms_make(const uint64_t *words, int nwords)
{
	Bitmapset  *b = NULL;

	for (int x = 0; x < 64 * nwords; x++)
		if ((words[x / 64] >> (x % 64)) & 1)
			b = bms_add_member(b, x);
	return b;
}

/* elog and CHECK_FOR_INTERRUPTS */
static char last_message[1024];
static int	last_level;
static int	n_interrupt_checks;

void
pg_standin_elog(int level, const char *fmt,...)
{
	va_list		ap;

	va_start(ap, fmt);
	vsnprintf(last_message, sizeof(last_message), fmt, ap);
	va_end(ap);
	last_level = level;
}

const char *
pg_standin_last_message(int *level)
{
	if (level)
		*level = last_level;
	return last_message;
}

void
pg_standin_check_for_interrupts(void)
{
	n_interrupt_checks++;
}

int
pg_standin_interrupt_checks(void)
{
	return n_interrupt_checks;
}

void
pg_standin_clear_log(void)
{
	last_message[0] = 0;
	last_level = 0;
	n_interrupt_checks = 0;
}

/* Sizes and field offsets of the planner-info structs, in a fixed order,
 * so that a binding can check its own mirror of them. */
#include "optimizer/gpuqo_planner_info.h"

int
pg_standin_layout(long *out, int cap)
{
	const long	v[] = {
		sizeof(VarInfo), offsetof(VarInfo, index),
		offsetof(GpuqoIndexInfo, unique), offsetof(GpuqoIndexInfo, rows), offsetof(GpuqoIndexInfo, cost),
		sizeof(EqClassInfo), offsetof(EqClassInfo, relids), offsetof(EqClassInfo, sels),
		offsetof(EqClassInfo, vars), offsetof(EqClassInfo, fk), offsetof(EqClassInfo, next),
		sizeof(BaseRelationC), offsetof(BaseRelationC, rows), offsetof(BaseRelationC, cost),
		offsetof(BaseRelationC, width),
		sizeof(GpuqoPlannerInfoC), offsetof(GpuqoPlannerInfoC, base_rels),
		offsetof(GpuqoPlannerInfoC, edge_table), offsetof(GpuqoPlannerInfoC, eq_classes),
		offsetof(GpuqoPlannerInfoC, n_eq_classes), offsetof(GpuqoPlannerInfoC, n_eq_class_vars),
		sizeof(QueryTreeC), offsetof(QueryTreeC, rows), offsetof(QueryTreeC, cost),
		offsetof(QueryTreeC, width), offsetof(QueryTreeC, left), offsetof(QueryTreeC, right),
	};
	const int	n = (int) (sizeof(v) / sizeof(v[0]));

	for (int i = 0; i < n && i < cap; i++)
		out[i] = v[i];
	return n;
}
