/*
 * Test stand-in for the part of Postgres's postgres.h that the gpuqo_run
 * shim uses: palloc/palloc0 (the stand-in's allocations form one arena, freed
 * by pg_standin_reset, as a memory-context reset frees a context), Max and
 * elog.  Only the test harness includes these headers; inside a Postgres
 * tree the real ones are used.
 */
#ifndef PG_STANDIN_POSTGRES_H
#define PG_STANDIN_POSTGRES_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef double Cost;

void *palloc(size_t size);
void *palloc0(size_t size);

#define Max(x, y) ((x) > (y) ? (x) : (y))

#define WARNING 19
#define ERROR 21
void pg_standin_elog(int level, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
#define elog(level, ...) pg_standin_elog(level, __VA_ARGS__)

#endif
