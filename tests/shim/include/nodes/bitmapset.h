/* Test stand-in for Postgres's nodes/bitmapset.h (64-bit words, member x in
 * bit x % 64 of word x / 64; the empty set is NULL). */
#ifndef PG_STANDIN_BITMAPSET_H
#define PG_STANDIN_BITMAPSET_H

#include <stdint.h>

typedef uint64_t bitmapword;

typedef struct Bitmapset
{
	int			nwords;
	bitmapword	words[];
} Bitmapset;

int			bms_next_member(const Bitmapset *a, int prevbit);
int			bms_num_members(const Bitmapset *a);
Bitmapset  *bms_add_member(Bitmapset *a, int x);

#endif
