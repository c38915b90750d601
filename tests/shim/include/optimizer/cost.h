/* Test stand-in for the cost GUCs of Postgres 12's optimizer/cost.h. */
#ifndef PG_STANDIN_COST_H
#define PG_STANDIN_COST_H

#include <stdbool.h>

extern double seq_page_cost;
extern double random_page_cost;
extern double cpu_tuple_cost;
extern double cpu_index_tuple_cost;
extern double cpu_operator_cost;
extern int	effective_cache_size;
extern Cost disable_cost;
extern bool enable_seqscan;
extern bool enable_indexscan;
extern bool enable_tidscan;
extern bool enable_sort;
extern bool enable_hashagg;
extern bool enable_nestloop;
extern bool enable_mergejoin;
extern bool enable_hashjoin;

#endif
