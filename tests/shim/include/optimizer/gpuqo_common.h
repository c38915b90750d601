/*
 * Test stand-in for the reference's src/include/optimizer/gpuqo_common.h:
 * the gpuqo GUC globals (:38-60), eqClassNSels (:62-67) and the statistics
 * types PathCost, GpuqoIndexInfo, VarStats, VarInfo (:81-113), with the
 * reference's field names, types and order (the shim reads them by name and
 * the Postgres side lays them out).
 */
#ifndef PG_STANDIN_GPUQO_COMMON_H
#define PG_STANDIN_GPUQO_COMMON_H

#include <stdbool.h>

extern int	gpuqo_algorithm;
extern int	gpuqo_scratchpad_size_mb;
extern int	gpuqo_max_memo_size_mb;
extern int	gpuqo_min_memo_size_mb;
extern int	gpuqo_n_parallel;
extern bool gpuqo_dpsub_filter_enable;
extern int	gpuqo_dpsub_filter_threshold;
extern int	gpuqo_dpsub_filter_cpu_enum_threshold;
extern int	gpuqo_dpsub_filter_keys_overprovisioning;
extern bool gpuqo_dpsub_ccc_enable;
extern bool gpuqo_dpsub_csg_enable;
extern int	gpuqo_dpsub_csg_threshold;
extern int	gpuqo_dpe_n_threads;
extern int	gpuqo_cpu_dpsub_parallel_chunk_size;
extern int	gpuqo_dpe_pairs_per_depbuf;
extern bool gpuqo_dpsub_tree_enable;
extern bool gpuqo_dpsub_bicc_enable;
extern bool gpuqo_spanning_tree_enable;
extern int	gpuqo_idp_n_iters;
extern int	gpuqo_idp_type;

/* selectivities of a class of `size` members: one per unordered pair */
static inline int
eqClassNSels(int size)
{
	return size * (size - 1) / 2;
}

typedef struct PathCost
{
	float		startup;
	float		total;
} PathCost;

typedef struct GpuqoIndexInfo
{
	bool		available;
	bool		unique;
	float		rows;
	PathCost	cost;
} GpuqoIndexInfo;

typedef struct VarStats
{
	float		stadistinct;
	float		stanullfrac;
	float		mcvfreq;
} VarStats;

typedef struct VarInfo
{
	VarStats	stats;
	GpuqoIndexInfo index;
} VarInfo;

#endif
