/*
 * gpuqo_run_amd.c -- drop-in replacement of the reference's gpuqo_run for
 * the Postgres backend (src/backend/optimizer/gpuqo_amd/ in a Postgres tree,
 * linked against libgpuqo_amd.so).
 *
 * Replaces
 *     extern "C" QueryTreeC *gpuqo_run(int gpuqo_algorithm, GpuqoPlannerInfoC *info)
 * declared at src/include/optimizer/gpuqo.h:23 and defined at
 * src/backend/optimizer/gpuqo/gpuqo_main_internal.cu:177-185, with the same
 * signature, the same GUC globals (src/include/optimizer/gpuqo_common.h:38-60,
 * optimizer/cost.h) and the same palloc'd QueryTreeC result, so gpuqo()
 * (gpuqo_main.c:548-594) and queryTree2Plan (gpuqo_main.c:199-249) stay as
 * they are.  Every allocation is a palloc in the caller's memory context,
 * released with it, as the reference's output tree is.
 *
 * The three steps are exported on their own (gpuqo_amd_query_from_info,
 * gpuqo_amd_config_from_gucs, gpuqo_amd_tree_from_plan) so that they can be
 * checked one by one.
 */
#include "postgres.h"
#include "miscadmin.h"
#include "nodes/bitmapset.h"
#include "optimizer/cost.h"
#include "optimizer/gpuqo_common.h"
#include "optimizer/gpuqo_planner_info.h"

#include "gpuqo_amd.h"

/* Relation i of the reference is member i+1 of a Bitmapset
 * (gpuqo_bfs_indexing.cu:33); in gq_query it is bit i of the words. */
static void
bms_to_words(const Bitmapset *s, uint64_t *w, int nw)
{
	int			x = -1;

	memset(w, 0, sizeof(uint64_t) * nw);
	while ((x = bms_next_member(s, x)) >= 0)
		if (x >= 1 && x - 1 < 64 * nw)
			w[(x - 1) / 64] |= (uint64_t) 1 << ((x - 1) % 64);
}

static Bitmapset *
words_to_bms(const uint64_t *w)
{
	Bitmapset  *s = NULL;

	for (int i = 0; i < 64 * GQ_MAX_WORDS; i++)
		if ((w[i / 64] >> (i % 64)) & 1)
			s = bms_add_member(s, i + 1);
	return s;
}

/*
 * The planner info as a gq_query (convertGpuqoPlannerInfo,
 * gpuqo_planner_info.cu:133-209): base relations and edge table in relation
 * order, equivalence classes in list order, members ascending, sels in
 * eqClassIndex order, one fk set and one VarInfo per member.  The arrays are
 * palloc'd.  Returns 0, or -1 for a malformed info.
 */
int
gpuqo_amd_query_from_info(const GpuqoPlannerInfoC *info, gq_query *q)
{
	const int	n = info->n_rels;
	const int	nw = (n + 63) / 64;
	int			e = 0,
				m = 0,
				so = 0;
	gq_base_rel *base;
	uint64_t   *edges,
			   *fks;
	int32_t    *moff,
			   *mem;
	float	   *sels;
	gq_var_info *vars;

	if (n < 1 || n > GQ_MAX_RELS || info->n_eq_classes < 0)
		return -1;
	base = palloc(sizeof(gq_base_rel) * n);
	edges = palloc0(sizeof(uint64_t) * n * nw);
	for (int i = 0; i < n; i++)
	{
		base[i].rows = info->base_rels[i].rows;
		base[i].tuples = info->base_rels[i].tuples;
		base[i].cost.startup = info->base_rels[i].cost.startup;
		base[i].cost.total = info->base_rels[i].cost.total;
		base[i].width = info->base_rels[i].width;
		bms_to_words(info->edge_table[i], edges + (size_t) i * nw, nw);
	}
	moff = palloc(sizeof(int32_t) * (info->n_eq_classes + 1));
	mem = palloc(sizeof(int32_t) * Max(1, info->n_eq_class_vars));
	sels = palloc(sizeof(float) * Max(1, info->n_eq_class_sels));
	fks = palloc0(sizeof(uint64_t) * nw * Max(1, info->n_eq_class_fks));
	vars = palloc0(sizeof(gq_var_info) * Max(1, info->n_eq_class_vars));
	for (const EqClassInfo *ec = info->eq_classes; ec != NULL; ec = ec->next, e++)
	{
		int			s = bms_num_members(ec->relids),
					x = -1,
					j = 0;

		if (e >= info->n_eq_classes || m + s > info->n_eq_class_vars ||
			so + s * (s - 1) / 2 > info->n_eq_class_sels)
			return -1;
		moff[e] = m;
		while ((x = bms_next_member(ec->relids, x)) >= 0)
		{
			const VarInfo *v = &ec->vars[j];

			mem[m + j] = x - 1;
			bms_to_words(ec->fk[j], fks + (size_t) (m + j) * nw, nw);
			vars[m + j].stadistinct = v->stats.stadistinct;
			vars[m + j].stanullfrac = v->stats.stanullfrac;
			vars[m + j].mcvfreq = v->stats.mcvfreq;
			vars[m + j].index_available = v->index.available ? 1 : 0;
			vars[m + j].index_unique = v->index.unique ? 1 : 0;
			vars[m + j].index_rows = v->index.rows;
			vars[m + j].index_cost.startup = v->index.cost.startup;
			vars[m + j].index_cost.total = v->index.cost.total;
			j++;
		}
		for (int k = 0; k < eqClassNSels(s); k++)
			sels[so + k] = ec->sels[k];
		so += eqClassNSels(s);
		m += s;
	}
	if (e != info->n_eq_classes)
		return -1;
	moff[e] = m;
	q->n_rels = n;
	q->n_words = nw;
	q->base_rels = base;
	q->edge_table = edges;
	q->n_eq_classes = info->n_eq_classes;
	q->ec_member_offsets = moff;
	q->ec_members = mem;
	q->ec_sels = sels;
	q->ec_fks = fks;
	q->ec_vars = vars;
	return 0;
}

/* The GUC globals (gpuqo_common.h:38-60, cost.h) as a gq_config. */
void
gpuqo_amd_config_from_gucs(int algorithm, gq_config *cfg)
{
	gq_config_default(cfg);
	cfg->gpuqo_algorithm = algorithm;
	cfg->gpuqo_dpsub_bicc = gpuqo_dpsub_bicc_enable;
	cfg->gpuqo_dpsub_filter = gpuqo_dpsub_filter_enable;
	cfg->gpuqo_dpsub_filter_threshold = gpuqo_dpsub_filter_threshold;
	cfg->gpuqo_dpsub_filter_cpu_enum_threshold = gpuqo_dpsub_filter_cpu_enum_threshold;
	cfg->gpuqo_dpsub_filter_keys_overprovisioning = gpuqo_dpsub_filter_keys_overprovisioning;
	cfg->gpuqo_dpsub_ccc = gpuqo_dpsub_ccc_enable;
	cfg->gpuqo_dpsub_csg = gpuqo_dpsub_csg_enable;
	cfg->gpuqo_dpsub_csg_threshold = gpuqo_dpsub_csg_threshold;
	cfg->gpuqo_dpsub_tree = gpuqo_dpsub_tree_enable;
	cfg->gpuqo_spanning_tree = gpuqo_spanning_tree_enable;
	cfg->gpuqo_n_parallel = gpuqo_n_parallel;
	cfg->gpuqo_scratchpad_size_mb = gpuqo_scratchpad_size_mb;
	cfg->gpuqo_min_memo_size_mb = gpuqo_min_memo_size_mb;
	cfg->gpuqo_max_memo_size_mb = gpuqo_max_memo_size_mb;
	cfg->gpuqo_idp_type = gpuqo_idp_type;
	cfg->gpuqo_idp_n_iters = gpuqo_idp_n_iters;
	cfg->gpuqo_dpe_n_threads = gpuqo_dpe_n_threads;
	cfg->gpuqo_cpu_dpsub_parallel_chunk_size = gpuqo_cpu_dpsub_parallel_chunk_size;
	cfg->gpuqo_dpe_pairs_per_depbuf = gpuqo_dpe_pairs_per_depbuf;
	cfg->seq_page_cost = seq_page_cost;
	cfg->random_page_cost = random_page_cost;
	cfg->cpu_tuple_cost = cpu_tuple_cost;
	cfg->cpu_index_tuple_cost = cpu_index_tuple_cost;
	cfg->cpu_operator_cost = cpu_operator_cost;
	cfg->disable_cost = disable_cost;
	cfg->effective_cache_size = (int32_t) effective_cache_size;
	cfg->work_mem = work_mem;
	cfg->enable_seqscan = enable_seqscan;
	cfg->enable_indexscan = enable_indexscan;
	cfg->enable_tidscan = enable_tidscan;
	cfg->enable_sort = enable_sort;
	cfg->enable_hashagg = enable_hashagg;
	cfg->enable_nestloop = enable_nestloop;
	cfg->enable_mergejoin = enable_mergejoin;
	cfg->enable_hashjoin = enable_hashjoin;
	/* the reference picks its cost function at build time (Makefile
	 * cost_function=postgres|simple|cout -> -DGPUQO_COST_FUNCTION=0|1|2);
	 * the same define selects it here */
#ifdef GPUQO_COST_FUNCTION
	cfg->cost_function = GPUQO_COST_FUNCTION;
#endif
}

static QueryTreeC *
node_to_tree(const gq_plan_node *nodes, int n_nodes, int i)
{
	QueryTreeC *qt;

	if (i < 0 || i >= n_nodes)
		return NULL;
	qt = palloc(sizeof(QueryTreeC));
	qt->id = words_to_bms(nodes[i].relids);
	qt->rows = nodes[i].rows;
	qt->cost.startup = nodes[i].cost.startup;
	qt->cost.total = nodes[i].cost.total;
	qt->width = nodes[i].width;
	qt->left = node_to_tree(nodes, n_nodes, nodes[i].left);
	qt->right = node_to_tree(nodes, n_nodes, nodes[i].right);
	return qt;
}

/* The plan (pre-order nodes) as a palloc'd QueryTreeC
 * (convertQueryTree, gpuqo_planner_info.cu:256-271): leaves have NULL
 * children. */
QueryTreeC *
gpuqo_amd_tree_from_plan(const gq_plan_out *out)
{
	return out->n_nodes > 0 ? node_to_tree(out->nodes, out->n_nodes, 0) : NULL;
}

/*
 * gpuqo_run: plan `info` with the algorithm and the GUCs in force.  Returns
 * NULL where the reference does (an algorithm it cannot run, no join tree),
 * after a WARNING with the library's message; an interrupt raised through
 * gq_interrupt() is re-raised with CHECK_FOR_INTERRUPTS().
 */
QueryTreeC *
gpuqo_run(int gpuqo_algorithm_arg, GpuqoPlannerInfoC *info)
{
	gq_config	cfg;
	gq_query	q;
	gq_plan_out out;
	gq_status	st;

	if (gpuqo_amd_query_from_info(info, &q) != 0)
	{
		elog(WARNING, "gpuqo_amd: malformed planner info");
		return NULL;
	}
	gpuqo_amd_config_from_gucs(gpuqo_algorithm_arg, &cfg);
	out.capacity = 2 * info->n_rels;
	out.n_nodes = 0;
	out.nodes = palloc(sizeof(gq_plan_node) * out.capacity);
	st = gq_plan(&q, &cfg, &out);
	if (st == GQ_ERR_INTERRUPTED)
		CHECK_FOR_INTERRUPTS();
	if (st != GQ_OK)
	{
		elog(WARNING, "gpuqo_amd: %s", gq_last_error());
		return NULL;
	}
	return gpuqo_amd_tree_from_plan(&out);
}
