/*
 * Test stand-in for the reference's
 * src/include/optimizer/gpuqo_planner_info.h:15-57: the planner info handed
 * to gpuqo_run and the QueryTreeC it returns, with the reference's field
 * names, types and order.  Relation i is member i+1 of each Bitmapset.
 */
#ifndef PG_STANDIN_GPUQO_PLANNER_INFO_H
#define PG_STANDIN_GPUQO_PLANNER_INFO_H

#include "nodes/bitmapset.h"

typedef struct EquivalenceClass EquivalenceClass;

typedef Bitmapset *EdgeMask;
typedef Bitmapset *RelationID;

typedef struct EqClassInfo
{
	EquivalenceClass *eclass;
	RelationID	relids;
	float	   *sels;			/* eqClassNSels(members), eqClassIndex order */
	VarInfo    *vars;			/* one per member, ascending */
	RelationID *fk;				/* one per member, ascending */
	struct EqClassInfo *next;
} EqClassInfo;

typedef struct BaseRelationC
{
	RelationID	id;
	float		rows;
	float		tuples;
	PathCost	cost;
	int			width;
} BaseRelationC;

typedef struct GpuqoPlannerInfoC
{
	int			n_rels;
	BaseRelationC *base_rels;
	EdgeMask   *edge_table;
	EqClassInfo *eq_classes;
	int			n_eq_classes;
	int			n_eq_class_sels;
	int			n_eq_class_fks;
	int			n_eq_class_vars;
} GpuqoPlannerInfoC;

typedef struct QueryTreeC
{
	RelationID	id;
	float		rows;
	PathCost	cost;
	int			width;
	struct QueryTreeC *left;
	struct QueryTreeC *right;
} QueryTreeC;

#endif
