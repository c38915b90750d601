/*
 * gq_host.h -- host-side planner model: query in the library's numbering
 * (relation i = bit i), the Remapper, plan trees, and the conversion into
 * the device blob.  Templates over the set width W (64-bit words).
 */
#ifndef GQ_HOST_H
#define GQ_HOST_H

#include <stdint.h>

#include <list>
#include <vector>

#include "gq_cost_model.h"
#include "gq_dp.h"
#include "gq_internal.h"

namespace gq {

template <int W>
using Set = SetN<W>;

template <int W>
inline bool subset_of(const Set<W> &a, const Set<W> &b) { return (a & b) == a; }   /* a ⊆ b */
template <int W>
inline bool intersects(const Set<W> &a, const Set<W> &b) { return !empty(a & b); }

/* GpuqoPlannerInfo<BitmapsetN> (gpuqo_planner_info.cuh:64-137) in vectors. */
template <int W>
struct HQuery {
    int n = 0;
    int n_iters = 0;
    Params p{};
    std::vector<float> rows, tuples, startup, total;
    std::vector<int> width;
    std::vector<uint8_t> composite;
    std::vector<Set<W>> adj;
    std::vector<Set<W>> ec_relids;
    std::vector<float> sels;          /* per class s(s-1)/2, concatenated */
    std::vector<Set<W>> fks;          /* per member                      */
    std::vector<gq_var_info> vars;    /* per member                      */
};

/* Arrays the cost model reads, derived from an HQuery. */
template <int W>
struct HostView {
    std::vector<float> rel_rows, rel_fksel, mem_nd, mem_avgfreq, mem_mcv, mem_idx_rows, mem_idx_startup,
        mem_idx_total, qualtab;
    std::vector<uint64_t> inc;
    std::vector<int> ec_moff, ec_soff, mem_rel;
    std::vector<uint8_t> ec_hasfk, mem_flags;
    View<Set<W>> v{};
    void build(const HQuery<W> &q);
};

/* Plan tree node (QueryTree, gpuqo_planner_info.cuh:55-62). */
template <int W>
struct QT {
    Set<W> id;
    float rows, startup, total;
    int width;
    QT *left = nullptr, *right = nullptr;
};

template <int W>
void free_tree(QT<W> *t) {
    if (!t) return;
    free_tree(t->left);
    free_tree(t->right);
    delete t;
}

/* One entry of a relabelling: relations `from` (a set in the outer
 * numbering) become relation `to`; `qt` carries a composite's plan
 * (remapper_transf_el_t, gpuqo_remapper.cuh:20-25). */
template <int W>
struct RemapEl {
    Set<W> from;
    int to;
    QT<W> *qt;
};

}  // namespace gq

#endif
