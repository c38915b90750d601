/*
 * gq_host.h -- host-side planner model: query in the library's numbering
 * (relation i = bit i), the Remapper, plan trees, and the conversion into
 * the device blob.  Templates over the set width W (64-bit words).
 */
#ifndef GQ_HOST_H
#define GQ_HOST_H

#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <list>
#include <mutex>
#include <thread>
#include <vector>

#include "gq_cost_model.h"
#include "gq_dp.h"
#include "gq_internal.h"

namespace gq {

/* f(0) .. f(n-1) on up to 16 host threads (inline below 32 items); f must
 * be safe to run concurrently for different indices.  An exception thrown
 * by f (or by starting a thread) stops the remaining items and is rethrown
 * here after every started thread has been joined. */
template <class F>
void parallel_for(int n, F f) {
    const unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1u, 16u);
    nt = std::min(nt, n / 16);
    if (n < 32 || nt <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next(0);
    std::exception_ptr err;
    std::mutex err_mu;
    auto work = [&]() {
        try {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= n) break;
                f(i);
            }
        } catch (...) {
            next.store(n);
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    try {
        for (int t = 1; t < nt; t++) th.emplace_back(work);
    } catch (...) {
        next.store(n);
        for (auto &x : th) x.join();
        throw;
    }
    work();
    for (auto &x : th) x.join();
    if (err) std::rethrow_exception(err);
}

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
    std::vector<float> rel_rows, rel_fksel, mem_nd, mem_inv_nd, mem_skew, mem_mcv, mem_idx_rows, mem_idx_startup,
        mem_idx_total, qualtab;
    std::vector<PairRec> pair_rec;
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
