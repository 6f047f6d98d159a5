// gt_centrality.hpp -- C++ host-side mirror of gt-algo's centrality surface,
// header-only, over the C ABI of gt_hip.h.
//
// Names, argument meaning and failure behaviour follow the reference:
//   gt::Plan                 crates/gt-core/src/par/plan.rs:55-101 (Plan::new, range)
//   gt::PowerIteration       crates/gt-algo/src/centrality.rs:79-99
//   gt::Unity / Constant     crates/gt-core/src/prop/dense.rs:373-460
//   gt::DenseProp            crates/gt-core/src/prop/dense.rs:88-127 (sized_for grows, never shrinks)
//   gt::AdjList              crates/gt-core/src/adj/list.rs:59-97, 258-280 (with_vertices, add_edge)
//   gt::Reversed             crates/gt-core/src/view/reversed.rs:109-140
//   gt::pagerank             crates/gt-algo/src/centrality.rs:234-248
//   gt::pagerank_with        crates/gt-algo/src/centrality.rs:261-384
//   gt::eigenvector          crates/gt-algo/src/centrality.rs:405-497
// A reference panic is a thrown gt::Panic carrying the same message.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gt_hip.h"

namespace gt {

struct Panic : std::runtime_error {
    gt_status status;
    Panic(gt_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(gt_status s) {
    if (s != GT_OK) throw Panic(s, gt_last_error());
}

class Plan {
  public:
    // Plan::new(n, grain): max(1, ceil(n / grain)) balanced chunks.
    static Plan make(uint64_t n, uint64_t grain) {
        if (grain == 0) throw Panic(GT_ERR_INVALID_ARG, "Plan::new: grain must be non-zero");
        uint64_t c = (n + grain - 1) / grain;
        return Plan(n, c ? c : 1);
    }
    uint64_t len() const { return n_; }
    uint64_t chunks() const { return chunks_; }
    std::pair<uint64_t, uint64_t> range(uint64_t k) const {
        if (k >= chunks_) throw Panic(GT_ERR_INVALID_ARG, "Plan::range: chunk out of range");
        uint64_t base = n_ / chunks_, rem = n_ % chunks_;
        uint64_t start = k * base + (k < rem ? k : rem);
        return {start, start + base + (k < rem ? 1 : 0)};
    }

  private:
    Plan(uint64_t n, uint64_t c) : n_(n), chunks_(c) {}
    uint64_t n_, chunks_;
};

struct PowerIteration {
    double epsilon;
    uint64_t max_iter;  // 0 = no cap
    Plan plan;
    gt_power_iteration c() const {
        return gt_power_iteration{epsilon, max_iter, gt_plan{plan.len(), plan.chunks()}};
    }
};

struct Unity {
    gt_prop c() const { return gt_prop{GT_MAP_UNITY, GT_F64, GT_HOST, 0, 1.0, nullptr, 0, 0}; }
};

struct Constant {
    double value;
    explicit Constant(double v) : value(v) {}
    gt_prop c() const { return gt_prop{GT_MAP_CONSTANT, GT_F64, GT_HOST, 0, value, nullptr, 0, 0}; }
};

// A dense f64 map owned by one graph, in host memory.
class DenseProp {
  public:
    explicit DenseProp(uint64_t graph) : graph_(graph) {}
    DenseProp(uint64_t graph, std::vector<double> data) : graph_(graph), data_(std::move(data)) {}
    uint64_t graph_id() const { return graph_; }
    size_t len() const { return data_.size(); }
    const std::vector<double> &as_slice() const { return data_; }
    std::vector<double> &as_mut_slice() { return data_; }
    // Grow to `bound` (new slots zero) and never shrink; returns the old length.
    size_t sized_for(uint64_t graph, size_t bound) {
        if (graph != graph_) throw Panic(GT_ERR_WRONG_GRAPH, "map does not belong to this graph");
        size_t had = data_.size();
        if (had < bound) data_.resize(bound, 0.0);
        return had;
    }
    gt_prop c() const {
        return gt_prop{GT_MAP_DENSE, GT_F64, GT_HOST, 0, 0.0, data_.data(), data_.size(), graph_};
    }

  private:
    uint64_t graph_;
    std::vector<double> data_;
};

inline uint64_t fresh_graph_id() {
    static uint64_t next = 1u << 20;
    return ++next;
}

// Edge i is the i-th add_edge; the device copy is built (on the GPU, with the
// push-swap in-half order) on first use and rebuilt after a mutation.
class AdjList {
  public:
    static AdjList with_vertices(uint32_t n) { return AdjList(n); }
    uint64_t graph_id() const { return id_; }
    uint32_t vertex_bound() const { return n_; }
    uint32_t num_vertices() const { return n_; }
    uint64_t edge_bound() const { return src_.size(); }
    uint64_t add_edge(uint32_t s, uint32_t t) {
        if (s >= n_ || t >= n_) throw Panic(GT_ERR_NO_SUCH_VERTEX, "no such vertex");
        src_.push_back(s);
        tgt_.push_back(t);
        dev_.reset();
        return src_.size() - 1;
    }
    gt_graph *device() const {
        if (!dev_) {
            gt_graph *g = nullptr;
            check(gt_graph_build_edges(id_, n_, src_.size(), src_.data(), tgt_.data(), GT_HOST, &g));
            dev_.reset(g, [](gt_graph *p) { gt_graph_free(p); });
        }
        return dev_.get();
    }
    gt_view view() const { return gt_view{GT_VIEW_DIRECTED, 0, nullptr, 0, nullptr, 0}; }

  private:
    explicit AdjList(uint32_t n) : id_(fresh_graph_id()), n_(n) {}
    uint64_t id_;
    uint32_t n_;
    std::vector<uint32_t> src_, tgt_;
    mutable std::shared_ptr<gt_graph> dev_;
};

// Rev<&AdjList>: in-halves and out-halves swap.
struct Reversed {
    const AdjList &g;
    uint64_t graph_id() const { return g.graph_id(); }
    uint32_t vertex_bound() const { return g.vertex_bound(); }
    uint32_t num_vertices() const { return g.num_vertices(); }
    gt_graph *device() const { return g.device(); }
    gt_view view() const { return gt_view{GT_VIEW_REVERSED, 0, nullptr, 0, nullptr, 0}; }
};

template <typename G, typename W, typename P>
size_t pagerank_with(const G &g, const W &weight, const P &pers, double damping,
                     const PowerIteration &params, DenseProp &rank) {
    if (rank.graph_id() != g.graph_id())
        throw Panic(GT_ERR_WRONG_GRAPH, "pagerank: rank map does not belong to this graph");
    const size_t had = rank.sized_for(g.graph_id(), g.vertex_bound());
    gt_prop w = weight.c(), p = pers.c();
    gt_view v = g.view();
    gt_power_iteration pi = params.c();
    gt_vertex_out out{rank.as_mut_slice().data(), rank.len(), had, GT_HOST, 0, rank.graph_id()};
    uint64_t iters = 0;
    check(gt_pagerank_f64(g.device(), &v, &w, &p, damping, &pi, &out, &iters));
    return iters;
}

template <typename G, typename W>
size_t pagerank(const G &g, const W &weight, double damping, const PowerIteration &params,
                DenseProp &rank) {
    const uint32_t n = g.num_vertices();
    return pagerank_with(g, weight, Constant(n == 0 ? 0.0 : 1.0 / n), damping, params, rank);
}

template <typename G, typename W>
std::pair<double, size_t> eigenvector(const G &g, const W &weight, const PowerIteration &params,
                                      DenseProp &x) {
    if (x.graph_id() != g.graph_id())
        throw Panic(GT_ERR_WRONG_GRAPH,
                    "eigenvector: centrality map does not belong to this graph");
    const size_t had = x.sized_for(g.graph_id(), g.vertex_bound());
    gt_prop w = weight.c();
    gt_view v = g.view();
    gt_power_iteration pi = params.c();
    gt_vertex_out out{x.as_mut_slice().data(), x.len(), had, GT_HOST, 0, x.graph_id()};
    double eig = 0.0;
    uint64_t iters = 0;
    check(gt_eigenvector_f64(g.device(), &v, &w, &pi, &out, &eig, &iters));
    return {eig, iters};
}

}  // namespace gt
