// The reference's centrality acceptance tests (crates/gt-algo/tests/u16_centrality.rs),
// restated against the C++ mirror (include/gt_centrality.hpp) and run on the GPU.
// Each test keeps the reference's name, inputs and tolerance.  Bit-identity
// across thread counts becomes bit-identity across repeated runs.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "gt_centrality.hpp"

using namespace gt;

static int failures = 0;
#define EXPECT(cond, ...)                                              \
    do {                                                               \
        if (!(cond)) {                                                 \
            std::printf("  FAIL %s:%d: %s\n    ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                                  \
            std::printf("\n");                                         \
            ++failures;                                                \
            return;                                                    \
        }                                                              \
    } while (0)

struct Rng {  // SplitMix64, u16_centrality.rs:58-72
    uint64_t s;
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
};

using Edges = std::vector<std::pair<uint32_t, uint32_t>>;

static AdjList build(uint32_t n, const Edges &e) {
    AdjList g = AdjList::with_vertices(n);
    for (auto &p : e) g.add_edge(p.first, p.second);
    return g;
}

static Edges thousand_edges() {  // u16_centrality.rs:93-107
    const uint32_t N = 1000;
    Rng r{0x243f6a8885a308d3ull};
    Edges e;
    for (uint32_t i = 0; i < N; ++i) e.push_back({i, (i + 1) % N});
    for (int k = 0; k < 4000; ++k) {
        uint32_t s = r.below(N);
        uint32_t t = r.below(N);
        e.push_back({s, t});
    }
    Edges kept;
    for (auto &p : e) {
        uint32_t i = p.first;
        bool sink = i >= 17 && (i - 17) % 20 == 0 && i < 17 + 20 * 50;
        if (!sink) kept.push_back(p);
    }
    return kept;
}

static PowerIteration params(uint64_t n, double eps, uint64_t max_iter) {
    return PowerIteration{eps, max_iter, Plan::make(n, 64)};
}

// u16_centrality.rs:138-175
static std::vector<double> dense_pagerank(uint32_t n, const Edges &e, const std::vector<double> &w,
                                          const std::vector<double> &pers, double d, int sweeps) {
    std::vector<double> deg(n, 0.0), m((size_t)n * n, 0.0);
    for (size_t k = 0; k < e.size(); ++k) deg[e[k].first] += w[k];
    for (size_t k = 0; k < e.size(); ++k)
        if (deg[e[k].first] != 0.0) m[(size_t)e[k].second * n + e[k].first] += w[k] / deg[e[k].first];
    std::vector<double> r = pers;
    for (int it = 0; it < sweeps; ++it) {
        double p_sink = 0.0;
        for (uint32_t i = 0; i < n; ++i)
            if (deg[i] == 0.0) p_sink += r[i];
        std::vector<double> nr(n);
        for (uint32_t t = 0; t < n; ++t) {
            double acc = p_sink * pers[t];
            for (uint32_t s = 0; s < n; ++s) acc += m[(size_t)t * n + s] * r[s];
            nr[t] = (1.0 - d) * pers[t] + d * acc;
        }
        r = nr;
    }
    return r;
}

static void pagerank_matches_a_dense_power_iteration() {
    Edges e = thousand_edges();
    AdjList g = build(1000, e);
    std::vector<double> ones(e.size(), 1.0), pers(1000, 1.0 / 1000);
    std::vector<double> want = dense_pagerank(1000, e, ones, pers, 0.85, 400);
    DenseProp rank(g.graph_id());
    size_t iters = pagerank(g, Unity{}, 0.85, params(1000, 1e-13, 10000), rank);
    EXPECT(iters >= 5 && iters < 10000, "converged in %zu sweeps", iters);
    double worst = 0.0, total = 0.0;
    for (int i = 0; i < 1000; ++i) {
        worst = std::fmax(worst, std::fabs(rank.as_slice()[i] - want[i]));
        total += rank.as_slice()[i];
    }
    EXPECT(worst < 1e-12, "worst absolute difference %.3e", worst);
    EXPECT(std::fabs(total - 1.0) < 1e-9, "total mass %.17g", total);
}

static void pagerank_is_bit_identical_across_runs() {
    AdjList g = build(1000, thousand_edges());
    DenseProp a(g.graph_id()), b(g.graph_id());
    size_t ia = pagerank(g, Unity{}, 0.85, params(1000, 1e-13, 10000), a);
    size_t ib = pagerank(g, Unity{}, 0.85, params(1000, 1e-13, 10000), b);
    EXPECT(ia == ib, "sweeps %zu vs %zu", ia, ib);
    for (int i = 0; i < 1000; ++i)
        EXPECT(a.as_slice()[i] == b.as_slice()[i], "vertex %d differs", i);
}

static void an_odd_number_of_sweeps_leaves_the_latest_iterate_in_the_callers_map() {
    AdjList g = build(2, {{0, 1}});
    DenseProp rank(g.graph_id());
    EXPECT(pagerank(g, Unity{}, 0.85, params(2, 1e-12, 1), rank) == 1, "one sweep");
    const auto &got = rank.as_slice();
    EXPECT(got[0] != 0.5, "the initial iterate was handed back");
    EXPECT(std::fabs(got[0] - 0.2875) < 1e-15, "rank[0] = %.17g", got[0]);
    EXPECT(std::fabs(got[1] - 0.7125) < 1e-15, "rank[1] = %.17g", got[1]);
    DenseProp rank2(g.graph_id());
    EXPECT(pagerank(g, Unity{}, 0.85, params(2, 1e-12, 2), rank2) == 2, "two sweeps");
    std::vector<double> want = dense_pagerank(2, {{0, 1}}, {1.0}, {0.5, 0.5}, 0.85, 2);
    for (int i = 0; i < 2; ++i)
        EXPECT(std::fabs(rank2.as_slice()[i] - want[i]) < 1e-15, "%.17g vs %.17g",
               rank2.as_slice()[i], want[i]);
}

static void a_zero_weight_predecessor_does_not_poison_the_vector() {
    AdjList g = build(3, {{0, 1}, {1, 2}});
    DenseProp w(g.graph_id(), {0.0, 1.0});
    DenseProp rank(g.graph_id());
    pagerank(g, w, 0.85, params(3, 1e-13, 1000), rank);
    for (double x : rank.as_slice()) EXPECT(std::isfinite(x), "rank = %g", x);
    AdjList alone = build(3, {{1, 2}});
    DenseProp want(alone.graph_id());
    pagerank(alone, Unity{}, 0.85, params(3, 1e-13, 1000), want);
    for (int i = 0; i < 3; ++i)
        EXPECT(std::fabs(rank.as_slice()[i] - want.as_slice()[i]) < 1e-12, "%.17g vs %.17g",
               rank.as_slice()[i], want.as_slice()[i]);
}

static void a_personalisation_vector_biases_the_answer() {
    Edges e = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    AdjList g = build(4, e);
    DenseProp flat(g.graph_id());
    pagerank(g, Unity{}, 0.85, params(4, 1e-14, 10000), flat);
    for (double x : flat.as_slice()) EXPECT(std::fabs(x - 0.25) < 1e-12, "%.17g", x);
    DenseProp pers(g.graph_id(), {0.7, 0.1, 0.1, 0.1});
    DenseProp biased(g.graph_id());
    pagerank_with(g, Unity{}, pers, 0.85, params(4, 1e-14, 10000), biased);
    const auto &got = biased.as_slice();
    EXPECT(got[0] > 0.25, "favoured vertex fell to %g", got[0]);
    EXPECT(got[0] > got[1] && got[1] > got[2] && got[2] > got[3], "ordering");
    double total = got[0] + got[1] + got[2] + got[3];
    EXPECT(std::fabs(total - 1.0) < 1e-9, "total %.17g", total);
    std::vector<double> want = dense_pagerank(4, e, {1, 1, 1, 1}, {0.7, 0.1, 0.1, 0.1}, 0.85, 400);
    for (int i = 0; i < 4; ++i) EXPECT(std::fabs(got[i] - want[i]) < 1e-12, "%d", i);
}

static void an_already_sized_map_is_the_initial_iterate() {
    AdjList g = build(2, {{0, 1}});
    DenseProp rank(g.graph_id(), {0.2875, 0.7125});
    pagerank(g, Unity{}, 0.85, params(2, 1e-12, 1), rank);
    double want0 = 0.15 * 0.5 + 0.85 * (0.7125 * 0.5);
    double want1 = 0.15 * 0.5 + 0.85 * (0.7125 * 0.5 + 0.2875);
    EXPECT(std::fabs(rank.as_slice()[0] - want0) < 1e-15, "%.17g", rank.as_slice()[0]);
    EXPECT(std::fabs(rank.as_slice()[1] - want1) < 1e-15, "%.17g", rank.as_slice()[1]);
}

static void a_plan_that_does_not_cover_the_index_space_is_rejected() {
    AdjList g = build(4, {{0, 1}});
    DenseProp rank(g.graph_id());
    try {
        pagerank(g, Unity{}, 0.85, params(3, 1e-6, 10), rank);
    } catch (const Panic &p) {
        EXPECT(std::string(p.what()).find("sweep: the plan must cover the output exactly") !=
                   std::string::npos,
               "message: %s", p.what());
        return;
    }
    EXPECT(false, "no panic");
}

static void eigenvector_satisfies_its_own_equation() {
    Rng r{0xfeedfacecafebeefull};
    const uint32_t n = 60;
    Edges e;
    for (uint32_t i = 0; i < n; ++i) e.push_back({i, (i + 1) % n});
    for (int k = 0; k < 240; ++k) {
        uint32_t s = r.below(n);
        uint32_t t = r.below(n);
        e.push_back({s, t});
    }
    AdjList g = build(n, e);
    std::vector<double> wv(e.size());
    for (size_t i = 0; i < e.size(); ++i) wv[i] = 1.0 + (double)(i % 5) * 0.25;
    DenseProp w(g.graph_id(), wv);
    DenseProp c(g.graph_id());
    auto [eig, iters] = eigenvector(g, w, params(n, 1e-14, 100000), c);
    EXPECT(iters > 1, "did not iterate");
    EXPECT(eig > 1.0, "eigenvalue %g", eig);
    std::vector<double> ac(n, 0.0);
    for (size_t k = 0; k < e.size(); ++k) ac[e[k].second] += wv[k] * c.as_slice()[e[k].first];
    double worst = 0.0, l2 = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        worst = std::fmax(worst, std::fabs(ac[i] - eig * c.as_slice()[i]));
        l2 += c.as_slice()[i] * c.as_slice()[i];
    }
    EXPECT(worst < 1e-10, "residual %.3e", worst);
    EXPECT(std::fabs(std::sqrt(l2) - 1.0) < 1e-12, "norm %.17g", std::sqrt(l2));
}

static void a_directed_cycle_has_unit_eigenvalue() {
    Edges e;
    for (uint32_t i = 0; i < 7; ++i) e.push_back({i, (i + 1) % 7});
    AdjList g = build(7, e);
    DenseProp c(g.graph_id());
    auto r1 = eigenvector(g, Unity{}, params(7, 1e-15, 1000), c);
    EXPECT(std::fabs(r1.first - 1.0) < 1e-12, "eigenvalue %.17g", r1.first);
    for (double x : c.as_slice()) EXPECT(std::fabs(x - 1.0 / std::sqrt(7.0)) < 1e-12, "%.17g", x);
    DenseProp c2(g.graph_id());
    auto r2 = eigenvector(g, Constant(2.5), params(7, 1e-15, 1000), c2);
    EXPECT(std::fabs(r2.first - 2.5) < 1e-12, "eigenvalue %.17g", r2.first);
}

static void an_edgeless_graph_has_a_zero_eigenvalue_and_no_nans() {
    AdjList g = build(5, {});
    DenseProp c(g.graph_id());
    auto [eig, iters] = eigenvector(g, Unity{}, params(5, 1e-12, 100), c);
    EXPECT(eig == 0.0, "eigenvalue %g", eig);
    EXPECT(iters >= 1, "iters %zu", iters);
    for (double x : c.as_slice()) EXPECT(x == 0.0, "c = %g", x);
}

static void an_odd_eigenvector_sweep_leaves_the_latest_iterate_in_the_callers_map() {
    AdjList g = build(3, {{0, 1}, {1, 2}});
    DenseProp c(g.graph_id());
    auto [eig, iters] = eigenvector(g, Unity{}, params(3, 1e-12, 1), c);
    EXPECT(iters == 1, "iters %zu", iters);
    EXPECT(std::fabs(eig - std::sqrt(2.0) / 3.0) < 1e-15, "eig %.17g", eig);
    double h = 1.0 / std::sqrt(2.0);
    EXPECT(std::fabs(c.as_slice()[0]) < 1e-15, "c[0] = %g", c.as_slice()[0]);
    EXPECT(std::fabs(c.as_slice()[1] - h) < 1e-15, "c[1] = %.17g", c.as_slice()[1]);
    EXPECT(std::fabs(c.as_slice()[2] - h) < 1e-15, "c[2] = %.17g", c.as_slice()[2]);
}

static void the_unity_weight_fast_path_is_the_general_path() {
    AdjList g = build(1000, thousand_edges());
    DenseProp ones(g.graph_id(), std::vector<double>(g.edge_bound(), 1.0));
    DenseProp a(g.graph_id()), b(g.graph_id());
    pagerank(g, Unity{}, 0.85, params(1000, 1e-13, 10000), a);
    pagerank(g, ones, 0.85, params(1000, 1e-13, 10000), b);
    for (int i = 0; i < 1000; ++i)
        EXPECT(a.as_slice()[i] == b.as_slice()[i], "vertex %d: %.17g vs %.17g", i, a.as_slice()[i],
               b.as_slice()[i]);
}

static void a_reversed_view_is_the_reversed_graph() {
    Edges e = thousand_edges(), r;
    for (auto &p : e) r.push_back({p.second, p.first});
    AdjList g = build(1000, e), h = build(1000, r);
    DenseProp a(g.graph_id()), b(h.graph_id());
    size_t ia = pagerank(Reversed{g}, Unity{}, 0.85, params(1000, 1e-13, 10000), a);
    size_t ib = pagerank(h, Unity{}, 0.85, params(1000, 1e-13, 10000), b);
    EXPECT(ia == ib || ia + 1 == ib || ib + 1 == ia, "sweeps %zu vs %zu", ia, ib);
    double l1 = 0;
    for (int i = 0; i < 1000; ++i) l1 += std::fabs(a.as_slice()[i] - b.as_slice()[i]);
    EXPECT(l1 < 1e-12, "L1 %.3e", l1);
}

int main() {
    std::vector<std::pair<const char *, std::function<void()>>> tests = {
        {"pagerank_matches_a_dense_power_iteration", pagerank_matches_a_dense_power_iteration},
        {"pagerank_is_bit_identical_across_runs", pagerank_is_bit_identical_across_runs},
        {"an_odd_number_of_sweeps_leaves_the_latest_iterate_in_the_callers_map",
         an_odd_number_of_sweeps_leaves_the_latest_iterate_in_the_callers_map},
        {"a_zero_weight_predecessor_does_not_poison_the_vector",
         a_zero_weight_predecessor_does_not_poison_the_vector},
        {"a_personalisation_vector_biases_the_answer", a_personalisation_vector_biases_the_answer},
        {"an_already_sized_map_is_the_initial_iterate", an_already_sized_map_is_the_initial_iterate},
        {"a_plan_that_does_not_cover_the_index_space_is_rejected",
         a_plan_that_does_not_cover_the_index_space_is_rejected},
        {"eigenvector_satisfies_its_own_equation", eigenvector_satisfies_its_own_equation},
        {"a_directed_cycle_has_unit_eigenvalue", a_directed_cycle_has_unit_eigenvalue},
        {"an_edgeless_graph_has_a_zero_eigenvalue_and_no_nans",
         an_edgeless_graph_has_a_zero_eigenvalue_and_no_nans},
        {"an_odd_eigenvector_sweep_leaves_the_latest_iterate_in_the_callers_map",
         an_odd_eigenvector_sweep_leaves_the_latest_iterate_in_the_callers_map},
        {"the_unity_weight_fast_path_is_the_general_path",
         the_unity_weight_fast_path_is_the_general_path},
        {"a_reversed_view_is_the_reversed_graph", a_reversed_view_is_the_reversed_graph},
    };
    for (auto &t : tests) {
        int before = failures;
        try {
            t.second();
        } catch (const std::exception &ex) {
            std::printf("  FAIL exception: %s\n", ex.what());
            ++failures;
        }
        std::printf("%s %s\n", failures == before ? "ok  " : "FAIL", t.first);
        std::fflush(stdout);
    }
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
