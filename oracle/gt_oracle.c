/*
 * gt_oracle.c -- CPU restatement of graph-tool-rs' PageRank and eigenvector
 * power iteration, used ONLY as test infrastructure: the checker that the HIP
 * library is compared against, and the timed CPU baseline.  Nothing in the
 * product links or calls this file.
 *
 * Compiled with -O2 -ffp-contract=off and no fast-math: Rust never contracts
 * a*b+c into an FMA, so restating the reference's operations in the same
 * order gives the same IEEE results.
 *
 * What each function follows (all paths relative to the reference tree):
 *   sm_next ............ crates/gt-algo/tests/u16_centrality.rs:58-72 (SplitMix64),
 *                        crates/gt-algo/benches/kernels.rs:49-73
 *   adj_insert_out ..... crates/gt-core/src/adj/block.rs:171-199 (O(1) push-swap)
 *   adj_insert_in ...... crates/gt-core/src/adj/block.rs:205-209
 *   oracle_adj_build ... crates/gt-core/src/adj/list.rs:258-280, 471-500
 *                        (out-half insert first, then in-half, edge id = call order)
 *   view_* ............. crates/gt-core/src/view/reversed.rs:109-140 (halves swap),
 *                        crates/gt-core/src/view/filtered.rs:116-124, 270-320
 *                        (keeps_incident = keep_edge(e) && keep_vertex(other))
 *   plan_range ......... crates/gt-core/src/par/plan.rs:55-101
 *   run_chunks ......... crates/gt-algo/src/centrality.rs:141-167 (sweep) and
 *                        crates/gt-core/src/par/reduce.rs:40-55 (det_reduce):
 *                        per-chunk partial, folded left to right from 0.0
 *   oracle_pagerank .... crates/gt-algo/src/centrality.rs:234-384
 *   oracle_eigenvector . crates/gt-algo/src/centrality.rs:405-497
 *   oracle_betweenness . crates/gt-algo/src/centrality.rs:500-689 (Brandes::run,
 *                        betweenness), crates/gt-core/src/view/undirected.rs:110-112
 *                        (undirected out_edges = the whole block)
 *
 * Pinning: every function above is checked by tests/test_oracle.py against the
 * reference's own known answers (u16_centrality.rs, the LAYOUT fixture of
 * u34_diff_adjacency.rs), against closed forms (directed cycles, the scipy
 * linear solve of the PageRank fixed point, dense eigenpairs) and against brute
 * force (a dense n x n power iteration).  oracle_pagerank_grad (a build
 * extension: implicit differentiation of the fixed point) is pinned by central
 * finite differences of oracle_pagerank and by the row-scale identity
 * sum_{e in out(s)} w_e dL/dw_e = 0.  oracle_betweenness is pinned by
 * tests/test_oracle_betweenness.py: the reference's own 100 seeded graphs of
 * u16_centrality.rs:707-770 against an independent restatement of its
 * reference_brandes helper (:628-690), the path closed form (:772-799), the
 * relabelling case (:801-831), the unreachable case (:870-887), thread-count
 * independence (:846-868) and brute force from the definition (all-pairs BFS
 * distances and path counts, sigma_st(v) = sigma_sv * sigma_vt).
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GT_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* SplitMix64 (u16_centrality.rs:58-72)                                      */
/* ------------------------------------------------------------------------ */

static inline uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t sm_next(uint64_t *s) {
    *s += 0x9e3779b97f4a7c15ULL;
    return sm_mix(*s);
}

/* 53-bit uniform in [0,1) from one SplitMix64 draw. */
static inline double sm_unit(uint64_t *s) {
    return (double)(sm_next(s) >> 11) * 0x1.0p-53;
}

/* Directed G(n, m): per edge src = next % n, then tgt = next % n
 * (kernels.rs:68-70 draws source before target). */
GT_EXPORT void oracle_gen_er(uint32_t n, uint64_t m, uint64_t seed,
                             uint32_t *src, uint32_t *tgt) {
    uint64_t s = seed;
    for (uint64_t i = 0; i < m; ++i) {
        src[i] = (uint32_t)(sm_next(&s) % n);
        tgt[i] = (uint32_t)(sm_next(&s) % n);
    }
}

/* R-MAT: edge i draws `scale` uniforms, one per bit level, most significant
 * level first; quadrant thresholds a, a+b, a+b+c; (row,col) = (src,tgt);
 * no label permutation, no noise. */
GT_EXPORT void oracle_gen_rmat(uint32_t scale, uint64_t m, double a, double b,
                               double c, uint64_t seed, uint32_t *src,
                               uint32_t *tgt) {
    uint64_t s = seed;
    const double ab = a + b, abc = a + b + c;
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            double r = sm_unit(&s);
            uint32_t bit = 1u << (scale - 1 - l);
            if (r < a) {
            } else if (r < ab) {
                v |= bit;
            } else if (r < abc) {
                u |= bit;
            } else {
                u |= bit;
                v |= bit;
            }
        }
        src[i] = u;
        tgt[i] = v;
    }
}

/* ------------------------------------------------------------------------ */
/* AdjList: one block per vertex, out-half [0,out_len) then in-half           */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t other;
    uint32_t idx;
} gt_entry; /* AdjEntry, 8 bytes (entry.rs:39-48) */

typedef struct {
    uint32_t n;
    uint64_t edge_bound;
    uint64_t *start;   /* n+1: block v occupies e[start[v] .. start[v+1]) */
    uint32_t *out_len; /* n */
    uint32_t *len;     /* n */
    gt_entry *e;
} gt_adj;

static void adj_insert_out(gt_adj *g, uint32_t v, gt_entry x) {
    uint64_t at = g->start[v] + g->out_len[v];
    uint64_t back = g->start[v] + g->len[v];
    if (g->out_len[v] < g->len[v]) {
        /* in-half non-empty: its first entry moves to the back */
        g->e[back] = g->e[at];
    }
    g->e[at] = x;
    g->len[v] += 1;
    g->out_len[v] += 1;
}

static void adj_insert_in(gt_adj *g, uint32_t v, gt_entry x) {
    g->e[g->start[v] + g->len[v]] = x;
    g->len[v] += 1;
}

GT_EXPORT void oracle_adj_free(gt_adj *g) {
    if (!g) return;
    free(g->start);
    free(g->out_len);
    free(g->len);
    free(g->e);
    free(g);
}

/* Build by calling add_edge(src[i], tgt[i]) for i = 0..m in order; edge i
 * gets id i.  Returns NULL if an endpoint is out of range. */
GT_EXPORT gt_adj *oracle_adj_build(uint32_t n, uint64_t m, const uint32_t *src,
                                   const uint32_t *tgt) {
    for (uint64_t i = 0; i < m; ++i)
        if (src[i] >= n || tgt[i] >= n) return NULL;
    gt_adj *g = (gt_adj *)calloc(1, sizeof(gt_adj));
    g->n = n;
    g->edge_bound = m;
    g->start = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    g->out_len = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    g->len = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    for (uint64_t i = 0; i < m; ++i) {
        g->start[src[i] + 1] += 1;
        g->start[tgt[i] + 1] += 1;
    }
    for (uint32_t v = 0; v < n; ++v) g->start[v + 1] += g->start[v];
    g->e = (gt_entry *)malloc((g->start[n] ? g->start[n] : 1) * sizeof(gt_entry));
    for (uint64_t i = 0; i < m; ++i) {
        gt_entry o = {tgt[i], (uint32_t)i};
        gt_entry in = {src[i], (uint32_t)i};
        adj_insert_out(g, src[i], o);
        adj_insert_in(g, tgt[i], in);
    }
    return g;
}

/* Flatten to CSC (in-halves) and CSR (out-halves), each in storage order. */
GT_EXPORT void oracle_adj_export(const gt_adj *g, uint32_t *in_off,
                                 uint32_t *in_other, uint32_t *in_eid,
                                 uint32_t *out_off, uint32_t *out_other,
                                 uint32_t *out_eid) {
    uint64_t pi = 0, po = 0;
    for (uint32_t v = 0; v < g->n; ++v) {
        in_off[v] = (uint32_t)pi;
        out_off[v] = (uint32_t)po;
        const gt_entry *b = g->e + g->start[v];
        for (uint32_t k = 0; k < g->out_len[v]; ++k, ++po) {
            out_other[po] = b[k].other;
            out_eid[po] = b[k].idx;
        }
        for (uint32_t k = g->out_len[v]; k < g->len[v]; ++k, ++pi) {
            in_other[pi] = b[k].other;
            in_eid[pi] = b[k].idx;
        }
    }
    in_off[g->n] = (uint32_t)pi;
    out_off[g->n] = (uint32_t)po;
}

/* Whole block of v (out-half then in-half), for layout fixtures. */
GT_EXPORT uint32_t oracle_adj_block(const gt_adj *g, uint32_t v,
                                    uint32_t *other, uint32_t *idx,
                                    uint32_t *out_len) {
    const gt_entry *b = g->e + g->start[v];
    for (uint32_t k = 0; k < g->len[v]; ++k) {
        other[k] = b[k].other;
        idx[k] = b[k].idx;
    }
    *out_len = g->out_len[v];
    return g->len[v];
}

/* ------------------------------------------------------------------------ */
/* Views: reversed swaps the halves; a mask filter keeps an incidence iff     */
/* keep_edge(e) && keep_vertex(other).                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    const gt_adj *g;
    int reversed;
    const uint8_t *vmask; /* NULL: all vertices kept */
    const uint8_t *emask; /* NULL: all edges kept */
} gt_view;

static inline const gt_entry *view_in(const gt_view *w, uint32_t v, uint32_t *cnt) {
    const gt_adj *g = w->g;
    const gt_entry *b = g->e + g->start[v];
    if (w->reversed) {
        *cnt = g->out_len[v];
        return b;
    }
    *cnt = g->len[v] - g->out_len[v];
    return b + g->out_len[v];
}

static inline const gt_entry *view_out(const gt_view *w, uint32_t v, uint32_t *cnt) {
    const gt_adj *g = w->g;
    const gt_entry *b = g->e + g->start[v];
    if (!w->reversed) {
        *cnt = g->out_len[v];
        return b;
    }
    *cnt = g->len[v] - g->out_len[v];
    return b + g->out_len[v];
}

static inline int view_keeps(const gt_view *w, gt_entry x) {
    if (w->emask && !w->emask[x.idx]) return 0;
    if (w->vmask && !w->vmask[x.other]) return 0;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Plan and the ordered fold                                                 */
/* ------------------------------------------------------------------------ */

static inline void plan_range(uint64_t n, uint64_t chunks, uint64_t k,
                              uint64_t *lo, uint64_t *hi) {
    uint64_t base = n / chunks, rem = n % chunks;
    uint64_t start = k * base + (k < rem ? k : rem);
    *lo = start;
    *hi = start + base + (k < rem ? 1 : 0);
}

static inline uint64_t plan_chunks_for(uint64_t n, uint64_t grain) {
    uint64_t c = (n + grain - 1) / grain;
    return c ? c : 1;
}

typedef double (*chunk_fn)(void *ctx, uint64_t lo, uint64_t hi);

typedef struct {
    chunk_fn f;
    void *ctx;
    uint64_t n, chunks;
    double *partials;
    atomic_ulong next;
} chunk_job;

static void *chunk_worker(void *arg) {
    chunk_job *j = (chunk_job *)arg;
    for (;;) {
        uint64_t k = atomic_fetch_add(&j->next, 1);
        if (k >= j->chunks) break;
        uint64_t lo, hi;
        plan_range(j->n, j->chunks, k, &lo, &hi);
        j->partials[k] = j->f(j->ctx, lo, hi);
    }
    return NULL;
}

/* Map every chunk of Plan{n, chunks} (any thread, any order), then fold the
 * partials left to right from 0.0: the result does not depend on `threads`. */
static double run_chunks(uint64_t n, uint64_t chunks, int threads, chunk_fn f,
                         void *ctx) {
    double *partials = (double *)malloc(chunks * sizeof(double));
    chunk_job j;
    j.f = f;
    j.ctx = ctx;
    j.n = n;
    j.chunks = chunks;
    j.partials = partials;
    atomic_init(&j.next, 0);
    if (threads <= 1 || chunks == 1) {
        chunk_worker(&j);
    } else {
        int t = threads;
        if ((uint64_t)t > chunks) t = (int)chunks;
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)t);
        for (int i = 1; i < t; ++i) pthread_create(&th[i], NULL, chunk_worker, &j);
        chunk_worker(&j);
        for (int i = 1; i < t; ++i) pthread_join(th[i], NULL);
        free(th);
    }
    double acc = 0.0;
    for (uint64_t k = 0; k < chunks; ++k) acc = acc + partials[k];
    free(partials);
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Property maps as the path sees them                                       */
/* ------------------------------------------------------------------------ */

enum { MAP_UNITY = 0, MAP_CONSTANT = 1, MAP_DENSE = 2 };

typedef struct {
    int kind;
    double value;
    const double *data;
} gt_map;

static inline double map_get(const gt_map *m, uint64_t k) {
    if (m->kind == MAP_UNITY) return 1.0;
    if (m->kind == MAP_CONSTANT) return m->value;
    return m->data[k];
}

/* ------------------------------------------------------------------------ */
/* PageRank (centrality.rs:261-384)                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    const gt_view *w;
    const uint8_t *active;
    const double *deg;
    const gt_map *weight;
    const gt_map *pers;
    const double *prev;
    double *next;
    double p_sink;
    double d;
    const uint32_t *sinks;
} pr_ctx;

static double pr_sink_chunk(void *c, uint64_t lo, uint64_t hi) {
    pr_ctx *x = (pr_ctx *)c;
    double acc = 0.0;
    for (uint64_t k = lo; k < hi; ++k) acc += x->prev[x->sinks[k]];
    return acc;
}

static double pr_sweep_chunk(void *c, uint64_t lo, uint64_t hi) {
    pr_ctx *x = (pr_ctx *)c;
    double local = 0.0;
    const double d = x->d;
    for (uint64_t i = lo; i < hi; ++i) {
        if (!x->active[i]) {
            x->next[i] = x->prev[i];
            continue;
        }
        double pv = map_get(x->pers, i);
        double r = x->p_sink * pv;
        uint32_t cnt;
        const gt_entry *in = view_in(x->w, (uint32_t)i, &cnt);
        for (uint32_t k = 0; k < cnt; ++k) {
            if (!view_keeps(x->w, in[k])) continue;
            uint32_t s = in[k].other;
            double ds = x->deg[s];
            if (ds != 0.0) r += (x->prev[s] * map_get(x->weight, in[k].idx)) / ds;
        }
        double nv = (1.0 - d) * pv + d * r;
        local += fabs(nv - x->prev[i]);
        x->next[i] = nv;
    }
    return local;
}

/* Returns the sweep count, or -1 if the plan does not cover the index space
 * (the reference panics "sweep: the plan must cover the output exactly").
 * rank[0..n) is the in/out iterate; slots [had, n) are seeded with pers. */
GT_EXPORT int64_t oracle_pagerank(const gt_adj *g, int reversed,
                                  const uint8_t *vmask, const uint8_t *emask,
                                  int wkind, double wval, const double *w,
                                  int pkind, double pval, const double *pers,
                                  double damping, double epsilon,
                                  uint64_t max_iter, uint64_t plan_len,
                                  uint64_t plan_chunks, double *rank,
                                  uint64_t had, int threads) {
    gt_view view = {g, reversed, vmask, emask};
    const uint32_t n = g->n;
    gt_map wm = {wkind, wval, w};
    gt_map pm = {pkind, pval, pers};

    uint8_t *active = (uint8_t *)malloc(n ? n : 1);
    for (uint32_t v = 0; v < n; ++v) active[v] = vmask ? (vmask[v] != 0) : 1;

    double *deg = (double *)calloc(n ? n : 1, sizeof(double));
    uint32_t *sinks = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t nsinks = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (!active[v]) continue;
        uint32_t cnt;
        const gt_entry *out = view_out(&view, v, &cnt);
        double k;
        if (wkind == MAP_UNITY) {
            uint64_t c = 0;
            for (uint32_t j = 0; j < cnt; ++j) c += view_keeps(&view, out[j]);
            k = (double)c;
        } else {
            double acc = 0.0;
            for (uint32_t j = 0; j < cnt; ++j)
                if (view_keeps(&view, out[j])) acc += map_get(&wm, out[j].idx);
            k = acc;
        }
        deg[v] = k;
        if (k == 0.0) sinks[nsinks++] = v;
    }

    for (uint64_t i = had; i < n; ++i) rank[i] = map_get(&pm, i);

    if (plan_len != n) {
        free(active);
        free(deg);
        free(sinks);
        return -1;
    }
    double *scratch = (double *)calloc(n ? n : 1, sizeof(double));
    double *cur = rank, *next = scratch;
    uint64_t grain = (plan_len + plan_chunks - 1) / plan_chunks;
    if (grain < 1) grain = 1;
    uint64_t sink_chunks = plan_chunks_for(nsinks, grain);

    double delta = epsilon + 1.0;
    uint64_t iter = 0;
    while (delta >= epsilon) {
        pr_ctx c = {&view, active, deg, &wm, &pm, cur, next, 0.0, damping, sinks};
        c.p_sink = run_chunks(nsinks, sink_chunks, threads, pr_sink_chunk, &c);
        delta = run_chunks(n, plan_chunks, threads, pr_sweep_chunk, &c);
        double *t = cur;
        cur = next;
        next = t;
        iter += 1;
        if (max_iter > 0 && iter == max_iter) break;
    }
    if (iter % 2 == 1) memcpy(next, cur, sizeof(double) * n);
    free(scratch);
    free(active);
    free(deg);
    free(sinks);
    return (int64_t)iter;
}

/* ------------------------------------------------------------------------ */
/* Eigenvector (centrality.rs:405-497)                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    const gt_view *w;
    const uint8_t *active;
    const gt_map *weight;
    const double *prev;
    double *next;
    double scale;
} ev_ctx;

static double ev_sweep_chunk(void *c, uint64_t lo, uint64_t hi) {
    ev_ctx *x = (ev_ctx *)c;
    double acc = 0.0;
    for (uint64_t i = lo; i < hi; ++i) {
        if (!x->active[i]) {
            x->next[i] = x->prev[i];
            continue;
        }
        double s = 0.0;
        uint32_t cnt;
        const gt_entry *in = view_in(x->w, (uint32_t)i, &cnt);
        for (uint32_t k = 0; k < cnt; ++k) {
            if (!view_keeps(x->w, in[k])) continue;
            s += map_get(x->weight, in[k].idx) * x->prev[in[k].other];
        }
        x->next[i] = s;
        acc += s * s;
    }
    return acc;
}

static double ev_norm_chunk(void *c, uint64_t lo, uint64_t hi) {
    ev_ctx *x = (ev_ctx *)c;
    double acc = 0.0;
    for (uint64_t i = lo; i < hi; ++i) {
        if (!x->active[i]) continue;
        x->next[i] /= x->scale;
        acc += fabs(x->next[i] - x->prev[i]);
    }
    return acc;
}

/* Returns the sweep count (or -1 on a plan mismatch); *eig gets the last norm. */
GT_EXPORT int64_t oracle_eigenvector(const gt_adj *g, int reversed,
                                     const uint8_t *vmask, const uint8_t *emask,
                                     int wkind, double wval, const double *w,
                                     double epsilon, uint64_t max_iter,
                                     uint64_t plan_len, uint64_t plan_chunks,
                                     double *x, uint64_t had, int threads,
                                     double *eig) {
    gt_view view = {g, reversed, vmask, emask};
    const uint32_t n = g->n;
    gt_map wm = {wkind, wval, w};
    uint8_t *active = (uint8_t *)malloc(n ? n : 1);
    uint64_t nv = 0;
    for (uint32_t v = 0; v < n; ++v) {
        active[v] = vmask ? (vmask[v] != 0) : 1;
        nv += active[v];
    }
    double uniform = nv == 0 ? 0.0 : 1.0 / (double)nv;
    for (uint64_t i = had; i < n; ++i) x[i] = uniform;
    *eig = 0.0;
    if (plan_len != n) {
        free(active);
        return -1;
    }
    double *scratch = (double *)calloc(n ? n : 1, sizeof(double));
    double *cur = x, *next = scratch;
    double norm = 0.0, delta = epsilon + 1.0;
    uint64_t iter = 0;
    while (delta >= epsilon) {
        ev_ctx c = {&view, active, &wm, cur, next, 1.0};
        double sq = run_chunks(n, plan_chunks, threads, ev_sweep_chunk, &c);
        norm = sqrt(sq);
        c.scale = norm == 0.0 ? 1.0 : norm;
        delta = run_chunks(n, plan_chunks, threads, ev_norm_chunk, &c);
        double *t = cur;
        cur = next;
        next = t;
        iter += 1;
        if (max_iter > 0 && iter == max_iter) break;
    }
    if (iter % 2 == 1) memcpy(next, cur, sizeof(double) * n);
    free(scratch);
    free(active);
    *eig = norm;
    return (int64_t)iter;
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the PageRank fixed point (build extension; no reference       */
/* counterpart).  With M[t,s] = w_e / deg_s over edges e: s -> t, sink      */
/* indicator q (deg == 0) and x = (1-d) p + d (M x + (q.x) p), the gradient */
/* of L = c.x through the fixed point is lambda^T dF, where                 */
/*   lambda = c + d M^T lambda + d (p.lambda) q                             */
/* (the transposed system, solved by the same power iteration), and        */
/*   dL/dw_e = d * x_s / deg_s * (lambda_t - mu_s),  mu = M^T lambda        */
/*   dL/dp   = ((1-d) + d (q.x)) lambda                                     */
/*   dL/dd   = lambda . (x - p) / d                                         */
/* An edge whose source has zero weighted out-degree gets gradient 0.       */
/* Pinned by central finite differences of oracle_pagerank and by the       */
/* row-scale identity sum_{e in out(s)} w_e dL/dw_e = 0.                    */
/* ------------------------------------------------------------------------ */

GT_EXPORT int64_t oracle_pagerank_grad(const gt_adj *g, int reversed, int wkind, double wval,
                                       const double *w, int pkind, double pval,
                                       const double *pers, double damping, const double *x,
                                       const double *cvec, double epsilon, uint64_t max_iter,
                                       double *lambda, double *grad_w, double *grad_p,
                                       double *grad_d) {
    gt_view view = {g, reversed, NULL, NULL};
    const uint32_t n = g->n;
    const double d = damping;
    gt_map wm = {wkind, wval, w};
    gt_map pm = {pkind, pval, pers};
    double *deg = (double *)calloc(n ? n : 1, sizeof(double));
    double *nl = (double *)calloc(n ? n : 1, sizeof(double));
    double *mu = (double *)calloc(n ? n : 1, sizeof(double));
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t cnt;
        const gt_entry *out = view_out(&view, v, &cnt);
        double k = 0.0;
        for (uint32_t j = 0; j < cnt; ++j) k += map_get(&wm, out[j].idx);
        deg[v] = wkind == MAP_UNITY ? (double)cnt : k;
    }
    /* The system is linear in cvec: stop when the L1 change drops below
     * epsilon * ||cvec||_1 (epsilon itself when cvec is zero). */
    double c1 = 0.0;
    for (uint32_t v = 0; v < n; ++v) {
        lambda[v] = cvec[v];
        c1 += fabs(cvec[v]);
    }
    const double thr = epsilon * (c1 > 0.0 ? c1 : 1.0);
    double delta = thr + 1.0;
    uint64_t iter = 0;
    while (delta >= thr) {
        double pl = 0.0;
        for (uint32_t v = 0; v < n; ++v) pl += map_get(&pm, v) * lambda[v];
        delta = 0.0;
        for (uint32_t s = 0; s < n; ++s) {
            uint32_t cnt;
            const gt_entry *out = view_out(&view, s, &cnt);
            double acc = 0.0;
            for (uint32_t j = 0; j < cnt; ++j) acc += map_get(&wm, out[j].idx) * lambda[out[j].other];
            double m = deg[s] != 0.0 ? acc / deg[s] : 0.0;
            double v = cvec[s] + d * m + (deg[s] == 0.0 ? d * pl : 0.0);
            delta += fabs(v - lambda[s]);
            nl[s] = v;
        }
        memcpy(lambda, nl, sizeof(double) * n);
        iter += 1;
        if (max_iter > 0 && iter == max_iter) break;
    }
    double p_sink = 0.0, gd = 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        uint32_t cnt;
        const gt_entry *out = view_out(&view, s, &cnt);
        double acc = 0.0;
        for (uint32_t j = 0; j < cnt; ++j) acc += map_get(&wm, out[j].idx) * lambda[out[j].other];
        mu[s] = deg[s] != 0.0 ? acc / deg[s] : 0.0;
        if (deg[s] == 0.0) p_sink += x[s];
    }
    for (uint64_t e = 0; e < g->edge_bound; ++e) grad_w[e] = 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        if (deg[s] == 0.0) continue;
        uint32_t cnt;
        const gt_entry *out = view_out(&view, s, &cnt);
        const double cs = d * x[s] / deg[s];
        for (uint32_t j = 0; j < cnt; ++j) grad_w[out[j].idx] = cs * (lambda[out[j].other] - mu[s]);
    }
    for (uint32_t v = 0; v < n; ++v) {
        grad_p[v] = ((1.0 - d) + d * p_sink) * lambda[v];
        gd += lambda[v] * (x[v] - map_get(&pm, v));
    }
    *grad_d = gd / d;
    free(deg);
    free(nl);
    free(mu);
    return (int64_t)iter;
}

/* ------------------------------------------------------------------------ */
/* Betweenness, Brandes (centrality.rs:500-689)                              */
/* ------------------------------------------------------------------------ */

/* out_edges(u) of the view: the out-half (the in-half when reversed); on an
 * undirected view the whole block, out-half first (undirected.rs:110-112).
 * Filtered incidences are dropped (filtered.rs:122-124). */
typedef struct {
    const gt_view *w;
    int undirected;
    const uint32_t *verts; /* pivots, in order */
    uint64_t nverts;
    uint64_t chunks;
    double *vb;            /* [chunks][n] partials */
    double *eb;            /* [chunks][edge_bound] partials */
    atomic_ulong next;
} bc_job;

/* One source (Brandes::run, centrality.rs:545-611): BFS examining vertices on
 * dequeue, self-loops skipped, every parallel edge a distinct shortest path;
 * then dependencies accumulated in reverse dequeue order. */
static void bc_source(const bc_job *j, uint32_t s, int64_t *dist, double *sigma, double *dep,
                      uint32_t *order, uint64_t *in_head, uint64_t *in_tail, uint64_t *in_next,
                      uint32_t *in_pred, uint32_t *in_edge, double *vb, double *eb) {
    const gt_view *w = j->w;
    const uint32_t n = w->g->n;
    for (uint32_t v = 0; v < n; ++v) {
        dist[v] = -1;
        sigma[v] = 0.0;
        dep[v] = 0.0;
        in_head[v] = UINT64_MAX;
    }
    uint64_t npred = 0;
    uint64_t head = 0, tail = 0;
    sigma[s] = 1.0;
    dist[s] = 0;
    order[tail++] = s;
    while (head < tail) {
        const uint32_t u = order[head++];
        const int64_t du = dist[u];
        const double su = sigma[u];
        for (int half = 0; half < (j->undirected ? 2 : 1); ++half) {
            uint32_t cnt;
            const gt_entry *es = half == 0 ? view_out(w, u, &cnt) : view_in(w, u, &cnt);
            for (uint32_t k = 0; k < cnt; ++k) {
                if (!view_keeps(w, es[k])) continue;
                const uint32_t x = es[k].other;
                if (x == u) continue;
                if (dist[x] < 0) {
                    dist[x] = du + 1;
                    sigma[x] = su;
                    order[tail++] = x;
                } else if (dist[x] == du + 1) {
                    sigma[x] += su;
                } else {
                    continue;
                }
                /* incoming[x].push((u, e)), kept in push order */
                in_pred[npred] = u;
                in_edge[npred] = es[k].idx;
                in_next[npred] = UINT64_MAX;
                if (in_head[x] == UINT64_MAX) in_head[x] = npred;
                else in_next[in_tail[x]] = npred;
                in_tail[x] = npred;
                ++npred;
            }
        }
    }
    for (uint64_t i = tail; i-- > 0;) {
        const uint32_t u = order[i];
        const double su = sigma[u];
        for (uint64_t p = in_head[u]; p != UINT64_MAX; p = in_next[p]) {
            const uint32_t v = in_pred[p];
            const double factor = (sigma[v] / su) * (1.0 + dep[u]);
            dep[v] += factor;
            eb[in_edge[p]] += factor;
        }
        if (u != s) vb[u] += dep[u];
    }
}

static void *bc_worker(void *arg) {
    bc_job *j = (bc_job *)arg;
    const gt_adj *g = j->w->g;
    const uint32_t n = g->n;
    const uint64_t ents = 2 * (g->start[n] + 1);
    int64_t *dist = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    double *sigma = (double *)malloc(sizeof(double) * (n + 1));
    double *dep = (double *)malloc(sizeof(double) * (n + 1));
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint64_t *in_head = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    uint64_t *in_tail = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    uint64_t *in_next = (uint64_t *)malloc(sizeof(uint64_t) * ents);
    uint32_t *in_pred = (uint32_t *)malloc(sizeof(uint32_t) * ents);
    uint32_t *in_edge = (uint32_t *)malloc(sizeof(uint32_t) * ents);
    for (;;) {
        const uint64_t k = atomic_fetch_add(&j->next, 1);
        if (k >= j->chunks) break;
        uint64_t lo, hi;
        plan_range(j->nverts, j->chunks, k, &lo, &hi);
        double *vb = j->vb + k * (uint64_t)n;
        double *eb = j->eb + k * g->edge_bound;
        for (uint64_t i = lo; i < hi; ++i)
            bc_source(j, j->verts[i], dist, sigma, dep, order, in_head, in_tail, in_next, in_pred,
                      in_edge, vb, eb);
    }
    free(dist);
    free(sigma);
    free(dep);
    free(order);
    free(in_head);
    free(in_tail);
    free(in_next);
    free(in_pred);
    free(in_edge);
    return NULL;
}

/* betweenness (centrality.rs:630-689): maps zero-filled, the pivot list cut by
 * Plan::new(len, ceil(len / 16)) (BETWEENNESS_CHUNKS, :107-115), one private
 * accumulator pair per chunk, partials summed in chunk order, both maps
 * halved on an undirected view.  pivots == NULL means every vertex of the
 * view, in index order (the reference's only mode); a pivot list is a build
 * extension used to time a bounded sample.  Returns 0, or -1 for a pivot out
 * of range. */
GT_EXPORT int oracle_betweenness(const gt_adj *g, int reversed, int undirected,
                                 const uint8_t *vmask, const uint8_t *emask,
                                 const uint32_t *pivots, uint64_t npivots, int threads,
                                 double *vb, double *eb) {
    const uint32_t n = g->n;
    gt_view view = {g, reversed, vmask, emask};
    uint32_t *verts = NULL;
    uint64_t nv = 0;
    if (pivots) {
        for (uint64_t i = 0; i < npivots; ++i)
            if (pivots[i] >= n) return -1;
        verts = (uint32_t *)malloc(sizeof(uint32_t) * (npivots + 1));
        memcpy(verts, pivots, sizeof(uint32_t) * npivots);
        nv = npivots;
    } else {
        verts = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
        for (uint32_t v = 0; v < n; ++v)
            if (!vmask || vmask[v]) verts[nv++] = v;
    }
    for (uint32_t v = 0; v < n; ++v) vb[v] = 0.0;
    for (uint64_t e = 0; e < g->edge_bound; ++e) eb[e] = 0.0;
    const uint64_t grain = nv / 16 + (nv % 16 != 0);
    bc_job j;
    j.w = &view;
    j.undirected = undirected;
    j.verts = verts;
    j.nverts = nv;
    j.chunks = plan_chunks_for(nv, grain ? grain : 1);
    j.vb = (double *)calloc(j.chunks * (uint64_t)n + 1, sizeof(double));
    j.eb = (double *)calloc(j.chunks * g->edge_bound + 1, sizeof(double));
    atomic_init(&j.next, 0);
    int t = threads < 1 ? 1 : threads;
    if ((uint64_t)t > j.chunks) t = (int)j.chunks;
    if (t <= 1) {
        bc_worker(&j);
    } else {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)t);
        for (int i = 1; i < t; ++i) pthread_create(&th[i], NULL, bc_worker, &j);
        bc_worker(&j);
        for (int i = 1; i < t; ++i) pthread_join(th[i], NULL);
        free(th);
    }
    for (uint64_t k = 0; k < j.chunks; ++k) {
        const double *pv = j.vb + k * (uint64_t)n;
        const double *pe = j.eb + k * g->edge_bound;
        for (uint32_t v = 0; v < n; ++v) vb[v] += pv[v];
        for (uint64_t e = 0; e < g->edge_bound; ++e) eb[e] += pe[e];
    }
    if (undirected) {
        for (uint32_t v = 0; v < n; ++v) vb[v] /= 2.0;
        for (uint64_t e = 0; e < g->edge_bound; ++e) eb[e] /= 2.0;
    }
    free(j.vb);
    free(j.eb);
    free(verts);
    return 0;
}
