// Host graph builder: sequential insertion (Algorithm 1) into the reference's native arrays.
//
// Follows numba_engine.py:287-419 (_insert_batch), :98-148 (_search_layer), :201-262
// (_select) and :265-284 (_set_links): greedy descent with ef = 1, a layer search with
// ef_construction on every layer the node joins, neighbour selection (simple or heuristic with
// the extend / keep-pruned options), reverse links appended in selection order and a row
// re-selected with the same strategy when it exceeds its cap. Stale ids beyond a row's count
// are left as the reference leaves them.
//
// Distances: float64, left to right, unfused (numba_engine.py:57-70). When every value is an
// 8-bit integer (0..255) the same sums are computed in 32-bit integers: every partial sum is an
// exact integer, so the float64 value is identical and the build is bit-identical, only faster.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/hnsw_mi355x.h"

namespace {

struct Cand {
    double d;
    int64_t id;
};

inline bool lt(const Cand &a, const Cand &b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

struct MinCmp {  // std heap comparator giving a min-heap on (d, id)
    bool operator()(const Cand &a, const Cand &b) const { return lt(b, a); }
};
struct MaxCmp {  // max-heap on (d, id)
    bool operator()(const Cand &a, const Cand &b) const { return lt(a, b); }
};

struct Builder {
    int64_t n;
    int dim, metric, M, efc, selection;
    bool extend, keep_pruned;
    const float *vec;
    int32_t *links0, *count0, *links_up, *count_up;
    const int64_t *upper;
    int64_t *state;
    int64_t dcount = 0;
    bool use_int = false;
    std::vector<uint8_t> u8;  // exact 8-bit copy when eligible
    std::vector<uint32_t> visited;
    uint32_t epoch = 0;
    std::vector<Cand> cand_heap, res_heap;

    int stride0() const { return 2 * M + 1; }
    int stride_up() const { return M + 1; }

    uint32_t next_epoch() {
        epoch += 1;
        if (epoch >= 4294967295u) {
            std::fill(visited.begin(), visited.end(), 0u);
            epoch = 1;
        }
        return epoch;
    }

    // Distance from node a's vector (as the float64 "query") to node b.
    double dist_nodes(int64_t a, int64_t b) const {
        if (use_int) {
            const uint8_t *x = u8.data() + (size_t)b * dim, *q = u8.data() + (size_t)a * dim;
            int32_t acc = 0;
            if (metric == HM_L2) {
                for (int i = 0; i < dim; ++i) {
                    int32_t d = (int32_t)x[i] - (int32_t)q[i];
                    acc += d * d;
                }
                return (double)acc;
            }
            for (int i = 0; i < dim; ++i) acc += (int32_t)x[i] * (int32_t)q[i];
            return metric == HM_IP ? -(double)acc : 1.0 - (double)acc;
        }
        const float *x = vec + (size_t)b * dim, *q = vec + (size_t)a * dim;
        double acc = 0.0;
        if (metric == HM_L2) {
            for (int i = 0; i < dim; ++i) {
                double d = (double)x[i] - (double)q[i];
                acc += d * d;
            }
            return acc;
        }
        for (int i = 0; i < dim; ++i) acc += (double)x[i] * (double)q[i];
        return metric == HM_IP ? -acc : 1.0 - acc;
    }

    int32_t *row(int64_t node, int layer, int32_t **count) {
        if (layer == 0) {
            *count = count0 + node;
            return links0 + (size_t)node * stride0();
        }
        int64_t r = upper[node] + layer - 1;
        *count = count_up + r;
        return links_up + (size_t)r * stride_up();
    }

    // _search_layer (numba_engine.py:98-148) from `entry` (sorted), returns the sorted results.
    std::vector<Cand> search_layer(int64_t q, const std::vector<Cand> &entry, int ef, int layer) {
        uint32_t ep = next_epoch();
        cand_heap.assign(entry.begin(), entry.end());
        std::make_heap(cand_heap.begin(), cand_heap.end(), MinCmp());
        res_heap.assign(entry.begin(), entry.end());
        std::make_heap(res_heap.begin(), res_heap.end(), MaxCmp());
        while ((int)res_heap.size() > ef) {
            std::pop_heap(res_heap.begin(), res_heap.end(), MaxCmp());
            res_heap.pop_back();
        }
        for (const Cand &c : entry) visited[c.id] = ep;
        while (!cand_heap.empty()) {
            std::pop_heap(cand_heap.begin(), cand_heap.end(), MinCmp());
            Cand top = cand_heap.back();
            cand_heap.pop_back();
            if (top.d > res_heap.front().d) break;
            int32_t *cnt;
            const int32_t *nb = row(top.id, layer, &cnt);
            double worst = res_heap.front().d;
            const int c = *cnt;
            for (int i = 0; i < c; ++i) {
                int64_t e = nb[i];
                if (visited[e] == ep) continue;
                visited[e] = ep;
                double d = dist_nodes(q, e);
                dcount += 1;
                if ((int)res_heap.size() < ef || d < worst) {
                    cand_heap.push_back({d, e});
                    std::push_heap(cand_heap.begin(), cand_heap.end(), MinCmp());
                    res_heap.push_back({d, e});
                    std::push_heap(res_heap.begin(), res_heap.end(), MaxCmp());
                    if ((int)res_heap.size() > ef) {
                        std::pop_heap(res_heap.begin(), res_heap.end(), MaxCmp());
                        res_heap.pop_back();
                    }
                    worst = res_heap.front().d;
                }
            }
        }
        std::vector<Cand> out(res_heap.begin(), res_heap.end());
        std::sort(out.begin(), out.end(), lt);
        return out;
    }

    // _select (numba_engine.py:201-262).
    std::vector<Cand> select(int64_t base, const std::vector<Cand> &candidates, int m, int layer) {
        if (selection == 0) {
            std::vector<Cand> out(candidates.begin(),
                                  candidates.begin() + std::min<size_t>(m, candidates.size()));
            return out;
        }
        std::vector<Cand> work;
        const std::vector<Cand> *walk = &candidates;
        if (extend) {
            uint32_t ep = next_epoch();
            for (const Cand &c : candidates) visited[c.id] = ep;
            visited[base] = ep;
            work = candidates;
            for (const Cand &c : candidates) {
                int32_t *cnt;
                const int32_t *adj = row(c.id, layer, &cnt);
                const int ac = *cnt;
                for (int i = 0; i < ac; ++i) {
                    int64_t a = adj[i];
                    if (visited[a] != ep) {
                        visited[a] = ep;
                        dcount += 1;
                        work.push_back({dist_nodes(base, a), a});
                    }
                }
            }
            std::sort(work.begin(), work.end(), lt);
            walk = &work;
        }
        std::vector<Cand> kept, discarded;
        for (const Cand &c : *walk) {
            if ((int)kept.size() >= m) break;
            bool good = true;
            for (const Cand &r : kept) {
                dcount += 1;
                if (!(c.d < dist_nodes(c.id, r.id))) {
                    good = false;
                    break;
                }
            }
            if (good) kept.push_back(c);
            else discarded.push_back(c);
        }
        if (keep_pruned) {
            for (const Cand &c : discarded) {
                if ((int)kept.size() >= m) break;
                kept.push_back(c);
            }
        }
        return kept;
    }

    void set_links(int64_t node, int layer, const std::vector<Cand> &chosen) {
        int32_t *cnt;
        int32_t *r = row(node, layer, &cnt);
        for (size_t i = 0; i < chosen.size(); ++i) r[i] = (int32_t)chosen[i].id;
        *cnt = (int32_t)chosen.size();
    }

    // _insert_batch body for one node (numba_engine.py:310-419).
    void insert(int64_t node, int64_t level) {
        if (state[0] < 0) {
            state[0] = node;
            state[1] = level;
            return;
        }
        const int64_t ep = state[0];
        std::vector<Cand> entry{{dist_nodes(node, ep), ep}};
        dcount += 1;
        const int64_t max_level = state[1];
        for (int64_t layer = max_level; layer > level; --layer)
            entry = search_layer(node, entry, 1, (int)layer);
        const int m0 = 2 * M;
        for (int64_t layer = std::min(level, max_level); layer >= 0; --layer) {
            std::vector<Cand> candidates = search_layer(node, entry, efc, (int)layer);
            std::vector<Cand> chosen = select(node, candidates, M, (int)layer);
            set_links(node, (int)layer, chosen);
            const int cap = layer == 0 ? m0 : M;
            for (const Cand &c : chosen) {
                const int64_t e = c.id;
                int32_t *cnt;
                int32_t *r = row(e, (int)layer, &cnt);
                r[*cnt] = (int32_t)node;
                *cnt += 1;
                if (*cnt > cap) {
                    std::vector<Cand> scored;
                    scored.reserve(*cnt);
                    for (int i = 0; i < *cnt; ++i) {
                        dcount += 1;
                        scored.push_back({dist_nodes(e, r[i]), r[i]});
                    }
                    std::sort(scored.begin(), scored.end(), lt);
                    std::vector<Cand> kept = select(e, scored, cap, (int)layer);
                    set_links(e, (int)layer, kept);
                }
            }
            entry = std::move(candidates);
        }
        if (level > max_level) {
            state[0] = node;
            state[1] = level;
        }
    }
};

}  // namespace

extern "C" int hm_build(int64_t start, int64_t n, int32_t dim, int32_t metric, int32_t M,
                        int32_t ef_construction, int32_t selection, int32_t extend_candidates,
                        int32_t keep_pruned, const float *vectors, const int64_t *levels,
                        int32_t *links0, int32_t *count0, const int64_t *upper_offset,
                        int32_t *links_up, int32_t *count_up, int64_t *state, int64_t *dcount) {
    if (start < 0 || n < start || dim < 1 || M < 2 || ef_construction < 1 || !state)
        return HM_EINVAL;
    Builder b;
    b.n = n;
    b.dim = dim;
    b.metric = metric;
    b.M = M;
    b.efc = ef_construction;
    b.selection = selection;
    b.extend = extend_candidates != 0;
    b.keep_pruned = keep_pruned != 0;
    b.vec = vectors;
    b.links0 = links0;
    b.count0 = count0;
    b.links_up = links_up;
    b.count_up = count_up;
    b.upper = upper_offset;
    b.state = state;
    b.visited.assign((size_t)(n > 0 ? n : 1), 0u);
    if (metric != HM_COSINE && n > 0) {
        const size_t total = (size_t)n * dim;
        bool ok = true;
        for (size_t i = 0; i < total && ok; ++i) {
            const float x = vectors[i];
            ok = x >= 0.0f && x <= 255.0f && x == floorf(x);
        }
        if (ok) {
            b.use_int = true;
            b.u8.resize(total);
            for (size_t i = 0; i < total; ++i) b.u8[i] = (uint8_t)vectors[i];
        }
    }
    for (int64_t node = start; node < n; ++node) b.insert(node, levels[node]);
    if (dcount) *dcount += b.dcount;
    return HM_OK;
}
