// Structural check of a native graph before it reaches the kernels, which do not bounds-check:
// the conditions and ValueError messages of the reference's native loader
// (numba_engine.py:878-927), in its order, over the hm_native_graph the caller would pass to
// hm_open. One pass over each link array; the upper-row ownership is checked by claiming rows
// instead of sorting the offsets.
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/hnsw_mi355x.h"

namespace hm {
int invalid(const std::string &msg);
}

namespace {

// Counts in [0, cap] and every used slot a node id; rows have `width` slots.
int check_rows(const int32_t *links, const int32_t *counts, int64_t nrows, int width, int cap,
               int64_t n, const char *where) {
    for (int64_t r = 0; r < nrows; ++r)
        if (counts[r] < 0 || counts[r] > cap)
            return hm::invalid(std::string(where) + ": link counts outside [0, " +
                               std::to_string(cap) + "]");
    for (int64_t r = 0; r < nrows; ++r) {
        const int32_t *row = links + r * (int64_t)width;
        for (int j = 0; j < counts[r]; ++j)
            if (row[j] < 0 || row[j] >= n)
                return hm::invalid(std::string(where) + ": a link points outside the graph");
    }
    return HM_OK;
}

}  // namespace

extern "C" int hm_check_graph(const hm_native_graph *g) {
    if (!g) return hm::invalid("null argument");
    const int64_t n = g->n, rows = g->rows;
    if (n < 0 || rows < 0 || g->M < 1) return hm::invalid("bad graph header");
    if (n == 0) {
        if (rows || g->entry_point != -1 || g->max_level != -1)
            return hm::invalid("an empty graph has no entry point and no layers");
        return HM_OK;
    }
    int64_t top = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (g->levels[i] < 0) return hm::invalid("negative node level");
        if (g->levels[i] > top) top = g->levels[i];
    }
    if (g->max_level != top || g->entry_point < 0 || g->entry_point >= n ||
        g->levels[g->entry_point] != top)
        return hm::invalid("the entry point is not a node of the top layer");
    const int m0 = 2 * g->M;
    int st = check_rows(g->links0, g->count0, n, m0 + 1, m0, n, "layer 0");
    if (st) return st;
    // Each upper node owns rows [offset, offset + level) of links_up; together they must tile
    // [0, rows) exactly. layer[r] records the layer of row r (0: not yet claimed).
    std::vector<int32_t> layer((size_t)rows, 0);
    int64_t claimed = 0;
    const char *tiling = "upper-layer rows of the nodes overlap or leave gaps";
    for (int64_t i = 0; i < n; ++i) {
        const int64_t lv = g->levels[i];
        if (lv == 0) continue;
        const int64_t off = g->upper_offset[i];
        if (off < 0 || off > rows - lv) return hm::invalid(tiling);
        for (int64_t j = 0; j < lv; ++j) {
            if (layer[(size_t)(off + j)] != 0) return hm::invalid(tiling);
            layer[(size_t)(off + j)] = (int32_t)(j + 1);
        }
        claimed += lv;
    }
    if (claimed != rows) return hm::invalid(tiling);
    st = check_rows(g->links_up, g->count_up, rows, g->M + 1, g->M, n, "upper layers");
    if (st) return st;
    for (int64_t r = 0; r < rows; ++r) {
        const int32_t *row = g->links_up + r * (int64_t)(g->M + 1);
        for (int j = 0; j < g->count_up[r]; ++j)
            if (g->levels[row[j]] < layer[(size_t)r])
                return hm::invalid("an upper-layer link points to a node absent from that layer");
    }
    return HM_OK;
}
