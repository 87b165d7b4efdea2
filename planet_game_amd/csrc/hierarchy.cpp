// Host construction of the aggregation-multigrid hierarchy structure and of
// the index maps the device kernels use.
//
// The aggregates, coarse patterns and entry targets are bit-identical to the
// reference's build_hierarchy (sim/planet/surface/heat_transport.cpp:73-200):
// they come from the same greedy, order-dependent algorithm on the
// unit-conductance graph, so they depend on the mesh level only and are
// shared by every instance of that level.
#include "internal.hpp"

#include <algorithm>
#include <stdexcept>

namespace psim {
namespace {

// heat_transport.cpp:73-108. Pass 1 claims, in row order, every row whose
// whole neighbourhood is still free (the row and all its neighbours form an
// aggregate); pass 2 gives each leftover row the aggregate of its most
// strongly coupled aggregated neighbour (largest -value, first on ties),
// else a new singleton aggregate.
int32_t aggregate_rows(const HostLevel& lv, std::vector<int32_t>& agg) {
    agg.assign(static_cast<size_t>(lv.n), -1);
    int32_t count = 0;
    for (int32_t r = 0; r < lv.n; ++r) {
        bool whole = agg[r] < 0;
        for (int32_t k = lv.row[r]; whole && k < lv.row[r + 1]; ++k) whole = agg[lv.col[k]] < 0;
        if (!whole) continue;
        agg[r] = count;
        for (int32_t k = lv.row[r]; k < lv.row[r + 1]; ++k) agg[lv.col[k]] = count;
        ++count;
    }
    for (int32_t r = 0; r < lv.n; ++r) {
        if (agg[r] >= 0) continue;
        double strongest = 0.0;
        int32_t pick = -1;
        for (int32_t k = lv.row[r]; k < lv.row[r + 1]; ++k) {
            const int32_t a = agg[lv.col[k]];
            if (a >= 0 && -lv.unit_off[k] > strongest) {
                strongest = -lv.unit_off[k];
                pick = a;
            }
        }
        agg[r] = pick >= 0 ? pick : count++;
    }
    return count;
}

// heat_transport.cpp:113-154: the coarse pattern is, per aggregate, the
// sorted set of other aggregates its rows touch; each fine entry lands on
// that slot or on the coarse diagonal.
HostLevel coarsen(HostLevel& fine) {
    HostLevel coarse;
    coarse.n = aggregate_rows(fine, fine.aggregate);
    std::vector<std::vector<int32_t>> touch(static_cast<size_t>(coarse.n));
    for (int32_t r = 0; r < fine.n; ++r)
        for (int32_t k = fine.row[r]; k < fine.row[r + 1]; ++k) {
            const int32_t a = fine.aggregate[r], b = fine.aggregate[fine.col[k]];
            if (a != b) touch[a].push_back(b);
        }
    coarse.row.assign(static_cast<size_t>(coarse.n) + 1, 0);
    for (int32_t a = 0; a < coarse.n; ++a) {
        auto& t = touch[a];
        std::sort(t.begin(), t.end());
        t.erase(std::unique(t.begin(), t.end()), t.end());
        coarse.col.insert(coarse.col.end(), t.begin(), t.end());
        coarse.row[a + 1] = static_cast<int32_t>(coarse.col.size());
    }
    coarse.nnz = static_cast<int32_t>(coarse.col.size());
    fine.target.assign(static_cast<size_t>(fine.nnz), -1);
    for (int32_t r = 0; r < fine.n; ++r)
        for (int32_t k = fine.row[r]; k < fine.row[r + 1]; ++k) {
            const int32_t a = fine.aggregate[r], b = fine.aggregate[fine.col[k]];
            if (a == b) continue;
            const auto lo = coarse.col.begin() + coarse.row[a], hi = coarse.col.begin() + coarse.row[a + 1];
            fine.target[k] = static_cast<int32_t>(std::lower_bound(lo, hi, b) - coarse.col.begin());
        }
    return coarse;
}

// Gather lists that reproduce coarsen_values (heat_transport.cpp:157-171):
// the reference accumulates, for rows in ascending order, the row's
// diagonal into its aggregate's diagonal and then each entry into its
// target. A coarse value is therefore the left-to-right sum of its
// contributions in ascending (row, entry) order, starting from 0.0.
void galerkin_maps(HostLevel& fine, const HostLevel& coarse) {
    std::vector<std::vector<int32_t>> d(static_cast<size_t>(coarse.n)), o(static_cast<size_t>(coarse.nnz));
    for (int32_t r = 0; r < fine.n; ++r) {
        d[fine.aggregate[r]].push_back(-r - 1);
        for (int32_t k = fine.row[r]; k < fine.row[r + 1]; ++k) {
            if (fine.target[k] < 0)
                d[fine.aggregate[r]].push_back(k);
            else
                o[fine.target[k]].push_back(k);
        }
    }
    const auto flatten = [](const std::vector<std::vector<int32_t>>& lists, std::vector<int32_t>& start,
                            std::vector<int32_t>& flat) {
        start.assign(lists.size() + 1, 0);
        flat.clear();
        for (size_t i = 0; i < lists.size(); ++i) {
            flat.insert(flat.end(), lists[i].begin(), lists[i].end());
            start[i + 1] = static_cast<int32_t>(flat.size());
        }
    };
    flatten(d, fine.g_diag_start, fine.g_diag);
    flatten(o, fine.g_off_start, fine.g_off);
}

// Galerkin values on the host (unit-conductance hierarchy), same order.
void galerkin_values(const HostLevel& fine, HostLevel& coarse) {
    coarse.unit_diag.assign(static_cast<size_t>(coarse.n), 0.0);
    coarse.unit_off.assign(static_cast<size_t>(coarse.nnz), 0.0);
    for (int32_t c = 0; c < coarse.n; ++c) {
        double s = 0.0;
        for (int32_t q = fine.g_diag_start[c]; q < fine.g_diag_start[c + 1]; ++q) {
            const int32_t code = fine.g_diag[q];
            s += code < 0 ? fine.unit_diag[-code - 1] : fine.unit_off[code];
        }
        coarse.unit_diag[c] = s;
    }
    for (int32_t t = 0; t < coarse.nnz; ++t) {
        double s = 0.0;
        for (int32_t q = fine.g_off_start[t]; q < fine.g_off_start[t + 1]; ++q) s += fine.unit_off[fine.g_off[q]];
        coarse.unit_off[t] = s;
    }
}

// Aggregate-major row order and the LDS tiles of the restriction sweep:
// consecutive coarse rows packed while their fine rows fit in one tile.
void restriction_tiles(HostLevel& fine, int32_t coarse_n) {
    fine.agg_start.assign(static_cast<size_t>(coarse_n) + 1, 0);
    for (int32_t r = 0; r < fine.n; ++r) ++fine.agg_start[fine.aggregate[r] + 1];
    for (int32_t c = 0; c < coarse_n; ++c) fine.agg_start[c + 1] += fine.agg_start[c];
    fine.agg_rows.assign(static_cast<size_t>(fine.n), 0);
    std::vector<int32_t> fill(fine.agg_start.begin(), fine.agg_start.end() - 1);
    for (int32_t r = 0; r < fine.n; ++r) fine.agg_rows[fill[fine.aggregate[r]]++] = r;   // ascending
    fine.tile_start.assign(1, 0);
    int32_t tile_first_row = 0;
    for (int32_t c = 0; c < coarse_n; ++c) {
        const int32_t size = fine.agg_start[c + 1] - fine.agg_start[c];
        if (size > tile_rows) throw std::logic_error("aggregate larger than a restriction tile");
        if (fine.agg_start[c + 1] - tile_first_row > tile_rows) {
            fine.tile_start.push_back(c);
            tile_first_row = fine.agg_start[c];
        }
    }
    fine.tile_start.push_back(coarse_n);
    const size_t down_tiles = fine.tile_start.size() - 1;
    fine.down_rows.assign(down_tiles * static_cast<size_t>(tile_rows), -1);
    for (size_t t = 0; t < down_tiles; ++t) {
        const int32_t first = fine.agg_start[fine.tile_start[t]], last = fine.agg_start[fine.tile_start[t + 1]];
        for (int32_t q = first; q < last; ++q)
            fine.down_rows[t * static_cast<size_t>(tile_rows) + static_cast<size_t>(q - first)] = fine.agg_rows[q];
    }
}

// Slot-major ELL columns of a level (padding -1).
void ell_columns(HostLevel& lv) {
    lv.width = 0;
    for (int32_t i = 0; i < lv.n; ++i) lv.width = std::max(lv.width, lv.row[i + 1] - lv.row[i]);
    lv.col_ell.assign(static_cast<size_t>(lv.width) * static_cast<size_t>(lv.n), -1);
    for (int32_t i = 0; i < lv.n; ++i)
        for (int32_t k = lv.row[i]; k < lv.row[i + 1]; ++k) lv.col_ell[ell_position(lv, i, k)] = lv.col[k];
}

// The Galerkin lists of `fine` in ELL positions (fine entries) and indexed
// by the coarse level's ELL positions; same summation order.
void ell_galerkin(HostLevel& fine, const HostLevel& coarse) {
    std::vector<int32_t> fine_row(static_cast<size_t>(fine.nnz));
    for (int32_t i = 0; i < fine.n; ++i)
        for (int32_t k = fine.row[i]; k < fine.row[i + 1]; ++k) fine_row[k] = i;
    fine.d_diag.resize(fine.g_diag.size());
    for (size_t q = 0; q < fine.g_diag.size(); ++q) {
        const int32_t code = fine.g_diag[q];
        fine.d_diag[q] = code < 0 ? code : ell_position(fine, fine_row[code], code);
    }
    const size_t slots = static_cast<size_t>(coarse.width) * static_cast<size_t>(coarse.n);
    std::vector<std::vector<int32_t>> lists(slots);
    for (int32_t c = 0; c < coarse.n; ++c)
        for (int32_t t = coarse.row[c]; t < coarse.row[c + 1]; ++t) {
            auto& list = lists[static_cast<size_t>(ell_position(coarse, c, t))];
            for (int32_t q = fine.g_off_start[t]; q < fine.g_off_start[t + 1]; ++q) {
                const int32_t k = fine.g_off[q];
                list.push_back(ell_position(fine, fine_row[k], k));
            }
        }
    fine.d_off_start.assign(slots + 1, 0);
    fine.d_off.clear();
    for (size_t p = 0; p < slots; ++p) {
        fine.d_off.insert(fine.d_off.end(), lists[p].begin(), lists[p].end());
        fine.d_off_start[p + 1] = static_cast<int32_t>(fine.d_off.size());
    }
}

// Sorted unique aggregates of the given rows and of their neighbours.
std::vector<int32_t> read_aggregates(const HostLevel& lv, const std::vector<int32_t>& rows) {
    std::vector<int32_t> out;
    for (const int32_t i : rows) {
        out.push_back(lv.aggregate[i]);
        for (int32_t k = lv.row[i]; k < lv.row[i + 1]; ++k) out.push_back(lv.aggregate[lv.col[k]]);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

// The lists of the merged prolongation phase on `fine` (two levels above
// the coarsest), whose aggregates are the rows of `mid`. Left empty when a
// tile would read more mid rows than a tile holds.
void mid_lists(HostLevel& fine, const HostLevel& mid) {
    const int32_t tiles = (fine.n + tile_rows - 1) / tile_rows;
    std::vector<std::vector<int32_t>> mids(static_cast<size_t>(tiles)), coarse(static_cast<size_t>(tiles));
    int32_t s1 = 0, s2 = 0;
    for (int32_t t = 0; t < tiles; ++t) {
        std::vector<int32_t> rows;
        for (int32_t i = t * tile_rows; i < std::min(fine.n, (t + 1) * tile_rows); ++i) rows.push_back(i);
        mids[t] = read_aggregates(fine, rows);
        coarse[t] = read_aggregates(mid, mids[t]);
        s1 = std::max(s1, static_cast<int32_t>(mids[t].size()));
        s2 = std::max(s2, static_cast<int32_t>(coarse[t].size()));
    }
    if (s1 > tile_rows) return;
    fine.mid_stride = s1;
    fine.mid_coarse_stride = s2;
    fine.mid_pad.assign(static_cast<size_t>(tiles) * static_cast<size_t>(s1), -1);
    fine.mid_coarse_pad.assign(static_cast<size_t>(tiles) * static_cast<size_t>(s2), -1);
    fine.agg_loc.assign(static_cast<size_t>(fine.n), -1);
    fine.agg_col_loc.assign(fine.col_ell.size(), -1);
    for (int32_t t = 0; t < tiles; ++t) {
        const auto& m = mids[t];
        std::copy(m.begin(), m.end(), fine.mid_pad.begin() + static_cast<std::ptrdiff_t>(t) * s1);
        std::copy(coarse[t].begin(), coarse[t].end(), fine.mid_coarse_pad.begin() + static_cast<std::ptrdiff_t>(t) * s2);
        const auto local = [&](int32_t a) {
            return static_cast<int32_t>(std::lower_bound(m.begin(), m.end(), a) - m.begin());
        };
        for (int32_t i = t * tile_rows; i < std::min(fine.n, (t + 1) * tile_rows); ++i) {
            fine.agg_loc[i] = local(fine.aggregate[i]);
            for (int32_t k = fine.row[i]; k < fine.row[i + 1]; ++k)
                fine.agg_col_loc[ell_position(fine, i, k)] = local(fine.aggregate[fine.col[k]]);
        }
    }
}

}  // namespace

uint64_t checked_graph_entries(uint64_t n, const uint64_t* offset) {
    if (offset[0] != 0) throw std::invalid_argument("graph offsets must start at 0");
    for (uint64_t i = 0; i < n; ++i)
        if (offset[i + 1] < offset[i]) throw std::invalid_argument("graph offsets must not decrease");
    if (offset[n] >= (1ULL << 31)) throw std::invalid_argument("graph has too many entries");
    return offset[n];
}

HostHierarchy build_hierarchy(const HostGraph& g) {
    if (static_cast<int32_t>(g.row.size()) != g.n + 1) throw std::invalid_argument("graph offsets do not match its size");
    HostHierarchy h;
    h.levels.emplace_back();
    HostLevel& fine = h.levels.back();
    fine.n = g.n;
    fine.nnz = g.row[g.n];
    fine.row = g.row;
    fine.col = g.col;
    fine.unit_diag = g.weight_sum;
    fine.unit_off.resize(g.weight.size());
    for (size_t k = 0; k < g.weight.size(); ++k) fine.unit_off[k] = -g.weight[k];
    // heat_transport.cpp:188-199: coarsen while above 400 rows and while
    // coarsening still shrinks the level.
    while (h.levels.back().n > coarsest_rows) {
        HostLevel coarse = coarsen(h.levels.back());
        HostLevel& top = h.levels.back();
        if (coarse.n >= top.n) {
            top.aggregate.clear();
            top.target.clear();
            break;
        }
        galerkin_maps(top, coarse);
        galerkin_values(top, coarse);
        restriction_tiles(top, coarse.n);
        h.levels.push_back(std::move(coarse));
    }
    for (HostLevel& lv : h.levels) ell_columns(lv);
    for (size_t j = 0; j + 1 < h.levels.size(); ++j) {
        HostLevel& lv = h.levels[j];
        lv.agg_ell.assign(lv.col_ell.size(), -1);
        for (size_t p = 0; p < lv.col_ell.size(); ++p)
            if (lv.col_ell[p] >= 0) lv.agg_ell[p] = lv.aggregate[static_cast<size_t>(lv.col_ell[p])];
    }
    for (size_t j = 0; j + 1 < h.levels.size(); ++j) ell_galerkin(h.levels[j], h.levels[j + 1]);
    if (h.levels.size() >= 2) {
        HostLevel& lv = h.levels[h.levels.size() - 2];
        const int32_t tiles = (lv.n + tile_rows - 1) / tile_rows;
        lv.need_start.assign(static_cast<size_t>(tiles) + 1, 0);
        lv.need.clear();
        for (int32_t t = 0; t < tiles; ++t) {
            std::vector<int32_t> rows;
            for (int32_t i = t * tile_rows; i < std::min(lv.n, (t + 1) * tile_rows); ++i) {
                rows.push_back(lv.aggregate[i]);
                for (int32_t k = lv.row[i]; k < lv.row[i + 1]; ++k) rows.push_back(lv.aggregate[lv.col[k]]);
            }
            std::sort(rows.begin(), rows.end());
            rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
            lv.need.insert(lv.need.end(), rows.begin(), rows.end());
            lv.need_start[t + 1] = static_cast<int32_t>(lv.need.size());
        }
        lv.need_stride = 0;
        for (int32_t t = 0; t < tiles; ++t) lv.need_stride = std::max(lv.need_stride, lv.need_start[t + 1] - lv.need_start[t]);
        lv.need_pad.assign(static_cast<size_t>(tiles) * static_cast<size_t>(lv.need_stride), -1);
        for (int32_t t = 0; t < tiles; ++t)
            for (int32_t q = lv.need_start[t]; q < lv.need_start[t + 1]; ++q)
                lv.need_pad[static_cast<size_t>(t) * static_cast<size_t>(lv.need_stride) +
                            static_cast<size_t>(q - lv.need_start[t])] = lv.need[q];
    }
    if (h.levels.size() >= 3) mid_lists(h.levels[h.levels.size() - 3], h.levels[h.levels.size() - 2]);
    return h;
}

}  // namespace psim
