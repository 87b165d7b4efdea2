// Host-side data structures of libpsim (not part of the C ABI).
#pragma once

#include "../../include/psim.h"

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace psim {

// Error text for psim_last_error(); set by every failing entry point.
void set_error(const std::string& message);

// Failures raised by the device half of the library.
struct DeviceFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceOutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NotConverged : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Maps exceptions to status codes (the C++ wrappers map them back).
template <class F> int guarded(F&& body) {
    try {
        body();
        return PSIM_OK;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return PSIM_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        set_error(e.what());
        return PSIM_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        set_error(e.what());
        return PSIM_INVALID_ARGUMENT;
    } catch (const std::logic_error& e) {
        set_error(e.what());
        return PSIM_LOGIC_ERROR;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return PSIM_OUT_OF_MEMORY;
    } catch (const DeviceOutOfMemory& e) {
        set_error(e.what());
        return PSIM_OUT_OF_MEMORY;
    } catch (const DeviceFailure& e) {
        set_error(e.what());
        return PSIM_DEVICE_ERROR;
    } catch (const NotConverged& e) {
        set_error(e.what());
        return PSIM_NOT_CONVERGED;
    } catch (const std::exception& e) {
        set_error(e.what());
        return PSIM_RUNTIME_ERROR;
    }
}

// The icosahedral Voronoi mesh in structure-of-arrays form, bit-identical to
// the reference's make_icosphere (sim/planet/mesh/icosphere.cpp:200-419),
// plus its transport graph (sim/planet/surface/heat_transport.cpp:433-452).
struct HostMesh {
    uint32_t level = 0;
    double radius = 0.0;
    int32_t cells = 0, edges = 0, directed = 0;
    std::vector<double> center, east, north;   // 3 per cell (x, y, z)
    std::vector<double> area;
    std::vector<uint32_t> edge_offset;
    std::vector<uint8_t> edge_count;
    std::vector<uint32_t> neighbor, edge;      // per directed (cell-edge) entry
    std::vector<float> normal_east, normal_north;
    std::vector<uint32_t> first, second;       // per edge
    std::vector<double> length, distance;
    // transport graph
    std::vector<int32_t> row;                  // cells + 1 CSR offsets
    std::vector<double> weight, weight_sum;
};

HostMesh build_icosphere(uint32_t level, double radius);

// A transport graph as the reference's TransportGraph holds it.
struct HostGraph {
    int32_t n = 0;
    std::vector<double> area;
    std::vector<int32_t> row;      // n + 1
    std::vector<int32_t> col;      // nnz
    std::vector<double> weight;    // nnz, l / d
    std::vector<double> weight_sum;
};

// One multigrid level's structure (heat_transport.cpp:54-67) and the maps
// the device kernels use: aggregate-major row order with LDS tiles for the
// fused residual/restriction sweep, and Galerkin gather lists that
// reproduce coarsen_values' summation order (heat_transport.cpp:157-171).
struct HostLevel {
    int32_t n = 0, nnz = 0;
    std::vector<int32_t> row, col;              // off-diagonal CSR
    std::vector<double> unit_diag, unit_off;    // unit-conductance Galerkin values
    // Everything below is empty on the coarsest level.
    std::vector<int32_t> aggregate;             // n: row -> coarse row
    std::vector<int32_t> target;                // nnz: coarse entry, or -1 = coarse diagonal
    std::vector<int32_t> agg_start, agg_rows;   // coarse CSR over fine rows (ascending)
    std::vector<int32_t> tile_start;            // coarse-row ranges packed into <= 256 fine rows
    // Galerkin gather lists of the NEXT level's values, in reference order:
    // for coarse diagonal c: codes g_diag[g_diag_start[c] ..) where code >= 0
    // is a fine off-diagonal entry and code < 0 is fine diagonal row -code-1;
    // for coarse entry t: fine entries g_off[g_off_start[t] ..).
    std::vector<int32_t> g_diag_start, g_diag;
    std::vector<int32_t> g_off_start, g_off;
    // Device layout: off-diagonals in slot-major ELL (entry k of row i, the
    // s-th of that row, lives at s * n + i; padding has column -1), and the
    // Galerkin lists with fine entries named by ELL position: the diagonal
    // lists keep the diagonal codes, the off-diagonal lists are indexed by
    // the NEXT level's ELL positions (empty for padding).
    int32_t width = 0;
    std::vector<int32_t> col_ell;
    std::vector<int32_t> agg_ell;                // aggregate of each ELL slot's column (-1 padding)
    std::vector<int32_t> d_diag;                 // g_diag in ELL positions
    std::vector<int32_t> d_off_start, d_off;     // per next-level ELL position
    // On the level just above the coarsest: for each 256-row tile, the
    // coarse rows its prolongation reads (aggregates of its rows and of
    // their neighbours), sorted.
    std::vector<int32_t> need_start, need;
    // The same lists padded to a common stride (-1 padding), so that a
    // tile's row ids are one load away: tile t's list at need_pad[t * need_stride ..).
    int32_t need_stride = 0;
    std::vector<int32_t> need_pad;
    // Rows of each restriction tile in aggregate-major order, padded to the
    // tile size (-1 padding): down tile t's rows at down_rows[t * tile_rows ..).
    std::vector<int32_t> down_rows;
    // On the level two above the coarsest, for the phase that prolongs from
    // the coarsest level through the next one in one sweep: per 256-row tile
    // the next level's rows its rows read (aggregates of its rows and of
    // their neighbours, sorted; mid_pad, stride mid_stride) and the coarsest
    // rows those read (mid_coarse_pad, stride mid_coarse_stride), both
    // padded with -1; and for every row (agg_loc) and ELL slot
    // (agg_col_loc) the position of its aggregate in its tile's list.
    int32_t mid_stride = 0, mid_coarse_stride = 0;
    std::vector<int32_t> mid_pad, mid_coarse_pad, agg_loc, agg_col_loc;
};

// ELL position of CSR entry k of row i.
inline int32_t ell_position(const HostLevel& lv, int32_t i, int32_t k) { return (k - lv.row[i]) * lv.n + i; }

struct HostHierarchy {
    std::vector<HostLevel> levels;
};

// heat_transport.cpp:175-200, on the graph's unit-conductance weights.
HostHierarchy build_hierarchy(const HostGraph& graph);

// A caller's CSR offsets of n rows: offset[0] == 0, never decreasing, and
// fewer than 2^31 entries in all (rows and entries are int32 on the
// device). Throws std::invalid_argument otherwise; returns the entry count.
uint64_t checked_graph_entries(uint64_t n, const uint64_t* offset);

HostGraph graph_of(const HostMesh& mesh);

// The agglomerated transport graph's groups (heat_transport.cpp:467-512):
// the mesh one level coarser, each fine cell's group (a coarse cell) and
// each group's area, the sum of its fine cells' areas in cell order. At
// subdivision 0 the groups are the cells themselves.
struct HostAgglomeration {
    HostMesh coarse;
    std::vector<uint32_t> group;   // per fine cell
    std::vector<double> area;      // per group
};
HostAgglomeration agglomerate(const HostMesh& fine);

constexpr int32_t tile_rows = 256;     // the reference's cell block size
constexpr int32_t coarsest_rows = 400;  // heat_transport.cpp:188

}  // namespace psim

struct psim_mesh;
namespace psim {
const HostMesh& host_mesh_of(const ::psim_mesh* m);
}  // namespace psim
