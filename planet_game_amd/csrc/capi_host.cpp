// C ABI: host-only entry points (mesh, hierarchy view, workload, errors).
#include "internal.hpp"
#include "workload.hpp"

#include "../../include/psim.h"

#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>

struct psim_mesh {
    psim::HostMesh mesh;
};

struct psim_hierarchy {
    psim::HostHierarchy h;
};

struct psim_agglomeration {
    psim_mesh coarse;
    std::vector<uint32_t> group;
    std::vector<double> area;
};

namespace psim {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& message) { g_last_error = message; }

}  // namespace psim

using psim::guarded;

extern "C" {

const char* psim_last_error(void) { return psim::g_last_error.c_str(); }
int psim_version(void) { return 1; }

int psim_mesh_create(uint32_t subdivision, double radius_m, psim_mesh** out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("null output handle");
        auto m = std::make_unique<psim_mesh>();
        m->mesh = psim::build_icosphere(subdivision, radius_m);
        *out = m.release();
    });
}

void psim_mesh_destroy(psim_mesh* mesh) { delete mesh; }

int psim_mesh_counts(const psim_mesh* m, uint64_t* cells, uint64_t* edges, uint64_t* directed) {
    return guarded([&] {
        if (m == nullptr) throw std::invalid_argument("null mesh");
        if (cells) *cells = static_cast<uint64_t>(m->mesh.cells);
        if (edges) *edges = static_cast<uint64_t>(m->mesh.edges);
        if (directed) *directed = static_cast<uint64_t>(m->mesh.directed);
    });
}

int psim_mesh_cells(const psim_mesh* m, double* center, double* east, double* north, double* area,
                    uint32_t* edge_offset, uint8_t* edge_count) {
    return guarded([&] {
        if (m == nullptr) throw std::invalid_argument("null mesh");
        const auto& h = m->mesh;
        const size_t n = static_cast<size_t>(h.cells);
        if (center) std::memcpy(center, h.center.data(), 3 * n * sizeof(double));
        if (east) std::memcpy(east, h.east.data(), 3 * n * sizeof(double));
        if (north) std::memcpy(north, h.north.data(), 3 * n * sizeof(double));
        if (area) std::memcpy(area, h.area.data(), n * sizeof(double));
        if (edge_offset) std::memcpy(edge_offset, h.edge_offset.data(), n * sizeof(uint32_t));
        if (edge_count) std::memcpy(edge_count, h.edge_count.data(), n);
    });
}

int psim_mesh_cell_edges(const psim_mesh* m, uint32_t* neighbor, uint32_t* edge, float* ne, float* nn) {
    return guarded([&] {
        if (m == nullptr) throw std::invalid_argument("null mesh");
        const auto& h = m->mesh;
        const size_t d = static_cast<size_t>(h.directed);
        if (neighbor) std::memcpy(neighbor, h.neighbor.data(), d * 4);
        if (edge) std::memcpy(edge, h.edge.data(), d * 4);
        if (ne) std::memcpy(ne, h.normal_east.data(), d * 4);
        if (nn) std::memcpy(nn, h.normal_north.data(), d * 4);
    });
}

int psim_mesh_edges(const psim_mesh* m, uint32_t* first, uint32_t* second, double* length, double* distance) {
    return guarded([&] {
        if (m == nullptr) throw std::invalid_argument("null mesh");
        const auto& h = m->mesh;
        const size_t e = static_cast<size_t>(h.edges);
        if (first) std::memcpy(first, h.first.data(), e * 4);
        if (second) std::memcpy(second, h.second.data(), e * 4);
        if (length) std::memcpy(length, h.length.data(), e * 8);
        if (distance) std::memcpy(distance, h.distance.data(), e * 8);
    });
}

int psim_mesh_graph(const psim_mesh* m, double* area, uint64_t* offset, uint64_t* neighbour, double* weight,
                    double* weight_sum) {
    return guarded([&] {
        if (m == nullptr) throw std::invalid_argument("null mesh");
        const auto& h = m->mesh;
        const size_t n = static_cast<size_t>(h.cells), d = static_cast<size_t>(h.directed);
        if (area) std::memcpy(area, h.area.data(), n * 8);
        if (offset)
            for (size_t i = 0; i <= n; ++i) offset[i] = static_cast<uint64_t>(h.row[i]);
        if (neighbour)
            for (size_t k = 0; k < d; ++k) neighbour[k] = h.neighbor[k];
        if (weight) std::memcpy(weight, h.weight.data(), d * 8);
        if (weight_sum) std::memcpy(weight_sum, h.weight_sum.data(), n * 8);
    });
}

int psim_agglomeration_create(const psim_mesh* mesh, psim_agglomeration** out) {
    return guarded([&] {
        if (mesh == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        psim::HostAgglomeration a = psim::agglomerate(mesh->mesh);
        auto result = std::make_unique<psim_agglomeration>();
        result->coarse.mesh = std::move(a.coarse);
        result->group = std::move(a.group);
        result->area = std::move(a.area);
        *out = result.release();
    });
}

void psim_agglomeration_destroy(psim_agglomeration* a) { delete a; }

int psim_agglomeration_counts(const psim_agglomeration* a, uint64_t* cells, uint64_t* groups) {
    return guarded([&] {
        if (a == nullptr) throw std::invalid_argument("null agglomeration");
        if (cells) *cells = a->group.size();
        if (groups) *groups = a->area.size();
    });
}

int psim_agglomeration_groups(const psim_agglomeration* a, uint32_t* group_of_cell, double* group_area_m2) {
    return guarded([&] {
        if (a == nullptr) throw std::invalid_argument("null agglomeration");
        if (group_of_cell) std::memcpy(group_of_cell, a->group.data(), a->group.size() * 4);
        if (group_area_m2) std::memcpy(group_area_m2, a->area.data(), a->area.size() * 8);
    });
}

const psim_mesh* psim_agglomeration_mesh(const psim_agglomeration* a) { return a ? &a->coarse : nullptr; }

int psim_hierarchy_create(uint64_t n, const uint64_t* offset, const uint64_t* neighbour, const double* weight,
                          const double* weight_sum, psim_hierarchy** out) {
    return guarded([&] {
        if (out == nullptr || offset == nullptr || weight_sum == nullptr) throw std::invalid_argument("null argument");
        if (n >= (1ULL << 30)) throw std::invalid_argument("graph too large");
        const uint64_t nnz = psim::checked_graph_entries(n, offset);
        if (nnz > 0 && (neighbour == nullptr || weight == nullptr)) throw std::invalid_argument("null argument");
        psim::HostGraph g;
        g.n = static_cast<int32_t>(n);
        g.row.resize(n + 1);
        for (uint64_t i = 0; i <= n; ++i) g.row[i] = static_cast<int32_t>(offset[i]);
        g.col.resize(nnz);
        g.weight.assign(weight, weight + nnz);
        for (uint64_t k = 0; k < nnz; ++k) {
            if (neighbour[k] >= n) throw std::invalid_argument("graph neighbour out of range");
            g.col[k] = static_cast<int32_t>(neighbour[k]);
        }
        g.weight_sum.assign(weight_sum, weight_sum + n);
        auto h = std::make_unique<psim_hierarchy>();
        h->h = psim::build_hierarchy(g);
        *out = h.release();
    });
}

void psim_hierarchy_destroy(psim_hierarchy* h) { delete h; }

uint64_t psim_hierarchy_depth(const psim_hierarchy* h) { return h ? h->h.levels.size() : 0; }

int psim_hierarchy_level_counts(const psim_hierarchy* h, uint64_t level, uint64_t* size, uint64_t* nnz) {
    return guarded([&] {
        if (h == nullptr || level >= h->h.levels.size()) throw std::invalid_argument("no such level");
        const auto& lv = h->h.levels[level];
        if (size) *size = static_cast<uint64_t>(lv.n);
        if (nnz) *nnz = static_cast<uint64_t>(lv.nnz);
    });
}

int psim_hierarchy_level(const psim_hierarchy* h, uint64_t level, uint64_t* offset, uint64_t* column,
                         uint64_t* aggregate, uint64_t* target, double* diagonal, double* value) {
    return guarded([&] {
        if (h == nullptr || level >= h->h.levels.size()) throw std::invalid_argument("no such level");
        const auto& lv = h->h.levels[level];
        if (offset)
            for (size_t i = 0; i < lv.row.size(); ++i) offset[i] = static_cast<uint64_t>(lv.row[i]);
        if (column)
            for (size_t k = 0; k < lv.col.size(); ++k) column[k] = static_cast<uint64_t>(lv.col[k]);
        if (aggregate)
            for (size_t i = 0; i < lv.aggregate.size(); ++i) aggregate[i] = static_cast<uint64_t>(lv.aggregate[i]);
        if (target)
            for (size_t k = 0; k < lv.target.size(); ++k)
                target[k] = lv.target[k] < 0 ? ~0ULL : static_cast<uint64_t>(lv.target[k]);
        if (diagonal) std::memcpy(diagonal, lv.unit_diag.data(), lv.unit_diag.size() * 8);
        if (value) std::memcpy(value, lv.unit_off.data(), lv.unit_off.size() * 8);
    });
}

uint64_t psim_instance_seed(uint64_t base_seed, uint64_t index) { return psim::instance_seed(base_seed, index); }

int psim_heat_instance(const psim_mesh* m, uint64_t seed, double k_base, double sigma, double* conductance,
                       double* capacity_area, uint8_t* land, double* temperature) {
    return guarded([&] {
        if (m == nullptr || !conductance || !capacity_area || !land || !temperature)
            throw std::invalid_argument("null argument");
        psim::make_heat_instance(m->mesh, seed, k_base, sigma, conductance, capacity_area, land, temperature);
    });
}

int psim_heat_forcing(const psim_mesh* m, uint64_t seed, uint32_t step, double* forcing) {
    return guarded([&] {
        if (m == nullptr || forcing == nullptr) throw std::invalid_argument("null argument");
        psim::make_heat_forcing(m->mesh, seed, step, forcing);
    });
}

}  // extern "C"

// Accessor for the device half of the library.
namespace psim {
const HostMesh& host_mesh_of(const psim_mesh* m) { return m->mesh; }
}  // namespace psim
