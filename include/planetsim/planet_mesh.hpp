// PlanetMesh and make_icosphere with the reference's interface
// (sim/planet/mesh/planet_mesh.hpp:15-125, cell_id.hpp:9-31,
// sim/core/math/vec3d.hpp:8-52), built by libpsim's host mesh generator
// (bit-identical to the reference's) and uploaded to the device on the
// first operator call.
#pragma once

#include "device.hpp"

#include <algorithm>
#include <compare>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <mutex>
#include <span>
#include <vector>

namespace planetsim {

struct Vec3d {
    double x = 0.0;
    double y = 0.0;
    double z = 0.0;
};

class CellId {
  public:
    using value_type = std::uint32_t;
    constexpr CellId() noexcept = default;
    explicit constexpr CellId(value_type v) noexcept : value_(v) {}
    [[nodiscard]] static constexpr CellId invalid() noexcept { return CellId{std::numeric_limits<value_type>::max()}; }
    [[nodiscard]] constexpr value_type value() const noexcept { return value_; }
    [[nodiscard]] constexpr std::size_t to_index() const noexcept { return value_; }
    [[nodiscard]] constexpr bool is_valid() const noexcept { return *this != invalid(); }
    friend constexpr bool operator==(CellId, CellId) noexcept = default;
    friend constexpr auto operator<=>(CellId, CellId) noexcept = default;

  private:
    value_type value_ = std::numeric_limits<value_type>::max();
};

class EdgeId {
  public:
    using value_type = std::uint32_t;
    constexpr EdgeId() noexcept = default;
    explicit constexpr EdgeId(value_type v) noexcept : value_(v) {}
    [[nodiscard]] static constexpr EdgeId invalid() noexcept { return EdgeId{static_cast<value_type>(-1)}; }
    [[nodiscard]] constexpr value_type value() const noexcept { return value_; }
    [[nodiscard]] constexpr std::size_t to_index() const noexcept { return value_; }
    [[nodiscard]] constexpr bool is_valid() const noexcept { return *this != invalid(); }
    friend constexpr bool operator==(EdgeId, EdgeId) noexcept = default;
    friend constexpr auto operator<=>(EdgeId, EdgeId) noexcept = default;

  private:
    value_type value_ = static_cast<value_type>(-1);
};

struct CellGeometry {
    Vec3d center_unit;
    Vec3d east_unit;
    Vec3d north_unit;
    double area_m2 = 0.0;
    CellId id;
    std::uint32_t source_vertex_index = 0;   // not exported by libpsim; 0
    std::uint32_t edge_offset = 0;
    std::uint8_t edge_count = 0;
    [[nodiscard]] bool is_pentagon() const noexcept { return edge_count == 5U; }
};

struct CellEdgeGeometry {
    CellId neighbor;
    EdgeId edge;
    float outward_normal_east = 0.0F;
    float outward_normal_north = 0.0F;
};

struct EdgeGeometry {
    CellId first_cell;
    CellId second_cell;
    double length_m = 0.0;
    double centroid_distance_m = 0.0;
};

struct CellBlock {
    std::uint32_t begin = 0;
    std::uint32_t end = 0;
};

inline constexpr std::size_t deterministic_cell_block_size = 256U;

class PlanetMesh;
[[nodiscard]] PlanetMesh make_icosphere(std::uint32_t subdivision, double radius_m);
namespace detail {
// A PlanetMesh copied out of a libpsim mesh handle.
[[nodiscard]] PlanetMesh mesh_from_handle(const psim_mesh* raw, std::uint32_t subdivision, double radius_m);
}  // namespace detail

class PlanetMesh {
  public:
    [[nodiscard]] std::uint32_t subdivision() const noexcept { return subdivision_; }
    [[nodiscard]] double radius_m() const noexcept { return radius_m_; }
    [[nodiscard]] std::size_t cell_count() const noexcept { return cells_.size(); }
    [[nodiscard]] std::size_t edge_count() const noexcept { return edges_.size(); }
    [[nodiscard]] std::size_t directed_edge_count() const noexcept { return cell_edges_.size(); }
    [[nodiscard]] const CellGeometry& cell(CellId id) const { return cells_.at(id.to_index()); }
    [[nodiscard]] const EdgeGeometry& edge(EdgeId id) const { return edges_.at(id.to_index()); }
    [[nodiscard]] std::span<const CellGeometry> cells() const noexcept { return cells_; }
    [[nodiscard]] std::span<const EdgeGeometry> edges() const noexcept { return edges_; }
    [[nodiscard]] std::span<const CellBlock> blocks() const noexcept { return blocks_; }
    [[nodiscard]] std::span<const CellEdgeGeometry> cell_edges(CellId id) const {
        const CellGeometry& c = cell(id);
        return {cell_edges_.data() + c.edge_offset, c.edge_count};
    }

    // The device copy used by the FV operators (uploaded once, shared by
    // copies of this mesh).
    [[nodiscard]] psim_fv* device() const {
        std::call_once(device_->once, [this] {
            const std::size_t n = cells_.size(), e = edges_.size(), d = cell_edges_.size();
            std::vector<double> area(n), length(e), distance(e);
            std::vector<std::uint32_t> offset(n), neighbor(d), edge(d), first(e), second(e);
            std::vector<std::uint8_t> count(n);
            std::vector<float> ne(d), nn(d);
            for (std::size_t c = 0; c < n; ++c) {
                area[c] = cells_[c].area_m2;
                offset[c] = cells_[c].edge_offset;
                count[c] = cells_[c].edge_count;
            }
            for (std::size_t k = 0; k < d; ++k) {
                neighbor[k] = cell_edges_[k].neighbor.value();
                edge[k] = cell_edges_[k].edge.value();
                ne[k] = cell_edges_[k].outward_normal_east;
                nn[k] = cell_edges_[k].outward_normal_north;
            }
            for (std::size_t k = 0; k < e; ++k) {
                first[k] = edges_[k].first_cell.value();
                second[k] = edges_[k].second_cell.value();
                length[k] = edges_[k].length_m;
                distance[k] = edges_[k].centroid_distance_m;
            }
            psim_fv* fv = nullptr;
            gpu::check(psim_fv_create_from_arrays(gpu::context(), n, e, area.data(), offset.data(), count.data(),
                                                  neighbor.data(), edge.data(), ne.data(), nn.data(), first.data(),
                                                  second.data(), length.data(), distance.data(), &fv));
            device_->fv.reset(fv);
        });
        return device_->fv.get();
    }

  private:
    friend PlanetMesh detail::mesh_from_handle(const psim_mesh* raw, std::uint32_t subdivision, double radius_m);
    struct DeviceCopy {
        std::once_flag once;
        struct Free {
            void operator()(psim_fv* f) const noexcept { psim_fv_destroy(f); }
        };
        std::unique_ptr<psim_fv, Free> fv;
    };
    std::uint32_t subdivision_ = 0;
    double radius_m_ = 0.0;
    std::vector<CellGeometry> cells_;
    std::vector<CellEdgeGeometry> cell_edges_;
    std::vector<EdgeGeometry> edges_;
    std::vector<CellBlock> blocks_;
    std::shared_ptr<DeviceCopy> device_ = std::make_shared<DeviceCopy>();
};

inline PlanetMesh make_icosphere(std::uint32_t subdivision, double radius_m) {
    psim_mesh* raw = nullptr;
    gpu::check(psim_mesh_create(subdivision, radius_m, &raw));
    std::unique_ptr<psim_mesh, void (*)(psim_mesh*)> handle(raw, psim_mesh_destroy);
    return detail::mesh_from_handle(raw, subdivision, radius_m);
}

inline PlanetMesh detail::mesh_from_handle(const psim_mesh* raw, std::uint32_t subdivision, double radius_m) {
    std::uint64_t n = 0, e = 0, d = 0;
    gpu::check(psim_mesh_counts(raw, &n, &e, &d));
    std::vector<double> center(3 * n), east(3 * n), north(3 * n), area(n), length(e), distance(e);
    std::vector<std::uint32_t> offset(n), neighbor(d), edge(d), first(e), second(e);
    std::vector<std::uint8_t> count(n);
    std::vector<float> ne(d), nn(d);
    gpu::check(psim_mesh_cells(raw, center.data(), east.data(), north.data(), area.data(), offset.data(), count.data()));
    gpu::check(psim_mesh_cell_edges(raw, neighbor.data(), edge.data(), ne.data(), nn.data()));
    gpu::check(psim_mesh_edges(raw, first.data(), second.data(), length.data(), distance.data()));
    PlanetMesh mesh;
    mesh.subdivision_ = subdivision;
    mesh.radius_m_ = radius_m;
    mesh.cells_.resize(n);
    for (std::size_t c = 0; c < n; ++c) {
        CellGeometry& g = mesh.cells_[c];
        g.center_unit = {center[3 * c], center[3 * c + 1], center[3 * c + 2]};
        g.east_unit = {east[3 * c], east[3 * c + 1], east[3 * c + 2]};
        g.north_unit = {north[3 * c], north[3 * c + 1], north[3 * c + 2]};
        g.area_m2 = area[c];
        g.id = CellId{static_cast<std::uint32_t>(c)};
        g.edge_offset = offset[c];
        g.edge_count = count[c];
    }
    mesh.cell_edges_.resize(d);
    for (std::size_t k = 0; k < d; ++k) mesh.cell_edges_[k] = {CellId{neighbor[k]}, EdgeId{edge[k]}, ne[k], nn[k]};
    mesh.edges_.resize(e);
    for (std::size_t k = 0; k < e; ++k) mesh.edges_[k] = {CellId{first[k]}, CellId{second[k]}, length[k], distance[k]};
    for (std::size_t begin = 0; begin < n; begin += deterministic_cell_block_size) {
        const std::size_t end = std::min(begin + deterministic_cell_block_size, static_cast<std::size_t>(n));
        mesh.blocks_.push_back({static_cast<std::uint32_t>(begin), static_cast<std::uint32_t>(end)});
    }
    return mesh;
}

}  // namespace planetsim
