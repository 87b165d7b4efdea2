// Finite-volume operators with the reference's signatures
// (sim/planet/operators/finite_volume.hpp:21-44), computed on the device by
// libpsim and bit-identical to the reference for float and double.
#pragma once

#include "field.hpp"
#include "planet_mesh.hpp"

#include <stdexcept>
#include <type_traits>

namespace planetsim {

namespace detail {
template <typename T> constexpr int dtype_of() {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "operators exist for float and double");
    return std::is_same_v<T, double> ? PSIM_F64 : PSIM_F32;
}
// finite_volume.cpp:10-19
template <typename F> void require_size(const F& field, std::size_t expected, const char* message) {
    if (field.size() != expected) throw std::invalid_argument(message);
}
}  // namespace detail

// +1 when `cell` is the edge's first cell (flux leaves it), -1 otherwise
// (finite_volume.cpp:23-32).
[[nodiscard]] inline double edge_orientation(const PlanetMesh& mesh, CellId cell, EdgeId edge) {
    const EdgeGeometry& g = mesh.edge(edge);
    if (g.first_cell == cell) return 1.0;
    if (g.second_cell == cell) return -1.0;
    throw std::invalid_argument("cell is not incident to edge");
}

template <typename T>
void divergence(const PlanetMesh& mesh, const EdgeField<T>& edge_normal_flux, Field2D<T>& result,
                std::size_t worker_count = 1U) {
    detail::require_size(edge_normal_flux, mesh.edge_count(), "edge field size does not match the mesh");
    detail::require_size(result, mesh.cell_count(), "cell field size does not match the mesh");
    gpu::check_workers(worker_count);
    gpu::check(psim_divergence(mesh.device(), detail::dtype_of<T>(), 1, edge_normal_flux.values().data(),
                               result.values().data(), PSIM_HOST));
}

template <typename T>
void gradient(const PlanetMesh& mesh, const Field2D<T>& scalar, Field2D<T>& east, Field2D<T>& north,
              std::size_t worker_count = 1U) {
    detail::require_size(scalar, mesh.cell_count(), "cell field size does not match the mesh");
    detail::require_size(east, mesh.cell_count(), "cell field size does not match the mesh");
    detail::require_size(north, mesh.cell_count(), "cell field size does not match the mesh");
    gpu::check_workers(worker_count);
    gpu::check(psim_gradient(mesh.device(), detail::dtype_of<T>(), 1, scalar.values().data(), east.values().data(),
                             north.values().data(), PSIM_HOST));
}

template <typename T>
void laplacian(const PlanetMesh& mesh, const Field2D<T>& scalar, Field2D<T>& result, std::size_t worker_count = 1U) {
    detail::require_size(scalar, mesh.cell_count(), "cell field size does not match the mesh");
    detail::require_size(result, mesh.cell_count(), "cell field size does not match the mesh");
    gpu::check_workers(worker_count);
    gpu::check(psim_laplacian(mesh.device(), detail::dtype_of<T>(), 1, scalar.values().data(), result.values().data(),
                              PSIM_HOST));
}

}  // namespace planetsim
