// Seeded heat-step workload generator (host).
#pragma once

#include "internal.hpp"

#include <cstdint>

namespace psim {

uint64_t instance_seed(uint64_t base, uint64_t index);
void make_heat_instance(const HostMesh& mesh, uint64_t seed, double k_base, double sigma, double* conductance,
                        double* capacity_area, uint8_t* land, double* temperature);
void make_heat_forcing(const HostMesh& mesh, uint64_t seed, uint32_t step, double* forcing);

}  // namespace psim
