// Everything of the drop-in layer: fields, mesh, FV operators, transport
// graph, GraphMultigrid, diffusion source and the implicit heat step.
#pragma once

#include "field.hpp"
#include "finite_volume.hpp"
#include "heat_transport.hpp"
#include "planet_mesh.hpp"
