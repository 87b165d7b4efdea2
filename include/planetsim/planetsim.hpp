// Everything of the drop-in layer: fields, mesh, FV operators, transport
// graph, GraphMultigrid, diffusion source, the implicit heat step and
// transports, and IncompleteLU.
#pragma once

#include "field.hpp"
#include "finite_volume.hpp"
#include "heat_transport.hpp"
#include "incomplete_lu.hpp"
#include "planet_mesh.hpp"
