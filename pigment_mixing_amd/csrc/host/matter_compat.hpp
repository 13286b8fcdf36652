// matter_compat.hpp — include in place of the reference's
// "simulation/simulation.hpp" (matter/src/simulation/simulation.hpp): the
// facade's Simulation and the types a caller of the reference uses, at global
// scope where the reference declares them.
#pragma once
#include "simulation.hpp"

using matter_amd::BC;
using matter_amd::FrameData;
using matter_amd::Latent;
using matter_amd::ObjectPlate;
using matter_amd::ObjectSpatula;
using matter_amd::Particles;
using matter_amd::PlasticModel;
using matter_amd::PlateType;
using matter_amd::Simulation;
using matter_amd::T;
using matter_amd::TM;
using matter_amd::TV;
