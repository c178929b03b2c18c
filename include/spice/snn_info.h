#pragma once
// Global network facts handed to every model functor (reference: spice/snn_info.h:8-14).
#include <spice/util/stdint.h>

namespace spice
{
struct snn_info
{
	int_ num_neurons = 0;
	// Optional per-network model parameters (0 = model default), e.g. the Poisson rate of a
	// parameter sweep. Not in the reference (which has only num_neurons); models that do not
	// read them behave exactly as there.
	float param[4] = { 0, 0, 0, 0 };
};
} // namespace spice
