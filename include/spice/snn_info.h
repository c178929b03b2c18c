#pragma once
// Global network facts handed to every model functor (reference: spice/snn_info.h:8-14).
#include <spice/util/stdint.h>

namespace spice
{
struct snn_info
{
	int_ num_neurons = 0;
};
} // namespace spice
