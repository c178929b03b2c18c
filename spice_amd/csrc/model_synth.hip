// Instantiates the gfx950 kernels and host driver for the model 'synth'.
#include "net.h"

#include <spice/models/synth.h>

namespace spice::hipx
{
net_iface * make_net_synth( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::synth>( p, std::move( g ), s );
}
} // namespace spice::hipx
