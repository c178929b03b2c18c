// Instantiates the gfx950 kernels and host drivers for the model 'synth'.
#include "batch.h"
#include "net.h"

#include <spice/models/synth.h>

namespace spice::hipx
{
net_iface * make_net_synth( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::synth>( p, std::move( g ), s );
}

batch_iface * make_batch_synth( std::vector<net_params> const & ps, std::vector<csr_dev> && gs, hipStream_t s )
{
	return make_batch<::spice::synth>( ps, std::move( gs ), s );
}

batch_iface * make_group_synth( std::vector<net_iface *> const & nets, hipStream_t s )
{
	return make_group<::spice::synth>( nets, s );
}
} // namespace spice::hipx
