// Instantiates the gfx950 kernels and host drivers for the model 'brunel_3pop_stdp_deferred'.
#include "batch.h"
#include "net.h"

#include <spice/models/brunel_3pop_stdp.h>

namespace spice::hipx
{
net_iface * make_net_brunel_3pop_stdp_deferred( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel_3pop_stdp_deferred>( p, std::move( g ), s );
}

batch_iface * make_batch_brunel_3pop_stdp_deferred( std::vector<net_params> const & ps, std::vector<csr_dev> && gs, hipStream_t s )
{
	return make_batch<::spice::brunel_3pop_stdp_deferred>( ps, std::move( gs ), s );
}

batch_iface * make_group_brunel_3pop_stdp_deferred( std::vector<net_iface *> const & nets, hipStream_t s )
{
	return make_group<::spice::brunel_3pop_stdp_deferred>( nets, s );
}
} // namespace spice::hipx
