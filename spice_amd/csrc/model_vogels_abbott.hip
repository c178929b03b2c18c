// Instantiates the gfx950 kernels and host drivers for the model 'vogels_abbott'.
#include "batch.h"
#include "net.h"

#include <spice/models/vogels_abbott.h>

namespace spice::hipx
{
net_iface * make_net_vogels_abbott( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::vogels_abbott>( p, std::move( g ), s );
}

batch_iface * make_batch_vogels_abbott( std::vector<net_params> const & ps, std::vector<csr_dev> && gs, hipStream_t s )
{
	return make_batch<::spice::vogels_abbott>( ps, std::move( gs ), s );
}

batch_iface * make_group_vogels_abbott( std::vector<net_iface *> const & nets, hipStream_t s )
{
	return make_group<::spice::vogels_abbott>( nets, s );
}
} // namespace spice::hipx
