// Instantiates the gfx950 kernels and host drivers for the model 'brunel_with_plasticity'.
#include "batch.h"
#include "net.h"

#include <spice/models/brunel_with_plasticity.h>

namespace spice::hipx
{
net_iface * make_net_brunel_with_plasticity( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel_with_plasticity>( p, std::move( g ), s );
}

batch_iface * make_batch_brunel_with_plasticity( std::vector<net_params> const &, std::vector<csr_dev> &&, hipStream_t )
{
	throw std::invalid_argument( "batches hold static (non-plastic) models only" );
}
} // namespace spice::hipx
