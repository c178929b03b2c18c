// Instantiates the gfx950 kernels and host drivers for the model 'brunel_3pop_stdp'.
#include "batch.h"
#include "net.h"

#include <spice/models/brunel_3pop_stdp.h>

namespace spice::hipx
{
net_iface * make_net_brunel_3pop_stdp( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel_3pop_stdp>( p, std::move( g ), s );
}

batch_iface * make_batch_brunel_3pop_stdp( std::vector<net_params> const &, std::vector<csr_dev> &&, hipStream_t )
{
	throw std::invalid_argument( "batches hold static (non-plastic) models only" );
}
} // namespace spice::hipx
