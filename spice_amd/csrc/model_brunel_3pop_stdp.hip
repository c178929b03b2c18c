// Instantiates the gfx950 kernels and host driver for the model 'brunel_3pop_stdp'.
#include "net.h"

#include <spice/models/brunel_3pop_stdp.h>

namespace spice::hipx
{
net_iface * make_net_brunel_3pop_stdp( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel_3pop_stdp>( p, std::move( g ), s );
}
} // namespace spice::hipx
