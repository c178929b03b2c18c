// Instantiates the gfx950 kernels and host driver for the model 'vogels_abbott'.
#include "net.h"

#include <spice/models/vogels_abbott.h>

namespace spice::hipx
{
net_iface * make_net_vogels_abbott( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::vogels_abbott>( p, std::move( g ), s );
}
} // namespace spice::hipx
