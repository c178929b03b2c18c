// Instantiates the gfx950 kernels and host driver for the model 'brunel_with_plasticity'.
#include "net.h"

#include <spice/models/brunel_with_plasticity.h>

namespace spice::hipx
{
net_iface * make_net_brunel_with_plasticity( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel_with_plasticity>( p, std::move( g ), s );
}
} // namespace spice::hipx
