// Instantiates the gfx950 kernels and host driver for the model 'brunel_3pop'.
#include "net.h"

#include <spice/models/brunel_3pop.h>

namespace spice::hipx
{
net_iface * make_net_brunel_3pop( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel_3pop>( p, std::move( g ), s );
}
} // namespace spice::hipx
