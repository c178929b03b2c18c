// Instantiates the gfx950 kernels and host driver for the model 'brunel'.
#include "net.h"

#include <spice/models/brunel.h>

namespace spice::hipx
{
net_iface * make_net_brunel( net_params const & p, csr_dev && g, hipStream_t s )
{
	return new hip_net<::spice::brunel>( p, std::move( g ), s );
}
} // namespace spice::hipx
