#pragma once
// Every model compiled into the library, indexed by spice_model_id (spice_hip.h).
#include <spice/models/brunel.h>
#include <spice/models/brunel_3pop.h>
#include <spice/models/brunel_3pop_stdp.h>
#include <spice/models/brunel_with_plasticity.h>
#include <spice/models/synth.h>
#include <spice/models/vogels_abbott.h>
#include <spice_hip.h>

#include <stdexcept>
#include <type_traits>

namespace spice
{
template <typename Model>
struct model_id;
template <>
struct model_id<vogels_abbott> : std::integral_constant<int, SPICE_MODEL_VOGELS_ABBOTT>
{
};
template <>
struct model_id<brunel> : std::integral_constant<int, SPICE_MODEL_BRUNEL>
{
};
template <>
struct model_id<brunel_with_plasticity> : std::integral_constant<int, SPICE_MODEL_BRUNEL_WITH_PLASTICITY>
{
};
template <>
struct model_id<synth> : std::integral_constant<int, SPICE_MODEL_SYNTH>
{
};
template <>
struct model_id<brunel_3pop> : std::integral_constant<int, SPICE_MODEL_BRUNEL_3POP>
{
};
template <>
struct model_id<brunel_3pop_stdp> : std::integral_constant<int, SPICE_MODEL_BRUNEL_3POP_STDP>
{
};
template <>
struct model_id<brunel_3pop_stdp_deferred> : std::integral_constant<int, SPICE_MODEL_BRUNEL_3POP_STDP_DEFERRED>
{
};

// Calls f(Model{}) for the model with the given id; throws std::invalid_argument otherwise.
template <typename F>
decltype( auto ) dispatch_model( int id, F && f )
{
	switch( id )
	{
	case SPICE_MODEL_VOGELS_ABBOTT: return f( vogels_abbott{} );
	case SPICE_MODEL_BRUNEL: return f( brunel{} );
	case SPICE_MODEL_BRUNEL_WITH_PLASTICITY: return f( brunel_with_plasticity{} );
	case SPICE_MODEL_SYNTH: return f( synth{} );
	case SPICE_MODEL_BRUNEL_3POP: return f( brunel_3pop{} );
	case SPICE_MODEL_BRUNEL_3POP_STDP: return f( brunel_3pop_stdp{} );
	case SPICE_MODEL_BRUNEL_3POP_STDP_DEFERRED: return f( brunel_3pop_stdp_deferred{} );
	default: throw std::invalid_argument( "unknown model id" );
	}
}
} // namespace spice
