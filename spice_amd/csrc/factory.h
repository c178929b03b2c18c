#pragma once
// Model dispatch for the C ABI and the server: per-model factories (one translation unit per
// model instantiates its kernels) and the layout -> generator plumbing shared by them.
#include "batch.h"
#include "graph.h"
#include "net.h"

#include <spice/util/layout.h>
#include <spice_hip.h>

#include <cmath>
#include <tuple>
#include <vector>

namespace spice::hipx
{
#define SPICE_MODEL_FACTORIES( m )                                                                            \
	net_iface * make_net_##m( net_params const &, csr_dev &&, hipStream_t );                                  \
	batch_iface * make_batch_##m( std::vector<net_params> const &, std::vector<csr_dev> &&, hipStream_t );    \
	batch_iface * make_group_##m( std::vector<net_iface *> const &, hipStream_t );
SPICE_MODEL_FACTORIES( vogels_abbott )
SPICE_MODEL_FACTORIES( brunel )
SPICE_MODEL_FACTORIES( brunel_with_plasticity )
SPICE_MODEL_FACTORIES( synth )
SPICE_MODEL_FACTORIES( brunel_3pop )
SPICE_MODEL_FACTORIES( brunel_3pop_stdp )
SPICE_MODEL_FACTORIES( brunel_3pop_stdp_deferred )
#undef SPICE_MODEL_FACTORIES

// Graph generation into caller-owned device memory behind spice_hip_generate_graph* (capi.hip).
int generate_graph_entry( void * ( *scratch )(), uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns,
                          uint64_t seed, int64_t * d_row_ptr, int32_t * d_col, uint64_t col_cap, uint64_t * out_nnz, uint64_t * out_width );

inline void require( bool c, char const * msg )
{
	if( !c ) throw std::invalid_argument( msg );
}

inline util::layout make_layout( uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns )
{
	require( npops >= 0 && nconns >= 0 && ( pops || npops == 0 ) && ( conns || nconns == 0 ), "invalid layout arguments" );
	std::vector<size_t> p( pops, pops + npops );
	std::vector<std::tuple<size_t, size_t, float>> c;
	for( int32_t i = 0; i < nconns; i++ ) c.emplace_back( conns[i].src_pop, conns[i].dst_pop, conns[i].p );
	return util::layout( p, c );
}

inline std::vector<util::gen::conn> gen_conns( util::layout const & l )
{
	std::vector<util::gen::conn> out;
	for( auto const & c : l.connections() )
	{
		util::gen::conn g;
		g.src_first = std::get<0>( c );
		g.src_last = std::get<1>( c );
		g.dst_first = std::get<2>( c );
		g.dst_last = std::get<3>( c );
		double const range = double( g.dst_last - g.dst_first );
		double const p = std::get<4>( c );
		g.mean = range * p;
		g.sd = std::sqrt( range * p * ( 1.0 - p ) );
		out.push_back( g );
	}
	return out;
}

inline hipStream_t new_stream()
{
	hipStream_t s = nullptr;
	SPICE_HIP( hipStreamCreateWithFlags( &s, hipStreamNonBlocking ) );
	return s;
}

// The net takes ownership of `s` once constructed; on failure `s` is destroyed here.
inline net_iface * make_net( net_params const & p, csr_dev && g, hipStream_t s )
{
	try
	{
		switch( p.model )
		{
		case SPICE_MODEL_VOGELS_ABBOTT: return make_net_vogels_abbott( p, std::move( g ), s );
		case SPICE_MODEL_BRUNEL: return make_net_brunel( p, std::move( g ), s );
		case SPICE_MODEL_BRUNEL_WITH_PLASTICITY: return make_net_brunel_with_plasticity( p, std::move( g ), s );
		case SPICE_MODEL_SYNTH: return make_net_synth( p, std::move( g ), s );
		case SPICE_MODEL_BRUNEL_3POP: return make_net_brunel_3pop( p, std::move( g ), s );
		case SPICE_MODEL_BRUNEL_3POP_STDP: return make_net_brunel_3pop_stdp( p, std::move( g ), s );
		case SPICE_MODEL_BRUNEL_3POP_STDP_DEFERRED: return make_net_brunel_3pop_stdp_deferred( p, std::move( g ), s );
		default: throw std::invalid_argument( "unknown model id" );
		}
	}
	catch( ... )
	{
		(void)hipStreamDestroy( s );
		throw;
	}
}

// The batch takes ownership of `s` once constructed; on failure `s` is destroyed here.
inline batch_iface * make_batch( int model, std::vector<net_params> const & ps, std::vector<csr_dev> && gs, hipStream_t s )
{
	try
	{
		switch( model )
		{
		case SPICE_MODEL_VOGELS_ABBOTT: return make_batch_vogels_abbott( ps, std::move( gs ), s );
		case SPICE_MODEL_BRUNEL: return make_batch_brunel( ps, std::move( gs ), s );
		case SPICE_MODEL_BRUNEL_WITH_PLASTICITY: return make_batch_brunel_with_plasticity( ps, std::move( gs ), s );
		case SPICE_MODEL_SYNTH: return make_batch_synth( ps, std::move( gs ), s );
		case SPICE_MODEL_BRUNEL_3POP: return make_batch_brunel_3pop( ps, std::move( gs ), s );
		case SPICE_MODEL_BRUNEL_3POP_STDP: return make_batch_brunel_3pop_stdp( ps, std::move( gs ), s );
		case SPICE_MODEL_BRUNEL_3POP_STDP_DEFERRED: return make_batch_brunel_3pop_stdp_deferred( ps, std::move( gs ), s );
		default: throw std::invalid_argument( "unknown model id" );
		}
	}
	catch( ... )
	{
		(void)hipStreamDestroy( s );
		throw;
	}
}

// A batch over nets owned elsewhere, all of model `model`; `s` stays the caller's.
inline batch_iface * make_group( int model, std::vector<net_iface *> const & nets, hipStream_t s )
{
	switch( model )
	{
	case SPICE_MODEL_VOGELS_ABBOTT: return make_group_vogels_abbott( nets, s );
	case SPICE_MODEL_BRUNEL: return make_group_brunel( nets, s );
	case SPICE_MODEL_BRUNEL_WITH_PLASTICITY: return make_group_brunel_with_plasticity( nets, s );
	case SPICE_MODEL_SYNTH: return make_group_synth( nets, s );
	case SPICE_MODEL_BRUNEL_3POP: return make_group_brunel_3pop( nets, s );
	case SPICE_MODEL_BRUNEL_3POP_STDP: return make_group_brunel_3pop_stdp( nets, s );
	case SPICE_MODEL_BRUNEL_3POP_STDP_DEFERRED: return make_group_brunel_3pop_stdp_deferred( nets, s );
	default: throw std::invalid_argument( "unknown model id" );
	}
}
} // namespace spice::hipx
