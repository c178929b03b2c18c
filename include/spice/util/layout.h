#pragma once
// Network connectivity descriptor: population sizes plus (src pop, dst pop, probability)
// connections, flattened to neuron-id ranges.
//
// Same interface and semantics as the reference's layout (spice/util/layout.h:24-57,
// spice/util/layout.cpp:51-110): validation errors throw std::invalid_argument, connections
// are sorted by (src pop, dst pop), duplicates are rejected, and max_degree() is the ELL
// row width: for each source population, mean + 3 sigma of the summed binomial degrees,
// rounded up to a multiple of 32. The multi-GPU helpers (static_load_balance, cut) are not
// part of the single-GPU path and are omitted.
#include <spice/util/stdint.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace spice::util
{
class layout
{
public:
	// ([src_first, src_last), [dst_first, dst_last), p)
	using edge = std::tuple<int_, int_, int_, int_, float>;

	layout( size_ num_neurons, float p )
	    : layout( check_n( num_neurons ), std::vector<std::tuple<size_, size_, float>>{ { 0, 0, p } } )
	{
	}

	layout( std::vector<size_> const & pops, std::vector<std::tuple<size_, size_, float>> connections )
	{
		require( !pops.empty(), "layout must contain at least 1 (non-empty) population" );
		for( auto pop : pops )
			require(
			    pop > 0 && pop <= size_( std::numeric_limits<int>::max() ), "invalid population size" );
		for( auto const & c : connections )
		{
			require(
			    std::get<0>( c ) < pops.size() && std::get<1>( c ) < pops.size(),
			    "invalid index in connections matrix" );
			require( std::get<2>( c ) >= 0.0f && std::get<2>( c ) <= 1.0f, "invalid connect. prob." );
		}

		std::sort( connections.begin(), connections.end(), []( auto const & a, auto const & b ) {
			return std::make_pair( std::get<0>( a ), std::get<1>( a ) ) <
			       std::make_pair( std::get<0>( b ), std::get<1>( b ) );
		} );
		for( size_ i = 1; i < connections.size(); i++ )
			require(
			    std::get<0>( connections[i] ) != std::get<0>( connections[i - 1] ) ||
			        std::get<1>( connections[i] ) != std::get<1>( connections[i - 1] ),
			    "no duplicate edges in connections list allowed" );

		std::vector<size_> first( pops.size() + 1, 0 );
		for( size_ i = 0; i < pops.size(); i++ ) first[i + 1] = first[i] + pops[i];
		_n = first.back();
		require( _n <= size_( std::numeric_limits<int>::max() ), "invalid node count" );

		for( auto const & c : connections )
		{
			auto const s = std::get<0>( c ), d = std::get<1>( c );
			_connections.emplace_back(
			    int_( first[s] ), int_( first[s + 1] ), int_( first[d] ), int_( first[d + 1] ), std::get<2>( c ) );
		}
		_max_degree = estimate_max_degree( _connections );
	}

	size_ size() const { return _n; }
	std::vector<edge> const & connections() const { return _connections; }
	size_ max_degree() const { return _max_degree; }

	// ELL row width: max over source populations of floor(m + 3 sqrt(s2)), m and s2 the
	// summed binomial mean and variance of the row degree, rounded up to a multiple of 32.
	static size_ estimate_max_degree( std::vector<edge> const & conns )
	{
		size_ result = 0;
		size_ i = 0;
		while( i < conns.size() )
		{
			int_ const src = std::get<0>( conns[i] );
			double m = 0.0, s2 = 0.0;
			for( ; i < conns.size() && std::get<0>( conns[i] ) == src; i++ )
			{
				double const range = double( std::get<3>( conns[i] ) - std::get<2>( conns[i] ) );
				double const p = std::get<4>( conns[i] );
				m += range * p;
				s2 += range * p * ( 1.0 - p );
			}
			result = std::max( result, size_( m + 3 * std::sqrt( s2 ) ) );
		}
		return ( result + 31 ) / 32 * 32;
	}

private:
	size_ _n = 0;
	std::vector<edge> _connections;
	size_ _max_degree = 0;

	static void require( bool cond, char const * msg )
	{
		if( !cond ) throw std::invalid_argument( std::string( "layout: " ) + msg );
	}
	static std::vector<size_> check_n( size_ n )
	{
		require( n > 0, "layout must contain at least 1 neuron" );
		return { n };
	}
};
} // namespace spice::util
