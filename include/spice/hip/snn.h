#pragma once
// spice::hip::snn<Model>: the MI355X (gfx950) backend behind the reference's C++ API, a
// drop-in beside cpu::snn / cuda::snn (reference: spice/cuda/snn.h:16-77). Header-only; all
// work goes through the C ABI of spice_hip.h (link with libspice_hip.so).
//
//   spice::hip::snn<spice::vogels_abbott> net( { 4000, 0.02f }, 0.0001f, 8 );
//   std::vector<int> spikes;
//   net.step( &spikes );
//
// Errors are rethrown as the reference's exception types.
#include <spice/models/all.h>
#include <spice/snn.h>
#include <spice/util/layout.h>
#include <spice_hip.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <typeinfo>
#include <vector>

namespace spice::hip
{
inline void check( int status )
{
	if( status == SPICE_OK ) return;
	std::string const msg = spice_last_error();
	switch( status )
	{
	case SPICE_ERR_INVALID_ARGUMENT: throw std::invalid_argument( msg );
	case SPICE_ERR_BAD_ALLOC: throw std::bad_alloc();
	case SPICE_ERR_BAD_CAST: throw std::bad_cast();
	default: throw std::runtime_error( msg );
	}
}

// Default graph/RNG seeds follow the reference's process-wide counter starting at 1337
// (spice/cuda/algorithm.cu:35-39); pass an explicit seed for reproducible networks.
inline std::uint64_t next_seed()
{
	static std::atomic<std::uint64_t> s{ 1337 };
	return s++;
}

template <typename Model>
class snn : public ::spice::snn<Model>
{
	using ntuple = typename Model::neuron::tuple_t;
	using stuple = typename Model::synapse::tuple_t;

public:
	snn( util::layout const & desc, float dt, int_ delay = 1, std::uint64_t seed = next_seed(), std::uint32_t flags = 0 )
	    : ::spice::snn<Model>( dt, delay )
	{
		// pass the already-flattened layout as one population per distinct range
		std::vector<std::uint64_t> pops;
		std::vector<spice_connection> conns;
		flatten( desc, pops, conns );
		check( spice_hip_create(
		    model_id<Model>::value, pops.data(), int32_t( pops.size() ), conns.data(), int32_t( conns.size() ), dt, delay, seed, flags, &_net ) );
	}

	snn( std::vector<int_> const & adj, size_ width, float dt, int_ delay = 1, std::uint64_t seed = next_seed(), std::uint32_t flags = 0 )
	    : ::spice::snn<Model>( dt, delay )
	{
		if( width == 0 || adj.size() % width ) throw std::invalid_argument( "adjacency size must be a multiple of width" );
		check( spice_hip_create_from_ell( model_id<Model>::value, adj.data(), adj.size() / width, width, dt, delay, seed, flags, &_net ) );
	}

	// Conversion constructor: copies graph, state and clock of any backend.
	explicit snn( ::spice::snn<Model> const & net, std::uint64_t seed = next_seed(), std::uint32_t flags = 0 )
	    : snn( net.adj().first, net.adj().second, net.dt(), net.delay(), seed, flags )
	{
		import_neurons( net.neurons() );
		if constexpr( Model::synapse::size > 0 ) import_synapses( net.synapses() );
		check( spice_hip_set_clock( _net, net.step_index(), net.simtime(), net.simtime_compensation() ) );
		this->_set_clock( net.step_index(), net.simtime(), net.simtime_compensation() );
	}

	snn( snn const & ) = delete;
	snn & operator=( snn const & ) = delete;
	~snn() override { spice_hip_destroy( _net ); }

	void step( std::vector<int> * out_spikes = nullptr ) override
	{
		int_ * d;
		uint_ * c;
		step( &d, &c, out_spikes );
	}

	// Device-pointer step (spice/cuda/snn.h:36-37): pointers stay valid until the next step.
	void step( int_ ** out_dspikes, uint_ ** out_dnum_spikes, std::vector<int> * out_spikes = nullptr )
	{
		this->_step( [&]( int_, float ) {
			check( spice_hip_step( _net, out_dspikes, out_dnum_spikes ) );
			if( out_spikes )
			{
				out_spikes->resize( num_neurons() );
				std::uint64_t n = 0;
				check( spice_hip_copy_spikes( _net, out_spikes->data(), out_spikes->size(), &n ) );
				out_spikes->resize( n );
			}
		} );
	}

	// n steps without readback, replayed from a captured HIP graph (extension).
	void run( long_ n )
	{
		check( spice_hip_run( _net, n ) );
		this->_advance( n );
	}
	void synchronize() { check( spice_hip_synchronize( _net ) ); }

	size_ num_neurons() const override { return query( spice_hip_num_neurons ); }
	size_ num_synapses() const override { return query( spice_hip_num_synapses ); }
	size_ num_edges() const { return query( spice_hip_num_edges ); }

	std::pair<std::vector<int>, size_> adj() const override
	{
		size_ const w = query( spice_hip_width );
		std::vector<int> e( num_neurons() * w );
		check( spice_hip_export_adj( _net, e.data(), e.size() ) );
		return { std::move( e ), w };
	}

	std::vector<ntuple> neurons() const override
	{
		std::vector<ntuple> out( num_neurons() );
		util::for_n<Model::neuron::size>( [&]( auto I ) {
			using T = std::tuple_element_t<I, ntuple>;
			std::vector<T> a( out.size() );
			check( spice_hip_export_neuron_attr( _net, int32_t( I ), a.data(), a.size() * sizeof( T ) ) );
			for( size_ i = 0; i < out.size(); i++ ) std::get<I>( out[i] ) = a[i];
		} );
		return out;
	}

	std::vector<stuple> synapses() const override
	{
		std::vector<stuple> out( Model::synapse::size > 0 ? num_synapses() : 0 );
		util::for_n<Model::synapse::size>( [&]( auto I ) {
			using T = std::tuple_element_t<I, stuple>;
			std::vector<T> a( out.size() );
			check( spice_hip_export_synapse_attr( _net, int32_t( I ), a.data(), a.size() * sizeof( T ), 1 ) );
			for( size_ i = 0; i < out.size(); i++ ) std::get<I>( out[i] ) = a[i];
		} );
		return out;
	}

	spice_net * handle() const { return _net; }

private:
	spice_net * _net = nullptr;

	template <typename F>
	size_ query( F f ) const
	{
		std::uint64_t v = 0;
		check( f( _net, &v ) );
		return size_( v );
	}

	void import_neurons( std::vector<ntuple> const & v )
	{
		util::for_n<Model::neuron::size>( [&]( auto I ) {
			using T = std::tuple_element_t<I, ntuple>;
			std::vector<T> a( v.size() );
			for( size_ i = 0; i < v.size(); i++ ) a[i] = std::get<I>( v[i] );
			check( spice_hip_import_neuron_attr( _net, int32_t( I ), a.data(), a.size() * sizeof( T ) ) );
		} );
	}
	void import_synapses( std::vector<stuple> const & v )
	{
		util::for_n<Model::synapse::size>( [&]( auto I ) {
			using T = std::tuple_element_t<I, stuple>;
			std::vector<T> a( v.size() );
			for( size_ i = 0; i < v.size(); i++ ) a[i] = std::get<I>( v[i] );
			check( spice_hip_import_synapse_attr( _net, int32_t( I ), a.data(), a.size() * sizeof( T ), 1 ) );
		} );
	}

	// Rebuilds a (pops, connections) description whose flattening equals desc's.
	static void flatten( util::layout const & desc, std::vector<std::uint64_t> & pops, std::vector<spice_connection> & conns )
	{
		std::vector<int_> cuts{ 0, int_( desc.size() ) };
		for( auto const & e : desc.connections() )
			for( int_ x : { std::get<0>( e ), std::get<1>( e ), std::get<2>( e ), std::get<3>( e ) } ) cuts.push_back( x );
		std::sort( cuts.begin(), cuts.end() );
		cuts.erase( std::unique( cuts.begin(), cuts.end() ), cuts.end() );
		for( size_t i = 0; i + 1 < cuts.size(); i++ ) pops.push_back( std::uint64_t( cuts[i + 1] - cuts[i] ) );
		auto pop_of = [&]( int_ first ) { return std::uint64_t( std::lower_bound( cuts.begin(), cuts.end(), first ) - cuts.begin() ); };
		for( auto const & e : desc.connections() )
		{
			// a flattened range may span several populations of the cut list only if the
			// original layout had them merged, which cannot happen: ranges are whole pops.
			for( std::uint64_t s = pop_of( std::get<0>( e ) ); s < pop_of( std::get<1>( e ) ); s++ )
				for( std::uint64_t d = pop_of( std::get<2>( e ) ); d < pop_of( std::get<3>( e ) ); d++ )
					conns.push_back( { s, d, std::get<4>( e ) } );
		}
	}
};
} // namespace spice::hip
