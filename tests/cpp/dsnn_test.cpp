// C++ drop-in test of spice::hip::snn<Model> through its public C++ API, mirroring the
// reference's GPU tests (spice_test/cuda/snn.d.cpp:72-157: dSNN.Ctor, dSNN.Step,
// dSNN.StepWithDelay) with the CPU checker standing in for cpu::snn.
// Exit code 0 = all checks passed. Prints one line per check.
#include "../../oracle/cpu_snn.h"
#include "../../oracle/host_graph.h"

#include <spice/hip/snn.h>
#include <spice/models/all.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <tuple>
#include <vector>

using namespace spice;

// cpu::snn stand-in: the CPU checker behind the same abstract interface.
template <typename Model>
class cpu_net : public spice::snn<Model>
{
public:
	cpu_net( util::layout const & desc, float dt, int_ delay = 1, std::uint64_t seed = 1 )
	    : spice::snn<Model>( dt, delay )
	    , _w( desc.max_degree() )
	{
		auto g = oracle::generate_csr( oracle::gen_conns( desc ), desc.size(), desc.max_degree(), seed );
		_net = std::make_unique<oracle::cpu_snn<Model>>( g.row_ptr, g.col, _w, dt, delay, seed, false, 40 );
	}
	void step( std::vector<int> * out ) override
	{
		this->_step( [&]( int_, float ) { _net->step( out ); } );
	}
	size_ num_neurons() const override { return size_( _net->num_neurons() ); }
	size_ num_synapses() const override { return num_neurons() * _w; }
	std::pair<std::vector<int>, size_> adj() const override
	{
		auto const & rp = _net->row_ptr();
		auto const & col = _net->col();
		std::vector<int> e( num_neurons() * _w, INT_MAX );
		for( size_ r = 0; r < num_neurons(); r++ ) std::copy( col.begin() + rp[r], col.begin() + rp[r + 1], e.begin() + r * _w );
		return { e, _w };
	}
	std::vector<typename Model::neuron::tuple_t> neurons() const override
	{
		std::vector<typename Model::neuron::tuple_t> out( num_neurons() );
		util::for_n<Model::neuron::size>( [&]( auto I ) {
			std::vector<std::tuple_element_t<I, typename Model::neuron::tuple_t>> a( out.size() );
			_net->export_neuron_attr( int( I ), a.data() );
			for( size_ i = 0; i < out.size(); i++ ) std::get<I>( out[i] ) = a[i];
		} );
		return out;
	}
	std::vector<typename Model::synapse::tuple_t> synapses() const override
	{
		std::vector<typename Model::synapse::tuple_t> out( Model::synapse::size ? num_synapses() : 0 );
		auto const & rp = _net->row_ptr();
		util::for_n<Model::synapse::size>( [&]( auto I ) {
			std::vector<std::tuple_element_t<I, typename Model::synapse::tuple_t>> a( _net->col().size() );
			_net->export_synapse_attr( int( I ), a.data() );
			for( size_ r = 0; r < num_neurons(); r++ )
				for( int64_t k = rp[r]; k < rp[r + 1]; k++ ) std::get<I>( out[r * _w + ( k - rp[r] )] ) = a[k];
		} );
		return out;
	}

private:
	size_ _w;
	std::unique_ptr<oracle::cpu_snn<Model>> _net;
};

static int failures = 0;
#define CHECK( cond, what )                                                  \
	do                                                                        \
	{                                                                         \
		bool const ok_ = ( cond );                                            \
		std::printf( "%s %s\n", ok_ ? "ok  " : "FAIL", what );                \
		if( !ok_ ) failures++;                                                \
	} while( 0 )

template <typename Cont>
static bool set_equal( Cont a, Cont b )
{
	std::sort( a.begin(), a.end() );
	std::sort( b.begin(), b.end() );
	return a == b;
}

// mean absolute difference per tuple <= thres (spice_test/cuda/snn.d.cpp:26-44)
template <typename Vec>
static bool close( Vec const & lhs, Vec const & rhs, double thres )
{
	if( lhs.size() != rhs.size() ) return false;
	constexpr size_t sz = std::tuple_size_v<typename Vec::value_type>;
	if( sz == 0 ) return true;
	for( size_t i = 0; i < lhs.size(); i++ )
	{
		double d = 0;
		util::for_n<sz>( [&]( auto I ) { d += std::abs( double( std::get<I>( lhs[i] ) ) - double( std::get<I>( rhs[i] ) ) ); } );
		if( d / sz > thres ) return false;
	}
	return true;
}

template <typename Model>
static void ctor_test( char const * name )
{
	char buf[256];
	{
		hip::snn<Model> d( { 100, 0.1f }, 0.0001f );
		std::snprintf( buf, sizeof buf, "%s Ctor: num_neurons/dt/delay", name );
		CHECK( d.num_neurons() == 100u && d.dt() == 0.0001f && d.delay() == 1, buf );
	}
	{
		hip::snn<Model> d( util::layout{ 100, 0.1f }, 0.0001f, 15, 7 );
		cpu_net<Model> h( { 100, 0.1f }, 0.0001f, 15, 7 );
		std::snprintf( buf, sizeof buf, "%s Ctor: same graph and init as the CPU checker", name );
		CHECK( d.delay() == 15 && d.adj() == h.adj() && close( h.neurons(), d.neurons(), 0.0 ) && close( h.synapses(), d.synapses(), 0.0 ), buf );
	}
	{
		cpu_net<Model> h( { 100, 0.1f }, 0.0001f, 15, 3 );
		hip::snn<Model> d( h );
		std::snprintf( buf, sizeof buf, "%s Conversion ctor: graph, neurons, synapses copied exactly", name );
		CHECK( d.num_neurons() == 100u && d.delay() == 15 && d.adj() == h.adj() && close( h.neurons(), d.neurons(), 0.0 ) &&
		           close( h.synapses(), d.synapses(), 0.0 ),
		       buf );
	}
}

static void step_test( int delay )
{
	char buf[256];
	cpu_net<vogels_abbott> h( { 4000, 0.02f }, 0.0001f, delay, 11 );
	hip::snn<vogels_abbott> d( h );
	std::vector<int> hs, ds;
	bool all = true;
	for( int i = 0; i < 1000; i++ )
	{
		h.step( &hs );
		d.step( &ds );
		all = all && set_equal( hs, ds );
	}
	std::snprintf( buf, sizeof buf, "vogels Step (delay %d): spike sets equal for 1000 steps", delay );
	CHECK( all, buf );
	std::snprintf( buf, sizeof buf, "vogels Step (delay %d): state close (1e-5)", delay );
	CHECK( close( h.neurons(), d.neurons(), 1e-5 ), buf );
}

int main()
{
	try
	{
		ctor_test<vogels_abbott>( "vogels_abbott" );
		ctor_test<brunel>( "brunel" );
		ctor_test<brunel_with_plasticity>( "brunel_with_plasticity" );
		step_test( 1 );
		step_test( 8 );
		{ // error behaviour: the reference's exception types
			bool thrown = false;
			try
			{
				hip::snn<brunel> bad( { 10, 0.1f }, 0.0001f, 0 );
			}
			catch( std::invalid_argument const & )
			{
				thrown = true;
			}
			CHECK( thrown, "delay 0 throws std::invalid_argument" );
		}
	}
	catch( std::exception const & e )
	{
		std::printf( "FAIL exception: %s\n", e.what() );
		return 2;
	}
	std::printf( "%d failure(s)\n", failures );
	return failures ? 1 : 0;
}
