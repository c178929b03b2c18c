// TEST INFRASTRUCTURE ONLY: a thin C ABI over the reference's own CPU simulator, compiled
// together with the reference's sources (see the `ref` target of oracle/Makefile; the
// sources are read where they lie and never copied). Used to pin the CPU checker and the
// generator's degree statistics against the real reference.
#include <spice/cpu/snn.h>
#include <spice/models/brunel.h>
#include <spice/models/brunel_with_plasticity.h>
#include <spice/models/synth.h>
#include <spice/models/vogels_abbott.h>
#include <spice/util/adj_list.h>
#include <spice/util/layout.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

using namespace spice;

static thread_local std::string g_err;

struct ref_net
{
	virtual ~ref_net() = default;
	virtual void step( std::vector<int> * out ) = 0;
	virtual std::pair<std::vector<int>, size_t> adj() const = 0;
	virtual size_t num_neurons() const = 0;
	virtual int neuron_attrs() const = 0;
	virtual int synapse_attrs() const = 0;
	virtual void export_neurons( int a, void * out ) const = 0;
	virtual void export_synapses( int a, void * out ) const = 0;
};

template <typename Model>
struct ref_net_t : ref_net
{
	cpu::snn<Model> net;
	ref_net_t( util::layout const & l, float dt, int delay )
	    : net( l, dt, delay )
	{
	}
	void step( std::vector<int> * out ) override { net.step( out ); }
	std::pair<std::vector<int>, size_t> adj() const override { return net.adj(); }
	size_t num_neurons() const override { return net.num_neurons(); }
	int neuron_attrs() const override { return int( Model::neuron::size ); }
	int synapse_attrs() const override { return int( Model::synapse::size ); }
	template <typename Vec>
	static void out_attr( Vec const & v, int a, void * out )
	{
		using Tuple = typename Vec::value_type;
		util::for_n<std::tuple_size_v<Tuple>>( [&]( auto I ) {
			if( int( I ) != a ) return;
			auto * o = static_cast<std::tuple_element_t<I, Tuple> *>( out );
			for( size_t i = 0; i < v.size(); i++ ) o[i] = std::get<I>( v[i] );
		} );
	}
	void export_neurons( int a, void * out ) const override { out_attr( net.neurons(), a, out ); }
	void export_synapses( int a, void * out ) const override { out_attr( net.synapses(), a, out ); }
};

static util::layout make_layout( uint64_t const * pops, int npops, uint64_t const * csrc, uint64_t const * cdst, float const * cp, int nconns )
{
	std::vector<size_t> p( pops, pops + npops );
	std::vector<std::tuple<size_t, size_t, float>> c;
	for( int i = 0; i < nconns; i++ ) c.emplace_back( csrc[i], cdst[i], cp[i] );
	return util::layout( p, c );
}

extern "C" {
const char * ref_last_error() { return g_err.c_str(); }

// model: 0 vogels_abbott, 1 brunel, 2 brunel_with_plasticity, 3 synth
int ref_create(
    int model, uint64_t const * pops, int npops, uint64_t const * csrc, uint64_t const * cdst, float const * cp, int nconns, float dt, int delay, ref_net ** out )
{
	try
	{
		auto l = make_layout( pops, npops, csrc, cdst, cp, nconns );
		switch( model )
		{
		case 0: *out = new ref_net_t<vogels_abbott>( l, dt, delay ); break;
		case 1: *out = new ref_net_t<brunel>( l, dt, delay ); break;
		case 2: *out = new ref_net_t<brunel_with_plasticity>( l, dt, delay ); break;
		case 3: *out = new ref_net_t<synth>( l, dt, delay ); break;
		default: g_err = "bad model"; return 1;
		}
		return 0;
	}
	catch( std::exception const & e )
	{
		g_err = e.what();
		return 1;
	}
}
void ref_destroy( ref_net * n ) { delete n; }
int ref_step( ref_net * n, int32_t * out, uint64_t cap, uint64_t * count )
{
	std::vector<int> s;
	n->step( &s );
	*count = s.size();
	if( s.size() > cap ) return 1;
	std::memcpy( out, s.data(), s.size() * 4 );
	return 0;
}
uint64_t ref_num_neurons( ref_net * n ) { return n->num_neurons(); }
uint64_t ref_width( ref_net * n ) { return n->adj().second; }
void ref_export_adj( ref_net * n, int32_t * out )
{
	auto a = n->adj();
	std::memcpy( out, a.first.data(), a.first.size() * 4 );
}
void ref_export_neurons( ref_net * n, int a, void * out ) { n->export_neurons( a, out ); }
void ref_export_synapses( ref_net * n, int a, void * out ) { n->export_synapses( a, out ); }

// layout flattening through the reference's own class (for known-answer comparisons)
int ref_layout(
    uint64_t const * pops, int npops, uint64_t const * csrc, uint64_t const * cdst, float const * cp, int nconns, int32_t * out5 /* nconns x 4 */, float * outp, uint64_t * n, uint64_t * maxdeg )
{
	try
	{
		auto l = make_layout( pops, npops, csrc, cdst, cp, nconns );
		size_t i = 0;
		for( auto const & e : l.connections() )
		{
			out5[4 * i + 0] = std::get<0>( e );
			out5[4 * i + 1] = std::get<1>( e );
			out5[4 * i + 2] = std::get<2>( e );
			out5[4 * i + 3] = std::get<3>( e );
			outp[i] = std::get<4>( e );
			i++;
		}
		*n = l.size();
		*maxdeg = l.max_degree();
		return 0;
	}
	catch( std::exception const & e )
	{
		g_err = e.what();
		return 1;
	}
}

// The reference's host graph generator (util/adj_list.cpp:49-91) into an ELL buffer.
int ref_generate( uint64_t const * pops, int npops, uint64_t const * csrc, uint64_t const * cdst, float const * cp, int nconns, int32_t * out )
{
	try
	{
		auto l = make_layout( pops, npops, csrc, cdst, cp, nconns );
		std::vector<int> e;
		util::adj_list::generate( l, e );
		std::memcpy( out, e.data(), e.size() * 4 );
		return 0;
	}
	catch( std::exception const & e )
	{
		g_err = e.what();
		return 1;
	}
}
}
