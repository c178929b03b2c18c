// TEST INFRASTRUCTURE ONLY (the CPU checker): C ABI over oracle/cpu_snn.h and
// oracle/host_graph.h for the Python tests and the CPU baseline leg of bench.py.
// Nothing in the product links or loads this library.
#include "cpu_snn.h"
#include "host_graph.h"

#include <spice/models/all.h>
#include <spice/util/detmath.h>
#include <spice/util/layout.h>
#include <spice/util/numeric.h>
#include <spice_hip.h>

#include <climits>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <typeinfo>

using namespace spice;

static thread_local std::string g_err;

// Checker-only flag: draw random numbers from the reference CPU backend's sequential
// xoshiro128+ stream seeded with `seed` (the reference seeds nets 1337, 1338, ...).
constexpr uint32_t ORACLE_FLAG_REF_RNG = 1u << 16;

template <typename F>
static int guard( F && f )
{
	try
	{
		f();
		return SPICE_OK;
	}
	catch( std::invalid_argument const & e )
	{
		g_err = e.what();
		return SPICE_ERR_INVALID_ARGUMENT;
	}
	catch( std::bad_alloc const & e )
	{
		g_err = e.what();
		return SPICE_ERR_BAD_ALLOC;
	}
	catch( std::bad_cast const & e )
	{
		g_err = e.what();
		return SPICE_ERR_BAD_CAST;
	}
	catch( std::exception const & e )
	{
		g_err = e.what();
		return SPICE_ERR_RUNTIME;
	}
}

static util::layout make_layout( uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns )
{
	std::vector<size_t> p( pops, pops + npops );
	std::vector<std::tuple<size_t, size_t, float>> c;
	for( int32_t i = 0; i < nconns; i++ ) c.emplace_back( conns[i].src_pop, conns[i].dst_pop, conns[i].p );
	return util::layout( p, c );
}

static oracle::net_base * make_net(
    int32_t model, std::vector<int64_t> rp, std::vector<int32_t> col, int64_t width, float dt, int32_t delay, uint64_t seed, uint32_t flags, int32_t shift )
{
	return dispatch_model( model, [&]( auto m ) -> oracle::net_base * {
		using M = decltype( m );
		return new oracle::cpu_snn<M>(
		    std::move( rp ),
		    std::move( col ),
		    width,
		    dt,
		    delay,
		    seed,
		    ( flags & SPICE_FLAG_FIXED_POINT ) != 0,
		    shift,
		    ( flags & ORACLE_FLAG_REF_RNG ) != 0 );
	} );
}

extern "C" {
const char * oracle_last_error( void ) { return g_err.c_str(); }

int oracle_layout_flatten(
    uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns, spice_edge * out, int32_t * nout, uint64_t * n, uint64_t * maxdeg )
{
	return guard( [&] {
		auto l = make_layout( pops, npops, conns, nconns );
		int32_t i = 0;
		for( auto const & e : l.connections() )
			out[i++] = { std::get<0>( e ), std::get<1>( e ), std::get<2>( e ), std::get<3>( e ), std::get<4>( e ) };
		*nout = i;
		*n = l.size();
		*maxdeg = l.max_degree();
	} );
}

// Host build of the generator: size query (row_ptr may be NULL) and fill.
int oracle_generate_csr(
    uint64_t const * pops,
    int32_t npops,
    spice_connection const * conns,
    int32_t nconns,
    uint64_t seed,
    int64_t * out_n,
    int64_t * out_width,
    int64_t * out_nnz,
    int64_t * row_ptr,
    int32_t * col )
{
	return guard( [&] {
		auto l = make_layout( pops, npops, conns, nconns );
		auto g = oracle::generate_csr( oracle::gen_conns( l ), int64_t( l.size() ), int64_t( l.max_degree() ), seed );
		*out_n = int64_t( l.size() );
		*out_width = int64_t( l.max_degree() );
		*out_nnz = int64_t( g.col.size() );
		if( row_ptr ) std::memcpy( row_ptr, g.row_ptr.data(), g.row_ptr.size() * 8 );
		if( col ) std::memcpy( col, g.col.data(), g.col.size() * 4 );
	} );
}

// Host build of rows [row_first, row_last) only: row_ptr gets the row_last - row_first + 1
// absolute offsets, col (NULL: offsets only) the targets of those rows (col_cap elements).
int oracle_generate_rows(
    uint64_t const * pops,
    int32_t npops,
    spice_connection const * conns,
    int32_t nconns,
    uint64_t seed,
    int64_t row_first,
    int64_t row_last,
    int64_t * row_ptr,
    int32_t * col,
    int64_t col_cap )
{
	return guard( [&] {
		auto l = make_layout( pops, npops, conns, nconns );
		if( row_first < 0 || row_first > row_last || row_last > int64_t( l.size() ) ) throw std::invalid_argument( "row range out of bounds" );
		auto g = oracle::generate_rows( oracle::gen_conns( l ), row_first, row_last, int64_t( l.max_degree() ), seed, col != nullptr );
		std::memcpy( row_ptr, g.row_ptr.data(), g.row_ptr.size() * 8 );
		if( !col ) return;
		if( int64_t( g.col.size() ) > col_cap ) throw std::invalid_argument( "column buffer too small" );
		std::memcpy( col, g.col.data(), g.col.size() * 4 );
	} );
}

int oracle_create(
    int32_t model,
    uint64_t const * pops,
    int32_t npops,
    spice_connection const * conns,
    int32_t nconns,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    int32_t shift,
    oracle::net_base ** out )
{
	return guard( [&] {
		auto l = make_layout( pops, npops, conns, nconns );
		auto g = oracle::generate_csr( oracle::gen_conns( l ), int64_t( l.size() ), int64_t( l.max_degree() ), seed );
		*out = make_net( model, std::move( g.row_ptr ), std::move( g.col ), int64_t( l.max_degree() ), dt, delay, seed, flags, shift );
	} );
}

int oracle_create_csr(
    int32_t model,
    int64_t const * row_ptr,
    int32_t const * col,
    int64_t n,
    int64_t width,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    int32_t shift,
    oracle::net_base ** out )
{
	return guard( [&] {
		std::vector<int64_t> rp( row_ptr, row_ptr + n + 1 );
		std::vector<int32_t> c( col, col + rp[n] );
		*out = make_net( model, std::move( rp ), std::move( c ), width, dt, delay, seed, flags, shift );
	} );
}

// ELL import: each row's entries up to the first INT_MAX (rows are sorted).
int oracle_create_ell(
    int32_t model, int32_t const * ell, int64_t n, int64_t width, float dt, int32_t delay, uint64_t seed, uint32_t flags, int32_t shift, oracle::net_base ** out )
{
	return guard( [&] {
		std::vector<int64_t> rp( n + 1, 0 );
		std::vector<int32_t> c;
		for( int64_t r = 0; r < n; r++ )
		{
			for( int64_t j = 0; j < width && ell[r * width + j] != INT_MAX; j++ ) c.push_back( ell[r * width + j] );
			rp[r + 1] = int64_t( c.size() );
		}
		*out = make_net( model, std::move( rp ), std::move( c ), width, dt, delay, seed, flags, shift );
	} );
}

int oracle_destroy( oracle::net_base * net )
{
	delete net;
	return SPICE_OK;
}

int oracle_step( oracle::net_base * net, int32_t * out, uint64_t cap, uint64_t * count )
{
	return guard( [&] {
		std::vector<int> s;
		net->step( out || count ? &s : nullptr );
		if( count ) *count = s.size();
		if( out )
		{
			if( s.size() > cap ) throw std::invalid_argument( "spike buffer too small" );
			std::memcpy( out, s.data(), s.size() * 4 );
		}
	} );
}

// Runs nsteps steps and records every step's spikes: ids of step k at
// ids[offsets[k] .. offsets[k+1]) (nsteps + 1 offsets).
int oracle_run_record( oracle::net_base * net, int64_t nsteps, int32_t * ids, uint64_t cap, uint64_t * offsets )
{
	return guard( [&] {
		std::vector<int> s;
		uint64_t o = 0;
		for( int64_t i = 0; i < nsteps; i++ )
		{
			offsets[i] = o;
			net->step( &s );
			if( o + s.size() > cap ) throw std::invalid_argument( "spike buffer too small" );
			std::memcpy( ids + o, s.data(), s.size() * 4 );
			o += s.size();
		}
		offsets[nsteps] = o;
	} );
}

int oracle_run( oracle::net_base * net, int64_t nsteps )
{
	return guard( [&] {
		for( int64_t i = 0; i < nsteps; i++ ) net->step( nullptr );
	} );
}

int oracle_info( oracle::net_base * net, int64_t * n, int64_t * width, int64_t * nnz, int32_t * nna, int32_t * nsa )
{
	*n = net->num_neurons();
	*width = net->width();
	*nnz = int64_t( net->col().size() );
	*nna = net->num_neuron_attrs();
	*nsa = net->num_synapse_attrs();
	return SPICE_OK;
}

int oracle_export_csr( oracle::net_base * net, int64_t * row_ptr, int32_t * col )
{
	std::memcpy( row_ptr, net->row_ptr().data(), net->row_ptr().size() * 8 );
	std::memcpy( col, net->col().data(), net->col().size() * 4 );
	return SPICE_OK;
}

int oracle_export_neuron_attr( oracle::net_base * net, int32_t a, void * out )
{
	return guard( [&] {
		if( a < 0 || a >= net->num_neuron_attrs() ) throw std::invalid_argument( "bad attribute" );
		net->export_neuron_attr( a, out );
	} );
}
int oracle_import_neuron_attr( oracle::net_base * net, int32_t a, void const * in )
{
	return guard( [&] {
		if( a < 0 || a >= net->num_neuron_attrs() ) throw std::invalid_argument( "bad attribute" );
		net->import_neuron_attr( a, in );
	} );
}
int oracle_export_synapse_attr( oracle::net_base * net, int32_t a, void * out )
{
	return guard( [&] {
		if( a < 0 || a >= net->num_synapse_attrs() ) throw std::invalid_argument( "bad attribute" );
		net->export_synapse_attr( a, out );
	} );
}
int oracle_import_synapse_attr( oracle::net_base * net, int32_t a, void const * in )
{
	return guard( [&] {
		if( a < 0 || a >= net->num_synapse_attrs() ) throw std::invalid_argument( "bad attribute" );
		net->import_synapse_attr( a, in );
	} );
}
int oracle_set_params( oracle::net_base * net, float const * p )
{
	net->set_params( p );
	return SPICE_OK;
}
int oracle_export_traces( oracle::net_base * net, float * x, float * y ) { return net->export_traces( x, y ) ? 0 : 1; }
int oracle_get_clock( oracle::net_base * net, int64_t * i, float * sum, float * c )
{
	net->get_clock( i, sum, c );
	return SPICE_OK;
}
int oracle_set_clock( oracle::net_base * net, int64_t i, float sum, float c )
{
	net->set_clock( i, sum, c );
	return SPICE_OK;
}
// Spikes in flight: counts[k] ids (concatenated in ids) emitted at step i - nsteps + k, i the
// next step to run; replaces the net's own record of the last nsteps (<= delay) steps.
int oracle_set_in_flight( oracle::net_base * net, int64_t nsteps, int64_t const * counts, int32_t const * ids )
{
	return guard( [&] {
		std::vector<int64_t> c( counts, counts + nsteps );
		int64_t total = 0;
		for( auto x : c ) total += x;
		net->set_in_flight( c, std::vector<int>( ids, ids + total ) );
	} );
}

// ---- primitives exposed for known-answer and accuracy tests ----
void oracle_philox( uint32_t const ctr[4], uint32_t const key[2], uint32_t out[4] )
{
	auto r = util::philox4x32_10( { ctr[0], ctr[1], ctr[2], ctr[3] }, key[0], key[1] );
	out[0] = r.x;
	out[1] = r.y;
	out[2] = r.z;
	out[3] = r.w;
}
void oracle_det_log( double const * x, double * y, int64_t n )
{
	for( int64_t i = 0; i < n; i++ ) y[i] = util::det::log( x[i] );
}
void oracle_det_sqrt( double const * x, double * y, int64_t n )
{
	for( int64_t i = 0; i < n; i++ ) y[i] = util::det::sqrt( x[i] );
}
void oracle_det_cos2pi( double const * x, double * y, int64_t n )
{
	for( int64_t i = 0; i < n; i++ ) y[i] = util::det::cos2pi( x[i] );
}
void oracle_exp_fix( uint32_t const * x, uint32_t * y, int64_t n )
{
	for( int64_t i = 0; i < n; i++ ) y[i] = util::gen::exp_fix_one( x[i] );
}
// Degrees of every (row, connection) pair, clipped to the row width, for distribution tests.
int oracle_degrees(
    uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns, uint64_t seed, int32_t * out_deg /* n x nedges, -1 = not covering */ )
{
	return guard( [&] {
		auto l = make_layout( pops, npops, conns, nconns );
		auto gc = oracle::gen_conns( l );
		int64_t const n = int64_t( l.size() ), w = int64_t( l.max_degree() );
		for( int64_t r = 0; r < n; r++ )
		{
			int64_t total = 0;
			for( uint32_t c = 0; c < gc.size(); c++ )
			{
				int32_t & o = out_deg[r * int64_t( gc.size() ) + c];
				if( r < gc[c].src_first || r >= gc[c].src_last )
				{
					o = -1;
					continue;
				}
				int64_t d = util::gen::sample_degree( seed, uint32_t( r ), c, gc[c] );
				if( d > w - total ) d = w - total;
				total += d;
				o = int32_t( d );
			}
		}
	} );
}
void oracle_kahan( float dt, int64_t nsteps, float * deltas, float * sums )
{
	util::kahan_sum<float> k;
	for( int64_t i = 0; i < nsteps; i++ )
	{
		deltas[i] = k.add( dt );
		sums[i] = float( k );
	}
}
}
