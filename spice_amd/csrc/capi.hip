// C ABI of the gfx950 simulator (declared in include/spice_hip.h).
#include "factory.h"
#include "serve.h"

#include <spice/models/all.h>
#include <spice/util/layout.h>
#include <spice_hip.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>
#include <new>
#include <string>
#include <typeinfo>


using namespace spice;
using namespace spice::hipx;

struct spice_net
{
	net_iface * impl;
	batch_iface * owner = nullptr; // set for nets borrowed from a batch
};

struct spice_batch
{
	batch_iface * impl;
	std::vector<spice_net *> handles; // borrowed per-net handles, released with the batch
};

struct spice_server
{
	std::int64_t chunk;
	std::unique_ptr<hip_exec> exec;
	std::unique_ptr<serve::scheduler<hip_exec>> sched; // declared last: stopped first
};

static thread_local std::string g_err;

template <typename F>
static int guard( F && f )
{
	try
	{
		g_err.clear();
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
		g_err = std::string( "bad_alloc: " ) + e.what();
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

// Queries and exports of a net borrowed from a batch first wait for the batch's stream.
static net_iface * impl( spice_net * n )
{
	require( n && n->impl, "null net handle" );
	if( n->owner ) n->owner->synchronize();
	return n->impl;
}

// Stepping entry points: a net inside a batch is stepped only through the batch.
static net_iface * stepping_impl( spice_net * n )
{
	require( n && n->impl, "null net handle" );
	require( n->owner == nullptr, "a net that belongs to a batch is stepped through spice_batch_run" );
	return n->impl;
}

static net_params params_of( int32_t model, spice_net_desc const & d, float dt, uint32_t flags )
{
	net_params p{ model, dt, d.delay, d.seed, flags, { d.params[0], d.params[1], d.params[2], d.params[3] } };
	return p;
}

// Generator scratch of spice_hip_generate_graph_noalloc: GRAPH_SCRATCH_BYTES in the code
// object's global segment, so every GPU the library runs on holds its own copy until the
// process ends. In host code `scratch` is a stand-in for the symbol, not a device address:
// kernels and copies get the per-device address that graph_scratch() looked up.
static __device__ __align__( 256 ) char scratch[GRAPH_SCRATCH_BYTES];

// This device's address of `scratch`, looked up once per device (with that device current,
// at its first no-allocation call) and kept; a fixed table, so the lookup allocates nothing.
static void * graph_scratch()
{
	constexpr int MAX_DEVICES = 64;
	static std::mutex m;
	static void * addr[MAX_DEVICES] = {};
	int dev = 0;
	SPICE_HIP( hipGetDevice( &dev ) );
	require( dev >= 0 && dev < MAX_DEVICES, "device id beyond the scratch table" );
	std::lock_guard<std::mutex> lk( m );
	if( !addr[dev] )
	{
		void * p = nullptr;
		SPICE_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
		require( p != nullptr, "hipGetSymbolAddress returned no address for the generator scratch" );
		addr[dev] = p;
	}
	return addr[dev];
}

namespace spice::hipx
{
// Graph generation into caller-owned device memory (spice_hip_generate_graph*). `scratch` null:
// the connection table and the scan's tile sums are allocated for the call; otherwise it
// returns this device's address of a preallocated GRAPH_SCRATCH_BYTES region, whose size
// limits the connection and neuron counts.
int generate_graph_entry( void * ( *scratch )(), uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns,
                          uint64_t seed, int64_t * d_row_ptr, int32_t * d_col, uint64_t col_cap, uint64_t * out_nnz, uint64_t * out_width )
{
	return guard( [&] {
		require( d_row_ptr != nullptr, "null row offset buffer" );
		auto l = make_layout( pops, npops, conns, nconns );
		auto const gc = gen_conns( l );
		int64_t const n = int64_t( l.size() ), width = int64_t( l.max_degree() );
		require( n < int64_t( INT_MAX ) - 1024, "too many neurons" );
		int64_t nnz = 0;
		if( scratch )
		{
			require( gc.size() * sizeof( util::gen::conn ) <= GRAPH_SCRATCH_CONN_BYTES,
			         "too many connections for the preallocated generator scratch (at most 1024)" );
			require( std::size_t( graph_scratch_tiles( n ) ) * 8 <= GRAPH_SCRATCH_BYTES - GRAPH_SCRATCH_CONN_BYTES,
			         "too many neurons for the preallocated generator scratch (at most 130,023,424)" );
			char * base = static_cast<char *>( scratch() );
			nnz = generate_into( gc, n, width, seed, d_row_ptr, d_col, col_cap, reinterpret_cast<util::gen::conn *>( base ),
			                     reinterpret_cast<int64_t *>( base + GRAPH_SCRATCH_CONN_BYTES ), nullptr );
		}
		else
		{
			dbuf<util::gen::conn> cs( std::max<std::size_t>( gc.size(), 1 ) );
			dbuf<int64_t> ts( std::size_t( graph_scratch_tiles( n ) ) );
			nnz = generate_into( gc, n, width, seed, d_row_ptr, d_col, col_cap, cs.data(), ts.data(), nullptr );
		}
		if( out_nnz ) *out_nnz = uint64_t( nnz );
		if( out_width ) *out_width = uint64_t( width );
		require( d_col == nullptr || uint64_t( nnz ) <= col_cap, "column buffer smaller than the graph (its size is in *out_nnz)" );
	} );
}
} // namespace spice::hipx

extern "C" {
const char * spice_last_error( void ) { return g_err.c_str(); }

const char * spice_version( void )
{
	return "spice_hip gfx950 (hipcc " __VERSION__ ", -O3 -ffp-contract=off, wave64)";
}

int spice_layout_flatten(
    uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns, spice_edge * out, int32_t * nout, uint64_t * n, uint64_t * maxdeg )
{
	return guard( [&] {
		require( out || nconns == 0, "null output" );
		auto l = make_layout( pops, npops, conns, nconns );
		int32_t i = 0;
		for( auto const & e : l.connections() )
			out[i++] = { std::get<0>( e ), std::get<1>( e ), std::get<2>( e ), std::get<3>( e ), std::get<4>( e ) };
		if( nout ) *nout = i;
		if( n ) *n = l.size();
		if( maxdeg ) *maxdeg = l.max_degree();
	} );
}

int spice_model_info( int32_t model, int32_t * nna, int32_t * nsa, int32_t * ntypes, int32_t * stypes )
{
	return guard( [&] {
		dispatch_model( model, [&]( auto m ) {
			using M = decltype( m );
			using Nn = typename M::neuron;
			using Ss = typename M::synapse;
			if( nna ) *nna = int32_t( Nn::size );
			if( nsa ) *nsa = int32_t( Ss::size );
			util::for_n<Nn::size>( [&]( auto I ) {
				if( ntypes ) ntypes[I] = std::is_floating_point_v<typename Nn::template type<I>> ? SPICE_ATTR_F32 : SPICE_ATTR_I32;
			} );
			util::for_n<Ss::size>( [&]( auto I ) {
				if( stypes ) stypes[I] = std::is_floating_point_v<typename Ss::template type<I>> ? SPICE_ATTR_F32 : SPICE_ATTR_I32;
			} );
			return 0;
		} );
	} );
}

int spice_hip_create(
    int32_t model,
    uint64_t const * pops,
    int32_t npops,
    spice_connection const * conns,
    int32_t nconns,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    spice_net ** out )
{
	return guard( [&] {
		require( out != nullptr, "null output handle" );
		auto l = make_layout( pops, npops, conns, nconns );
		require( dt > 0.0f, "dt must be positive" );
		require( delay >= 1, "delay must be >= 1" );
		hipStream_t s = new_stream();
		csr_dev g;
		try
		{
			g = generate_csr( gen_conns( l ), int64_t( l.size() ), int64_t( l.max_degree() ), seed, s );
		}
		catch( ... )
		{
			(void)hipStreamDestroy( s );
			throw;
		}
		*out = new spice_net{ make_net( { model, dt, delay, seed, flags, { 0, 0, 0, 0 } }, std::move( g ), s ) };
	} );
}

int spice_hip_generate_graph( uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns, uint64_t seed,
                              int64_t * d_row_ptr, int32_t * d_col, uint64_t col_cap, uint64_t * out_nnz, uint64_t * out_width )
{
	return generate_graph_entry( nullptr, pops, npops, conns, nconns, seed, d_row_ptr, d_col, col_cap, out_nnz, out_width );
}

int spice_hip_generate_graph_noalloc( uint64_t const * pops, int32_t npops, spice_connection const * conns, int32_t nconns,
                                      uint64_t seed, int64_t * d_row_ptr, int32_t * d_col, uint64_t col_cap, uint64_t * out_nnz,
                                      uint64_t * out_width )
{
	return generate_graph_entry( &graph_scratch, pops, npops, conns, nconns, seed, d_row_ptr, d_col, col_cap, out_nnz, out_width );
}

int spice_hip_create_desc( int32_t model, spice_net_desc const * d, float dt, uint32_t flags, spice_net ** out )
{
	return guard( [&] {
		require( out != nullptr && d != nullptr, "null argument" );
		auto l = make_layout( d->pops, d->npops, d->conns, d->nconns );
		require( dt > 0.0f, "dt must be positive" );
		require( d->delay >= 1, "delay must be >= 1" );
		hipStream_t s = new_stream();
		csr_dev g;
		try
		{
			g = generate_csr( gen_conns( l ), int64_t( l.size() ), int64_t( l.max_degree() ), d->seed, s );
		}
		catch( ... )
		{
			(void)hipStreamDestroy( s );
			throw;
		}
		*out = new spice_net{ make_net( params_of( model, *d, dt, flags ), std::move( g ), s ) };
	} );
}

int spice_batch_create( int32_t model, int32_t nnets, spice_net_desc const * descs, float dt, uint32_t flags, spice_batch ** out )
{
	return guard( [&] {
		require( out != nullptr && descs != nullptr && nnets > 0, "invalid batch arguments" );
		require( dt > 0.0f, "dt must be positive" );
		std::vector<net_params> ps;
		std::vector<util::layout> ls;
		for( int32_t k = 0; k < nnets; k++ )
		{
			ls.push_back( make_layout( descs[k].pops, descs[k].npops, descs[k].conns, descs[k].nconns ) );
			require( descs[k].delay >= 1, "delay must be >= 1" );
			ps.push_back( params_of( model, descs[k], dt, flags ) );
		}
		hipStream_t s = new_stream();
		std::vector<csr_dev> gs;
		try
		{
			for( int32_t k = 0; k < nnets; k++ )
				gs.push_back( generate_csr( gen_conns( ls[k] ), int64_t( ls[k].size() ), int64_t( ls[k].max_degree() ), descs[k].seed, s ) );
		}
		catch( ... )
		{
			(void)hipStreamDestroy( s );
			throw;
		}
		auto * b = new spice_batch{ make_batch( model, ps, std::move( gs ), s ), {} };
		for( int k = 0; k < b->impl->size(); k++ ) b->handles.push_back( new spice_net{ b->impl->net( k ), b->impl } );
		*out = b;
	} );
}

int spice_batch_destroy( spice_batch * b )
{
	return guard( [&] {
		if( !b ) return;
		for( auto * h : b->handles ) delete h;
		delete b->impl;
		delete b;
	} );
}

int spice_batch_run( spice_batch * b, int64_t n )
{
	return guard( [&] {
		require( b && b->impl, "null batch handle" );
		b->impl->run( n );
	} );
}

int spice_batch_synchronize( spice_batch * b )
{
	return guard( [&] {
		require( b && b->impl, "null batch handle" );
		b->impl->synchronize();
	} );
}

int spice_batch_size( spice_batch * b, int32_t * n )
{
	return guard( [&] {
		require( b && b->impl && n, "null argument" );
		*n = b->impl->size();
	} );
}

int spice_batch_net( spice_batch * b, int32_t i, spice_net ** out )
{
	return guard( [&] {
		require( b && b->impl && out, "null argument" );
		require( i >= 0 && i < int32_t( b->handles.size() ), "net index out of range" );
		*out = b->handles[i];
	} );
}

int spice_hip_create_from_csr(
    int32_t model,
    int64_t const * row_ptr,
    int32_t const * col,
    uint64_t n,
    uint64_t width,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    spice_net ** out )
{
	return guard( [&] {
		require( out != nullptr, "null output handle" );
		require( row_ptr != nullptr && n > 0 && n < uint64_t( INT_MAX ), "invalid node count" );
		require( row_ptr[0] == 0, "row_ptr[0] must be 0" );
		for( uint64_t r = 0; r < n; r++ )
		{
			int64_t const len = row_ptr[r + 1] - row_ptr[r];
			require( len >= 0 && uint64_t( len ) <= width, "row longer than the adjacency width" );
		}
		require( col != nullptr || row_ptr[n] == 0, "null column array" );
		// every row sorted by target: the delivery finds a row's segment inside a target tile by
		// binary search (the tile offsets), so an unsorted row would put targets outside it
		for( uint64_t r = 0; r < n; r++ )
			for( int64_t e = row_ptr[r]; e < row_ptr[r + 1]; e++ )
			{
				require( col[e] >= 0 && uint64_t( col[e] ) < n, "neighbor id out of range" );
				require( e == row_ptr[r] || col[e] >= col[e - 1], "adjacency rows must be sorted by neighbor id" );
			}
		require( dt > 0.0f && delay >= 1, "dt must be positive and delay >= 1" );
		hipStream_t s = new_stream();
		csr_dev g;
		try
		{
			g = upload_csr( row_ptr, col, int64_t( n ), int64_t( width ), s );
		}
		catch( ... )
		{
			(void)hipStreamDestroy( s );
			throw;
		}
		*out = new spice_net{ make_net( { model, dt, delay, seed, flags, { 0, 0, 0, 0 } }, std::move( g ), s ) };
	} );
}

int spice_hip_create_from_ell(
    int32_t model, int32_t const * edges, uint64_t n, uint64_t width, float dt, int32_t delay, uint64_t seed, uint32_t flags, spice_net ** out )
{
	std::vector<int64_t> rp;
	std::vector<int32_t> col;
	int const st = guard( [&] {
		require( width > 0, "width must be positive" );
		require( width % 32 == 0, "width must be a multiple of 32" ); // reference: cuda/snn.d.cpp:103-107
		require( edges != nullptr && n > 0, "invalid adjacency" );
		rp.assign( n + 1, 0 );
		for( uint64_t r = 0; r < n; r++ )
		{
			for( uint64_t j = 0; j < width && edges[r * width + j] != INT_MAX; j++ ) col.push_back( edges[r * width + j] );
			rp[r + 1] = int64_t( col.size() );
		}
	} );
	if( st != SPICE_OK ) return st;
	return spice_hip_create_from_csr( model, rp.data(), col.data(), n, width, dt, delay, seed, flags, out );
}

int spice_hip_destroy( spice_net * net )
{
	return guard( [&] {
		if( !net ) return;
		require( net->owner == nullptr, "a net that belongs to a batch is released with the batch" );
		delete net->impl;
		delete net;
	} );
}

int spice_hip_step( spice_net * net, int32_t ** dspk, uint32_t ** dcnt )
{
	return guard( [&] { stepping_impl( net )->step( dspk, dcnt ); } );
}
int spice_hip_copy_spikes( spice_net * net, int32_t * out, uint64_t cap, uint64_t * count )
{
	return guard( [&] {
		require( count != nullptr, "null count" );
		impl( net )->copy_spikes( out, cap, count );
	} );
}
int spice_hip_run( spice_net * net, int64_t n )
{
	return guard( [&] { stepping_impl( net )->run( n ); } );
}
int spice_hip_run_profiled( spice_net * net, int64_t n, double * kernel_ms )
{
	return guard( [&] {
		require( kernel_ms != nullptr, "null output" );
		stepping_impl( net )->run_profiled( n, kernel_ms );
	} );
}
int spice_hip_synchronize( spice_net * net )
{
	return guard( [&] { impl( net )->synchronize(); } );
}

int spice_hip_num_neurons( spice_net * net, uint64_t * out )
{
	return guard( [&] { *out = uint64_t( impl( net )->num_neurons() ); } );
}
int spice_hip_num_synapses( spice_net * net, uint64_t * out )
{
	return guard( [&] { *out = uint64_t( impl( net )->num_neurons() * impl( net )->width() ); } );
}
int spice_hip_width( spice_net * net, uint64_t * out )
{
	return guard( [&] { *out = uint64_t( impl( net )->width() ); } );
}
int spice_hip_num_edges( spice_net * net, uint64_t * out )
{
	return guard( [&] { *out = uint64_t( impl( net )->num_edges() ); } );
}
int spice_hip_dt( spice_net * net, float * out )
{
	return guard( [&] { *out = impl( net )->params().dt; } );
}
int spice_hip_delay( spice_net * net, int32_t * out )
{
	return guard( [&] { *out = impl( net )->params().delay; } );
}
int spice_hip_model( spice_net * net, int32_t * out )
{
	return guard( [&] { *out = impl( net )->params().model; } );
}
int spice_hip_get_clock( spice_net * net, int64_t * step, float * sum, float * c )
{
	return guard( [&] { impl( net )->get_clock( step, sum, c ); } );
}
int spice_hip_set_clock( spice_net * net, int64_t step, float sum, float c )
{
	return guard( [&] { stepping_impl( net )->set_clock( step, sum, c ); } );
}
int spice_hip_get_stats( spice_net * net, spice_stats * out )
{
	return guard( [&] {
		require( out != nullptr, "null output" );
		impl( net )->get_stats( out );
	} );
}
int spice_hip_reset_stats( spice_net * net )
{
	return guard( [&] { impl( net )->reset_stats(); } );
}

namespace
{
// SplitMix64 output function (S. Vigna's published finalizer): a pseudo-random 64-bit value
// per access index.
__device__ inline std::uint64_t mix64( std::uint64_t z )
{
	z = ( z ^ ( z >> 30 ) ) * 0xbf58476d1ce4e5b9ull;
	z = ( z ^ ( z >> 27 ) ) * 0x94d049bb133111ebull;
	return z ^ ( z >> 31 );
}

// `count` 4-byte accesses at pseudo-random places of a[0 .. n): a load, and with rmw the store
// of the changed value after it -- the plasticity post pass's access to a weight, without its
// index and trace gathers. 4 accesses per thread in flight.
__global__ __launch_bounds__( 256 ) void k_probe_lines( float * a, std::uint64_t n, std::uint64_t count, std::uint64_t seed, int rmw )
{
	constexpr int U = 4;
	std::uint64_t const nthreads = std::uint64_t( gridDim.x ) * blockDim.x;
	float sink = 0.0f;
	for( std::uint64_t i0 = std::uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x; i0 < count; i0 += U * nthreads )
	{
		std::uint64_t idx[U];
		float x[U];
#pragma unroll
		for( int u = 0; u < U; u++ ) idx[u] = mix64( i0 + u * nthreads + seed ) % n;
#pragma unroll
		for( int u = 0; u < U; u++ ) x[u] = i0 + u * nthreads < count ? a[idx[u]] : 0.0f;
#pragma unroll
		for( int u = 0; u < U; u++ )
		{
			if( i0 + u * nthreads >= count ) continue;
			if( rmw )
				a[idx[u]] = x[u] + 1.0f;
			else
				sink += x[u];
		}
	}
	if( sink < 0.0f ) a[0] = sink; // never (the array holds non-negative values): keeps the loads
}
} // namespace

int spice_hip_probe_random_lines( uint64_t bytes, uint64_t accesses, int32_t rmw, int32_t reps, double * out_us )
{
	return guard( [&] {
		require( out_us != nullptr, "null output" );
		require( bytes >= 64 && bytes <= ( std::uint64_t( 64 ) << 30 ), "bytes must be in [64, 64 GiB]" );
		require( accesses >= 1 && accesses <= ( std::uint64_t( 1 ) << 32 ) && reps >= 1 && reps <= 10000, "invalid access or repetition count" );
		std::uint64_t const n = bytes / 4;
		dbuf<float> a( n );
		hipEvent_t e0 = nullptr, e1 = nullptr;
		try
		{
			SPICE_HIP( hipMemset( a.data(), 0, a.bytes() ) );
			SPICE_HIP( hipEventCreate( &e0 ) );
			SPICE_HIP( hipEventCreate( &e1 ) );
			unsigned const grid = unsigned( 8 * device_cu_count() );
			SPICE_LAUNCH( k_probe_lines<<<grid, 256>>>( a.data(), n, accesses, 0, rmw ) ); // warm
			SPICE_HIP( hipEventRecord( e0, nullptr ) );
			for( int r = 0; r < reps; r++ ) SPICE_LAUNCH( k_probe_lines<<<grid, 256>>>( a.data(), n, accesses, std::uint64_t( r + 1 ) << 40, rmw ) );
			SPICE_HIP( hipEventRecord( e1, nullptr ) );
			SPICE_HIP( hipEventSynchronize( e1 ) );
			float ms = 0.0f;
			SPICE_HIP( hipEventElapsedTime( &ms, e0, e1 ) );
			*out_us = double( ms ) * 1e3 / reps;
		}
		catch( ... )
		{
			if( e0 ) (void)hipEventDestroy( e0 );
			if( e1 ) (void)hipEventDestroy( e1 );
			throw;
		}
		(void)hipEventDestroy( e0 );
		(void)hipEventDestroy( e1 );
	} );
}
int spice_hip_export_adj( spice_net * net, int32_t * out, uint64_t cap )
{
	return guard( [&] {
		require( out != nullptr, "null output" );
		impl( net )->export_adj( out, cap );
	} );
}
int spice_hip_export_csr( spice_net * net, int64_t * row_ptr, int32_t * col )
{
	return guard( [&] {
		require( row_ptr != nullptr && ( col != nullptr || impl( net )->num_edges() == 0 ), "null output" );
		impl( net )->export_csr( row_ptr, col );
	} );
}
int spice_hip_export_csr_rows( spice_net * net, uint64_t row_first, uint64_t row_last, int64_t * row_ptr, int32_t * col, uint64_t col_cap )
{
	return guard( [&] {
		require( row_ptr != nullptr, "null output" );
		uint64_t const n = uint64_t( impl( net )->num_neurons() );
		require( row_first <= row_last && row_last <= n, "row range out of bounds" );
		impl( net )->export_csr_rows( int64_t( row_first ), int64_t( row_last ), row_ptr, col, col_cap );
	} );
}
int spice_hip_export_neuron_attr( spice_net * net, int32_t a, void * out, uint64_t bytes )
{
	return guard( [&] {
		require( out != nullptr, "null output" );
		impl( net )->export_neuron_attr( a, out, bytes );
	} );
}
int spice_hip_import_neuron_attr( spice_net * net, int32_t a, void const * in, uint64_t bytes )
{
	return guard( [&] {
		require( in != nullptr, "null input" );
		impl( net )->import_neuron_attr( a, in, bytes );
	} );
}
int spice_hip_export_synapse_attr( spice_net * net, int32_t a, void * out, uint64_t bytes, int32_t ell )
{
	return guard( [&] {
		require( out != nullptr || bytes == 0, "null output" );
		impl( net )->export_synapse_attr( a, out, bytes, ell != 0 );
	} );
}
int spice_hip_import_synapse_attr( spice_net * net, int32_t a, void const * in, uint64_t bytes, int32_t ell )
{
	return guard( [&] {
		require( in != nullptr || bytes == 0, "null input" );
		impl( net )->import_synapse_attr( a, in, bytes, ell != 0 );
	} );
}

// ---- serving ----
static job_spec job_spec_of( spice_job_desc const & j )
{
	require( j.dt > 0.0f, "dt must be positive" );
	require( j.net.delay >= 1, "delay must be >= 1" );
	require( j.steps >= 1, "a job runs at least one step" );
	require( ( j.flags & ~std::uint32_t( SPICE_FLAG_FIXED_POINT ) ) == 0, "unknown flags" );
	util::layout const l = make_layout( j.net.pops, j.net.npops, j.net.conns, j.net.nconns );
	require( l.size() >= 1, "a job needs at least one neuron" );
	job_spec s;
	s.p = params_of( j.model, j.net, j.dt, j.flags );
	bool const plastic = dispatch_model( j.model, []( auto m ) { return is_plastic_v<decltype( m )>; } );
	require( !plastic || j.net.delay <= HISTORY - 1, "plastic models support delays of at most 255 steps" );
	s.conns = gen_conns( l );
	s.n = std::int64_t( l.size() );
	s.width = std::int64_t( l.max_degree() );
	return s;
}

static std::uint64_t job_estimate( job_spec const & s, std::int64_t chunk )
{
	return estimate_job_bytes( s.p.model, s.n, s.width, s.p.delay, ( s.p.flags & SPICE_FLAG_FIXED_POINT ) != 0, chunk, s.p.dt );
}

static serve::scheduler<hip_exec> & sched_of( spice_server * sv )
{
	require( sv && sv->sched, "null server handle" );
	return *sv->sched;
}

int spice_server_create( uint64_t budget_bytes, int64_t chunk_steps, spice_spike_sink sink, void * user, spice_server ** out )
{
	return guard( [&] {
		require( out != nullptr, "null output handle" );
		require( budget_bytes > 0 && chunk_steps >= 1, "budget and chunk length must be positive" );
		auto sv = std::make_unique<spice_server>();
		sv->chunk = chunk_steps;
		sv->exec = std::make_unique<hip_exec>( sink, user, chunk_steps );
		sv->sched = std::make_unique<serve::scheduler<hip_exec>>( *sv->exec, budget_bytes, chunk_steps );
		*out = sv.release();
	} );
}

int spice_server_destroy( spice_server * sv )
{
	return guard( [&] {
		if( !sv ) return;
		sv->sched.reset();
		sv->exec.reset();
		delete sv;
	} );
}

int spice_server_estimate( spice_job_desc const * job, int64_t chunk_steps, uint64_t * out_bytes )
{
	return guard( [&] {
		require( job && out_bytes && chunk_steps >= 1, "invalid arguments" );
		*out_bytes = job_estimate( job_spec_of( *job ), chunk_steps );
	} );
}

int spice_server_submit( spice_server * sv, spice_job_desc const * job, uint64_t * out_id )
{
	return guard( [&] {
		auto & sc = sched_of( sv );
		require( job && out_id, "null argument" );
		std::string const f = sc.fatal();
		if( !f.empty() ) throw std::runtime_error( "server stopped after an error: " + f );
		job_spec s = job_spec_of( *job );
		std::uint64_t const est = job_estimate( s, sv->chunk );
		*out_id = sc.submit( std::move( s ), job->steps, est );
	} );
}

int spice_server_wait( spice_server * sv, uint64_t id, double timeout_s, spice_job_result * out, int32_t * out_finished )
{
	return guard( [&] {
		auto & sc = sched_of( sv );
		serve::job_record r;
		bool const fin = sc.wait( id, timeout_s, r );
		if( out_finished ) *out_finished = fin ? 1 : 0;
		if( out )
		{
			*out = spice_job_result{};
			out->id = r.id;
			out->status = r.status;
			out->steps = r.steps;
			out->done = r.done;
			out->synapses = r.synapses;
			out->device_bytes = r.bytes;
			out->reserved_bytes = r.est_bytes;
			out->spikes = r.spikes;
			out->events = r.events;
			out->plastic_events = r.plastic_events;
			out->digest = r.digest;
			out->t_submit = r.t_submit;
			out->t_admit = r.t_admit;
			out->t_done = r.t_done;
		}
	} );
}

int spice_server_cancel( spice_server * sv, uint64_t id )
{
	return guard( [&] { sched_of( sv ).cancel( id ); } );
}

int spice_server_drain( spice_server * sv )
{
	return guard( [&] { sched_of( sv ).drain(); } );
}

int spice_server_get_info( spice_server * sv, spice_server_info * out )
{
	return guard( [&] {
		auto & sc = sched_of( sv );
		require( out != nullptr, "null output" );
		serve::server_info const i = sc.info();
		*out = { i.budget,    i.reserved, i.peak_reserved, i.queued,  i.resident, i.finished, i.chunks,
		         i.rebuilds,  i.steps_launched, i.busy_s,  sc.now(), i.launch_s, i.admit_s,  i.wait_s,
		         i.collect_s };
	} );
}

int spice_server_job_error( spice_server * sv, uint64_t id, char * buf, uint64_t cap )
{
	return guard( [&] {
		auto & sc = sched_of( sv );
		require( buf != nullptr && cap > 0, "null output" );
		serve::job_record r;
		sc.wait( id, 0.0, r );
		std::size_t const n = std::min<std::size_t>( r.error.size(), std::size_t( cap - 1 ) );
		std::memcpy( buf, r.error.data(), n );
		buf[n] = 0;
	} );
}

// ---- launch shapes and the autotuner ----
int spice_hip_get_launch( spice_net * net, spice_launch * out )
{
	return guard( [&] {
		require( out != nullptr, "null output" );
		launch_cfg const c = impl( net )->get_launch();
		*out = { c.tile_w, c.flush_slots, c.flush_grid, c.recv_block, c.recv_parts, { 0, 0, 0 } };
	} );
}

int spice_hip_set_launch( spice_net * net, spice_launch const * in )
{
	return guard( [&] {
		require( in != nullptr, "null input" );
		stepping_impl( net )->set_launch( { in->tile_w, in->flush_slots, in->flush_grid, in->recv_block, in->recv_parts } );
	} );
}

namespace
{
using tune_key = std::tuple<int, std::int64_t, std::int64_t, int, std::uint32_t>;
std::mutex g_tune_mutex;
std::map<tune_key, std::pair<launch_cfg, double>> g_tune_cache;
} // namespace

int spice_hip_autotune( spice_net * net, int64_t warmup, int64_t steps, spice_launch * out, double * out_us )
{
	return guard( [&] {
		net_iface * n = stepping_impl( net );
		require( warmup >= 0 && steps >= 1, "warmup >= 0 and steps >= 1 required" );
		net_params const & p = n->params();
		tune_key const key{ p.model, n->num_neurons(), n->num_edges(), p.delay, p.flags };
		std::pair<launch_cfg, double> best;
		bool hit = false;
		{
			std::lock_guard<std::mutex> g( g_tune_mutex );
			auto it = g_tune_cache.find( key );
			if( it != g_tune_cache.end() )
			{
				best = it->second;
				hit = true;
			}
		}
		if( !hit )
		{
			std::unique_ptr<net_iface> c( n->clone() );
			c->run( warmup );
			c->synchronize();
			// wall time per step (ms) of `steps` graph-replayed steps under cfg: a step's
			// plasticity work runs beside the next step's delivery, so shapes are evaluated together
			auto time_of = [&]( launch_cfg cfg ) {
				c->set_launch( cfg );
				c->run( 32 ); // capture + warm the graph
				c->synchronize();
				auto const t0 = std::chrono::steady_clock::now();
				c->run( steps );
				c->synchronize();
				return std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now() - t0 ).count() / double( steps );
			};
			launch_cfg cfg = c->get_launch();
			double best_t = time_of( cfg );
			std::vector<std::pair<double, launch_cfg>> timed{ { best_t, cfg } };
			std::int64_t const N = n->num_neurons();
			launch_cfg const start = cfg;
			auto trial = [&]( launch_cfg const & x ) {
				double const t = time_of( x );
				timed.push_back( { t, x } );
				if( t < best_t )
				{
					best_t = t;
					cfg = x;
				}
			};
			// delivery: tile width x workgroup size, then the split of each tile's spikes
			for( int bs : { 256, 512, 1024 } )
				for( int tw : { 128, 256, 512, 1024, 2048, 4096, 8192 } )
				{
					std::int64_t const tiles = ( N + tw - 1 ) / tw;
					if( ( tw == start.tile_w && bs == start.recv_block ) || N * ( tiles + 1 ) * 4 > ( std::int64_t( 4 ) << 30 ) ||
					    tw > n->max_tile_w() )
						continue; // the start shape, pivots above 4 GB, or wider than the net's tiles
					launch_cfg x = start;
					x.tile_w = tw;
					x.recv_block = bs;
					trial( x );
				}
			{
				launch_cfg const base = cfg;
				for( int parts : { 2, 3, 4, 6 } )
					for( int bs : { 256, 512, 1024 } )
					{
						launch_cfg x = base;
						x.recv_parts = parts;
						x.recv_block = bs;
						trial( x );
					}
			}
			if( n->is_plastic() )
			{
				int cus = 0;
				{
					int dev = 0;
					SPICE_HIP( hipGetDevice( &dev ) );
					SPICE_HIP( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) );
				}
				launch_cfg const base = cfg;
				for( int fs : { 64, 128, 256, 512, 1024 } )
					for( int fg : { 1, 2, 4, 8, 16 } )
					{
						launch_cfg x = base;
						x.flush_slots = fs;
						x.flush_grid = fg * cus;
						if( x.flush_slots == base.flush_slots && x.flush_grid == base.flush_grid ) continue;
						trial( x );
					}
			}
			// the copy keeps stepping, so each shape was timed on other activity (a bursting
			// network's step time moves by 2x and more): the three fastest shapes are timed
			// again, in turn, three times each, and the fastest in total is kept
			std::sort( timed.begin(), timed.end(), []( auto const & a, auto const & b ) { return a.first < b.first; } );
			std::vector<launch_cfg> top;
			for( auto const & tx : timed )
			{
				launch_cfg const & x = tx.second;
				bool seen = false;
				for( launch_cfg const & y : top )
					seen = seen || ( x.tile_w == y.tile_w && x.flush_slots == y.flush_slots && x.flush_grid == y.flush_grid &&
					                 x.recv_block == y.recv_block && x.recv_parts == y.recv_parts );
				if( !seen ) top.push_back( x );
				if( top.size() == 3 ) break;
			}
			if( top.size() > 1 )
			{
				std::vector<double> sum( top.size(), 0.0 );
				for( int r = 0; r < 3; r++ )
					for( std::size_t i = 0; i < top.size(); i++ ) sum[i] += time_of( top[i] );
				std::size_t const k = std::size_t( std::min_element( sum.begin(), sum.end() ) - sum.begin() );
				cfg = top[k];
				best_t = sum[k] / 3.0;
			}
			best = { cfg, best_t * 1e3 };
			std::lock_guard<std::mutex> g( g_tune_mutex );
			g_tune_cache[key] = best;
		}
		n->set_launch( best.first );
		if( out ) *out = { best.first.tile_w, best.first.flush_slots, best.first.flush_grid, best.first.recv_block, best.first.recv_parts, { 0, 0, 0 } };
		if( out_us ) *out_us = best.second;
	} );
}
}
