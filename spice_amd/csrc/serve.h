#pragma once
// GPU executor of the serving loop (policy: scheduler.h). Every admitted job is an ordinary
// hip_net with a spike log; the resident jobs of one (model, accumulation mode) share one
// batch, rebuilt whenever the resident set changes, and the batches of different models run
// on their own streams, so one chunk of the serving loop is one batched launch sequence per
// model, all overlapping on the device. After each chunk every job's spikes (per step) go to
// the caller's sink.
#include "factory.h"
#include "scheduler.h"

#include <spice/models/all.h>

#include <algorithm>
#include <map>
#include <memory>
#include <utility>
#include <vector>

namespace spice::hipx
{
struct job_spec
{
	net_params p{};
	std::vector<util::gen::conn> conns;
	std::int64_t n = 0, width = 0;
};

// splitmix64 finalizer: the digest of a spike stream is the wrapping sum of
// mix64(step << 32 | id) over its spikes, independent of the order within a step.
inline std::uint64_t mix64( std::uint64_t x )
{
	x ^= x >> 30;
	x *= 0xbf58476d1ce4e5b9ull;
	x ^= x >> 27;
	x *= 0x94d049bb133111ebull;
	x ^= x >> 31;
	return x;
}

// Spike-log capacity of a job for chunks of `chunk` steps: one spike per neuron and 1.6 ms
// of simulated time (the models' refractory periods allow at most one per 2 ms), at least
// 1/16 spike per neuron and step, plus slack.
inline std::uint64_t spike_log_capacity( std::int64_t n, std::int64_t chunk, float dt )
{
	double const f = std::min( 1.0, std::max( 1.0 / 16, double( dt ) / 1.6e-3 ) );
	return std::uint64_t( double( n ) * double( chunk ) * f ) + 4096;
}

// Upper bound of a job's device memory (admission reservation): every buffer of a hip_net
// with E <= N * width edges, plus the spike log and per-allocation slack.
inline std::uint64_t estimate_job_bytes( int model, std::int64_t n, std::int64_t width, int delay, bool fixed, std::int64_t chunk, float dt )
{
	return dispatch_model( model, [&]( auto m ) -> std::uint64_t {
		using M = decltype( m );
		using Nn = typename M::neuron;
		using Ss = typename M::synapse;
		std::uint64_t const N = std::uint64_t( n ), E = N * std::uint64_t( width );
		std::uint64_t const cus = 256;
		std::uint64_t const tile_w = std::min<std::uint64_t>( TILE_MAX, std::max<std::uint64_t>( 256, ( ( N + 2 * cus - 1 ) / ( 2 * cus ) + 63 ) / 64 * 64 ) );
		std::uint64_t const ntiles = ( N + tile_w - 1 ) / tile_w;
		std::uint64_t b = 8 * ( N + 1 ) + 4 * E + 4 * N;                 // CSR (+ generator degrees)
		b += N * Nn::size_in_bytes + N * Nn::size * ( fixed ? 8 : 4 );    // neurons + accumulators
		b += E * Ss::size_in_bytes;                                        // synapses
		b += 4 * N * std::uint64_t( delay + 2 ) + 4 * N * ( ntiles + 1 ); // ring, pivots
		if constexpr( is_plastic_v<M> )
		{
			b += 8 * HWORDS * N + 36 * N;
			// eager: traces, flags and target lists by parity, by-target index (at most every
			// edge) with its offsets and build temporaries (the sort's second buffer and work space)
			if constexpr( has_neuron_traces_v<Ss> ) b += 17 * E + 48 * N + ( 64 << 20 );
		}
		b += 8 * NSTAT * ( N / 256 + 8 * cus + ntiles + 1 );             // statistics slots
		b += 4 * spike_log_capacity( n, chunk, dt ) + 8 * std::uint64_t( chunk ) + 64;
		b += 64 * 64 * 1024;                                               // allocation granularity
		return b;
	} );
}

class hip_exec
{
public:
	using spec = job_spec;
	struct handle
	{
		std::unique_ptr<net_iface> net;
		int model = 0;
		bool fixed = false;
	};

	hip_exec( spice_spike_sink sink, void * user, std::int64_t chunk_max )
	    : _sink( sink )
	    , _user( user )
	    , _chunk_max( chunk_max )
	{
		SPICE_HIP( hipGetDevice( &_dev ) );
	}

	~hip_exec()
	{
		_batches.clear();
		for( auto & kv : _streams ) (void)hipStreamDestroy( kv.second );
		if( _off ) (void)hipHostFree( _off );
		if( _ids ) (void)hipHostFree( _ids );
	}

	std::unique_ptr<handle> create( spec const & s, serve::job_record & r )
	{
		SPICE_HIP( hipSetDevice( _dev ) );
		hipStream_t st = new_stream();
		csr_dev g;
		try
		{
			g = generate_csr( s.conns, s.n, s.width, s.p.seed, st );
		}
		catch( ... )
		{
			(void)hipStreamDestroy( st );
			throw;
		}
		auto h = std::make_unique<handle>();
		h->net.reset( make_net( s.p, std::move( g ), st ) );
		h->model = s.p.model;
		h->fixed = ( s.p.flags & SPICE_FLAG_FIXED_POINT ) != 0;
		h->net->enable_spike_log( spike_log_capacity( s.n, _chunk_max, s.p.dt ), _chunk_max );
		h->net->synchronize();
		r.bytes = h->net->device_bytes();
		r.synapses = std::uint64_t( h->net->num_edges() );
		return h;
	}

	void run( std::vector<handle *> const & members, bool changed, std::int64_t nsteps )
	{
		SPICE_HIP( hipSetDevice( _dev ) );
		if( changed )
		{
			_batches.clear();
			std::map<std::pair<int, bool>, std::vector<net_iface *>> groups;
			for( auto * h : members ) groups[{ h->model, h->fixed }].push_back( h->net.get() );
			for( auto & kv : groups )
			{
				hipStream_t & st = _streams[kv.first];
				if( !st ) st = new_stream();
				_batches.emplace_back( make_group( kv.first.first, kv.second, st ) );
			}
		}
		for( auto & b : _batches )
		{
			b->begin_spike_log( nsteps );
			b->run( nsteps );
		}
	}

	void wait()
	{
		for( auto & b : _batches ) b->synchronize();
	}

	void collect( handle & h, serve::job_record & r, std::int64_t nsteps )
	{
		std::uint64_t const cap = spike_log_capacity( h.net->num_neurons(), _chunk_max, h.net->params().dt );
		reserve( nsteps, cap );
		std::uint64_t const total = h.net->read_spike_log( _off, _ids, cap );
		if( total > cap ) throw std::runtime_error( "spike log overflow (more than one spike per neuron and 1.6 ms)" );
		std::int64_t const t0 = h.net->host_step() - nsteps;
		std::uint64_t d = 0;
		for( std::int64_t k = 0; k < nsteps; k++ )
		{
			std::uint64_t const hi = std::uint64_t( t0 + k ) << 32;
			for( std::uint64_t j = _off[k]; j < _off[k + 1]; j++ ) d += mix64( hi | std::uint32_t( _ids[j] ) );
		}
		r.digest += d;
		r.spikes += total;
		if( _sink ) _sink( _user, r.id, t0, nsteps, _off, _ids );
	}

	void finish( handle & h, serve::job_record & r )
	{
		spice_stats st{};
		h.net->get_stats( &st );
		r.events = st.events;
		r.plastic_events = st.plastic_events;
	}

private:
	spice_spike_sink _sink;
	void * _user;
	std::int64_t _chunk_max;
	int _dev = 0;
	std::map<std::pair<int, bool>, hipStream_t> _streams;
	std::vector<std::unique_ptr<batch_iface>> _batches;
	std::uint64_t * _off = nullptr;
	std::int32_t * _ids = nullptr;
	std::uint64_t _off_n = 0, _ids_n = 0;

	// pinned host staging for the log copies
	void reserve( std::int64_t nsteps, std::uint64_t cap )
	{
		if( std::uint64_t( nsteps + 1 ) > _off_n )
		{
			std::uint64_t const n = std::uint64_t( std::max<std::int64_t>( nsteps + 1, _chunk_max + 1 ) );
			void * p = nullptr;
			SPICE_HIP( hipHostMalloc( &p, n * 8, hipHostMallocDefault ) );
			if( _off ) (void)hipHostFree( _off );
			_off = static_cast<std::uint64_t *>( p );
			_off_n = n;
		}
		if( cap > _ids_n )
		{
			void * p = nullptr;
			SPICE_HIP( hipHostMalloc( &p, cap * 4, hipHostMallocDefault ) );
			if( _ids ) (void)hipHostFree( _ids );
			_ids = static_cast<std::int32_t *>( p );
			_ids_n = cap;
		}
	}
};
} // namespace spice::hipx
