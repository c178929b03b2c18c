#pragma once
// Host side of one simulated network on one GPU: owns every device buffer, the HIP stream
// and the captured step graphs, and launches the kernels of kernels.h.
//
// Reference counterpart: cuda::snn<Model> (spice/cuda/snn.h:16-77, spice/cuda/snn.d.cpp).
#include "common.h"
#include "dview.h"
#include "graph.h"
#include "graphs.h"
#include "kernels.h"

#include <spice/models/model.h>
#include <spice/util/graphgen.h>
#include <spice/util/numeric.h>
#include <spice_hip.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace spice::hipx
{
struct net_params
{
	int model;
	float dt;
	int delay;
	std::uint64_t seed;
	std::uint32_t flags;
	float param[4];
};

struct launch_cfg
{
	std::int32_t tile_w = 0, flush_slots = 0, flush_grid = 0, recv_block = 0, recv_parts = 0;
};

// Model-independent interface used by the C ABI.
struct net_iface
{
	virtual ~net_iface() = default;
	virtual void step( std::int32_t ** dspk, std::uint32_t ** dcnt ) = 0;
	virtual void run( std::int64_t n ) = 0;
	virtual void run_profiled( std::int64_t n, double * kernel_ms ) = 0;
	virtual void copy_spikes( std::int32_t * out, std::uint64_t cap, std::uint64_t * count ) = 0;
	virtual void synchronize() = 0;
	virtual net_params const & params() const = 0;
	virtual std::int64_t num_neurons() const = 0;
	virtual std::int64_t width() const = 0;
	virtual std::int64_t num_edges() const = 0;
	virtual void get_clock( std::int64_t * step, float * sum, float * c ) const = 0;
	virtual void set_clock( std::int64_t step, float sum, float c ) = 0;
	virtual void get_stats( spice_stats * out ) = 0;
	virtual void reset_stats() = 0;
	virtual void export_csr( std::int64_t * row_ptr, std::int32_t * col ) = 0;
	// rows [r0, r1): r1 - r0 + 1 absolute offsets, and (col != nullptr) their target ids
	virtual void export_csr_rows( std::int64_t r0, std::int64_t r1, std::int64_t * row_ptr, std::int32_t * col, std::uint64_t col_cap ) = 0;
	virtual void export_adj( std::int32_t * ell, std::uint64_t cap ) = 0;
	virtual void export_neuron_attr( int a, void * out, std::uint64_t bytes ) = 0;
	virtual void import_neuron_attr( int a, void const * in, std::uint64_t bytes ) = 0;
	virtual void export_synapse_attr( int a, void * out, std::uint64_t bytes, bool ell ) = 0;
	virtual void import_synapse_attr( int a, void const * in, std::uint64_t bytes, bool ell ) = 0;

	// Spike log (served jobs). enable_spike_log allocates a log of `cap` ids for windows of at
	// most `max_steps` steps (before the first run); begin_spike_log opens a window over the
	// next `nsteps` steps; once the net has run exactly those steps, read_spike_log copies the
	// window out: off[0..nsteps] (off[k] = first id of step t0+k, off[nsteps] = total) and the
	// ids, and returns the total (> cap: the log overflowed, ids past cap were dropped). No step
	// of the net may be in flight during begin/read.
	virtual void enable_spike_log( std::uint64_t cap, std::int64_t max_steps ) = 0;
	virtual void begin_spike_log( std::int64_t nsteps ) = 0;
	virtual std::uint64_t read_spike_log( std::uint64_t * off, std::int32_t * ids, std::uint64_t cap ) = 0;
	virtual std::uint64_t device_bytes() const = 0;
	virtual std::int64_t host_step() const = 0;

	// Launch shapes (see hip_net::apply_launch) and a copy of the net (graph, neuron and
	// synapse state as stored; clock and plasticity bookkeeping start afresh) for tuning them.
	virtual launch_cfg get_launch() const = 0;
	virtual void set_launch( launch_cfg c ) = 0;
	virtual net_iface * clone() const = 0;
	virtual bool is_plastic() const = 0;
	virtual int max_tile_w() const = 0; // widest tile_w set_launch accepts
};

inline int device_cu_count()
{
	static int cus = [] {
		int dev = 0;
		SPICE_HIP( hipGetDevice( &dev ) );
		hipDeviceProp_t p;
		SPICE_HIP( hipGetDeviceProperties( &p, dev ) );
		return p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
	}();
	return cus;
}

template <typename Model>
class hip_net final : public net_iface
{
	using N = typename Model::neuron;
	using S = typename Model::synapse;
	static constexpr bool plastic = is_plastic_v<Model>;
	static constexpr bool traced = plastic && has_neuron_traces_v<S>;
	static constexpr bool nodefer = eager_nodefer_v<Model>; // every eager event applied by the passes

public:
	hip_net( net_params const & p, csr_dev && g, hipStream_t stream )
	    : _p( p )
	    , _g( std::move( g ) )
	    , _s( stream )
	{
		if( !( p.dt > 0.0f ) ) throw std::invalid_argument( "dt must be positive" );
		if( p.flags & ~std::uint32_t( SPICE_FLAG_FIXED_POINT ) ) throw std::invalid_argument( "unknown flags" );
		if( p.delay < 1 ) throw std::invalid_argument( "delay must be >= 1" );
		if( plastic && p.delay > HISTORY - 1 )
			throw std::invalid_argument( "plastic models support delays of at most 255 steps" );
		if( _g.n < 1 || _g.n > INT_MAX - 1024 ) throw std::invalid_argument( "invalid neuron count" );
		static_assert( N::size <= MAXA && S::size <= MAXA, "at most 16 neuron/synapse attributes" );
		// the update of step T runs beside the delivery of step T + 1 (see k_step_overlap)
		_overlap = overlap_capable_v<Model> && p.delay >= 2;
		allocate();
		initialize();
	}

	~hip_net() override
	{
		if( _s ) (void)hipStreamSynchronize( _s );
		_graphs.clear();
		if( _s ) (void)hipStreamDestroy( _s );
	}

	// ---- simulation ----
	void step( std::int32_t ** dspk, std::uint32_t ** dcnt ) override
	{
		std::int64_t const t = _step;
		launch_steps( t, 1 );
		tick();
		int const slot = int( t % ( _p.delay + 2 ) );
		if( dspk ) *dspk = _ring.data() + std::int64_t( slot ) * _g.n;
		if( dcnt ) *dcnt = _counts.data() + slot;
	}

	// Replays captured step graphs (all of them captured at the first run; see graphs.h).
	void run( std::int64_t n ) override
	{
		if( n < 0 ) throw std::invalid_argument( "negative step count" );
		if( n == 0 ) return;
		if( _graphs.empty() ) _graphs.capture( _s, [&]( int t, int len ) { launch_steps( t, len ); } );
		_graphs.run( _s, _step, n, [&] { tick(); } );
	}

	// Runs n steps eagerly, each kernel with HIP events attached to its dispatch on the net's
	// stream, and adds the summed kernel durations (ms) to kernel_ms[0..2]: receive (delivery,
	// with the previous step's plasticity), update, and 0; overlapped steps: the first delivery,
	// the last update (with the passes of the step before it), and the n - 1 overlapped step
	// kernels; fused steps: 0, 0 and the n step kernels.
	void run_profiled( std::int64_t n, double * kernel_ms ) override
	{
		if( n < 0 ) throw std::invalid_argument( "negative step count" );
		if( _fused )
		{
			if( fixed() )
				run_profiled_fused<true>( n, kernel_ms );
			else
				run_profiled_fused<false>( n, kernel_ms );
			return;
		}
		if( _overlap )
		{
			if( fixed() )
				run_profiled_overlap<true>( n, kernel_ms );
			else
				run_profiled_overlap<false>( n, kernel_ms );
			return;
		}
		constexpr int batch = 128, ev_per_step = 4;
		synchronize(); // profiled steps run serially on the main stream
		std::vector<hipEvent_t> ev( batch * ev_per_step, nullptr );
		try
		{
			for( auto & e : ev ) SPICE_HIP( hipEventCreate( &e ) );
			for( std::int64_t done = 0; done < n; )
			{
				int const k = int( std::min<std::int64_t>( batch, n - done ) );
				for( int i = 0; i < k; i++ )
				{
					hipEvent_t * e = &ev[i * ev_per_step];
					if( fixed() )
						launch_step_events<true>( e, _vs[_step & ( CLOCK_SLOTS - 1 )] );
					else
						launch_step_events<false>( e, _vs[_step & ( CLOCK_SLOTS - 1 )] );
					tick();
				}
				SPICE_HIP( hipStreamSynchronize( _s ) );
				for( int i = 0; i < k; i++ )
				{
					hipEvent_t * e = &ev[i * ev_per_step];
					for( int q = 0; q < 2; q++ ) // kernel_ms[2] stays 0: plasticity runs inside the delivery kernel
					{
						float ms = 0;
						SPICE_HIP( hipEventElapsedTime( &ms, e[2 * q], e[2 * q + 1] ) );
						kernel_ms[q] += ms;
					}
				}
				done += k;
			}
		}
		catch( ... )
		{
			for( auto & e : ev )
				if( e ) (void)hipEventDestroy( e );
			throw;
		}
		for( auto & e : ev ) (void)hipEventDestroy( e );
	}

	void copy_spikes( std::int32_t * out, std::uint64_t cap, std::uint64_t * count ) override
	{
		if( _step == 0 )
		{
			*count = 0;
			return;
		}
		int const slot = int( ( _step - 1 ) % ( _p.delay + 2 ) );
		std::uint32_t c = 0;
		SPICE_HIP( hipMemcpyAsync( &c, _counts.data() + slot, 4, hipMemcpyDeviceToHost, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) );
		*count = c;
		if( out && c )
		{
			if( c > cap ) throw std::invalid_argument( "spike buffer too small" );
			SPICE_HIP( hipMemcpyAsync( out, _ring.data() + std::int64_t( slot ) * _g.n, std::size_t( c ) * 4, hipMemcpyDeviceToHost, _s ) );
			SPICE_HIP( hipStreamSynchronize( _s ) );
		}
	}

	void synchronize() override { SPICE_HIP( hipStreamSynchronize( _s ) ); }

	// ---- queries ----
	net_params const & params() const override { return _p; }
	std::int64_t num_neurons() const override { return _g.n; }
	std::int64_t width() const override { return _g.width; }
	std::int64_t num_edges() const override { return _g.nnz; }

	void get_clock( std::int64_t * step, float * sum, float * c ) const override
	{
		*step = _step;
		*sum = float( _clock );
		*c = _clock.compensation();
	}

	// Moves the clock (conversion constructor / resume). Spikes in flight and the lazy
	// plasticity bookkeeping are reset: synapse state is taken as current through step-1.
	void set_clock( std::int64_t step, float sum, float c ) override
	{
		if( step < 0 || step > INT_MAX - 2 ) throw std::invalid_argument( "invalid step index" );
		synchronize();
		ctl_t ctl{ step, sum, c, std::int32_t( step % ( _p.delay + 2 ) ) };
		SPICE_HIP( hipMemcpy( _ctl.data() + ( step & ( CLOCK_SLOTS - 1 ) ), &ctl, sizeof( ctl ), hipMemcpyHostToDevice ) );
		SPICE_HIP( hipMemset( _counts.data(), 0, _counts.bytes() ) );
		if constexpr( plastic )
		{
			SPICE_HIP( hipMemset( _hring.data(), 0, _hring.bytes() ) );
			SPICE_HIP( hipMemset( _fl_count.data(), 0, _fl_count.bytes() ) );
			reset_caught();
			if constexpr( traced )
				reset_eager();
			else
			{
				SPICE_HIP( hipMemset( _last_spk.data(), 0xff, _last_spk.bytes() ) );
				std::vector<std::int32_t> synced( _g.n, std::int32_t( step - 1 ) );
				SPICE_HIP( hipMemcpy( _synced.data(), synced.data(), _synced.bytes(), hipMemcpyHostToDevice ) );
				SPICE_HIP( hipMemset( _pend.data(), 0, _pend.bytes() ) );
			}
		}
		_step = step;
		_clock.set( sum, c );
		_dt_last = 0.0f;
	}

	void get_stats( spice_stats * out ) override
	{
		synchronize();
		std::vector<unsigned long long> b( _bstats.size() );
		SPICE_HIP( hipMemcpyAsync( b.data(), _bstats.data(), _bstats.bytes(), hipMemcpyDeviceToHost, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) );
		unsigned long long t[NSTAT] = {};
		for( std::size_t i = 0; i < b.size(); i++ ) t[i % NSTAT] += b[i];
		*out = { std::uint64_t( _step - _stats_step0 ), t[ST_SPIKES], t[ST_EVENTS], t[ST_PLASTIC], t[ST_FLUSH_ROWS],
		         t[ST_FLUSH_SYN], t[ST_FLUSH_SLOTS], t[ST_FLUSH_WRITTEN] };
	}
	void reset_stats() override
	{
		synchronize();
		SPICE_HIP( hipMemsetAsync( _bstats.data(), 0, _bstats.bytes(), _s ) );
		synchronize();
		_stats_step0 = _step;
	}

	// ---- export / import ----
	void export_csr( std::int64_t * row_ptr, std::int32_t * col ) override
	{
		synchronize();
		SPICE_HIP( hipMemcpy( row_ptr, _g.row_ptr.data(), ( _g.n + 1 ) * 8, hipMemcpyDeviceToHost ) );
		if( _g.nnz ) SPICE_HIP( hipMemcpy( col, _g.col.data(), _g.nnz * 4, hipMemcpyDeviceToHost ) );
	}

	void export_csr_rows( std::int64_t r0, std::int64_t r1, std::int64_t * row_ptr, std::int32_t * col, std::uint64_t col_cap ) override
	{
		synchronize();
		SPICE_HIP( hipMemcpy( row_ptr, _g.row_ptr.data() + r0, std::size_t( r1 - r0 + 1 ) * 8, hipMemcpyDeviceToHost ) );
		if( !col ) return;
		std::int64_t const cnt = row_ptr[r1 - r0] - row_ptr[0];
		if( std::uint64_t( cnt ) > col_cap ) throw std::invalid_argument( "column buffer too small for the row range" );
		if( cnt ) SPICE_HIP( hipMemcpy( col, _g.col.data() + row_ptr[0], std::size_t( cnt ) * 4, hipMemcpyDeviceToHost ) );
	}

	void export_adj( std::int32_t * ell, std::uint64_t cap ) override
	{
		if( cap < std::uint64_t( _g.n * _g.width ) ) throw std::invalid_argument( "adjacency buffer too small" );
		std::vector<std::int64_t> rp( _g.n + 1 );
		std::vector<std::int32_t> col( _g.nnz );
		export_csr( rp.data(), col.data() );
		for( std::int64_t r = 0; r < _g.n; r++ )
		{
			std::int64_t const len = rp[r + 1] - rp[r];
			std::copy( col.begin() + rp[r], col.begin() + rp[r + 1], ell + r * _g.width );
			std::fill( ell + r * _g.width + len, ell + ( r + 1 ) * _g.width, INT_MAX );
		}
	}

	void export_neuron_attr( int a, void * out, std::uint64_t bytes ) override
	{
		check_attr( a, int( N::size ) );
		if( bytes != std::uint64_t( _g.n ) * attr_bytes<N>( a ) ) throw std::invalid_argument( "buffer size mismatch" );
		synchronize();
		SPICE_HIP( hipMemcpy( out, _na[a].data(), bytes, hipMemcpyDeviceToHost ) );
	}
	void import_neuron_attr( int a, void const * in, std::uint64_t bytes ) override
	{
		check_attr( a, int( N::size ) );
		if( bytes != std::uint64_t( _g.n ) * attr_bytes<N>( a ) ) throw std::invalid_argument( "buffer size mismatch" );
		synchronize();
		SPICE_HIP( hipMemcpy( _na[a].data(), in, bytes, hipMemcpyHostToDevice ) );
	}

	void export_synapse_attr( int a, void * out, std::uint64_t bytes, bool ell ) override
	{
		check_attr( a, int( S::size ) );
		std::size_t const sz = attr_bytes<S>( a );
		if( bytes != ( ell ? std::uint64_t( _g.n * _g.width ) : std::uint64_t( _g.nnz ) ) * sz )
			throw std::invalid_argument( "buffer size mismatch" );
		sync_synapses();
		if( !ell )
		{
			if( _g.nnz ) SPICE_HIP( hipMemcpy( out, _sa[a].data(), bytes, hipMemcpyDeviceToHost ) );
			return;
		}
		std::vector<char> csr( _g.nnz * sz );
		if( _g.nnz ) SPICE_HIP( hipMemcpy( csr.data(), _sa[a].data(), csr.size(), hipMemcpyDeviceToHost ) );
		std::vector<std::int64_t> rp( _g.n + 1 );
		SPICE_HIP( hipMemcpy( rp.data(), _g.row_ptr.data(), rp.size() * 8, hipMemcpyDeviceToHost ) );
		char * o = static_cast<char *>( out );
		std::memset( o, 0, bytes );
		for( std::int64_t r = 0; r < _g.n; r++ )
			std::memcpy( o + r * _g.width * sz, csr.data() + rp[r] * sz, ( rp[r + 1] - rp[r] ) * sz );
	}

	void import_synapse_attr( int a, void const * in, std::uint64_t bytes, bool ell ) override
	{
		check_attr( a, int( S::size ) );
		std::size_t const sz = attr_bytes<S>( a );
		if( bytes != ( ell ? std::uint64_t( _g.n * _g.width ) : std::uint64_t( _g.nnz ) ) * sz )
			throw std::invalid_argument( "buffer size mismatch" );
		sync_synapses();
		if( !ell )
		{
			if( _g.nnz ) SPICE_HIP( hipMemcpy( _sa[a].data(), in, bytes, hipMemcpyHostToDevice ) );
			return;
		}
		std::vector<std::int64_t> rp( _g.n + 1 );
		SPICE_HIP( hipMemcpy( rp.data(), _g.row_ptr.data(), rp.size() * 8, hipMemcpyDeviceToHost ) );
		std::vector<char> csr( _g.nnz * sz );
		char const * src = static_cast<char const *>( in );
		for( std::int64_t r = 0; r < _g.n; r++ )
			std::memcpy( csr.data() + rp[r] * sz, src + r * _g.width * sz, ( rp[r + 1] - rp[r] ) * sz );
		if( _g.nnz ) SPICE_HIP( hipMemcpy( _sa[a].data(), csr.data(), csr.size(), hipMemcpyHostToDevice ) );
	}

	// ---- spike log ----
	void enable_spike_log( std::uint64_t cap, std::int64_t max_steps ) override
	{
		if( !_graphs.empty() || _slog_ctl.size() ) throw std::logic_error( "the spike log is enabled once, before the first run" );
		if( max_steps < 1 ) throw std::invalid_argument( "spike log window must hold at least one step" );
		_slog_ctl.resize( 1 );
		_slog_ids.resize( std::max<std::uint64_t>( cap, 1 ) );
		_slog_off.resize( std::size_t( max_steps ) );
		_slog_cap = cap;
		_slog_max = max_steps;
		slog_t const w{ 0, 0, cap }; // empty window
		SPICE_HIP( hipMemcpy( _slog_ctl.data(), &w, sizeof( w ), hipMemcpyHostToDevice ) );
		for( dview * v : { &_v, &_vs[0], &_vs[1], &_vs[2], &_vs[3] } )
		{
			v->slog = _slog_ctl.data();
			v->slog_ids = _slog_ids.data();
			v->slog_off = _slog_off.data();
		}
	}

	void begin_spike_log( std::int64_t nsteps ) override
	{
		if( !_slog_ctl.size() ) throw std::logic_error( "spike log not enabled" );
		if( nsteps < 0 || nsteps > _slog_max ) throw std::invalid_argument( "spike log window too long" );
		slog_t const w{ _step, nsteps, _slog_cap };
		SPICE_HIP( hipMemcpy( _slog_ctl.data(), &w, sizeof( w ), hipMemcpyHostToDevice ) );
		note_spike_log( nsteps );
	}

	// Host bookkeeping of a window whose device side a batch opened (k_batch_log_window).
	void note_spike_log( std::int64_t nsteps )
	{
		if( !_slog_ctl.size() ) return;
		if( nsteps < 0 || nsteps > _slog_max ) throw std::invalid_argument( "spike log window too long" );
		_slog_t0 = _step;
		_slog_n = nsteps;
	}

	std::uint64_t read_spike_log( std::uint64_t * off, std::int32_t * ids, std::uint64_t cap ) override
	{
		if( !_slog_ctl.size() ) throw std::logic_error( "spike log not enabled" );
		if( _step != _slog_t0 + _slog_n ) throw std::logic_error( "spike log window not completed" );
		off[0] = 0;
		if( _slog_n == 0 ) return 0;
		SPICE_HIP( hipMemcpy( off, _slog_off.data(), std::size_t( _slog_n ) * 8, hipMemcpyDeviceToHost ) );
		std::uint32_t last = 0;
		int const slot = int( ( _step - 1 ) % ( _p.delay + 2 ) );
		SPICE_HIP( hipMemcpy( &last, _counts.data() + slot, 4, hipMemcpyDeviceToHost ) );
		std::uint64_t const total = off[_slog_n - 1] + last;
		off[_slog_n] = total;
		std::uint64_t const n = std::min( { total, _slog_cap, cap } );
		if( n ) SPICE_HIP( hipMemcpy( ids, _slog_ids.data(), std::size_t( n ) * 4, hipMemcpyDeviceToHost ) );
		return total;
	}

	std::uint64_t device_bytes() const override
	{
		std::uint64_t b = _g.row_ptr.bytes() + _g.col.bytes();
		for( int a = 0; a < MAXA; a++ ) b += _na[a].bytes() + _acc[a].bytes() + _sa[a].bytes();
		b += _ring.bytes() + _counts.bytes() + _hring.bytes() + _last_spk.bytes() + _synced.bytes() + _pend.bytes() +
		     _fl_id.bytes() + _fl_synced.bytes() + _fl_pend.bytes() + _hasp.bytes() + _pl_lo.bytes() + _pl_hi.bytes() +
		     _fl_count.bytes() + _xtr.bytes() + _ytr.bytes() + _fl8.bytes() + _po_id.bytes() + _po_count.bytes() +
		     _tp_ptr.bytes() + _caught.bytes() + _tp_ent.bytes() + _ctl.bytes() + _bstats.bytes() +
		     _pivots.bytes() + _slog_ctl.bytes() + _slog_ids.bytes() + _slog_off.bytes();
		return b;
	}

	// ---- launch shapes ----
	launch_cfg get_launch() const override { return _lc; }
	void set_launch( launch_cfg c ) override
	{
		synchronize();
		apply_launch( c );
	}
	bool is_plastic() const override { return plastic; }
	bool overlapped() const { return _overlap; }
	// widest delivery tile: wide tiles for the two-kernel step of static models
	int max_tile_w() const override { return !plastic && !_overlap ? TILE_WIDE : TILE_MAX; }

	net_iface * clone() const override
	{
		// the original is only read (its lazily updated synapses are copied as they are)
		const_cast<hip_net *>( this )->synchronize();
		hipStream_t s = nullptr;
		SPICE_HIP( hipStreamCreateWithFlags( &s, hipStreamNonBlocking ) );
		try
		{
			csr_dev g;
			g.n = _g.n;
			g.nnz = _g.nnz;
			g.width = _g.width;
			g.row_ptr.resize( _g.row_ptr.size() );
			g.col.resize( _g.col.size() );
			SPICE_HIP( hipMemcpyAsync( g.row_ptr.data(), _g.row_ptr.data(), _g.row_ptr.bytes(), hipMemcpyDeviceToDevice, s ) );
			SPICE_HIP( hipMemcpyAsync( g.col.data(), _g.col.data(), _g.col.bytes(), hipMemcpyDeviceToDevice, s ) );
			SPICE_HIP( hipStreamSynchronize( s ) );
			auto * c = new hip_net( _p, std::move( g ), s );
			s = nullptr; // owned by c
			for( int a = 0; a < MAXA; a++ )
			{
				if( _na[a].bytes() ) SPICE_HIP( hipMemcpy( c->_na[a].data(), _na[a].data(), _na[a].bytes(), hipMemcpyDeviceToDevice ) );
				if( _sa[a].bytes() ) SPICE_HIP( hipMemcpy( c->_sa[a].data(), _sa[a].data(), _sa[a].bytes(), hipMemcpyDeviceToDevice ) );
			}
			c->apply_launch( _lc );
			return c;
		}
		catch( ... )
		{
			if( s ) (void)hipStreamDestroy( s );
			throw;
		}
	}

	// ---- used by hip_batch (batched stepping of many nets) ----
	dview const & view( int slot ) const { return _vs[slot & ( CLOCK_SLOTS - 1 )]; }
	int num_tiles() const { return _v.nact; } // the delivery's tiles (those with incoming synapses)
	int num_receive_parts() const { return _v.rparts; }
	int num_update_blocks() const { return _grid_update; }
	// flush workgroups of this net inside a batch: about two per expected work item wave
	// (lazy: rows reach their flush about every FLUSH_AGE steps, and once after each arrival
	// of their spike; eager: arriving rows and spiking targets), at most the net's own grid
	int num_batch_flush_blocks() const
	{
		std::int64_t const items = traced ? ( _g.n / 256 + 1 ) * ( _v.fparts + _v.pparts ) : ( _g.n / FLUSH_AGE + _g.n / 256 + 1 ) * _v.fparts;
		return int( std::clamp<std::int64_t>( ( items + 1 ) / 2, 1, _grid_flush ) );
	}
	bool fixed() const { return ( _p.flags & SPICE_FLAG_FIXED_POINT ) != 0; }
	std::int64_t host_step() const override { return _step; }
	void advance_host( std::int64_t n )
	{
		for( std::int64_t k = 0; k < n; k++ ) tick();
	}

private:
	net_params _p;
	csr_dev _g;
	hipStream_t _s = nullptr;
	step_graphs _graphs;

	dbuf<char> _na[MAXA], _acc[MAXA], _sa[MAXA];
	dbuf<std::int32_t> _ring;
	dbuf<std::uint32_t> _counts;
	dbuf<unsigned long long> _hring;
	dbuf<std::int32_t> _last_spk;
	dbuf<std::int32_t> _synced, _pend, _fl_id, _fl_synced, _fl_pend;
	dbuf<std::uint8_t> _hasp;
	dbuf<std::int32_t> _pl_lo, _pl_hi;
	dbuf<std::uint32_t> _fl_count;
	dbuf<float> _xtr, _ytr;                    // eager: [2][N] traces by step parity
	dbuf<std::uint8_t> _fl8;                   // eager: [2][N] step flags
	dbuf<std::int32_t> _po_id;                 // eager: [2][N] spiking targets
	dbuf<std::uint32_t> _po_count;             // eager: [4]
	dbuf<std::int64_t> _tp_ptr;                // eager: by-target index offsets
	dbuf<std::int64_t> _caught;                // plastic: step whose plasticity a synchronization did
	dbuf<tp_entry> _tp_ent;                    // eager: by-target index entries
	std::int64_t _caught_host = -1;
	std::int64_t _tp_total = 0, _tp_targets = 0;
	dbuf<ctl_t> _ctl;
	dbuf<unsigned long long> _bstats;
	std::int64_t _stats_step0 = 0;
	dview _v{};                // step views: _vs[p] reads clock slot p
	dview _vs[CLOCK_SLOTS]{};

	int _grid_update = 0, _grid_flush = 0, _grid_tiles = 0;
	int _grid_upd = 0; // workgroups of this net's own update kernel (_upd_block threads)
	int _upd_block = SPICE_UPDATE_BLOCK;
	int _grid_plast = 0; // plasticity workgroups of the delivery kernel
	bool _overlap = false; // steps run as k_step_overlap kernels
	bool _fused = false;   // steps run as k_step_fused kernels (decided by the launch shape)
	launch_cfg _lc{};
	dbuf<std::int32_t> _pivots;
	dbuf<std::int32_t> _tile_ids;
	std::int64_t _step = 0;
	util::kahan_sum<float> _clock;
	float _dt_last = 0.0f; // compensated dt of the last step (what its kernels used)

	// the host's copy of the clock advances with every enqueued step, as the device's does
	void tick()
	{
		_step++;
		_dt_last = _clock.add( _p.dt );
	}

	dbuf<slog_t> _slog_ctl;
	dbuf<std::int32_t> _slog_ids;
	dbuf<std::uint64_t> _slog_off;
	std::uint64_t _slog_cap = 0;
	std::int64_t _slog_max = 0, _slog_t0 = 0, _slog_n = 0;

	template <typename Decl>
	static std::size_t attr_bytes( int a )
	{
		std::size_t r = 0;
		util::for_n<Decl::size>( [&]( auto I ) {
			if( int( I ) == a ) r = sizeof( typename Decl::template type<I> );
		} );
		return r;
	}
	static void check_attr( int a, int n )
	{
		if( a < 0 || a >= n ) throw std::invalid_argument( "attribute index out of range" );
	}

	void allocate()
	{
		std::int64_t const n = _g.n;
		util::for_n<N::size>( [&]( auto I ) {
			using A = typename N::template type<I>;
			_na[I].resize( n * sizeof( A ) );
			if constexpr( ( ( receive_mask_v<N> >> I ) & 1u ) != 0 )
			{
				std::size_t const asz = ( std::is_floating_point_v<A> && true ) ? ( fixed() ? 8 : sizeof( A ) ) : sizeof( A );
				_acc[I].resize( 2 * n * asz ); // by step parity
				SPICE_HIP( hipMemsetAsync( _acc[I].data(), 0, _acc[I].bytes(), _s ) );
			}
		} );
		util::for_n<S::size>( [&]( auto I ) {
			using A = typename S::template type<I>;
			_sa[I].resize( std::max<std::int64_t>( _g.nnz, 1 ) * sizeof( A ) );
		} );
		_ring.resize( std::size_t( _p.delay + 2 ) * n );
		_counts.resize( _p.delay + 2 );
		SPICE_HIP( hipMemsetAsync( _counts.data(), 0, _counts.bytes(), _s ) );
		if constexpr( plastic )
		{
			_hring.resize( std::size_t( HWORDS ) * n );
			_hasp.resize( n );
			_pl_lo.resize( n );
			_pl_hi.resize( n );
			_fl_id.resize( 2 * n ); // lists of steps of both parities
			_fl_count.resize( CLOCK_SLOTS );
			_caught.resize( 1 );
			SPICE_HIP( hipMemsetAsync( _hring.data(), 0, _hring.bytes(), _s ) );
			SPICE_HIP( hipMemsetAsync( _fl_count.data(), 0, _fl_count.bytes(), _s ) );
			reset_caught();
			if constexpr( traced )
			{
				_xtr.resize( 2 * n );
				_ytr.resize( 2 * n );
				_fl8.resize( 2 * n );
				_po_id.resize( 2 * n );
				_po_count.resize( CLOCK_SLOTS );
				_tp_ptr.resize( n + 1 );
				reset_eager();
			}
			else
			{
				_last_spk.resize( n );
				_synced.resize( n );
				_pend.resize( n );
				_fl_synced.resize( 2 * n );
				_fl_pend.resize( 2 * n );
				SPICE_HIP( hipMemsetAsync( _last_spk.data(), 0xff, _last_spk.bytes(), _s ) ); // -1
				SPICE_HIP( hipMemsetAsync( _pend.data(), 0, _pend.bytes(), _s ) );
				std::vector<std::int32_t> synced( n, -1 );
				SPICE_HIP( hipMemcpyAsync( _synced.data(), synced.data(), _synced.bytes(), hipMemcpyHostToDevice, _s ) );
				SPICE_HIP( hipStreamSynchronize( _s ) ); // `synced` is a host temporary
			}
		}
		_ctl.resize( CLOCK_SLOTS );
		SPICE_HIP( hipMemsetAsync( _ctl.data(), 0, _ctl.bytes(), _s ) );

		dview & v = _v;
		v.N = std::int32_t( n );
		v.D = _p.delay;
		v.width = std::int32_t( std::min<std::int64_t>( _g.width, INT_MAX ) );
		v.shift = 40;
		v.overlap = _overlap ? 1 : 0;
		v.DT = _p.dt;
		v.flags = _p.flags;
		v.seed = _p.seed;
		for( int k = 0; k < 4; k++ ) v.param[k] = _p.param[k];
		v.row_ptr = _g.row_ptr.data();
		v.col = _g.col.data();
		for( int a = 0; a < MAXA; a++ )
		{
			v.na[a] = _na[a].data();
			v.acc[a] = _acc[a].data();
			v.sa[a] = _sa[a].data();
		}
		v.ring = _ring.data();
		v.counts = _counts.data();
		v.hring = _hring.data();
		v.last_spk = _last_spk.data();
		v.synced = _synced.data();
		v.pend = _pend.data();
		v.hasp = _hasp.data();
		v.pl_lo = _pl_lo.data();
		v.pl_hi = _pl_hi.data();
		v.fl_id = _fl_id.data();
		v.fl_synced = _fl_synced.data();
		v.fl_pend = _fl_pend.data();
		v.fl_count = _fl_count.data();
		v.xtr = _xtr.data();
		v.ytr = _ytr.data();
		v.fl8 = _fl8.data();
		v.po_id = _po_id.data();
		v.po_count = _po_count.data();
		v.tp_ptr = _tp_ptr.data();
		v.caught = _caught.data();
		v.ctl = _ctl.data();
		v.ctl_next = _ctl.data() + 1;

		if constexpr( traced ) build_by_target();
		_grid_update = int( ( n + 255 ) / 256 );
		_upd_block = n >= UPDATE_LARGE_N ? UPDATE_BLOCK_LARGE : SPICE_UPDATE_BLOCK;
		_grid_upd = int( ( n + _upd_block - 1 ) / _upd_block );
		apply_launch( launch_cfg{} );
	}

	// Eager plasticity state at a clock (re)start: zero traces and flags, no pending lists.
	void reset_eager()
	{
		for( auto * b : { &_xtr, &_ytr } ) SPICE_HIP( hipMemsetAsync( b->data(), 0, b->bytes(), _s ) );
		SPICE_HIP( hipMemsetAsync( _fl8.data(), 0, _fl8.bytes(), _s ) );
		SPICE_HIP( hipMemsetAsync( _po_count.data(), 0, _po_count.bytes(), _s ) );
	}

	// Nothing synchronized yet (plastic models).
	void reset_caught()
	{
		std::int64_t const none = -1;
		SPICE_HIP( hipMemcpyAsync( _caught.data(), &none, 8, hipMemcpyHostToDevice, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) ); // `none` is a host temporary
		_caught_host = -1;
	}

	// By-target index of the plastic synapses: count per target, prefix sums, fill.
	void build_by_target()
	{
		std::int64_t const n = _g.n;
		unsigned const rows_grid = unsigned( ( n + 3 ) / 4 ); // one wave per row
		dbuf<std::uint32_t> cnt( n );
		SPICE_HIP( hipMemsetAsync( cnt.data(), 0, cnt.bytes(), _s ) );
		SPICE_LAUNCH( k_tp_count<Model><<<rows_grid, 256, 0, _s>>>( _v, cnt.data() ) );
		SPICE_LAUNCH( k_tp_scan<<<1, 1024, 0, _s>>>( cnt.data(), _tp_ptr.data(), n ) );
		std::int64_t total = 0;
		SPICE_HIP( hipMemcpyAsync( &total, _tp_ptr.data() + n, 8, hipMemcpyDeviceToHost, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) );
		_tp_total = total;
		_tp_ent.resize( std::size_t( std::max<std::int64_t>( total, 1 ) ) );
		_v.tp_ent = _tp_ent.data();
		if( total > 0 )
		{
			dbuf<unsigned long long> cursor( n );
			SPICE_HIP( hipMemcpyAsync( cursor.data(), _tp_ptr.data(), 8 * n, hipMemcpyDeviceToDevice, _s ) );
			SPICE_LAUNCH( k_tp_fill<Model><<<rows_grid, 256, 0, _s>>>( _v, cursor.data(), _tp_ent.data() ) );
			SPICE_HIP( hipStreamSynchronize( _s ) ); // before the temporaries are freed
		}
		if( total > 1 && total <= INT_MAX )
		{
			// each target's list sorted by source: neighbouring entries of a post-pass wave then
			// gather neighbouring traces, flags and row starts (fewer cache lines), and the index
			// does not depend on the order of the fill's atomics. Radix sort on the source bits
			// of the entries read as 64-bit words (the sort counts items in int: an index of more
			// than 2^31 entries keeps the fill's order, which is equally correct).
			int const bits = 64 - __builtin_clzll( static_cast<unsigned long long>( std::max<std::int64_t>( n - 1, 1 ) ) );
			auto const * keys = reinterpret_cast<unsigned long long const *>( _tp_ent.data() );
			dbuf<unsigned long long> sorted{ std::size_t( total ) };
			std::size_t tb = 0;
			SPICE_HIP( hipcub::DeviceSegmentedRadixSort::SortKeys( nullptr, tb, keys, sorted.data(), int( total ), int( n ), _tp_ptr.data(),
			                                                       _tp_ptr.data() + 1, 32, 32 + bits, _s ) );
			dbuf<char> temp( std::max<std::size_t>( tb, 1 ) );
			SPICE_HIP( hipcub::DeviceSegmentedRadixSort::SortKeys( temp.data(), tb, keys, sorted.data(), int( total ), int( n ), _tp_ptr.data(),
			                                                       _tp_ptr.data() + 1, 32, 32 + bits, _s ) );
			SPICE_HIP( hipMemcpyAsync( _tp_ent.data(), sorted.data(), std::size_t( total ) * 8, hipMemcpyDeviceToDevice, _s ) );
			SPICE_HIP( hipStreamSynchronize( _s ) ); // before the temporaries are freed
		}
		// targets with incoming plastic synapses (the mean in-degree sets the post-pass split)
		std::vector<std::int64_t> ptr( n + 1 );
		SPICE_HIP( hipMemcpy( ptr.data(), _tp_ptr.data(), ptr.size() * 8, hipMemcpyDeviceToHost ) );
		_tp_targets = 0;
		for( std::int64_t i = 0; i < n; i++ ) _tp_targets += ptr[i + 1] > ptr[i];
	}

	// Launch shapes (zero fields: the defaults). tile_w: targets per delivery tile (the
	// receive workgroup's LDS accumulators; ~2 tiles per CU by default); flush_slots: target
	// slots per flush work item (one wave), so a flushed row of length L is split into about
	// L / flush_slots items; flush_grid: persistent flush workgroups (8 per CU by default).
	launch_cfg default_launch() const
	{
		int const cus = device_cu_count();
		std::int64_t const n = _g.n;
		launch_cfg c;
		c.tile_w = int( std::min<std::int64_t>( TILE_MAX, std::max<std::int64_t>( 256, ( ( n + 2 * cus - 1 ) / ( 2 * cus ) + 63 ) / 64 * 64 ) ) );
		c.flush_slots = 256;
		double const mean_row = n > 0 ? double( _g.nnz ) / double( n ) : 0.0;
		std::int64_t const fparts = std::max( 1, int( mean_row / c.flush_slots + 0.5 ) );
		c.flush_grid = int( std::min<std::int64_t>( 8 * cus, std::max<std::int64_t>( 1, ( n * fparts + 3 ) / 4 ) ) );
		// small nets: few events per tile, so more and smaller delivery workgroups
		c.recv_block = n <= 65536 ? 256 : 1024;
		c.recv_parts = 1;
		return c;
	}

	void apply_launch( launch_cfg c )
	{
		launch_cfg const d = default_launch();
		if( c.tile_w == 0 ) c.tile_w = d.tile_w;
		if( c.flush_slots == 0 ) c.flush_slots = d.flush_slots;
		if( c.flush_grid == 0 ) c.flush_grid = d.flush_grid;
		if( c.recv_block == 0 ) c.recv_block = d.recv_block;
		if( c.recv_parts == 0 ) c.recv_parts = d.recv_parts;
		if( c.recv_block != 256 && c.recv_block != 512 && c.recv_block != 1024 )
			throw std::invalid_argument( "recv_block must be 256, 512 or 1024" );
		if( c.recv_parts < 1 || c.recv_parts > 16 ) throw std::invalid_argument( "recv_parts must be in [1, 16]" );
		int const cus = device_cu_count();
		if( c.tile_w < 64 || c.tile_w > max_tile_w() || c.tile_w % 64 )
			throw std::invalid_argument( max_tile_w() > TILE_MAX ? "tile_w must be a multiple of 64 in [64, 8192]"
			                                                    : "tile_w must be a multiple of 64 in [64, 1024] (wider tiles: models without plasticity that run two kernels per step)" );
		if( c.flush_slots < 64 || c.flush_slots > ( 1 << 20 ) || c.flush_slots % 64 )
			throw std::invalid_argument( "flush_slots must be a multiple of 64 in [64, 2^20]" );
		if( c.flush_grid < 1 || c.flush_grid > 64 * cus ) throw std::invalid_argument( "flush_grid must be in [1, 64 x CUs]" );
		std::int64_t const n = _g.n;
		dview & v = _v;
		bool const retile = c.tile_w != v.tile_w;
		if( retile )
		{
			std::int32_t const ntiles = std::int32_t( ( n + c.tile_w - 1 ) / c.tile_w );
			_pivots.resize( std::size_t( n ) * ( ntiles + 1 ) );
			v.tile_w = c.tile_w;
			v.ntiles = ntiles;
			v.pivots = _pivots.data();
			SPICE_LAUNCH( k_pivots<<<unsigned( ( n + 3 ) / 4 ), 256, 0, _s>>>( _v ) );
			find_active_tiles(); // (synchronizes)
		}
		double const mean_row = n > 0 ? double( _g.nnz ) / double( n ) : 0.0;
		v.fparts = std::max( 1, int( mean_row / c.flush_slots + 0.5 ) );
		double const mean_in = _tp_targets > 0 ? double( _tp_total ) / double( _tp_targets ) : 0.0;
		v.pparts = std::max( 1, int( mean_in / c.flush_slots + 0.5 ) );
		v.rparts = c.recv_parts;
		_grid_flush = c.flush_grid;
		_grid_tiles = std::max( 1, std::min( v.nact * v.rparts, ( 16 * 1024 / c.recv_block ) * cus ) );
		_grid_plast = plastic ? std::max( 1, _grid_flush * 256 / c.recv_block ) : 0; // the same waves as flush_grid
		_lc = c;
		v.upd_slot0 = upd_slot0();
		// one kernel per step (k_step_fused): static models whose steps do not overlap, one
		// delivery work item per tile, every tile a delivery tile, tiles of at most FUSE_CHUNKS
		// workgroup widths
		_fused = !plastic && !_overlap && c.recv_parts == 1 && v.nact == v.ntiles && c.tile_w <= TILE_MAX &&
		         c.tile_w <= FUSE_CHUNKS * c.recv_block;
		v.fused = _fused ? 1 : 0;
		// statistics slots: one per workgroup of the widest step kernel (receive: per work item;
		// eager plasticity: after those, one per plasticity workgroup or synchronization block;
		// the overlapped step's update blocks, or the fused step's FUSE_CHUNKS per tile, after
		// those), and enough for the update kernel's blocks
		std::size_t const slots =
		    std::size_t( std::max( { _grid_update, _grid_upd, 2 * _grid_flush, 2 * _grid_tiles,
		                             upd_slot0() + _grid_update, upd_slot0() + FUSE_CHUNKS * v.ntiles } ) ) *
		    NSTAT; // (update blocks of >= 256 neurons)
		if( slots > _bstats.size() )
		{
			std::vector<unsigned long long> keep( NSTAT, 0ull );
			if( _bstats.size() )
			{
				std::vector<unsigned long long> b( _bstats.size() );
				SPICE_HIP( hipMemcpyAsync( b.data(), _bstats.data(), _bstats.bytes(), hipMemcpyDeviceToHost, _s ) );
				SPICE_HIP( hipStreamSynchronize( _s ) );
				for( std::size_t i = 0; i < b.size(); i++ ) keep[i % NSTAT] += b[i];
			}
			_bstats.resize( slots );
			SPICE_HIP( hipMemsetAsync( _bstats.data(), 0, _bstats.bytes(), _s ) );
			SPICE_HIP( hipMemcpyAsync( _bstats.data(), keep.data(), NSTAT * 8, hipMemcpyHostToDevice, _s ) );
			SPICE_HIP( hipStreamSynchronize( _s ) ); // `keep` is a host temporary
			v.bstats = _bstats.data();
		}
		for( int p = 0; p < CLOCK_SLOTS; p++ )
		{
			_vs[p] = _v;
			_vs[p].ctl = _ctl.data() + p;
			_vs[p].ctl_next = _ctl.data() + ( ( p + 1 ) & ( CLOCK_SLOTS - 1 ) );
			_vs[p].ctl_prev = _ctl.data() + ( ( p + CLOCK_SLOTS - 1 ) & ( CLOCK_SLOTS - 1 ) );
		}
		SPICE_HIP( hipStreamSynchronize( _s ) );
		_graphs.clear(); // the captured launches hold the old shapes
	}

	// The delivery's tiles: those that some row has a target in (a population that no
	// connection targets gets no delivery workgroups); at least one, so that every step has a
	// delivery work item. Runs right after the pivots are built, before the views are copied.
	void find_active_tiles()
	{
		dview & v = _v;
		dbuf<std::uint8_t> flag( std::size_t( v.ntiles ) );
		SPICE_HIP( hipMemsetAsync( flag.data(), 0, flag.bytes(), _s ) );
		SPICE_LAUNCH( k_tile_active<<<unsigned( ( _g.n + 3 ) / 4 ), 256, 0, _s>>>( v, flag.data() ) );
		std::vector<std::uint8_t> f( std::size_t( v.ntiles ) );
		SPICE_HIP( hipMemcpyAsync( f.data(), flag.data(), f.size(), hipMemcpyDeviceToHost, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) );
		std::vector<std::int32_t> ids;
		for( std::int32_t t = 0; t < v.ntiles; t++ )
			if( f[std::size_t( t )] ) ids.push_back( t );
		if( ids.empty() ) ids.push_back( 0 );
		_tile_ids.resize( ids.size() );
		SPICE_HIP( hipMemcpy( _tile_ids.data(), ids.data(), ids.size() * 4, hipMemcpyHostToDevice ) );
		v.tile_ids = _tile_ids.data();
		v.nact = std::int32_t( ids.size() );
	}

	void initialize()
	{
		SPICE_LAUNCH( k_init_neurons<Model><<<_grid_update, 256, 0, _s>>>( _v ) );
		if constexpr( plastic )
		{
			SPICE_LAUNCH( k_init_synapses<Model><<<unsigned( ( _g.n + 3 ) / 4 ), 256, 0, _s>>>( _v ) );
		}
		SPICE_HIP( hipStreamSynchronize( _s ) );
	}

	// First statistics slot of the overlapped step's update blocks (after the delivery work
	// items and the plasticity workgroups).
	int upd_slot0() const { return _v.ntiles * _v.rparts + std::max( _grid_plast, _grid_flush ); }

	// Enqueues steps t .. t + len - 1. Two kernels per step, or (overlapped steps) the delivery
	// of step t, then one k_step_overlap per step boundary (update of T, delivery of T + 1,
	// passes of T - 1), then the last update with the passes of the step before it: len + 1
	// kernels, and as with the two-kernel step the passes of the last step are left to the
	// next delivery kernel (or a synchronization).
	void launch_steps( std::int64_t t, std::int64_t len )
	{
		if( _fused )
		{
			for( std::int64_t k = 0; k < len; k++ )
			{
				if( fixed() )
					launch_fused<true>( _vs[( t + k ) & ( CLOCK_SLOTS - 1 )], nullptr, nullptr );
				else
					launch_fused<false>( _vs[( t + k ) & ( CLOCK_SLOTS - 1 )], nullptr, nullptr );
			}
			return;
		}
		if( !_overlap )
		{
			for( std::int64_t k = 0; k < len; k++ ) launch_step( t + k );
			return;
		}
		for( std::int64_t k = 0; k <= len; k++ )
		{
			if( fixed() )
				launch_overlap_kernel<true>( t, len, k, nullptr, nullptr );
			else
				launch_overlap_kernel<false>( t, len, k, nullptr, nullptr );
		}
	}

	// Kernel k (0 .. len) of the overlapped sequence of steps t .. t + len - 1 (see
	// launch_steps), with HIP events attached to its dispatch when start is not null.
	template <bool F>
	void launch_overlap_kernel( std::int64_t t, std::int64_t len, std::int64_t k, hipEvent_t start, hipEvent_t stop )
	{
		int const bs = _lc.recv_block;
		if( k == 0 )
		{
			// delivery of step t only: step t - 1's passes run beside the first update
			dview const & v = _vs[t & ( CLOCK_SLOTS - 1 )];
			std::int32_t const tiles = _grid_tiles;
			auto go = [&]( auto b ) {
				constexpr int BS = decltype( b )::value;
				if( start )
					launch_timed( reinterpret_cast<void const *>( &k_receive_tiled<Model, F, BS> ), unsigned( tiles ), BS, start, stop, v, tiles );
				else
					SPICE_LAUNCH( ( k_receive_tiled<Model, F, BS><<<unsigned( tiles ), BS, 0, _s>>>( v, tiles ) ) );
			};
			bs == 256 ? go( std::integral_constant<int, 256>{} ) : bs == 512 ? go( std::integral_constant<int, 512>{} ) : go( std::integral_constant<int, 1024>{} );
			return;
		}
		std::int64_t const T = t + k - 1; // this kernel updates step T
		dview const & v = _vs[T & ( CLOCK_SLOTS - 1 )];
		std::int32_t const nrecv = k < len ? _grid_tiles : 0; // the last kernel delivers nothing
		std::int32_t const nplast = _grid_plast, s0 = upd_slot0();
		unsigned const grid = unsigned( nrecv + nplast + int( ( _g.n + bs - 1 ) / bs ) );
		auto go = [&]( auto b ) {
			constexpr int BS = decltype( b )::value;
			if( start )
				launch_timed( reinterpret_cast<void const *>( &k_step_overlap<Model, F, BS> ), grid, BS, start, stop, v, nrecv, nplast, s0 );
			else
				SPICE_LAUNCH( ( k_step_overlap<Model, F, BS><<<grid, BS, 0, _s>>>( v, nrecv, nplast, s0 ) ) );
		};
		bs == 256 ? go( std::integral_constant<int, 256>{} ) : bs == 512 ? go( std::integral_constant<int, 512>{} ) : go( std::integral_constant<int, 1024>{} );
	}

	// run_profiled for overlapped steps: kernel 0 -> kernel_ms[0], kernels 1 .. n - 1 ->
	// kernel_ms[2], kernel n -> kernel_ms[1].
	template <bool F>
	void run_profiled_overlap( std::int64_t n, double * kernel_ms )
	{
		if( n == 0 ) return;
		constexpr int batch = 128;
		synchronize();
		std::int64_t const t0 = _step;
		std::vector<hipEvent_t> ev( 2 * batch, nullptr );
		try
		{
			for( auto & e : ev ) SPICE_HIP( hipEventCreate( &e ) );
			for( std::int64_t done = 0; done <= n; )
			{
				int const k = int( std::min<std::int64_t>( batch, n + 1 - done ) );
				for( int i = 0; i < k; i++ ) launch_overlap_kernel<F>( t0, n, done + i, ev[2 * i], ev[2 * i + 1] );
				SPICE_HIP( hipStreamSynchronize( _s ) );
				for( int i = 0; i < k; i++ )
				{
					float ms = 0;
					SPICE_HIP( hipEventElapsedTime( &ms, ev[2 * i], ev[2 * i + 1] ) );
					std::int64_t const idx = done + i;
					kernel_ms[idx == 0 ? 0 : idx == n ? 1 : 2] += ms;
				}
				done += k;
			}
		}
		catch( ... )
		{
			for( auto & e : ev )
				if( e ) (void)hipEventDestroy( e );
			throw;
		}
		for( auto & e : ev ) (void)hipEventDestroy( e );
		for( std::int64_t i = 0; i < n; i++ ) tick();
	}

	// Enqueues the kernels of step `t` (t mod 4 selects the clock slot): delivery (plastic
	// models: with step t - 1's plasticity work in the same kernel), then the neuron update.
	void launch_step( std::int64_t t )
	{
		dview const & v = _vs[t & ( CLOCK_SLOTS - 1 )];
		if( fixed() )
		{
			launch_receive<true>( v );
			launch_update<true>( v, nullptr, nullptr );
		}
		else
		{
			launch_receive<false>( v );
			launch_update<false>( v, nullptr, nullptr );
		}
	}

	// the neuron update kernel of the two-kernel step (workgroups of _upd_block threads), with
	// HIP events attached to its dispatch when start is not null
	template <bool F>
	void launch_update( dview const & v, hipEvent_t start, hipEvent_t stop )
	{
		if( _upd_block == UPDATE_BLOCK_LARGE )
		{
			if( start )
				launch_timed( reinterpret_cast<void const *>( &k_update<Model, F, UPDATE_BLOCK_LARGE> ), unsigned( _grid_upd ), UPDATE_BLOCK_LARGE, start, stop, v );
			else
				SPICE_LAUNCH( ( k_update<Model, F, UPDATE_BLOCK_LARGE><<<_grid_upd, UPDATE_BLOCK_LARGE, 0, _s>>>( v ) ) );
		}
		else
		{
			if( start )
				launch_timed( reinterpret_cast<void const *>( &k_update<Model, F> ), unsigned( _grid_upd ), SPICE_UPDATE_BLOCK, start, stop, v );
			else
				SPICE_LAUNCH( ( k_update<Model, F><<<_grid_upd, SPICE_UPDATE_BLOCK, 0, _s>>>( v ) ) );
		}
	}

	// the fused step kernel (static models), with HIP events attached to its dispatch when
	// start is not null
	template <bool F>
	void launch_fused( dview const & v, hipEvent_t start, hipEvent_t stop )
	{
		if constexpr( !plastic )
		{
			std::int32_t const tiles = _grid_tiles, s0 = upd_slot0();
			auto go = [&]( auto b ) {
				constexpr int BS = decltype( b )::value;
				if( start )
					launch_timed( reinterpret_cast<void const *>( &k_step_fused<Model, F, BS> ), unsigned( tiles ), BS, start, stop, v, tiles, s0 );
				else
					SPICE_LAUNCH( ( k_step_fused<Model, F, BS><<<unsigned( tiles ), BS, 0, _s>>>( v, tiles, s0 ) ) );
			};
			int const bs = _lc.recv_block;
			bs == 256 ? go( std::integral_constant<int, 256>{} ) : bs == 512 ? go( std::integral_constant<int, 512>{} ) : go( std::integral_constant<int, 1024>{} );
		}
	}

	// run_profiled for fused steps: every step's kernel -> kernel_ms[2]
	template <bool F>
	void run_profiled_fused( std::int64_t n, double * kernel_ms )
	{
		constexpr int batch = 128;
		synchronize(); // profiled steps run serially on the main stream
		std::vector<hipEvent_t> ev( 2 * batch, nullptr );
		try
		{
			for( auto & e : ev ) SPICE_HIP( hipEventCreate( &e ) );
			for( std::int64_t done = 0; done < n; )
			{
				int const k = int( std::min<std::int64_t>( batch, n - done ) );
				for( int i = 0; i < k; i++ )
				{
					launch_fused<F>( _vs[_step & ( CLOCK_SLOTS - 1 )], ev[2 * i], ev[2 * i + 1] );
					tick();
				}
				SPICE_HIP( hipStreamSynchronize( _s ) );
				for( int i = 0; i < k; i++ )
				{
					float ms = 0;
					SPICE_HIP( hipEventElapsedTime( &ms, ev[2 * i], ev[2 * i + 1] ) );
					kernel_ms[2] += ms;
				}
				done += k;
			}
		}
		catch( ... )
		{
			for( auto & e : ev )
				if( e ) (void)hipEventDestroy( e );
			throw;
		}
		for( auto & e : ev ) (void)hipEventDestroy( e );
	}

	// delivery workgroups, plus (plastic models) the previous step's plasticity workgroups
	template <bool F>
	void launch_receive( dview const & v )
	{
		launch_receive_events<F>( v, nullptr, nullptr );
	}

	// the delivery kernel at the launch shape's workgroup size and tile capacity (wide tiles:
	// static models only), with HIP events attached to its dispatch when start is not null
	template <bool F>
	void launch_receive_events( dview const & v, hipEvent_t start, hipEvent_t stop )
	{
		unsigned const grid = unsigned( _grid_tiles + _grid_plast );
		std::int32_t const tiles = _grid_tiles;
		auto go = [&]( auto bs, auto tm ) {
			constexpr int BS = decltype( bs )::value, TM = decltype( tm )::value;
			if( start )
				launch_timed( reinterpret_cast<void const *>( &k_receive_tiled<Model, F, BS, TM> ), grid, BS, start, stop, v, tiles );
			else
				SPICE_LAUNCH( ( k_receive_tiled<Model, F, BS, TM><<<grid, BS, 0, _s>>>( v, tiles ) ) );
		};
		auto by_block = [&]( auto tm ) {
			if( _lc.recv_block == 256 )
				go( std::integral_constant<int, 256>{}, tm );
			else if( _lc.recv_block == 512 )
				go( std::integral_constant<int, 512>{}, tm );
			else
				go( std::integral_constant<int, 1024>{}, tm );
		};
		if constexpr( !plastic )
			if( _lc.tile_w > TILE_MAX )
			{
				by_block( std::integral_constant<int, TILE_WIDE>{} );
				return;
			}
		by_block( std::integral_constant<int, TILE_MAX>{} );
	}

	// Launches `fn` on the net's stream with `start`/`stop` attached to the dispatch itself, so
	// that their elapsed time is the kernel's own duration (no launch gap, unlike events
	// recorded between launches). args: the kernel's exact parameter types.
	template <typename... A>
	void launch_timed( void const * fn, unsigned grid, unsigned block, hipEvent_t start, hipEvent_t stop, A... a )
	{
		void * args[] = { static_cast<void *>( &a )... };
		SPICE_HIP( hipExtLaunchKernel( fn, dim3( grid ), dim3( block ), args, 0, _s, start, stop, 0 ) );
	}

	// e[0], e[1] time the delivery kernel, e[2], e[3] the neuron update
	template <bool F>
	void launch_step_events( hipEvent_t * e, dview const & v )
	{
		launch_receive_events<F>( v, e[0], e[1] );
		launch_update<F>( v, e[2], e[3] );
	}

	// Brings all plastic synapses up to date through the last completed step.
	void sync_synapses()
	{
		synchronize();
		if constexpr( plastic )
		{
			if( _step > 0 && _caught_host != _step - 1 )
			{
				// the last step's plasticity work, which the next delivery kernel would run, then
				// (eager) its events of the rows arriving next or (lazy) every row's replay up to it;
				// recorded in `caught` so the next delivery kernel skips the former
				std::int64_t const now = _step - 1;
				dview const & v = _vs[_step & ( CLOCK_SLOTS - 1 )]; // a view of step now + 1
				std::int64_t const slot0 = std::int64_t( v.ntiles ) * v.rparts;
				if constexpr( traced )
				{
					if( fixed() )
						SPICE_LAUNCH( k_stdp<Model, true><<<_grid_flush, 256, 0, _s>>>( v, now, slot0 ) );
					else
						SPICE_LAUNCH( k_stdp<Model, false><<<_grid_flush, 256, 0, _s>>>( v, now, slot0 ) );
					// deferred events of the rows arriving next (none when the passes apply all)
					if constexpr( !nodefer ) SPICE_LAUNCH( k_sync_eager<Model><<<unsigned( ( _g.n + 3 ) / 4 ), 256, 0, _s>>>( _v, now ) );
				}
				else
				{
					if( fixed() )
						SPICE_LAUNCH( k_flush<Model, true><<<_grid_flush, 256, 0, _s>>>( v, now, slot0 ) );
					else
						SPICE_LAUNCH( k_flush<Model, false><<<_grid_flush, 256, 0, _s>>>( v, now, slot0 ) );
					// dt of the last completed step (the step that moved the clock to its state)
					SPICE_LAUNCH( k_sync_all<Model><<<unsigned( ( _g.n + 3 ) / 4 ), 256, 0, _s>>>( _v, std::int32_t( now ), _last_dt() ) );
				}
				SPICE_HIP( hipMemcpyAsync( _caught.data(), &now, 8, hipMemcpyHostToDevice, _s ) );
				SPICE_HIP( hipStreamSynchronize( _s ) ); // `now` is a host temporary
				_caught_host = now;
			}
		}
		synchronize();
	}

	// Compensated dt of step _step - 1 (the last enqueued step), as its kernels computed it
	// from the device clock; the nominal dt before any step (no replay needs it then).
	float _last_dt() const { return _step > 0 && _dt_last > 0.0f ? _dt_last : _p.dt; }
};
} // namespace spice::hipx
