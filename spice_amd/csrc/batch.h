#pragma once
// Batched stepping of many independent networks of one model (parameter sweeps, served
// jobs). Every net keeps its own buffers, clock and statistics (it is an ordinary hip_net);
// the batch adds device tables that map the workgroups of ONE launch per step phase to
// (net, local tile), (net, local neuron block) and, for plastic models, (net, local flush
// block), and replays captured HIP graphs of those launches (graphs.h). Per-net results are
// identical to stepping the nets one by one: the kernels run the same per-net bodies and all
// randomness is keyed on (net seed, neuron, step). Nets may sit at different steps (served
// jobs join a running mix): each net's clock slot is chosen from its own step parity.
#include "net.h"

#include <memory>
#include <vector>

namespace spice::hipx
{
struct batch_iface
{
	virtual ~batch_iface() = default;
	virtual void run( std::int64_t n ) = 0;
	virtual void synchronize() = 0;
	virtual int size() const = 0;
	virtual net_iface * net( int i ) = 0;
	// opens the spike-log window of every net (with a log) over the next nsteps steps
	virtual void begin_spike_log( std::int64_t nsteps ) = 0;
};

template <typename Model>
class hip_batch final : public batch_iface
{
	static constexpr bool plastic = is_plastic_v<Model>;
	// delivery workgroup size: most tiles of a sweep belong to small nets and carry few events,
	// so more, smaller workgroups shorten the launch's critical path
	static constexpr int recv_block = SPICE_BATCH_RECV_BLOCK;

public:
	// `nets` must outlive the batch unless they are also handed over in `owned`. The batch
	// destroys `stream` at the end if `own_stream`.
	hip_batch( std::vector<hip_net<Model> *> nets, std::vector<std::unique_ptr<hip_net<Model>>> && owned, hipStream_t stream, bool own_stream )
	    : _nets( std::move( nets ) )
	    , _owned( std::move( owned ) )
	    , _s( stream )
	    , _own_stream( own_stream )
	{
		if( _nets.empty() ) throw std::invalid_argument( "empty batch" );
		_fixed = _nets[0]->fixed();
		for( auto * n : _nets )
			if( n->fixed() != _fixed ) throw std::invalid_argument( "all nets of a batch use the same accumulation mode" );
		// Many nets already fill the GPU, so then a batched net does not need many small
		// delivery tiles: each net gets about max(64, 64 x CUs / nets) tiles (fewer workgroups,
		// each with a fixed staging cost). On a sweep of two 256-net batches 64 tiles per net
		// beat 16, 32, 128 and 256; a batch of a few nets keeps its nets' own tiles. Fixed-point
		// results do not depend on the tile width (fp32 sums are reordered, as by any launch
		// shape); a net keeps wider tiles it already has, and keeps the new width afterwards.
		std::int64_t const tiles_per_net = std::max<std::int64_t>( 64, std::int64_t( SPICE_BATCH_ITEMS_PER_CU ) * device_cu_count() / std::int64_t( _nets.size() ) );
		for( auto * n : _nets )
		{
			launch_cfg c = n->get_launch();
			int const tw = int( std::clamp<std::int64_t>( ( n->num_neurons() / tiles_per_net + 63 ) / 64 * 64, 64, TILE_MAX ) );
			if( tw > c.tile_w || c.tile_w > TILE_MAX ) // (the batched kernels hold tiles of up to TILE_MAX)
			{
				c.tile_w = c.tile_w > TILE_MAX ? TILE_MAX : tw;
				n->set_launch( c );
			}
		}
		std::vector<std::int32_t> tn, tl, bn, bl, fn, fl, fc;
		std::vector<dview> vs[CLOCK_SLOTS];
		for( std::size_t k = 0; k < _nets.size(); k++ )
		{
			hip_net<Model> & n = *_nets[k];
			for( int t = 0; t < n.num_tiles() * n.num_receive_parts(); t++ ) // (tile, part) work items
			{
				tn.push_back( std::int32_t( k ) );
				tl.push_back( t );
			}
			for( int b = 0; b < n.num_update_blocks(); b++ )
			{
				bn.push_back( std::int32_t( k ) );
				bl.push_back( b );
			}
			if constexpr( plastic )
			{
				int const nf = n.num_batch_flush_blocks();
				for( int b = 0; b < nf; b++ )
				{
					fn.push_back( std::int32_t( k ) );
					fl.push_back( b );
					fc.push_back( nf );
				}
			}
			// batch step t runs net step host_step + t: clock slot (host_step + t) mod 4
			// (the batch runs two kernels per step, or overlapped ones, never a net's fused step:
			// its delivery writes the spike log's step offsets, which the update then reads)
			for( int p = 0; p < CLOCK_SLOTS; p++ )
			{
				dview w = n.view( int( ( n.host_step() + p ) & ( CLOCK_SLOTS - 1 ) ) );
				w.fused = 0;
				vs[p].push_back( w );
			}
		}
		_ntiles = std::int32_t( tn.size() );
		_nblocks = std::int32_t( bn.size() );
		_nflush = std::int32_t( fn.size() );
		// (the overlapped kernel's update workgroups have the delivery's size: 256 threads, the
		// batched update blocks' size)
		_overlap = recv_block == 256;
		for( auto * n : _nets ) _overlap = _overlap && n->overlapped();
		upload( _tile_net, tn );
		upload( _tile_local, tl );
		upload( _block_net, bn );
		upload( _block_local, bl );
		upload( _flush_net, fn );
		upload( _flush_local, fl );
		upload( _flush_count, fc );
		for( int p = 0; p < CLOCK_SLOTS; p++ ) upload( _views[p], vs[p] );
		_grid_tiles = std::min( _ntiles, ( recv_block == 256 ? SPICE_BATCH_GRID_PER_CU : SPICE_BATCH_GRID_PER_CU / 4 ) * device_cu_count() );
		for( auto * n : _nets ) n->synchronize();
		SPICE_HIP( hipStreamSynchronize( _s ) );
	}

	~hip_batch() override
	{
		if( _s ) (void)hipStreamSynchronize( _s );
		_graphs.clear();
		if( _s && _own_stream ) (void)hipStreamDestroy( _s );
	}

	void run( std::int64_t n ) override
	{
		if( n < 0 ) throw std::invalid_argument( "negative step count" );
		if( n == 0 ) return;
		if( _graphs.empty() ) _graphs.capture( _s, [&]( int t, int len ) { launch_steps( t, len ); } );
		_graphs.run( _s, _step, n, [&] { _step++; } );
		for( auto * net : _nets ) net->advance_host( n );
	}

	void begin_spike_log( std::int64_t nsteps ) override
	{
		for( auto * net : _nets ) net->note_spike_log( nsteps ); // validates nsteps first
		std::int32_t const n = std::int32_t( _nets.size() );
		SPICE_LAUNCH( k_batch_log_window<<<unsigned( ( n + 255 ) / 256 ), 256, 0, _s>>>( _views[_step & ( CLOCK_SLOTS - 1 )].data(), n, nsteps ) );
	}

	void synchronize() override
	{
		SPICE_HIP( hipStreamSynchronize( _s ) );
	}
	int size() const override { return int( _nets.size() ); }
	net_iface * net( int i ) override
	{
		if( i < 0 || i >= size() ) throw std::invalid_argument( "net index out of range" );
		return _nets[i];
	}

private:
	std::vector<hip_net<Model> *> _nets;
	std::vector<std::unique_ptr<hip_net<Model>>> _owned;
	hipStream_t _s = nullptr;
	bool _own_stream = false;
	step_graphs _graphs;
	bool _fixed = false;
	bool _overlap = false; // every net runs overlapped steps: so does the batch
	std::int64_t _step = 0;
	std::int32_t _ntiles = 0, _nblocks = 0, _nflush = 0;
	int _grid_tiles = 0;
	dbuf<std::int32_t> _tile_net, _tile_local, _block_net, _block_local, _flush_net, _flush_local, _flush_count;
	dbuf<dview> _views[CLOCK_SLOTS];

	template <typename T>
	void upload( dbuf<T> & d, std::vector<T> const & h )
	{
		d.resize( std::max<std::size_t>( h.size(), 1 ) );
		if( !h.empty() ) SPICE_HIP( hipMemcpyAsync( d.data(), h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) );
	}

	// Same schedules as hip_net::launch_steps: two kernels per step (delivery with step t - 1's
	// plasticity, update), or the overlapped sequence (delivery of step t, one overlapped
	// kernel per step boundary, the last update).
	void launch_steps( std::int64_t t, std::int64_t len )
	{
		for( std::int64_t k = 0; k < len; k++ )
		{
			if( _overlap ) break;
			if( _fixed )
				launch_step_t<true>( t + k );
			else
				launch_step_t<false>( t + k );
		}
		if( !_overlap ) return;
		for( std::int64_t k = 0; k <= len; k++ )
		{
			if( _fixed )
				launch_overlap_t<true>( t, len, k );
			else
				launch_overlap_t<false>( t, len, k );
		}
	}

	template <bool F>
	void launch_overlap_t( std::int64_t t, std::int64_t len, std::int64_t k )
	{
		if( k == 0 )
		{
			// delivery of step t only: step t - 1's plasticity runs beside the first update
			dview const * v = _views[t & ( CLOCK_SLOTS - 1 )].data();
			SPICE_LAUNCH( k_batch_receive<Model, F, recv_block><<<unsigned( _grid_tiles ), recv_block, 0, _s>>>(
			    v, _tile_net.data(), _tile_local.data(), _ntiles, _grid_tiles, _flush_net.data(), _flush_local.data(), _flush_count.data() ) );
			return;
		}
		dview const * v = _views[( t + k - 1 ) & ( CLOCK_SLOTS - 1 )].data(); // updates step t + k - 1
		std::int32_t const grecv = k < len ? _grid_tiles : 0, nplast = plastic ? _nflush : 0;
		unsigned const grid = unsigned( grecv + nplast + _nblocks );
		SPICE_LAUNCH( k_batch_step_overlap<Model, F, recv_block><<<grid, recv_block, 0, _s>>>(
		    v, _tile_net.data(), _tile_local.data(), _ntiles, grecv, _flush_net.data(), _flush_local.data(), _flush_count.data(), nplast,
		    _block_net.data(), _block_local.data() ) );
	}

	template <bool F>
	void launch_step_t( std::int64_t t )
	{
		dview const * v = _views[t & ( CLOCK_SLOTS - 1 )].data();
		unsigned const grid = unsigned( _grid_tiles + ( plastic ? _nflush : 0 ) );
		SPICE_LAUNCH( k_batch_receive<Model, F, recv_block><<<grid, recv_block, 0, _s>>>( v, _tile_net.data(), _tile_local.data(), _ntiles, _grid_tiles,
		                                                                                    _flush_net.data(), _flush_local.data(), _flush_count.data() ) );
		SPICE_LAUNCH( k_batch_update<Model, F><<<_nblocks, 256, 0, _s>>>( v, _block_net.data(), _block_local.data() ) );
	}

};

// Builds the nets of a batch (each gets its own stream for setup) and the batch, which owns
// them and `s`.
template <typename Model>
batch_iface * make_batch( std::vector<net_params> const & ps, std::vector<csr_dev> && gs, hipStream_t s )
{
	std::vector<std::unique_ptr<hip_net<Model>>> nets;
	std::vector<hip_net<Model> *> raw;
	for( std::size_t k = 0; k < ps.size(); k++ )
	{
		hipStream_t ns = nullptr;
		SPICE_HIP( hipStreamCreateWithFlags( &ns, hipStreamNonBlocking ) );
		try
		{
			nets.emplace_back( new hip_net<Model>( ps[k], std::move( gs[k] ), ns ) );
		}
		catch( ... )
		{
			(void)hipStreamDestroy( ns );
			throw;
		}
		raw.push_back( nets.back().get() );
	}
	return new hip_batch<Model>( std::move( raw ), std::move( nets ), s, true );
}

// A batch over nets owned elsewhere (all of model `Model`); `s` stays the caller's.
template <typename Model>
batch_iface * make_group( std::vector<net_iface *> const & nets, hipStream_t s )
{
	std::vector<hip_net<Model> *> raw;
	for( auto * n : nets )
	{
		auto * h = dynamic_cast<hip_net<Model> *>( n );
		if( !h ) throw std::invalid_argument( "batch nets must all be of the batch's model" );
		raw.push_back( h );
	}
	return new hip_batch<Model>( std::move( raw ), {}, s, false );
}
} // namespace spice::hipx
