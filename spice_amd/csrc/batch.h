#pragma once
// Batched stepping of many independent networks of one model (parameter sweeps, served
// jobs). Every net keeps its own buffers, clock and statistics (it is an ordinary hip_net);
// the batch adds device tables that map the workgroups of ONE launch per step phase to
// (net, local tile) and (net, local neuron block), and replays a captured HIP graph of those
// launches. Per-net results are identical to stepping the nets one by one: the kernels run the
// same per-net bodies and all randomness is keyed on (net seed, neuron, step).
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
};

template <typename Model>
class hip_batch final : public batch_iface
{
	static constexpr int graph_steps = 32;

public:
	hip_batch( std::vector<std::unique_ptr<hip_net<Model>>> && nets, hipStream_t stream )
	    : _nets( std::move( nets ) )
	    , _s( stream )
	{
		static_assert( !is_plastic_v<Model>, "batches hold static (non-plastic) models" );
		if( _nets.empty() ) throw std::invalid_argument( "empty batch" );
		_fixed = ( _nets[0]->params().flags & SPICE_FLAG_FIXED_POINT ) != 0;
		for( auto & n : _nets )
		{
			if( n->host_step() != 0 ) throw std::invalid_argument( "batch nets must start at step 0" );
			if( ( ( n->params().flags & SPICE_FLAG_FIXED_POINT ) != 0 ) != _fixed )
				throw std::invalid_argument( "all nets of a batch use the same accumulation mode" );
		}
		std::vector<std::int32_t> tn, tl, bn, bl;
		std::vector<dview> v0, v1;
		for( std::size_t k = 0; k < _nets.size(); k++ )
		{
			for( int t = 0; t < _nets[k]->num_tiles(); t++ )
			{
				tn.push_back( std::int32_t( k ) );
				tl.push_back( t );
			}
			for( int b = 0; b < _nets[k]->num_update_blocks(); b++ )
			{
				bn.push_back( std::int32_t( k ) );
				bl.push_back( b );
			}
			v0.push_back( _nets[k]->view( 0 ) );
			v1.push_back( _nets[k]->view( 1 ) );
		}
		_ntiles = std::int32_t( tn.size() );
		_nblocks = std::int32_t( bn.size() );
		upload( _tile_net, tn );
		upload( _tile_local, tl );
		upload( _block_net, bn );
		upload( _block_local, bl );
		upload( _views[0], v0 );
		upload( _views[1], v1 );
		_grid_tiles = std::min( _ntiles, 16 * device_cu_count() );
		for( auto & n : _nets ) n->synchronize();
		SPICE_HIP( hipStreamSynchronize( _s ) );
	}

	~hip_batch() override
	{
		if( _exec ) (void)hipGraphExecDestroy( _exec );
		if( _s ) (void)hipStreamSynchronize( _s );
		if( _s ) (void)hipStreamDestroy( _s );
	}

	void run( std::int64_t n ) override
	{
		if( n < 0 ) throw std::invalid_argument( "negative step count" );
		std::int64_t done = 0;
		auto eager = [&] {
			launch_step( _step++ );
			done++;
		};
		if( ( _step & 1 ) && done < n ) eager();
		if( n - done >= graph_steps )
		{
			if( !_exec ) capture();
			for( ; done + graph_steps <= n; done += graph_steps )
			{
				SPICE_HIP( hipGraphLaunch( _exec, _s ) );
				_step += graph_steps;
			}
		}
		while( done < n ) eager();
		for( auto & net : _nets ) net->advance_host( n );
	}

	void synchronize() override { SPICE_HIP( hipStreamSynchronize( _s ) ); }
	int size() const override { return int( _nets.size() ); }
	net_iface * net( int i ) override
	{
		if( i < 0 || i >= size() ) throw std::invalid_argument( "net index out of range" );
		return _nets[i].get();
	}

private:
	std::vector<std::unique_ptr<hip_net<Model>>> _nets;
	hipStream_t _s = nullptr;
	hipGraphExec_t _exec = nullptr;
	bool _fixed = false;
	std::int64_t _step = 0;
	std::int32_t _ntiles = 0, _nblocks = 0;
	int _grid_tiles = 0;
	dbuf<std::int32_t> _tile_net, _tile_local, _block_net, _block_local;
	dbuf<dview> _views[2];

	template <typename T>
	void upload( dbuf<T> & d, std::vector<T> const & h )
	{
		d.resize( std::max<std::size_t>( h.size(), 1 ) );
		if( !h.empty() ) SPICE_HIP( hipMemcpyAsync( d.data(), h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice, _s ) );
		SPICE_HIP( hipStreamSynchronize( _s ) );
	}

	void launch_step( std::int64_t t )
	{
		if( _fixed )
			launch_step_t<true>( t );
		else
			launch_step_t<false>( t );
	}

	template <bool F>
	void launch_step_t( std::int64_t t )
	{
		dview const * v = _views[t & 1].data();
		SPICE_LAUNCH( k_batch_receive<Model, F><<<_grid_tiles, TBS, 0, _s>>>( v, _tile_net.data(), _tile_local.data(), _ntiles ) );
		SPICE_LAUNCH( k_batch_update<Model, F><<<_nblocks, 256, 0, _s>>>( v, _block_net.data(), _block_local.data() ) );
	}

	void capture()
	{
		hipGraph_t g = nullptr;
		SPICE_HIP( hipStreamBeginCapture( _s, hipStreamCaptureModeThreadLocal ) );
		try
		{
			for( int i = 0; i < graph_steps; i++ ) launch_step( i );
		}
		catch( ... )
		{
			(void)hipStreamEndCapture( _s, &g );
			if( g ) (void)hipGraphDestroy( g );
			throw;
		}
		SPICE_HIP( hipStreamEndCapture( _s, &g ) );
		hipError_t const e = hipGraphInstantiate( &_exec, g, nullptr, nullptr, 0 );
		(void)hipGraphDestroy( g );
		check( e, "hipGraphInstantiate" );
	}
};

// Builds the nets of a batch (each gets its own stream for setup) and the batch itself.
template <typename Model>
batch_iface * make_batch( std::vector<net_params> const & ps, std::vector<csr_dev> && gs, hipStream_t s )
{
	std::vector<std::unique_ptr<hip_net<Model>>> nets;
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
	}
	return new hip_batch<Model>( std::move( nets ), s );
}
} // namespace spice::hipx
