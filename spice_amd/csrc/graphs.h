#pragma once
// Captured HIP graphs of a step sequence, and their replay with a bounded number of graph
// launches in flight.
//
// The step kernels read every step-dependent value from device memory, but the clock slot
// they read is fixed per launch (slot t mod 4 for step t, see dview.h), so a graph of L steps
// replays correctly from any step whose clock phase equals the phase it was captured at.
// Kept: one 1-step graph per phase (any step) and graphs of 4, 8, 16 and 32 steps captured
// at phase 0. A run of n steps therefore replays graphs only: 1-step graphs up to the next
// multiple of 4, the largest block graphs that fit, 1-step graphs for the rest.
//
// Replay keeps at most INFLIGHT graph launches queued on the stream: before launch i the host
// waits for launch i - INFLIGHT (an event recorded after each launch). A long run (100,000
// steps are 3,125 launches of up to 64 kernels) thus never queues more than 64 x 64 dispatches
// ahead of the GPU, whatever the caller asks for, while a chunk of up to ~2,000 steps is
// enqueued without the host waiting (so callers that drive several streams, like the server's
// per-model batches, get all their chunks queued at once and running side by side).
#include "common.h"
#include "dview.h"

#include <cstdint>

namespace spice::hipx
{
class step_graphs
{
public:
	static constexpr int BLOCKS = 4;   // block graphs of 4, 8, 16 and 32 steps
	static_assert( BLOCKS == CLOCK_SLOTS, "clear() walks both graph sets with one bound" );
	static constexpr int INFLIGHT = 64; // graph launches queued ahead of the GPU, at most

	step_graphs() = default;
	step_graphs( step_graphs const & ) = delete;
	step_graphs & operator=( step_graphs const & ) = delete;
	~step_graphs() { clear(); }

	bool empty() const { return _blk[BLOCKS - 1] == nullptr; }

	// Destroys every graph and event. The caller has synchronized the stream.
	void clear()
	{
		for( hipGraphExec_t * set : { _one, _blk } )
			for( int k = 0; k < CLOCK_SLOTS; k++ )
			{
				if( set[k] ) (void)hipGraphExecDestroy( set[k] );
				set[k] = nullptr;
			}
		for( auto & e : _ev )
		{
			if( e ) (void)hipEventDestroy( e );
			e = nullptr;
		}
		_launched = 0;
	}

	// launch(t, len) enqueues on `s` the kernels of len steps from a step of clock phase t & 3.
	template <typename Launch>
	void capture( hipStream_t s, Launch && launch )
	{
		try
		{
			for( int p = 0; p < CLOCK_SLOTS; p++ ) _one[p] = capture_one( s, [&] { launch( p, 1 ); } );
			for( int k = 0; k < BLOCKS; k++ ) _blk[k] = capture_one( s, [&] { launch( 0, 4 << k ); } );
			for( auto & e : _ev ) SPICE_HIP( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
		}
		catch( ... )
		{
			clear();
			throw;
		}
	}

	// Enqueues n steps starting at step t on `s`; tick() is called once per step enqueued.
	template <typename Tick>
	void run( hipStream_t s, std::int64_t t, std::int64_t n, Tick && tick )
	{
		while( n > 0 )
		{
			hipGraphExec_t g = _one[t & ( CLOCK_SLOTS - 1 )];
			int len = 1;
			if( ( t & ( CLOCK_SLOTS - 1 ) ) == 0 && n >= 4 )
			{
				int k = BLOCKS - 1;
				while( ( 4 << k ) > n ) k--;
				g = _blk[k];
				len = 4 << k;
			}
			launch( s, g );
			for( int i = 0; i < len; i++ ) tick();
			t += len;
			n -= len;
		}
	}

private:
	hipGraphExec_t _one[CLOCK_SLOTS] = {};
	hipGraphExec_t _blk[BLOCKS] = {};
	hipEvent_t _ev[INFLIGHT] = {};
	std::uint64_t _launched = 0;

	void launch( hipStream_t s, hipGraphExec_t g )
	{
		hipEvent_t & e = _ev[_launched % INFLIGHT];
		if( _launched >= INFLIGHT ) SPICE_HIP( hipEventSynchronize( e ) ); // launch _launched - INFLIGHT is done
		SPICE_HIP( hipGraphLaunch( g, s ) );
		SPICE_HIP( hipEventRecord( e, s ) );
		_launched++;
	}

	template <typename F>
	static hipGraphExec_t capture_one( hipStream_t s, F && enqueue )
	{
		hipGraph_t g = nullptr;
		SPICE_HIP( hipStreamBeginCapture( s, hipStreamCaptureModeThreadLocal ) );
		try
		{
			enqueue();
		}
		catch( ... )
		{
			(void)hipStreamEndCapture( s, &g );
			if( g ) (void)hipGraphDestroy( g );
			throw;
		}
		SPICE_HIP( hipStreamEndCapture( s, &g ) );
		hipGraphExec_t x = nullptr;
		hipError_t const e = hipGraphInstantiate( &x, g, nullptr, nullptr, 0 );
		(void)hipGraphDestroy( g );
		check( e, "hipGraphInstantiate" );
		return x;
	}
};
} // namespace spice::hipx
