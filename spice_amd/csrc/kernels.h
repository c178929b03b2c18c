#pragma once
// Model-templated step kernels for gfx950 (subsystems b, c, d).
//
// Per step T (step index read from device memory, see dview.h), two kernels on one stream, or
// (overlapped steps) one kernel per step boundary that holds the work of both:
//   k_receive_tiled<M> : delivers the spikes emitted at T - D. One workgroup per (target tile,
//                  part of the step's spike list): the spikes' row segments inside the tile
//                  are walked as one flat index space and M::neuron::receive accumulates
//                  into the tile's LDS accumulators (fp32, or int64 fixed point), which then
//                  go to the global ones. Plastic models bring each delivered plastic synapse
//                  up to date first (lazy replay, or the previous step's deferred events for
//                  neuron-trace synapses), and extra workgroups of the same kernel do step
//                  T - 1's plasticity work (eager pre/post passes, or the lazy flush), which
//                  touches no row delivered at T.
//   k_update<M>  : one thread per neuron: add accumulated inputs, run M::neuron::update,
//                  store, and append spikes to the ring with one atomic per workgroup
//                  (ballot + popcount + mbcnt). Plastic models also record the spike
//                  history and queue the rows (and, eagerly, spiking targets) for step T's
//                  plasticity work. Block 0 writes step T + 1's clock and empties its counters,
//                  so the same launches are valid at every step and are captured once into a
//                  HIP graph.
//   k_step_overlap<M>: for models whose delivery may run before its targets' update (receive()
//                  split into a gate the update tests, or reading nothing of the target; for
//                  plastic models, eager with refractory endpoints) and a delay of at least 2:
//                  the update of step T, the delivery of step T + 1 and the plasticity passes
//                  of step T - 1 in one kernel, one kernel per step boundary.
//
// Reference counterparts: _process_neurons (spice/cuda/algorithm.cu:181-230),
// _process_spikes HNDL_SPKS / UPDT_SYNS (algorithm.cu:238-313), _gather (algorithm.cu:340-384),
// the per-step sequence of cuda::snn::step (spice/cuda/snn.d.cpp:233-289).
#include "common.h"
#include "dview.h"

#include <spice/models/model.h>
#include <spice/util/meta.h>
#include <spice/util/philox.h>

#include <cstdint>
#include <tuple>
#include <type_traits>
#include <utility>

namespace spice::hipx
{
constexpr int HISTORY = 256;            // spike-history window (steps) of plastic models
constexpr int HWORDS = HISTORY / 64;     // 64-bit words per neuron in the history ring
// Lazy plasticity (per-synapse contract): rows are brought up to date at least every
// FLUSH_AGE steps, before their window leaves the history. An event-free stretch is merged
// into one update(n) whose decay pow(d, n) differs from n per-step products by up to ~n/2
// ulp, which also keeps the stretches this short.
constexpr int FLUSH_AGE = HISTORY - 1;

template <typename T>
__device__ inline T ld( void const * base, std::int64_t i )
{
	return static_cast<T const *>( base )[i];
}
template <typename T>
__device__ inline void st( void * base, std::int64_t i, T x )
{
	static_cast<T *>( base )[i] = x;
}

// Streaming accesses (data touched once per step: synapse rows and scattered weights) with
// the non-temporal hint, so they do not evict the small per-neuron arrays that the same
// kernels gather from repeatedly. SPICE_NT bits: 1 post pass, 2 pre pass, 4 delivery.
#ifndef SPICE_NT
#define SPICE_NT 0
#endif
template <bool NT, typename T>
__device__ inline T lds( T const * p )
{
	if constexpr( NT )
		return __builtin_nontemporal_load( p );
	else
		return *p;
}
template <bool NT, typename T>
__device__ inline void sts( T * p, T x )
{
	if constexpr( NT )
		__builtin_nontemporal_store( x, p );
	else
		*p = x;
}

// ---- iterators handed to model functors ------------------------------------------------

// State held in registers (neuron update, synapse replay).
template <typename Tuple>
struct reg_iter
{
	Tuple & s;
	std::uint32_t i;
	template <int I>
	__device__ auto & get() const
	{
		return std::get<I>( s );
	}
	__device__ uint_ id() const { return i; }
};

// Receive target: reads see the neuron state (end of the previous step); atomic_add through
// the backend goes to the accumulator of attribute I.
template <typename T, int I>
struct in_proxy
{
	dview const * v;
	std::int32_t j;
	__device__ operator T() const { return ld<T>( v->na[I], j ); }
};

template <typename Neuron>
struct recv_iter
{
	dview const * v;
	std::int32_t j;
	template <int I>
	__device__ in_proxy<typename Neuron::template type<I>, I> get() const
	{
		return { v, j };
	}
	__device__ uint_ id() const { return uint_( j ); }
};

struct no_syn_iter
{
};

// ---- device backend (reference counterpart: spice/cuda/backend.cuh:10-55) -----------------
template <bool FIXED>
struct dev_backend
{
	util::philox_stream rng;
	int shift;

	__device__ float rand() { return rng.rand(); }

	template <typename T, int I>
	__device__ void atomic_add( in_proxy<T, I> p, typename util::identity<T>::type x ) const
	{
		if constexpr( std::is_floating_point_v<T> )
		{
			if constexpr( FIXED )
			{
				double const scaled = double( x ) * double( 1ull << shift );
				atomicAdd( static_cast<unsigned long long *>( p.v->acc[I] ) + p.j, (unsigned long long)( long long )__builtin_rint( scaled ) );
			}
			else
				unsafeAtomicAdd( static_cast<float *>( p.v->acc[I] ) + p.j, float( x ) );
		}
		else
			atomicAdd( static_cast<T *>( p.v->acc[I] ) + p.j, x );
	}
	template <typename T>
	__device__ static void atomic_add( T & var, typename util::identity<T>::type x )
	{
		atomicAdd( &var, x );
	}

	template <typename T>
	__device__ static T min( T x, T hi )
	{
		return x < hi ? x : hi;
	}
	template <typename T>
	__device__ static T max( T lo, T x )
	{
		return x < lo ? lo : x;
	}
	template <typename T>
	__device__ static T clamp( T x, T lo, T hi )
	{
		return min( hi, max( lo, x ) );
	}
	template <typename T>
	__device__ static T exp( T x )
	{
		return expf( x );
	}
	__device__ static float pow( float x, float y ) { return powf( x, y ); }
	__device__ static float fma( float a, float b, float c ) { return fmaf( a, b, c ); }
};

template <bool FIXED>
__device__ inline dev_backend<FIXED> make_backend( dview const & v, std::uint32_t a, std::uint32_t b, std::uint32_t c, std::uint32_t tag )
{
	return { util::philox_stream( v.seed, a, b, c, tag ), v.shift };
}

__device__ inline snn_info info_of( dview const & v )
{
	snn_info i;
	i.num_neurons = v.N;
	for( int k = 0; k < 4; k++ ) i.param[k] = v.param[k];
	return i;
}
__device__ inline float step_dt( dview const & v ) { return v.DT - v.ctl->c; }

// Spike ring entry of step t (t >= -D): D + 2 entries, so that the update of step T clears the
// entry of step T + 1 (last read by the delivery of step T - 1) while step T's delivery may
// still read the entry of step T - D.
__device__ inline int ring_slot( dview const & v, std::int64_t t ) { return int( ( t + v.D + 2 ) % ( v.D + 2 ) ); }
// The same for step ctl->step + k, -D <= k <= 1, from the clock's entry (no division).
__device__ inline int ring_at( dview const & v, int k )
{
	int const r = v.ctl->rslot + k, R = v.D + 2;
	return r < 0 ? r + R : r >= R ? r - R : r;
}

// Adds per-lane counts to statistics slot `slot` (a slot is written by one workgroup at a
// time): wave reduction, then the waves' totals through LDS, then one plain read-add-write
// by thread 0. Every thread of the block must call it.
template <int K>
__device__ inline void block_add_stats( dview const & v, std::int64_t slot, unsigned long long const ( &x )[K], int const ( &idx )[K] )
{
	__shared__ unsigned long long part[16][K]; // blockDim.x <= 1024: at most 16 waves
	int const w = threadIdx.x >> 6;
#pragma unroll
	for( int k = 0; k < K; k++ )
	{
		unsigned long long const t = wave_sum( x[k] );
		if( lane_id() == 0 ) part[w][k] = t;
	}
	__syncthreads();
	if( threadIdx.x == 0 )
	{
		int const nw = int( blockDim.x + 63 ) >> 6;
#pragma unroll
		for( int k = 0; k < K; k++ )
		{
			unsigned long long t = 0;
			for( int i = 0; i < nw; i++ ) t += part[i][k];
			if( t ) v.bstats[slot * NSTAT + idx[k]] += t;
		}
	}
}

// ---- state load/store -------------------------------------------------------------------
template <typename Decl, bool NT = false>
__device__ inline void load_tuple( typename Decl::tuple_t & s, void * const * arr, std::int64_t i )
{
	util::for_n<Decl::size>( [&]( auto I ) {
		using T = typename Decl::template type<I>;
		std::get<I>( s ) = lds<NT>( static_cast<T const *>( arr[I] ) + i );
	} );
}
template <typename Decl, bool NT = false>
__device__ inline void store_tuple( typename Decl::tuple_t const & s, void * const * arr, std::int64_t i )
{
	util::for_n<Decl::size>( [&]( auto I ) {
		using T = typename Decl::template type<I>;
		sts<NT>( static_cast<T *>( arr[I] ) + i, std::get<I>( s ) );
	} );
}

// Stores the attributes of `after` that differ (bit for bit) from `before`, the values loaded
// for index i. An event that leaves a synapse unchanged -- a potentiation of a weight clamped at
// its maximum, a depression of one clamped at zero -- then writes nothing, and a memory line
// none of whose synapses changed is never dirtied: for the scattered post-pass weights that
// saves the line's write-back. Equal results: an unchanged value is stored by not storing it.
// Returns whether anything was stored.
template <typename Decl, bool NT = false>
__device__ inline bool store_changed( typename Decl::tuple_t const & before, typename Decl::tuple_t const & after, void * const * arr,
                                      std::int64_t i )
{
	bool any = false;
	util::for_n<Decl::size>( [&]( auto I ) {
		using T = typename Decl::template type<I>;
		using B = std::conditional_t<sizeof( T ) == 8, std::uint64_t, std::uint32_t>;
		static_assert( sizeof( T ) == sizeof( B ), "synapse attributes are 4 or 8 bytes" );
		if( __builtin_bit_cast( B, std::get<I>( after ) ) != __builtin_bit_cast( B, std::get<I>( before ) ) )
		{
			sts<NT>( static_cast<T *>( arr[I] ) + i, std::get<I>( after ) );
			any = true;
		}
	} );
	return any;
}

// ---- spike history and lazy plasticity --------------------------------------------------
// Spike history (plastic models): bit (t mod HISTORY) of ring word ((t mod HISTORY) / 64) of
// neuron j is set iff j spiked at step t, valid for the last HISTORY steps. Lazy plasticity
// also keeps last_spk[j], j's latest spike.

__device__ inline bool spiked_at( dview const & v, std::int32_t j, std::int32_t e )
{
	std::int64_t const w = ( e & ( HISTORY - 1 ) ) >> 6;
	return ( v.hring[w * v.N + j] >> ( e & 63 ) ) & 1ull;
}

// History words of neuron j (independent loads, issued together).
struct hwords
{
	unsigned long long w[HWORDS];
};
__device__ inline hwords load_hwords( dview const & v, std::int32_t j )
{
	hwords h;
#pragma unroll
	for( int w = 0; w < HWORDS; w++ ) h.w[w] = v.hring[std::int64_t( w ) * v.N + j];
	return h;
}

// f(e) for every step e in [first, last] at which the neuron with history words h spiked,
// in increasing order (last - first < HISTORY).
template <typename F>
__device__ inline void for_each_spike( hwords const & h, std::int32_t first, std::int32_t last, F && f )
{
	for( std::int32_t e = first; e <= last; )
	{
		int const idx = e & ( HISTORY - 1 );
		int const b0 = idx & 63;
		int const cnt = ( 64 - b0 ) < ( last - e + 1 ) ? ( 64 - b0 ) : ( last - e + 1 );
		unsigned long long word = 0;
#pragma unroll
		for( int w = 0; w < HWORDS; w++ )
			if( w == ( idx >> 6 ) ) word = h.w[w];
		unsigned long long bits = word >> b0;
		if( cnt < 64 ) bits &= ( 1ull << cnt ) - 1ull;
		while( bits )
		{
			int const k = __builtin_ctzll( bits );
			f( e + k );
			bits &= bits - 1ull;
		}
		e += cnt;
	}
}

template <typename F>
__device__ inline void for_each_spike( dview const & v, std::int32_t j, std::int32_t first, std::int32_t last, F && f )
{
	if( first > last ) return;
	for_each_spike( load_hwords( v, j ), first, last, f );
}

// Applies, in time order, the plasticity of every step in (synced, now] to one synapse
// src -> j: the pending pre event at synced+1 (if pend) and j's post spikes. Event-free
// stretches are merged into the next call's nsteps (the contract's n-step form) and a
// trailing event-free stretch is applied with pre = post = false. Equivalent to the CPU
// checker's dense per-step updates (reference counterpart: algorithm.cu:283-294).
template <typename Syn, typename Iter, typename Bak>
__device__ inline void replay(
    Iter syn, dview const & v, std::int32_t j, std::int32_t now, std::int32_t synced, std::int32_t pend, float dt, snn_info info, Bak & bak )
{
	if( now <= synced ) return;
	std::int32_t t = synced, first = synced + 1;
	std::int32_t const ls = v.last_spk[j];
	if( pend )
	{
		bool const post = ls >= first && spiked_at( v, j, first );
		Syn::update( syn, 1, true, post, dt, info, bak );
		t = first++;
	}
	for_each_spike( v, j, first, ls < now ? ls : now, [&]( std::int32_t e ) {
		Syn::update( syn, e - t, false, true, dt, info, bak );
		t = e;
	} );
	if( t < now ) Syn::update( syn, now - t, false, false, dt, info, bak );
}

// ---- eager plasticity (neuron-trace synapses) ----------------------------------------------
// Trace synapses change only at events, and an event needs only the two neurons' traces of
// that step, so every event is applied in (or right after) the step it happens, and the
// synapse state is current at all times. Step T's event of synapse s -> d (pre: s's spike
// arrives at T, post: d spiked at T; one on_event call with both flags) is applied by exactly
// one of:
//   * the pre pass of step T: rows whose spike arrives at T and not at T + 1;
//   * the post pass of step T: for d spiked at T, its incoming plastic synapses (by-target
//     index) whose source's spike arrives neither at T nor at T + 1;
//   * the delivery of step T + 1 (deferred): rows whose spike arrives at T + 1, before the
//     delivery reads them.
// The two passes of step T run inside the delivery kernel of step T + 1, beside its delivery:
// they touch no row that delivery reads, and they read only step T's traces and flags (arrays
// by step parity), which the update of step T + 1 writes only after that kernel.

//
// A synapse type promising refractory endpoints (refractory_steps >= 2, see model.h) needs no
// deferred events: the pre pass of step T takes every row whose spike arrives at T, the post
// pass every other source, and the delivery only reads weights. A weight it reads may still
// get an event of step T - 1 or T applied beside it, but only one whose target does not take
// this delivery's input.

// flag bits of fl8
constexpr std::uint8_t F8_SPIKED = 1, F8_ARRIVED = 2, F8_ARRIVES_NEXT = 4;

// Eager plasticity with every event applied by the passes of its step (no deferred events).
template <typename Model>
inline constexpr bool eager_nodefer_v =
    is_plastic_v<Model> && has_neuron_traces_v<typename Model::synapse> && refractory_steps_v<typename Model::synapse> >= 2;

// The update of step T may run beside the delivery of step T + 1 (given a delay of at least 2
// steps): delivery reads nothing that the update writes -- no target state (receive() split
// into a gate the update tests, or declared to read nothing) and, for plastic models, no weight
// whose pending events matter.
template <typename Model>
inline constexpr bool overlap_capable_v =
    input_ahead_of_update_v<typename Model::neuron> && ( !is_plastic_v<Model> || eager_nodefer_v<Model> );

// ---- kernels -------------------------------------------------------------------------------

template <typename Model>
__global__ __launch_bounds__( 256 ) void k_init_neurons( dview v )
{
	using N = typename Model::neuron;
	std::int64_t const i = std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
	if( i >= v.N ) return;
	typename N::tuple_t s{};
	auto bak = make_backend<false>( v, std::uint32_t( i ), 0u, 0u, util::TAG_NEURON_INIT );
	N::init( reg_iter<typename N::tuple_t>{ s, std::uint32_t( i ) }, info_of( v ), bak );
	store_tuple<N>( s, v.na, i );
}

// One wave per source row: synapse init + "row has a plastic synapse" flag.
template <typename Model>
__global__ __launch_bounds__( 256 ) void k_init_synapses( dview v )
{
	using S = typename Model::synapse;
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N ) return;
	snn_info const info = info_of( v );
	std::int64_t const lo = v.row_ptr[r], hi = v.row_ptr[r + 1];
	std::int64_t first = hi, last = lo - 1;
	for( std::int64_t e = lo + lane_id(); e < hi; e += 64 )
	{
		std::int32_t const j = v.col[e];
		typename S::tuple_t s{};
		auto bak = make_backend<false>(
		    v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), 0u, util::TAG_SYNAPSE_INIT );
		S::init( reg_iter<typename S::tuple_t>{ s, std::uint32_t( e ) }, int_( r ), j, info, bak );
		store_tuple<S>( s, v.sa, e );
		if( S::plastic( int_( r ), j, info ) )
		{
			first = e < first ? e : first;
			last = e;
		}
	}
	first = wave_min( first );
	last = wave_max( last );
	if( lane_id() == 0 )
	{
		v.hasp[r] = last >= first;
		v.pl_lo[r] = last >= first ? std::int32_t( first - lo ) : 0;
		v.pl_hi[r] = last >= first ? std::int32_t( last + 1 - lo ) : 0;
	}
}

// ---- target-tiled spike delivery -----------------------------------------------------------
#ifndef SPICE_UPDATE_BLOCK
#define SPICE_UPDATE_BLOCK 256 // threads per neuron-update workgroup of a single net (64..256)
#endif
// update_block's per-wave compaction counters hold at most 16 waves
static_assert( SPICE_UPDATE_BLOCK % 64 == 0 && SPICE_UPDATE_BLOCK >= 64 && SPICE_UPDATE_BLOCK <= 1024,
               "SPICE_UPDATE_BLOCK must be a multiple of 64 threads, at most 1024" );
// Large networks update in workgroups of this many threads: each workgroup reserves its
// spikes' ring slots with one atomic on the step's counter, and at N = 800,640 (the reference's
// synth network) 3,128 workgroups of 256 serialized on it (29 us per update; 12.5 us with 1024).
constexpr int UPDATE_BLOCK_LARGE = 1024;
constexpr std::int64_t UPDATE_LARGE_N = 131072; // from this many neurons
#ifndef SPICE_BATCH_ITEMS_PER_CU
#define SPICE_BATCH_ITEMS_PER_CU 64
#endif
#ifndef SPICE_BATCH_GRID_PER_CU
#define SPICE_BATCH_GRID_PER_CU 64 // batched delivery workgroups (of 256 threads) per CU, at most
#endif
#ifndef SPICE_BATCH_RECV_BLOCK
#define SPICE_BATCH_RECV_BLOCK 256
#endif
constexpr int TILE_MAX = 1024; // targets per tile (LDS accumulators)
// Wide tiles (static models, two-kernel step only: a second instantiation of the delivery kernel,
// so the other kernels keep the smaller LDS footprint; up to 157 KB of LDS for two 8-byte
// accumulators and 1,024-thread workgroups). A sparse network's rows have few targets in each
// 1,024-neuron tile (the reference's synth network: ~1.6), and every delivery workgroup stages
// all of the step's spikes: wider tiles stage each spike in fewer workgroups.
constexpr int TILE_WIDE = 8192;
// Delivery workgroups have 256 or 1024 threads (a launch shape): they stage that many spikes
// per pass in LDS.

// Per row r and tile t: offset (within the row) of the first target >= t * tile_w, for
// t = 0 .. ntiles (the last entry is the row length). One wave per row, binary searches.
static inline __global__ __launch_bounds__( 256 ) void k_pivots( dview v )
{
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N ) return;
	std::int64_t const lo = v.row_ptr[r];
	std::int32_t const len = std::int32_t( v.row_ptr[r + 1] - lo );
	std::int32_t const * c = v.col + lo;
	for( int t = lane_id(); t <= v.ntiles; t += 64 )
	{
		std::int64_t const x = std::int64_t( t ) * v.tile_w;
		std::int32_t a = 0, b = len; // first index with c[index] >= x
		while( a < b )
		{
			std::int32_t const m = ( a + b ) >> 1;
			if( c[m] < x )
				a = m + 1;
			else
				b = m;
		}
		v.pivots[r * ( v.ntiles + 1 ) + t] = a;
	}
}

// flag[t] = 1 for every tile t that some row has a target in. One wave per row.
static inline __global__ __launch_bounds__( 256 ) void k_tile_active( dview v, std::uint8_t * flag )
{
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N ) return;
	std::int32_t const * pv = v.pivots + r * ( v.ntiles + 1 );
	for( int t = lane_id(); t < v.ntiles; t += 64 )
		if( pv[t + 1] > pv[t] ) flag[t] = 1;
}

// Attributes of the receiving neurons that the delivering workgroup copies into LDS: those
// that receive() declares it reads (receive_reads), of at most 4 bytes.
template <typename Neuron, std::size_t... I>
constexpr std::uint32_t staged_mask_of( std::index_sequence<I...> )
{
	return ( 0u | ... | ( ( ( receive_reads_v<Neuron> >> I ) & 1u ) != 0 && sizeof( typename Neuron::template type<int( I )> ) <= 4 ? 1u << I : 0u ) );
}
template <typename Neuron>
constexpr std::uint32_t staged_mask()
{
	return staged_mask_of<Neuron>( std::make_index_sequence<std::size_t( Neuron::size )>{} );
}

// Receive target inside a tile: reads see the neuron state (from the tile's LDS copy when the
// attribute is staged), backend.atomic_add goes to the tile's LDS accumulator of attribute I.
template <typename T, int I, typename A, bool InMask, bool Staged = false>
struct tile_proxy
{
	dview const * v;
	std::int32_t j;
	A * lds; // accumulator of attribute I for this tile, indexed by j - tile start
	std::int32_t t0;
	std::uint32_t sval; // Staged: attribute I's value (bits) from the tile's LDS copy
	__device__ operator T() const
	{
		if constexpr( Staged )
		{
			using B = std::conditional_t<sizeof( T ) == 1, std::uint8_t, std::conditional_t<sizeof( T ) == 2, std::uint16_t, std::uint32_t>>;
			return __builtin_bit_cast( T, B( sval ) );
		}
		else
			return ld<T>( v->na[I], j );
	}
};

// Accumulator type of a received attribute of type T: fp32, or int64 fixed point, for floating
// point attributes; T itself otherwise.
template <typename T, bool FIXED>
using acc_type = std::conditional_t<std::is_floating_point_v<T>, std::conditional_t<FIXED, long long, float>, T>;

// LDS word of the tile accumulators: 4 bytes unless an accumulator needs 8.
template <typename Neuron, bool FIXED, std::size_t... I>
constexpr bool acc_needs_8( std::index_sequence<I...> )
{
	return ( false || ... || ( ( ( receive_mask_v<Neuron> >> I ) & 1u ) != 0 && sizeof( acc_type<typename Neuron::template type<int( I )>, FIXED> ) > 4 ) );
}
template <typename Neuron, bool FIXED>
using acc_word = std::conditional_t<acc_needs_8<Neuron, FIXED>( std::make_index_sequence<std::size_t( Neuron::size )>{} ), unsigned long long, std::uint32_t>;

template <typename Neuron, bool FIXED, int TM = TILE_MAX>
struct tile_iter
{
	dview const * v;
	std::int32_t j;
	std::int32_t t0;
	acc_word<Neuron, FIXED> * lds; // [nacc][TM] accumulator slots
	std::uint32_t const * stage;   // [nstaged][TM] 4-byte slots
	template <int I>
	__device__ auto get() const
	{
		using T = typename Neuron::template type<I>;
		constexpr std::uint32_t mask = receive_mask_v<Neuron>;
		constexpr std::uint32_t smask = staged_mask<Neuron>();
		constexpr int a = __builtin_popcount( mask & ( ( 1u << I ) - 1u ) );
		constexpr int b = __builtin_popcount( smask & ( ( 1u << I ) - 1u ) );
		using A = acc_type<T, FIXED>;
		constexpr bool staged = ( ( smask >> I ) & 1u ) != 0;
		std::uint32_t sv = 0;
		if constexpr( staged ) sv = stage[b * TM + ( j - t0 )];
		return tile_proxy<T, I, A, ( ( mask >> I ) & 1u ) != 0, staged>{ v, j, reinterpret_cast<A *>( lds + a * TM ), t0, sv };
	}
	__device__ uint_ id() const { return uint_( j ); }
};

// The same target, with the staged attribute values already read (sval[k]: the k-th staged
// attribute): the delivery of static models reads them for all entries of a pass at once.
template <typename Neuron, bool FIXED, int TM = TILE_MAX>
struct tile_iter_vals
{
	static constexpr int NS = __builtin_popcount( staged_mask<Neuron>() ) > 0 ? __builtin_popcount( staged_mask<Neuron>() ) : 1;
	dview const * v;
	std::int32_t j;
	std::int32_t t0;
	acc_word<Neuron, FIXED> * lds;
	std::uint32_t sval[NS];
	template <int I>
	__device__ auto get() const
	{
		using T = typename Neuron::template type<I>;
		constexpr std::uint32_t mask = receive_mask_v<Neuron>;
		constexpr std::uint32_t smask = staged_mask<Neuron>();
		constexpr int a = __builtin_popcount( mask & ( ( 1u << I ) - 1u ) );
		constexpr int b = __builtin_popcount( smask & ( ( 1u << I ) - 1u ) );
		using A = acc_type<T, FIXED>;
		constexpr bool staged = ( ( smask >> I ) & 1u ) != 0;
		return tile_proxy<T, I, A, ( ( mask >> I ) & 1u ) != 0, staged>{ v, j, reinterpret_cast<A *>( lds + a * TM ), t0,
		                                                                  staged ? sval[b] : 0u };
	}
	__device__ uint_ id() const { return uint_( j ); }
};

template <bool FIXED>
struct tile_backend : dev_backend<FIXED>
{
	using dev_backend<FIXED>::atomic_add;
	template <typename T, int I, typename A, bool InMask, bool Staged>
	__device__ void atomic_add( tile_proxy<T, I, A, InMask, Staged> p, typename util::identity<T>::type x ) const
	{
		static_assert( InMask, "receive() adds to a neuron attribute that the model's receive_mask does not list" );
		A * dst = p.lds + ( p.j - p.t0 );
		if constexpr( std::is_floating_point_v<T> && FIXED )
			atomicAdd( reinterpret_cast<unsigned long long *>( dst ),
			           (unsigned long long)( long long )__builtin_rint( double( x ) * double( 1ull << this->shift ) ) );
		else
			atomicAdd( dst, A( x ) );
	}
};

// Exclusive scan of one 64-bit value per thread over the block (a multiple of 64 threads):
// returns this thread's exclusive prefix and writes the block total to *total (LDS). Contains
// barriers.
__device__ inline std::uint64_t block_exclusive_scan( std::uint64_t x, std::uint64_t * wsum, std::uint64_t * total )
{
	std::uint64_t const incl = wave_inclusive_scan( x );
	int const tid = threadIdx.x;
	if( lane_id() == 63 ) wsum[tid >> 6] = incl;
	__syncthreads();
	std::uint64_t wb = 0;
	for( int w = 0; w < ( tid >> 6 ); w++ ) wb += wsum[w];
	if( tid == int( blockDim.x ) - 1 ) *total = wb + incl;
	__syncthreads();
	return wb + incl - x;
}

// For each u: the largest k[u] in [0, cnt) with pre[k[u]] <= f[u] (pre non-decreasing,
// pre[0] = 0 <= f[u], cnt >= 1). Binary lifting with step sizes from the largest power of two
// below cnt down to 1, branch-free: the U searches advance together, so each step's U LDS
// reads are in flight at once (log2(cnt) dependent reads in all, not U times that).
template <int U>
__device__ inline void segments_of( std::int32_t const * pre, std::int32_t cnt, std::int32_t const ( &f )[U], std::int32_t ( &k )[U] )
{
#pragma unroll
	for( int u = 0; u < U; u++ ) k[u] = 0;
	for( std::int32_t st = cnt > 1 ? std::int32_t( 1u << ( 31 - __builtin_clz( std::uint32_t( cnt - 1 ) ) ) ) : 0; st > 0; st >>= 1 )
	{
		std::int32_t p[U];
#pragma unroll
		for( int u = 0; u < U; u++ )
		{
			std::int32_t const m = k[u] + st;
			p[u] = pre[m < cnt ? m : cnt - 1];
		}
#pragma unroll
		for( int u = 0; u < U; u++ )
		{
			std::int32_t const m = k[u] + st;
			k[u] = ( m < cnt && p[u] <= f[u] ) ? m : k[u];
		}
	}
}

// One delivered event: receive(), or for a neuron with a split receive() its ungated half
// deliver() (the neuron update then tests the gate, accepts(), on the state receive() would
// have read; a target that does not accept drops the whole step's input).
template <typename N, typename... A>
__device__ inline void deliver_event( A &&... a )
{
	if constexpr( gated_receive_v<N> )
		N::deliver( static_cast<A &&>( a )... );
	else
		N::receive( static_cast<A &&>( a )... );
}

// Delivers the spikes of step T - D, one workgroup per tile of tile_w target neurons:
// the workgroup stages a chunk of spikes (id, segment of the row inside the tile, and for
// plastic models the row's lazy state) in LDS, forms the segments' prefix sums, and walks
// the concatenated segments as one flat index space (2 or 4 entries per thread in flight; an
// entry's segment found from a bitmap of segment starts, or for very large chunks by binary
// search). receive() accumulates into
// LDS with LDS atomics; at the end the tile's sums go to the global accumulators with plain
// stores: one workgroup owns each tile, so no global atomics are needed. AHEAD: the kernel's
// clock is step T - 1's (the overlapped step kernel delivers the next step beside the update).
// KEEP (the fused step, one work item per tile): the sums stay in LDS for the same workgroup's
// update of the tile; the returned pointer addresses them ([attribute][TM], as stored below).
template <typename Model, bool FIXED, int BS, bool AHEAD = false, int TM = TILE_MAX, bool KEEP = false>
__device__ inline void const * receive_tile( dview const & v, std::int32_t work )
{
	// work item = (active tile, part): the tile's spike list is split into v.rparts parts,
	// each delivered by its own workgroup (more workgroups in flight for few, large tiles);
	// tiles without incoming synapses get no work items
	std::int32_t const act = work / v.rparts;
	std::int32_t const part = work - act * v.rparts;
	std::int32_t const tile = v.tile_ids[act];
	using N = typename Model::neuron;
	using S = typename Model::synapse;
	constexpr bool plastic = is_plastic_v<Model>;
	constexpr std::uint32_t rmask = receive_mask_v<N>;
	constexpr int nacc = __builtin_popcount( rmask ) > 0 ? __builtin_popcount( rmask ) : 1;

	// a split receive() reads nothing of its target: no attribute is staged
	constexpr std::uint32_t smask = gated_receive_v<N> ? 0u : staged_mask<N>();
	constexpr int nstaged = __builtin_popcount( smask );
	using word = acc_word<N, FIXED>;
	__shared__ word acc[nacc * TM];
	__shared__ std::uint32_t stage[nstaged > 0 ? nstaged * TM : 1];
	// a chunk's non-empty segments (spike rows inside the tile), compacted: one record (row-store
	// offset, source) read per event, the first flat event index, and for finding an event's
	// segment without a search (chunks of at most 32 * SW events) a bitmap of the flat indices
	// where segments start, each word paired with the segment holding its first event
	constexpr int SW = BS;
	struct alignas( 16 ) seg_rec
	{
		std::int64_t off; // row-store index of the segment's entry f is off + f
		std::int32_t src;
		std::int32_t pad;
	};
	__shared__ seg_rec seg[BS];
	__shared__ std::int32_t seg_pre[BS];
	__shared__ std::uint64_t seg_map[SW]; // low half: start bitmap word, high half: its first event's segment
	__shared__ std::uint64_t seg_total;
	constexpr bool eager = plastic && has_neuron_traces_v<S>;
	constexpr bool lazy = plastic && !eager;
	constexpr bool deferred = eager && !eager_nodefer_v<Model>; // delivery applies step T - 1's events
	// record and target id of each entry loaded without branches (entries past the chunk's end
	// read its last entry and are ignored)
	constexpr bool flat = !plastic || ( eager && !deferred );
	__shared__ std::int32_t seg_synced[lazy ? BS : 1];
	__shared__ std::int32_t seg_pend[lazy || deferred ? BS : 1]; // deferred: the row's spike also arrived at T - 1
	__shared__ float seg_xf[deferred ? BS : 1];                  // deferred: the row's pre trace after T - 1
	__shared__ std::uint8_t tile_f8[deferred ? TM : 1];    // deferred: the tile's step T - 1 flags
	__shared__ float tile_y[deferred ? TM : 1];            // deferred: the tile's post traces after T - 1
	__shared__ std::uint64_t wsum[BS / 64];

	std::int64_t const T = v.ctl->step + ( AHEAD ? 1 : 0 );
	int const slot = ring_at( v, ( AHEAD ? 1 : 0 ) - v.D ); // ring entry of step T - D
	std::int32_t const nall = T < v.D ? 0 : std::int32_t( v.counts[slot] );
	std::int32_t const s_beg = std::int32_t( std::int64_t( nall ) * part / v.rparts );
	std::int32_t const nspk = std::int32_t( std::int64_t( nall ) * ( part + 1 ) / v.rparts ) - s_beg;
	if( !v.overlap && !v.fused && v.slog && work == 0 && threadIdx.x == 0 )
	{
		// spike log: step T's ids start where step T-1's (final by now) end
		std::int64_t const k = T - v.slog->t0;
		if( k >= 0 && k < v.slog->nsteps )
			v.slog_off[k] = k == 0 ? 0ull : v.slog_off[k - 1] + v.counts[ring_at( v, ( AHEAD ? 1 : 0 ) - 1 )];
	}
	std::int32_t const * spk = v.ring + std::int64_t( slot ) * v.N + s_beg;
	float const dt = step_dt( v );
	snn_info const info = info_of( v );
	int const tid = threadIdx.x;
	unsigned long long ev = 0, pev = 0;
	// deferred: step T - 1's events of the delivered rows were left to this delivery, unless a
	// synchronization (state export) has applied them already
	bool const catch_up = deferred && T >= 1 && *v.caught != T - 1;
	std::int64_t const prev = ( ( T - 1 ) & 1 ) * std::int64_t( v.N ); // step T - 1's arrays
	{
		std::int32_t const t0 = tile * v.tile_w;
		std::int32_t const tw = ( v.N - t0 ) < v.tile_w ? ( v.N - t0 ) : v.tile_w;
		for( int x = tid; x < nacc * TM; x += BS ) acc[x] = word( 0 );
		if constexpr( deferred )
			if( catch_up ) // what step T - 1's deferred events read of their targets
				for( int x = tid; x < tw; x += BS )
				{
					tile_f8[x] = v.fl8[prev + t0 + x];
					tile_y[x] = v.ytr[prev + t0 + x];
				}
		// the tile's values of the attributes receive() reads (they do not change during
		// delivery: the neuron update runs after it; a split receive() reads none)
		util::for_n<N::size>( [&]( auto I ) {
			if constexpr( ( ( smask >> I ) & 1u ) != 0 )
			{
				using T = typename N::template type<I>;
				using B = std::conditional_t<sizeof( T ) == 1, std::uint8_t, std::conditional_t<sizeof( T ) == 2, std::uint16_t, std::uint32_t>>;
				constexpr int b = __builtin_popcount( smask & ( ( 1u << I ) - 1u ) );
				for( int x = tid; x < tw; x += BS ) stage[b * TM + x] = std::uint32_t( __builtin_bit_cast( B, ld<T>( v.na[I], t0 + x ) ) );
			}
		} );
		__syncthreads();

		for( std::int32_t c0 = 0; c0 < nspk; c0 += BS )
		{
			std::int32_t const cnt = ( nspk - c0 ) < BS ? ( nspk - c0 ) : BS;
			std::int32_t len = 0, src = 0, pend = 0, synced = 0;
			std::int64_t beg = 0;
			float xf = 0.0f;
			if( tid < cnt )
			{
				src = spk[c0 + tid];
				std::int32_t const * pv = v.pivots + std::int64_t( src ) * ( v.ntiles + 1 ) + tile;
				std::int32_t const a = pv[0], b = pv[1];
				beg = v.row_ptr[src] + a;
				len = b - a;
				if constexpr( deferred )
				{
					pend = catch_up && ( v.fl8[prev + src] & F8_ARRIVED );
					xf = catch_up ? v.xtr[prev + src] : 0.0f;
				}
				else if constexpr( lazy )
				{
					synced = v.synced[src];
					pend = v.pend[src];
				}
			}
			for( int w = tid; w < SW; w += BS ) seg_map[w] = 0ull; // ordered by the scan's barriers
			// one exclusive block scan of (non-empty segments, events): segment c of the chunk
			// holds the flat events [pre, pre + len)
			std::uint64_t const sc = block_exclusive_scan( ( std::uint64_t( len > 0 ) << 32 ) | std::uint32_t( len ), wsum, &seg_total );
			std::int32_t const pre = std::int32_t( std::uint32_t( sc ) ), c = std::int32_t( sc >> 32 );
			std::int32_t const total = std::int32_t( std::uint32_t( seg_total ) ), nseg = std::int32_t( seg_total >> 32 );
			bool const bitmap = total <= 32 * SW;
			if( len > 0 )
			{
				seg[c] = seg_rec{ beg - pre, src, 0 };
				seg_pre[c] = pre;
				if constexpr( deferred )
				{
					seg_pend[c] = pend;
					seg_xf[c] = xf;
				}
				else if constexpr( lazy )
				{
					seg_synced[c] = synced;
					seg_pend[c] = pend;
				}
				if( bitmap )
				{
					auto * const words = reinterpret_cast<std::uint32_t *>( seg_map ); // little endian: [2w] bits, [2w + 1] segment
					atomicOr( &words[2 * ( pre >> 5 )], 1u << ( pre & 31 ) );
					for( std::int32_t w = ( pre + 31 ) >> 5; ( w << 5 ) < pre + len; w++ ) words[2 * w + 1] = std::uint32_t( c );
				}
			}
			__syncthreads();

			// the flat walk: 4 events per thread in flight while more than two workgroups' worth
			// remain, then 2 (a short last pass, or a small tile, runs fewer idle searches)
			auto walk = [&]( auto uu, std::int32_t f0 ) {
				constexpr int U = decltype( uu )::value;
				std::int32_t kk[U], ff[U], sv[U];
				std::int64_t e[U];
				std::int32_t j[U];
				bool ok[U];
#pragma unroll
				for( int u = 0; u < U; u++ )
				{
					ff[u] = f0 + tid + BS * u;
					ok[u] = ff[u] < total;
				}
				if( bitmap )
				{
					// the segment holding the word's first event, plus the segments that start in
					// the word after it and at or before the event
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						std::int32_t const f = ok[u] ? ff[u] : total - 1;
						std::uint32_t const upto = ( 2u << ( f & 31 ) ) - 1u;
						std::uint64_t const m = seg_map[f >> 5];
						kk[u] = std::int32_t( m >> 32 ) + __builtin_popcount( std::uint32_t( m ) & upto & ~1u );
					}
				}
				else
					segments_of<U>( seg_pre, nseg, ff, kk );
				if constexpr( flat )
				{
					// all U records, then all U target ids, without branches: an entry past the
					// chunk's end reads the chunk's last entry (a valid address) and is ignored
					// (models whose delivery applies plasticity keep the branches: the extra live
					// registers cost them a wave per SIMD)
					seg_rec r[U];
#pragma unroll
					for( int u = 0; u < U; u++ ) r[u] = seg[kk[u]];
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						e[u] = r[u].off + ( ok[u] ? ff[u] : total - 1 );
						sv[u] = r[u].src;
					}
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						std::int32_t const jj = lds<( SPICE_NT & 4 ) != 0>( v.col + e[u] );
						j[u] = ok[u] ? jj : 0;
					}
				}
				else
				{
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						seg_rec const r = seg[kk[u]];
						e[u] = r.off + ff[u];
						sv[u] = r.src;
						j[u] = ok[u] ? lds<( SPICE_NT & 4 ) != 0>( v.col + e[u] ) : 0;
					}
				}
				if constexpr( deferred )
				{
					// neuron-trace synapses: step T - 1's event of each plastic synapse (the row's
					// spike arrived at T - 1, or the target spiked at T - 1), then delivery. The
					// state, the target's flags and (for a pre event) its post trace are loaded
					// for all u before any event is applied.
					typename S::tuple_t sy[U];
					std::uint8_t f8[U];
					float yv[U];
#pragma unroll
					for( int u = 0; u < U; u++ )
						if( ok[u] ) load_tuple<S, ( SPICE_NT & 4 ) != 0>( sy[u], v.sa, e[u] );
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						bool const pl = ok[u] && catch_up && S::plastic( sv[u], j[u], info );
						f8[u] = pl ? std::uint8_t( tile_f8[j[u] - t0] | ( seg_pend[kk[u]] ? 0x80 : 0 ) ) : std::uint8_t( 0 );
						yv[u] = pl && seg_pend[kk[u]] ? tile_y[j[u] - t0] : 0.0f;
					}
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						if( !ok[u] ) continue;
						std::int32_t const s = sv[u];
						tile_backend<FIXED> bak{ make_backend<FIXED>( v, std::uint32_t( s ), std::uint32_t( j[u] ), std::uint32_t( T ), util::TAG_RECEIVE ) };
						tile_iter<N, FIXED, TM> dst{ &v, j[u], t0, acc, stage };
						reg_iter<typename S::tuple_t> itr{ sy[u], std::uint32_t( e[u] ) };
						bool const pre = ( f8[u] & 0x80 ) != 0, post = ( f8[u] & F8_SPIKED ) != 0;
						if( pre || post )
						{
							// the event of step T - 1, keyed like the pass that would have applied it
							auto sbak = make_backend<FIXED>( v, std::uint32_t( e[u] ), std::uint32_t( std::uint64_t( e[u] ) >> 32 ),
							                                 std::uint32_t( T - 1 ), util::TAG_SYNAPSE_UPDATE );
							typename S::tuple_t const before = sy[u];
							S::on_event( itr, pre, post, post ? seg_xf[kk[u]] : 0.0f, yv[u], info, sbak );
							store_changed<S, ( SPICE_NT & 4 ) != 0>( before, sy[u], v.sa, e[u] );
							pev++;
						}
						deliver_event<N>( s, dst, itr, info, bak );
					}
				}
				else if constexpr( eager )
				{
					// neuron-trace synapses without deferred events: the synapse state of all U
					// entries (valid addresses, see above), then delivery
					typename S::tuple_t sy[U];
#pragma unroll
					for( int u = 0; u < U; u++ ) load_tuple<S, ( SPICE_NT & 4 ) != 0>( sy[u], v.sa, e[u] );
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						if( !ok[u] ) continue;
						std::int32_t const s = sv[u];
						tile_backend<FIXED> bak{ make_backend<FIXED>( v, std::uint32_t( s ), std::uint32_t( j[u] ), std::uint32_t( T ), util::TAG_RECEIVE ) };
						tile_iter<N, FIXED, TM> dst{ &v, j[u], t0, acc, stage };
						reg_iter<typename S::tuple_t> itr{ sy[u], std::uint32_t( e[u] ) };
						deliver_event<N>( s, dst, itr, info, bak );
					}
				}
				else if constexpr( !plastic )
				{
					// the staged attribute values of all U targets first (one LDS round), then
					// receive() for each entry
					using iter_v = tile_iter_vals<N, FIXED, TM>;
					std::uint32_t svals[U][iter_v::NS];
#pragma unroll
					for( int u = 0; u < U; u++ )
#pragma unroll
						for( int b = 0; b < iter_v::NS; b++ )
							svals[u][b] = nstaged > 0 ? stage[b * TM + ( ok[u] ? j[u] - t0 : 0 )] : 0u;
#pragma unroll
					for( int u = 0; u < U; u++ )
					{
						if( !ok[u] ) continue;
						std::int32_t const s = sv[u];
						tile_backend<FIXED> bak{ make_backend<FIXED>( v, std::uint32_t( s ), std::uint32_t( j[u] ), std::uint32_t( T ), util::TAG_RECEIVE ) };
						iter_v dst{ &v, j[u], t0, acc, {} };
#pragma unroll
						for( int b = 0; b < iter_v::NS; b++ ) dst.sval[b] = svals[u][b];
						deliver_event<N>( s, dst, no_syn_iter{}, info, bak );
					}
				}
				else
				{
#pragma unroll
				for( int u = 0; u < U; u++ )
				{
					if( !ok[u] ) continue;
					std::int32_t const s = sv[u];
					tile_backend<FIXED> bak{ make_backend<FIXED>( v, std::uint32_t( s ), std::uint32_t( j[u] ), std::uint32_t( T ), util::TAG_RECEIVE ) };
					tile_iter<N, FIXED, TM> dst{ &v, j[u], t0, acc, stage };
					if constexpr( plastic )
					{
						typename S::tuple_t sy;
						load_tuple<S>( sy, v.sa, e[u] );
						reg_iter<typename S::tuple_t> itr{ sy, std::uint32_t( e[u] ) };
						if( S::plastic( s, j[u], info ) )
						{
							std::int32_t const synced = seg_synced[kk[u]], pend = seg_pend[kk[u]];
							replay<S>( itr, v, j[u], std::int32_t( T - 1 ), synced, pend, dt, info, bak );
							store_tuple<S>( sy, v.sa, e[u] );
							pev++;
						}
						deliver_event<N>( s, dst, itr, info, bak );
					}
					else
						deliver_event<N>( s, dst, no_syn_iter{}, info, bak );
				}
				}
			};
			std::int32_t f0 = 0;
			for( ; total - f0 > 2 * BS; f0 += 4 * BS ) walk( std::integral_constant<int, 4>{}, f0 );
			for( ; f0 < total; f0 += 2 * BS ) walk( std::integral_constant<int, 2>{}, f0 );
			if( tid == 0 ) ev += std::uint64_t( total );
			__syncthreads(); // segments are overwritten by the next chunk
		}

		// the tile's sums -> step T's global accumulators (zero here). One part: this workgroup
		// owns them, plain stores; several parts: atomic adds (integer and thus exact in
		// fixed-point mode; in fp32 mode the parts' order varies, as LDS atomic order already does)
		bool const owned = v.rparts == 1;
		std::int64_t const apar = ( T & 1 ) * std::int64_t( v.N );
		util::for_n<N::size>( [&]( auto I ) {
			if constexpr( !KEEP && ( ( rmask >> I ) & 1u ) != 0 )
			{
				using A = acc_type<typename N::template type<I>, FIXED>;
				constexpr int a = __builtin_popcount( rmask & ( ( 1u << I ) - 1u ) );
				A const * src = reinterpret_cast<A const *>( acc + a * TM );
				for( int x = tid; x < tw; x += BS )
				{
					A const val = src[x];
					if( val == A( 0 ) ) continue;
					if( owned )
						st<A>( v.acc[I], apar + t0 + x, val );
					else if constexpr( std::is_same_v<A, long long> )
						atomicAdd( reinterpret_cast<unsigned long long *>( v.acc[I] ) + apar + t0 + x, static_cast<unsigned long long>( val ) );
					else
						atomicAdd( static_cast<A *>( v.acc[I] ) + apar + t0 + x, val );
				}
			}
		} );
		__syncthreads();
	}
	block_add_stats<2>( v, work, { ev, pev }, { ST_EVENTS, ST_PLASTIC } );
	return acc;
}

template <typename Model, bool FIXED>
__device__ inline void stdp_block( dview const & v, std::int64_t T, std::int64_t blk, std::int64_t nblk, std::int64_t slot0 );
template <typename Model, bool FIXED>
__device__ inline void flush_block( dview const & v, std::int64_t T, std::int64_t blk, std::int64_t nblk, std::int64_t slot0 );

// Plasticity runs inside the delivery kernel: the plasticity work of step T - 1 (eager pre
// and post passes, or the lazy flush of the rows queued at T - 1) touches no row that step T
// delivers, so workgroups beyond the delivery ones do it while delivery runs, and the next
// kernel in stream order (the update of step T, then the delivery of step T + 1 with step T's
// plasticity) sees it done. Skipped when a synchronization has done it already
// (v.caught = T - 1).
template <typename Model, bool FIXED>
__device__ inline void plasticity_prev_block( dview const & v, std::int64_t blk, std::int64_t nblk )
{
	std::int64_t const T = v.ctl->step;
	if( T < 1 || *v.caught == T - 1 ) return;
	std::int64_t const slot0 = std::int64_t( v.ntiles ) * v.rparts;
	if constexpr( has_neuron_traces_v<typename Model::synapse> )
		stdp_block<Model, FIXED>( v, T - 1, blk, nblk, slot0 );
	else
		flush_block<Model, FIXED>( v, T - 1, blk, nblk, slot0 );
}

// Workgroups [0, grecv) deliver (work items round robin); plastic models give the workgroups
// [grecv, gridDim.x) to step T - 1's plasticity work.
template <typename Model, bool FIXED, int BS, int TM = TILE_MAX>
__global__ __launch_bounds__( BS ) void k_receive_tiled( dview v, std::int32_t grecv )
{
	if( std::int32_t( blockIdx.x ) < grecv )
	{
		std::int32_t const nwork = v.nact * v.rparts;
		for( std::int32_t w = blockIdx.x; w < nwork; w += grecv ) receive_tile<Model, FIXED, BS, false, TM>( v, w );
	}
	else if constexpr( is_plastic_v<Model> )
		plasticity_prev_block<Model, FIXED>( v, blockIdx.x - grecv, gridDim.x - grecv );
}

template <typename Model, bool FIXED>
__device__ inline void update_block( dview const & v, std::int64_t blk, std::int64_t slot );

// The overlapped step (overlap_capable_v models, delay >= 2), clock of step T: workgroups
// [0, nrecv) deliver step T + 1, [nrecv, nrecv + nplast) run step T - 1's plasticity passes,
// the rest update the neurons of step T (blockDim.x neurons each; statistics slots from
// upd_slot0). None reads what another writes: the accumulators, plasticity lists, traces and
// flags are kept by step parity, the delivery reads ring entries of earlier steps and no target
// state (the update tests receive()'s gate), and the passes touch only weights whose targets
// do not take that delivery's input. With nrecv = 0 the kernel ends a run of overlapped steps
// (the last update, beside the passes of the step before it).
template <typename Model, bool FIXED, int BS>
__global__ __launch_bounds__( BS ) void k_step_overlap( dview v, std::int32_t nrecv, std::int32_t nplast, std::int32_t upd_slot0 )
{
	std::int32_t const b = blockIdx.x;
	if( b < nrecv )
	{
		std::int32_t const nwork = v.nact * v.rparts;
		for( std::int32_t w = b; w < nwork; w += nrecv ) receive_tile<Model, FIXED, BS, true>( v, w );
	}
	else if( b < nrecv + nplast )
	{
		if constexpr( is_plastic_v<Model> ) plasticity_prev_block<Model, FIXED>( v, b - nrecv, nplast );
	}
	else
		update_block<Model, FIXED>( v, b - nrecv - nplast, upd_slot0 + ( b - nrecv - nplast ) );
}

// Batched delivery: global work item gt belongs to net tile_net[gt] (local work item
// tile_local[gt] = tile * rparts + part); the net's view is read from device memory. Plastic
// models: workgroup grecv + b runs plasticity block block_local[b] of the block_count[b]
// blocks of net block_net[b].
template <typename Model, bool FIXED, int BS>
__global__ __launch_bounds__( BS ) void k_batch_receive( dview const * views, std::int32_t const * tile_net, std::int32_t const * tile_local,
                                                         std::int32_t ntiles, std::int32_t grecv, std::int32_t const * block_net,
                                                         std::int32_t const * block_local, std::int32_t const * block_count )
{
	if( std::int32_t( blockIdx.x ) < grecv )
	{
		for( std::int32_t gt = blockIdx.x; gt < ntiles; gt += grecv )
			receive_tile<Model, FIXED, BS>( views[tile_net[gt]], tile_local[gt] );
	}
	else if constexpr( is_plastic_v<Model> )
	{
		std::int32_t const b = blockIdx.x - grecv;
		plasticity_prev_block<Model, FIXED>( views[block_net[b]], block_local[b], block_count[b] );
	}
}

// Batched overlapped step (every net of the batch overlapped): k_step_overlap over the nets'
// tables -- workgroups [0, grecv) deliver each net's next step (work items round robin),
// [grecv, grecv + nplast) run plasticity block flush_local[b] of the flush_count[b] blocks of
// net flush_net[b], the rest update neuron block upd_local[u] of net upd_net[u] (256 neurons).
// grecv = 0: the last update of a run, beside the passes of the step before it.
template <typename Model, bool FIXED, int BS>
__global__ __launch_bounds__( BS ) void k_batch_step_overlap( dview const * views, std::int32_t const * tile_net, std::int32_t const * tile_local,
                                                              std::int32_t ntiles, std::int32_t grecv, std::int32_t const * flush_net,
                                                              std::int32_t const * flush_local, std::int32_t const * flush_count, std::int32_t nplast,
                                                              std::int32_t const * upd_net, std::int32_t const * upd_local )
{
	std::int32_t const b = blockIdx.x;
	if( b < grecv )
	{
		for( std::int32_t gt = b; gt < ntiles; gt += grecv ) receive_tile<Model, FIXED, BS, true>( views[tile_net[gt]], tile_local[gt] );
	}
	else if( b < grecv + nplast )
	{
		if constexpr( is_plastic_v<Model> )
			plasticity_prev_block<Model, FIXED>( views[flush_net[b - grecv]], flush_local[b - grecv], flush_count[b - grecv] );
	}
	else
	{
		std::int32_t const u = b - grecv - nplast;
		dview const & v = views[upd_net[u]];
		update_block<Model, FIXED>( v, upd_local[u], v.upd_slot0 + upd_local[u] );
	}
}



// Neuron update of neurons [first, first + min(count, blockDim.x)) (blockDim.x <= 1024) at step
// T; statistics go to slot `slot`. The range that starts at neuron 0 also clears the next
// step's counters and writes the next step's clock.
// lds (the fused step): the inputs are the tile sums a KEEP delivery left in LDS for neurons
// lds_t0 ..; [attribute][lds_stride words]) instead of the global accumulators, which stay zero.
template <typename Model, bool FIXED>
__device__ inline void update_range( dview const & v, std::int64_t first, std::int32_t count, std::int64_t slot, void const * lds = nullptr,
                                     std::int32_t lds_t0 = 0, std::int32_t lds_stride = 0 )
{
	using N = typename Model::neuron;
	using S = typename Model::synapse;
	constexpr bool plastic = is_plastic_v<Model>;
	constexpr bool eager = plastic && has_neuron_traces_v<S>;
	constexpr std::uint32_t rmask = receive_mask_v<N>;

	std::int64_t const T = v.ctl->step;
	int const wslot = ring_at( v, 0 );
	if( first == 0 && threadIdx.x == 0 )
	{
		v.counts[ring_at( v, 1 )] = 0; // step T - D - 1's entry (delivered at T - 1); refilled at T + 1
		if constexpr( plastic ) v.fl_count[( T + 1 ) & 3] = 0; // step T - 1's flush may still read slot (T - 1) & 3
		if constexpr( eager ) v.po_count[( T + 1 ) & 3] = 0;
		// next step's clock: exactly kahan_sum::add on the host
		ctl_t const c = *v.ctl;
		float const y = v.DT - c.c;
		float const t = c.sum + y;
		v.ctl_next->c = ( t - c.sum ) - y;
		v.ctl_next->sum = t;
		v.ctl_next->step = c.step + 1;
		v.ctl_next->rslot = c.rslot + 1 == v.D + 2 ? 0 : c.rslot + 1;
	}
	float const dt = step_dt( v );
	snn_info const info = info_of( v );
	std::int64_t const i = first + threadIdx.x;
	bool spiked = false, flush = false, post = false;
	std::int32_t fl_synced = 0, fl_pend = 0;

	if( std::int32_t( threadIdx.x ) < count && i < v.N )
	{
		typename N::tuple_t s;
		load_tuple<N>( s, v.na, i );
		// the plastic bookkeeping this neuron's step reads does not depend on the update: load
		// it now, with the state (after the state's store the loads would wait for the update)
		std::int64_t const N64 = v.N;
		std::int64_t const hcur = ( ( T & ( HISTORY - 1 ) ) >> 6 ) * N64 + i; // step T's history word
		unsigned long long h_arr = 0, h_next = 0, h_now = 0;
		float x_prev = 0.0f, y_prev = 0.0f;
		std::uint8_t has_plastic = 0;
		std::int64_t in_lo = 0, in_hi = 0;
		std::int32_t synced0 = 0, pend0 = 0;
		if constexpr( plastic )
		{
			if( T >= v.D ) h_arr = v.hring[( ( ( T - v.D ) & ( HISTORY - 1 ) ) >> 6 ) * N64 + i];
			if( !eager_nodefer_v<Model> && T + 1 >= v.D && v.D != 1 ) h_next = v.hring[( ( ( T + 1 - v.D ) & ( HISTORY - 1 ) ) >> 6 ) * N64 + i];
			h_now = v.hring[hcur];
			has_plastic = v.hasp[i];
			if constexpr( eager )
			{
				std::int64_t const prv = N64 - ( T & 1 ) * N64;
				x_prev = v.xtr[prv + i];
				y_prev = v.ytr[prv + i];
				in_lo = v.tp_ptr[i];
				in_hi = v.tp_ptr[i + 1];
			}
			else
			{
				synced0 = v.synced[i];
				pend0 = v.pend[i];
			}
		}
		// step T's inputs (accumulators of step parity T & 1); a split receive()'s gate is tested
		// here, on the state the delivery's receive() would have read: a target that does not
		// accept drops the step's input
		bool open = true;
		if constexpr( gated_receive_v<N> ) open = N::accepts( reg_iter<typename N::tuple_t>{ s, std::uint32_t( i ) } );
		std::int64_t const ai = ( T & 1 ) * N64 + i;
		util::for_n<N::size>( [&]( auto I ) {
			if constexpr( ( ( rmask >> I ) & 1u ) != 0 )
			{
				using A = typename N::template type<I>;
				using AC = acc_type<A, FIXED>;
				AC a;
				if( lds )
				{
					constexpr int slotI = __builtin_popcount( rmask & ( ( 1u << I ) - 1u ) );
					a = reinterpret_cast<AC const *>( static_cast<acc_word<N, FIXED> const *>( lds ) + slotI * lds_stride )[i - lds_t0];
				}
				else
				{
					a = ld<AC>( v.acc[I], ai );
					st<AC>( v.acc[I], ai, AC( 0 ) );
				}
				if constexpr( std::is_floating_point_v<A> && FIXED )
				{
					if( open ) std::get<I>( s ) = std::get<I>( s ) + float( a ) * __builtin_ldexpf( 1.0f, -v.shift );
				}
				else
				{
					if( open ) std::get<I>( s ) = A( std::get<I>( s ) + a );
				}
			}
		} );

		auto bak = make_backend<FIXED>(
		    v, std::uint32_t( i ), std::uint32_t( T ), std::uint32_t( std::uint64_t( T ) >> 32 ), util::TAG_NEURON_UPDATE );
		spiked = N::update( reg_iter<typename N::tuple_t>{ s, std::uint32_t( i ) }, dt, info, bak );
		store_tuple<N>( s, v.na, i );

		if constexpr( plastic )
		{
			// this neuron's spike of step T - D reaches its synapses now
			bool const arrived = T >= v.D && ( ( h_arr >> ( ( T - v.D ) & 63 ) ) & 1ull );
			// and its spike of step T + 1 - D at the next step (recorded already unless D = 1); only
			// models with deferred events (or lazy plasticity) leave such rows to that delivery
			bool const arrives_next = !eager_nodefer_v<Model> && T + 1 >= v.D &&
			                          ( v.D == 1 ? spiked : ( ( h_next >> ( ( T + 1 - v.D ) & 63 ) ) & 1ull ) != 0 );
			{ // record this step in the history ring (overwriting step T - HISTORY)
				unsigned long long const bit = 1ull << ( T & 63 );
				unsigned long long const nw = spiked ? ( h_now | bit ) : ( h_now & ~bit );
				if( nw != h_now ) v.hring[hcur] = nw;
			}
			if constexpr( eager )
			{
				// traces after step T (the dense checker's per-step recurrence) and the step's
				// flags, into the arrays of step parity T & 1
				std::int64_t const cur = ( T & 1 ) * N64;
				float const d = S::trace_decay( dt );
				v.xtr[cur + i] = fmaf( x_prev, d, arrived ? 1.0f : 0.0f );
				v.ytr[cur + i] = fmaf( y_prev, d, spiked ? 1.0f : 0.0f );
				v.fl8[cur + i] = std::uint8_t( ( spiked ? F8_SPIKED : 0 ) | ( arrived ? F8_ARRIVED : 0 ) | ( arrives_next ? F8_ARRIVES_NEXT : 0 ) );
				// pre pass: a row whose spike arrives now, unless (deferred events) it arrives again
				// at T + 1 and that delivery applies this step's events of the row; post pass: a
				// spiking neuron with incoming plastic synapses
				flush = arrived && !arrives_next && has_plastic;
				post = spiked && in_hi > in_lo;
			}
			else
			{
				std::int32_t synced = synced0, pend = pend0;
				bool dirty = false;
				if( arrived )
				{
					synced = std::int32_t( T - 1 );
					pend = 1;
					dirty = true;
				}
				if( spiked ) v.last_spk[i] = std::int32_t( T );
				// Rows are flushed before their window overflows; a row never processed yet gets
				// its first flush at step (i mod age), which spreads flushes evenly over steps.
				constexpr int age = FLUSH_AGE;
				// Rows go to this step's flush list (and are current through T afterwards) when
				//  * their spike arrived now: the pending pre event of step T is applied right
				//    away, or
				//  * their history window is full.
				// A row whose spike arrives at T + 1 is left to that delivery instead: the flush
				// of step T may run while step T + 1 delivers, so their rows must differ.
				if( has_plastic && !arrives_next && ( pend || T - synced >= age || ( synced < 0 && T >= i % age ) ) )
				{
					flush = true;
					fl_synced = synced;
					fl_pend = pend;
					synced = std::int32_t( T );
					pend = 0;
					dirty = true;
				}
				if( dirty )
				{
					v.synced[i] = synced;
					v.pend[i] = pend;
				}
			}
		}
	}

	// Spike compaction: ballot + popcount per wave, prefix over the block's waves in LDS, one
	// atomic per block on the step's counter, mbcnt gives each spiking lane its slot.
	__shared__ unsigned wcount[16][3]; // at most 1024 threads
	__shared__ unsigned bbase[3];
	__shared__ std::uint64_t lbase; // overlapped step: where step T's ids start in the spike log
	int const w = threadIdx.x >> 6;
	std::uint64_t const m = ballot( spiked );
	std::uint64_t const fm = plastic ? ballot( flush ) : 0ull;
	std::uint64_t const pm = eager ? ballot( post ) : 0ull;
	if( lane_id() == 0 )
	{
		wcount[w][0] = unsigned( __popcll( m ) );
		wcount[w][1] = unsigned( __popcll( fm ) );
		wcount[w][2] = unsigned( __popcll( pm ) );
	}
	__syncthreads();
	if( threadIdx.x == 0 )
	{
		// (the spike log's loads go out before the counter atomics: they do not depend on them)
		if( v.slog )
		{
			// where step T's ids start in the log: written by this step's delivery kernel, or
			// (the overlapped step, whose delivery of step T ran beside step T - 1's update; the
			// fused step, whose delivery may not have written it yet) here, from step T - 1's
			// offset and count, both final (written by an earlier kernel; this update clears
			// the count of step T + 1, not of T - 1)
			std::int64_t const k = T - v.slog->t0;
			std::uint64_t base = 0;
			if( k >= 0 && k < v.slog->nsteps )
			{
				if( !v.overlap && !v.fused )
					base = v.slog_off[k];
				else
				{
					base = k == 0 ? 0ull : v.slog_off[k - 1] + v.counts[ring_at( v, -1 )];
					if( first == 0 ) v.slog_off[k] = base;
				}
			}
			lbase = base;
		}
		unsigned t0 = 0, t1 = 0, t2 = 0;
		for( int k = 0; k < int( blockDim.x >> 6 ); k++ )
		{
			unsigned const a = wcount[k][0], b = wcount[k][1], c = wcount[k][2];
			wcount[k][0] = t0;
			wcount[k][1] = t1;
			wcount[k][2] = t2;
			t0 += a;
			t1 += b;
			t2 += c;
		}
		bbase[0] = t0 ? atomicAdd( &v.counts[wslot], t0 ) : 0u;
		if constexpr( plastic ) bbase[1] = t1 ? atomicAdd( &v.fl_count[T & 3], t1 ) : 0u;
		if constexpr( eager ) bbase[2] = t2 ? atomicAdd( &v.po_count[T & 3], t2 ) : 0u;
		if( t0 ) v.bstats[slot * NSTAT + ST_SPIKES] += t0;
	}
	__syncthreads();
	if( spiked )
	{
		unsigned const o = bbase[0] + wcount[w][0] + lanes_below( m );
		v.ring[std::int64_t( wslot ) * v.N + o] = std::int32_t( i );
		if( v.slog )
		{
			std::int64_t const k = T - v.slog->t0;
			if( k >= 0 && k < v.slog->nsteps )
			{
				std::uint64_t const p = lbase + o;
				if( p < v.slog->cap ) v.slog_ids[p] = std::int32_t( i );
			}
		}
	}
	if constexpr( plastic )
	{
		if( flush )
		{
			std::int64_t const o = ( T & 1 ) * std::int64_t( v.N ) + bbase[1] + wcount[w][1] + lanes_below( fm );
			v.fl_id[o] = std::int32_t( i );
			if constexpr( !eager )
			{
				v.fl_synced[o] = fl_synced;
				v.fl_pend[o] = fl_pend;
			}
		}
	}
	if constexpr( eager )
	{
		if( post ) v.po_id[( T & 1 ) * std::int64_t( v.N ) + bbase[2] + wcount[w][2] + lanes_below( pm )] = std::int32_t( i );
	}
}

// Neuron update of block blk (blockDim.x neurons) at step T.
template <typename Model, bool FIXED>
__device__ inline void update_block( dview const & v, std::int64_t blk, std::int64_t slot )
{
	update_range<Model, FIXED>( v, blk * blockDim.x, std::int32_t( blockDim.x ), slot );
}

template <typename Model, bool FIXED, int BS = SPICE_UPDATE_BLOCK>
__global__ __launch_bounds__( BS ) void k_update( dview v )
{
	update_block<Model, FIXED>( v, blockIdx.x, blockIdx.x );
}

// The fused step (static models whose steps do not overlap, one delivery work item per tile,
// every tile a delivery tile, tiles of at most FUSE_CHUNKS workgroup widths): one kernel per
// step. The workgroup that delivers step T's arrivals to a tile holds all of the tile's inputs
// (its LDS sums, which never go to the global accumulators), so it goes on to update the
// tile's neurons at step T from them. No workgroup waits for another, and no delivery reads what an
// update of the same step writes (other tiles' neurons, the ring entry of step T, the counters
// of steps T and T + 1). Several work items per tile would need a device-scope fence per item
// to hand the sums to the tile's last one (with one L2 per XCD, such a fence writes back and
// invalidates the XCD's L2): measured 4.5-7x slower per step than two kernels. Update
// statistics go to slots from upd_slot0 (FUSE_CHUNKS per tile).
constexpr int FUSE_CHUNKS = 4;
template <typename Model, bool FIXED, int BS>
__global__ __launch_bounds__( BS ) void k_step_fused( dview v, std::int32_t grecv, std::int32_t upd_slot0 )
{
	for( std::int32_t w = blockIdx.x; w < v.nact; w += grecv )
	{
		void const * sums = receive_tile<Model, FIXED, BS, false, TILE_MAX, true>( v, w );
		std::int32_t const tile = v.tile_ids[w];
		std::int32_t const t0 = tile * v.tile_w;
		std::int32_t const tw = ( v.N - t0 ) < v.tile_w ? ( v.N - t0 ) : v.tile_w;
		for( std::int32_t c = 0; c < tw; c += BS )
			update_range<Model, FIXED>( v, t0 + c, tw - c, upd_slot0 + std::int64_t( tile ) * FUSE_CHUNKS + c / BS, sums, t0, TILE_MAX );
	}
}

// Batched neuron update: workgroup b updates block block_local[b] of net block_net[b].
template <typename Model, bool FIXED>
__global__ __launch_bounds__( 256 ) void k_batch_update( dview const * views, std::int32_t const * block_net, std::int32_t const * block_local )
{
	update_block<Model, FIXED>( views[block_net[blockIdx.x]], block_local[blockIdx.x], block_local[blockIdx.x] );
}

// Eager plasticity of step T (neuron-trace synapses): the pre pass and the post pass as one
// index space of work items, (row, part) of the pre rows then (target, part) of the spiking
// targets, taken round robin by the waves of workgroups blk, blk + nblk, ... (of this net).
//   pre  item: a slice of the row's plastic slots; every plastic synapse gets step T's event
//              (pre, post = target spiked at T) with the row's pre trace and the target's post
//              trace after step T.
//   post item: a slice of the target's incoming plastic synapses (by-target index); a
//              synapse whose source's spike arrives at T or T + 1 is left to the pre pass or
//              to the next delivery, the others get a post event with the source's pre trace.
// Both kinds write disjoint synapses (a synapse is in a pre row iff its source arrived at T).
// Statistics go to slots slot0 + blk.
#ifndef SPICE_POST_UNROLL
#define SPICE_POST_UNROLL 4
#endif
template <typename Model, bool FIXED>
__device__ inline void stdp_block( dview const & v, std::int64_t T, std::int64_t blk, std::int64_t nblk, std::int64_t slot0 )
{
	using S = typename Model::synapse;
	constexpr int U = 4;                  // pre pass: synapses per lane per iteration in flight
	constexpr int UP = SPICE_POST_UNROLL; // post pass: by-target entries per lane per iteration
	constexpr bool NTPOST = ( SPICE_NT & 1 ) != 0, NTPRE = ( SPICE_NT & 2 ) != 0;
	std::int64_t const cur = ( T & 1 ) * std::int64_t( v.N ); // step T's lists, traces and flags
	std::int64_t const npre = std::int64_t( v.fl_count[T & 3] ) * v.fparts;
	std::int64_t const nitems = npre + std::int64_t( v.po_count[T & 3] ) * v.pparts;
	snn_info const info = info_of( v );
	int const lane = lane_id();
	std::int64_t const wave0 = uniform64( ( blk * blockDim.x + threadIdx.x ) >> 6 );
	std::int64_t const nwaves = nblk * ( blockDim.x >> 6 );
	unsigned long long touched = 0, written = 0, words = 0, rows = 0;
	for( std::int64_t it = wave0; it < nitems; it += nwaves )
	{
		if( it < npre )
		{
			std::int64_t const f = it / v.fparts;
			std::int64_t const p = it - f * v.fparts;
			std::int32_t const s = uniform( v.fl_id[cur + f] );
			std::int64_t const r0 = v.row_ptr[s];
			std::int64_t const lo = r0 + v.pl_lo[s], hi = r0 + v.pl_hi[s];
			std::int64_t const chunk = ( ( hi - lo + v.fparts - 1 ) / v.fparts + 63 ) & ~std::int64_t( 63 );
			std::int64_t const beg = lo + p * chunk;
			std::int64_t const end = beg + chunk < hi ? beg + chunk : hi;
			float const xs = v.xtr[cur + s];
			for( std::int64_t base = beg; base < end; base += 64 * U )
			{
				std::int32_t j[U];
				bool ok[U];
#pragma unroll
				for( int u = 0; u < U; u++ )
				{
					std::int64_t const e = base + lane + 64 * u;
					ok[u] = e < end;
					j[u] = ok[u] ? lds<NTPRE>( v.col + e ) : 0;
				}
				typename S::tuple_t sy[U];
				std::uint8_t f8[U];
				float yv[U];
#pragma unroll
				for( int u = 0; u < U; u++ )
				{
					ok[u] = ok[u] && S::plastic( s, j[u], info );
					if( ok[u] )
					{
						load_tuple<S, NTPRE>( sy[u], v.sa, base + lane + 64 * u );
						f8[u] = v.fl8[cur + j[u]];
						yv[u] = v.ytr[cur + j[u]];
					}
				}
#pragma unroll
				for( int u = 0; u < U; u++ )
				{
					if( !ok[u] ) continue;
					std::int64_t const e = base + lane + 64 * u;
					auto bak = make_backend<FIXED>(
					    v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), std::uint32_t( T ), util::TAG_SYNAPSE_UPDATE );
					reg_iter<typename S::tuple_t> itr{ sy[u], std::uint32_t( e ) };
					bool const post = ( f8[u] & F8_SPIKED ) != 0;
					typename S::tuple_t const before = sy[u];
					S::on_event( itr, true, post, post ? xs : 0.0f, yv[u], info, bak );
					written += store_changed<S, NTPRE>( before, sy[u], v.sa, e ) ? 1 : 0;
					touched++;
				}
			}
			if( lane == 0 )
			{
				if( end > beg ) words += std::uint64_t( end - beg ); // target ids
				if( p == 0 ) rows++;
			}
		}
		else
		{
			std::int64_t const q = it - npre;
			std::int64_t const f = q / v.pparts;
			std::int64_t const p = q - f * v.pparts;
			std::int32_t const d = uniform( v.po_id[cur + f] );
			std::int64_t const lo = v.tp_ptr[d], hi = v.tp_ptr[d + 1];
			std::int64_t const chunk = ( ( hi - lo + v.pparts - 1 ) / v.pparts + 63 ) & ~std::int64_t( 63 );
			std::int64_t const beg = lo + p * chunk;
			std::int64_t const end = beg + chunk < hi ? beg + chunk : hi;
			for( std::int64_t base = beg; base < end; base += 64 * UP )
			{
				tp_entry te[UP];
				bool ok[UP];
#pragma unroll
				for( int u = 0; u < UP; u++ )
				{
					std::int64_t const k = base + lane + 64 * u;
					ok[u] = k < end;
					te[u] = ok[u] ? v.tp_ent[k] : tp_entry{ 0, 0 };
				}
				// the source's flags, row start and trace: independent loads, issued together
				std::int64_t e[UP];
				float xv[UP];
				std::uint8_t f8[UP];
#pragma unroll
				for( int u = 0; u < UP; u++ )
				{
					f8[u] = v.fl8[cur + te[u].src];
					e[u] = v.row_ptr[te[u].src] + te[u].off;
					xv[u] = v.xtr[cur + te[u].src];
				}
#pragma unroll
				for( int u = 0; u < UP; u++ ) ok[u] = ok[u] && !( f8[u] & ( F8_ARRIVED | F8_ARRIVES_NEXT ) );
				typename S::tuple_t sy[UP];
#pragma unroll
				for( int u = 0; u < UP; u++ )
					if( ok[u] ) load_tuple<S, NTPOST>( sy[u], v.sa, e[u] );
#pragma unroll
				for( int u = 0; u < UP; u++ )
				{
					if( !ok[u] ) continue;
					auto bak = make_backend<FIXED>(
					    v, std::uint32_t( e[u] ), std::uint32_t( std::uint64_t( e[u] ) >> 32 ), std::uint32_t( T ), util::TAG_SYNAPSE_UPDATE );
					reg_iter<typename S::tuple_t> itr{ sy[u], std::uint32_t( e[u] ) };
					typename S::tuple_t const before = sy[u];
					S::on_event( itr, false, true, xv[u], 0.0f, info, bak ); // post only: y is not read
					written += store_changed<S, NTPOST>( before, sy[u], v.sa, e[u] ) ? 1 : 0;
					touched++;
				}
			}
			if( lane == 0 )
			{
				if( end > beg ) words += 2 * std::uint64_t( end - beg ); // (source, offset) entries
				if( p == 0 ) rows++;
			}
		}
	}
	block_add_stats<4>( v, slot0 + blk, { touched, written, words, rows }, { ST_FLUSH_SYN, ST_FLUSH_WRITTEN, ST_FLUSH_SLOTS, ST_FLUSH_ROWS } );
}

// Eager plasticity of step `T` on its own (a synchronization before a state export applies
// the passes of the last completed step that the next delivery kernel would otherwise run).
template <typename Model, bool FIXED>
__global__ __launch_bounds__( 256 ) void k_stdp( dview v, std::int64_t T, std::int64_t slot0 )
{
	stdp_block<Model, FIXED>( v, T, blockIdx.x, gridDim.x, slot0 );
}

// Row-major flush of the rows queued by the update of step T (lazy plasticity of the
// per-synapse contract): work item = (row, part); waves of workgroups blk, blk + nblk, ...
// (of this net) take items round robin. v is a view of step T + 1 (its ctl_prev is step T's
// clock). Statistics go to slots slot0 + blk.
template <typename Model, bool FIXED>
__device__ inline void flush_block( dview const & v, std::int64_t T, std::int64_t blk, std::int64_t nblk, std::int64_t slot0 )
{
	using S = typename Model::synapse;
	static_assert( !has_neuron_traces_v<S>, "neuron-trace synapses use the eager passes" );
	constexpr int U = 4; // slots per lane per iteration: independent loads in flight
	std::int64_t const nitems = std::int64_t( v.fl_count[T & 3] ) * v.fparts;
	std::int64_t const fb = ( T & 1 ) * std::int64_t( v.N ); // step T's flush list
	float const dt = v.DT - v.ctl_prev->c;                    // step T's compensated dt
	snn_info const info = info_of( v );
	int const lane = lane_id();
	std::int64_t const wave0 = uniform64( ( blk * blockDim.x + threadIdx.x ) >> 6 );
	std::int64_t const nwaves = nblk * ( blockDim.x >> 6 );
	unsigned long long touched = 0, slots = 0, rows = 0;
	for( std::int64_t it = wave0; it < nitems; it += nwaves )
	{
		std::int64_t const f = it / v.fparts;
		std::int64_t const p = it - f * v.fparts;
		std::int32_t const s = uniform( v.fl_id[fb + f] );
		std::int32_t const synced = uniform( v.fl_synced[fb + f] );
		std::int32_t const pend = uniform( v.fl_pend[fb + f] );
		std::int64_t const r0 = v.row_ptr[s];
		std::int64_t const lo = r0 + v.pl_lo[s], hi = r0 + v.pl_hi[s];
		std::int64_t const chunk = ( ( hi - lo + v.fparts - 1 ) / v.fparts + 63 ) & ~std::int64_t( 63 );
		std::int64_t const beg = lo + p * chunk;
		std::int64_t const end = beg + chunk < hi ? beg + chunk : hi;
		for( std::int64_t base = beg; base < end; base += 64 * U )
		{
			std::int32_t j[U];
			bool ok[U];
#pragma unroll
			for( int u = 0; u < U; u++ )
			{
				std::int64_t const e = base + lane + 64 * u;
				ok[u] = e < end;
				j[u] = ok[u] ? v.col[e] : 0;
			}
#pragma unroll
			for( int u = 0; u < U; u++ ) ok[u] = ok[u] && S::plastic( s, j[u], info );
#pragma unroll
			for( int u = 0; u < U; u++ )
			{
				if( !ok[u] ) continue;
				std::int64_t const e = base + lane + 64 * u;
				typename S::tuple_t sy;
				load_tuple<S>( sy, v.sa, e );
				auto bak = make_backend<FIXED>(
				    v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), std::uint32_t( T ), util::TAG_SYNAPSE_UPDATE );
				reg_iter<typename S::tuple_t> itr{ sy, std::uint32_t( e ) };
				replay<S>( itr, v, j[u], std::int32_t( T ), synced, pend, dt, info, bak );
				store_tuple<S>( sy, v.sa, e );
				touched++;
			}
		}
		if( lane == 0 )
		{
			if( end > beg ) slots += std::uint64_t( end - beg );
			if( p == 0 ) rows++;
		}
	}
	block_add_stats<4>( v, slot0 + blk, { touched, touched, slots, rows }, { ST_FLUSH_SYN, ST_FLUSH_WRITTEN, ST_FLUSH_SLOTS, ST_FLUSH_ROWS } );
}

// The lazy flush of step `T` on its own (a synchronization before a state export runs the
// flush of the last completed step that the next delivery kernel would otherwise run); v is
// a view of step T + 1.
template <typename Model, bool FIXED>
__global__ __launch_bounds__( 256 ) void k_flush( dview v, std::int64_t T, std::int64_t slot0 )
{
	flush_block<Model, FIXED>( v, T, blockIdx.x, gridDim.x, slot0 );
}

// Opens the spike-log window of every net of a batch: the next nsteps steps, starting at the
// net's current step (read from its clock slot in `views`).
static __global__ __launch_bounds__( 256 ) void k_batch_log_window( dview const * views, std::int32_t n, std::int64_t nsteps )
{
	std::int32_t const k = std::int32_t( blockIdx.x * blockDim.x + threadIdx.x );
	if( k >= n ) return;
	dview const & v = views[k];
	if( !v.slog ) return;
	v.slog->t0 = v.ctl->step;
	v.slog->nsteps = nsteps;
}

// Brings every row with plastic synapses up to date through step `now` (= last completed
// step); one wave per row. Used before exporting synapse state (lazy plasticity).
template <typename Model>
__global__ __launch_bounds__( 256 ) void k_sync_all( dview v, std::int32_t now, float dt )
{
	using S = typename Model::synapse;
	static_assert( !has_neuron_traces_v<S>, "neuron-trace synapses are synchronized by k_sync_eager" );
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N || !v.hasp[r] ) return;
	snn_info const info = info_of( v );
	std::int32_t const synced = v.synced[r], pend = v.pend[r];
	std::int64_t const r0 = v.row_ptr[r];
	for( std::int64_t e = r0 + v.pl_lo[r] + lane_id(); e < r0 + v.pl_hi[r]; e += 64 )
	{
		std::int32_t const j = v.col[e];
		if( !S::plastic( std::int32_t( r ), j, info ) ) continue;
		typename S::tuple_t sy;
		load_tuple<S>( sy, v.sa, e );
		auto bak = make_backend<false>( v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), std::uint32_t( now ), util::TAG_SYNAPSE_UPDATE );
		reg_iter<typename S::tuple_t> it{ sy, std::uint32_t( e ) };
		replay<S>( it, v, j, now, synced, pend, dt, info, bak );
		store_tuple<S>( sy, v.sa, e );
	}
	if( lane_id() == 0 )
	{
		v.synced[r] = now;
		v.pend[r] = 0;
	}
}

// Eager plasticity: applies the events of step `now` (= last completed step) that were left
// to the next delivery -- those of the rows whose spike arrives at now + 1 -- exactly as that
// delivery would; one wave per row of the arriving list. The host then records `now` in
// v.caught so the delivery skips them. Used before exporting synapse state.
template <typename Model>
__global__ __launch_bounds__( 256 ) void k_sync_eager( dview v, std::int64_t now )
{
	using S = typename Model::synapse;
	std::int64_t const x = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	std::int64_t const T = now + 1; // the step whose delivery would apply them
	if( T < v.D ) return;
	int const slot = ring_slot( v, T - v.D ); // spikes of step T - D
	if( x >= std::int64_t( v.counts[slot] ) ) return;
	std::int32_t const r = v.ring[std::int64_t( slot ) * v.N + x];
	if( !v.hasp[r] ) return;
	snn_info const info = info_of( v );
	std::int64_t const prev = ( now & 1 ) * std::int64_t( v.N );
	bool const pre = ( v.fl8[prev + r] & F8_ARRIVED ) != 0;
	float const xs = v.xtr[prev + r];
	std::int64_t const r0 = v.row_ptr[r];
	for( std::int64_t e = r0 + v.pl_lo[r] + lane_id(); e < r0 + v.pl_hi[r]; e += 64 )
	{
		std::int32_t const j = v.col[e];
		if( !S::plastic( r, j, info ) ) continue;
		bool const post = ( v.fl8[prev + j] & F8_SPIKED ) != 0;
		if( !pre && !post ) continue;
		typename S::tuple_t sy;
		load_tuple<S>( sy, v.sa, e );
		auto bak = make_backend<false>( v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), std::uint32_t( now ), util::TAG_SYNAPSE_UPDATE );
		reg_iter<typename S::tuple_t> it{ sy, std::uint32_t( e ) };
		typename S::tuple_t const before = sy;
		S::on_event( it, pre, post, post ? xs : 0.0f, pre ? v.ytr[prev + j] : 0.0f, info, bak );
		store_changed<S>( before, sy, v.sa, e );
	}
}

// ---- by-target index of the plastic synapses (eager plasticity) ----------------------------
// Counts every target's incoming plastic synapses; one wave per source row.
template <typename Model>
__global__ __launch_bounds__( 256 ) void k_tp_count( dview v, std::uint32_t * cnt )
{
	using S = typename Model::synapse;
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N ) return;
	snn_info const info = info_of( v );
	for( std::int64_t e = v.row_ptr[r] + lane_id(); e < v.row_ptr[r + 1]; e += 64 )
	{
		std::int32_t const j = v.col[e];
		if( S::plastic( std::int32_t( r ), j, info ) ) atomicAdd( &cnt[j], 1u );
	}
}

// ptr[0..n] = exclusive prefix sums of cnt[0..n-1]; one workgroup of 1024 threads, each
// summing a contiguous slice (n is a neuron count: this runs once per network).
static inline __global__ __launch_bounds__( 1024 ) void k_tp_scan( std::uint32_t const * cnt, std::int64_t * ptr, std::int64_t n )
{
	__shared__ std::int64_t part[1024];
	int const t = threadIdx.x;
	std::int64_t const per = ( n + 1023 ) / 1024;
	std::int64_t const a = t * per < n ? t * per : n, b = a + per < n ? a + per : n;
	std::int64_t sum = 0;
	for( std::int64_t i = a; i < b; i++ ) sum += cnt[i];
	part[t] = sum;
	__syncthreads();
	if( t == 0 )
	{
		std::int64_t run = 0;
		for( int k = 0; k < 1024; k++ )
		{
			std::int64_t const x = part[k];
			part[k] = run;
			run += x;
		}
		ptr[n] = run;
	}
	__syncthreads();
	std::int64_t run = part[t];
	for( std::int64_t i = a; i < b; i++ )
	{
		ptr[i] = run;
		run += cnt[i];
	}
}

// Fills the index: every plastic synapse appends (source, offset in the source's row) to its
// target's list (cursor: a copy of ptr), in arbitrary order; the host then sorts each list by
// source.
template <typename Model>
__global__ __launch_bounds__( 256 ) void k_tp_fill( dview v, unsigned long long * cursor, tp_entry * ent )
{
	using S = typename Model::synapse;
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N ) return;
	snn_info const info = info_of( v );
	std::int64_t const r0 = v.row_ptr[r];
	for( std::int64_t e = r0 + lane_id(); e < v.row_ptr[r + 1]; e += 64 )
	{
		std::int32_t const j = v.col[e];
		if( S::plastic( std::int32_t( r ), j, info ) )
		{
			unsigned long long const k = atomicAdd( &cursor[j], 1ull );
			ent[k] = tp_entry{ std::int32_t( e - r0 ), std::int32_t( r ) };
		}
	}
}
} // namespace spice::hipx
