#pragma once
// Model-templated step kernels for gfx950 (subsystems b, c, d).
//
// Per step (step index T read from device memory, see dview.h):
//   k_receive<M> : deliver the spikes emitted at T - D. Work items are (spike, part): a
//                  wave64 walks a contiguous slice of the spike's CSR row (coalesced target
//                  reads) and calls M::neuron::receive for each target through a proxy
//                  iterator that routes backend.atomic_add into per-attribute accumulators
//                  (fp32 atomics, or int64 fixed point). Plastic models first bring each
//                  plastic synapse up to date (lazy STDP replay) and then receive reads it.
//   k_update<M>  : one thread per neuron: add accumulated inputs, run M::neuron::update,
//                  store, and append spikes to the ring with one atomic per wave
//                  (ballot + popcount + mbcnt). Plastic models also shift the per-neuron
//                  spike history and queue rows whose history window is full.
//   k_flush<M>   : plastic models only: bring queued rows up to date.
// The last kernel of a step advances the device clock (last-block-done), so the same
// launches are valid at every step and can be captured once into a HIP graph.
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

namespace spice::hipx
{
constexpr int HISTORY = 64; // post-spike history window (steps) of lazy plasticity
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
	return i;
}
__device__ inline float step_dt( dview const & v ) { return v.DT - v.ctl->c; }

// Called by every thread at the end of the last kernel of a step: the last block to finish
// advances the device clock exactly as kahan_sum::add does on the host.
__device__ inline void finish_step( dview const & v )
{
	__syncthreads();
	if( threadIdx.x == 0 )
	{
		__threadfence();
		unsigned const prev = atomicAdd( &v.ctl->done, 1u );
		if( prev == gridDim.x - 1 )
		{
			__threadfence();
			ctl_t * c = v.ctl;
			float const y = v.DT - c->c;
			float const t = c->sum + y;
			c->c = ( t - c->sum ) - y;
			c->sum = t;
			c->step = c->step + 1;
			c->done = 0;
			v.stats->steps += 1;
		}
	}
}

// ---- state load/store -------------------------------------------------------------------
template <typename Decl>
__device__ inline void load_tuple( typename Decl::tuple_t & s, void * const * arr, std::int64_t i )
{
	util::for_n<Decl::size>( [&]( auto I ) {
		using T = typename Decl::template type<I>;
		std::get<I>( s ) = ld<T>( arr[I], i );
	} );
}
template <typename Decl>
__device__ inline void store_tuple( typename Decl::tuple_t const & s, void * const * arr, std::int64_t i )
{
	util::for_n<Decl::size>( [&]( auto I ) {
		using T = typename Decl::template type<I>;
		st<T>( arr[I], i, std::get<I>( s ) );
	} );
}

// ---- lazy plasticity ---------------------------------------------------------------------
// Applies, in time order, the plasticity of every step in (synced, now] to one synapse:
// the pending pre event at synced+1 (if pend) and the target's post spikes, given as
// history bits (bit b <=> target spiked at step now - b). Event-free stretches are merged
// into the next call's nsteps (the contract's n-step form), and a trailing event-free
// stretch is applied with pre = post = false. Equivalent to the CPU checker's dense
// per-step updates (reference counterpart: algorithm.cu:283-294).
template <typename Syn, typename Iter, typename Bak>
__device__ inline void replay(
    Iter syn, unsigned long long hist, std::int32_t now, std::int32_t synced, std::int32_t pend, float dt, snn_info info, Bak & bak )
{
	std::int32_t const n = now - synced;
	if( n <= 0 ) return;
	unsigned long long h = n >= 64 ? hist : ( hist & ( ( 1ull << n ) - 1ull ) );
	std::int32_t t = synced;
	if( pend )
	{
		int const b = n - 1;
		bool const post = ( h >> b ) & 1ull;
		h &= ~( 1ull << b );
		Syn::update( syn, 1, true, post, dt, info, bak );
		t = synced + 1;
	}
	while( h )
	{
		int const b = 63 - __clzll( h );
		std::int32_t const e = now - b;
		Syn::update( syn, e - t, false, true, dt, info, bak );
		t = e;
		h &= ~( 1ull << b );
	}
	if( t < now ) Syn::update( syn, now - t, false, false, dt, info, bak );
}

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
	bool anyp = false;
	for( std::int64_t e = lo + lane_id(); e < hi; e += 64 )
	{
		std::int32_t const j = v.col[e];
		typename S::tuple_t s{};
		auto bak = make_backend<false>(
		    v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), 0u, util::TAG_SYNAPSE_INIT );
		S::init( reg_iter<typename S::tuple_t>{ s, std::uint32_t( e ) }, int_( r ), j, info, bak );
		store_tuple<S>( s, v.sa, e );
		anyp |= S::plastic( int_( r ), j, info );
	}
	std::uint64_t const m = ballot( anyp );
	if( lane_id() == 0 ) v.hasp[r] = m != 0;
}

template <typename Model, bool FIXED>
__global__ __launch_bounds__( 256 ) void k_receive( dview v )
{
	using N = typename Model::neuron;
	using S = typename Model::synapse;
	constexpr bool plastic = is_plastic_v<Model>;

	std::int64_t const T = v.ctl->step;
	if( T < v.D ) return;
	int const D1 = v.D + 1;
	int const slot = int( ( T + 1 ) % D1 ); // ring entry of step T - D
	std::int64_t const nitems = std::int64_t( v.counts[slot] ) * v.parts;
	std::int32_t const * spk = v.ring + std::int64_t( slot ) * v.N;
	float const dt = step_dt( v );
	snn_info const info = info_of( v );
	int const lane = lane_id();
	std::int64_t const wave0 = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	std::int64_t const nwaves = std::int64_t( gridDim.x ) * ( blockDim.x >> 6 );

	unsigned long long ev = 0, pev = 0;
	for( std::int64_t it = wave0; it < nitems; it += nwaves )
	{
		std::int64_t const k = it / v.parts;
		std::int64_t const p = it - k * v.parts;
		std::int32_t const s = uniform( spk[k] );
		std::int64_t const lo = v.row_ptr[s], hi = v.row_ptr[s + 1];
		std::int64_t const chunk = ( ( hi - lo + v.parts - 1 ) / v.parts + 63 ) & ~std::int64_t( 63 );
		std::int64_t const beg = lo + p * chunk;
		std::int64_t const end = beg + chunk < hi ? beg + chunk : hi;
		if( beg >= end ) continue;
		ev += std::uint64_t( end - beg );

		if constexpr( plastic )
		{
			std::int32_t const synced = v.synced[s];
			std::int32_t const pend = v.pend[s];
			for( std::int64_t e = beg + lane; e < end; e += 64 )
			{
				std::int32_t const j = v.col[e];
				typename S::tuple_t sy;
				load_tuple<S>( sy, v.sa, e );
				auto bak = make_backend<FIXED>( v, std::uint32_t( s ), std::uint32_t( j ), std::uint32_t( T ), util::TAG_RECEIVE );
				if( S::plastic( s, j, info ) )
				{
					replay<S>( reg_iter<typename S::tuple_t>{ sy, std::uint32_t( e ) }, v.hist[j], std::int32_t( T - 1 ), synced, pend, dt, info, bak );
					store_tuple<S>( sy, v.sa, e );
					pev++;
				}
				N::receive( s, recv_iter<N>{ &v, j }, reg_iter<typename S::tuple_t>{ sy, std::uint32_t( e ) }, info, bak );
			}
		}
		else
		{
			for( std::int64_t e = beg + lane; e < end; e += 64 )
			{
				std::int32_t const j = v.col[e];
				auto bak = make_backend<FIXED>( v, std::uint32_t( s ), std::uint32_t( j ), std::uint32_t( T ), util::TAG_RECEIVE );
				N::receive( s, recv_iter<N>{ &v, j }, no_syn_iter{}, info, bak );
			}
		}
	}
	if( lane == 0 && ev ) atomicAdd( &v.stats->events, ev );
	if constexpr( plastic )
	{
		// plastic event counts are per lane
		unsigned long long const tot = wave_sum( pev );
		if( lane == 0 && tot ) atomicAdd( &v.stats->plastic_events, tot );
	}
}

template <typename Model, bool FIXED>
__global__ __launch_bounds__( 256 ) void k_update( dview v )
{
	using N = typename Model::neuron;
	constexpr bool plastic = is_plastic_v<Model>;
	constexpr std::uint32_t rmask = receive_mask_v<N>;

	std::int64_t const T = v.ctl->step;
	int const D1 = v.D + 1;
	int const wslot = int( T % D1 );
	if( blockIdx.x == 0 && threadIdx.x == 0 )
	{
		v.counts[( T + 1 ) % D1] = 0; // consumed by this step's receive; refilled at T + 1
		if constexpr( plastic ) v.fl_count[( T + 1 ) & 1] = 0;
	}
	float const dt = step_dt( v );
	snn_info const info = info_of( v );
	std::int64_t const i = std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
	bool spiked = false, flush = false;
	std::int32_t fl_synced = 0, fl_pend = 0;

	if( i < v.N )
	{
		typename N::tuple_t s;
		load_tuple<N>( s, v.na, i );
		util::for_n<N::size>( [&]( auto I ) {
			if constexpr( ( ( rmask >> I ) & 1u ) != 0 )
			{
				using A = typename N::template type<I>;
				if constexpr( std::is_floating_point_v<A> && FIXED )
				{
					long long const a = ld<long long>( v.acc[I], i );
					st<long long>( v.acc[I], i, 0 );
					std::get<I>( s ) = std::get<I>( s ) + float( a ) * __builtin_ldexpf( 1.0f, -v.shift );
				}
				else
				{
					A const a = ld<A>( v.acc[I], i );
					st<A>( v.acc[I], i, A( 0 ) );
					std::get<I>( s ) = A( std::get<I>( s ) + a );
				}
			}
		} );

		auto bak = make_backend<FIXED>(
		    v, std::uint32_t( i ), std::uint32_t( T ), std::uint32_t( std::uint64_t( T ) >> 32 ), util::TAG_NEURON_UPDATE );
		spiked = N::update( reg_iter<typename N::tuple_t>{ s, std::uint32_t( i ) }, dt, info, bak );
		store_tuple<N>( s, v.na, i );

		if constexpr( plastic )
		{
			unsigned long long h = v.hist[i];
			std::int32_t synced = v.synced[i], pend = v.pend[i];
			bool dirty = false;
			if( ( h >> ( v.D - 1 ) ) & 1ull ) // this neuron's spike was delivered this step
			{
				synced = std::int32_t( T - 1 );
				pend = 1;
				dirty = true;
			}
			v.hist[i] = ( h << 1 ) | ( spiked ? 1ull : 0ull );
			if( v.hasp[i] && T - synced >= FLUSH_AGE )
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

	// spike compaction: one atomic per wave
	std::uint64_t const m = ballot( spiked );
	if( m )
	{
		unsigned base = 0;
		if( lane_id() == 0 )
		{
			base = atomicAdd( &v.counts[wslot], unsigned( __popcll( m ) ) );
			atomicAdd( &v.stats->spikes, (unsigned long long)__popcll( m ) );
		}
		base = __builtin_amdgcn_readfirstlane( base );
		if( spiked ) v.ring[std::int64_t( wslot ) * v.N + base + lanes_below( m )] = std::int32_t( i );
	}
	if constexpr( plastic )
	{
		std::uint64_t const fm = ballot( flush );
		if( fm )
		{
			unsigned base = 0;
			if( lane_id() == 0 ) base = atomicAdd( &v.fl_count[T & 1], unsigned( __popcll( fm ) ) );
			base = __builtin_amdgcn_readfirstlane( base );
			if( flush )
			{
				unsigned const o = base + lanes_below( fm );
				v.fl_id[o] = std::int32_t( i );
				v.fl_synced[o] = fl_synced;
				v.fl_pend[o] = fl_pend;
			}
		}
	}
	else
		finish_step( v );
}

template <typename Model, bool FIXED>
__global__ __launch_bounds__( 256 ) void k_flush( dview v )
{
	using S = typename Model::synapse;
	std::int64_t const T = v.ctl->step;
	std::int64_t const nitems = std::int64_t( v.fl_count[T & 1] ) * v.fparts;
	float const dt = step_dt( v );
	snn_info const info = info_of( v );
	int const lane = lane_id();
	std::int64_t const wave0 = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	std::int64_t const nwaves = std::int64_t( gridDim.x ) * ( blockDim.x >> 6 );
	unsigned long long touched = 0;
	for( std::int64_t it = wave0; it < nitems; it += nwaves )
	{
		std::int64_t const f = it / v.fparts;
		std::int64_t const p = it - f * v.fparts;
		std::int32_t const s = uniform( v.fl_id[f] );
		std::int32_t const synced = uniform( v.fl_synced[f] );
		std::int32_t const pend = uniform( v.fl_pend[f] );
		std::int64_t const lo = v.row_ptr[s], hi = v.row_ptr[s + 1];
		std::int64_t const chunk = ( ( hi - lo + v.fparts - 1 ) / v.fparts + 63 ) & ~std::int64_t( 63 );
		std::int64_t const beg = lo + p * chunk;
		std::int64_t const end = beg + chunk < hi ? beg + chunk : hi;
		for( std::int64_t e = beg + lane; e < end; e += 64 )
		{
			std::int32_t const j = v.col[e];
			if( !S::plastic( s, j, info ) ) continue;
			typename S::tuple_t sy;
			load_tuple<S>( sy, v.sa, e );
			auto bak = make_backend<FIXED>(
			    v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), std::uint32_t( T ), util::TAG_SYNAPSE_UPDATE );
			replay<S>( reg_iter<typename S::tuple_t>{ sy, std::uint32_t( e ) }, v.hist[j], std::int32_t( T ), synced, pend, dt, info, bak );
			store_tuple<S>( sy, v.sa, e );
			touched++;
		}
		if( lane == 0 && p == 0 ) atomicAdd( &v.stats->flush_rows, 1ull );
	}
	unsigned long long const tot = wave_sum( touched );
	if( lane == 0 && tot ) atomicAdd( &v.stats->flush_synapses, tot );
	finish_step( v );
}

// Brings every row with plastic synapses up to date through step `now` (= last completed
// step); one wave per row. Used before exporting synapse state.
template <typename Model>
__global__ __launch_bounds__( 256 ) void k_sync_all( dview v, std::int32_t now, float dt )
{
	using S = typename Model::synapse;
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= v.N || !v.hasp[r] ) return;
	snn_info const info = info_of( v );
	std::int32_t const synced = v.synced[r], pend = v.pend[r];
	for( std::int64_t e = v.row_ptr[r] + lane_id(); e < v.row_ptr[r + 1]; e += 64 )
	{
		std::int32_t const j = v.col[e];
		if( !S::plastic( std::int32_t( r ), j, info ) ) continue;
		typename S::tuple_t sy;
		load_tuple<S>( sy, v.sa, e );
		auto bak = make_backend<false>( v, std::uint32_t( e ), std::uint32_t( std::uint64_t( e ) >> 32 ), std::uint32_t( now ), util::TAG_SYNAPSE_UPDATE );
		replay<S>( reg_iter<typename S::tuple_t>{ sy, std::uint32_t( e ) }, v.hist[j], now, synced, pend, dt, info, bak );
		store_tuple<S>( sy, v.sa, e );
	}
	if( lane_id() == 0 )
	{
		v.synced[r] = now;
		v.pend[r] = 0;
	}
}
} // namespace spice::hipx
