#pragma once
// TEST INFRASTRUCTURE ONLY (the CPU checker). Not part of the product; nothing under
// spice_amd/ or include/ may use it.
//
// Single-threaded CPU restatement of the reference simulator cpu::snn
// (spice/cpu/snn.h:17-49, spice/cpu/snn.h.cpp:73-181, spice/snn.cpp:38-44):
//   step i, dt = kahan.add(DT):
//     1. if i >= delay: for every spike s of step i - delay (ascending id), for every
//        synapse of s (ascending target): Model::neuron::receive(s, dst, syn)   (fp32 adds)
//     2. for every neuron n (ascending): spiked = Model::neuron::update(n, dt); append
//     3. plastic models: for every synapse (src -> dst) with plastic(src, dst):
//        Model::synapse::update(syn, 1, pre = spiked(src, i - delay), post = spiked(dst, i), dt)
// Differences from the reference, by design of the new build:
//   * randomness comes from Philox keyed on (seed, neuron, step) (and analogous keys for
//     init / receive), so the GPU draws exactly the same numbers;
//   * the adjacency is CSR (ELL export keeps the reference's row-width padding);
//   * optional fixed-point mode: receive() contributions are summed per target attribute
//     as int64 in units of 2^-shift and added to the state before update(); the GPU's
//     fixed-point mode is defined identically, so it is bit-exact against this.
#include <spice/models/model.h>
#include <spice/util/meta.h>
#include <spice/util/numeric.h>
#include <spice/util/philox.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <tuple>
#include <type_traits>
#include <vector>

namespace spice::oracle
{
// Sequential xoshiro128+ stream seeded through the SplitMix64 finalizer: the random source of
// the reference's CPU backend (spice/cpu/backend.h:23,52; spice/util/random.h:18-73), written
// from the published algorithms (D. Blackman, S. Vigna, "Scrambled linear pseudorandom number
// generators", 2018; SplitMix64 mixing function). Used only in the checker's reference-RNG
// mode, which reproduces the reference cpu::snn trajectories of its Poisson models exactly,
// so the checker can be pinned against the reference build.
struct xoshiro128p
{
	std::uint32_t s[4];
	static std::uint64_t mix( std::uint64_t x )
	{
		x = ( x ^ ( x >> 30 ) ) * 0xbf58476d1ce4e5b9ull;
		x = ( x ^ ( x >> 27 ) ) * 0x94d049bb133111ebull;
		return x ^ ( x >> 31 );
	}
	explicit xoshiro128p( std::uint64_t seed = 1 )
	{
		std::uint64_t h = mix( seed );
		s[0] = std::uint32_t( h );
		s[1] = std::uint32_t( h >> 32 );
		h = mix( h );
		s[2] = std::uint32_t( h );
		s[3] = std::uint32_t( h >> 32 );
	}
	std::uint32_t operator()()
	{
		std::uint32_t const r = s[0] + s[3];
		std::uint32_t const t = s[1] << 9;
		s[2] ^= s[0];
		s[3] ^= s[1];
		s[1] ^= s[2];
		s[0] ^= s[3];
		s[2] ^= t;
		s[3] = ( s[3] << 11 ) | ( s[3] >> 21 );
		return r;
	}
};

// Receive-target proxy for fixed-point mode: reads see the state, atomic_add accumulates.
template <typename T, int I>
struct fix_proxy
{
	T const * state;
	std::int64_t * acc;
	operator T() const { return *state; }
};

struct cpu_backend
{
	util::philox_stream rng;
	int shift = 40;
	xoshiro128p * seq = nullptr; // reference-RNG mode: one sequential stream per net

	float rand() { return seq ? util::u01_left( ( *seq )() ) : rng.rand(); }

	template <typename T>
	static void atomic_add( T & var, typename util::identity<T>::type val )
	{
		var += val;
	}
	template <typename T, int I>
	void atomic_add( fix_proxy<T, I> p, typename util::identity<T>::type val ) const
	{
		if constexpr( std::is_floating_point_v<T> )
			*p.acc += std::llrint( double( val ) * std::ldexp( 1.0, shift ) );
		else
			*p.acc += std::int64_t( val );
	}

	template <typename T>
	static T min( T x, T hi )
	{
		return std::min( x, hi );
	}
	template <typename T>
	static T max( T lo, T x )
	{
		return std::max( lo, x );
	}
	template <typename T>
	static T clamp( T x, T lo, T hi )
	{
		return min( hi, max( lo, x ) );
	}
	template <typename T>
	static T exp( T x )
	{
		return std::exp( x );
	}
	static float pow( float x, float y ) { return std::pow( x, y ); }
	static float fma( float a, float b, float c ) { return std::fma( a, b, c ); }
};

template <typename Tuple>
struct elem_iter
{
	Tuple * data;
	std::int64_t i;
	template <int I>
	auto & get() const
	{
		return std::get<I>( data[i] );
	}
	uint_ id() const { return uint_( i ); }
};

template <typename Tuple>
struct fix_iter
{
	Tuple const * data;
	std::vector<std::int64_t> * acc; // one array per attribute
	std::int64_t i;
	template <int I>
	fix_proxy<std::tuple_element_t<I, Tuple>, I> get() const
	{
		return { &std::get<I>( data[i] ), &acc[I][i] };
	}
	uint_ id() const { return uint_( i ); }
};

// Model-independent interface so one C ABI can drive every model.
struct net_base
{
	virtual ~net_base() = default;
	virtual void step( std::vector<int> * out ) = 0;
	virtual std::int64_t num_neurons() const = 0;
	virtual std::int64_t width() const = 0;
	virtual std::vector<std::int64_t> const & row_ptr() const = 0;
	virtual std::vector<std::int32_t> const & col() const = 0;
	virtual int num_neuron_attrs() const = 0;
	virtual int num_synapse_attrs() const = 0;
	virtual size_t neuron_attr_size( int a ) const = 0;
	virtual size_t synapse_attr_size( int a ) const = 0;
	virtual void export_neuron_attr( int a, void * out ) const = 0;
	virtual void import_neuron_attr( int a, void const * in ) = 0;
	virtual void export_synapse_attr( int a, void * out ) const = 0; // CSR order
	virtual void import_synapse_attr( int a, void const * in ) = 0;
	virtual void get_clock( std::int64_t * i, float * sum, float * c ) const = 0;
	virtual bool export_traces( float * x, float * y ) const = 0;
	virtual void set_params( float const * p ) = 0;
	virtual void set_clock( std::int64_t i, float sum, float c ) = 0;
	virtual void set_in_flight( std::vector<std::int64_t> const & counts, std::vector<int> const & ids ) = 0;
};

template <typename Model>
class cpu_snn : public net_base
{
	using N = typename Model::neuron;
	using S = typename Model::synapse;
	using ntuple = typename N::tuple_t;
	using stuple = typename S::tuple_t;

public:
	// Graph given as CSR with the ELL row width it was built for.
	cpu_snn(
	    std::vector<std::int64_t> row_ptr,
	    std::vector<std::int32_t> col,
	    std::int64_t width,
	    float dt,
	    int delay,
	    std::uint64_t seed,
	    bool fixed,
	    int shift,
	    bool ref_rng = false )
	    : _row_ptr( std::move( row_ptr ) )
	    , _col( std::move( col ) )
	    , _n( std::int64_t( _row_ptr.size() ) - 1 )
	    , _width( width )
	    , _dt( dt )
	    , _delay( delay )
	    , _seed( seed )
	    , _fixed( fixed )
	    , _shift( shift )
	    , _ref_rng( ref_rng )
	    , _seq( seed )
	{
		if( !( dt > 0.0f ) ) throw std::invalid_argument( "dt must be positive" );
		if( delay < 1 ) throw std::invalid_argument( "delay must be >= 1" );
		_info.num_neurons = int_( _n );
		_neurons.resize( _n );
		for( std::int64_t i = 0; i < _n; i++ )
		{
			cpu_backend bak{ { _seed, std::uint32_t( i ), 0u, 0u, util::TAG_NEURON_INIT }, _shift, seq() };
			N::init( elem_iter<ntuple>{ _neurons.data(), i }, _info, bak );
		}
		if constexpr( S::size > 0 )
		{
			_synapses.resize( _col.size() );
			for( std::int64_t s = 0; s < _n; s++ )
				for( std::int64_t k = _row_ptr[s]; k < _row_ptr[s + 1]; k++ )
				{
					cpu_backend bak{
					    { _seed, std::uint32_t( k ), std::uint32_t( std::uint64_t( k ) >> 32 ), 0u, util::TAG_SYNAPSE_INIT },
					    _shift,
					    seq() };
					S::init( elem_iter<stuple>{ _synapses.data(), k }, int_( s ), _col[k], _info, bak );
				}
			_flags.assign( delay + 1, std::vector<char>( _n, 0 ) );
			if constexpr( has_neuron_traces_v<S> )
			{
				_x.assign( _n, 0.0f );
				_y.assign( _n, 0.0f );
			}
		}
		for( auto & a : _acc ) a.assign( _fixed ? _n : 0, 0 );
	}

	void step( std::vector<int> * out ) override
	{
		std::int64_t const i = _i++;
		float const dt = _clock.add( _dt );
		std::int64_t const D = _delay;

		// 1. receive the spikes emitted at step i - delay
		if( _counts.size() >= size_t( D ) )
		{
			std::int64_t const cnt = _counts.front();
			for( std::int64_t x = 0; x < cnt; x++ )
			{
				int_ const s = _ids[x];
				for( std::int64_t k = _row_ptr[s]; k < _row_ptr[s + 1]; k++ )
				{
					int_ const dst = _col[k];
					cpu_backend bak{ { _seed, std::uint32_t( s ), std::uint32_t( dst ), std::uint32_t( i ), util::TAG_RECEIVE }, _shift, seq() };
					elem_iter<stuple const> syn{ _synapses.data(), k };
					if( _fixed )
						N::receive( s, fix_iter<ntuple>{ _neurons.data(), _acc, dst }, syn, _info, bak );
					else
						N::receive( s, elem_iter<ntuple>{ _neurons.data(), dst }, syn, _info, bak );
				}
			}
			_ids.erase( _ids.begin(), _ids.begin() + cnt );
			_counts.erase( _counts.begin() );
		}

		// 2. update neurons (fixed-point inputs are applied first)
		std::int64_t const before = std::int64_t( _ids.size() );
		for( std::int64_t n = 0; n < _n; n++ )
		{
			if( _fixed ) apply_fixed( n );
			cpu_backend bak{
			    { _seed, std::uint32_t( n ), std::uint32_t( i ), std::uint32_t( std::uint64_t( i ) >> 32 ), util::TAG_NEURON_UPDATE },
			    _shift,
			    seq() };
			bool const spiked = N::update( elem_iter<ntuple>{ _neurons.data(), n }, dt, _info, bak );
			if constexpr( S::size > 0 ) _flags[i % ( D + 1 )][n] = spiked;
			if( spiked ) _ids.push_back( int( n ) );
		}
		_counts.push_back( std::int64_t( _ids.size() ) - before );
		if( out ) out->assign( _ids.begin() + before, _ids.end() );

		// 3. dense plasticity
		if constexpr( S::size > 0 && has_neuron_traces_v<S> )
		{
			// neuron-local traces, then every plastic synapse with this step's flags
			auto const & pre = _flags[( i + 1 ) % ( D + 1 )];
			auto const & post = _flags[i % ( D + 1 )];
			float const d = S::trace_decay( dt );
			for( std::int64_t n = 0; n < _n; n++ )
			{
				_x[n] = std::fma( _x[n], d, pre[n] ? 1.0f : 0.0f );
				_y[n] = std::fma( _y[n], d, post[n] ? 1.0f : 0.0f );
			}
			for( std::int64_t s = 0; s < _n; s++ )
				for( std::int64_t k = _row_ptr[s]; k < _row_ptr[s + 1]; k++ )
				{
					int_ const dst = _col[k];
					if( !S::plastic( int_( s ), dst, _info ) ) continue;
					cpu_backend bak{
					    { _seed, std::uint32_t( k ), std::uint32_t( std::uint64_t( k ) >> 32 ), std::uint32_t( i ), util::TAG_SYNAPSE_UPDATE },
					    _shift,
					    seq() };
					S::on_event( elem_iter<stuple>{ _synapses.data(), k }, pre[s] != 0, post[dst] != 0, _x[s], _y[dst], _info, bak );
				}
		}
		else if constexpr( S::size > 0 )
		{
			auto const & pre = _flags[( i + 1 ) % ( D + 1 )];
			auto const & post = _flags[i % ( D + 1 )];
			for( std::int64_t s = 0; s < _n; s++ )
				for( std::int64_t k = _row_ptr[s]; k < _row_ptr[s + 1]; k++ )
				{
					int_ const dst = _col[k];
					if( !S::plastic( int_( s ), dst, _info ) ) continue;
					cpu_backend bak{
					    { _seed, std::uint32_t( k ), std::uint32_t( std::uint64_t( k ) >> 32 ), std::uint32_t( i ), util::TAG_SYNAPSE_UPDATE },
					    _shift,
					    seq() };
					S::update( elem_iter<stuple>{ _synapses.data(), k }, 1, pre[s] != 0, post[dst] != 0, dt, _info, bak );
				}
		}
	}

	std::int64_t num_neurons() const override { return _n; }
	std::int64_t width() const override { return _width; }
	std::vector<std::int64_t> const & row_ptr() const override { return _row_ptr; }
	std::vector<std::int32_t> const & col() const override { return _col; }
	int num_neuron_attrs() const override { return int( N::size ); }
	int num_synapse_attrs() const override { return int( S::size ); }
	size_t neuron_attr_size( int a ) const override { return attr_size<ntuple>( a ); }
	size_t synapse_attr_size( int a ) const override { return attr_size<stuple>( a ); }

	void export_neuron_attr( int a, void * out ) const override { copy_out( _neurons, a, out ); }
	void import_neuron_attr( int a, void const * in ) override { copy_in( _neurons, a, in ); }
	void export_synapse_attr( int a, void * out ) const override { copy_out( _synapses, a, out ); }
	void import_synapse_attr( int a, void const * in ) override { copy_in( _synapses, a, in ); }

	void set_params( float const * p ) override
	{
		for( int k = 0; k < 4; k++ ) _info.param[k] = p[k];
	}

	bool export_traces( float * x, float * y ) const override
	{
		if( _x.empty() ) return false;
		std::copy( _x.begin(), _x.end(), x );
		std::copy( _y.begin(), _y.end(), y );
		return true;
	}

	void get_clock( std::int64_t * i, float * sum, float * c ) const override
	{
		*i = _i;
		*sum = float( _clock );
		*c = _clock.compensation();
	}
	void set_clock( std::int64_t i, float sum, float c ) override
	{
		_i = i;
		_clock.set( sum, c );
	}

	// The spikes of the last counts.size() (<= delay) steps, oldest first: what the next
	// steps deliver (and, for plastic models, their pre flags).
	void set_in_flight( std::vector<std::int64_t> const & counts, std::vector<int> const & ids ) override
	{
		std::int64_t const k = std::int64_t( counts.size() );
		if( k > _delay ) throw std::invalid_argument( "at most `delay` steps are in flight" );
		for( int x : ids )
			if( x < 0 || x >= _n ) throw std::invalid_argument( "spike id out of range" );
		_counts = counts;
		_ids = ids;
		if constexpr( S::size > 0 )
		{
			std::size_t o = 0;
			for( std::int64_t s = 0; s < k; s++ )
			{
				std::int64_t const step = _i - k + s;
				auto & f = _flags[std::size_t( ( step % ( _delay + 1 ) + _delay + 1 ) % ( _delay + 1 ) )];
				std::fill( f.begin(), f.end(), 0 );
				for( std::int64_t q = 0; q < counts[s]; q++ ) f[ids[o + q]] = 1;
				o += std::size_t( counts[s] );
			}
		}
	}

private:
	std::vector<std::int64_t> _row_ptr;
	std::vector<std::int32_t> _col;
	std::int64_t _n, _width;
	float _dt;
	int _delay;
	std::uint64_t _seed;
	bool _fixed;
	int _shift;
	bool _ref_rng;
	xoshiro128p _seq;
	snn_info _info;

	xoshiro128p * seq() { return _ref_rng ? &_seq : nullptr; }

	std::vector<ntuple> _neurons;
	std::vector<stuple> _synapses;
	std::vector<int> _ids;
	std::vector<std::int64_t> _counts;
	std::vector<std::vector<char>> _flags;
	std::vector<float> _x, _y; // neuron-local pre/post traces (trace synapses)
	std::vector<std::int64_t> _acc[N::size > 0 ? N::size : 1];

	std::int64_t _i = 0;
	util::kahan_sum<float> _clock;

	void apply_fixed( std::int64_t n )
	{
		util::for_n<N::size>( [&]( auto I ) {
			if constexpr( ( ( receive_mask_v<N> >> I ) & 1u ) != 0 )
			{
				auto & x = std::get<I>( _neurons[n] );
				using T = std::decay_t<decltype( x )>;
				std::int64_t const a = _acc[I][n];
				_acc[I][n] = 0;
				if constexpr( std::is_floating_point_v<T> )
					x = x + float( a ) * std::ldexp( 1.0f, -_shift );
				else
					x = T( x + a );
			}
		} );
	}

	template <typename Tuple>
	static size_t attr_size( int a )
	{
		size_t r = 0;
		util::for_n<std::tuple_size_v<Tuple>>( [&]( auto I ) {
			if( int( I ) == a ) r = sizeof( std::tuple_element_t<I, Tuple> );
		} );
		return r;
	}
	template <typename Tuple>
	static void copy_out( std::vector<Tuple> const & v, int a, void * out )
	{
		util::for_n<std::tuple_size_v<Tuple>>( [&]( auto I ) {
			if( int( I ) != a ) return;
			auto * o = static_cast<std::tuple_element_t<I, Tuple> *>( out );
			for( size_t i = 0; i < v.size(); i++ ) o[i] = std::get<I>( v[i] );
		} );
	}
	template <typename Tuple>
	static void copy_in( std::vector<Tuple> & v, int a, void const * in )
	{
		util::for_n<std::tuple_size_v<Tuple>>( [&]( auto I ) {
			if( int( I ) != a ) return;
			auto const * p = static_cast<std::tuple_element_t<I, Tuple> const *>( in );
			for( size_t i = 0; i < v.size(); i++ ) std::get<I>( v[i] ) = p[i];
		} );
	}
};
} // namespace spice::oracle
