#pragma once
// Compile-time attribute lists for neuron and synapse state.
//
// A model declares its per-neuron / per-synapse state as a list of attribute types, e.g.
// neuron<float, int_> for (V, Twait). The framework stores each attribute in its own array
// (structure of arrays in HBM) and hands model functors an iterator whose get<I>() yields
// attribute I of one element. Same contract as the reference's type_list / get<I>
// (spice/util/meta.h:124-171); the implementation here is independent.
#include <spice/util/host_defines.h>
#include <spice/util/stdint.h>

#include <tuple>
#include <type_traits>
#include <utility>

namespace spice::util
{
template <typename... Ts>
struct type_list
{
	using tuple_t = std::tuple<Ts...>;
	using ptuple_t = std::tuple<Ts *...>;
	using cptuple_t = std::tuple<Ts const *...>;

	static constexpr size_ size = sizeof...( Ts );
	static constexpr size_ size_in_bytes = ( size_( 0 ) + ... + sizeof( Ts ) );

	template <size_ I>
	using type = std::tuple_element_t<I, tuple_t>;

	template <size_ I>
	static constexpr size_ ith_size_in_bytes()
	{
		return sizeof( type<I> );
	}
};

// Access attribute I through any iterator that provides get<I>(). Returns whatever the
// iterator returns: a reference for plain state, a proxy object for receive targets.
template <int_ I, typename Iter>
HYBRID decltype( auto ) get( Iter && it )
{
	return std::forward<Iter>( it ).template get<I>();
}

// Calls f(std::integral_constant<size_, I>) for I = 0 .. N-1 (unrolled at compile time).
template <size_ N, typename Fn, size_ I = 0>
HYBRID inline void for_n( Fn && f )
{
	if constexpr( I < N )
	{
		f( std::integral_constant<size_, I>() );
		for_n<N, Fn, I + 1>( std::forward<Fn>( f ) );
	}
}

template <typename T>
struct identity
{
	using type = T;
};
} // namespace spice::util
