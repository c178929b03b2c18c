#pragma once
// The model plug-in contract.
//
// A model is a struct with a nested `neuron` type and, optionally, a nested `synapse`
// type. Each derives from neuron<Ts...> / synapse<Ts...>, which declares the per-element
// state attributes, and overrides any of these static functors (same contract as the
// reference, spice/models/model.h:10-58, so reference models compile unchanged):
//
//   neuron::init(it, info, backend)                    once per neuron
//   neuron::update(it, dt, info, backend) -> bool      once per neuron per step; true = spike
//   neuron::receive(src, dst_it, syn_it, info, backend) once per delivered synaptic event
//   synapse::init(syn_it, src, dst, info, backend)     once per synapse
//   synapse::update(syn_it, nsteps, pre, post, dt, info, backend)
//                                                      plasticity over nsteps steps; the
//                                                      last step carries the pre/post flags
//   synapse::plastic(src, dst, info) -> bool           which synapses update() applies to
//
// The backend offers rand() in [0, 1), atomic_add(target, value), min, max, clamp, exp, pow.
//
// Contract points that a parallel implementation fixes (and the CPU checker follows):
//  * receive() may only modify neuron state through backend.atomic_add, and reads of the
//    target neuron see its state as of the end of the previous step (inputs delivered in
//    the same step are not visible to each other);
//  * synapse::update(n, pre, post) must equal n-1 event-free steps followed by one step
//    with the given flags (this is what lets plasticity be applied lazily).
//
// Optional traits: a neuron type may declare
//     static constexpr std::uint32_t receive_mask = (1u << A) | (1u << B);
// naming the attributes receive() adds to (without it every attribute gets an input
// accumulator), and
//     static constexpr std::uint32_t receive_reads = 1u << C;
// naming the attributes receive() reads from its target (without it, none are assumed: reads
// then go to device memory for every delivered event; with it, the delivering workgroup keeps
// a copy of its tile's values of those attributes in LDS). Neither changes results.
#include <spice/snn_info.h>
#include <spice/util/host_defines.h>
#include <spice/util/meta.h>

#include <cstdint>
#include <type_traits>

namespace spice
{
template <typename... Ts>
struct neuron : util::type_list<Ts...>
{
	template <typename Iter, typename Backend>
	HYBRID static void init( Iter, snn_info, Backend & )
	{
	}

	template <typename Iter, typename Backend>
	HYBRID static bool update( Iter, float, snn_info, Backend & )
	{
		return false;
	}

	template <typename Iter, typename SynIter, typename Backend>
	HYBRID static void receive( int_, Iter, SynIter &&, snn_info, Backend & )
	{
	}
};

template <typename... Ts>
struct synapse : util::type_list<Ts...>
{
	template <typename Iter, typename Backend>
	HYBRID static void init( Iter, int_, int_, snn_info, Backend & )
	{
	}

	template <typename Iter, typename Backend>
	HYBRID static void update( Iter &&, int_, bool, bool, float, snn_info, Backend & )
	{
	}

	HYBRID static bool plastic( int_, int_, snn_info ) { return true; }
};

// Synapse whose plasticity is driven by two neuron-local traces (pair-based STDP and its
// relatives). The framework keeps, per neuron,
//   x(t) = fma(x(t-1), trace_decay(dt), arrived(t))   pre trace: a spike of this neuron
//                                                      reaches its synapses at step t
//   y(t) = fma(y(t-1), trace_decay(dt), spiked(t))    post trace
// (both start at 0), and every step, for every plastic synapse src -> dst, calls
//   on_event(syn, pre = arrived_src(t), post = spiked_dst(t), x_src(t), y_dst(t), info, backend)
// The synapse state holds only what the rule needs per synapse (typically the weight).
// on_event must leave the synapse unchanged when pre and post are both false; that is what
// lets the GPU apply it only at events (eagerly, in or right after the event's step) with
// results identical to the dense per-step definition. It reads x only when post is set and y
// only when pre is set (the GPU passes 0 for a trace the event does not read, so it is never
// loaded). The synapse::update of the generic contract is not used.
template <typename... Ts>
struct trace_synapse : synapse<Ts...>
{
	static constexpr bool neuron_traces = true;

	HYBRID static float trace_decay( float dt ) { return 1.0f - dt * ( 1.0f / 0.02f ); }

	template <typename Iter, typename Backend>
	HYBRID static void on_event( Iter &&, bool, bool, float, float, snn_info, Backend & )
	{
	}
};

namespace detail
{
template <typename S, typename = void>
struct has_neuron_traces : std::false_type
{
};
template <typename S>
struct has_neuron_traces<S, std::void_t<decltype( S::neuron_traces )>> : std::bool_constant<S::neuron_traces>
{
};
} // namespace detail

// True for synapse types derived from trace_synapse<...>.
template <typename Synapse>
inline constexpr bool has_neuron_traces_v = detail::has_neuron_traces<Synapse>::value;

struct model
{
	struct neuron : ::spice::neuron<>
	{
	};
	struct synapse : ::spice::synapse<>
	{
	};
};

namespace detail
{
template <typename N, typename = void>
struct receive_mask_of
{
	static constexpr std::uint32_t value = N::size >= 32 ? 0xffffffffu : ( ( 1u << N::size ) - 1u );
};
template <typename N>
struct receive_mask_of<N, std::void_t<decltype( N::receive_mask )>>
{
	static constexpr std::uint32_t value = N::receive_mask;
};
} // namespace detail

// Bit I set <=> neuron attribute I can be the target of receive()'s atomic_add.
template <typename Neuron>
inline constexpr std::uint32_t receive_mask_v = detail::receive_mask_of<Neuron>::value;

namespace detail
{
template <typename N, typename = void>
struct receive_reads_of
{
	static constexpr std::uint32_t value = 0;
};
template <typename N>
struct receive_reads_of<N, std::void_t<decltype( N::receive_reads )>>
{
	static constexpr std::uint32_t value = N::receive_reads;
};
} // namespace detail

// Bit I set <=> receive() reads neuron attribute I of its target.
template <typename Neuron>
inline constexpr std::uint32_t receive_reads_v = detail::receive_reads_of<Neuron>::value;

template <typename Model>
inline constexpr bool is_plastic_v = Model::synapse::size > 0;
} // namespace spice
