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
//
// Optional split of receive() into a gate and an ungated input: a neuron type may declare
//     static constexpr bool gated_receive = true;
//     template <typename Iter> static bool accepts( Iter n );   reads only receive_reads
//     static void deliver( src, dst_it, syn_it, info, backend ); reads nothing of the target
// with receive() == (accepts(dst) ? deliver() : nothing). The GPU may then add a step's input
// before the target's previous update has finished and leave the gate to the neuron update,
// which sees the state receive() would have seen: with a delay of at least 2 steps the
// delivery of step T + 1 runs beside the update of step T. A neuron declaring
//     static constexpr std::uint32_t receive_reads = 0;
// (receive() reads nothing of its target) allows the same without a gate.
//
// Optional trait of a neuron-trace synapse (trace_synapse below):
//     static constexpr int refractory_steps = R;
// promises that a neuron that is an endpoint of a plastic synapse and spikes at step t neither
// spikes nor accepts input at steps t + 1 .. t + R. With R >= 2 a plastic synapse's events of
// the two steps before an arrival of its source's spike concern a target that does not take
// that spike's input (a post event), or cannot happen (a pre event), so the delivery may read
// the weight while those events are still being applied; the GPU then applies every event in
// the plasticity passes of its step and none in the delivery.
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

namespace detail
{
template <typename N, typename = void>
struct declares_reads : std::false_type
{
};
template <typename N>
struct declares_reads<N, std::void_t<decltype( N::receive_reads )>> : std::true_type
{
};
template <typename N, typename = void>
struct gated_receive_of : std::false_type
{
};
template <typename N>
struct gated_receive_of<N, std::void_t<decltype( N::gated_receive )>> : std::bool_constant<N::gated_receive>
{
};
template <typename S, typename = void>
struct refractory_steps_of
{
	static constexpr int value = 0;
};
template <typename S>
struct refractory_steps_of<S, std::void_t<decltype( S::refractory_steps )>>
{
	static constexpr int value = S::refractory_steps;
};
} // namespace detail

// The neuron splits receive() into accepts() and deliver().
template <typename Neuron>
inline constexpr bool gated_receive_v = detail::gated_receive_of<Neuron>::value;

// Delivery may run before the target's previous update is done: receive() is split into a gate
// and an ungated input, or is declared to read nothing of its target.
template <typename Neuron>
inline constexpr bool input_ahead_of_update_v =
    gated_receive_v<Neuron> || ( detail::declares_reads<Neuron>::value && receive_reads_v<Neuron> == 0 );

// Refractory steps promised for the endpoints of plastic neuron-trace synapses (0: none).
template <typename Synapse>
inline constexpr int refractory_steps_v = detail::refractory_steps_of<Synapse>::value;

template <typename Model>
inline constexpr bool is_plastic_v = Model::synapse::size > 0;
} // namespace spice
