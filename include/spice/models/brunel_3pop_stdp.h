#pragma once
// brunel_3pop with additive pair-based STDP on excitatory -> excitatory synapses.
//
// Neuron-local pre/post traces (tau = 20 ms; decay (1 - dt/tau) per step, +1 per spike
// arrival / emission, see trace_synapse in model.h) drive the weight of each plastic synapse:
//   W <- clamp(W + post * A+ wmax x_pre - pre * A- wmax y_post, 0, wmax),
// A+ = 0.01, A- = 0.0105, wmax = 3e-4 (in the 1e-4 = "standard excitatory weight" units of
// the reference's plastic model; receive() adds W * 5e4 / N volts, see brunel_3pop.h).
// W starts at 1e-4 for Poisson/excitatory sources and -5e-4 for inhibitory ones (those stay
// fixed: only E -> E synapses are plastic). Per synapse only W is stored.
#include <spice/models/brunel_3pop.h>

namespace spice
{
struct brunel_3pop_stdp : model
{
	enum syn_attr
	{
		W
	};

	struct neuron : brunel_3pop::neuron
	{
		// both halves of the split receive() are redefined (the base's deliver() has fixed weights)
		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void deliver( int_, Iter dst, SynIter && syn, snn_info info, Backend & bak )
		{
			using util::get;
			bak.atomic_add( get<V>( dst ), get<W>( syn ) * 50'000 / info.num_neurons );
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_ src, Iter dst, SynIter && syn, snn_info info, Backend & bak )
		{
			if( accepts( dst ) ) deliver( src, dst, syn, info, bak );
		}
	};

	struct synapse : ::spice::trace_synapse<float>
	{
		// plastic synapses join excitatory LIF neurons: after a spike at t, Twait = 20 closes the
		// gate for the inputs of t + 1 .. t + 20 and the next spike comes at t + 20 at the earliest
		static constexpr int refractory_steps = 19;

		template <typename Iter, typename Backend>
		HYBRID static void init( Iter syn, int_ src, int_, snn_info info, Backend & )
		{
			util::get<W>( syn ) = src < brunel_3pop::num_exc_end( info ) ? 0.0001f : -0.0005f;
		}

		HYBRID static float trace_decay( float dt ) { return 1.0f - dt * ( 1.0f / 0.02f ); }

		template <typename Iter, typename Backend>
		HYBRID static void on_event( Iter && syn, bool const pre, bool const post, float const x, float const y, snn_info, Backend & bak )
		{
			using util::get;
			float const Aplus = 0.01f * 0.0003f;
			float const Aminus = 0.0105f * 0.0003f;
			float const wmax = 0.0003f;
			float const w = get<W>( syn );
			get<W>( syn ) = bak.clamp( w + post * Aplus * x - pre * Aminus * y, 0.0f, wmax );
		}

		HYBRID static bool plastic( int_ const src, int_ const dst, snn_info const info )
		{
			int_ const np = brunel_3pop::num_poisson( info );
			int_ const ne = brunel_3pop::num_exc_end( info );
			return src >= np && src < ne && dst >= np && dst < ne;
		}
	};
};

// The same model without the refractory promise: the GPU cannot rely on a plastic synapse's
// target ignoring the input of a delivery while that synapse's previous event is still being
// applied, so each delivery applies the previous step's events of the rows it reads itself
// (deferred events), and the update of a step runs after its delivery. Identical results; kept
// so that both schedules stay exercised.
struct brunel_3pop_stdp_deferred : brunel_3pop_stdp
{
	struct synapse : brunel_3pop_stdp::synapse
	{
		static constexpr int refractory_steps = 0;
	};
};
} // namespace spice
