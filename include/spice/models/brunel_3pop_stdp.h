#pragma once
// brunel_3pop with additive pair-based STDP on excitatory -> excitatory synapses.
//
// Each synapse carries its weight W and pre/post traces. Over each step the traces decay by
// (1 - dt / 20 ms) and jump by 1 on a pre (spike arrival) / post (target spike) event; then
//   W <- clamp(W + post * A+ wmax Zpre - pre * A- wmax Zpost, 0, wmax),
// A+ = 0.01, A- = 0.0105, wmax = 3e-4 (in the 1e-4 = "standard excitatory weight" units of
// the reference's plastic model; receive() adds W * 5e4 / N volts, see brunel_3pop.h). Traces start at 0;
// W starts at 1e-4 for Poisson/excitatory sources and -5e-4 for inhibitory ones (those stay
// fixed: only E -> E synapses are plastic).
#include <spice/models/brunel_3pop.h>

namespace spice
{
struct brunel_3pop_stdp : model
{
	enum syn_attr
	{
		W,
		Zpre,
		Zpost
	};

	struct neuron : brunel_3pop::neuron
	{
		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_, Iter dst, SynIter && syn, snn_info info, Backend & bak )
		{
			using util::get;
			if( get<Twait>( dst ) <= 0 ) bak.atomic_add( get<V>( dst ), get<W>( syn ) * 50'000 / info.num_neurons );
		}
	};

	struct synapse : ::spice::synapse<float, float, float>
	{
		template <typename Iter, typename Backend>
		HYBRID static void init( Iter syn, int_ src, int_, snn_info info, Backend & )
		{
			using util::get;
			get<W>( syn ) = src < brunel_3pop::num_exc_end( info ) ? 0.0001f : -0.0005f;
			get<Zpre>( syn ) = 0.0f;
			get<Zpost>( syn ) = 0.0f;
		}

		template <typename Iter, typename Backend>
		HYBRID static void update(
		    Iter && syn, int_ const nsteps, bool const pre, bool const post, float const dt, snn_info, Backend & bak )
		{
			using util::get;
			float const TstdpInv = 1.0f / 0.02f;
			float const Aplus = 0.01f * 0.0003f;
			float const Aminus = 0.0105f * 0.0003f;
			float const wmax = 0.0003f;
			float const decay = bak.pow( 1 - dt * TstdpInv, nsteps );

			float const zpre = bak.fma( get<Zpre>( syn ), decay, pre );
			float const zpost = bak.fma( get<Zpost>( syn ), decay, post );
			get<Zpre>( syn ) = zpre;
			get<Zpost>( syn ) = zpost;
			float const w = get<W>( syn );
			get<W>( syn ) = bak.clamp( w + post * Aplus * zpre - pre * Aminus * zpost, 0.0f, wmax );
		}

		HYBRID static bool plastic( int_ const src, int_ const dst, snn_info const info )
		{
			int_ const np = brunel_3pop::num_poisson( info );
			int_ const ne = brunel_3pop::num_exc_end( info );
			return src >= np && src < ne && dst >= np && dst < ne;
		}
	};
};
} // namespace spice
