#pragma once
// Brunel network with plastic synapses ("brunel+"): the neuron dynamics of brunel.h, with a
// per-synapse weight W and per-synapse traces Zpre / Zpost. Synapses from the LIF
// excitatory population onto any neuron below 0.9 N are plastic. Each step, traces decay by
// (1 - dt / 20 ms) and jump by the pre/post spike flags, then
//   W <- clamp(W - pre * 0.0202 W exp(-Zpost/dt) + post * 0.01 (1 - W) exp(-Zpre/dt), 0, 3e-4).
// Equations and constants are those of the reference model
// (spice/models/brunel_with_plasticity.h:9-137) and must stay identical.
#include <spice/models/model.h>

namespace spice
{
struct brunel_with_plasticity : model
{
	enum neuron_attr
	{
		V,
		Twait
	};
	enum syn_attr
	{
		W,
		Zpre,
		Zpost
	};

	struct neuron : ::spice::neuron<float, int_>
	{
		static constexpr std::uint32_t receive_mask = 1u << V;
		static constexpr std::uint32_t receive_reads = 1u << Twait;

		template <typename Iter, typename Backend>
		HYBRID static void init( Iter n, snn_info, Backend & )
		{
			using util::get;
			get<V>( n ) = 0.0f;
			get<Twait>( n ) = 0;
		}

		template <typename Iter, typename Backend>
		HYBRID static bool update( Iter n, float const dt, snn_info info, Backend & bak )
		{
			using util::get;
			float const TmemInv = 1.0 / 0.02;
			float const Vrest = 0.0;
			int_ const Tref = 20;
			float const Vthres = 0.02f;

			if( n.id() < static_cast<uint_>( info.num_neurons / 2 ) ) // Poisson source
			{
				float const rate = 20;
				return bak.rand() < ( rate * dt );
			}
			if( --get<Twait>( n ) <= 0 )
			{
				if( get<V>( n ) > Vthres )
				{
					get<V>( n ) = Vrest;
					get<Twait>( n ) = Tref;
					return true;
				}
				float const v = get<V>( n );
				get<V>( n ) = v + ( Vrest - v ) * ( dt * TmemInv );
			}
			return false;
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_, Iter dst, SynIter && syn, snn_info info, Backend & bak )
		{
			using util::get;
			if( get<Twait>( dst ) <= 0 ) bak.atomic_add( get<V>( dst ), get<W>( syn ) * 20'000 / info.num_neurons );
		}
	};

	struct synapse : ::spice::synapse<float, float, float>
	{
		template <typename Iter, typename Backend>
		HYBRID static void init( Iter syn, int_ src, int_, snn_info info, Backend & )
		{
			using util::get;
			int_ const nexc = static_cast<int_>( 0.9f * info.num_neurons );
			get<W>( syn ) = src < nexc ? 0.0001f : -0.0005f;
			get<Zpre>( syn ) = 1.0f;
			get<Zpost>( syn ) = 1.0f;
		}

		template <typename Iter, typename Backend>
		HYBRID static void update(
		    Iter && syn, int_ const nsteps, bool const pre, bool const post, float const dt, snn_info, Backend & bak )
		{
			using util::get;
			float const TstdpInv = 1.0f / 0.02f;
			float const dtInv = 1.0f / dt;
			float const decay = bak.pow( 1 - dt * TstdpInv, nsteps );

			get<Zpre>( syn ) = bak.fma( get<Zpre>( syn ), decay, pre );
			get<Zpost>( syn ) = bak.fma( get<Zpost>( syn ), decay, post );

			float const w = get<W>( syn );
			get<W>( syn ) = bak.clamp(
			    w - pre * 0.0202f * w * bak.exp( -get<Zpost>( syn ) * dtInv ) +
			        post * 0.01f * ( 1.0f - w ) * bak.exp( -get<Zpre>( syn ) * dtInv ),
			    0.0f,
			    0.0003f );
		}

		HYBRID static bool plastic( int_ const src, int_ const dst, snn_info const info )
		{
			int_ const npoisson = info.num_neurons / 2;
			int_ const nexc = 9 * info.num_neurons / 10;
			return src >= npoisson && src < nexc && dst < nexc;
		}
	};
};
} // namespace spice
