#pragma once
// Brunel (2000) style network: the first half of the neurons are 20 Hz Poisson sources,
// the rest are LIF neurons (tau 20 ms, threshold 20 mV, reset to rest, refractory 20 steps),
// the first 90 % of all neurons excitatory. Weights scale as 2e4 / N and are only applied
// to non-refractory targets. Equations and constants are those of the reference model
// (spice/models/brunel.h:9-82) and must stay identical for results to match it.
#include <spice/models/model.h>

namespace spice
{
struct brunel : model
{
	struct neuron : ::spice::neuron<float, int_>
	{
		enum attr
		{
			V,
			Twait
		};
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
			float const TmemInv = 1.0 / 0.02; // 1/s
			float const Vrest = 0.0;          // V
			int_ const Tref = 20;             // steps
			float const Vthres = 0.02f;       // V

			if( n.id() < static_cast<uint_>( info.num_neurons / 2 ) ) // Poisson source
			{
				float const rate = 20; // Hz
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

		// receive() = accepts() ? deliver() : nothing (split so that the GPU may test the gate
		// in the neuron update)
		static constexpr bool gated_receive = true;

		template <typename Iter>
		HYBRID static bool accepts( Iter n )
		{
			return util::get<Twait>( n ) <= 0;
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void deliver( int_ src, Iter dst, SynIter &&, snn_info info, Backend & bak )
		{
			int_ const nexc = static_cast<int_>( 0.9f * info.num_neurons );
			float const Wex = 0.0001f * 20'000 / info.num_neurons;  // V
			float const Win = -0.0005f * 20'000 / info.num_neurons; // V
			bak.atomic_add( util::get<V>( dst ), src < nexc ? Wex : Win );
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_ src, Iter dst, SynIter && syn, snn_info info, Backend & bak )
		{
			if( accepts( dst ) ) deliver( src, dst, syn, info, bak );
		}
	};
};
} // namespace spice
