#pragma once
// Three-population Brunel network: a 20 Hz Poisson input population (the first N/5
// neurons; snn_info::param[0] > 0 sets another rate, e.g. for parameter sweeps), then excitatory (4/5 of the rest) and inhibitory LIF populations. LIF: fp32 V,
// tau 20 ms, threshold 20 mV, reset to 10 mV, 20 refractory steps (2 ms at dt = 0.1 ms).
// Fixed fp32 weights scaling as 1/N like the reference's (+5/N V from Poisson and excitatory
// sources, -25/N V from inhibitory ones: the reference's 1 : -5 ratio), applied to
// non-refractory targets. The constant is 5e4 / N rather than the reference's 2e4 / N: with
// a 20 % Poisson population at p = 0.1 the latter gives a mean external drive of
// 0.02 N * 0.1 * 20 Hz * (2/N V) * 20 ms = 16 mV, below the 20 mV threshold, and the LIF
// populations stay silent at every N; 5e4 / N restores the 40 mV mean external drive of the
// reference's own Brunel topology (half Poisson, samples/main.cpp:1392-1396).
// Intended topology: all population pairs connected at p = 0.1, delay 15 steps.
#include <spice/models/model.h>

namespace spice
{
struct brunel_3pop : model
{
	HYBRID static int_ num_poisson( snn_info info ) { return info.num_neurons / 5; }
	HYBRID static int_ num_exc_end( snn_info info )
	{
		int_ const np = num_poisson( info );
		return np + ( info.num_neurons - np ) * 4 / 5;
	}

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
			float const TmemInv = 1.0f / 0.02f; // 1/s
			float const Vrest = 0.0f;           // V
			float const Vreset = 0.01f;         // V
			float const Vthres = 0.02f;         // V
			int_ const Tref = 20;               // steps

			if( n.id() < static_cast<uint_>( num_poisson( info ) ) )
			{
				float const rate = info.param[0] > 0.0f ? info.param[0] : 20.0f; // Hz (param[0] overrides)
				return bak.rand() < ( rate * dt );
			}
			if( --get<Twait>( n ) <= 0 )
			{
				if( get<V>( n ) > Vthres )
				{
					get<V>( n ) = Vreset;
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
			float const Wex = 0.0001f * 50'000 / info.num_neurons;  // V
			float const Win = -0.0005f * 50'000 / info.num_neurons; // V
			bak.atomic_add( util::get<V>( dst ), src < num_exc_end( info ) ? Wex : Win );
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_ src, Iter dst, SynIter && syn, snn_info info, Backend & bak )
		{
			if( accepts( dst ) ) deliver( src, dst, syn, info, bak );
		}
	};
};
} // namespace spice
