#pragma once
// Vogels & Abbott (2005) conductance-based LIF network (COBA).
//
// Equations and constants are those of the reference model (spice/models/vogels_abbott.h:9-90)
// and must stay identical for results to match it: Euler step of
//   tau_m dV/dt = (V_rest - V) + g_E (E_E - V) + g_I (E_I - V) + I_bg,
// exponential conductance decay tau_E = 5 ms / tau_I = 10 ms, refractory 50 steps, the first
// 80 % of neurons excitatory, weights scaled by 16e6 / N^2.
#include <spice/models/model.h>

namespace spice
{
struct vogels_abbott : model
{
	struct neuron : ::spice::neuron<float, float, float, int_>
	{
		enum attr
		{
			V,
			Gex,
			Gin,
			Twait
		};
		static constexpr std::uint32_t receive_mask = ( 1u << Gex ) | ( 1u << Gin );
		static constexpr std::uint32_t receive_reads = 0; // receive() reads nothing of its target

		template <typename Iter, typename Backend>
		HYBRID static void init( Iter n, snn_info, Backend & )
		{
			using util::get;
			get<V>( n ) = -0.06f;
			get<Gex>( n ) = 0.0f;
			get<Gin>( n ) = 0.0f;
			get<Twait>( n ) = 0;
		}

		template <typename Iter, typename Backend>
		HYBRID static bool update( Iter n, float const dt, snn_info, Backend & )
		{
			using util::get;
			int_ const Tref = 50;                // steps
			float const Vrest = -0.06f;          // V
			float const Vthres = -0.05f;         // V
			float const TmemInv = 1.0f / 0.02f;  // 1/s
			float const Eex = 0.0f;              // V
			float const Ein = -0.08f;            // V
			float const Ibg = 0.02f;             // V
			float const TexInv = 1.0f / 0.005f;  // 1/s
			float const TinInv = 1.0f / 0.01f;   // 1/s

			bool spiked = false;
			if( --get<Twait>( n ) <= 0 )
			{
				if( get<V>( n ) > Vthres )
				{
					get<V>( n ) = Vrest;
					get<Twait>( n ) = Tref;
					spiked = true;
				}
				else
				{
					float const v = get<V>( n );
					get<V>( n ) = v + ( ( Vrest - v ) + get<Gex>( n ) * ( Eex - v ) +
					                    get<Gin>( n ) * ( Ein - v ) + Ibg ) *
					                      ( dt * TmemInv );
				}
			}
			get<Gex>( n ) -= get<Gex>( n ) * ( dt * TexInv );
			get<Gin>( n ) -= get<Gin>( n ) * ( dt * TinInv );
			return spiked;
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_ src, Iter dst, SynIter &&, snn_info info, Backend & bak )
		{
			using util::get;
			float const Wex = 0.4f * 16'000'000 / ( (long_)info.num_neurons * info.num_neurons ); // S
			float const Win = 5.1f * 16'000'000 / ( (long_)info.num_neurons * info.num_neurons ); // S
			int_ const nexc = static_cast<int_>( 0.8f * info.num_neurons );
			if( src < nexc )
				bak.atomic_add( get<Gex>( dst ), Wex );
			else
				bak.atomic_add( get<Gin>( dst ), Win );
		}
	};
};
} // namespace spice
