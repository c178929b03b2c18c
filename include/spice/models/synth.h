#pragma once
// Synthetic benchmark model: one population, every neuron fires with probability 0.005 per
// step, each delivered spike increments an integer counter at the target
// (reference: spice/models/synth.h:12-42).
#include <spice/models/model.h>

namespace spice
{
struct synth : model
{
	struct neuron : ::spice::neuron<int>
	{
		enum attr
		{
			N
		};

		static constexpr std::uint32_t receive_reads = 0; // receive() reads nothing of its target

		template <typename Iter, typename Backend>
		HYBRID static void init( Iter n, snn_info, Backend & )
		{
			util::get<N>( n ) = 0;
		}

		template <typename Iter, typename Backend>
		HYBRID static bool update( Iter, float, snn_info, Backend & bak )
		{
			return bak.rand() < 0.005f;
		}

		template <typename Iter, typename SynIter, typename Backend>
		HYBRID static void receive( int_, Iter dst, SynIter &&, snn_info, Backend & bak )
		{
			bak.atomic_add( util::get<N>( dst ), 1 );
		}
	};
};
} // namespace spice
