#pragma once
// Kahan-compensated accumulation of the simulation clock.
//
// Every step adds dt to the simulated time; add() returns the compensated increment that
// is actually passed to the models (reference: spice/util/numeric.h:8-33,
// spice/snn.cpp:38-44). The gfx950 kernels advance an identical copy of this state on
// the device with the same float operations, so host and device agree bit for bit.
#include <spice/util/host_defines.h>

namespace spice::util
{
template <typename Prec>
class kahan_sum
{
public:
	// Adds delta; returns the increment actually applied (delta - compensation).
	HYBRID Prec add( Prec delta )
	{
		Prec const y = delta - _c;
		Prec const t = _sum + y;
		_c = ( t - _sum ) - y;
		_sum = t;
		return y;
	}

	HYBRID operator Prec() const { return _sum; }
	HYBRID Prec compensation() const { return _c; }
	HYBRID void set( Prec sum, Prec c )
	{
		_sum = sum;
		_c = c;
	}

private:
	Prec _c = 0;
	Prec _sum = 0;
};

// (i mod size) for i >= -size (reference: spice/util/circular_buffer.h:13-20).
template <typename Int>
HYBRID Int circidx( Int i, Int size )
{
	return ( i + size ) % size;
}
} // namespace spice::util
