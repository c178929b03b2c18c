#pragma once
// Deterministic double-precision math shared by the host build and the gfx950 build.
//
// Every routine is a fixed sequence of IEEE-754 operations that are correctly rounded on
// both x86-64 and gfx950 (add, sub, mul, div, fma, exact bit manipulation), with each
// multiply-add written as an explicit fma so that compiler contraction settings cannot
// change a result. The graph generator uses these, which is what makes the generated
// adjacency bit-identical between the host and the device. No vendor math library code
// is reproduced; the methods are textbook:
//   * log: exponent/mantissa split, mantissa reduced to [sqrt(1/2), sqrt(2)), then
//     ln(f) = 2 atanh(s), s = (f-1)/(f+1), via its odd power series.
//   * sqrt: Newton-Raphson (Heron) iteration from an exponent-halved initial guess.
//   * cos(2 pi u): quadrant reduction to |theta| <= pi/4 and Taylor polynomials.
#include <spice/util/host_defines.h>

#include <cmath>
#include <cstdint>

namespace spice::util::det
{
HYBRID inline std::uint64_t bits( double x ) { return __builtin_bit_cast( std::uint64_t, x ); }
HYBRID inline double from_bits( std::uint64_t u ) { return __builtin_bit_cast( double, u ); }

HYBRID inline double fma_( double a, double b, double c )
{
#if defined( SPICE_DEVICE_CODE ) && defined( __HIP_DEVICE_COMPILE__ )
	return __builtin_fma( a, b, c );
#else
	return std::fma( a, b, c );
#endif
}

// Natural logarithm of a positive, normal, finite double.
HYBRID inline double log( double x )
{
	constexpr double LN2_HI = 6.93147180369123816490e-01; // high bits of ln 2
	constexpr double LN2_LO = 1.90821492927058770002e-10; // ln 2 - LN2_HI
	constexpr double SQRT2 = 1.41421356237309504880;

	std::uint64_t const b = bits( x );
	double e = double( int( ( b >> 52 ) & 0x7ff ) - 1023 );
	double f = from_bits( ( b & 0x000fffffffffffffull ) | 0x3ff0000000000000ull ); // [1, 2)
	if( f > SQRT2 )
	{
		f = f * 0.5; // exact
		e = e + 1.0;
	}
	double const s = ( f - 1.0 ) / ( f + 1.0 );
	double const s2 = s * s;
	// atanh(s)/s = sum_k s^(2k) / (2k+1), |s| <= 0.1716, 12 terms reach < 2^-60.
	double p = 1.0 / 23.0;
	p = fma_( p, s2, 1.0 / 21.0 );
	p = fma_( p, s2, 1.0 / 19.0 );
	p = fma_( p, s2, 1.0 / 17.0 );
	p = fma_( p, s2, 1.0 / 15.0 );
	p = fma_( p, s2, 1.0 / 13.0 );
	p = fma_( p, s2, 1.0 / 11.0 );
	p = fma_( p, s2, 1.0 / 9.0 );
	p = fma_( p, s2, 1.0 / 7.0 );
	p = fma_( p, s2, 1.0 / 5.0 );
	p = fma_( p, s2, 1.0 / 3.0 );
	p = fma_( p, s2, 1.0 );
	double const lnf = ( 2.0 * s ) * p;
	return fma_( e, LN2_HI, fma_( e, LN2_LO, lnf ) );
}

// Square root of a non-negative finite double (exactly 0 for 0).
HYBRID inline double sqrt( double a )
{
	if( !( a > 0.0 ) ) return 0.0;
	// Initial guess: halve the biased exponent (within a factor ~1.5 of the root).
	double y = from_bits( ( bits( a ) >> 1 ) + 0x1ff8000000000000ull );
	for( int i = 0; i < 7; ++i ) y = 0.5 * ( y + a / y );
	return y;
}

// cos(2 pi u) for u in [0, 1].
HYBRID inline double cos2pi( double u )
{
	constexpr double PI_2 = 1.57079632679489661923; // pi / 2
	double const t = 4.0 * u;                       // exact
	double const q = std::floor( t + 0.5 );         // nearest quadrant, t + 0.5 exact here
	double const r = t - q;                         // [-0.5, 0.5], exact
	double const th = r * PI_2;                     // |th| <= pi/4
	double const t2 = th * th;
	// cos: 1 - t2/2! + t2^2/4! - ... (to t2^8)
	double c = 1.0 / 20922789888000.0;       // 1/16!
	c = fma_( c, -t2, 1.0 / 87178291200.0 ); // 1/14!
	c = fma_( c, -t2, 1.0 / 479001600.0 );   // 1/12!
	c = fma_( c, -t2, 1.0 / 3628800.0 );     // 1/10!
	c = fma_( c, -t2, 1.0 / 40320.0 );       // 1/8!
	c = fma_( c, -t2, 1.0 / 720.0 );         // 1/6!
	c = fma_( c, -t2, 1.0 / 24.0 );          // 1/4!
	c = fma_( c, -t2, 1.0 / 2.0 );           // 1/2!
	c = fma_( c, -t2, 1.0 );
	// sin: th (1 - t2/3! + t2^2/5! - ...) (to t2^8)
	double s = 1.0 / 355687428096000.0;      // 1/17!
	s = fma_( s, -t2, 1.0 / 1307674368000.0 ); // 1/15!
	s = fma_( s, -t2, 1.0 / 6227020800.0 );  // 1/13!
	s = fma_( s, -t2, 1.0 / 39916800.0 );    // 1/11!
	s = fma_( s, -t2, 1.0 / 362880.0 );      // 1/9!
	s = fma_( s, -t2, 1.0 / 5040.0 );        // 1/7!
	s = fma_( s, -t2, 1.0 / 120.0 );         // 1/5!
	s = fma_( s, -t2, 1.0 / 6.0 );           // 1/3!
	s = fma_( s, -t2, 1.0 );
	s = s * th;
	int const quad = int( q ) & 3;
	return quad == 0 ? c : quad == 1 ? -s : quad == 2 ? -c : s;
}
} // namespace spice::util::det
