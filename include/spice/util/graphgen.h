#pragma once
// Random-graph generator primitives, shared by the gfx950 kernel and its host build.
//
// Semantics follow the reference generator (spice/util/adj_list.cpp:49-91,
// spice/util/random.h:117-156): for every source row and every connection (src pop ->
// dst pop, p) covering it,
//   degree  = clamp(round(N(range*p, range*p*(1-p))), 0, range), clipped so the row's total
//             never exceeds the ELL width (layout max_degree);
//   targets = degree strictly increasing ids in [dst_first, dst_last), drawn as exponential
//             spacings: T_k = E_0 + ... + E_k (E ~ Exp(1)), id_k = first +
//             floor(T_k * (range - degree) / T_deg) + k.
// What differs by design: randomness is Philox keyed on (seed, row, connection, index), so
// every row is generated independently and in parallel; the exponentials are carried in
// 24-bit fixed point and summed in uint64, so the prefix sums are exact and independent of
// summation order (a wave scan on the device equals a sequential loop on the host); and all
// floating-point steps use the deterministic routines in detmath.h.
#include <spice/util/detmath.h>
#include <spice/util/host_defines.h>
#include <spice/util/philox.h>

#include <cstdint>

namespace spice::util::gen
{
// One flattened connection as the generator sees it. mean/sd are computed once on the
// host (double precision) so no device square root enters the result.
struct conn
{
	std::int32_t src_first, src_last, dst_first, dst_last;
	double mean; // range * p
	double sd;   // sqrt(range * p * (1 - p))
};

// Degree of (row, connection c) before width clipping.
HYBRID inline std::int32_t sample_degree( std::uint64_t seed, std::uint32_t row, std::uint32_t c, conn const & cn )
{
	std::int32_t const range = cn.dst_last - cn.dst_first;
	if( !( cn.sd > 0.0 ) )
	{
		// p == 0 or p == 1 (or range == 0): the degree is exact.
		double const m = cn.mean;
		std::int32_t d = std::int32_t( std::rint( m ) );
		return d < 0 ? 0 : d > range ? range : d;
	}
	philox_u4 const r = philox( seed, { row, c, 0u, TAG_GRAPH_DEGREE } );
	// Box-Muller: z = sqrt(-2 ln u1) cos(2 pi u2), u1 in (0, 1], u2 in [0, 1).
	double const u1 = ( double( r.x ) + 1.0 ) * ( 1.0 / 4294967296.0 );
	double const u2 = double( r.y ) * ( 1.0 / 4294967296.0 );
	double const z = det::sqrt( -2.0 * det::log( u1 ) ) * det::cos2pi( u2 );
	double const x = std::rint( det::fma_( z, cn.sd, cn.mean ) );
	if( !( x > 0.0 ) ) return 0;
	if( x >= double( range ) ) return range;
	return std::int32_t( x );
}

// Four consecutive Exp(1) draws k = 4*blk .. 4*blk+3 of (row, c), as floor(E * 2^24).
// E = -ln(U), U = (x + 1) / 2^32 in (0, 1], so 0 <= E <= 22.2 and the value fits 32 bits.
HYBRID inline std::uint32_t exp_fix_one( std::uint32_t x )
{
	// -ln((x+1) 2^-32) = 32 ln 2 - ln(x+1); computed as -ln(m * 2^-32) with m = x+1 exact.
	double const u = ( double( x ) + 1.0 ) * ( 1.0 / 4294967296.0 ); // exact
	double const e = -det::log( u );
	double const f = e * 16777216.0; // exact scaling
	return f > 0.0 ? std::uint32_t( f ) : 0u;
}

HYBRID inline void exp_fix4( std::uint64_t seed, std::uint32_t row, std::uint32_t c, std::uint32_t blk, std::uint32_t out[4] )
{
	philox_u4 const r = philox( seed, { blk, row, c, TAG_GRAPH_POS } );
	out[0] = exp_fix_one( r.x );
	out[1] = exp_fix_one( r.y );
	out[2] = exp_fix_one( r.z );
	out[3] = exp_fix_one( r.w );
}

// Target offset (within the connection's dst range) of draw k, given the exact prefix sum
// T_k, the total T_deg and R = range - degree.
HYBRID inline std::int32_t position( std::uint64_t tk, std::uint64_t ttotal, std::int32_t R, std::int32_t k )
{
	if( ttotal == 0 ) return k;
	double const q = ( double( tk ) * double( R ) ) / double( ttotal );
	return std::int32_t( q ) + k; // q >= 0: truncation == floor
}
} // namespace spice::util::gen
