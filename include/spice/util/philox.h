#pragma once
// Philox4x32-10 counter-based random number generator.
//
// Written from the published description: J. K. Salmon, M. A. Moraes, R. O. Dror,
// D. E. Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11, section 5.1
// (Philox-4x32, multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85, 10 rounds). Being counter-based, every random number in the
// simulator is a pure function of (seed, counter). The host build and the gfx950 build
// therefore produce identical streams, and no RNG state is stored per thread.
//
// The reference uses sequential xoshiro streams instead (spice/util/random.h:33-116); the
// counter-based design is what lets host and device agree bit for bit.
#include <spice/util/host_defines.h>

#include <cstdint>

namespace spice::util
{
struct philox_u4
{
	std::uint32_t x, y, z, w;
};

HYBRID inline void _philox_mulhilo( std::uint32_t a, std::uint32_t b, std::uint32_t & hi, std::uint32_t & lo )
{
	std::uint64_t const p = std::uint64_t( a ) * std::uint64_t( b );
	hi = std::uint32_t( p >> 32 );
	lo = std::uint32_t( p );
}

// One Philox4x32-10 block: 128-bit counter, 64-bit key -> 128 random bits.
HYBRID inline philox_u4 philox4x32_10( philox_u4 ctr, std::uint32_t k0, std::uint32_t k1 )
{
	constexpr std::uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
	constexpr std::uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined( SPICE_DEVICE_CODE )
#pragma unroll
#endif
	for( int r = 0; r < 10; ++r )
	{
		if( r > 0 )
		{
			k0 += W0;
			k1 += W1;
		}
		std::uint32_t hi0, lo0, hi1, lo1;
		_philox_mulhilo( M0, ctr.x, hi0, lo0 );
		_philox_mulhilo( M1, ctr.z, hi1, lo1 );
		ctr = { hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0 };
	}
	return ctr;
}

HYBRID inline philox_u4 philox( std::uint64_t seed, philox_u4 ctr )
{
	return philox4x32_10( ctr, std::uint32_t( seed ), std::uint32_t( seed >> 32 ) );
}

// Uniform float in [0, 1) from the top 24 bits (same mapping as the reference's
// uniform_left_inc, spice/util/random.h:120-123).
HYBRID inline float u01_left( std::uint32_t x ) { return float( x >> 8 ) * ( 1.0f / 16777216.0f ); }

// Stream tags: the last counter word separates independent uses of one seed.
enum rng_tag : std::uint32_t
{
	TAG_GRAPH_DEGREE = 1,
	TAG_GRAPH_POS = 2,
	TAG_NEURON_INIT = 3,
	TAG_NEURON_UPDATE = 4,
	TAG_SYNAPSE_INIT = 5,
	TAG_RECEIVE = 6,
	TAG_SYNAPSE_UPDATE = 7,
};

// Per-call-site random stream handed to model functors as backend.rand(). The stream is
// identified by three 32-bit words (e.g. neuron id and step) and a tag; successive draws
// walk a 24-bit block index stored next to the tag. Host and device share this class.
class philox_stream
{
public:
	HYBRID philox_stream( std::uint64_t seed, std::uint32_t a, std::uint32_t b, std::uint32_t c, std::uint32_t tag )
	    : _seed( seed )
	    , _a( a )
	    , _b( b )
	    , _c( c )
	    , _tag( tag )
	{
	}

	HYBRID std::uint32_t next()
	{
		if( _i == 0 ) _buf = philox( _seed, { _a, _b, _c, _tag | ( _blk++ << 8 ) } );
		std::uint32_t const r = _i == 0 ? _buf.x : _i == 1 ? _buf.y : _i == 2 ? _buf.z : _buf.w;
		_i = ( _i + 1 ) & 3;
		return r;
	}

	// @return uniform float in [0, 1)
	HYBRID float rand() { return u01_left( next() ); }

private:
	std::uint64_t _seed;
	std::uint32_t _a, _b, _c, _tag;
	std::uint32_t _blk = 0;
	std::uint32_t _i = 0;
	philox_u4 _buf{};
};
} // namespace spice::util
