#pragma once
// HIP runtime plumbing for the gfx950 library: error mapping, RAII device buffers and
// wave64 helpers built on amdgcn builtins.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

namespace spice::hipx
{
// Reference counterpart: success_or_throw / call (spice/cuda/util/error.h:19-67).
inline void check( hipError_t e, char const * what )
{
	if( e == hipSuccess ) return;
	(void)hipGetLastError(); // clear sticky error state, as the reference does (error.h:33)
	if( e == hipErrorOutOfMemory ) throw std::bad_alloc();
	throw std::runtime_error( std::string( what ) + ": " + hipGetErrorString( e ) );
}
#define SPICE_HIP( call ) ::spice::hipx::check( ( call ), #call )
#define SPICE_LAUNCH( ... )                                           \
	do                                                                \
	{                                                                 \
		__VA_ARGS__;                                                  \
		::spice::hipx::check( hipGetLastError(), "kernel launch" );   \
	} while( 0 )

// Owning device allocation (reference counterpart: dbuffer, spice/cuda/util/dbuffer.h:17-90).
template <typename T>
class dbuf
{
public:
	dbuf() = default;
	explicit dbuf( size_t n ) { resize( n ); }
	~dbuf() { release(); }
	dbuf( dbuf const & ) = delete;
	dbuf & operator=( dbuf const & ) = delete;
	dbuf( dbuf && o ) noexcept
	    : _p( std::exchange( o._p, nullptr ) )
	    , _n( std::exchange( o._n, 0 ) )
	{
	}
	dbuf & operator=( dbuf && o ) noexcept
	{
		release();
		_p = std::exchange( o._p, nullptr );
		_n = std::exchange( o._n, 0 );
		return *this;
	}

	void resize( size_t n )
	{
		release();
		if( n == 0 ) return;
		void * p = nullptr;
		hipError_t e = hipMalloc( &p, n * sizeof( T ) );
		if( e != hipSuccess )
		{
			(void)hipGetLastError();
			throw std::bad_alloc();
		}
		_p = static_cast<T *>( p );
		_n = n;
	}
	void release()
	{
		if( _p ) (void)hipFree( _p );
		_p = nullptr;
		_n = 0;
	}

	T * data() const { return _p; }
	size_t size() const { return _n; }
	size_t bytes() const { return _n * sizeof( T ); }

private:
	T * _p = nullptr;
	size_t _n = 0;
};

// ---- wave64 helpers ----------------------------------------------------------------------

__device__ inline int lane_id() { return int( threadIdx.x & 63u ); }

__device__ inline std::uint64_t ballot( bool p ) { return __builtin_amdgcn_ballot_w64( p ); }

// Number of set bits of `mask` below this lane.
__device__ inline int lanes_below( std::uint64_t mask )
{
	return int( __builtin_amdgcn_mbcnt_hi( std::uint32_t( mask >> 32 ), __builtin_amdgcn_mbcnt_lo( std::uint32_t( mask ), 0u ) ) );
}

__device__ inline std::uint32_t shfl_u32( std::uint32_t x, int src_lane )
{
	return std::uint32_t( __builtin_amdgcn_ds_bpermute( src_lane << 2, int( x ) ) );
}
__device__ inline std::uint64_t shfl_u64( std::uint64_t x, int src_lane )
{
	return std::uint64_t( shfl_u32( std::uint32_t( x ), src_lane ) ) |
	       ( std::uint64_t( shfl_u32( std::uint32_t( x >> 32 ), src_lane ) ) << 32 );
}
__device__ inline int uniform( int x ) { return __builtin_amdgcn_readfirstlane( x ); }
__device__ inline std::int64_t uniform64( std::int64_t x )
{
	std::uint32_t lo = __builtin_amdgcn_readfirstlane( std::uint32_t( x ) );
	std::uint32_t hi = __builtin_amdgcn_readfirstlane( std::uint32_t( std::uint64_t( x ) >> 32 ) );
	return std::int64_t( ( std::uint64_t( hi ) << 32 ) | lo );
}

// Inclusive prefix sum over the 64 lanes of a wave (Hillis-Steele on ds_bpermute).
__device__ inline std::uint64_t wave_inclusive_scan( std::uint64_t x )
{
	int const l = lane_id();
#pragma unroll
	for( int d = 1; d < 64; d <<= 1 )
	{
		std::uint64_t const y = shfl_u64( x, l - d );
		if( l >= d ) x += y;
	}
	return x;
}
__device__ inline std::uint64_t wave_sum( std::uint64_t x ) { return shfl_u64( wave_inclusive_scan( x ), 63 ); }

// Butterfly min / max over the 64 lanes (every lane gets the result).
__device__ inline std::int64_t wave_min( std::int64_t x )
{
	int const l = lane_id();
#pragma unroll
	for( int d = 1; d < 64; d <<= 1 )
	{
		std::int64_t const y = std::int64_t( shfl_u64( std::uint64_t( x ), l ^ d ) );
		x = y < x ? y : x;
	}
	return x;
}
__device__ inline std::int64_t wave_max( std::int64_t x )
{
	int const l = lane_id();
#pragma unroll
	for( int d = 1; d < 64; d <<= 1 )
	{
		std::int64_t const y = std::int64_t( shfl_u64( std::uint64_t( x ), l ^ d ) );
		x = y > x ? y : x;
	}
	return x;
}
} // namespace spice::hipx
