// Random-graph generation on gfx950 (subsystem a).
//
// Semantics: include/spice/util/graphgen.h (the reference's generator,
// spice/util/adj_list.cpp:49-91, re-keyed on Philox). Work split:
//   k_degrees : one thread per source row, sums the clipped per-connection degrees;
//   scan      : exclusive scan of row degrees -> int64 row offsets (CSR, > 2^32 edges OK);
//   k_targets : one wave64 per source row. For each covering connection the wave makes two
//               passes over the degree+1 exponential draws, 4 consecutive draws per lane
//               (one Philox block): pass 1 reduces the exact uint64 total, pass 2 forms the
//               exact prefix sums with a wave inclusive scan plus a running carry and writes
//               each lane's 4 targets (a wave stores 1 KB contiguous per round).
// Everything is integer except the final double-precision scaling, so the result equals the
// sequential host build bit for bit.
#include "graph.h"

namespace spice::hipx
{
using util::gen::conn;

static __global__ __launch_bounds__( 256 ) void k_degrees(
    conn const * conns, int nconns, std::int64_t n, std::int64_t width, std::uint64_t seed, std::int32_t * deg )
{
	std::int64_t const r = std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
	if( r >= n ) return;
	std::int64_t total = 0;
	for( int c = 0; c < nconns; c++ )
	{
		conn const cn = conns[c];
		if( r < cn.src_first || r >= cn.src_last ) continue;
		std::int64_t d = util::gen::sample_degree( seed, std::uint32_t( r ), std::uint32_t( c ), cn );
		if( d > width - total ) d = width - total;
		total += d;
	}
	deg[r] = std::int32_t( total );
}

static __global__ __launch_bounds__( 256 ) void k_targets(
    conn const * conns, int nconns, std::int64_t n, std::int64_t width, std::uint64_t seed, std::int64_t const * row_ptr, std::int32_t * col )
{
	std::int64_t const r = uniform64( ( std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> 6 );
	if( r >= n ) return;
	int const lane = lane_id();
	std::int64_t off = row_ptr[r];
	std::int64_t total = 0;
	for( int c = 0; c < nconns; c++ )
	{
		conn const cn = conns[c];
		if( r < cn.src_first || r >= cn.src_last ) continue;
		std::int64_t d = util::gen::sample_degree( seed, std::uint32_t( r ), std::uint32_t( c ), cn );
		if( d > width - total ) d = width - total;
		total += d;
		if( d == 0 ) continue;
		std::int32_t const R = ( cn.dst_last - cn.dst_first ) - std::int32_t( d );
		std::int64_t const nblk = ( d + 1 + 3 ) / 4;

		// pass 1: T_d = sum of F_0 .. F_d
		std::uint64_t part = 0;
		for( std::int64_t b = lane; b < nblk; b += 64 )
		{
			std::uint32_t f[4];
			util::gen::exp_fix4( seed, std::uint32_t( r ), std::uint32_t( c ), std::uint32_t( b ), f );
#pragma unroll
			for( int q = 0; q < 4; q++ )
				if( 4 * b + q <= d ) part += f[q];
		}
		std::uint64_t const ttotal = wave_sum( part );

		// pass 2: T_k and targets
		std::uint64_t carry = 0;
		for( std::int64_t b0 = 0; b0 < nblk; b0 += 64 )
		{
			std::int64_t const b = b0 + lane;
			std::uint32_t f[4] = { 0, 0, 0, 0 };
			if( b < nblk ) util::gen::exp_fix4( seed, std::uint32_t( r ), std::uint32_t( c ), std::uint32_t( b ), f );
			std::uint64_t s[4];
			s[0] = f[0];
			s[1] = s[0] + f[1];
			s[2] = s[1] + f[2];
			s[3] = s[2] + f[3];
			std::uint64_t const incl = wave_inclusive_scan( s[3] );
			std::uint64_t const excl = carry + incl - s[3];
#pragma unroll
			for( int q = 0; q < 4; q++ )
			{
				std::int64_t const k = 4 * b + q;
				if( k < d )
					col[off + k] = cn.dst_first + util::gen::position( excl + s[q], ttotal, R, std::int32_t( k ) );
			}
			carry += shfl_u64( incl, 63 );
		}
		off += d;
	}
}

// ---- exclusive scan int32 -> int64 (3 phases) --------------------------------------------
constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 4, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Block-level exclusive scan of a tile; writes tile-local prefixes and the tile total. Each
// thread reads its items before writing the same positions, so in == out is allowed.
template <typename In>
static __global__ __launch_bounds__( SCAN_BLOCK ) void k_scan_tiles( In const * in, std::int64_t * out, std::int64_t n, std::int64_t * tile_sums )
{
	__shared__ std::uint64_t wsum[SCAN_BLOCK / 64];
	std::int64_t const base = std::int64_t( blockIdx.x ) * SCAN_TILE + std::int64_t( threadIdx.x ) * SCAN_ITEMS;
	std::uint64_t v[SCAN_ITEMS];
	std::uint64_t t = 0;
#pragma unroll
	for( int q = 0; q < SCAN_ITEMS; q++ )
	{
		v[q] = ( base + q < n ) ? std::uint64_t( std::int64_t( in[base + q] ) ) : 0;
		t += v[q];
	}
	std::uint64_t const incl = wave_inclusive_scan( t );
	int const w = threadIdx.x >> 6;
	if( lane_id() == 63 ) wsum[w] = incl;
	__syncthreads();
	std::uint64_t wbase = 0;
	for( int i = 0; i < w; i++ ) wbase += wsum[i];
	std::uint64_t run = wbase + incl - t;
#pragma unroll
	for( int q = 0; q < SCAN_ITEMS; q++ )
	{
		if( base + q < n ) out[base + q] = std::int64_t( run );
		run += v[q];
	}
	if( threadIdx.x == SCAN_BLOCK - 1 ) tile_sums[blockIdx.x] = std::int64_t( run );
}

// Single block: exclusive scan of the tile totals (in place), total into out_total.
static __global__ __launch_bounds__( 1024 ) void k_scan_tile_sums( std::int64_t * tile_sums, std::int64_t ntiles, std::int64_t * out_total )
{
	__shared__ std::uint64_t wsum[16];
	__shared__ std::uint64_t carry_s;
	if( threadIdx.x == 0 ) carry_s = 0;
	__syncthreads();
	for( std::int64_t b0 = 0; b0 < ntiles; b0 += 1024 )
	{
		std::int64_t const i = b0 + threadIdx.x;
		std::uint64_t const x = i < ntiles ? std::uint64_t( tile_sums[i] ) : 0;
		std::uint64_t const incl = wave_inclusive_scan( x );
		int const w = threadIdx.x >> 6;
		if( lane_id() == 63 ) wsum[w] = incl;
		__syncthreads();
		std::uint64_t wbase = carry_s;
		for( int k = 0; k < w; k++ ) wbase += wsum[k];
		if( i < ntiles ) tile_sums[i] = std::int64_t( wbase + incl - x );
		__syncthreads();
		if( threadIdx.x == 1023 ) carry_s = wbase + incl;
		__syncthreads();
	}
	if( threadIdx.x == 0 ) *out_total = std::int64_t( carry_s );
}

static __global__ __launch_bounds__( 256 ) void k_scan_add( std::int64_t * out, std::int64_t n, std::int64_t const * tile_sums )
{
	std::int64_t const i = std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
	if( i < n ) out[i] += tile_sums[i / SCAN_TILE];
}

// Exclusive scan with the per-tile sums in `sums` (ceil(n / SCAN_TILE) entries, at least 1);
// out[n] = total. Enqueued on s.
template <typename In>
static void exclusive_scan_to_i64( In const * in, std::int64_t * out, std::int64_t n, std::int64_t * sums, hipStream_t s )
{
	std::int64_t const ntiles = ( n + SCAN_TILE - 1 ) / SCAN_TILE;
	if( n > 0 )
	{
		SPICE_LAUNCH( k_scan_tiles<In><<<dim3( unsigned( ntiles ) ), SCAN_BLOCK, 0, s>>>( in, out, n, sums ) );
	}
	SPICE_LAUNCH( k_scan_tile_sums<<<1, 1024, 0, s>>>( sums, ntiles, out + n ) );
	if( n > 0 )
	{
		SPICE_LAUNCH( k_scan_add<<<dim3( unsigned( ( n + 255 ) / 256 ) ), 256, 0, s>>>( out, n, sums ) );
	}
}

void exclusive_scan_i32_to_i64( std::int32_t const * in, std::int64_t * out, std::int64_t n, hipStream_t s )
{
	std::int64_t const ntiles = ( n + SCAN_TILE - 1 ) / SCAN_TILE;
	dbuf<std::int64_t> sums( ntiles > 0 ? ntiles : 1 );
	exclusive_scan_to_i64( in, out, n, sums.data(), s );
	SPICE_HIP( hipStreamSynchronize( s ) ); // `sums` is released on return
}

csr_dev generate_csr( std::vector<conn> const & conns, std::int64_t n, std::int64_t width, std::uint64_t seed, hipStream_t s )
{
	csr_dev g;
	g.n = n;
	g.width = width;
	g.row_ptr.resize( n + 1 );
	dbuf<conn> dconns( conns.size() > 0 ? conns.size() : 1 );
	if( !conns.empty() )
		SPICE_HIP( hipMemcpyAsync( dconns.data(), conns.data(), conns.size() * sizeof( conn ), hipMemcpyHostToDevice, s ) );
	dbuf<std::int32_t> deg( n > 0 ? n : 1 );
	if( n > 0 )
		SPICE_LAUNCH( k_degrees<<<dim3( unsigned( ( n + 255 ) / 256 ) ), 256, 0, s>>>(
		    dconns.data(), int( conns.size() ), n, width, seed, deg.data() ) );
	exclusive_scan_i32_to_i64( deg.data(), g.row_ptr.data(), n, s );
	SPICE_HIP( hipMemcpy( &g.nnz, g.row_ptr.data() + n, sizeof( std::int64_t ), hipMemcpyDeviceToHost ) );
	g.col.resize( g.nnz > 0 ? g.nnz : 1 );
	if( n > 0 && g.nnz > 0 )
	{
		std::int64_t const blocks = ( n + 3 ) / 4; // 4 waves (rows) per block
		SPICE_LAUNCH( k_targets<<<dim3( unsigned( blocks ) ), 256, 0, s>>>(
		    dconns.data(), int( conns.size() ), n, width, seed, g.row_ptr.data(), g.col.data() ) );
	}
	SPICE_HIP( hipStreamSynchronize( s ) );
	return g;
}

// ---- graph into caller-owned device memory --------------------------------------------------
// Row lengths go straight into row_ptr[0 .. n) (int64) and are scanned in place, so the only
// scratch is the connection table and the scan's per-tile sums.
static __global__ __launch_bounds__( 256 ) void k_degrees_i64(
    conn const * conns, int nconns, std::int64_t n, std::int64_t width, std::uint64_t seed, std::int64_t * len )
{
	std::int64_t const r = std::int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
	if( r >= n ) return;
	std::int64_t total = 0;
	for( int c = 0; c < nconns; c++ )
	{
		conn const cn = conns[c];
		if( r < cn.src_first || r >= cn.src_last ) continue;
		std::int64_t d = util::gen::sample_degree( seed, std::uint32_t( r ), std::uint32_t( c ), cn );
		if( d > width - total ) d = width - total;
		total += d;
	}
	len[r] = total;
}

std::int64_t graph_scratch_tiles( std::int64_t n ) { return n > 0 ? ( n + SCAN_TILE - 1 ) / SCAN_TILE : 1; }

std::int64_t generate_into( std::vector<conn> const & conns, std::int64_t n, std::int64_t width, std::uint64_t seed, std::int64_t * row_ptr,
                            std::int32_t * col, std::uint64_t col_cap, conn * conn_scratch, std::int64_t * tile_scratch, hipStream_t s )
{
	// conn_scratch and tile_scratch need no clearing: the copy below and the scan write every
	// entry they later read
	if( !conns.empty() )
		SPICE_HIP( hipMemcpyAsync( conn_scratch, conns.data(), conns.size() * sizeof( conn ), hipMemcpyHostToDevice, s ) );
	if( n > 0 )
		SPICE_LAUNCH( k_degrees_i64<<<dim3( unsigned( ( n + 255 ) / 256 ) ), 256, 0, s>>>( conn_scratch, int( conns.size() ), n, width, seed, row_ptr ) );
	exclusive_scan_to_i64( row_ptr, row_ptr, n, tile_scratch, s );
	std::int64_t nnz = 0;
	SPICE_HIP( hipMemcpyAsync( &nnz, row_ptr + n, sizeof( nnz ), hipMemcpyDeviceToHost, s ) );
	SPICE_HIP( hipStreamSynchronize( s ) ); // conns (host) copied, nnz read
	if( col && std::uint64_t( nnz ) <= col_cap && n > 0 && nnz > 0 )
	{
		SPICE_LAUNCH( k_targets<<<dim3( unsigned( ( n + 3 ) / 4 ) ), 256, 0, s>>>( conn_scratch, int( conns.size() ), n, width, seed, row_ptr, col ) );
		SPICE_HIP( hipStreamSynchronize( s ) );
	}
	return nnz;
}

csr_dev upload_csr( std::int64_t const * row_ptr, std::int32_t const * col, std::int64_t n, std::int64_t width, hipStream_t s )
{
	csr_dev g;
	g.n = n;
	g.width = width;
	g.nnz = row_ptr[n];
	g.row_ptr.resize( n + 1 );
	g.col.resize( g.nnz > 0 ? g.nnz : 1 );
	SPICE_HIP( hipMemcpyAsync( g.row_ptr.data(), row_ptr, ( n + 1 ) * sizeof( std::int64_t ), hipMemcpyHostToDevice, s ) );
	if( g.nnz > 0 )
		SPICE_HIP( hipMemcpyAsync( g.col.data(), col, g.nnz * sizeof( std::int32_t ), hipMemcpyHostToDevice, s ) );
	SPICE_HIP( hipStreamSynchronize( s ) );
	return g;
}
} // namespace spice::hipx
