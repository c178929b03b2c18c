#pragma once
// TEST INFRASTRUCTURE ONLY (the CPU checker). Not part of the product; nothing under
// spice_amd/ or include/ may use it.
//
// Host build of the random-graph generator: composes the shared primitives of
// include/spice/util/graphgen.h sequentially, row by row, connection by connection, draw by
// draw. The device kernel evaluates the same primitives in parallel (wave scans), so the two
// CSR graphs must be bit-identical. Semantics follow the reference's host generator
// (spice/util/adj_list.cpp:49-91): per row, per covering connection, degree clipped to the
// remaining ELL width, then sorted targets from exponential spacings.
#include <spice/util/graphgen.h>
#include <spice/util/layout.h>

#include <cmath>
#include <cstdint>
#include <vector>

namespace spice::oracle
{
inline std::vector<util::gen::conn> gen_conns( util::layout const & desc )
{
	std::vector<util::gen::conn> out;
	for( auto const & c : desc.connections() )
	{
		util::gen::conn g;
		g.src_first = std::get<0>( c );
		g.src_last = std::get<1>( c );
		g.dst_first = std::get<2>( c );
		g.dst_last = std::get<3>( c );
		double const range = double( g.dst_last - g.dst_first );
		double const p = std::get<4>( c );
		g.mean = range * p;
		g.sd = std::sqrt( range * p * ( 1.0 - p ) );
		out.push_back( g );
	}
	return out;
}

struct csr
{
	std::vector<std::int64_t> row_ptr;
	std::vector<std::int32_t> col;
};

// Length of row r: the sum over covering connections of the sampled degrees, each clipped to
// what is left of the row width.
inline std::int64_t row_length( std::vector<util::gen::conn> const & conns, std::int64_t r, std::int64_t width, std::uint64_t seed )
{
	std::int64_t total = 0;
	for( std::uint32_t c = 0; c < conns.size(); c++ )
	{
		if( r < conns[c].src_first || r >= conns[c].src_last ) continue;
		std::int64_t d = util::gen::sample_degree( seed, std::uint32_t( r ), c, conns[c] );
		if( d > width - total ) d = width - total;
		total += d;
	}
	return total;
}

// Writes the targets of row r to out[0 .. row_length): per covering connection, the sorted
// targets of its exponential spacings.
inline void row_targets( std::vector<util::gen::conn> const & conns, std::int64_t r, std::int64_t width, std::uint64_t seed, std::int32_t * out )
{
	std::int64_t off = 0, total = 0;
	for( std::uint32_t c = 0; c < conns.size(); c++ )
	{
		auto const & cn = conns[c];
		if( r < cn.src_first || r >= cn.src_last ) continue;
		std::int64_t d = util::gen::sample_degree( seed, std::uint32_t( r ), c, cn );
		if( d > width - total ) d = width - total;
		total += d;
		std::int32_t const range = cn.dst_last - cn.dst_first;
		std::int32_t const R = range - std::int32_t( d );
		// F_0 .. F_d (d + 1 draws), 4 per Philox block.
		std::vector<std::uint32_t> f( ( d + 1 + 3 ) / 4 * 4 );
		for( std::uint32_t b = 0; b * 4 < f.size(); b++ )
			util::gen::exp_fix4( seed, std::uint32_t( r ), c, b, &f[b * 4] );
		std::uint64_t ttotal = 0;
		for( std::int64_t k = 0; k <= d; k++ ) ttotal += f[k];
		std::uint64_t tk = 0;
		for( std::int64_t k = 0; k < d; k++ )
		{
			tk += f[k];
			out[off + k] = cn.dst_first + util::gen::position( tk, ttotal, R, std::int32_t( k ) );
		}
		off += d;
	}
}

// Generates the CSR adjacency of `desc` for `seed` (rows clipped to `width` entries).
inline csr generate_csr(
    std::vector<util::gen::conn> const & conns, std::int64_t n, std::int64_t width, std::uint64_t seed )
{
	csr g;
	g.row_ptr.assign( n + 1, 0 );
	for( std::int64_t r = 0; r < n; r++ ) g.row_ptr[r + 1] = g.row_ptr[r] + row_length( conns, r, width, seed );
	g.col.resize( g.row_ptr[n] );
	for( std::int64_t r = 0; r < n; r++ ) row_targets( conns, r, width, seed, g.col.data() + g.row_ptr[r] );
	return g;
}

// Rows [r0, r1) only (graphs too large to build whole on the host): the absolute offsets
// row_ptr[r0] .. row_ptr[r1] (every earlier row's length is summed: lengths are cheap, one
// Philox block per row and connection) and the targets of those rows.
inline csr generate_rows(
    std::vector<util::gen::conn> const & conns, std::int64_t r0, std::int64_t r1, std::int64_t width, std::uint64_t seed, bool cols = true )
{
	csr g;
	std::int64_t base = 0;
	for( std::int64_t r = 0; r < r0; r++ ) base += row_length( conns, r, width, seed );
	g.row_ptr.assign( r1 - r0 + 1, base );
	for( std::int64_t r = r0; r < r1; r++ ) g.row_ptr[r - r0 + 1] = g.row_ptr[r - r0] + row_length( conns, r, width, seed );
	if( !cols ) return g; // offsets only: no targets are drawn
	g.col.resize( g.row_ptr[r1 - r0] - base );
	for( std::int64_t r = r0; r < r1; r++ ) row_targets( conns, r, width, seed, g.col.data() + ( g.row_ptr[r - r0] - base ) );
	return g;
}
} // namespace spice::oracle
