#pragma once
// Device graph construction (model independent).
#include "common.h"

#include <spice/util/graphgen.h>

#include <cstdint>
#include <vector>

namespace spice::hipx
{
struct csr_dev
{
	dbuf<std::int64_t> row_ptr; // N + 1
	dbuf<std::int32_t> col;     // E
	std::int64_t n = 0, nnz = 0, width = 0;
};

// Builds the CSR adjacency for the flattened connections on the device:
// degrees (one thread per row) -> exclusive scan -> targets (one wave per row).
csr_dev generate_csr( std::vector<util::gen::conn> const & conns, std::int64_t n, std::int64_t width, std::uint64_t seed, hipStream_t s );

// Uploads a host CSR adjacency.
csr_dev upload_csr( std::int64_t const * row_ptr, std::int32_t const * col, std::int64_t n, std::int64_t width, hipStream_t s );

// Exclusive scan of n int32 values into n + 1 int64 values (out[n] = total).
void exclusive_scan_i32_to_i64( std::int32_t const * in, std::int64_t * out, std::int64_t n, hipStream_t s );

// Generates the graph into caller-owned device memory: row_ptr (n + 1 entries) always, col only
// if it is non-null and holds the graph (nnz <= col_cap). conn_scratch holds conns.size()
// entries, tile_scratch graph_scratch_tiles(n). Synchronous on s; returns nnz.
std::int64_t generate_into( std::vector<util::gen::conn> const & conns, std::int64_t n, std::int64_t width, std::uint64_t seed,
                            std::int64_t * row_ptr, std::int32_t * col, std::uint64_t col_cap, util::gen::conn * conn_scratch,
                            std::int64_t * tile_scratch, hipStream_t s );
std::int64_t graph_scratch_tiles( std::int64_t n );

// A preallocated generator scratch region: GRAPH_SCRATCH_BYTES of device memory, the first
// GRAPH_SCRATCH_CONN_BYTES for the connection table, the rest for the scan's tile sums.
constexpr std::size_t GRAPH_SCRATCH_BYTES = std::size_t( 1 ) << 20;
constexpr std::size_t GRAPH_SCRATCH_CONN_BYTES = std::size_t( 1024 ) * sizeof( util::gen::conn );
} // namespace spice::hipx
