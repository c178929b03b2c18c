#pragma once
// Device-side view of one network: plain pointers and sizes passed by value to every
// kernel. All step-dependent values (step index, Kahan clock, ring slot) are read from
// device memory (ctl), so the kernels of a step take identical arguments every step and a
// captured HIP graph can be replayed.
//
// HBM layout of one net:
//   row_ptr  int64[N+1], col int32[E]           CSR adjacency, rows sorted by target
//   na[a]    attribute a of every neuron          structure of arrays
//   acc[a]   receive accumulators of attribute a  [2][N] by step parity: fp32 / int64 (fixed
//                        point) / int32
//   sa[a]    attribute a of every synapse         CSR order (plastic models only)
//   ring     int32[(D+2) * N], counts uint32[D+2] spike ids of the last D+2 steps (entry t mod
//                        (D+2): step T's update clears step T+1's entry while step T's
//                        delivery may still read step T-D's)
//   hring    uint64[HWORDS * N]  spike-history ring: bit (t mod H) of word (t mod H)/64 of
//                        neuron i set iff i spiked at step t (H = 256 steps, plastic models)
//   hasp     uint8[N]    row s has at least one plastic synapse
//   pl_lo/hi int32[N]    first / one-past-last plastic slot of row s (offsets in the row)
//   pivots   int32[N * (ntiles+1)]  per row and target tile: offset of the tile's first target
//   tile_ids int32[nact] the tiles that some row has targets in (at least one tile)
// lazy plasticity (per-synapse contract):
//   last_spk int32[N]    latest spike step
//   synced   int32[N]    row s is up to date through this step
//   pend     int32[N]    row s has an undelivered pre event at synced+1
//   fl_*     int32[2][N] rows queued at this step (spike arrived, or history window full),
//                        with their (synced, pend) at that time; lists by step parity,
//                        counters [4] by step mod 4
// eager plasticity (neuron-trace synapses), all [2][N] arrays by step parity:
//   xtr/ytr  f32[2][N]   neuron-local pre/post traces after step T
//   fl8      u8[2][N]    bit 0: spiked at T, bit 1: its spike arrives at T, bit 2: at T + 1
//                        (bit 2 only for models whose deliveries apply deferred events)
//   fl_id    int32[2][N] rows whose spike arrived at T (and, with deferred events, does not
//                        arrive at T + 1)
//   po_id    int32[2][N] neurons that spiked at T and have incoming plastic synapses
//   tp_ptr   int64[N+1], tp_ent {src, offset in src's row}[P]: the P plastic synapses by target
// both:
//   caught   int64       step whose plasticity work (and, eager, deferred events of the rows
//                        arriving next) a synchronization already did
//   slog_ids int32[cap]  spike log (served jobs): ids of steps t0 .. t0+nsteps-1, step by step
//   slog_off u64[nsteps] offset of step t0+k in slog_ids
#include <cstdint>

namespace spice::hipx
{
constexpr int MAXA = 16;

// Device clock. Four slots: the kernels of step T read slot T & 3 and the neuron-update
// kernel writes step T + 1's clock into slot (T + 1) & 3, so the clock advances without
// atomics or an extra kernel, the plasticity work of step T (which runs in step T + 1's
// delivery kernel) still reads its own slot (ctl_prev), and a captured graph of a multiple of
// 4 steps replays.
constexpr int CLOCK_SLOTS = 4;
struct ctl_t
{
	std::int64_t step; // index of the step being executed
	float sum;         // Kahan sum of dt over completed steps
	float c;           // Kahan compensation
	std::int32_t rslot; // spike ring entry of this step: step mod (D + 2)
};

// Device counters live in per-block slots: block b of any step kernel adds its block-reduced
// counts to bstats[b * NSTAT + k] with a plain read-add-write (no other block touches the
// slot, kernels of a stream never overlap), so no same-address atomics are needed. The host
// sums the slots when it reads statistics.
enum stat_index
{
	ST_SPIKES = 0,
	ST_EVENTS,
	ST_PLASTIC,
	ST_FLUSH_ROWS,
	ST_FLUSH_SYN,
	ST_FLUSH_SLOTS,
	ST_FLUSH_WRITTEN, // plastic synapses whose state the plasticity passes wrote back
	NSTAT = 8
};

// Spike log window of a net (device memory, set by the host between runs): steps
// t0 .. t0 + nsteps - 1 append their spike ids to the log, at most cap ids in all.
struct slog_t
{
	std::int64_t t0;
	std::int64_t nsteps;
	std::uint64_t cap;
};

// One plastic synapse of the by-target index: its source row and its offset in that row
// (offset first: read as one little-endian 64-bit word the source is the high half, so the
// word orders entries by source).
struct alignas( 8 ) tp_entry
{
	std::int32_t off;
	std::int32_t src;
};

struct dview
{
	std::int32_t N, D, width, shift;
	float DT;
	std::int32_t fparts; // work items per flushed row
	std::int32_t pparts; // work items per spiking target (eager plasticity)
	std::int32_t rparts; // delivery workgroups per tile (the step's spike list split between them)
	std::int32_t overlap; // 1: the update of step T runs beside the delivery of T + 1 (it writes
	                      // the spike log's step offsets then; otherwise the delivery does)
	std::int32_t upd_slot0; // first statistics slot of the update blocks of an overlapped kernel
	std::int32_t fused;     // 1: one kernel per step, each tile updated by its delivery workgroup
	                        // (the update then writes the spike log's step offsets)
	std::uint32_t flags;
	std::uint64_t seed;
	float param[4]; // per-net model parameters (snn_info::param)

	std::int64_t const * row_ptr;
	std::int32_t const * col;

	void * na[MAXA];
	void * acc[MAXA];
	void * sa[MAXA];

	std::int32_t * ring;
	std::uint32_t * counts;

	unsigned long long * hring; // [HWORDS][N] spike-history ring
	std::int32_t * last_spk;    // [N] latest spike step (-1: none)
	std::int32_t * synced;
	std::int32_t * pend;
	std::uint8_t * hasp;
	std::int32_t * pl_lo; // plastic slots of row r lie in [row_ptr[r] + pl_lo[r], row_ptr[r] + pl_hi[r])
	std::int32_t * pl_hi;
	std::int32_t * fl_id;
	std::int32_t * fl_synced;
	std::int32_t * fl_pend;
	std::uint32_t * fl_count; // [4]
	float * xtr;
	float * ytr;
	std::uint8_t * fl8;
	std::int32_t * po_id;
	std::uint32_t * po_count; // [4]
	std::int64_t const * tp_ptr;
	tp_entry const * tp_ent;
	std::int64_t * caught;

	ctl_t * ctl;      // this step's clock
	ctl_t * ctl_next; // the next step's clock (written by k_update)
	ctl_t * ctl_prev; // the previous step's clock (its dt for the lazy flush of that step)
	unsigned long long * bstats; // [nblocks_max * NSTAT]

	std::int32_t * pivots; // [N][ntiles + 1]: offset in row r of its first target >= t * tile_w
	std::int32_t ntiles, tile_w;
	std::int32_t * tile_ids; // [nact]: the tiles with incoming synapses (the delivery's tiles)
	std::int32_t nact;

	slog_t * slog; // nullptr: no spike log
	std::int32_t * slog_ids;
	std::uint64_t * slog_off;
};
} // namespace spice::hipx
