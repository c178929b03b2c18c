#pragma once
// Device-side view of one network: plain pointers and sizes passed by value to every
// kernel. All step-dependent values (step index, Kahan clock, ring slot) are read from
// device memory (ctl), so the kernels of a step take identical arguments every step and a
// captured HIP graph can be replayed.
//
// HBM layout of one net:
//   row_ptr  int64[N+1], col int32[E]           CSR adjacency, rows sorted by target
//   na[a]    attribute a of every neuron          structure of arrays
//   acc[a]   receive accumulator of attribute a   fp32 / int64 (fixed point) / int32
//   sa[a]    attribute a of every synapse         CSR order (plastic models only)
//   ring     int32[(D+1) * N], counts uint32[D+1] spike ids of the last D+1 steps
//   hist     uint64[N]   bit b = spiked b steps ago     (plastic models)
//   synced   int32[N]    row s is up to date through this step
//   pend     int32[N]    row s has an undelivered pre event at synced+1
//   hasp     uint8[N]    row s has at least one plastic synapse
//   flush_*  int32[N]    rows whose history window is full this step, with their
//                        (synced, pend) at that time; counters double-buffered by step parity
#include <cstdint>

namespace spice::hipx
{
constexpr int MAXA = 16;

struct ctl_t
{
	std::int64_t step; // index of the step being executed / next to execute
	float sum;         // Kahan sum of dt over completed steps
	float c;           // Kahan compensation
	std::uint32_t done;
	std::uint32_t pad;
};

struct stats_t
{
	unsigned long long steps, spikes, events, plastic_events, flush_rows, flush_synapses;
};

struct dview
{
	std::int32_t N, D, width, shift;
	float DT;
	std::int32_t parts;  // work items per delivered spike row (receive)
	std::int32_t fparts; // work items per flushed row
	std::uint32_t flags;
	std::uint64_t seed;

	std::int64_t const * row_ptr;
	std::int32_t const * col;

	void * na[MAXA];
	void * acc[MAXA];
	void * sa[MAXA];

	std::int32_t * ring;
	std::uint32_t * counts;

	unsigned long long * hist;
	std::int32_t * synced;
	std::int32_t * pend;
	std::uint8_t * hasp;
	std::int32_t * fl_id;
	std::int32_t * fl_synced;
	std::int32_t * fl_pend;
	std::uint32_t * fl_count; // [2]

	ctl_t * ctl;
	stats_t * stats;
};
} // namespace spice::hipx
