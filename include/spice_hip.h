/*
 * spice_hip.h - C ABI of the MI355X (gfx950) single-GPU spiking-network simulator.
 *
 * This is the drop-in boundary beside the reference's cuda::snn<Model>. The reference's
 * backend is a C++ template (spice/cuda/snn.h:16-77) that drives a templated kernel layer
 * (spice/cuda/algorithm.h:15-72) explicitly instantiated per model
 * (spice/cuda/algorithm.cu:490-747). A C ABI cannot carry templates, so the model is a
 * run-time id, the network is an opaque handle that owns every device buffer, and all
 * arguments are plain pointers and sizes. include/spice/hip/snn.h wraps this ABI back into
 * the reference's C++ API (spice::snn<Model>).
 *
 * Errors: every function returns a spice_status. On failure spice_last_error() returns a
 * thread-local message; the C++ wrapper rethrows the reference's exception types
 * (std::invalid_argument, std::bad_alloc, std::runtime_error, std::bad_cast).
 *
 * Threading: a net is not thread-safe; different nets may be driven from different host
 * threads. All work of one net goes to the net's own HIP stream and is asynchronous unless
 * stated otherwise.
 */
#ifndef SPICE_HIP_H
#define SPICE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Models compiled into the library. 0-3 are the reference's models
 * (spice/models/{vogels_abbott,brunel,brunel_with_plasticity,synth}.h); 4-5 are the
 * three-population Brunel variants of the benchmark configurations; 6 is model 5 without its
 * promise of refractory plastic endpoints (the same dynamics, run on the two-kernel step whose
 * deliveries apply the previous step's events of the rows they read). */
enum spice_model_id
{
	SPICE_MODEL_VOGELS_ABBOTT = 0,
	SPICE_MODEL_BRUNEL = 1,
	SPICE_MODEL_BRUNEL_WITH_PLASTICITY = 2,
	SPICE_MODEL_SYNTH = 3,
	SPICE_MODEL_BRUNEL_3POP = 4,
	SPICE_MODEL_BRUNEL_3POP_STDP = 5,
	SPICE_MODEL_BRUNEL_3POP_STDP_DEFERRED = 6,
	SPICE_NUM_MODELS = 7
};

enum spice_status
{
	SPICE_OK = 0,
	SPICE_ERR_INVALID_ARGUMENT = 1, /* reference: spice_assert -> std::invalid_argument (spice/util/assert.cpp:22-35) */
	SPICE_ERR_BAD_ALLOC = 2,        /* reference: cuda_malloc -> std::bad_alloc (spice/cuda/util/memory.cpp:9-15) */
	SPICE_ERR_RUNTIME = 3,          /* reference: success_or_throw -> std::runtime_error (spice/cuda/util/error.h:19-53) */
	SPICE_ERR_BAD_CAST = 4          /* reference: narrow -> std::bad_cast (spice/util/type_traits.h:115) */
};

enum spice_flag
{
	/* Sum receive() contributions per target attribute as int64 fixed point (units of
	 * 2^-40) instead of fp32 atomics. Integer addition is associative, so results do not
	 * depend on atomic order and match the CPU checker's fixed-point mode bit for bit. */
	SPICE_FLAG_FIXED_POINT = 1u << 0
	/* other bits are rejected (SPICE_ERR_INVALID_ARGUMENT) */
};

/* Attribute type codes reported by spice_model_info. */
enum spice_attr_type
{
	SPICE_ATTR_F32 = 0,
	SPICE_ATTR_I32 = 1
};

typedef struct spice_net spice_net;

/* (src population, dst population, connection probability): one entry of the reference's
 * layout connectivity list (spice/util/layout.h:31-33). */
typedef struct
{
	uint64_t src_pop;
	uint64_t dst_pop;
	float p;
} spice_connection;

/* Flattened connection ([src_first, src_last), [dst_first, dst_last), p): the reference's
 * layout::edge (spice/util/layout.h:27-28). */
typedef struct
{
	int32_t src_first, src_last, dst_first, dst_last;
	float p;
} spice_edge;

/* One network of a batch (or a single network with per-net model parameters): layout as
 * (pops, conns), its delay, its seed and up to 4 model parameters (snn_info::param; 0 = the
 * model's default, e.g. brunel_3pop's Poisson rate in Hz). */
typedef struct
{
	const uint64_t * pops;
	int32_t npops;
	const spice_connection * conns;
	int32_t nconns;
	int32_t delay;
	uint64_t seed;
	float params[4];
} spice_net_desc;

typedef struct spice_batch spice_batch;

/* Counters accumulated on the device since creation (or the last reset). */
typedef struct
{
	uint64_t steps;          /* completed steps */
	uint64_t spikes;         /* spikes emitted */
	uint64_t events;         /* synaptic events delivered (real edges of delivered spikes) */
	uint64_t plastic_events; /* plastic synapses changed when brought up to date on delivery */
	/* Plasticity passes. Lazy (per-synapse contract): rows flushed (spike arrived or history
	 * window full), plastic synapses changed, row slots scanned. Eager (neuron-trace
	 * synapses): pre-pass rows + post-pass targets, synapses changed, 4-byte index words read
	 * (one per pre-pass slot, two per post-pass entry). */
	uint64_t flush_rows;
	uint64_t flush_synapses;
	uint64_t flush_slots;
	/* plastic synapses the passes wrote back: eager, only those whose state the event changed
	 * (an unchanged synapse is read, not written); lazy, every replayed synapse */
	uint64_t flush_written;
} spice_stats;

/* Thread-local message of the last failed call on this thread ("" if none). */
const char * spice_last_error( void );

/* Library build string (compiler, target, flags). */
const char * spice_version( void );

/* ---- host utilities ------------------------------------------------------------------- */

/* Validates and flattens a population/connection description exactly as the reference's
 * layout constructor does (spice/util/layout.cpp:51-110). out_edges must hold nconns
 * entries. Replaces util::layout::{layout, connections, size, max_degree}. */
int spice_layout_flatten(
    const uint64_t * pops,
    int32_t npops,
    const spice_connection * conns,
    int32_t nconns,
    spice_edge * out_edges,
    int32_t * out_nedges,
    uint64_t * out_num_neurons,
    uint64_t * out_max_degree );

/* Number and types of a model's neuron/synapse attributes (the reference's
 * Model::neuron::tuple_t / Model::synapse::tuple_t, spice/util/meta.h:130-171). The type
 * arrays must hold 16 entries. */
int spice_model_info(
    int32_t model,
    int32_t * n_neuron_attrs,
    int32_t * n_synapse_attrs,
    int32_t * neuron_attr_types,
    int32_t * synapse_attr_types );

/* ---- construction --------------------------------------------------------------------- */

/* Replaces cuda::snn<Model>::snn(layout const&, float dt, int delay) (spice/cuda/snn.h:20-25,
 * spice/cuda/snn.d.cpp:58-87): builds the random graph on the device (Philox generator keyed
 * by `seed`), allocates and initialises state. The graph is CSR with rows clipped to the
 * layout's max_degree, so it exports to the reference's ELL format unchanged. */
int spice_hip_create(
    int32_t model,
    const uint64_t * pops,
    int32_t npops,
    const spice_connection * conns,
    int32_t nconns,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    spice_net ** out_net );

/* spice_hip_create from a descriptor (adds per-net model parameters). */
int spice_hip_create_desc( int32_t model, const spice_net_desc * desc, float dt, uint32_t flags, spice_net ** out_net );

/* Replaces cuda::snn<Model>::snn(std::vector<int> const& adj, size_t width, float dt, int delay)
 * (spice/cuda/snn.h:26-32, spice/cuda/snn.d.cpp:89-121): imports an ELL adjacency
 * (num_neurons x width, each row sorted, padded with INT_MAX). width must be a multiple of 32. */
int spice_hip_create_from_ell(
    int32_t model,
    const int32_t * edges,
    uint64_t num_neurons,
    uint64_t width,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    spice_net ** out_net );

/* Same as above from a CSR adjacency (row_ptr has num_neurons + 1 entries); `width` is the
 * ELL width reported by num_synapses()/adj() and must be >= every row length. */
int spice_hip_create_from_csr(
    int32_t model,
    const int64_t * row_ptr,
    const int32_t * col,
    uint64_t num_neurons,
    uint64_t width,
    float dt,
    int32_t delay,
    uint64_t seed,
    uint32_t flags,
    spice_net ** out_net );

/* Replaces the reference's kernel-layer generate_adj_list (spice/cuda/algorithm.h:15-20,
 * spice/cuda/algorithm.cu:397-443): generates the random graph of a layout (the generator of
 * spice_hip_create, same seed -> same graph) into caller-owned DEVICE memory, as CSR:
 * d_row_ptr receives num_neurons + 1 int64 offsets; d_col (NULL: offsets only) receives the
 * nnz int32 targets if nnz <= col_cap (elements), else the call fails with
 * SPICE_ERR_INVALID_ARGUMENT after setting *out_nnz (so a caller can size col and call again).
 * *out_width: the layout's ELL width. The generator's scratch (connection table, scan partial
 * sums) is allocated for the call. Synchronous; runs on the null stream. */
int spice_hip_generate_graph(
    const uint64_t * pops,
    int32_t npops,
    const spice_connection * conns,
    int32_t nconns,
    uint64_t seed,
    int64_t * d_row_ptr,
    int32_t * d_col,
    uint64_t col_cap,
    uint64_t * out_nnz,
    uint64_t * out_width );

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results and errors as spice_hip_generate_graph, but the generator's scratch is a fixed
 * 1 MiB region in the library's own device memory (one per GPU, held until the process ends)
 * instead of memory allocated for the call; the call allocates no device memory. Limits set
 * by that region: at most 1024 flattened connections and 130,023,424 neurons (larger inputs
 * fail with SPICE_ERR_INVALID_ARGUMENT). All calls on one GPU share the region: calls that
 * overlap on one GPU (from several host threads) race and must be serialised by the caller.
 * Synchronous; runs on the null stream. */
int spice_hip_generate_graph_noalloc(
    const uint64_t * pops,
    int32_t npops,
    const spice_connection * conns,
    int32_t nconns,
    uint64_t seed,
    int64_t * d_row_ptr,
    int32_t * d_col,
    uint64_t col_cap,
    uint64_t * out_nnz,
    uint64_t * out_width );

/* Releases every device buffer of the net (the reference's dbuffer RAII,
 * spice/cuda/util/dbuffer.h:17-90). */
int spice_hip_destroy( spice_net * net );

/* ---- simulation ------------------------------------------------------------------------ */

/* Replaces cuda::snn<Model>::step(int** out_dspikes, uint** out_dnum_spikes, ...)
 * (spice/cuda/snn.h:36-37, spice/cuda/snn.d.cpp:233-289): enqueues one step. The returned
 * device pointers (either may be NULL) address this step's spike ids and spike count; they
 * stay valid until the next step is enqueued. Asynchronous. */
int spice_hip_step( spice_net * net, int32_t ** out_dspikes, uint32_t ** out_dcount );

/* Copies the spike ids of the most recent step to host memory (synchronous), as
 * cuda::snn<Model>::step(std::vector<int>*) does (spice/cuda/snn.d.cpp:280-287).
 * Spike order is unspecified (atomic append), as in the reference. */
int spice_hip_copy_spikes( spice_net * net, int32_t * out, uint64_t cap, uint64_t * out_count );

/* Enqueues nsteps steps without spike readback, replaying a captured HIP graph of the
 * step sequence. Equivalent to nsteps calls of spice_hip_step. Asynchronous. */
int spice_hip_run( spice_net * net, int64_t nsteps );

/* Enqueues nsteps steps without graph capture, each kernel launched with start and stop HIP
 * events attached to its dispatch (hipExtLaunchKernel: the kernel's own duration, without the
 * launch gap) on the net's stream, waits, and adds the summed per-kernel device times in
 * milliseconds to kernel_ms[0..2]. Two-kernel steps: kernel_ms[0] spike delivery (for plastic
 * models including the previous step's plasticity work, which runs in the same kernel),
 * kernel_ms[1] neuron update, kernel_ms[2] 0. Overlapped steps (models whose delivery may run
 * before the target's update, delay >= 2): kernel_ms[0] the first step's delivery, kernel_ms[1]
 * the last step's update, kernel_ms[2] the nsteps - 1 kernels that each run the update of a
 * step beside the delivery of the next. Fused steps (static models whose steps do not overlap,
 * launch shapes with one delivery workgroup per tile of at most four workgroup widths, every
 * tile receiving input): kernel_ms[0] = kernel_ms[1] = 0, kernel_ms[2] the nsteps kernels that
 * each deliver a step's input to a tile and update the tile's neurons. Same results as
 * spice_hip_run. Used for roofline figures. */
int spice_hip_run_profiled( spice_net * net, int64_t nsteps, double * kernel_ms );

/* Waits for all work of the net. */
int spice_hip_synchronize( spice_net * net );

/* ---- batches: many independent networks stepped together -------------------------------- */
/* No counterpart in the reference (it steps one network per cuda::snn). A batch holds nnets
 * networks of one model (sizes, delays, seeds and parameters may differ; dt and flags are
 * shared) and steps all of them with one launch per step phase (a captured HIP graph).
 * Each network's results equal those of the same network stepped alone. */
int spice_batch_create( int32_t model, int32_t nnets, const spice_net_desc * descs, float dt, uint32_t flags, spice_batch ** out_batch );
int spice_batch_destroy( spice_batch * batch );
/* Enqueues nsteps steps of every network in the batch (asynchronous). */
int spice_batch_run( spice_batch * batch, int64_t nsteps );
int spice_batch_synchronize( spice_batch * batch );
int spice_batch_size( spice_batch * batch, int32_t * out_nnets );
/* Borrowed handle of network i, valid while the batch lives: every query / export / spike
 * readback of this header works on it (they wait for the batch first); stepping it alone,
 * setting its clock or destroying it is rejected. */
int spice_batch_net( spice_batch * batch, int32_t i, spice_net ** out_net );

/* ---- launch shapes and the autotuner --------------------------------------------------- */
/* No counterpart in the reference (its launch shapes are compile-time constants). tile_w:
 * targets per spike-delivery tile (multiple of 64, 64..1024; up to 8192 for models without
 * plasticity that run two kernels per step -- delay 1, or receive() reading target state without
 * a split gate); flush_slots: synapses per
 * plasticity work item (a lazy flush's row slice, or an eager pass's row or target slice;
 * multiple of 64); flush_grid: plasticity workgroups (of 256 threads' worth);
 * recv_block: threads per delivery workgroup. A zero field selects the default. Launch shapes
 * never change results. */
typedef struct spice_launch
{
	int32_t tile_w;
	int32_t flush_slots;
	int32_t flush_grid;
	int32_t recv_block; /* delivery workgroup size: 256, 512 or 1024 threads */
	int32_t recv_parts; /* delivery workgroups per tile, 1..16 (the step's spikes split between them) */
	int32_t reserved[3];
} spice_launch;

int spice_hip_get_launch( spice_net * net, spice_launch * out );
int spice_hip_set_launch( spice_net * net, const spice_launch * in );
/* Times candidate launch shapes on a copy of the network (warmup steps first, then `steps`
 * graph-replayed steps per candidate, wall time per step: plasticity overlaps delivery, so
 * shapes are evaluated together; tile width first, then delivery parts, then the plasticity
 * shape; the three fastest are then timed again in turn, three times each, since the copy's
 * activity drifts between candidates),
 * applies the fastest to the network and returns it with its step time (us). The network
 * itself does not advance. Results are cached per (model, neurons, edges, delay, flags) for
 * the life of the process. */
int spice_hip_autotune( spice_net * net, int64_t warmup, int64_t steps, spice_launch * out, double * out_us_per_step );

/* ---- serving: a request loop over device-resident simulation jobs ------------------------ */
/* No counterpart in the reference. A server admits submitted jobs in arrival order while the
 * sum of their device-memory reservations fits its budget, keeps each admitted job's graph
 * and state on the device until it has run its steps, and steps all resident jobs together
 * in chunks (one batched launch sequence per model and accumulation mode, the models on
 * their own streams); jobs arriving while others run join at the next chunk. After every
 * chunk the spikes of each job are handed to the sink, on the server's worker thread:
 * offsets[k] .. offsets[k+1] index the ids of step first_step + k (nsteps + 1 offsets);
 * the arrays are valid during the call only. A job's spikes and statistics equal those of
 * the same network run alone (bit-identical with SPICE_FLAG_FIXED_POINT). */
typedef struct spice_server spice_server;
typedef void ( *spice_spike_sink )(
    void * user, uint64_t job, int64_t first_step, int64_t nsteps, const uint64_t * offsets, const int32_t * ids );

typedef struct spice_job_desc
{
	int32_t model;      /* spice_model_id */
	spice_net_desc net; /* populations, connections, delay, seed, params (read at submit) */
	float dt;
	uint32_t flags;
	int64_t steps; /* duration in steps (>= 1) */
} spice_job_desc;

typedef enum spice_job_status
{
	SPICE_JOB_QUEUED = 0,
	SPICE_JOB_RUNNING = 1,
	SPICE_JOB_DONE = 2,
	SPICE_JOB_FAILED = 3,
	SPICE_JOB_CANCELLED = 4
} spice_job_status;

typedef struct spice_job_result
{
	uint64_t id;
	int32_t status; /* spice_job_status */
	int32_t reserved0;
	int64_t steps, done; /* requested / simulated */
	uint64_t synapses, device_bytes, reserved_bytes;
	uint64_t spikes, events, plastic_events;
	/* wrapping sum over the streamed spikes of mix64(step << 32 | id), mix64 = the splitmix64
	 * finalizer: equal digests <=> (with overwhelming probability) equal spike streams */
	uint64_t digest;
	double t_submit, t_admit, t_done; /* seconds since the server started */
} spice_job_result;

typedef struct spice_server_info
{
	uint64_t budget, reserved, peak_reserved;
	int64_t queued, resident, finished, chunks, rebuilds, steps_launched;
	double busy_s, now;
	/* the worker's wall time by cycle phase (seconds): enqueueing chunks (batch rebuilds and
	 * graph captures included), building admitted jobs, waiting for chunks, and collecting
	 * spikes and retiring jobs (the device has no chunk in flight during the last) */
	double launch_s, admit_s, wait_s, collect_s;
} spice_server_info;

/* budget_bytes: device-memory admission budget; chunk_steps: longest chunk (also the spike
 * log window). sink may be NULL. */
int spice_server_create( uint64_t budget_bytes, int64_t chunk_steps, spice_spike_sink sink, void * user, spice_server ** out );
/* Cancels queued and running jobs, waits for the worker and frees everything. */
int spice_server_destroy( spice_server * server );
/* Device-memory reservation of a job for a server with chunks of chunk_steps steps. */
int spice_server_estimate( const spice_job_desc * job, int64_t chunk_steps, uint64_t * out_bytes );
/* Rejects (SPICE_ERR_INVALID_ARGUMENT) a job whose reservation exceeds the budget. */
int spice_server_submit( spice_server * server, const spice_job_desc * job, uint64_t * out_id );
/* Waits up to timeout_s seconds (negative: no limit) for the job to finish; *out_finished
 * tells whether it did, *out gets its record either way. */
int spice_server_wait( spice_server * server, uint64_t id, double timeout_s, spice_job_result * out, int32_t * out_finished );
int spice_server_cancel( spice_server * server, uint64_t id );
/* Waits until no job is queued or resident. */
int spice_server_drain( spice_server * server );
int spice_server_get_info( spice_server * server, spice_server_info * out );
/* Error message of a failed job (empty otherwise), NUL-terminated, truncated to cap. */
int spice_server_job_error( spice_server * server, uint64_t id, char * buf, uint64_t cap );

/* ---- queries ---------------------------------------------------------------------------- */

/* num_neurons() / num_synapses() (= num_neurons x ELL width, the reference's count,
 * spice/util/adj_list.cpp:97) / ELL width / number of real edges / dt() / delay()
 * (spice/snn.h:24-28). */
int spice_hip_num_neurons( spice_net * net, uint64_t * out );
int spice_hip_num_synapses( spice_net * net, uint64_t * out );
int spice_hip_width( spice_net * net, uint64_t * out );
int spice_hip_num_edges( spice_net * net, uint64_t * out );
int spice_hip_dt( spice_net * net, float * out );
int spice_hip_delay( spice_net * net, int32_t * out );
int spice_hip_model( spice_net * net, int32_t * out );

/* Step counter and Kahan-summed simulated time (sum, compensation) of the base class
 * spice::snn (spice/snn.h:40-44, spice/util/numeric.h:8-33). */
int spice_hip_get_clock( spice_net * net, int64_t * step, float * sum, float * c );
int spice_hip_set_clock( spice_net * net, int64_t step, float sum, float c );

int spice_hip_get_stats( spice_net * net, spice_stats * out );
int spice_hip_reset_stats( spice_net * net );

/* No counterpart in the reference: a measurement of the device, not of a network. Times
 * `accesses` 4-byte accesses at pseudo-random places of a fresh `bytes`-byte device array (a
 * load each, and with rmw != 0 the store of the changed value after it: the plasticity post
 * pass's access to a weight), averaged over `reps` launches after one untimed launch; returns
 * the microseconds per launch. Allocates and frees the array (synchronous). The rate of random
 * 64-byte memory lines it yields bounds the kernels whose traffic is such lines. */
int spice_hip_probe_random_lines( uint64_t bytes, uint64_t accesses, int32_t rmw, int32_t reps, double * out_us );

/* ---- state export / import (synchronous) ------------------------------------------------ */

/* adj(): ELL edges (num_neurons x width, INT_MAX-padded) (spice/snn.h:32,
 * spice/cuda/snn.d.cpp:308-312). `cap` is in elements. */
int spice_hip_export_adj( spice_net * net, int32_t * out, uint64_t cap );

/* CSR adjacency: row_ptr (num_neurons + 1) and col (num_edges). */
int spice_hip_export_csr( spice_net * net, int64_t * row_ptr, int32_t * col );

/* Rows [row_first, row_last) of the CSR adjacency, for graphs too large to copy whole
 * (1e10 synapses are 40 GB): row_ptr receives the row_last - row_first + 1 absolute offsets
 * row_ptr[row_first] .. row_ptr[row_last]; col (NULL: offsets only) receives the
 * row_ptr[row_last] - row_ptr[row_first] target ids of those rows, col_cap is its size in
 * elements (SPICE_ERR_INVALID_ARGUMENT if too small). Synchronous. No counterpart in the
 * reference, whose adj() copies the whole ELL array (spice/cuda/snn.d.cpp:308-312). */
int spice_hip_export_csr_rows(
    spice_net * net, uint64_t row_first, uint64_t row_last, int64_t * row_ptr, int32_t * col, uint64_t col_cap );

/* neurons(): one attribute at a time, num_neurons elements (spice/snn.h:33,
 * spice/util/meta.h:212-233). `bytes` must equal num_neurons * attribute size. */
int spice_hip_export_neuron_attr( spice_net * net, int32_t attr, void * out, uint64_t bytes );
int spice_hip_import_neuron_attr( spice_net * net, int32_t attr, const void * in, uint64_t bytes );

/* synapses(): one attribute at a time. ell_order != 0: num_neurons x width slots in the
 * reference's ELL slot order (src * width + j, zero padding) (spice/snn.h:34); ell_order == 0:
 * num_edges entries in CSR order. Plastic synapses are brought up to date first. */
int spice_hip_export_synapse_attr( spice_net * net, int32_t attr, void * out, uint64_t bytes, int32_t ell_order );
int spice_hip_import_synapse_attr( spice_net * net, int32_t attr, const void * in, uint64_t bytes, int32_t ell_order );

#ifdef __cplusplus
}
#endif

#endif /* SPICE_HIP_H */
