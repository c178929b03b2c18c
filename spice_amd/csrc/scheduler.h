#pragma once
// Continuous-batching job scheduler of the serving loop (no HIP dependency: the executor does
// all device work, so the policy is testable with a host-only executor).
//
// Jobs are admitted in arrival order (FIFO, no overtaking) while the sum of the admitted
// jobs' device-memory reservations stays within the budget; admitted jobs stay resident
// until they have run their steps. The worker thread repeats one cycle:
//   1. launch a chunk of steps for every resident job at once (the executor folds them into
//      batched launches), the chunk being the shortest remaining job capped at chunk_max,
//   2. while the chunk runs, build the device state of newly admitted jobs (they join the
//      next chunk),
//   3. wait for the chunk, hand every job's spikes of the chunk to the executor's sink,
//      retire finished or cancelled jobs and release their reservations.
// A job's steps are independent of which other jobs share its chunks (the executor
// guarantees per-job results equal to running the job alone).
//
// Executor interface (Exec):
//   using spec = ...; using handle = ...;
//   std::unique_ptr<handle> create( spec const &, job_record & );  // may throw; sets bytes
//   void run( std::vector<handle *> const & members, bool changed, std::int64_t nsteps ); // async
//   void wait();
//   void collect( handle &, job_record &, std::int64_t nsteps ); // after wait(): spikes -> sink
//   void finish( handle &, job_record & );                        // final counters
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace spice::serve
{
enum job_status : int
{
	JOB_QUEUED = 0,
	JOB_RUNNING = 1,
	JOB_DONE = 2,
	JOB_FAILED = 3,
	JOB_CANCELLED = 4
};

struct job_record
{
	std::uint64_t id = 0;
	int status = JOB_QUEUED;
	std::int64_t steps = 0;      // requested
	std::int64_t done = 0;       // simulated so far
	std::uint64_t est_bytes = 0; // admission reservation
	std::uint64_t bytes = 0;     // device bytes actually held
	std::uint64_t synapses = 0;
	std::uint64_t spikes = 0, events = 0, plastic_events = 0;
	std::uint64_t digest = 0; // order-independent digest of the streamed (step, id) pairs
	double t_submit = 0, t_admit = 0, t_done = 0;
	bool cancel = false;
	std::string error;
};

struct server_info
{
	std::uint64_t budget = 0, reserved = 0, peak_reserved = 0;
	std::int64_t queued = 0, resident = 0, finished = 0, chunks = 0, rebuilds = 0;
	std::int64_t steps_launched = 0; // sum over chunks of the chunk length
	double busy_s = 0;                // wall time with a chunk in flight
	// where the worker's time goes, per cycle phase: enqueueing the chunk (including batch
	// rebuilds and graph captures), building admitted jobs, waiting for the chunk, and
	// collecting its spikes and retiring jobs (the device idles in the last)
	double launch_s = 0, admit_s = 0, wait_s = 0, collect_s = 0;
};

template <typename Exec>
class scheduler
{
	using spec = typename Exec::spec;
	using handle = typename Exec::handle;

public:
	scheduler( Exec & ex, std::uint64_t budget, std::int64_t chunk_max )
	    : _ex( ex )
	    , _chunk_max( chunk_max )
	    , _t0( std::chrono::steady_clock::now() )
	{
		if( budget == 0 ) throw std::invalid_argument( "memory budget must be positive" );
		if( chunk_max < 1 ) throw std::invalid_argument( "chunk length must be positive" );
		_info.budget = budget;
		_thread = std::thread( [this] { loop(); } );
	}

	~scheduler()
	{
		{
			std::lock_guard<std::mutex> g( _m );
			_stop = true;
			for( auto & kv : _jobs )
				if( kv.second.status == JOB_QUEUED || kv.second.status == JOB_RUNNING ) kv.second.cancel = true;
		}
		_cv.notify_all();
		_thread.join();
	}

	scheduler( scheduler const & ) = delete;
	scheduler & operator=( scheduler const & ) = delete;

	double now() const { return std::chrono::duration<double>( std::chrono::steady_clock::now() - _t0 ).count(); }

	std::uint64_t submit( spec s, std::int64_t steps, std::uint64_t est_bytes )
	{
		if( steps < 1 ) throw std::invalid_argument( "a job runs at least one step" );
		std::lock_guard<std::mutex> g( _m );
		if( est_bytes > _info.budget ) throw std::invalid_argument( "job does not fit the device memory budget" );
		if( _stop ) throw std::logic_error( "server is shutting down" );
		std::uint64_t const id = ++_next_id;
		job_record & r = _jobs[id];
		r.id = id;
		r.steps = steps;
		r.est_bytes = est_bytes;
		r.t_submit = now();
		_queue.push_back( { id, std::move( s ) } );
		_info.queued++;
		_cv.notify_all();
		return id;
	}

	// Waits until job `id` has finished (done, failed or cancelled) or timeout_s passed
	// (negative: no limit). Returns whether it finished; `out` gets its record either way.
	bool wait( std::uint64_t id, double timeout_s, job_record & out )
	{
		std::unique_lock<std::mutex> g( _m );
		auto it = _jobs.find( id );
		if( it == _jobs.end() ) throw std::invalid_argument( "unknown job id" );
		auto fin = [&] { return it->second.status >= JOB_DONE; };
		if( timeout_s < 0 )
			_done_cv.wait( g, fin );
		else
			_done_cv.wait_for( g, std::chrono::duration<double>( timeout_s ), fin );
		out = it->second;
		return fin();
	}

	void cancel( std::uint64_t id )
	{
		std::lock_guard<std::mutex> g( _m );
		auto it = _jobs.find( id );
		if( it == _jobs.end() ) throw std::invalid_argument( "unknown job id" );
		job_record & r = it->second;
		if( r.status >= JOB_DONE ) return;
		r.cancel = true;
		if( r.status == JOB_QUEUED )
			for( auto q = _queue.begin(); q != _queue.end(); ++q )
				if( q->id == id )
				{
					_queue.erase( q );
					_info.queued--;
					_info.finished++;
					r.status = JOB_CANCELLED;
					r.t_done = now();
					_done_cv.notify_all();
					break;
				}
		_cv.notify_all();
	}

	void drain()
	{
		std::unique_lock<std::mutex> g( _m );
		_done_cv.wait( g, [&] { return _queue.empty() && _info.resident == 0 && _creating == 0; } );
	}

	server_info info() const
	{
		std::lock_guard<std::mutex> g( _m );
		return _info;
	}

	// Exception message of the worker thread if it died (empty otherwise).
	std::string fatal() const
	{
		std::lock_guard<std::mutex> g( _m );
		return _fatal;
	}

private:
	struct pending
	{
		std::uint64_t id;
		spec s;
	};
	struct resident
	{
		std::uint64_t id;
		std::unique_ptr<handle> h;
	};

	Exec & _ex;
	std::int64_t const _chunk_max;
	std::chrono::steady_clock::time_point const _t0;
	mutable std::mutex _m;
	std::condition_variable _cv, _done_cv;
	std::deque<pending> _queue;
	std::map<std::uint64_t, job_record> _jobs;
	std::vector<resident> _res; // worker thread only
	server_info _info;
	std::uint64_t _next_id = 0;
	std::int64_t _creating = 0; // admitted, device state being built
	bool _stop = false;
	std::string _fatal;
	std::thread _thread;

	// Under the lock: the queue head if it fits the budget (cancelled heads are retired).
	bool next_admission( pending & out )
	{
		while( !_queue.empty() )
		{
			job_record & r = _jobs[_queue.front().id];
			if( r.cancel )
			{
				r.status = JOB_CANCELLED;
				r.t_done = now();
				_queue.pop_front();
				_info.queued--;
				_info.finished++;
				_done_cv.notify_all();
				continue;
			}
			if( _info.reserved + r.est_bytes > _info.budget ) return false;
			_info.reserved += r.est_bytes;
			_info.peak_reserved = std::max( _info.peak_reserved, _info.reserved );
			out = std::move( _queue.front() );
			_queue.pop_front();
			_info.queued--;
			_creating++;
			return true;
		}
		return false;
	}

	// Builds the device state of every admissible queued job. Returns whether any joined.
	bool admit()
	{
		bool any = false;
		for( ;; )
		{
			pending p;
			job_record rec;
			{
				std::lock_guard<std::mutex> g( _m );
				if( !next_admission( p ) ) return any;
				rec = _jobs[p.id];
			}
			std::unique_ptr<handle> h;
			std::string err;
			try
			{
				h = _ex.create( p.s, rec );
			}
			catch( std::exception const & e )
			{
				err = e.what();
			}
			std::lock_guard<std::mutex> g( _m );
			_creating--;
			job_record & r = _jobs[p.id];
			if( !h )
			{
				r.status = JOB_FAILED;
				r.error = err.empty() ? "job creation failed" : err;
				r.t_done = now();
				_info.reserved -= r.est_bytes;
				_info.finished++;
				_done_cv.notify_all();
				continue;
			}
			r.bytes = rec.bytes;
			r.synapses = rec.synapses;
			if( r.bytes > r.est_bytes ) // keep the reservation an upper bound of what is held
			{
				_info.reserved += r.bytes - r.est_bytes;
				r.est_bytes = r.bytes;
				_info.peak_reserved = std::max( _info.peak_reserved, _info.reserved );
			}
			r.status = JOB_RUNNING;
			r.t_admit = now();
			_info.resident++;
			_res.push_back( { p.id, std::move( h ) } );
			any = true;
		}
	}

	void retire( std::size_t k, int status, std::string const & err )
	{
		std::lock_guard<std::mutex> g( _m );
		job_record & r = _jobs[_res[k].id];
		r.status = status;
		if( !err.empty() ) r.error = err;
		_info.reserved -= r.est_bytes;
		_info.resident--;
		_info.finished++;
		_res[k].h.reset(); // releases the job's device memory
		r.t_done = now();
		_done_cv.notify_all();
	}

	void loop()
	{
		try
		{
			run_loop();
		}
		catch( std::exception const & e )
		{
			std::lock_guard<std::mutex> g( _m );
			_fatal = e.what();
			for( auto & kv : _jobs )
				if( kv.second.status < JOB_DONE )
				{
					kv.second.status = JOB_FAILED;
					kv.second.error = std::string( "server failed: " ) + e.what();
				}
			_queue.clear();
			_info.queued = 0;
			_info.resident = 0;
			_res.clear();
			_done_cv.notify_all();
		}
	}

	void run_loop()
	{
		bool changed = true;
		for( ;; )
		{
			// chunk over the resident jobs: the shortest remaining job bounds it
			std::vector<handle *> members;
			std::vector<std::int64_t> remaining;
			std::int64_t chunk = _chunk_max;
			{
				std::lock_guard<std::mutex> g( _m );
				if( _stop && _queue.empty() && _res.empty() ) return;
				for( auto & x : _res )
				{
					job_record const & r = _jobs[x.id];
					chunk = std::min( chunk, r.steps - r.done );
				}
			}
			for( auto & x : _res ) members.push_back( x.h.get() );
			auto const tb = std::chrono::steady_clock::now();
			if( !members.empty() )
			{
				_ex.run( members, changed, chunk );
				changed = false;
			}
			auto const tl = std::chrono::steady_clock::now();
			// build newly admitted jobs while the chunk runs
			std::size_t const nrun = _res.size();
			if( admit() ) changed = true;
			auto const ta = std::chrono::steady_clock::now();
			if( members.empty() )
			{
				if( nrun == _res.size() && _res.empty() )
				{
					std::unique_lock<std::mutex> g( _m );
					_cv.wait_for( g, std::chrono::milliseconds( 50 ), [&] { return _stop || !_queue.empty(); } );
				}
				continue;
			}
			_ex.wait();
			auto const tw = std::chrono::steady_clock::now();
			{
				std::lock_guard<std::mutex> g( _m );
				_info.chunks++;
				_info.steps_launched += chunk;
				_info.busy_s += std::chrono::duration<double>( tw - tb ).count();
				_info.launch_s += std::chrono::duration<double>( tl - tb ).count();
				_info.admit_s += std::chrono::duration<double>( ta - tl ).count();
				_info.wait_s += std::chrono::duration<double>( tw - ta ).count();
			}
			// spikes of the chunk, then retirements (only the jobs that ran in it)
			std::vector<std::size_t> gone;
			for( std::size_t k = 0; k < nrun; k++ )
			{
				job_record snap;
				{
					std::lock_guard<std::mutex> g( _m );
					snap = _jobs[_res[k].id];
				}
				std::string err;
				try
				{
					_ex.collect( *_res[k].h, snap, chunk );
				}
				catch( std::exception const & e )
				{
					err = e.what();
				}
				snap.done += chunk;
				bool const fin = snap.done >= snap.steps;
				if( err.empty() && ( fin || snap.cancel ) )
				{
					try
					{
						_ex.finish( *_res[k].h, snap );
					}
					catch( std::exception const & e )
					{
						err = e.what();
					}
				}
				bool cancel = false;
				{
					std::lock_guard<std::mutex> g( _m );
					job_record & r = _jobs[_res[k].id];
					r.done = snap.done;
					r.spikes = snap.spikes;
					r.events = snap.events;
					r.plastic_events = snap.plastic_events;
					r.digest = snap.digest;
					cancel = r.cancel;
				}
				if( !err.empty() )
				{
					retire( k, JOB_FAILED, err );
					gone.push_back( k );
				}
				else if( fin )
				{
					retire( k, JOB_DONE, "" );
					gone.push_back( k );
				}
				else if( cancel )
				{
					retire( k, JOB_CANCELLED, "" );
					gone.push_back( k );
				}
			}
			if( !gone.empty() )
			{
				for( auto it = gone.rbegin(); it != gone.rend(); ++it ) _res.erase( _res.begin() + std::ptrdiff_t( *it ) );
				changed = true;
			}
			{
				std::lock_guard<std::mutex> g( _m );
				if( changed ) _info.rebuilds++;
				_info.collect_s += std::chrono::duration<double>( std::chrono::steady_clock::now() - tw ).count();
			}
		}
	}
};
} // namespace spice::serve
