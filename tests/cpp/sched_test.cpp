// Host-only test of the serving scheduler (spice_amd/csrc/scheduler.h) with a fake executor
// that records what it is asked to do: FIFO admission under the memory budget, exact step
// counts in chunks of at most chunk_max steps, failures, cancellation, drain and concurrent
// submitters. Prints one line per check and exits non-zero on the first failure.
#include "scheduler.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

using namespace spice::serve;

#define CHECK( c )                                                                \
	do                                                                            \
	{                                                                             \
		if( !( c ) )                                                              \
		{                                                                         \
			std::printf( "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c );            \
			std::exit( 1 );                                                       \
		}                                                                         \
	} while( 0 )

struct fake_exec
{
	struct spec
	{
		std::uint64_t bytes = 0;
		bool fail = false;
		int step_us = 0; // simulated cost of one step
	};
	struct handle
	{
		fake_exec * ex;
		spec s;
		std::int64_t steps = 0;
		~handle()
		{
			std::lock_guard<std::mutex> g( ex->m );
			ex->live -= s.bytes;
		}
	};

	std::mutex m;
	std::uint64_t live = 0, peak = 0;
	std::int64_t max_chunk = 0, runs = 0;
	std::set<std::size_t> member_counts;

	std::unique_ptr<handle> create( spec const & s, job_record & r )
	{
		if( s.fail ) throw std::runtime_error( "forced creation failure" );
		std::lock_guard<std::mutex> g( m );
		live += s.bytes;
		peak = std::max( peak, live );
		r.bytes = s.bytes;
		return std::unique_ptr<handle>( new handle{ this, s, 0 } );
	}
	void run( std::vector<handle *> const & members, bool, std::int64_t n )
	{
		int us = 0;
		for( auto * h : members )
		{
			h->steps += n;
			us = std::max( us, h->s.step_us );
		}
		{
			std::lock_guard<std::mutex> g( m );
			max_chunk = std::max( max_chunk, n );
			runs++;
			member_counts.insert( members.size() );
		}
		if( us ) std::this_thread::sleep_for( std::chrono::microseconds( us * n ) );
	}
	void wait() {}
	void collect( handle & h, job_record & r, std::int64_t n )
	{
		r.spikes += std::uint64_t( n );
		r.digest += std::uint64_t( h.steps );
	}
	void finish( handle & h, job_record & r ) { r.events = std::uint64_t( h.steps ); }
};

static void test_fifo_budget()
{
	fake_exec ex;
	{
		scheduler<fake_exec> s( ex, 100, 64 );
		auto a = s.submit( { 60, false, 20 }, 500, 60 );
		auto b = s.submit( { 60, false, 0 }, 100, 60 ); // must wait for a
		auto c = s.submit( { 30, false, 0 }, 100, 30 ); // fits beside a, but may not overtake b
		job_record ra, rb, rc;
		CHECK( s.wait( a, -1, ra ) && s.wait( b, -1, rb ) && s.wait( c, -1, rc ) );
		CHECK( ra.status == JOB_DONE && rb.status == JOB_DONE && rc.status == JOB_DONE );
		CHECK( rb.t_admit >= ra.t_done );
		CHECK( rc.t_admit >= rb.t_admit );
		CHECK( ra.done == 500 && rb.done == 100 && rc.done == 100 );
		CHECK( ra.events == 500 && rb.events == 100 && rc.events == 100 ); // exact step counts
		CHECK( ra.spikes == 500 );
		CHECK( s.info().peak_reserved <= 100 );
		CHECK( s.info().reserved == 0 );
	}
	CHECK( ex.peak <= 100 && ex.live == 0 );
	CHECK( ex.max_chunk <= 64 );
	std::printf( "ok fifo_budget\n" );
}

static void test_reject_fail_cancel()
{
	fake_exec ex;
	scheduler<fake_exec> s( ex, 100, 50 );
	bool threw = false;
	try
	{
		s.submit( { 101, false, 0 }, 10, 101 );
	}
	catch( std::invalid_argument const & )
	{
		threw = true;
	}
	CHECK( threw );
	threw = false;
	try
	{
		s.submit( { 1, false, 0 }, 0, 1 );
	}
	catch( std::invalid_argument const & )
	{
		threw = true;
	}
	CHECK( threw );
	auto f = s.submit( { 40, true, 0 }, 10, 40 );
	job_record rf;
	CHECK( s.wait( f, -1, rf ) && rf.status == JOB_FAILED && rf.error == "forced creation failure" );
	auto run = s.submit( { 70, false, 100 }, 1000000, 70 );
	auto queued = s.submit( { 70, false, 0 }, 10, 70 );
	std::this_thread::sleep_for( std::chrono::milliseconds( 50 ) );
	s.cancel( queued );
	job_record rq;
	CHECK( s.wait( queued, 5.0, rq ) && rq.status == JOB_CANCELLED && rq.done == 0 );
	s.cancel( run );
	job_record rr;
	CHECK( s.wait( run, 30.0, rr ) && rr.status == JOB_CANCELLED && rr.done > 0 && rr.done < 1000000 && rr.done % 50 == 0 );
	s.drain();
	CHECK( s.info().reserved == 0 && s.info().resident == 0 );
	bool timed_out = false;
	auto slow = s.submit( { 10, false, 1000 }, 200, 10 );
	job_record rs;
	timed_out = !s.wait( slow, 0.001, rs );
	CHECK( timed_out && rs.status <= JOB_RUNNING );
	CHECK( s.wait( slow, -1, rs ) && rs.status == JOB_DONE && rs.done == 200 );
	std::printf( "ok reject_fail_cancel\n" );
}

static void test_concurrent_submitters()
{
	fake_exec ex;
	scheduler<fake_exec> s( ex, 1000, 32 );
	std::vector<std::uint64_t> ids[4];
	std::vector<std::thread> th;
	for( int t = 0; t < 4; t++ )
		th.emplace_back( [&, t] {
			for( int k = 0; k < 25; k++ ) ids[t].push_back( s.submit( { std::uint64_t( 10 + k ), false, 5 }, 7 + k * 13, 10 + k ) );
		} );
	for( auto & x : th ) x.join();
	s.drain();
	for( int t = 0; t < 4; t++ )
		for( int k = 0; k < 25; k++ )
		{
			job_record r;
			CHECK( s.wait( ids[t][k], 0.0, r ) && r.status == JOB_DONE && r.done == 7 + k * 13 && r.events == std::uint64_t( 7 + k * 13 ) );
		}
	CHECK( s.info().finished == 100 && s.info().reserved == 0 );
	CHECK( ex.member_counts.size() > 1 ); // jobs joined and left the running mix
	CHECK( *ex.member_counts.rbegin() > 1 );
	std::printf( "ok concurrent_submitters\n" );
}

static void test_destroy_cancels_outstanding()
{
	fake_exec ex;
	std::uint64_t id;
	{
		scheduler<fake_exec> s( ex, 100, 16 );
		id = s.submit( { 50, false, 200 }, 100000000, 50 );
		s.submit( { 60, false, 0 }, 10, 60 );
		std::this_thread::sleep_for( std::chrono::milliseconds( 20 ) );
	}
	CHECK( ex.live == 0 );
	(void)id;
	std::printf( "ok destroy_cancels_outstanding\n" );
}

int main()
{
	test_fifo_budget();
	test_reject_fail_cancel();
	test_concurrent_submitters();
	test_destroy_cancels_outstanding();
	std::printf( "ALL OK\n" );
	return 0;
}
