#pragma once
// Abstract network interface shared by every backend: the reference's spice::snn<Model>
// (spice/snn.h:16-45, spice/snn.cpp:29-44). Owns dt, delay, the step counter and the
// Kahan-summed simulated time; _step() hands the backend the step index and the compensated
// dt. Protected accessors expose the clock so a backend can be constructed from another
// network (conversion constructor).
#include <spice/snn_info.h>
#include <spice/util/numeric.h>
#include <spice/util/stdint.h>

#include <functional>
#include <limits>
#include <stdexcept>
#include <utility>
#include <vector>

namespace spice
{
template <typename Model>
class snn
{
public:
	virtual ~snn() = default;

	virtual void step( std::vector<int> * out_spikes = nullptr ) = 0;

	virtual size_ num_neurons() const = 0;
	virtual size_ num_synapses() const = 0;
	float dt() const { return _dt; }
	int_ delay() const { return _delay; }
	snn_info info() const { return { static_cast<int_>( num_neurons() ) }; }

	// (ELL edges padded with INT_MAX, width)
	virtual std::pair<std::vector<int>, size_> adj() const = 0;
	virtual std::vector<typename Model::neuron::tuple_t> neurons() const = 0;
	virtual std::vector<typename Model::synapse::tuple_t> synapses() const = 0;

	// Clock of the base class: steps done, simulated time, Kahan compensation.
	long_ step_index() const { return _i; }
	float simtime() const { return float( _simtime ); }
	float simtime_compensation() const { return _simtime.compensation(); }

protected:
	explicit snn( float dt, int_ delay = 1 )
	    : _dt( dt )
	    , _delay( delay )
	{
		if( !( dt >= 0.0f ) ) throw std::invalid_argument( "snn: dt must be >= 0" );
		if( delay < 1 ) throw std::invalid_argument( "snn: delay must be >= 1" );
	}

	void _step( std::function<void( int_, float )> const & impl )
	{
		if( _i >= std::numeric_limits<int_>::max() ) throw std::invalid_argument( "snn: step counter overflow" );
		impl( int_( _i++ ), _simtime.add( _dt ) );
	}

	// Advances the base clock by n steps without calling a backend (multi-step runs).
	void _advance( long_ n )
	{
		for( long_ k = 0; k < n; k++ )
		{
			_i++;
			_simtime.add( _dt );
		}
	}

	void _set_clock( long_ i, float sum, float c )
	{
		_i = i;
		_simtime.set( sum, c );
	}

private:
	float _dt;
	int_ _delay;
	long_ _i = 0;
	util::kahan_sum<float> _simtime;
};
} // namespace spice
