#pragma once
// HYBRID marks functions compiled for both host and device. Model functors and the
// deterministic generator primitives are HYBRID so that the host build (CPU checker,
// host graph builder) and the gfx950 build share one definition
// (reference counterpart: spice/util/host_defines.h:3-7).
#if defined( __HIPCC__ )
#define HYBRID __host__ __device__
#define SPICE_DEVICE_CODE 1
#else
#define HYBRID
#endif
