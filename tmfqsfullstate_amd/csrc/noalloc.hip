// tmfqs_hip_expectation_device: the expectation value of a diagonal
// observable over a state in the caller's own device memory, enqueued on the
// caller's stream, with no device or pinned allocation and no
// synchronisation -- for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// It runs the same kernels as tmfqs_hip_expectation (k_observable, then the
// fixed-order k_finish), so both give the same bits.  Their only scratch is
// one FP64 partial sum per workgroup of k_observable, at most kReduceBlocks
// of them: here a file-scope device array (8 KiB per GPU, held until the
// process ends).  Each workgroup writes its own partial before k_finish reads
// it, so nothing has to start at zero.  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU (on different streams, or from
// different host threads without ordering) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>

#include "engine.h"
#include "tmfqs_hip.h"

using namespace tmfqs_hip;

namespace {

constexpr unsigned kScratchDoubles = kReduceBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

// The array's device address on each GPU (each GPU holds its own copy),
// resolved once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

double *scratch_on_current_device(int device) {
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if(!g_scratch_ptr[device]) {
		void *p = nullptr;
		check_device(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		if(!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
		g_scratch_ptr[device] = static_cast<double *>(p);
	}
	return g_scratch_ptr[device];
}

}  // namespace

extern "C" int tmfqs_hip_expectation_device(int device, void *stream, int precision, const void *amps, unsigned n,
                                            const tmfqs_hip_diag_term *terms, size_t nterms, double *out) {
	try {
		if(precision != TMFQS_HIP_COMPLEX128 && precision != TMFQS_HIP_COMPLEX64)
			throw std::invalid_argument("tmfqs_hip_expectation_device: precision must be TMFQS_HIP_COMPLEX128 or TMFQS_HIP_COMPLEX64");
		if(!amps || !out) throw std::invalid_argument("tmfqs_hip_expectation_device: null state or output pointer");
		if(n > 63u) throw std::invalid_argument("Number of qubits exceeds supported range for state indexing");
		int count = 0;
		check_device(hipGetDeviceCount(&count), "hipGetDeviceCount");
		if(device < 0 || device >= count || device >= kMaxDevices)
			throw std::invalid_argument("tmfqs_hip_expectation_device: no such device");
		const ObsParams O = observable_params(n, terms, nterms);
		// k_observable runs at most kReduceBlocks workgroups, one partial each
		const unsigned long long N = 1ull << n;
		const unsigned long long blocks = (N + 255) / 256 < kReduceBlocks ? (N + 255) / 256 : kReduceBlocks;
		if(blocks > kScratchDoubles)
			throw std::invalid_argument("tmfqs_hip_expectation_device: reduction needs more partials than the scratch array holds");
		check_device(hipSetDevice(device), "hipSetDevice");
		double *part = scratch_on_current_device(device);
		check_device(launch_observable(Launch{static_cast<hipStream_t>(stream), precision}, amps, nullptr, n, O, part, out),
		             "k_observable");
		return TMFQS_HIP_OK;
	} catch(...) {
		return status_of_current_exception();
	}
}
