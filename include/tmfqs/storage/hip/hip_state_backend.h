// Device storage backend: an IStateBackend over the C-ABI of libtmfqs_hip.so
// (include/tmfqs_hip.h).  Follows the reference's backend factory pattern
// (src/tmfqs/storage/blosc/blosc_state_backend.cpp:10-21).
#ifndef TMFQS_STORAGE_HIP_HIP_STATE_BACKEND_H
#define TMFQS_STORAGE_HIP_HIP_STATE_BACKEND_H

#include <memory>

#include "tmfqs/config/register_config.h"
#include "tmfqs/storage/i_state_backend.h"

struct tmfqs_hip_state;
struct tmfqs_hip_config;

namespace tmfqs {

/** True when a HIP device is present. */
bool isHipStateBackendAvailable();
/** A device backend, or nullptr when no HIP device is present. */
std::unique_ptr<IStateBackend> createHipStateBackend(const RegisterConfig &cfg);
/** The C-ABI handle behind a backend created by createHipStateBackend (nullptr otherwise). */
tmfqs_hip_state *hipBackendHandle(IStateBackend *backend);
/** Throws the exception a tmfqs_hip_* status stands for (nothing for TMFQS_HIP_OK). */
void raiseHipStatus(int status);
/** The C-ABI configuration of HipConfig's device settings. */
tmfqs_hip_config toHipConfig(const HipConfig &c);
/** A chunked register in host memory behind the same interface
 *  (StorageStrategyKind::HipChunked), or nullptr when no HIP device is present. */
std::unique_ptr<IStateBackend> createHipChunkedStateBackend(const RegisterConfig &cfg);

}  // namespace tmfqs

#endif
