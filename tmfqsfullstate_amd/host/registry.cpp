// Strategy registry and factory (reference:
// src/tmfqs/storage/factory/storage_strategy_registry.cpp:13-115 and
// state_backend_factory.cpp).  This library ships the device backend only:
// Dense, Blosc and Zfp report unavailable and constructing them throws
// std::runtime_error exactly as an unavailable codec backend does in the
// reference.  Auto resolves to the device backend, the only one available in
// this build (the reference's size rule would otherwise pick Dense).
#include <stdexcept>

#include "tmfqs/core/state_space.h"
#include "tmfqs/storage/factory/state_backend_factory.h"
#include "tmfqs/storage/factory/storage_strategy_registry.h"
#include "tmfqs/storage/hip/hip_state_backend.h"

namespace tmfqs {

StorageStrategyKind StorageStrategyRegistry::resolve(unsigned int numQubits, const RegisterConfig &cfg) {
	if(cfg.strategy != StorageStrategyKind::Auto) return cfg.strategy;
	if(isHipStateBackendAvailable()) return StorageStrategyKind::Hip;
	const size_t bytes = checkedAmplitudeElementCount(numQubits) * sizeof(double);
	if(bytes < cfg.autoThresholdBytes) return StorageStrategyKind::Dense;
	if(isAvailable(StorageStrategyKind::Blosc)) return StorageStrategyKind::Blosc;
	if(isAvailable(StorageStrategyKind::Zfp)) return StorageStrategyKind::Zfp;
	return StorageStrategyKind::Dense;
}

bool StorageStrategyRegistry::isAvailable(StorageStrategyKind kind) {
	switch(kind) {
		case StorageStrategyKind::Hip:
		case StorageStrategyKind::HipChunked: return isHipStateBackendAvailable();
		case StorageStrategyKind::Auto: return true;
		default: return false;
	}
}

std::vector<std::string> StorageStrategyRegistry::listAvailable() {
	std::vector<std::string> names;
	if(isHipStateBackendAvailable()) {
		names.push_back("hip");
		names.push_back("hip-chunked");
	}
	names.push_back("auto");
	return names;
}

std::string StorageStrategyRegistry::toString(StorageStrategyKind kind) {
	switch(kind) {
		case StorageStrategyKind::Dense: return "dense";
		case StorageStrategyKind::Blosc: return "blosc";
		case StorageStrategyKind::Zfp: return "zfp";
		case StorageStrategyKind::Auto: return "auto";
		case StorageStrategyKind::Hip: return "hip";
		case StorageStrategyKind::HipChunked: return "hip-chunked";
	}
	return "unknown";
}

RegisterConfig StorageStrategyRegistry::tuneConfig(unsigned int, const RegisterConfig &cfg, StorageStrategyKind) { return cfg; }

BackendSelection StorageStrategyRegistry::createSelection(unsigned int numQubits, const RegisterConfig &cfg) {
	BackendSelection sel;
	sel.strategy = resolve(numQubits, cfg);
	sel.config = tuneConfig(numQubits, cfg, sel.strategy);
	sel.config.strategy = sel.strategy;
	switch(sel.strategy) {
		case StorageStrategyKind::Hip:
			sel.backend = createHipStateBackend(sel.config);
			if(!sel.backend) throw std::runtime_error("Hip backend requested but no HIP device is available");
			break;
		case StorageStrategyKind::HipChunked:
			sel.backend = createHipChunkedStateBackend(sel.config);
			if(!sel.backend) throw std::runtime_error("HipChunked backend requested but no HIP device is available");
			break;
		case StorageStrategyKind::Auto: throw std::runtime_error("StorageStrategyRegistry internal error: unresolved auto strategy");
		default:
			throw std::runtime_error(toString(sel.strategy) +
			                         " backend requested but unavailable in this build (use StorageStrategyKind::Hip)");
	}
	return sel;
}

std::unique_ptr<IStateBackend> StateBackendFactory::create(unsigned int n, const RegisterConfig &cfg) { return createSelection(n, cfg).backend; }
BackendSelection StateBackendFactory::createSelection(unsigned int n, const RegisterConfig &cfg) {
	return StorageStrategyRegistry::createSelection(n, cfg);
}

}  // namespace tmfqs
