// Convenience wrappers (reference: include/tmfqs/storage/state_backend_factory.h).
#ifndef TMFQS_STORAGE_STATE_BACKEND_FACTORY_H
#define TMFQS_STORAGE_STATE_BACKEND_FACTORY_H

#include "tmfqs/storage/factory/state_backend_factory.h"
#include "tmfqs/storage/factory/storage_strategy_registry.h"

namespace tmfqs {

inline StorageStrategyKind resolveStorageStrategy(unsigned int n, const RegisterConfig &cfg) { return StorageStrategyRegistry::resolve(n, cfg); }
inline BackendSelection createBackendSelection(unsigned int n, const RegisterConfig &cfg) { return StateBackendFactory::createSelection(n, cfg); }
inline std::unique_ptr<IStateBackend> createBackend(unsigned int n, const RegisterConfig &cfg) { return StateBackendFactory::create(n, cfg); }
inline bool isStrategyAvailable(StorageStrategyKind kind) { return StorageStrategyRegistry::isAvailable(kind); }
inline std::vector<std::string> listAvailableStrategies() { return StorageStrategyRegistry::listAvailable(); }
inline std::string storageStrategyToString(StorageStrategyKind kind) { return StorageStrategyRegistry::toString(kind); }

}  // namespace tmfqs

#endif
