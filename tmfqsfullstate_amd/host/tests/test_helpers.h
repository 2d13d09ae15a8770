// Support for the C++ test drivers: a driver registers named cases and runs
// them in order.  A failed CHECK records where and what and lets the case go
// on, so one run lists every failure; an exception escaping a case fails that
// case with its message.  The exit status is non-zero when anything failed.
#pragma once

#include <cmath>
#include <functional>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "tmfqsfs.h"

namespace hipcheck {

struct Outcome {
	int checks = 0;
	std::vector<std::string> failures;
};

inline Outcome &outcome() {
	static Outcome o;
	return o;
}

inline void record(bool ok, const char *what, const char *file, int line) {
	Outcome &o = outcome();
	++o.checks;
	if(!ok) o.failures.push_back(std::string(file) + ":" + std::to_string(line) + ": " + what);
}

inline bool near(double a, double b, double tol = 1e-10) { return std::fabs(a - b) < tol; }

// True when fn throws exactly an Exc (or a class derived from it).
template <typename Exc, typename Fn>
bool raises(Fn &&fn) {
	try {
		fn();
	} catch(const Exc &) {
		return true;
	} catch(...) {
		return false;
	}
	return false;
}

inline tmfqs::RegisterConfig deviceConfig() {
	tmfqs::RegisterConfig cfg;
	cfg.strategy = tmfqs::StorageStrategyKind::Hip;
	return cfg;
}

// Every amplitude and the total probability of a and b within tol.
inline bool sameState(const tmfqs::QuantumRegister &a, const tmfqs::QuantumRegister &b, double tol) {
	if(a.stateCount() != b.stateCount()) return false;
	const tmfqs::AmplitudesVector x = a.amplitudes(), y = b.amplitudes();
	for(size_t i = 0; i < x.size(); ++i)
		if(!near(x[i], y[i], tol)) return false;
	return near(a.totalProbability(), b.totalProbability(), tol);
}

class Suite {
	public:
		explicit Suite(std::string title) : title_(std::move(title)) {}

		Suite &add(const char *name, std::function<void()> fn) {
			cases_.emplace_back(name, std::move(fn));
			return *this;
		}

		int run() {
			for(auto &[name, fn] : cases_) {
				const size_t before = outcome().failures.size();
				try {
					fn();
				} catch(const std::exception &e) {
					outcome().failures.push_back(name + ": uncaught exception: " + e.what());
				}
				const bool ok = outcome().failures.size() == before;
				std::cout << (ok ? "  ok    " : "  FAIL  ") << name << "\n";
			}
			for(const std::string &f : outcome().failures) std::cerr << "FAILED " << f << "\n";
			std::cout << title_ << ": " << cases_.size() << " cases, " << outcome().checks << " checks, "
			          << outcome().failures.size() << " failures\n";
			return outcome().failures.empty() ? 0 : 1;
		}

	private:
		std::string title_;
		std::vector<std::pair<std::string, std::function<void()>>> cases_;
};

}  // namespace hipcheck

#define CHECK(expr) ::hipcheck::record(static_cast<bool>(expr), #expr, __FILE__, __LINE__)
