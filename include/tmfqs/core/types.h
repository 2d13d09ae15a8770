// Core value types of the public API (reference: include/tmfqs/core/types.h:13-156).
#ifndef TMFQS_CORE_TYPES_H
#define TMFQS_CORE_TYPES_H

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <utility>
#include <vector>

namespace tmfqs {

/** Index of a qubit; qubit 0 is the most significant bit of a basis index. */
using QubitIndex = unsigned int;
/** Index of a computational basis state (64-bit: registers beyond 31 qubits work). */
using StateIndex = std::uint64_t;

/** Complex amplitude. */
struct Amplitude {
	double real;
	double imag;
};

/** Interleaved amplitudes [re0, im0, re1, im1, ...]. */
using AmplitudesVector = std::vector<double>;

/** Ordered list of qubit targets. */
class QubitList {
	public:
		QubitList() = default;
		QubitList(std::initializer_list<QubitIndex> init) : values_(init) {}
		explicit QubitList(std::vector<QubitIndex> values) : values_(std::move(values)) {}
		size_t size() const noexcept { return values_.size(); }
		bool empty() const noexcept { return values_.empty(); }
		void reserve(size_t n) { values_.reserve(n); }
		void push_back(QubitIndex v) { values_.push_back(v); }
		QubitIndex operator[](size_t i) const { return values_[i]; }
		QubitIndex &operator[](size_t i) { return values_[i]; }
		std::vector<QubitIndex>::const_iterator begin() const noexcept { return values_.begin(); }
		std::vector<QubitIndex>::const_iterator end() const noexcept { return values_.end(); }
		std::vector<QubitIndex>::iterator begin() noexcept { return values_.begin(); }
		std::vector<QubitIndex>::iterator end() noexcept { return values_.end(); }
		const std::vector<QubitIndex> &values() const noexcept { return values_; }

	private:
		std::vector<QubitIndex> values_;
};

/** List of basis-state indices. */
class BasisStateList {
	public:
		BasisStateList() = default;
		BasisStateList(std::initializer_list<StateIndex> init) : values_(init) {}
		explicit BasisStateList(std::vector<StateIndex> values) : values_(std::move(values)) {}
		size_t size() const noexcept { return values_.size(); }
		bool empty() const noexcept { return values_.empty(); }
		void reserve(size_t n) { values_.reserve(n); }
		void push_back(StateIndex v) { values_.push_back(v); }
		StateIndex operator[](size_t i) const { return values_[i]; }
		StateIndex &operator[](size_t i) { return values_[i]; }
		std::vector<StateIndex>::const_iterator begin() const noexcept { return values_.begin(); }
		std::vector<StateIndex>::const_iterator end() const noexcept { return values_.end(); }
		std::vector<StateIndex>::iterator begin() noexcept { return values_.begin(); }
		std::vector<StateIndex>::iterator end() noexcept { return values_.end(); }
		const std::vector<StateIndex> &values() const noexcept { return values_; }

	private:
		std::vector<StateIndex> values_;
};

}  // namespace tmfqs

#endif
