// Circuit IR and builders (reference: src/tmfqs/circuit/circuit.cpp).
#include "tmfqs/circuit/circuit.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <type_traits>

#include "tmfqs/core/constants.h"
#include "tmfqs/core/state_space.h"

namespace tmfqs {
namespace circuit {

Circuit::Circuit(unsigned int qubits) : qubits_(qubits) {}
unsigned int Circuit::qubitCount() const noexcept { return qubits_; }
const std::vector<Operation> &Circuit::operations() const noexcept { return operations_; }
size_t Circuit::operationCount() const noexcept { return operations_.size(); }

Circuit &Circuit::h(QubitIndex t) { return append(Hadamard{t}); }
Circuit &Circuit::x(QubitIndex t) { return append(PauliX{t}); }
Circuit &Circuit::cx(QubitIndex c, QubitIndex t) { return append(ControlledNot{c, t}); }
Circuit &Circuit::controlledPhase(QubitIndex c, QubitIndex t, double theta) { return append(ControlledPhaseShift{c, t, theta}); }
Circuit &Circuit::swap(QubitIndex a, QubitIndex b) { return append(Swap{a, b}); }
Circuit &Circuit::gate(QuantumGate g, QubitList targets) { return append(MatrixGate{std::move(g), std::move(targets)}); }
Circuit &Circuit::phaseFlip(StateIndex s) { return append(PhaseFlipBasisState{s}); }
Circuit &Circuit::inversionAboutMean(bool materialized) { return append(InversionAboutMean{materialized}); }

Circuit &Circuit::append(const Operation &op) {
	operations_.push_back(op);
	return *this;
}

Circuit &Circuit::append(Operation &&op) {
	operations_.push_back(std::move(op));
	return *this;
}

std::string operationName(const Operation &operation) {
	return std::visit(
	    [](const auto &op) -> std::string {
		    using Op = std::decay_t<decltype(op)>;
		    if constexpr(std::is_same<Op, Hadamard>::value) return "h";
		    if constexpr(std::is_same<Op, PauliX>::value) return "x";
		    if constexpr(std::is_same<Op, ControlledNot>::value) return "cx";
		    if constexpr(std::is_same<Op, ControlledPhaseShift>::value) return "controlled_phase";
		    if constexpr(std::is_same<Op, Swap>::value) return "swap";
		    if constexpr(std::is_same<Op, MatrixGate>::value) return "gate";
		    if constexpr(std::is_same<Op, PhaseFlipBasisState>::value) return "phase_flip_basis_state";
		    if constexpr(std::is_same<Op, InversionAboutMean>::value) return "inversion_about_mean";
		    return "unknown";
	    },
	    operation);
}

// reference: circuit.cpp:87-100: for each target, H then controlled phases
// pi/2^(control-target) from every later qubit; then mirror swaps.
void appendQft(Circuit &c) {
	const unsigned int n = c.qubitCount();
	for(unsigned int t = 0; t < n; ++t) {
		c.h(t);
		for(unsigned int ctl = t + 1u; ctl < n; ++ctl) c.controlledPhase(ctl, t, kPi / static_cast<double>(StateIndex{1u} << (ctl - t)));
	}
	for(unsigned int l = 0; l < n / 2u; ++l) c.swap(l, n - l - 1u);
}

Circuit makeQft(unsigned int qubits) {
	Circuit c(qubits);
	appendQft(c);
	return c;
}

// reference: circuit.cpp:108-116.
unsigned int groverIterationCount(StateIndex stateCount, size_t markedCount) {
	if(markedCount == 0u) throw std::invalid_argument("Grover circuit requires at least one marked state");
	const double ideal = (kPi / 4.0) * std::sqrt(static_cast<double>(stateCount) / static_cast<double>(markedCount));
	return static_cast<unsigned int>(std::floor(ideal));
}

// reference: circuit.cpp:118-141 (sorted unique marks; range check inside the iteration loop).
void appendGrover(Circuit &c, const GroverCircuitOptions &options) {
	std::vector<StateIndex> marked = options.markedStates.values();
	if(marked.empty()) throw std::invalid_argument("Grover circuit requires at least one marked state");
	std::sort(marked.begin(), marked.end());
	marked.erase(std::unique(marked.begin(), marked.end()), marked.end());
	for(unsigned int q = 0; q < c.qubitCount(); ++q) c.h(q);
	const StateIndex N = checkedStateCount(c.qubitCount());
	const unsigned int iterations = groverIterationCount(N, marked.size());
	for(unsigned int it = 0; it < iterations; ++it) {
		for(StateIndex s : marked) {
			if(s >= N) throw std::invalid_argument("Grover marked state is out of range");
			c.phaseFlip(s);
		}
		c.inversionAboutMean(options.materializedDiffusion);
	}
}

Circuit makeGrover(unsigned int qubits, GroverCircuitOptions options) {
	Circuit c(qubits);
	appendGrover(c, options);
	return c;
}

Circuit makeRingLayers(unsigned int qubits, unsigned int layers, unsigned int seed) {
	if(qubits < 2) throw std::invalid_argument("ring layers need at least two qubits");
	Circuit c(qubits);
	std::mt19937 gen(seed);
	std::uniform_real_distribution<double> dist(0.0, 2.0 * std::acos(-1.0));
	for(unsigned int l = 0; l < layers; ++l) {
		for(unsigned int q = 0; q < qubits; ++q) c.h(q);
		for(unsigned int i = 0; i < qubits; ++i) c.controlledPhase(i, (i + 1) % qubits, dist(gen));
	}
	return c;
}

}  // namespace circuit
}  // namespace tmfqs
