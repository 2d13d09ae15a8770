// Gate matrices (reference: src/tmfqs/gates/quantum_gate.cpp).  Products are
// formed in FP64 on the host; the device receives the composed matrix.
#include "tmfqs/gates/quantum_gate.h"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "tmfqs/core/constants.h"
#include "tmfqs/core/math.h"

namespace tmfqs {

QuantumGate::QuantumGate(unsigned int dimension)
    : matrix_(static_cast<size_t>(dimension) * dimension, Amplitude{0.0, 0.0}), dimension_(dimension) {}

Amplitude *QuantumGate::operator[](unsigned int i) {
	if(i >= dimension_) throw std::out_of_range("QuantumGate row index out of range");
	return matrix_.data() + static_cast<size_t>(i) * dimension_;
}

const Amplitude *QuantumGate::operator[](unsigned int i) const {
	if(i >= dimension_) throw std::out_of_range("QuantumGate row index out of range");
	return matrix_.data() + static_cast<size_t>(i) * dimension_;
}

unsigned int QuantumGate::dimension() const noexcept { return dimension_; }

QuantumGate QuantumGate::operator*(Amplitude x) const {
	QuantumGate out(dimension_);
	for(size_t i = 0; i < matrix_.size(); ++i) out.matrix_[i] = amplitudeMultiply(matrix_[i], x);
	return out;
}

QuantumGate operator*(Amplitude x, const QuantumGate &U) { return U * x; }

// reference: quantum_gate.cpp:59-77 (i-k-j loop, accumulate left*right[k][j]).
QuantumGate QuantumGate::operator*(const QuantumGate &qg) const {
	if(qg.dimension_ != dimension_) throw std::invalid_argument("QuantumGate dimensions differ in multiplication");
	QuantumGate out(dimension_);
	const size_t d = dimension_;
	for(size_t i = 0; i < d; ++i)
		for(size_t k = 0; k < d; ++k) {
			const Amplitude l = matrix_[i * d + k];
			for(size_t j = 0; j < d; ++j)
				out.matrix_[i * d + j] = amplitudeAdd(out.matrix_[i * d + j], amplitudeMultiply(l, qg.matrix_[k * d + j]));
		}
	return out;
}

std::ostream &operator<<(std::ostream &os, const QuantumGate &qg) {
	for(unsigned int i = 0; i < qg.dimension(); ++i) {
		for(unsigned int j = 0; j < qg.dimension(); ++j) os << qg[i][j].real << " " << qg[i][j].imag << "\t";
		os << "\n";
	}
	return os;
}

void QuantumGate::printQuantumGate() const { std::cout << *this; }

QuantumGate QuantumGate::Identity(unsigned int d) {
	QuantumGate g(d);
	for(unsigned int i = 0; i < d; ++i) g[i][i].real = 1.0;
	return g;
}

QuantumGate QuantumGate::Hadamard() {
	QuantumGate g(2);
	const double c = 1.0 / std::sqrt(2.0);
	g[0][0].real = c;
	g[0][1].real = c;
	g[1][0].real = c;
	g[1][1].real = -c;
	return g;
}

QuantumGate QuantumGate::PauliX() {
	QuantumGate g(2);
	g[0][1].real = 1.0;
	g[1][0].real = 1.0;
	return g;
}

QuantumGate QuantumGate::PauliY() {
	QuantumGate g(2);
	g[0][1].imag = -1.0;
	g[1][0].imag = 1.0;
	return g;
}

QuantumGate QuantumGate::PauliZ() {
	QuantumGate g(2);
	g[0][0].real = 1.0;
	g[1][1].real = -1.0;
	return g;
}

QuantumGate QuantumGate::PhaseShift(double theta) {
	QuantumGate g(2);
	g[0][0].real = 1.0;
	const Amplitude p = complexExp({0.0, theta});
	g[1][1] = p;
	return g;
}

QuantumGate QuantumGate::PiOverEight() { return PhaseShift(kPi / 4.0); }

QuantumGate QuantumGate::ControlledNot() {
	QuantumGate g(4);
	g[0][0].real = 1.0;
	g[1][1].real = 1.0;
	g[2][3].real = 1.0;
	g[3][2].real = 1.0;
	return g;
}

QuantumGate QuantumGate::Toffoli() {
	QuantumGate g = Identity(8);
	g[6][6].real = 0.0;
	g[7][7].real = 0.0;
	g[6][7].real = 1.0;
	g[7][6].real = 1.0;
	return g;
}

QuantumGate QuantumGate::ControlledPhaseShift(double theta) {
	QuantumGate g(4);
	for(unsigned int i = 0; i < 3; ++i) g[i][i].real = 1.0;
	g[3][3] = complexExp({0.0, theta});
	return g;
}

QuantumGate QuantumGate::Swap() {
	QuantumGate g(4);
	g[0][0].real = 1.0;
	g[1][2].real = 1.0;
	g[2][1].real = 1.0;
	g[3][3].real = 1.0;
	return g;
}

QuantumGate QuantumGate::Ising(double theta) {
	QuantumGate g(4);
	double s, c;
	__builtin_sincos(theta / 2.0, &s, &c);  // one sincos call, as the reference's build emits
	for(unsigned int i = 0; i < 4; ++i) {
		g[i][i].real = c;
		g[i][3 - i].imag = -s;
	}
	return g;
}

}  // namespace tmfqs
