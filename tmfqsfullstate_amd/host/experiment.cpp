// Comparison metrics and run reports (reference: src/tmfqs/experiment/comparison.cpp:12-99,
// report.cpp:39-95).  compareStates reads each state once in bulk instead of
// one virtual amplitude() call per basis state; the metrics and their
// accumulation order are the reference's.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "tmfqs/experiment/comparison.h"
#include "tmfqs/experiment/report.h"
#include "tmfqs/register/quantum_register.h"
#include "tmfqs/storage/factory/storage_strategy_registry.h"
#include "tmfqs/storage/hip/hip_state_backend.h"
#include "tmfqs_hip.h"

namespace tmfqs {
namespace experiment {

namespace {

// Both states are simulation results held on one device without a pending
// affine overlay: compare them there (tmfqs_hip_compare).
bool compareOnDevice(const state::QuantumState &reference, const state::QuantumState &candidate, StateComparison &m) {
	const QuantumRegister *a = reference.registerPtr();
	const QuantumRegister *b = candidate.registerPtr();
	if(!a || !b) return false;
	tmfqs_hip_state *ha = hipBackendHandle(a->backend());
	tmfqs_hip_state *hb = hipBackendHandle(b->backend());
	if(!ha || !hb || !a->overlayIsIdentity() || !b->overlayIsIdentity()) return false;
	tmfqs_hip_comparison c{};
	if(tmfqs_hip_compare(ha, hb, &c) != TMFQS_HIP_OK) return false;
	m.bitwiseEqual = c.bitwise_equal != 0;
	m.maxAbsAmplitudeError = c.max_abs_amplitude_error;
	m.maxAbsComponentError = c.max_abs_component_error;
	m.rmseAmplitude = c.rmse_amplitude;
	m.relL2 = c.rel_l2;
	m.maxAbsProbabilityError = c.max_abs_probability_error;
	m.totalProbabilityDiff = c.total_probability_diff;
	m.worstState = c.worst_state;
	return true;
}

}  // namespace

StateComparison compareStates(const state::QuantumState &reference, const state::QuantumState &candidate) {
	if(reference.stateCount() != candidate.stateCount()) throw std::invalid_argument("Compared states must have the same state count");
	StateComparison dm;
	if(compareOnDevice(reference, candidate, dm)) return dm;
	const AmplitudesVector ra = reference.amplitudes();
	const AmplitudesVector ca = candidate.amplitudes();
	StateComparison m;
	m.bitwiseEqual = true;
	double diffNormSq = 0.0, refNormSq = 0.0, refTotal = 0.0, gotTotal = 0.0;
	const StateIndex N = reference.stateCount();
	for(StateIndex s = 0; s < N; ++s) {
		const double rr = ra[2 * s], ri = ra[2 * s + 1], gr = ca[2 * s], gi = ca[2 * s + 1];
		if(rr != gr || ri != gi) m.bitwiseEqual = false;
		const double dr = gr - rr, di = gi - ri;
		const double d2 = dr * dr + di * di;
		const double ae = std::sqrt(d2);
		if(ae > m.maxAbsAmplitudeError) {
			m.maxAbsAmplitudeError = ae;
			m.worstState = s;
		}
		m.maxAbsComponentError = std::max(m.maxAbsComponentError, std::max(std::abs(dr), std::abs(di)));
		diffNormSq += d2;
		refNormSq += rr * rr + ri * ri;
		const double rp = rr * rr + ri * ri, gp = gr * gr + gi * gi;
		refTotal += rp;
		gotTotal += gp;
		m.maxAbsProbabilityError = std::max(m.maxAbsProbabilityError, std::abs(gp - rp));
	}
	m.rmseAmplitude = std::sqrt(diffNormSq / static_cast<double>(N));
	m.relL2 = refNormSq == 0.0 ? 0.0 : std::sqrt(diffNormSq / refNormSq);
	m.totalProbabilityDiff = std::abs(gotTotal - refTotal);
	return m;
}

std::string toCsvRow(const std::string &circuit, const std::string &scenario, unsigned int qubits, const StateComparison &c) {
	std::ostringstream o;
	o << std::setprecision(17) << circuit << ',' << scenario << ',' << qubits << ',' << (c.bitwiseEqual ? "true" : "false") << ','
	  << c.maxAbsAmplitudeError << ',' << c.maxAbsComponentError << ',' << c.relL2 << ',' << c.rmseAmplitude << ','
	  << c.maxAbsProbabilityError << ',' << c.totalProbabilityDiff << ',' << c.worstState;
	return o.str();
}

std::string toJson(const StateComparison &c) {
	std::ostringstream o;
	o << std::setprecision(17) << "{\"bitwiseEqual\":" << (c.bitwiseEqual ? "true" : "false") << ",\"maxAbsAmplitudeError\":" << c.maxAbsAmplitudeError
	  << ",\"maxAbsComponentError\":" << c.maxAbsComponentError << ",\"relL2\":" << c.relL2 << ",\"rmseAmplitude\":" << c.rmseAmplitude
	  << ",\"maxAbsProbabilityError\":" << c.maxAbsProbabilityError << ",\"totalProbabilityDiff\":" << c.totalProbabilityDiff
	  << ",\"worstState\":" << c.worstState << "}";
	return o.str();
}

namespace {

std::string jsonEscape(const std::string &v) {
	std::string o;
	for(char ch : v) {
		switch(ch) {
			case '\\': o += "\\\\"; break;
			case '"': o += "\\\""; break;
			case '\n': o += "\\n"; break;
			case '\r': o += "\\r"; break;
			case '\t': o += "\\t"; break;
			default: o += ch;
		}
	}
	return o;
}

void writeText(const std::string &path, const std::string &content) {
	std::ofstream out(path);
	if(!out) throw std::runtime_error("Failed to open report output: " + path);
	out << content;
}

}  // namespace

std::string toJson(const RunReport &r) {
	std::ostringstream o;
	o << std::setprecision(17) << "{\n  \"qubits\": " << r.qubits << ",\n  \"operation_count\": " << r.operationCount
	  << ",\n  \"strategy\": \"" << StorageStrategyRegistry::toString(r.strategy) << "\",\n  \"execution_seconds\": " << r.executionSeconds
	  << ",\n  \"operations\": [";
	for(size_t i = 0; i < r.operations.size(); ++i) {
		const auto &op = r.operations[i];
		o << (i == 0 ? "\n" : ",\n") << "    {\"index\": " << op.index << ", \"name\": \"" << jsonEscape(op.name) << "\", \"seconds\": " << op.seconds
		  << "}";
	}
	if(!r.operations.empty()) o << "\n  ";
	o << "]\n}\n";
	return o.str();
}

std::string toCsv(const RunReport &r) {
	std::ostringstream o;
	o << std::setprecision(17) << "qubits,strategy,total_operations,total_seconds,operation_index,operation_name,operation_seconds\n";
	const std::string head = std::to_string(r.qubits) + "," + StorageStrategyRegistry::toString(r.strategy) + "," + std::to_string(r.operationCount) + ",";
	if(r.operations.empty()) {
		o << head << r.executionSeconds << ",,,\n";
		return o.str();
	}
	for(const auto &op : r.operations) o << head << r.executionSeconds << ',' << op.index << ',' << op.name << ',' << op.seconds << '\n';
	return o.str();
}

void writeJsonReport(const RunReport &r, const std::string &path) { writeText(path, toJson(r)); }
void writeCsvReport(const RunReport &r, const std::string &path) { writeText(path, toCsv(r)); }

}  // namespace experiment
}  // namespace tmfqs
