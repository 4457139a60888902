// Plaquette::CliffordCircuit<Precision> -- drop-in host class for the
// reference's plaquette_simulator/src/CliffordCircuit.hpp:16-151: reads the
// circuit text file (one operation per line, `NAME q...`), owns a
// CliffordState of (num_qubits, batch_size) with the default seed, and
// Execute() applies the operations in order.
//
// The reference issues one kernel launch per listed qubit (pair); here the
// whole file is parsed into one op list (pq_circuit_parse) and Execute() runs
// it as a program -- one persistent kernel by default, or one launch per op
// (PQ_RUN_PLAIN), or a hipGraph replay (PQ_RUN_GRAPH).  Results, records and
// record ids are those of the same operations issued one by one.
//
// Errors: an unsupported operation name (a blank line's empty name included)
// throws std::runtime_error("Unsupported operation: <name>") from the
// constructor, as the reference does (:44-46).  A file that cannot be opened
// reads as an empty circuit (the reference's std::ifstream fails silently).
// Deviation: qubit indices are checked when Execute() runs (the reference
// reads out of bounds, e.g. Test_CliffordCircuit.hpp:58 measures qubit 3 of 3)
// and an odd CX/CZ qubit count is rejected when the file is read.
#pragma once

#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "CliffordState.hpp"
#include "plaquette_amd.h"

namespace Plaquette {

template <typename Precision> class CliffordCircuit {
 public:
  CliffordCircuit(const std::string &file_name, std::size_t num_qubits, std::size_t batch_size)
      : state_(num_qubits, batch_size) {
    std::ifstream in(file_name);
    std::stringstream ss;
    if (in) ss << in.rdbuf();
    ReadOperations(ss.str());
  }

  auto &GetCliffordState() { return state_; }
  auto GetNumQubits() const { return state_.GetNumQubits(); }
  auto GetBatchSize() const { return state_.GetBatchSize(); }
  /// Lines read (CliffordCircuit.hpp:127-129).
  auto GetNumOperations() const { return num_lines_; }
  /// Engine ops the lines expand to (one per qubit, or per CX/CZ pair).
  const std::vector<pq_op> &GetOps() const { return ops_; }

  /// Apply every operation once (CliffordCircuit.hpp:131-135).  M lines
  /// append new records after those the state already holds.
  void Execute(int form = PQ_RUN_PERSISTENT) {
    if (ops_.empty()) return;
    std::vector<pq_op> ops = ops_;
    const std::size_t base = state_.GetNumMeasurements();
    for (pq_op &o : ops)
      if (o.code == PQ_OP_M) o.slot += static_cast<uint32_t>(base);
    state_.RunProgram(ops, form);
  }

  /// Wait for the state's queued work.
  void Synchronize() { detail::check(pq_synchronize(state_.Handle()), "synchronize"); }

  /// One line per operation: its name, then each qubit followed by a space.
  void Print() const {
    std::cout << "Circuit:" << std::endl;
    std::istringstream lines(text_);
    std::string line;
    while (std::getline(lines, line)) {
      std::istringstream ls(line);
      std::string name;
      ls >> name;
      std::cout << name << " ";
      std::size_t q = 0;
      while (ls >> q) std::cout << q << " ";
      std::cout << std::endl;
    }
  }

 private:
  void ReadOperations(const std::string &text) {
    uint64_t nops = 0, nlines = 0;
    if (pq_circuit_parse(text.data(), text.size(), 0, nullptr, 0, &nops, &nlines, nullptr) != PQ_OK)
      throw std::runtime_error(pq_last_error());
    ops_.resize(nops);
    if (pq_circuit_parse(text.data(), text.size(), 0, ops_.data(), ops_.size(), &nops, &nlines,
                         nullptr) != PQ_OK)
      throw std::runtime_error(pq_last_error());
    num_lines_ = nlines;
    text_ = text;
  }

  CliffordState<Precision> state_;
  std::vector<pq_op> ops_;
  std::size_t num_lines_ = 0;
  std::string text_;
};

}  // namespace Plaquette
