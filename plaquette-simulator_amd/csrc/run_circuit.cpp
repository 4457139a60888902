// run_circuit: time one circuit file on the GPU, as the reference's
// benchmarks/run_circuit.cpp:10-32 times CliffordCircuit::Execute.
//
//   run_circuit <file_name> <n_qubits> <batch_size> [plain|persistent|graph|frame]
//
// Prints "<file_name> <seconds>".  The timed region is Execute() plus the
// wait for its kernels (the form defaults to one persistent kernel; `plain`
// issues one launch per op like the reference; `frame` runs X/Z once on the
// host, which the ring circuits allow: they have no flagged gates).
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

#include "CliffordCircuit.hpp"

int main(int argc, char *argv[]) {
  using namespace Plaquette;
  if (argc != 4 && argc != 5) {
    std::cout << "Usage: " << argv[0] << " <file_name> <n_qubits> <batch_size> [plain|persistent|graph|frame]"
              << std::endl;
    return 1;
  }
  const std::string file_name = argv[1];
  const std::size_t num_qubits = std::stoul(argv[2]);
  const std::size_t batch_size = std::stoul(argv[3]);
  int form = PQ_RUN_PERSISTENT;
  if (argc == 5) {
    const std::string f = argv[4];
    if (f == "plain") form = PQ_RUN_PLAIN;
    else if (f == "graph") form = PQ_RUN_GRAPH;
    else if (f == "frame") form = PQ_RUN_FRAME;
    else if (f != "persistent") {
      std::cerr << "unknown form " << f << std::endl;
      return 1;
    }
  }
  try {
    CliffordCircuit<int> circuit(file_name, num_qubits, batch_size);
    circuit.Synchronize();
    const auto t1 = std::chrono::high_resolution_clock::now();
    circuit.Execute(form);
    circuit.Synchronize();
    const auto t2 = std::chrono::high_resolution_clock::now();
    std::cout << file_name << " " << std::chrono::duration<double>(t2 - t1).count() << std::endl;
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
