// Plaquette::CliffordState<Precision> -- drop-in host class for the reference's
// plaquette_simulator/src/CliffordState.hpp:24-689, backed by the gfx950
// bit-plane engine through the C ABI in plaquette_amd.h.
//
// Same constructors, method names, argument meaning and error behaviour.
// Where the reference exposes Kokkos types (GetX/GetZ/GetR return Views, the
// flagged gates take a View<int*>, Measurement holds device Views), this
// class exposes engine handles instead: DevicePlanes (the bit-plane tableau)
// and RecordPlane (one bit plane of a measurement record, dereferenceable
// from Measurement::measurement_results_device as in the reference's tests).
// Deviations, all deliberate:
//  * qubit indices are range-checked (the reference reads out of bounds);
//  * ResetCliffordState re-initialises the tableau (the reference's extent
//    {B, 2n+1, 2n+1} writes past the x/z views);
//  * re-measuring into an existing record refreshes its host copy (the
//    reference's CopyToHost keeps the first copy forever);
//  * CheckSum returns the exact total as `long` precision-independent.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "Error.hpp"
#include "plaquette_amd.h"

namespace Plaquette {

namespace detail {
inline void check(int rc, const char *what) {
  if (rc != PQ_OK) {
    std::string m = std::string(what) + ": " + pq_last_error();
    throw Exception("Error in Plaquette: " + m);
  }
}
}  // namespace detail

/// Bit-plane tableau handle (replaces the Kokkos Views of GetX/GetZ/GetR).
struct DevicePlanes {
  void *words = nullptr;  ///< device base, layout word[w][col][row]
  std::size_t num_qubits = 0, batch_size = 0, word_groups = 0, rows = 0;
};

/// One bit plane (W words) of a measurement record on the device.
struct RecordPlane {
  pq_state state = nullptr;
  std::size_t record = 0;
  int which = 0;  ///< 0 = results, 1 = determined
};

template <class Precision> class CliffordState {
 public:
  /// Measurement record (CliffordState.hpp:32-58).
  struct Measurement {
    std::size_t qubit_measured = 0;
    std::vector<Precision> measurement_results_host;
    std::vector<Precision> measurement_determined_host;
    std::unique_ptr<RecordPlane> measurement_results_device;
    std::unique_ptr<RecordPlane> measurement_determined_device;
    std::size_t batch_size = 0;

    void CopyToHost() {
      if (!measurement_results_host.empty() || batch_size == 0) return;
      std::vector<int32_t> res(batch_size), det(batch_size);
      detail::check(pq_get_record(measurement_results_device->state,
                                  measurement_results_device->record, res.data(),
                                  det.data()),
                    "CopyToHost");
      measurement_results_host.assign(res.begin(), res.end());
      measurement_determined_host.assign(det.begin(), det.end());
    }
  };

  CliffordState() = delete;

  /// CliffordState.hpp:97-139.
  CliffordState(std::size_t num_qubits, std::size_t batch_size, int seed = -1,
                bool init_to_zero_state = true) {
    detail::check(pq_create(num_qubits, batch_size, seed, init_to_zero_state ? 1 : 0,
                            &h_),
                  "CliffordState");
  }

  /// A sampling state (no reference counterpart): only the phase column per
  /// 64-shot group is kept; it runs frame-form programs from |0...0> and
  /// keeps their records (pq_create_sampling).
  struct Sampling {};
  CliffordState(Sampling, std::size_t num_qubits, std::size_t batch_size, int seed = -1) {
    detail::check(pq_create_sampling(num_qubits, batch_size, seed, &h_), "CliffordState(sampling)");
  }

  /// CliffordState.hpp:153-159.
  CliffordState(std::vector<Precision> &x, std::vector<Precision> &z,
                std::vector<Precision> &r, std::size_t num_qubits,
                std::size_t batch_size, long int seed = -1)
      : CliffordState(num_qubits, batch_size, (int)seed, false) {
    HostToDevice(x, z, r);
  }

  /// CliffordState.hpp:168-174: copies the tableau and the seed.
  CliffordState(const CliffordState &other) {
    detail::check(pq_clone(other.h_, &h_), "CliffordState(copy)");
  }
  CliffordState &operator=(const CliffordState &) = delete;

  ~CliffordState() {
    if (h_) pq_destroy(h_);
  }

  [[nodiscard]] DevicePlanes GetX() const { return planes(); }
  [[nodiscard]] DevicePlanes GetZ() const { return planes(); }
  [[nodiscard]] DevicePlanes GetR() const { return planes(); }
  [[nodiscard]] pq_state Handle() const { return h_; }

  /// CliffordState.hpp:207-231.
  void HostToDevice(std::vector<Precision> &x, std::vector<Precision> &z,
                    std::vector<Precision> &r) {
    const std::size_t n = GetNumQubits(), B = GetBatchSize(), R = 2 * n + 1;
    PLAQUETTE_ABORT_IF(x.size() != B * R * n || z.size() != B * R * n ||
                           r.size() != B * R,
                       "HostToDevice: vector sizes do not match the state");
    std::vector<int32_t> xi(x.begin(), x.end()), zi(z.begin(), z.end()),
        ri(r.begin(), r.end());
    detail::check(pq_host_to_device(h_, xi.data(), zi.data(), ri.data()), "HostToDevice");
  }

  /// CliffordState.hpp:238-259: flat [b][row][col] vectors.
  auto DeviceToHost() {
    const std::size_t n = GetNumQubits(), B = GetBatchSize(), R = 2 * n + 1;
    const bool empty = (n == 0 || B == 0);
    std::vector<int32_t> x(empty ? 0 : B * R * n), z(x.size()), r(empty ? 0 : B * R);
    detail::check(pq_device_to_host(h_, x.data(), z.data(), r.data()), "DeviceToHost");
    return std::make_tuple(std::vector<Precision>(x.begin(), x.end()),
                           std::vector<Precision>(z.begin(), z.end()),
                           std::vector<Precision>(r.begin(), r.end()));
  }

  /// CliffordState.hpp:268-274 (takes the source state).
  void DeviceToDevice(const CliffordState &other) {
    detail::check(pq_copy_from(h_, other.h_), "DeviceToDevice");
  }

  std::string GetKokkosBackend() const { return pq_backend_name(); }
  std::size_t GetNumQubits() const { return info().n; }
  std::size_t GetBatchSize() const { return info().B; }
  long int GetSeed() const { return (long int)info().seed; }

  void ResetCliffordState() { detail::check(pq_reset(h_), "ResetCliffordState"); }

  /// CliffordState.hpp:327-338.
  long CheckSum(int starting_batch_id = 0, int ending_batch_id = -1) {
    int64_t v = 0;
    detail::check(pq_checksum(h_, starting_batch_id, ending_batch_id, &v), "CheckSum");
    return (long)v;
  }

  // ---- flagged gates (CliffordState.hpp:348-372) ----
  void ApplyHadamardGateWithApplyFlag(std::size_t q, const RecordPlane &flag) {
    flagged(PQ_OP_H_FLAG, q, flag);
  }
  void ApplyPhaseGateWithApplyFlag(std::size_t q, const RecordPlane &flag) {
    flagged(PQ_OP_S_FLAG, q, flag);
  }
  void ApplyHadamardGateWithApplyFlag(std::size_t q, const std::vector<Precision> &flag) {
    flagged_host(PQ_OP_H_FLAG, q, flag);
  }
  void ApplyPhaseGateWithApplyFlag(std::size_t q, const std::vector<Precision> &flag) {
    flagged_host(PQ_OP_S_FLAG, q, flag);
  }
  /// Conditional Pauli (X, Y or Z code) on the shots a record flags.
  void ApplyPauliWithApplyFlag(uint32_t code, std::size_t q, const RecordPlane &flag) {
    flagged(code, q, flag);
  }
  void ApplyPauliXGateWithApplyFlag(std::size_t q, const RecordPlane &flag) {
    flagged(PQ_OP_X_FLAG, q, flag);
  }
  void ApplyPauliYGateWithApplyFlag(std::size_t q, const RecordPlane &flag) {
    flagged(PQ_OP_Y_FLAG, q, flag);
  }
  void ApplyPauliZGateWithApplyFlag(std::size_t q, const RecordPlane &flag) {
    flagged(PQ_OP_Z_FLAG, q, flag);
  }
  void ApplyPauliXGateWithApplyFlag(std::size_t q, const std::vector<Precision> &flag) {
    flagged_host(PQ_OP_X_FLAG, q, flag);
  }
  void ApplyPauliYGateWithApplyFlag(std::size_t q, const std::vector<Precision> &flag) {
    flagged_host(PQ_OP_Y_FLAG, q, flag);
  }
  void ApplyPauliZGateWithApplyFlag(std::size_t q, const std::vector<Precision> &flag) {
    flagged_host(PQ_OP_Z_FLAG, q, flag);
  }
  /// CX / CZ on the flagged shots (the H/S flag API of :348-372 extended to
  /// the two-qubit gates; no reference counterpart).
  void ApplyControlNotGateWithApplyFlag(std::size_t c, std::size_t t, const RecordPlane &flag) {
    flagged2(PQ_OP_CX_FLAG, c, t, flag);
  }
  void ApplyControlPhaseGateWithApplyFlag(std::size_t c, std::size_t t,
                                          const RecordPlane &flag) {
    flagged2(PQ_OP_CZ_FLAG, c, t, flag);
  }
  void ApplyControlNotGateWithApplyFlag(std::size_t c, std::size_t t,
                                        const std::vector<Precision> &flag) {
    flagged2_host(PQ_OP_CX_FLAG, c, t, flag);
  }
  void ApplyControlPhaseGateWithApplyFlag(std::size_t c, std::size_t t,
                                          const std::vector<Precision> &flag) {
    flagged2_host(PQ_OP_CZ_FLAG, c, t, flag);
  }

  // ---- gates applied per shot with a probability ----
  // The per-shot form of the reference's WithProb functors
  // (GateFunctors.hpp:38-75, 96-124, 172-202, 256-288, 330-357, 379-406):
  // each shot applies the gate iff frand(0,1) < prob, one draw per shot.
  void ApplyHadamardGateWithProb(std::size_t q, float prob) { with_prob(PQ_OP_H, q, 0, prob); }
  void ApplyPhaseGateWithProb(std::size_t q, float prob) { with_prob(PQ_OP_S, q, 0, prob); }
  void ApplyPauliXGateWithProb(std::size_t q, float prob) { with_prob(PQ_OP_X, q, 0, prob); }
  void ApplyPauliYGateWithProb(std::size_t q, float prob) { with_prob(PQ_OP_Y, q, 0, prob); }
  void ApplyPauliZGateWithProb(std::size_t q, float prob) { with_prob(PQ_OP_Z, q, 0, prob); }
  void ApplyControlNotGateWithProb(std::size_t c, std::size_t t, float prob) {
    with_prob(PQ_OP_CX, c, t, prob);
  }
  void ApplyControlPhaseGateWithProb(std::size_t c, std::size_t t, float prob) {
    with_prob(PQ_OP_CZ, c, t, prob);
  }

  // ---- gates (CliffordState.hpp:379-531) ----
  void ApplyHadamardGateToQubits(const std::vector<std::size_t> &t) { each1(PQ_OP_H, t); }
  void ApplyHadamardGate(std::size_t q) { gate(PQ_OP_H, q); }
  void ApplyControlNotGateToQubits(const std::vector<std::size_t> &t) { each2(PQ_OP_CX, t); }
  void ApplyControlNotGate(std::size_t c, std::size_t t) { gate(PQ_OP_CX, c, t); }
  void ApplyControlPhaseGateToQubits(const std::vector<std::size_t> &t) { each2(PQ_OP_CZ, t); }
  void ApplyControlPhaseGate(std::size_t c, std::size_t t) { gate(PQ_OP_CZ, c, t); }
  void ApplyPhaseGateToQubits(const std::vector<std::size_t> &t) { each1(PQ_OP_S, t); }
  void ApplyPhaseGate(std::size_t q) { gate(PQ_OP_S, q); }
  void ApplyPauliXGateToQubits(const std::vector<std::size_t> &t) { each1(PQ_OP_X, t); }
  void ApplyPauliXGate(std::size_t q) { gate(PQ_OP_X, q); }
  void ApplyPauliYGateToQubits(const std::vector<std::size_t> &t) { each1(PQ_OP_Y, t); }
  void ApplyPauliYGate(std::size_t q) { gate(PQ_OP_Y, q); }
  void ApplyPauliZGateToQubits(const std::vector<std::size_t> &t) { each1(PQ_OP_Z, t); }
  void ApplyPauliZGate(std::size_t q) { gate(PQ_OP_Z, q); }

  /// Per-shot Pauli channel: X, Y, Z with probabilities px, py, pz.
  void ApplyPauliNoise(std::size_t q, double px, double py, double pz) {
    auto th = [](double p) -> uint32_t {
      const double v = p * 4294967296.0;
      return v <= 0 ? 0u : (v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v);
    };
    detail::check(pq_noise(h_, q, th(px), th(px + py), th(px + py + pz)), "ApplyPauliNoise");
  }

  // ---- measurement (CliffordState.hpp:543-659) ----
  auto MeasureQubits(const std::vector<std::size_t> &target_qubits, float bias = 0.5) {
    for (std::size_t q : target_qubits) MeasureQubit(q, bias);
  }

  std::size_t MeasureQubit(std::size_t target_qubit, float bias = 0.5,
                           std::optional<std::size_t> measurement_id = std::nullopt,
                           const std::vector<Precision> *coins = nullptr) {
    int64_t rec = -1;
    if (measurement_id.has_value()) {
      PLAQUETTE_ABORT_IF(measurement_id.value() >= measurements_.size(),
                         "Measurement id is out of range.");
      rec = (int64_t)measurement_id.value();
    }
    std::vector<int32_t> c;
    if (coins) c.assign(coins->begin(), coins->end());
    uint64_t id = 0;
    detail::check(pq_measure(h_, target_qubit, bias, rec, coins ? c.data() : nullptr, &id),
                  "MeasureQubit");
    if (rec < 0) {
      measurements_.emplace_back();
      Measurement &m = measurements_.back();
      m.batch_size = GetBatchSize();
      m.measurement_results_device.reset(new RecordPlane{h_, (std::size_t)id, 0});
      m.measurement_determined_device.reset(new RecordPlane{h_, (std::size_t)id, 1});
    }
    Measurement &m = measurements_[id];
    m.qubit_measured = target_qubit;
    m.measurement_results_host.clear();
    m.measurement_determined_host.clear();
    return (std::size_t)id;
  }

  /// CliffordState.hpp:635-648.
  auto GetMeasurement(std::size_t measurement_id, std::size_t batch_index) {
    std::optional<std::pair<int, int>> result;
    if (measurement_id < measurements_.size() && batch_index < GetBatchSize()) {
      int32_t res = 0, det = 0;
      detail::check(pq_get_measurement(h_, measurement_id, batch_index, &res, &det),
                    "GetMeasurement");
      result = std::make_pair((int)res, (int)det);
    }
    return result;
  }

  /// CliffordState.hpp:655-659.
  auto &GetMeasurement(std::size_t measurement_id) {
    PLAQUETTE_ABORT_IF_NOT(measurement_id < measurements_.size(),
                           "Measurement id or batch index is out of range.");
    return measurements_[measurement_id];
  }

  std::size_t GetNumMeasurements() const { return measurements_.size(); }

  /// Run a whole op list (plain, persistent or graph form; see plaquette_amd.h).
  void RunProgram(const std::vector<pq_op> &ops, int form = PQ_RUN_PERSISTENT) {
    detail::check(pq_run_program(h_, ops.data(), ops.size(), form), "RunProgram");
    SyncRecords();
  }

  /// Mirror record metadata after programs created records.
  void SyncRecords() {
    uint64_t nrec = 0;
    detail::check(pq_info(h_, nullptr, nullptr, nullptr, &nrec), "info");
    while (measurements_.size() < nrec) {
      const std::size_t id = measurements_.size();
      measurements_.emplace_back();
      Measurement &m = measurements_.back();
      m.batch_size = GetBatchSize();
      m.measurement_results_device.reset(new RecordPlane{h_, id, 0});
      m.measurement_determined_device.reset(new RecordPlane{h_, id, 1});
    }
    for (auto &m : measurements_) {
      m.measurement_results_host.clear();
      m.measurement_determined_host.clear();
    }
  }

 private:
  struct Info {
    uint64_t n, B;
    int64_t seed;
  };
  Info info() const {
    Info i{};
    detail::check(pq_info(h_, &i.n, &i.B, &i.seed, nullptr), "info");
    return i;
  }
  DevicePlanes planes() const {
    DevicePlanes p;
    uint64_t W = 0, R = 0;
    detail::check(pq_device_handles(h_, &p.words, nullptr, &W, &R, nullptr), "planes");
    p.num_qubits = GetNumQubits();
    p.batch_size = GetBatchSize();
    p.word_groups = W;
    p.rows = R;
    return p;
  }
  void gate(uint32_t code, std::size_t q0, std::size_t q1 = 0) {
    detail::check(pq_gate(h_, code, q0, q1), "gate");
  }
  void each1(uint32_t code, const std::vector<std::size_t> &t) {
    for (std::size_t q : t) gate(code, q);
  }
  void each2(uint32_t code, const std::vector<std::size_t> &t) {
    PLAQUETTE_ABORT_IF(t.size() % 2 != 0, "two-qubit gate list needs pairs");
    for (std::size_t i = 0; i + 1 < t.size(); i += 2) gate(code, t[i], t[i + 1]);
  }
  void flagged(uint32_t code, std::size_t q, const RecordPlane &flag) {
    PLAQUETTE_ABORT_IF(flag.state != h_, "flag record belongs to another state");
    PLAQUETTE_ABORT_IF(flag.which != 0, "flag must be a results plane");
    detail::check(pq_gate_flag_record(h_, code, q, flag.record), "flagged gate");
  }
  void flagged_host(uint32_t code, std::size_t q, const std::vector<Precision> &flag) {
    PLAQUETTE_ABORT_IF(flag.size() != GetBatchSize(), "flag vector needs batch_size entries");
    std::vector<int32_t> f(flag.begin(), flag.end());
    detail::check(pq_gate_flag_host(h_, code, q, f.data()), "flagged gate");
  }
  void flagged2(uint32_t code, std::size_t c, std::size_t t, const RecordPlane &flag) {
    PLAQUETTE_ABORT_IF(flag.state != h_, "flag record belongs to another state");
    PLAQUETTE_ABORT_IF(flag.which != 0, "flag must be a results plane");
    detail::check(pq_gate2_flag_record(h_, code, c, t, flag.record), "flagged gate");
  }
  void flagged2_host(uint32_t code, std::size_t c, std::size_t t,
                     const std::vector<Precision> &flag) {
    PLAQUETTE_ABORT_IF(flag.size() != GetBatchSize(), "flag vector needs batch_size entries");
    std::vector<int32_t> f(flag.begin(), flag.end());
    detail::check(pq_gate2_flag_host(h_, code, c, t, f.data()), "flagged gate");
  }
  void with_prob(uint32_t code, std::size_t q0, std::size_t q1, float prob) {
    detail::check(pq_gate_prob(h_, code, q0, q1, prob), "gate with probability");
  }

  pq_state h_ = nullptr;
  std::vector<Measurement> measurements_;
};

}  // namespace Plaquette
