// pybind11 module `plaquette_simulator_bindings`: the reference's Python
// surface (plaquette_simulator/src/Bindings.cpp:11-31, same class and method
// names) over the gfx950 engine, plus programs, records and timing helpers.
#include <hip/hip_runtime_api.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>

#include "CliffordState.hpp"
#include "plaquette_amd.h"

namespace py = pybind11;
using State = Plaquette::CliffordState<int>;

namespace {

int form_of(const py::object &f) {
  if (py::isinstance<py::int_>(f)) return f.cast<int>();
  const std::string s = f.cast<std::string>();
  if (s == "plain") return PQ_RUN_PLAIN;
  if (s == "persistent") return PQ_RUN_PERSISTENT;
  if (s == "graph") return PQ_RUN_GRAPH;
  if (s == "frame") return PQ_RUN_FRAME;
  throw std::invalid_argument("form must be 'plain', 'persistent' or 'graph'");
}

const pq_op *ops_ptr(const py::array_t<uint32_t, py::array::c_style | py::array::forcecast> &a,
                     uint64_t *n) {
  if (a.size() == 0) {
    *n = 0;
    return nullptr;
  }
  if (a.ndim() != 2 || a.shape(1) != 8)
    throw std::invalid_argument("ops must be an integer array of shape (N, 8)");
  *n = (uint64_t)a.shape(0);
  return reinterpret_cast<const pq_op *>(a.data());
}

void check(int rc, const char *what) {
  if (rc != PQ_OK) throw Plaquette::Exception(std::string(what) + ": " + pq_last_error());
}

struct Program {
  pq_program p = nullptr;
  pq_state s = nullptr;
  Program(State &st, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ops) {
    uint64_t n = 0;
    const pq_op *o = ops_ptr(ops, &n);
    s = st.Handle();
    check(pq_program_create(s, o, n, &p), "Program");
  }
  ~Program() {
    if (p) {
      pq_synchronize(s);
      pq_program_destroy(p);
    }
  }
};

struct Batch {
  pq_batch b = nullptr;
  std::vector<py::object> keep;  // member states and programs stay alive
  Batch() { check(pq_batch_create(&b), "Batch"); }
  ~Batch() {
    if (b) {
      pq_batch_wait(b);
      pq_batch_destroy(b);
    }
  }
};

struct Server {
  pq_server s = nullptr;
  Server(uint64_t budget, uint32_t slots) { check(pq_server_create(budget, slots, &s), "Server"); }
  ~Server() {
    if (s) {
      py::gil_scoped_release nogil;
      pq_server_destroy(s);
    }
  }
};

// Zero-copy export of device buffers in the DLPack exchange format (the
// DLManagedTensor of dlpack.h, restated here: a C struct layout shared by
// every framework that reads it).  Device type 10 is kDLROCM, so torch,
// CuPy and others on ROCm see HIP device memory.
namespace dlpack {
struct Device {
  int32_t device_type, device_id;
};
struct DataType {
  uint8_t code, bits;
  uint16_t lanes;
};
struct Tensor {
  void *data;
  Device device;
  int32_t ndim;
  DataType dtype;
  int64_t *shape, *strides;
  uint64_t byte_offset;
};
struct ManagedTensor {
  Tensor dl_tensor;
  void *manager_ctx;
  void (*deleter)(ManagedTensor *);
};
constexpr int32_t kROCm = 10;
constexpr uint8_t kUInt = 1;

// Owns the shape and a reference to the Python object whose memory it is.
struct Ctx {
  int64_t shape[3];
  py::object owner;
};

void deleter(ManagedTensor *t) {
  if (!t) return;
  {
    py::gil_scoped_acquire gil;
    delete static_cast<Ctx *>(t->manager_ctx);
  }
  delete t;
}

// A "dltensor" capsule over `ptr` (uint64 words, C order, shape dims).
py::capsule export_words(void *ptr, int device, const std::vector<int64_t> &dims,
                         py::object owner) {
  auto *ctx = new Ctx();
  for (size_t i = 0; i < dims.size(); ++i) ctx->shape[i] = dims[i];
  ctx->owner = std::move(owner);
  auto *t = new ManagedTensor();
  t->dl_tensor.data = ptr;
  t->dl_tensor.device = Device{kROCm, device};
  t->dl_tensor.ndim = (int32_t)dims.size();
  t->dl_tensor.dtype = DataType{kUInt, 64, 1};
  t->dl_tensor.shape = ctx->shape;
  t->dl_tensor.strides = nullptr;  // compact row-major
  t->dl_tensor.byte_offset = 0;
  t->manager_ctx = ctx;
  t->deleter = deleter;
  // a capsule nobody consumed frees the tensor; a consumer renames it
  // "used_dltensor" and calls the deleter itself
  return py::capsule(t, "dltensor", [](PyObject *cap) {
    if (PyCapsule_IsValid(cap, "dltensor")) {
      auto *mt = static_cast<ManagedTensor *>(PyCapsule_GetPointer(cap, "dltensor"));
      if (mt && mt->deleter) mt->deleter(mt);
    } else {
      PyErr_Clear();
    }
  });
}
}  // namespace dlpack

}  // namespace

PYBIND11_MODULE(plaquette_simulator_bindings, m) {
  m.doc() = "MI355X (gfx950) batched stabilizer-tableau engine";
  py::register_exception<Plaquette::Exception>(m, "PlaquetteError", PyExc_RuntimeError);

  py::class_<State>(m, "CliffordState")
      .def(py::init<size_t, size_t, int, bool>(), py::arg("num_qubits"),
           py::arg("batch_size"), py::arg("seed") = -1, py::arg("init_to_zero") = true)
      .def_static(
          "from_host",
          [](std::vector<int> x, std::vector<int> z, std::vector<int> r, size_t n, size_t B,
             long seed) { return new State(x, z, r, n, B, seed); },
          py::arg("x"), py::arg("z"), py::arg("r"), py::arg("num_qubits"),
          py::arg("batch_size"), py::arg("seed") = -1, py::return_value_policy::take_ownership)
      .def("copy", [](const State &s) { return new State(s); },
           py::return_value_policy::take_ownership)
      // --- the reference's bound methods (Bindings.cpp:15-31) ---
      .def("get_num_qubits", &State::GetNumQubits)
      .def("get_batch_size", &State::GetBatchSize)
      .def("get_seed", &State::GetSeed)
      .def("reset_clifford_state", &State::ResetCliffordState)
      .def(
          "measure_qubit",
          [](State &s, size_t q, float bias, std::optional<size_t> id,
             std::optional<std::vector<int>> coins) {
            return s.MeasureQubit(q, bias, id, coins ? &*coins : nullptr);
          },
          py::arg("target_qubit"), py::arg("bias") = 0.5f,
          py::arg("measurement_id") = py::none(), py::arg("coins") = py::none())
      .def("get_kokkos_backend", &State::GetKokkosBackend)
      .def("get_measurement",
           (std::optional<std::pair<int, int>>(State::*)(size_t, size_t)) &
               State::GetMeasurement)
      .def("apply_hadamard_gate", &State::ApplyHadamardGate)
      .def("apply_control_not_gate", &State::ApplyControlNotGate)
      .def("apply_control_phase_gate", &State::ApplyControlPhaseGate)
      .def("apply_phase_gate", &State::ApplyPhaseGate)
      .def("apply_pauli_x_gate", &State::ApplyPauliXGate)
      .def("apply_pauli_z_gate", &State::ApplyPauliZGate)
      .def("device_to_host", &State::DeviceToHost)
      // --- superset ---
      .def("apply_pauli_y_gate", &State::ApplyPauliYGate)
      .def(
          "apply_hadamard_gate_with_apply_flag",
          [](State &s, size_t q, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyHadamardGateWithApplyFlag(q, *s.GetMeasurement(flag.cast<size_t>())
                                                        .measurement_results_device);
            else
              s.ApplyHadamardGateWithApplyFlag(q, flag.cast<std::vector<int>>());
          },
          py::arg("target_qubit"), py::arg("flag"),
          "flag: a measurement id (its result bits) or a per-shot int sequence")
      .def(
          "apply_phase_gate_with_apply_flag",
          [](State &s, size_t q, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyPhaseGateWithApplyFlag(q, *s.GetMeasurement(flag.cast<size_t>())
                                                    .measurement_results_device);
            else
              s.ApplyPhaseGateWithApplyFlag(q, flag.cast<std::vector<int>>());
          },
          py::arg("target_qubit"), py::arg("flag"))
      .def(
          "apply_pauli_with_apply_flag",
          [](State &s, const std::string &kind, size_t q, size_t id) {
            uint32_t code = kind == "X" ? PQ_OP_X_FLAG
                            : kind == "Y" ? PQ_OP_Y_FLAG
                            : kind == "Z" ? PQ_OP_Z_FLAG
                                          : 0;
            if (!code) throw std::invalid_argument("kind must be X, Y or Z");
            s.ApplyPauliWithApplyFlag(code, q, *s.GetMeasurement(id).measurement_results_device);
          },
          py::arg("kind"), py::arg("target_qubit"), py::arg("measurement_id"))
      .def(
          "apply_pauli_x_gate_with_apply_flag",
          [](State &s, size_t q, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyPauliXGateWithApplyFlag(
                  q, *s.GetMeasurement(flag.cast<size_t>()).measurement_results_device);
            else
              s.ApplyPauliXGateWithApplyFlag(q, flag.cast<std::vector<int>>());
          },
          py::arg("target_qubit"), py::arg("flag"))
      .def(
          "apply_pauli_y_gate_with_apply_flag",
          [](State &s, size_t q, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyPauliYGateWithApplyFlag(
                  q, *s.GetMeasurement(flag.cast<size_t>()).measurement_results_device);
            else
              s.ApplyPauliYGateWithApplyFlag(q, flag.cast<std::vector<int>>());
          },
          py::arg("target_qubit"), py::arg("flag"))
      .def(
          "apply_pauli_z_gate_with_apply_flag",
          [](State &s, size_t q, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyPauliZGateWithApplyFlag(
                  q, *s.GetMeasurement(flag.cast<size_t>()).measurement_results_device);
            else
              s.ApplyPauliZGateWithApplyFlag(q, flag.cast<std::vector<int>>());
          },
          py::arg("target_qubit"), py::arg("flag"))
      .def(
          "apply_control_not_gate_with_apply_flag",
          [](State &s, size_t c, size_t t, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyControlNotGateWithApplyFlag(
                  c, t, *s.GetMeasurement(flag.cast<size_t>()).measurement_results_device);
            else
              s.ApplyControlNotGateWithApplyFlag(c, t, flag.cast<std::vector<int>>());
          },
          py::arg("control_qubit"), py::arg("target_qubit"), py::arg("flag"))
      .def(
          "apply_control_phase_gate_with_apply_flag",
          [](State &s, size_t c, size_t t, const py::object &flag) {
            if (py::isinstance<py::int_>(flag))
              s.ApplyControlPhaseGateWithApplyFlag(
                  c, t, *s.GetMeasurement(flag.cast<size_t>()).measurement_results_device);
            else
              s.ApplyControlPhaseGateWithApplyFlag(c, t, flag.cast<std::vector<int>>());
          },
          py::arg("control_qubit"), py::arg("target_qubit"), py::arg("flag"))
      .def("apply_hadamard_gate_with_prob", &State::ApplyHadamardGateWithProb,
           py::arg("target_qubit"), py::arg("prob"))
      .def("apply_phase_gate_with_prob", &State::ApplyPhaseGateWithProb,
           py::arg("target_qubit"), py::arg("prob"))
      .def("apply_pauli_x_gate_with_prob", &State::ApplyPauliXGateWithProb,
           py::arg("target_qubit"), py::arg("prob"))
      .def("apply_pauli_y_gate_with_prob", &State::ApplyPauliYGateWithProb,
           py::arg("target_qubit"), py::arg("prob"))
      .def("apply_pauli_z_gate_with_prob", &State::ApplyPauliZGateWithProb,
           py::arg("target_qubit"), py::arg("prob"))
      .def("apply_control_not_gate_with_prob", &State::ApplyControlNotGateWithProb,
           py::arg("control_qubit"), py::arg("target_qubit"), py::arg("prob"))
      .def("apply_control_phase_gate_with_prob", &State::ApplyControlPhaseGateWithProb,
           py::arg("control_qubit"), py::arg("target_qubit"), py::arg("prob"))
      .def("apply_pauli_noise", &State::ApplyPauliNoise, py::arg("target_qubit"),
           py::arg("px"), py::arg("py"), py::arg("pz"))
      .def("check_sum", &State::CheckSum, py::arg("starting_batch_id") = 0,
           py::arg("ending_batch_id") = -1)
      .def("device_to_host_numpy",
           [](State &s) {
             const size_t n = s.GetNumQubits(), B = s.GetBatchSize(), R = 2 * n + 1;
             const bool empty = n == 0 || B == 0;
             py::array_t<int32_t> x({empty ? 0 : B, R, n}), z({empty ? 0 : B, R, n}),
                 r({empty ? 0 : B, R});
             check(pq_device_to_host(s.Handle(), x.mutable_data(), z.mutable_data(),
                                     r.mutable_data()),
                   "device_to_host");
             return py::make_tuple(x, z, r);
           })
      .def("host_to_device",
           [](State &s, std::vector<int> x, std::vector<int> z, std::vector<int> r) {
             s.HostToDevice(x, z, r);
           })
      .def("device_to_device", &State::DeviceToDevice)
      // DLPack capsules over the device bit-planes (no copy).  The state's
      // queued work is finished first; the capsule keeps the state alive.
      // planes: uint64 [W][2n+1][2n+1] = word[w][col][row] (col < n: x, then
      // z, then r).  records: uint64 [nrec][2][W] (results, determined); the
      // record arena moves when records are added, so take a new capsule
      // after measurements that create records.
      .def("planes_dlpack",
           [](py::object self) {
             State &s = self.cast<State &>();
             void *tab = nullptr;
             uint64_t W = 0, R = 0;
             check(pq_device_handles(s.Handle(), &tab, nullptr, &W, &R, nullptr), "planes_dlpack");
             check(pq_synchronize(s.Handle()), "planes_dlpack");
             int dev = 0;
             (void)hipGetDevice(&dev);
             return dlpack::export_words(tab, dev, {(int64_t)W, (int64_t)R, (int64_t)R}, self);
           })
      .def("compare",
           [](State &s, State &other) {
             uint64_t dp = 0, dr = 0;
             check(pq_compare(s.Handle(), other.Handle(), &dp, &dr), "compare");
             return py::make_tuple(dp, dr);
           },
           "(tableau words, record words) in which the two states differ (on the device)")
      .def("read_records",
           [](State &s) {
             void *recs = nullptr;
             uint64_t W = 0, nrec = 0;
             check(pq_device_handles(s.Handle(), nullptr, &recs, &W, nullptr, nullptr),
                   "read_records");
             check(pq_info(s.Handle(), nullptr, nullptr, nullptr, &nrec), "read_records");
             check(pq_synchronize(s.Handle()), "read_records");
             py::array_t<uint64_t> a({(size_t)nrec, (size_t)2, (size_t)W});
             if (nrec && W &&
                 hipMemcpy(a.mutable_data(), recs, nrec * 2 * W * 8, hipMemcpyDeviceToHost) !=
                     hipSuccess)
               throw Plaquette::Exception("read_records: device copy failed");
             return a;
           })
      .def("records_dlpack",
           [](py::object self) {
             State &s = self.cast<State &>();
             void *recs = nullptr;
             uint64_t W = 0, nrec = 0;
             check(pq_device_handles(s.Handle(), nullptr, &recs, &W, nullptr, nullptr),
                   "records_dlpack");
             check(pq_info(s.Handle(), nullptr, nullptr, nullptr, &nrec), "records_dlpack");
             check(pq_synchronize(s.Handle()), "records_dlpack");
             int dev = 0;
             (void)hipGetDevice(&dev);
             return dlpack::export_words(recs, dev, {(int64_t)nrec, 2, (int64_t)W}, self);
           })
      .def("get_record",
           [](State &s, size_t id) {
             const size_t B = s.GetBatchSize();
             py::array_t<int32_t> res(B), det(B);
             check(pq_get_record(s.Handle(), id, res.mutable_data(), det.mutable_data()),
                   "get_record");
             return py::make_tuple(res, det);
           })
      .def("read_planes",
           [](State &s) {
             uint64_t W = 0, R = 0;
             check(pq_device_handles(s.Handle(), nullptr, nullptr, &W, &R, nullptr), "planes");
             const size_t n = s.GetNumQubits();
             py::array_t<uint64_t> a({(size_t)W, 2 * n + 1, (size_t)R});
             check(pq_read_planes(s.Handle(), a.mutable_data()), "read_planes");
             return a;
           })
      .def_property_readonly("num_records", &State::GetNumMeasurements)
      // --- diagnostics mirroring the reference's test-only functors ---
      .def("row_multiply",
           [](State &s, size_t i, size_t k) {
             check(pq_row_multiply(s.Handle(), i, k), "row_multiply");
           })
      .def("row_product_sign",
           [](State &s, size_t i, size_t k) {
             py::array_t<int32_t> out(s.GetBatchSize());
             check(pq_row_product_sign(s.Handle(), i, k, out.mutable_data()), "row_product_sign");
             return out;
           })
      .def("measure_determined",
           [](State &s, size_t q) {
             py::array_t<int32_t> out(s.GetBatchSize());
             check(pq_measure_determined(s.Handle(), q, out.mutable_data()), "measure_determined");
             return out;
           })
      .def(
          "run_program",
          [](State &s, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ops,
             const py::object &form) {
            uint64_t n = 0;
            const pq_op *o = ops_ptr(ops, &n);
            check(pq_run_program(s.Handle(), o, n, form_of(form)), "run_program");
            s.SyncRecords();
          },
          py::arg("ops"), py::arg("form") = "persistent")
      .def("synchronize", [](State &s) { check(pq_synchronize(s.Handle()), "synchronize"); })
      .def("sync_records", &State::SyncRecords)
      .def("event_record",
           [](State &s, int slot) { check(pq_event_record(s.Handle(), slot), "event_record"); })
      .def("event_elapsed_ms", [](State &s, int a, int b) {
        float ms = 0;
        check(pq_event_elapsed_ms(s.Handle(), a, b, &ms), "event_elapsed_ms");
        return ms;
      });

  py::class_<Program>(m, "Program")
      .def(py::init<State &, py::array_t<uint32_t, py::array::c_style | py::array::forcecast>>(),
           py::arg("state"), py::arg("ops"), py::keep_alive<1, 2>())
      .def(
          "run",
          [](Program &p, State &s, const py::object &form) {
            // any state of the shape the program was compiled for
            check(pq_program_run(s.Handle(), p.p, form_of(form)), "Program.run");
          },
          py::arg("state"), py::arg("form") = "persistent")
      .def("info",
           [](Program &p) {
             uint64_t g = 0, mm = 0, nz = 0, ni = 0;
             check(pq_program_info(p.p, &g, &mm, &nz, &ni), "info");
             py::dict d;
             d["gates"] = g;
             d["measurements"] = mm;
             d["noise_ops"] = nz;
             d["instructions"] = ni;
             uint64_t ns = 0, nt = 0;
             check(pq_program_layout(p.p, &ns, &nt), "info");
             d["segments"] = ns;
             d["tiled_segments"] = nt;
             return d;
           })

      .def("set_launch", [](Program &p, uint32_t threads, uint32_t blocks) {
        check(pq_program_set_launch(p.p, threads, blocks), "set_launch");
      })
      .def("set_slices",
           [](Program &p, uint32_t slices) {
             check(pq_program_set_slices(p.p, slices), "set_slices");
           })
      .def("set_slice_wait",
           [](Program &p, uint32_t polls) {
             check(pq_program_set_slice_wait(p.p, polls), "set_slice_wait");
           })
      .def(
          "slice_status",
          [](Program &p, State &s) {
            uint32_t used = 1, failed = 0;
            check(pq_program_slice_status(s.Handle(), p.p, &used, &failed), "slice_status");
            return py::make_tuple(used, failed);
          },
          py::arg("state"))
      .def("set_census",
           [](Program &p, bool on) { check(pq_program_set_census(p.p, on ? 1 : 0), "set_census"); })
      .def("read_census",
           [](Program &p) {
             uint64_t c[8];
             check(pq_program_read_census(p.s, p.p, c), "read_census");
             py::dict d;
             d["random"] = c[0];
             d["determined"] = c[1];
             d["random_rows"] = c[2];
             d["determined_steps"] = c[3];
             d["ticks_gates"] = c[4];
             d["ticks_pivot_check"] = c[5];
             d["ticks_determined"] = c[6];
             d["ticks_general"] = c[7];
             return d;
           })
      .def("set_timing",
           [](Program &p, bool on) { check(pq_program_set_timing(p.p, on ? 1 : 0), "set_timing"); })
      .def("kernel_ms", [](Program &p) {
        float ms = 0;
        uint64_t k = 0;
        check(pq_program_kernel_ms(p.p, &ms, &k), "kernel_ms");
        return py::make_tuple(ms, k);
      });

  py::class_<Batch>(m, "Batch")
      .def(py::init<>())
      .def(
          "add",
          [](Batch &b, State &st, Program &p) {
            check(pq_batch_add(b.b, st.Handle(), p.p), "Batch.add");
            b.keep.push_back(py::cast(&st));
            b.keep.push_back(py::cast(&p));
          },
          py::arg("state"), py::arg("program"))
      .def(
          "run",
          [](Batch &b, const py::object &form, bool wait) {
            check(pq_batch_run(b.b, form_of(form)), "Batch.run");
            if (wait) check(pq_batch_wait(b.b), "Batch.wait");
          },
          py::arg("form") = "graph", py::arg("wait") = true)
      .def("wait", [](Batch &b) { check(pq_batch_wait(b.b), "Batch.wait"); })
      .def(
          "set_launch",
          [](Batch &b, uint32_t threads, uint32_t blocks, uint64_t lds) {
            check(pq_batch_set_launch(b.b, threads, blocks, lds), "Batch.set_launch");
          },
          py::arg("threads") = 0, py::arg("blocks") = 0, py::arg("lds_budget") = 0)
      .def(
          "autotune",
          [](Batch &b, int reps) {
            uint32_t t = 0;
            uint64_t l = 0;
            float ms = 0;
            check(pq_batch_autotune(b.b, reps, &t, &l, &ms), "Batch.autotune");
            uint32_t wt = 1;
            check(pq_batch_word_tile(b.b, &wt), "Batch.autotune");
            py::dict d;
            d["threads"] = t;
            d["lds_budget"] = l;
            d["word_tile"] = wt;
            d["best_ms"] = ms;
            return d;
          },
          py::arg("reps") = 2)
      .def("info",
           [](Batch &b) {
             uint64_t mem = 0, items = 0, lds = 0;
             uint32_t t = 0, bl = 0;
             check(pq_batch_info(b.b, &mem, &items, &t, &bl, &lds), "Batch.info");
             uint32_t wt = 1;
             check(pq_batch_word_tile(b.b, &wt), "Batch.info");
             py::dict d;
             d["members"] = mem;
             d["items"] = items;
             d["threads"] = t;
             d["blocks"] = bl;
             d["lds_budget"] = lds;
             d["word_tile"] = wt;
             return d;
           })
      .def("set_word_tile",
           [](Batch &b, uint32_t words) {
             check(pq_batch_set_word_tile(b.b, words), "Batch.set_word_tile");
           })
      .def("set_timing",
           [](Batch &b, bool on) { check(pq_batch_set_timing(b.b, on ? 1 : 0), "Batch.set_timing"); })
      .def("last_ms", [](Batch &b) {
        float ms = 0;
        check(pq_batch_last_ms(b.b, &ms), "Batch.last_ms");
        return ms;
      });

  py::class_<Server>(m, "Server")
      .def(py::init<uint64_t, uint32_t>(), py::arg("budget_bytes"), py::arg("slots") = 8)
      .def(
          "submit",
          [](Server &sv, uint64_t n, uint64_t B, int64_t seed,
             py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ops) {
            uint64_t nops = 0, t = 0;
            const pq_op *o = ops_ptr(ops, &nops);
            check(pq_server_submit(sv.s, n, B, seed, o, nops, &t), "Server.submit");
            return t;
          },
          py::arg("num_qubits"), py::arg("batch_size"), py::arg("seed"), py::arg("ops"))
      .def("status",
           [](Server &sv, uint64_t t) {
             int st = 0;
             double lat = 0, q = 0;
             check(pq_server_status(sv.s, t, &st, &lat, &q), "Server.status");
             return py::make_tuple(st, lat, q);
           })
      .def("result",
           [](Server &sv, uint64_t t) {
             uint64_t nrec = 0, W = 0;
             check(pq_server_result(sv.s, t, nullptr, 0, &nrec, &W), "Server.result");
             py::array_t<uint64_t> a({(size_t)nrec, (size_t)2, (size_t)W});
             check(pq_server_result(sv.s, t, a.mutable_data(), nrec * 2 * W, &nrec, &W),
                   "Server.result");
             return a;
           })
      .def("release", [](Server &sv, uint64_t t) { check(pq_server_release(sv.s, t), "Server.release"); })
      .def(
          "set_policy",
          [](Server &sv, const std::string &policy, double age_limit_s) {
            int p = -1;
            if (policy == "fcfs") p = PQ_SERVE_FCFS;
            if (policy == "smallest_first") p = PQ_SERVE_SMALLEST_FIRST;
            if (p < 0) throw std::runtime_error("Server.set_policy: 'fcfs' or 'smallest_first'");
            check(pq_server_set_policy(sv.s, p, age_limit_s), "Server.set_policy");
          },
          py::arg("policy"), py::arg("age_limit_s") = 0.0)
      .def(
          "set_form",
          [](Server &sv, const std::string &form) {
            int f = -1;
            if (form == "persistent") f = PQ_RUN_PERSISTENT;
            if (form == "frame") f = PQ_RUN_FRAME;
            if (f < 0) throw std::runtime_error("Server.set_form: 'persistent' or 'frame'");
            check(pq_server_set_form(sv.s, f), "Server.set_form");
          },
          py::arg("form"),
          "'frame': requests whose ops allow it run in the frame form (plans shared)")
      .def("request_bytes",
           [](Server &sv, uint64_t n, uint64_t B,
              py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ops) {
             uint64_t k = 0, bytes = 0;
             const pq_op *o = ops_ptr(ops, &k);
             check(pq_server_request_bytes(sv.s, n, B, o, k, &bytes), "Server.request_bytes");
             return bytes;
           })
      .def("frame_stats",
           [](Server &sv) {
             uint64_t runs = 0, builds = 0, hits = 0;
             check(pq_server_frame_stats(sv.s, &runs, &builds, &hits), "Server.frame_stats");
             py::dict d;
             d["frame_runs"] = runs;
             d["plan_builds"] = builds;
             d["plan_hits"] = hits;
             return d;
           })
      .def("drain",
           [](Server &sv) {
             py::gil_scoped_release nogil;
             check(pq_server_drain(sv.s), "Server.drain");
           })
      .def("stats", [](Server &sv) {
        uint64_t done = 0, peak = 0, queued = 0;
        uint32_t maxr = 0;
        check(pq_server_stats(sv.s, &done, &peak, &maxr, &queued), "Server.stats");
        py::dict d;
        d["completed"] = done;
        d["peak_bytes"] = peak;
        d["max_running"] = maxr;
        d["queued"] = queued;
        return d;
      });

  m.def(
      "parse_circuit",
      [](const std::string &text, uint32_t first_record) {
        uint64_t nops = 0, nlines = 0, used = 0;
        if (pq_circuit_parse(text.data(), text.size(), first_record, nullptr, 0, &nops, &nlines,
                             &used) != PQ_OK)
          throw std::runtime_error(pq_last_error());
        py::array_t<uint32_t> a({(size_t)nops, (size_t)8});
        if (pq_circuit_parse(text.data(), text.size(), first_record,
                             reinterpret_cast<pq_op *>(a.mutable_data()), nops, &nops, &nlines,
                             &used) != PQ_OK)
          throw std::runtime_error(pq_last_error());
        return py::make_tuple(a, nlines, used);
      },
      py::arg("text"), py::arg("first_record") = 0,
      "Parse CliffordCircuit text into (ops uint32[nops, 8], lines, qubits used).");
  m.def(
      "frame_plan",
      [](uint64_t n, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ops) {
        size_t nops = 0;
        const pq_op *o = ops_ptr(ops, &nops);
        uint64_t info[16];
        check(pq_frame_plan(n, o, nops, info, nullptr, nullptr, nullptr, nullptr, nullptr),
              "frame_plan");
        py::array_t<uint32_t> stream(info[0]), chunks({(size_t)info[1], (size_t)4}),
            preload({(size_t)info[2], (size_t)2});
        const size_t R = 2 * n + 1;
        py::array_t<uint64_t> x({R, (size_t)info[3]}), z({R, (size_t)info[3]});
        check(pq_frame_plan(n, o, nops, info, stream.mutable_data(), chunks.mutable_data(),
                            preload.mutable_data(), x.mutable_data(), z.mutable_data()),
              "frame_plan");
        py::dict d;
        const char *names[16] = {"stream_words", "chunks_n", "preload_n", "nw", "nref",
                                 "lds_rec", "lds_const", "lds_noise", "nb_words", "sb_words",
                                 "lds_bytes", "random", "determined", "final_identity", "nbufs",
                                 "split_chunk"};
        for (int i = 0; i < 16; ++i) d[names[i]] = info[i];
        d["stream"] = stream;
        d["chunks"] = chunks;
        d["preload"] = preload;
        d["x"] = x;
        d["z"] = z;
        return d;
      },
      py::arg("num_qubits"), py::arg("ops"),
      "Frame-form plan of an op list (host only): op stream, chunk table, final shared X/Z");
  m.def(
      "sampling_state",
      [](size_t n, size_t B, int seed) { return new State(State::Sampling{}, n, B, seed); },
      py::arg("num_qubits"), py::arg("batch_size"), py::arg("seed") = -1,
      py::return_value_policy::take_ownership,
      "A CliffordState that keeps only its phase column: frame-form programs from |0...0>, "
      "records and reset only (W*(2n+1) words on the device instead of W*(2n+1)^2)");
  m.def("set_window_qubits",
        [](uint32_t q) { check(pq_set_window_qubits(q), "set_window_qubits"); });
  m.def("coin_threshold", &pq_coin_threshold);
  m.def("draw", &pq_draw);
  m.def("backend_name", &pq_backend_name);
  m.def("device_count", []() {
    int c = 0;
    pq_device_count(&c);
    return c;
  });
  m.attr("OP_H") = PQ_OP_H;
  m.attr("OP_S") = PQ_OP_S;
  m.attr("OP_X") = PQ_OP_X;
  m.attr("OP_Y") = PQ_OP_Y;
  m.attr("OP_Z") = PQ_OP_Z;
  m.attr("OP_CX") = PQ_OP_CX;
  m.attr("OP_CZ") = PQ_OP_CZ;
  m.attr("OP_H_FLAG") = PQ_OP_H_FLAG;
  m.attr("OP_S_FLAG") = PQ_OP_S_FLAG;
  m.attr("OP_NOISE") = PQ_OP_NOISE;
  m.attr("OP_X_FLAG") = PQ_OP_X_FLAG;
  m.attr("OP_Y_FLAG") = PQ_OP_Y_FLAG;
  m.attr("OP_Z_FLAG") = PQ_OP_Z_FLAG;
  m.attr("OP_CX_FLAG") = PQ_OP_CX_FLAG;
  m.attr("OP_CZ_FLAG") = PQ_OP_CZ_FLAG;
  m.attr("FLAG_RANDOM") = PQ_FLAG_RANDOM;
  m.attr("OP_M") = PQ_OP_M;
}
