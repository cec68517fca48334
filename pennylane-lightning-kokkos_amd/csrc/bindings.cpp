// pybind11 module `lightning_kokkos_qubit_ops` over the C ABI of plk_amd.h.
//
// Mirrors the class and method names of the reference module
// (pennylane_lightning_kokkos/src/bindings/Bindings.cpp:50-968) so that
// callers which dispatch on names -- `_serialize.py:294` and `apply_kokkos`
// (lightning_kokkos.py:378) use getattr on the state class -- keep working.
// Every call goes through the C ABI; this file includes no HIP header.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <complex>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "plk_amd.h"

namespace py = pybind11;

namespace {

struct EngineError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void chk(int rc) {
    if (rc != PLK_OK) {
        throw EngineError(plk_last_error());
    }
}

// Gate methods of the reference state class (Bindings.cpp:109-432).
const char *const kGates[] = {
    "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "SWAP", "CSWAP", "Toffoli",
    "CY", "CZ", "PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "Rot", "CRX", "CRY",
    "CRZ", "CRot", "IsingXX", "IsingXY", "IsingYY", "IsingZZ", "MultiRZ", "SingleExcitation",
    "SingleExcitationMinus", "SingleExcitationPlus", "DoubleExcitation",
    "DoubleExcitationMinus", "DoubleExcitationPlus"};

// InitializationSettings (Bindings.cpp:854-967): kept for API compatibility;
// device_id selects the GPU, the other fields are recorded only.
struct Settings {
    int num_threads = 0, device_id = 0, disable_warnings = 0, print_configuration = 0,
        tune_internals = 0, tools_help = 0;
    std::string map_device_id_by, tools_libs, tools_args;
    bool h_num_threads = false, h_device_id = false, h_map = false, h_disable = false,
         h_print = false, h_tune = false, h_libs = false, h_help = false, h_args = false;
};

std::string wires_str(const std::vector<size_t> &w) {
    std::ostringstream os;
    os << '[';
    for (size_t i = 0; i < w.size(); i++) {
        os << w[i] << (i + 1 < w.size() ? ", " : "");
    }
    os << ']';
    return os.str();
}

template <class PrecisionT> struct Traits;
template <> struct Traits<float> {
    static constexpr int dtype = PLK_C64;
    static constexpr const char *suffix = "C64";
};
template <> struct Traits<double> {
    static constexpr int dtype = PLK_C128;
    static constexpr const char *suffix = "C128";
};

template <class PrecisionT> struct StateVector {
    using C = std::complex<PrecisionT>;
    plk_state *h = nullptr;
    int n = 0;
    StateVector(size_t num_qubits, int device) {
        chk(plk_state_create(static_cast<int>(num_qubits), Traits<PrecisionT>::dtype, device, &h));
        n = static_cast<int>(num_qubits);
    }
    ~StateVector() {
        if (h) {
            plk_state_destroy(h);
        }
    }
    StateVector(const StateVector &) = delete;
    StateVector &operator=(const StateVector &) = delete;
    size_t length() const { return size_t{1} << n; }
    void gate(const std::string &name, const std::vector<size_t> &wires, bool inv,
              const std::vector<double> &params) {
        if (name == "Identity") {
            return;
        }
        chk(plk_apply(h, name.c_str(), wires.data(), wires.size(), inv ? 1 : 0, params.data(),
                      params.size()));
    }
    void matrix(const std::vector<size_t> &wires, const C *m, size_t len, bool inv) {
        std::vector<double> d(2 * len);
        for (size_t i = 0; i < len; i++) {
            d[2 * i] = m[i].real();
            d[2 * i + 1] = m[i].imag();
        }
        if ((size_t{1} << (2 * wires.size())) != len) {
            throw EngineError("matrix size does not match the number of wires");
        }
        chk(plk_apply_matrix(h, d.data(), wires.data(), wires.size(), inv ? 1 : 0));
    }
};

// Observable handle with what __repr__ / get_wires / __eq__ need.
struct Observable {
    std::shared_ptr<plk_obs> h;
    std::string kind;
    std::string repr;
    std::string key;
    std::vector<size_t> wires;
    virtual ~Observable() = default;
};

std::shared_ptr<plk_obs> own(plk_obs *o) {
    return std::shared_ptr<plk_obs>(o, [](plk_obs *p) { plk_obs_destroy(p); });
}

std::vector<size_t> obs_wires(plk_obs *o) {
    std::vector<size_t> w(1024);
    size_t cnt = 0;
    chk(plk_obs_wires(o, w.data(), w.size(), &cnt));
    w.resize(std::min(cnt, w.size()));
    return w;
}

// One observable base type per precision, as in the reference module.
template <class PrecisionT> struct ObservableP : Observable {};
template <class PrecisionT> struct Named : ObservableP<PrecisionT> {};
template <class PrecisionT> struct Hermitian : ObservableP<PrecisionT> {};
template <class PrecisionT> struct TensorProd : ObservableP<PrecisionT> {};
template <class PrecisionT> struct Hamiltonian : ObservableP<PrecisionT> {};
template <class PrecisionT> struct SparseHamiltonian : ObservableP<PrecisionT> {};

template <class PrecisionT> struct OpsStruct {
    std::shared_ptr<plk_ops> h;
    std::vector<std::string> names;
    std::vector<std::vector<double>> params;
    std::vector<bool> inverses;
};

template <class PrecisionT> struct AdjointJacobian {};

template <class T> std::vector<double> as_doubles(const py::array_t<T, py::array::c_style | py::array::forcecast> &a) {
    auto b = a.request();
    const T *p = static_cast<const T *>(b.ptr);
    return std::vector<double>(p, p + b.size);
}

template <class PrecisionT>
std::vector<double> complex_doubles(
    const py::array_t<std::complex<PrecisionT>, py::array::c_style | py::array::forcecast> &a) {
    auto b = a.request();
    const auto *p = static_cast<const std::complex<PrecisionT> *>(b.ptr);
    std::vector<double> d(2 * static_cast<size_t>(b.size));
    for (py::ssize_t i = 0; i < b.size; i++) {
        d[2 * i] = p[i].real();
        d[2 * i + 1] = p[i].imag();
    }
    return d;
}

// CSR entry points by index width.
int csr_expval(plk_state *sv, const double *d, const int32_t *i, const int32_t *p, size_t nnz,
               size_t rows, double *out) {
    return plk_expval_csr_i32(sv, d, i, p, nnz, rows, out);
}
int csr_expval(plk_state *sv, const double *d, const int64_t *i, const int64_t *p, size_t nnz,
               size_t rows, double *out) {
    return plk_expval_csr(sv, d, i, p, nnz, rows, out);
}
int obs_sparse(const double *d, const int32_t *i, const int32_t *p, size_t nnz, size_t rows,
               const size_t *w, size_t nw, plk_obs **out) {
    return plk_obs_sparse_i32(d, i, p, nnz, rows, w, nw, out);
}
int obs_sparse(const double *d, const int64_t *i, const int64_t *p, size_t nnz, size_t rows,
               const size_t *w, size_t nw, plk_obs **out) {
    return plk_obs_sparse(d, i, p, nnz, rows, w, nw, out);
}

template <class PrecisionT> void bind_precision(py::module_ &m) {
    using SV = StateVector<PrecisionT>;
    using C = std::complex<PrecisionT>;
    using np_c = py::array_t<C, py::array::c_style | py::array::forcecast>;
    // CSR index width per precision, as the reference's serialiser sends it
    // (_serialize.py:136-154): int32 for complex64, int64 for complex128.
    using IndexT = std::conditional_t<std::is_same_v<PrecisionT, float>, int32_t, int64_t>;
    using np_idx = py::array_t<IndexT, py::array::c_style | py::array::forcecast>;
    using np_r = py::array_t<double, py::array::c_style | py::array::forcecast>;
    const std::string sfx = Traits<PrecisionT>::suffix;

    auto cls = py::class_<SV>(m, ("LightningKokkos_" + sfx).c_str());
    cls.def(py::init([](size_t n) { return new SV(n, 0); }))
        .def(py::init([](size_t n, const Settings &s) { return new SV(n, s.device_id); }))
        .def(py::init([](const np_c &arr) {
            auto b = arr.request();
            size_t len = static_cast<size_t>(b.size);
            size_t n = 0;
            while ((size_t{1} << n) < len) {
                n++;
            }
            auto *sv = new SV(n, 0);
            chk(plk_state_upload(sv->h, b.ptr, len));
            return sv;
        }))
        .def(py::init([](const np_c &arr, const Settings &s) {
            auto b = arr.request();
            size_t len = static_cast<size_t>(b.size);
            size_t n = 0;
            while ((size_t{1} << n) < len) {
                n++;
            }
            auto *sv = new SV(n, s.device_id);
            chk(plk_state_upload(sv->h, b.ptr, len));
            return sv;
        }))
        .def("setBasisState", [](SV &sv, size_t index) { chk(plk_state_set_basis(sv.h, index)); },
             "Create Basis State on Device.")
        .def("setStateVector",
             [](SV &sv, const std::vector<size_t> &indices, const np_c &state) {
                 auto b = state.request();
                 if (static_cast<size_t>(b.size) != indices.size()) {
                     throw EngineError("setStateVector: indices and values differ in length");
                 }
                 chk(plk_state_set_values(sv.h, indices.data(), b.ptr, indices.size()));
             },
             "Set State Vector on device with values and their corresponding indices")
        .def("Identity", [](SV &, const std::vector<size_t> &, bool, const std::vector<double> &) {},
             "Apply the Identity gate.");
    for (const char *g : kGates) {
        const std::string name(g);
        cls.def(g,
                [name](SV &sv, const std::vector<size_t> &wires, bool adjoint,
                       const std::vector<double> &params) { sv.gate(name, wires, adjoint, params); },
                ("Apply the " + name + " gate.").c_str());
    }
    cls.def("apply",
            [](SV &sv, const std::vector<std::string> &names,
               const std::vector<std::vector<size_t>> &wires, const std::vector<bool> &adjoints,
               const std::vector<std::vector<double>> &params) {
                if (names.size() != wires.size()) {
                    throw EngineError("Incompatible number of ops and wires");
                }
                if (names.size() != adjoints.size()) {
                    throw EngineError("Incompatible number of ops and adjoints");
                }
                for (size_t i = 0; i < names.size(); i++) {
                    sv.gate(names[i], wires[i], adjoints[i],
                            i < params.size() ? params[i] : std::vector<double>{});
                }
            })
        .def("apply",
             [](SV &sv, const std::vector<std::string> &names,
                const std::vector<std::vector<size_t>> &wires, const std::vector<bool> &adjoints) {
                 if (names.size() != wires.size()) {
                     throw EngineError("Incompatible number of ops and wires");
                 }
                 if (names.size() != adjoints.size()) {
                     throw EngineError("Incompatible number of ops and adjoints");
                 }
                 for (size_t i = 0; i < names.size(); i++) {
                     sv.gate(names[i], wires[i], adjoints[i], {});
                 }
             })
        .def("apply",
             [](SV &sv, const std::string &name, const std::vector<size_t> &wires, bool inv,
                const std::vector<std::vector<double>> &, const np_c &matrix) {
                 auto b = matrix.request();
                 if (b.size == 0) {
                     sv.gate(name, wires, inv, {});
                 } else {
                     sv.matrix(wires, static_cast<const C *>(b.ptr), static_cast<size_t>(b.size), inv);
                 }
             },
             "Apply operation via the gate matrix")
        .def("applyGenerator",
             [](SV &sv, const std::string &name, const std::vector<size_t> &wires, bool adj,
                const std::vector<double> &) {
                 double s = 0;
                 chk(plk_apply_generator(sv.h, name.c_str(), wires.data(), wires.size(), adj ? 1 : 0,
                                         &s));
                 return static_cast<PrecisionT>(s);
             })
        .def("ExpectationValue",
             [](SV &sv, const std::string &name, const std::vector<size_t> &wires,
                const std::vector<double> &, const np_c &matrix) {
                 // MeasuresKokkos.hpp:80-101: named observables use their
                 // kernel (wires reversed when a matrix is also given), any
                 // other name the matrix.
                 static const char *named[] = {"Identity", "PauliX", "PauliY", "PauliZ", "Hadamard"};
                 auto b = matrix.request();
                 double out = 0;
                 for (const char *nm : named) {
                     if (name == nm) {
                         std::vector<size_t> w = wires;
                         if (b.size) {
                             std::reverse(w.begin(), w.end());
                         }
                         chk(plk_expval_named(sv.h, name.c_str(), w.data(), w.size(), &out));
                         return out;
                     }
                 }
                 auto d = complex_doubles<PrecisionT>(matrix);
                 chk(plk_expval_matrix(sv.h, d.data(), wires.data(), wires.size(), &out));
                 return out;
             },
             "Calculate the expectation value of the given observable.")
        .def("ExpectationValue",
             [](SV &sv, const std::vector<std::string> &, const std::vector<size_t> &wires,
                const std::vector<std::vector<double>> &, const np_c &matrix) {
                 auto d = complex_doubles<PrecisionT>(matrix);
                 double out = 0;
                 chk(plk_expval_matrix(sv.h, d.data(), wires.data(), wires.size(), &out));
                 return out;
             },
             "Calculate the expectation value of the given observable.")
        .def("ExpectationValue",
             [](SV &sv, const std::vector<size_t> &wires, const np_c &matrix) {
                 auto d = complex_doubles<PrecisionT>(matrix);
                 double out = 0;
                 chk(plk_expval_matrix(sv.h, d.data(), wires.data(), wires.size(), &out));
                 return out;
             },
             "Calculate the expectation value of the given observable.")
        .def("ExpectationValue",
             [](SV &sv, const np_c &data, const np_idx &indices, const np_idx &indptr) {
                 auto d = complex_doubles<PrecisionT>(data);
                 double out = 0;
                 if (indptr.size() == 0) {
                     throw EngineError("CSR indptr is empty");
                 }
                 chk(csr_expval(sv.h, d.data(), indices.data(), indptr.data(),
                                static_cast<size_t>(indices.size()),
                                static_cast<size_t>(indptr.size()) - 1, &out));
                 return out;
             },
             "Calculate the expectation value of the given observable.")
        .def("probs",
             [](SV &sv, const std::vector<size_t> &wires) {
                 const size_t m = wires.empty() ? static_cast<size_t>(sv.n) : wires.size();
                 std::vector<double> p(size_t{1} << m);
                 chk(plk_probs(sv.h, wires.data(), wires.size(), p.data()));
                 py::array_t<PrecisionT> out(static_cast<py::ssize_t>(p.size()));
                 auto *o = out.mutable_data();
                 for (size_t i = 0; i < p.size(); i++) {
                     o[i] = static_cast<PrecisionT>(p[i]);
                 }
                 return out;
             })
        .def("GenerateSamples",
             [](SV &sv, size_t num_wires, size_t num_shots) {
                 // The reference draws from a pool seeded with 5374857
                 // (MeasuresKokkos.hpp:551); the same seed feeds SplitMix64 here.
                 py::array_t<size_t> out({static_cast<py::ssize_t>(num_shots),
                                          static_cast<py::ssize_t>(num_wires)});
                 if (num_wires != static_cast<size_t>(sv.n)) {
                     throw EngineError("GenerateSamples: num_wires must equal the qubit count");
                 }
                 if (num_shots > 0) {
                     chk(plk_generate_samples(sv.h, num_shots, 5374857, out.mutable_data()));
                 }
                 return out;
             })
        .def("DeviceToHost",
             [](SV &sv, np_c &host) {
                 auto b = host.request(true);
                 if (b.size) {
                     chk(plk_state_download(sv.h, b.ptr, static_cast<size_t>(b.size)));
                 }
             },
             "Synchronize data from the GPU device to host.")
        .def("HostToDevice",
             [](SV &sv, const np_c &host) {
                 auto b = host.request();
                 if (b.size) {
                     chk(plk_state_upload(sv.h, b.ptr, static_cast<size_t>(b.size)));
                 }
             },
             "Synchronize data from the host device to GPU.")
        .def("numQubits", [](const SV &sv) { return static_cast<size_t>(sv.n); })
        .def("dataLength", [](const SV &sv) { return sv.length(); })
        .def("resetKokkos", [](SV &sv) { chk(plk_state_reset(sv.h)); });

    // ---- observables (Bindings.cpp:591-736) ------------------------------
    using OB = ObservableP<PrecisionT>;
    using ObsPtr = std::shared_ptr<OB>;
    py::class_<OB, ObsPtr>(m, ("ObservableKokkos_" + sfx).c_str(), py::module_local());
    auto common = [](auto &c) {
        c.def("__repr__", [](const OB &o) { return o.repr; })
            .def("get_wires", [](const OB &o) { return o.wires; }, "Get wires of observables")
            .def("__eq__", [](const OB &self, py::handle other) {
                if (!py::isinstance<OB>(other)) {
                    return false;
                }
                const auto &o = other.cast<const OB &>();
                return self.kind == o.kind && self.key == o.key;
            }, "Compare two observables");
    };
    {
        using N = Named<PrecisionT>;
        py::class_<N, std::shared_ptr<N>, OB> c(m, ("NamedObsKokkos_" + sfx).c_str(),
                                                        py::module_local());
        c.def(py::init([](const std::string &name, const std::vector<size_t> &wires) {
            auto o = std::make_shared<N>();
            plk_obs *h = nullptr;
            chk(plk_obs_named(name.c_str(), wires.data(), wires.size(), &h));
            o->h = own(h);
            o->kind = "Named";
            o->repr = name + wires_str(wires);
            o->key = o->repr;
            o->wires = wires;
            return o;
        }));
        common(c);
    }
    {
        using H = Hermitian<PrecisionT>;
        py::class_<H, std::shared_ptr<H>, OB> c(m, ("HermitianObsKokkos_" + sfx).c_str(),
                                                        py::module_local());
        c.def(py::init([](const np_c &matrix, const std::vector<size_t> &wires) {
            auto o = std::make_shared<H>();
            auto d = complex_doubles<PrecisionT>(matrix);
            plk_obs *h = nullptr;
            chk(plk_obs_hermitian(d.data(), wires.data(), wires.size(), &h));
            o->h = own(h);
            o->kind = "Hermitian";
            std::ostringstream os;
            os << "Hermitian" << std::hash<std::string>{}(std::string(
                                     reinterpret_cast<const char *>(d.data()), d.size() * sizeof(double)));
            o->repr = os.str();
            o->key = o->repr + wires_str(wires);
            o->wires = wires;
            return o;
        }));
        common(c);
    }
    {
        using TP = TensorProd<PrecisionT>;
        py::class_<TP, std::shared_ptr<TP>, OB> c(m, ("TensorProdObsKokkos_" + sfx).c_str(),
                                                          py::module_local());
        c.def(py::init([](const std::vector<ObsPtr> &obs) {
            auto o = std::make_shared<TP>();
            std::vector<const plk_obs *> hs;
            std::string rep, key;
            for (size_t i = 0; i < obs.size(); i++) {
                hs.push_back(obs[i]->h.get());
                rep += obs[i]->repr + (i + 1 < obs.size() ? " @ " : "");
                key += obs[i]->kind + ":" + obs[i]->key + ";";
            }
            plk_obs *h = nullptr;
            chk(plk_obs_tensor(hs.data(), hs.size(), &h));
            o->h = own(h);
            o->kind = "TensorProd";
            o->repr = rep;
            o->key = key;
            o->wires = obs_wires(h);
            return o;
        }));
        common(c);
    }
    {
        using HM = Hamiltonian<PrecisionT>;
        py::class_<HM, std::shared_ptr<HM>, OB> c(m, ("HamiltonianKokkos_" + sfx).c_str(),
                                                          py::module_local());
        c.def(py::init([](const np_r &coeffs, const std::vector<ObsPtr> &obs) {
            auto o = std::make_shared<HM>();
            auto cf = as_doubles<double>(coeffs);
            if (cf.size() != obs.size()) {
                throw EngineError("Hamiltonian: coefficients and terms differ in length");
            }
            std::vector<const plk_obs *> hs;
            std::ostringstream rep;
            std::ostringstream key;
            rep << "Hamiltonian: { 'coeffs' : [";
            for (size_t i = 0; i < cf.size(); i++) {
                rep << cf[i] << (i + 1 < cf.size() ? ", " : "");
            }
            rep << "], 'observables' : [";
            for (size_t i = 0; i < obs.size(); i++) {
                hs.push_back(obs[i]->h.get());
                rep << obs[i]->repr << (i + 1 < obs.size() ? ", " : "");
                key.precision(17);
                key << cf[i] << "*" << obs[i]->kind << ":" << obs[i]->key << ";";
            }
            rep << "]}";
            plk_obs *h = nullptr;
            chk(plk_obs_hamiltonian(cf.data(), hs.data(), hs.size(), &h));
            o->h = own(h);
            o->kind = "Hamiltonian";
            o->repr = rep.str();
            o->key = key.str();
            o->wires = obs_wires(h);
            return o;
        }));
        common(c);
    }
    {
        using SH = SparseHamiltonian<PrecisionT>;
        py::class_<SH, std::shared_ptr<SH>, OB> c(
            m, ("SparseHamiltonianKokkos_" + sfx).c_str(), py::module_local());
        c.def(py::init([](const np_c &data, const np_idx &indices, const np_idx &indptr,
                          const std::vector<size_t> &wires) {
            auto o = std::make_shared<SH>();
            auto d = complex_doubles<PrecisionT>(data);
            if (d.size() / 2 != static_cast<size_t>(indices.size())) {
                throw EngineError("SparseHamiltonian: data and indices differ in length");
            }
            plk_obs *h = nullptr;
            chk(obs_sparse(d.data(), indices.data(), indptr.data(),
                           static_cast<size_t>(indices.size()),
                           indptr.size() == 0 ? 0 : static_cast<size_t>(indptr.size()) - 1,
                           wires.data(), wires.size(), &h));
            o->h = own(h);
            o->kind = "SparseHamiltonian";
            std::ostringstream rep;
            rep << "SparseHamiltonian: {\n'data' : ";
            for (size_t i = 0; i < d.size() / 2; i++) {
                rep << "(" << d[2 * i] << "," << d[2 * i + 1] << ")";
            }
            rep << ",\n'indices' : ";
            for (py::ssize_t i = 0; i < indices.size(); i++) {
                rep << indices.data()[i];
            }
            rep << ",\n'indptr' : ";
            for (py::ssize_t i = 0; i < indptr.size(); i++) {
                rep << indptr.data()[i];
            }
            rep << "\n}";
            o->repr = rep.str();
            o->key = o->repr;
            o->wires = wires;
            return o;
        }));
        common(c);
    }

    // ---- operations and adjoint (Bindings.cpp:741-821) --------------------
    using Ops = OpsStruct<PrecisionT>;
    auto make_ops = [](const std::vector<std::string> &names,
                       const std::vector<std::vector<double>> &params,
                       const std::vector<std::vector<size_t>> &wires, const std::vector<bool> &inverses,
                       const std::vector<std::vector<double>> &mats_interleaved) {
        const size_t n = names.size();
        if (params.size() != n || wires.size() != n || inverses.size() != n) {
            throw EngineError("operation lists must have equal lengths");
        }
        std::vector<const char *> cn;
        std::vector<double> pf, mf;
        std::vector<size_t> po{0}, wf, wo{0}, mo{0};
        std::vector<int> inv;
        for (size_t i = 0; i < n; i++) {
            cn.push_back(names[i].c_str());
            pf.insert(pf.end(), params[i].begin(), params[i].end());
            po.push_back(pf.size());
            wf.insert(wf.end(), wires[i].begin(), wires[i].end());
            wo.push_back(wf.size());
            inv.push_back(inverses[i] ? 1 : 0);
            if (i < mats_interleaved.size()) {
                mf.insert(mf.end(), mats_interleaved[i].begin(), mats_interleaved[i].end());
            }
            mo.push_back(mf.size() / 2);
        }
        pf.push_back(0);
        wf.push_back(0);
        mf.push_back(0);
        mf.push_back(0);
        plk_ops *h = nullptr;
        chk(plk_ops_create(n, cn.data(), pf.data(), po.data(), wf.data(), wo.data(), inv.data(),
                           mf.data(), mo.data(), &h));
        Ops o;
        o.h = std::shared_ptr<plk_ops>(h, [](plk_ops *p) { plk_ops_destroy(p); });
        o.names = names;
        o.params = params;
        o.inverses = inverses;
        return o;
    };
    py::class_<Ops>(m, ("OpsStructKokkos_" + sfx).c_str(), py::module_local())
        .def(py::init([make_ops](const std::vector<std::string> &names,
                                 const std::vector<std::vector<double>> &params,
                                 const std::vector<std::vector<size_t>> &wires,
                                 const std::vector<bool> &inverses,
                                 const std::vector<std::vector<std::complex<double>>> &mats) {
            std::vector<std::vector<double>> mi;
            for (auto &mm : mats) {
                std::vector<double> v;
                for (auto c : mm) {
                    v.push_back(c.real());
                    v.push_back(c.imag());
                }
                mi.push_back(v);
            }
            return make_ops(names, params, wires, inverses, mi);
        }))
        .def("__repr__", [](const Ops &ops) {
            std::ostringstream os;
            os << "Operations: [";
            for (size_t i = 0; i < ops.names.size(); i++) {
                os << "{'name': " << ops.names[i] << ", 'params': [";
                for (size_t k = 0; k < ops.params[i].size(); k++) {
                    os << ops.params[i][k] << (k + 1 < ops.params[i].size() ? ", " : "");
                }
                os << "], 'inv': " << (ops.inverses[i] ? 1 : 0) << "}";
                if (i + 1 < ops.names.size()) {
                    os << ",";
                }
            }
            os << "]";
            return os.str();
        });

    using Adj = AdjointJacobian<PrecisionT>;
    py::class_<Adj>(m, ("AdjointJacobianKokkos_" + sfx).c_str(), py::module_local())
        .def(py::init<>())
        .def("create_ops_list",
             [make_ops](Adj &, const std::vector<std::string> &names, const std::vector<np_r> &params,
                        const std::vector<std::vector<size_t>> &wires,
                        const std::vector<bool> &inverses, const std::vector<np_c> &matrices) {
                 std::vector<std::vector<double>> p, mi;
                 for (auto &a : params) {
                     p.push_back(as_doubles<double>(a));
                 }
                 for (auto &a : matrices) {
                     mi.push_back(complex_doubles<PrecisionT>(a));
                 }
                 return make_ops(names, p, wires, inverses, mi);
             })
        .def("adjoint_jacobian",
             [](Adj &, const SV &sv, const std::vector<ObsPtr> &observables, const Ops &ops,
                const std::vector<size_t> &trainable) {
                 std::vector<const plk_obs *> hs;
                 for (auto &o : observables) {
                     hs.push_back(o->h.get());
                 }
                 const size_t nobs = hs.size(), ntp = trainable.size();
                 py::array_t<double> jac({nobs, ntp});
                 std::vector<double> buf(std::max<size_t>(1, nobs * ntp));
                 {
                     py::gil_scoped_release release;
                     chk(plk_adjoint_jacobian(sv.h, hs.data(), nobs, ops.h.get(), trainable.data(),
                                              ntp, 0, buf.data()));
                 }
                 std::copy(buf.begin(), buf.begin() + nobs * ntp, jac.mutable_data());
                 return jac;
             });
}

} // namespace

PYBIND11_MODULE(lightning_kokkos_qubit_ops, m) {
    py::register_exception<EngineError>(m, "PLException");
    bind_precision<float>(m);
    bind_precision<double>(m);

    auto config = []() {
        std::vector<char> buf(16384);
        chk(plk_print_configuration(buf.data(), buf.size()));
        return std::string(buf.data());
    };
    m.def("kokkos_start", []() {}, "No-op: the HIP runtime initialises on first use.");
    m.def("kokkos_end", []() {}, "No-op: resources are released with the objects.");
    m.def("print_configuration", config, "Device and build configuration.");
    m.def("kokkos_config_info", [config]() {
        py::dict d;
        d["Backend"] = py::dict(py::arg("Parallel") = std::string("HIP (gfx950)"));
        d["Configuration"] = config();
        return d;
    }, "Device and build configuration as a dictionary.");

    py::class_<Settings>(m, "InitializationSettings")
        .def(py::init<>())
#define PLK_FIELD(NAME, FIELD, FLAG, TYPE)                                                    \
    .def("get_" NAME, [](const Settings &s) { return s.FIELD; })                             \
        .def("has_" NAME, [](const Settings &s) { return s.FLAG; })                          \
        .def("set_" NAME, [](Settings &s, TYPE v) -> Settings & {                            \
            s.FIELD = v;                                                                     \
            s.FLAG = true;                                                                   \
            return s;                                                                        \
        }, py::return_value_policy::reference_internal)
        PLK_FIELD("num_threads", num_threads, h_num_threads, int)
        PLK_FIELD("device_id", device_id, h_device_id, int)
        PLK_FIELD("map_device_id_by", map_device_id_by, h_map, const std::string &)
        PLK_FIELD("disable_warnings", disable_warnings, h_disable, int)
        PLK_FIELD("print_configuration", print_configuration, h_print, int)
        PLK_FIELD("tune_internals", tune_internals, h_tune, int)
        PLK_FIELD("tools_libs", tools_libs, h_libs, const std::string &)
        PLK_FIELD("tools_help", tools_help, h_help, int)
        PLK_FIELD("tools_args", tools_args, h_args, const std::string &)
#undef PLK_FIELD
        .def("__repr__", [](const Settings &s) {
            std::ostringstream os;
            os << "InitializationSettings:\n"
               << "num_threads = " << s.num_threads << '\n'
               << "device_id = " << s.device_id << '\n'
               << "map_device_id_by = " << s.map_device_id_by << '\n'
               << "disable_warnings = " << s.disable_warnings << '\n'
               << "print_configuration = " << s.print_configuration << '\n'
               << "tune_internals = " << s.tune_internals << '\n'
               << "tools_libs = " << s.tools_libs << '\n'
               << "tools_help = " << s.tools_help << '\n'
               << "tools_args = " << s.tools_args;
            return os.str();
        });
}
