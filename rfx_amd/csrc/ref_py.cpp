// pybind11 module rfx_amd._rfxref: the reference-signature C++ entry points
// of rfx_ref.hpp for Python callers and tests (numpy arrays in and out).

#include <hip/hip_runtime_api.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>

#include "rfx_ref.hpp"

namespace py = pybind11;
using rfx::ref::integer_t;

namespace {

using I32 = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using F64F = py::array_t<double, py::array::f_style | py::array::forcecast>;

void need(bool ok, const char* what) {
  if (!ok) throw std::invalid_argument(what);
}

// prox: n x n float64 Fortran-ordered (column-major, as the reference), updated in place.
// nnode defaults to the length of nodestatus (the reference's callers pass maxnode).
void per_tree_proximity(bool gpu, I32 nodestatus, I32 nodextr, I32 nin, py::array_t<double> prox, integer_t nnode) {
  const integer_t n = static_cast<integer_t>(nodextr.size());
  need(nin.size() == n && prox.ndim() == 2 && prox.shape(0) == n && prox.shape(1) == n, "shape mismatch");
  need((prox.flags() & py::array::f_style) != 0 && prox.writeable(), "prox must be a writable Fortran-ordered array");
  if (nnode <= 0) nnode = static_cast<integer_t>(nodestatus.size());
  need(nnode <= nodestatus.size(), "nnode exceeds nodestatus");
  auto f = gpu ? &rfx::ref::gpu_proximity : &rfx::ref::cpu_proximity;
  f(nodestatus.data(), nodextr.data(), nin.data(), n, nnode, prox.mutable_data(), nullptr, nullptr, nullptr, nullptr,
    nullptr, nullptr);
}

py::array_t<uint16_t> upper_triangle_fp16(I32 nodestatus, I32 nodextr, I32 nin) {
  const integer_t n = static_cast<integer_t>(nodextr.size());
  need(nin.size() == n, "shape mismatch");
  const size_t len = static_cast<size_t>(n) * (n + 1) / 2;
  void* d = nullptr;
  if (hipMalloc(&d, 2 * len) != hipSuccess) throw std::runtime_error("hipMalloc failed");
  py::array_t<uint16_t> out(len);
  try {
    rfx::ref::gpu_proximity_upper_triangle_fp16(nodestatus.data(), nodextr.data(), nin.data(), n,
                                                static_cast<integer_t>(nodestatus.size()),
                                                static_cast<__half*>(d), nullptr, nullptr, nullptr, nullptr,
                                                nullptr, nullptr);
    if (hipMemcpy(out.mutable_data(), d, 2 * len, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  (void)hipFree(d);
  return out;
}

F64F finishprox(I32 nout, F64F prox) {
  const integer_t n = static_cast<integer_t>(nout.size());
  need(prox.ndim() == 2 && prox.shape(0) == n && prox.shape(1) == n, "shape mismatch");
  F64F out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(n)});
  rfx::ref::cpu_finishprox(n, nout.data(), const_cast<double*>(prox.data()), out.mutable_data());
  return out;
}

// nin, nodextr: [T, n] int32; returns the n x n float64 RF-GAP matrix.
F64F proximity_rfgap(I32 nin, I32 nodextr) {
  need(nin.ndim() == 2 && nodextr.ndim() == 2 && nin.shape(0) == nodextr.shape(0) &&
           nin.shape(1) == nodextr.shape(1),
       "nin and nodextr must be [T, n] arrays of one shape");
  const integer_t T = static_cast<integer_t>(nin.shape(0)), n = static_cast<integer_t>(nin.shape(1));
  F64F out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(n)});
  rfx::ref::gpu_proximity_rfgap(T, n, nin.data(), nodextr.data(), out.mutable_data());
  return out;
}

py::array_t<float> outlier_scores(F64F prox) {
  need(prox.ndim() == 2 && prox.shape(0) == prox.shape(1), "prox must be square");
  const integer_t n = static_cast<integer_t>(prox.shape(0));
  std::vector<float> s = rfx::ref::compute_outlier_scores(prox.data(), n);
  py::array_t<float> out(static_cast<py::ssize_t>(n));
  std::copy(s.begin(), s.end(), out.mutable_data());
  return out;
}

py::array_t<double> compute_mds_3d(F64F prox, py::object oob_counts, bool use_rfgap) {
  need(prox.ndim() == 2 && prox.shape(0) == prox.shape(1), "prox must be square");
  const integer_t n = static_cast<integer_t>(prox.shape(0));
  I32 oc;
  const integer_t* ocp = nullptr;
  if (!oob_counts.is_none()) {
    oc = oob_counts.cast<I32>();
    need(oc.size() == n, "oob_counts length");
    ocp = oc.data();
  }
  std::vector<double> c = rfx::ref::compute_mds_3d_cpu(prox.data(), n, true, ocp, use_rfgap);
  py::array_t<double> out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(3)});
  std::copy(c.begin(), c.end(), out.mutable_data());
  return out;
}

class LowRank {
 public:
  LowRank(integer_t n, integer_t rank, int32_t format) : m_(n, rank, static_cast<rfx::ref::FactorFormat>(format)) {
    if (!m_.initialize()) throw std::runtime_error("initialize failed");
  }
  void add_tree(I32 nodestatus, I32 nodextr, I32 nin) {
    need(nodextr.size() == m_.get_n_samples() && nin.size() == m_.get_n_samples(), "shape mismatch");
    m_.add_tree_contribution(nodestatus.data(), nodextr.data(), nin.data(),
                             static_cast<integer_t>(nodestatus.size()));
  }
  void finalize() { m_.finalize_accumulation(0, true); }
  py::tuple factors() {
    double *A = nullptr, *B = nullptr;
    integer_t r = 0;
    if (!rfx::ref::get_lowrank_factors_host(&m_, &A, &B, &r, m_.get_n_samples()))
      throw std::runtime_error("get_lowrank_factors_host failed");
    const py::ssize_t n = m_.get_n_samples();
    // column-major n x r -> numpy Fortran arrays
    F64F a({n, static_cast<py::ssize_t>(r)}), b({n, static_cast<py::ssize_t>(r)});
    std::copy(A, A + n * r, a.mutable_data());
    std::copy(B, B + n * r, b.mutable_data());
    delete[] A;
    delete[] B;
    return py::make_tuple(a, b, r);
  }
  py::array_t<double> mds(integer_t k) {
    std::vector<double> c = rfx::ref::compute_mds_from_factors_host(&m_, k);
    if (c.empty()) throw std::runtime_error("compute_mds_from_factors_host failed");
    py::array_t<double> out({static_cast<py::ssize_t>(m_.get_n_samples()), static_cast<py::ssize_t>(k)});
    std::copy(c.begin(), c.end(), out.mutable_data());
    return out;
  }
  std::vector<double> eigenvalues() const { return m_.top_eigenvalues(); }
  // reconstruct_proximity_matrix_host: n x n float64 (Fortran order)
  F64F reconstruct() {
    double* P = nullptr;
    if (!rfx::ref::reconstruct_proximity_matrix_host(&m_, &P, m_.get_n_samples()))
      throw std::runtime_error("reconstruct_proximity_matrix_host failed");
    const py::ssize_t n = m_.get_n_samples();
    F64F out({n, n});
    std::copy(P, P + n * n, out.mutable_data());
    delete[] P;
    return out;
  }
  // reconstruct_full_matrix in another storage level, into device memory, copied back
  F64F reconstruct_level(int32_t level) {
    const py::ssize_t n = m_.get_n_samples();
    void* d = nullptr;
    if (hipMalloc(&d, 8 * static_cast<size_t>(n) * n) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    F64F out({n, n});
    try {
      m_.reconstruct_full_matrix(static_cast<double*>(d), static_cast<integer_t>(n),
                                 static_cast<rfx::ref::FactorFormat>(level));
      if (hipMemcpy(out.mutable_data(), d, 8 * static_cast<size_t>(n) * n, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
    } catch (...) {
      (void)hipFree(d);
      throw;
    }
    (void)hipFree(d);
    return out;
  }
  void add_fp16_triangle(py::array_t<uint16_t> tri) {
    m_.add_tree_contribution_incremental_upper_triangle_fp16(reinterpret_cast<const __half*>(tri.data()),
                                                             m_.get_n_samples());
  }

 private:
  rfx::ref::LowRankProximityMatrix m_;
};

}  // namespace

PYBIND11_MODULE(_rfxref, m) {
  m.doc() = "Reference-signature C++ entry points of the MI355X proximity path";
  m.def("gpu_proximity",
        [](I32 s, I32 x, I32 b, py::array_t<double> p, integer_t nnode) { per_tree_proximity(true, s, x, b, p, nnode); },
        py::arg("nodestatus"), py::arg("nodextr"), py::arg("nin"), py::arg("prox"), py::arg("nnode") = 0,
        "rf::gpu_proximity semantics, accumulated into prox (n x n, column-major)");
  m.def("cpu_proximity",
        [](I32 s, I32 x, I32 b, py::array_t<double> p, integer_t nnode) { per_tree_proximity(false, s, x, b, p, nnode); },
        py::arg("nodestatus"), py::arg("nodextr"), py::arg("nin"), py::arg("prox"), py::arg("nnode") = 0,
        "rf::cpu_proximity semantics, accumulated into prox (n x n, column-major)");
  m.def("set_casewise", [](bool on) { rfx::ref::g_config.use_casewise = on; }, "rfx::ref::g_config.use_casewise");
  m.def("get_casewise", []() { return rfx::ref::g_config.use_casewise; });
  m.def("upper_triangle_fp16", &upper_triangle_fp16, "one tree's OOB fold as IEEE half bits, packed upper");
  m.def("finishprox", &finishprox, "rf::cpu_finishprox semantics");
  m.def("proximity_rfgap", &proximity_rfgap, "rf::gpu_proximity_rfgap semantics over all trees");
  m.def("outlier_scores", &outlier_scores, "RandomForest::compute_outlier_scores semantics");
  m.def("compute_mds_3d", &compute_mds_3d, py::arg("prox"), py::arg("oob_counts") = py::none(),
        py::arg("use_rfgap") = false);
  py::class_<LowRank>(m, "LowRankProximityMatrix")
      .def(py::init<integer_t, integer_t, int32_t>())
      .def("add_tree", &LowRank::add_tree)
      .def("finalize", &LowRank::finalize)
      .def("factors", &LowRank::factors)
      .def("mds", &LowRank::mds, py::arg("k") = 3)
      .def("eigenvalues", &LowRank::eigenvalues)
      .def("reconstruct", &LowRank::reconstruct)
      .def("reconstruct_level", &LowRank::reconstruct_level)
      .def("add_fp16_triangle", &LowRank::add_fp16_triangle);
}
