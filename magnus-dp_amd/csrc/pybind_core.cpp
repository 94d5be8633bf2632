// pybind_core.cpp -- the `_core` extension module of the MI355X build.
//
// Same functions, argument names, defaults and error texts as the
// reference's pybind module (src/api/pybind.cpp:437-548).  Host (numpy)
// arrays are staged through device memory and computed by
// libmagnus_hip.so's C ABI on the default stream; the GIL is released while
// the device works (pybind.cpp:171-188).  Computation is FP64: float32 input
// is computed in float64 and returned as float32; complex dtypes are not
// supported by this build.
#include <hip/hip_runtime.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstring>
#include <limits>
#include <optional>
#include <string>
#include <vector>

#include "magnus_capi.h"

namespace py = pybind11;

namespace {

using F64Array = py::array_t<double, py::array::c_style | py::array::forcecast>;

[[noreturn]] void raise_status(int rc) {
    const std::string msg = magnus_last_error();
    if (rc == MAGNUS_EINVAL) throw py::value_error(msg);
    if (rc == MAGNUS_ERANGE) throw py::index_error(msg);
    throw std::runtime_error(msg);
}

void check(int rc) {
    if (rc != MAGNUS_OK) raise_status(rc);
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct DeviceBuffer {
    double* ptr = nullptr;
    size_t count = 0;
    explicit DeviceBuffer(size_t n) : count(n) {
        if (n) check_hip(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(double)), "hipMalloc");
    }
    ~DeviceBuffer() {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void upload(const double* src) {
        if (count) check_hip(hipMemcpy(ptr, src, count * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void download(double* dst) const {
        if (count) check_hip(hipMemcpy(dst, ptr, count * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
};

// dtype policy: float64 native, float32 computed in float64, everything else TypeError
enum class Kind { F64, F32 };

Kind dtype_kind(const py::array& data) {
    const py::dtype dt = data.dtype();
    if (dt.is(py::dtype::of<double>())) return Kind::F64;
    if (dt.is(py::dtype::of<float>())) return Kind::F32;
    if (dt.is(py::dtype("complex64")) || dt.is(py::dtype("complex128")))
        throw py::type_error("magnus (MI355X build) supports dtypes float32 and float64; complex dtypes are not supported");
    throw py::type_error("magnus only supports dtypes float32, float64, complex64, and complex128");
}

py::array finish(const F64Array& a, Kind k) {
    if (k == Kind::F32) return a.attr("astype")("float32");
    return a;
}

F64Array as_f64(const py::handle& obj, const char* what) {
    F64Array t = F64Array::ensure(obj);
    if (!t) throw py::type_error(std::string("could not convert ") + what + " to the selected numeric dtype");
    return t;
}

int32_t resolve_op(const std::string& op) {
    int32_t v = 0;
    check(magnus_resolve_op(op.c_str(), &v));
    return v;
}

magnus_problem problem(size_t n, size_t dim, size_t samples, double t0, double tf, int32_t op, int32_t mat,
                       int32_t integ, int32_t gl) {
    magnus_problem p{};
    p.n = n;
    p.dim = dim;
    p.samples = samples;
    p.t0 = t0;
    p.tf = tf;
    p.op = op;
    p.matrix_backend = mat;
    p.integrator = integ;
    p.gl_backend = gl;
    return p;
}

// matrix_input_shape (composer/backends.hpp:108-122)
std::pair<size_t, size_t> matrix_shape(size_t n, const py::array& a) {
    if (n == 0) throw py::value_error("n must be at least 1");
    if (a.ndim() != 3) throw py::value_error("data must have shape (samples, dim, dim)");
    const size_t samples = (size_t)a.shape(0), rows = (size_t)a.shape(1), cols = (size_t)a.shape(2);
    if (samples < 2) throw py::value_error("data must contain at least two samples");
    if (rows == 0 || cols != rows) throw py::value_error("data must contain square matrices");
    return {samples, rows};
}

// spacecurve_input_shape (composer/backends.hpp:124-137)
size_t spacecurve_shape(size_t n, const py::array& a) {
    if (n == 0) throw py::value_error("n must be at least 1");
    if (a.ndim() != 2) throw py::value_error("data must have shape (samples, 3)");
    const size_t samples = (size_t)a.shape(0);
    if (samples < 2) throw py::value_error("data must contain at least two samples");
    if ((size_t)a.shape(1) != 3) throw py::value_error("data must have shape (samples, 3)");
    return samples;
}

void expect_shape(const py::array& a, const std::vector<size_t>& shape, const char* msg) {
    if ((size_t)a.ndim() != shape.size()) throw py::value_error(msg);
    for (size_t i = 0; i < shape.size(); ++i)
        if ((size_t)a.shape((py::ssize_t)i) != shape[i]) throw py::value_error(msg);
}

std::vector<py::ssize_t> carry_shape(size_t n, size_t samples, size_t dim) {
    return {(py::ssize_t)((n + 3) / 2), (py::ssize_t)(n - 1), (py::ssize_t)samples, (py::ssize_t)dim,
            (py::ssize_t)dim};
}

py::tuple names(const char* (*fn)(int32_t), bool with_auto) {
    std::vector<std::string> v;
    for (int32_t i = 0; fn(i); ++i) v.emplace_back(fn(i));
    if (with_auto) v.emplace_back("Auto");
    py::tuple t(v.size());
    for (size_t i = 0; i < v.size(); ++i) t[i] = py::str(v[i]);
    return t;
}

// --------------------------------------------------------------------------
py::object compute(size_t n, py::array data, double t0, double tf, const std::string& op,
                   const std::string& matrix_backend, const std::string& integrator, const std::string& gl_backend,
                   bool record_vjp) {
    const Kind kind = dtype_kind(data);
    F64Array in = as_f64(data, "data");
    auto [samples, dim] = matrix_shape(n, in);
    const int32_t opi = resolve_op(op);
    int32_t mat = 0, integ = 0, gl = 0;
    check(magnus_resolve_gl_backend(gl_backend.c_str(), &gl));
    check(magnus_resolve_integrator(integrator.c_str(), &integ));
    check(magnus_resolve_matrix_backend(matrix_backend.c_str(), &mat));
    magnus_problem p = problem(n, dim, samples, t0, tf, opi, mat, integ, gl);
    check(magnus_validate(&p, 0, 1));

    std::vector<py::ssize_t> out_shape =
        opi == MAGNUS_OP_MANY ? std::vector<py::ssize_t>{(py::ssize_t)n, (py::ssize_t)dim, (py::ssize_t)dim}
                              : std::vector<py::ssize_t>{(py::ssize_t)dim, (py::ssize_t)dim};
    F64Array out(out_shape);
    F64Array carry = record_vjp ? F64Array(carry_shape(n, samples, dim)) : F64Array();

    DeviceBuffer d_in(samples * dim * dim), d_out(magnus_output_elems(&p));
    std::optional<DeviceBuffer> d_carry;
    if (record_vjp) d_carry.emplace(magnus_carry_elems(&p));
    d_in.upload(in.data());
    int rc;
    {
        py::gil_scoped_release release;
        rc = magnus_forward_f64(&p, 1, d_in.ptr, d_out.ptr, record_vjp ? d_carry->ptr : nullptr, nullptr);
        if (rc == MAGNUS_OK) {
            d_out.download(out.mutable_data());
            if (record_vjp) d_carry->download(carry.mutable_data());
        }
    }
    check(rc);
    if (record_vjp) return py::make_tuple(finish(out, kind), finish(carry, kind));
    return finish(out, kind);
}

py::object compute_sc(size_t n, py::array data, double t0, double tf, const std::string& op,
                      const std::string& integrator, const std::string& gl_backend, bool record_vjp) {
    const Kind kind = dtype_kind(data);
    F64Array in = as_f64(data, "data");
    const size_t samples = spacecurve_shape(n, in);
    const int32_t opi = resolve_op(op);
    int32_t integ = 0, gl = 0;
    check(magnus_resolve_gl_backend(gl_backend.c_str(), &gl));
    check(magnus_resolve_integrator(integrator.c_str(), &integ));
    magnus_problem p = problem(n, 2, samples, t0, tf, opi, MAGNUS_MAT_SPACECURVE, integ, gl);
    check(magnus_validate(&p, 1, 1));

    std::vector<py::ssize_t> out_shape = opi == MAGNUS_OP_MANY
                                             ? std::vector<py::ssize_t>{(py::ssize_t)n, 3}
                                             : std::vector<py::ssize_t>{3};
    F64Array out(out_shape);
    F64Array carry = record_vjp ? F64Array(carry_shape(n, samples, 2)) : F64Array();
    DeviceBuffer d_in(samples * 3), d_out(magnus_sc_output_elems(&p));
    std::optional<DeviceBuffer> d_carry;
    if (record_vjp) d_carry.emplace(magnus_sc_carry_elems(&p));
    d_in.upload(in.data());
    int rc;
    {
        py::gil_scoped_release release;
        rc = magnus_sc_forward_f64(&p, 1, d_in.ptr, d_out.ptr, record_vjp ? d_carry->ptr : nullptr, nullptr);
        if (rc == MAGNUS_OK) {
            d_out.download(out.mutable_data());
            if (record_vjp) d_carry->download(carry.mutable_data());
        }
    }
    check(rc);
    if (record_vjp) return py::make_tuple(finish(out, kind), finish(carry, kind));
    return finish(out, kind);
}

py::array compute_vjp(size_t n, py::array data, py::array cotangent, py::object vjp_data, double t0, double tf,
                      const std::string& op, const std::string& matrix_backend, const std::string& integrator) {
    const Kind kind = dtype_kind(data);
    F64Array in = as_f64(data, "data");
    F64Array cot = as_f64(cotangent, "cotangent");
    auto [samples, dim] = matrix_shape(n, in);
    const int32_t opi = resolve_op(op);
    if (opi == MAGNUS_OP_MANY) expect_shape(cot, {n, dim, dim}, "many output must have shape (n, dim, dim)");
    else expect_shape(cot, {dim, dim}, "one/sum output must have shape (dim, dim)");
    std::optional<F64Array> carry;
    if (!vjp_data.is_none()) {
        carry.emplace(as_f64(vjp_data, "VJP data"));
        std::vector<size_t> cs;
        for (auto v : carry_shape(n, samples, dim)) cs.push_back((size_t)v);
        expect_shape(*carry, cs, "VJP data has an invalid shape");
    }
    int32_t mat = 0, integ = 0;
    check(magnus_resolve_integrator(integrator.c_str(), &integ));
    check(magnus_resolve_matrix_backend(matrix_backend.c_str(), &mat));
    magnus_problem p = problem(n, dim, samples, t0, tf, opi, mat, integ, MAGNUS_AUTO);
    check(magnus_validate(&p, 0, 0));

    F64Array out({(py::ssize_t)samples, (py::ssize_t)dim, (py::ssize_t)dim});
    DeviceBuffer d_in(samples * dim * dim), d_cot(magnus_output_elems(&p)), d_out(samples * dim * dim);
    std::optional<DeviceBuffer> d_carry;
    d_in.upload(in.data());
    d_cot.upload(cot.data());
    if (carry) {
        d_carry.emplace(magnus_carry_elems(&p));
        d_carry->upload(carry->data());
    }
    int rc;
    {
        py::gil_scoped_release release;
        rc = magnus_vjp_f64(&p, 1, d_in.ptr, d_cot.ptr, carry ? d_carry->ptr : nullptr, d_out.ptr, nullptr);
        if (rc == MAGNUS_OK) d_out.download(out.mutable_data());
    }
    check(rc);
    return finish(out, kind);
}

py::array compute_sc_vjp(size_t n, py::array data, py::array cotangent, py::object vjp_data, double t0,
                         double tf, const std::string& op, const std::string& integrator) {
    const Kind kind = dtype_kind(data);
    F64Array in = as_f64(data, "data");
    F64Array cot = as_f64(cotangent, "cotangent");
    const size_t samples = spacecurve_shape(n, in);
    const int32_t opi = resolve_op(op);
    if (opi == MAGNUS_OP_MANY) expect_shape(cot, {n, 3}, "many spacecurve output must have shape (n, 3)");
    else expect_shape(cot, {3}, "one/sum spacecurve output must have shape (3,)");
    std::optional<F64Array> carry;
    if (!vjp_data.is_none()) {
        carry.emplace(as_f64(vjp_data, "VJP data"));
        std::vector<size_t> cs;
        for (auto v : carry_shape(n, samples, 2)) cs.push_back((size_t)v);
        expect_shape(*carry, cs, "VJP data has an invalid shape");
    }
    int32_t integ = 0;
    check(magnus_resolve_integrator(integrator.c_str(), &integ));
    magnus_problem p = problem(n, 2, samples, t0, tf, opi, MAGNUS_MAT_SPACECURVE, integ, MAGNUS_AUTO);
    check(magnus_validate(&p, 1, 0));

    F64Array out({(py::ssize_t)samples, (py::ssize_t)3});
    DeviceBuffer d_in(samples * 3), d_cot(magnus_sc_output_elems(&p)), d_out(samples * 3);
    std::optional<DeviceBuffer> d_carry;
    d_in.upload(in.data());
    d_cot.upload(cot.data());
    if (carry) {
        d_carry.emplace(magnus_sc_carry_elems(&p));
        d_carry->upload(carry->data());
    }
    int rc;
    {
        py::gil_scoped_release release;
        rc = magnus_sc_vjp_f64(&p, 1, d_in.ptr, d_cot.ptr, carry ? d_carry->ptr : nullptr, d_out.ptr, nullptr);
        if (rc == MAGNUS_OK) d_out.download(out.mutable_data());
    }
    check(rc);
    return finish(out, kind);
}

// _replace_gl_table (pybind.cpp:61-116)
void replace_gl_table(size_t max_order, py::array weights, py::array nodes) {
    if (max_order == 0) throw py::value_error("max_order must be at least 1");
    size_t a = max_order, b = max_order + 1;
    if (a % 2 == 0) a /= 2;
    else b /= 2;
    if (a != 0 && b > std::numeric_limits<size_t>::max() / a) throw py::value_error("max_order is too large");
    const size_t expected = a * b;
    auto copy = [&](py::array arr, const char* name) {
        F64Array t = F64Array::ensure(arr);
        if (!t) throw py::type_error("GL data must be convertible to float64");
        if (t.ndim() != 1 || (size_t)t.shape(0) != expected)
            throw py::value_error(std::string("GL ") + name + " values must have shape (" + std::to_string(expected) +
                                  ",)");
        return std::vector<double>(t.data(), t.data() + expected);
    };
    std::vector<double> w = copy(weights, "weights");
    std::vector<double> x = copy(nodes, "nodes");
    check(magnus_replace_gl_table(max_order, w.data(), x.data()));
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "Python bindings for the magnus MI355X (HIP) library";
    m.def("max_order", []() { return (size_t)magnus_max_order(); },
          "Return the max available Gauss-Legendre order. Max magnus order is then any n satisfying (n + 3) / 2 <= "
          "max_order().");
    m.def("numeric_backends", []() { return names(magnus_numeric_backend_name, false); },
          "Return the available numeric backend dispatch names.");
    m.def("matrix_backends", []() { return names(magnus_matrix_backend_name, true); },
          "Return the available matrix backend dispatch names.");
    m.def("integrators", []() { return names(magnus_integrator_name, true); },
          "Return the available integrator dispatch names.");
    m.def("gl_backends", []() { return names(magnus_gl_backend_name, true); },
          "Return the available Gauss-Legendre execution backend dispatch names.");
    m.def("ops", []() { return names(magnus_op_name, false); }, "Return the available kernel operation dispatch names.");
    m.def("_replace_gl_table", &replace_gl_table, py::arg("max_order"), py::arg("weights"), py::arg("nodes"),
          "Replace the process-global Gauss-Legendre table from flattened float64 weights and nodes.");
    m.def("compute", &compute, py::arg("n"), py::arg("data"), py::arg("t0"), py::arg("tf"), py::arg("op") = "sum",
          py::arg("matrix_backend") = "Auto", py::arg("integrator") = "Auto", py::arg("gl_backend") = "Auto",
          py::arg("record_vjp") = false, "Compute a Magnus operation from sampled matrix data.");
    m.def("compute_sc", &compute_sc, py::arg("n"), py::arg("data"), py::arg("t0"), py::arg("tf"),
          py::arg("op") = "sum", py::arg("integrator") = "Auto", py::arg("gl_backend") = "Auto",
          py::arg("record_vjp") = false, "Compute a SpaceCurve Magnus operation from sampled 3-vector data.");
    m.def("compute_vjp", &compute_vjp, py::arg("n"), py::arg("data"), py::arg("cotangent"),
          py::arg("vjp_data") = py::none(), py::arg("t0") = 0.0, py::arg("tf") = 1.0, py::arg("op") = "sum",
          py::arg("matrix_backend") = "Auto", py::arg("integrator") = "Auto",
          "Compute the VJP of a Magnus operation with respect to sampled matrix data.");
    m.def("compute_sc_vjp", &compute_sc_vjp, py::arg("n"), py::arg("data"), py::arg("cotangent"),
          py::arg("vjp_data") = py::none(), py::arg("t0") = 0.0, py::arg("tf") = 1.0, py::arg("op") = "sum",
          py::arg("integrator") = "Auto",
          "Compute the VJP of a SpaceCurve Magnus operation with respect to sampled vector data.");
    m.def("last_kernel", []() { return std::string(magnus_last_kernel()); },
          "Kernel family used by the last call on this thread (diagnostics).");
}
