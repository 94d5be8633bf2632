// kernels.hpp -- argument blocks and launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace magnus {

constexpr int kArgMaxQ = 64;

// Entry-parallel forward (kernels_ep.hip).  One job = (instance, GL node).
struct EpFwdArgs {
    const double* in;    // [B][S][m_in]
    double* carry;       // [B][Q][n-1][S][m] or nullptr
    double* pingpong;    // [jobs][2][S][m] level buffers when carry == nullptr
    double* partial;     // [jobs][n_out][m]
    double* omega1;      // [B][m]
    long long S;
    long long jobs;
    int n, d, m, m_in, Q, Qj, n_out, op;
    int sc;       // su(2) 4-vector product
    int outputs;  // 0: record the carry only (VJP::none)
    int tpj, jpw; // threads per job, jobs per workgroup
    double dt;
    double w[kArgMaxQ];
    double x[kArgMaxQ];
};

struct FwdReduceArgs {
    const double* partial;  // [B][Qj][n_out][m]
    const double* omega1;   // [B][m]
    double* out;            // [B][n_out][m_out]
    long long B;
    int n, m, m_out, n_out, Qj, op, sc;
};

struct EpVjpArgs {
    const double* in;     // [B][S][m_in]
    const double* cot;    // [B][n_out][m_cot]
    const double* carry;  // [B][Q][n-1][S][m]
    double* gbuf;         // [jobs][2][S][m]
    double* dAjob;        // [jobs][S][m]
    long long S;
    long long jobs;
    int n, d, m, m_in, m_cot, Q, n_out, op;
    int sc;
    int bt_fused;  // Fixed2/SpaceCurve bar_total = temp * (1 + x); generic = temp + temp * x
    int tpj, jpw;
    double dt;
    double w[kArgMaxQ];
    double x[kArgMaxQ];
};

struct VjpReduceArgs {
    const double* dAjob;  // [B][Qj][S][m]
    const double* cot;    // [B][n_out][m_cot]
    double* dA;           // [B][S][m_out]
    long long B;
    long long S;
    int n, m, m_out, m_cot, n_out, Qj, op, sc;
    double dt;
};

// Per-kernel timing hooks (implemented by the C-ABI layer).  When timing is
// enabled, timer_begin records a start event on `s` and returns a handle that
// timer_end closes; both are no-ops (nullptr) otherwise.
void* timer_begin(const char* name, hipStream_t s);
void timer_end(void* handle);

hipError_t launch_ep_forward(int integ, const EpFwdArgs& a, int block, long long grid, int ept, hipStream_t s);
hipError_t launch_ep_vjp(int integ, const EpVjpArgs& a, int block, long long grid, int ept, hipStream_t s);
hipError_t launch_fwd_reduce(const FwdReduceArgs& a, hipStream_t s);
hipError_t launch_vjp_reduce(int integ, const VjpReduceArgs& a, hipStream_t s);

}  // namespace magnus
