// Host build of the right-hand side's arithmetic helpers
// (eobgpu_amd/csrc/eob_fastmath.hpp).  log_pos writes every multiply-add as
// fma() with contraction off, so this host build rounds as the device build does;
// the division and square-root helpers start here from single-precision
// estimates (coarser than the hardware's), so this host build tests their refinement.
#include "../eobgpu_amd/csrc/eob_fastmath.hpp"

extern "C" void probe_log_pos(const double* x, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = eob::log_pos(x[i]);
}
extern "C" void probe_qdiv(const double* a, const double* b, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = eob::qdiv(a[i], b[i]);
}
extern "C" void probe_qrcp(const double* b, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = eob::qrcp(b[i]);
}
extern "C" void probe_qsqrt(const double* x, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = eob::qsqrt(x[i]);
}
extern "C" void probe_cbrt_pos(const double* x, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = eob::cbrt_pos(x[i]);
}
extern "C" void probe_qsqrt_recip(const double* x, double* r, double* inv, int n) {
    for (int i = 0; i < n; ++i) eob::qsqrt_recip(x[i], r[i], inv[i]);
}
