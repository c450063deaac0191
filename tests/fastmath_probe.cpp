// Host build of the device path's log_pos (eobgpu_amd/csrc/eob_fastmath.hpp):
// the header writes every multiply-add as fma() and switches contraction off,
// so this build rounds exactly as the device build does.
#include "../eobgpu_amd/csrc/eob_fastmath.hpp"

extern "C" void probe_log_pos(const double* x, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = eob::log_pos(x[i]);
}
