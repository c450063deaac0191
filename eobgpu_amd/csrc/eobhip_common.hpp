// Host-side plumbing shared by the C-ABI translation units: error reporting and
// launch helpers.  No entry point aborts the process; failures return a status
// and leave a message for eobhip_last_error().
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdarg>
#include "../../include/eobhip.h"

namespace eobhip_detail {

void set_error(const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return EOBHIP_EHIP;
    }
    return EOBHIP_OK;
}

inline int hip_status(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? EOBHIP_ENOMEM : EOBHIP_EHIP;
    }
    return EOBHIP_OK;
}

inline unsigned blocks_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// Row-batched real FFTs through cached hipFFT plans (spline_kernels.hip); rows go
// through in chunks of kFftChunk and then power-of-two chunks.
// inverse = 0: n rows of N doubles -> n rows of N/2+1 complex (D2Z, unnormalised);
// inverse = 1: n rows of N/2+1 complex -> n rows of N doubles (Z2D, unnormalised).
constexpr int kFftChunk = 64;
int fft_exec(int inverse, const void* in, void* out, int rows, int N, void* stream);
int fft_rows(int inverse, const void* in, void* out, int n, int N, void* stream);

}  // namespace eobhip_detail

#define EOBHIP_REQUIRE(cond, ...)                        \
    do {                                                 \
        if (!(cond)) {                                   \
            eobhip_detail::set_error(__VA_ARGS__);       \
            return EOBHIP_EINVAL;                        \
        }                                                \
    } while (0)
