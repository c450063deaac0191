// Projection of the geocentre polarizations onto detectors: antenna patterns,
// arrival-time delays and the detector-frame signal.
//
// Reference: BBHWaveformTD.get_detector_information (eob/seobnrv4phm.py:597-634)
// and the "detector_fd" projection of BBHWaveformTD.__call__ (:742-767).  The
// detector tensors and vertices are the caller's (the reference reads bilby's
// interferometer definitions), and so is the Greenwich mean sidereal time of
// each event (the reference takes it from lal).
#include <hip/hip_runtime.h>

#include "eobhip_common.hpp"

using namespace eobhip_detail;

namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.141592653589793;
constexpr double kSpeedOfLight = 299792458.0;   // m/s

// Per (binary b): theta = pi/2 - dec, phi = ra - gmst, the wave frame
// u = (cos phi cos theta, cos theta sin phi, -sin theta), v = (-sin phi, cos phi, 0),
// m = -u sin psi - v cos psi, n = -u cos psi + v sin psi,
// e+ = m m - n n, ex = m n + n m; F+ = D : e+, Fx = D : ex per detector tensor D;
// the delay omega . vertex / c with omega = (sin theta cos phi, sin theta sin phi, cos theta).
__global__ __launch_bounds__(kBlock) void k_detector_response(const double* ra, const double* dec, const double* gmst, const double* psi,
                                                              int n, const double* tensors, const double* vertices, int nifo,
                                                              double* Fp, double* Fc, double* tshift) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        const double phi = ra[b] - gmst[b], theta = 0.5 * kPi - dec[b], ps = psi[b];
        double sph, cph, sth, cth, sps, cps;
        sincos(phi, &sph, &cph);
        sincos(theta, &sth, &cth);
        sincos(ps, &sps, &cps);
        const double u[3] = {cph * cth, cth * sph, -sth};
        const double v[3] = {-sph, cph, 0.0};
        double m[3], nn[3];
        for (int i = 0; i < 3; ++i) {
            m[i] = -u[i] * sps - v[i] * cps;
            nn[i] = -u[i] * cps + v[i] * sps;
        }
        const double om[3] = {sth * cph, sth * sph, cth};
        for (int d = 0; d < nifo; ++d) {
            const double* D = tensors + 9 * d;
            double fp = 0.0, fc = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    fp += D[3 * i + j] * (m[i] * m[j] - nn[i] * nn[j]);
                    fc += D[3 * i + j] * (m[i] * nn[j] + nn[i] * m[j]);
                }
            const double* V = vertices + 3 * d;
            Fp[(size_t)d * n + b] = fp;
            Fc[(size_t)d * n + b] = fc;
            tshift[(size_t)d * n + b] = (om[0] * V[0] + om[1] * V[1] + om[2] * V[2]) / kSpeedOfLight;
        }
    }
}

// out[b][d][k] = (F+_db h+~_bk + Fx_db hx~_bk) exp(-2 pi i f_k dt_db), f_k = k df,
// zero below k_min (the reference zeroes the templates below bilby's start bin).
// grid (blocks, binaries, ceil(nifo / kDetGroup)); a thread reads h+~ and hx~ of
// its bins once for up to kDetGroup interferometers, and takes bins k0, k0 + S,
// k0 + 2S, ... (S = the grid's threads per binary, so a wavefront's accesses are
// consecutive).  The phase e^{-2 pi i f_k dt} comes from one sincos for the
// thread's first bin and one for the step S df, then a product per bin
// (about F / S products, each within an ulp or two).
constexpr int kDetGroup = 4;
constexpr int kDetBinsPerThread = 32;
__global__ __launch_bounds__(kBlock) void k_detector_signal(const double2* __restrict__ hp, const double2* __restrict__ hc,
                                                            const double* __restrict__ Fp, const double* __restrict__ Fc,
                                                            const double* __restrict__ tshift, int n, int nifo, int F, double df,
                                                            int k_min, double2* __restrict__ out) {
    const int b = blockIdx.y, d0 = blockIdx.z * kDetGroup;
    const int nd = nifo - d0 < kDetGroup ? nifo - d0 : kDetGroup;
    const long long S = (long long)gridDim.x * blockDim.x;
    const long long k0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double fp[kDetGroup], fc[kDetGroup], er[kDetGroup], ei[kDetGroup], rr[kDetGroup], ri[kDetGroup];
#pragma unroll
    for (int d = 0; d < kDetGroup; ++d) {
        fp[d] = fc[d] = 0.0;
        er[d] = rr[d] = 1.0;
        ei[d] = ri[d] = 0.0;
        if (d < nd) {
            const size_t q = (size_t)(d0 + d) * n + b;
            fp[d] = Fp[q];
            fc[d] = Fc[q];
            const double ts = tshift[q];
            sincos(-2.0 * kPi * ((double)k0 * df) * ts, &ei[d], &er[d]);
            sincos(-2.0 * kPi * ((double)S * df) * ts, &ri[d], &rr[d]);
        }
    }
    const double2* p = hp + (size_t)b * F;
    const double2* c = hc + (size_t)b * F;
    for (long long k = k0; k < F; k += S) {
        const double2 pk = p[k], ck = c[k];
#pragma unroll
        for (int d = 0; d < kDetGroup; ++d) {
            if (d < nd) {
                double2 r = make_double2(0.0, 0.0);
                if (k >= k_min) {
                    const double sr = fp[d] * pk.x + fc[d] * ck.x, si = fp[d] * pk.y + fc[d] * ck.y;
                    r = make_double2(sr * er[d] - si * ei[d], sr * ei[d] + si * er[d]);
                }
                out[((size_t)b * nifo + d0 + d) * F + k] = r;
                const double t = er[d] * rr[d] - ei[d] * ri[d];
                ei[d] = er[d] * ri[d] + ei[d] * rr[d];
                er[d] = t;
            }
        }
    }
}

}  // namespace

extern "C" {

int eobhip_detector_response(const double* ra, const double* dec, const double* gmst, const double* psi, int n, const double* tensors,
                             const double* vertices, int nifo, double* Fplus, double* Fcross, double* time_shift, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && nifo >= 0, "bad sizes");
    if (n == 0 || nifo == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_detector_response, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), ra, dec, gmst, psi, n,
                       tensors, vertices, nifo, Fplus, Fcross, time_shift);
    return check_launch("detector_response");
}

int eobhip_detector_signal(const double* hp_fd, const double* hc_fd, const double* Fplus, const double* Fcross, const double* time_shift,
                           int n, int nifo, int F, double df, int k_min, double* out, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && nifo >= 0 && F >= 0 && k_min >= 0, "bad sizes");
    EOBHIP_REQUIRE(n <= 65535 && nifo <= 65535, "n and nifo must be <= 65535");
    if (n == 0 || nifo == 0 || F == 0) return EOBHIP_OK;
    const long long span = (long long)kBlock * kDetBinsPerThread;
    const int bx = (int)((F + span - 1) / span);
    const int bz = (nifo + kDetGroup - 1) / kDetGroup;
    hipLaunchKernelGGL(k_detector_signal, dim3(bx, n, bz), dim3(kBlock), 0, as_stream(stream), reinterpret_cast<const double2*>(hp_fd),
                       reinterpret_cast<const double2*>(hc_fd), Fplus, Fcross, time_shift, n, nifo, F, df, k_min,
                       reinterpret_cast<double2*>(out));
    return check_launch("detector_signal");
}

}  // extern "C"
