// Device-side building blocks: per-binary data, the lane-group ODE right-hand
// side, the DOP853 attempted step and the per-binary integration state machine
// shared by the persistent (work-stealing) and the plain (one launch per step)
// integrators.
//
// Lane groups.  A binary is integrated by a group of G consecutive lanes of one
// wavefront (G = 1, 4, 8, 16, 32 or 64).  Inside the right-hand side
//   - the 13 pieces of H's u-dependence are evaluated on different lanes and
//     gathered; H, dH/dr, dH/dpr, dH/dpphi and dH/dpphi at pr = 0 follow in
//     closed form on every lane (the reference evaluates two full gradients
//     and a third H per RHS, src/EOB.cu:3102, :2612, :2619);
//   - the 35 flux modes are spread over the G lanes and summed across them;
//   - lane j < 4 owns state component j of the DOP853 stages.
// With G = 1 one lane does all of it (the reference's lane-per-binary layout).
#pragma once
#include "eob_ham.hpp"
#include "eob_flux.hpp"

namespace eob {

struct BinData {
    HamCoef h;
    WaveCoef w;
    double eta_f;     // eta as the flux sees it (clamped to 1/4 within 1e-4)
    double nu;        // mu / M as RR_force computes it
    double flux_scale;   // 1 / (8 pi nu): the flux sum's 1/(8 pi) and RR_force's 1/nu as one product
    PieceCoef pc;     // H's u-dependence in the common piece form (piece_setup)
};

constexpr int kNAdd = 80;   // rows of additionalArgs

// DOP853 tableau (Hairer & Wanner; same published coefficients as
// scipy.integrate DOP853 and src/DOPR853.cu:11-185)
__constant__ static double dC[12] = {0.0, 0.05260015195876773, 0.0789002279381516, 0.1183503419072274, 0.2816496580927726, 0.3333333333333333, 0.25, 0.3076923076923077, 0.6512820512820513, 0.6, 0.8571428571428571, 1.0};
#define DOP_A                                                                                                                                   \
    {{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                                                                              \
     {0.05260015195876773, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                                                              \
     {0.0197250569845379, 0.0591751709536137, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                                               \
     {0.02958758547680685, 0.0, 0.08876275643042054, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                                             \
     {0.2413651341592667, 0.0, -0.8845494793282861, 0.924834003261792, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                                \
     {0.037037037037037035, 0.0, 0.0, 0.17082860872947386, 0.12546768756682242, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                           \
     {0.037109375, 0.0, 0.0, 0.17025221101954405, 0.06021653898045596, -0.017578125, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                           \
     {0.03709200011850479, 0.0, 0.0, 0.17038392571223998, 0.10726203044637328, -0.015319437748624402, 0.008273789163814023, 0.0, 0.0, 0.0, 0.0, 0.0}, \
     {0.6241109587160757, 0.0, 0.0, -3.3608926294469414, -0.868219346841726, 27.59209969944671, 20.154067550477894, -43.48988418106996, 0.0, 0.0, 0.0, 0.0}, \
     {0.47766253643826434, 0.0, 0.0, -2.4881146199716677, -0.590290826836843, 21.230051448181193, 15.279233632882423, -33.28821096898486, -0.020331201708508627, 0.0, 0.0, 0.0}, \
     {-0.9371424300859873, 0.0, 0.0, 5.186372428844064, 1.0914373489967295, -8.149787010746927, -18.52006565999696, 22.739487099350505, 2.4936055526796523, -3.0467644718982196, 0.0, 0.0}, \
     {2.273310147516538, 0.0, 0.0, -10.53449546673725, -2.0008720582248625, -17.9589318631188, 27.94888452941996, -2.8589982771350235, -8.87285693353063, 12.360567175794303, 0.6433927460157636, 0.0}}
constexpr double kB1 = 0.054293734116568765, kB6 = 4.450312892752409, kB7 = 1.8915178993145003, kB8 = -5.801203960010585,
                 kB9 = 0.3111643669578199, kB10 = -0.1521609496625161, kB11 = 0.20136540080403034, kB12 = 0.04471061572777259;
constexpr double kE1 = 0.01312004499419488, kE6 = -1.2251564463762044, kE7 = -0.4957589496572502, kE8 = 1.6643771824549864,
                 kE9 = -0.35032884874997366, kE10 = 0.3341791187130175, kE11 = 0.08192320648511571, kE12 = -0.022355307863886294;
constexpr double kBHH1 = 0.2440944881889764, kBHH2 = 0.7338466882816118, kBHH3 = 0.022058823529411766;
constexpr double kSafe = 0.9, kAlpha = 0.125, kMinScale = 1.0 / 3.0, kMaxScale = 6.0;

// The tableau by source stage for a stage loop that runs (not unrolled):
// dStage.a[j][i] = A[j + 1 + i][j] (weight of k_j in stage j + 1 + i; 0 past the
// last stage), dStage.b[j] and dStage.e[j] the weights of k_j in the 8th-order
// solution and the 5th-order error estimate.
struct StageTable {
    double a[12][11];
    double b[12];
    double e[12];
};
constexpr StageTable make_stage_table() {
    constexpr double A[12][12] = DOP_A;
    StageTable t{};
    for (int j = 0; j < 12; ++j)
        for (int i = 0; i < 11; ++i) t.a[j][i] = (j + 1 + i < 12) ? A[j + 1 + i][j] : 0.0;
    for (int j = 0; j < 12; ++j) t.b[j] = t.e[j] = 0.0;
    t.b[0] = kB1; t.b[5] = kB6; t.b[6] = kB7; t.b[7] = kB8; t.b[8] = kB9; t.b[9] = kB10; t.b[10] = kB11; t.b[11] = kB12;
    t.e[0] = kE1; t.e[5] = kE6; t.e[6] = kE7; t.e[7] = kE8; t.e[8] = kE9; t.e[9] = kE10; t.e[10] = kE11; t.e[11] = kE12;
    return t;
}
__constant__ static StageTable dStage = make_stage_table();

// ------------------------------------------------------------ group helpers
template <int G>
__device__ __forceinline__ double gshfl(double v, int src) {
    if constexpr (G == 1) return v;
    else return __shfl(v, src, G);
}
template <int G>
__device__ __forceinline__ int gshfl_i(int v, int src) {
    if constexpr (G == 1) return v;
    else return __shfl(v, src, G);
}
// the value of lane `lane` of the wavefront (a scalar register read)
__device__ __forceinline__ double read_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// lane permutation of a double within rows of 16 lanes (DPP)
template <int CTRL>
__device__ __forceinline__ double dpp_perm(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// lane permutation of a double across rows (DPP row broadcasts); rows outside
// ROWS keep 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_bcast(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes of the wavefront, the same on every lane: xor-1 and
// xor-2 swaps inside quads, the half-row and row mirrors give every lane its
// row's sum; row_bcast:15 adds row 0 into row 1 and row 2 into row 3,
// row_bcast:31 adds rows 0+1 into row 3, and lane 63's total is read out:
// (r3 + r2) + (r1 + r0).
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_perm<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_perm<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_perm<0x141>(v);   // row_half_mirror
    v += dpp_perm<0x140>(v);   // row_mirror
    v += dpp_bcast<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_bcast<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return read_lane(v, 63);
}

// Sum over lanes 0..3 of a 64-lane group whose other lanes hold 0 (the four
// state components): two swaps inside the quad, then lane 0 -- the same
// additions, in the same order, as wave_sum makes for these values.
__device__ __forceinline__ double quad0_sum(double v) {
    v += dpp_perm<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_perm<0x4E>(v);    // quad_perm [2,3,0,1]
    return read_lane(v, 0);
}

template <int G>
__device__ __forceinline__ double gsum(double v) {
    if constexpr (G == 64) {
        return wave_sum(v);
    } else if constexpr (G > 1) {
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, G);
    }
    return v;
}

// one tangent evaluation; td selects the direction (0 r, 1 pr, 2 pphi, 3 pphi at pr = 0)
__device__ __forceinline__ Dual ham_tangent(const HamCoef& hc, double r, double pr, double L, int td) {
    const Dual R(r, td == 0 ? 1.0 : 0.0);
    const Dual PR(td == 3 ? 0.0 : pr, td == 1 ? 1.0 : 0.0);
    const Dual LL(L, td >= 2 ? 1.0 : 0.0);
    return ham_eval<Dual>(hc, R, PR, LL);
}

// lane q of the group
template <int G>
__device__ __forceinline__ double gread(double v, int q) {
    if constexpr (G == 64) return read_lane(v, q);
    else return gshfl<G>(v, q);
}

// For a 64-lane group: lane 63, which owns no state component, no flux mode
// and no piece, evaluates H's expression at P = pr^2 = 0, and dH/dpphi at
// pr = 0 is read from it; the other lanes evaluate it at the state's pr.
// Lane 63's H and derivatives therefore differ from the others', and the
// integrator reads group-wide values (k1 for the stopping criterion) from
// lane 0.
struct LaneP0 {
    static constexpr bool kLaneP0 = true;
    static constexpr int kLane = 63;
    int gl;
    __device__ double P(double p) const { return gl == kLane ? 0.0 : p; }
    __device__ double omc(double dHdL) const { return read_lane(dHdL, kLane); }
};

// Per-wavefront LDS slots for the piece gather of a 64-lane group: the 13
// pieces' (value, d/du) and one slot the spare lanes write.  Every kernel that
// integrates runs blocks of at most 256 threads (4 wavefronts).
constexpr int kMaxWavesPerBlock = 4;
__device__ __forceinline__ double2* piece_slots() {
    __shared__ double2 slots[kMaxWavesPerBlock][kNumPieces + 1];
    return slots[threadIdx.x >> 6];
}

// Per-wavefront LDS copy of the last first-stage k of a 64-lane group (a
// rejected step retries from the same state, whose k1 is then not recomputed).
__device__ __forceinline__ double* k1_slots() {
    __shared__ double slots[kMaxWavesPerBlock][4];
    return slots[threadIdx.x >> 6];
}

// H and its derivatives at one state.  The 13 pieces of H's u-dependence are
// evaluated on different lanes of the group at once (lane q takes piece q; a
// group of fewer than 13 lanes takes several per lane), gathered, and
// combined in closed form on every lane.
// For G >= 16, cf is this lane's piece (lane q < 13 takes piece q, the others
// any piece); smaller groups read the pieces from pcs.
template <int G>
__device__ __forceinline__ HamPack ham_pack(const HamCoef& hc, const PieceCoef& pcs, double r, double pr, double L, int gl,
                                            const double (*cf)[3] = nullptr) {
    // sqrt(r) and 1/sqrt(r) in one call: su = sqrt(u), u = su^2, and
    // hs = d su/du = su r / 2 = sqrt(r) / 2
    double sr, su;
    qsqrt_recip(r, sr, su);
    const double u = su * su, lr = log_pos(r), hs = 0.5 * sr;
    double v[kNumPieces], d[kNumPieces];
    if constexpr (G == 64) {
        // gathered through this wavefront's LDS slots: one write, then 13
        // same-address (broadcast) reads into vector registers; a wavefront's
        // LDS accesses complete in order, so no barrier is involved
        double pv, pd;
        piece_eval(cf, u, r, lr, su, hs, pv, pd);
        double2* slot = piece_slots();
        slot[gl < kNumPieces ? gl : kNumPieces] = make_double2(pv, pd);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < kNumPieces; ++k) {
            const double2 t = slot[k];
            v[k] = t.x;
            d[k] = t.y;
        }
    } else if constexpr (G >= 16) {
        double pv, pd;
        piece_eval(cf, u, r, lr, su, hs, pv, pd);
#pragma unroll
        for (int k = 0; k < kNumPieces; ++k) {
            v[k] = gread<G>(pv, k);
            d[k] = gread<G>(pd, k);
        }
    } else {
        constexpr int PER = (kNumPieces + G - 1) / G;
        double pv[PER], pd[PER];
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            int q = gl + t * G;
            if (q >= kNumPieces) q = 0;
            piece_eval(pcs.c[q], u, r, lr, su, hs, pv[t], pd[t]);
        }
#pragma unroll
        for (int k = 0; k < kNumPieces; ++k) {
            v[k] = gshfl<G>(pv[k / G], k % G);
            d[k] = gshfl<G>(pd[k / G], k % G);
        }
    }
    if constexpr (G == 64) return ham_combine(hc, r, u, pr, L, v, d, LaneP0{gl});
    else return ham_combine(hc, r, u, pr, L, v, d);
}

// Flux sum 8 pi F (F before the 1/nu and 1/Omega factors of RR_force),
// src/EOB.cu:2506-2560, with the modes split over the group.  omega is the Omega the flux uses.  For
// G >= 35, zl is this lane's mode (lane i < 35 takes mode i, the others any mode).
template <int G>
__device__ __forceinline__ double group_flux(const ModeCoef* __restrict__ mc, double eta_f, double r, double L, double Hreal,
                                             double omega, double omc, int gl, const ModeCoef* zl = nullptr) {
    const double vPhiIn = qrcp(omc * omc * (r * r * r));
    double v, cb_in;
    if constexpr (G >= 4) {
        // the two cube roots in one call, on two lanes
        const double cr = cbrt_pos((gl & 1) ? vPhiIn : omega);
        v = gread<G>(cr, 0);
        cb_in = gread<G>(cr, 1);
    } else {
        v = cbrt_pos(omega);
        cb_in = cbrt_pos(vPhiIn);
    }
    const double logv = log_pos(v);
    const double vPhi = r * cb_in * omega;
    const double vPhi2 = vPhi * vPhi;
    const double Heff = qdiv(Hreal * Hreal - 1.0, 2.0 * eta_f) + 1.0;
    double part = 0.0;
    if constexpr (G >= kNumModes) {
        // one mode per lane; the spare lanes evaluate some mode and drop it
        const double fm = flux_mode(*zl, v, logv, Hreal, omega, Heff, L, vPhi2);
        part = (gl < kNumModes) ? fm : 0.0;
    } else {
        for (int i = gl; i < kNumModes; i += G) part += flux_mode(mc[i], v, logv, Hreal, omega, Heff, L, vPhi2);
    }
    return gsum<G>(part);
}

// k = [dH/dpr, dH/dpphi, -dH/dr + Fr, Fphi] (src/EOB.cu:3047-3144), valid on every lane of the group.
template <int G>
__device__ __forceinline__ void group_rhs(const BinData& bd, const ModeCoef* __restrict__ mc, double r, double pr, double L, int gl,
                                          double k[4]) {
    HamPack p;
    double flux, fscale;
    if constexpr (G >= kNumModes) {
        // this lane's piece and mode coefficients and the binary's scalars are
        // read before the piece gather writes LDS (the empty asm keeps the
        // reads above it; otherwise they are read again after that write)
        double cf[kPieceDeg][3];
        const double (*src)[3] = bd.pc.c[gl < kNumPieces ? gl : 0];
#pragma unroll
        for (int k = 0; k < kPieceDeg; ++k) {
            cf[k][0] = src[k][0]; cf[k][1] = src[k][1]; cf[k][2] = src[k][2];
        }
        const ModeCoef zl = mc[gl < kNumModes ? gl : kNumModes - 1];
        const HamCoef hc = bd.h;
        const double nu = bd.nu, eta_f = bd.eta_f;
        fscale = bd.flux_scale;
        asm volatile("" ::: "memory");
        p = ham_pack<G>(hc, bd.pc, r, pr, L, gl, cf);
        flux = group_flux<G>(mc, eta_f, r, L, nu * p.H, p.dHdL, p.omc, gl, &zl);
    } else {
        p = ham_pack<G>(bd.h, bd.pc, r, pr, L, gl, bd.pc.c[gl < kNumPieces ? gl : 0]);
        flux = group_flux<G>(mc, bd.eta_f, r, L, bd.nu * p.H, p.dHdL, p.omc, gl);
        fscale = bd.flux_scale;
    }
    const double omega = p.dHdL;   // the ODE overwrites omega with dH/dpphi before RR_force
    flux *= fscale;
    const double f_over_om = qdiv(flux, omega);
    k[0] = p.dHdpr;
    k[1] = p.dHdL;
    k[2] = -p.dHdr + (qdiv(-pr, L) * f_over_om);
    k[3] = -0.0 + (-f_over_om);
}

// ---------------------------------------------------------------- (2,2) mode
struct Cplx {
    double re, im;
};
EOB_HD Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
EOB_HD Cplx cexp_(double re, double im) {
    const double e = exp(re);
    return {e * cos(im), e * sin(im)};
}
EOB_HD Cplx cdiv(Cplx a, Cplx b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
EOB_HD Cplx clog_(Cplx z) { return {log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }

// Gamma(z) for Re z > 0 by shifting to Re z >= 10 and the Stirling series of
// ln Gamma (Bernoulli terms through B_16); only exp(ln Gamma) is used, so the
// branch of the logarithm is immaterial.
EOB_HD Cplx cgamma(Cplx z) {
    Cplx prod = {1.0, 0.0};
    while (z.re < 10.0) {
        prod = cmul(prod, z);
        z.re += 1.0;
    }
    const Cplx iz = cdiv({1.0, 0.0}, z);
    const Cplx iz2 = cmul(iz, iz);
    const double b[8] = {1.0 / 12, -1.0 / 360, 1.0 / 1260, -1.0 / 1680, 1.0 / 1188, -691.0 / 360360, 1.0 / 156, -3617.0 / 122400};
    Cplx ser = {b[7], 0.0};
    for (int k = 6; k >= 0; --k) {
        ser = cmul(ser, iz2);
        ser.re += b[k];
    }
    ser = cmul(ser, iz);
    const Cplx lz = clog_(z);
    const Cplx t = cmul({z.re - 0.5, z.im}, lz);
    const double lre = t.re - z.re + 0.9189385332046728 + ser.re;   // + ln(2 pi)/2
    const double lim = t.im - z.im + ser.im;
    const Cplx g = cexp_(lre, lim);
    return cdiv(g, prod);
}

// h_22 on one knot, src/EOB.cu:1236-1613 for (l, m) = (2, 2).
// v = Omega^(1/3), Hreal = nu H, vPhiIn = 1/(omega_circ^2 r^3); mc22 is flux mode 1.
EOB_HD Cplx hlm22_knot(const ModeCoef& mc22, const WaveCoef& wc, double eta, double r, double phi, double v, double Hreal,
                       double vPhiIn) {
    if (eta > 0.25 && eta < 0.25 + 1e-4) eta = 0.25;
    const double v2 = v * v;
    const double Omega = v2 * v;
    const double vh3 = Hreal * Omega;
    const double vh = cbrt(vh3);
    const double el = kEulerGamma + mc22.log2m + log(v);
    double vPhi = r * cbrt(vPhiIn);
    vPhi *= Omega;
    const double vPhi2 = vPhi * vPhi;
    // Newtonian multipole prefix * vPhi2 * Y_2^{-2}(pi/2, phi), Y_2^{-2} = sqrt(15/(32 pi)) e^{-2 i phi}
    const double yamp = 0.3862742020231896;   // sqrt(15 / (32 pi))
    const Cplx ylm = {yamp * cos(-2.0 * phi), yamp * sin(-2.0 * phi)};
    Cplx hN = cmul({wc.pf22_re, wc.pf22_im}, ylm);
    hN.re *= vPhi2; hN.im *= vPhi2;
    const double Slm = (Hreal * Hreal - 1.0) / (2.0 * eta) + 1.0;
    const double k = 2.0 * Omega;
    const double hk = Hreal * k;
    const Cplx G = cgamma({3.0, -2.0 * hk});
    // exp(pi hk + 2 i hk ln(4k/sqrt(e))) * Gamma(3 - 2 i hk) / 2!
    Cplx T = cmul(cexp_(kPi * hk, 2.0 * hk * log(4.0 * k / 1.6487212707001282)), G);
    T.re *= 0.5; T.im *= 0.5;
    const double delta = vh3 * (wc.delta22vh3 + vh3 * (wc.delta22vh6 + vh * vh * (wc.delta22vh9 * vh))) + wc.delta22v5 * v * v2 * v2
                         + wc.delta22v6 * v2 * v2 * v2 + wc.delta22v8 * v2 * v2 * v2 * v2;
    double acc = mc22.c[10];
    acc = acc * v;
    acc = acc * v + (mc22.c[9] + mc22.cl10 * el);
    acc = acc * v + mc22.c[8];
    acc = acc * v + (mc22.c[7] + mc22.cl8 * el);
    acc = acc * v + (mc22.c[6] + mc22.cl7 * el);
    acc = acc * v + (mc22.c[5] + mc22.cl6 * el);
    acc = acc * v + mc22.c[4];
    acc = acc * v + mc22.c[3];
    acc = acc * v + mc22.c[2];
    acc = acc * v + mc22.c[1];
    acc = acc * v + mc22.c[0];
    const double rho = 1.0 + acc * v;
    const double amp = Slm * rho * rho;
    Cplx h = cmul(T, {cos(delta), sin(delta)});
    h.re *= amp; h.im *= amp;
    return cmul(h, hN);
}

}  // namespace eob
