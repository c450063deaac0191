// Per-binary DOP853 integration state machine (DOPR853.integrate +
// StoppingCriterion + SEOBNRv4PHM.run_trajectory semantics), advanced one
// attempted step per call of `iterate`.
//
// Reference: eob/new_Dormand_horiz.py:424-778 (driver), eob/seobnrv4phm.py:61-150
// (stop codes) and :2072-2136 (two passes), src/DOPR853.cu:357-762 (stages,
// error norm, controller).
//
// Semantics kept from the reference, per binary:
//   - stages accumulate a_sq k_q term by term from 0 (src/DOPR853.cu:381-516);
//   - error norm with 4th powers and signed max (src/DOPR853.cu:586-616);
//   - controller alpha = 1/8, safe = 0.9, scale in [1/3, 6], no growth right
//     after a reject, errOld clamped to [1e-4, 1e300];
//   - a pass ends when x >= tmax, step_num >= max_step - 1, or the stop flag is
//     set; the stop flag starts true and is replaced by (stop code != 0) at each
//     accepted step, so a pass whose first attempt is rejected ends there;
//   - stop codes 2 (r<1), 3 (r<6, dr/dt>0), 4 (r<6, dpr/dt>0), 5 (r<6 and
//     omega falling for the last 4 stored steps), stored omegas start at 0 in
//     every pass;
//   - fix_step: h is reset to hInit after every attempt, hInit halves on a reject;
//   - pass 2 restarts at the row nearest t[num_steps-1] - step_back, minus one.
// Each binary is integrated exactly as the reference integrates a batch that
// contains only that binary (its StoppingCriterion indexes the omega history of
// other lanes when the batch is compacted; a batch of one has no other lanes).
#pragma once
#include "eob_device.hpp"

namespace eob {

struct IntegParams {
    const double* cond;       // [4][nbin] initial state
    const BinData* bd;        // [nbin]
    const ModeCoef* mc;       // [nbin][35]
    double* dense;            // [nbin][max_step][5]: t, r, phi, pr, pphi
    int* num_steps;           // [nbin] points stored (pass 2 steps + 1, or steps + 1)
    int* stop_info;           // [nbin]
    int* pass1_steps;         // [nbin]
    unsigned long long* attempts;  // [nbin] attempted steps (both passes)
    int* queue;               // work counter (persistent mode)
    int nbin, max_step;
    double rtol, atol, hInit, step_back, fine_step, tmax;
    int two_pass;             // 1: run_trajectory; 0: one DOPR853.integrate pass
    int fix_step;             // first (or only) pass at fixed step hInit (the reference: off)
    unsigned long long max_attempts;  // per binary, both passes; ends runaway step-size loops
    const double* x0;         // optional restart: start time per binary (with step0)
    const int* step0;         // optional restart: row the initial state occupies
    int use_stop;             // StoppingCriterion on (the reference driver without one uses stop code 0)
    int* done;                // optional [nbin], host memory mapped for the device: 1 once the binary's outputs are written
};

// err^kAlpha for the controller's err^-kAlpha (src/DOPR853.cu:675-762 calls
// pow): with kAlpha = 1/8 this is three square roots (about a third of pow's
// instructions, within 2 ulp of the exact power like pow itself)
static_assert(kAlpha == 0.125, "root8 is err^(1/8)");
__device__ __forceinline__ double root8(double err) { return sqrt(sqrt(sqrt(err))); }

template <int G>
struct CompCount {
    static constexpr int NC = (G >= 4) ? 1 : 4 / G;
};

// state of one group; scalars are identical on every lane of the group
template <int G>
struct GState {
    static constexpr int NC = CompCount<G>::NC;
    int b, pass, step_num, sn1, code;
    int attempted, just_acc, stop_flag, prev_reject;
    int fail;                    // 1: non-finite error or step, 2: attempt cap reached
    int k1_ok;                   // 64-lane groups: k1 of the current state is in k1_slots()
    double x, h, err_old, h_init;
    double om0, om1, om2, om3;   // last four stored omegas, om3 newest
    double y[NC];                // owned state components
    unsigned long long att;
};

// component index owned by lane gl in slot q
template <int G>
__device__ __forceinline__ int comp_of(int gl, int q) {
    if constexpr (G >= 4) return gl;
    else return gl + G * q;
}

template <int G>
__device__ __forceinline__ void gather_state(const GState<G>& s, int gl, double yf[4]) {
    if constexpr (G == 1) {
        yf[0] = s.y[0]; yf[1] = s.y[1]; yf[2] = s.y[2]; yf[3] = s.y[3];
    } else if constexpr (G >= 4) {
#pragma unroll
        for (int c = 0; c < 4; ++c) yf[c] = gread<G>(s.y[0], c);
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int owner = c % G, slot = c / G;
            double mine = 0.0;
#pragma unroll
            for (int q = 0; q < GState<G>::NC; ++q) if (q == slot) mine = s.y[q];
            yf[c] = gshfl<G>(mine, owner);
        }
    }
}

__device__ __forceinline__ double* dense_row(const IntegParams& P, int b, int row) {
    return P.dense + ((size_t)b * P.max_step + row) * 5;
}

template <int G>
__device__ __forceinline__ void write_row(const IntegParams& P, const GState<G>& s, int gl) {
    // rows count from the end for a negative step number (numpy indexing, reached
    // by a restart from a negative row); a restart from row max_step - 1 would
    // store past the last row (the reference raises IndexError there), and such
    // rows are dropped
    int rr = s.step_num;
    if (rr < 0) rr += P.max_step;
    if (rr < 0 || rr >= P.max_step) return;
    double* row = dense_row(P, s.b, rr);
#pragma unroll
    for (int q = 0; q < GState<G>::NC; ++q) {
        const int c = comp_of<G>(gl, q);
        if (c < 4) row[1 + c] = s.y[q];
    }
    if (gl == 0) row[0] = s.x;
}

template <int G>
__device__ __forceinline__ void start_pass(GState<G>& s, int pass, double h0) {
    s.pass = pass;
    s.h = h0;
    s.h_init = h0;
    s.err_old = 1e-4;
    s.prev_reject = 0;
    s.attempted = 0;
    s.just_acc = 0;
    s.stop_flag = 1;
    s.code = 0;
    s.om0 = s.om1 = s.om2 = s.om3 = 0.0;
    s.k1_ok = 0;
}

// Wait here for this wavefront's outstanding vector-memory loads.  The state a
// binary starts from comes from memory on the paths that (re)start it; without
// this wait the compiler waits for those loads at the top of every iteration,
// and on this architecture that wait also covers the rows stored since, so
// every attempted step would stall until the previous step's stores are
// acknowledged (slow while other kernels load the memory system).
__device__ __forceinline__ void wait_loads() {
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0); expcnt and lgkmcnt left at their maxima
}

template <int G>
__device__ __forceinline__ void begin_binary(const IntegParams& P, GState<G>& s, int b, int gl) {
    s.b = b;
    s.sn1 = 0;
    s.att = 0ull;
    s.fail = 0;
#pragma unroll
    for (int q = 0; q < GState<G>::NC; ++q) {
        const int c = comp_of<G>(gl, q);
        s.y[q] = (c < 4) ? P.cond[(size_t)c * P.nbin + b] : 0.0;
    }
    start_pass(s, 1, P.hInit);
    if (P.step0) {
        // continue an earlier integration from row step0 (its row is already stored)
        s.step_num = P.step0[b];
        s.x = P.x0[b];
    } else {
        s.step_num = 0;
        s.x = 0.0;
        write_row(P, s, gl);
    }
    wait_loads();
}

// relaxed device-coherent load (rows written earlier by this group)
__device__ __forceinline__ double coherent_load(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// end of pass 1: walk back step_back and set up the fixed-step pass
template <int G>
__device__ __forceinline__ void restart_pass2(const IntegParams& P, GState<G>& s, int gl) {
    __threadfence();
    s.sn1 = s.step_num;
    int last = s.step_num - 1;
    if (last < 0) last = 0;
    const double t_stop = coherent_load(dense_row(P, s.b, last));
    const double td = t_stop - P.step_back;
    // argmin over rows 0..step_num (unwritten rows hold t = 0 and tie with row 0)
    double best_v = 1e308;
    int best_i = 0x7fffffff;
    const int last_row = (s.step_num < P.max_step - 1) ? s.step_num : P.max_step - 1;
    for (int row = gl; row <= last_row; row += G) {
        const double d = fabs(coherent_load(dense_row(P, s.b, row)) - td);
        if (d < best_v) { best_v = d; best_i = row; }
    }
    if constexpr (G > 1) {
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1) {
            const double ov = __shfl_xor(best_v, off, G);
            const int oi = __shfl_xor(best_i, off, G);
            if (ov < best_v || (ov == best_v && oi < best_i)) { best_v = ov; best_i = oi; }
        }
    }
    if (best_i > last_row) best_i = 0;   // every |t - t_desired| was NaN
    int row = best_i - 1;
    bool zero_state = false;
    if (row < 0) {
        // Python index -1: the last row, which is still zero unless pass 1 filled it
        row = P.max_step - 1;
        zero_state = (s.step_num < P.max_step - 1);
    }
    const double* rp = dense_row(P, s.b, row);
    s.x = zero_state ? 0.0 : coherent_load(rp);
#pragma unroll
    for (int q = 0; q < GState<G>::NC; ++q) {
        const int c = comp_of<G>(gl, q);
        s.y[q] = (c < 4 && !zero_state) ? coherent_load(rp + 1 + c) : 0.0;
    }
    s.step_num = best_i - 1;
    start_pass(s, 2, P.fine_step);
    wait_loads();
}

template <int G>
__device__ __forceinline__ void finish_binary(const IntegParams& P, GState<G>& s, int gl) {
    if (gl == 0) {
        P.num_steps[s.b] = s.step_num + 1;
        P.stop_info[s.b] = s.code;
        P.pass1_steps[s.b] = P.two_pass ? s.sn1 : s.step_num;
        P.attempts[s.b] = s.att;
    }
    if (P.done) {
        // the rows and the four outputs above reach memory before the flag: a
        // system-scope release fence by the whole wavefront (it writes the
        // cache's dirty lines back and, unlike a full fence, leaves the cache's
        // contents valid for the other wavefronts), then lane 0's system-scope
        // store, which goes through to host memory
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        if (gl == 0) __hip_atomic_store(P.done + s.b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    s.pass = 0;
}

// Per stage, the twelve weights the stage loop needs after the right-hand side
// (dStage.a[st][0..9], dStage.b[st], dStage.e[st]), for a copy in LDS: read
// at the top of a stage, their latency hides behind the right-hand side (from
// the constant table the compiler loads them just before their use and waits).
constexpr int kStageRow = 12;
constexpr int kStageRowsBytes = 12 * kStageRow * 8;
__device__ __forceinline__ void stage_rows_copy(double* dst, int lane) {
    for (int i = lane; i < 12 * kStageRow; i += 64) {
        const int st = i / kStageRow, c = i % kStageRow;
        dst[i] = (c < 10) ? dStage.a[st][c] : ((c == 10) ? dStage.b[st] : dStage.e[st]);
    }
}

// One attempted DOP853 step given k1 at the current state.  rows: the stage
// weights in LDS (stage_rows_copy) or null for the constant table.
template <int G, bool kRows = false>
__device__ __forceinline__ void attempt_step(const IntegParams& P, GState<G>& s, const BinData& bd, const ModeCoef* __restrict__ mc,
                                             const double k1f[4], int gl, bool fixed, const double* rows = nullptr) {
    constexpr int NC = GState<G>::NC;
    // Stages run in a loop that is not unrolled, so the right-hand side is
    // emitted once (a compact body that stays in the instruction cache).  Each
    // stage's k is folded into running sums as soon as it is known:
    // acc[i] = sum over the stages so far of A[next + i][.] k, plus the 8th-order
    // (bsum) and error (esum) combinations; the additions happen in the same
    // order as sum_j A[st][j] k_j, so only the bookkeeping changes.
    double acc[11][NC], bsum[NC], esum[NC], k0[NC], k8[NC], k11[NC];
    const double h = s.h;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int c = comp_of<G>(gl, q);
        const double kq = k1f[c < 4 ? c : 0];
        k0[q] = kq;
        k8[q] = 0.0;
        k11[q] = 0.0;
#pragma unroll
        for (int i = 0; i < 11; ++i) acc[i][q] = dStage.a[0][i] * kq;
        bsum[q] = kB1 * kq;
        esum[q] = kE1 * kq;
    }
#pragma unroll 1
    for (int st = 1; st < 12; ++st) {
        // the LDS weights are read here, the constant table's where they are used
        // (hoisting those as well costs the register-heavy narrow groups spills)
        double wa[10] = {}, wb = 0.0, we = 0.0;
        if constexpr (kRows) {
            const double* r = rows + st * kStageRow;
#pragma unroll
            for (int i = 0; i < 10; ++i) wa[i] = r[i];
            wb = r[10];
            we = r[11];
        }
        double arg[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) arg[q] = s.y[q] + h * acc[0][q];
        double af[4];
        if constexpr (G == 1) {
            af[0] = arg[0]; af[1] = arg[1]; af[2] = arg[2]; af[3] = arg[3];
        } else if constexpr (G >= 4) {
#pragma unroll
            for (int c = 0; c < 4; ++c) af[c] = gread<G>(arg[0], c);
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double mine = 0.0;
#pragma unroll
                for (int q = 0; q < NC; ++q) if (q == c / G) mine = arg[q];
                af[c] = gshfl<G>(mine, c % G);
            }
        }
        double kk[4];
        group_rhs<G>(bd, mc, af[0], af[2], af[3], gl, kk);
        const double bj = kRows ? wb : dStage.b[st], ej = kRows ? we : dStage.e[st];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const int c = comp_of<G>(gl, q);
            const double kj = kk[c < 4 ? c : 0];
#pragma unroll
            for (int i = 0; i < 10; ++i) acc[i][q] = acc[i + 1][q] + (kRows ? wa[i] : dStage.a[st][i]) * kj;
            acc[10][q] = 0.0;
            bsum[q] = bsum[q] + bj * kj;
            esum[q] = esum[q] + ej * kj;
            if (st == 8) k8[q] = kj;
            if (st == 11) k11[q] = kj;
        }
    }
    // error norm (src/DOPR853.cu:586-616); stage 11 lives in k2, stage 12 in k3
    double e5 = 0.0, e3 = 0.0, ynew[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int c = comp_of<G>(gl, q);
        const double temp = bsum[q];
        const double yo = s.y[q];
        const double yn = yo + h * temp;
        ynew[q] = yn;
        const double mx = (yo > yn) ? yo : yn;
        const double sk = 1.0 / (P.atol + P.rtol * mx);
        const double t3 = (temp - kBHH1 * k0[q] - kBHH2 * k8[q] - kBHH3 * k11[q]) * sk;
        const double t5 = esum[q] * sk;
        if (c < 4) {
            const double t3s = t3 * t3, t5s = t5 * t5;
            e3 += t3s * t3s;
            e5 += t5s * t5s;
        }
    }
    if constexpr (G == 64) {
        e3 = quad0_sum(e3);
        e5 = quad0_sum(e5);
    } else {
        e3 = gsum<G>(e3);
        e5 = gsum<G>(e5);
    }
    double denom = e5 + 0.01 * e3;
    denom = (denom < 0.0) ? 1.0 : denom;
    const double err = fabs(h) * e5 * sqrt(1.0 / (4.0 * denom));
    // controller (src/DOPR853.cu:675-762)
    const double h_old = h;
    bool ok;
    if (err <= 1.0) {
        double scale;
        if (err == 0.0) scale = kMaxScale;
        else {
            scale = kSafe / root8(err);
            scale = (scale > kMaxScale) ? kMaxScale : ((scale < kMinScale) ? kMinScale : scale);
        }
        if (s.prev_reject && scale > 1.0) scale = 1.0;
        s.h = h * scale;
        double e = err;
        e = (e < 1e-4) ? 1e-4 : ((e > 1e300) ? 1e300 : e);
        s.err_old = e;
        s.prev_reject = 0;
        ok = true;
    } else {
        s.prev_reject = 1;
        double t = kSafe / root8(err);
        t = (t < kMinScale) ? kMinScale : ((t > 1e300) ? 1e300 : t);
        s.h = h * t;
        ok = false;
    }
    if (fixed) {
        if (!ok) s.h_init *= 0.5;
        s.h = s.h_init;
    }
    s.attempted = 1;
    s.att += 1ull;
    // a non-finite error or step would make the reference loop forever; end the binary instead
    if (!(err >= 0.0) || !(fabs(s.h) < 1e300) || s.h == 0.0) s.fail = 1;
    else if (s.att >= P.max_attempts) s.fail = 2;
    if (ok) {
        s.k1_ok = 0;
        s.x = s.x + h_old;
#pragma unroll
        for (int q = 0; q < NC; ++q) s.y[q] = ynew[q];
        s.step_num += 1;
        write_row(P, s, gl);
        s.just_acc = 1;
    } else {
        s.just_acc = 0;
    }
}

// One driver iteration for a group.  Returns false when the group needs a new
// binary (persistent mode refills; plain mode marks it finished).
// The tables come from HBM (P.bd, P.mc) or, when the caller staged them in LDS
// for its binary (lds_bd / lds_mc non-null), from there; the LDS copy of the
// stage weights (stage_rows_copy) then follows the modes' table.
template <int G, bool kLds = false>
__device__ __forceinline__ void iterate(const IntegParams& P, GState<G>& s, int gl, bool& need_new,
                                        const BinData* lds_bd = nullptr, const ModeCoef* lds_mc = nullptr) {
    need_new = false;
    if (s.pass != 1 && s.pass != 2) return;
    const BinData& bd = kLds ? *lds_bd : P.bd[s.b];
    const ModeCoef* mc = kLds ? lds_mc : P.mc + (size_t)s.b * kNumModes;
    double yf[4];
    gather_state<G>(s, gl, yf);
    double k1[4];
    if constexpr (G == 64) {
        // after a rejected attempt the state is unchanged, and so is its k1
        double* k1s = k1_slots();
        if (s.k1_ok) {
            k1[0] = k1s[0]; k1[1] = k1s[1]; k1[2] = k1s[2]; k1[3] = k1s[3];
        } else {
            group_rhs<G>(bd, mc, yf[0], yf[2], yf[3], gl, k1);
            // lane 63 evaluated H at pr = 0 (LaneP0): the group's k1 is lane 0's
#pragma unroll
            for (int c = 0; c < 4; ++c) k1[c] = read_lane(k1[c], 0);
            k1s[0] = k1[0]; k1s[1] = k1[1]; k1s[2] = k1[2]; k1s[3] = k1[3];   // every lane stores the same values
            s.k1_ok = 1;
        }
    } else {
        group_rhs<G>(bd, mc, yf[0], yf[2], yf[3], gl, k1);
    }
    if (s.attempted) {
        if (s.just_acc) {
            // StoppingCriterion at the accepted state (eob/seobnrv4phm.py:82-150)
            const double r = yf[0], drdt = k1[0], om = k1[1], dprdt = k1[2];
            int code = 0;
            if (r < 1.0) code = 2;
            else if (r < 6.0 && drdt > 0.0) code = 3;
            else if (r < 6.0 && dprdt > 0.0) code = 4;
            else if (s.step_num > 5) {
                const bool third = (s.om1 - s.om0 < 0.0) && (s.om2 - s.om1 < 0.0) && (s.om3 - s.om2 < 0.0);
                if (r < 6.0 && om < s.om3 && third) code = 5;
            }
            s.om0 = s.om1; s.om1 = s.om2; s.om2 = s.om3; s.om3 = om;
            if (!P.use_stop) code = 0;
            s.code = code;
            s.stop_flag = (code != 0);
        }
        const bool term = (s.x >= P.tmax) || (s.step_num >= P.max_step - 1) || s.stop_flag || s.fail;
        if (term) {
            if (s.fail) {
                s.code = -s.fail;
                if (s.pass == 1) s.sn1 = s.step_num;
                finish_binary<G>(P, s, gl);
                need_new = true;
            } else if (s.pass == 1 && P.two_pass) {
                restart_pass2<G>(P, s, gl);
            } else {
                finish_binary<G>(P, s, gl);
                need_new = true;
            }
            return;   // k1 is stale for the new state
        }
    }
    // pass 2 is always fixed-step; fix_step makes the first (or only) pass fixed-step too
    const bool fixed = (s.pass == 2) || P.fix_step;
    if constexpr (kLds) attempt_step<G, true>(P, s, bd, mc, k1, gl, fixed, reinterpret_cast<const double*>(mc + kNumModes));
    else attempt_step<G>(P, s, bd, mc, k1, gl, fixed);
}

}  // namespace eob
