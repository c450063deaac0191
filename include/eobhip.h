/*
 * eobhip — MI355X (gfx950) HIP implementation of EOBGPU's batched SEOBNR hot
 * path behind a C ABI.
 *
 * Conventions
 *   - Every pointer argument is a device pointer unless its comment says "host".
 *   - Layouts are the reference's: state arrays [4][numSys] flattened j*numSys+i,
 *     additionalArgs [80][numSys] (rows 0..9 m1, m2, chi1, chi2, omega0, K, d5,
 *     dSO, dSS, h22_calib; rows 10..79 Re/Im of the 35 Newtonian prefixes),
 *     complex arrays are interleaved (re, im) doubles.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).  Calls
 *     are asynchronous on that stream unless stated otherwise.
 *   - Every entry point returns 0 (EOBHIP_OK) or an error code; the message of the
 *     last failure on the calling thread is available from eobhip_last_error().
 *     Nothing calls exit().
 */
#ifndef EOBHIP_H
#define EOBHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EOBHIP_OK 0
#define EOBHIP_EINVAL 1
#define EOBHIP_EHIP 2
#define EOBHIP_ENOMEM 3
#define EOBHIP_EUNSUPPORTED 4

const char* eobhip_last_error(void);
int eobhip_version(void);

/* The DOP853 tableau compiled into the library (host out, 12+144+12+12+3 doubles:
 * c, a (row-major 12x12), b, e5, bhh). No reference counterpart; lets tests pin
 * the constants against a published table. */
int eobhip_dop853_tableau(double* out_host);

/* One wavefront that waits `cycles` shader clocks on `stream` (at most 4e9).
 * No reference counterpart: eobgpu_amd.pipeline.concurrent_streams launches it
 * to find streams whose kernels run concurrently (streams beyond the process's
 * hardware queues share one, and work behind a long kernel in a shared queue
 * waits for it). */
int eobhip_spin(long long cycles, void* stream);

/* ---------------------------------------------------------------- eob.pyx */
/* evaluate_Ham_align_AD_wrap, include/EOB.hh:11; src/eob.pyx:92-102 */
int eobhip_evaluate_Ham_align_AD(double* out, const double* r, const double* phi, const double* pr, const double* pphi,
                                 const double* m_1, const double* m_2, const double* chi_1, const double* chi_2,
                                 double K, double d5, double dSO, double dSS, int numSys, void* stream);
/* grad_Ham_align_AD_wrap, include/EOB.hh:8; src/eob.pyx:72-79.  grad_out [4][numSys] */
int eobhip_grad_Ham_align_AD(const double* arg, double* grad_out, const double* additionalArgs, int numSys, void* stream);
/* hessian_Ham_align_AD_wrap, include/EOB.hh:9; src/eob.pyx:81-89.  hess_out [16][numSys], (a*4+b) */
int eobhip_hessian_Ham_align_AD(const double* arg, double* hess_out, const double* additionalArgs, int numSys, void* stream);
/* IC_cons_wrap, include/EOB.hh:13; src/eob.pyx:94-101.  res [2][numSys] = (dH/dpphi - omega0, dH/dr);
 * args [4][numSys] is overwritten with (r, 0, 0, pphi) as the reference does */
int eobhip_IC_cons(double* res, const double* x, double* args, const double* additionalArgs, double* grad_out, int numSys,
                   void* stream);

/* root_find_all_wrap, include/EOB.hh:6; src/eob.pyx:51-58; src/EOB.cu:2853-3031.
 * (r, pphi) solving IC_cons = 0 (dH/dpphi = omega0 from additionalArgs row 4, dH/dr = 0
 * at pr = 0) by a discrete Newton iteration from x0In; xOut, x0In [2][numBinAll];
 * n = 2, num_args = 4, num_add_args = 80 (argsIn is not read: IC_cons rebuilds the state). */
int eobhip_root_find_all(double* xOut, const double* x0In, const double* argsIn, const double* additionalArgsIn, int max_iter,
                         double err, int numBinAll, int n, int num_args, int num_add_args, void* stream);

/* ------------------------------------------- eob.pyx SEOBNRv5Class (opaque handle) */
typedef struct eobhip_seobnr* eobhip_seobnr_t;
/* SEOBNRv5::SEOBNRv5 / ~SEOBNRv5, include/EOB.hh:372-374 */
int eobhip_seobnr_create(eobhip_seobnr_t* out_handle);
int eobhip_seobnr_destroy(eobhip_seobnr_t h);
/* SEOBNRv5::update_information, include/EOB.hh:377; src/EOB.cu:3355-3374.
 * All per-binary inputs [numSys]; newtonian_prefixes_hlms is kept by pointer
 * (complex [5][numSys] for modes (2,2),(2,1),(3,3),(4,4),(5,5)) like the reference.
 * Returns EOBHIP_EINVAL if already allocated (the reference prints and returns). */
int eobhip_seobnr_update_information(eobhip_seobnr_t h, const double* m1, const double* m2, const double* eta,
                                     const double* tplspin, const double* chi_S, const double* chi_A, int use_hm, int numSys,
                                     const double* newtonian_prefixes_hlms, void* stream);
/* SEOBNRv5::deallocate_information, include/EOB.hh:376 */
int eobhip_seobnr_deallocate_information(eobhip_seobnr_t h);
/* SEOBNRv5::RR_force_wrap, include/EOB.hh:378.  force_out [2][numSys]; omega taken from additionalArgs row 4 */
int eobhip_seobnr_RR_force(eobhip_seobnr_t h, double* force_out, double* grad_out, const double* args,
                           const double* additionalArgs, int numSys, const int* hCoeffs_index, void* stream);
/* SEOBNRv5::IC_diss_wrap, include/EOB.hh:381; src/EOB.cu:2666-2710, 3410-3421.
 * out[numSys] = F_phi / (dpphi/dr) - dH/dpr at (args r, pr[i], args pphi); like the reference
 * it writes pr into args row 2 and dH/dpphi into additionalArgs row 4, the gradient into
 * grad_out [4][numSys], the Hessian into hess_out [16][numSys], the force into force_out
 * [2][numSys] and the pr = 0 gradient into grad_temp_force [4][numSys]. */
int eobhip_seobnr_IC_diss(eobhip_seobnr_t h, double* out, const double* pr, double* args, double* additionalArgs, double* grad_out,
                          double* grad_temp_force, double* hess_out, double* force_out, int numSys, void* stream);
/* SEOBNRv5::root_find_scalar_all_wrap, include/EOB.hh:379; src/EOB.cu:2712-2834, 3398-3408.
 * pr_res[numBinAll] with IC_diss(pr) = 0 by false position from start_bounds [2][numBinAll];
 * argsIn [4][numBinAll] (r, -, -, pphi) and additionalArgsIn [80][numBinAll] are read only. */
int eobhip_seobnr_root_find_scalar_all(eobhip_seobnr_t h, double* pr_res, const double* start_bounds, const double* argsIn,
                                       const double* additionalArgsIn, int max_iter, double err, int numBinAll, int num_args,
                                       int num_add_args, void* stream);
/* SEOBNRv5::ODE_Ham_align_AD_wrap, include/EOB.hh:380; src/EOB.cu:3486-3496.  k [4][numSys] */
int eobhip_seobnr_ODE_Ham_align_AD(eobhip_seobnr_t h, const double* x, const double* arg, double* k,
                                   const double* additionalArgs, int numSys, const int* hCoeffs_index, void* stream);
/* SEOBNRv5::EOBGetSpinFactorizedWaveform_wrap, include/EOB.hh:384-401, for (l, m) = (2, 2);
 * other modes return EOBHIP_EUNSUPPORTED.  Knot arrays [numSys][traj_length]; out complex. */
int eobhip_seobnr_EOBGetSpinFactorizedWaveform(eobhip_seobnr_t h, double* out, const double* r, const double* phi,
                                               const double* pr, const double* pphi, const double* v, const double* Hreal,
                                               const double* eta, const double* vPhiInput, int l, int m,
                                               const double* newtonian_prefix, double h22_calib, int numSys, int traj_length,
                                               const int* hCoeffs_index, void* stream);

/* ---------------------------------------------------------------- dopr853.pyx */
/* dormandPrinceSteps_wrap, include/DOPR853.hh:6-27; src/DOPR853.cu:357-517.
 * The 12 stages; the stage-11 result goes to k2, stage 12 to k3. */
int eobhip_dormandPrinceSteps(double* x, double* xCurrent, double* solOld, double* h, double* arg, double* additionalArgs,
                              double* k1, double* k2, double* k3, double* k4, double* k5, double* k6, double* k7,
                              double* k8, double* k9, double* k10, double* ak_term_buffer, int nargs, int numEq,
                              int num_add_args, eobhip_seobnr_t ode, const int* hCoeffs_index, void* stream);
/* error_wrap, include/DOPR853.hh:29-43; src/DOPR853.cu:539-618 (rtol = atol = 1e-12 as in the reference) */
int eobhip_error(double* err, const double* solOld, double* solNew, const double* h, const double* k1, const double* k2,
                 const double* k3, const double* k6, const double* k7, const double* k8, const double* k9,
                 const double* k10, int numEq, int nargs, void* stream);
/* same with explicit tolerances (no reference counterpart: the reference hard-codes them, src/DOPR853.cu:7-8) */
int eobhip_error_tol(double* err, const double* solOld, double* solNew, const double* h, const double* k1, const double* k2,
                     const double* k3, const double* k6, const double* k7, const double* k8, const double* k9,
                     const double* k10, int numEq, int nargs, double rtol, double atol, void* stream);
/* controllerSuccess_wrap, include/DOPR853.hh:45-54; src/DOPR853.cu:675-762 (1-byte bool flags) */
int eobhip_controllerSuccess(unsigned char* flagSuccess, const double* err, double* errOld, unsigned char* previousReject,
                             double* h, const double* x, int numEq, int nargs, void* stream);

/* ---------------------------------------------------------------- tdinterp.pyx */
/* interpolate_TD (interpolate_TD_wrap), include/interpolate.hh:13-15; src/interpolate.cu:269-308.
 * t_arr [maxlen][numBinAll], propArrays/B/diag/... [maxlen][nsubs*numBinAll]; outputs alias:
 * upper_diag <- c1, diag <- c2, lower_diag <- c3 */
int eobhip_interpolate_TD(const double* t_arr, const double* propArrays, double* B, double* upper_diag, double* diag,
                          double* lower_diag, const int* lengths, int numBinAll, int nsubs, void* stream);
/* all_in_one_likelihood_wrap, include/interpolate.hh:17; src/interpolate.cu:616-628.
 * temp_sum complex [ceil(data_length/1024)][num_bin_all][nchannels] partial sums */
int eobhip_all_in_one_likelihood(double* temp_sum, const double* hp_fd, const double* hc_fd, const double* data,
                                 const double* psd, const double* Fplus, const double* Fcross, const double* time_shift,
                                 double df, int num_bin_all, int nchannels, int data_length, void* stream);

/* ------------------------------------------------ batched pipeline (MI355X design)
 * No single reference counterpart: these compose SEOBNRv4PHM.setup_class
 * (eob/seobnrv4phm.py:1922-1964), run_trajectory (:2072-2136) with
 * DOPR853.integrate (eob/new_Dormand_horiz.py:424-778), get_hlms (:2138-2221)
 * and the spline resample + likelihood (:2319-2372, :772-823) into device-resident
 * batch kernels. */

/* bytes per binary of the two device tables produced by eobhip_bin_setup */
size_t eobhip_bindata_bytes(void);
size_t eobhip_modes_bytes(void);
/* per-binary Hamiltonian + flux tables from additionalArgs [80][n] */
int eobhip_bin_setup(const double* additionalArgs, int n, int use_hm, void* bindata, void* modes, void* stream);
/* additionalArgs [80][n] built on the device from masses m1, m2 (any common unit;
 * scaled to M = 1) and aligned spins chi1, chi2 [n], as SEOBNRv4PHM.setup_class
 * builds it on the host (eob/seobnrv4phm.py:1922-1964: rows 0-3 the scaled masses
 * and spins, row 4 omega0 = 0, rows 5-9 K, d5, dSO, dSS, h22_calib, rows 10-79 the
 * Newtonian prefixes, Damour-Iyer-Nagar 2008 Eqs. 5-7), then the tables of
 * eobhip_bin_setup, in one launch; bit for bit the host's
 * eobgpu_amd.pipeline.make_additional_args.  Its reference counterpart is the
 * host code of setup_class. */
int eobhip_params_setup(const double* m1, const double* m2, const double* chi1, const double* chi2, int n, int use_hm, double K,
                        double d5, double dSO, double dSS, double h22_calib, double* additionalArgs, void* bindata, void* modes,
                        void* stream);
/* The parameter-independent factors of the 35 Newtonian prefixes the device uses
 * (host out, complex interleaved [35]); no reference counterpart (tests compare
 * them with the host's). */
int eobhip_newton_prefix_constants(double* out_host);

/* Circular-orbit start at r0: pphi0 from dH/dr = 0, omega0 = dH/dpphi, and pr0 by the
 * reference's IC_diss false position on [-3e-2, 0] (src/EOB.cu:2666-2834). */
int eobhip_initial_conditions(const double* r0, const void* bindata, const void* modes, int n, double* pphi0, double* pr0,
                              double* omega0, int max_iter, double err, void* stream);
/* eobhip_initial_conditions written as the integrator's initial state:
 * cond [4][n] = (r0, 0, pr0, pphi0) and omega0[n], in one launch.  No
 * reference counterpart (the reference builds condBound from the three
 * arrays, eob/seobnrv4phm.py:2086); a pass starts with it. */
int eobhip_initial_state(const double* r0, const void* bindata, const void* modes, int n, double* cond, double* omega0, int max_iter,
                         double err, void* stream);

/* Integrate n binaries.  mode 0: one persistent work-stealing kernel (groups of `group`
 * lanes pull binaries from an atomic queue); mode 1: one kernel launch per attempted
 * step with on-device compaction of the active list (the reference's batching model).
 * two_pass = 1 gives run_trajectory semantics (adaptive pass, walk back step_back,
 * fixed-step pass of fine_step); two_pass = 0 one integrate() pass (fix_step as given).
 * x0/step0 (both NULL, or both [n]): continue from row step0 at time x0 (integrate()'s
 * step_num/denseOutput restart, eob/new_Dormand_horiz.py:504-520) with condBound holding that row;
 * step0 in [-max_step, max_step - 2] (negative rows count from the end, as numpy's; rows
 * past max_step - 1 are not stored);
 * use_stop = 0 runs without the StoppingCriterion (stop code 0);
 * dense [n][max_step][5] (t, r, phi, pr, pphi); num_steps[n] = stored points;
 * scratch: eobhip_integrate_scratch_bytes(n) bytes of device memory.
 * iterations_out (host, may be NULL): driver iterations (plain mode) or 0. */
size_t eobhip_integrate_scratch_bytes(int n);
int eobhip_integrate(const double* condBound, const void* bindata, const void* modes, int n, int max_step, double rtol,
                     double atol, double hInit, double step_back, double fine_step, double tmax, int two_pass, int fix_step,
                     const double* x0, const int* step0, int use_stop, int mode, int group, int block, int grid_groups,
                     double* dense, int* num_steps, int* stop_info, int* pass1_steps, unsigned long long* attempts,
                     void* scratch, long long* iterations_out, void* stream);

/* eobhip_integrate in the persistent mode with a completion flag per binary:
 * done[b] becomes 1 once binary b's rows, num_steps, stop_info, pass1_steps
 * and attempts are in memory (a system-scope release, then the store), so a
 * host thread polling done can start the binary's downstream work (knots,
 * splines, likelihood) on another stream while the rest of the batch still
 * integrates.  done: n ints in host memory the device can address (pinned,
 * e.g. hipHostMalloc), zeroed by the caller before the launch.  No counterpart
 * in the reference, whose driver returns after the whole batch
 * (eob/new_Dormand_horiz.py:424-778); arguments as eobhip_integrate without
 * mode (persistent) and iterations_out. */
int eobhip_integrate_signal(const double* condBound, const void* bindata, const void* modes, int n, int max_step, double rtol,
                            double atol, double hInit, double step_back, double fine_step, double tmax, int two_pass, int fix_step,
                            const double* x0, const int* step0, int use_stop, int group, int block, int grid_groups, double* dense,
                            int* num_steps, int* stop_info, int* pass1_steps, unsigned long long* attempts, void* scratch, int* done,
                            void* stream);

/* eobhip_integrate with the reference's batch semantics of the stopping criterion:
 * DOPR853.integrate's loop (eob/new_Dormand_horiz.py:585-775) with
 * StoppingCriterion.__call__ (eob/seobnrv4phm.py:82-150) applied to the binaries
 * accepted in each iteration in ascending order, whose third check reads the omega
 * history at the binary's position in that list and which stores omegas only for
 * the binaries past step 5 while any is; run_trajectory's two passes are
 * batch-synchronous (eob/seobnrv4phm.py:2086-2123).  A binary's result then depends
 * on its batch (eobhip_integrate treats each binary as a batch of one).  One launch
 * per attempted step plus a one-block criterion launch and a 4-byte copy.
 * old_omegas [max_step][n] (StoppingCriterion.old_omegas): zeroed at the start of each
 * pass and left with the last pass's history.  Other arguments as eobhip_integrate
 * (block 256, no grid_groups). */
int eobhip_integrate_batch_coupled(const double* condBound, const void* bindata, const void* modes, int n, int max_step,
                                   double rtol, double atol, double hInit, double step_back, double fine_step, double tmax,
                                   int two_pass, int fix_step, const double* x0, const int* step0, int use_stop, int group,
                                   double* dense, int* num_steps, int* stop_info, int* pass1_steps, unsigned long long* attempts,
                                   double* old_omegas, void* scratch, long long* iterations_out, void* stream);

/* (2,2) mode on the trajectory knots of every binary (get_hlms for (2,2)):
 * knots [n][max_step][3] = (phi, Re, Im) of h22 * exp(2 i phi); rows >= num_steps untouched */
int eobhip_hlm22_knots(const double* dense, const int* num_steps, const void* bindata, const void* modes, int n,
                       int max_step, double* knots, void* stream);

/* Not-a-knot splines of the three knot series (phi and Re, Im of h22 e^{2 i phi}) per
 * binary (CubicSplineInterpolantTD / interpolate_TD semantics, src/interpolate.cu:11-308),
 * one wavefront per binary.  coef [n][max_step][12] = (y, c1, c2, c3) x 3;
 * scratch: eobhip_spline_scratch_doubles(n, max_step) doubles (binary b's part starts at
 * b * eobhip_spline_scratch_doubles(1, max_step)). */
size_t eobhip_spline_scratch_doubles(int n, int max_step);
/* The same splines for any number of series: knots [n][max_step][nser],
 * coef [n][max_step][4 nser], 1 <= nser <= 32; scratch:
 * eobhip_spline_series_scratch_doubles(n, max_step, nser) doubles (binary b's part
 * starts at b * eobhip_spline_series_scratch_doubles(1, max_step, nser)). */
size_t eobhip_spline_series_scratch_doubles(int n, int max_step, int nser);
int eobhip_spline_build_series(const double* dense, const double* knots, const int* num_steps, int n, int max_step, int nser,
                               double* coef, double* scratch, void* stream);
int eobhip_spline_build_batched(const double* dense, const double* knots, const int* num_steps, int n, int max_step,
                                double* coef, double* scratch, void* stream);
/* h+ = amp * Re[h22(t_j)] on t_j = t_0 + j dT_M (dT_M [n] in units of M), j < N; zero past the
 * last knot; nsamp[n] = samples covered by the knots (eob/seobnrv4phm.py:2319-2372). */
int eobhip_resample_hplus(const double* dense, const double* coef, const int* num_steps, const double* amp,
                          const double* dT_M, int n, int max_step, int N, double* hplus, int* nsamp, void* stream);
/* eobhip_resample_hplus with the cross polarization as well: h+ + i hx = amp * h22(t_j),
 * the real and imaginary parts of the reference's template_channels, which it
 * transforms as the two polarizations (eob/seobnrv4phm.py:720, 742-744); hcross [n][N]. */
/* Higher modes of a batch (get_hlms for ells, mms, eob/seobnrv4phm.py:2138-2221, through
 * EOBGetSpinFactorizedWaveform, src/EOB.cu:1236-1613): the waveform tables of nm <= 15
 * modes per binary (wtab: n * nm * eobhip_wave_mode_bytes() bytes) from the batch's
 * additional args and tables (use_hm as at setup), then the modes on the knots,
 * knots [n][max_step][2 nm + 1] = Re(h_j e^{i m_j phi}), Im(...), phi -- the series
 * compute_full_waveform splines (:2336-2356). */
size_t eobhip_wave_mode_bytes(void);
int eobhip_wave_tables(const double* add, int n, int use_hm, const void* bindata, const void* modes, int nm, const int* ls,
                       const int* ms, void* wtab, void* stream);
int eobhip_hlm_modes_knots(const double* dense, const int* num_steps, const void* bindata, const void* wtab, const double* add, int n,
                           int max_step, int nm, double* knots, void* stream);
/* The mode sum on the uniform grid from splines of those 2 nm + 1 series (coef
 * [n][max_step][4 (2 nm + 1)], eobhip_spline_build_series): h = amp sum_j S_j(t_j)
 * e^{-i m_j phi(t_j)}, hplus = Re h, hcross = Im h [n][N] (the template's real and
 * imaginary parts, eob/seobnrv4phm.py:2366-2372, 720, 742-744); ms: host array, 1 <= m <= 8. */
int eobhip_resample_modes(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M, int n,
                          int max_step, int N, int nm, const int* ms, double* hplus, double* hcross, int* nsamp, void* stream);
/* TDInterp_wrap2 / TDInterp2 (src/tdinterp.pyx:60-75, src/interpolate.cu:441-549):
 * templateChannels complex [numBinAll][data_length] = sum over modes of (re + i im)
 * e^{-i m phi} at dataTime[i], i < lengths[b], intervals inds [numBinAll][max_length];
 * propArrays, c1, c2, c3 [numBinAll][2 numModes + 1][old_length] (Re modes, Im modes,
 * phi); ts [numBinAll][old_length]; ls, ms, inds, lengths, old_lengths device int32.
 * legacy_stride 1 indexes the series of binary b with stride old_lengths[b] as the
 * reference does (src/interpolate.cu:521-527; right only when every binary has
 * old_length knots); 0 uses the padded stride old_length. */
int eobhip_TDInterp2(void* templateChannels, const double* dataTime, const double* ts, const double* propArrays, const double* c1,
                     const double* c2, const double* c3, int old_length, const int* old_lengths, int data_length, int numBinAll,
                     int numModes, const int* ls, const int* ms, const int* inds, const int* lengths, int max_length, int legacy_stride,
                     void* stream);
int eobhip_resample_hpc(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M, int n,
                        int max_step, int N, double* hplus, double* hcross, int* nsamp, void* stream);

/* ------------------------------------------------ detector projection
 * BBHWaveformTD.get_detector_information (eob/seobnrv4phm.py:597-634): per binary b,
 * theta = pi/2 - dec, phi = ra - gmst; Fplus, Fcross, time_shift [nifo][n] from the
 * detector tensors [nifo][3][3] and vertices [nifo][3] (metres); gmst [n] is the
 * caller's (the reference computes it with lal). */
int eobhip_detector_response(const double* ra, const double* dec, const double* gmst, const double* psi, int n, const double* tensors,
                             const double* vertices, int nifo, double* Fplus, double* Fcross, double* time_shift, void* stream);
/* The "detector_fd" signal of BBHWaveformTD.__call__ (eob/seobnrv4phm.py:742-767):
 * out [n][nifo][F] complex = (F+ hp~ + Fx hc~) exp(-2 pi i f dt), f = k df, zero for k < k_min;
 * hp_fd, hc_fd complex [n][F]. */
int eobhip_detector_signal(const double* hp_fd, const double* hc_fd, const double* Fplus, const double* Fcross, const double* time_shift,
                           int n, int nifo, int F, double df, int k_min, double* out, void* stream);

/* Row-wise real-to-complex FFT, n rows of N doubles -> n rows of N/2+1 complex (hipFFT D2Z;
 * like the reference's rfft without a dt factor, eob/seobnrv4phm.py:742-744). */
int eobhip_rfft_batched(const double* in, double* out, int n, int N, void* stream);
/* log L_b = -1/2 * 4 df * sum_k |d_k - h_b,k|^2 / S(f_k) with the analytic
 * S(f) = (f/f0)^-4 + 1 + (f/f0)^2 (1/S(0) = 0): the reduction of all_in_one_likelihood
 * plus the host finish of get_ll (eob/seobnrv4phm.py:772-823) in two device passes.
 * partial: eobhip_fd_likelihood_partial_doubles(n, F) doubles. */
int eobhip_fd_likelihood(const double* htilde, const double* dtilde, double* partial, int n, int F, double df, double f0,
                         double* logl, void* stream);
size_t eobhip_fd_likelihood_partial_doubles(int n, int F);
/* eobhip_fd_likelihood without a caller-provided buffer and without allocating:
 * the partial sums go to a static device array of the library (4 MiB per GPU,
 * eobhip_fd_likelihood_static_max_partials() doubles, held until the process
 * ends), so n * ceil(F / 2048) may not exceed that count (EOBHIP_EINVAL
 * otherwise).  Calls that overlap on one GPU share the array and race.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). */
size_t eobhip_fd_likelihood_static_max_partials(void);
int eobhip_fd_likelihood_static(const double* htilde, const double* dtilde, int n, int F, double df, double f0, double* logl,
                                void* stream);

/* ------------------------------------------------ likelihood from the support
 * The same log L as eobhip_fd_likelihood on rfft(h+) (the reference's rfft +
 * all_in_one_likelihood + get_ll finish, src/interpolate.cu:553-628,
 * eob/seobnrv4phm.py:742-807), computed as -2 df (DD - 2 sum_j h_j g_j + hh)
 * with the data-only terms DD, g and the per-M weight bank set up once and hh
 * from one M-point FFT of h on its support (M = power of two >= 2 x support,
 * at most N).  N is a power of two; M ranges over 2^m_min .. N.
 *
 * Set-up for one data set: dtilde complex [N/2+1] -> g [N], dd [1], bank
 * [eobhip_lik_bank_doubles(N, m_min_log2)]; scratch eobhip_lik_setup_scratch_bytes(N). */
size_t eobhip_lik_bank_doubles(int N, int m_min_log2);
size_t eobhip_lik_setup_scratch_bytes(int N);
int eobhip_lik_setup(const double* dtilde, int N, double df, double f0, int m_min_log2, double* g, double* dd, double* bank,
                     void* scratch, void* stream);
/* Downstream gather: for j < k, binary idx[j] (< n_src) of a batch into slot j of
 * downstream buffers of the same layouts: its dense rows below num_steps, num_steps,
 * its bindata and modes tables (bindata_bytes and modes_bytes per binary, multiples
 * of 8), amp and pos, and dT_M_out[j] = (1 / (total_mass[idx[j]] * mtsun)) * dt.  No
 * reference counterpart: eobgpu_amd.pipeline.StreamedPass collects the binaries
 * that have finished integrating with one launch. */
int eobhip_gather_binaries(const long long* idx, int k, int n_src, int max_step, const double* dense, const int* num_steps,
                           const void* bindata, size_t bindata_bytes, const void* modes, size_t modes_bytes, const double* total_mass,
                           const double* amp, const long long* pos, double dt, double mtsun, double* dense_out, int* num_steps_out,
                           void* bindata_out, void* modes_out, double* dT_M_out, double* amp_out, long long* pos_out, void* stream);

/* Per template: support length nsamp[n] (as eobhip_resample_hplus counts it) and log2 M in mlog2[n]. */
int eobhip_lik_sizes(const double* dense, const int* num_steps, const double* dT_M, int n, int max_step, int N, int m_min_log2,
                     int* nsamp, int* mlog2, void* stream);
/* Partial-sum counts per template for the largest log2 M of the batch (host outputs). */
int eobhip_lik_partial_blocks(int max_mlog2, int* gx_resample, int* gx_hh);
/* h on j < M_b at hbuf + rowoff[b] (zero for nsamp_b <= j < M_b) and per-block partials of
 * sum_j h_j g_j in cross_partial [n][gx_resample]. */
int eobhip_lik_resample(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M,
                        const int* nsamp, const int* mlog2, const long long* rowoff, int n, int max_step, int max_mlog2,
                        const double* g, double* hbuf, double* cross_partial, void* stream);
/* X (complex, per template M_b/2+1 bins at X + xoff[b], from eobhip_rfft_batched on hbuf) ->
 * logl[n]; hh_partial [n][gx_hh] scratch; terms (optional, [n][2]) = (sum h g, hh). */
int eobhip_lik_finish(const double* X, const long long* xoff, const int* mlog2, int n, int max_mlog2, int m_min_log2,
                      const double* bank, const double* cross_partial, double* hh_partial, const double* dd, double df, double* logl,
                      double* terms, void* stream);

/* ---------------------------------------------------- merger-ringdown attachment
 * The fits of the NQC and ringdown attachment, evaluated elementwise on device
 * arrays of n: which selects the function (EOBHIP_FIT_*), (l, m) the mode, x1..x4
 * its arguments in the reference's order (unused ones may be null):
 *   DT22       EOBGetNRSpinPeakDeltaTv4(2, 2, m1, m2, chi1, chi2)  eob/seobnrv4phm.py:878
 *   AMP/ADOT/ADDOT  EOBGetNRSpinPeak{Amplitude,ADot,ADDot}V4(l, m, m1, m2, chiS, chiA)  :945, :1078, :1162
 *   OMEGA/OMEGADOT  EOBGetNRSpinPeak{Omega,OmegaDot}V4(l, m, eta, a)  :1261, :1363
 *   C1F/C2F/D1F/D2F EOBCalculateRD{Amplitude,Phase}Coefficient{1,2}(l, m, eta, chi)  eob/eobstuff.py:44-322
 *   FINAL_MASS compute_final_mass_SEOBNRv4(m1, m2, a1, a2)  eob/eobstuff.py:6
 *   FINAL_SPIN bbh_final_spin_non_precessing_HBR2016(m1, m2, chi1, chi2, "M3J4")  (pesummary; called at
 *              eob/seobnrv4phm.py:2459): the published HBR2016 fit
 *   ISCO_RADIUS the Kerr ISCO radius of spin x1 (lalsimulation's SimRadiusKerrISCO, eob/eobstuff.py:15) */
#define EOBHIP_FIT_DT22 0
#define EOBHIP_FIT_AMP 1
#define EOBHIP_FIT_ADOT 2
#define EOBHIP_FIT_ADDOT 3
#define EOBHIP_FIT_OMEGA 4
#define EOBHIP_FIT_OMEGADOT 5
#define EOBHIP_FIT_C1F 6
#define EOBHIP_FIT_C2F 7
#define EOBHIP_FIT_D1F 8
#define EOBHIP_FIT_D2F 9
#define EOBHIP_FIT_FINAL_MASS 10
#define EOBHIP_FIT_FINAL_SPIN 11
#define EOBHIP_FIT_ISCO_RADIUS 12
#define EOBHIP_FIT_COUNT 13
int eobhip_imr_fit(int which, int l, int m, const double* x1, const double* x2, const double* x3, const double* x4, int n, double* out,
                   void* stream);
/* Kerr QNM frequency M omega_lm0 (e^{-i omega t}, Im < 0; omega complex [n]) of holes of spin af[n]:
 * Leaver's continued fractions, replacing lalsimulation's SimIMREOBGenerateQNMFreqV2fromFinal
 * behind compute_QNM (eob/eobstuff.py:32-41, which returns conj(omega) / M_f).  The fundamental
 * mode of l = max(2, |m|) <= 5; NaN for |af| > 0.9999.  The first call on a device fills a
 * table of the ringdown modes' roots (199 spins) and waits for it. */
int eobhip_kerr_qnm(int l, int m, const double* af, int n, double* omega, void* stream);
/* The attachment's parameters for n binaries and nm <= 8 modes (ls, ms host arrays; each
 * mode one of (2,2), (2,1), (3,3), (4,4), (5,5)), compute_full_waveform
 * (eob/seobnrv4phm.py:2308-2519) per binary: masses and spins m1, m2, chi1, chi2 [n] (any
 * mass unit), the trajectory (dense, num_steps), the mode splines coef [n][max_step][4 (2 nm + 1)]
 * (eobhip_spline_build_series over eobhip_hlm_modes_knots' series) and the splines of r and pr
 * rp_coef [n][max_step][8], grid spacing dT_M [n] (units of M).  Outputs: qnm [n][nm] complex
 * (lalsimulation's convention omega_r + i / tau, units of M), nqc [n][nm][5] = a1 a2 a3 b1 b2
 * (EOBCalculateNQCCoefficientsV4_freeattach, :1465-1683), rd [n][nm][eobhip_imr_rd_doubles()]
 * (attach values amp, damp, phi, omega; c1c c2c c1f c2f d1c d1f d2f; omega_r, 1/tau; eta),
 * binf [n][eobhip_imr_bin_doubles()] (t_peak, t_match, dt, M_f, a_f, DeltaT22), bini
 * [n][eobhip_imr_bin_ints()] (grid samples, attachment sample, ringdown samples: this
 * binary's own 30 damping times, where the reference pads every binary to the batch's longest). */
int eobhip_imr_params(const double* m1, const double* m2, const double* chi1, const double* chi2, const double* dense,
                      const int* num_steps, const double* coef, const double* rp_coef, const double* dT_M, int n, int max_step, int nm,
                      const int* ls, const int* ms, double* qnm, double* nqc, double* rd, double* binf, int* bini, void* stream);
/* The IMR template on t_j = t_0 + j dT (j < N): amp [n] times the sum over modes of the
 * NQC-corrected modes (EOBNonQCCorrection, :1686-1700) up to the attachment sample and of the
 * ringdown (compute_MR_mode_free, :1779-1844) for the ringdown samples after it, 0 beyond;
 * hplus = Re, hcross = Im [n][N] (BBHWaveformTD's template_channels, :720); nsamp [n] the
 * nonzero length.  The parameters are eobhip_imr_params' outputs. */
int eobhip_imr_resample(const double* dense, const int* num_steps, const double* coef, const double* rp_coef, const double* amp,
                        const double* dT_M, int n, int max_step, int N, int nm, const int* ls, const int* ms, const double* nqc,
                        const double* rd, const double* binf, const int* bini, double* hplus, double* hcross, int* nsamp, void* stream);
size_t eobhip_imr_rd_doubles(void);
size_t eobhip_imr_bin_doubles(void);
size_t eobhip_imr_bin_ints(void);

#ifdef __cplusplus
}
#endif
#endif /* EOBHIP_H */
