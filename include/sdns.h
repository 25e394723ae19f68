/* C-ABI of libsdns_hip.so: the MI355X (gfx950) triply periodic Navier-Stokes
 * RK4 hot path of spectralDNS, for ctypes / any FFI.
 *
 * Conventions
 *  - Arrays are device pointers (HIP), C order, caller-owned:
 *      spectral  (ncomp, N0, N1, N2/2+1) complex  (float2 / double2)
 *      physical  (ncomp, N0, N1, N2)     real     (float / double)
 *  - Forward transforms are scaled by 1/(N0 N1 N2); backward transforms are
 *    unscaled (shenfun convention used by spectralDNS/solvers/NS.py).
 *  - Wavenumbers k_i = (2 pi / L_i) * fftfreq(N_i)*N_i (axes 0, 1) and
 *    rfftfreq (axis 2), as NS.py:37-41.
 *  - Every call enqueues work on the plan's HIP stream and returns without
 *    synchronising.  A plan is not thread-safe; use one plan per thread.
 *  - Return value 0 on success, a negative SDNS_ERR_* code otherwise; the
 *    message of the last failure on the calling thread is sdns_last_error().
 *    No C++ exception crosses this interface and no call falls back to the CPU.
 */
#ifndef SDNS_H
#define SDNS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdns_plan sdns_plan;

enum {
  SDNS_OK = 0,
  SDNS_ERR_INVALID = -1,     /* bad argument */
  SDNS_ERR_UNSUPPORTED = -2, /* size / option without a compiled kernel */
  SDNS_ERR_HIP = -3,         /* HIP runtime error */
  SDNS_ERR_NOMEM = -4        /* device allocation failed */
};
enum { SDNS_SINGLE = 0, SDNS_DOUBLE = 1 };                                   /* spectralinit.py:25-29 */
enum { SDNS_DEALIAS_NONE = 0, SDNS_DEALIAS_23 = 1, SDNS_DEALIAS_32 = 2 };    /* config.py:181-183 */
enum { SDNS_MODE_DIRECT = 0, SDNS_MODE_OUTPUT_MASK = 1 };                   /* 2/3-rule variants */

/* Build the spaces of one box: replaces get_context()'s TensorProductSpace /
 * VectorSpace / get_dealiased / mask / K / K2 / K_over_K2 set-up
 * (spectralDNS/solvers/NS.py:12-48).  hip_stream may be NULL (null stream).
 * dealias_mode selects the 2/3-rule variant: SDNS_MODE_DIRECT truncates
 * |k_i| >= N_i/3 on the inputs of the dealiased backward transforms
 * (shenfun dealias_direct), SDNS_MODE_OUTPUT_MASK keeps |k_i| < 2/3(N_i/2+1)
 * (cpp/spectralDNS.cpp:122-129, spectralDNS3D_short.py:44-46). */
int sdns_plan_create(sdns_plan** plan, const int32_t N[3], const double L[3], int precision,
                     int dealias, int mask_nyquist, int dealias_mode, void* hip_stream);
int sdns_plan_destroy(sdns_plan* plan);
/* Schedule variant of a pass: pass ids 0..4 as in sdns_timing_read, 5 for the
 * plain transforms; variant 0 is the default schedule, 1 the alternative
 * points-per-thread factorisation where one is compiled (otherwise variant 0
 * runs).  The z pass (id 2) has variants 2 and 3 too (its LDS-assisted scheme
 * with schedule 0 / 1) and 4 and 5 (rows staged HBM -> LDS by direct loads,
 * schedule 0 / 1; where the rows do not fit in LDS these run variant 2 / 3).
 * The x-forward + RK pass (id 4) has variants 2 and 3 too: schedules 0 / 1
 * whose plain RK stages under the 3/2 rule use kernels with the kept x rows
 * fixed at compile time (x lines of 96 to 768 points, arrays within 2 GiB;
 * other cases run as variants 0 / 1).
 * Set by the autotuner (spectraldns_amd/autotune.py). */
int sdns_plan_set_variant(sdns_plan* plan, int pass, int variant);
/* Rebind the plan to another HIP stream (NULL = null stream). */
int sdns_plan_set_stream(sdns_plan* plan, void* hip_stream);
/* Device workspace the plan needs; sdns_set_workspace hands it caller memory
 * (at least sdns_workspace_bytes, 256-byte aligned).  Without it the plan
 * allocates its own on first use. */
size_t sdns_workspace_bytes(const sdns_plan* plan);
int sdns_set_workspace(sdns_plan* plan, void* dptr, size_t bytes);

/* VT.forward / VT.backward of ncomp components (set_velocity, get_velocity,
 * get_curl, get_pressure, get_divergence: NS.py:86-110). */
int sdns_forward(sdns_plan* plan, const void* u_phys, void* u_hat, int ncomp);
int sdns_backward(sdns_plan* plan, const void* u_hat, void* u_phys, int ncomp);

/* Kernel plug-ins of the reference's @optimizer layer
 * (spectralDNS/optimization/__init__.py:12-55):
 *  cross1: c = a x b on npoints physical points per component (maths/cross.py:13-25)
 *  cross2: c_hat = 1j*(K x a_hat) (maths/cross.py:27-32)
 *  add_pressure_diffusion: NS.py:203-217 in place on rhs; P_hat may be NULL
 *  divergence: div_hat = 1j*(K . u_hat) (NS.py:106-110)
 *  mul_minus_i: out = -1j*a on n complex points (NS.py:96-99)
 *  mask_nyquist: zero the Nyquist planes of a (3, N0, N1, Nh) array (NS.py:253-254) */
int sdns_cross1(sdns_plan* plan, void* c, const void* a, const void* b, size_t npoints);
int sdns_cross2(sdns_plan* plan, void* c_hat, const void* a_hat);
int sdns_add_pressure_diffusion(sdns_plan* plan, void* rhs, const void* u_hat, double nu, void* p_hat);
int sdns_divergence(sdns_plan* plan, void* div_hat, const void* u_hat);
int sdns_mul_minus_i(sdns_plan* plan, void* out, const void* a, size_t n);
/* Velocity-vorticity solver (spectralDNS/solvers/VV.py):
 *  cross2_k_over_k2: u_hat = 1j*((K/K^2) x w_hat), K^2 = 0 taken as 1
 *    (compute_velocity, VV.py:57-73: cross2(v_hat, K_over_K2, w_hat))
 *  add_linear: rhs -= nu*K2*w_hat, then rhs += source when not NULL
 *    (add_linear, VV.py:108-113), in place on rhs */
int sdns_cross2_k_over_k2(sdns_plan* plan, void* u_hat, const void* w_hat);
int sdns_add_linear(sdns_plan* plan, void* rhs, const void* w_hat, double nu, const void* source);
/* The VV right-hand side after the convection passes, in one launch:
 * rhs = 1j*(K x v_hat) (Conv, VV.py:104-106), Nyquist planes zeroed when the
 * plan masks them (VV.py:143-144), then rhs -= nu*K2*w_hat and rhs += source
 * when not NULL (add_linear, VV.py:110-115); v_hat = F(u x w). */
int sdns_vv_rhs_tail(sdns_plan* plan, void* rhs, const void* v_hat, const void* w_hat, double nu,
                     const void* source);
/* MHD solver in Elsasser form (spectralDNS/solvers/MHD.py), any dealiasing
 * (the 3/2 rule runs the products on the padded grid of
 * sdns_backward_dealiased / sdns_forward_dealiased):
 *  dealias_truncate: out = a with the modes outside the plan's kept input
 *    range zeroed, 3 components (the input truncation of shenfun's dealiased
 *    backward transform with dealias_direct, as VMp.backward(ub_hat) in
 *    Conv, MHD.py:140-146); for a 3/2-rule plan the kept input modes are
 *    |k| < N/2 (the Nyquist modes zeroed)
 *  elsasser_products: zz[3i+j] = (u_i + b_i)*(u_j - b_j) on npoints physical
 *    points per component, ub = (u0, u1, u2, b0, b1, b2) (divergenceConvection,
 *    MHD.py:118-130, before its forward transforms)
 *  set_elsasser: c_hat[0:3] = -1j*sum_j K_j (zz_ij + zz_ji)/2 and
 *    c_hat[3:6] = 1j*sum_j K_j (zz_ji - zz_ij)/2 from the 9 spectral
 *    products (set_Elsasser, MHD.py:104-113) */
int sdns_dealias_truncate(sdns_plan* plan, void* out, const void* a);
/* Backward transform of the dealiased space Tp / VTp / VMp (NS.py:29-32,
 * MHD.py:30-32; VMp.backward in Conv, MHD.py:140-146), ncomp <= 3:
 *  - 2/3 rule (dealias_direct) and no dealiasing: dealias_truncate followed
 *    by sdns_backward, in the transform's own passes (each pass reads only the
 *    kept modes of its axis); u is (ncomp, N0, N1, N2);
 *  - 3/2 rule: the padded transform; u is (ncomp, M0, M1, M2), M = 3N/2, the
 *    modes |k| < N/2 of every axis placed in the padded spectrum (the Nyquist
 *    modes dropped), the other bins zero. */
int sdns_backward_dealiased(sdns_plan* plan, const void* u_hat, void* u, int ncomp);
/* Forward transform of the dealiased space (Tp.forward in divergenceConvection,
 * MHD.py:118-130, and the Standard/Divergence/Skewed forms, NS.py:138-176),
 * ncomp <= 3: sdns_forward for the 2/3 rule and no dealiasing; for the 3/2
 * rule the forward transform of the padded grid (ncomp, M0, M1, M2), scaled by
 * 1/(M0 M1 M2), truncated to the modes -N/2..N/2-1 (axes 0, 1) and 0..N/2
 * (axis 2) of u_hat. */
int sdns_forward_dealiased(sdns_plan* plan, const void* u, void* u_hat, int ncomp);
int sdns_elsasser_products(sdns_plan* plan, void* zz, const void* ub, size_t npoints);
int sdns_set_elsasser(sdns_plan* plan, void* c_hat, const void* zz_hat);
/* The Standard, Divergence and Skewed convection forms of the NS solver
 * (spectralDNS/solvers/NS.py:138-199), any dealiasing (physical arrays on the
 * plan's grid, or on the padded grid for the 3/2 rule):
 *  gradient_truncated: out[j] = 1j*K_j*a on the kept modes, 0 elsewhere, for
 *    one spectral component a (the input of Tp.backward(1j*K[j]*U_hat[i]) in
 *    standard_convection, NS.py:138-145); the kept modes are those of
 *    sdns_dealias_truncate
 *  dot3: out = u0*g0 + u1*g1 + u2*g2 on npoints physical points
 *    (np.sum(u_dealias*gradUi, 0))
 *  uu_products: uu = (u0u0, u0u1, u0u2, u1u1, u1u2, u2u2) on npoints
 *    physical points (the products of divergence_convection, NS.py:147-165)
 *  div_uu: out = scale * 1j*sum_j K_j F(u_i u_j) from the six spectral
 *    products, with the reference's grouping of the sums (scale -1 gives the
 *    Divergence Conv, NS.py:180-185) */
int sdns_gradient_truncated(sdns_plan* plan, void* out, const void* a_hat);
int sdns_dot3(sdns_plan* plan, void* out, const void* u, const void* g, size_t npoints);
int sdns_uu_products(sdns_plan* plan, void* uu, const void* u, size_t npoints);
int sdns_div_uu(sdns_plan* plan, void* out, const void* uu_hat, double scale);
int sdns_mask_nyquist(sdns_plan* plan, void* a_hat);

/* Conv of getConvection('Vortex') (NS.py:191-198): rhs = VTp.forward(u x curl u)
 * on the dealiased space, without mask, pressure or diffusion. */
int sdns_convection(sdns_plan* plan, const void* u_hat, void* rhs);

/* F_k(u_dealias x w_dealias) with the vorticity given: out = VTp.forward(
 * VTp.backward(u_hat) x VTp.backward(w_hat)) on the dealiased space, without
 * mask (Conv of the velocity-vorticity solver before its cross2,
 * spectralDNS/solvers/VV.py:91-99: compute_velocity, VTp.backward(w_hat),
 * Cross).  w_hat need not be a curl: no relation between u_hat and w_hat is
 * assumed.  out may not alias u_hat or w_hat. */
int sdns_convection_uw(sdns_plan* plan, const void* u_hat, const void* w_hat, void* out);

/* ComputeRHS (NS.py:219-261) with the Vortex convection (NS.py:191-198):
 * dealiased u x curl(u), Nyquist mask, pressure and diffusion, + source
 * (source and p_hat may be NULL). */
int sdns_compute_rhs(sdns_plan* plan, const void* u_hat, void* rhs, double nu, const void* source,
                     void* p_hat);

/* nsteps classical RK4 steps in place on u_hat (integrators.py:149-159,
 * coefficients in the working precision :185-186), step s using dt_seq[s]
 * (host array, the dt sequence of the solve() loop).  diag, when not NULL,
 * is a device array of 2*nsteps doubles receiving after every step the
 * kinetic energy 0.5*sum w|u_hat|^2 and the dissipation nu*sum w|K x u_hat|^2
 * (FP64 reductions; w = 1 on kz = 0, N2/2 and 2 elsewhere).
 * u_lo (single precision only; ignored in double): low-order part of the
 * state for compensated accumulation.  The state is then u_hat + u_lo, the
 * RK4 sum u2 is carried as an exact (hi, lo) pair with TwoSum updates, and
 * after each step u_hat = fl(u_hat + u_lo) and u_lo holds the remainder;
 * start it at zero.  NULL = plain single-precision arithmetic.  p_hat, when
 * not NULL, receives the pressure of the last RHS evaluation, as the
 * reference leaves P_hat after a step. */
int sdns_rk4_steps(sdns_plan* plan, void* u_hat, void* u_lo, const double* dt_seq, int nsteps, double nu,
                   double* diag, void* p_hat);
/* sdns_rk4_steps with a Source added to every stage's right-hand side in
 * every mode (rhs += Source after pressure and diffusion, NS.py:256-258, as
 * the reference's RK4 calls ComputeRHS per stage, integrators.py:149-159).
 * source: device array shaped like u_hat (NULL = none); held fixed through
 * the steps, and it may not alias u_hat or u_lo. */
int sdns_rk4_steps_source(sdns_plan* plan, void* u_hat, void* u_lo, const double* dt_seq, int nsteps, double nu,
                          const void* source, double* diag, void* p_hat);

/* Kinetic energy and dissipation of u_hat in FP64 into the device array out[2]. */
int sdns_energy(sdns_plan* plan, const void* u_hat, double nu, double* out);

/* Array updates of the ForwardEuler, AB2 and BS5 (adaptiveRK) integrators
 * (spectralDNS/maths/integrators.py:66-75 stage and solution sums, :161-175
 * u0 += rhs*dt and AB2), which run with sdns_compute_rhs as their RHS:
 * out = c[0]*x[0] + c[1]*x[1] + ... + c[nterms-1]*x[nterms-1] over
 * state-shaped (3, N0, N1, N2/2+1) complex arrays, 1 <= nterms <= 10.  The
 * coefficients are rounded to the working precision, each product is rounded
 * before it is added and the terms are added left to right (NumPy's order
 * for `u = x0; u += c1*x1; ...`).  out may be one of the x (in place). */
int sdns_lincomb(sdns_plan* plan, void* out, int nterms, const void* const* x, const double* c);
/* Embedded error estimate of adaptiveRK with errnorm "2"
 * (integrators.py:77-89): out[k] (device, FP64) = sum over the modes of
 * component k of |err_k / sc_k|^2 with sc = atol + max(|u0|, |u1|)*rtol in
 * the working precision; the caller takes max_k sqrt(out[k]) / sqrt(N0 N1 Nh). */
int sdns_error_norm(sdns_plan* plan, const void* err, const void* u0, const void* u1, double atol, double rtol,
                    double* out);

/* Large-scale forcing (demo/Isotropic.py:154-184) in place on u_hat: zero the
 * mean mode, then scale the modes with |K|^2 <= kf^2 by
 * alpha = sqrt((target - E_upper) / E_lower) so that energy_fourier(u_hat)
 * equals target; the energies are FP64 reductions and alpha is computed on the
 * device (no host round trip).  sums (device, 2 doubles, may be NULL) receives
 * (energy_fourier before scaling, E_lower). */
int sdns_force(sdns_plan* plan, void* u_hat, double target, double kf, double* sums);
/* sdns_force on the compensated single-precision state u_hat + u_lo of
 * sdns_rk4_steps (same forcing, Isotropic.py:154-184): the energies are FP64
 * sums over u_hat + u_lo, and each forced mode is scaled as the FP64 value
 * alpha*(u_hat + u_lo), then split again into u_hat = fl(value) and u_lo =
 * the remainder.  SDNS_ERR_INVALID for a double-precision plan. */
int sdns_force_compensated(sdns_plan* plan, void* u_hat, void* u_lo, double target, double kf, double* sums);

/* Shell sums of the energy spectrum (demo/Isotropic.py:88-115): out (device,
 * 2*nbins doubles) receives, for shells z = 1..nbins, the sum over modes with
 * z - 1/2 < |K| <= z + 1/2 of uiui = sum_c w |u_hat_c|^2 (out[z-1]) and the
 * number of such modes (out[nbins + z - 1]). */
int sdns_spectrum(sdns_plan* plan, const void* u_hat, int nbins, double* out);
/* sdns_spectrum that allocates nothing: no counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight diagnostics loops).  The per-block partial sums go to a
 * static device array of the library (4 MiB on every GPU it runs on, held
 * until the process ends) instead of the plan's workspace, so the plan's
 * workspace is not allocated by this call.  nbins is at most 8192
 * (SDNS_ERR_INVALID beyond).  Same results as sdns_spectrum up to the order
 * of the FP64 additions.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race. */
int sdns_spectrum_noalloc(sdns_plan* plan, const void* u_hat, int nbins, double* out);

/* Per-pass GPU time from HIP events recorded on the plan's stream around
 * every pass launch (benchmark support).  sdns_timing_enable(plan, 1) starts
 * recording; sdns_timing_read waits for the recorded events, writes the summed
 * milliseconds and launch counts per pass id (0: x-inverse, 1: y-inverse +
 * curl, 2: z pass + cross product, 3: y-forward, 4: x-forward + projection +
 * RK stage, 5: other launches) and clears the record. */
int sdns_timing_enable(sdns_plan* plan, int enable);
int sdns_timing_read(sdns_plan* plan, double ms[6], int count[6]);
/* Pass ids recorded while timing is enabled (bit i = pass id i; default 63,
 * every pass).  Fewer events keep the timed launches back to back. */
int sdns_timing_passes(sdns_plan* plan, int mask);

/* ------------------------------------------------------------------------
 * Batched ensembles of independent boxes.  No counterpart in the reference,
 * which advances one box per process (spectralDNS/__init__.py:69-123); each
 * box of a batch takes exactly the RK4 step of sdns_rk4_steps
 * (integrators.py:149-159 over ComputeRHS, NS.py:219-261) with its own N, L,
 * nu and dt.  One step of the whole ensemble is 20 pass launches, each
 * covering every box, plus one reduction, captured once as a hipGraph on the
 * batch's own stream and replayed per step.
 *
 * sdns_batch_create: nbox boxes, N and L of box i at N[3i..3i+2], L[3i..3i+2];
 * precision, dealias (2/3 or 3/2 rule), mask_nyquist and dealias_mode as in
 * sdns_plan_create and shared by all boxes.  Every dealiased line length
 * (3N/2 under the 3/2 rule) must have a batched kernel: x and y in
 * {16, 24, 32, 48, 64, 96, 128}, z/2 in {8, 12, 16, 24, 32, 48, 64}; otherwise
 * SDNS_ERR_UNSUPPORTED.  hip_stream is the caller's stream (NULL = null
 * stream): sdns_batch_rk4_steps orders its work after what the caller has
 * enqueued there and makes that stream wait for its completion. */
typedef struct sdns_batch sdns_batch;
int sdns_batch_create(sdns_batch** batch, int nbox, const int32_t* N, const double* L, int precision,
                      int dealias, int mask_nyquist, int dealias_mode, void* hip_stream);
int sdns_batch_destroy(sdns_batch* batch);
/* u_hat[i]: device pointer of box i's state (3, N0, N1, N2/2+1), caller-owned;
 * advanced in place.  The batch keeps the pointers until they are set again. */
int sdns_batch_set_state(sdns_batch* batch, void* const* u_hat);
/* Viscosity and time step of every box (host arrays of nbox doubles; nu and
 * dt are rounded to the working precision as config.py:123-127 does). */
int sdns_batch_set_params(sdns_batch* batch, const double* nu, const double* dt);
/* 1 (default): replay the captured step graph; 0: enqueue the launches. */
int sdns_batch_use_graph(sdns_batch* batch, int enable);
int sdns_batch_set_stream(sdns_batch* batch, void* hip_stream);
/* nsteps RK4 steps of every box.  diag, when not NULL, is a device array of
 * 2*nbox doubles receiving (kinetic energy, dissipation) of every box after
 * the last step, in FP64, with the definitions of sdns_rk4_steps. */
int sdns_batch_rk4_steps(sdns_batch* batch, int nsteps, double* diag);
/* Schedule of one batched pass for one line length (no counterpart in the
 * reference, which has no batched path): pass 0..4 (x-inverse, y-inverse +
 * curl, z pass, y-forward, x-forward + RK), length = the line length of that
 * pass (3N/2 under the 3/2 rule; N2/2 or 3N2/4 for the z pass), variant 0
 * (default) or 1 (half the points per thread: twice the threads per line and
 * half the columns or lines per block, where that line plan exists; otherwise
 * the same as 0).  Lays the work lists out again and drops the captured step;
 * the results do not depend on the variant beyond rounding.  Set by the
 * ensemble autotuner (spectraldns_amd/autotune.py). */
int sdns_batch_set_variant(sdns_batch* batch, int pass, int length, int variant);
/* The single-box plan of box i (owned by the batch; for diagnostics such as
 * sdns_energy or sdns_spectrum on that box). */
int sdns_batch_box_plan(sdns_batch* batch, int i, sdns_plan** plan);

/* ------------------------------------------------------------------------
 * Learned subgrid closure.  No counterpart in the reference (its solvers run
 * resolved DNS only); the coarse-grid step it serves is the reference's RK4
 * (integrators.py:149-159) with a Source term (NS.py:256-258) recomputed from
 * every stage: Source = P(-i k_j tau_ij), tau predicted from the filtered
 * velocity by a network of three periodic 3x3x3 conv layers 3 -> 32 -> 32 -> 6
 * (bias, SiLU after layers 1 and 2), bf16 weights and activations, FP32
 * accumulation on the matrix cores.
 *
 * Parameters are one FP32 device vector of sdns_closure_nparams() entries:
 * W1[27][32][3] b1[32] W2[27][32][32] b2[32] W3[27][6][32] b3[6], W[tap][cout][cin]
 * with tap = 9 (dz+1) + 3 (dy+1) + (dx+1) applied to the input at the point
 * (i0+dz, i1+dy, i2+dx) (periodic).  tau components: xx, yy, zz, xy, xz, yz.
 * Physical fields are planar FP32 device arrays [c][N0][N1][N2].
 * Grid: N2 % 32 == 0 and N0*N1*N2 % 4096 == 0. */
typedef struct sdns_closure sdns_closure;
size_t sdns_closure_nparams(void);
int sdns_closure_create(sdns_closure** closure, const int32_t N[3], void* hip_stream);
int sdns_closure_destroy(sdns_closure* closure);
int sdns_closure_set_stream(sdns_closure* closure, void* hip_stream);
/* Convert the FP32 parameter vector to the packed bf16 kernel layouts; call
 * again after every parameter update. */
int sdns_closure_set_params(sdns_closure* closure, const float* params);
/* tau (6 fields) = out_scale * net(in_scale * u) for u (3 fields). */
int sdns_closure_forward(sdns_closure* closure, const float* u_phys, float in_scale, float out_scale,
                         float* tau_phys);
/* One training evaluation: loss = mean over points and the 6 components of
 * (net(in_scale u) - tau / out_scale)^2 into *loss (device double, may be
 * NULL) and its gradient w.r.t. every parameter into grad (FP32, nparams). */
int sdns_closure_train_step(sdns_closure* closure, const float* u_phys, const float* tau_phys, float in_scale,
                            float out_scale, float* grad, double* loss);
/* Adam update (Kingma & Ba) of n FP32 parameters at step t >= 1 (device arrays). */
int sdns_adam_step(float* params, const float* grad, float* m, float* v, size_t n, float lr, float beta1,
                   float beta2, float eps, int t, void* hip_stream);
/* Box filter by r of the physical velocity u (3 fields, plan's grid and
 * precision): ubar (3 fields) = means over non-overlapping r^3 blocks, tau (6
 * fields) = bar(u_i u_j) - ubar_i ubar_j with FP64 sums. */
int sdns_box_filter(sdns_plan* plan, const void* u_phys, int r, void* ubar, void* tau);
/* sdns_rk4_steps on a single-precision plan of the closure's grid with the
 * closure's Source added to every stage's right-hand side.  per_stage = 0:
 * the Source is evaluated once per step from the step's initial state and held
 * through the four stages, as the reference holds Source during a step
 * (it is set between steps by the update hook); per_stage = 1: evaluated from
 * every stage state.  Evaluation: the state's velocity is transformed back,
 * passed through the network, tau transformed forward. */
int sdns_closure_rk4_steps(sdns_closure* closure, sdns_plan* plan, void* u_hat, const double* dt_seq, int nsteps,
                           double nu, float in_scale, float out_scale, int per_stage, double* diag);

/* Human-readable message of the last failing call on this thread. */
const char* sdns_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SDNS_H */
