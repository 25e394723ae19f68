// Launcher declarations: one explicit instantiation per working precision
// lives in sdns_f32.hip / sdns_f64.hip; the C-ABI layer (sdns_abi.hip) calls
// these with runtime line lengths.  Every launcher returns 0, or -1 when the
// length has no compiled kernel.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "sdns_lineplan.hpp"
#include "sdns_types.hpp"

// Line lengths with compiled kernels: strided (x, y) complex lines and the
// half-length complex lines of the z-axis real transforms.
#define SDNS_STRIDED_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768) X(1024) X(1536)
#define SDNS_HALF_SIZES(X) \
  X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(36) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(768)

namespace sdns {

// `var` selects the compiled schedule (points per thread) of the length;
// variants that do not exist for a length fall back to variant 0.
template <typename Rt> int launch_p1(int M, int var, const P1Args<Rt>& a, hipStream_t s);
template <typename Rt> int launch_p2(int M, int var, const P2Args<Rt>& a, hipStream_t s);
template <typename Rt> int launch_p3(int n, int var, const P3Args<Rt>& a, hipStream_t s);
// P5: returns the number of column tiles in *nblk (entries of the partial sums);
// var 2 and 3 run schedules 0 and 1 with the 3/2-rule stage kernels where they apply
template <typename Rt> int launch_p5(int M, int var, int mode, const P5Args<Rt>& a, hipStream_t s, int* nblk);
// dir = -1 forward, +1 inverse
template <typename Rt> int launch_strided(int M, int var, int dir, int nfields, const StridedArgs<Rt>& a,
                                          hipStream_t s);
// dir = -1 R2C, +1 C2R
template <typename Rt> int launch_z(int n, int var, int dir, int nfields, const ZArgs<Rt>& a, hipStream_t s);

// elementwise helpers on the spectral grid [3][N0][N1][Nh]
template <typename Rt> struct EwArgs {
  void* out;
  const void* a;
  const void* b;
  void* aux;          // pressure output / second output
  const Rt *kx, *ky, *kz;
  int N0, N1, Nh, N2;
  long n;             // points per component
  Rt nu;
  int mask_nyquist;
  int keep[3];        // largest |k| index kept per axis by the dealiasing (EW_SGS_FORCE)
};
// EW_SGS_FORCE: out[3] = P(-i k_j a_ij) for the symmetric tensor a[6] (xx, yy, zz, xy, xz, yz),
// P the projection k (k . f) / k^2 removed, modes outside keep[] and Nyquist planes zeroed.
enum EwOp { EW_CROSS1 = 0, EW_CROSS2 = 1, EW_PRESSURE_DIFFUSION = 2, EW_DIVERGENCE = 3,
            EW_MUL_MINUS_I = 4, EW_MASK_NYQUIST = 5, EW_SGS_FORCE = 6, EW_CROSS2_KOK = 7, EW_ADD_LINEAR = 8,
            EW_TRUNCATE = 9, EW_SET_ELSASSER = 10, EW_ELSASSER_PRODUCTS = 11, EW_GRAD_TRUNC = 12, EW_DOT3 = 13,
            EW_UU_PRODUCTS = 14, EW_DIV_UU = 15, EW_VV_TAIL = 16 };
template <typename Rt> int launch_elementwise(int op, const EwArgs<Rt>& a, hipStream_t s);
// FP64 (energy_fourier(u_hat), energy_fourier(curl_hat)) partial sums
template <typename Rt> int launch_energy(const EwArgs<Rt>& a, double* partials, int nblk, hipStream_t s);
int launch_sum_partials(const double* partials, int nblk, double* out, double e_scale, double o_scale, hipStream_t s);
// Large-scale forcing (demo/Isotropic.py:154-184): partial sums of (energy_fourier(u),
// energy_fourier(u on |K|^2 <= kf^2)) excluding the mean mode, then the in-place rescale of the
// low modes by alpha = sqrt((target - upper)/lower) read from sums[0..1] on the device.
template <typename Rt> int launch_force_reduce(const EwArgs<Rt>& a, Rt kf2, double* partials, int nblk, hipStream_t s);
template <typename Rt> int launch_force_scale(const EwArgs<Rt>& a, Rt kf2, const double* sums, double target,
                                              hipStream_t s);
// Shell spectrum (demo/Isotropic.py:88-115): per-block partial sums of uiui and mode counts for
// shells z = 1..nb (z = #{bins b_i = i + 1/2 : b_i < |K|}), out layout [nblk][2][nb].
template <typename Rt> int launch_spectrum(const EwArgs<Rt>& a, int nb, double* partials, int nblk, hipStream_t s);
int launch_sum_rows(const double* partials, int nblk, int width, double* out, hipStream_t s);
// Same column sums, one block per column (for few columns over many rows).
int launch_sum_cols(const double* partials, int nblk, int width, double* out, hipStream_t s);

// ------------------------------------------------------------------ integrator updates
// Linear combinations of state-shaped complex arrays, out = c[0] x[0] + c[1] x[1] + ...
// (maths/integrators.py:66-75,126-139,161-175): each product is rounded before it is
// added, the terms are added left to right; out may be one of the x.
constexpr int SDNS_LIN_MAX = 10;
template <typename Rt> struct LinArgs {
  void* out;
  const void* x[SDNS_LIN_MAX];
  Rt c[SDNS_LIN_MAX];
  int nterms;
  long n;             // complex elements per array
};
template <typename Rt> int launch_lincomb(const LinArgs<Rt>& a, hipStream_t s);
// Embedded-error norm of adaptiveRK (integrators.py:77-89): FP64 partial sums per block and
// component of |err / sc|^2, sc = atol + max(|u0|, |u1|) rtol in the working precision;
// layout [nblk][3].
template <typename Rt> int launch_errnorm(const void* err, const void* u0, const void* u1, Rt atol, Rt rtol,
                                          long n, double* partials, int nblk, hipStream_t s);

// ------------------------------------------------------------------ batched ensembles
// Line lengths of the batched pass kernels (strided x/y lines; half-length z
// lines).  Every block of a batched launch has the same size whatever the
// box, so the tile per length is fixed: T = threads * E / M columns for the
// strided passes, LPB = zthreads * E / n lines for the z pass.
#define SDNS_BATCH_STRIDED(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128)
#define SDNS_BATCH_HALF(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64)

// Schedule variant v of a length: v = 0 is the default; v = 1 halves the
// points per thread (twice the threads per line, half the columns or lines
// per block, so no more LDS) where that line plan exists, and is the default
// elsewhere.  A box's length reaches the kernels as m + BATCH_VSTRIDE * v.
constexpr int BATCH_VSTRIDE = 8192;
constexpr int batch_E0(int M) {
  switch (M) {
    case 16: return 4;  case 24: return 6;  case 32: return 8;  case 48: return 12;
    case 64: return 8;  case 96: return 12; case 128: return 16;
    default: return 0;
  }
}
constexpr int batch_zE0(int n) {
  switch (n) {
    case 8: return 8; case 12: return 6; case 16: return 8; case 24: return 6;
    case 32: return 8; case 48: return 6; case 64: return 8;
    default: return 0;
  }
}
constexpr int batch_halved(int M, int e0) { return e0 >= 4 && num_passes(M, e0 / 2) > 0 ? e0 / 2 : e0; }
constexpr int batch_E(int M, int v = 0) { return v ? batch_halved(M, batch_E0(M)) : batch_E0(M); }
constexpr int batch_zE(int n, int v = 0) { return v ? batch_halved(n, batch_zE0(n)) : batch_zE0(n); }
template <typename Rt> constexpr int batch_threads() { return sizeof(Rt) == 4 ? 256 : 128; }
template <typename Rt> constexpr int batch_zthreads() { return sizeof(Rt) == 4 ? 128 : 64; }
template <typename Rt> constexpr int batch_T(int M, int v = 0) {
  return batch_E(M, v) ? batch_threads<Rt>() * batch_E(M, v) / M : 0;
}
template <typename Rt> constexpr int batch_LPB(int n, int v = 0) {
  return batch_zE(n, v) ? batch_zthreads<Rt>() * batch_zE(n, v) / n : 0;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// Every column tile of T columns at element offsets hi * hs + lo (lo < cdiv),
// rows rs and fields fs apart, covers whole 128-byte lines of `p`: then the
// strided kernels stream the buffer (StridedArgs::stream_in / stream_out).
template <typename Rt> inline bool tiles_whole_lines(const void* p, int T, long cdiv, long hs, long rs, long fs) {
  const long esz = 2 * (long)sizeof(Rt);
  return T > 0 && T * esz % 128 == 0 && cdiv % T == 0 && hs * esz % 128 == 0 && rs * esz % 128 == 0 &&
         fs * esz % 128 == 0 && reinterpret_cast<uintptr_t>(p) % 128 == 0;
}
// LDS (complex elements) of a strided batch kernel compiled for lengths up
// to lmax: tile + twiddle table
template <typename Rt> constexpr int batch_strided_lds(int lmax) {
  int m = 0;
#define SDNS_X(M) if (M <= lmax) m = cmax(m, M * cmax(batch_T<Rt>(M, 0), batch_T<Rt>(M, 1)) + M);
  SDNS_BATCH_STRIDED(SDNS_X)
#undef SDNS_X
  return m;
}
// LDS of the z batch kernel: three line slots (stride zslot) per line + two tables
template <typename Rt> constexpr int batch_z_lds(int lmax) {
  int m = 0;
#define SDNS_X(n)                                                                                         \
  if (n <= lmax)                                                                                          \
    for (int v = 0; v < 2; ++v)                                                                           \
      m = cmax(m, 3 * batch_LPB<Rt>(n, v) * z3_slot(n, n / batch_zE(n, v), (int)sizeof(Rt) * 2) + 2 * n);
  SDNS_BATCH_HALF(SDNS_X)
#undef SDNS_X
  return m;
}

// args / len / work are device arrays: one argument block and one line length
// per box, one work item per block (nwork blocks, a multiple of 8); lmax is
// the largest line length of the ensemble for the pass.
template <typename Rt> int launch_batch_p1(int lmax, const P1Args<Rt>* args, const int* len, const BatchWork* work,
                                           int nwork, hipStream_t s);
template <typename Rt> int launch_batch_p2(int lmax, const P2Args<Rt>* args, const int* len, const BatchWork* work,
                                           int nwork, hipStream_t s);
template <typename Rt> int launch_batch_p3(int lmax, const P3Args<Rt>* args, const int* len, const BatchWork* work,
                                           int nwork, hipStream_t s);
template <typename Rt> int launch_batch_fwd(int lmax, const StridedArgs<Rt>* args, const int* len,
                                            const BatchWork* work, int nwork, int nfields, hipStream_t s);
template <typename Rt> int launch_batch_p5(int lmax, int mode, const P5Args<Rt>* args, const int* len,
                                           const BatchWork* work, int nwork, hipStream_t s);

}  // namespace sdns
