// Kernel argument blocks shared by the host orchestration (sdns_abi.hip) and
// the pass kernels (sdns_kernels.hpp).  Everything is plain data passed by
// value as kernel arguments.
//
// Buffer layouts (complex elements of the working precision):
//   state  u_hat           [3][N0][N1][Nh]        Nh = N2/2+1, kz contiguous
//   W1 (x-inverse out)     [5][M0][nky][nkz]      fields u, v, w, X(kx w), X(kx v)
//   W2 (y-inverse out)     [6][M0][M1][nkz]       fields u, v, w, wx, wy, wz
//   W3 (z pass out)        [3][M0][M1][nkzo]      F_z(u x w)
//   W4 (y-forward out)     [3][M0][nkyo][nkzo]
// nky/nkz are the kept input modes of the dealiased backward transform,
// nkyo/nkzo the kept output modes of the dealiased forward transform, stored
// compactly in fftfreq order (k >= 0 first, then the negative modes).  The kz
// dimension of the work buffers is padded to nkzp / nkzop (multiples of 16
// elements) so that every row starts on a 128-byte line.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace sdns {

// Map between FFT bins of a length-M transform and stored rows.  Kept modes
// are -neg..pos; mode k lives in bin (k mod M) and in row (k >= 0 ? k : base + k),
// where base is the stored length (N for the uncompacted state, neg+pos+1 for
// compact work buffers).
struct RowMap {
  int neg, pos, base;
};

template <typename Rt> struct P1Args {  // x-inverse of u_hat + kx products
  const void* u;      // state [3][N0][N1][Nh]
  const void* w;      // given vorticity [3][N0][N1][Nh], or null: then W1 holds
                      // the six plain x-inverses u, v, w, wx, wy, wz (no kx products)
  void* w1;           // [5][M0][Q] ([6][M0][Q] with a given vorticity)
  const void* tw;     // forward twiddles of length M0
  const Rt* kx;       // wavenumbers of axis 0 by stored index
  int N0, N1, Nh;
  RowMap x;           // kept input modes of axis 0 (base = N0)
  int y_neg, y_pos;   // kept input modes of axis 1
  int nkz, nkzp;      // kept input kz count, padded kz stride of W1
  int Q;              // nky * nkzp columns (= W1 row stride)
};

template <typename Rt> struct P2Args {  // y-inverse + curl assembly
  const void* w1;
  void* w2;           // [6][M0][M1][nkz]
  const void* tw;     // forward twiddles of length M1
  const Rt* ky;       // wavenumbers of axis 1 by stored index
  const Rt* kz;       // wavenumbers of axis 2
  int N1, M0;
  RowMap y;           // kept input modes of axis 1 (base = nky)
  int nkz, nkzp;
  int Q;              // M0 * nkzp
};

template <typename Rt> struct P3Args {  // z C2R x6, cross product, z R2C x3
  const void* w2;
  void* w3;           // [3][L][nkzo]
  const void* tw;     // forward twiddles of length n = M2/2
  const void* rtw;    // exp(-2 pi i p / M2), p in [0, n)
  int nkz, nkzp, nkzo, nkzop;
  int L;              // M0 * M1 lines
};

// Generic batched strided-axis transform: columns q = hi*cdiv + lo live at
// offset hi*hi_stride + lo; row r of a column at offset r*row_stride; field f
// (blockIdx.y) at offset f*field_stride.
template <typename Rt> struct StridedArgs {
  const void* in;
  void* out;
  const void* tw;
  int Q, cdiv, cvalid;  // column q valid if q < Q and (q mod cdiv) < cvalid
  long in_hi, in_row, in_field;
  long out_hi, out_row, out_field;
  RowMap rin, rout;
  Rt scale;           // applied to every output (1 = none)
  int stream_in, stream_out;  // non-temporal loads / stores (set by the launcher
                              // when every column tile covers whole 128-byte lines)
};

// z-axis real transforms on contiguous lines: C2R (spectral lines of nk_in
// bins -> M2 reals) or R2C (M2 reals -> nk_out bins).  Field f via blockIdx.y.
template <typename Rt> struct ZArgs {
  const void* in;
  void* out;
  const void* tw;     // length n = M2/2
  const void* rtw;    // exp(-2 pi i p / M2)
  int L;
  int nk;             // spectral bins kept (input of C2R / output of R2C)
  long in_line, in_field, out_line, out_field;  // element strides
  Rt scale;
};

enum P5Mode { P5_RHS = 0, P5_RK0 = 1, P5_RK12 = 2, P5_RK3 = 3 };

template <typename Rt> struct P5Args {  // x-forward + mask + pressure + diffusion + RK4 stage
  const void* w4;     // [3][M0][nkyo][nkzo]
  void* A;            // step base u1 (and final result)
  void* B;            // stage value u0
  void* Cacc;         // accumulator u2
  void* Alo;          // compensated mode: low part of u1 / of the result (caller's u_lo)
  void* Clo;          // compensated mode: low part of u2
  void* rhs;          // P5_RHS: rhs output
  void* phat;         // P5_RHS: pressure output (may be null)
  const void* source; // P5_RHS: added source (may be null)
  double* partials;   // P5_RK3: per-block (energy, enstrophy) sums (may be null)
  const void* tw;     // forward twiddles of length M0
  const Rt *kx, *ky, *kz;
  int N0, N1, Nh, N2;
  RowMap x;           // kept output modes of axis 0
  int y_neg, y_pos;   // kept output modes of axis 1
  int nkzo, nkzop;    // kept output kz count, padded kz stride of W4
  int mask_nyquist;
  int conv_only;      // P5_RHS: write the scaled convection term only
  Rt nu, scale, adt, bdt;
  int Q;              // N1 * Nh state columns
};

// One block of a batched pass: box index (-1 for padding blocks) and the
// block's tile within that box.
struct BatchWork {
  int box, tile;
};

// LDS stride (complex elements) of one z line owned by P threads; the staged
// spectral line needs n + 1 entries.  With P >= 16 a line fills at least a
// quarter wave, and a stride of 16 modulo 32 puts the two lines of a fp32
// half-wave 32 banks apart, so their contiguous accesses share no bank.
constexpr int zslot(int n, int P) { return P >= 16 ? n + 1 + ((16 - (n + 1) % 32) + 32) % 32 : n + 1; }

// Lines per block of the LDS-assisted z scheme (three slots of stride sl per
// line within 40 KB), and the slot stride it uses: the padded stride when it
// keeps the lines per block and either the blocks that fit in a CU's 160 KB of
// LDS (twiddle tables included) or at least five of them.  Measured on the
// 256^3 z passes: padding gains 16% at n = 128 (6 -> 5 blocks per CU) and
// loses 15% at n = 192 (4 -> 3 blocks per CU).
constexpr int z3_lines(int n, int P, int esz, int sl) {
  int L = 256 / P;
  if (L < 1) L = 1;
  while (L > 1 && 3 * L * sl * esz > 40960) L /= 2;
  return L;
}
// The single-box kernel keeps the real-transform twiddles in LDS for FP32
// and reads them from global memory for FP64 (the batched kernel keeps both
// tables in LDS).
constexpr bool z3_rtw_in_lds(int esz) { return esz == 8; }
constexpr int z3_blocks_per_cu(int n, int P, int esz, int sl) {
  return 163840 / ((3 * z3_lines(n, P, esz, sl) * sl + (z3_rtw_in_lds(esz) ? 2 * n : n)) * esz);
}
constexpr int z3_slot(int n, int P, int esz) {
  return z3_lines(n, P, esz, zslot(n, P)) == z3_lines(n, P, esz, n + 1) &&
                 (z3_blocks_per_cu(n, P, esz, zslot(n, P)) >= z3_blocks_per_cu(n, P, esz, n + 1) ||
                  z3_blocks_per_cu(n, P, esz, zslot(n, P)) >= 5)
             ? zslot(n, P)
             : n + 1;
}

}  // namespace sdns
