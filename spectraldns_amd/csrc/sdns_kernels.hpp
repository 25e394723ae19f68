// Pass kernels of the fused pseudo-spectral RK4 step.
//
// One RHS evaluation (spectralDNS/solvers/NS.py:219-261, Vortex form
// NS.py:191-198) is five launches that each stream HBM once:
//   P1 x-inverse of u_hat (+ kx*w_hat, kx*v_hat for the curl)  [strided, x]
//   P2 y-inverse and curl assembly i(K x u_hat)                [strided, y]
//   P3 z C2R of 6 fields, u x curl(u), z R2C of 3 fields        [contiguous, z]
//   P4 y-forward                                                 [strided, y]
//   P5 x-forward, Nyquist mask, pressure projection, diffusion,
//      RK4 stage update (integrators.py:149-159) and, on the last
//      stage, the FP64 energy / enstrophy partial sums           [strided, x]
// Physical space never touches HBM.  Dealiasing is folded into the row maps:
// truncated modes are neither read nor written, padded bins are zeros.
//
// Every kernel takes E (points per thread) as a template parameter so the
// launcher can pick one of the compiled schedules per line length.
#pragma once
#include <cstdint>
#include <type_traits>

#include "sdns_device.hpp"
#include "sdns_types.hpp"

namespace sdns {

// ------------------------------------------------------------------ tiling
// Columns per block: at least 128 contiguous bytes per row (one cache line)
// and about 256 threads, at most 1024 threads, LDS tile M*T <= 64 KB, and T*P
// a multiple of 64 so that every block is made of whole waves (the P5 energy
// reduction shuffles across full 64-lane waves).
constexpr int strided_T(int M, int P, int esz) {
  int T0 = 256 / P;
  if (T0 < 128 / esz) T0 = 128 / esz;
  if (T0 * P > 1024) T0 = 1024 / P;
  if (T0 < 1) T0 = 1;
  if (T0 > 64) T0 = 64;
  for (int T = T0; T >= 1; --T)
    if ((T * P) % 64 == 0 && M * T * esz <= 65536) return T;
  for (int T = 64; T >= 1; --T)
    if ((T * P) % 64 == 0 && M * T * esz <= 65536 && T * P <= 1024) return T;
  return 0;
}
template <typename Rt, int M, int E_> struct SCfg {
  static constexpr int E = E_;
  static constexpr int P = M / E;
  static constexpr int T = strided_T(M, P, (int)sizeof(cx<Rt>));
  static constexpr int THREADS = T * P;
  static_assert(T > 0 && THREADS <= 1024, "no tile fits");
};


constexpr int z_lpb(int n, int P, int esz) {
  int L = 256 / P;
  if (L < 1) L = 1;
  while (L > 1 && L * zslot(n, P) * esz > 32768) L /= 2;
  return L;
}
template <typename Rt, int n, int E_> struct ZCfg {
  static constexpr int E = E_;
  static constexpr int P = n / E;
  static constexpr int LPB = z_lpb(n, P, (int)sizeof(cx<Rt>));
  static constexpr int THREADS = LPB * P;
};

__device__ __forceinline__ int bin_row(int p, int M, RowMap m) {
  if (p <= m.pos) return p;
  if (p >= M - m.neg) return m.base + (p - M);
  return -1;
}
template <typename C> __device__ __forceinline__ C czero() { return mk<C>(0, 0); }

// Occupancy target (minimum waves per SIMD, the second __launch_bounds__
// argument) from the number of complex values a thread keeps live: about
// 1.6 registers per live 32-bit word plus fixed overhead, out of 512.  It caps
// the register budget so that several blocks share a CU without spilling.
constexpr int occ_for(int live_cx, int esz) {
  const int regs = live_cx * (esz / 4) * 2 * 16 / 10 + 48;
  const int w = 512 / regs;
  return w < 1 ? 1 : (w > 4 ? 4 : w);
}
template <typename Rt> constexpr int esz_of() { return (int)sizeof(Rt); }
constexpr int stage_k(int n, int P) { return (n + 1 + P - 1) / P; }

// Column tile of this block.  Blocks are dispatched round-robin over the 8
// XCDs; consecutive tiles go to blocks of the same XCD so that neighbouring
// tiles, which share 128-byte lines of rows whose length is not a multiple
// of 16 elements, meet in the same L2.  Launch 8*ceil(ntiles/8) blocks;
// returns -1 for the padding blocks.
__device__ __forceinline__ int xcd_tile(int ntiles) {
  const int per = (ntiles + 7) >> 3;
  const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  return t < ntiles ? t : -1;
}
inline int xcd_grid(int ntiles) { return ((ntiles + 7) / 8) * 8; }

// Copy a twiddle table (cnt complex) into LDS; the caller synchronises.
template <typename C>
__device__ __forceinline__ void table_to_lds(C* dst, const C* __restrict__ src, int cnt) {
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
}

// ------------------------------------------------------------------ P1
// blockIdx.y selects the component: y = 0 -> w (fields 2 and 3 = X(kx w)),
// y = 1 -> v (fields 1 and 4 = X(kx v)), y = 2 -> u (field 0).  With a given
// vorticity (a.w) there are six blocks rows y = 0..5, each the plain
// x-inverse of one field: u, v, w from a.u, then wx, wy, wz from a.w.
template <typename Rt, int M, int E, int T>
__device__ __forceinline__ void p1_body(const P1Args<Rt>& a, int tile, int yb, cx<Rt>* lds, const cx<Rt>* tw) {
  using C = cx<Rt>;
  constexpr int P = M / E;
  using F = LineFFT<Rt, M, E, +1>;
  const int c = threadIdx.x % T, t = threadIdx.x / T;
  const int q = tile * T + c;  // compact column iyc * nkzp + kz
  const int iyc = q / a.nkzp, kzi = q - iyc * a.nkzp;
  const bool valid = q < a.Q && kzi < a.nkz;
  const int nky = a.y_neg + a.y_pos + 1;
  const int iy = iyc <= a.y_pos ? iyc : a.N1 - (nky - iyc);
  const bool given = a.w != nullptr;
  const int comp = given ? yb % 3 : 2 - yb;
  const long plane = (long)a.N1 * a.Nh;
  const C* __restrict__ U = (const C*)(given && yb >= 3 ? a.w : a.u) + (long)comp * a.N0 * plane +
                            (valid ? (long)iy * a.Nh + kzi : 0);
  C v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int r = bin_row(t + P * e, M, a.x);
    v[e] = (valid && r >= 0) ? U[r * plane] : czero<C>();
  }
  __syncthreads();
  C* __restrict__ out = (C*)a.w1 + q;
  const long orow = a.Q, ofield = (long)M * a.Q;
  C* line = lds + c;
  auto store = [&](int f, C(&x)[E]) {
    if (valid) {
#pragma unroll
      for (int e = 0; e < E; ++e) out[f * ofield + (long)(t + P * e) * orow] = x[e];
    }
  };
  if (given) {
    F::run(v, t, line, T, tw);
    store(yb, v);
    return;
  }
  if (comp > 0) {
    C x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int r = bin_row(t + P * e, M, a.x);
      x[e] = cscale(v[e], r >= 0 ? a.kx[r] : Rt(0));
    }
    F::run(x, t, line, T, tw);
    store(comp == 2 ? 3 : 4, x);
  }
  F::run(v, t, line, T, tw);
  store(comp, v);
}

// Waves per SIMD the x-inverse kernel is compiled for: the register estimate,
// raised to 3 for FP32 when three blocks fit in a CU's LDS (the 24-point
// schedule of the 384-point lines then fits 168 VGPRs with 16 bytes of spills
// per lane; at 2 waves per SIMD it waits on its stores).
template <typename Rt, int M, int E> constexpr int p1_occ() {
  constexpr int o = occ_for(2 * E, esz_of<Rt>());
  constexpr int lds = (M * SCfg<Rt, M, E>::T + M) * 2 * esz_of<Rt>();
  return sizeof(Rt) == 4 && o < 3 && 163840 / lds >= 3 && 3 * SCfg<Rt, M, E>::THREADS <= 4 * 64 * 3 ? 3 : o;
}
// D: the launcher's tile divisor for small grids (T / D columns per block).
template <typename Rt, int M, int E, int D = 1>
__global__ void __launch_bounds__((SCfg<Rt, M, E>::THREADS / D), (p1_occ<Rt, M, E>())) k_p1(P1Args<Rt> a) {
  using C = cx<Rt>;
  constexpr int T = SCfg<Rt, M, E>::T / D;
  __shared__ C lds[M * T];
  __shared__ C tw[M];
  const int tile = xcd_tile((a.Q + T - 1) / T);
  if (tile < 0) return;
  table_to_lds(tw, (const C*)a.tw, M);
  p1_body<Rt, M, E, T>(a, tile, blockIdx.y, lds, tw);
}

// ------------------------------------------------------------------ P2
// curl(u)^ = i (K x u_hat) (maths/cross.py:27-32) assembled after the
// x-inverse: wx = i(ky W - kz V), wy = i(kz U - X(kx W)), wz = i(X(kx V) - ky U).
// blockIdx.y = 0: V, W -> wx, V, W ; blockIdx.y = 1: U, X(kx W), X(kx V) -> wy, wz, U.
template <typename Rt, int M, int E, int T>
__device__ __forceinline__ void p2_body(const P2Args<Rt>& a, int tile, int yb, cx<Rt>* lds, const cx<Rt>* tw) {
  using C = cx<Rt>;
  constexpr bool STREAM = T * (int)sizeof(C) >= 128;  // tiles of whole 128-byte lines
  constexpr int P = M / E;
  using F = LineFFT<Rt, M, E, +1>;
  const int c = threadIdx.x % T, t = threadIdx.x / T;
  const int q = tile * T + c;  // column x * nkzp + kz
  const int x = q / a.nkzp, kzi = q - x * a.nkzp;
  const bool valid = q < a.Q && kzi < a.nkz;
  const int nky = a.y.base;
  const long ifield = (long)a.M0 * nky * a.nkzp, ofield = (long)a.M0 * M * a.nkzp;
  const C* __restrict__ in = (const C*)a.w1 + (valid ? (long)x * nky * a.nkzp + kzi : 0);
  C* __restrict__ out = (C*)a.w2 + (long)x * M * a.nkzp + kzi;
  const int rs = a.nkzp;
  const Rt kzv = a.kz[valid ? kzi : 0];
  C* line = lds + c;
  auto row = [&](int e) { return valid ? bin_row(t + P * e, M, a.y) : -1; };
  auto kyv = [&](int e) {
    const int p = t + P * e;
    const int k = p <= a.y.pos ? p : p - M;
    return row(e) >= 0 ? a.ky[k >= 0 ? k : a.N1 + k] : Rt(0);
  };
  auto load = [&](int f, C(&v)[E]) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int r = row(e);
      v[e] = r >= 0 ? ld_pass<STREAM>(in + f * ifield + (long)r * rs) : czero<C>();
    }
  };
  auto store = [&](int f, C(&v)[E]) {
    if (valid) {
#pragma unroll
      for (int e = 0; e < E; ++e) st_pass<STREAM>(out + f * ofield + (long)(t + P * e) * rs, v[e]);
    }
  };
  if (yb == 0) {
    C A[E], B[E], om[E];
    load(1, A);  // V
    load(2, B);  // W
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const Rt ky = kyv(e);
      om[e] = cmuli<1>(mk<C>(ky * B[e].x - kzv * A[e].x, ky * B[e].y - kzv * A[e].y));
    }
    F::run(om, t, line, T, tw);
    store(3, om);
    F::run(A, t, line, T, tw);
    store(1, A);
    F::run(B, t, line, T, tw);
    store(2, B);
  } else {
    C A[E], B[E];
    load(0, A);  // U
    load(3, B);  // X(kx W)
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e)
      B[e] = cmuli<1>(mk<C>(kzv * A[e].x - B[e].x, kzv * A[e].y - B[e].y));
    F::run(B, t, line, T, tw);
    store(4, B);
    load(4, B);  // X(kx V)
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const Rt ky = kyv(e);
      B[e] = cmuli<1>(mk<C>(B[e].x - ky * A[e].x, B[e].y - ky * A[e].y));
    }
    F::run(B, t, line, T, tw);
    store(5, B);
    F::run(A, t, line, T, tw);
    store(0, A);
  }
}

template <typename Rt, int M, int E, int D = 1>
__global__ void __launch_bounds__((SCfg<Rt, M, E>::THREADS / D), (occ_for(3 * E, esz_of<Rt>()))) k_p2(P2Args<Rt> a) {
  using C = cx<Rt>;
  constexpr int T = SCfg<Rt, M, E>::T / D;
  __shared__ C lds[M * T];
  __shared__ C tw[M];
  const int tile = xcd_tile((a.Q + T - 1) / T);
  if (tile < 0) return;
  table_to_lds(tw, (const C*)a.tw, M);
  p2_body<Rt, M, E, T>(a, tile, blockIdx.y, lds, tw);
}

// ------------------------------------------------------------------ z-line real transforms
// Length-M2 real line <-> half-length complex FFT (n = M2/2):
//  C2R: Z[p] = (X[p] + conj X[n-p]) + i e^{+2 pi i p/M2} (X[p] - conj X[n-p]),
//       z = IFFT_n(Z) = x[2m] + i x[2m+1]   (imaginary parts of X[0], X[n] ignored,
//       as numpy/FFTW c2r do)
//  R2C: Z = FFT_n(x[2m] + i x[2m+1]),
//       X[k] = (Z[k] + conj Z[n-k])/2 - (i/2) e^{-2 pi i k/M2} (Z[k] - conj Z[n-k]).

// Z[t + P e] from a spectral line staged in LDS (my[0..n], zero above nk).
template <typename Rt, int n, int E>
__device__ __forceinline__ void c2r_pre_lds(cx<Rt> (&z)[E], const cx<Rt>* my, const cx<Rt>* __restrict__ rtw,
                                            int t) {
  using C = cx<Rt>;
  constexpr int P = n / E;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + P * e;
    C xa = my[p];
    C xb = my[n - p];
    if (p == 0) { xa.y = 0; xb.y = 0; }
    const C ev = mk<C>(xa.x + xb.x, xa.y - xb.y);
    const C d = mk<C>(xa.x - xb.x, xa.y + xb.y);
    const C od = cmulc(d, rtw[p]);
    z[e] = mk<C>(ev.x - od.y, ev.y + od.x);
  }
}

// c2r pre-processing of a staged line: the thread's own bins X[p], p = t + P e,
// come from its staging registers r[e]; only the partners X[n - p] are read
// from the line's LDS area.  Bins at or above NKI are zero on every line (the
// dealiased input keeps fewer bins), so elements whose own bin or partner is
// at or above NKI for every thread are compile-time zeros.
template <typename Rt, int n, int E, int NKI, int K>
__device__ __forceinline__ void c2r_pre_reg(cx<Rt> (&z)[E], const cx<Rt> (&r)[K], const cx<Rt>* my,
                                            const cx<Rt>* __restrict__ rtw, int t) {
  using C = cx<Rt>;
  constexpr int P = n / E;
  static_assert(K >= E, "staging registers hold the thread's E bins");
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + P * e;
    C xa = (P * e >= NKI) ? czero<C>() : r[e];
    C xb = (n - (P * e + P - 1) >= NKI) ? czero<C>() : my[n - p];
    if (p == 0) { xa.y = 0; xb.y = 0; }
    const C ev = mk<C>(xa.x + xb.x, xa.y - xb.y);
    const C d = mk<C>(xa.x - xb.x, xa.y + xb.y);
    const C od = cmulc(d, rtw[p]);
    z[e] = mk<C>(ev.x - od.y, ev.y + od.x);
  }
}
// Number of staged bins per thread (t + P k, k < this) that c2r_pre_reg reads
// as partners: up to index n - P e_lo, e_lo the first element with a partner
// below NKI.
constexpr int stage_kept(int n, int P, int NKI, int K) {
  int elo = 0;
  while (n - (P * elo + P - 1) >= NKI) ++elo;
  const int k = (n - P * elo) / P + 1;
  return k < K ? k : K;
}

// Bin k of the real transform from the half-length FFT values A = Z[k] and
// Bz = Z[(n - k) mod n], w = rtw[k].
template <typename C> __device__ __forceinline__ C r2c_bin(C A, C Bz, C w) {
  using Rt = decltype(A.x);
  const C s = mk<C>(A.x + Bz.x, A.y - Bz.y);   // A + conj(B)
  const C d = mk<C>(A.x - Bz.x, A.y + Bz.y);   // A - conj(B)
  const C wd = cmul(w, d);
  return mk<C>(Rt(0.5) * (s.x + wd.y), Rt(0.5) * (s.y - wd.x));
}

// Post-process of the forward half-length FFT held in z (natural positions
// t + P e); `my` is this line's LDS area.  Writes kept bins k < nk (and the
// Nyquist bin n when nk > n).
template <typename Rt, int n, int E, bool WAVE = false>
__device__ __forceinline__ void r2c_post(cx<Rt> (&z)[E], cx<Rt>* my, cx<Rt>* __restrict__ out, int nk, bool valid,
                                         const cx<Rt>* __restrict__ rtw, int t, Rt scale) {
  using C = cx<Rt>;
  constexpr int P = n / E;
#pragma unroll
  for (int e = 0; e < E; ++e) my[t + P * e] = z[e];
  lds_barrier<WAVE>();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + P * e;
    const C X = r2c_bin(z[e], my[(n - k) % n], rtw[k]);
    if (valid && t < nk - P * e) st_stream(out + k, cscale(X, scale));
    if (valid && k == 0 && nk > n) st_stream(out + n, mk<C>((z[e].x - z[e].y) * scale, Rt(0)));
  }
  lds_barrier<WAVE>();
}

// r2c_post writing through a buffer descriptor: `base` is the byte offset of
// this line in the descriptor's span (lines outside the span are dropped by
// the range check); unit scale.
// NKO: output bins at or above NKO are never needed (compile-time bound of nk),
// so their elements are skipped and only the partners of the needed ones are
// written to LDS.
template <typename Rt, int n, int E, bool WAVE = false, int NKO = n + 1>
__device__ __forceinline__ void r2c_post_buf(cx<Rt> (&z)[E], cx<Rt>* my, __amdgpu_buffer_rsrc_t rs, int base, int nk,
                                             const cx<Rt>* __restrict__ rtw, int t) {
  using C = cx<Rt>;
  constexpr int P = n / E, ES = (int)sizeof(C);
  constexpr int EHI = (NKO + P - 1) / P < E ? (NKO + P - 1) / P : E;  // elements with a needed bin
  constexpr int PLO = n - P * EHI + 1;                                 // lowest partner index read (k > 0)
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (e == 0 || P * e + P - 1 >= PLO) my[t + P * e] = z[e];
  lds_barrier<WAVE>();
#pragma unroll
  for (int e = 0; e < EHI; ++e) {
    const int k = t + P * e;
    const C X = r2c_bin(z[e], my[(n - k) % n], rtw[k]);
    buf_store_nt<C>(X, rs, t < nk - P * e ? base + ES * k : BUF_OOB);
    if (k == 0 && nk > n) buf_store_nt<C>(mk<C>(z[e].x - z[e].y, Rt(0)), rs, base + ES * n);
  }
  lds_barrier<WAVE>();
}

// Staged load of one spectral line: the P threads of a line read bins
// t, t+P, ... (coalesced) into registers; stage_store puts them into LDS with
// zeros above nk (up to the Nyquist bin n).
template <typename Rt, int n, int P>
struct LineStage {
  static constexpr int K = (n + 1 + P - 1) / P;
  cx<Rt> r[K];
  __device__ __forceinline__ void load(const cx<Rt>* __restrict__ in, int nk, bool valid, int t) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      // (t < nk - P k rather than t + P k < nk: the bound is uniform, so the
      // compare needs no per-element index register)
      r[k] = (valid && t < nk - P * k) ? in[t + P * k] : czero<cx<Rt>>();
    }
  }
  // Through a buffer descriptor: `base` = byte offset of this line in the
  // descriptor's span; bins >= nk and lines outside the span read zero.
  __device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t rs, int base, int nk, int t) {
    constexpr int ES = (int)sizeof(cx<Rt>);
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = buf_load<cx<Rt>>(rs, t < nk - P * k ? base + ES * (t + P * k) : BUF_OOB);
  }
  __device__ __forceinline__ void store(cx<Rt>* my, int t) const { store_first<K>(my, t); }
  // the first KS staged bins only (the partners that c2r_pre_reg reads)
  template <int KS> __device__ __forceinline__ void store_first(cx<Rt>* my, int t) const {
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int i = t + P * k;
      if (i <= n) my[i] = r[k];
    }
  }
};

// ------------------------------------------------------------------ P3
template <typename Rt, int n, int E, int D = 1>
__global__ void __launch_bounds__((ZCfg<Rt, n, E>::THREADS / D), (occ_for(6 * E + stage_k(n, n / E), esz_of<Rt>()))) k_p3(P3Args<Rt> a) {
  using C = cx<Rt>;
  using Cfg = ZCfg<Rt, n, E>;
  constexpr int P = Cfg::P, LPB = Cfg::LPB / D;
  constexpr bool WL = 64 % P == 0;  // a line's threads share one wave
  using Fi = LineFFT<Rt, n, E, +1, WL>;
  using Ff = LineFFT<Rt, n, E, -1, WL>;
  __shared__ C lds[LPB * zslot(n, P)];
  __shared__ C tw[n], rtw[n];
  table_to_lds(tw, (const C*)a.tw, n);
  table_to_lds(rtw, (const C*)a.rtw, n);
  __syncthreads();
  const int t = threadIdx.x % P, l = threadIdx.x / P;
  const int line0 = blockIdx.x * LPB;
  C* my = lds + l * zslot(n, P);
  const int nl = min(LPB, a.L - line0);
  const long ifield = (long)a.L * a.nkzp;
  const C* inb = (const C*)a.w2 + (long)line0 * a.nkzp;
  const long ibytes = (long)nl * a.nkzp * (long)sizeof(C);
  const int ibase = l * a.nkzp * (int)sizeof(C);
  LineStage<Rt, n, P> st;
  st.load_buf(buf_rsrc(inb, ibytes), ibase, a.nkz, t);
  C ph[6][E];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    st.store(my, t);
    lds_barrier<WL>();
    c2r_pre_lds<Rt, n, E>(ph[f], my, rtw, t);
    lds_barrier<WL>();
    if (f < 5) st.load_buf(buf_rsrc(inb + (f + 1) * ifield, ibytes), ibase, a.nkz, t);  // next field in flight during the FFT
    Fi::run(ph[f], t, my, 1, tw);
  }
  // u x w on the packed points (real part: x[2m], imaginary part: x[2m+1]);
  // the products overwrite the velocity registers
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const C u0 = ph[0][e], u1 = ph[1][e], u2 = ph[2][e];
    const C w0 = ph[3][e], w1 = ph[4][e], w2 = ph[5][e];
    ph[0][e] = mk<C>(u1.x * w2.x - u2.x * w1.x, u1.y * w2.y - u2.y * w1.y);
    ph[1][e] = mk<C>(u2.x * w0.x - u0.x * w2.x, u2.y * w0.y - u0.y * w2.y);
    ph[2][e] = mk<C>(u0.x * w1.x - u1.x * w0.x, u0.y * w1.y - u1.y * w0.y);
  }
  const long ofield = (long)a.L * a.nkzop;
  C* outb = (C*)a.w3 + (long)line0 * a.nkzop;
  const long obytes = (long)nl * a.nkzop * (long)sizeof(C);
  const int obase = l * a.nkzop * (int)sizeof(C);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Ff::run(ph[i], t, my, 1, tw);
    r2c_post_buf<Rt, n, E, WL>(ph[i], my, buf_rsrc(outb + i * ofield, obytes), obase, a.nkzo, rtw, t);
  }
}

// P3, LDS-assisted scheme: the three velocity lines stay in registers, the
// first two vorticity lines are parked in LDS (one slot per line each) and
// only the last vorticity line is held in registers, so 4E complex values per
// thread are live instead of 6E.  Three LDS slots per line serve as staging,
// FFT exchange and parking space.
constexpr int z3_lpb(int n, int P, int esz) { return z3_lines(n, P, esz, z3_slot(n, P, esz)); }
template <typename Rt, int n, int E_> struct ZBCfg {
  static constexpr int E = E_;
  static constexpr int P = n / E;
  static constexpr int LPB = z3_lpb(n, P, (int)sizeof(cx<Rt>));
  static constexpr int THREADS = LPB * P;
};

// NKI / NKO: compile-time bounds of the input bins held (a.nkz <= NKI) and of
// the output bins needed (a.nkzo <= NKO); n + 1 for the generic pass.
template <typename Rt, int n, int E, int LPB, int NKI = n + 1, int NKO = n + 1>
__device__ __forceinline__ void p3b_body(const P3Args<Rt>& a, int line0, cx<Rt>* lds, const cx<Rt>* tw,
                                         const cx<Rt>* rtw) {
  using C = cx<Rt>;
  constexpr int P = n / E, SL = z3_slot(n, P, (int)sizeof(cx<Rt>));
  constexpr int KS = stage_kept(n, P, NKI, LineStage<Rt, n, P>::K);
  constexpr bool WL = 64 % P == 0;  // a line's threads share one wave
  // the bank-conflict-free first exchange of the 192-point lines (12 points per thread)
  constexpr bool SWZ = n == 192 && E == 12;
  using Fi = LineFFT<Rt, n, E, +1, WL, SWZ>;
  using Ff = LineFFT<Rt, n, E, -1, WL, SWZ>;
  __syncthreads();  // twiddle tables written by the caller
  const int t = threadIdx.x % P, l = threadIdx.x / P;
  C* s0 = lds + (3 * l) * SL;
  C* s1 = s0 + SL;
  C* s2 = s1 + SL;
  // buffer descriptors over this block's lines (lines past a.L read zero and
  // are not written)
  const int nl = min(LPB, a.L - line0);
  const long ifield = (long)a.L * a.nkzp;
  const C* inb = (const C*)a.w2 + (long)line0 * a.nkzp;
  const long ibytes = (long)nl * a.nkzp * (long)sizeof(C);
  const int ibase = l * a.nkzp * (int)sizeof(C);
  LineStage<Rt, n, P> st;
  st.load_buf(buf_rsrc(inb, ibytes), ibase, a.nkz, t);
  C U[3][E], W[E];
  // velocity lines -> registers (staging and exchange in slot 0)
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    st.template store_first<KS>(s0, t);
    lds_barrier<WL>();
    c2r_pre_reg<Rt, n, E, NKI>(U[f], st.r, s0, rtw, t);
    lds_barrier<WL>();
    st.load_buf(buf_rsrc(inb + (f + 1) * ifield, ibytes), ibase, a.nkz, t);
    Fi::run(U[f], t, s0, 1, tw);
  }
  // w0 -> slot 1, w1 -> slot 2 (each slot is its own staging and exchange area)
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    C* sl = f == 0 ? s1 : s2;
    st.template store_first<KS>(sl, t);
    lds_barrier<WL>();
    c2r_pre_reg<Rt, n, E, NKI>(W, st.r, sl, rtw, t);
    lds_barrier<WL>();
    st.load_buf(buf_rsrc(inb + (4 + f) * ifield, ibytes), ibase, a.nkz, t);
    Fi::run(W, t, sl, 1, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) sl[t + P * e] = W[e];
  }
  // w2 -> registers (slot 0)
  st.template store_first<KS>(s0, t);
  lds_barrier<WL>();
  c2r_pre_reg<Rt, n, E, NKI>(W, st.r, s0, rtw, t);
  lds_barrier<WL>();
  Fi::run(W, t, s0, 1, tw);
  lds_barrier<WL>();
  // u x w on the packed points; the products overwrite the velocity registers
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + P * e;
    const C u0 = U[0][e], u1 = U[1][e], u2 = U[2][e];
    const C w0 = s1[p], w1 = s2[p], w2 = W[e];
    U[0][e] = mk<C>(u1.x * w2.x - u2.x * w1.x, u1.y * w2.y - u2.y * w1.y);
    U[1][e] = mk<C>(u2.x * w0.x - u0.x * w2.x, u2.y * w0.y - u0.y * w2.y);
    U[2][e] = mk<C>(u0.x * w1.x - u1.x * w0.x, u0.y * w1.y - u1.y * w0.y);
  }
  lds_barrier<WL>();
  const long ofield = (long)a.L * a.nkzop;
  C* outb = (C*)a.w3 + (long)line0 * a.nkzop;
  const long obytes = (long)nl * a.nkzop * (long)sizeof(C);
  const int obase = l * a.nkzop * (int)sizeof(C);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    C* sl = i == 0 ? s0 : (i == 1 ? s1 : s2);
    Ff::run(U[i], t, sl, 1, tw);
    r2c_post_buf<Rt, n, E, WL, NKO>(U[i], sl, buf_rsrc(outb + i * ofield, obytes), obase, a.nkzo, rtw, t);
  }
}

template <typename Rt, int n, int E, int NKI = n + 1, int NKO = n + 1>
__global__ void __launch_bounds__((ZBCfg<Rt, n, E>::THREADS), (occ_for(4 * E + stage_k(n, n / E), esz_of<Rt>()))) k_p3b(P3Args<Rt> a) {
  using C = cx<Rt>;
  constexpr int LPB = ZBCfg<Rt, n, E>::LPB;
  __shared__ C lds[3 * LPB * z3_slot(n, n / E, (int)sizeof(cx<Rt>))];
  // FP64: the real-transform twiddles are read from global memory (a 3 KB
  // table that stays in L1) so that one more block fits in a CU's LDS
  constexpr bool RTW_LDS = z3_rtw_in_lds((int)sizeof(cx<Rt>));
  __shared__ C tw[n], rtw_s[RTW_LDS ? n : 1];
  table_to_lds(tw, (const C*)a.tw, n);
  if constexpr (RTW_LDS) table_to_lds(rtw_s, (const C*)a.rtw, n);
  const C* rtw = RTW_LDS ? rtw_s : (const C*)a.rtw;
  p3b_body<Rt, n, E, LPB, NKI, NKO>(a, blockIdx.x * LPB, lds, tw, rtw);
}

// ------------------------------------------------------------------ z pass, LDS-DMA scheme
// Scheme C: at the start of a block the six spectral z-line rows of all its
// LPB lines are copied HBM -> LDS by global_load_lds_dwordx4 (no registers
// involved, so every row of the block is in flight at once; a block's lines
// are consecutive rows, so each field is one contiguous chunk).  Then, per
// line, the six C2R transforms run from the staged rows, w0 and w1 are parked
// in LDS over the consumed velocity rows, the cross product is formed in
// registers and the three R2C transforms go back to HBM.  Registers hold 4E
// complex values and no prefetch.  LDS (dynamic, bytes from zc_lds_bytes):
//   [6][LPB][nkzp] staging | [LPB][zslot(n, P)] exchange | w0/w1 slots if they do
//   not fit over the staged u rows | tw[n] | rtw[n].

// Copy `bytes` (a multiple of 16) from global src to LDS dst: 16 bytes per
// lane, a wave fills 1 KB of LDS per instruction.  blockDim.x % 64 == 0.
template <typename C>
__device__ __forceinline__ void lds_dma_copy(C* dst, const C* src, int bytes) {
  typedef __attribute__((address_space(3))) void lds_void;
  typedef __attribute__((address_space(1))) void glb_void;
  const int units = bytes >> 4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int u0 = wave * 64; u0 < units; u0 += nw * 64) {
    const int u = u0 + lane;
    if (u < units)
      __builtin_amdgcn_global_load_lds((glb_void*)((const char*)src + (long)u * 16),
                                       (lds_void*)((char*)dst + (long)u0 * 16), 16, 0, 0);
  }
}

// Z[t + P e] of a staged spectral row my[0..nk) (entries at and above nk are zero).
template <typename Rt, int n, int E>
__device__ __forceinline__ void c2r_pre_row(cx<Rt> (&z)[E], const cx<Rt>* my, int nk, const cx<Rt>* rtw, int t) {
  using C = cx<Rt>;
  constexpr int P = n / E;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + P * e;
    C xa = p < nk ? my[p] : czero<C>();
    C xb = n - p < nk ? my[n - p] : czero<C>();
    if (p == 0) { xa.y = 0; xb.y = 0; }
    const C ev = mk<C>(xa.x + xb.x, xa.y - xb.y);
    const C d = mk<C>(xa.x - xb.x, xa.y + xb.y);
    const C od = cmulc(d, rtw[p]);
    z[e] = mk<C>(ev.x - od.y, ev.y + od.x);
  }
}

struct ZcLayout {
  int stage, xch, om, tw, rtw, total;  // offsets / size in complex elements
};
__host__ __device__ inline ZcLayout zc_layout(int n, int P, int lpb, int nkzp) {
  ZcLayout z;
  z.stage = 0;
  z.xch = 6 * lpb * nkzp;
  const int slot = zslot(n, P);
  const bool fits = 2 * slot <= 3 * nkzp;  // w0, w1 over the staged u rows of the same block
  z.om = fits ? 0 : z.xch + lpb * slot;
  z.tw = z.xch + lpb * slot + (fits ? 0 : 2 * lpb * slot);
  z.rtw = z.tw + n;
  z.total = z.rtw + n;
  return z;
}

// Occupancy target 2: a block's staged rows take up to 64 KB of LDS, so at most
// two blocks share a CU whatever the register budget.
template <typename Rt, int n, int E, int LPB>
__global__ void __launch_bounds__(LPB * (n / E), 2) k_p3c(P3Args<Rt> a) {
  using C = cx<Rt>;
  constexpr int P = n / E, SL = zslot(n, P);
  static_assert((LPB * P) % 64 == 0, "whole waves");
  constexpr bool WL = 64 % P == 0;
  using Fi = LineFFT<Rt, n, E, +1, WL>;
  using Ff = LineFFT<Rt, n, E, -1, WL>;
  extern __shared__ __align__(16) unsigned char zc_smem[];
  C* sm = (C*)zc_smem;
  const ZcLayout Z = zc_layout(n, P, LPB, a.nkzp);
  C* st = sm + Z.stage;
  C* tw = sm + Z.tw;
  C* rtw = sm + Z.rtw;
  const int line0 = blockIdx.x * LPB;
  const int nl = min(LPB, a.L - line0);
  const long ifield = (long)a.L * a.nkzp;
  const int chunk = nl * a.nkzp;  // complex elements per field
  const C* __restrict__ in = (const C*)a.w2 + (long)line0 * a.nkzp;
#pragma unroll
  for (int f = 0; f < 6; ++f) lds_dma_copy(st + f * LPB * a.nkzp, in + f * ifield, chunk * (int)sizeof(C));
  table_to_lds(tw, (const C*)a.tw, n);
  table_to_lds(rtw, (const C*)a.rtw, n);
  __asm__ __volatile__("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  const int t = threadIdx.x % P, l = threadIdx.x / P;
  const int line = line0 + l;
  const bool valid = line < a.L;
  const long li = valid ? line : 0;
  C* xl = sm + Z.xch + l * SL;
  C* o0 = sm + Z.om + (2 * l) * SL;
  C* o1 = o0 + SL;
  auto row = [&](int f) { return (const C*)(st + (f * LPB + l) * a.nkzp); };
  C U[3][E], W[E];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    c2r_pre_row<Rt, n, E>(U[f], row(f), a.nkz, rtw, t);
    Fi::run(U[f], t, xl, 1, tw);
  }
  __syncthreads();  // every line's u rows are consumed before w0/w1 may overwrite them
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    c2r_pre_row<Rt, n, E>(W, row(3 + f), a.nkz, rtw, t);
    Fi::run(W, t, xl, 1, tw);
    C* om = f == 0 ? o0 : o1;
#pragma unroll
    for (int e = 0; e < E; ++e) om[t + P * e] = W[e];
  }
  c2r_pre_row<Rt, n, E>(W, row(5), a.nkz, rtw, t);
  Fi::run(W, t, xl, 1, tw);
  lds_barrier<WL>();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + P * e;
    const C u0 = U[0][e], u1 = U[1][e], u2 = U[2][e];
    const C w0 = o0[p], w1 = o1[p], w2 = W[e];
    U[0][e] = mk<C>(u1.x * w2.x - u2.x * w1.x, u1.y * w2.y - u2.y * w1.y);
    U[1][e] = mk<C>(u2.x * w0.x - u0.x * w2.x, u2.y * w0.y - u0.y * w2.y);
    U[2][e] = mk<C>(u0.x * w1.x - u1.x * w0.x, u0.y * w1.y - u1.y * w0.y);
  }
  lds_barrier<WL>();
  C* __restrict__ out = (C*)a.w3 + li * a.nkzop;
  const long ofield = (long)a.L * a.nkzop;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Ff::run(U[i], t, xl, 1, tw);
    r2c_post<Rt, n, E, WL>(U[i], xl, out + i * ofield, a.nkzo, valid, rtw, t, Rt(1));
  }
}

// ------------------------------------------------------------------ plain z transforms
template <typename Rt, int n, int E, int D = 1>
__global__ void __launch_bounds__((ZCfg<Rt, n, E>::THREADS / D), (occ_for(E + stage_k(n, n / E), esz_of<Rt>()))) k_z_c2r(ZArgs<Rt> a) {
  using C = cx<Rt>;
  using Cfg = ZCfg<Rt, n, E>;
  constexpr int P = Cfg::P, LPB = Cfg::LPB / D;
  constexpr bool WL = 64 % P == 0;
  using Fi = LineFFT<Rt, n, E, +1, WL>;
  __shared__ C lds[LPB * zslot(n, P)];
  __shared__ C tw[n], rtw[n];
  table_to_lds(tw, (const C*)a.tw, n);
  table_to_lds(rtw, (const C*)a.rtw, n);
  __syncthreads();
  const int t = threadIdx.x % P, l = threadIdx.x / P;
  const int line = blockIdx.x * LPB + l;
  const bool valid = line < a.L;
  const long li = valid ? line : 0;
  C* my = lds + l * zslot(n, P);
  const C* in = (const C*)a.in + blockIdx.y * a.in_field + li * a.in_line;
  LineStage<Rt, n, P> st;
  st.load(in, a.nk, valid, t);
  st.store(my, t);
  lds_barrier<WL>();
  C z[E];
  c2r_pre_lds<Rt, n, E>(z, my, rtw, t);
  lds_barrier<WL>();
  Fi::run(z, t, my, 1, tw);
  if (valid) {
    C* out = (C*)a.out + blockIdx.y * a.out_field + li * a.out_line;  // real line viewed as n complex pairs
#pragma unroll
    for (int e = 0; e < E; ++e) st_stream(out + t + P * e, cscale(z[e], a.scale));
  }
}

template <typename Rt, int n, int E, int D = 1>
__global__ void __launch_bounds__((ZCfg<Rt, n, E>::THREADS / D), (occ_for(E + stage_k(n, n / E), esz_of<Rt>()))) k_z_r2c(ZArgs<Rt> a) {
  using C = cx<Rt>;
  using Cfg = ZCfg<Rt, n, E>;
  constexpr int P = Cfg::P, LPB = Cfg::LPB / D;
  constexpr bool WL = 64 % P == 0;
  using Ff = LineFFT<Rt, n, E, -1, WL>;
  __shared__ C lds[LPB * zslot(n, P)];
  __shared__ C tw[n], rtw[n];
  table_to_lds(tw, (const C*)a.tw, n);
  table_to_lds(rtw, (const C*)a.rtw, n);
  __syncthreads();
  const int t = threadIdx.x % P, l = threadIdx.x / P;
  const int line = blockIdx.x * LPB + l;
  const bool valid = line < a.L;
  const long li = valid ? line : 0;
  C* my = lds + l * zslot(n, P);
  const C* in = (const C*)a.in + blockIdx.y * a.in_field + li * a.in_line;
  C z[E];
#pragma unroll
  for (int e = 0; e < E; ++e) z[e] = valid ? in[t + P * e] : czero<C>();
  lds_barrier<WL>();
  Ff::run(z, t, my, 1, tw);
  C* out = (C*)a.out + blockIdx.y * a.out_field + li * a.out_line;
  r2c_post<Rt, n, E, WL>(z, my, out, a.nk, valid, rtw, t, a.scale);
}

// ------------------------------------------------------------------ generic strided
template <typename Rt, int M, int E, int DIR, int T>
__device__ __forceinline__ void strided_body(const StridedArgs<Rt>& a, int tile, int yb, cx<Rt>* lds,
                                             const cx<Rt>* tw) {
  using C = cx<Rt>;
  constexpr int P = M / E;
  using F = LineFFT<Rt, M, E, DIR>;
  const int c = threadIdx.x % T, t = threadIdx.x / T;
  const int q = tile * T + c;
  const int hi = q / a.cdiv, lo = q - hi * a.cdiv;
  const bool valid = q < a.Q && lo < a.cvalid;
  const C* __restrict__ in = (const C*)a.in + yb * a.in_field + (valid ? hi * a.in_hi + lo : 0);
  C* __restrict__ out = (C*)a.out + yb * a.out_field + hi * a.out_hi + lo;
  C v[E];
  auto load = [&](auto stream) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int r = bin_row(t + P * e, M, a.rin);
      v[e] = (valid && r >= 0) ? ld_pass<decltype(stream)::value>(in + (long)r * a.in_row) : czero<C>();
    }
  };
  if (a.stream_in) load(std::true_type{});
  else load(std::false_type{});
  __syncthreads();
  F::run(v, t, lds + c, T, tw);
  auto store = [&](auto stream) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int r = bin_row(t + P * e, M, a.rout);
      if (r >= 0) st_pass<decltype(stream)::value>(out + (long)r * a.out_row, cscale(v[e], a.scale));
    }
  };
  if (valid) {
    if (a.stream_out) store(std::true_type{});
    else store(std::false_type{});
  }
}

template <typename Rt, int M, int E, int DIR, int D = 1>
__global__ void __launch_bounds__((SCfg<Rt, M, E>::THREADS / D), (occ_for(E, esz_of<Rt>()))) k_strided(StridedArgs<Rt> a) {
  using C = cx<Rt>;
  constexpr int T = SCfg<Rt, M, E>::T / D;
  __shared__ C lds[M * T];
  __shared__ C tw[M];
  const int tile = xcd_tile((a.Q + T - 1) / T);
  if (tile < 0) return;
  table_to_lds(tw, (const C*)a.tw, M);
  strided_body<Rt, M, E, DIR, T>(a, tile, blockIdx.y, lds, tw);
}

// ------------------------------------------------------------------ P5
// Knuth's TwoSum: returns s = fl(a + b) and sets err so that s + err = a + b
// exactly (no contraction inside).
template <typename Rt> __device__ __forceinline__ Rt two_sum(Rt a, Rt b, Rt& err) {
#pragma clang fp contract(off)
  const Rt s = a + b;
  const Rt bp = s - a;
  err = (a - (s - bp)) + (b - bp);
  return s;
}

// w |u|^2 and w |K x u|^2 of one mode (components x0, y0, x1, y1, x2, y2) in FP64.
template <typename Rt>
__device__ __forceinline__ void energy_terms(const double (&v)[6], Rt K0, Rt K1, Rt K2, double w, double& esum,
                                             double& osum) {
  esum += w * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
  const double k0 = K0, k1 = K1, k2 = K2;
  const double ax = k1 * v[4] - k2 * v[2], ay = k1 * v[5] - k2 * v[3];
  const double bx = k2 * v[0] - k0 * v[4], by = k2 * v[1] - k0 * v[5];
  const double cx = k0 * v[2] - k1 * v[0], cy = k0 * v[3] - k1 * v[1];
  osum += w * (ax * ax + ay * ay + bx * bx + by * by + cx * cx + cy * cy);
}

// Block sum of the per-thread FP64 (energy, enstrophy) sums of a P5 tile into
// partials[2 tile], partials[2 tile + 1]; every thread of the block calls it.
template <int TH> __device__ __forceinline__ void p5_partials(double esum, double osum, double* partials, int tile) {
  __shared__ double red[2][TH / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    esum += __shfl_down(esum, off, 64);
    osum += __shfl_down(osum, off, 64);
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) { red[0][wave] = esum; red[1][wave] = osum; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0, s1 = 0;
    for (int w = 0; w < TH / 64; ++w) { s0 += red[0][w]; s1 += red[1][w]; }
    partials[2 * tile] = s0;
    partials[2 * tile + 1] = s1;
  }
}

// COMP: compensated accumulation of the RK4 sum u2 = u1 + sum a dt rhs in a
// (hi, lo) pair of single-precision fields; the state is u = hi + lo with hi
// the caller's u_hat and lo the caller's u_lo.  Stage values stay plain.
template <typename Rt, int M, int E, int MODE, bool COMP, int T>
__device__ __forceinline__ void p5_body(const P5Args<Rt>& a, int tile, cx<Rt>* lds, const cx<Rt>* tw) {
  using C = cx<Rt>;
  constexpr bool STREAM = T * (int)sizeof(C) >= 128;  // tiles of whole 128-byte lines
  constexpr int P = M / E, TH = T * P;
  using F = LineFFT<Rt, M, E, -1>;
  const int c = threadIdx.x % T, t = threadIdx.x / T;
  const int q = tile * T + c;
  const bool valid = q < a.Q;
  const int qq = valid ? q : 0;
  const int iy = qq / a.Nh, kz = qq - iy * a.Nh;
  const int ky = iy < a.N1 / 2 ? iy : iy - a.N1;
  const int nkyo = a.y_neg + a.y_pos + 1;
  const bool colconv = valid && ky >= -a.y_neg && ky <= a.y_pos && kz < a.nkzo;
  const int kyc = ky >= 0 ? ky : nkyo + ky;
  const long wrow = (long)nkyo * a.nkzop, wfield = (long)M * wrow;
  const C* __restrict__ w4 = (const C*)a.w4 + (colconv ? (long)kyc * a.nkzop + kz : 0);
  C X0[E], X1[E], X2[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const long o = (long)(t + P * e) * wrow;
    X0[e] = colconv ? w4[o] : czero<C>();
    X1[e] = colconv ? w4[wfield + o] : czero<C>();
    X2[e] = colconv ? w4[2 * wfield + o] : czero<C>();
  }
  __syncthreads();
  C* line = lds + c;
  F::run(X0, t, line, T, tw);
  F::run(X1, t, line, T, tw);
  F::run(X2, t, line, T, tw);

  const long plane = (long)a.N1 * a.Nh, fs = (long)a.N0 * plane;
  const Rt K1 = a.ky[iy], K2c = a.kz[kz];
  const bool nyq_col = a.mask_nyquist && (iy == a.N1 / 2 || kz == a.N2 / 2);
  double esum = 0.0, osum = 0.0;
  // Elements go in chunks of G: all loads of a chunk are issued before any of
  // its stores, so G elements' state is in flight at once instead of one
  // round trip per array and element.  An element's rows are written only by
  // that element, so no load of a chunk can see a store of the same chunk.
  constexpr int G = (E >= 24 || TH >= 1024 || MODE == P5_RHS) ? 1 : 2;
  static_assert(E % G == 0, "chunk must divide the points per thread");
  const C* ust = (const C*)(MODE == P5_RK0 || MODE == P5_RHS ? a.A : a.B);
  const C* src = (const C*)a.source;
  const C* Ain = (const C*)a.A;
  const C* Cin = (const C*)a.Cacc;
  const C* Lin = (const C*)(MODE == P5_RK0 ? a.Alo : a.Clo);
  constexpr bool LOAD_A = MODE == P5_RK12;
  constexpr bool LOAD_C = MODE == P5_RK12 || MODE == P5_RK3;
  constexpr bool LOAD_L = COMP && MODE != P5_RHS;
#pragma unroll
  for (int e0 = 0; e0 < E; e0 += G) {
    int ixs[G];
    bool keeps[G];
    C su[G][3], sr[G][3], xa[G][3], xc[G][3], xl[G][3];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int p = t + P * (e0 + g);
      int ix;
      bool kept;
      if (M == a.N0) {
        ix = p;
        const int k = p < a.N0 / 2 ? p : p - a.N0;
        kept = k >= -a.x.neg && k <= a.x.pos;
      } else {
        ix = bin_row(p, M, a.x);
        kept = true;
      }
      ixs[g] = valid ? ix : -1;
      keeps[g] = kept;
      if (ixs[g] < 0) continue;
      if constexpr (MODE == P5_RHS) {
        if (a.conv_only) continue;
      }
      const long off = (long)ix * plane + q;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        su[g][i] = ust[i * fs + off];
        if (src) sr[g][i] = src[i * fs + off];
        if constexpr (LOAD_A) xa[g][i] = Ain[i * fs + off];
        if constexpr (LOAD_C) xc[g][i] = Cin[i * fs + off];
        if constexpr (LOAD_L) xl[g][i] = Lin[i * fs + off];
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int e = e0 + g;
      const int ix = ixs[g];
      const bool kept = keeps[g];
      if (ix < 0) continue;
      const long off = (long)ix * plane + q;
      if constexpr (MODE == P5_RHS) {
        if (a.conv_only) {
          const bool cv = colconv && kept;
          C* rhs = (C*)a.rhs;
          st_pass<STREAM>(rhs + off, cv ? cscale(X0[e], a.scale) : czero<C>());
          st_pass<STREAM>(rhs + fs + off, cv ? cscale(X1[e], a.scale) : czero<C>());
          st_pass<STREAM>(rhs + 2 * fs + off, cv ? cscale(X2[e], a.scale) : czero<C>());
          continue;
        }
      }
      const bool conv = colconv && kept && !(nyq_col || (a.mask_nyquist && ix == a.N0 / 2));
      C r0 = conv ? cscale(X0[e], a.scale) : czero<C>();
      C r1 = conv ? cscale(X1[e], a.scale) : czero<C>();
      C r2 = conv ? cscale(X2[e], a.scale) : czero<C>();
      const Rt K0 = a.kx[ix];
      Rt k2 = K0 * K0;
      k2 += K1 * K1;
      k2 += K2c * K2c;
      const Rt den = k2 == Rt(0) ? Rt(1) : k2;
      const Rt o0 = K0 / den, o1 = K1 / den, o2 = K2c / den;
      // P_hat = sum_i rhs_i K_i/K^2 ; rhs_i -= P_hat K_i ; rhs -= (nu K^2) u_hat  (NS.py:208-215)
      C ph = mk<C>(r0.x * o0 + r1.x * o1, r0.y * o0 + r1.y * o1);
      ph = mk<C>(ph.x + r2.x * o2, ph.y + r2.y * o2);
      r0 = mk<C>(r0.x - ph.x * K0, r0.y - ph.y * K0);
      r1 = mk<C>(r1.x - ph.x * K1, r1.y - ph.y * K1);
      r2 = mk<C>(r2.x - ph.x * K2c, r2.y - ph.y * K2c);
      const Rt z = a.nu * k2;
      const C s0 = su[g][0], s1 = su[g][1], s2 = su[g][2];
      r0 = mk<C>(r0.x - z * s0.x, r0.y - z * s0.y);
      r1 = mk<C>(r1.x - z * s1.x, r1.y - z * s1.y);
      r2 = mk<C>(r2.x - z * s2.x, r2.y - z * s2.y);
      if (src) {  // rhs += Source (NS.py:256-258), every mode
        r0 = cadd(r0, sr[g][0]); r1 = cadd(r1, sr[g][1]); r2 = cadd(r2, sr[g][2]);
      }
      if constexpr (MODE == P5_RHS) {
        C* rhs = (C*)a.rhs;
        st_pass<STREAM>(rhs + off, r0); st_pass<STREAM>(rhs + fs + off, r1); st_pass<STREAM>(rhs + 2 * fs + off, r2);
        if (a.phat) ((C*)a.phat)[off] = ph;
      } else {
        C* A = (C*)a.A;
        C* B = (C*)a.B;
        C* Cc = (C*)a.Cacc;
        const Rt bdt = a.bdt, adt = a.adt;
        auto upd = [](C base, C r, Rt s) { return mk<C>(base.x + s * r.x, base.y + s * r.y); };
        if constexpr (!COMP) {
          if constexpr (MODE == P5_RK0) {
            // u0 = u1 + b dt rhs ; u2 = u1 + a dt rhs  (u1 = u2 = u0 at stage 0)
            st_pass<STREAM>(B + (off), upd(s0, r0, bdt)); st_pass<STREAM>(B + (fs + off), upd(s1, r1, bdt)); st_pass<STREAM>(B + (2 * fs + off), upd(s2, r2, bdt));
            st_pass<STREAM>(Cc + (off), upd(s0, r0, adt)); st_pass<STREAM>(Cc + (fs + off), upd(s1, r1, adt)); st_pass<STREAM>(Cc + (2 * fs + off), upd(s2, r2, adt));
          } else if constexpr (MODE == P5_RK12) {
            st_pass<STREAM>(B + (off), upd(xa[g][0], r0, bdt)); st_pass<STREAM>(B + (fs + off), upd(xa[g][1], r1, bdt)); st_pass<STREAM>(B + (2 * fs + off), upd(xa[g][2], r2, bdt));
            st_pass<STREAM>(Cc + (off), upd(xc[g][0], r0, adt)); st_pass<STREAM>(Cc + (fs + off), upd(xc[g][1], r1, adt));
            st_pass<STREAM>(Cc + (2 * fs + off), upd(xc[g][2], r2, adt));
          } else {
            const C n0 = upd(xc[g][0], r0, adt), n1 = upd(xc[g][1], r1, adt), n2 = upd(xc[g][2], r2, adt);
            st_pass<STREAM>(A + (off), n0); st_pass<STREAM>(A + (fs + off), n1); st_pass<STREAM>(A + (2 * fs + off), n2);
            if (a.phat) ((C*)a.phat)[off] = ph;
            if (a.partials) {
              const double w = (kz == 0 || kz == a.N2 / 2) ? 1.0 : 2.0;
              const double v[6] = {n0.x, n0.y, n1.x, n1.y, n2.x, n2.y};
              energy_terms(v, K0, K1, K2c, w, esum, osum);
            }
          }
        } else {
          C* Alo = (C*)a.Alo;
          C* Clo = (C*)a.Clo;
          const C rr[3] = {r0, r1, r2};
          if constexpr (MODE == P5_RK0) {
            const C ss[3] = {s0, s1, s2};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const long o = i * fs + off;
              st_pass<STREAM>(B + (o), upd(ss[i], rr[i], bdt));
              const Rt yx = adt * rr[i].x, yy = adt * rr[i].y;
              Rt ex, ey;
              const Rt hx = two_sum(ss[i].x, yx, ex), hy = two_sum(ss[i].y, yy, ey);
              const C l = xl[g][i];
              st_pass<STREAM>(Cc + (o), mk<C>(hx, hy));
              Clo[o] = mk<C>(l.x + ex, l.y + ey);
            }
          } else if constexpr (MODE == P5_RK12) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const long o = i * fs + off;
              st_pass<STREAM>(B + (o), upd(xa[g][i], rr[i], bdt));
              const C c = xc[g][i], l = xl[g][i];
              const Rt yx = adt * rr[i].x, yy = adt * rr[i].y;
              Rt ex, ey;
              const Rt hx = two_sum(c.x, yx, ex), hy = two_sum(c.y, yy, ey);
              st_pass<STREAM>(Cc + (o), mk<C>(hx, hy));
              Clo[o] = mk<C>(l.x + ex, l.y + ey);
            }
          } else {
            double v[6];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const long o = i * fs + off;
              const C c = xc[g][i], l = xl[g][i];
              const Rt yx = adt * rr[i].x, yy = adt * rr[i].y;
              Rt ex, ey;
              const Rt hx = two_sum(c.x, yx, ex), hy = two_sum(c.y, yy, ey);
              const Rt lx = l.x + ex, ly = l.y + ey;
              Rt rx, ry;  // renormalise: hi = fl(h + l), lo = exact remainder
              const Rt nx = two_sum(hx, lx, rx), ny = two_sum(hy, ly, ry);
              st_pass<STREAM>(A + (o), mk<C>(nx, ny));
              Alo[o] = mk<C>(rx, ry);
              v[2 * i] = (double)nx + (double)rx;
              v[2 * i + 1] = (double)ny + (double)ry;
            }
            if (a.phat) ((C*)a.phat)[off] = ph;
            if (a.partials) energy_terms(v, K0, K1, K2c, (kz == 0 || kz == a.N2 / 2) ? 1.0 : 2.0, esum, osum);
          }
        }
      }
    }
  }
  if constexpr (MODE == P5_RK3) {
    if (a.partials) p5_partials<TH>(esum, osum, a.partials, tile);
  }
}

// P5 stages under the 3/2 rule (M = 3 N0 / 2 bins, kept output modes
// -N0/2..N0/2-1).  The kept bins are p < M/3 (row p) and p >= 2M/3 (row
// p - M/3); with p = t + P e that is e < E/3 and e >= 2E/3 on every lane, and
// kept element j (e = j, or j + E/3 past the first third) lands in row
// t + P j.  Known at compile time, these leave the update loop without
// branches, and it is software-pipelined: the state rows of the next chunk of
// G elements are loaded before the current chunk is computed and stored (an
// element's rows are written only by that element, so in program order every
// load precedes the stores it could meet).  Every access is a buffer access:
// one descriptor per array, the lane's byte offset (the same for every
// element) and a uniform offset of component and element in a scalar
// register, so no access needs a 64-bit address of its own.  W4 entries of
// columns that carry no kept mode get an offset past the descriptor and read
// as zero.  Plain RK stages with no source and no pressure output; the
// per-mode arithmetic is p5_body's.  The host checks that the three
// components of a state array and of W4 span at most BUF_OOB bytes
// (p5k_fits).
template <typename Rt, int M, int E, int MODE, int T>
__device__ __forceinline__ void p5k_body(const P5Args<Rt>& a, int tile, cx<Rt>* lds, const cx<Rt>* tw) {
  using C = cx<Rt>;
  static_assert(M % 3 == 0 && E % 3 == 0 && MODE != P5_RHS, "3/2-rule RK stages");
  constexpr int P = M / E, TH = T * P, E3 = E / 3, NK = 2 * E3;
  constexpr int CS = (int)sizeof(C);
  // state arrays read per element: RK0 u1; RK1-2 u0, u1, u2; RK3 u0, u2
  constexpr int NL = MODE == P5_RK12 ? 3 : (MODE == P5_RK3 ? 2 : 1);
  constexpr int G = MODE == P5_RK12 ? 1 : 2;  // spill-free at E = 24 (256 VGPRs)
  static_assert(NK % G == 0, "chunk must divide the kept elements");
  using F = LineFFT<Rt, M, E, -1>;
  const int c = threadIdx.x % T, t = threadIdx.x / T;
  const int q = tile * T + c;
  const bool valid = q < a.Q;
  const int qq = valid ? q : 0;
  const int iy = qq / a.Nh, kz = qq - iy * a.Nh;
  const int ky = iy < a.N1 / 2 ? iy : iy - a.N1;
  const int nkyo = a.y_neg + a.y_pos + 1;
  const bool colconv = valid && ky >= -a.y_neg && ky <= a.y_pos && kz < a.nkzo;
  const int kyc = ky >= 0 ? ky : nkyo + ky;
  const int wrow = nkyo * a.nkzop, wfield = M * wrow;
  const auto rw = buf_rsrc(a.w4, 3L * wfield * CS);
  const int wo = colconv ? (t * wrow + kyc * a.nkzop + kz) * CS : BUF_OOB;
  C X0[E], X1[E], X2[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    X0[e] = buf_load_so<C>(rw, wo, P * e * wrow * CS);
    X1[e] = buf_load_so<C>(rw, wo, (wfield + P * e * wrow) * CS);
    X2[e] = buf_load_so<C>(rw, wo, (2 * wfield + P * e * wrow) * CS);
  }
  __syncthreads();
  C* line = lds + c;
  F::run(X0, t, line, T, tw);
  F::run(X1, t, line, T, tw);
  F::run(X2, t, line, T, tw);

  const int plane = a.N1 * a.Nh, fs = a.N0 * plane;
  const Rt K1 = a.ky[iy], K2c = a.kz[kz];
  const bool colok = colconv && !(a.mask_nyquist && (iy == a.N1 / 2 || kz == a.N2 / 2));
  const int nyq_ix = a.mask_nyquist ? a.N0 / 2 : -1;
  const double wz = (kz == 0 || kz == a.N2 / 2) ? 1.0 : 2.0;
  double esum = 0.0, osum = 0.0;
  const long span = 3L * fs * CS;
  const auto r0p = buf_rsrc(MODE == P5_RK0 ? a.A : a.B, span);
  const auto r1p = buf_rsrc(MODE == P5_RK12 ? a.A : a.Cacc, span);
  const auto r2p = buf_rsrc(a.Cacc, span);
  const auto rA = buf_rsrc(a.A, span);
  const auto rB = buf_rsrc(a.B, span);
  const auto rC = buf_rsrc(a.Cacc, span);
  const Rt bdt = a.bdt, adt = a.adt;
  const int so = (t * plane + q) * CS;  // row t of the column
  if (valid) {
    C cur[G][NL][3], nxt[G][NL][3];
    Rt kcur[G], knxt[G];
    auto load = [&](C(&d)[G][NL][3], Rt(&kd)[G], int j0) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int eo = P * (j0 + g) * plane;
        kd[g] = a.kx[t + P * (j0 + g)];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          d[g][0][i] = buf_load_so<C>(r0p, so, (i * fs + eo) * CS);
          if constexpr (NL > 1) d[g][1][i] = buf_load_so<C>(r1p, so, (i * fs + eo) * CS);
          if constexpr (NL > 2) d[g][2][i] = buf_load_so<C>(r2p, so, (i * fs + eo) * CS);
        }
      }
    };
    load(cur, kcur, 0);
#pragma unroll
    for (int j0 = 0; j0 < NK; j0 += G) {
      if (j0 + G < NK) load(nxt, knxt, j0 + G);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int j = j0 + g, e = j < E3 ? j : j + E3;
        const int ix = t + P * j;
        const int eo = P * j * plane;
        const bool conv = colok & (ix != nyq_ix);
        C r0 = conv ? cscale(X0[e], a.scale) : czero<C>();
        C r1 = conv ? cscale(X1[e], a.scale) : czero<C>();
        C r2 = conv ? cscale(X2[e], a.scale) : czero<C>();
        const Rt K0 = kcur[g];
        Rt k2 = K0 * K0;
        k2 += K1 * K1;
        k2 += K2c * K2c;
        const Rt den = k2 == Rt(0) ? Rt(1) : k2;
        const Rt o0 = K0 / den, o1 = K1 / den, o2 = K2c / den;
        // P_hat = sum_i rhs_i K_i/K^2 ; rhs_i -= P_hat K_i ; rhs -= (nu K^2) u_hat  (NS.py:208-215)
        C ph = mk<C>(r0.x * o0 + r1.x * o1, r0.y * o0 + r1.y * o1);
        ph = mk<C>(ph.x + r2.x * o2, ph.y + r2.y * o2);
        r0 = mk<C>(r0.x - ph.x * K0, r0.y - ph.y * K0);
        r1 = mk<C>(r1.x - ph.x * K1, r1.y - ph.y * K1);
        r2 = mk<C>(r2.x - ph.x * K2c, r2.y - ph.y * K2c);
        const Rt z = a.nu * k2;
        const C s0 = cur[g][0][0], s1 = cur[g][0][1], s2 = cur[g][0][2];
        r0 = mk<C>(r0.x - z * s0.x, r0.y - z * s0.y);
        r1 = mk<C>(r1.x - z * s1.x, r1.y - z * s1.y);
        r2 = mk<C>(r2.x - z * s2.x, r2.y - z * s2.y);
        auto upd = [](C base, C r, Rt s) { return mk<C>(base.x + s * r.x, base.y + s * r.y); };
        const int o0s = eo * CS, o1s = (fs + eo) * CS, o2s = (2 * fs + eo) * CS;
        if constexpr (MODE == P5_RK0) {
          // u0 = u1 + b dt rhs ; u2 = u1 + a dt rhs  (u1 = u2 = u0 at stage 0)
          buf_store_so_nt<C>(upd(s0, r0, bdt), rB, so, o0s);
          buf_store_so_nt<C>(upd(s1, r1, bdt), rB, so, o1s);
          buf_store_so_nt<C>(upd(s2, r2, bdt), rB, so, o2s);
          buf_store_so_nt<C>(upd(s0, r0, adt), rC, so, o0s);
          buf_store_so_nt<C>(upd(s1, r1, adt), rC, so, o1s);
          buf_store_so_nt<C>(upd(s2, r2, adt), rC, so, o2s);
        } else if constexpr (MODE == P5_RK12) {
          buf_store_so_nt<C>(upd(cur[g][1][0], r0, bdt), rB, so, o0s);
          buf_store_so_nt<C>(upd(cur[g][1][1], r1, bdt), rB, so, o1s);
          buf_store_so_nt<C>(upd(cur[g][1][2], r2, bdt), rB, so, o2s);
          buf_store_so_nt<C>(upd(cur[g][2][0], r0, adt), rC, so, o0s);
          buf_store_so_nt<C>(upd(cur[g][2][1], r1, adt), rC, so, o1s);
          buf_store_so_nt<C>(upd(cur[g][2][2], r2, adt), rC, so, o2s);
        } else {
          const C n0 = upd(cur[g][1][0], r0, adt), n1 = upd(cur[g][1][1], r1, adt), n2 = upd(cur[g][1][2], r2, adt);
          buf_store_so_nt<C>(n0, rA, so, o0s);
          buf_store_so_nt<C>(n1, rA, so, o1s);
          buf_store_so_nt<C>(n2, rA, so, o2s);
          const double v[6] = {n0.x, n0.y, n1.x, n1.y, n2.x, n2.y};
          energy_terms(v, K0, K1, K2c, wz, esum, osum);
        }
      }
      if (j0 + G < NK) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          kcur[g] = knxt[g];
#pragma unroll
          for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int i = 0; i < 3; ++i) cur[g][l][i] = nxt[g][l][i];
        }
      }
    }
  }
  if constexpr (MODE == P5_RK3) {
    if (a.partials) p5_partials<TH>(esum, osum, a.partials, tile);
  }
}

// Waves per SIMD that whole blocks of `threads` fill at a register-bound
// occupancy of `occ` (a CU holds 4 occ waves): an 8-wave block at occ 3 gets
// one block per CU, as at occ 2, so the larger register budget costs nothing.
constexpr int occ_blocks(int occ, int threads) {
  const int bw = threads / 64;
  const int w = (4 * occ) / bw * bw / 4;
  return w < 1 ? 1 : w;
}
// FP32 kernels are held to at least 2 waves per SIMD (256 registers): past
// that the compiler would move into the AGPR half of the register file and
// halve the occupancy.
template <typename Rt, int M, int E> constexpr int p5_occ() {
  const int o = occ_blocks(occ_for(3 * E, esz_of<Rt>()), SCfg<Rt, M, E>::THREADS);
  return sizeof(Rt) == 4 && o < 2 ? 2 : o;
}
template <typename Rt, int M, int E, int MODE, bool COMP = false, int D = 1>
__global__ void __launch_bounds__((SCfg<Rt, M, E>::THREADS / D), (p5_occ<Rt, M, E>())) k_p5(P5Args<Rt> a) {
  using C = cx<Rt>;
  constexpr int T = SCfg<Rt, M, E>::T / D;
  __shared__ C lds[M * T];
  __shared__ C tw[M];
  const int tile = xcd_tile((a.Q + T - 1) / T);
  if (tile < 0) return;
  table_to_lds(tw, (const C*)a.tw, M);
  p5_body<Rt, M, E, MODE, COMP, T>(a, tile, lds, tw);
}
// The 3/2-rule RK stages (p5k_body).
template <typename Rt, int M, int E, int MODE>
__global__ void __launch_bounds__((SCfg<Rt, M, E>::THREADS), (p5_occ<Rt, M, E>())) k_p5k(P5Args<Rt> a) {
  using C = cx<Rt>;
  constexpr int T = SCfg<Rt, M, E>::T;
  __shared__ C lds[M * T];
  __shared__ C tw[M];
  const int tile = xcd_tile((a.Q + T - 1) / T);
  if (tile < 0) return;
  table_to_lds(tw, (const C*)a.tw, M);
  p5k_body<Rt, M, E, MODE, T>(a, tile, lds, tw);
}

}  // namespace sdns
