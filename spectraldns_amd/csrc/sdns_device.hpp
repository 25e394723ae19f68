// Device building blocks: complex arithmetic, compile-time trigonometry,
// small in-register DFTs and the register/LDS Stockham line FFT used by every
// pass kernel.
//
// Line FFT scheme.  A line of M complex points is owned by P = M/E threads;
// thread t holds the E points at positions t + P*e (e = 0..E-1) in registers,
// both on entry and on exit (natural order).  Each pass of radix R performs
// E/R radix-R butterflies per thread in the Stockham autosort formulation of
// Govindaraju et al., "High Performance Discrete Fourier Transforms on
// Graphics Processors" (SC'08): butterfly j reads positions j + k*M/R, applies
// the twiddles W^(k*(j mod Ns)) with W = exp(-+2 pi i/(Ns R)), runs an R-point
// DFT, and writes positions (j/Ns)*Ns*R + (j mod Ns) + k*Ns.  Because
// j = t + P*b, the positions a thread reads in every pass are again
// t + P*e, so only the exchange between passes goes through LDS, and the last
// pass leaves X[t + P*e] in v[e].  All radices divide E, and E, M and the
// radix sequence are compile-time constants, so the whole line stays in VGPRs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "sdns_lineplan.hpp"

namespace sdns {

template <typename R> struct cx_of;
template <> struct cx_of<float> { using type = float2; };
template <> struct cx_of<double> { using type = double2; };
template <typename R> using cx = typename cx_of<R>::type;

template <typename C> __device__ __forceinline__ C mk(decltype(C::x) a, decltype(C::x) b) { C r; r.x = a; r.y = b; return r; }
// Load of a pass input that is read once, where neighbouring tiles do not
// share its 128-byte lines: non-temporal.
template <typename C> __device__ __forceinline__ C ld_stream(const C* p) {
  typedef decltype(C::x) V __attribute__((ext_vector_type(2)));
  const V w = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  C r;
  r.x = w.x;
  r.y = w.y;
  return r;
}
// Store of a pass output that the next pass reads once: non-temporal.
template <typename C> __device__ __forceinline__ void st_stream(C* p, C v) {
  typedef decltype(C::x) V __attribute__((ext_vector_type(2)));
  V w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
}
// Pass inputs and outputs of column-tiled kernels: non-temporal when the
// block's tile covers whole 128-byte lines (STREAM); narrower tiles share
// lines with their neighbours through L2 and stay cached.
template <bool STREAM, typename C> __device__ __forceinline__ C ld_pass(const C* p) {
  if constexpr (STREAM) return ld_stream(p);
  else return *p;
}
template <bool STREAM, typename C> __device__ __forceinline__ void st_pass(C* p, C v) {
  if constexpr (STREAM) st_stream(p, v);
  else *p = v;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return mk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return mk<C>(a.x * s, a.y * s); }
// multiply by +i (S=+1) or -i (S=-1)
template <int S, typename C> __device__ __forceinline__ C cmuli(C a) {
  if constexpr (S > 0) return mk<C>(-a.y, a.x);
  else return mk<C>(a.y, -a.x);
}

// ---------------------------------------------------------------- compile-time trig
// cos/sin(2 pi m / n) evaluated at compile time: exact integer reduction to a
// quadrant, complement to [0, pi/4], then a Taylor series (|x| <= pi/4, 26 terms).
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 14; ++i) { term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0)); sum += term; }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) { term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i)); sum += term; }
  return sum;
}
struct cs_pair { double c, s; };
constexpr cs_pair cexp2pi(long m, long n) {
  m %= n; if (m < 0) m += n;
  long a = 4 * m;
  long q = a / n, rem = a - q * n;  // angle = (pi/2)(q + rem/n)
  double c = 0, s = 0;
  if (2 * rem <= n) { double phi = (kPi / 2) * (double)rem / (double)n; c = taylor_cos(phi); s = taylor_sin(phi); }
  else { double phi = (kPi / 2) * (double)(n - rem) / (double)n; c = taylor_sin(phi); s = taylor_cos(phi); }
  switch (q) {
    case 0: return {c, s};
    case 1: return {-s, c};
    case 2: return {-c, -s};
    default: return {s, -c};
  }
}

// ---------------------------------------------------------------- small DFTs
// DIR = -1: forward (exp(-2 pi i jk/R)), DIR = +1: inverse (unscaled).
template <int DIR, typename C> __device__ __forceinline__ void dft2(C& a, C& b) {
  C t = a; a = cadd(t, b); b = csub(t, b);
}
template <int DIR, typename C> __device__ __forceinline__ void dft3(C& a0, C& a1, C& a2) {
  using R = decltype(C::x);
  constexpr R h = R(0.866025403784438646763723170752936183);  // sqrt(3)/2
  C s = cadd(a1, a2), d = csub(a1, a2);
  C m = mk<C>(a0.x - R(0.5) * s.x, a0.y - R(0.5) * s.y);
  a0 = cadd(a0, s);
  C r = cmuli<DIR>(cscale(d, h));  // DIR * i * sqrt(3)/2 * d
  a1 = cadd(m, r);
  a2 = csub(m, r);
}
template <int DIR, typename C> __device__ __forceinline__ void dft4(C& a0, C& a1, C& a2, C& a3) {
  C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
  C t3 = cmuli<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2); a2 = csub(t0, t2); a1 = cadd(t1, t3); a3 = csub(t1, t3);
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dft(C (&v)[R]);

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0..N-1.
template <typename F, int... Is>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// Cooley-Tukey split R = R1*R2 in registers: n = n1*R2 + n2, k = k1 + R1*k2,
// X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_R^(n2 k1) sum_n1 x[n1 R2 + n2] W_R1^(n1 k1).
template <int R1, int R2, int DIR, typename C> __device__ __forceinline__ void dft_ct(C (&v)[R1 * R2]) {
  using Rt = decltype(C::x);
  constexpr int R = R1 * R2;
  C y[R2][R1];
  sfor<R2>([&](auto n2c) {
    constexpr int n2 = decltype(n2c)::value;
    C a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[n1 * R2 + n2];
    dft<R1, DIR>(a);
    sfor<R1>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      if constexpr (n2 * k1 != 0) {
        constexpr long m = (((long)DIR * n2 * k1) % R + R) % R;
        if constexpr ((4 * m) % R == 0) {
          // quarter turns i^q: a swap and a sign instead of a complex product
          constexpr int q = (int)(4 * m / R);
          if constexpr (q == 1) a[k1] = cmuli<1>(a[k1]);
          else if constexpr (q == 2) a[k1] = mk<C>(-a[k1].x, -a[k1].y);
          else if constexpr (q == 3) a[k1] = cmuli<-1>(a[k1]);
        } else {
          constexpr cs_pair w = cexp2pi(m, R);
          a[k1] = cmul(a[k1], mk<C>((Rt)w.c, (Rt)w.s));
        }
      }
      y[n2][k1] = a[k1];
    });
  });
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    C b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2][k1];
    dft<R2, DIR>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}

// Prime-factor (Good-Thomas) split R = R1*R2 with gcd(R1, R2) = 1: input
// n = (n1 R2 + n2 R1) mod R, output k = (k1 R2 t2 + k2 R1 t1) mod R with
// t2 = R2^-1 mod R1 and t1 = R1^-1 mod R2; then
// X[k] = sum_n2 W_R2^(n2 k2) sum_n1 x[n] W_R1^(n1 k1), with no twiddles
// between the two stages.
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}
template <int R1, int R2, int DIR, typename C> __device__ __forceinline__ void dft_pfa(C (&v)[R1 * R2]) {
  constexpr int R = R1 * R2, t2 = inv_mod(R2 % R1, R1), t1 = inv_mod(R1 % R2, R2);
  C y[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    C a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[(n1 * R2 + n2 * R1) % R];
    dft<R1, DIR>(a);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) y[n2][k1] = a[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    C b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2][k1];
    dft<R2, DIR>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[(k1 * R2 * t2 + k2 * R1 * t1) % R] = b[k2];
  }
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dft(C (&v)[R]) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<DIR>(v[0], v[1]);
  } else if constexpr (R == 3) {
    dft3<DIR>(v[0], v[1], v[2]);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 6) {
    dft_pfa<2, 3, DIR>(v);
  } else if constexpr (R == 8) {
    dft_ct<2, 4, DIR>(v);
  } else if constexpr (R == 12) {
    dft_pfa<3, 4, DIR>(v);
  } else if constexpr (R == 16) {
    dft_ct<4, 4, DIR>(v);
  } else if constexpr (R == 24) {
    dft_pfa<3, 8, DIR>(v);
  } else {
    static_assert(R == 0, "unsupported radix");
  }
}


// Points per thread.  Variant 0 is the default schedule; variant 1 is the
// alternative the autotuner may pick (same as variant 0 where no other
// factorisation fits).  Strided (x, y) complex lines:
constexpr int strided_E(int M, int v = 0) {
  switch (M) {
    case 4: return 4;   case 6: return 6;   case 8: return 8;    case 12: return 12;
    case 16: return 16; case 24: return 24; case 32: return 8;   case 36: return 12;
    case 48: return 12; case 64: return v ? 4 : 8;  case 72: return 12; case 96: return v ? 24 : 12;
    case 128: return v ? 8 : 16; case 144: return 12; case 192: return v ? 24 : 12; case 256: return v ? 8 : 16;
    case 288: return 12; case 384: return v ? 24 : 12; case 512: return v ? 8 : 16; case 576: return 12;
    case 768: return v ? 24 : 12; case 1024: return v ? 8 : 16; case 1536: return v ? 24 : 12;
    default: return 0;
  }
}
// Half-length complex lines of the z-axis real transforms:
constexpr int zline_E(int n, int v = 0) {
  switch (n) {
    case 4: return 4;   case 6: return 6;   case 8: return 8;   case 12: return 12;
    case 16: return 8;  case 24: return 12; case 32: return v ? 2 : 8; case 36: return 12;
    case 48: return 12; case 64: return 8;  case 72: return 12; case 96: return 12;
    case 128: return v ? 16 : 8; case 144: return 12; case 192: return v ? 6 : 12; case 256: return v ? 8 : 16;
    case 288: return 12; case 384: return v ? 24 : 12; case 512: return v ? 8 : 16; case 768: return 12;
    default: return 0;
  }
}
// Number of distinct schedule variants compiled for a length.
constexpr int strided_variants(int M) { return strided_E(M, 1) != strided_E(M, 0) ? 2 : 1; }
constexpr int zline_variants(int n) { return zline_E(n, 1) != zline_E(n, 0) ? 2 : 1; }

// Barrier of an LDS exchange.  WAVE: every thread that exchanges data is in
// the calling wave; LDS operations of one wave complete in issue order, so a
// wave-level barrier (no s_barrier) and wave-scope fences that stop the
// compiler from moving LDS accesses across it suffice.  Otherwise the block
// barrier.
template <bool WAVE> __device__ __forceinline__ void lds_barrier() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ------------------------------------------------------------------ buffer access
// Raw buffer loads and stores (stride 0, byte offsets).  An access whose
// offset is at or beyond the descriptor's byte count reads zeros / writes
// nothing, so a masked element costs a select of its offset instead of a
// branch.  A descriptor covers one block's span of lines (at most 2^31 - 1
// bytes); BUF_OOB is an offset beyond every such span.
constexpr int BUF_OOB = 0x7ff00000;
typedef unsigned int buf_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int buf_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, long bytes) {
  const int nb = bytes > 0x7fffffffL ? 0x7fffffff : (int)bytes;
  // word 3 = 0x00020000: 32-bit data format, the gfx9 raw-buffer setting
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nb, 0x00020000);
}
template <typename C> __device__ __forceinline__ C buf_load(__amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(C) == 8) {
    return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  } else {
    static_assert(sizeof(C) == 16, "complex of float or double");
    return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  }
}
// With a uniform part of the offset in a scalar register (soff): the lane's
// voff + soff addresses the element; the descriptor must cover voff + soff.
// The store is non-temporal (the RK stage values are read once, by the next
// stage).
template <typename C> __device__ __forceinline__ C buf_load_so(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(C) == 8) {
    return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    static_assert(sizeof(C) == 16, "complex of float or double");
    return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
}
template <typename C> __device__ __forceinline__ void buf_store_so_nt(C x, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(C) == 8) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(buf_u32x2, x), r, voff, soff, 2);
  } else {
    static_assert(sizeof(C) == 16, "complex of float or double");
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(buf_u32x4, x), r, voff, soff, 2);
  }
}
// Non-temporal variant (cache policy nt) for pass outputs read once by the next pass.
template <typename C> __device__ __forceinline__ void buf_store_nt(C x, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(C) == 8) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(buf_u32x2, x), r, off, 0, 2);
  } else {
    static_assert(sizeof(C) == 16, "complex of float or double");
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(buf_u32x4, x), r, off, 0, 2);
  }
}

// tw[m] = exp(-2 pi i m / M), m in [0, M): the forward twiddle table of the
// line length; the inverse direction uses its conjugate.  WAVE: the P threads
// of a line are in one wave (exchanges need no block barrier).
//
// First-exchange layout: with P = 16 threads per line and a first radix R0 = E
// = 12 (the 192-point lines), the first pass writes element p = 12 t + k,
// which puts the threads t, t + 4, t + 8, t + 12 of a line on the same LDS
// banks.  SWZ stores element p of that exchange at p ^ ((p / 48) & 3): for the
// writes (t = 4a + c) the XOR is a = t / 4 and the banks of the 16 threads
// become distinct; for the reads p = t + 16 e it is the compile-time
// (e / 3) & 3, so both sides keep immediate offsets from per-thread bases.
template <typename Rt, int M, int E, int DIR, bool WAVE = false, bool SWZ = false> struct LineFFT {
  using C = cx<Rt>;
  static constexpr int P = M / E;
  static constexpr int NP = num_passes(M, E);
  static_assert(E > 0 && M % E == 0 && NP > 0, "unsupported line length");
  static_assert(!SWZ || (P == 16 && radix_at(M, E, 0) == 12 && E == 12 && NP > 1),
                "the first-exchange layout is derived for 16 threads and a first radix of 12");

  // lds: this line's exchange area, element p at lds[p * es].  Every thread
  // of the block (of the wave, if WAVE) must call run() (it contains barriers
  // when NP > 1).
  __device__ static __forceinline__ void run(C (&v)[E], int t, C* lds, int es, const C* __restrict__ tw) {
    pass<0, 1>(v, t, lds, es, tw);
  }

  template <int PASS, int NS>
  __device__ static __forceinline__ void pass(C (&v)[E], int t, C* lds, int es, const C* __restrict__ tw) {
    constexpr int R = radix_at(M, E, PASS);
    constexpr int NB = E / R;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      C x[R];
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] = v[b + k * NB];
      if constexpr (NS > 1) {
        const int jm = (t + P * b) % NS;
        constexpr int step = M / (NS * R);
#pragma unroll
        for (int k = 1; k < R; ++k) {
          const C w = tw[jm * k * step];
          x[k] = (DIR < 0) ? cmul(x[k], w) : cmulc(x[k], w);
        }
      }
      dft<R, DIR>(x);
#pragma unroll
      for (int k = 0; k < R; ++k) v[b + k * NB] = x[k];
    }
    if constexpr (SWZ && PASS == 0) {
      // p = 12 t + k stored at 12 t + (k ^ a), a = (t / 4) & 3 (12 t is a
      // multiple of 4, so the XOR stays in the low two bits of k); read back
      // at 16 e + (t ^ ((e / 3) & 3))
      const int a = (t >> 2) & 3;
      C* wbase[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) wbase[c] = lds + (12 * t + (c ^ a)) * es;
#pragma unroll
      for (int k = 0; k < R; ++k) wbase[k & 3][4 * (k >> 2) * es] = v[k];
      lds_barrier<WAVE>();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[(16 * e + (t ^ ((e / 3) & 3))) * es];
      lds_barrier<WAVE>();
      pass<PASS + 1, NS * R>(v, t, lds, es, tw);
    } else if constexpr (PASS + 1 < NP) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int j = t + P * b;
        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
        for (int k = 0; k < R; ++k) lds[(base + k * NS) * es] = v[b + k * NB];
      }
      lds_barrier<WAVE>();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[(t + P * e) * es];
      lds_barrier<WAVE>();
      pass<PASS + 1, NS * R>(v, t, lds, es, tw);
    }
  }
};

}  // namespace sdns
