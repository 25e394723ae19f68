// Launcher bodies (size dispatch) and elementwise kernels.  Included once per
// working precision by sdns_f32.hip / sdns_f64.hip.
#pragma once
#include <type_traits>

#include "sdns_kernels.hpp"
#include "sdns_launch.hpp"

namespace sdns {

// ---------------------------------------------------------------- launch helpers
// Small grids: a launch of few blocks (a 64^3 box's passes have 43-66 column
// tiles of 32 columns, its z pass 64 blocks of 64 lines) leaves most CUs
// idle, so the launcher divides the tile (columns, or lines per block of the
// z kernels) by D in {1, 2, 4}, doubling D while the launch has fewer than
// 512 blocks.  D > 1 is compiled for lines of up to 128 points (z: 64) and
// only while a block stays whole waves.
constexpr int max_div(int len, int lim, int T, int threads) {
  int d = 1;
  while (2 * d <= 4 && len <= lim && T % (2 * d) == 0 && threads / (2 * d) >= 64 && threads % (128 * d) == 0) d *= 2;
  return d;
}
inline int tile_div(long blocks1, int maxd) {
  int d = 1;
  while (2 * d <= maxd && blocks1 * d < 512) d *= 2;
  return d;
}
// Calls go(std::integral_constant<int, D>) for the runtime divisor d (1, 2, 4 <= MD).
template <int MD, typename F> void with_div(int d, F go) {
  if constexpr (MD >= 4) {
    if (d == 4) return go(std::integral_constant<int, 4>{});
  }
  if constexpr (MD >= 2) {
    if (d == 2) return go(std::integral_constant<int, 2>{});
  }
  go(std::integral_constant<int, 1>{});
}
template <typename Rt, int M, int E> void p1_go(const P1Args<Rt>& a, hipStream_t s) {
  using Cfg = SCfg<Rt, M, E>;
  constexpr int MD = max_div(M, 128, Cfg::T, Cfg::THREADS);
  const int ny = a.w ? 6 : 3;
  with_div<MD>(tile_div((long)ny * ((a.Q + Cfg::T - 1) / Cfg::T), MD), [&](auto dd) {
    constexpr int D = decltype(dd)::value, T = Cfg::T / D;
    hipLaunchKernelGGL((k_p1<Rt, M, E, D>), dim3(xcd_grid((a.Q + T - 1) / T), ny), dim3(Cfg::THREADS / D), 0, s, a);
  });
}
template <typename Rt, int M, int E> void p2_go(const P2Args<Rt>& a, hipStream_t s) {
  using Cfg = SCfg<Rt, M, E>;
  constexpr int MD = max_div(M, 128, Cfg::T, Cfg::THREADS);
  with_div<MD>(tile_div(2L * ((a.Q + Cfg::T - 1) / Cfg::T), MD), [&](auto dd) {
    constexpr int D = decltype(dd)::value, T = Cfg::T / D;
    hipLaunchKernelGGL((k_p2<Rt, M, E, D>), dim3(xcd_grid((a.Q + T - 1) / T), 2), dim3(Cfg::THREADS / D), 0, s, a);
  });
}
template <typename Rt, int n, int E> void p3_go(const P3Args<Rt>& a, hipStream_t s) {
  using Cfg = ZCfg<Rt, n, E>;
  constexpr int MD = max_div(n, 64, Cfg::LPB, Cfg::THREADS);
  with_div<MD>(tile_div((a.L + Cfg::LPB - 1) / Cfg::LPB, MD), [&](auto dd) {
    constexpr int D = decltype(dd)::value, LPB = Cfg::LPB / D;
    hipLaunchKernelGGL((k_p3<Rt, n, E, D>), dim3((a.L + LPB - 1) / LPB), dim3(Cfg::THREADS / D), 0, s, a);
  });
}
// The LDS-assisted z pass, specialised (for n >= 64) on the dealiased bin
// counts: the 3/2 rule holds 2n/3 input bins and needs 2n/3 + 1 outputs, the
// 2/3 rule holds ceil(2n/3) inputs; any other plan gets the generic kernel.
template <typename Rt, int n, int E> void p3b_go(const P3Args<Rt>& a, hipStream_t s) {
  using Cfg = ZBCfg<Rt, n, E>;
  const dim3 g((a.L + Cfg::LPB - 1) / Cfg::LPB), b(Cfg::THREADS);
  if constexpr (n >= 64) {
    constexpr int K32 = 2 * n / 3, K23 = (2 * n + 2) / 3;
    if (n % 3 == 0 && a.nkz <= K32 && a.nkzo <= K32 + 1) {
      hipLaunchKernelGGL((k_p3b<Rt, n, E, K32, K32 + 1>), g, b, 0, s, a);
      return;
    }
    if (a.nkz <= K23) {
      hipLaunchKernelGGL((k_p3b<Rt, n, E, K23, n + 1>), g, b, 0, s, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_p3b<Rt, n, E>), g, b, 0, s, a);
}
// Scheme C (LDS-DMA staging) with LPB lines per block, if its block is whole
// waves and its LDS fits in 64 KB; false otherwise.
template <typename Rt, int n, int E, int LPB> bool p3c_try(const P3Args<Rt>& a, hipStream_t s) {
  constexpr int TH = LPB * (n / E);
  if constexpr (n < 32 || TH % 64 != 0 || TH > 512) {
    return false;
  } else {
    const size_t bytes = (size_t)zc_layout(n, n / E, LPB, a.nkzp).total * sizeof(cx<Rt>);
    if (bytes > 65536) return false;
    hipLaunchKernelGGL((k_p3c<Rt, n, E, LPB>), dim3((a.L + LPB - 1) / LPB), dim3(TH), bytes, s, a);
    return true;
  }
}
template <typename Rt, int n, int E> void p3c_go(const P3Args<Rt>& a, hipStream_t s) {
  if (p3c_try<Rt, n, E, 8>(a, s)) return;
  if (p3c_try<Rt, n, E, 4>(a, s)) return;
  p3b_go<Rt, n, E>(a, s);  // rows too long for the LDS budget: LDS-assisted scheme
}
// The 3/2-rule P5 stages (k_p5k) are compiled for the line lengths of the
// padded grids from 96 to 768 points; schedule variants 2 and 3 run them when
// the stage is a plain RK stage of the 3/2 rule whose arrays a buffer
// descriptor reaches (otherwise, and for variants 0 and 1, p5_body runs).
constexpr bool p5k_built(int M, int E) { return M % 3 == 0 && E % 3 == 0 && M >= 96 && M <= 768; }
template <typename Rt> bool p5k_fits(int M, int mode, const P5Args<Rt>& a) {
  if (mode == P5_RHS || a.Alo || a.source || a.phat || a.conv_only) return false;
  if (3 * a.N0 != 2 * M || a.x.neg != a.N0 / 2 || a.x.pos != a.N0 / 2 - 1 || a.x.base != a.N0) return false;
  const long esz = 2 * (long)sizeof(Rt);
  const long state = 3L * a.N0 * a.N1 * a.Nh * esz;
  const long w4 = 3L * M * (a.y_neg + a.y_pos + 1) * a.nkzop * esz;
  return state <= BUF_OOB && w4 <= BUF_OOB;
}
template <typename Rt, int M, int E> void p5_go(int mode, const P5Args<Rt>& a, hipStream_t s, int* nblk, bool fixed_rows) {
  using Cfg = SCfg<Rt, M, E>;
  const int nt = (a.Q + Cfg::T - 1) / Cfg::T;
  if (nblk) *nblk = nt;
  const dim3 g(xcd_grid(nt)), b(Cfg::THREADS);
  if constexpr (p5k_built(M, E)) {
    if (fixed_rows && p5k_fits(M, mode, a)) {
      switch (mode) {
        case P5_RK0: hipLaunchKernelGGL((k_p5k<Rt, M, E, P5_RK0>), g, b, 0, s, a); break;
        case P5_RK12: hipLaunchKernelGGL((k_p5k<Rt, M, E, P5_RK12>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_p5k<Rt, M, E, P5_RK3>), g, b, 0, s, a); break;
      }
      return;
    }
  }
  if constexpr (sizeof(Rt) == 4) {
    if (a.Alo && mode != P5_RHS) {
      switch (mode) {
        case P5_RK0: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RK0, true>), g, b, 0, s, a); break;
        case P5_RK12: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RK12, true>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RK3, true>), g, b, 0, s, a); break;
      }
      return;
    }
  }
  constexpr int MD = max_div(M, 128, Cfg::T, Cfg::THREADS);
  with_div<MD>(tile_div(nt, MD), [&](auto dd) {
    constexpr int D = decltype(dd)::value, T = Cfg::T / D;
    const int ntd = (a.Q + T - 1) / T;
    if (nblk) *nblk = ntd;
    const dim3 gd(xcd_grid(ntd)), bd(Cfg::THREADS / D);
    switch (mode) {
      case P5_RHS: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RHS, false, D>), gd, bd, 0, s, a); break;
      case P5_RK0: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RK0, false, D>), gd, bd, 0, s, a); break;
      case P5_RK12: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RK12, false, D>), gd, bd, 0, s, a); break;
      default: hipLaunchKernelGGL((k_p5<Rt, M, E, P5_RK3, false, D>), gd, bd, 0, s, a); break;
    }
  });
}
template <typename Rt, int M, int E> void strided_go(int dir, int nfields, const StridedArgs<Rt>& a0, hipStream_t s) {
  using Cfg = SCfg<Rt, M, E>;
  const int nt = (a0.Q + Cfg::T - 1) / Cfg::T;
  constexpr int MD = max_div(M, 128, Cfg::T, Cfg::THREADS);
  with_div<MD>(tile_div((long)nfields * nt, MD), [&](auto dd) {
    constexpr int D = decltype(dd)::value, T = Cfg::T / D;
    StridedArgs<Rt> a = a0;
    a.stream_in = tiles_whole_lines<Rt>(a.in, T, a.cdiv, a.in_hi, a.in_row, a.in_field);
    a.stream_out = tiles_whole_lines<Rt>(a.out, T, a.cdiv, a.out_hi, a.out_row, a.out_field);
    const dim3 gd(xcd_grid((a.Q + T - 1) / T), nfields), bd(Cfg::THREADS / D);
    if (dir < 0) hipLaunchKernelGGL((k_strided<Rt, M, E, -1, D>), gd, bd, 0, s, a);
    else hipLaunchKernelGGL((k_strided<Rt, M, E, 1, D>), gd, bd, 0, s, a);
  });
}
template <typename Rt, int n, int E> void z_go(int dir, int nfields, const ZArgs<Rt>& a, hipStream_t s) {
  using Cfg = ZCfg<Rt, n, E>;
  constexpr int MD = max_div(n, 64, Cfg::LPB, Cfg::THREADS);
  with_div<MD>(tile_div((long)nfields * ((a.L + Cfg::LPB - 1) / Cfg::LPB), MD), [&](auto dd) {
    constexpr int D = decltype(dd)::value, LPB = Cfg::LPB / D;
    const dim3 g((a.L + LPB - 1) / LPB, nfields), b(Cfg::THREADS / D);
    if (dir < 0) hipLaunchKernelGGL((k_z_r2c<Rt, n, E, D>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_z_c2r<Rt, n, E, D>), g, b, 0, s, a);
  });
}

// (variant 1 of a length with a single schedule instantiates the same kernel as variant 0)
template <typename Rt> int launch_p1(int M, int var, const P1Args<Rt>& a, hipStream_t s) {
  switch (M) {
#define SDNS_X(m)                                                        \
  case m:                                                                \
    if (var) p1_go<Rt, m, strided_E(m, 1)>(a, s);                        \
    else p1_go<Rt, m, strided_E(m, 0)>(a, s);                            \
    return 0;
    SDNS_STRIDED_SIZES(SDNS_X)
#undef SDNS_X
    default: return -1;
  }
}

template <typename Rt> int launch_p2(int M, int var, const P2Args<Rt>& a, hipStream_t s) {
  switch (M) {
#define SDNS_X(m)                                                        \
  case m:                                                                \
    if (var) p2_go<Rt, m, strided_E(m, 1)>(a, s);                        \
    else p2_go<Rt, m, strided_E(m, 0)>(a, s);                            \
    return 0;
    SDNS_STRIDED_SIZES(SDNS_X)
#undef SDNS_X
    default: return -1;
  }
}

// P3 variants: 0 and 1 = all six lines in registers with schedule 0 / 1;
// 2 and 3 = the LDS-assisted scheme with schedule 0 / 1; 4 and 5 = the
// LDS-DMA staging scheme with schedule 0 / 1.
template <typename Rt> int launch_p3(int n, int var, const P3Args<Rt>& a, hipStream_t s) {
  switch (n) {
#define SDNS_X(m)                                                        \
  case m:                                                                \
    if (var == 5) p3c_go<Rt, m, zline_E(m, 1)>(a, s);                    \
    else if (var == 4) p3c_go<Rt, m, zline_E(m, 0)>(a, s);               \
    else if (var == 3) p3b_go<Rt, m, zline_E(m, 1)>(a, s);               \
    else if (var == 2) p3b_go<Rt, m, zline_E(m, 0)>(a, s);               \
    else if (var == 1) p3_go<Rt, m, zline_E(m, 1)>(a, s);                \
    else p3_go<Rt, m, zline_E(m, 0)>(a, s);                              \
    return 0;
    SDNS_HALF_SIZES(SDNS_X)
#undef SDNS_X
    default: return -1;
  }
}

template <typename Rt> int launch_p5(int M, int var, int mode, const P5Args<Rt>& a, hipStream_t s, int* nblk) {
  switch (M) {
#define SDNS_X(m)                                                        \
  case m:                                                                \
    if (var & 1) p5_go<Rt, m, strided_E(m, 1)>(mode, a, s, nblk, var >= 2); \
    else p5_go<Rt, m, strided_E(m, 0)>(mode, a, s, nblk, var >= 2);      \
    return 0;
    SDNS_STRIDED_SIZES(SDNS_X)
#undef SDNS_X
    default: return -1;
  }
}

template <typename Rt> int launch_strided(int M, int var, int dir, int nfields, const StridedArgs<Rt>& a,
                                          hipStream_t s) {
  switch (M) {
#define SDNS_X(m)                                                        \
  case m:                                                                \
    if (var) strided_go<Rt, m, strided_E(m, 1)>(dir, nfields, a, s);     \
    else strided_go<Rt, m, strided_E(m, 0)>(dir, nfields, a, s);         \
    return 0;
    SDNS_STRIDED_SIZES(SDNS_X)
#undef SDNS_X
    default: return -1;
  }
}

template <typename Rt> int launch_z(int n, int var, int dir, int nfields, const ZArgs<Rt>& a, hipStream_t s) {
  switch (n) {
#define SDNS_X(m)                                                        \
  case m:                                                                \
    if (var) z_go<Rt, m, zline_E(m, 1)>(dir, nfields, a, s);             \
    else z_go<Rt, m, zline_E(m, 0)>(dir, nfields, a, s);                 \
    return 0;
    SDNS_HALF_SIZES(SDNS_X)
#undef SDNS_X
    default: return -1;
  }
}

// ------------------------------------------------------------------ elementwise
// Spectral point index i -> (ix, iy, kz) of the [N0][N1][Nh] grid.
template <typename Rt, int OP>
__global__ void __launch_bounds__(256) k_elementwise(EwArgs<Rt> a) {
  using C = cx<Rt>;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const long n = a.n;
  if constexpr (OP == EW_CROSS1) {  // real c = a x b (maths/cross.py:13-25)
    const Rt* A = (const Rt*)a.a;
    const Rt* B = (const Rt*)a.b;
    Rt* O = (Rt*)a.out;
    const Rt a0 = A[i], a1 = A[n + i], a2 = A[2 * n + i];
    const Rt b0 = B[i], b1 = B[n + i], b2 = B[2 * n + i];
    O[i] = a1 * b2 - a2 * b1;
    O[n + i] = a2 * b0 - a0 * b2;
    O[2 * n + i] = a0 * b1 - a1 * b0;
    return;
  } else if constexpr (OP == EW_ELSASSER_PRODUCTS) {  // solvers/MHD.py:118-130 on n physical points
    const Rt* A = (const Rt*)a.a;  // u0 u1 u2 b0 b1 b2
    Rt* O = (Rt*)a.out;            // z0_i * z1_j at field 3 i + j
    Rt z0[3], z1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const Rt u = A[c * n + i], bb = A[(3 + c) * n + i];
      z0[c] = u + bb;
      z1[c] = u - bb;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) O[(3 * r + c) * n + i] = z0[r] * z1[c];
    return;
  } else if constexpr (OP == EW_DOT3) {  // out = sum(u*g, 0) on n points (NS.py:138-145)
    const Rt* U = (const Rt*)a.a;
    const Rt* G = (const Rt*)a.b;
    Rt s = U[i] * G[i];
    s += U[n + i] * G[n + i];
    s += U[2 * n + i] * G[2 * n + i];
    ((Rt*)a.out)[i] = s;
    return;
  } else if constexpr (OP == EW_UU_PRODUCTS) {  // u_i u_j, i <= j, on n points (NS.py:147-165)
    const Rt* U = (const Rt*)a.a;
    Rt* O = (Rt*)a.out;
    const Rt u0 = U[i], u1 = U[n + i], u2 = U[2 * n + i];
    O[i] = u0 * u0;
    O[n + i] = u0 * u1;
    O[2 * n + i] = u0 * u2;
    O[3 * n + i] = u1 * u1;
    O[4 * n + i] = u1 * u2;
    O[5 * n + i] = u2 * u2;
    return;
  } else if constexpr (OP == EW_MUL_MINUS_I) {  // out = -1j * a (NS.py:96-99)
    const C v = ((const C*)a.a)[i];
    ((C*)a.out)[i] = mk<C>(v.y, -v.x);
    return;
  } else {
    const int kz = (int)(i % a.Nh);
    const long r = i / a.Nh;
    const int iy = (int)(r % a.N1);
    const int ix = (int)(r / a.N1);
    const Rt K0 = a.kx[ix], K1 = a.ky[iy], K2 = a.kz[kz];
    const C* A = (const C*)a.a;
    C* O = (C*)a.out;
    if constexpr (OP == EW_CROSS2) {  // c = 1j*(K x a) (maths/cross.py:27-32)
      const C a0 = A[i], a1 = A[n + i], a2 = A[2 * n + i];
      const C c0 = mk<C>(K1 * a2.x - K2 * a1.x, K1 * a2.y - K2 * a1.y);
      const C c1 = mk<C>(K2 * a0.x - K0 * a2.x, K2 * a0.y - K0 * a2.y);
      const C c2 = mk<C>(K0 * a1.x - K1 * a0.x, K0 * a1.y - K1 * a0.y);
      O[i] = cmuli<1>(c0);
      O[n + i] = cmuli<1>(c1);
      O[2 * n + i] = cmuli<1>(c2);
    } else if constexpr (OP == EW_CROSS2_KOK) {  // c = 1j*((K/K^2) x a), K^2 = 0 -> 1 (solvers/VV.py:71-72)
      Rt k2 = K0 * K0;
      k2 += K1 * K1;
      k2 += K2 * K2;
      const Rt den = k2 == Rt(0) ? Rt(1) : k2;
      const Rt o0 = K0 / den, o1 = K1 / den, o2 = K2 / den;
      const C a0 = A[i], a1 = A[n + i], a2 = A[2 * n + i];
      const C c0 = mk<C>(o1 * a2.x - o2 * a1.x, o1 * a2.y - o2 * a1.y);
      const C c1 = mk<C>(o2 * a0.x - o0 * a2.x, o2 * a0.y - o0 * a2.y);
      const C c2 = mk<C>(o0 * a1.x - o1 * a0.x, o0 * a1.y - o1 * a0.y);
      O[i] = cmuli<1>(c0);
      O[n + i] = cmuli<1>(c1);
      O[2 * n + i] = cmuli<1>(c2);
    } else if constexpr (OP == EW_VV_TAIL) {  // VV.py:104-146 after F(u x w): cross2, mask, add_linear
      // a = v_hat = F(u x w), b = w_hat, aux = Source (may be null); out = rhs
      const C v0 = A[i], v1 = A[n + i], v2 = A[2 * n + i];
      const C c0 = cmuli<1>(mk<C>(K1 * v2.x - K2 * v1.x, K1 * v2.y - K2 * v1.y));
      const C c1 = cmuli<1>(mk<C>(K2 * v0.x - K0 * v2.x, K2 * v0.y - K0 * v2.y));
      const C c2 = cmuli<1>(mk<C>(K0 * v1.x - K1 * v0.x, K0 * v1.y - K1 * v0.y));
      const bool nyq = a.mask_nyquist && (ix == a.N0 / 2 || iy == a.N1 / 2 || kz == a.N2 / 2);
      Rt k2 = K0 * K0;
      k2 += K1 * K1;
      k2 += K2 * K2;
      const Rt z = a.nu * k2;
      const C* W = (const C*)a.b;
      const C cc[3] = {c0, c1, c2};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const C w = W[c * n + i];
        C r = nyq ? czero<C>() : cc[c];
        r = mk<C>(r.x - z * w.x, r.y - z * w.y);
        if (a.aux) {
          const C src = ((const C*)a.aux)[c * n + i];
          r = mk<C>(r.x + src.x, r.y + src.y);
        }
        O[c * n + i] = r;
      }
    } else if constexpr (OP == EW_ADD_LINEAR) {  // rhs -= nu*K2*w_hat; rhs += Source (VV.py:108-113)
      Rt k2 = K0 * K0;
      k2 += K1 * K1;
      k2 += K2 * K2;
      const Rt z = a.nu * k2;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const C w = A[c * n + i];
        C r = O[c * n + i];
        r = mk<C>(r.x - z * w.x, r.y - z * w.y);
        if (a.b) {
          const C src = ((const C*)a.b)[c * n + i];
          r = mk<C>(r.x + src.x, r.y + src.y);
        }
        O[c * n + i] = r;
      }
    } else if constexpr (OP == EW_GRAD_TRUNC) {  // out[j] = 1j*K_j*a on the kept modes (NS.py:141-142)
      const int kx = ix < a.N0 / 2 ? ix : ix - a.N0, ky = iy < a.N1 / 2 ? iy : iy - a.N1;
      const bool keep = (kx < 0 ? -kx : kx) <= a.keep[0] && (ky < 0 ? -ky : ky) <= a.keep[1] && kz <= a.keep[2];
      const C v = A[i];
      const Rt Kv[3] = {K0, K1, K2};
#pragma unroll
      for (int j = 0; j < 3; ++j) O[j * n + i] = keep ? mk<C>(-Kv[j] * v.y, Kv[j] * v.x) : czero<C>();
    } else if constexpr (OP == EW_DIV_UU) {  // divergence_convection (NS.py:147-165) times a.nu (+-1)
      // a = F(u0u0, u0u1, u0u2, u1u1, u1u2, u2u2)
      const C a00 = A[i], a01 = A[n + i], a02 = A[2 * n + i], a11 = A[3 * n + i], a12 = A[4 * n + i],
              a22 = A[5 * n + i];
      C s0 = mk<C>(K0 * a00.x, K0 * a00.y);
      s0 = mk<C>(s0.x + K1 * a01.x, s0.y + K1 * a01.y);
      s0 = mk<C>(s0.x + K2 * a02.x, s0.y + K2 * a02.y);
      const C r0 = cmuli<1>(s0);
      const C p1 = mk<C>(-K0 * a01.y, K0 * a01.x), q1 = mk<C>(-K1 * a11.y, K1 * a11.x),
              t1 = mk<C>(-K2 * a12.y, K2 * a12.x);
      const C r1 = mk<C>(p1.x + (q1.x + t1.x), p1.y + (q1.y + t1.y));
      const C p2 = mk<C>(-K0 * a02.y, K0 * a02.x), q2 = mk<C>(-K1 * a12.y, K1 * a12.x),
              t2 = mk<C>(-K2 * a22.y, K2 * a22.x);
      const C r2 = mk<C>(p2.x + (q2.x + t2.x), p2.y + (q2.y + t2.y));
      O[i] = mk<C>(a.nu * r0.x, a.nu * r0.y);
      O[n + i] = mk<C>(a.nu * r1.x, a.nu * r1.y);
      O[2 * n + i] = mk<C>(a.nu * r2.x, a.nu * r2.y);
    } else if constexpr (OP == EW_TRUNCATE) {  // 2/3-rule input truncation of a dealiased backward
      const int kx = ix < a.N0 / 2 ? ix : ix - a.N0, ky = iy < a.N1 / 2 ? iy : iy - a.N1;
      const bool keep = (kx < 0 ? -kx : kx) <= a.keep[0] && (ky < 0 ? -ky : ky) <= a.keep[1] && kz <= a.keep[2];
#pragma unroll
      for (int c = 0; c < 3; ++c) O[c * n + i] = keep ? A[c * n + i] : czero<C>();
    } else if constexpr (OP == EW_SET_ELSASSER) {  // solvers/MHD.py:104-113, a = ZZ_hat[3][3], out = c[6]
      const Rt Kv[3] = {K0, K1, K2};
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        C sp = czero<C>(), sm = czero<C>();
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const C zrj = A[(3 * r + j) * n + i], zjr = A[(3 * j + r) * n + i];
          const C p = mk<C>(Kv[j] * (zrj.x + zjr.x), Kv[j] * (zrj.y + zjr.y));
          const C m = mk<C>(Kv[j] * (zjr.x - zrj.x), Kv[j] * (zjr.y - zrj.y));
          sp = j == 0 ? p : mk<C>(sp.x + p.x, sp.y + p.y);
          sm = j == 0 ? m : mk<C>(sm.x + m.x, sm.y + m.y);
        }
        O[r * n + i] = mk<C>(sp.y / Rt(2), -sp.x / Rt(2));        // -1j*s/2
        O[(3 + r) * n + i] = mk<C>(-sm.y / Rt(2), sm.x / Rt(2));  // 1j*s/2
      }
    } else if constexpr (OP == EW_DIVERGENCE) {  // 1j*(K . a) (NS.py:106-110)
      const C a0 = A[i], a1 = A[n + i], a2 = A[2 * n + i];
      const C d = mk<C>(K0 * a0.x + K1 * a1.x + K2 * a2.x, K0 * a0.y + K1 * a1.y + K2 * a2.y);
      O[i] = cmuli<1>(d);
    } else if constexpr (OP == EW_MASK_NYQUIST) {  // in place on out (NS.py:253-254)
      if (ix == a.N0 / 2 || iy == a.N1 / 2 || kz == a.N2 / 2) {
        O[i] = czero<C>();
        O[n + i] = czero<C>();
        O[2 * n + i] = czero<C>();
      }
    } else if constexpr (OP == EW_PRESSURE_DIFFUSION) {  // NS.py:203-217 on rhs = out, u_hat = a
      Rt k2 = K0 * K0;
      k2 += K1 * K1;
      k2 += K2 * K2;
      const Rt den = k2 == Rt(0) ? Rt(1) : k2;
      const Rt o0 = K0 / den, o1 = K1 / den, o2 = K2 / den;
      C r0 = O[i], r1 = O[n + i], r2 = O[2 * n + i];
      C ph = mk<C>(r0.x * o0 + r1.x * o1, r0.y * o0 + r1.y * o1);
      ph = mk<C>(ph.x + r2.x * o2, ph.y + r2.y * o2);
      r0 = mk<C>(r0.x - ph.x * K0, r0.y - ph.y * K0);
      r1 = mk<C>(r1.x - ph.x * K1, r1.y - ph.y * K1);
      r2 = mk<C>(r2.x - ph.x * K2, r2.y - ph.y * K2);
      const Rt z = a.nu * k2;
      const C s0 = A[i], s1 = A[n + i], s2 = A[2 * n + i];
      O[i] = mk<C>(r0.x - z * s0.x, r0.y - z * s0.y);
      O[n + i] = mk<C>(r1.x - z * s1.x, r1.y - z * s1.y);
      O[2 * n + i] = mk<C>(r2.x - z * s2.x, r2.y - z * s2.y);
      if (a.aux) ((C*)a.aux)[i] = ph;
    } else if constexpr (OP == EW_SGS_FORCE) {
      const int kx = ix < a.N0 / 2 ? ix : ix - a.N0, ky = iy < a.N1 / 2 ? iy : iy - a.N1;
      const bool keep = (kx < 0 ? -kx : kx) <= a.keep[0] && (ky < 0 ? -ky : ky) <= a.keep[1] && kz <= a.keep[2] &&
                        !(a.mask_nyquist && (ix == a.N0 / 2 || iy == a.N1 / 2 || kz == a.N2 / 2));
      const C t00 = A[i], t11 = A[n + i], t22 = A[2 * n + i], t01 = A[3 * n + i], t02 = A[4 * n + i],
              t12 = A[5 * n + i];
      // s_i = k_j a_ij ; f = -i s
      const C s0 = mk<C>(K0 * t00.x + K1 * t01.x + K2 * t02.x, K0 * t00.y + K1 * t01.y + K2 * t02.y);
      const C s1 = mk<C>(K0 * t01.x + K1 * t11.x + K2 * t12.x, K0 * t01.y + K1 * t11.y + K2 * t12.y);
      const C s2 = mk<C>(K0 * t02.x + K1 * t12.x + K2 * t22.x, K0 * t02.y + K1 * t12.y + K2 * t22.y);
      Rt k2 = K0 * K0;
      k2 += K1 * K1;
      k2 += K2 * K2;
      const Rt den = k2 == Rt(0) ? Rt(1) : k2;
      const C ks = mk<C>((K0 * s0.x + K1 * s1.x + K2 * s2.x) / den, (K0 * s0.y + K1 * s1.y + K2 * s2.y) / den);
      const C p0 = mk<C>(s0.x - K0 * ks.x, s0.y - K0 * ks.y);
      const C p1 = mk<C>(s1.x - K1 * ks.x, s1.y - K1 * ks.y);
      const C p2 = mk<C>(s2.x - K2 * ks.x, s2.y - K2 * ks.y);
      O[i] = keep ? mk<C>(p0.y, -p0.x) : czero<C>();
      O[n + i] = keep ? mk<C>(p1.y, -p1.x) : czero<C>();
      O[2 * n + i] = keep ? mk<C>(p2.y, -p2.x) : czero<C>();
    }
  }
}

template <typename Rt> int launch_elementwise(int op, const EwArgs<Rt>& a, hipStream_t s) {
  const dim3 g((unsigned)((a.n + 255) / 256)), b(256);
  switch (op) {
    case EW_CROSS1: hipLaunchKernelGGL((k_elementwise<Rt, EW_CROSS1>), g, b, 0, s, a); return 0;
    case EW_CROSS2: hipLaunchKernelGGL((k_elementwise<Rt, EW_CROSS2>), g, b, 0, s, a); return 0;
    case EW_PRESSURE_DIFFUSION: hipLaunchKernelGGL((k_elementwise<Rt, EW_PRESSURE_DIFFUSION>), g, b, 0, s, a); return 0;
    case EW_DIVERGENCE: hipLaunchKernelGGL((k_elementwise<Rt, EW_DIVERGENCE>), g, b, 0, s, a); return 0;
    case EW_MUL_MINUS_I: hipLaunchKernelGGL((k_elementwise<Rt, EW_MUL_MINUS_I>), g, b, 0, s, a); return 0;
    case EW_MASK_NYQUIST: hipLaunchKernelGGL((k_elementwise<Rt, EW_MASK_NYQUIST>), g, b, 0, s, a); return 0;
    case EW_SGS_FORCE: hipLaunchKernelGGL((k_elementwise<Rt, EW_SGS_FORCE>), g, b, 0, s, a); return 0;
    case EW_CROSS2_KOK: hipLaunchKernelGGL((k_elementwise<Rt, EW_CROSS2_KOK>), g, b, 0, s, a); return 0;
    case EW_ADD_LINEAR: hipLaunchKernelGGL((k_elementwise<Rt, EW_ADD_LINEAR>), g, b, 0, s, a); return 0;
    case EW_TRUNCATE: hipLaunchKernelGGL((k_elementwise<Rt, EW_TRUNCATE>), g, b, 0, s, a); return 0;
    case EW_SET_ELSASSER: hipLaunchKernelGGL((k_elementwise<Rt, EW_SET_ELSASSER>), g, b, 0, s, a); return 0;
    case EW_ELSASSER_PRODUCTS: hipLaunchKernelGGL((k_elementwise<Rt, EW_ELSASSER_PRODUCTS>), g, b, 0, s, a); return 0;
    case EW_GRAD_TRUNC: hipLaunchKernelGGL((k_elementwise<Rt, EW_GRAD_TRUNC>), g, b, 0, s, a); return 0;
    case EW_DOT3: hipLaunchKernelGGL((k_elementwise<Rt, EW_DOT3>), g, b, 0, s, a); return 0;
    case EW_UU_PRODUCTS: hipLaunchKernelGGL((k_elementwise<Rt, EW_UU_PRODUCTS>), g, b, 0, s, a); return 0;
    case EW_DIV_UU: hipLaunchKernelGGL((k_elementwise<Rt, EW_DIV_UU>), g, b, 0, s, a); return 0;
    case EW_VV_TAIL: hipLaunchKernelGGL((k_elementwise<Rt, EW_VV_TAIL>), g, b, 0, s, a); return 0;
    default: return -1;
  }
}

// energy_fourier of u_hat and of curl_hat = i K x u_hat, FP64 partial sums per block
template <typename Rt>
__global__ void __launch_bounds__(256) k_energy(EwArgs<Rt> a, double* __restrict__ part) {
  using C = cx<Rt>;
  const long n = a.n;
  double es = 0, os = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int kz = (int)(i % a.Nh);
    const long r = i / a.Nh;
    const int iy = (int)(r % a.N1);
    const int ix = (int)(r / a.N1);
    const double k0 = a.kx[ix], k1 = a.ky[iy], k2 = a.kz[kz];
    const C* A = (const C*)a.a;
    const C u0 = A[i], u1 = A[n + i], u2 = A[2 * n + i];
    const double w = (kz == 0 || kz == a.N2 / 2) ? 1.0 : 2.0;
    const double x0 = u0.x, y0 = u0.y, x1 = u1.x, y1 = u1.y, x2 = u2.x, y2 = u2.y;
    es += w * (x0 * x0 + y0 * y0 + x1 * x1 + y1 * y1 + x2 * x2 + y2 * y2);
    const double ax = k1 * x2 - k2 * x1, ay = k1 * y2 - k2 * y1;
    const double bx = k2 * x0 - k0 * x2, by = k2 * y0 - k0 * y2;
    const double cxv = k0 * x1 - k1 * x0, cyv = k0 * y1 - k1 * y0;
    os += w * (ax * ax + ay * ay + bx * bx + by * by + cxv * cxv + cyv * cyv);
  }
  __shared__ double red[2][4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    es += __shfl_down(es, off, 64);
    os += __shfl_down(os, off, 64);
  }
  if (threadIdx.x % 64 == 0) { red[0][threadIdx.x / 64] = es; red[1][threadIdx.x / 64] = os; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

template <typename Rt> int launch_energy(const EwArgs<Rt>& a, double* partials, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_energy<Rt>), dim3(nblk), dim3(256), 0, s, a, partials);
  return 0;
}

// ------------------------------------------------------------------ forcing
template <typename Rt>
__global__ void __launch_bounds__(256) k_force_reduce(EwArgs<Rt> a, Rt kf2, double* __restrict__ part) {
  using C = cx<Rt>;
  const long n = a.n;
  double s_all = 0, s_low = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    if (i == 0) continue;  // the mean mode is set to zero first (Isotropic.py:163-164)
    const int kz = (int)(i % a.Nh);
    const long r = i / a.Nh;
    const int iy = (int)(r % a.N1);
    const int ix = (int)(r / a.N1);
    const Rt K0 = a.kx[ix], K1 = a.ky[iy], K2 = a.kz[kz];
    Rt k2 = K0 * K0;
    k2 += K1 * K1;
    k2 += K2 * K2;
    const C* A = (const C*)a.a;
    const C* Lo = (const C*)a.b;  // low part of a compensated state, or null
    double v[6] = {A[i].x, A[i].y, A[n + i].x, A[n + i].y, A[2 * n + i].x, A[2 * n + i].y};
    if (Lo) {
      v[0] += Lo[i].x; v[1] += Lo[i].y; v[2] += Lo[n + i].x;
      v[3] += Lo[n + i].y; v[4] += Lo[2 * n + i].x; v[5] += Lo[2 * n + i].y;
    }
    const double w = (kz == 0 || kz == a.N2 / 2) ? 1.0 : 2.0;
    const double e = w * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    s_all += e;
    if (k2 <= kf2) s_low += e;
  }
  __shared__ double red[2][4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s_all += __shfl_down(s_all, off, 64);
    s_low += __shfl_down(s_low, off, 64);
  }
  if (threadIdx.x % 64 == 0) { red[0][threadIdx.x / 64] = s_all; red[1][threadIdx.x / 64] = s_low; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

template <typename Rt>
__global__ void __launch_bounds__(256) k_force_scale(EwArgs<Rt> a, Rt kf2, const double* __restrict__ sums,
                                                    double target) {
  using C = cx<Rt>;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const long n = a.n;
  C* U = (C*)a.out;
  C* Lo = (C*)a.aux;  // low part of a compensated state, or null
  if (i == 0) {
    U[0] = czero<C>(); U[n] = czero<C>(); U[2 * n] = czero<C>();
    if (Lo) { Lo[0] = czero<C>(); Lo[n] = czero<C>(); Lo[2 * n] = czero<C>(); }
    return;
  }
  const int kz = (int)(i % a.Nh);
  const long r = i / a.Nh;
  const int iy = (int)(r % a.N1);
  const int ix = (int)(r / a.N1);
  const Rt K0 = a.kx[ix], K1 = a.ky[iy], K2 = a.kz[kz];
  Rt k2 = K0 * K0;
  k2 += K1 * K1;
  k2 += K2 * K2;
  if (!(k2 <= kf2)) return;
  const double lower = sums[1], upper = sums[0] - sums[1];
  if (Lo) {
    // the pair (hi, lo) scaled as one FP64 value, then split again
    const double al = sqrt((target - upper) / lower);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long j = c * n + i;
      const double x = al * ((double)U[j].x + (double)Lo[j].x), y = al * ((double)U[j].y + (double)Lo[j].y);
      const C h = mk<C>((Rt)x, (Rt)y);
      U[j] = h;
      Lo[j] = mk<C>((Rt)(x - (double)h.x), (Rt)(y - (double)h.y));
    }
    return;
  }
  const Rt alpha = (Rt)sqrt((target - upper) / lower);
  U[i] = cscale(U[i], alpha);
  U[n + i] = cscale(U[n + i], alpha);
  U[2 * n + i] = cscale(U[2 * n + i], alpha);
}

template <typename Rt> int launch_force_reduce(const EwArgs<Rt>& a, Rt kf2, double* partials, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_force_reduce<Rt>), dim3(nblk), dim3(256), 0, s, a, kf2, partials);
  return 0;
}
template <typename Rt> int launch_force_scale(const EwArgs<Rt>& a, Rt kf2, const double* sums, double target,
                                              hipStream_t s) {
  hipLaunchKernelGGL((k_force_scale<Rt>), dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a, kf2, sums, target);
  return 0;
}

// ------------------------------------------------------------------ spectrum
template <typename Rt>
__global__ void __launch_bounds__(256) k_spectrum(EwArgs<Rt> a, int nb, double* __restrict__ part) {
  using C = cx<Rt>;
  extern __shared__ double hist[];  // [2][nb]: uiui sums, counts
  for (int b = threadIdx.x; b < 2 * nb; b += 256) hist[b] = 0.0;
  __syncthreads();
  const long n = a.n;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int kz = (int)(i % a.Nh);
    const long r = i / a.Nh;
    const int iy = (int)(r % a.N1);
    const int ix = (int)(r / a.N1);
    const Rt K0 = a.kx[ix], K1 = a.ky[iy], K2 = a.kz[kz];
    Rt k2 = K0 * K0;
    k2 += K1 * K1;
    k2 += K2 * K2;
    const double x = (double)sqrt(k2);
    // np.digitize(x, bins=[0.5, 1.5, ...], right=True): number of bins strictly below x
    int z = x > 0.5 ? (int)ceil(x - 0.5) : 0;
    if (z < 1 || z > nb) continue;
    const C* A = (const C*)a.a;
    const C u0 = A[i], u1 = A[n + i], u2 = A[2 * n + i];
    const double w = (kz == 0 || kz == a.Nh - 1) ? 1.0 : 2.0;
    const double e = w * ((double)u0.x * u0.x + (double)u0.y * u0.y + (double)u1.x * u1.x + (double)u1.y * u1.y +
                          (double)u2.x * u2.x + (double)u2.y * u2.y);
    atomicAdd(&hist[z - 1], e);
    atomicAdd(&hist[nb + z - 1], 1.0);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 2 * nb; b += 256) part[(long)blockIdx.x * 2 * nb + b] = hist[b];
}

template <typename Rt> int launch_spectrum(const EwArgs<Rt>& a, int nb, double* partials, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_spectrum<Rt>), dim3(nblk), dim3(256), 2 * nb * sizeof(double), s, a, nb, partials);
  return 0;
}

// ------------------------------------------------------------------ integrator updates
template <typename Rt> __global__ void __launch_bounds__(256) k_lincomb(LinArgs<Rt> a) {
#pragma clang fp contract(off)
  using C = cx<Rt>;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (long)gridDim.x * 256) {
    const C x0 = ((const C*)a.x[0])[i];
    C s = mk<C>(a.c[0] * x0.x, a.c[0] * x0.y);
#pragma unroll
    for (int j = 1; j < SDNS_LIN_MAX; ++j) {
      if (j < a.nterms) {
        const C xj = ((const C*)a.x[j])[i];
        const Rt px = a.c[j] * xj.x, py = a.c[j] * xj.y;
        s = mk<C>(s.x + px, s.y + py);
      }
    }
    ((C*)a.out)[i] = s;
  }
}

template <typename Rt> int launch_lincomb(const LinArgs<Rt>& a, hipStream_t s) {
  if (a.nterms < 1 || a.nterms > SDNS_LIN_MAX) return -1;
  const long nblk = std::min<long>((a.n + 255) / 256, 8192);
  hipLaunchKernelGGL((k_lincomb<Rt>), dim3((unsigned)std::max<long>(nblk, 1)), dim3(256), 0, s, a);
  return 0;
}

template <typename Rt>
__global__ void __launch_bounds__(256) k_errnorm(const void* err, const void* u0, const void* u1, Rt atol, Rt rtol,
                                                 long n, double* __restrict__ part) {
#pragma clang fp contract(off)
  using C = cx<Rt>;
  const int comp = blockIdx.y;
  const C* E = (const C*)err + comp * n;
  const C* A = (const C*)u0 + comp * n;
  const C* B = (const C*)u1 + comp * n;
  double acc = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const C a = A[i], b = B[i], e = E[i];
    const Rt ma = hypot(a.x, a.y), mb = hypot(b.x, b.y);
    const Rt sc = atol + (ma > mb ? ma : mb) * rtol;
    const Rt m = hypot(e.x / sc, e.y / sc);
    acc += (double)(m * m);
  }
  __shared__ double red[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[3 * blockIdx.x + comp] = red[0] + red[1] + red[2] + red[3];
}

template <typename Rt> int launch_errnorm(const void* err, const void* u0, const void* u1, Rt atol, Rt rtol,
                                          long n, double* partials, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_errnorm<Rt>), dim3(nblk, 3), dim3(256), 0, s, err, u0, u1, atol, rtol, n, partials);
  return 0;
}

}  // namespace sdns

#define SDNS_INSTANTIATE(Rt)                                                                     \
  namespace sdns {                                                                               \
  template int launch_p1<Rt>(int, int, const P1Args<Rt>&, hipStream_t);                          \
  template int launch_p2<Rt>(int, int, const P2Args<Rt>&, hipStream_t);                          \
  template int launch_p3<Rt>(int, int, const P3Args<Rt>&, hipStream_t);                          \
  template int launch_p5<Rt>(int, int, int, const P5Args<Rt>&, hipStream_t, int*);               \
  template int launch_strided<Rt>(int, int, int, int, const StridedArgs<Rt>&, hipStream_t);      \
  template int launch_z<Rt>(int, int, int, int, const ZArgs<Rt>&, hipStream_t);                  \
  template int launch_elementwise<Rt>(int, const EwArgs<Rt>&, hipStream_t);                      \
  template int launch_energy<Rt>(const EwArgs<Rt>&, double*, int, hipStream_t);                  \
  template int launch_force_reduce<Rt>(const EwArgs<Rt>&, Rt, double*, int, hipStream_t);        \
  template int launch_force_scale<Rt>(const EwArgs<Rt>&, Rt, const double*, double, hipStream_t); \
  template int launch_spectrum<Rt>(const EwArgs<Rt>&, int, double*, int, hipStream_t);            \
  template int launch_lincomb<Rt>(const LinArgs<Rt>&, hipStream_t);                               \
  template int launch_errnorm<Rt>(const void*, const void*, const void*, Rt, Rt, long, double*, int, hipStream_t); \
  }
