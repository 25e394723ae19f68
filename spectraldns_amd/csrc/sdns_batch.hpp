// Batched uneven-size ensembles: every pass of the RK4 step is ONE launch
// over all boxes of the ensemble.  Each block reads its (box, tile) work item,
// loads that box's argument block and runs the pass body compiled for the
// box's line length (the same bodies as the single-box kernels).  The block
// size does not depend on the box, so the tile per length is fixed (see
// batch_T / batch_LPB).  LMAX is the largest line length compiled into a
// kernel: an ensemble of small boxes gets a kernel whose register budget is
// not set by the long-line schedules.
#pragma once
#include "sdns_kernels.hpp"
#include "sdns_launch.hpp"

namespace sdns {

// Cases of the strided-pass length switch: line length m with schedule
// variant 0 (case m) and, where it differs, variant 1 (case m + BATCH_VSTRIDE),
// tile T.
#define SDNS_BATCH_CASE_V(m, V, ...)                          \
  case m + V * BATCH_VSTRIDE:                                 \
    if constexpr (m <= LMAX && (V == 0 || batch_E(m, 1) != batch_E(m, 0))) { \
      constexpr int E = batch_E(m, V), T = batch_T<Rt>(m, V); \
      C* tw = smem + m * T;                                   \
      table_to_lds(tw, (const C*)a.tw, m);                    \
      __VA_ARGS__;                                            \
    }                                                         \
    break;
#define SDNS_BATCH_CASE(m, ...) SDNS_BATCH_CASE_V(m, 0, __VA_ARGS__) SDNS_BATCH_CASE_V(m, 1, __VA_ARGS__)

template <typename Rt, int LMAX>
__global__ void __launch_bounds__(batch_threads<Rt>())
    k_batch_p1(const P1Args<Rt>* __restrict__ args, const int* __restrict__ len, const BatchWork* __restrict__ work) {
  using C = cx<Rt>;
  __shared__ C smem[batch_strided_lds<Rt>(LMAX)];
  const BatchWork w = work[blockIdx.x];
  if (w.box < 0) return;
  const P1Args<Rt> a = args[w.box];
  switch (len[w.box]) {
#define SDNS_X(m) SDNS_BATCH_CASE(m, p1_body<Rt, m, E, T>(a, w.tile, blockIdx.y, smem, tw))
    SDNS_BATCH_STRIDED(SDNS_X)
#undef SDNS_X
    default: break;
  }
}

template <typename Rt, int LMAX>
__global__ void __launch_bounds__(batch_threads<Rt>())
    k_batch_p2(const P2Args<Rt>* __restrict__ args, const int* __restrict__ len, const BatchWork* __restrict__ work) {
  using C = cx<Rt>;
  __shared__ C smem[batch_strided_lds<Rt>(LMAX)];
  const BatchWork w = work[blockIdx.x];
  if (w.box < 0) return;
  const P2Args<Rt> a = args[w.box];
  switch (len[w.box]) {
#define SDNS_X(m) SDNS_BATCH_CASE(m, p2_body<Rt, m, E, T>(a, w.tile, blockIdx.y, smem, tw))
    SDNS_BATCH_STRIDED(SDNS_X)
#undef SDNS_X
    default: break;
  }
}

template <typename Rt, int LMAX>
__global__ void __launch_bounds__(batch_threads<Rt>())
    k_batch_fwd(const StridedArgs<Rt>* __restrict__ args, const int* __restrict__ len,
                const BatchWork* __restrict__ work) {
  using C = cx<Rt>;
  __shared__ C smem[batch_strided_lds<Rt>(LMAX)];
  const BatchWork w = work[blockIdx.x];
  if (w.box < 0) return;
  const StridedArgs<Rt> a = args[w.box];
  switch (len[w.box]) {
#define SDNS_X(m) SDNS_BATCH_CASE(m, strided_body<Rt, m, E, -1, T>(a, w.tile, blockIdx.y, smem, tw))
    SDNS_BATCH_STRIDED(SDNS_X)
#undef SDNS_X
    default: break;
  }
}

template <typename Rt, int MODE, int LMAX>
__global__ void __launch_bounds__(batch_threads<Rt>())
    k_batch_p5(const P5Args<Rt>* __restrict__ args, const int* __restrict__ len, const BatchWork* __restrict__ work) {
  using C = cx<Rt>;
  __shared__ C smem[batch_strided_lds<Rt>(LMAX)];
  const BatchWork w = work[blockIdx.x];
  if (w.box < 0) return;
  const P5Args<Rt> a = args[w.box];
  switch (len[w.box]) {
#define SDNS_X(m) SDNS_BATCH_CASE(m, p5_body<Rt, m, E, MODE, false, T>(a, w.tile, smem, tw))
    SDNS_BATCH_STRIDED(SDNS_X)
#undef SDNS_X
    default: break;
  }
}
#undef SDNS_BATCH_CASE

// z pass: C2R of six fields, u x curl(u), R2C of three (the LDS-assisted
// scheme of the single-box z pass), LPB lines per block.
template <typename Rt, int LMAX>
__global__ void __launch_bounds__(batch_zthreads<Rt>())
    k_batch_p3(const P3Args<Rt>* __restrict__ args, const int* __restrict__ len, const BatchWork* __restrict__ work) {
  using C = cx<Rt>;
  __shared__ C smem[batch_z_lds<Rt>(LMAX)];
  const BatchWork w = work[blockIdx.x];
  if (w.box < 0) return;
  const P3Args<Rt> a = args[w.box];
  switch (len[w.box]) {
#define SDNS_ZV(n, V)                                                    \
  case n + V * BATCH_VSTRIDE:                                            \
    if constexpr (n <= LMAX && (V == 0 || batch_zE(n, 1) != batch_zE(n, 0))) { \
      constexpr int E = batch_zE(n, V), LPB = batch_LPB<Rt>(n, V);       \
      C* tw = smem + 3 * LPB * z3_slot(n, n / E, (int)sizeof(C));        \
      C* rtw = tw + n;                                                   \
      table_to_lds(tw, (const C*)a.tw, n);                               \
      table_to_lds(rtw, (const C*)a.rtw, n);                             \
      p3b_body<Rt, n, E, LPB>(a, w.tile * LPB, smem, tw, rtw);           \
    }                                                                    \
    break;
#define SDNS_X(n) SDNS_ZV(n, 0) SDNS_ZV(n, 1)
    SDNS_BATCH_HALF(SDNS_X)
#undef SDNS_X
#undef SDNS_ZV
    default: break;
  }
}

// ------------------------------------------------------------------ launchers
// Strided kernels are compiled for LMAX = 64 and 128, the z kernel for 32 and
// 64; `lmax` is the largest length of the ensemble for that pass.
template <typename Rt> int launch_batch_p1(int lmax, const P1Args<Rt>* args, const int* len, const BatchWork* work,
                                           int nwork, hipStream_t s) {
  const dim3 g(nwork, 3), b(batch_threads<Rt>());
  if (lmax <= 64) hipLaunchKernelGGL((k_batch_p1<Rt, 64>), g, b, 0, s, args, len, work);
  else hipLaunchKernelGGL((k_batch_p1<Rt, 128>), g, b, 0, s, args, len, work);
  return 0;
}
template <typename Rt> int launch_batch_p2(int lmax, const P2Args<Rt>* args, const int* len, const BatchWork* work,
                                           int nwork, hipStream_t s) {
  const dim3 g(nwork, 2), b(batch_threads<Rt>());
  if (lmax <= 64) hipLaunchKernelGGL((k_batch_p2<Rt, 64>), g, b, 0, s, args, len, work);
  else hipLaunchKernelGGL((k_batch_p2<Rt, 128>), g, b, 0, s, args, len, work);
  return 0;
}
template <typename Rt> int launch_batch_p3(int lmax, const P3Args<Rt>* args, const int* len, const BatchWork* work,
                                           int nwork, hipStream_t s) {
  const dim3 g(nwork), b(batch_zthreads<Rt>());
  if (lmax <= 32) hipLaunchKernelGGL((k_batch_p3<Rt, 32>), g, b, 0, s, args, len, work);
  else hipLaunchKernelGGL((k_batch_p3<Rt, 64>), g, b, 0, s, args, len, work);
  return 0;
}
template <typename Rt> int launch_batch_fwd(int lmax, const StridedArgs<Rt>* args, const int* len,
                                            const BatchWork* work, int nwork, int nfields, hipStream_t s) {
  const dim3 g(nwork, nfields), b(batch_threads<Rt>());
  if (lmax <= 64) hipLaunchKernelGGL((k_batch_fwd<Rt, 64>), g, b, 0, s, args, len, work);
  else hipLaunchKernelGGL((k_batch_fwd<Rt, 128>), g, b, 0, s, args, len, work);
  return 0;
}
template <typename Rt, int LMAX>
void batch_p5_go(int mode, const P5Args<Rt>* args, const int* len, const BatchWork* work, int nwork, hipStream_t s) {
  const dim3 g(nwork), b(batch_threads<Rt>());
  switch (mode) {
    case P5_RK0: hipLaunchKernelGGL((k_batch_p5<Rt, P5_RK0, LMAX>), g, b, 0, s, args, len, work); break;
    case P5_RK12: hipLaunchKernelGGL((k_batch_p5<Rt, P5_RK12, LMAX>), g, b, 0, s, args, len, work); break;
    default: hipLaunchKernelGGL((k_batch_p5<Rt, P5_RK3, LMAX>), g, b, 0, s, args, len, work); break;
  }
}
template <typename Rt> int launch_batch_p5(int lmax, int mode, const P5Args<Rt>* args, const int* len,
                                           const BatchWork* work, int nwork, hipStream_t s) {
  if (mode != P5_RK0 && mode != P5_RK12 && mode != P5_RK3) return -1;
  if (lmax <= 64) batch_p5_go<Rt, 64>(mode, args, len, work, nwork, s);
  else batch_p5_go<Rt, 128>(mode, args, len, work, nwork, s);
  return 0;
}

}  // namespace sdns

#define SDNS_BATCH_INSTANTIATE(Rt)                                                                                   \
  namespace sdns {                                                                                                   \
  template int launch_batch_p1<Rt>(int, const P1Args<Rt>*, const int*, const BatchWork*, int, hipStream_t);          \
  template int launch_batch_p2<Rt>(int, const P2Args<Rt>*, const int*, const BatchWork*, int, hipStream_t);          \
  template int launch_batch_p3<Rt>(int, const P3Args<Rt>*, const int*, const BatchWork*, int, hipStream_t);          \
  template int launch_batch_fwd<Rt>(int, const StridedArgs<Rt>*, const int*, const BatchWork*, int, int, hipStream_t); \
  template int launch_batch_p5<Rt>(int, int, const P5Args<Rt>*, const int*, const BatchWork*, int, hipStream_t);     \
  }
