// Learned subgrid closure kernels (declared in sdns_closure.hpp).
//
// The 3x3x3 periodic convolution is an implicit GEMM on the bf16 matrix cores:
// v_mfma_f32_16x16x32_bf16 multiplies A (16 x 32) by B (32 x 16) into a 16 x 16
// FP32 tile.  Lane l (0..63) holds A[l % 16][8 (l / 16) + j], B[8 (l / 16) + j][l % 16]
// (j = 0..7) and D[4 (l / 16) + i][l % 16] (i = 0..3).  In the forward
// product A rows are 16 consecutive output points of one z-row, K is the input
// channel index (32 channels per tap, or 4 taps x 8 channels for 8-channel
// inputs) and B columns are output channels; channels-last storage makes
// every A and B fragment one 16-byte load per lane.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sdns_closure.hpp"

namespace sdns {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ------------------------------------------------------------------ convolution
// Block: 4 waves, 64 output points (4 x 16 consecutive points of one z-row).
template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  constexpr int NB = COUT / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const long p0 = (long)blockIdx.x * 64 + wave * 16;  // first point of this wave
  const int i2_0 = (int)(p0 % a.n2);
  const long r = p0 / a.n2;
  const int i1 = (int)(r % a.n1), i0 = (int)(r / a.n1);
  const int i2 = i2_0 + m;  // this lane's A row (point)
  f32x4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8* __restrict__ X = (const bf16x8*)a.x;
  const bf16x8* __restrict__ W = (const bf16x8*)a.w;
  if constexpr (CIN == 32) {
    for (int tz = 0; tz < 3; ++tz) {
      const int j0 = wrap(i0 + tz - 1, a.n0);
      for (int ty = 0; ty < 3; ++ty) {
        const long row = ((long)j0 * a.n1 + wrap(i1 + ty - 1, a.n1)) * a.n2;
#pragma unroll
        for (int tx = 0; tx < 3; ++tx) {
          const int tap = 9 * tz + 3 * ty + tx;
          const bf16x8 av = X[(row + wrap(i2 + tx - 1, a.n2)) * 4 + g];
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const bf16x8 bv = W[((long)tap * COUT + b * 16 + m) * 4 + g];
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[b], 0, 0, 0);
          }
        }
      }
    }
  } else {
    static_assert(CIN == 8, "8 or 32 input channels");
    // K = 32 = 4 taps x 8 channels: k-group g of MFMA s is tap 4 s + g
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      const int tap = 4 * s + g;
      bf16x8 av = {}, bv[NB];
      if (tap < 27) {
        const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
        const long q = ((long)wrap(i0 + tz - 1, a.n0) * a.n1 + wrap(i1 + ty - 1, a.n1)) * a.n2 + wrap(i2 + tx - 1, a.n2);
        av = X[q];
#pragma unroll
        for (int b = 0; b < NB; ++b) bv[b] = W[(long)tap * COUT + b * 16 + m];
      } else {
#pragma unroll
        for (int b = 0; b < NB; ++b) bv[b] = bf16x8{};
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[b], acc[b], 0, 0, 0);
    }
  }
  // epilogue: D[4 g + i][m] -> point p0 + 4 g + i, channel 16 b + m
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int co = b * 16 + m;
    const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long pt = p0 + 4 * g + i;
      const float v = acc[b][i] + bias;
      if constexpr (EPI == CONV_EPI_SILU) {
        const __bf16 zb = (__bf16)v;
        if (a.z) a.z[pt * COUT + co] = zb;
        a.a[pt * COUT + co] = (__bf16)silu_f((float)zb);
      } else if constexpr (EPI == CONV_EPI_LINEAR) {
        if (co < a.ncout) a.y[pt * a.ystride + co] = v;
      } else {
        a.a[pt * COUT + co] = (__bf16)(v * dsilu_f((float)a.zprev[pt * COUT + co]));
      }
    }
  }
}

// LDS-tiled convolution.  A block owns 4 consecutive y-rows x tzn = min(N2, 64)
// consecutive z points over CONV_XP consecutive x-planes, and marches along
// x: the layer's weights (all 27 taps) stay in LDS, and the periodic input
// halo is a rolling window of x-planes (6 rows x (tzn + 2) points each) in 4
// LDS slots, the plane after next loaded into registers while the current
// output plane is computed and stored into its free slot afterwards.  Wave w
// computes row w of the current plane: tzn / 16 point tiles x COUT / 16
// channel tiles, each weight fragment read once per K-step and used for every
// point tile.  The MFMA sequence per output is the one of k_conv (taps in the
// same order, same K split), so the results are the same bit for bit.
// Needs N0 % 4 == 0, N1 % 4 == 0, N2 % 16 == 0 and N2 <= 64 or N2 % 64 == 0.
constexpr int CONV_ROWS = 4, CONV_XP = 4, CONV_SLOTS = 4;
inline int conv_tzn(int n2) { return n2 <= 64 ? n2 : 64; }
template <int CIN, int COUT> constexpr int conv_w_bytes() { return 27 * COUT * CIN * 2; }
template <int CIN> constexpr int conv_plane_chunks(int tzn) { return (CONV_ROWS + 2) * (tzn + 2) * CIN / 8; }

// A fragment (16 points x 8 channels per lane) of point tile t and the B
// fragments of K-step ks for output row w; pl[tz] is the LDS slot of input
// plane i0 - 1 + tz.  CIN = 32: one tap per K-step; CIN = 8: taps 4 ks ..
// 4 ks + 3, one per lane group g (lanes past the last tap multiply zeros).
template <int CIN, int COUT, int MT>
__device__ __forceinline__ void conv_frags(int ks, const bf16x8* wl, const bf16x8* const (&pl)[3], int w, int m,
                                           int g, bf16x8 (&av)[MT], bf16x8 (&bv)[COUT / 16]) {
  constexpr int NB = COUT / 16, npt = 16 * MT + 2;
  if constexpr (CIN == 32) {
    const int tz = ks / 9, ty = (ks / 3) % 3, tx = ks % 3;
#pragma unroll
    for (int b = 0; b < NB; ++b) bv[b] = wl[(ks * COUT + b * 16 + m) * 4 + g];
    const bf16x8* h = pl[tz] + ((w + ty) * npt + m + tx) * 4 + g;
#pragma unroll
    for (int t = 0; t < MT; ++t) av[t] = h[64 * t];
  } else {
    const bool ok = 4 * ks + g < 27;
    const int tap = ok ? 4 * ks + g : 26;
    const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
#pragma unroll
    for (int b = 0; b < NB; ++b) bv[b] = ok ? wl[tap * COUT + b * 16 + m] : bf16x8{};
    const bf16x8* h = pl[tz] + (w + ty) * npt + m + tx;
#pragma unroll
    for (int t = 0; t < MT; ++t) av[t] = ok ? h[16 * t] : bf16x8{};
  }
}

template <int CIN, int COUT, int EPI, int MT>
__global__ void __launch_bounds__(256) k_conv_lds(ConvArgs a) {
  constexpr int NB = COUT / 16, CG = CIN / 8, KS = CIN == 32 ? 27 : 7;
  constexpr int tzn = 16 * MT, npt = tzn + 2, PC = conv_plane_chunks<CIN>(tzn);
  constexpr int PR = (PC + 255) / 256;  // prefetch registers (16-byte chunks) per thread
  extern __shared__ __align__(16) unsigned char conv_smem[];
  bf16x8* wl = (bf16x8*)conv_smem;
  bf16x8* hl = (bf16x8*)(conv_smem + conv_w_bytes<CIN, COUT>());
  const int n0 = a.n0, n1 = a.n1, n2 = a.n2;
  const int nzb = n2 / tzn, nrb = n1 / CONV_ROWS;
  int blk = blockIdx.x;
  const int zb = blk % nzb;
  blk /= nzb;
  const int rb = blk % nrb, xc = blk / nrb;
  const int i1b = rb * CONV_ROWS, z0 = zb * tzn, x0 = xc * CONV_XP;
  const bf16x8* __restrict__ X = (const bf16x8*)a.x;
  const bf16x8* __restrict__ W = (const bf16x8*)a.w;
  // chunk c of an input plane: (row, point, channel group), row-major
  auto load_plane = [&](int j, bf16x8 (&r)[PR]) {
    const long pbase = (long)wrap(x0 + j, n0) * n1;
#pragma unroll
    for (int e = 0; e < PR; ++e) {
      const int c = threadIdx.x + 256 * e;
      if (c < PC) {
        const int cg = c % CG, q = c / CG, pt = q % npt, row = q / npt;
        r[e] = X[((pbase + wrap(i1b + row - 1, n1)) * n2 + wrap(z0 + pt - 1, n2)) * CG + cg];
      }
    }
  };
  auto store_plane = [&](int j, const bf16x8 (&r)[PR]) {
    bf16x8* dst = hl + ((j + 1) % CONV_SLOTS) * PC;
#pragma unroll
    for (int e = 0; e < PR; ++e) {
      const int c = threadIdx.x + 256 * e;
      if (c < PC) dst[c] = r[e];
    }
  };
  for (int c = threadIdx.x; c < 27 * COUT * CG; c += 256) wl[c] = W[c];
  bf16x8 pre[PR];
  for (int j = -1; j <= 1; ++j) {  // planes x0 - 1, x0, x0 + 1
    load_plane(j, pre);
    store_plane(j, pre);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  for (int k = 0; k < CONV_XP; ++k) {
    if (k + 2 <= CONV_XP) load_plane(k + 2, pre);  // in flight during this plane's MFMAs
    const bf16x8* const pl[3] = {hl + (k % CONV_SLOTS) * PC, hl + ((k + 1) % CONV_SLOTS) * PC,
                                 hl + ((k + 2) % CONV_SLOTS) * PC};
    f32x4 acc[MT][NB];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[t][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 av[MT], bv[NB], an[MT], bn[NB];
    conv_frags<CIN, COUT, MT>(0, wl, pl, w, m, g, av, bv);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if (ks + 1 < KS) conv_frags<CIN, COUT, MT>(ks + 1, wl, pl, w, m, g, an, bn);
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[t][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv[b], acc[t][b], 0, 0, 0);
      if (ks + 1 < KS) {
#pragma unroll
        for (int t = 0; t < MT; ++t) av[t] = an[t];
#pragma unroll
        for (int b = 0; b < NB; ++b) bv[b] = bn[b];
      }
    }
    // epilogue: D[4 g + i][m] of tile t -> point (x0 + k, i1b + w, z0 + 16 t + 4 g + i), channel 16 b + m
    const long prow = ((long)(x0 + k) * n1 + i1b + w) * n2 + z0;
#pragma unroll
    for (int t = 0; t < MT; ++t) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int co = b * 16 + m;
        const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const long pt = prow + 16 * t + 4 * g + i;
          const float v = acc[t][b][i] + bias;
          if constexpr (EPI == CONV_EPI_SILU) {
            const __bf16 zb = (__bf16)v;
            if (a.z) a.z[pt * COUT + co] = zb;
            a.a[pt * COUT + co] = (__bf16)silu_f((float)zb);
          } else if constexpr (EPI == CONV_EPI_LINEAR) {
            if (co < a.ncout) a.y[pt * a.ystride + co] = v;
          } else {
            a.a[pt * COUT + co] = (__bf16)(v * dsilu_f((float)a.zprev[pt * COUT + co]));
          }
        }
      }
    }
    // the slot of plane k + 2 held plane k - 2, which no wave reads any more
    if (k + 2 <= CONV_XP) store_plane(k + 2, pre);
    __syncthreads();
  }
}

template <int CIN, int COUT, int EPI, int MT> int conv_lds_go(const ConvArgs& a, hipStream_t s) {
  const size_t bytes = (size_t)conv_w_bytes<CIN, COUT>() + (size_t)CONV_SLOTS * conv_plane_chunks<CIN>(16 * MT) * 16;
  // dynamic LDS beyond 64 KB must be allowed for the kernel (on the current device)
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv_lds<CIN, COUT, EPI, MT>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return -1;
  const unsigned nblk = (unsigned)((long)(a.n0 / CONV_XP) * (a.n1 / CONV_ROWS) * (a.n2 / (16 * MT)));
  hipLaunchKernelGGL((k_conv_lds<CIN, COUT, EPI, MT>), dim3(nblk), dim3(256), bytes, s, a);
  return 0;
}
template <int CIN, int COUT, int EPI> int conv_lds_go(const ConvArgs& a, hipStream_t s) {
  switch (conv_tzn(a.n2) / 16) {
    case 1: return conv_lds_go<CIN, COUT, EPI, 1>(a, s);
    case 2: return conv_lds_go<CIN, COUT, EPI, 2>(a, s);
    case 3: return conv_lds_go<CIN, COUT, EPI, 3>(a, s);
    default: return conv_lds_go<CIN, COUT, EPI, 4>(a, s);
  }
}

int launch_conv(int cin, int cout, int epi, const ConvArgs& a, hipStream_t s) {
  const long P = (long)a.n0 * a.n1 * a.n2;
  if (a.n2 % 16 || P % 64) return -1;
  // 32-channel inputs: the LDS-tiled kernel; 8-channel inputs (a quarter of the
  // MFMAs of a 32-channel layer) run faster on k_conv's many small blocks
  const bool tiled = cin == 32 && a.n0 % CONV_XP == 0 && a.n1 % CONV_ROWS == 0 && (a.n2 <= 64 || a.n2 % 64 == 0);
  const dim3 grid((unsigned)(P / 64)), blk(256);
#define SDNS_CONV(CI, CO, EP)                                                                       \
  if (cin == CI && cout == CO && epi == EP) {                                                       \
    if (tiled) return conv_lds_go<CI, CO, EP>(a, s);                                                \
    hipLaunchKernelGGL((k_conv<CI, CO, EP>), grid, blk, 0, s, a);                                   \
    return 0;                                                                                       \
  }
  SDNS_CONV(8, 32, CONV_EPI_SILU)      // layer 1 forward
  SDNS_CONV(32, 32, CONV_EPI_SILU)     // layer 2 forward
  SDNS_CONV(32, 16, CONV_EPI_LINEAR)   // layer 3 forward
  SDNS_CONV(8, 32, CONV_EPI_DSILU)     // layer 3 input gradient -> layer 2 pre-activation gradient
  SDNS_CONV(32, 32, CONV_EPI_DSILU)    // layer 2 input gradient -> layer 1 pre-activation gradient
#undef SDNS_CONV
  return -1;
}

// ------------------------------------------------------------------ weight gradient
// dW[tap][co][ci] = sum_p g[p][co] x[p + d(tap)][ci]: M = co, N = ci, K = points.
// Block (chunk, tap): 4 waves over the chunk's 4096 points in K-steps of 32
// points (one z-row segment); wave partial tiles are summed in LDS.
constexpr int WG_CHUNK = 4096;
template <int GC, int XC>
__global__ void __launch_bounds__(256) k_wgrad(const __bf16* __restrict__ gp, const __bf16* __restrict__ xp,
                                               float* __restrict__ part, int n0, int n1, int n2) {
  constexpr int MB = (GC + 15) / 16, NBK = (XC + 15) / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int tap = blockIdx.y;
  const int tz = tap / 9 - 1, ty = (tap / 3) % 3 - 1, tx = tap % 3 - 1;
  f32x4 acc[MB][NBK];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NBK; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const long c0 = (long)blockIdx.x * WG_CHUNK;
  for (int step = wave; step < WG_CHUNK / 32; step += 4) {
    const long p = c0 + step * 32;  // 32 points of one z-row (n2 % 32 == 0)
    const int i2 = (int)(p % n2);
    const long r = p / n2;
    const int i1 = (int)(r % n1), i0 = (int)(r / n1);
    const long srow = ((long)wrap(i0 + tz, n0) * n1 + wrap(i1 + ty, n1)) * n2;
    bf16x8 av[MB], bv[NBK];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const long pj = p + 8 * g + j;
      const long qj = srow + wrap(i2 + 8 * g + j + tx, n2);
#pragma unroll
      for (int b = 0; b < MB; ++b) {
        const int co = b * 16 + m;
        av[b][j] = co < GC ? gp[pj * GC + co] : (__bf16)0.f;
      }
#pragma unroll
      for (int b = 0; b < NBK; ++b) {
        const int ci = b * 16 + m;
        bv[b][j] = ci < XC ? xp[qj * XC + ci] : (__bf16)0.f;
      }
    }
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NBK; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
  }
  // sum the four waves: lane tiles to LDS, wave 0 adds
  __shared__ float red[3][64 * MB * NBK * 4];
  if (wave > 0) {
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NBK; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) red[wave - 1][((i * NBK + j) * 4 + k) * 64 + lane] = acc[i][j][k];
  }
  __syncthreads();
  if (wave == 0) {
    float* out = part + ((long)blockIdx.x * 27 + tap) * GC * XC;
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NBK; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int idx = ((i * NBK + j) * 4 + k) * 64 + lane;
          const float v = acc[i][j][k] + red[0][idx] + red[1][idx] + red[2][idx];
          const int co = i * 16 + 4 * g + k, ci = j * 16 + m;  // D[row = co][col = ci]
          if (co < GC && ci < XC) out[co * XC + ci] = v;
        }
  }
}

int launch_wgrad(int gc, int xc, const __bf16* g, const __bf16* x, float* part, int n0, int n1, int n2,
                 hipStream_t s) {
  const long P = (long)n0 * n1 * n2;
  if (P % WG_CHUNK || n2 % 32) return -1;
  const int nchunk = (int)(P / WG_CHUNK);
  const dim3 grid(nchunk, 27), blk(256);
  if (gc == 32 && xc == 8) hipLaunchKernelGGL((k_wgrad<32, 8>), grid, blk, 0, s, g, x, part, n0, n1, n2);
  else if (gc == 32 && xc == 32) hipLaunchKernelGGL((k_wgrad<32, 32>), grid, blk, 0, s, g, x, part, n0, n1, n2);
  else if (gc == 8 && xc == 32) hipLaunchKernelGGL((k_wgrad<8, 32>), grid, blk, 0, s, g, x, part, n0, n1, n2);
  else return -1;
  return nchunk;
}

// grad[tap][co][ci] (canonical, real sizes) = sum over chunks, fixed order
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float* __restrict__ part, int nchunk, int gc, int xc,
                                                      int cor, int cir, float* __restrict__ grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int n = 27 * cor * cir;
  if (i >= n) return;
  const int ci = i % cir, co = (i / cir) % cor, tap = i / (cir * cor);
  const long off = ((long)tap * gc + co) * xc + ci;
  float sum = 0.f;
  for (int c = 0; c < nchunk; ++c) sum += part[(long)c * 27 * gc * xc + off];
  grad[i] = sum;
}

int launch_wgrad_reduce(const float* part, int nchunk, int gc, int xc, int cout_real, int cin_real, float* grad,
                        hipStream_t s) {
  const int n = 27 * cout_real * cin_real;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((n + 255) / 256), dim3(256), 0, s, part, nchunk, gc, xc, cout_real,
                     cin_real, grad);
  return 0;
}

// ------------------------------------------------------------------ bias gradient
// 256 blocks of partial column sums, then one block sums them in order.
template <typename T>
__global__ void __launch_bounds__(256) k_bgrad_part(const T* __restrict__ gp, int gc, int creal, long n,
                                                    float* __restrict__ part) {
  __shared__ float s[256];
  for (int c = 0; c < creal; ++c) {
    float a = 0.f;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) a += (float)gp[p * gc + c];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * creal + c] = s[0];
    __syncthreads();
  }
}
__global__ void k_bgrad_sum(const float* __restrict__ part, int nblk, int creal, float* __restrict__ grad) {
  const int c = threadIdx.x;
  if (c >= creal) return;
  float a = 0.f;
  for (int b = 0; b < nblk; ++b) a += part[b * creal + c];
  grad[c] = a;
}
int launch_bgrad_bf16(const __bf16* g, int gc, int creal, long npts, float* part, float* grad, hipStream_t s) {
  hipLaunchKernelGGL((k_bgrad_part<__bf16>), dim3(256), dim3(256), 0, s, g, gc, creal, npts, part);
  hipLaunchKernelGGL(k_bgrad_sum, dim3(1), dim3(64), 0, s, (const float*)part, 256, creal, grad);
  return 0;
}
int launch_bgrad_f32(const float* g, int gc, int creal, long npts, float* part, float* grad, hipStream_t s) {
  hipLaunchKernelGGL((k_bgrad_part<float>), dim3(256), dim3(256), 0, s, g, gc, creal, npts, part);
  hipLaunchKernelGGL(k_bgrad_sum, dim3(1), dim3(64), 0, s, (const float*)part, 256, creal, grad);
  return 0;
}

// ------------------------------------------------------------------ loss
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                              float* __restrict__ gy, __bf16* __restrict__ gyb, long n,
                                              double* __restrict__ part) {
  __shared__ double s[256];
  double acc = 0.0;
  const float sc = 2.0f / (float)(6.0 * (double)n);
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const float d = c < 6 ? y[p * 8 + c] - t[p * 8 + c] : 0.f;
      acc += (double)d * d;
      const float gv = sc * d;
      gy[p * 8 + c] = gv;
      gyb[p * 8 + c] = (__bf16)gv;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_loss_sum(const double* __restrict__ part, int nblk, long n, double* __restrict__ loss) {
  if (threadIdx.x) return;
  double a = 0.0;
  for (int b = 0; b < nblk; ++b) a += part[b];
  *loss = a / (6.0 * (double)n);
}
int launch_loss(const float* y, const float* t, float* gy, __bf16* gyb, long npts, double* part, double* loss,
                hipStream_t s) {
  hipLaunchKernelGGL(k_loss, dim3(256), dim3(256), 0, s, y, t, gy, gyb, npts, part);
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(64), 0, s, (const double*)part, 256, npts, loss);
  return 0;
}

// ------------------------------------------------------------------ layout conversions
__global__ void k_pack_input(const float* __restrict__ u, float scale, __bf16* __restrict__ x, long n) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  bf16x8 v = {};
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = (__bf16)(u[c * n + p] * scale);
  ((bf16x8*)x)[p] = v;
}
__global__ void k_pack_target(const float* __restrict__ tau, float scale, float* __restrict__ t, long n) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) t[p * 8 + c] = c < 6 ? tau[c * n + p] * scale : 0.f;
}
__global__ void k_unpack_output(const float* __restrict__ y, float scale, float* __restrict__ tau, long n) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) tau[c * n + p] = y[p * 8 + c] * scale;
}
int launch_pack_input(const float* u, float scale, __bf16* x, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, u, scale, x, n);
  return 0;
}
int launch_pack_target(const float* tau, float scale, float* t, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_target, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tau, scale, t, n);
  return 0;
}
int launch_unpack_output(const float* y, float scale, float* tau, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack_output, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, scale, tau, n);
  return 0;
}

// Canonical FP32 parameters -> packed bf16 layouts (one thread per packed entry).
__global__ void __launch_bounds__(256) k_pack_weights(const float* __restrict__ P, ClosureWeights w) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  // segment sizes
  constexpr int S1 = 27 * 32 * 8, S2 = 27 * 32 * 32, S3 = 27 * 16 * 32, S4 = 27 * 32 * 8, S5 = 27 * 32 * 32;
  if (i < S1) {  // w1f[tap][co][ci<8] = W1[tap][co][ci<3]
    const int ci = i % 8, co = (i / 8) % 32, tap = i / 256;
    w.w1f[i] = (__bf16)(ci < 3 ? P[CL_OW1 + (tap * 32 + co) * 3 + ci] : 0.f);
    return;
  }
  int j = i - S1;
  if (j < S2) {
    w.w2f[j] = (__bf16)P[CL_OW2 + j];
    return;
  }
  j -= S2;
  if (j < S3) {  // w3f[tap][co<16][ci] = W3[tap][co<6][ci]
    const int ci = j % 32, co = (j / 32) % 16, tap = j / 512;
    w.w3f[j] = (__bf16)(co < 6 ? P[CL_OW3 + (tap * 6 + co) * 32 + ci] : 0.f);
    return;
  }
  j -= S3;
  if (j < S4) {  // w3b[tap'][c2 (32)][c3 < 8] = W3[26 - tap'][c3 < 6][c2]
    const int c3 = j % 8, c2 = (j / 8) % 32, tp = j / 256;
    w.w3b[j] = (__bf16)(c3 < 6 ? P[CL_OW3 + ((26 - tp) * 6 + c3) * 32 + c2] : 0.f);
    return;
  }
  j -= S4;
  if (j < S5) {  // w2b[tap'][ci][co] = W2[26 - tap'][co][ci]
    const int co = j % 32, ci = (j / 32) % 32, tp = j / 1024;
    w.w2b[j] = (__bf16)P[CL_OW2 + ((26 - tp) * 32 + co) * 32 + ci];
    return;
  }
  j -= S5;
  if (j < 32) { w.b1[j] = P[CL_OB1 + j]; return; }
  j -= 32;
  if (j < 32) { w.b2[j] = P[CL_OB2 + j]; return; }
  j -= 32;
  if (j < 16) w.b3[j] = j < 6 ? P[CL_OB3 + j] : 0.f;
}
int launch_pack_weights(const float* params, const ClosureWeights& w, hipStream_t s) {
  const int n = 27 * 32 * 8 + 27 * 32 * 32 + 27 * 16 * 32 + 27 * 32 * 8 + 27 * 32 * 32 + 32 + 32 + 16;
  hipLaunchKernelGGL(k_pack_weights, dim3((n + 255) / 256), dim3(256), 0, s, params, w);
  return 0;
}

// ------------------------------------------------------------------ Adam
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
int launch_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, int t,
                hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1,
                     c2);
  return 0;
}

// ------------------------------------------------------------------ box filter
// One thread per coarse point; FP64 sums over the r^3 fine points.
template <typename R>
__global__ void __launch_bounds__(256) k_box_filter(const R* __restrict__ u, int N0, int N1, int N2, int r,
                                                    R* __restrict__ ub, R* __restrict__ tau) {
  const int n0 = N0 / r, n1 = N1 / r, n2 = N2 / r;
  const long nc = (long)n0 * n1 * n2, nf = (long)N0 * N1 * N2;
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= nc) return;
  const int c2 = (int)(q % n2), c1 = (int)((q / n2) % n1), c0 = (int)(q / ((long)n2 * n1));
  double s[3] = {0, 0, 0}, ss[6] = {0, 0, 0, 0, 0, 0};
  for (int a = 0; a < r; ++a)
    for (int b = 0; b < r; ++b) {
      const long row = ((long)(c0 * r + a) * N1 + (c1 * r + b)) * N2 + (long)c2 * r;
      for (int c = 0; c < r; ++c) {
        const double x = u[row + c], y = u[nf + row + c], z = u[2 * nf + row + c];
        s[0] += x; s[1] += y; s[2] += z;
        ss[0] += x * x; ss[1] += y * y; ss[2] += z * z; ss[3] += x * y; ss[4] += x * z; ss[5] += y * z;
      }
    }
  const double inv = 1.0 / ((double)r * r * r);
  const double m0 = s[0] * inv, m1 = s[1] * inv, m2 = s[2] * inv;
  ub[q] = (R)m0; ub[nc + q] = (R)m1; ub[2 * nc + q] = (R)m2;
  tau[q] = (R)(ss[0] * inv - m0 * m0);
  tau[nc + q] = (R)(ss[1] * inv - m1 * m1);
  tau[2 * nc + q] = (R)(ss[2] * inv - m2 * m2);
  tau[3 * nc + q] = (R)(ss[3] * inv - m0 * m1);
  tau[4 * nc + q] = (R)(ss[4] * inv - m0 * m2);
  tau[5 * nc + q] = (R)(ss[5] * inv - m1 * m2);
}
int launch_box_filter(const float* u, int N0, int N1, int N2, int r, float* ubar, float* tau, hipStream_t s) {
  if (r < 1 || N0 % r || N1 % r || N2 % r) return -1;
  const long nc = (long)(N0 / r) * (N1 / r) * (N2 / r);
  hipLaunchKernelGGL((k_box_filter<float>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, u, N0, N1, N2, r,
                     ubar, tau);
  return 0;
}
int launch_box_filter_f64(const double* u, int N0, int N1, int N2, int r, double* ubar, double* tau, hipStream_t s) {
  if (r < 1 || N0 % r || N1 % r || N2 % r) return -1;
  const long nc = (long)(N0 / r) * (N1 / r) * (N2 / r);
  hipLaunchKernelGGL((k_box_filter<double>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, u, N0, N1, N2, r,
                     ubar, tau);
  return 0;
}

}  // namespace sdns
