// Learned subgrid closure: argument blocks and launchers of the kernels in
// sdns_closure.hip.
//
// Network: three periodic 3x3x3 cross-correlation layers 3 -> 32 -> 32 -> 6
// with bias and SiLU after layers 1 and 2, on channels-last fields
// (N0, N1, N2, C) of the coarse grid.  Storage: bf16 weights, inputs and
// layer outputs; FP32 accumulation (MFMA 16x16x32 bf16); FP32 network output,
// loss gradient and master parameters.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sdns {

// canonical parameter vector (FP32): W1[27][32][3] b1[32] W2[27][32][32] b2[32] W3[27][6][32] b3[6]
constexpr int CL_C1 = 32, CL_C2 = 32, CL_OUT = 6, CL_IN = 3;
constexpr int CL_OW1 = 0;
constexpr int CL_OB1 = CL_OW1 + 27 * CL_C1 * CL_IN;
constexpr int CL_OW2 = CL_OB1 + CL_C1;
constexpr int CL_OB2 = CL_OW2 + 27 * CL_C2 * CL_C1;
constexpr int CL_OW3 = CL_OB2 + CL_C2;
constexpr int CL_OB3 = CL_OW3 + 27 * CL_OUT * CL_C2;
constexpr int CL_NPARAM = CL_OB3 + CL_OUT;

// device layouts of the packed bf16 weights (per layer, [27][cout][cin] padded)
struct ClosureWeights {
  __bf16* w1f;   // [27][32][8]   layer 1 forward (cin 3 padded to 8)
  __bf16* w2f;   // [27][32][32]  layer 2 forward
  __bf16* w3f;   // [27][16][32]  layer 3 forward (cout 6 padded to 16)
  __bf16* w3b;   // [27][32][8]   layer 3 input gradient: tap flipped, (ci, co) swapped, co 6 padded to 8
  __bf16* w2b;   // [27][32][32]  layer 2 input gradient: tap flipped, transposed
  float* b1;     // [32]
  float* b2;     // [32]
  float* b3;     // [16] (6 real)
};

enum { CONV_EPI_SILU = 0, CONV_EPI_LINEAR = 1, CONV_EPI_DSILU = 2 };

struct ConvArgs {
  const __bf16* x;      // input [P][cin]
  const __bf16* w;      // [27][cout][cin]
  const float* bias;    // [cout] or null
  __bf16* z;            // SILU: pre-activation out [P][cout] (null: not stored)
  __bf16* a;            // SILU: activation out; DSILU: gradient out [P][cout]
  float* y;             // LINEAR: FP32 out [P][ystride], channels < ncout written
  const __bf16* zprev;  // DSILU: stored pre-activation [P][cout] whose SiLU' multiplies the result
  int n0, n1, n2;
  int ystride, ncout;
};

// cin in {8, 32}, cout in {16, 32}; N2 % 16 == 0 and N0*N1*N2 % 64 == 0.
// 32-channel inputs are LDS-tiled (weights and a rolling window of input
// planes staged per block) when N0 % 4 == 0, N1 % 4 == 0 and N2 <= 64 or
// N2 % 64 == 0.  Returns nonzero if the shape or the launch set-up
// is not supported.
int launch_conv(int cin, int cout, int epi, const ConvArgs& a, hipStream_t s);
// Weight-gradient partials: part[chunk][27][gc][xc] = sum over the chunk's
// points p of g[p][co] x[p + d(tap)][ci]; gc, xc in {8, 32}; chunk = 4096
// points (N0*N1*N2 % 4096 == 0, N2 % 32 == 0).  Returns the number of chunks.
int launch_wgrad(int gc, int xc, const __bf16* g, const __bf16* x, float* part, int n0, int n1, int n2,
                 hipStream_t s);
// Sum the partials over chunks into the canonical gradient W[tap][co < cout_real][ci < cin_real].
int launch_wgrad_reduce(const float* part, int nchunk, int gc, int xc, int cout_real, int cin_real, float* grad,
                        hipStream_t s);
// Bias gradient: grad[c] = sum_p g[p][c] for c < creal (g bf16 or FP32 with row stride gc).
int launch_bgrad_bf16(const __bf16* g, int gc, int creal, long npts, float* part, float* grad, hipStream_t s);
int launch_bgrad_f32(const float* g, int gc, int creal, long npts, float* part, float* grad, hipStream_t s);
// Loss and its gradient: y, t FP32 [P][8] (6 real); gy = 2 (y - t) / (6 P) as FP32
// and bf16 [P][8] (pad 0); loss (device double) = mean (y - t)^2.
int launch_loss(const float* y, const float* t, float* gy, __bf16* gyb, long npts, double* part, double* loss,
                hipStream_t s);
// Planar FP32 fields [c][P] * scale <-> channels-last.
int launch_pack_input(const float* u, float scale, __bf16* x, long npts, hipStream_t s);        // [3][P] -> bf16 [P][8]
int launch_pack_target(const float* tau, float scale, float* t, long npts, hipStream_t s);      // [6][P] -> f32 [P][8]
int launch_unpack_output(const float* y, float scale, float* tau, long npts, hipStream_t s);    // f32 [P][8] -> [6][P]
// Canonical FP32 parameters -> bf16 device layouts.
int launch_pack_weights(const float* params, const ClosureWeights& w, hipStream_t s);
// Adam (Kingma & Ba 2015) on n FP32 parameters, step t >= 1.
int launch_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, int t,
                hipStream_t s);
// Box filter by r of a physical FP32 field u[3][N0][N1][N2] into ubar[3][n][..]
// and tau[6][n][..] (tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j), FP64 sums).
int launch_box_filter(const float* u, int N0, int N1, int N2, int r, float* ubar, float* tau, hipStream_t s);
int launch_box_filter_f64(const double* u, int N0, int N1, int N2, int r, double* ubar, double* tau, hipStream_t s);

}  // namespace sdns
