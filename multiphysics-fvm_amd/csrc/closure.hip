// Learned flux-correction closure: a 3-layer 3x3 CNN on MFMA (BF16 operands,
// FP32 accumulation) and its use inside the SSPRK3 step.
//
// No counterpart in the reference.  The network maps the normalised
// primitives (rho, v1, v2, p) of a 2D periodic Cartesian state to
// corrections of the x1- and x2-face fluxes of mass, momentum 1, 2 and
// energy.  The corrected right-hand side is k + kc with
// kc = -div(dF) (conservative by construction); whenever the corrected stage
// state has rho <= 0 or p <= 0 in any cell, the stage keeps the plain
// MUSCL/HLL result (positivity fallback).
//
// Tensors are channels-last [batch][n2][n1][C] (C contiguous), periodic in
// x1 and x2.  A convolution is an implicit GEMM per 16-pixel row segment:
// D[co][pixel] = sum_k W[co][k] * X[k][pixel] with k = (tap, ci), evaluated
// with v_mfma_f32_16x16x16bf16_1k.  Operand layout of that instruction
// (64 lanes, g = lane / 16, c = lane % 16):
//   A: lane holds A[row c][k = 4g .. 4g+3]     (4 BF16)
//   B: lane holds B[k = 4g .. 4g+3][col c]     (4 BF16)
//   D: lane holds D[row 4g .. 4g+3][col c]     (4 FP32)
// so with channels-last tensors each operand fragment is one 8-byte load and
// each result fragment one 8- or 16-byte store.
#include <stdint.h>

#include "common.h"

namespace mfvm {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// FP32 -> BF16, round to nearest even (NaN stays NaN).
MFVM_DEV uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
MFVM_DEV float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kTap = 9;

template <int N> struct Round16 { static constexpr int v = (N + 15) / 16 * 16; };

enum { CNN_FWD_RELU = 0, CNN_FWD_LINEAR = 1, CNN_BWD_DATA = 2 };

MFVM_DEV int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// The 4 network input features of the cell at offset o of a 2D state
// (rho, v1, v2, p from the conserved fields; p as the reference closure with
// v3), shifted, scaled and rounded to BF16.
template <class R>
MFVM_DEV bf16x4 cnn_feat4(long long fs, double gm1, const R* u, long long o, const float* shift, const float* scale) {
  const double rho = double(u[o]);
  const double v1 = double(u[o + fs]) / rho, v2 = double(u[o + 2 * fs]) / rho;
  const double v3 = double(u[o + 3 * fs]) / rho;
  const double p = gm1 * (double(u[o + 4 * fs]) - 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3));
  const double q[4] = {rho, v1, v2, p};
  bf16x4 h;
#pragma unroll
  for (int f = 0; f < 4; ++f) h[f] = (short)f2bf((float(q[f]) - shift[f]) * scale[f]);
  return h;
}

// FEAT (first layer of the closure's inference): the input fragment of a
// pixel is computed from the float64 state by cnn_feat4 instead of read from
// a feature tensor -- the same values, one launch fewer.
struct FeatSrc {
  Layout lay;
  const double* U;
  double gm1;
  const float* shift;
  const float* scale;
};

// One 3x3 convolution (cross-correlation, periodic padding).
//   FWD_RELU:   y = relu(conv(x, w) + b)            (BF16 out)
//   FWD_LINEAR: y = conv(x, w) + b                  (FP32 out)
//   BWD_DATA:   y = conv(x, w) * (mask > 0)         (BF16 out; w in the
//               transposed-flipped layout, x = output gradient)
// w: [9][round16(COUT)][CIN] BF16 (rows >= COUT are zero).  CIN % 4 == 0,
// n1 % 16 == 0.  One wave per 16-pixel row segment (grid-stride).
template <int CIN, int COUT, int MODE, bool FEAT = false>
__global__ __launch_bounds__(256) void cnn_conv_kernel(int batch, int n1, int n2, const uint16_t* __restrict__ x,
                                                      const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                      const uint16_t* __restrict__ mask, void* __restrict__ y,
                                                      FeatSrc fsrc) {
  static_assert(!FEAT || (CIN == 4 && MODE == CNN_FWD_RELU), "features feed the first layer only");
  constexpr int COUTP = Round16<COUT>::v;
  constexpr int NCB = COUTP / 16;
  constexpr int KP = Round16<kTap * CIN>::v;
  constexpr int NKS = KP / 16;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const long long nseg = (long long)batch * n2 * (n1 / 16);
  const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long long nwave = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long seg = wave; seg < nseg; seg += nwave) {
    const int xs = (int)(seg % (n1 / 16)) * 16;
    const long long row = seg / (n1 / 16);
    const int yy = (int)(row % n2);
    const long long b = row / n2;
    const uint16_t* xb = FEAT ? x : x + b * (long long)n2 * n1 * CIN;
    f32x4 acc[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NKS; ++s) {
      const int k0 = 16 * s + 4 * g;
      const int tap = k0 / CIN, ci0 = k0 % CIN;
      bf16x4 bv = {0, 0, 0, 0};
      if (tap < kTap) {
        const int px = wrap(xs + c + (tap % 3) - 1, n1), py = wrap(yy + (tap / 3) - 1, n2);
        if constexpr (FEAT)
          bv = cnn_feat4<double>(fsrc.lay.fs, fsrc.gm1, fsrc.U, fsrc.lay.at(px + 1, py + 1, 1), fsrc.shift, fsrc.scale);
        else
          bv = *reinterpret_cast<const bf16x4*>(xb + ((long long)py * n1 + px) * CIN + ci0);
      }
      const int tw = tap < kTap ? tap : kTap - 1;
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const bf16x4 av = *reinterpret_cast<const bf16x4*>(w + ((long long)tw * COUTP + cb * 16 + c) * CIN + ci0);
        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[cb], 0, 0, 0);
      }
    }
    const long long pix = (b * n2 + yy) * (long long)n1 + xs + c;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int co0 = cb * 16 + 4 * g;
      if (co0 >= COUT) continue;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = acc[cb][r];
      if (MODE == CNN_FWD_LINEAR) {
        float4 o;
        o.x = v[0] + bias[co0];
        o.y = v[1] + bias[co0 + 1];
        o.z = v[2] + bias[co0 + 2];
        o.w = v[3] + bias[co0 + 3];
        *reinterpret_cast<float4*>(static_cast<float*>(y) + pix * COUT + co0) = o;
      } else {
        if (MODE == CNN_FWD_RELU) {
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = fmaxf(v[r] + bias[co0 + r], 0.f);
        } else {
          const bf16x4 m = *reinterpret_cast<const bf16x4*>(mask + pix * COUT + co0);
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = bf2f((uint16_t)m[r]) > 0.f ? v[r] : 0.f;
        }
        bf16x4 o;
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = (short)f2bf(v[r]);
        *reinterpret_cast<bf16x4*>(static_cast<uint16_t*>(y) + pix * COUT + co0) = o;
      }
    }
  }
}

// Weight and bias gradients of one convolution, accumulated (atomically)
// into gw[cout][cin][3][3] and gb[cout] (FP32):
//   gw[co][ci][tap] += sum_p gy[p][co] * x[p + off(tap)][ci],  gb[co] += sum_p gy[p][co].
// MFMA with M = co, N = ci, K = 16 pixels of one row segment.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void cnn_wgrad_kernel(int batch, int n1, int n2, const uint16_t* __restrict__ x,
                                                       const uint16_t* __restrict__ gy, float* __restrict__ gw,
                                                       float* __restrict__ gb) {
  constexpr int NCB = Round16<COUT>::v / 16;
  constexpr int NIB = Round16<CIN>::v / 16;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const long long nseg = (long long)batch * n2 * (n1 / 16);
  const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long long nwave = (long long)gridDim.x * (blockDim.x >> 6);
  f32x4 acc[kTap][NCB][NIB];
#pragma unroll
  for (int t = 0; t < kTap; ++t)
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
      for (int ib = 0; ib < NIB; ++ib) acc[t][cb][ib] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[NCB];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb) bsum[cb] = 0.f;
  for (long long seg = wave; seg < nseg; seg += nwave) {
    const int xs = (int)(seg % (n1 / 16)) * 16;
    const long long row = seg / (n1 / 16);
    const int yy = (int)(row % n2);
    const long long b = row / n2;
    const uint16_t* xb = x + b * (long long)n2 * n1 * CIN;
    const long long p0 = (b * n2 + yy) * (long long)n1 + xs;  // first pixel of the segment
    bf16x4 av[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int co = cb * 16 + c;
#pragma unroll
      for (int e = 0; e < 4; ++e) av[cb][e] = co < COUT ? (short)gy[(p0 + 4 * g + e) * COUT + co] : (short)0;
      if (g == 0 && co < COUT) {
        float s = 0.f;
        for (int q = 0; q < 16; ++q) s += bf2f(gy[(p0 + q) * COUT + co]);
        bsum[cb] += s;
      }
    }
#pragma unroll
    for (int t = 0; t < kTap; ++t) {
      const int py = wrap(yy + (t / 3) - 1, n2);
#pragma unroll
      for (int ib = 0; ib < NIB; ++ib) {
        const int ci = ib * 16 + c;
        bf16x4 bv;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int px = wrap(xs + 4 * g + e + (t % 3) - 1, n1);
          bv[e] = ci < CIN ? (short)xb[((long long)py * n1 + px) * CIN + ci] : (short)0;
        }
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
          acc[t][cb][ib] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[cb], bv, acc[t][cb][ib], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kTap; ++t)
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
      for (int ib = 0; ib < NIB; ++ib) {
        const int ci = ib * 16 + c;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = cb * 16 + 4 * g + r;
          if (co < COUT && ci < CIN) atomicAdd(gw + ((long long)co * CIN + ci) * kTap + t, acc[t][cb][ib][r]);
        }
      }
  if (g == 0) {
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int co = cb * 16 + c;
      if (co < COUT) atomicAdd(gb + co, bsum[cb]);
    }
  }
}

// FP32 master weights w[cout][cin][3][3] -> BF16 forward layout
// wf[9][round16(cout)][cin] and transposed-flipped backward layout
// wb[9][round16(cin)][cout] with wb[t][ci][co] = w[co][ci][8 - t].
__global__ void cnn_pack_kernel(int cin, int cout, const float* __restrict__ w, uint16_t* __restrict__ wf,
                                uint16_t* __restrict__ wb) {
  const int coutp = (cout + 15) / 16 * 16, cinp = (cin + 15) / 16 * 16;
  const int nf = kTap * coutp * cin, nb = kTap * cinp * cout;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nf + nb; t += gridDim.x * blockDim.x) {
    if (t < nf) {
      const int ci = t % cin, co = (t / cin) % coutp, tap = t / (cin * coutp);
      wf[t] = co < cout ? f2bf(w[((long long)co * cin + ci) * kTap + tap]) : 0;
    } else {
      const int u = t - nf;
      const int co = u % cout, ci = (u / cout) % cinp, tap = u / (cout * cinp);
      wb[u] = ci < cin ? f2bf(w[((long long)co * cin + ci) * kTap + (kTap - 1 - tap)]) : 0;
    }
  }
}

// Adam (FP32 master weights): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// w -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).
__global__ void cnn_adam_kernel(long long n, float* __restrict__ w, const float* __restrict__ g,
                                float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Network input: x[b][j][i][:] = ((rho, v1, v2, p) - shift) * scale as BF16
// from the ghosted SoA state of a 2D grid (one x3 plane), batch stride bs
// elements between states.
template <class R>
__global__ void cnn_features_kernel(Layout lay, double gm1, int batch, long long bs, const R* __restrict__ U,
                                    const float* __restrict__ shift, const float* __restrict__ scale,
                                    uint16_t* __restrict__ x) {
  const long long n = (long long)batch * lay.n1 * lay.n2;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % lay.n1);
    const int j = (int)((t / lay.n1) % lay.n2);
    const long long b = t / ((long long)lay.n1 * lay.n2);
    const R* u = U + b * bs;
    *reinterpret_cast<bf16x4*>(x + t * 4) = cnn_feat4<R>(lay.fs, gm1, u, lay.at(i + 1, j + 1, 1), shift, scale);
  }
}

// Correction of the right-hand side from the network output y (flux
// corrections dF = y * out_scale at the x1 face i-1/2 (channels 0-3) and the
// x2 face j-1/2 (channels 4-7) of cell (i, j)):
//   kc(i,j) = -((dF1(i+1,j) - dF1(i,j)) * inv_dx1 + (dF2(i,j+1) - dF2(i,j)) * inv_dx2).
MFVM_DEV void closure_kc(const float* y, int n1, int n2, long long base, int i, int j, const float* osc,
                         float inv1, float inv2, float kc[4]) {
  const long long c0 = base + (long long)j * n1 + i;
  const long long cx = base + (long long)j * n1 + wrap(i + 1, n1);
  const long long cy = base + (long long)wrap(j + 1, n2) * n1 + i;
#pragma unroll
  for (int f = 0; f < 4; ++f)
    kc[f] = -(((y[cx * 8 + f] - y[c0 * 8 + f]) * inv1 + (y[cy * 8 + 4 + f] - y[c0 * 8 + 4 + f]) * inv2) * osc[f]);
}

// Training loss and its gradient with respect to y:
//   L += sum_f wf[f] * (kc_f - target_f)^2 / count over all cells,
//   gy = dL/dy (BF16), by the adjoint of kc (periodic).
__global__ void cnn_loss_kernel(int batch, int n1, int n2, float inv1, float inv2, const float* __restrict__ y,
                                const float* __restrict__ target, const float* __restrict__ osc,
                                const float* __restrict__ wf, float inv_count, uint16_t* __restrict__ gy,
                                float* __restrict__ loss) {
  __shared__ float red[256];
  const long long n = (long long)batch * n1 * n2;
  float part = 0.f;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n1);
    const int j = (int)((t / n1) % n2);
    const long long base = (t / ((long long)n1 * n2)) * (long long)n1 * n2;
    // gk(c) = dL/dkc at cell c
    float gk[3][4];  // (i,j), (i-1,j), (i,j-1)
    const int ci[3] = {i, wrap(i - 1, n1), i};
    const int cj[3] = {j, j, wrap(j - 1, n2)};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float kc[4];
      closure_kc(y, n1, n2, base, ci[q], cj[q], osc, inv1, inv2, kc);
      const long long c = base + (long long)cj[q] * n1 + ci[q];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const float r = kc[f] - target[c * 4 + f];
        gk[q][f] = 2.f * wf[f] * r * inv_count;
        if (q == 0) part += wf[f] * r * r * inv_count;
      }
    }
    // kc(i) = -((y1(i+1) - y1(i)) inv1 + ...) osc  =>  dL/dy1(i) = osc inv1 (gk(i) - gk(i-1))
    bf16x4 g1, g2;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      g1[f] = (short)f2bf(osc[f] * inv1 * (gk[0][f] - gk[1][f]));
      g2[f] = (short)f2bf(osc[f] * inv2 * (gk[0][f] - gk[2][f]));
    }
    *reinterpret_cast<bf16x4*>(gy + t * 8) = g1;
    *reinterpret_cast<bf16x4*>(gy + t * 8 + 4) = g2;
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Corrected stage state T = out + coef * dt * kc (fields rho, m1, m2, E;
// momentum 3 unchanged) on the interior; flag |= 1 where T has rho <= 0 or
// p <= 0 (or is not finite).
template <class R>
__global__ void closure_apply_kernel(Layout lay, double gm1, const R* __restrict__ out, const float* __restrict__ y,
                                     const float* __restrict__ osc, float inv1, float inv2, double coef,
                                     const double* __restrict__ dt, R* __restrict__ T, int* __restrict__ flag) {
  const int n1 = lay.n1, n2 = lay.n2;
  const long long n = (long long)n1 * n2;
  const double cd = coef * *dt;
  int bad = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n1), j = (int)(t / n1);
    float kc[4];
    closure_kc(y, n1, n2, 0, i, j, osc, inv1, inv2, kc);
    const long long o = lay.at(i + 1, j + 1, 1);
    const int fld[4] = {0, 1, 2, 4};
    double v[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) v[q] = double(out[o + q * lay.fs]);
#pragma unroll
    for (int f = 0; f < 4; ++f) v[fld[f]] += cd * double(kc[f]);
    const double rho = v[0];
    const double p = gm1 * (v[4] - 0.5 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) / rho);
    if (!(rho > 0.0) || !(p > 0.0)) bad = 1;
#pragma unroll
    for (int q = 0; q < 5; ++q) T[o + q * lay.fs] = R(v[q]);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// out <- T on the interior unless the flag of this stage is set (then the
// plain result stays and the stage is counted).
template <class R>
__global__ void closure_select_kernel(Layout lay, const R* __restrict__ T, R* __restrict__ out,
                                      const int* __restrict__ flag, int* __restrict__ count) {
  if (*flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(count, 1);
    return;
  }
  const long long n = (long long)lay.n1 * lay.n2;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < 5 * n; t += (long long)gridDim.x * blockDim.x) {
    const int q = (int)(t / n);
    const long long r = t % n;
    const long long o = lay.at((int)(r % lay.n1) + 1, (int)(r / lay.n1) + 1, 1) + q * lay.fs;
    out[o] = T[o];
  }
}

// ---------------------------------------------------------------------------
// Fused inference: features, the three convolutions and the corrected stage
// state of one stage in one launch.
// ---------------------------------------------------------------------------
// One workgroup (4 waves) per CT x CT tile of cells (CT = 16, or 8 on grids
// below 256^2, whose 16 x 16 tiles would leave most CUs idle).  The tile's corrected
// state needs the network output y at its cells and one more column / row
// (the faces i+1/2, j+1/2 of its last cells), so layer 3 is evaluated on
// (CT+1)^2 pixels, layer 2 on (CT+3)^2, layer 1 on (CT+5)^2 and the features
// on (CT+7)^2 (periodic), each layer's output held in LDS (channels-last,
// BF16; layer 3 in FP32).  The convolutions are implicit GEMMs on the gfx950
// v_mfma_f32_16x16x32_bf16 (16 output channels x 16 pixels per instruction,
// K = 32 = one tap of 32 input channels, or 8 taps of the 4 features; lane l
// holds A[l%16][8(l/16) .. +7], B[8(l/16) .. +7][l%16], D[4(l/16) .. +3][l%16]).
// Each pixel's activations and y are those of the layer kernels above (BF16
// rounding after the ReLU, FP32 accumulation); only the order of the FP32
// sums inside a 32-wide K block differs.  The corrected state and the
// positivity flag follow closure_apply_kernel.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
// CT: tile width (16, or 8 where 16 would give fewer than 256 workgroups)
template <int CT> struct NetTile {
  static constexpr int W0 = CT + 7, W1 = CT + 5, W2 = CT + 3, W3 = CT + 1;  // region widths
  // LDS: the regions in BF16 elements, each starting 16-byte aligned (the
  // layer reads are 16-byte loads); layer 3 (FP32) last
  static constexpr int NF0 = (W0 * W0 * 4 + 7) / 8 * 8, NY1 = W1 * W1 * 32, NY2 = W2 * W2 * 32;
  static constexpr size_t lds() { return (size_t)(NF0 + NY1 + NY2) * 2 + (size_t)W3 * W3 * 8 * 4; }
};

MFVM_DEV bf16x8 ld_bf16x8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p)); }
MFVM_DEV bf16x8 two_bf16x4(const uint16_t* p, bool okp, const uint16_t* q, bool okq) {
  const uint2 a = okp ? *reinterpret_cast<const uint2*>(p) : uint2{0u, 0u};
  const uint2 b = okq ? *reinterpret_cast<const uint2*>(q) : uint2{0u, 0u};
  return __builtin_bit_cast(bf16x8, uint4{a.x, a.y, b.x, b.y});
}

// One 3x3 layer over an LDS region: input win x win pixels of CIN channels,
// output (win-2)^2 pixels of COUT channels (ReLU -> BF16, or linear FP32).
// The wave's weight fragments (every K step and output block) are read once
// into registers, then reused for each 16-pixel tile the wave takes.
template <int CIN, int COUT, bool RELU>
MFVM_DEV void net_layer(const uint16_t* __restrict__ xin, int win, const uint16_t* __restrict__ w,
                        const float* __restrict__ bias, void* yout) {
  constexpr int COUTP = Round16<COUT>::v, NMB = COUTP / 16;
  constexpr int NKS = CIN == 4 ? 2 : kTap;  // K steps of 32
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15, wave = threadIdx.x >> 6;
  const int wout = win - 2, npix = wout * wout, ntile = (npix + 15) / 16;
  bf16x8 av[NKS][NMB];
#pragma unroll
  for (int s = 0; s < NKS; ++s) {
#pragma unroll
    for (int mb = 0; mb < NMB; ++mb) {
      const int co = mb * 16 + c;
      if constexpr (CIN == 4) {
        // k = 32 s + 8 g + e: taps 8 s + 2 g and 8 s + 2 g + 1, 4 features each
        const int t0 = 8 * s + 2 * g, t1 = t0 + 1;
        av[s][mb] = two_bf16x4(w + ((long long)min(t0, 8) * COUTP + co) * 4, t0 < kTap,
                               w + ((long long)min(t1, 8) * COUTP + co) * 4, t1 < kTap);
      } else {
        // k = 32 s + 8 g + e: tap s, input channels 8 g .. 8 g + 7
        av[s][mb] = ld_bf16x8(w + ((long long)s * COUTP + co) * CIN + 8 * g);
      }
    }
  }
  float bs[NMB][4];
#pragma unroll
  for (int mb = 0; mb < NMB; ++mb)
#pragma unroll
    for (int r = 0; r < 4; ++r) bs[mb][r] = mb * 16 + 4 * g + r < COUT ? bias[mb * 16 + 4 * g + r] : 0.f;
  for (int t = wave; t < ntile; t += 4) {
    const int p = min(16 * t + c, npix - 1);  // pixels past the region repeat the last (not stored)
    const int ox = p % wout, oy = p / wout;
    f32x4 acc[NMB];
#pragma unroll
    for (int mb = 0; mb < NMB; ++mb) acc[mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NKS; ++s) {
      bf16x8 bv;
      if constexpr (CIN == 4) {
        const int t0 = 8 * s + 2 * g, t1 = t0 + 1;
        const uint16_t* x0 = xin + ((oy + min(t0, 8) / 3) * win + ox + min(t0, 8) % 3) * 4;
        const uint16_t* x1 = xin + ((oy + min(t1, 8) / 3) * win + ox + min(t1, 8) % 3) * 4;
        bv = two_bf16x4(x0, t0 < kTap, x1, t1 < kTap);
      } else {
        bv = ld_bf16x8(xin + ((oy + s / 3) * win + ox + s % 3) * CIN + 8 * g);
      }
#pragma unroll
      for (int mb = 0; mb < NMB; ++mb) acc[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[s][mb], bv, acc[mb], 0, 0, 0);
    }
    if (16 * t + c >= npix) continue;
#pragma unroll
    for (int mb = 0; mb < NMB; ++mb) {
      const int co0 = mb * 16 + 4 * g;
      if (co0 >= COUT) continue;
      if constexpr (RELU) {
        bf16x4 o;
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = (short)f2bf(fmaxf(acc[mb][r] + bs[mb][r], 0.f));
        *reinterpret_cast<bf16x4*>(static_cast<uint16_t*>(yout) + p * COUT + co0) = o;
      } else {
        float4 o;
        o.x = acc[mb][0] + bs[mb][0];
        o.y = acc[mb][1] + bs[mb][1];
        o.z = acc[mb][2] + bs[mb][2];
        o.w = acc[mb][3] + bs[mb][3];
        *reinterpret_cast<float4*>(static_cast<float*>(yout) + p * COUT + co0) = o;
      }
    }
  }
}

// yg != NULL: the network output y of the tile's cells is also written there
// ([n2][n1][8] FP32, as cnn_conv's), for checking against the layer kernels.
template <int CT>
__global__ __launch_bounds__(256) void closure_net_apply_kernel(
    Layout lay, double gm1, const double* __restrict__ U, const double* __restrict__ out, const float* __restrict__ shift,
    const float* __restrict__ scale, const uint16_t* __restrict__ w0, const float* __restrict__ b0,
    const uint16_t* __restrict__ w1, const float* __restrict__ b1, const uint16_t* __restrict__ w2,
    const float* __restrict__ b2, const float* __restrict__ osc, float inv1, float inv2, double coef,
    const double* __restrict__ dt, double* __restrict__ T, int* __restrict__ flag, float* __restrict__ yg) {
  using NT = NetTile<CT>;
  constexpr int CW0 = NT::W0, CW1 = NT::W1, CW2 = NT::W2, CW3 = NT::W3;
  extern __shared__ double smem_raw[];
  uint16_t* F0 = reinterpret_cast<uint16_t*>(smem_raw);  // [CW0^2][4]
  uint16_t* Y1 = F0 + NT::NF0;                           // [CW1^2][32]
  uint16_t* Y2 = Y1 + NT::NY1;                           // [CW2^2][32]
  float* Y3 = reinterpret_cast<float*>(Y2 + NT::NY2);    // [CW3^2][8]
  const int n1 = lay.n1, n2 = lay.n2;
  const int i0 = blockIdx.x * CT, j0 = blockIdx.y * CT;  // 0-based interior
  // features of the stage input (cnn_feat4), periodic
  for (int e = threadIdx.x; e < CW0 * CW0; e += blockDim.x) {
    const int x = wrap(i0 - 3 + e % CW0, n1), y = wrap(j0 - 3 + e / CW0, n2);
    *reinterpret_cast<bf16x4*>(F0 + e * 4) = cnn_feat4<double>(lay.fs, gm1, U, lay.at(x + 1, y + 1, 1), shift, scale);
  }
  __syncthreads();
  net_layer<4, 32, true>(F0, CW0, w0, b0, Y1);
  __syncthreads();
  net_layer<32, 32, true>(Y1, CW1, w1, b1, Y2);
  __syncthreads();
  net_layer<32, 8, false>(Y2, CW2, w2, b2, Y3);
  __syncthreads();
  // corrected state of the tile's cells (closure_kc, closure_apply_kernel)
  const int tx = threadIdx.x % CT, ty = threadIdx.x / CT;
  const int i = i0 + tx, j = j0 + ty;
  int bad = 0;
  if (ty < CT && i < n1 && j < n2) {
    const float* y0 = Y3 + (ty * CW3 + tx) * 8;
    const float* yx = y0 + 8;
    const float* yy = y0 + CW3 * 8;
    float kc[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) kc[f] = -(((yx[f] - y0[f]) * inv1 + (yy[4 + f] - y0[4 + f]) * inv2) * osc[f]);
    const double cd = coef * *dt;
    const long long o = lay.at(i + 1, j + 1, 1);
    const int fld[4] = {0, 1, 2, 4};
    double v[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) v[q] = out[o + q * lay.fs];
#pragma unroll
    for (int f = 0; f < 4; ++f) v[fld[f]] += cd * double(kc[f]);
    const double rho = v[0];
    const double p = gm1 * (v[4] - 0.5 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) / rho);
    if (!(rho > 0.0) || !(p > 0.0)) bad = 1;
#pragma unroll
    for (int q = 0; q < 5; ++q) T[o + q * lay.fs] = v[q];
    if (yg) {
#pragma unroll
      for (int f = 0; f < 8; ++f) yg[((long long)j * n1 + i) * 8 + f] = y0[f];
    }
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void fill_int_kernel(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

static unsigned grid_waves(long long nseg) {
  long long blocks = (nseg + 3) / 4;
  if (blocks > 4096) blocks = 4096;
  return (unsigned)(blocks < 1 ? 1 : blocks);
}

template <int CIN, int COUT>
static hipError_t conv_t(int mode, int batch, int n1, int n2, const void* x, const void* w, const float* b,
                         const void* mask, void* y, hipStream_t s) {
  const unsigned g = grid_waves((long long)batch * n2 * (n1 / 16));
  const uint16_t* xs = static_cast<const uint16_t*>(x);
  const uint16_t* ws = static_cast<const uint16_t*>(w);
  const uint16_t* ms = static_cast<const uint16_t*>(mask);
  if (mode == CNN_FWD_RELU)
    hipLaunchKernelGGL((cnn_conv_kernel<CIN, COUT, CNN_FWD_RELU>), dim3(g), dim3(256), 0, s, batch, n1, n2, xs, ws, b,
                       ms, y, FeatSrc{});
  else if (mode == CNN_FWD_LINEAR)
    hipLaunchKernelGGL((cnn_conv_kernel<CIN, COUT, CNN_FWD_LINEAR>), dim3(g), dim3(256), 0, s, batch, n1, n2, xs, ws,
                       b, ms, y, FeatSrc{});
  else
    hipLaunchKernelGGL((cnn_conv_kernel<CIN, COUT, CNN_BWD_DATA>), dim3(g), dim3(256), 0, s, batch, n1, n2, xs, ws,
                       b, ms, y, FeatSrc{});
  return hipGetLastError();
}

// Supported (cin, cout) of a call: the network's layers and their transposes.
hipError_t cnn_conv(int mode, int cin, int cout, int batch, int n1, int n2, const void* x, const void* w,
                    const float* b, const void* mask, void* y, hipStream_t s) {
  if (cin == 4 && cout == 32) return conv_t<4, 32>(mode, batch, n1, n2, x, w, b, mask, y, s);
  if (cin == 32 && cout == 32) return conv_t<32, 32>(mode, batch, n1, n2, x, w, b, mask, y, s);
  if (cin == 32 && cout == 8) return conv_t<32, 8>(mode, batch, n1, n2, x, w, b, mask, y, s);
  if (cin == 8 && cout == 32) return conv_t<8, 32>(mode, batch, n1, n2, x, w, b, mask, y, s);
  return hipErrorInvalidValue;
}

// First layer of the closure's inference on one float64 2D state: features
// and the 4 -> 32 ReLU convolution in one launch (cnn_features + cnn_conv).
hipError_t cnn_conv_features(const mfvm_grid* g, double gamma, const void* U, const float* shift, const float* scale,
                             const void* w, const float* b, void* y, hipStream_t s) {
  const FeatSrc f{make_layout(g->n[0], g->n[1], g->n[2]), static_cast<const double*>(U), gamma - 1.0, shift, scale};
  const unsigned gr = grid_waves((long long)g->n[1] * (g->n[0] / 16));
  hipLaunchKernelGGL((cnn_conv_kernel<4, 32, CNN_FWD_RELU, true>), dim3(gr), dim3(256), 0, s, 1, g->n[0], g->n[1],
                     nullptr, static_cast<const uint16_t*>(w), b, nullptr, y, f);
  return hipGetLastError();
}

bool cnn_shape_ok(int mode, int cin, int cout) {
  if (mode == CNN_BWD_DATA) return (cin == 32 && cout == 32) || (cin == 8 && cout == 32);
  return (cin == 4 && cout == 32) || (cin == 32 && cout == 32) || (cin == 32 && cout == 8);
}

hipError_t cnn_wgrad(int cin, int cout, int batch, int n1, int n2, const void* x, const void* gy, float* gw,
                     float* gb, hipStream_t s) {
  const unsigned g = grid_waves((long long)batch * n2 * (n1 / 16) / 8);
  const uint16_t* xs = static_cast<const uint16_t*>(x);
  const uint16_t* gs = static_cast<const uint16_t*>(gy);
  if (cin == 4 && cout == 32)
    hipLaunchKernelGGL((cnn_wgrad_kernel<4, 32>), dim3(g), dim3(256), 0, s, batch, n1, n2, xs, gs, gw, gb);
  else if (cin == 32 && cout == 32)
    hipLaunchKernelGGL((cnn_wgrad_kernel<32, 32>), dim3(g), dim3(256), 0, s, batch, n1, n2, xs, gs, gw, gb);
  else if (cin == 32 && cout == 8)
    hipLaunchKernelGGL((cnn_wgrad_kernel<32, 8>), dim3(g), dim3(256), 0, s, batch, n1, n2, xs, gs, gw, gb);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t cnn_pack(int cin, int cout, const float* w, void* wf, void* wb, hipStream_t s) {
  hipLaunchKernelGGL(cnn_pack_kernel, dim3(64), dim3(256), 0, s, cin, cout, w, static_cast<uint16_t*>(wf),
                     static_cast<uint16_t*>(wb));
  return hipGetLastError();
}

hipError_t cnn_adam(long long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                    float eps, int t, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  hipLaunchKernelGGL(cnn_adam_kernel, dim3(blocks < 1 ? 1 : blocks), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}

template <class R>
hipError_t cnn_features(const mfvm_grid* g, double gamma, int batch, long long bs, const void* U, const float* shift,
                        const float* scale, void* x, hipStream_t s) {
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const long long n = (long long)batch * g->n[0] * g->n[1];
  const unsigned blocks = (unsigned)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);
  hipLaunchKernelGGL(cnn_features_kernel<R>, dim3(blocks), dim3(256), 0, s, lay, gamma - 1.0, batch, bs,
                     static_cast<const R*>(U), shift, scale, static_cast<uint16_t*>(x));
  return hipGetLastError();
}
template hipError_t cnn_features<double>(const mfvm_grid*, double, int, long long, const void*, const float*,
                                         const float*, void*, hipStream_t);
template hipError_t cnn_features<float>(const mfvm_grid*, double, int, long long, const void*, const float*,
                                        const float*, void*, hipStream_t);

hipError_t cnn_loss(int batch, int n1, int n2, float inv1, float inv2, const float* y, const float* target,
                    const float* osc, const float* wf, void* gy, float* loss, hipStream_t s) {
  const long long n = (long long)batch * n1 * n2;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);
  hipLaunchKernelGGL(cnn_loss_kernel, dim3(blocks), dim3(256), 0, s, batch, n1, n2, inv1, inv2, y, target, osc, wf,
                     1.0f / (float)n, static_cast<uint16_t*>(gy), loss);
  return hipGetLastError();
}

template <class R>
hipError_t closure_apply(const mfvm_grid* g, double gamma, const void* out, const float* y, const float* osc,
                         double coef, const double* dt, void* T, int* flag, hipStream_t s) {
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const long long n = (long long)g->n[0] * g->n[1];
  const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  const float inv1 = (float)(1.0 / g->lval[0]), inv2 = (float)(1.0 / g->lval[1]);
  hipLaunchKernelGGL(closure_apply_kernel<R>, dim3(blocks), dim3(256), 0, s, lay, gamma - 1.0,
                     static_cast<const R*>(out), y, osc, inv1, inv2, coef, dt, static_cast<R*>(T), flag);
  return hipGetLastError();
}
template <class R>
hipError_t closure_select(const mfvm_grid* g, const void* T, void* out, const int* flag, int* count,
                          hipStream_t s) {
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const long long n = 5LL * g->n[0] * g->n[1];
  const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  hipLaunchKernelGGL(closure_select_kernel<R>, dim3(blocks), dim3(256), 0, s, lay, static_cast<const R*>(T),
                     static_cast<R*>(out), flag, count);
  return hipGetLastError();
}
template hipError_t closure_apply<double>(const mfvm_grid*, double, const void*, const float*, const float*, double,
                                          const double*, void*, int*, hipStream_t);
template hipError_t closure_select<double>(const mfvm_grid*, const void*, void*, const int*, int*, hipStream_t);

// Features, the three layers and the corrected stage state T of one stage in
// one launch (closure_net_apply_kernel); the tile grid covers the 2D grid.
hipError_t closure_net_apply(const mfvm_grid* g, double gamma, const void* U, const void* out, const float* shift,
                             const float* scale, const void* const wf[3], const float* const b[3], const float* osc,
                             double coef, const double* dt, void* T, int* flag, float* yg, hipStream_t s) {
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const float inv1 = (float)(1.0 / g->lval[0]), inv2 = (float)(1.0 / g->lval[1]);
  const bool big = (long long)((g->n[0] + 15) / 16) * ((g->n[1] + 15) / 16) >= 256;
#define MFVM_NET_LAUNCH(CTV)                                                                                    \
  hipLaunchKernelGGL((closure_net_apply_kernel<CTV>), dim3((g->n[0] + CTV - 1) / CTV, (g->n[1] + CTV - 1) / CTV), \
                     dim3(256), NetTile<CTV>::lds(), s, lay, gamma - 1.0, static_cast<const double*>(U),         \
                     static_cast<const double*>(out), shift, scale, static_cast<const uint16_t*>(wf[0]), b[0],   \
                     static_cast<const uint16_t*>(wf[1]), b[1], static_cast<const uint16_t*>(wf[2]), b[2], osc,  \
                     inv1, inv2, coef, dt, static_cast<double*>(T), flag, yg)
  if (big)
    MFVM_NET_LAUNCH(16);
  else
    MFVM_NET_LAUNCH(8);
#undef MFVM_NET_LAUNCH
  return hipGetLastError();
}

hipError_t fill_int(int* p, long long n, int v, hipStream_t s) {
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace mfvm
