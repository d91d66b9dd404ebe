// Diagnostics of one ghosted state: conserved totals, extrema, a positivity
// monitor and errors against a second (reference) state, reduced in float64
// over the interior cells.  The reference framework has no run-time monitor;
// these reductions let a long GPU run be checked without copying the state
// to the host.
//
// Reduction order is fixed: every block reduces the cells of a fixed
// grid-stride set (by a fixed butterfly within each wave, then the waves in
// order), and one block sums the per-block partials in a fixed order.  The
// results are therefore identical from run to run for the same input.
//
// The pressure of the positivity monitor is the reference's ideal-gas closure
// expression, operand for operand: P = ((ga-1)*rho) * (E - (vx1^2+vx2^2+vx3^2)/2)
// with vx = m/rho and E = rhoE/rho (navier_stokes.py:94-100).

#include "common.h"

namespace mfvm {

constexpr int kDiagBlock = 256;
constexpr int kDiagMaxBlocks = 1280;  // 5 per CU: one resident round at 5 waves per SIMD
constexpr int kDiagCellsPerThread = 4;

struct DiagArgs {
  Layout lay;
  Geo geo;
  unsigned n1, n2, ncells;
  const void* U;
  const float* lo;
  const double* ref;
  double gamma;  // > 0: Navier-Stokes positivity monitor (5 fields)
  double* partials;
};

// Combine of one slot kind: 0 sum, 1 min, 2 max.
MFVM_DEV double d_comb(int op, double a, double b) {
  return op == 0 ? a + b : (op == 1 ? nmin(a, b) : nmax(a, b));
}
MFVM_DEV double d_ident(int op) { return op == 0 ? 0.0 : (op == 1 ? __builtin_inf() : -__builtin_inf()); }

template <int NV> constexpr int diag_nout() { return 5 * NV + 3; }

// Slots: [0,NV) sum U*V, [NV,2NV) min U, [2NV,3NV) max U, [3NV,4NV) sum |U-ref|*V,
// [4NV,5NV) max |U-ref|, 5NV total V, 5NV+1 min P, 5NV+2 count of cells
// with !(rho > 0) or !(P > 0).
template <int NV> MFVM_DEV constexpr int slot_op(int s) {
  return s < NV ? 0 : s < 2 * NV ? 1 : s < 3 * NV ? 2 : s < 4 * NV ? 0 : s < 5 * NV ? 2 : s == 5 * NV + 1 ? 1 : 0;
}

// Reduce every slot of v over the block in a fixed order (a butterfly in
// each wave, then the waves in order) with one barrier; thread s < S writes
// slot s to out.  sh holds (block / 64) * S doubles.
template <int S, int NV>
MFVM_DEV void block_reduce_all(double (&v)[S], double* sh, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double x = v[s];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = d_comb(slot_op<NV>(s), x, __shfl_xor(x, m, 64));
    if (lane == 0) sh[wv * S + s] = x;
  }
  __syncthreads();
  if (threadIdx.x < S) {
    const int s = threadIdx.x, op = slot_op<NV>(s);
    double r = sh[s];
    for (int w = 1; w < kDiagBlock / 64; ++w) r = d_comb(op, r, sh[w * S + s]);
    out[s] = r;
  }
}

// One cell's values: float64 of U (+ lo), the reference values and V.
template <int NV, class R, bool UNIF>
MFVM_DEV void diag_load(const DiagArgs& a, unsigned c, double (&u)[NV], double (&rf)[NV], double& V) {
  const R* U = static_cast<const R*>(a.U);
  const long long fs = a.lay.fs;
  const unsigned i = c % a.n1, r = c / a.n1;
  const unsigned j = r % a.n2, k = r / a.n2;
  const long long o = a.lay.at(i + 1, j + 1, k + 1);
  V = g_vol<UNIF, double>(a.geo, i + 1, j + 1, k + 1);
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    u[q] = double(U[o + q * fs]);
    if (a.lo) u[q] = u[q] + double(a.lo[o + q * fs]);
    rf[q] = a.ref ? a.ref[o + q * fs] : 0.0;
  }
}

template <int NV>
MFVM_DEV void diag_acc(const DiagArgs& a, const double (&u)[NV], const double (&rf)[NV], double V,
                       double (&acc)[diag_nout<NV>()]) {
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    acc[q] = acc[q] + u[q] * V;
    acc[NV + q] = nmin(acc[NV + q], u[q]);
    acc[2 * NV + q] = nmax(acc[2 * NV + q], u[q]);
    if (a.ref) {
      const double e = fabs(u[q] - rf[q]);
      acc[3 * NV + q] = acc[3 * NV + q] + e * V;
      acc[4 * NV + q] = nmax(acc[4 * NV + q], e);
    }
  }
  acc[5 * NV] = acc[5 * NV] + V;
  if constexpr (NV == 5) {
    if (a.gamma > 0.0) {
      const double rho = u[0];
      const double v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho, E = u[4] / rho;
      const double P = ((a.gamma - 1.0) * rho) * (E - (((v1 * v1) + (v2 * v2)) + (v3 * v3)) / 2.0);
      acc[5 * NV + 1] = nmin(acc[5 * NV + 1], P);
      if (!(rho > 0.0) || !(P > 0.0)) acc[5 * NV + 2] = acc[5 * NV + 2] + 1.0;
    }
  }
}

// Each thread takes the cells c, c + stride, ... (grid-stride, x1 fastest
// across the lanes), then the block reduces.
template <int NV, class R, bool UNIF>
__global__ __launch_bounds__(kDiagBlock) void diag_kernel(DiagArgs a) {
  constexpr int S = diag_nout<NV>();
  __shared__ double sh[(kDiagBlock / 64) * S];
  double acc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = d_ident(slot_op<NV>(s));
  const unsigned stride = gridDim.x * kDiagBlock;
  for (unsigned c = blockIdx.x * kDiagBlock + threadIdx.x; c < a.ncells; c += stride) {
    double u[NV], rf[NV], V;
    diag_load<NV, R, UNIF>(a, c, u, rf, V);
    diag_acc<NV>(a, u, rf, V, acc);
  }
  block_reduce_all<S, NV>(acc, sh, a.partials + (long long)blockIdx.x * S);
}

// One block: thread t combines the partial rows t, t + block, ... (each row
// read whole), then the block reduces; the L1 sums are divided by the total
// volume and the slots a call did not ask for are zeroed.
template <int NV>
__global__ __launch_bounds__(kDiagBlock) void diag_finalize_kernel(const double* partials, int nblk, int has_ref,
                                                                   int has_p, double* out) {
  constexpr int S = diag_nout<NV>();
  __shared__ double sh[(kDiagBlock / 64) * S];
  __shared__ double res[S];
  double acc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = d_ident(slot_op<NV>(s));
  for (int b = threadIdx.x; b < nblk; b += kDiagBlock) {
    double row[S];
#pragma unroll
    for (int s = 0; s < S; ++s) row[s] = partials[(long long)b * S + s];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = d_comb(slot_op<NV>(s), acc[s], row[s]);
  }
  block_reduce_all<S, NV>(acc, sh, res);
  __syncthreads();
  if (threadIdx.x < S) {
    const int s = threadIdx.x;
    double v = res[s];
    if (s >= 3 * NV && s < 4 * NV) v = has_ref ? v / res[5 * NV] : 0.0;
    else if (s >= 4 * NV && s < 5 * NV) v = has_ref ? v : 0.0;
    else if (s > 5 * NV) v = has_p ? v : 0.0;
    out[s] = v;
  }
}

long long diag_nblocks(const mfvm_grid* g) {
  const long long nc = (long long)g->n[0] * g->n[1] * g->n[2];
  long long nb = (nc + kDiagBlock * kDiagCellsPerThread - 1) / (kDiagBlock * kDiagCellsPerThread);
  return nb < 1 ? 1 : (nb > kDiagMaxBlocks ? kDiagMaxBlocks : nb);
}

template <int NV, class R>
static hipError_t diag_launch_t(const mfvm_grid* g, const DiagArgs& a, int nblk, double* out, hipStream_t s) {
  const int tok = t_begin(MFVM_TK_OTHER, s);
  if (g->uniform) hipLaunchKernelGGL((diag_kernel<NV, R, true>), dim3(nblk), dim3(kDiagBlock), 0, s, a);
  else hipLaunchKernelGGL((diag_kernel<NV, R, false>), dim3(nblk), dim3(kDiagBlock), 0, s, a);
  t_end(tok, s);
  hipLaunchKernelGGL((diag_finalize_kernel<NV>), dim3(1), dim3(kDiagBlock), 0, s, a.partials, nblk,
                     a.ref != nullptr, a.gamma > 0.0, out);
  return hipGetLastError();
}

template <int NV>
static hipError_t diag_launch_nv(const mfvm_grid* g, int dtype, const DiagArgs& a, int nblk, double* out,
                                 hipStream_t s) {
  return dtype == MFVM_F64 ? diag_launch_t<NV, double>(g, a, nblk, out, s)
                           : diag_launch_t<NV, float>(g, a, nblk, out, s);
}

// Argument checks are the caller's (mfvm_diagnostics).
hipError_t diag_launch(const mfvm_grid* g, int nvar, int dtype, const void* U, const void* lo, const void* ref,
                       double gamma, double* partials, double* out, hipStream_t s) {
  DiagArgs a;
  a.lay = make_layout(g->n[0], g->n[1], g->n[2]);
  a.geo = make_geo(g);
  a.n1 = g->n[0];
  a.n2 = g->n[1];
  a.ncells = (unsigned)((long long)g->n[0] * g->n[1] * g->n[2]);
  a.U = U;
  a.lo = static_cast<const float*>(lo);
  a.ref = static_cast<const double*>(ref);
  a.gamma = gamma;
  a.partials = partials;
  const int nblk = (int)diag_nblocks(g);
  switch (nvar) {
    case 1: return diag_launch_nv<1>(g, dtype, a, nblk, out, s);
    case 2: return diag_launch_nv<2>(g, dtype, a, nblk, out, s);
    case 3: return diag_launch_nv<3>(g, dtype, a, nblk, out, s);
    case 4: return diag_launch_nv<4>(g, dtype, a, nblk, out, s);
    case 5: return diag_launch_nv<5>(g, dtype, a, nblk, out, s);
    default: return diag_launch_nv<6>(g, dtype, a, nblk, out, s);
  }
}

}  // namespace mfvm
