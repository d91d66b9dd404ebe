// Compressible Navier-Stokes kernels: closure, fused RK stage, CFL reduction.
//
// Reference: src/physics/navier_stokes.py and src/numerics/* of
// multiphysics-fvm.  Every expression keeps the reference operation order
// (see the line citations), which makes the float64 instantiation bitwise
// equal to the NumPy reference.
//
// Stage kernel design (gfx950):
//  * one thread per (x1, x2) column of a BX x BY tile; the block marches in
//    x3 over a chunk of planes, keeping the x3 stencil (three planes of the
//    reconstruction fields, the carried left face state, the previous x3
//    face flux and the in-plane partial divergence) in registers;
//  * per plane, each thread reconstructs its own cell along x1 and x2 once
//    (limited slopes shared by both faces of the cell); right-face states are
//    exchanged through LDS, fluxes F.S are computed once per face and
//    exchanged through LDS again for the divergence;
//  * tile-edge cells (the left/right neighbours of the tile in x1 and x2)
//    are reconstructed by helper threads, and their faces computed once;
//  * the stage update (Euler / RK3 / SSPRK3 / RK4 forms) is fused into the
//    epilogue of each cell, and the last stage of a step also reduces the CFL
//    rates of the new state (wavefront shuffles, then LDS, then one partial
//    per block).
#include "ns_core.h"

namespace mfvm {

// Inviscid closure: P on the interior and the face ghosts (edges and corners
// are never read by the RHS).  T stays +-0 and is not written.
template <class R>
__global__ void ns_closure_p_kernel(Layout lay, NSPar p, const R* __restrict__ U, R* __restrict__ C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  const int n1 = lay.n1, n2 = lay.n2, n3 = lay.n3;
  if (i > n1 + 1 || j > n2 + 1) return;
  const int nb = (i == 0 || i == n1 + 1) + (j == 0 || j == n2 + 1) + (k == 0 || k == n3 + 1);
  if (nb > 1) return;
  const long long fs = lay.fs, c = lay.at(i, j, k);
  R rho, v[3];
  load_v(U, c, fs, rho, v);
  C[c] = closure_p(p, rho, v[0], v[1], v[2], U[c + 4 * fs]);
}

// P2 (the closure kernel): uniform power-of-two geometry (set_pow2).

// WP: also write P (the float32 stage passes form P from U themselves).
template <bool UNIF, class R, bool P2, bool WP>
__global__ __launch_bounds__(256) void ns_closure_visc_kernel(Layout lay, Geo geo, NSPar p,
                                                              const R* __restrict__ U,
                                                              R* __restrict__ C, int kz, int remap) {
  if constexpr (P2) set_pow2(geo);
  closure_march_body<true, UNIF, R, WP>(lay, geo, p, U, C, kz, tile_id(remap != 0));
}

// The fused stage kernels take SPEC (spec_args) as the split ones do.
template <int LIM, int RS, bool UNIF, class R, int SPEC>
__global__ __launch_bounds__(256) void ns_stage_kernel(StageArgs a) {
  spec_args(a, SPEC);
  ns_stage_body<LIM, RS, UNIF, R, false, false>(a, grid_tile(a), ptr_src<R>(a));
}

// Compensated FP32 stage: half the registers of the FP64 kernel, two waves per SIMD.
template <int LIM, int RS, bool UNIF, int SPEC>
__global__ __launch_bounds__(256, 2) void ns_stage_comp_kernel(StageArgs a) {
  spec_args(a, SPEC);
  ns_stage_body<LIM, RS, UNIF, float, true, false>(a, grid_tile(a), ptr_src<float>(a));
}

// Split stage (a.xy != NULL): x1/x2 faces (ns_xyrow_kernel), then x3 faces,
// sources and update (ns_z_kernel / ns_z_comp_kernel).  SPEC: see spec_args.

// Row-march form of the x1/x2 pass (ns_xyrow_body): XR_WAVES independent
// waves per workgroup, one (strip, row chunk, plane) task each; BUF:
// buffer-resource reads (arrays below 4 GiB).
template <int LIM, int RS, bool UNIF, class R, bool BUF, int SPEC>
MFVM_DEV void ns_xyrow_entry(StageArgs& a) {
  spec_args(a, SPEC);
  // consecutive workgroups (neighbouring row chunks of a plane, whose halo
  // rows overlap) on one XCD, sharing its L2
  // (the wave index through readfirstlane: the compiler then knows the task,
  // and with it the march's rows and loop exits, to be uniform per wave, and
  // keeps the loop in scalar branches instead of exec masks)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / XR_W);
  const long long task = (long long)xcd_block_1d(a.remap != 0) * XR_WAVES + wave;
  // (P formed from U: the closure before a split stage leaves it out)
  if constexpr (BUF) ns_xyrow_body<LIM, RS, UNIF, R>(a, buf_src<R, true>(a), task);
  else ns_xyrow_body<LIM, RS, UNIF, R>(a, ptr_src<R, true>(a), task);
}
template <int LIM, int RS, bool UNIF, class R, bool BUF, int SPEC>
__global__ __launch_bounds__(XR_W * XR_WAVES) void ns_xyrow_kernel(StageArgs a) {
  ns_xyrow_entry<LIM, RS, UNIF, R, BUF, SPEC>(a);
}
// The float32 pass at four waves per SIMD (128 VGPRs, no spills): 1.200 ->
// 1.130 ms per compensated RT 256^3 step against the 3 waves it gets unbounded.
template <int LIM, int RS, bool UNIF, bool BUF, int SPEC>
__global__ __launch_bounds__(XR_W * XR_WAVES) __attribute__((amdgpu_waves_per_eu(4)))
void ns_xyrow_f32_kernel(StageArgs a) {
  ns_xyrow_entry<LIM, RS, UNIF, float, BUF, SPEC>(a);
}

// (Buffer-resource reads measure the same here as pointer reads.)
template <int LIM, int RS, bool UNIF, int SPEC>
__global__ __launch_bounds__(256) void ns_z_kernel(StageArgs a) {
  spec_args(a, SPEC);
  ns_stage_body<LIM, RS, UNIF, double, false, true>(a, grid_tile(a), ptr_src<double, true>(a));
}

// FIN: the final stage (float64 accumulation and CFL partials), held to four
// waves per SIMD (128 VGPRs, 2 spilled): 0.73 -> 0.62 ms at RT 256^3 against
// the three its 133 VGPRs give unbounded.  The other stages leave that code
// out and fit four (114 VGPRs).
template <int LIM, int RS, bool UNIF, bool FIN, int SPEC>
__global__ __launch_bounds__(256, FIN ? 4 : 3) void ns_z_comp_kernel(StageArgs a) {
  spec_args(a, SPEC);
  ns_stage_body<LIM, RS, UNIF, float, true, true, PtrSrc<float>, FIN ? 1 : 0, true>(a, grid_tile(a), ptr_src<float>(a));
}

// Standalone CFL partials (navier_stokes.py:512-583), same grid as the stage.
template <bool UNIF, class R>
__global__ __launch_bounds__(256) void ns_cfl_kernel(Layout lay, Geo geo, NSPar p, const R* U, const R* lo,
                                                    double* partials, int kz) {
  TileCtx t;
  t.x = blockIdx.x;
  t.y = blockIdx.y;
  t.z = blockIdx.z;
  t.bx = blockDim.x;
  t.by = blockDim.y;
  t.slot = partials + 2 * hw_block_id();
  cfl_tile_body<UNIF, R>(lay, geo, p, U, lo, kz, t);
}

// A learned-closure step's last selection fused with the CFL partials that
// follow it: unless *flag is set, the corrected state T replaces the
// interior of U; the partials are those of the selected state (T, or U when
// the flag is set, whose interior then stays and the stage is counted).  The
// copy into U never meets a read of U: the partials then read T.
template <bool UNIF, class R>
__global__ __launch_bounds__(256) void ns_select_cfl_kernel(Layout lay, Geo geo, NSPar p, const R* __restrict__ T,
                                                           R* U, const int* __restrict__ flag,
                                                           int* __restrict__ count, double* partials, int kz) {
  const bool keep = *flag != 0;
  TileCtx t;
  t.x = blockIdx.x;
  t.y = blockIdx.y;
  t.z = blockIdx.z;
  t.bx = blockDim.x;
  t.by = blockDim.y;
  t.slot = partials + 2 * hw_block_id();
  if (!keep) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int i = 1 + t.x * t.bx + tid % t.bx, j = 1 + t.y * t.by + tid / t.bx;
    const int kb = 1 + t.z * kz, ke = min(kb + kz, lay.n3 + 1);
    if (i <= lay.n1 && j <= lay.n2)
      for (int k = kb; k < ke; ++k) {
        const long long o = lay.at(i, j, k);
#pragma unroll
        for (int q = 0; q < 5; ++q) U[o + q * lay.fs] = T[o + q * lay.fs];
      }
  } else if (hw_block_id() == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    atomicAdd(count, 1);
  }
  cfl_tile_body<UNIF, R>(lay, geo, p, keep ? U : T, nullptr, kz, t);
}

__global__ void ns_dt_finalize_kernel(const double* partials, long long nblk, double cfl,
                                      double smax, double* dt_out) {
  const double dt = dt_reduce_block(partials, nblk, cfl, smax);
  if (threadIdx.x == 0) *dt_out = dt;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

static dim3 ns_grid(const mfvm_grid* g, const Tile& t) {
  return dim3((g->n[0] + t.bx - 1) / t.bx, (g->n[1] + t.by - 1) / t.by, (g->n[2] + t.kz - 1) / t.kz);
}

long long ns_nblocks(const mfvm_grid* g, const mfvm_ns_params* p) {
  const Tile t = ns_tile(g, p);
  const dim3 gr = ns_grid(g, t);
  return (long long)gr.x * gr.y * gr.z;
}

constexpr int CLOSURE_KZ_F64 = 32, CLOSURE_KZ_F32 = 16;

// x3 planes per closure workgroup: the defaults above; MFVM_CLOSURE_KZ
// overrides both (1 to 64).
static int closure_kz(bool f32) {
  static const int env = [] {
    const char* e = getenv("MFVM_CLOSURE_KZ");
    const int v = e ? atoi(e) : 0;
    return v < 0 ? 0 : (v > 64 ? 64 : v);
  }();
  return env > 0 ? env : (f32 ? CLOSURE_KZ_F32 : CLOSURE_KZ_F64);
}

// with_p = false: P is left out -- the float32 stage passes and the float64
// split passes form it from U -- and an inviscid closure launches nothing.
template <class R>
hipError_t launch_closure(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, void* C,
                          hipStream_t s, bool with_p) {
  if (!with_p && p->mu == 0.0) return hipSuccess;
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const Geo geo = make_geo(g);
  const NSPar q = make_ns_par(p);
  const R* u = static_cast<const R*>(U);
  R* c = static_cast<R*>(C);
  const dim3 blk(64, 4, 1);
  const int tk = t_begin(MFVM_TK_CLOSURE, s);
  if (!q.visc) {
    const dim3 grd((g->n[0] + 2 + 63) / 64, (g->n[1] + 2 + 3) / 4, g->n[2] + 2);
    hipLaunchKernelGGL((ns_closure_p_kernel<R>), grd, blk, 0, s, lay, q, u, c);
  } else {
    // x3 planes per workgroup: the float32 kernel (no prefetch) measures
    // 13% faster with chunks of 16 than of 32 at 256^3
    const int kzz = closure_kz(sizeof(R) == sizeof(float));
    const int kz = g->n[2] < kzz ? g->n[2] : kzz;
    const dim3 grd((g->n[0] + 63) / 64, (g->n[1] + 3) / 4, (g->n[2] + kz - 1) / kz);
    const size_t lds = 3 * (4 + 2) * (64 + 2) * sizeof(R);
    const int remap = xcd_remap_enabled();
#define MFVM_CLOSURE_VISC(WPV)                                                                        \
    if (g->uniform && pow2_geo(geo))                                                                   \
      hipLaunchKernelGGL((ns_closure_visc_kernel<true, R, true, WPV>), grd, blk, lds, s, lay, geo, q, u, c, kz,  \
                         remap);                                                                       \
    else if (g->uniform)                                                                               \
      hipLaunchKernelGGL((ns_closure_visc_kernel<true, R, false, WPV>), grd, blk, lds, s, lay, geo, q, u, c, kz, \
                         remap);                                                                       \
    else                                                                                               \
      hipLaunchKernelGGL((ns_closure_visc_kernel<false, R, false, WPV>), grd, blk, lds, s, lay, geo, q, u, c, kz, \
                         remap);
    if (with_p) {
      MFVM_CLOSURE_VISC(true)
    } else {
      MFVM_CLOSURE_VISC(false)
    }
#undef MFVM_CLOSURE_VISC
  }
  t_end(tk, s);
  return hipGetLastError();
}

// Rows per chunk of the row-march x1/x2 pass: 16 (FP64 and FP32 both 1-6%
// faster than 24 or 32 at 256^3), halved (down to 2) while the pass would
// have fewer than 4096 wave tasks -- a 2D 256^2 grid has 64 of 16 rows, which
// leave most CUs idle; MFVM_XYROW_ROWS overrides (2 to XR_ROWS).
static int xyrow_rows(long long strips, int n2, long long planes) {
  static const int env = [] {
    const char* e = getenv("MFVM_XYROW_ROWS");
    const int v = e ? atoi(e) : 0;
    return v <= 0 ? 0 : (v < 2 ? 2 : (v > XR_ROWS ? XR_ROWS : v));
  }();
  if (env) return env;
  int r = 16;
  while (r > 2 && strips * ((n2 + r - 1) / r) * planes < 4096) r /= 2;
  return r;
}

template <int LIM, int RS, bool UNIF, class R, int SPEC>
static void xyrow_launch(const StageArgs& a0, bool buf, hipStream_t s) {
  StageArgs a = a0;
  const long long ns = (a.lay.n1 + XR_W - 1) / XR_W;
  const int m1 = a.xm1 > 0 ? a.xm1 : a.lay.n3 + 1;
  a.kz = xyrow_rows(ns, a.lay.n2, m1 - a.xm0);
  const long long nc = (a.lay.n2 + a.kz - 1) / a.kz;
  const long long ntask = ns * nc * (m1 - a.xm0);
  const dim3 g((unsigned)((ntask + XR_WAVES - 1) / XR_WAVES)), b(XR_W * XR_WAVES);
  const size_t lds = (size_t)XR_WAVES * XR_LDS_ELEMS * sizeof(R);
  if constexpr (sizeof(R) == sizeof(float)) {
    if (buf) hipLaunchKernelGGL((ns_xyrow_f32_kernel<LIM, RS, UNIF, true, SPEC>), g, b, lds, s, a);
    else hipLaunchKernelGGL((ns_xyrow_f32_kernel<LIM, RS, UNIF, false, SPEC>), g, b, lds, s, a);
  } else {
    if (buf) hipLaunchKernelGGL((ns_xyrow_kernel<LIM, RS, UNIF, R, true, SPEC>), g, b, lds, s, a);
    else hipLaunchKernelGGL((ns_xyrow_kernel<LIM, RS, UNIF, R, false, SPEC>), g, b, lds, s, a);
  }
}

template <int LIM, int RS, class R>
static void xyrow_any(const StageArgs& a, int spec, bool unif, bool buf, hipStream_t s) {
  if (spec == 2) xyrow_launch<LIM, RS, true, R, 2>(a, buf, s);
  else if (spec == 1) xyrow_launch<LIM, RS, true, R, 1>(a, buf, s);
  else if (unif) xyrow_launch<LIM, RS, true, R, 0>(a, buf, s);
  else xyrow_launch<LIM, RS, false, R, 0>(a, buf, s);
}

// The x3 pass: float64, or the compensated float32 form (final stage or not).
template <int LIM, int RS, bool UNIF, int SPEC>
static void z_launch(bool comp, bool fin, dim3 g, dim3 b, size_t red, hipStream_t s, const StageArgs& a) {
  if (comp && fin) hipLaunchKernelGGL((ns_z_comp_kernel<LIM, RS, UNIF, true, SPEC>), g, b, red, s, a);
  else if (comp) hipLaunchKernelGGL((ns_z_comp_kernel<LIM, RS, UNIF, false, SPEC>), g, b, red, s, a);
  else hipLaunchKernelGGL((ns_z_kernel<LIM, RS, UNIF, SPEC>), g, b, red, s, a);
}

template <int LIM, int RS, class R>
static hipError_t launch_stage_t(const StageArgs& a, bool unif, dim3 grd, dim3 blk, size_t lds, hipStream_t s) {
  const bool comp = a.comp != MFVM_COMP_NONE;
  if (a.xy && comp == (sizeof(R) == sizeof(float))) {
    // split stage: x1/x2 faces (timed as the stage), then x3 faces, sources,
    // update and CFL (timed as "other")
    const size_t red = 2 * sizeof(double) * ((blk.x * blk.y + 63) / 64);
    StageArgs ax = a;
    ax.partials = nullptr;
    const bool p2 = pow2_geo(a.geo);
    const int spec = unif && p2 ? (a.p.visc ? 2 : 1) : 0;
    const bool buf = buf_src_ok(a.lay, sizeof(R));
    const bool fin = a.comp >= MFVM_COMP_FINAL_SET;
    const int tx = t_begin(MFVM_TK_STAGE, s);
    xyrow_any<LIM, RS, R>(ax, spec, unif, buf, s);
    t_end(tx, s);
    const int tz = t_begin(MFVM_TK_OTHER, s);
    if (spec == 2) z_launch<LIM, RS, true, 2>(comp, fin, grd, blk, red, s, a);
    else if (spec == 1) z_launch<LIM, RS, true, 1>(comp, fin, grd, blk, red, s, a);
    else if (unif) z_launch<LIM, RS, true, 0>(comp, fin, grd, blk, red, s, a);
    else z_launch<LIM, RS, false, 0>(comp, fin, grd, blk, red, s, a);
    t_end(tz, s);
    return hipGetLastError();
  }
  const int tk = t_begin(MFVM_TK_STAGE, s);
  const bool p2 = pow2_geo(a.geo);
  const int spec = unif && p2 ? (a.p.visc ? 2 : 1) : 0;
  if (a.comp != MFVM_COMP_NONE) {
    if (spec == 2)
      hipLaunchKernelGGL((ns_stage_comp_kernel<LIM, RS, true, 2>), grd, blk, lds, s, a);
    else if (spec == 1)
      hipLaunchKernelGGL((ns_stage_comp_kernel<LIM, RS, true, 1>), grd, blk, lds, s, a);
    else if (unif)
      hipLaunchKernelGGL((ns_stage_comp_kernel<LIM, RS, true, 0>), grd, blk, lds, s, a);
    else
      hipLaunchKernelGGL((ns_stage_comp_kernel<LIM, RS, false, 0>), grd, blk, lds, s, a);
  } else if (spec == 2)
    hipLaunchKernelGGL((ns_stage_kernel<LIM, RS, true, R, 2>), grd, blk, lds, s, a);
  else if (spec == 1)
    hipLaunchKernelGGL((ns_stage_kernel<LIM, RS, true, R, 1>), grd, blk, lds, s, a);
  else if (unif)
    hipLaunchKernelGGL((ns_stage_kernel<LIM, RS, true, R, 0>), grd, blk, lds, s, a);
  else
    hipLaunchKernelGGL((ns_stage_kernel<LIM, RS, false, R, 0>), grd, blk, lds, s, a);
  t_end(tk, s);
  return hipGetLastError();
}

template <class R>
hipError_t launch_stage(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_stage* st,
                        hipStream_t s) {
  const Tile t = ns_tile(g, p);
  StageArgs a;
  a.lay = make_layout(g->n[0], g->n[1], g->n[2]);
  a.il = make_ilayout(g->n[0], g->n[1], g->n[2]);
  a.geo = make_geo(g);
  a.p = make_ns_par(p);
  a.U = st->U;
  a.C = st->C;
  a.U0 = st->U0;
  a.out = st->out;
  a.kacc = st->kacc;
  a.update = st->update;
  a.acc = st->acc;
  a.c0 = st->c0;
  a.c1 = st->c1;
  a.acc_coef = st->acc_coef;
  a.dt = st->dt;
  a.dt_dev = st->dt_dev;
  a.partials = st->partials;
  a.kz = t.kz;
  a.remap = xcd_remap_enabled();
  a.lo = st->lo;
  a.dinc = st->dinc;
  a.snext = st->snext;
  a.comp = st->comp;
  a.xy = st->xy;
  const dim3 blk(t.bx, t.by, 1);
  const dim3 grd = ns_grid(g, t);
  const size_t lds = stage_lds_bytes(t.bx, t.by, sizeof(R));
#define MFVM_DISPATCH_RS(LIMV)                                                     \
  return p->riemann == MFVM_RS_HLL ? launch_stage_t<LIMV, MFVM_RS_HLL, R>(a, g->uniform, grd, blk, lds, s) \
                                   : launch_stage_t<LIMV, MFVM_RS_RUSANOV, R>(a, g->uniform, grd, blk, lds, s);
  switch (p->limiter) {
    case MFVM_LIM_NONE: MFVM_DISPATCH_RS(MFVM_LIM_NONE)
    case MFVM_LIM_MINMOD: MFVM_DISPATCH_RS(MFVM_LIM_MINMOD)
    case MFVM_LIM_MC: MFVM_DISPATCH_RS(MFVM_LIM_MC)
    default: MFVM_DISPATCH_RS(MFVM_LIM_VANLEER)
  }
#undef MFVM_DISPATCH_RS
}

template <class R>
hipError_t launch_cfl_partials(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, const void* lo,
                               double* partials, hipStream_t s) {
  const Tile t = ns_tile(g, p);
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const Geo geo = make_geo(g);
  const NSPar q = make_ns_par(p);
  const dim3 blk(t.bx, t.by, 1);
  const dim3 grd = ns_grid(g, t);
  const size_t red = 2 * sizeof(double) * ((t.bx * t.by + 63) / 64);
  if (g->uniform)
    hipLaunchKernelGGL((ns_cfl_kernel<true, R>), grd, blk, red, s, lay, geo, q, static_cast<const R*>(U),
                       static_cast<const R*>(lo), partials, t.kz);
  else
    hipLaunchKernelGGL((ns_cfl_kernel<false, R>), grd, blk, red, s, lay, geo, q, static_cast<const R*>(U),
                       static_cast<const R*>(lo), partials, t.kz);
  return hipGetLastError();
}

// ns_select_cfl_kernel: launch_cfl_partials of the selected state (float64).
hipError_t launch_select_cfl_partials(const mfvm_grid* g, const mfvm_ns_params* p, const void* T, void* U,
                                      const int* flag, int* count, double* partials, hipStream_t s) {
  const Tile t = ns_tile(g, p);
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const Geo geo = make_geo(g);
  const NSPar q = make_ns_par(p);
  const dim3 blk(t.bx, t.by, 1);
  const dim3 grd = ns_grid(g, t);
  const size_t red = 2 * sizeof(double) * ((t.bx * t.by + 63) / 64);
  const double* tt = static_cast<const double*>(T);
  double* u = static_cast<double*>(U);
  if (g->uniform)
    hipLaunchKernelGGL((ns_select_cfl_kernel<true, double>), grd, blk, red, s, lay, geo, q, tt, u, flag, count,
                       partials, t.kz);
  else
    hipLaunchKernelGGL((ns_select_cfl_kernel<false, double>), grd, blk, red, s, lay, geo, q, tt, u, flag, count,
                       partials, t.kz);
  return hipGetLastError();
}

template <class R>
hipError_t launch_cfl(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, const void* lo, double cfl,
                      double* partials, double* dt_out, hipStream_t s) {
  const hipError_t e = launch_cfl_partials<R>(g, p, U, lo, partials, s);
  if (e != hipSuccess) return e;
  const long long nblk = ns_nblocks(g, p);
  hipLaunchKernelGGL(ns_dt_finalize_kernel, dim3(1), dim3(1024), 0, s, partials, nblk, cfl, g->smax,
                     dt_out);
  return hipGetLastError();
}

// dt of the next step appended to a history: dt = min(cfl dt, t_end - t),
// hist[*step] = dt, *dt_cur = dt, *t_cur += dt, ++*step (replayable from a
// captured graph; t_end = +inf: no clipping).  clear != NULL: clear[0..nclear)
// are set to 0 as well (per-step flags of the next step, one launch fewer).
__global__ void ns_dt_append_kernel(const double* partials, long long nblk, double cfl, double smax, double t_end,
                                    double* hist, double* dt_cur, double* t_cur, int* step, int* clear,
                                    int nclear) {
  double dt = dt_reduce_block(partials, nblk, cfl, smax);
  if (threadIdx.x == 0) {
    const double t = *t_cur;
    if (t_end - t < dt) dt = t_end > t ? t_end - t : 0.0;
    const int n = *step;
    hist[n] = dt;
    *dt_cur = dt;
    *t_cur = t + dt;
    *step = n + 1;
  }
  if (clear && (int)threadIdx.x < nclear) clear[threadIdx.x] = 0;
}

hipError_t launch_dt_finalize(const mfvm_grid* g, const mfvm_ns_params* p, const double* partials,
                              double cfl, double* dt_out, hipStream_t s) {
  const long long nblk = ns_nblocks(g, p);
  hipLaunchKernelGGL(ns_dt_finalize_kernel, dim3(1), dim3(1024), 0, s, partials, nblk, cfl, g->smax,
                     dt_out);
  return hipGetLastError();
}

hipError_t launch_dt_append(const mfvm_grid* g, const mfvm_ns_params* p, const double* partials, double cfl,
                            double t_end, double* hist, double* dt_cur, double* t_cur, int* step, int* clear,
                            int nclear, hipStream_t s) {
  if (nclear > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ns_dt_append_kernel, dim3(1), dim3(1024), 0, s, partials, ns_nblocks(g, p), cfl, g->smax,
                     t_end, hist, dt_cur, t_cur, step, clear, nclear);
  return hipGetLastError();
}

template hipError_t launch_closure<double>(const mfvm_grid*, const mfvm_ns_params*, const void*, void*, hipStream_t, bool);
template hipError_t launch_closure<float>(const mfvm_grid*, const mfvm_ns_params*, const void*, void*, hipStream_t, bool);
template hipError_t launch_stage<double>(const mfvm_grid*, const mfvm_ns_params*, const mfvm_stage*, hipStream_t);
template hipError_t launch_stage<float>(const mfvm_grid*, const mfvm_ns_params*, const mfvm_stage*, hipStream_t);
template hipError_t launch_cfl<double>(const mfvm_grid*, const mfvm_ns_params*, const void*, const void*, double, double*, double*, hipStream_t);
template hipError_t launch_cfl<float>(const mfvm_grid*, const mfvm_ns_params*, const void*, const void*, double, double*, double*, hipStream_t);
template hipError_t launch_cfl_partials<double>(const mfvm_grid*, const mfvm_ns_params*, const void*, const void*, double*, hipStream_t);

}  // namespace mfvm
