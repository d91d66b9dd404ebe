// Compressible Navier-Stokes device code shared by the single-grid kernels
// (ns.hip) and the batched ensemble kernels (ensemble.hip): closure,
// reconstruction, fluxes, sources, the fused RK stage body and the CFL
// reduction.
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
#pragma once

#include <cstdlib>

#include "common.h"

namespace mfvm {


enum { C_P = 0, C_T11 = 1, C_T22 = 2, C_T33 = 3, C_T12 = 4, C_T13 = 5, C_T23 = 6 };
#define NS_NC 7   // closure fields
#define NQ 9      // reconstruction fields per direction: U(5), P, row d of T

// Closure field holding T[d][e] (symmetric: T21 is T12 etc., navier_stokes.py:137-139).
MFVM_DEV constexpr int trow(int d, int e) {
  return d == e ? (d == 0 ? C_T11 : (d == 1 ? C_T22 : C_T33))
                : ((d + e == 1) ? C_T12 : ((d + e == 2) ? C_T13 : C_T23));
}

struct NSPar {
  double gamma, gm1, mu, g[3];
  int visc;  // mu != 0: viscous fluxes and closure (T == +-0 otherwise)
  int grav;  // any g != 0: gravity source (+-0 otherwise)
};

struct StageArgs {
  Layout lay;
  ILayout il;
  Geo geo;
  NSPar p;
  const void* U;
  const void* C;
  const void* U0;
  void* out;
  void* kacc;
  int update, acc;
  double c0, c1, acc_coef, dt;
  const double* dt_dev;
  double* partials;
  int kz;
  int remap;  // XCD-aware tile order (tile_id)
  // split stages: interior [5] in-plane divergence written by the x1/x2
  // kernel and read by the x3 kernel (in the stage's precision R)
  void* xy;
  // compensated FP32 stages (MFVM_COMP_*): lo = compensation of U0 (ghosted),
  // dinc = increment over U0 (interior), snext = next stage state (ghosted)
  void* lo;
  void* dinc;
  void* snext;
  int comp;
  // The row-march x1/x2 pass covers planes [xm0, xm1) (xm1 = 0: to the last).
  int xm0 = 1, xm1 = 0;
};

// ---------------------------------------------------------------------------
// Closure (navier_stokes.py:44-156): P on interior + face ghosts, T interior.
// ---------------------------------------------------------------------------
// :100  P = (ga-1)*rho*(E - (v1**2+v2**2+v3**2)/2), from the cell's U.
template <class R>
MFVM_DEV R closure_p_e(const NSPar& p, R rho, R v1, R v2, R v3, R E) {
  return (R(p.gm1) * rho) * (E - (((v1 * v1) + (v2 * v2)) + (v3 * v3)) / R(2));
}
template <class R>
MFVM_DEV R closure_p(const NSPar& p, R rho, R v1, R v2, R v3, R rhoE) {
  return closure_p_e(p, rho, v1, v2, v3, qdiv(rhoE, rho));
}

// :94-97 primitive velocity of the cell at offset o.
template <class R>
MFVM_DEV void load_v(const R* __restrict__ U, long long o, long long fs, R& rho, R v[3]) {
  rho = U[o];
  v[0] = qdiv(U[o + fs], rho);
  v[1] = qdiv(U[o + 2 * fs], rho);
  v[2] = qdiv(U[o + 3 * fs], rho);
}

// The five conserved values of the cell at offset o (zeros when !ok).
template <class R>
MFVM_DEV void load_u5(const R* __restrict__ U, long long o, long long fs, bool ok, R u[5]) {
#pragma unroll
  for (int f = 0; f < 5; ++f) u[f] = ok ? U[o + f * fs] : R(0);
}

// :111-139 Stokes tensor from the centred differences dv[a][b] = v_a(+b) -
// v_a(-b) and the cell's own velocity; writes the six unique entries.
template <bool UNIF, class R>
MFVM_DEV void closure_t(const Geo& geo, const NSPar& p, int i, int j, int k, const R dv[3][3],
                        const R v[3], R* __restrict__ C, long long c, long long fs) {
  R g[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    if (geo.lpow2[b]) {
#pragma unroll
      for (int a = 0; a < 3; ++a) g[a][b] = dv[a][b] * R(0.5 * geo.linv[b]);
    } else {
      const R twoL = R(2) * g_len<UNIF, R>(geo, b, i, j, k);
#pragma unroll
      for (int a = 0; a < 3; ++a) g[a][b] = qdiv(dv[a][b], twoL);
    }
  }
  const R v1 = v[0], v2 = v[1], v3 = v[2];
  const R mu = R(p.mu);
  const R two3 = R(2.0 / 3.0);
  R t11, t22, t33, t12, t13, t23;
  if (UNIF || !geo.curv) {
    // Cartesian: every Lame term is exactly +-0 and drops out of the sums.
    const R divv = (g[0][0] + g[1][1]) + g[2][2];
    t11 = mu * ((R(2) * g[0][0]) - (two3 * divv));
    t22 = mu * ((R(2) * g[1][1]) - (two3 * divv));
    t33 = mu * ((R(2) * g[2][2]) - (two3 * divv));
    t12 = mu * (g[0][1] + g[1][0]);
    t13 = mu * (g[0][2] + g[2][0]);
    t23 = mu * (g[1][2] + g[2][1]);
  } else {
    const int a = i - 1, b = j - 1, e = k - 1;
    const R h1 = met<R>(geo, MFVM_MET_H1, a, b, e), h2 = met<R>(geo, MFVM_MET_H2, a, b, e),
            h3 = met<R>(geo, MFVM_MET_H3, a, b, e);
    const R d1h2 = met<R>(geo, MFVM_MET_D1H2, a, b, e), d1h3 = met<R>(geo, MFVM_MET_D1H3, a, b, e);
    const R d2h1 = met<R>(geo, MFVM_MET_D2H1, a, b, e), d2h3 = met<R>(geo, MFVM_MET_D2H3, a, b, e);
    const R d3h1 = met<R>(geo, MFVM_MET_D3H1, a, b, e), d3h2 = met<R>(geo, MFVM_MET_D3H2, a, b, e);
    // :128
    const R divv = (((((g[0][0] + g[1][1]) + g[2][2]) + (v1 * (d1h2 + d1h3)) / h1) +
                     (v2 * (d2h1 + d2h3)) / h2) +
                    (v3 * (d3h1 + d3h2)) / h3);
    // :131-136
    t11 = mu * ((R(2) * g[0][0]) - (two3 * divv));
    t22 = mu * ((R(2) * (g[1][1] + (v1 * d1h2) / h1)) - (two3 * divv));
    t33 = mu * ((R(2) * ((g[2][2] + (v1 * d1h3) / h1) + (v2 * d2h3) / h2)) - (two3 * divv));
    t12 = mu * (((g[0][1] + g[1][0]) - (v2 * d1h2) / h1) + (v1 * d2h1) / h2);
    t13 = mu * (((g[0][2] + g[2][0]) - (v3 * d1h3) / h1) + (v1 * d3h1) / h3);
    t23 = mu * (((g[1][2] + g[2][1]) - (v3 * d2h3) / h2) + (v2 * d3h2) / h3);
  }
  C[c + C_T11 * fs] = t11;
  C[c + C_T22 * fs] = t22;
  C[c + C_T33 * fs] = t33;
  C[c + C_T12 * fs] = t12;
  C[c + C_T13 * fs] = t13;
  C[c + C_T23 * fs] = t23;
}

// Closure marching in x3.  One thread per (x1, x2) column of a BX x BY tile
// (the workgroup shape) over a chunk of kz planes.  WP = false: P is not
// written (the float32 stage passes form it from U where they read it).  Each velocity is computed
// once: the column keeps v at k-1, k, k+1 in registers, and plane k of the
// tile plus a one-cell ring is exchanged through LDS for the x1 and x2
// differences.  The ring cells are computed by helper threads; where a ring
// cell is a face ghost of the domain, its helper also writes its P.  Without
// viscosity only P is formed (on the same cells).  LDS: 3 (BY+2)(BX+2) R.
template <bool VISC, bool UNIF, class R, bool WP = true>
MFVM_DEV void closure_march_body(const Layout& lay, const Geo& geo, const NSPar& p,
                                 const R* __restrict__ U, R* __restrict__ C, int kz, TileId tl) {
  if constexpr (!WP && !VISC) return;
  extern __shared__ double smem_raw[];
  R* vs = reinterpret_cast<R*>(smem_raw);  // [3][BY+2][BX+2]
  const int BX = blockDim.x, BY = blockDim.y;
  const int sx = BX + 2, splane = (BY + 2) * sx;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
  const int n1 = lay.n1, n2 = lay.n2, n3 = lay.n3;
  const long long fs = lay.fs;
  const int i0 = 1 + tl.x * BX, j0 = 1 + tl.y * BY;
  const int kb = 1 + tl.z * kz, ke = min(kb + kz, n3 + 1);
  const int ilast = min(i0 + BX - 1, n1), jlast = min(j0 + BY - 1, n2);
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i <= n1 && j <= n2;

  // Ring cell of this helper thread (h < 2*BY + 2*BX), in global and LDS
  // coordinates; hv = false when the ring slot lies beyond the tile's cells.
  int hi = 0, hj = 0, hli = 0, hlj = 0;
  bool hv = false;
  {
    const int h = tid;
    if (h < BY) {
      hj = j0 + h; hi = i0 - 1; hli = 0; hlj = h + 1; hv = hj <= jlast;
    } else if (h < 2 * BY) {
      hj = j0 + h - BY; hi = ilast + 1; hli = hi - i0 + 1; hlj = h - BY + 1; hv = hj <= jlast;
    } else if (h < 2 * BY + BX) {
      hi = i0 + h - 2 * BY; hj = j0 - 1; hli = hi - i0 + 1; hlj = 0; hv = hi <= ilast;
    } else if (h < 2 * BY + 2 * BX) {
      hi = i0 + h - 2 * BY - BX; hj = jlast + 1; hli = hi - i0 + 1; hlj = hj - j0 + 1; hv = hi <= ilast;
    }
  }
  const bool hghost = hv && (hi == 0 || hi == n1 + 1 || hj == 0 || hj == n2 + 1);

  R vm[3] = {R(0), R(0), R(0)}, vc[3] = {R(0), R(0), R(0)}, vn[3] = {R(0), R(0), R(0)};
  if (own) {
    R rho;
    const long long om = lay.at(i, j, kb - 1), oc = lay.at(i, j, kb);
    if (VISC || (WP && kb - 1 == 0)) {
      load_v(U, om, fs, rho, vm);
      if (WP && kb - 1 == 0) C[om] = closure_p(p, rho, vm[0], vm[1], vm[2], U[om + 4 * fs]);
    }
    load_v(U, oc, fs, rho, vc);
    if (WP) C[oc] = closure_p(p, rho, vc[0], vc[1], vc[2], U[oc + 4 * fs]);
  }
  // PF (float64): the U values of plane k+1 of the own column (un) and of
  // plane k of the ring cell (uh) are loaded one iteration ahead, so that
  // their latency overlaps the previous plane's work (9% faster).  The
  // float32 kernel runs more waves per SIMD and measures slower with it.
  constexpr bool PF = sizeof(R) == sizeof(double);
  const bool hload = hv && (VISC || (WP && hghost));
  auto own_load = [&](int k) { return own && (VISC || k + 1 < ke || k + 1 == n3 + 1); };
  R un[5], uh[5];
  if (PF) {
    load_u5(U, lay.at(i, j, kb + 1), fs, kb < ke && own_load(kb), un);
    load_u5(U, lay.at(hi, hj, kb), fs, kb < ke && hload, uh);
  }
  for (int k = kb; k < ke; ++k) {
    if (VISC) __syncthreads();  // the previous plane's LDS reads are done
    if (VISC && own) {
      const int l = (ty + 1) * sx + (tx + 1);
#pragma unroll
      for (int a = 0; a < 3; ++a) vs[a * splane + l] = vc[a];
    }
    if (hload) {
      const long long oh = lay.at(hi, hj, k);
      if (!PF) load_u5(U, oh, fs, true, uh);
      const R rho = uh[0];
      R vh[3];
      vh[0] = qdiv(uh[1], rho);
      vh[1] = qdiv(uh[2], rho);
      vh[2] = qdiv(uh[3], rho);
      if (VISC) {
        const int l = hlj * sx + hli;
#pragma unroll
        for (int a = 0; a < 3; ++a) vs[a * splane + l] = vh[a];
      }
      if (WP && hghost) C[oh] = closure_p(p, rho, vh[0], vh[1], vh[2], uh[4]);
      if (PF) load_u5(U, lay.at(hi, hj, k + 1), fs, k + 1 < ke, uh);
    }
    const bool pnext = k + 1 < ke || k + 1 == n3 + 1;
    if (own_load(k)) {
      const long long on = lay.at(i, j, k + 1);
      if (!PF) load_u5(U, on, fs, true, un);
      const R rho = un[0];
      vn[0] = qdiv(un[1], rho);
      vn[1] = qdiv(un[2], rho);
      vn[2] = qdiv(un[3], rho);
      if (WP && pnext) C[on] = closure_p(p, rho, vn[0], vn[1], vn[2], un[4]);
      if (PF) load_u5(U, lay.at(i, j, k + 2), fs, k + 1 < ke && own_load(k + 1), un);
    }
    if constexpr (VISC) {
      __syncthreads();
      if (own) {
        const int l = (ty + 1) * sx + (tx + 1);
        R dv[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const R* va = vs + a * splane;
          dv[a][0] = va[l + 1] - va[l - 1];
          dv[a][1] = va[l + sx] - va[l - sx];
          dv[a][2] = vn[a] - vm[a];
        }
        closure_t<UNIF, R>(geo, p, i, j, k, dv, vc, C, lay.at(i, j, k), fs);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          vm[a] = vc[a];
          vc[a] = vn[a];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Reconstruction and face flux
// ---------------------------------------------------------------------------

// Reads of the stored quantities at ghosted offset o: u(o, f) of the five
// conserved fields, c(o, f) of the closure (P and the Stokes tensor), both
// field-major with a field stride of fs elements.
//
// PU: the pressure is not read from the closure but formed from the cell's U,
// as the closure forms it (navier_stokes.py:94-100; p_from_u): four quotients
// over rho and a few products instead of a read (8 B per cell in float64).
// Bit-identical to the closure's P, so the float64 passes may take either.
// The float32 stage passes and the float64 split passes take it; the closure
// before them then leaves P out.
template <class R, class S>
MFVM_DEV R p_from_u(const S& src, long long o) {
  const R rho = src.u(o, 0);
  const R n4[4] = {src.u(o, 1), src.u(o, 2), src.u(o, 3), src.u(o, 4)};
  R q4[4];
  qdiv_by<4, R>(n4, rho, q4);
  return closure_p_e(src.par, rho, q4[0], q4[1], q4[2], q4[3]);
}
template <class R, bool PU = sizeof(R) == sizeof(float)>
struct PtrSrc {
  const R* __restrict__ U;
  const R* __restrict__ C;
  long long fs;
  NSPar par;
  MFVM_DEV R u(long long o, int f) const { return U[o + f * fs]; }
  MFVM_DEV R c(long long o, int f) const {
    if (PU && f == C_P) return p_from_u<R>(*this, o);
    return C[o + f * fs];
  }
};

// The same reads through buffer resources: one 32-bit byte offset per cell
// and the field offset as the scalar offset operand, so that the fields of a
// cell need no per-field address arithmetic and no 64-bit base per field
// (fewer scalar registers, fewer of them spilled).  Valid while every array
// is below 4 GiB (buf_src_ok).
template <class R, bool PU = sizeof(R) == sizeof(float)>
struct BufSrc {
  __amdgpu_buffer_rsrc_t U, C;
  unsigned fsb;  // field stride in bytes
  NSPar par;
  MFVM_DEV R ld(__amdgpu_buffer_rsrc_t r, long long o, int f) const {
    const int vo = (int)((unsigned)o * (unsigned)sizeof(R));
    const int so = (int)((unsigned)f * fsb);
    if constexpr (sizeof(R) == 8)
      return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    else
      return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
  }
  MFVM_DEV R u(long long o, int f) const { return ld(U, o, f); }
  MFVM_DEV R c(long long o, int f) const {
    if (PU && f == C_P) return p_from_u<R>(*this, o);
    return ld(C, o, f);
  }
};

// Stage kernels specialised by SPEC (host-checked): 1 -- uniform power-of-
// two geometry (set_pow2); 2 -- that and a viscous flow (mu != 0), so the
// viscous terms are compiled in without a flag test.  0: both read at run
// time.  Called on the kernel's own copy of the arguments.
MFVM_DEV void spec_args(StageArgs& a, int spec) {
  if (spec >= 1) set_pow2(a.geo);
  if (spec == 2) a.p.visc = 1;
}

template <class R, bool PU = sizeof(R) == sizeof(float)>
MFVM_DEV PtrSrc<R, PU> ptr_src(const StageArgs& a) {
  return PtrSrc<R, PU>{static_cast<const R*>(a.U), static_cast<const R*>(a.C), a.lay.fs, a.p};
}

// Buffer resource over [p, p + 4 GiB): raw (stride 0), the gfx9 data-format
// word; the range is not the array's, the callers keep offsets inside it.
MFVM_DEV __amdgpu_buffer_rsrc_t buf_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, -1, 0x00020000);
}

template <class R, bool PU = sizeof(R) == sizeof(float)>
MFVM_DEV BufSrc<R, PU> buf_src(const StageArgs& a) {
  return BufSrc<R, PU>{buf_rsrc(a.U), buf_rsrc(a.C), (unsigned)(a.lay.fs * (long long)sizeof(R)), a.p};
}

template <int D, class R, class S>
MFVM_DEV void load_q(const S& src, long long o, bool visc, R q[NQ]) {
#pragma unroll
  for (int f = 0; f < 5; ++f) q[f] = src.u(o, f);
  q[5] = src.c(o, 0);
  if (visc) {
#pragma unroll
    for (int e = 0; e < 3; ++e) q[6 + e] = src.c(o, trow(D, e));
  } else {
#pragma unroll
    for (int e = 0; e < 3; ++e) q[6 + e] = R(0);
  }
}

// Cell length along D at ghosted cell (i,j,k) (interior only).
template <int D, bool UNIF, class R>
MFVM_DEV R cell_len(const Geo& g, int i, int j, int k) {
  if (g.lpow2[D]) return R(g.lval[D]);
  return g_len<UNIF, R>(g, D, i, j, k);
}

// reconstructors.py:115-148 for one cell: limited slope from (qm, q, qp),
// face states q + s/2 (left side of the cell's upper face) and q - s/2.
template <int D, int LIM, bool UNIF, class R>
MFVM_DEV void slope_states(const Geo& g, int i, int j, int k, int nf, const R qm[NQ],
                           const R q[NQ], const R qp[NQ], R lo[NQ], R hi[NQ]) {
  const R len = cell_len<D, UNIF, R>(g, i, j, k);
  // float32 over a power-of-two length: every limiter is positively
  // homogeneous of degree one and scaling by a power of two is exact, so
  // limiter(db / L, df / L) * L is formed as limiter(db, df) (equal but for
  // the underflow range of the limiter's products).  float64 keeps the
  // reference's operations (slope_limiters.py, reconstructors.py:120-148).
  const bool raw = sizeof(R) == sizeof(float) && g.lpow2[D];
#pragma unroll
  for (int f = 0; f < NQ; ++f) {
    if (f < nf) {
      const R db = q[f] - qm[f], df = qp[f] - q[f];
      const R b = raw ? db : div_len<R>(g, D, db, len);
      const R fw = raw ? df : div_len<R>(g, D, df, len);
      const R s = raw ? limit<LIM, R>(b, fw) : limit<LIM, R>(b, fw) * len;
      hi[f] = q[f] + s / R(2);
      lo[f] = q[f] - s / R(2);
    } else {
      hi[f] = q[f];
      lo[f] = q[f];
    }
  }
}

// States of ghosted cell (i,j,k) along D, loading the stencil from memory.
// hi: state on the cell's upper face (reference fields_L), lo: on its lower
// face (fields_R).  Ghost cells carry zero slope (reconstructors.py:127-132).
template <int D, int LIM, bool UNIF, class R, class S>
MFVM_DEV void cell_states(const StageArgs& a, const S& src, int i, int j, int k, R lo[NQ], R hi[NQ]) {
  const long long st = D == 0 ? 1 : (D == 1 ? a.lay.sy : a.lay.sz);
  const long long o = a.lay.at(i, j, k);
  R q[NQ];
  load_q<D, R>(src, o, a.p.visc, q);
  const int cd = D == 0 ? i : (D == 1 ? j : k);
  const int nd = D == 0 ? a.lay.n1 : (D == 1 ? a.lay.n2 : a.lay.n3);
  if (LIM == MFVM_LIM_NONE || cd == 0 || cd == nd + 1) {
#pragma unroll
    for (int f = 0; f < NQ; ++f) lo[f] = hi[f] = q[f];
    return;
  }
  R qm[NQ], qp[NQ];
  load_q<D, R>(src, o - st, a.p.visc, qm);
  load_q<D, R>(src, o + st, a.p.visc, qp);
  slope_states<D, LIM, UNIF, R>(a.geo, i, j, k, a.p.visc ? 9 : 6, qm, q, qp, lo, hi);
}

// Normal flux F.S through a D-face from left state l and right state r:
// convective flux (navier_stokes.py:195-212), Riemann solver
// (riemann_solvers.py), centred viscous flux (navier_stokes.py:259-275,
// diffusive_solvers.py:28), total F = Fc + Fd (:680-682), times the area.
template <int D, int RS, class R>
MFVM_DEV void face_flux(const NSPar& p, const R l[NQ], const R r[NQ], R area, R fsout[5]) {
  const R rl = l[0], rr = r[0];
  const R Pl = l[5], Pr = r[5];
  // The five quotients over each side's density: velocity, specific energy
  // and gamma P / rho for the sound speed (:309).
  R ql[5], qr[5];
  {
    const R nl[5] = {l[1], l[2], l[3], l[4], R(p.gamma) * Pl};
    const R nr[5] = {r[1], r[2], r[3], r[4], R(p.gamma) * Pr};
    qdiv_by<5, R>(nl, rl, ql);
    qdiv_by<5, R>(nr, rr, qr);
  }
  const R vl[3] = {ql[0], ql[1], ql[2]};
  const R vr[3] = {qr[0], qr[1], qr[2]};
  const R El = ql[3], Er = qr[3];
  R fl[5], fr[5];
  fl[0] = rl * vl[D];
  fr[0] = rr * vr[D];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    fl[1 + e] = (rl * vl[e]) * vl[D];
    fr[1 + e] = (rr * vr[e]) * vr[D];
    if (e == D) {
      fl[1 + e] = fl[1 + e] + Pl;
      fr[1 + e] = fr[1 + e] + Pr;
    }
  }
  fl[4] = ((rl * El) + Pl) * vl[D];
  fr[4] = ((rr * Er) + Pr) * vr[D];
  // :309  cs = sqrt(ga*P/rho)
  const R csl = qsqrt(ql[4]);
  const R csr = qsqrt(qr[4]);
  R F[5];
  if constexpr (RS == MFVM_RS_HLL) {
    // Wave set {vn, vn-cs, vn+cs}: its min is vn-cs and its max vn+cs.
    // (float32: the hardware min / max, equal but for NaN operands.)
    R smin, smax;
    if constexpr (sizeof(R) == sizeof(float)) {
      smin = fminf(vl[D] - csl, vr[D] - csr);
      smax = fmaxf(vl[D] + csl, vr[D] + csr);
    } else {
      smin = nmin(vl[D] - csl, vr[D] - csr);
      smax = nmax(vl[D] + csl, vr[D] + csr);
    }
    if (smin >= R(0)) {
#pragma unroll
      for (int q = 0; q < 5; ++q) F[q] = fl[q];
    } else if (smax <= R(0)) {
#pragma unroll
      for (int q = 0; q < 5; ++q) F[q] = fr[q];
    } else {
      const R den = smax - smin;
      const R sm = smin * smax;
      R nm[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) nm[q] = (smax * fl[q] - smin * fr[q]) + sm * (r[q] - l[q]);
      qdiv_by<5, R>(nm, den, F);
    }
  } else {
    // max |{vn, vn-cs, vn+cs}| is |vn| + cs, rounded once.
    const R s = nmax(fabs(vl[D]) + csl, fabs(vr[D]) + csr);
#pragma unroll
    for (int q = 0; q < 5; ++q) F[q] = ((fl[q] + fr[q]) / R(2)) - ((s * (r[q] - l[q])) / R(2));
  }
  if (p.visc) {
    // row D of T: l[6..8] = T_D1, T_D2, T_D3
    const R el = ((-l[6]) * vl[0] - l[7] * vl[1]) - l[8] * vl[2];
    const R er = ((-r[6]) * vr[0] - r[7] * vr[1]) - r[8] * vr[2];
#pragma unroll
    for (int e = 0; e < 3; ++e) F[1 + e] = F[1 + e] + ((-l[6 + e]) + (-r[6 + e])) / R(2);
    F[4] = F[4] + (el + er) / R(2);
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) fsout[q] = F[q] * area;
}

// Source terms at interior cell (i,j,k) (navier_stokes.py:358-509), given
// the 12 stored quantities of the cell: U(5), P, T11,T22,T33,T12,T13,T23.
template <bool UNIF, class R>
MFVM_DEV void cell_source(const Geo& g, const NSPar& p, int i, int j, int k, const R* u,
                          const R* c, R src[5], bool& any) {
  any = false;
#pragma unroll
  for (int q = 0; q < 5; ++q) src[q] = R(0);
  if ((UNIF || !g.curv) && !p.grav) return;
  any = true;
  const R rho = u[0];
  const R n3[3] = {u[1], u[2], u[3]};
  R q3[3];
  ieee_div_by<3, R>(n3, rho, q3);
  const R v1 = q3[0], v2 = q3[1], v3 = q3[2];
  R sg1 = R(0), sg2 = R(0), sg3 = R(0);
  if (!UNIF && g.curv) {
    const int a = i - 1, b = j - 1, e = k - 1;
    const R h1 = met<R>(g, MFVM_MET_H1, a, b, e), h2 = met<R>(g, MFVM_MET_H2, a, b, e),
            h3 = met<R>(g, MFVM_MET_H3, a, b, e);
    const R d1h = met<R>(g, MFVM_MET_D1H, a, b, e), d1h2 = met<R>(g, MFVM_MET_D1H2, a, b, e),
            d1h3 = met<R>(g, MFVM_MET_D1H3, a, b, e);
    const R d2h = met<R>(g, MFVM_MET_D2H, a, b, e), d2h1 = met<R>(g, MFVM_MET_D2H1, a, b, e),
            d2h3 = met<R>(g, MFVM_MET_D2H3, a, b, e);
    const R d3h = met<R>(g, MFVM_MET_D3H, a, b, e), d3h1 = met<R>(g, MFVM_MET_D3H1, a, b, e),
            d3h2 = met<R>(g, MFVM_MET_D3H2, a, b, e);
    const R P = c[C_P];
    const R T11 = c[C_T11], T22 = c[C_T22], T33 = c[C_T33];
    const R T12 = c[C_T12], T13 = c[C_T13], T23 = c[C_T23];
    // :439-441
    const R c1 = ((((P * d1h) / h1 + (((rho * v2) * v2) * d1h2) / h1) + (((rho * v3) * v3) * d1h3) / h1) -
                  (((rho * v1) * v2) * d2h1) / h2) -
                 (((rho * v1) * v3) * d3h1) / h3;
    const R c2 = ((((P * d2h) / h2 + (((rho * v1) * v1) * d2h1) / h2) + (((rho * v3) * v3) * d2h3) / h2) -
                  (((rho * v2) * v1) * d1h2) / h1) -
                 (((rho * v2) * v3) * d3h2) / h3;
    const R c3 = ((((P * d3h) / h3 + (((rho * v1) * v1) * d3h1) / h3) + (((rho * v2) * v2) * d3h2) / h3) -
                  (((rho * v3) * v1) * d1h3) / h1) -
                 (((rho * v3) * v2) * d2h3) / h2;
    // :446-448
    const R e1 = (((T12 * d2h1) / h2 + (T13 * d3h1) / h3) - (T22 * d1h2) / h1) - (T33 * d1h3) / h1;
    const R e2 = (((T12 * d1h2) / h1 + (T23 * d3h2) / h3) - (T11 * d2h1) / h2) - (T33 * d2h3) / h2;
    const R e3 = (((T13 * d1h3) / h1 + (T23 * d2h3) / h2) - (T11 * d3h1) / h3) - (T22 * d3h2) / h3;
    sg1 = c1 + e1;
    sg2 = c2 + e2;
    sg3 = c3 + e3;
  }
  // :500-504 gravity; src = sg + sp (:691)
  const R g1 = R(p.g[0]), g2 = R(p.g[1]), g3 = R(p.g[2]);
  src[0] = R(0);
  src[1] = sg1 + rho * g1;
  src[2] = sg2 + rho * g2;
  src[3] = sg3 + rho * g3;
  src[4] = R(0) + rho * (((g1 * v1) + (g2 * v2)) + (g3 * v3));
}

// CFL rates of one cell (navier_stokes.py:549-575): max_d (|v_d|+cs)/L_d and mu/rho.
template <bool UNIF, class R>
MFVM_DEV void cell_cfl(const Geo& g, const NSPar& p, int i, int j, int k, const R u[5], R& rate,
                       R& nu) {
  const R rho = u[0];
  // v, E and mu / rho: five quotients over rho (IEEE, shared reciprocal)
  const R n5[5] = {u[1], u[2], u[3], u[4], R(p.mu)};
  R q5[5];
  ieee_div_by<5, R>(n5, rho, q5);
  const R v1 = q5[0], v2 = q5[1], v3 = q5[2];
  const R E = q5[3];
  const R P = (R(p.gm1) * rho) * (E - (((v1 * v1) + (v2 * v2)) + (v3 * v3)) / R(2));
  const R cs = sqrt((R(p.gamma) * P) / rho);
  const R a1 = fabs(v1) + cs, a2 = fabs(v2) + cs, a3 = fabs(v3) + cs;
  const R r1 = div_len<R>(g, 0, a1, cell_len<0, UNIF, R>(g, i, j, k));
  const R r2 = div_len<R>(g, 1, a2, cell_len<1, UNIF, R>(g, i, j, k));
  const R r3 = div_len<R>(g, 2, a3, cell_len<2, UNIF, R>(g, i, j, k));
  rate = nmax(nmax(r1, r2), r3);
  nu = q5[4];
}

// Block max of (rate, nu) -> partials[2*block], NaN-propagating.
template <class R>
MFVM_DEV void block_max2(R rate, R nu, double* slot, double* sred) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    rate = nmax(rate, __shfl_xor(rate, o, 64));
    nu = nmax(nu, __shfl_xor(nu, o, 64));
  }
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();
  if ((tid & 63) == 0) {
    sred[2 * (tid >> 6)] = double(rate);
    sred[2 * (tid >> 6) + 1] = double(nu);
  }
  __syncthreads();
  if (tid == 0) {
    double a = sred[0], b = sred[1];
    for (int w = 1; w < nw; ++w) {
      a = nmax(a, sred[2 * w]);
      b = nmax(b, sred[2 * w + 1]);
    }
    slot[0] = a;
    slot[1] = b;
  }
}

// Linear id of this workgroup in its launch grid.
MFVM_DEV long long hw_block_id() {
  return (long long)(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
}

// ---------------------------------------------------------------------------
// Fused stage kernel
// ---------------------------------------------------------------------------

// The tile one workgroup advances: tile coordinates (x1, x2 tile and x3
// chunk), tile shape (bx * by threads take part, bx fastest), and where the
// tile's CFL maxima go (NULL: no CFL reduction).
struct TileCtx {
  int x, y, z;
  int bx, by;
  double* slot;
};

// Tile of a one-workgroup-per-tile launch over one grid.
MFVM_DEV TileCtx grid_tile(const StageArgs& a) {
  const TileId t = tile_id(a.remap != 0);
  TileCtx c;
  c.x = t.x;
  c.y = t.y;
  c.z = t.z;
  c.bx = blockDim.x;
  c.by = blockDim.y;
  c.slot = a.partials ? a.partials + 2 * hw_block_id() : nullptr;
  return c;
}
// x1 and x2 faces of plane m of a tile: states of the own cells and of the
// tile-edge cells, F.S of every face once, and the in-plane part of the
// divergence accm = ((F1(i+1/2) - F1(i-1/2)) + F2(j+1/2)) - F2(j-1/2)
// (navier_stokes.py:705-709).  Every thread of the workgroup must call it (it
// holds two barriers); cu: the 12 stored quantities of the own cell.
template <int LIM, int RS, bool UNIF, class R, class S>
MFVM_DEV void inplane_div(const StageArgs& a, const S& src, R* sm, int BX, int BY, int tx, int ty, int tid,
                          int i0, int j0, int i, int j, bool own, int m, const R cu[12], R accm[5]) {
  const int n1 = a.lay.n1, n2 = a.lay.n2;
  const bool visc = a.p.visc != 0;
  const int wx = BX + 1;
  const int nrx = BY * wx, nry = (BY + 1) * BX;
  // LDS: RX[9][BY][BX+1] lower-face states of the x1 faces, LX[9][BY]
  // upper-face states of the tile-edge cells, RY/LY likewise for x2, then the
  // face fluxes FX[5][BY][BX+1] / FY[5][BY+1][BX] in space of their own.  A
  // plane takes two barriers (states written, fluxes written): the next
  // plane writes its states once every thread is past the second one, after
  // its last state read, and its fluxes once every thread is past the first
  // one, after its last flux read.
  R* RX = sm;
  R* LX = RX + NQ * nrx;
  R* RY = LX + NQ * BY;
  R* LY = RY + NQ * nry;
  R* FX = LY + NQ * BX;
  R* FY = FX + 5 * nrx;
  const int NT = BX * BY;
  R hx[NQ], hy[NQ];
  if (j <= n2 && i <= n1 + 1) {
    R lo[NQ];
    if (i <= n1) {
      // own cell: centre from the z-window, neighbours from memory
      const long long o = a.lay.at(i, j, m);
      R q[NQ];
#pragma unroll
      for (int f = 0; f < 6; ++f) q[f] = cu[f];
      q[6] = cu[5 + C_T11];
      q[7] = cu[5 + C_T12];
      q[8] = cu[5 + C_T13];
      if (LIM == MFVM_LIM_NONE) {
#pragma unroll
        for (int f = 0; f < NQ; ++f) lo[f] = hx[f] = q[f];
      } else {
        R qm[NQ], qp[NQ];
        load_q<0, R>(src, o - 1, visc, qm);
        load_q<0, R>(src, o + 1, visc, qp);
        slope_states<0, LIM, UNIF, R>(a.geo, i, j, m, visc ? 9 : 6, qm, q, qp, lo, hx);
      }
    } else {
      cell_states<0, LIM, UNIF, R>(a, src, i, j, m, lo, hx);  // x1 ghost: raw
    }
#pragma unroll
    for (int f = 0; f < NQ; ++f) RX[f * nrx + ty * wx + tx] = lo[f];
  }
  if (i <= n1 && j <= n2 + 1) {
    R lo[NQ];
    if (j <= n2) {
      const long long o = a.lay.at(i, j, m);
      R q[NQ];
#pragma unroll
      for (int f = 0; f < 6; ++f) q[f] = cu[f];
      q[6] = cu[5 + C_T12];
      q[7] = cu[5 + C_T22];
      q[8] = cu[5 + C_T23];
      if (LIM == MFVM_LIM_NONE) {
#pragma unroll
        for (int f = 0; f < NQ; ++f) lo[f] = hy[f] = q[f];
      } else {
        R qm[NQ], qp[NQ];
        load_q<1, R>(src, o - a.lay.sy, visc, qm);
        load_q<1, R>(src, o + a.lay.sy, visc, qp);
        slope_states<1, LIM, UNIF, R>(a.geo, i, j, m, visc ? 9 : 6, qm, q, qp, lo, hy);
      }
    } else {
      cell_states<1, LIM, UNIF, R>(a, src, i, j, m, lo, hy);  // x2 ghost: raw
    }
#pragma unroll
    for (int f = 0; f < NQ; ++f) RY[f * nry + ty * BX + tx] = lo[f];
  }
  // helpers, phase 1: [0, 2BX) x2 tile-edge cells, [2BX, 2BX+2BY) x1 tile-edge cells
  for (int t = tid; t < 2 * (BX + BY); t += NT) {
    R lo[NQ], hi[NQ];
    if (t < 2 * BX) {
      const int c = t % BX, side = t / BX, ii = i0 + c;
      if (ii <= n1) {
        if (side == 0) {
          cell_states<1, LIM, UNIF, R>(a, src, ii, j0 - 1, m, lo, hi);
#pragma unroll
          for (int f = 0; f < NQ; ++f) LY[f * BX + c] = hi[f];
        } else if (j0 + BY <= n2 + 1) {
          cell_states<1, LIM, UNIF, R>(a, src, ii, j0 + BY, m, lo, hi);
#pragma unroll
          for (int f = 0; f < NQ; ++f) RY[f * nry + BY * BX + c] = lo[f];
        }
      }
    } else {
      const int u = t - 2 * BX, r = u % BY, side = u / BY, jj = j0 + r;
      if (jj <= n2) {
        if (side == 0) {
          cell_states<0, LIM, UNIF, R>(a, src, i0 - 1, jj, m, lo, hi);
#pragma unroll
          for (int f = 0; f < NQ; ++f) LX[f * BY + r] = hi[f];
        } else if (i0 + BX <= n1 + 1) {
          cell_states<0, LIM, UNIF, R>(a, src, i0 + BX, jj, m, lo, hi);
#pragma unroll
          for (int f = 0; f < NQ; ++f) RX[f * nrx + r * wx + BX] = lo[f];
        }
      }
    }
  }
  __syncthreads();
  // phase 2: fluxes into registers (own x1 and x2 faces; one tile-edge face
  // for each of the last BX + BY threads), then stored over the states
  R fx[5], fy[5], fh[5];
  if (own) {
    R rr[NQ];
#pragma unroll
    for (int f = 0; f < NQ; ++f) rr[f] = RX[f * nrx + ty * wx + tx + 1];
    face_flux<0, RS, R>(a.p, hx, rr, g_area<UNIF, R>(a.geo, 0, i, j - 1, m - 1), fx);
#pragma unroll
    for (int f = 0; f < NQ; ++f) rr[f] = RY[f * nry + (ty + 1) * BX + tx];
    face_flux<1, RS, R>(a.p, hy, rr, g_area<UNIF, R>(a.geo, 1, i - 1, j, m - 1), fy);
  }
  // helper faces: u in [0, BX) faces j0-1/2, [BX, BX+BY) faces i0-1/2
  const int u = NT - 1 - tid;
  int hkind = 0, hidx = 0;
  if (u < BX + BY) {
    R ll[NQ], rr[NQ];
    if (u < BX) {
      const int c = u, ii = i0 + c;
      if (ii <= n1) {
#pragma unroll
        for (int f = 0; f < NQ; ++f) {
          ll[f] = LY[f * BX + c];
          rr[f] = RY[f * nry + c];
        }
        face_flux<1, RS, R>(a.p, ll, rr, g_area<UNIF, R>(a.geo, 1, ii - 1, j0 - 1, m - 1), fh);
        hkind = 1;
        hidx = c;
      }
    } else {
      const int r = u - BX, jj = j0 + r;
      if (jj <= n2) {
#pragma unroll
        for (int f = 0; f < NQ; ++f) {
          ll[f] = LX[f * BY + r];
          rr[f] = RX[f * nrx + r * wx];
        }
        face_flux<0, RS, R>(a.p, ll, rr, g_area<UNIF, R>(a.geo, 0, i0 - 1, jj - 1, m - 1), fh);
        hkind = 2;
        hidx = r * wx;
      }
    }
  }
  if (own) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      FX[q * nrx + ty * wx + tx + 1] = fx[q];
      FY[q * nry + (ty + 1) * BX + tx] = fy[q];
    }
  }
  if (hkind == 1) {
#pragma unroll
    for (int q = 0; q < 5; ++q) FY[q * nry + hidx] = fh[q];
  } else if (hkind == 2) {
#pragma unroll
    for (int q = 0; q < 5; ++q) FX[q * nrx + hidx] = fh[q];
  }
  __syncthreads();
  if (own) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const R ax = FX[q * nrx + ty * wx + tx + 1] - FX[q * nrx + ty * wx + tx];
      accm[q] = (ax + FY[q * nry + (ty + 1) * BX + tx]) - FY[q * nry + ty * BX + tx];
    }
  }
}

// ZONLY: the in-plane part of the divergence comes from a.xy (written by the
// x1/x2 kernel of a split stage) instead of being computed here.  src reads
// a.U and a.C (PtrSrc or BufSrc).
// FIN (compensated stages): 2 -- whether this is the final stage is read
// from a.comp; 0 / 1 -- known at compile time (a.comp must agree), which
// leaves the float64 accumulation and CFL code out of the other stages.
template <int LIM, int RS, bool UNIF, class R, bool COMP, bool ZONLY, class S, int FIN = 2, bool FC = false>
MFVM_DEV void ns_stage_body(const StageArgs& a, const TileCtx& tl, const S& src) {
  extern __shared__ double smem_raw[];
  R* sm = reinterpret_cast<R*>(smem_raw);
  const int BX = tl.bx, BY = tl.by;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int tx = tid % BX, ty = tid / BX;
  const int n1 = a.lay.n1, n2 = a.lay.n2, n3 = a.lay.n3;
  const long long fs = a.lay.fs;
  const int i0 = 1 + tl.x * BX, j0 = 1 + tl.y * BY;
  const int i = i0 + tx, j = j0 + ty;
  const int kb = 1 + tl.z * a.kz;
  const int ke = min(kb + a.kz, n3 + 1);
  const bool own = (i <= n1) && (j <= n2);
  // Every lane marches: a lane past the grid's edge reads and computes the
  // nearest interior column instead (its results are never stored), so that
  // no state carried from plane to plane depends on a branch.
  const int ic = min(i, n1), jc = min(j, n2);
  const bool visc = a.p.visc != 0;
  // LATE (x3-only pass on a uniform Cartesian grid): the in-plane divergence
  // and the source of cell m-1 are read and formed when it is finalized
  // (the source needs only U there), instead of being carried a plane.
  constexpr bool LATE = ZONLY && UNIF;
  constexpr bool EARLY = ZONLY && !COMP && LIM != MFVM_LIM_NONE;

  double dt = a.dt_dev ? *a.dt_dev : a.dt;

  // z-window of the own column: prev (ZS of m-1), cur (all of m), next (all of m+1)
  R zp[NQ], cu[12], nx[12];
  R lz_prev[NQ];            // upper-face state of cell m-1 (left state of face m-1/2)
  R fsz_prev[5];            // F.S of face m-3/2
  R acc_prev[5], src_prev[5];
  bool src_any_prev = false;
  double rate_max = 0.0, nu_max = 0.0;

  // The x3-only pass on a Cartesian grid reads T11, T22, T12 nowhere (the x3
  // states take row 3 of T, the gravity source only U): it skips them.
  const bool cart = UNIF || !a.geo.curv;
  auto load12 = [&](int k, R* w) {
    const long long o = a.lay.at(ic, jc, k);
#pragma unroll
    for (int f = 0; f < 5; ++f) w[f] = src.u(o, f);
#pragma unroll
    for (int f = 0; f < NS_NC; ++f) {
      const bool row12 = f == C_T11 || f == C_T22 || f == C_T12;
      w[5 + f] = (f == 0 || (visc && !(ZONLY && cart && row12))) ? src.c(o, f) : R(0);
    }
  };
  // ZS of a 12-vector: U, P, T13, T23, T33
  auto zs = [&](const R* w, R* z) {
#pragma unroll
    for (int f = 0; f < 6; ++f) z[f] = w[f];
    z[6] = w[5 + C_T13];
    z[7] = w[5 + C_T23];
    z[8] = w[5 + C_T33];
  };

  {
    R w0[12];
    load12(kb - 1, w0);
    zs(w0, zp);
    load12(kb, cu);
    if (kb - 1 == 0 || LIM == MFVM_LIM_NONE) {
#pragma unroll
      for (int f = 0; f < NQ; ++f) lz_prev[f] = zp[f];
    } else {
      R wm[12], zm[NQ], zc[NQ], lo[NQ];
      load12(kb - 2, wm);
      zs(wm, zm);
      zs(cu, zc);
      slope_states<2, LIM, UNIF, R>(a.geo, ic, jc, kb - 1, visc ? 9 : 6, zm, zp, zc, lo, lz_prev);
    }
    // The x3-only pass has no barriers between planes: it loads each plane
    // one iteration ahead, so that the loads overlap the previous plane's work.
    if (ZONLY && kb + 1 <= n3 + 1) load12(kb + 1, nx);
  }

  // Two forms of the march over planes, equal in results.  FC (the
  // compensated x3-only pass): two parts per plane, closing the previous
  // plane before the next plane's loads; it also leaves the final stage few
  // enough registers for four waves per SIMD (with both, 1.57 -> 1.42 ms of
  // x3 passes per RT 256^3 step).  Otherwise one loop body per plane: the
  // face/close form instantiated in a file also changed the compiler's
  // allocation of that file's other kernels (the ensemble's x1/x2 pass,
  // 102 -> 150 VGPRs).
  if constexpr (FC) {
    // The march over planes, in two parts per plane m: face(m), the x3 face
    // m-1/2 and the in-plane work of plane m (it issues the loads of plane m+2
    // and, EARLY, the reads of cell m-1); close(m), which finalizes cell m-1
    // (its stores) and rotates the window.
    R fsz[5];
    R lz_cur[NQ];
    R nn[12];  // ZONLY: plane m+2, loaded in face(m)
    R accm[5], srcm[5];
    bool src_any = false;
    // EARLY: the reads of cell m-1, which close(m) finalizes -- U0, kacc and,
    // in the LATE form, the in-plane divergence -- are issued at the start of
    // face(m), before the x3 face's work, so that their latency overlaps it (the
    // float64 x3-only pass stays at 2 waves per SIMD).  The other forms issue
    // them at the finalize (early reads would cost the compensated x3 pass a
    // wave per SIMD).
    R fxy[5], f0[5], f1[5];
    auto xy_reads = [&](int k) {
      const long long oi = a.il.at(ic, jc, k);
  #pragma unroll
      for (int q = 0; q < 5; ++q) fxy[q] = static_cast<const R*>(a.xy)[oi + q * a.il.fs];
    };
    // (U0 == U, as in the first stage of a step: U0 of cell m-1 is the
    // window's zp, read already)
    const bool u0_own = a.U0 == a.U;
    auto base_reads = [&](int k) {
      const long long oc = a.lay.at(ic, jc, k), oi = a.il.at(ic, jc, k);
      const bool need0 = !u0_own && (a.update == MFVM_UPD_BASE || a.update == MFVM_UPD_SSP || a.update == MFVM_UPD_ACC6);
      const bool needk = a.update == MFVM_UPD_ACC6 || a.acc == MFVM_ACC_ADD;
  #pragma unroll
      for (int q = 0; q < 5; ++q) {
        f0[q] = need0 ? static_cast<const R*>(a.U0)[oc + q * fs] : R(0);
        f1[q] = needk ? static_cast<const R*>(a.kacc)[oi + q * a.il.fs] : R(0);
      }
    };
    auto face = [&](int m) {
      if constexpr (EARLY) {
        if (m > kb) {
          if constexpr (LATE) xy_reads(m - 1);
          base_reads(m - 1);
        }
      }
      {
        R zc[NQ], lo[NQ];
        zs(cu, zc);
        if constexpr (ZONLY) {
          if (m + 2 <= n3 + 1 && m + 1 <= ke) load12(m + 2, nn);
        } else {
          if (m + 1 <= n3 + 1) load12(m + 1, nx);
        }
        if (m == n3 + 1 || LIM == MFVM_LIM_NONE) {
  #pragma unroll
          for (int f = 0; f < NQ; ++f) lo[f] = lz_cur[f] = zc[f];
        } else {
          R zn[NQ];
          zs(nx, zn);
          slope_states<2, LIM, UNIF, R>(a.geo, ic, jc, m, visc ? 9 : 6, zp, zc, zn, lo, lz_cur);
        }
        const R s3 = g_area<UNIF, R>(a.geo, 2, ic - 1, jc - 1, m - 1);
        face_flux<2, RS, R>(a.p, lz_prev, lo, s3, fsz);
      }

      // ---- in-plane faces of plane m (all threads take part in the barriers) ----
      // Phase 1: x1 and x2 states of the own cells and of the tile-edge cells;
      // phase 2: F.S of every x1 and x2 face of the tile, each computed once;
      // phase 3: flux differences.  Helper tasks go to the first threads in
      // phase 1 and to the last threads in phase 2, so that no wave takes two.
      src_any = false;
      if (!LATE && m < ke) {
        if constexpr (ZONLY) {
          const long long oi = a.il.at(ic, jc, m);
  #pragma unroll
          for (int q = 0; q < 5; ++q) accm[q] = static_cast<const R*>(a.xy)[oi + q * a.il.fs];
        } else {
          inplane_div<LIM, RS, UNIF, R>(a, src, sm, BX, BY, tx, ty, tid, i0, j0, i, j, own, m, cu, accm);
        }
        cell_source<UNIF, R>(a.geo, a.p, ic, jc, m, cu, cu + 5, srcm, src_any);
      }

    };
    auto close = [&](int m) {
      // ---- finalize cell m-1 (stores from the lanes of interior columns) --------
      if (m > kb) {
        const int k = m - 1;
        R kr[5];
        R vol;
        const bool vp2 = a.geo.vpow2 != 0;
        if (!vp2) vol = g_vol<UNIF, R>(a.geo, ic, jc, k);
        if constexpr (LATE) {
          if constexpr (EARLY) {
  #pragma unroll
            for (int q = 0; q < 5; ++q) acc_prev[q] = fxy[q];
          } else {
            const long long oi = a.il.at(ic, jc, k);
  #pragma unroll
            for (int q = 0; q < 5; ++q) acc_prev[q] = static_cast<const R*>(a.xy)[oi + q * a.il.fs];
          }
          // (uniform Cartesian: the source reads U only, never the closure)
          cell_source<UNIF, R>(a.geo, a.p, ic, jc, k, zp, zp + 5, src_prev, src_any_prev);
        }
  #pragma unroll
        for (int q = 0; q < 5; ++q) {
          const R x = -((acc_prev[q] + fsz[q]) - fsz_prev[q]);
          kr[q] = vp2 ? x * R(a.geo.vinv) : qdiv(x, vol);
          if (src_any_prev) kr[q] = kr[q] + src_prev[q];
        }
        const long long oc = a.lay.at(ic, jc, k);
        const long long oi = a.il.at(ic, jc, k);
        R* out = static_cast<R*>(a.out);
        R* ka = static_cast<R*>(a.kacc);
        // The cell's reads of U0, kacc, lo and dinc are all issued before its
        // first store (the arrays do not alias the outputs), so that they take
        // one memory round trip instead of one per field.
        if constexpr (!COMP) {
          R nu_[5], u0v[5], kav[5];
          const R dtr = R(dt);
          if constexpr (EARLY) {
  #pragma unroll
            for (int q = 0; q < 5; ++q) {
              u0v[q] = u0_own ? zp[q] : f0[q];
              kav[q] = f1[q];
            }
          } else {
            const R* U0 = static_cast<const R*>(a.U0);
            const bool need0 = a.update == MFVM_UPD_BASE || a.update == MFVM_UPD_SSP || a.update == MFVM_UPD_ACC6;
            const bool needk = a.update == MFVM_UPD_ACC6 || a.acc == MFVM_ACC_ADD;
  #pragma unroll
            for (int q = 0; q < 5; ++q) {
              u0v[q] = u0_own ? zp[q] : (need0 ? U0[oc + q * fs] : R(0));
              kav[q] = needk ? ka[oi + q * a.il.fs] : R(0);
            }
          }
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            const R uq = zp[q];  // stage input U at cell m-1
            R o = uq;
            switch (a.update) {
              case MFVM_UPD_EULER: o = uq + dtr * kr[q]; break;
              case MFVM_UPD_BASE: o = u0v[q] + R(a.c0 * dt) * kr[q]; break;
              case MFVM_UPD_SSP: o = (R(a.c0) * u0v[q]) + (R(a.c1) * (uq + dtr * kr[q])); break;
              case MFVM_UPD_ACC6: o = u0v[q] + R(dt / 6.0) * (kav[q] + kr[q]); break;
              default: break;
            }
            nu_[q] = o;
            if (own) {
              if (a.update != MFVM_UPD_NONE) out[oc + q * fs] = o;
              if (a.acc == MFVM_ACC_SET) ka[oi + q * a.il.fs] = kr[q];
              else if (a.acc == MFVM_ACC_ADD) ka[oi + q * a.il.fs] = kav[q] + R(a.acc_coef) * kr[q];
            }
          }
          if (a.partials) {
            R rate, nu;
            cell_cfl<UNIF, R>(a.geo, a.p, ic, jc, k, nu_, rate, nu);
            rate_max = own ? nmax(rate_max, double(rate)) : rate_max;
            nu_max = own ? nmax(nu_max, double(nu)) : nu_max;
          }
        } else {
          // Compensated update: U0 = hi + lo (two FP32 words).  Stage states
          // are the FP32 values hi + d with an FP32 increment d (lo, below half
          // an ulp of hi, changes such a value by at most its own rounding, so
          // the other stages do not read it); the last stage adds its
          // increment to hi + lo in FP64 and splits the sum back into (hi, lo).
          // The SSP middle stage forms its state as its input (hi + d1, the
          // window's zp) + (d2 - d1), without reading hi: within about an ulp
          // of hi + d2, the FP32 state's own precision (the increments, and so
          // the FP64 accumulation, are unchanged).
          R* lo = static_cast<R*>(a.lo);
          R* dv = static_cast<R*>(a.dinc);
          R* sn = static_cast<R*>(a.snext);
          const bool fin = FIN == 2 ? a.comp >= MFVM_COMP_FINAL_SET : FIN == 1;
          const bool needd = a.comp == MFVM_COMP_SSP || a.comp == MFVM_COMP_FINAL_SSP;
          const bool needk = a.comp == MFVM_COMP_FINAL_ACC6 || a.acc == MFVM_ACC_ADD;
          const R* U0 = static_cast<const R*>(a.U0);
          double nw[5];
          R hv[5], lv[5], dvv[5], kav[5];
          const bool needh = fin || a.comp != MFVM_COMP_SSP;
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            const long long og = oc + q * fs, ii = oi + q * a.il.fs;
            hv[q] = (u0_own || !needh) ? zp[q] : U0[og];
            lv[q] = fin ? lo[og] : R(0);
            dvv[q] = needd ? dv[ii] : R(0);
            kav[q] = needk ? ka[ii] : R(0);
          }
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            const R kq = kr[q];
            const long long og = oc + q * fs, ii = oi + q * a.il.fs;
            const R h = hv[q], l = lv[q];
            if (!fin) {
              R d, nxt;
              if (a.comp == MFVM_COMP_SET) {
                d = R(a.c0 * dt) * kq;
                nxt = h + d;
              } else {
                d = R(a.c0) * (dvv[q] + R(dt) * kq);
                nxt = h + (d - dvv[q]);  // h: the stage input (see above)
              }
              if (own) {
                dv[ii] = d;
                sn[og] = nxt;
              }
            } else {
              double inc;
              if (a.comp == MFVM_COMP_FINAL_SET) inc = (a.c0 * dt) * double(kq);
              else if (a.comp == MFVM_COMP_FINAL_SSP) inc = a.c0 * (double(dvv[q]) + dt * double(kq));
              else inc = (dt / 6.0) * (double(kav[q]) + double(kq));
              const double v = (double(h) + double(l)) + inc;
              const R nh = R(v);
              if (own) {
                out[og] = nh;
                lo[og] = R(v - double(nh));
              }
              nw[q] = v;
            }
            if (own) {
              if (a.acc == MFVM_ACC_SET) ka[ii] = kq;
              else if (a.acc == MFVM_ACC_ADD) ka[ii] = kav[q] + R(a.acc_coef) * kq;
            }
          }
          if (fin && a.partials) {
            double rate, nu;
            cell_cfl<UNIF, double>(a.geo, a.p, ic, jc, k, nw, rate, nu);
            rate_max = own ? nmax(rate_max, rate) : rate_max;
            nu_max = own ? nmax(nu_max, nu) : nu_max;
          }
        }
      }

      // ---- rotate the window -------------------------------------------------------
      {
  #pragma unroll
        for (int q = 0; q < 5; ++q) fsz_prev[q] = fsz[q];
  #pragma unroll
        for (int f = 0; f < NQ; ++f) lz_prev[f] = lz_cur[f];
        if (!LATE && m < ke) {
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            acc_prev[q] = accm[q];
            src_prev[q] = srcm[q];
          }
          src_any_prev = src_any;
        }
        zs(cu, zp);
  #pragma unroll
        for (int f = 0; f < 12; ++f) cu[f] = nx[f];
        if constexpr (ZONLY) {
  #pragma unroll
          for (int f = 0; f < 12; ++f) nx[f] = nn[f];
        }
      }
    };
    if constexpr (COMP) {
      // close(m-1) before face(m): no store is issued between a plane's loads
      // and the end of the trip, where the carried window is handed over, so
      // that hand-over waits for those loads only, not for the stores to
      // complete (compensated x3 pass at RT 256^3: 1.57 -> 1.53 ms per step;
      // the float64 pass measured 1% slower in this order and keeps the other).
      face(kb);
      for (int m = kb + 1; m <= ke; ++m) {
        close(m - 1);
        face(m);
      }
      close(ke);
    } else {
      for (int m = kb; m <= ke; ++m) {
        face(m);
        close(m);
      }
    }

  } else {
    for (int m = kb; m <= ke; ++m) {
      // ---- x3 face m-1/2 -------------------------------------------------------
      R fsz[5];
      R lz_cur[NQ];
      R nn[12];  // ZONLY: plane m+2, loaded now for the next iteration
      // EARLY: the reads of the cell finalized in this iteration (m-1) -- U0,
      // kacc and, in the LATE form, the in-plane divergence -- are issued here,
      // before the x3 face's work, so that their latency overlaps it (the
      // float64 x3-only pass stays at 2 waves per SIMD).  The other forms issue
      // them at the finalize (early reads would cost the compensated x3 pass a
      // wave per SIMD).
      R fxy[5], f0[5], f1[5];
      auto xy_reads = [&](int k) {
        const long long oi = a.il.at(ic, jc, k);
  #pragma unroll
        for (int q = 0; q < 5; ++q) fxy[q] = static_cast<const R*>(a.xy)[oi + q * a.il.fs];
      };
      // (U0 == U, as in the first stage of a step: U0 of cell m-1 is the
      // window's zp, read already)
      const bool u0_own = a.U0 == a.U;
      auto base_reads = [&](int k) {
        const long long oc = a.lay.at(ic, jc, k), oi = a.il.at(ic, jc, k);
        const bool need0 = !u0_own && (a.update == MFVM_UPD_BASE || a.update == MFVM_UPD_SSP || a.update == MFVM_UPD_ACC6);
        const bool needk = a.update == MFVM_UPD_ACC6 || a.acc == MFVM_ACC_ADD;
  #pragma unroll
        for (int q = 0; q < 5; ++q) {
          f0[q] = need0 ? static_cast<const R*>(a.U0)[oc + q * fs] : R(0);
          f1[q] = needk ? static_cast<const R*>(a.kacc)[oi + q * a.il.fs] : R(0);
        }
      };
      if constexpr (EARLY) {
        if (m > kb) {
          if constexpr (LATE) xy_reads(m - 1);
          base_reads(m - 1);
        }
      }
      {
        R zc[NQ], lo[NQ];
        zs(cu, zc);
        if constexpr (ZONLY) {
          if (m + 2 <= n3 + 1 && m + 1 <= ke) load12(m + 2, nn);
        } else {
          if (m + 1 <= n3 + 1) load12(m + 1, nx);
        }
        if (m == n3 + 1 || LIM == MFVM_LIM_NONE) {
  #pragma unroll
          for (int f = 0; f < NQ; ++f) lo[f] = lz_cur[f] = zc[f];
        } else {
          R zn[NQ];
          zs(nx, zn);
          slope_states<2, LIM, UNIF, R>(a.geo, ic, jc, m, visc ? 9 : 6, zp, zc, zn, lo, lz_cur);
        }
        const R s3 = g_area<UNIF, R>(a.geo, 2, ic - 1, jc - 1, m - 1);
        face_flux<2, RS, R>(a.p, lz_prev, lo, s3, fsz);
      }

      // ---- in-plane faces of plane m (all threads take part in the barriers) ----
      // Phase 1: x1 and x2 states of the own cells and of the tile-edge cells;
      // phase 2: F.S of every x1 and x2 face of the tile, each computed once;
      // phase 3: flux differences.  Helper tasks go to the first threads in
      // phase 1 and to the last threads in phase 2, so that no wave takes two.
      R accm[5], srcm[5];
      bool src_any = false;
      if (!LATE && m < ke) {
        if constexpr (ZONLY) {
          const long long oi = a.il.at(ic, jc, m);
  #pragma unroll
          for (int q = 0; q < 5; ++q) accm[q] = static_cast<const R*>(a.xy)[oi + q * a.il.fs];
        } else {
          inplane_div<LIM, RS, UNIF, R>(a, src, sm, BX, BY, tx, ty, tid, i0, j0, i, j, own, m, cu, accm);
        }
        cell_source<UNIF, R>(a.geo, a.p, ic, jc, m, cu, cu + 5, srcm, src_any);
      }

      // ---- finalize cell m-1 (stores from the lanes of interior columns) --------
      if (m > kb) {
        const int k = m - 1;
        R kr[5];
        R vol;
        const bool vp2 = a.geo.vpow2 != 0;
        if (!vp2) vol = g_vol<UNIF, R>(a.geo, ic, jc, k);
        if constexpr (LATE) {
          if constexpr (EARLY) {
  #pragma unroll
            for (int q = 0; q < 5; ++q) acc_prev[q] = fxy[q];
          } else {
            const long long oi = a.il.at(ic, jc, k);
  #pragma unroll
            for (int q = 0; q < 5; ++q) acc_prev[q] = static_cast<const R*>(a.xy)[oi + q * a.il.fs];
          }
          // (uniform Cartesian: the source reads U only, never the closure)
          cell_source<UNIF, R>(a.geo, a.p, ic, jc, k, zp, zp + 5, src_prev, src_any_prev);
        }
  #pragma unroll
        for (int q = 0; q < 5; ++q) {
          const R x = -((acc_prev[q] + fsz[q]) - fsz_prev[q]);
          kr[q] = vp2 ? x * R(a.geo.vinv) : qdiv(x, vol);
          if (src_any_prev) kr[q] = kr[q] + src_prev[q];
        }
        const long long oc = a.lay.at(ic, jc, k);
        const long long oi = a.il.at(ic, jc, k);
        R* out = static_cast<R*>(a.out);
        R* ka = static_cast<R*>(a.kacc);
        // The cell's reads of U0, kacc, lo and dinc are all issued before its
        // first store (the arrays do not alias the outputs), so that they take
        // one memory round trip instead of one per field.
        if constexpr (!COMP) {
          R nu_[5], u0v[5], kav[5];
          const R dtr = R(dt);
          if constexpr (EARLY) {
  #pragma unroll
            for (int q = 0; q < 5; ++q) {
              u0v[q] = u0_own ? zp[q] : f0[q];
              kav[q] = f1[q];
            }
          } else {
            const R* U0 = static_cast<const R*>(a.U0);
            const bool need0 = a.update == MFVM_UPD_BASE || a.update == MFVM_UPD_SSP || a.update == MFVM_UPD_ACC6;
            const bool needk = a.update == MFVM_UPD_ACC6 || a.acc == MFVM_ACC_ADD;
  #pragma unroll
            for (int q = 0; q < 5; ++q) {
              u0v[q] = u0_own ? zp[q] : (need0 ? U0[oc + q * fs] : R(0));
              kav[q] = needk ? ka[oi + q * a.il.fs] : R(0);
            }
          }
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            const R uq = zp[q];  // stage input U at cell m-1
            R o = uq;
            switch (a.update) {
              case MFVM_UPD_EULER: o = uq + dtr * kr[q]; break;
              case MFVM_UPD_BASE: o = u0v[q] + R(a.c0 * dt) * kr[q]; break;
              case MFVM_UPD_SSP: o = (R(a.c0) * u0v[q]) + (R(a.c1) * (uq + dtr * kr[q])); break;
              case MFVM_UPD_ACC6: o = u0v[q] + R(dt / 6.0) * (kav[q] + kr[q]); break;
              default: break;
            }
            nu_[q] = o;
            if (own) {
              if (a.update != MFVM_UPD_NONE) out[oc + q * fs] = o;
              if (a.acc == MFVM_ACC_SET) ka[oi + q * a.il.fs] = kr[q];
              else if (a.acc == MFVM_ACC_ADD) ka[oi + q * a.il.fs] = kav[q] + R(a.acc_coef) * kr[q];
            }
          }
          if (a.partials) {
            R rate, nu;
            cell_cfl<UNIF, R>(a.geo, a.p, ic, jc, k, nu_, rate, nu);
            rate_max = own ? nmax(rate_max, double(rate)) : rate_max;
            nu_max = own ? nmax(nu_max, double(nu)) : nu_max;
          }
        } else {
          // Compensated update: U0 = hi + lo (two FP32 words).  Stage states
          // are the FP32 values hi + d with an FP32 increment d (lo, below half
          // an ulp of hi, changes such a value by at most its own rounding, so
          // the other stages do not read it); the last stage adds its
          // increment to hi + lo in FP64 and splits the sum back into (hi, lo).
          // The SSP middle stage forms its state as its input (hi + d1, the
          // window's zp) + (d2 - d1), without reading hi: within about an ulp
          // of hi + d2, the FP32 state's own precision (the increments, and so
          // the FP64 accumulation, are unchanged).
          R* lo = static_cast<R*>(a.lo);
          R* dv = static_cast<R*>(a.dinc);
          R* sn = static_cast<R*>(a.snext);
          const bool fin = FIN == 2 ? a.comp >= MFVM_COMP_FINAL_SET : FIN == 1;
          const bool needd = a.comp == MFVM_COMP_SSP || a.comp == MFVM_COMP_FINAL_SSP;
          const bool needk = a.comp == MFVM_COMP_FINAL_ACC6 || a.acc == MFVM_ACC_ADD;
          const R* U0 = static_cast<const R*>(a.U0);
          double nw[5];
          R hv[5], lv[5], dvv[5], kav[5];
          const bool needh = fin || a.comp != MFVM_COMP_SSP;
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            const long long og = oc + q * fs, ii = oi + q * a.il.fs;
            hv[q] = (u0_own || !needh) ? zp[q] : U0[og];
            lv[q] = fin ? lo[og] : R(0);
            dvv[q] = needd ? dv[ii] : R(0);
            kav[q] = needk ? ka[ii] : R(0);
          }
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            const R kq = kr[q];
            const long long og = oc + q * fs, ii = oi + q * a.il.fs;
            const R h = hv[q], l = lv[q];
            if (!fin) {
              R d, nxt;
              if (a.comp == MFVM_COMP_SET) {
                d = R(a.c0 * dt) * kq;
                nxt = h + d;
              } else {
                d = R(a.c0) * (dvv[q] + R(dt) * kq);
                nxt = h + (d - dvv[q]);  // h: the stage input (see above)
              }
              if (own) {
                dv[ii] = d;
                sn[og] = nxt;
              }
            } else {
              double inc;
              if (a.comp == MFVM_COMP_FINAL_SET) inc = (a.c0 * dt) * double(kq);
              else if (a.comp == MFVM_COMP_FINAL_SSP) inc = a.c0 * (double(dvv[q]) + dt * double(kq));
              else inc = (dt / 6.0) * (double(kav[q]) + double(kq));
              const double v = (double(h) + double(l)) + inc;
              const R nh = R(v);
              if (own) {
                out[og] = nh;
                lo[og] = R(v - double(nh));
              }
              nw[q] = v;
            }
            if (own) {
              if (a.acc == MFVM_ACC_SET) ka[ii] = kq;
              else if (a.acc == MFVM_ACC_ADD) ka[ii] = kav[q] + R(a.acc_coef) * kq;
            }
          }
          if (fin && a.partials) {
            double rate, nu;
            cell_cfl<UNIF, double>(a.geo, a.p, ic, jc, k, nw, rate, nu);
            rate_max = own ? nmax(rate_max, rate) : rate_max;
            nu_max = own ? nmax(nu_max, nu) : nu_max;
          }
        }
      }

      // ---- rotate the window -------------------------------------------------------
      {
  #pragma unroll
        for (int q = 0; q < 5; ++q) fsz_prev[q] = fsz[q];
  #pragma unroll
        for (int f = 0; f < NQ; ++f) lz_prev[f] = lz_cur[f];
        if (!LATE && m < ke) {
  #pragma unroll
          for (int q = 0; q < 5; ++q) {
            acc_prev[q] = accm[q];
            src_prev[q] = srcm[q];
          }
          src_any_prev = src_any;
        }
        zs(cu, zp);
  #pragma unroll
        for (int f = 0; f < 12; ++f) cu[f] = nx[f];
        if constexpr (ZONLY) {
  #pragma unroll
          for (int f = 0; f < 12; ++f) nx[f] = nn[f];
        }
      }
    }

  }

  if (tl.slot) block_max2<double>(rate_max, nu_max, tl.slot, reinterpret_cast<double*>(smem_raw));
}

// ---------------------------------------------------------------------------
// Row-march x1/x2 pass
// ---------------------------------------------------------------------------
// One wave advances a strip of 64 x1 cells (one per lane) through a chunk of
// up to XR_ROWS x2 rows of one plane, marching along x2: the x2 stencil
// window, the upper x2 state of the previous row and the previous x2 face
// flux stay in registers, so each cell's x2 states and each x2 face flux are
// formed once (two extra x2 states per chunk).  Along x1 each lane forms its
// own cell's states; the lower state goes to the left neighbour and the face
// flux to the right neighbour through this wave's LDS, with no workgroup
// barrier (a wave's LDS operations complete in issue order).  The two faces
// at the strip's ends are prepared for every row of the chunk at once: the
// flux through face i0-1/2 and the lower state of cell i0+64.
constexpr int XR_W = 64;      // x1 cells per strip (one wave)
constexpr int XR_ROWS = 32;   // most x2 rows per chunk
constexpr int XR_WAVES = 4;   // independent waves per workgroup

// LDS of one wave (elements of R): the x1 exchange area X[NQ][64] -- each
// row's x1 fields, then lower x1 states, then face fluxes (every lane's next
// write depends on its reads of the previous contents) -- and per row of the
// chunk the strip-end lower states EX[XR_ROWS][NQ], fluxes EF[XR_ROWS][5]
// and x1 fields EQ[XR_ROWS][2][NQ] of cells i0-1 and i0+64.
constexpr int XR_LDS_ELEMS = NQ * XR_W + XR_ROWS * NQ + XR_ROWS * 5 + XR_ROWS * 2 * NQ;

// This wave's LDS writes are visible to its other lanes' later reads: the
// compiler keeps the LDS accesses on either side in program order, and the
// LDS performs one wave's accesses in issue order, so no wait for the writes
// to complete is needed (an lgkmcnt(0) wait here measured 1% slower).
MFVM_DEV void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The 11 stored quantities of cell o read by the x1/x2 pass: U (5), P, then
// T11, T12, T13 (row 1) and T22, T23 (row 2 without T21 = T12).
template <class R, class S>
MFVM_DEV void load_xy11(const S& src, long long o, bool visc, R w[11]) {
#pragma unroll
  for (int f = 0; f < 5; ++f) w[f] = src.u(o, f);
  w[5] = src.c(o, C_P);
  w[6] = visc ? src.c(o, C_T11) : R(0);
  w[7] = visc ? src.c(o, C_T12) : R(0);
  w[8] = visc ? src.c(o, C_T13) : R(0);
  w[9] = visc ? src.c(o, C_T22) : R(0);
  w[10] = visc ? src.c(o, C_T23) : R(0);
}
// Reconstruction fields along x1 / x2 from the 11 quantities.
template <class R>
MFVM_DEV void q_x1(const R w[11], R q[NQ]) {
#pragma unroll
  for (int f = 0; f < 9; ++f) q[f] = w[f];
}
template <class R>
MFVM_DEV void q_x2(const R w[11], R q[NQ]) {
#pragma unroll
  for (int f = 0; f < 6; ++f) q[f] = w[f];
  q[6] = w[7];
  q[7] = w[9];
  q[8] = w[10];
}

// x2 states of ghosted cell (i, j, m) from its x2 window (wm, w, wp: the
// reconstruction fields of rows j-1, j, j+1); ghost rows are unsloped.
template <int LIM, bool UNIF, class R>
MFVM_DEV void x2_states(const StageArgs& a, int i, int j, int m, const R wm[NQ], const R w[NQ], const R wp[NQ],
                        R lo[NQ], R hi[NQ]) {
  if (LIM == MFVM_LIM_NONE || j == 0 || j == a.lay.n2 + 1) {
#pragma unroll
    for (int f = 0; f < NQ; ++f) lo[f] = hi[f] = w[f];
    return;
  }
  slope_states<1, LIM, UNIF, R>(a.geo, i, j, m, a.p.visc ? 9 : 6, wm, w, wp, lo, hi);
}

// One wave's strip: x1 cells i0 .. i0+63, rows jb .. je-1 of plane m; the
// in-plane divergence of every own cell goes to a.xy.  lds: this wave's
// XR_LDS_ELEMS elements.  Every lane runs the same instructions: lanes past
// the strip's last interior cell read the x1 ghost column instead (their
// results are never stored), so no state carried from row to row depends on
// a branch.
template <int LIM, int RS, bool UNIF, class R, class S>
MFVM_DEV void ns_xyrow_strip(const StageArgs& a, const S& src, R* __restrict__ lds, int i0, int jb, int je, int m) {
  R* XR = lds;                    // [NQ][64]
  R* XF = lds;                    // [5][64], over XR
  R* EX = XR + NQ * XR_W;         // [XR_ROWS][NQ]
  R* EF = EX + XR_ROWS * NQ;      // [XR_ROWS][5]
  R* EQ = EF + XR_ROWS * 5;       // [XR_ROWS][2][NQ]
  R* XQ = lds;                    // [NQ][64], x1 fields, over XR
  const int n1 = a.lay.n1, n2 = a.lay.n2;
  const int l = threadIdx.x & (XR_W - 1);
  const int i = i0 + l;
  const bool own = i <= n1;             // interior cell of this lane
  const int ic = min(i, n1 + 1);        // column read by this lane
  const int ig = min(i, n1);            // column of its metric factors (a valid cell)
  const bool visc = a.p.visc != 0;
  const int nr = je - jb;

  // ---- strip ends, every row of the chunk ---------------------------------------
  {
    // pass 1: lanes [0, 32) the upper x1 state of cell (i0-1, jb+l), lanes
    // [32, 64) the lower x1 state of cell (i0+64, jb+l-32) (rows and the
    // column clamped into the array; only the chunk's rows are kept)
    const bool left = l < 32;
    const int r = left ? l : l - 32;
    const int cj = min(jb + r, je - 1);
    const int ci = left ? i0 - 1 : min(i0 + XR_W, n1 + 1);
    R lo[NQ], hi[NQ], qe[NQ];
    cell_states<0, LIM, UNIF, R>(a, src, ci, cj, m, lo, hi);
    load_q<0, R>(src, a.lay.at(ci, cj, m), visc, qe);
    if (r < nr) {
#pragma unroll
      for (int f = 0; f < NQ; ++f) EQ[(2 * r + (left ? 0 : 1)) * NQ + f] = qe[f];
    }
    if (!left && r < nr) {
#pragma unroll
      for (int f = 0; f < NQ; ++f) EX[r * NQ + f] = lo[f];
    }
    // pass 2: the lower x1 state of cell (i0, jb+l) and the flux through
    // face i0-1/2 (kept from lanes [0, nr))
    R lo0[NQ], hi0[NQ], F[5];
    cell_states<0, LIM, UNIF, R>(a, src, i0, cj, m, lo0, hi0);
    face_flux<0, RS, R>(a.p, hi, lo0, g_area<UNIF, R>(a.geo, 0, i0 - 1, cj - 1, m - 1), F);
    if (left && r < nr) {
#pragma unroll
      for (int q = 0; q < 5; ++q) EF[r * 5 + q] = F[q];
    }
  }
  wave_lds_sync();

  // ---- x2 march start: states of row jb-1 -----------------------------------------
  // rows j-1 (x2 fields), j, j+1 and j+2 (loaded two rows ahead of use)
  R wm[NQ], cw[11], cn[11], cnn[11], hi2p[NQ], Gp[5], axp[5];
  {
    R c11[11];
    load_xy11<R>(src, a.lay.at(ic, jb - 1, m), visc, c11);
    q_x2<R>(c11, wm);
    load_xy11<R>(src, a.lay.at(ic, jb, m), visc, cw);
    load_xy11<R>(src, a.lay.at(ic, min(jb + 1, n2 + 1), m), visc, cn);
    load_xy11<R>(src, a.lay.at(ic, min(jb + 2, n2 + 1), m), visc, cnn);
    if (jb - 1 == 0 || LIM == MFVM_LIM_NONE) {
#pragma unroll
      for (int f = 0; f < NQ; ++f) hi2p[f] = wm[f];
    } else {
      R wmm[NQ], w0[NQ], lo[NQ];
      load_xy11<R>(src, a.lay.at(ic, jb - 2, m), visc, c11);
      q_x2<R>(c11, wmm);
      q_x2<R>(cw, w0);
      x2_states<LIM, UNIF, R>(a, ig, jb - 1, m, wmm, wm, w0, lo, hi2p);
    }
  }
  const long long ifs = a.il.fs;
  R* xy = static_cast<R*>(a.xy);
  // x2 part of row j: the states of row j, face j-1/2 and the divergence of
  // row j-1 (for j > jb).
  auto x2_row = [&](int j, const R (&c0)[11], const R (&c1)[11]) {
    {
      R w[NQ], wp[NQ], lo2[NQ], hi2[NQ], G[5];
      q_x2<R>(c0, w);
      q_x2<R>(c1, wp);  // row j+1 (unused when row j is the ghost row)
      x2_states<LIM, UNIF, R>(a, ig, j, m, wm, w, wp, lo2, hi2);
      face_flux<1, RS, R>(a.p, hi2p, lo2, g_area<UNIF, R>(a.geo, 1, ig - 1, j - 1, m - 1), G);
      if (own && j > jb) {
        const long long oi = a.il.at(i, j - 1, m);
#pragma unroll
        for (int q = 0; q < 5; ++q) xy[oi + q * ifs] = (axp[q] + G[q]) - Gp[q];
      }
#pragma unroll
      for (int q = 0; q < 5; ++q) Gp[q] = G[q];
#pragma unroll
      for (int f = 0; f < NQ; ++f) {
        hi2p[f] = hi2[f];
        wm[f] = w[f];
      }
    }
  };
  // x1 faces of row j: states, exchange through LDS, flux differences.
  auto x1_row = [&](int j, const R (&c0)[11]) {
    {
      wave_lds_sync();  // the previous row's reads of XF are done
      R q[NQ], lo1[NQ], hi1[NQ];
      q_x1<R>(c0, q);
      if (LIM == MFVM_LIM_NONE) {
#pragma unroll
        for (int f = 0; f < NQ; ++f) lo1[f] = hi1[f] = q[f];
      } else {
        // x1 neighbours from the adjacent lanes (the strip ends from EQ); the
        // ghost column's neighbours are itself: zero differences, zero slopes
        // over a valid cell length, raw states
#pragma unroll
        for (int f = 0; f < NQ; ++f) XQ[f * XR_W + l] = q[f];
        wave_lds_sync();
        R qm[NQ], qp[NQ];
        const bool in = i <= n1;
        const R* pm = !in ? XQ + l : (l == 0 ? EQ + 2 * (j - jb) * NQ : XQ + l - 1);
        const R* pp = !in ? XQ + l : (l == XR_W - 1 ? EQ + (2 * (j - jb) + 1) * NQ : XQ + l + 1);
        const int sm_ = (in && l == 0) ? 1 : XR_W;
        const int sp_ = (in && l == XR_W - 1) ? 1 : XR_W;
#pragma unroll
        for (int f = 0; f < NQ; ++f) {
          qm[f] = pm[f * sm_];
          qp[f] = pp[f * sp_];
        }
        slope_states<0, LIM, UNIF, R>(a.geo, ig, j, m, visc ? 9 : 6, qm, q, qp, lo1, hi1);
      }
#pragma unroll
      for (int f = 0; f < NQ; ++f) XR[f * XR_W + l] = lo1[f];
      wave_lds_sync();
      R rr[NQ], F[5];
      const R* src_r = l == XR_W - 1 ? EX + (j - jb) * NQ : XR + l + 1;
      const int st = l == XR_W - 1 ? 1 : XR_W;
#pragma unroll
      for (int f = 0; f < NQ; ++f) rr[f] = src_r[f * st];
      face_flux<0, RS, R>(a.p, hi1, rr, g_area<UNIF, R>(a.geo, 0, ig, j - 1, m - 1), F);
#pragma unroll
      for (int q2 = 0; q2 < 5; ++q2) XF[q2 * XR_W + l] = F[q2];
      wave_lds_sync();
      const R* src_l = l == 0 ? EF + (j - jb) * 5 : XF + l - 1;
      const int sl = l == 0 ? 1 : XR_W;
#pragma unroll
      for (int q2 = 0; q2 < 5; ++q2) axp[q2] = F[q2] - src_l[q2 * sl];
    }
  };
  // Rows j (c0) and j+1 (c1) are read, row j+2 (c2) is in flight; row j+3
  // then loads into c0.
  auto step = [&](int j, R (&c0)[11], const R (&c1)[11]) {
    x2_row(j, c0, c1);
    x1_row(j, c0);
    load_xy11<R>(src, a.lay.at(ic, min(j + 3, n2 + 1), m), visc, c0);
  };
  if constexpr (sizeof(R) == sizeof(double)) {
    // float64: the rows of the chunk three at a time, the three carried rows
    // rotating their roles (no copies, and a row loaded here is first read two
    // rows later); then the last rows one at a time; then the x2 part of row
    // je (the row after the chunk), which closes the divergence of row je-1.
    // Every branch here is wave-uniform.
    int j = jb;
    for (; j + 3 <= je; j += 3) {
      step(j, cw, cn);
      step(j + 1, cn, cnn);
      step(j + 2, cnn, cw);
    }
    for (; j < je; ++j) {
      step(j, cw, cn);
#pragma unroll
      for (int f = 0; f < 11; ++f) {
        const R t = cw[f];
        cw[f] = cn[f];
        cn[f] = cnn[f];
        cnn[f] = t;
      }
    }
    x2_row(je, cw, cn);
  } else {
    // float32 (four waves per SIMD, 128 VGPRs): two rows per trip with the
    // carried rows copied, which holds fewer of them live than the unrolled
    // rotation above (that one spills at 128).
    auto row = [&](int j) -> bool {
      x2_row(j, cw, cn);
      if (j == je) return true;
      x1_row(j, cw);
#pragma unroll
      for (int f = 0; f < 11; ++f) {
        cw[f] = cn[f];
        cn[f] = cnn[f];
      }
      load_xy11<R>(src, a.lay.at(ic, min(j + 3, n2 + 1), m), visc, cnn);
      return false;
    };
    for (int j = jb;; j += 2) {
      if (row(j) || row(j + 1)) break;
    }
  }
}

// x1/x2 pass of a split stage in row-march form: workgroups of XR_WAVES
// independent waves, one (strip, row chunk, plane) task per wave; rows per
// chunk a.kz (<= XR_ROWS).
template <int LIM, int RS, bool UNIF, class R, class S>
MFVM_DEV void ns_xyrow_body(const StageArgs& a, const S& src, long long task) {
  extern __shared__ double smem_raw[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / XR_W);
  R* lds = reinterpret_cast<R*>(smem_raw) + wave * XR_LDS_ELEMS;
  const int ns = (a.lay.n1 + XR_W - 1) / XR_W;
  const int rows = a.kz;
  const int nc = (a.lay.n2 + rows - 1) / rows;
  const int m1 = a.xm1 > 0 ? a.xm1 : a.lay.n3 + 1;
  const long long ntask = (long long)ns * nc * (m1 - a.xm0);
  if (task >= ntask) return;
  const int s = (int)(task % ns);
  const long long rest = task / ns;
  const int c = (int)(rest % nc);
  const int m = a.xm0 + (int)(rest / nc);
  const int jb = 1 + c * rows;
  const int je = min(jb + rows, a.lay.n2 + 1);
  ns_xyrow_strip<LIM, RS, UNIF, R>(a, src, lds, 1 + s * XR_W, jb, je, m);
}

// CFL maxima of one tile (navier_stokes.py:512-583): rate_c and nu over the
// tile's cells, reduced to the tile's slot.  lo != NULL: compensated FP32
// state, rates in FP64 of hi + lo.
template <bool UNIF, class R>
MFVM_DEV void cfl_tile_body(const Layout& lay, const Geo& geo, const NSPar& p, const R* U, const R* lo,
                            int kz, const TileCtx& tl) {
  extern __shared__ double smem_raw[];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int i = 1 + tl.x * tl.bx + tid % tl.bx;
  const int j = 1 + tl.y * tl.by + tid / tl.bx;
  const int kb = 1 + tl.z * kz, ke = min(kb + kz, lay.n3 + 1);
  double rate_max = 0.0, nu_max = 0.0;
  if (i <= lay.n1 && j <= lay.n2) {
    for (int k = kb; k < ke; ++k) {
      const long long o = lay.at(i, j, k);
      if (lo) {
        double u[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) u[q] = double(U[o + q * lay.fs]) + double(lo[o + q * lay.fs]);
        double rate, nu;
        cell_cfl<UNIF, double>(geo, p, i, j, k, u, rate, nu);
        rate_max = nmax(rate_max, rate);
        nu_max = nmax(nu_max, nu);
      } else {
        R u[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) u[q] = U[o + q * lay.fs];
        R rate, nu;
        cell_cfl<UNIF, R>(geo, p, i, j, k, u, rate, nu);
        rate_max = nmax(rate_max, double(rate));
        nu_max = nmax(nu_max, double(nu));
      }
    }
  }
  block_max2<double>(rate_max, nu_max, tl.slot, smem_raw);
}

// dt = cfl / (max rate_c + (2*max nu) * smax) over nblk partial pairs
// (navier_stokes.py:571-583), reduced by one workgroup of <= 1024 threads.
MFVM_DEV double dt_reduce_block(const double* partials, long long nblk, double cfl, double smax) {
  __shared__ double sa[1024], sb[1024];
  double a = 0.0, b = 0.0;
  for (long long t = threadIdx.x; t < nblk; t += blockDim.x) {
    a = nmax(a, partials[2 * t]);
    b = nmax(b, partials[2 * t + 1]);
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] = nmax(sa[threadIdx.x], sa[threadIdx.x + s]);
      sb[threadIdx.x] = nmax(sb[threadIdx.x], sb[threadIdx.x + s]);
    }
    __syncthreads();
  }
  return cfl / (sa[0] + (2.0 * sb[0]) * smax);
}

// ---------------------------------------------------------------------------
// Host-side helpers shared by the launchers
// ---------------------------------------------------------------------------

struct Tile {
  int bx, by, kz;
};

inline Tile ns_tile(const mfvm_grid* g, const mfvm_ns_params* p) {
  Tile t;
  t.bx = p->tile[0] > 0 ? p->tile[0] : 64;
  t.by = p->tile[1] > 0 ? p->tile[1] : 4;
  t.kz = p->tile[2] > 0 ? p->tile[2] : (g->n[2] < 32 ? g->n[2] : 32);
  return t;
}

// LDS bytes of the stage body for a bx x by tile.
inline size_t stage_lds_bytes(int bx, int by, size_t rsize) {
  const size_t nrx = (size_t)by * (bx + 1), nry = (size_t)(by + 1) * bx;
  size_t lds = (NQ * nrx + NQ * by + NQ * nry + NQ * bx + 5 * nrx + 5 * nry) * rsize;
  const size_t red = 2 * sizeof(double) * ((bx * by + 63) / 64);
  return lds < red ? red : lds;
}

inline NSPar make_ns_par(const mfvm_ns_params* p) {
  NSPar q;
  q.gamma = p->gamma;
  q.gm1 = p->gamma - 1.0;
  q.mu = p->mu;
  q.g[0] = p->g[0];
  q.g[1] = p->g[1];
  q.g[2] = p->g[2];
  q.visc = p->mu != 0.0;
  q.grav = (p->g[0] != 0.0) || (p->g[1] != 0.0) || (p->g[2] != 0.0);
  return q;
}

// BufSrc reads are valid: U (5 fields) and C (NS_NC fields) below 4 GiB.
inline bool buf_src_ok(const Layout& lay, size_t rsize) {
  return (unsigned long long)NS_NC * (unsigned long long)lay.fs * rsize < (1ull << 32);
}

}  // namespace mfvm
