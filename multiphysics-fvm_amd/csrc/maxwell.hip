// Maxwell kernels, B/D form with divergence cleaning (reference
// src/physics/maxwell.py).
//
// maxwell_res_kernel: residuals rB = div B and rD = div D - rho on interior
// cells (maxwell.py:491-508), with the edge padding of maxwell.py:511-512
// written into the face ghosts.
// maxwell_stage_kernel (and its split form maxwell_xy_kernel +
// maxwell_z_kernel): a 64 x 4 tile of (x1, x2) columns marching in x3;
// MUSCL/PC states of the six fields on both faces along each direction, curl
// flux of H = B/mu and E = D/eps (maxwell.py:137-181), Rusanov or HLL with
// the light-speed wave set, geometric source (:366-371), Ohm current
// (:415-420), cleaning source (:514-528), divergence and the fused stage
// update.  Each face flux is computed once.
#include "common.h"

namespace mfvm {

struct MXPar {
  double mu, ep, inv_mu, inv_ep;  // inverses used when mu / eps are powers of two
  int mu_pow2, ep_pow2;
  double c;                        // 1/sqrt(mu*eps)
  double negsigma, rho, negchiB, negchiD;
};

struct MXArgs {
  Layout lay;
  ILayout il;
  Geo geo;
  MXPar p;
  const void* U;
  const void* Rr;  // residuals [2][ghosted]
  StageIO io;
  void* xy;        // split stage: interior [6] in-plane flux differences
  const void* sigf;  // per-cell conductivity (interior layout, R) or NULL: p.negsigma
};

template <class R>
MFVM_DEV R mx_H(const MXPar& p, R b) { return p.mu_pow2 ? b * R(p.inv_mu) : b / R(p.mu); }
template <class R>
MFVM_DEV R mx_E(const MXPar& p, R d) { return p.ep_pow2 ? d * R(p.inv_ep) : d / R(p.ep); }

// ---------------------------------------------------------------------------
// Residuals
// ---------------------------------------------------------------------------
// div B and div D of interior cell (i,j,k) (maxwell.py:491-508): centred
// differences over 2 L_d, plus on curvilinear grids the scale-factor terms
// (F_d (dh_d' + dh_d'') / h_d).  The two fields are formed together: the
// metrics once for both, and each denominator's two quotients share one
// reciprocal refinement.
template <class R>
MFVM_DEV void mx_div2(const Layout& lay, const Geo& g, const R* U, int i, int j, int k, R& dB, R& dD) {
  const long long o = lay.at(i, j, k), fs = lay.fs;
  const long long st[3] = {1, lay.sy, lay.sz};
  R gb[3], gd[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const R nb = U[o + a * fs + st[a]] - U[o + a * fs - st[a]];
    const R nd = U[o + (3 + a) * fs + st[a]] - U[o + (3 + a) * fs - st[a]];
    if (g.lpow2[a]) {
      gb[a] = nb * R(0.5 * g.linv[a]);
      gd[a] = nd * R(0.5 * g.linv[a]);
    } else {
      const R n[2] = {nb, nd};
      R q[2];
      ieee_div_by<2, R>(n, R(2) * met<R>(g, MFVM_MET_L1 + a, i - 1, j - 1, k - 1), q);
      gb[a] = q[0];
      gd[a] = q[1];
    }
  }
  dB = (gb[0] + gb[1]) + gb[2];
  dD = (gd[0] + gd[1]) + gd[2];
  if (g.curv) {
    const int a = i - 1, b = j - 1, e = k - 1;
    const R m1 = met<R>(g, MFVM_MET_D1H2, a, b, e) + met<R>(g, MFVM_MET_D1H3, a, b, e);
    const R m2 = met<R>(g, MFVM_MET_D2H1, a, b, e) + met<R>(g, MFVM_MET_D2H3, a, b, e);
    const R m3 = met<R>(g, MFVM_MET_D3H1, a, b, e) + met<R>(g, MFVM_MET_D3H2, a, b, e);
    const R h[3] = {met<R>(g, MFVM_MET_H1, a, b, e), met<R>(g, MFVM_MET_H2, a, b, e), met<R>(g, MFVM_MET_H3, a, b, e)};
    const R mm[3] = {m1, m2, m3};
    R tb[3], td[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const R n[2] = {U[o + c * fs] * mm[c], U[o + (3 + c) * fs] * mm[c]};
      R q[2];
      ieee_div_by<2, R>(n, h[c], q);
      tb[c] = q[0];
      td[c] = q[1];
    }
    dB = ((dB + tb[0]) + tb[1]) + tb[2];
    dD = ((dD + td[0]) + td[1]) + td[2];
  }
}

// rhof: per-cell charge density (interior layout) or NULL: p.rho everywhere.
// A workgroup row is one wave (blockDim.x = 64), so j is wave-uniform and the
// x2 and x3 metric factors are scalar loads.
template <class R>
__global__ void maxwell_res_kernel(Layout lay, Geo geo, MXPar p, const R* __restrict__ U, R* __restrict__ Rr,
                                   const R* __restrict__ rhof) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int k = blockIdx.z;
  const int n1 = lay.n1, n2 = lay.n2, n3 = lay.n3;
  if (i > n1 + 1 || j > n2 + 1) return;
  const int nb = (i == 0 || i == n1 + 1) + (j == 0 || j == n2 + 1) + (k == 0 || k == n3 + 1);
  if (nb > 1) return;
  // edge padding: a face ghost takes the value of its adjacent interior cell
  const int ii = i == 0 ? 1 : (i == n1 + 1 ? n1 : i);
  const int jj = j == 0 ? 1 : (j == n2 + 1 ? n2 : j);
  const int kk = k == 0 ? 1 : (k == n3 + 1 ? n3 : k);
  R rB, dD;
  mx_div2<R>(lay, geo, U, ii, jj, kk, rB, dD);
  const long long ic = (long long)(kk - 1) * lay.n2 * lay.n1 + (long long)(jj - 1) * lay.n1 + (ii - 1);
  const R rD = dD - (rhof ? rhof[ic] : R(p.rho));
  const long long o = lay.at(i, j, k);
  Rr[o] = rB;
  Rr[o + lay.fs] = rD;
}

// ---------------------------------------------------------------------------
// Stage
// ---------------------------------------------------------------------------

// Curl flux along D of one state (B[3], D[3]) -> F[6].
template <int D, class R>
MFVM_DEV void mx_flux(const MXPar& p, const R* u, R* F) {
  const R H1 = mx_H<R>(p, u[0]), H2 = mx_H<R>(p, u[1]), H3 = mx_H<R>(p, u[2]);
  const R E1 = mx_E<R>(p, u[3]), E2 = mx_E<R>(p, u[4]), E3 = mx_E<R>(p, u[5]);
  if (D == 0) {
    F[0] = R(0); F[1] = -E3; F[2] = E2;
    F[3] = R(0); F[4] = H3; F[5] = -H2;
  } else if (D == 1) {
    F[0] = E3; F[1] = R(0); F[2] = -E1;
    F[3] = -H3; F[4] = R(0); F[5] = H1;
  } else {
    F[0] = -E2; F[1] = E1; F[2] = R(0);
    F[3] = H2; F[4] = -H1; F[5] = R(0);
  }
}

template <int D, int RS, class R>
MFVM_DEV void mx_face(const MXPar& p, const R* l, const R* r, R area, R* fs) {
  R fl[6], fr[6];
  mx_flux<D, R>(p, l, fl);
  mx_flux<D, R>(p, r, fr);
  const R c = R(p.c);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    R F;
    if (RS == MFVM_RS_RUSANOV) {
      F = ((fl[q] + fr[q]) / R(2)) - ((c * (r[q] - l[q])) / R(2));
    } else {
      const R smin = -c, smax = c;
      F = ((smax * fl[q] - smin * fr[q]) + (smin * smax) * (r[q] - l[q])) / (smax - smin);
    }
    fs[q] = (F + R(0)) * area;
  }
}

// MUSCL states of one cell along D from its stencil (qm, q, qp): limited
// slope s = limiter(back, front) * L with back/front = difference / L, then
// hi = q + s/2 (upper face), lo = q - s/2 (reconstructors.py:115-148).
template <int D, int LIM, class R>
MFVM_DEV void mx_slopes(const Geo& geo, int i, int j, int k, const R* qm, const R* q, const R* qp, R lo[6],
                        R hi[6]) {
  const R len = geo.lpow2[D] ? R(1.0 / geo.linv[D]) : met<R>(geo, MFVM_MET_L1 + D, i - 1, j - 1, k - 1);
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const R b = div_len<R>(geo, D, q[f] - qm[f], len);
    const R w = div_len<R>(geo, D, qp[f] - q[f], len);
    const R s = limit<LIM, R>(b, w) * len;
    hi[f] = q[f] + s / R(2);
    lo[f] = q[f] - s / R(2);
  }
}

// States of ghosted cell (i,j,k) along D, loading its stencil; ghost cells
// carry zero slope (reconstructors.py:127-132).
template <int D, int LIM, class R>
MFVM_DEV void mx_states(const MXArgs& a, const R* U, int i, int j, int k, R lo[6], R hi[6]) {
  const long long fs = a.lay.fs;
  const long long st = D == 0 ? 1 : (D == 1 ? a.lay.sy : a.lay.sz);
  const long long o = a.lay.at(i, j, k);
  R q[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) q[f] = U[o + f * fs];
  const int c = D == 0 ? i : (D == 1 ? j : k);
  const int n = D == 0 ? a.lay.n1 : (D == 1 ? a.lay.n2 : a.lay.n3);
  if (LIM == MFVM_LIM_NONE || c < 1 || c > n) {
#pragma unroll
    for (int f = 0; f < 6; ++f) lo[f] = hi[f] = q[f];
    return;
  }
  R qm[6], qp[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    qm[f] = U[o + f * fs - st];
    qp[f] = U[o + f * fs + st];
  }
  mx_slopes<D, LIM, R>(a.geo, i, j, k, qm, q, qp, lo, hi);
}

// Right-hand side and stage update of interior cell (i,j,k) from its flux
// differences acc = ((((F1+ - F1-) + F2+) - F2-) + F3+) - F3- and its values
// u: rhs = -acc/V + ((sg + sp) + sc) (maxwell.py:708, 722-727).
template <class R>
MFVM_DEV void mx_finish(const MXArgs& a, const R* __restrict__ Rr, int i, int j, int k, const R acc[6],
                        const R u[6]) {
  const long long o = a.lay.at(i, j, k), fs = a.lay.fs;
  // geometric source (exactly +-0 on Cartesian grids)
  R sg[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
  if (a.geo.curv) {
    const int x = i - 1, y = j - 1, z = k - 1;
    const R h1 = met<R>(a.geo, MFVM_MET_H1, x, y, z), h2 = met<R>(a.geo, MFVM_MET_H2, x, y, z),
            h3 = met<R>(a.geo, MFVM_MET_H3, x, y, z);
    const R d1h2 = met<R>(a.geo, MFVM_MET_D1H2, x, y, z), d1h3 = met<R>(a.geo, MFVM_MET_D1H3, x, y, z);
    const R d2h1 = met<R>(a.geo, MFVM_MET_D2H1, x, y, z), d2h3 = met<R>(a.geo, MFVM_MET_D2H3, x, y, z);
    const R d3h1 = met<R>(a.geo, MFVM_MET_D3H1, x, y, z), d3h2 = met<R>(a.geo, MFVM_MET_D3H2, x, y, z);
    const R H1 = mx_H<R>(a.p, u[0]), H2 = mx_H<R>(a.p, u[1]), H3 = mx_H<R>(a.p, u[2]);
    const R E1 = mx_E<R>(a.p, u[3]), E2 = mx_E<R>(a.p, u[4]), E3 = mx_E<R>(a.p, u[5]);
    // tensors t[a][b] with 0.0 on the diagonal (maxwell.py:146-165)
    const R tb[3][3] = {{R(0), E3, -E2}, {-E3, R(0), E1}, {E2, -E1, R(0)}};
    const R td[3][3] = {{R(0), -H3, H2}, {H3, R(0), -H1}, {-H2, H1, R(0)}};
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const R(*t)[3] = w == 0 ? tb : td;
      sg[3 * w + 0] = ((((t[0][1] * d2h1) / h2) + ((t[0][2] * d3h1) / h3)) - ((t[1][1] * d1h2) / h1)) -
                      ((t[2][2] * d1h3) / h1);
      sg[3 * w + 1] = ((((t[1][0] * d1h2) / h1) + ((t[1][2] * d3h2) / h3)) - ((t[0][0] * d2h1) / h2)) -
                      ((t[2][2] * d2h3) / h2);
      sg[3 * w + 2] = ((((t[2][0] * d1h3) / h1) + ((t[2][1] * d2h3) / h2)) - ((t[0][0] * d3h1) / h3)) -
                      ((t[1][1] * d3h2) / h3);
    }
  }
  // cleaning source: -chi * centred gradient of the edge-padded residual
  const long long st[3] = {1, a.lay.sy, a.lay.sz};
  R sc[6];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const R len2 = a.geo.lpow2[d] ? R(0) : R(2) * met<R>(a.geo, MFVM_MET_L1 + d, i - 1, j - 1, k - 1);
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const R dr = Rr[o + w * fs + st[d]] - Rr[o + w * fs - st[d]];
      const R gr = a.geo.lpow2[d] ? dr * R(0.5 * a.geo.linv[d]) : dr / len2;
      sc[3 * w + d] = R(w == 0 ? a.p.negchiB : a.p.negchiD) * gr;
    }
  }
  R kr[6];
  R vol = a.geo.vpow2 ? R(0) : met<R>(a.geo, MFVM_MET_V, i - 1, j - 1, k - 1);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    // J = -sigma E (maxwell.py:85-87); an array sigma is negated per cell
    const R ns = a.sigf ? -static_cast<const R*>(a.sigf)[a.il.at(i, j, k)] : R(a.p.negsigma);
    const R sp = q < 3 ? R(0) : ns * mx_E<R>(a.p, u[q]);
    const R src = (sg[q] + sp) + sc[q];
    const R x = -acc[q];
    kr[q] = (a.geo.vpow2 ? x * R(a.geo.vinv) : x / vol) + src;
  }
  stage_store<R>(a.io, 6, fs, a.il.fs, o, a.il.at(i, j, k), u, kr, nullptr);
}

// Tile of the stage kernel: 64 x 4 (x1, x2) columns, one thread each.
constexpr int MX_BX = 64, MX_BY = 4;

// LDS of the in-plane exchange: RX[6][BY][BX+1] lower-face x1 states (the
// last column: the cell beyond the tile), LX[6][BY] upper-face x1 states of
// the cells before the tile, RY[6][BY+1][BX] / LY[6][BX] likewise for x2,
// then the face fluxes FX[6][BY][BX+1] and FY[6][BY+1][BX] in space of their
// own, so that a plane takes two workgroup barriers (states written, fluxes
// written): the next plane writes its states once every thread is past the
// second barrier, after its last state read, and its fluxes once every
// thread is past its first barrier, after its last flux read.
constexpr int MX_NRX = MX_BY * (MX_BX + 1), MX_NRY = (MX_BY + 1) * MX_BX;
constexpr size_t mx_lds_elems() { return 6 * (MX_NRX + MX_BY + MX_NRY + MX_BX) + 6 * (MX_NRX + MX_NRY); }

// In-plane flux differences of plane m of the tile:
// acc = ((F1(i+1/2) - F1(i-1/2)) + F2(j+1/2)) - F2(j-1/2).  Every thread of
// the workgroup calls it (two barriers); cu: the own cell's values.
template <int LIM, int RS, class R>
MFVM_DEV void mx_inplane(const MXArgs& a, const R* __restrict__ U, R* sm, int i0, int j0, int m, const R cu[6],
                         R acc[6]) {
  constexpr int BX = MX_BX, BY = MX_BY, NT = MX_BX * MX_BY, wx = MX_BX + 1;
  constexpr int nrx = MX_NRX, nry = MX_NRY;
  R* RX = sm;
  R* LX = RX + 6 * nrx;
  R* RY = LX + 6 * BY;
  R* LY = RY + 6 * nry;
  R* FX = LY + 6 * BX;
  R* FY = FX + 6 * nrx;
  const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), tid = ty * BX + tx;
  const int n1 = a.lay.n1, n2 = a.lay.n2;
  const long long fs = a.lay.fs;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i <= n1 && j <= n2;
  R hx[6], hy[6];
  if (j <= n2 && i <= n1 + 1) {
    R lo[6];
    if (i <= n1) {
      if (LIM == MFVM_LIM_NONE) {
#pragma unroll
        for (int f = 0; f < 6; ++f) lo[f] = hx[f] = cu[f];
      } else {
        const long long o = a.lay.at(i, j, m);
        R qm[6], qp[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          qm[f] = U[o + f * fs - 1];
          qp[f] = U[o + f * fs + 1];
        }
        mx_slopes<0, LIM, R>(a.geo, i, j, m, qm, cu, qp, lo, hx);
      }
    } else {
      mx_states<0, LIM, R>(a, U, i, j, m, lo, hx);  // x1 ghost: raw
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) RX[f * nrx + ty * wx + tx] = lo[f];
  }
  if (i <= n1 && j <= n2 + 1) {
    R lo[6];
    if (j <= n2) {
      if (LIM == MFVM_LIM_NONE) {
#pragma unroll
        for (int f = 0; f < 6; ++f) lo[f] = hy[f] = cu[f];
      } else {
        const long long o = a.lay.at(i, j, m);
        R qm[6], qp[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          qm[f] = U[o + f * fs - a.lay.sy];
          qp[f] = U[o + f * fs + a.lay.sy];
        }
        mx_slopes<1, LIM, R>(a.geo, i, j, m, qm, cu, qp, lo, hy);
      }
    } else {
      mx_states<1, LIM, R>(a, U, i, j, m, lo, hy);  // x2 ghost: raw
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) RY[f * nry + ty * BX + tx] = lo[f];
  }
  // helpers: [0, 2BX) x2 tile-edge cells, [2BX, 2BX+2BY) x1 tile-edge cells
  if (tid < 2 * (BX + BY)) {
    const int t = tid;
    R lo[6], hi[6];
    if (t < 2 * BX) {
      const int c = t % BX, side = t / BX, ii = i0 + c;
      if (ii <= n1) {
        if (side == 0) {
          mx_states<1, LIM, R>(a, U, ii, j0 - 1, m, lo, hi);
#pragma unroll
          for (int f = 0; f < 6; ++f) LY[f * BX + c] = hi[f];
        } else if (j0 + BY <= n2 + 1) {
          mx_states<1, LIM, R>(a, U, ii, j0 + BY, m, lo, hi);
#pragma unroll
          for (int f = 0; f < 6; ++f) RY[f * nry + BY * BX + c] = lo[f];
        }
      }
    } else {
      const int u = t - 2 * BX, r = u % BY, side = u / BY, jj = j0 + r;
      if (jj <= n2) {
        if (side == 0) {
          mx_states<0, LIM, R>(a, U, i0 - 1, jj, m, lo, hi);
#pragma unroll
          for (int f = 0; f < 6; ++f) LX[f * BY + r] = hi[f];
        } else if (i0 + BX <= n1 + 1) {
          mx_states<0, LIM, R>(a, U, i0 + BX, jj, m, lo, hi);
#pragma unroll
          for (int f = 0; f < 6; ++f) RX[f * nrx + r * wx + BX] = lo[f];
        }
      }
    }
  }
  __syncthreads();
  // fluxes: the own cell's upper x1 and x2 faces; the tile's lower-edge
  // faces by the last BX + BY threads
  R fx[6], fy[6], fh[6];
  if (own) {
    R rr[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) rr[f] = RX[f * nrx + ty * wx + tx + 1];
    mx_face<0, RS, R>(a.p, hx, rr, met<R>(a.geo, MFVM_MET_S1, i, j - 1, m - 1), fx);
#pragma unroll
    for (int f = 0; f < 6; ++f) rr[f] = RY[f * nry + (ty + 1) * BX + tx];
    mx_face<1, RS, R>(a.p, hy, rr, met<R>(a.geo, MFVM_MET_S2, i - 1, j, m - 1), fy);
  }
  const int u = NT - 1 - tid;
  int hkind = 0, hidx = 0;
  if (u < BX + BY) {
    R ll[6], rr[6];
    if (u < BX) {
      const int c = u, ii = i0 + c;
      if (ii <= n1) {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          ll[f] = LY[f * BX + c];
          rr[f] = RY[f * nry + c];
        }
        mx_face<1, RS, R>(a.p, ll, rr, met<R>(a.geo, MFVM_MET_S2, ii - 1, j0 - 1, m - 1), fh);
        hkind = 1;
        hidx = c;
      }
    } else {
      const int r = u - BX, jj = j0 + r;
      if (jj <= n2) {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          ll[f] = LX[f * BY + r];
          rr[f] = RX[f * nrx + r * wx];
        }
        mx_face<0, RS, R>(a.p, ll, rr, met<R>(a.geo, MFVM_MET_S1, i0 - 1, jj - 1, m - 1), fh);
        hkind = 2;
        hidx = r * wx;
      }
    }
  }
  if (own) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      FX[q * nrx + ty * wx + tx + 1] = fx[q];
      FY[q * nry + (ty + 1) * BX + tx] = fy[q];
    }
  }
  if (hkind == 1) {
#pragma unroll
    for (int q = 0; q < 6; ++q) FY[q * nry + hidx] = fh[q];
  } else if (hkind == 2) {
#pragma unroll
    for (int q = 0; q < 6; ++q) FX[q * nrx + hidx] = fh[q];
  }
  __syncthreads();
  if (own) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const R ax = FX[q * nrx + ty * wx + tx + 1] - FX[q * nrx + ty * wx + tx];
      acc[q] = (ax + FY[q * nry + (ty + 1) * BX + tx]) - FY[q * nry + ty * BX + tx];
    }
  }
}

// Stage body: one workgroup per 64 x 4 tile and chunk of kz planes of x3,
// one thread per (x1, x2) column.  The x3 states come from a register window
// over planes m-1, m, m+1 of the column; the x1/x2 states of plane m are
// exchanged through LDS (mx_inplane), or, ZONLY, their flux differences are
// read from a.xy (written by maxwell_xy_kernel).  Every face flux is
// computed once, with the per-face arithmetic and the summation order of
// maxwell.py:680-727.
template <int LIM, int RS, class R, bool ZONLY>
MFVM_DEV void mx_stage_body(const MXArgs& a, int kz, int remap) {
  extern __shared__ double smem_raw[];
  R* sm = reinterpret_cast<R*>(smem_raw);
  const TileId tl = tile_id(remap != 0);
  const int n1 = a.lay.n1, n2 = a.lay.n2, n3 = a.lay.n3;
  const long long fs = a.lay.fs;
  const int i0 = 1 + tl.x * MX_BX, j0 = 1 + tl.y * MX_BY;
  const int i = i0 + threadIdx.x, j = j0 + __builtin_amdgcn_readfirstlane(threadIdx.y);  // a row is one wave
  const int kb = 1 + tl.z * kz, ke = min(kb + kz, n3 + 1);
  const bool own = i <= n1 && j <= n2;
  const R* __restrict__ U = static_cast<const R*>(a.U);
  const R* __restrict__ Rr = static_cast<const R*>(a.Rr);
  auto load6 = [&](int k, R* w) {
    const long long o = a.lay.at(i, j, k);
#pragma unroll
    for (int f = 0; f < 6; ++f) w[f] = U[o + f * fs];
  };

  // window of the own column: zp (plane m-1), cu (m), nx (m+1); lz_prev: the
  // upper x3 state of cell m-1; fz_prev: F.S of face m-3/2
  R zp[6], cu[6], nx[6], lz_prev[6], fz_prev[6], acc_prev[6];
  if (own) {
    load6(kb - 1, zp);
    load6(kb, cu);
    if (kb - 1 == 0 || LIM == MFVM_LIM_NONE) {
#pragma unroll
      for (int f = 0; f < 6; ++f) lz_prev[f] = zp[f];
    } else {
      R zm[6], lo[6];
      load6(kb - 2, zm);
      mx_slopes<2, LIM, R>(a.geo, i, j, kb - 1, zm, zp, cu, lo, lz_prev);
    }
  }
  for (int m = kb; m <= ke; ++m) {
    // x3 face m-1/2: left = upper state of cell m-1, right = lower state of cell m
    R fz[6], lz_cur[6];
    if (own) {
      if (m + 1 <= n3 + 1) load6(m + 1, nx);
      R lo[6];
      if (m == n3 + 1 || LIM == MFVM_LIM_NONE) {
#pragma unroll
        for (int f = 0; f < 6; ++f) lo[f] = lz_cur[f] = cu[f];
      } else {
        mx_slopes<2, LIM, R>(a.geo, i, j, m, zp, cu, nx, lo, lz_cur);
      }
      mx_face<2, RS, R>(a.p, lz_prev, lo, met<R>(a.geo, MFVM_MET_S3, i - 1, j - 1, m - 1), fz);
    }
    R accm[6];
    if (m < ke) {
      if constexpr (ZONLY) {
        if (own) {
          const R* xy = static_cast<const R*>(a.xy);
          const long long oi = a.il.at(i, j, m);
#pragma unroll
          for (int q = 0; q < 6; ++q) accm[q] = xy[oi + q * a.il.fs];
        }
      } else {
        mx_inplane<LIM, RS, R>(a, U, sm, i0, j0, m, cu, accm);
      }
    }
    // cell m-1: its x3 difference completes the divergence
    if (own && m > kb) {
      R acc[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) acc[q] = (acc_prev[q] + fz[q]) - fz_prev[q];
      mx_finish<R>(a, Rr, i, j, m - 1, acc, zp);
    }
    if (own) {
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        fz_prev[f] = fz[f];
        lz_prev[f] = lz_cur[f];
        acc_prev[f] = accm[f];
        zp[f] = cu[f];
        cu[f] = nx[f];
      }
    }
  }
}

template <int LIM, int RS, class R>
__global__ __launch_bounds__(MX_BX * MX_BY) void maxwell_stage_kernel(MXArgs a, int kz, int remap) {
  mx_stage_body<LIM, RS, R, false>(a, kz, remap);
}

// Split stage (float64).  The fused float64 body holds the x3 window, the
// x1/x2 exchange and the hoisted metric factors at once: over 256 VGPRs,
// one wave per SIMD.  Each split pass fits two waves per SIMD with a small
// spill, which measures 1.2x faster at 128x128x256.  (The float32 fused body
// already runs two waves and is faster than the split.)  x1/x2 pass: the
// in-plane flux differences of every cell of the tile's planes, into a.xy.
template <int LIM, int RS, class R>
__global__ __launch_bounds__(MX_BX * MX_BY, 2) void maxwell_xy_kernel(MXArgs a, int kz, int remap) {
  extern __shared__ double smem_raw[];
  R* sm = reinterpret_cast<R*>(smem_raw);
  const TileId tl = tile_id(remap != 0);
  const int i0 = 1 + tl.x * MX_BX, j0 = 1 + tl.y * MX_BY;
  const int i = i0 + threadIdx.x, j = j0 + __builtin_amdgcn_readfirstlane(threadIdx.y);  // a row is one wave
  const int kb = 1 + tl.z * kz, ke = min(kb + kz, a.lay.n3 + 1);
  const bool own = i <= a.lay.n1 && j <= a.lay.n2;
  const R* __restrict__ U = static_cast<const R*>(a.U);
  R* __restrict__ xy = static_cast<R*>(a.xy);
  for (int m = kb; m < ke; ++m) {
    R cu[6], acc[6];
    if (own) {
      const long long o = a.lay.at(i, j, m);
#pragma unroll
      for (int f = 0; f < 6; ++f) cu[f] = U[o + f * a.lay.fs];
    }
    mx_inplane<LIM, RS, R>(a, U, sm, i0, j0, m, cu, acc);
    if (own) {
      const long long oi = a.il.at(i, j, m);
#pragma unroll
      for (int q = 0; q < 6; ++q) xy[oi + q * a.il.fs] = acc[q];
    }
  }
}

// x3 pass: x3 faces, sources and the stage update.
template <int LIM, int RS, class R>
__global__ __launch_bounds__(MX_BX * MX_BY, 2) void maxwell_z_kernel(MXArgs a, int kz, int remap) {
  mx_stage_body<LIM, RS, R, true>(a, kz, remap);
}

// dt = cfl / (max_d c/Lmin_d + (2 chi_max) smax)   (maxwell.py:579-598)
__global__ void maxwell_dt_kernel(double c, double l1, double l2, double l3, double chimax, double smax,
                                  double cfl, double* dt_out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double rc = nmax(nmax(c / l1, c / l2), c / l3);
    *dt_out = cfl / (rc + (2.0 * chimax) * smax);
  }
}

static MXPar make_mx(const mfvm_maxwell_params* q) {
  MXPar p;
  int e;
  p.mu = q->mu;
  p.ep = q->ep;
  p.mu_pow2 = frexp(q->mu, &e) == 0.5;
  p.ep_pow2 = frexp(q->ep, &e) == 0.5;
  p.inv_mu = p.mu_pow2 ? 1.0 / q->mu : 0.0;
  p.inv_ep = p.ep_pow2 ? 1.0 / q->ep : 0.0;
  p.c = 1.0 / sqrt(q->mu * q->ep);
  p.negsigma = -q->sigma;
  p.rho = q->rho;
  p.negchiB = -q->chi_b;
  p.negchiD = -q->chi_d;
  return p;
}

template <class R>
hipError_t maxwell_launch_res(const mfvm_grid* g, const mfvm_maxwell_params* q, const void* U, void* Rr,
                              hipStream_t s) {
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const dim3 blk(64, 4, 1);
  const dim3 grd((g->n[0] + 2 + 63) / 64, (g->n[1] + 2 + 3) / 4, g->n[2] + 2);
  const int tk = t_begin(MFVM_TK_CLOSURE, s);
  hipLaunchKernelGGL(maxwell_res_kernel<R>, grd, blk, 0, s, lay, make_geo(g), make_mx(q),
                     static_cast<const R*>(U), static_cast<R*>(Rr), static_cast<const R*>(q->rho_field));
  t_end(tk, s);
  return hipGetLastError();
}

// x3 planes per workgroup: the fused stage and the x3 pass of the split
// stage (MX_KZ: 32 measures 0.7% faster than 16 and 1.6% faster than 8 at
// 128x128x256, the march's set-up spread over more planes), the x1/x2 pass
// (MX_XY_KZ: it keeps nothing from one plane to the next, so short chunks
// only add workgroups).
constexpr int MX_KZ = 32, MX_XY_KZ = 8;

template <int LIM, int RS, class R>
static hipError_t mx_launch_t(const MXArgs& a, hipStream_t s) {
  const int kz = a.lay.n3 < MX_KZ ? a.lay.n3 : MX_KZ;
  const dim3 blk(MX_BX, MX_BY, 1);
  const dim3 grd((a.lay.n1 + MX_BX - 1) / MX_BX, (a.lay.n2 + MX_BY - 1) / MX_BY, (a.lay.n3 + kz - 1) / kz);
  const size_t lds = mx_lds_elems() * sizeof(R);
  const int remap = xcd_remap_enabled();
  if (a.xy && sizeof(R) == sizeof(double)) {
    const int kx = a.lay.n3 < MX_XY_KZ ? a.lay.n3 : MX_XY_KZ;
    const dim3 gx(grd.x, grd.y, (a.lay.n3 + kx - 1) / kx);
    const int tx = t_begin(MFVM_TK_STAGE, s);
    hipLaunchKernelGGL((maxwell_xy_kernel<LIM, RS, R>), gx, blk, lds, s, a, kx, remap);
    t_end(tx, s);
    const int tz = t_begin(MFVM_TK_OTHER, s);
    hipLaunchKernelGGL((maxwell_z_kernel<LIM, RS, R>), grd, blk, 0, s, a, kz, remap);
    t_end(tz, s);
    return hipGetLastError();
  }
  const int tk = t_begin(MFVM_TK_STAGE, s);
  hipLaunchKernelGGL((maxwell_stage_kernel<LIM, RS, R>), grd, blk, lds, s, a, kz, remap);
  t_end(tk, s);
  return hipGetLastError();
}

template <class R>
hipError_t maxwell_launch_stage(const mfvm_grid* g, const mfvm_maxwell_params* q, const void* U,
                                const void* Rr, void* xy, const StageIO& io, hipStream_t s) {
  MXArgs a;
  a.xy = xy;
  a.lay = make_layout(g->n[0], g->n[1], g->n[2]);
  a.il = make_ilayout(g->n[0], g->n[1], g->n[2]);
  a.geo = make_geo(g);
  a.p = make_mx(q);
  a.U = U;
  a.Rr = Rr;
  a.io = io;
  a.sigf = q->sigma_field;
#define MX_RS(L)                                                                          \
  return q->riemann == MFVM_RS_HLL ? mx_launch_t<L, MFVM_RS_HLL, R>(a, s) : mx_launch_t<L, MFVM_RS_RUSANOV, R>(a, s);
  switch (q->limiter) {
    case MFVM_LIM_NONE: MX_RS(MFVM_LIM_NONE)
    case MFVM_LIM_MINMOD: MX_RS(MFVM_LIM_MINMOD)
    case MFVM_LIM_MC: MX_RS(MFVM_LIM_MC)
    default: MX_RS(MFVM_LIM_VANLEER)
  }
#undef MX_RS
}

hipError_t maxwell_launch_dt(const mfvm_grid* g, const mfvm_maxwell_params* q, double cfl, double* dt_out,
                             hipStream_t s) {
  const double c = 1.0 / sqrt(q->mu * q->ep);
  const double chimax = q->chi_b > q->chi_d ? q->chi_b : q->chi_d;  // Python max(chi_B, chi_D)
  hipLaunchKernelGGL(maxwell_dt_kernel, dim3(1), dim3(64), 0, s, c, g->lmin[0], g->lmin[1], g->lmin[2], chimax,
                     g->smax, cfl, dt_out);
  return hipGetLastError();
}

template hipError_t maxwell_launch_res<double>(const mfvm_grid*, const mfvm_maxwell_params*, const void*, void*, hipStream_t);
template hipError_t maxwell_launch_res<float>(const mfvm_grid*, const mfvm_maxwell_params*, const void*, void*, hipStream_t);
template hipError_t maxwell_launch_stage<double>(const mfvm_grid*, const mfvm_maxwell_params*, const void*, const void*, void*, const StageIO&, hipStream_t);
template hipError_t maxwell_launch_stage<float>(const mfvm_grid*, const mfvm_maxwell_params*, const void*, const void*, void*, const StageIO&, hipStream_t);

}  // namespace mfvm
