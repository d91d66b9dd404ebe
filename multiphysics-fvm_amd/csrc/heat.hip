// Heat equation kernels (reference src/physics/heat.py).
//
// One thread per interior cell of a 64 x 4 tile of a plane.  The Fourier
// flux phi_d = -k dT/dx_d is a centred difference of T = e/(rho cp) on
// interior cells and exactly zero on ghost cells (the reference zero-pads the
// closure, heat.py:95); only phi_d is reconstructed along d (the face normal
// picks that component).  The face flux is the centred mean of the two
// reconstructed values (diffusive_solvers.py:28), F = 0.0 + Fd
// (heat.py:446-448), and rhs = -sum(F.S)/V + q (heat.py:471).  T of the tile
// and a 3-cell ring is staged in LDS and phi_1 / phi_2 are formed once per
// position; the x3 stencil is read from memory per thread.
#include "common.h"

namespace mfvm {

struct HeatArgs {
  Layout lay;
  ILayout il;
  Geo geo;
  const void* U;
  StageIO io;
  double inv_rcp;   // unused when rcp_pow2 == 0
  int rcp_pow2;     // rho*cp is a power of two: T = e * (1/(rho cp)) exactly
  double rcp;       // rho*cp
  double negk;      // -k
  double q;         // volumetric source
  const void* qf;   // per-cell source (interior layout, R) or NULL: q everywhere
};

template <class R>
MFVM_DEV R heat_T(const HeatArgs& a, const R* U, long long o) {
  return a.rcp_pow2 ? U[o] * R(a.inv_rcp) : U[o] / R(a.rcp);
}

// phi_d at ghosted cell (i,j,k): zero outside the interior along d.
template <int D, class R>
MFVM_DEV R heat_phi(const HeatArgs& a, const R* U, int i, int j, int k) {
  const int c = D == 0 ? i : (D == 1 ? j : k);
  const int n = D == 0 ? a.lay.n1 : (D == 1 ? a.lay.n2 : a.lay.n3);
  if (c < 1 || c > n) return R(0);
  const long long st = D == 0 ? 1 : (D == 1 ? a.lay.sy : a.lay.sz);
  const long long o = a.lay.at(i, j, k);
  const R dT = heat_T<R>(a, U, o + st) - heat_T<R>(a, U, o - st);
  R g;
  if (a.geo.lpow2[D]) g = dT * R(0.5 * a.geo.linv[D]);
  else g = dT / (R(2) * met<R>(a.geo, MFVM_MET_L1 + D, i - 1, j - 1, k - 1));
  return R(a.negk) * g;  // heat.py:85-92
}

template <int D, int LIM, class R>
MFVM_DEV R heat_slope(const HeatArgs& a, const R* U, int i, int j, int k, R pm, R p, R pp) {
  const int c = D == 0 ? i : (D == 1 ? j : k);
  const int n = D == 0 ? a.lay.n1 : (D == 1 ? a.lay.n2 : a.lay.n3);
  if (LIM == MFVM_LIM_NONE || c < 1 || c > n) return R(0);
  const R len = a.geo.lpow2[D] ? R(1.0 / a.geo.linv[D]) : met<R>(a.geo, MFVM_MET_L1 + D, i - 1, j - 1, k - 1);
  const R b = div_len<R>(a.geo, D, p - pm, len);
  const R f = div_len<R>(a.geo, D, pp - p, len);
  return limit<LIM, R>(b, f) * len;
}

// Face fluxes F.S on the lower and upper D-faces of interior cell (i,j,k)
// from phi_D at the cells c-2 .. c+2 along D (ph; zero where unused).
template <int D, int LIM, class R>
MFVM_DEV void heat_faces(const HeatArgs& a, const R* U, int i, int j, int k, const R ph[5], R& fs_lo, R& fs_hi) {
  auto at = [&](int t, int& x, int& y, int& z) {
    x = i + (D == 0 ? t : 0);
    y = j + (D == 1 ? t : 0);
    z = k + (D == 2 ? t : 0);
  };
  int x, y, z;
  R s[3];
#pragma unroll
  for (int t = -1; t <= 1; ++t) {
    at(t, x, y, z);
    s[t + 1] = heat_slope<D, LIM, R>(a, U, x, y, z, ph[t + 1], ph[t + 2], ph[t + 3]);
  }
  // lower face: L from cell c-1 (upper side), R from cell c (lower side)
  const R lo_l = ph[1] + s[0] / R(2), lo_r = ph[2] - s[1] / R(2);
  const R hi_l = ph[2] + s[1] / R(2), hi_r = ph[3] - s[2] / R(2);
  const R f_lo = R(0) + (lo_l + lo_r) / R(2);
  const R f_hi = R(0) + (hi_l + hi_r) / R(2);
  const int m = MFVM_MET_S1 + D;
  if (D == 0) {
    fs_lo = f_lo * met<R>(a.geo, m, i - 1, j - 1, k - 1);
    fs_hi = f_hi * met<R>(a.geo, m, i, j - 1, k - 1);
  } else if (D == 1) {
    fs_lo = f_lo * met<R>(a.geo, m, i - 1, j - 1, k - 1);
    fs_hi = f_hi * met<R>(a.geo, m, i - 1, j, k - 1);
  } else {
    fs_lo = f_lo * met<R>(a.geo, m, i - 1, j - 1, k - 1);
    fs_hi = f_hi * met<R>(a.geo, m, i - 1, j - 1, k);
  }
}

// Whether phi_D at offset t from cell index c along D enters the faces of c.
template <int LIM>
MFVM_DEV bool heat_ph_used(int c, int t, int n) {
  const int cc = c + t;
  return !(cc < 0 || cc > n + 1 || (LIM == MFVM_LIM_NONE && (t == -2 || t == 2)));
}

// The same with phi_D computed from memory.
template <int D, int LIM, class R>
MFVM_DEV void heat_dir(const HeatArgs& a, const R* U, int i, int j, int k, R& fs_lo, R& fs_hi) {
  const int c = D == 0 ? i : (D == 1 ? j : k);
  const int n = D == 0 ? a.lay.n1 : (D == 1 ? a.lay.n2 : a.lay.n3);
  R ph[5];
#pragma unroll
  for (int t = -2; t <= 2; ++t)
    ph[t + 2] = heat_ph_used<LIM>(c, t, n)
                    ? heat_phi<D, R>(a, U, i + (D == 0 ? t : 0), j + (D == 1 ? t : 0), k + (D == 2 ? t : 0))
                    : R(0);
  heat_faces<D, LIM, R>(a, U, i, j, k, ph, fs_lo, fs_hi);
}

// One workgroup per 64 x 4 tile of a plane.  T of the tile and a 3-cell
// ring is staged in LDS, then phi_1 and phi_2 at every position the tile's
// x1 / x2 faces use, each once; the x3 faces read their column from memory.
constexpr int HT_BX = 64, HT_BY = 4, HT_H = 3;
static_assert(HT_BX == 64, "one wave per tile row");
constexpr int HT_TX = HT_BX + 2 * HT_H, HT_TY = HT_BY + 2 * HT_H;  // T tile
constexpr int HT_P1X = HT_BX + 4, HT_P2Y = HT_BY + 4;              // phi_1 / phi_2 tiles
constexpr size_t ht_lds_elems() { return HT_TX * HT_TY + HT_P1X * HT_BY + HT_BX * HT_P2Y; }

template <int LIM, class R>
__global__ __launch_bounds__(HT_BX * HT_BY) void heat_stage_kernel(HeatArgs a) {
  extern __shared__ double smem_raw[];
  R* Tt = reinterpret_cast<R*>(smem_raw);  // [HT_TY][HT_TX], origin (i0-3, j0-3)
  R* P1 = Tt + HT_TX * HT_TY;              // [HT_BY][HT_P1X], origin (i0-2, j0)
  R* P2 = P1 + HT_P1X * HT_BY;             // [HT_P2Y][HT_BX], origin (i0, j0-2)
  const int n1 = a.lay.n1, n2 = a.lay.n2;
  const int i0 = 1 + blockIdx.x * HT_BX, j0 = 1 + blockIdx.y * HT_BY;
  // A wave is one x2 row of the tile (HT_BX == 64): ty is wave-uniform, so
  // the row's metric factors are read through the scalar cache.
  const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), tid = ty * HT_BX + tx;
  const int i = i0 + tx, j = j0 + ty;
  const int k = 1 + blockIdx.z;
  const R* U = static_cast<const R*>(a.U);
  for (int e = tid; e < HT_TX * HT_TY; e += HT_BX * HT_BY) {
    const int ii = i0 - HT_H + e % HT_TX, jj = j0 - HT_H + e / HT_TX;
    Tt[e] = (ii >= 0 && ii <= n1 + 1 && jj >= 0 && jj <= n2 + 1) ? heat_T<R>(a, U, a.lay.at(ii, jj, k)) : R(0);
  }
  __syncthreads();
  auto T = [&](int ii, int jj) { return Tt[(jj - (j0 - HT_H)) * HT_TX + (ii - (i0 - HT_H))]; };
  // phi_D (heat.py:85-92): zero outside the interior along D
  for (int e = tid; e < HT_P1X * HT_BY + HT_BX * HT_P2Y; e += HT_BX * HT_BY) {
    if (e < HT_P1X * HT_BY) {
      const int ii = i0 - 2 + e % HT_P1X, jj = j0 + e / HT_P1X;
      R v = R(0);
      if (ii >= 1 && ii <= n1 && jj <= n2) {
        const R dT = T(ii + 1, jj) - T(ii - 1, jj);
        const R g = a.geo.lpow2[0] ? dT * R(0.5 * a.geo.linv[0])
                                   : dT / (R(2) * met<R>(a.geo, MFVM_MET_L1, ii - 1, jj - 1, k - 1));
        v = R(a.negk) * g;
      }
      P1[e] = v;
    } else {
      const int f = e - HT_P1X * HT_BY;
      const int ii = i0 + f % HT_BX, jj = j0 - 2 + f / HT_BX;
      R v = R(0);
      if (jj >= 1 && jj <= n2 && ii <= n1) {
        const R dT = T(ii, jj + 1) - T(ii, jj - 1);
        const R g = a.geo.lpow2[1] ? dT * R(0.5 * a.geo.linv[1])
                                   : dT / (R(2) * met<R>(a.geo, MFVM_MET_L2, ii - 1, jj - 1, k - 1));
        v = R(a.negk) * g;
      }
      P2[f] = v;
    }
  }
  __syncthreads();
  if (i > n1 || j > n2) return;
  R ph[5];
  R l1, h1, l2, h2, l3, h3;
#pragma unroll
  for (int t = -2; t <= 2; ++t) ph[t + 2] = heat_ph_used<LIM>(i, t, n1) ? P1[ty * HT_P1X + tx + 2 + t] : R(0);
  heat_faces<0, LIM, R>(a, U, i, j, k, ph, l1, h1);
#pragma unroll
  for (int t = -2; t <= 2; ++t) ph[t + 2] = heat_ph_used<LIM>(j, t, n2) ? P2[(ty + 2 + t) * HT_BX + tx] : R(0);
  heat_faces<1, LIM, R>(a, U, i, j, k, ph, l2, h2);
  heat_dir<2, LIM, R>(a, U, i, j, k, l3, h3);
  const R x = -(((((h1 - l1) + h2) - l2) + h3) - l3);
  R kr = a.geo.vpow2 ? x * R(a.geo.vinv) : x / met<R>(a.geo, MFVM_MET_V, i - 1, j - 1, k - 1);
  kr = kr + (a.qf ? static_cast<const R*>(a.qf)[a.il.at(i, j, k)] : R(a.q));
  const long long oc = a.lay.at(i, j, k);
  const R u = U[oc];
  stage_store<R>(a.io, 1, a.lay.fs, a.il.fs, oc, a.il.at(i, j, k), &u, &kr, nullptr);
}

// dt = cfl / (0.0 + (2*nu)*smax), nu = k/(rho cp)   (heat.py:338-349)
__global__ void heat_dt_kernel(double nu, double smax, double cfl, double* dt_out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dt_out = cfl / (0.0 + (2.0 * nu) * smax);
}

// dts[1..n] = dts[0]: the heat dt depends on the parameters and the grid
// only, so every step's value is the one computed for the first.
__global__ void heat_dt_fill_kernel(double* dts, int n) {
  const double v = dts[0];
  for (int s = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x); s <= n; s += (int)(gridDim.x * blockDim.x))
    dts[s] = v;
}

template <int LIM, class R>
static hipError_t heat_launch_t(const HeatArgs& a, hipStream_t s) {
  const dim3 blk(HT_BX, HT_BY, 1);
  const dim3 grd((a.lay.n1 + HT_BX - 1) / HT_BX, (a.lay.n2 + HT_BY - 1) / HT_BY, a.lay.n3);
  const int tk = t_begin(MFVM_TK_STAGE, s);
  hipLaunchKernelGGL((heat_stage_kernel<LIM, R>), grd, blk, ht_lds_elems() * sizeof(R), s, a);
  t_end(tk, s);
  return hipGetLastError();
}

template <class R>
hipError_t heat_launch_stage(const mfvm_grid* g, const mfvm_heat_params* p, const void* U,
                             const StageIO& io, hipStream_t s) {
  HeatArgs a;
  a.lay = make_layout(g->n[0], g->n[1], g->n[2]);
  a.il = make_ilayout(g->n[0], g->n[1], g->n[2]);
  a.geo = make_geo(g);
  a.U = U;
  a.io = io;
  a.rcp = p->rho_cp;
  int e = 0;
  const double m = frexp(p->rho_cp, &e);
  a.rcp_pow2 = (m == 0.5);
  a.inv_rcp = a.rcp_pow2 ? 1.0 / p->rho_cp : 0.0;
  a.negk = -p->conductivity;
  a.q = p->source;
  a.qf = p->source_field;
  switch (p->limiter) {
    case MFVM_LIM_NONE: return heat_launch_t<MFVM_LIM_NONE, R>(a, s);
    case MFVM_LIM_MINMOD: return heat_launch_t<MFVM_LIM_MINMOD, R>(a, s);
    case MFVM_LIM_MC: return heat_launch_t<MFVM_LIM_MC, R>(a, s);
    default: return heat_launch_t<MFVM_LIM_VANLEER, R>(a, s);
  }
}

hipError_t heat_launch_dt_fill(double* dts, int n, hipStream_t s) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(heat_dt_fill_kernel, dim3(1), dim3(256), 0, s, dts, n);
  return hipGetLastError();
}

hipError_t heat_launch_dt(const mfvm_grid* g, const mfvm_heat_params* p, double cfl, double* dt_out,
                          hipStream_t s) {
  const double nu = p->conductivity / p->rho_cp;
  hipLaunchKernelGGL(heat_dt_kernel, dim3(1), dim3(64), 0, s, nu, g->smax, cfl, dt_out);
  return hipGetLastError();
}

template hipError_t heat_launch_stage<double>(const mfvm_grid*, const mfvm_heat_params*, const void*, const StageIO&, hipStream_t);
template hipError_t heat_launch_stage<float>(const mfvm_grid*, const mfvm_heat_params*, const void*, const StageIO&, hipStream_t);

}  // namespace mfvm
