// Heat equation kernels (reference src/physics/heat.py).
//
// One thread per interior column of a 64 x 4 tile.  The Fourier
// flux phi_d = -k dT/dx_d is a centred difference of T = e/(rho cp) on
// interior cells and exactly zero on ghost cells (the reference zero-pads the
// closure, heat.py:95); only phi_d is reconstructed along d (the face normal
// picks that component).  The face flux is the centred mean of the two
// reconstructed values (diffusive_solvers.py:28), F = 0.0 + Fd
// (heat.py:446-448), and rhs = -sum(F.S)/V + q (heat.py:471).  Workgroups
// march along x3 (heat_march_kernel).
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

// x3-marching stage kernel.  One workgroup per 64 x 4 tile of the (x1, x2)
// plane and a chunk of kz planes (tiles in XCD-contiguous order, so that the
// workgroups sharing halo rows share an L2); a wave is one x2 row of the
// tile.  Per plane, T of the tile and an H-cell ring is in LDS (two buffers:
// the next plane's tile is read from memory one plane ahead, while this
// plane is computed), phi_1 and phi_2 are formed once per position from it,
// and the x3 Fourier fluxes of the own column are carried from plane to
// plane in registers (one new U of the column per plane, also read one plane
// ahead).  Every metric is (a*b)*c over 1D factors: the plane-invariant a*b
// products of the positions a thread uses are formed once, before the march,
// and the x3 factors c of the chunk's planes are staged in LDS.  The phi
// values are double-buffered as well, so that the march runs one workgroup
// barrier per plane: phi of a plane is formed in the phase that computes the
// cells of the plane before it.
// HW: half-width of the phi stencil of a face pair (1 without slopes, 2 with);
// H = HW + 1: T ring.
constexpr int HT_BX = 64, HT_BY = 4, HT_KZ = 16;  // tile, most planes per workgroup
static_assert(HT_BX == 64, "one wave per tile row");
template <int LIM> struct HeatTile {
  static constexpr int HW = LIM == MFVM_LIM_NONE ? 1 : 2, H = HW + 1;
  static constexpr int TX = HT_BX + 2 * H, TY = HT_BY + 2 * H, NT = TX * TY;  // T tile
  static constexpr int SLOTS = (NT + HT_BX * HT_BY - 1) / (HT_BX * HT_BY);    // tile values per thread
  static constexpr int NTP = SLOTS * HT_BX * HT_BY;  // T buffer, padded: every slot stores
  static constexpr int P1X = HT_BX + 2 * HW, P2Y = HT_BY + 2 * HW;           // phi_1 / phi_2 tiles
  // x3 factors c[k] of the seven metrics the march uses, for the planes of
  // one chunk (at most HT_KZ) and the HW + 1 (below) / HW - 1 (above) more
  // that phi_3 and the x3 faces reach
  static constexpr int NCP = HT_KZ + 2 * HW, NCT = 7 * NCP;
  // one phi buffer: P1 and HT_BX spare elements (the lanes of the wave that
  // forms the phi_1 edge positions beyond those positions write there), then
  // P2
  static constexpr int PB = P1X * HT_BY + HT_BX + HT_BX * P2Y;
  // LDS bytes: the factor table (double), then, in R, the two T buffers and
  // the two phi buffers
  static constexpr size_t lds_bytes(size_t rs) { return NCT * sizeof(double) + (2 * NTP + 2 * PB) * rs; }
};

// The three factor rows of metric m: a (x1), b (x2), c (x3).
struct FacRows {
  const double *a, *b, *c;
};
MFVM_DEV FacRows fac_rows(const Geo& g, int m) {
  const double* f = g.fac + (long long)(3 * m) * g.fstride;
  return FacRows{f, f + g.fstride, f + 2 * g.fstride};
}

// One phi_D position of the tile: LDS offsets of its T centre and of its
// value, whether it lies in the interior along D (else phi_D = 0), and the
// a*b product of L_D there.
struct PhiTask {
  int t, dst;
  bool in;
  double ab;
};

// Reconstructed face pair of one cell along D from phi_D at offsets -HW..HW
// (ph[HW + t]) and, with slopes, the cell lengths at offsets -1..1 (heat.py:
// 95-110 via reconstructors.py, diffusive_solvers.py:28, heat.py:446-448):
// F = 0.0 + (L + R) / 2 on the lower (f_lo) and upper (f_hi) face.  Without
// slopes the states are the phi values themselves (adding a zero slope
// changes nothing after the 0.0 + of the flux).
template <int D, int LIM, class R>
MFVM_DEV void heat_face_pair(const Geo& g, int c, int n, const R* ph, const R* len, R& f_lo, R& f_hi) {
  if constexpr (LIM == MFVM_LIM_NONE) {
    f_lo = R(0) + (ph[0] + ph[1]) / R(2);
    f_hi = R(0) + (ph[1] + ph[2]) / R(2);
  } else {
    R s[3];
#pragma unroll
    for (int t = -1; t <= 1; ++t) {
      const int cc = c + t;
      if (cc < 1 || cc > n) {
        s[t + 1] = R(0);
      } else {
        const R b = div_len<R>(g, D, ph[t + 2] - ph[t + 1], len[t + 1]);
        const R f = div_len<R>(g, D, ph[t + 3] - ph[t + 2], len[t + 1]);
        s[t + 1] = limit<LIM, R>(b, f) * len[t + 1];
      }
    }
    const R lo_l = ph[1] + s[0] / R(2), lo_r = ph[2] - s[1] / R(2);
    const R hi_l = ph[2] + s[1] / R(2), hi_r = ph[3] - s[2] / R(2);
    f_lo = R(0) + (lo_l + lo_r) / R(2);
    f_hi = R(0) + (hi_l + hi_r) / R(2);
  }
}

// FAST (host-checked): rho*cp is a power of two, an Euler update without an
// accumulator, a uniform source -- the runtime tests of those leave the loop.
template <int LIM, class R, bool FAST>
__global__ __launch_bounds__(HT_BX * HT_BY) void heat_march_kernel(HeatArgs a, int kz, int remap) {
  using TL = HeatTile<LIM>;
  constexpr int HW = TL::HW, H = TL::H, TX = TL::TX, NT = TL::NT, NTP = TL::NTP, SLOTS = TL::SLOTS;
  constexpr int P1X = TL::P1X, NTH = HT_BX * HT_BY, NCP = TL::NCP, PB = TL::PB;
  constexpr bool SL = LIM != MFVM_LIM_NONE;
  extern __shared__ double smem_raw[];
  double* ct = smem_raw;                       // [7][NCP] x3 factors, from c-index kb-HW-1
  R* Tb = reinterpret_cast<R*>(ct + TL::NCT);  // [2][NTP]: [TY][TX], origin (i0-H, j0-H)
  // two phi buffers, each P1 then P2 (PB elements)
  R* P1 = Tb + 2 * NTP;                        // [HT_BY][P1X], origin (i0-HW, j0), then HT_BX spare
  R* P2 = P1 + P1X * HT_BY + HT_BX;            // [P2Y][HT_BX], origin (i0, j0-HW)
  const Geo& g = a.geo;
  const int n1 = a.lay.n1, n2 = a.lay.n2, n3 = a.lay.n3;
  const long long sz = a.lay.sz;
  const int sy = (int)a.lay.sy;
  const TileId tl = tile_id(remap != 0);
  const int i0 = 1 + tl.x * HT_BX, j0 = 1 + tl.y * HT_BY;
  // ty is wave-uniform (a wave is one tile row)
  const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), tid = ty * HT_BX + tx;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i <= n1 && j <= n2;
  const int ic = min(i, n1), jc = min(j, n2);  // column of this lane (an interior one)
  const int kb = 1 + tl.z * kz, ke = min(kb + kz, n3 + 1);
  const R* U = static_cast<const R*>(a.U);
  const R inv = R(a.inv_rcp), rcp = R(a.rcp), negk = R(a.negk);
  const double dt = FAST ? (a.io.dt_dev ? *a.io.dt_dev : a.io.dt) : 0.0;
  auto toT = [&](R u) { return (FAST || a.rcp_pow2) ? u * inv : u / rcp; };  // heat.py:85 T = e/(rho cp)

  // ---- plane-invariant set-up ------------------------------------------------
  // tile values of this thread: plane-relative offsets, valid inside the array
  int toff[SLOTS];
  bool tval[SLOTS];
#pragma unroll
  for (int s = 0; s < SLOTS; ++s) {
    const int e = tid + s * NTH;
    const int ii = i0 - H + e % TX, jj = j0 - H + e / TX;
    tval[s] = e < NT && ii >= 0 && ii <= n1 + 1 && jj >= 0 && jj <= n2 + 1;
    toff[s] = tval[s] ? jj * sy + ii : 0;  // (a valid element, read and not used)
  }
  // phi tasks: the own positions of phi_1 and phi_2, and one position of the
  // tile's left/right (phi_1) or lower/upper (phi_2) edge strips
  const FacRows fl1 = fac_rows(g, MFVM_MET_L1), fl2 = fac_rows(g, MFVM_MET_L2);
  auto task1 = [&](int c, int r) {  // phi_1 at P1 column c, row r
    const int ii = i0 - HW + c, jj = j0 + r;
    PhiTask q;
    q.t = (r + H) * TX + (c + 1);
    q.dst = r * P1X + c;
    q.in = ii >= 1 && ii <= n1 && jj <= n2;
    q.ab = fl1.a[min(max(ii, 1), n1) - 1] * fl1.b[min(jj, n2) - 1];
    return q;
  };
  auto task2 = [&](int c, int r) {  // phi_2 at P2 column c, row r
    const int ii = i0 + c, jj = j0 - HW + r;
    PhiTask q;
    q.t = (r + 1) * TX + (c + H);
    q.dst = r * HT_BX + c;
    q.in = jj >= 1 && jj <= n2 && ii <= n1;
    q.ab = fl2.a[min(ii, n1) - 1] * fl2.b[min(max(jj, 1), n2) - 1];
    return q;
  };
  const PhiTask p1o = task1(tx + HW, ty), p2o = task2(tx, ty + HW);
  // edge positions: phi_1 by the lanes of the last wave (lanes past them
  // write the spare elements), phi_2 by whole waves
  constexpr int NE1 = 2 * HW * HT_BY, NE2 = 2 * HW * HT_BX;
  static_assert(NE1 <= HT_BX && NE2 % HT_BX == 0, "edge positions per wave");
  const bool e1 = ty == HT_BY - 1, e2 = ty < NE2 / HT_BX;  // wave-uniform
  PhiTask p1e, p2e;
  {
    const int side = tx / (HW * HT_BY) % 2, rem = tx % (HW * HT_BY);
    p1e = task1(side ? HT_BX + HW + rem % HW : rem % HW, (rem / HW) % HT_BY);
    if (tx >= NE1) p1e.dst = P1X * HT_BY + tx;
    const int s2 = tid / (HW * HT_BX), rm2 = tid % (HW * HT_BX);
    p2e = task2(rm2 % HT_BX, s2 ? HT_BY + HW + (rm2 / HT_BX) % HW : (rm2 / HT_BX) % HW);
  }
  // own column: a*b of the face areas, the volume and L3 (and, with slopes,
  // of L1 / L2 at the x1 / x2 neighbours)
  const FacRows fs1 = fac_rows(g, MFVM_MET_S1), fs2 = fac_rows(g, MFVM_MET_S2), fs3 = fac_rows(g, MFVM_MET_S3);
  const FacRows fv = fac_rows(g, MFVM_MET_V), fl3 = fac_rows(g, MFVM_MET_L3);
  const double s1lo = fs1.a[ic - 1] * fs1.b[jc - 1], s1hi = fs1.a[ic] * fs1.b[jc - 1];
  const double s2lo = fs2.a[ic - 1] * fs2.b[jc - 1], s2hi = fs2.a[ic - 1] * fs2.b[jc];
  const double s3ab = fs3.a[ic - 1] * fs3.b[jc - 1], vab = fv.a[ic - 1] * fv.b[jc - 1];
  const double l3ab = fl3.a[ic - 1] * fl3.b[jc - 1];
  double l1ab[3], l2ab[3];
  if constexpr (SL) {
#pragma unroll
    for (int t = -1; t <= 1; ++t) {
      l1ab[t + 1] = fl1.a[min(max(ic + t, 1), n1) - 1] * fl1.b[jc - 1];
      l2ab[t + 1] = fl2.a[ic - 1] * fl2.b[min(max(jc + t, 1), n2) - 1];
    }
  }
  // x3 factors of this chunk's planes (slots: S1, S2, S3, V, L1, L2, L3),
  // indices clamped into each row (a clamped entry is never used); the c row
  // of S3 (faces) has n3 + 1 entries
  const int cb = kb - HW - 1;
  for (int e = tid; e < TL::NCT; e += NTH) {
    const int sl = e / NCP, kk = cb + e % NCP;
    const int met = sl == 0 ? MFVM_MET_S1 : sl == 1 ? MFVM_MET_S2 : sl == 2 ? MFVM_MET_S3 : sl == 3 ? MFVM_MET_V
                  : sl == 4 ? MFVM_MET_L1 : sl == 5 ? MFVM_MET_L2 : MFVM_MET_L3;
    ct[e] = fac_rows(g, met).c[min(max(kk, 0), sl == 2 ? n3 : n3 - 1)];
  }
  enum { C_S1 = 0, C_S2 = 1, C_S3 = 2, C_V = 3, C_L1 = 4, C_L2 = 5, C_L3 = 6 };
  auto cf = [&](int sl, int k) { return ct[sl * NCP + (k - cb)]; };  // c[k] of slot sl
  __syncthreads();

  R tn[SLOTS];
  auto tile_load = [&](int k) {
    const R* Up = U + (long long)k * sz;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) tn[s] = Up[toff[s]];
  };
  auto tile_store = [&](R* dst) {
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) dst[tid + s * NTH] = tval[s] ? toT(tn[s]) : R(0);
  };
  const int col = jc * sy + ic;
  // own column: U at plane k (zero outside the array; never used there)
  auto col_load = [&](int k) { return (k >= 0 && k <= n3 + 1) ? (U + (long long)k * sz)[col] : R(0); };
  // phi_3 at plane k of the own column from U at k-1 and k+1 (heat.py:85-92):
  // zero outside the interior along x3
  auto phi3 = [&](int k, R um, R up) {
    if (k < 1 || k > n3) return R(0);
    const R dT = toT(up) - toT(um);
    const R gr = g.lpow2[2] ? dT * R(0.5 * g.linv[2]) : dT / (R(2) * R(l3ab * cf(C_L3, k - 1)));
    return negk * gr;
  };
  // phi_D of one task (heat.py:85-92).  It is formed on every lane (its T
  // reads and metric are valid everywhere) and selected, so that the
  // positions take no divergent branch.
  auto phi_task = [&](const PhiTask& q, const R* T, int st, bool pw, double linv, double c, R* P) {
    const R dT = T[q.t + st] - T[q.t - st];
    const R gr = pw ? dT * R(0.5 * linv) : dT / (R(2) * R(q.ab * c));
    P[q.dst] = q.in ? negk * gr : R(0);
  };

  // ---- prologue: tile of plane kb in Tb[0], of kb+1 in flight; phi_3 at
  // kb-HW .. kb+HW-1 and the column U at the planes the next phi_3 needs
  tile_load(kb);
  R ph3[2 * HW + 1];
  R ua = col_load(kb - HW - 1), ub = col_load(kb - HW);
  R uc;
#pragma unroll
  for (int t = 0; t < 2 * HW; ++t) {
    uc = col_load(kb - HW + 1 + t);
    ph3[t + 1] = phi3(kb - HW + t, ua, uc);  // slot t+1: shifted to t at the first plane
    ua = ub;
    ub = uc;
  }
  // ua = U(kb+HW-1), ub = U(kb+HW); uc: U(kb+HW+1), read one plane ahead
  uc = col_load(kb + HW + 1);
  R u_own = col_load(kb);
  tile_store(Tb);
  if (kb + 1 < ke) tile_load(kb + 1);

  // phi_1 / phi_2 at every position this tile's faces use, of plane m from
  // the T buffer bf into the phi buffers bf
  auto phis = [&](int m, int bf) {
    const R* T = Tb + bf * NTP;
    R* Q1 = P1 + bf * PB;
    R* Q2 = P2 + bf * PB;
    const bool pw1 = g.lpow2[0] != 0, pw2 = g.lpow2[1] != 0;
    const double c1 = cf(C_L1, m - 1), c2 = cf(C_L2, m - 1);
    phi_task(p1o, T, 1, pw1, g.linv[0], c1, Q1);
    phi_task(p2o, T, TX, pw2, g.linv[1], c2, Q2);
    if (e1) phi_task(p1e, T, 1, pw1, g.linv[0], c1, Q1);
    if (e2) phi_task(p2e, T, TX, pw2, g.linv[1], c2, Q2);
  };
  // the own cell of plane m from the phi buffers bf, and its update
  auto faces = [&](int m, int bf) {
    const R* Q1 = P1 + bf * PB;
    const R* Q2 = P2 + bf * PB;
    // x3: phi_3 of plane m+HW from the column; the window shifts by one
#pragma unroll
    for (int t = 0; t < 2 * HW; ++t) ph3[t] = ph3[t + 1];
    ph3[2 * HW] = phi3(m + HW, ua, uc);
    ua = ub;
    ub = uc;
    const R u = u_own;
    if (m + 1 < ke) {
      uc = col_load(m + HW + 2);
      u_own = col_load(m + 1);
    }
    // faces (positions outside the interior along D hold phi = 0 in P1 / P2)
    R ph[2 * HW + 1], len[3];
    R l1, h1, l2, h2, l3, h3;
#pragma unroll
    for (int t = -HW; t <= HW; ++t) ph[t + HW] = Q1[ty * P1X + tx + HW + t];
    if constexpr (SL) {
#pragma unroll
      for (int t = 0; t < 3; ++t) len[t] = g.lpow2[0] ? R(1.0 / g.linv[0]) : R(l1ab[t] * cf(C_L1, m - 1));
    }
    heat_face_pair<0, LIM, R>(g, ic, n1, ph, len, l1, h1);
    l1 = l1 * R(s1lo * cf(C_S1, m - 1));
    h1 = h1 * R(s1hi * cf(C_S1, m - 1));
#pragma unroll
    for (int t = -HW; t <= HW; ++t) ph[t + HW] = Q2[(ty + HW + t) * HT_BX + tx];
    if constexpr (SL) {
#pragma unroll
      for (int t = 0; t < 3; ++t) len[t] = g.lpow2[1] ? R(1.0 / g.linv[1]) : R(l2ab[t] * cf(C_L2, m - 1));
    }
    heat_face_pair<1, LIM, R>(g, jc, n2, ph, len, l2, h2);
    l2 = l2 * R(s2lo * cf(C_S2, m - 1));
    h2 = h2 * R(s2hi * cf(C_S2, m - 1));
#pragma unroll
    for (int t = 0; t <= 2 * HW; ++t) ph[t] = ph3[t];
    if constexpr (SL) {
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const int kk = min(max(m - 1 + t, 1), n3);
        len[t] = g.lpow2[2] ? R(1.0 / g.linv[2]) : R(l3ab * cf(C_L3, kk - 1));
      }
    }
    heat_face_pair<2, LIM, R>(g, m, n3, ph, len, l3, h3);
    l3 = l3 * R(s3ab * cf(C_S3, m - 1));
    h3 = h3 * R(s3ab * cf(C_S3, m));
    const R x = -(((((h1 - l1) + h2) - l2) + h3) - l3);
    R kr = g.vpow2 ? x * R(g.vinv) : x / R(vab * cf(C_V, m - 1));
    if constexpr (FAST) {
      // time_integrators.py:10-47 (Euler): U + dt * rhs
      kr = kr + R(a.q);
      if (own) static_cast<R*>(a.io.out)[a.lay.at(i, j, m)] = u + R(dt) * kr;
    } else if (own) {
      kr = kr + (a.qf ? static_cast<const R*>(a.qf)[a.il.at(i, j, m)] : R(a.q));
      stage_store<R>(a.io, 1, a.lay.fs, a.il.fs, a.lay.at(i, j, m), a.il.at(i, j, m), &u, &kr, nullptr);
    }
  };

  // One workgroup barrier per plane: phase m forms phi of plane m (T buffer
  // and phi buffers (m - kb) & 1), stores the tile of plane m+1 into the
  // other T buffer (whose plane m-1 the previous phase read) and computes the
  // cells of plane m-1 from the other phi buffers (which the previous phase
  // wrote; the phase before read them last).
  __syncthreads();  // Tb[0] (plane kb) written
  phis(kb, 0);
  if (kb + 1 < ke) {
    tile_store(Tb + NTP);
    if (kb + 2 < ke) tile_load(kb + 2);
  }
  __syncthreads();
  for (int m = kb + 1; m < ke; ++m) {
    const int cur = (m - kb) & 1;
    phis(m, cur);
    if (m + 1 < ke) {
      tile_store(Tb + (cur ^ 1) * NTP);  // plane m+1
      if (m + 2 < ke) tile_load(m + 2);
    }
    faces(m - 1, cur ^ 1);
    __syncthreads();
  }
  faces(ke - 1, (ke - 1 - kb) & 1);
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

// Planes per workgroup: HT_KZ (the march's set-up spread over more planes),
// or fewer where the grid would then have fewer than 1024 workgroups (four
// per CU); MFVM_HEAT_KZ (1 to HT_KZ) overrides.
static int heat_kz(const Layout& lay) {
  static const int env = [] {
    const char* e = getenv("MFVM_HEAT_KZ");
    return e ? atoi(e) : 0;
  }();
  if (env > 0) return env < HT_KZ ? env : HT_KZ;
  const long long tiles = (long long)((lay.n1 + HT_BX - 1) / HT_BX) * ((lay.n2 + HT_BY - 1) / HT_BY);
  long long kz = tiles * lay.n3 / 1024;
  return (int)(kz < 1 ? 1 : (kz > HT_KZ ? HT_KZ : kz));
}

template <int LIM, class R>
static hipError_t heat_launch_t(const HeatArgs& a, hipStream_t s) {
  const dim3 blk(HT_BX, HT_BY, 1);
  const int kz = heat_kz(a.lay);
  const dim3 grd((a.lay.n1 + HT_BX - 1) / HT_BX, (a.lay.n2 + HT_BY - 1) / HT_BY, (a.lay.n3 + kz - 1) / kz);
  const int tk = t_begin(MFVM_TK_STAGE, s);
  const size_t lds = HeatTile<LIM>::lds_bytes(sizeof(R));
  const bool fast = a.rcp_pow2 && !a.qf && a.io.update == MFVM_UPD_EULER && a.io.acc == MFVM_ACC_NONE;
  if (fast)
    hipLaunchKernelGGL((heat_march_kernel<LIM, R, true>), grd, blk, lds, s, a, kz, xcd_remap_enabled());
  else
    hipLaunchKernelGGL((heat_march_kernel<LIM, R, false>), grd, blk, lds, s, a, kz, xcd_remap_enabled());
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
