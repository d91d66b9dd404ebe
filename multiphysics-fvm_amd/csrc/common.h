// Shared device helpers for the finite-volume kernels (gfx950).
//
// Floating-point contract: every kernel is compiled with -ffp-contract=off and
// without fast-math, so in float64 each +, -, *, / and sqrt is one IEEE-754
// operation rounded to nearest, evaluated in the operand order written here.
// The FP64 path therefore reproduces NumPy's float64 results of the reference
// expressions bit for bit.  (The float32 instantiations divide and take square
// roots with the hardware approximations, see qdiv / qsqrt.)  The helpers below restate NumPy's ufunc semantics
// (sign, minimum, maximum) including NaN propagation and tie selection.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "../../include/mfvm.h"

#define MFVM_DEV __device__ __forceinline__
#define MFVM_MAX_BCVARS 6

namespace mfvm {

// np.sign: +1, -1, 0 (for +-0), NaN for NaN.
template <class R> MFVM_DEV R nsign(R x) {
  return x > R(0) ? R(1) : (x < R(0) ? R(-1) : (x == R(0) ? R(0) : x));
}

// np.minimum / np.maximum as NumPy's SIMD loops evaluate them: NaN in either
// operand propagates, and on a tie the second operand is returned.
template <class R> MFVM_DEV R nmin(R a, R b) { return (a < b || a != a) ? a : b; }
template <class R> MFVM_DEV R nmax(R a, R b) { return (a > b || a != a) ? a : b; }

// Division and square root of the stage arithmetic.  float64: IEEE (the
// bit-exact path).  float32 (the FP32 and FP32-compensated modes, which are
// held to an error bar, not to bit parity): the hardware reciprocal and
// square root (about 1 ulp), an eighth of the instructions of the IEEE
// sequences.
template <class R> MFVM_DEV R qdiv(R a, R b) { return a / b; }
template <> MFVM_DEV float qdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <class R> MFVM_DEV R qsqrt(R a) { return sqrt(a); }
template <> MFVM_DEV float qsqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }

// q[i] = n[i] / d for i < K, each the correctly rounded quotient.
//
// float64: the gfx9 division sequence, as the compiler emits it for a / b
// (v_div_scale of the denominator and of the numerator, v_rcp, four FMAs
// refining the reciprocal, the quotient and its residual, v_div_fmas,
// v_div_fixup), with the reciprocal refinement of d formed once for all K
// quotients instead of once per quotient.  The scaled denominator
// v_div_scale(d, d, n) depends on n only for a zero n or an extreme exponent
// of n or n / d; where every lane's K scaled denominators equal the one
// formed with n = d, each quotient here is the one its own sequence gives.
// A zero n passes too: v_div_fixup returns +-0 (or NaN for d = 0 or NaN)
// for it whatever the rest of the sequence computed.  Any other lane sends
// the wave through K plain divisions.
template <int K, class R>
MFVM_DEV void qdiv_by(const R (&n)[K], R d, R (&q)[K]);

// The same quotients, correctly rounded in float32 as well (IEEE float
// division there, where qdiv_by takes the hardware reciprocal).
template <int K, class R>
MFVM_DEV void ieee_div_by(const R (&n)[K], R d, R (&q)[K]) {
  if constexpr (sizeof(R) == sizeof(float)) {
#pragma unroll
    for (int i = 0; i < K; ++i) q[i] = n[i] / d;
  } else {
    qdiv_by<K, R>(n, d, q);
  }
}

template <int K, class R>
MFVM_DEV void qdiv_by(const R (&n)[K], R d, R (&q)[K]) {
  if constexpr (sizeof(R) == sizeof(float)) {
#pragma unroll
    for (int i = 0; i < K; ++i) q[i] = qdiv(n[i], d);
  } else {
    bool fl;
    const double ds = __builtin_amdgcn_div_scale(d, d, false, &fl);
    bool ok = true;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const double dsi = __builtin_amdgcn_div_scale(n[i], d, false, &fl);
      ok = ok & ((dsi == ds) | (n[i] == 0.0));
    }
    if (__builtin_amdgcn_ballot_w64(!ok) == 0) {
      const double y0 = __builtin_amdgcn_rcp(ds);
      const double e0 = __builtin_fma(-ds, y0, 1.0);
      const double y1 = __builtin_fma(y0, e0, y0);
      const double e1 = __builtin_fma(-ds, y1, 1.0);
      const double y2 = __builtin_fma(y1, e1, y1);
#pragma unroll
      for (int i = 0; i < K; ++i) {
        bool sc;
        const double ns = __builtin_amdgcn_div_scale(n[i], d, true, &sc);
        const double q0 = ns * y2;
        const double r = __builtin_fma(-ds, q0, ns);
        q[i] = __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(r, y2, q0, sc), d, n[i]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < K; ++i) q[i] = n[i] / d;
    }
  }
}

// Limited slope, reference slope_limiters.py.  LIM: 1 minmod, 2 MC, 3 van Leer.
//
// The reference multiplies a sign-agreement factor by a magnitude.  When both
// slopes are nonzero with one sign the factor is exactly +-1, so the product
// is the magnitude with the sign of b (copysign: bit-identical).  MC's 2-to-1
// split keeps the reference's +-1/3 factor (slope_limiters.py:66-71).  In
// every other case the factor is 0, +-1/2 or +-2/3 (or NaN for a NaN input)
// and the magnitude is 0 unless the factor is 0, so the reference value is
// +-0 for a finite magnitude and NaN for a NaN or infinite one (0 * inf):
// the magnitude times 0.  Each form below equals the reference for every
// input, NaN and infinities included, up to the sign of a zero.
template <int LIM, class R> MFVM_DEV R limit(R b, R f) {
  if constexpr (LIM == MFVM_LIM_MINMOD) {
    // slope_limiters.py:30-37
    const bool same = ((b > R(0)) & (f > R(0))) | ((b < R(0)) & (f < R(0)));
    const R least = nmin(fabs(b), fabs(f));
    return same ? copysign(least, b) : least * R(0);
  } else if constexpr (LIM == MFVM_LIM_MC) {
    // slope_limiters.py:60-71 (sign sum over three candidates divided by 3)
    const bool same = ((b > R(0)) & (f > R(0))) | ((b < R(0)) & (f < R(0)));
    const R c = R(0.5) * (b + f);
    const R least = nmin(fabs(c), nmin(fabs(R(2) * b), fabs(R(2) * f)));
    const bool opp = ((b > R(0)) & (f < R(0))) | ((b < R(0)) & (f > R(0)));
    const R third = copysign(R(1.0 / 3.0) * least, c);
    return same ? copysign(least, b) : (opp ? third : least * R(0));
  } else if constexpr (sizeof(R) == sizeof(float)) {
    // float32 van Leer: (2b)|f| * rcp(den) when b f > 0, else 0.  For equal
    // signs this is the reference magnitude (2b)f/den with the sign of b, as
    // the float64 branch below forms it, in fewer instructions; a NaN input
    // gives 0 here.  The hardware reciprocal of a denominator below 2^-126 is
    // inf; fields that start at exactly zero (Maxwell's B) pass through that
    // range, so the denominator is raised to 2^-100: the harmonic mean is
    // below half of it, so both it and the true value are then below 2^-101.
    const R den = fmaxf(fabsf(b) + fabsf(f), 0x1p-100f);
    const R harm = ((b + b) * fabsf(f)) * __builtin_amdgcn_rcpf(den);
    return (b * f > R(0)) ? harm : R(0);
  } else {
    // slope_limiters.py:94-103: the sign factor times the harmonic mean
    // num / den (0 where den is 0), num = (2 b) f.  Read off num: num > 0
    // means one sign and both nonzero, the factor +-1 with the sign of b, so
    // the value is copysign(num / den, b); otherwise the factor is 0 or +-1/2
    // (or NaN) times 0 or num / den, which is +-0 for a finite num and NaN for
    // a NaN or infinite one (0 * -inf where the product overflows): num * 0.
    // One comparison instead of the four that the sign test takes.
    const R den = fabs(b) + fabs(f);
    const R num = (R(2) * b) * f;
    const R harm = qdiv(num, den);
    return num > R(0) ? copysign(harm, b) : num * R(0);
  }
}

// Geometry: every metric is M[i,j,k] = (a[i] * b[j]) * c[k] over 1D factors
// stored as rows of a fixed stride; uniform Cartesian grids use constants.
struct Geo {
  const double* fac;
  int fstride;
  int curv;             // 0 cartesian (Lame terms vanish), 1 curvilinear
  int lpow2[3];         // L_d is one power of two everywhere
  int vpow2;
  double linv[3];       // exact reciprocal of L_d when lpow2
  double lval[3];       // L_d when uniform
  double vinv, vval;
  double sval[3];       // S_d when uniform
};

// i, j, k are 0-based interior (or face) indices of the metric's own array.
// The factor rows are read-only while any kernel runs, so they are read
// through the constant address space: where an index is wave-uniform the
// compiler may then take a scalar load (its value in an SGPR).
typedef const __attribute__((address_space(4))) double ConstF64;
template <class R> MFVM_DEV R met(const Geo& g, int m, int i, int j, int k) {
  ConstF64* f = (ConstF64*)(g.fac + (long long)(3 * m) * g.fstride);
  return R((f[i] * f[g.fstride + j]) * f[2 * g.fstride + k]);
}

// Cell length L_d, face area S_d and volume V; UNIF: constants of a uniform
// Cartesian grid (exactly the reference's values), else factor products.
// (i, j, k) are ghosted cell indices; face indices are 0-based face numbers.
template <bool UNIF, class R> MFVM_DEV R g_len(const Geo& g, int d, int i, int j, int k) {
  if (UNIF) return R(g.lval[d]);
  return met<R>(g, MFVM_MET_L1 + d, i - 1, j - 1, k - 1);
}
template <bool UNIF, class R> MFVM_DEV R g_area(const Geo& g, int d, int fi, int fj, int fk) {
  if (UNIF) return R(g.sval[d]);
  return met<R>(g, MFVM_MET_S1 + d, fi, fj, fk);
}
template <bool UNIF, class R> MFVM_DEV R g_vol(const Geo& g, int i, int j, int k) {
  if (UNIF) return R(g.vval);
  return met<R>(g, MFVM_MET_V, i - 1, j - 1, k - 1);
}

// x / L_d at cell (i,j,k) (0-based interior): exact reciprocal multiply when
// L_d is a uniform power of two (bitwise identical), true division otherwise.
template <class R> MFVM_DEV R div_len(const Geo& g, int d, R x, R len) {
  return g.lpow2[d] ? x * R(g.linv[d]) : qdiv(x, len);
}

// Tile coordinates of this workgroup.  The dispatcher hands workgroups to the
// 8 XCDs round-robin by linear id, and each XCD has its own L2.  With
// `remap`, the linear id is renumbered so that XCD x runs one contiguous range
// of tiles (x1 fastest, then x2, then x3): tiles whose halo rows overlap then
// run on the same XCD at the same time and share those rows through its L2.
// The map is a bijection on [0, nblocks) for any grid size.
struct TileId {
  int x, y, z;
};
MFVM_DEV TileId tile_id(bool remap) {
  TileId t;
  if (!remap) {
    t.x = blockIdx.x;
    t.y = blockIdx.y;
    t.z = blockIdx.z;
    return t;
  }
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  const unsigned w = (blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x;
  const unsigned q = nb >> 3, r = nb & 7u, x = w & 7u;
  const unsigned L = x * q + min(x, r) + (w >> 3);
  t.x = L % gx;
  t.y = (L / gx) % gy;
  t.z = L / (gx * gy);
  return t;
}

// The same renumbering for a one-dimensional grid: XCD x runs a contiguous
// range of block numbers.
MFVM_DEV unsigned xcd_block_1d(bool remap) {
  const unsigned w = blockIdx.x;
  if (!remap) return w;
  const unsigned nb = gridDim.x;
  const unsigned q = nb >> 3, r = nb & 7u, x = w & 7u;
  return x * q + min(x, r) + (w >> 3);
}

// XCD-aware tile order (tile_id) on the host side; MFVM_XCD_REMAP=0 selects
// the plain order.
inline int xcd_remap_enabled() {
  static const int on = [] {
    const char* e = getenv("MFVM_XCD_REMAP");
    return e ? atoi(e) : 1;
  }();
  return on;
}

// Ghosted SoA addressing: field-major, x1 fastest.
struct Layout {
  int n1, n2, n3;        // interior sizes
  long long sy, sz, fs;  // strides of j, k and of one field (elements)
  MFVM_DEV long long at(int i, int j, int k) const { return (long long)k * sz + (long long)j * sy + i; }
};

// Interior-only SoA addressing (rhs / accumulator buffers).
struct ILayout {
  long long sy, sz, fs;
  MFVM_DEV long long at(int i, int j, int k) const {  // 1-based ghosted indices in
    return (long long)(k - 1) * sz + (long long)(j - 1) * sy + (i - 1);
  }
};

inline Layout make_layout(int n1, int n2, int n3) {
  Layout L;
  L.n1 = n1; L.n2 = n2; L.n3 = n3;
  L.sy = n1 + 2;
  L.sz = (long long)(n1 + 2) * (n2 + 2);
  L.fs = L.sz * (n3 + 2);
  return L;
}

inline ILayout make_ilayout(int n1, int n2, int n3) {
  ILayout L;
  L.sy = n1;
  L.sz = (long long)n1 * n2;
  L.fs = L.sz * n3;
  return L;
}

Geo make_geo(const mfvm_grid* g);

// Every cell length and the volume are one power of two.
inline bool pow2_geo(const Geo& g) { return g.lpow2[0] && g.lpow2[1] && g.lpow2[2] && g.vpow2; }

// In a kernel specialised for pow2_geo grids (checked on the host): the
// flags become compile-time constants once the copy is inlined, so the exact
// reciprocal multiplies apply without a test and a branch at each use.
MFVM_DEV void set_pow2(Geo& g) {
  g.lpow2[0] = g.lpow2[1] = g.lpow2[2] = 1;
  g.vpow2 = 1;
}

// Runge-Kutta stage update of one value (time_integrators.py), k = rhs.
//   EULER: U + dt*k                 BASE: U0 + (c0*dt)*k
//   SSP:   c0*U0 + c1*(U + dt*k)     ACC6: U0 + (dt/6)*(kacc + k)
template <class R>
MFVM_DEV R stage_value(int mode, R u, R u0, R k, R ka, double c0, double c1, double dt) {
  switch (mode) {
    case MFVM_UPD_EULER: return u + R(dt) * k;
    case MFVM_UPD_BASE: return u0 + R(c0 * dt) * k;
    case MFVM_UPD_SSP: return (R(c0) * u0) + (R(c1) * (u + R(dt) * k));
    case MFVM_UPD_ACC6: return u0 + R(dt / 6.0) * (ka + k);
    default: return u;
  }
}

// Common stage arguments of every physics (update and accumulator modes).
struct StageIO {
  const void* U0;
  void* out;
  void* kacc;
  int update, acc;
  double c0, c1, acc_coef, dt;
  const double* dt_dev;
};

// Apply the stage update and accumulator of `nv` fields at one cell.
template <class R>
MFVM_DEV void stage_store(const StageIO& io, int nv, long long fs, long long ifs, long long oc,
                          long long oi, const R* u, const R* k, R* newu) {
  const double dt = io.dt_dev ? *io.dt_dev : io.dt;
  R* out = static_cast<R*>(io.out);
  R* ka = static_cast<R*>(io.kacc);
  const R* U0 = static_cast<const R*>(io.U0);
  for (int q = 0; q < nv; ++q) {
    const bool need0 = io.update == MFVM_UPD_BASE || io.update == MFVM_UPD_SSP || io.update == MFVM_UPD_ACC6;
    const R u0 = need0 ? U0[oc + q * fs] : R(0);
    const R kav = io.update == MFVM_UPD_ACC6 ? ka[oi + q * ifs] : R(0);
    const R o = stage_value<R>(io.update, u[q], u0, k[q], kav, io.c0, io.c1, dt);
    if (newu) newu[q] = o;
    if (io.update != MFVM_UPD_NONE) out[oc + q * fs] = o;
    if (io.acc == MFVM_ACC_SET) ka[oi + q * ifs] = k[q];
    else if (io.acc == MFVM_ACC_ADD) ka[oi + q * ifs] = ka[oi + q * ifs] + R(io.acc_coef) * k[q];
  }
}

// Event timing around launches (mfvm_timing).  t_begin returns a token that
// t_end closes on the same stream; both are no-ops while timing is off.
int t_begin(int kind, hipStream_t s);
void t_end(int token, hipStream_t s);

}  // namespace mfvm
