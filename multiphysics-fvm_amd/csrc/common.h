// Shared device helpers for the finite-volume kernels (gfx950).
//
// Floating-point contract: every kernel is compiled with -ffp-contract=off and
// without fast-math, so each +, -, *, / and sqrt is one IEEE-754 operation
// rounded to nearest, evaluated in the operand order written here.  The FP64
// path therefore reproduces NumPy's float64 results of the reference
// expressions bit for bit.  The helpers below restate NumPy's ufunc semantics
// (sign, minimum, maximum) including NaN propagation and tie selection.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Limited slope, reference slope_limiters.py.  LIM: 1 minmod, 2 MC, 3 van Leer.
template <int LIM, class R> MFVM_DEV R limit(R b, R f) {
  if constexpr (LIM == MFVM_LIM_MINMOD) {
    // slope_limiters.py:30-37
    R agree = (nsign(b) + nsign(f)) / R(2);
    R least = nmin(fabs(b), fabs(f));
    return agree * least;
  } else if constexpr (LIM == MFVM_LIM_MC) {
    // slope_limiters.py:60-71 (sign sum over three candidates divided by 3)
    R c = R(0.5) * (b + f);
    R bb = R(2) * b;
    R ff = R(2) * f;
    R agree = ((nsign(c) + nsign(bb)) + nsign(ff)) / R(3);
    R least = nmin(fabs(c), nmin(fabs(bb), fabs(ff)));
    return agree * least;
  } else {
    // slope_limiters.py:94-103 (harmonic mean on denom > 0, else 0)
    R agree = (nsign(b) + nsign(f)) / R(2);
    R denom = fabs(b) + fabs(f);
    R harm = denom > R(0) ? ((R(2) * b) * f) / denom : R(0);
    return agree * harm;
  }
}

// Geometry: every metric is M[i,j,k] = (a[i] * b[j]) * c[k] over 1D factors.
struct Geo {
  const double* fac;
  int off[MFVM_NMET][3];
  int n1, n2, n3;
  int curv;             // 0 cartesian (Lame terms vanish), 1 curvilinear
  int lpow2[3];         // L_d is one power of two everywhere
  double linv[3];       // its exact reciprocal
  int vpow2;
  double vinv;
};

// i, j, k are 0-based interior (or face) indices of the metric's own array.
template <class R> MFVM_DEV R met(const Geo& g, int m, int i, int j, int k) {
  const double* f = g.fac;
  return R((f[g.off[m][0] + i] * f[g.off[m][1] + j]) * f[g.off[m][2] + k]);
}

// x / L_d at cell (i,j,k) (0-based interior): exact reciprocal multiply when
// L_d is a uniform power of two (bitwise identical), true division otherwise.
template <class R> MFVM_DEV R div_len(const Geo& g, int d, R x, R len) {
  return g.lpow2[d] ? x * R(g.linv[d]) : x / len;
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

// Event timing around launches (mfvm_timing).  t_begin returns a token that
// t_end closes on the same stream; both are no-ops while timing is off.
int t_begin(int kind, hipStream_t s);
void t_end(int token, hipStream_t s);

}  // namespace mfvm
