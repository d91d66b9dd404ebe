// Ghost-cell rules shared by the single-grid fill (bc.hip) and the batched
// ensemble fill (ensemble.hip): src/solver.py:13-61 with
// src/grid/boundary_conditions.py.
#pragma once

#include "common.h"

namespace mfvm {

// Ghost rules packed 2 bits per (field, side): field v < 5 in code[0] at bit
// 12 v + 2 side, field 5 in code[1].  (Scalars rather than a runtime-indexed
// array keep the kernel argument out of scratch memory.)
struct BCArgs {
  Layout lay;
  int nvar;
  unsigned long long code[2];
};

inline void set_bc_rule(BCArgs& a, int var, int side, int rule) {
  const int w = var < 5 ? 0 : 1, sh = var < 5 ? 12 * var + 2 * side : 2 * side;
  a.code[w] = (a.code[w] & ~(3ull << sh)) | ((unsigned long long)(rule & 3) << sh);
}

MFVM_DEV int bc_rule(const BCArgs& a, int var, int side) {
  return var < 5 ? (int)((a.code[0] >> (12 * var + 2 * side)) & 3ull) : (int)((a.code[1] >> (2 * side)) & 3ull);
}

// Face-ghost value along axis D at ghosted (i,j,k), whose D index is 0 or
// n_D + 1 and whose other indices are interior: the adjacent interior cell
// (neumann), the opposite one (periodic) or minus the adjacent one
// (reflecting) -- boundary_conditions.py:31, 58, 85.
template <int D, class R>
MFVM_DEV R fg(const BCArgs& a, const R* f, int var, int side, int i, int j, int k) {
  const int rule = bc_rule(a, var, 2 * D + side);
  const int n = D == 0 ? a.lay.n1 : (D == 1 ? a.lay.n2 : a.lay.n3);
  const int src = rule == MFVM_BC_PERIODIC ? (side == 0 ? n : 1) : (side == 0 ? 1 : n);
  if (D == 0) i = src;
  else if (D == 1) j = src;
  else k = src;
  const R v = f[a.lay.at(i, j, k)];
  return rule == MFVM_BC_REFLECTING ? -v : v;
}

// Value of any ghost-shell cell of field f (solver.py:36-61): face ghosts by
// their side's rule; an edge is the mean of its two neighbouring face ghosts
// (that of the later axis first); a corner is the mean of its three
// neighbouring edges.  Edges and corners recompute their face ghosts from
// the interior, so no second pass is needed.
template <class R>
MFVM_DEV R shell_value(const BCArgs& a, int var, const R* f, int i, int j, int k) {
  const int n1 = a.lay.n1, n2 = a.lay.n2, n3 = a.lay.n3;
  const bool gi = i == 0 || i == n1 + 1, gj = j == 0 || j == n2 + 1, gk = k == 0 || k == n3 + 1;
  const int si = i == 0 ? 0 : 1, sj = j == 0 ? 0 : 1, sk = k == 0 ? 0 : 1;
  const int ii = i == 0 ? 1 : n1, jj = j == 0 ? 1 : n2, kk = k == 0 ? 1 : n3;  // one cell inwards
  if (gi && gj && gk) {
    const R e0 = R(0.5) * (fg<2, R>(a, f, var, sk, ii, jj, k) + fg<1, R>(a, f, var, sj, ii, j, kk));
    const R e1 = R(0.5) * (fg<2, R>(a, f, var, sk, ii, jj, k) + fg<0, R>(a, f, var, si, i, jj, kk));
    const R e2 = R(0.5) * (fg<1, R>(a, f, var, sj, ii, j, kk) + fg<0, R>(a, f, var, si, i, jj, kk));
    return ((e0 + e1) + e2) / R(3);
  }
  if (gi && gj) return R(0.5) * (fg<1, R>(a, f, var, sj, ii, j, k) + fg<0, R>(a, f, var, si, i, jj, k));
  if (gi && gk) return R(0.5) * (fg<2, R>(a, f, var, sk, ii, j, k) + fg<0, R>(a, f, var, si, i, j, kk));
  if (gj && gk) return R(0.5) * (fg<2, R>(a, f, var, sk, i, jj, k) + fg<1, R>(a, f, var, sj, i, j, kk));
  if (gi) return fg<0, R>(a, f, var, si, i, j, k);
  if (gj) return fg<1, R>(a, f, var, sj, i, j, k);
  return fg<2, R>(a, f, var, sk, i, j, k);
}

// Ghosted (i, j, k) of shell cell t, t in [0, nshell): first the two x1
// ghost layers (all j, k), then the x2 ghost layers (interior i, all k), then
// the x3 ghost layers (interior i, j).
MFVM_DEV void shell_cell(const Layout& lay, long long t, int& i, int& j, int& k) {
  const long long N2 = lay.n2 + 2, N3 = lay.n3 + 2;
  const long long sA = 2 * N2 * N3, sB = 2 * (long long)lay.n1 * N3;
  if (t < sA) {
    const long long r = t % (N2 * N3);
    i = t < N2 * N3 ? 0 : lay.n1 + 1;
    j = (int)(r % N2);
    k = (int)(r / N2);
  } else if (t < sA + sB) {
    const long long u = t - sA;
    const long long r = u % ((long long)lay.n1 * N3);
    j = u < (long long)lay.n1 * N3 ? 0 : lay.n2 + 1;
    i = 1 + (int)(r % lay.n1);
    k = (int)(r / lay.n1);
  } else {
    const long long u = t - sA - sB;
    const long long plane = (long long)lay.n1 * lay.n2;
    const long long r = u % plane;
    k = u < plane ? 0 : lay.n3 + 1;
    i = 1 + (int)(r % lay.n1);
    j = 1 + (int)(r / lay.n1);
  }
}

// The same cell in 32-bit arithmetic, for shells below 2^31 cells (one
// unsigned division per coordinate pair instead of several 64-bit ones).
MFVM_DEV void shell_cell32(const Layout& lay, unsigned t, int& i, int& j, int& k) {
  const unsigned n1 = lay.n1, n2 = lay.n2, n3 = lay.n3;
  const unsigned N2 = n2 + 2, N3 = n3 + 2;
  const unsigned pA = N2 * N3, pB = n1 * N3;
  if (t < 2 * pA) {
    const bool hi = t >= pA;
    const unsigned r = hi ? t - pA : t, q = r / N2;
    i = hi ? (int)n1 + 1 : 0;
    j = (int)(r - q * N2);
    k = (int)q;
  } else if (t < 2 * pA + 2 * pB) {
    const unsigned u = t - 2 * pA;
    const bool hi = u >= pB;
    const unsigned r = hi ? u - pB : u, q = r / n1;
    j = hi ? (int)n2 + 1 : 0;
    i = 1 + (int)(r - q * n1);
    k = (int)q;
  } else {
    const unsigned u = t - 2 * pA - 2 * pB, plane = n1 * n2;
    const bool hi = u >= plane;
    const unsigned r = hi ? u - plane : u, q = r / n1;
    k = hi ? (int)n3 + 1 : 0;
    i = 1 + (int)(r - q * n1);
    j = 1 + (int)q;
  }
}

inline long long shell_count(int n1, int n2, int n3) {
  return (long long)(n1 + 2) * (n2 + 2) * (n3 + 2) - (long long)n1 * n2 * n3;
}

}  // namespace mfvm
