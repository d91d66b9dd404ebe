// Ghost-cell rules shared by the single-grid fill (bc.hip) and the batched
// ensemble fill (ensemble.hip): src/solver.py:13-61 with
// src/grid/boundary_conditions.py.
#pragma once

#include "common.h"

namespace mfvm {

struct BCArgs {
  Layout lay;
  int nvar;
  int bc[MFVM_MAX_BCVARS][6];
};

// Face-ghost value at ghosted (i,j,k) lying on exactly the ghost layer of
// `axis` (side 0 = low, 1 = high); the other two indices are interior.
template <class R>
MFVM_DEV R face_ghost(const BCArgs& a, const R* f, int rule, int axis, int side, int i, int j, int k) {
  const int n = axis == 0 ? a.lay.n1 : (axis == 1 ? a.lay.n2 : a.lay.n3);
  int adj = side == 0 ? 1 : n;       // adjacent interior cell
  int opp = side == 0 ? n : 1;       // opposite interior cell (periodic)
  int src = rule == MFVM_BC_PERIODIC ? opp : adj;
  int ii = i, jj = j, kk = k;
  if (axis == 0) ii = src;
  else if (axis == 1) jj = src;
  else kk = src;
  const R v = f[a.lay.at(ii, jj, kk)];
  return rule == MFVM_BC_REFLECTING ? -v : v;
}

// Value of any ghost-shell cell of field f (solver.py:36-61 semantics).
template <class R>
MFVM_DEV R shell_value(const BCArgs& a, int var, const R* f, int i, int j, int k) {
  const int n[3] = {a.lay.n1, a.lay.n2, a.lay.n3};
  const int idx[3] = {i, j, k};
  int gax[3], gside[3], ng = 0;
  for (int d = 0; d < 3; ++d) {
    if (idx[d] == 0 || idx[d] == n[d] + 1) {
      gax[ng] = d;
      gside[ng] = idx[d] == 0 ? 0 : 1;
      ++ng;
    }
  }
  auto fg = [&](int d, int side, int x, int y, int z) -> R {
    return face_ghost<R>(a, f, a.bc[var][2 * d + side], d, side, x, y, z);
  };
  // move index along axis d one cell inwards
  auto inward = [&](int d, int* p) { p[d] = (p[d] == 0) ? 1 : n[d]; };
  if (ng == 1) return fg(gax[0], gside[0], i, j, k);
  if (ng == 2) {
    // edge: 0.5*(f[moved along first ghost axis] + f[moved along second])
    const int p = gax[0], q = gax[1];
    int a1[3] = {i, j, k}, a2[3] = {i, j, k};
    inward(p, a1);  // now a face ghost of axis q
    inward(q, a2);  // now a face ghost of axis p
    return R(0.5) * (fg(q, gside[1], a1[0], a1[1], a1[2]) + fg(p, gside[0], a2[0], a2[1], a2[2]));
  }
  // corner: (e[moved along x1] + e[moved along x2] + e[moved along x3]) / 3
  R e[3];
  for (int d = 0; d < 3; ++d) {
    int c[3] = {i, j, k};
    inward(d, c);  // an edge ghost whose ghost axes are the other two
    const int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
    int a1[3] = {c[0], c[1], c[2]}, a2[3] = {c[0], c[1], c[2]};
    inward(p, a1);
    inward(q, a2);
    e[d] = R(0.5) * (fg(q, gside[q == 2 ? 2 : (q == 1 ? 1 : 0)], a1[0], a1[1], a1[2]) +
                     fg(p, gside[p], a2[0], a2[1], a2[2]));
  }
  return ((e[0] + e[1]) + e[2]) / R(3);
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

inline long long shell_count(int n1, int n2, int n3) {
  return (long long)(n1 + 2) * (n2 + 2) * (n3 + 2) - (long long)n1 * n2 * n3;
}

}  // namespace mfvm
