// Ghost-cell fill (src/solver.py:13-61 with src/grid/boundary_conditions.py).
//
// One launch fills the whole ghost shell of every field.  Face ghosts come
// from the interior by the side's rule (neumann: adjacent cell, periodic:
// opposite interior cell, reflecting: minus the adjacent cell).  The
// reference then averages the 12 ghost edges from two face ghosts and the 8
// corners from three edge ghosts; here every edge/corner thread recomputes
// the face ghosts it needs from the interior, which yields the same values
// bit for bit without a second pass.  Only interior cells of src are read and
// only ghost cells of dst are written, so src == dst is race free.
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

template <class R>
__global__ void bc_kernel(BCArgs a, const R* __restrict__ src, R* __restrict__ dst, long long nshell) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int var = blockIdx.y;
  if (t >= nshell) return;
  const long long N1 = a.lay.n1 + 2, N2 = a.lay.n2 + 2, N3 = a.lay.n3 + 2;
  // shell enumeration: i in {0, n1+1} (all j,k); then i interior, j in {0,n2+1}; then k ghosts
  int i, j, k;
  const long long sA = 2 * N2 * N3, sB = 2 * (long long)a.lay.n1 * N3;
  if (t < sA) {
    const long long r = t % (N2 * N3);
    i = t < N2 * N3 ? 0 : a.lay.n1 + 1;
    j = (int)(r % N2);
    k = (int)(r / N2);
  } else if (t < sA + sB) {
    const long long u = t - sA;
    const long long r = u % ((long long)a.lay.n1 * N3);
    j = u < (long long)a.lay.n1 * N3 ? 0 : a.lay.n2 + 1;
    i = 1 + (int)(r % a.lay.n1);
    k = (int)(r / a.lay.n1);
  } else {
    const long long u = t - sA - sB;
    const long long plane = (long long)a.lay.n1 * a.lay.n2;
    const long long r = u % plane;
    k = u < plane ? 0 : a.lay.n3 + 1;
    i = 1 + (int)(r % a.lay.n1);
    j = 1 + (int)(r / a.lay.n1);
  }
  const R* f = src + var * a.lay.fs;
  dst[var * a.lay.fs + a.lay.at(i, j, k)] = shell_value<R>(a, var, f, i, j, k);
}

template <class R>
__global__ void copy_interior_kernel(Layout lay, const R* __restrict__ src, R* __restrict__ dst) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z % lay.n3;
  const int var = blockIdx.z / lay.n3;
  if (i > lay.n1) return;
  const long long o = var * lay.fs + lay.at(i, j, k);
  dst[o] = src[o];
}

template <class R>
hipError_t launch_bcs(const mfvm_grid* g, int nvar, const int32_t* bc, const void* src, void* dst,
                      hipStream_t s) {
  BCArgs a;
  a.lay = make_layout(g->n[0], g->n[1], g->n[2]);
  a.nvar = nvar;
  for (int v = 0; v < nvar; ++v)
    for (int d = 0; d < 6; ++d) a.bc[v][d] = bc[v * 6 + d];
  const long long N1 = g->n[0] + 2, N2 = g->n[1] + 2, N3 = g->n[2] + 2;
  const long long nshell = N1 * N2 * N3 - (long long)g->n[0] * g->n[1] * g->n[2];
  const dim3 grd((unsigned)((nshell + 255) / 256), nvar, 1);
  const int tk = t_begin(MFVM_TK_BC, s);
  hipLaunchKernelGGL(bc_kernel<R>, grd, dim3(256), 0, s, a, static_cast<const R*>(src),
                     static_cast<R*>(dst), nshell);
  t_end(tk, s);
  return hipGetLastError();
}

template <class R>
hipError_t launch_copy_interior(const mfvm_grid* g, int nvar, const void* src, void* dst, hipStream_t s) {
  const Layout lay = make_layout(g->n[0], g->n[1], g->n[2]);
  const dim3 grd((g->n[0] + 255) / 256, g->n[1], g->n[2] * nvar);
  hipLaunchKernelGGL(copy_interior_kernel<R>, grd, dim3(256), 0, s, lay, static_cast<const R*>(src),
                     static_cast<R*>(dst));
  return hipGetLastError();
}

template hipError_t launch_bcs<double>(const mfvm_grid*, int, const int32_t*, const void*, void*, hipStream_t);
template hipError_t launch_bcs<float>(const mfvm_grid*, int, const int32_t*, const void*, void*, hipStream_t);
template hipError_t launch_copy_interior<double>(const mfvm_grid*, int, const void*, void*, hipStream_t);
template hipError_t launch_copy_interior<float>(const mfvm_grid*, int, const void*, void*, hipStream_t);

}  // namespace mfvm
