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
#include "bc_core.h"

namespace mfvm {

// I32: the shell has fewer than 2^31 cells (32-bit index arithmetic).  One
// thread per ghost cell and field (the ensemble's fill, over many small
// shells, takes one thread per cell for all fields instead; a single grid's
// shell measured 4% faster with the fields spread over threads).
template <class R, bool I32>
__global__ void bc_kernel(BCArgs a, const R* __restrict__ src, R* __restrict__ dst, long long nshell) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int var = blockIdx.y;
  if (t >= nshell) return;
  int i, j, k;
  if constexpr (I32) shell_cell32(a.lay, (unsigned)t, i, j, k);
  else shell_cell(a.lay, t, i, j, k);
  const R* f = src + var * a.lay.fs;
  dst[var * a.lay.fs + a.lay.at(i, j, k)] = shell_value<R>(a, var, f, i, j, k);
}

// A learned-closure stage's selection fused with the ghost fill of the next
// stage: unless *flag is set, the corrected state T replaces the interior of
// out; either way the ghost shell of out is filled from the selected
// interior (T, or out itself when the flag is set, whose interior then
// stays and the stage is counted).  The interior writes never meet the ghost
// fill's reads: those come from T while the interior is written.
template <class R>
__global__ void select_bc_kernel(BCArgs a, const R* __restrict__ T, R* out, const int* __restrict__ flag,
                                 int* __restrict__ count, unsigned nshell, unsigned nint) {
  const bool keep = *flag != 0;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  const int var = blockIdx.y;
  if (keep && t == 0 && var == 0) atomicAdd(count, 1);
  const long long fo = var * a.lay.fs;
  if (t < nshell) {
    int i, j, k;
    shell_cell32(a.lay, t, i, j, k);
    out[fo + a.lay.at(i, j, k)] = shell_value<R>(a, var, (keep ? out : T) + fo, i, j, k);
  } else if (!keep && t - nshell < nint) {
    const unsigned r = t - nshell, n1 = (unsigned)a.lay.n1, n12 = n1 * (unsigned)a.lay.n2;
    const long long o = fo + a.lay.at((int)(r % n1) + 1, (int)(r % n12 / n1) + 1, (int)(r / n12) + 1);
    out[o] = T[o];
  }
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
  a.code[0] = a.code[1] = 0;
  for (int v = 0; v < nvar; ++v)
    for (int d = 0; d < 6; ++d) set_bc_rule(a, v, d, bc[v * 6 + d]);
  const long long nshell = shell_count(g->n[0], g->n[1], g->n[2]);
  const dim3 grd((unsigned)((nshell + 255) / 256), nvar, 1);
  const int tk = t_begin(MFVM_TK_BC, s);
  if (nshell < (1ll << 31))
    hipLaunchKernelGGL((bc_kernel<R, true>), grd, dim3(256), 0, s, a, static_cast<const R*>(src),
                       static_cast<R*>(dst), nshell);
  else
    hipLaunchKernelGGL((bc_kernel<R, false>), grd, dim3(256), 0, s, a, static_cast<const R*>(src),
                       static_cast<R*>(dst), nshell);
  t_end(tk, s);
  return hipGetLastError();
}

// select_bc_kernel over nvar fields; false where the shell and the interior
// together hold 2^31 or more cells (the caller then selects and fills apart).
template <class R>
bool launch_select_bcs(const mfvm_grid* g, int nvar, const int32_t* bc, const void* T, void* out, const int* flag,
                       int* count, hipStream_t s, hipError_t* err) {
  const long long nshell = shell_count(g->n[0], g->n[1], g->n[2]);
  const long long nint = (long long)g->n[0] * g->n[1] * g->n[2];
  if (nshell + nint >= (1ll << 31)) return false;
  BCArgs a;
  a.lay = make_layout(g->n[0], g->n[1], g->n[2]);
  a.nvar = nvar;
  a.code[0] = a.code[1] = 0;
  for (int v = 0; v < nvar; ++v)
    for (int d = 0; d < 6; ++d) set_bc_rule(a, v, d, bc[v * 6 + d]);
  const dim3 grd((unsigned)((nshell + nint + 255) / 256), nvar, 1);
  const int tk = t_begin(MFVM_TK_BC, s);
  hipLaunchKernelGGL(select_bc_kernel<R>, grd, dim3(256), 0, s, a, static_cast<const R*>(T), static_cast<R*>(out),
                     flag, count, (unsigned)nshell, (unsigned)nint);
  t_end(tk, s);
  *err = hipGetLastError();
  return true;
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
template bool launch_select_bcs<double>(const mfvm_grid*, int, const int32_t*, const void*, void*, const int*, int*,
                                        hipStream_t, hipError_t*);
template hipError_t launch_copy_interior<double>(const mfvm_grid*, int, const void*, void*, hipStream_t);
template hipError_t launch_copy_interior<float>(const mfvm_grid*, int, const void*, void*, hipStream_t);

}  // namespace mfvm
