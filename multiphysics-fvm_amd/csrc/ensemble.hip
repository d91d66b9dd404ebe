// Batched ensemble of Navier-Stokes instances (mfvm_ens_advance).
//
// Every member advances with exactly the arithmetic of the single-grid path
// (the same device bodies: closure_march_body, ns_stage_body, ns_xyrow_strip,
// cfl_tile_body, shell_value), so each member's result equals its own mfvm_ns_advance /
// mfvm_ns_comp_advance run bit for bit.  What changes is the launch shape:
// each phase of a stage (ghost fill, closure, the fused stage or the split
// x1/x2 and x3 passes) is ONE launch over a work list holding the tiles of
// all members, so a step costs the same number of launches for 64 members as
// for one, and small members no longer leave the GPU idle.  The work list order, the stage tile of each member and
// persistent vs one-workgroup-per-tile launches are options (picked by the
// autotuner in mfvm.ensemble); the steps can be captured into a hipGraph and
// replayed.
//
// Device table (caller-owned, mfvm_ens_table_bytes):
//   EnsMem[n] | StageArgs[nstages][n] | stage work | closure work |
//   x1/x2 row-march work | dt[n] | step[n]
// dt[m] is the member's current dt; step[m] counts the dts written to
// m.dts (the finalize kernel appends one per step), so one captured step is
// valid for every replay.
#include <algorithm>
#include <cstring>
#include <vector>

#include "bc_core.h"
#include "ns_core.h"

namespace mfvm {

struct EnsMem {
  Layout lay;
  ILayout il;
  Geo geo;
  BCArgs bca;
  int bx, by, kz;           // stage tile
  int ntx, nty, ntz;        // stage tiles per axis
  int cnx, cny, cnz, ckz;   // closure tiles (64 x 4 x ckz)
  long long nshell;         // ghost cells per field
  long long bc_off;         // first ghost-fill thread (ghost cell) of this member
  long long ntiles;         // stage tiles
  void* buf[3];             // stage states: U (hi), S1 (Sa), S2 (Sb)
  void* C;
  const void* lo;           // compensated: compensation words of U
  double* partials;
  double* dts;
  double* dt_cur;
  int* step;
  double smax;
};

struct EnsWork {
  int m, t;  // member, tile of the member
};

constexpr int kEnsBlock = 256;

// Work item of list index v.  With remap, XCD x (which receives the indices
// v = x mod 8, see tile_id) gets one contiguous range of the list.
MFVM_DEV long long ens_item(long long v, long long n, bool remap) {
  if (!remap) return v;
  const unsigned long long q = (unsigned long long)n >> 3, r = (unsigned long long)n & 7ull;
  const unsigned long long x = (unsigned long long)v & 7ull;
  return (long long)(x * q + (x < r ? x : r) + ((unsigned long long)v >> 3));
}

// One thread per ghost cell of every member, for all five fields.  I32: fewer
// than 2^31 ghost cells in all (32-bit index arithmetic).  The member is
// found once per wave, by a binary search on the wave's first thread; a wave
// that straddles members takes them one after another.  The member index
// stays wave-uniform, so every read of the member table is a scalar load.
template <class R, bool I32>
__global__ __launch_bounds__(kEnsBlock) void ens_bc_kernel(const EnsMem* __restrict__ mem, int nm, int src_sel,
                                                           int dst_sel, long long total) {
  const long long t = (long long)blockIdx.x * kEnsBlock + threadIdx.x;
  const long long t0 = (long long)blockIdx.x * kEnsBlock + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
  if (t0 >= total) return;
  int mb = 0, hi = nm - 1;
  while (mb < hi) {
    const int mid = (mb + hi + 1) >> 1;
    if (mem[mid].bc_off <= t0) mb = mid;
    else hi = mid - 1;
  }
  for (;;) {
    const EnsMem& e = mem[mb];
    const long long end = mb + 1 < nm ? mem[mb + 1].bc_off : total;
    if (t >= e.bc_off && t < end) {
      int i, j, k;
      if constexpr (I32) shell_cell32(e.lay, (unsigned)(t - e.bc_off), i, j, k);
      else shell_cell(e.lay, t - e.bc_off, i, j, k);
      const R* src = static_cast<const R*>(e.buf[src_sel]);
      R* dst = static_cast<R*>(e.buf[dst_sel]);
      const long long o = e.lay.at(i, j, k);
#pragma unroll
      for (int var = 0; var < 5; ++var)
        dst[var * e.lay.fs + o] = shell_value<R>(e.bca, var, src + var * e.lay.fs, i, j, k);
    }
    if (t0 + 63 < end || mb + 1 >= nm || end >= total) break;
    ++mb;
  }
}

// P2 / SPEC (these kernels): every member's geometry is uniform powers of
// two (set_pow2); SPEC 2 also a viscous flow (spec_args).
template <bool VISC, bool UNIF, class R, bool P2 = false>
__global__ __launch_bounds__(kEnsBlock) void ens_closure_kernel(const EnsMem* __restrict__ mem, NSPar p,
                                                                const EnsWork* __restrict__ work,
                                                                long long nwork, int sel, int remap) {
  for (long long v = blockIdx.x; v < nwork; v += gridDim.x) {
    const EnsWork w = work[ens_item(v, nwork, remap != 0)];
    const EnsMem& e = mem[w.m];
    TileId tl;
    tl.x = w.t % e.cnx;
    tl.y = (w.t / e.cnx) % e.cny;
    tl.z = w.t / (e.cnx * e.cny);
    Geo geo = e.geo;
    if constexpr (P2) set_pow2(geo);
    // (float32 members: no P, their stage forms it from U)
    closure_march_body<VISC, UNIF, R, sizeof(R) == sizeof(double)>(e.lay, geo, p, static_cast<const R*>(e.buf[sel]),
                                                                   static_cast<R*>(e.C), e.ckz, tl);
    __syncthreads();  // LDS of this tile is read before the next tile writes it
  }
}

template <int LIM, int RS, bool UNIF, class R, bool COMP, int SPEC = 0>
MFVM_DEV void ens_stage_loop(const EnsMem* __restrict__ mem, const StageArgs* __restrict__ st,
                             const EnsWork* __restrict__ work, long long nwork, int remap) {
  for (long long v = blockIdx.x; v < nwork; v += gridDim.x) {
    const EnsWork w = work[ens_item(v, nwork, remap != 0)];
    const EnsMem& e = mem[w.m];
    StageArgs a = st[w.m];
    spec_args(a, SPEC);
    TileCtx tc;
    tc.bx = e.bx;
    tc.by = e.by;
    tc.x = w.t % e.ntx;
    tc.y = (w.t / e.ntx) % e.nty;
    tc.z = w.t / (e.ntx * e.nty);
    tc.slot = a.partials ? a.partials + 2 * (long long)w.t : nullptr;
    ns_stage_body<LIM, RS, UNIF, R, COMP, false>(a, tc, ptr_src<R>(a));
    __syncthreads();
  }
}

template <int LIM, int RS, bool UNIF, int SPEC>
__global__ __launch_bounds__(kEnsBlock) void ens_stage_kernel(const EnsMem* __restrict__ mem,
                                                              const StageArgs* __restrict__ st,
                                                              const EnsWork* __restrict__ work,
                                                              long long nwork, int remap) {
  ens_stage_loop<LIM, RS, UNIF, double, false, SPEC>(mem, st, work, nwork, remap);
}

template <int LIM, int RS, bool UNIF, int SPEC>
__global__ __launch_bounds__(kEnsBlock, 2) void ens_stage_comp_kernel(const EnsMem* __restrict__ mem,
                                                                      const StageArgs* __restrict__ st,
                                                                      const EnsWork* __restrict__ work,
                                                                      long long nwork, int remap) {
  ens_stage_loop<LIM, RS, UNIF, float, true, SPEC>(mem, st, work, nwork, remap);
}

// Split stage over the work lists: the x1/x2 pass of every member, then the x3
// pass over the members' tiles (as ns_xyrow_kernel / ns_z_kernel of the
// single-grid path).  The x1/x2 pass is the row march (ns_xyrow_strip) over a
// work list of wave tasks: task t of member m is strip t mod ns, row chunk
// (t / ns) mod nc (ENS_XR_ROWS rows) and plane t / (ns nc) + 1.  Float64
// reads P from the closure, float32 forms it from U.
constexpr int ENS_XR_ROWS = 16;
inline long long ens_xr_tasks(const mfvm_ens_member& m) {
  const long long ns = (m.grid.n[0] + XR_W - 1) / XR_W, nc = (m.grid.n[1] + ENS_XR_ROWS - 1) / ENS_XR_ROWS;
  return ns * nc * m.grid.n[2];
}
// BUF: every member's arrays are below 4 GiB (buf_src_ok): buffer-resource reads.
template <int LIM, int RS, bool UNIF, class R, int SPEC, bool BUF>
__global__ __launch_bounds__(XR_W * XR_WAVES, sizeof(R) == sizeof(float) ? 4 : 2) void ens_xyrow_kernel(
    const EnsMem* __restrict__ mem, const StageArgs* __restrict__ st, const EnsWork* __restrict__ work,
    long long nwork, int remap) {
  extern __shared__ double smem_raw[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / XR_W);
  const long long v = (long long)blockIdx.x * XR_WAVES + wave;
  if (v >= nwork) return;
  const EnsWork w = work[ens_item(v, nwork, remap != 0)];
  const EnsMem& e = mem[w.m];
  StageArgs a = st[w.m];
  spec_args(a, SPEC);
  const int ns = (e.lay.n1 + XR_W - 1) / XR_W, nc = (e.lay.n2 + ENS_XR_ROWS - 1) / ENS_XR_ROWS;
  const int sx = w.t % ns, c = (w.t / ns) % nc, m = 1 + w.t / (ns * nc);
  const int jb = 1 + c * ENS_XR_ROWS, je = min(jb + ENS_XR_ROWS, e.lay.n2 + 1);
  R* lds = reinterpret_cast<R*>(smem_raw) + wave * XR_LDS_ELEMS;
  if constexpr (BUF) ns_xyrow_strip<LIM, RS, UNIF, R>(a, buf_src<R>(a), lds, 1 + sx * XR_W, jb, je, m);
  else ns_xyrow_strip<LIM, RS, UNIF, R>(a, ptr_src<R>(a), lds, 1 + sx * XR_W, jb, je, m);
}

template <int LIM, int RS, bool UNIF, class R, bool COMP, int FIN = 2, int SPEC = 0>
MFVM_DEV void ens_z_loop(const EnsMem* __restrict__ mem, const StageArgs* __restrict__ st,
                         const EnsWork* __restrict__ work, long long nwork, int remap) {
  for (long long v = blockIdx.x; v < nwork; v += gridDim.x) {
    const EnsWork w = work[ens_item(v, nwork, remap != 0)];
    const EnsMem& e = mem[w.m];
    StageArgs a = st[w.m];
    spec_args(a, SPEC);
    TileCtx tc;
    tc.bx = e.bx;
    tc.by = e.by;
    tc.x = w.t % e.ntx;
    tc.y = (w.t / e.ntx) % e.nty;
    tc.z = w.t / (e.ntx * e.nty);
    tc.slot = a.partials ? a.partials + 2 * (long long)w.t : nullptr;
    ns_stage_body<LIM, RS, UNIF, R, COMP, true, PtrSrc<R>, FIN>(a, tc, ptr_src<R>(a));
    __syncthreads();
  }
}

template <int LIM, int RS, bool UNIF, int SPEC>
__global__ __launch_bounds__(kEnsBlock) void ens_z_kernel(const EnsMem* __restrict__ mem,
                                                          const StageArgs* __restrict__ st,
                                                          const EnsWork* __restrict__ work, long long nwork,
                                                          int remap) {
  ens_z_loop<LIM, RS, UNIF, double, false, 2, SPEC>(mem, st, work, nwork, remap);
}

// FIN: the final stage (float64 accumulation, CFL partials) or not, as in
// ns_z_comp_kernel.
template <int LIM, int RS, bool UNIF, bool FIN, int SPEC>
__global__ __launch_bounds__(kEnsBlock, FIN ? 2 : 3) void ens_z_comp_kernel(const EnsMem* __restrict__ mem,
                                                                            const StageArgs* __restrict__ st,
                                                                            const EnsWork* __restrict__ work,
                                                                            long long nwork, int remap) {
  ens_z_loop<LIM, RS, UNIF, float, true, FIN ? 1 : 0, SPEC>(mem, st, work, nwork, remap);
}

template <bool UNIF, class R>
__global__ __launch_bounds__(kEnsBlock) void ens_cfl_kernel(const EnsMem* __restrict__ mem, NSPar p,
                                                            const EnsWork* __restrict__ work,
                                                            long long nwork, int remap) {
  for (long long v = blockIdx.x; v < nwork; v += gridDim.x) {
    const EnsWork w = work[ens_item(v, nwork, remap != 0)];
    const EnsMem& e = mem[w.m];
    TileCtx tc;
    tc.bx = e.bx;
    tc.by = e.by;
    tc.x = w.t % e.ntx;
    tc.y = (w.t / e.ntx) % e.nty;
    tc.z = w.t / (e.ntx * e.nty);
    tc.slot = e.partials + 2 * (long long)w.t;
    cfl_tile_body<UNIF, R>(e.lay, e.geo, p, static_cast<const R*>(e.buf[0]), static_cast<const R*>(e.lo),
                           e.kz, tc);
    __syncthreads();
  }
}

// One workgroup per member: dt of the member's next step from its partials.
__global__ __launch_bounds__(kEnsBlock) void ens_finalize_kernel(const EnsMem* __restrict__ mem, double cfl) {
  const EnsMem& e = mem[blockIdx.x];
  const double dt = dt_reduce_block(e.partials, e.ntiles, cfl, e.smax);
  if (threadIdx.x == 0) {
    const int s = *e.step;
    e.dts[s] = dt;
    *e.dt_cur = dt;
    *e.step = s + 1;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct EnsSeq {
  int in, out;  // buffer selectors (0 U/hi, 1 S1/Sa, 2 S2/Sb; out -1: none)
  int update, acc, comp;
  double c0, c1, ac;
};

// Stage sequences of time_integrators.py (as in mfvm_ns_step / mfvm_ns_comp_advance).
static int ens_sequence(int integrator, bool comp, EnsSeq* s, int* last_in) {
  if (!comp) {
    switch (integrator) {
      case MFVM_INT_RK3:  // time_integrators.py:87-103
        s[0] = {0, 1, MFVM_UPD_BASE, MFVM_ACC_SET, MFVM_COMP_NONE, 0.5, 0, 0};
        s[1] = {1, 2, MFVM_UPD_BASE, MFVM_ACC_ADD, MFVM_COMP_NONE, 1.0, 0, 4.0};
        s[2] = {2, 0, MFVM_UPD_ACC6, MFVM_ACC_NONE, MFVM_COMP_NONE, 0, 0, 0};
        *last_in = 2;
        return 3;
      case MFVM_INT_SSPRK3:  // :143-159
        s[0] = {0, 1, MFVM_UPD_EULER, MFVM_ACC_NONE, MFVM_COMP_NONE, 0, 0, 0};
        s[1] = {1, 2, MFVM_UPD_SSP, MFVM_ACC_NONE, MFVM_COMP_NONE, 0.75, 0.25, 0};
        s[2] = {2, 0, MFVM_UPD_SSP, MFVM_ACC_NONE, MFVM_COMP_NONE, 1.0 / 3.0, 2.0 / 3.0, 0};
        *last_in = 2;
        return 3;
      case MFVM_INT_RK4:  // :199-221
        s[0] = {0, 1, MFVM_UPD_BASE, MFVM_ACC_SET, MFVM_COMP_NONE, 0.5, 0, 0};
        s[1] = {1, 2, MFVM_UPD_BASE, MFVM_ACC_ADD, MFVM_COMP_NONE, 0.5, 0, 2.0};
        s[2] = {2, 1, MFVM_UPD_BASE, MFVM_ACC_ADD, MFVM_COMP_NONE, 1.0, 0, 2.0};
        s[3] = {1, 0, MFVM_UPD_ACC6, MFVM_ACC_NONE, MFVM_COMP_NONE, 0, 0, 0};
        *last_in = 1;
        return 4;
      default:
        return 0;
    }
  }
  switch (integrator) {
    case MFVM_INT_RK3:
      s[0] = {0, 1, MFVM_UPD_NONE, MFVM_ACC_SET, MFVM_COMP_SET, 0.5, 0, 0};
      s[1] = {1, 2, MFVM_UPD_NONE, MFVM_ACC_ADD, MFVM_COMP_SET, 1.0, 0, 4.0};
      s[2] = {2, -1, MFVM_UPD_NONE, MFVM_ACC_NONE, MFVM_COMP_FINAL_ACC6, 0, 0, 0};
      *last_in = 2;
      return 3;
    case MFVM_INT_SSPRK3:
      s[0] = {0, 1, MFVM_UPD_NONE, MFVM_ACC_NONE, MFVM_COMP_SET, 1.0, 0, 0};
      s[1] = {1, 2, MFVM_UPD_NONE, MFVM_ACC_NONE, MFVM_COMP_SSP, 0.25, 0, 0};
      s[2] = {2, -1, MFVM_UPD_NONE, MFVM_ACC_NONE, MFVM_COMP_FINAL_SSP, 2.0 / 3.0, 0, 0};
      *last_in = 2;
      return 3;
    case MFVM_INT_RK4:
      s[0] = {0, 1, MFVM_UPD_NONE, MFVM_ACC_SET, MFVM_COMP_SET, 0.5, 0, 0};
      s[1] = {1, 2, MFVM_UPD_NONE, MFVM_ACC_ADD, MFVM_COMP_SET, 0.5, 0, 2.0};
      s[2] = {2, 1, MFVM_UPD_NONE, MFVM_ACC_ADD, MFVM_COMP_SET, 1.0, 0, 2.0};
      s[3] = {1, -1, MFVM_UPD_NONE, MFVM_ACC_NONE, MFVM_COMP_FINAL_ACC6, 0, 0, 0};
      *last_in = 1;
      return 4;
    default:
      return 0;
  }
}

int ens_nstages(int integrator) {
  EnsSeq s[4];
  int last;
  return ens_sequence(integrator, false, s, &last);
}

static void member_tiles(const mfvm_ens_member& m, int* bx, int* by, int* kz, int nt[3]) {
  const int* n = m.grid.n;
  *bx = m.tile[0] > 0 ? m.tile[0] : 64;
  *by = m.tile[1] > 0 ? m.tile[1] : 4;
  *kz = m.tile[2] > 0 ? m.tile[2] : (n[2] < 32 ? n[2] : 32);
  nt[0] = (n[0] + *bx - 1) / *bx;
  nt[1] = (n[1] + *by - 1) / *by;
  nt[2] = (n[2] + *kz - 1) / *kz;
}

long long ens_member_tiles(const mfvm_ens_member* m) {
  int bx, by, kz, nt[3];
  member_tiles(*m, &bx, &by, &kz, nt);
  return (long long)nt[0] * nt[1] * nt[2];
}

static void closure_tiles(const mfvm_ens_member& m, int nt[3], int* kz) {
  const int* n = m.grid.n;
  *kz = n[2] < 32 ? n[2] : 32;
  nt[0] = (n[0] + 63) / 64;
  nt[1] = (n[1] + 3) / 4;
  nt[2] = (n[2] + *kz - 1) / *kz;
}

struct EnsLayout {
  size_t off_mem, off_st, off_sw, off_cw, off_xw, off_dt, off_step, bytes;
  long long nsw, ncw, nxw;
};

static EnsLayout ens_layout(const mfvm_ens_member* m, int n, int nst) {
  EnsLayout L;
  L.nsw = 0;
  L.ncw = 0;
  L.nxw = 0;
  for (int i = 0; i < n; ++i) {
    L.nsw += ens_member_tiles(&m[i]);
    L.nxw += ens_xr_tasks(m[i]);
    int nt[3], kz;
    closure_tiles(m[i], nt, &kz);
    L.ncw += (long long)nt[0] * nt[1] * nt[2];
  }
  size_t o = 0;
  L.off_mem = o;
  o = align256(o + sizeof(EnsMem) * n);
  L.off_st = o;
  o = align256(o + sizeof(StageArgs) * n * nst);
  L.off_sw = o;
  o = align256(o + sizeof(EnsWork) * L.nsw);
  L.off_cw = o;
  o = align256(o + sizeof(EnsWork) * L.ncw);
  L.off_xw = o;
  o = align256(o + sizeof(EnsWork) * L.nxw);
  L.off_dt = o;
  o = align256(o + sizeof(double) * n);
  L.off_step = o;
  o = align256(o + sizeof(int) * n);
  L.bytes = o;
  return L;
}

long long ens_table_bytes(const mfvm_ens_member* m, int n, int integrator) {
  const int nst = ens_nstages(integrator);
  if (nst == 0) return -1;
  return (long long)ens_layout(m, n, nst).bytes;
}

// Work list of the members' tiles in the requested order.
static void ens_worklist(const std::vector<long long>& ntiles, const std::vector<double>& cost, int order,
                         std::vector<EnsWork>& out) {
  const int n = (int)ntiles.size();
  out.clear();
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  if (order == MFVM_ENS_ORDER_COST)
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return cost[a] != cost[b] ? cost[a] > cost[b] : ntiles[a] > ntiles[b];
    });
  if (order == MFVM_ENS_ORDER_INTERLEAVE) {
    long long mx = 0;
    for (long long t : ntiles) mx = std::max(mx, t);
    for (long long t = 0; t < mx; ++t)
      for (int i = 0; i < n; ++i)
        if (t < ntiles[i]) out.push_back({i, (int)t});
    return;
  }
  for (int i : ord)
    for (long long t = 0; t < ntiles[i]; ++t) out.push_back({i, (int)t});
}

struct EnsRun {
  int n, nst, last_in;
  bool comp, unif, visc;
  int lim, rs;
  NSPar par;
  double cfl;
  EnsMem* d_mem;
  StageArgs* d_st;
  EnsWork* d_sw;
  EnsWork* d_cw;
  EnsWork* d_xw;  // row-march wave tasks of the x1/x2 pass
  long long nsw, ncw, nxw, bc_total;
  size_t lds_stage;
  int persist_blocks;  // 0: one workgroup per item
  int remap;
  bool split;          // every member has an xy scratch: x1/x2 pass + x3 pass
  bool p2;             // every member's geometry is uniform powers of two (pow2_geo)
  bool buf;            // every member's arrays are below 4 GiB (buf_src_ok)
  int in_sel[4];
};

static unsigned grid_for(long long nitems, int persist_blocks) {
  if (persist_blocks > 0 && nitems > persist_blocks) return (unsigned)persist_blocks;
  return (unsigned)nitems;
}

template <int LIM, int RS, bool UNIF, int SPEC>
static void launch_ens_stage_t(const EnsRun& r, const StageArgs* st, bool fin, hipStream_t s) {
  const unsigned g = grid_for(r.nsw, r.persist_blocks);
  const dim3 grd(g), blk(kEnsBlock);
  if (r.split) {
    const dim3 xg((unsigned)((r.nxw + XR_WAVES - 1) / XR_WAVES)), xb(XR_W * XR_WAVES);
#define MFVM_ENS_XYROW(RT)                                                                                   \
    if (r.buf)                                                                                               \
      hipLaunchKernelGGL((ens_xyrow_kernel<LIM, RS, UNIF, RT, SPEC, true>), xg, xb,                            \
                         (size_t)XR_WAVES * XR_LDS_ELEMS * sizeof(RT), s, r.d_mem, st, r.d_xw, r.nxw, r.remap); \
    else                                                                                                     \
      hipLaunchKernelGGL((ens_xyrow_kernel<LIM, RS, UNIF, RT, SPEC, false>), xg, xb,                           \
                         (size_t)XR_WAVES * XR_LDS_ELEMS * sizeof(RT), s, r.d_mem, st, r.d_xw, r.nxw, r.remap);
    if (r.comp) {
      MFVM_ENS_XYROW(float)
      if (fin)
        hipLaunchKernelGGL((ens_z_comp_kernel<LIM, RS, UNIF, true, SPEC>), grd, blk, r.lds_stage, s, r.d_mem, st,
                           r.d_sw, r.nsw, r.remap);
      else
        hipLaunchKernelGGL((ens_z_comp_kernel<LIM, RS, UNIF, false, SPEC>), grd, blk, r.lds_stage, s, r.d_mem, st,
                           r.d_sw, r.nsw, r.remap);
    } else {
      MFVM_ENS_XYROW(double)
      hipLaunchKernelGGL((ens_z_kernel<LIM, RS, UNIF, SPEC>), grd, blk, r.lds_stage, s, r.d_mem, st, r.d_sw, r.nsw,
                         r.remap);
    }
#undef MFVM_ENS_XYROW
    return;
  }
  if (r.comp)
    hipLaunchKernelGGL((ens_stage_comp_kernel<LIM, RS, UNIF, SPEC>), grd, blk, r.lds_stage, s, r.d_mem, st, r.d_sw,
                       r.nsw, r.remap);
  else
    hipLaunchKernelGGL((ens_stage_kernel<LIM, RS, UNIF, SPEC>), grd, blk, r.lds_stage, s, r.d_mem, st, r.d_sw, r.nsw,
                       r.remap);
}

template <int LIM, int RS>
static void launch_ens_stage_u(const EnsRun& r, const StageArgs* st, bool fin, hipStream_t s) {
  if (r.unif && r.p2 && r.visc) launch_ens_stage_t<LIM, RS, true, 2>(r, st, fin, s);
  else if (r.unif && r.p2) launch_ens_stage_t<LIM, RS, true, 1>(r, st, fin, s);
  else if (r.unif) launch_ens_stage_t<LIM, RS, true, 0>(r, st, fin, s);
  else launch_ens_stage_t<LIM, RS, false, 0>(r, st, fin, s);
}

template <int LIM>
static void launch_ens_stage_l(const EnsRun& r, const StageArgs* st, bool fin, hipStream_t s) {
  if (r.rs == MFVM_RS_HLL) launch_ens_stage_u<LIM, MFVM_RS_HLL>(r, st, fin, s);
  else launch_ens_stage_u<LIM, MFVM_RS_RUSANOV>(r, st, fin, s);
}

static void launch_ens_stage(const EnsRun& r, int q, hipStream_t s) {
  const StageArgs* st = r.d_st + (size_t)q * r.n;
  const bool fin = q == r.nst - 1;  // the compensated sequences finish in the last stage
  switch (r.lim) {
    case MFVM_LIM_NONE: launch_ens_stage_l<MFVM_LIM_NONE>(r, st, fin, s); break;
    case MFVM_LIM_MINMOD: launch_ens_stage_l<MFVM_LIM_MINMOD>(r, st, fin, s); break;
    case MFVM_LIM_MC: launch_ens_stage_l<MFVM_LIM_MC>(r, st, fin, s); break;
    default: launch_ens_stage_l<MFVM_LIM_VANLEER>(r, st, fin, s); break;
  }
}

template <class R>
static void launch_ens_bc(const EnsRun& r, int src, int dst, hipStream_t s) {
  const unsigned g = (unsigned)((r.bc_total + kEnsBlock - 1) / kEnsBlock);
  if (r.bc_total < (1ll << 31))
    hipLaunchKernelGGL((ens_bc_kernel<R, true>), dim3(g), dim3(kEnsBlock), 0, s, r.d_mem, r.n, src, dst, r.bc_total);
  else
    hipLaunchKernelGGL((ens_bc_kernel<R, false>), dim3(g), dim3(kEnsBlock), 0, s, r.d_mem, r.n, src, dst, r.bc_total);
}

template <class R>
static void launch_ens_closure(const EnsRun& r, int sel, hipStream_t s) {
  const unsigned g = grid_for(r.ncw, r.persist_blocks);
  const size_t lds = 3 * (4 + 2) * (64 + 2) * sizeof(R);
  const dim3 blk(64, 4, 1);
  if (r.visc) {
    if (r.unif && r.p2)
      hipLaunchKernelGGL((ens_closure_kernel<true, true, R, true>), dim3(g), blk, lds, s, r.d_mem, r.par, r.d_cw,
                         r.ncw, sel, r.remap);
    else if (r.unif)
      hipLaunchKernelGGL((ens_closure_kernel<true, true, R>), dim3(g), blk, lds, s, r.d_mem, r.par, r.d_cw, r.ncw,
                         sel, r.remap);
    else
      hipLaunchKernelGGL((ens_closure_kernel<true, false, R>), dim3(g), blk, lds, s, r.d_mem, r.par, r.d_cw, r.ncw,
                         sel, r.remap);
  } else {
    hipLaunchKernelGGL((ens_closure_kernel<false, true, R>), dim3(g), blk, 0, s, r.d_mem, r.par, r.d_cw, r.ncw,
                       sel, r.remap);
  }
}

template <class R>
static void launch_ens_cfl(const EnsRun& r, hipStream_t s) {
  const unsigned g = grid_for(r.nsw, r.persist_blocks);
  const size_t red = 2 * sizeof(double) * (kEnsBlock / 64);
  if (r.unif)
    hipLaunchKernelGGL((ens_cfl_kernel<true, R>), dim3(g), dim3(kEnsBlock), red, s, r.d_mem, r.par, r.d_sw, r.nsw,
                       r.remap);
  else
    hipLaunchKernelGGL((ens_cfl_kernel<false, R>), dim3(g), dim3(kEnsBlock), red, s, r.d_mem, r.par, r.d_sw, r.nsw,
                       r.remap);
  hipLaunchKernelGGL(ens_finalize_kernel, dim3(r.n), dim3(kEnsBlock), 0, s, r.d_mem, r.cfl);
}

// One step of every member: per stage a ghost fill, the closure and the
// fused stage; then each member's dt of the next step.
static hipError_t ens_step(const EnsRun& r, hipStream_t s) {
  for (int q = 0; q < r.nst; ++q) {
    const int in = r.in_sel[q];
    if (r.comp) {
      launch_ens_bc<float>(r, in, in, s);
      launch_ens_closure<float>(r, in, s);
    } else {
      launch_ens_bc<double>(r, in, in, s);
      launch_ens_closure<double>(r, in, s);
    }
    launch_ens_stage(r, q, s);
  }
  hipLaunchKernelGGL(ens_finalize_kernel, dim3(r.n), dim3(kEnsBlock), 0, s, r.d_mem, r.cfl);
  return hipGetLastError();
}

// Builds the device table (synchronous upload) and runs nsteps.  Arguments
// are checked by the caller (mfvm_ens_advance).
hipError_t ens_advance(const mfvm_ens_member* m, int n, const mfvm_ns_params* p, const mfvm_ens_opts* o,
                       int integrator, bool comp, double cfl, int nsteps, void* table, hipStream_t s) {
  EnsSeq seq[4];
  int last_in = 0;
  const int nst = ens_sequence(integrator, comp, seq, &last_in);
  const EnsLayout L = ens_layout(m, n, nst);
  char* base = static_cast<char*>(table);

  EnsRun r;
  r.n = n;
  r.nst = nst;
  r.last_in = last_in;
  r.comp = comp;
  r.lim = p->limiter;
  r.rs = p->riemann;
  r.par = make_ns_par(p);
  r.visc = r.par.visc != 0;
  r.cfl = cfl;
  r.unif = true;
  r.p2 = true;
  r.buf = true;
  for (int i = 0; i < n; ++i) {
    r.unif = r.unif && m[i].grid.uniform;
    r.p2 = r.p2 && pow2_geo(make_geo(&m[i].grid));
    r.buf = r.buf && buf_src_ok(make_layout(m[i].grid.n[0], m[i].grid.n[1], m[i].grid.n[2]), comp ? 4 : 8);
  }
  r.d_mem = reinterpret_cast<EnsMem*>(base + L.off_mem);
  r.d_st = reinterpret_cast<StageArgs*>(base + L.off_st);
  r.d_sw = reinterpret_cast<EnsWork*>(base + L.off_sw);
  r.d_cw = reinterpret_cast<EnsWork*>(base + L.off_cw);
  r.d_xw = reinterpret_cast<EnsWork*>(base + L.off_xw);
  r.nsw = L.nsw;
  r.ncw = L.ncw;
  r.nxw = L.nxw;
  r.remap = o->remap ? 1 : 0;
  r.split = true;
  for (int i = 0; i < n; ++i) r.split = r.split && m[i].xy != nullptr;
  for (int q = 0; q < nst; ++q) r.in_sel[q] = seq[q].in;
  double* d_dt = reinterpret_cast<double*>(base + L.off_dt);
  int* d_step = reinterpret_cast<int*>(base + L.off_step);

  const size_t rsize = comp ? sizeof(float) : sizeof(double);
  std::vector<char> host(L.bytes, 0);
  EnsMem* hm = reinterpret_cast<EnsMem*>(host.data() + L.off_mem);
  StageArgs* hs = reinterpret_cast<StageArgs*>(host.data() + L.off_st);
  std::vector<long long> st_tiles(n), cl_tiles(n), xr_tasks(n);
  std::vector<double> st_cost(n), cl_cost(n), xr_cost(n);
  long long bc_off = 0;
  size_t lds = 0;
  for (int i = 0; i < n; ++i) {
    const mfvm_ens_member& mm = m[i];
    EnsMem e;
    e.lay = make_layout(mm.grid.n[0], mm.grid.n[1], mm.grid.n[2]);
    e.il = make_ilayout(mm.grid.n[0], mm.grid.n[1], mm.grid.n[2]);
    e.geo = make_geo(&mm.grid);
    e.bca.lay = e.lay;
    e.bca.nvar = 5;
    e.bca.code[0] = e.bca.code[1] = 0;
    for (int v = 0; v < 5; ++v)
      for (int d = 0; d < 6; ++d) set_bc_rule(e.bca, v, d, mm.bc[v][d]);
    int nt[3];
    member_tiles(mm, &e.bx, &e.by, &e.kz, nt);
    e.ntx = nt[0];
    e.nty = nt[1];
    e.ntz = nt[2];
    int ct[3];
    closure_tiles(mm, ct, &e.ckz);
    e.cnx = ct[0];
    e.cny = ct[1];
    e.cnz = ct[2];
    e.nshell = shell_count(mm.grid.n[0], mm.grid.n[1], mm.grid.n[2]);
    e.bc_off = bc_off;
    bc_off += e.nshell;
    e.ntiles = (long long)nt[0] * nt[1] * nt[2];
    e.buf[0] = mm.U;
    e.buf[1] = mm.S1;
    e.buf[2] = mm.S2;
    e.C = mm.C;
    e.lo = comp ? mm.lo : nullptr;
    e.partials = mm.partials;
    e.dts = mm.dts;
    e.dt_cur = d_dt + i;
    e.step = d_step + i;
    e.smax = mm.grid.smax;
    hm[i] = e;
    st_tiles[i] = e.ntiles;
    st_cost[i] = e.kz;
    cl_tiles[i] = (long long)ct[0] * ct[1] * ct[2];
    cl_cost[i] = e.ckz;
    xr_tasks[i] = ens_xr_tasks(mm);
    xr_cost[i] = 1.0;
    lds = std::max(lds, stage_lds_bytes(e.bx, e.by, rsize));
    for (int q = 0; q < nst; ++q) {
      StageArgs a;
      memset(&a, 0, sizeof(a));
      a.lay = e.lay;
      a.il = e.il;
      a.geo = e.geo;
      a.p = r.par;
      a.U = e.buf[seq[q].in];
      a.C = mm.C;
      a.U0 = mm.U;
      a.kacc = mm.kacc;
      a.update = seq[q].update;
      a.acc = seq[q].acc;
      a.c0 = seq[q].c0;
      a.c1 = seq[q].c1;
      a.acc_coef = seq[q].ac;
      a.dt = 0.0;
      a.dt_dev = e.dt_cur;
      a.partials = (q == nst - 1) ? mm.partials : nullptr;
      a.kz = e.kz;
      a.remap = 0;
      a.comp = seq[q].comp;
      a.xy = r.split ? mm.xy : nullptr;
      if (comp) {
        const bool fin = seq[q].comp >= MFVM_COMP_FINAL_SET;
        a.out = fin ? mm.U : nullptr;
        a.lo = mm.lo;
        a.dinc = mm.dinc;
        a.snext = fin ? nullptr : e.buf[seq[q].out];
      } else {
        a.out = e.buf[seq[q].out];
        a.lo = nullptr;
        a.dinc = nullptr;
        a.snext = nullptr;
      }
      hs[(size_t)q * n + i] = a;
    }
  }
  r.bc_total = bc_off;
  r.lds_stage = lds;
  std::vector<EnsWork> sw, cw, xw;
  ens_worklist(st_tiles, st_cost, o->order, sw);
  ens_worklist(cl_tiles, cl_cost, o->order, cw);
  ens_worklist(xr_tasks, xr_cost, o->order, xw);
  memcpy(host.data() + L.off_sw, sw.data(), sizeof(EnsWork) * sw.size());
  memcpy(host.data() + L.off_cw, cw.data(), sizeof(EnsWork) * cw.size());
  memcpy(host.data() + L.off_xw, xw.data(), sizeof(EnsWork) * xw.size());

  int dev = 0, ncu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  r.persist_blocks = o->persist > 0 ? ((o->persist * ncu + 7) / 8) * 8 : 0;

  // Upload (the host copy must outlive the transfer: synchronous).
  if ((e = hipMemcpyAsync(table, host.data(), L.bytes, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;

  if (comp) launch_ens_cfl<float>(r, s);
  else launch_ens_cfl<double>(r, s);
  if ((e = hipGetLastError()) != hipSuccess) return e;

  int done = 0;
  if (o->graph > 0 && nsteps > 0) {
    const int chunk = std::min(o->graph, nsteps);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
    hipError_t ce = hipSuccess;
    for (int c = 0; c < chunk && ce == hipSuccess; ++c) ce = ens_step(r, s);
    e = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) e = ce;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int rep = 0; e == hipSuccess && rep < nsteps / chunk; ++rep) {
      e = hipGraphLaunch(exec, s);
      done += chunk;
    }
    const hipError_t se = hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return e;
    if (se != hipSuccess) return se;
  }
  for (; done < nsteps; ++done)
    if ((e = ens_step(r, s)) != hipSuccess) return e;
  // Leave the ghosts as the reference does: the fill of the last stage input.
  if (nsteps > 0) {
    if (comp) launch_ens_bc<float>(r, last_in, 0, s);
    else launch_ens_bc<double>(r, last_in, 0, s);
  }
  return hipGetLastError();
}

}  // namespace mfvm
