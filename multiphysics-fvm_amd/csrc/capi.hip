// C-ABI of libmfvm.so: argument checks, error reporting and stage sequencing.
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace mfvm {

// Defined in ns.hip / bc.hip / heat.hip / maxwell.hip.
template <class R>
hipError_t launch_closure(const mfvm_grid*, const mfvm_ns_params*, const void*, void*, hipStream_t, bool with_p = true);
template <class R> hipError_t launch_stage(const mfvm_grid*, const mfvm_ns_params*, const mfvm_stage*, hipStream_t);
template <class R> hipError_t launch_cfl(const mfvm_grid*, const mfvm_ns_params*, const void*, const void*, double, double*, double*, hipStream_t);
hipError_t launch_dt_finalize(const mfvm_grid*, const mfvm_ns_params*, const double*, double, double*, hipStream_t);
long long ns_nblocks(const mfvm_grid*, const mfvm_ns_params*);
template <class R> hipError_t launch_bcs(const mfvm_grid*, int, const int32_t*, const void*, void*, hipStream_t);
hipError_t launch_select_cfl_partials(const mfvm_grid* g, const mfvm_ns_params* p, const void* T, void* U,
                                      const int* flag, int* count, double* partials, hipStream_t s);
template <class R>
bool launch_select_bcs(const mfvm_grid*, int, const int32_t*, const void*, void*, const int*, int*, hipStream_t,
                       hipError_t*);
template <class R> hipError_t launch_copy_interior(const mfvm_grid*, int, const void*, void*, hipStream_t);
template <class R> hipError_t heat_launch_stage(const mfvm_grid*, const mfvm_heat_params*, const void*, const StageIO&, hipStream_t);
hipError_t heat_launch_dt(const mfvm_grid*, const mfvm_heat_params*, double, double*, hipStream_t);
hipError_t heat_launch_dt_fill(double*, int, hipStream_t);
template <class R> hipError_t maxwell_launch_res(const mfvm_grid*, const mfvm_maxwell_params*, const void*, void*, hipStream_t);
template <class R> hipError_t maxwell_launch_stage(const mfvm_grid*, const mfvm_maxwell_params*, const void*, const void*, void*, const StageIO&, hipStream_t);
hipError_t maxwell_launch_dt(const mfvm_grid*, const mfvm_maxwell_params*, double, double*, hipStream_t);
template <class R> hipError_t launch_cfl_partials(const mfvm_grid*, const mfvm_ns_params*, const void*, const void*, double*, hipStream_t);
hipError_t launch_dt_append(const mfvm_grid*, const mfvm_ns_params*, const double*, double, double, double*, double*,
                            double*, int*, int*, int, hipStream_t);
// Defined in closure.hip.
hipError_t cnn_conv(int mode, int cin, int cout, int batch, int n1, int n2, const void* x, const void* w,
                    const float* b, const void* mask, void* y, hipStream_t s);
bool cnn_shape_ok(int mode, int cin, int cout);
hipError_t closure_net_apply(const mfvm_grid* g, double gamma, const void* U, const void* out, const float* shift,
                             const float* scale, const void* const wf[3], const float* const b[3], const float* osc,
                             double coef, const double* dt, void* T, int* flag, float* yg, hipStream_t s);
hipError_t cnn_conv_features(const mfvm_grid* g, double gamma, const void* U, const float* shift, const float* scale,
                             const void* w, const float* b, void* y, hipStream_t s);
hipError_t cnn_wgrad(int cin, int cout, int batch, int n1, int n2, const void* x, const void* gy, float* gw,
                     float* gb, hipStream_t s);
hipError_t cnn_pack(int cin, int cout, const float* w, void* wf, void* wb, hipStream_t s);
hipError_t cnn_adam(long long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                    float eps, int t, hipStream_t s);
template <class R> hipError_t cnn_features(const mfvm_grid* g, double gamma, int batch, long long bs, const void* U,
                                           const float* shift, const float* scale, void* x, hipStream_t s);
hipError_t cnn_loss(int batch, int n1, int n2, float inv1, float inv2, const float* y, const float* target,
                    const float* osc, const float* wf, void* gy, float* loss, hipStream_t s);
template <class R> hipError_t closure_apply(const mfvm_grid* g, double gamma, const void* out, const float* y,
                                            const float* osc, double coef, const double* dt, void* T, int* flag,
                                            hipStream_t s);
template <class R> hipError_t closure_select(const mfvm_grid* g, const void* T, void* out, const int* flag,
                                             int* count, hipStream_t s);
hipError_t fill_int(int* p, long long n, int v, hipStream_t s);
// Defined in diag.hip.
long long diag_nblocks(const mfvm_grid* g);
hipError_t diag_launch(const mfvm_grid* g, int nvar, int dtype, const void* U, const void* lo, const void* ref,
                       double gamma, double* partials, double* out, hipStream_t s);
// Defined in ensemble.hip.
int ens_nstages(int integrator);
long long ens_member_tiles(const mfvm_ens_member* m);
long long ens_table_bytes(const mfvm_ens_member* m, int n, int integrator);
hipError_t ens_advance(const mfvm_ens_member* m, int n, const mfvm_ns_params* p, const mfvm_ens_opts* o,
                       int integrator, bool comp, double cfl, int nsteps, void* table, hipStream_t s);

static thread_local std::string g_err;

// ---- event timing ------------------------------------------------------------
struct TimedLaunch {
  int kind;
  hipEvent_t a, b;
};
static thread_local bool g_timing = false;
static thread_local std::vector<TimedLaunch> g_launches;
static thread_local std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int t_begin(int kind, hipStream_t s) {
  if (!g_timing) return -1;
  TimedLaunch t;
  t.kind = kind;
  t.a = take_event();
  t.b = take_event();
  if (!t.a || !t.b) return -1;
  (void)hipEventRecord(t.a, s);
  g_launches.push_back(t);
  return (int)g_launches.size() - 1;
}

void t_end(int token, hipStream_t s) {
  if (token < 0 || token >= (int)g_launches.size()) return;
  (void)hipEventRecord(g_launches[token].b, s);
}

static void t_recycle() {
  for (auto& t : g_launches) {
    g_event_pool.push_back(t.a);
    g_event_pool.push_back(t.b);
  }
  g_launches.clear();
}

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

static int hip_check(hipError_t e, const char* where) {
  if (e == hipSuccess) return MFVM_OK;
  return fail(MFVM_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

Geo make_geo(const mfvm_grid* g) {
  Geo o;
  o.fac = g->fac;
  o.fstride = g->fstride;
  o.curv = g->coord != MFVM_CART;
  for (int d = 0; d < 3; ++d) {
    o.lpow2[d] = g->lpow2[d];
    o.linv[d] = g->lpow2[d] ? 1.0 / g->lval[d] : 0.0;
    o.lval[d] = g->lval[d];
    o.sval[d] = g->sval[d];
  }
  o.vpow2 = g->vpow2;
  o.vinv = g->vpow2 ? 1.0 / g->vval : 0.0;
  o.vval = g->vval;
  return o;
}

static int check_grid(const mfvm_grid* g) {
  if (!g) return fail(MFVM_E_ARG, "grid descriptor is NULL");
  for (int d = 0; d < 3; ++d)
    if (g->n[d] < 1) return fail(MFVM_E_ARG, "grid sizes must be >= 1");
  if (g->coord < MFVM_CART || g->coord > MFVM_SPH)
    return fail(MFVM_E_ARG,
                "Unknown coordinate system. Expected one of: cartesian, cylindrical, spherical.");
  if (!g->fac) return fail(MFVM_E_ARG, "grid metric factors are NULL");
  const int mx = g->n[0] > g->n[1] ? (g->n[0] > g->n[2] ? g->n[0] : g->n[2]) : (g->n[1] > g->n[2] ? g->n[1] : g->n[2]);
  if (g->fstride < mx + 1) return fail(MFVM_E_ARG, "grid factor stride must be >= max(n)+1");
  if (g->uniform && g->coord != MFVM_CART) return fail(MFVM_E_ARG, "uniform grids must be Cartesian");
  return MFVM_OK;
}

static int check_bc(int nvar, const int32_t* bc) {
  if (nvar < 1 || nvar > MFVM_MAX_BCVARS) return fail(MFVM_E_ARG, "nvar out of range");
  if (!bc) return fail(MFVM_E_ARG, "bc table is NULL");
  for (int i = 0; i < nvar * 6; ++i)
    if (bc[i] < MFVM_BC_NEUMANN || bc[i] > MFVM_BC_REFLECTING)
      return fail(MFVM_E_ARG, "unknown boundary condition code");
  return MFVM_OK;
}

static int check_ns(const mfvm_grid* g, const mfvm_ns_params* p) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (!p) return fail(MFVM_E_ARG, "params descriptor is NULL");
  if (p->dtype != MFVM_F64 && p->dtype != MFVM_F32) return fail(MFVM_E_ARG, "dtype must be 0 (f64) or 1 (f32)");
  if (p->limiter < MFVM_LIM_NONE || p->limiter > MFVM_LIM_VANLEER)
    return fail(MFVM_E_UNSUP, "unsupported slope limiter / reconstructor");
  if (p->riemann != MFVM_RS_HLL && p->riemann != MFVM_RS_RUSANOV)
    return fail(MFVM_E_UNSUP, "unsupported Riemann solver");
  const int bx = p->tile[0] > 0 ? p->tile[0] : 64, by = p->tile[1] > 0 ? p->tile[1] : 4;
  if (bx < 2 || by < 1 || bx * by > 256 || (bx * by) % 64 != 0 || bx + by > bx * by)
    return fail(MFVM_E_ARG, "stage tile must have 2 <= x1 threads, x1*x2 threads a multiple of 64 and <= 256, "
                            "and x1 + x2 threads <= x1*x2");
  if (p->tile[2] < 0) return fail(MFVM_E_ARG, "stage tile x3 planes must be >= 0");
  return check_bc(5, &p->bc[0][0]);
}

}  // namespace mfvm

using namespace mfvm;

extern "C" {

const char* mfvm_last_error(void) { return g_err.c_str(); }

int32_t mfvm_abi_version(void) { return 1; }

int mfvm_timing(int32_t mode, double* ms, int64_t* count) {
  if (mode == 0 || mode == 1) {
    if (!g_launches.empty()) {
      (void)hipDeviceSynchronize();
      t_recycle();
    }
    g_timing = mode == 1;
    return MFVM_OK;
  }
  if (mode != 2 || !ms || !count) return fail(MFVM_E_ARG, "mfvm_timing: bad arguments");
  for (int k = 0; k < MFVM_TK_N; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (auto& t : g_launches) {
    if (hipEventSynchronize(t.b) != hipSuccess) return fail(MFVM_E_HIP, "mfvm_timing: event sync failed");
    float e = 0.f;
    if (hipEventElapsedTime(&e, t.a, t.b) != hipSuccess) return fail(MFVM_E_HIP, "mfvm_timing: elapsed failed");
    ms[t.kind] += e;
    count[t.kind] += 1;
  }
  t_recycle();
  return MFVM_OK;
}

int mfvm_apply_bcs(const mfvm_grid* g, int32_t nvar, const int32_t* bc, int32_t dtype, const void* src,
                   void* dst, void* stream) {
  int rc = check_grid(g);
  if (!rc) rc = check_bc(nvar, bc);
  if (rc) return rc;
  if (!src || !dst) return fail(MFVM_E_ARG, "apply_bcs: NULL buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == MFVM_F64) return hip_check(launch_bcs<double>(g, nvar, bc, src, dst, s), "apply_bcs");
  if (dtype == MFVM_F32) return hip_check(launch_bcs<float>(g, nvar, bc, src, dst, s), "apply_bcs");
  return fail(MFVM_E_ARG, "dtype must be 0 (f64) or 1 (f32)");
}

int mfvm_copy_interior(const mfvm_grid* g, int32_t nvar, int32_t dtype, const void* src, void* dst,
                       void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (nvar < 1 || !src || !dst) return fail(MFVM_E_ARG, "copy_interior: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == MFVM_F64) return hip_check(launch_copy_interior<double>(g, nvar, src, dst, s), "copy_interior");
  if (dtype == MFVM_F32) return hip_check(launch_copy_interior<float>(g, nvar, src, dst, s), "copy_interior");
  return fail(MFVM_E_ARG, "dtype must be 0 (f64) or 1 (f32)");
}

int mfvm_ns_closure(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, void* C, void* stream) {
  int rc = check_ns(g, p);
  if (rc) return rc;
  if (!U || !C) return fail(MFVM_E_ARG, "ns_closure: NULL buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(p->dtype == MFVM_F64 ? launch_closure<double>(g, p, U, C, s)
                                        : launch_closure<float>(g, p, U, C, s),
                   "ns_closure");
}

int mfvm_ns_stage(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_stage* st, void* stream) {
  int rc = check_ns(g, p);
  if (rc) return rc;
  if (!st || !st->U || !st->C) return fail(MFVM_E_ARG, "ns_stage: NULL input");
  if (st->comp != MFVM_COMP_NONE) {
    if (st->comp < MFVM_COMP_SET || st->comp > MFVM_COMP_FINAL_ACC6) return fail(MFVM_E_ARG, "ns_stage: bad comp mode");
    if (p->dtype != MFVM_F32) return fail(MFVM_E_ARG, "ns_stage: compensated stages need dtype f32");
    const bool fin = st->comp >= MFVM_COMP_FINAL_SET;
    if (!st->U0 || !st->lo || !st->dinc) return fail(MFVM_E_ARG, "ns_stage: compensated stage needs U0, lo, dinc");
    if (fin && !st->out) return fail(MFVM_E_ARG, "ns_stage: compensated final stage needs out");
    if (!fin && !st->snext) return fail(MFVM_E_ARG, "ns_stage: compensated stage needs snext");
    if (fin && st->out == st->U) return fail(MFVM_E_ARG, "ns_stage: output must not alias the stencil input");
    if (!fin && st->snext == st->U) return fail(MFVM_E_ARG, "ns_stage: snext must not alias the stencil input");
    if ((st->acc != MFVM_ACC_NONE || st->comp == MFVM_COMP_FINAL_ACC6) && !st->kacc)
      return fail(MFVM_E_ARG, "ns_stage: NULL accumulator");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return hip_check(launch_stage<float>(g, p, st, s), "ns_stage");
  }
  if (st->update < MFVM_UPD_NONE || st->update > MFVM_UPD_ACC6) return fail(MFVM_E_ARG, "ns_stage: bad update mode");
  if (st->acc < MFVM_ACC_NONE || st->acc > MFVM_ACC_ADD) return fail(MFVM_E_ARG, "ns_stage: bad accumulator mode");
  if (st->update != MFVM_UPD_NONE && !st->out) return fail(MFVM_E_ARG, "ns_stage: NULL output");
  if ((st->update == MFVM_UPD_BASE || st->update == MFVM_UPD_SSP || st->update == MFVM_UPD_ACC6) && !st->U0)
    return fail(MFVM_E_ARG, "ns_stage: NULL base state");
  if ((st->acc != MFVM_ACC_NONE || st->update == MFVM_UPD_ACC6) && !st->kacc)
    return fail(MFVM_E_ARG, "ns_stage: NULL accumulator");
  if (st->update != MFVM_UPD_NONE && st->out == st->U)
    return fail(MFVM_E_ARG, "ns_stage: output must not alias the stencil input");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(p->dtype == MFVM_F64 ? launch_stage<double>(g, p, st, s) : launch_stage<float>(g, p, st, s),
                   "ns_stage");
}

int mfvm_ns_rhs(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, void* C, void* rhs, void* stream) {
  int rc = mfvm_ns_closure(g, p, U, C, stream);
  if (rc) return rc;
  mfvm_stage st = {};
  st.U = U;
  st.C = C;
  st.kacc = rhs;
  st.update = MFVM_UPD_NONE;
  st.acc = MFVM_ACC_SET;
  return mfvm_ns_stage(g, p, &st, stream);
}

int64_t mfvm_ns_partials_size(const mfvm_grid* g, const mfvm_ns_params* p) {
  if (check_ns(g, p)) return -1;
  return 2 * ns_nblocks(g, p);
}

int mfvm_ns_dt(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, double cfl, double* partials,
               double* dt_out, void* stream) {
  int rc = check_ns(g, p);
  if (rc) return rc;
  if (!U || !partials || !dt_out) return fail(MFVM_E_ARG, "ns_dt: NULL buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(p->dtype == MFVM_F64 ? launch_cfl<double>(g, p, U, nullptr, cfl, partials, dt_out, s)
                                        : launch_cfl<float>(g, p, U, nullptr, cfl, partials, dt_out, s),
                   "ns_dt");
}

}  // extern "C"

namespace mfvm {
// ---- allocation-free CFL step --------------------------------------------------
// The CFL partials of mfvm_ns_dt_noalloc live in this file-scope device array
// (one copy per GPU, held until the process ends).  Host code never passes the
// array's name to a kernel: it takes the device address once per device with
// hipGetSymbolAddress and keeps it in ns_dt_scratch_ptr[device].
static __device__ __align__(256) double scratch[2 * MFVM_NS_DT_NOALLOC_MAX_BLOCKS];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* ns_dt_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with it current).
static int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_check(e, "ns_dt_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(MFVM_E_ARG, "ns_dt_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!ns_dt_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "ns_dt_noalloc: hipGetSymbolAddress");
    ns_dt_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = ns_dt_scratch_ptr[dev];
  return MFVM_OK;
}

}  // namespace mfvm

extern "C" {

int mfvm_ns_dt_noalloc(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, double cfl, double* dt_out,
                       void* stream) {
  int rc = check_ns(g, p);
  if (rc) return rc;
  if (!U || !dt_out) return fail(MFVM_E_ARG, "ns_dt_noalloc: NULL buffer");
  if (ns_nblocks(g, p) > MFVM_NS_DT_NOALLOC_MAX_BLOCKS)
    return fail(MFVM_E_ARG, "ns_dt_noalloc: the grid needs " + std::to_string(ns_nblocks(g, p)) +
                                " CFL partial slots, above the fixed maximum of " +
                                std::to_string(MFVM_NS_DT_NOALLOC_MAX_BLOCKS) + " (use mfvm_ns_dt)");
  double* partials = nullptr;
  if ((rc = scratch_address(&partials))) return rc;
  // Every workgroup of the partials kernel writes its own pair of slots and
  // the finalize reads only those, so nothing needs clearing between calls.
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(p->dtype == MFVM_F64 ? launch_cfl<double>(g, p, U, nullptr, cfl, partials, dt_out, s)
                                        : launch_cfl<float>(g, p, U, nullptr, cfl, partials, dt_out, s),
                   "ns_dt_noalloc");
}

// Stage sequence of one integrator call (time_integrators.py) for any
// physics: per stage a ghost fill of the stage input, the physics'
// pre-pass (closure / residual), and one fused stage launch.  dt from dt_dev
// when non-NULL; CFL partials of the new state when partials != NULL.
struct PhysOps {
  int nvar;
  int dtype;
  const int32_t* bc;
  void* ctx;
  int (*pre)(void* ctx, const void* in, void* C, void* stream);
  int (*stage)(void* ctx, const void* in, const void* C, const StageIO& io, double* partials, void* stream);
};

static int step_impl(const mfvm_grid* g, const PhysOps& ph, const mfvm_ns_work* w, int32_t integrator,
                     const double* dt_dev, double dt, double* partials, bool fix_ghosts, void* stream) {
  int rc;
  struct S {
    void* in;
    void* out;
    int update, acc;
    double c0, c1, ac;
  };
  S seq[4];
  int ns = 0;
  void* last_in = nullptr;
  switch (integrator) {
    case MFVM_INT_EULER:  // time_integrators.py:43-47
      seq[0] = {w->U, w->S1, MFVM_UPD_EULER, MFVM_ACC_NONE, 0, 0, 0};
      ns = 1;
      break;
    case MFVM_INT_RK3:  // :87-103
      seq[0] = {w->U, w->S1, MFVM_UPD_BASE, MFVM_ACC_SET, 0.5, 0, 0};
      seq[1] = {w->S1, w->S2, MFVM_UPD_BASE, MFVM_ACC_ADD, 1.0, 0, 4.0};
      seq[2] = {w->S2, w->U, MFVM_UPD_ACC6, MFVM_ACC_NONE, 0, 0, 0};
      ns = 3;
      last_in = w->S2;
      break;
    case MFVM_INT_SSPRK3:  // :143-159
      seq[0] = {w->U, w->S1, MFVM_UPD_EULER, MFVM_ACC_NONE, 0, 0, 0};
      seq[1] = {w->S1, w->S2, MFVM_UPD_SSP, MFVM_ACC_NONE, 0.75, 0.25, 0};
      seq[2] = {w->S2, w->U, MFVM_UPD_SSP, MFVM_ACC_NONE, 1.0 / 3.0, 2.0 / 3.0, 0};
      ns = 3;
      last_in = w->S2;
      break;
    default:  // rk4 :199-221
      seq[0] = {w->U, w->S1, MFVM_UPD_BASE, MFVM_ACC_SET, 0.5, 0, 0};
      seq[1] = {w->S1, w->S2, MFVM_UPD_BASE, MFVM_ACC_ADD, 0.5, 0, 2.0};
      seq[2] = {w->S2, w->S1, MFVM_UPD_BASE, MFVM_ACC_ADD, 1.0, 0, 2.0};
      seq[3] = {w->S1, w->U, MFVM_UPD_ACC6, MFVM_ACC_NONE, 0, 0, 0};
      ns = 4;
      last_in = w->S1;
      break;
  }
  for (int q = 0; q < ns; ++q) {
    if ((rc = mfvm_apply_bcs(g, ph.nvar, ph.bc, ph.dtype, seq[q].in, seq[q].in, stream))) return rc;
    if (ph.pre && (rc = ph.pre(ph.ctx, seq[q].in, w->C, stream))) return rc;
    StageIO io;
    io.U0 = w->U;
    io.out = seq[q].out;
    io.kacc = w->kacc;
    io.update = seq[q].update;
    io.acc = seq[q].acc;
    io.c0 = seq[q].c0;
    io.c1 = seq[q].c1;
    io.acc_coef = seq[q].ac;
    io.dt = dt;
    io.dt_dev = dt_dev;
    if ((rc = ph.stage(ph.ctx, seq[q].in, w->C, io, (q == ns - 1) ? partials : nullptr, stream))) return rc;
  }
  if (integrator == MFVM_INT_EULER && (rc = mfvm_copy_interior(g, ph.nvar, ph.dtype, w->S1, w->U, stream)))
    return rc;
  // Leave the ghosts as the reference does: apply_bcs of the last stage input.
  if (fix_ghosts && last_in && (rc = mfvm_apply_bcs(g, ph.nvar, ph.bc, ph.dtype, last_in, w->U, stream)))
    return rc;
  return MFVM_OK;
}

static int check_work(const mfvm_ns_work* w, int32_t integrator, bool need_c) {
  if (!w || !w->U || !w->S1 || !w->S2 || !w->kacc || (need_c && !w->C))
    return fail(MFVM_E_ARG, "work: NULL buffer");
  if (integrator < MFVM_INT_EULER || integrator > MFVM_INT_RK4) return fail(MFVM_E_ARG, "unknown integrator");
  return MFVM_OK;
}

// ---- Navier-Stokes ops ----
struct NSCtx {
  const mfvm_grid* g;
  const mfvm_ns_params* p;
  void* xy;  // split-stage scratch or NULL
};
static int ns_pre(void* c, const void* in, void* C, void* stream) {
  NSCtx* x = static_cast<NSCtx*>(c);
  // the float32 stage and the float64 split stage form P from U: the closure leaves it out
  if (x->p->dtype == MFVM_F32)
    return hip_check(launch_closure<float>(x->g, x->p, in, C, static_cast<hipStream_t>(stream), false), "ns_closure");
  if (x->xy)
    return hip_check(launch_closure<double>(x->g, x->p, in, C, static_cast<hipStream_t>(stream), false), "ns_closure");
  return mfvm_ns_closure(x->g, x->p, in, C, stream);
}
static int ns_stage_op(void* c, const void* in, const void* C, const StageIO& io, double* partials, void* stream) {
  NSCtx* x = static_cast<NSCtx*>(c);
  mfvm_stage st = {};
  st.U = in;
  st.C = C;
  st.U0 = io.U0;
  st.out = io.out;
  st.kacc = io.kacc;
  st.update = io.update;
  st.acc = io.acc;
  st.c0 = io.c0;
  st.c1 = io.c1;
  st.acc_coef = io.acc_coef;
  st.dt = io.dt;
  st.dt_dev = io.dt_dev;
  st.partials = partials;
  st.xy = x->xy;
  return mfvm_ns_stage(x->g, x->p, &st, stream);
}

int mfvm_ns_step(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w, int32_t integrator,
                 double dt, void* stream) {
  int rc = check_ns(g, p);
  if (!rc) rc = check_work(w, integrator, true);
  if (rc) return rc;
  NSCtx x{g, p, w->xy};
  PhysOps ph{5, p->dtype, &p->bc[0][0], &x, ns_pre, ns_stage_op};
  return step_impl(g, ph, w, integrator, nullptr, dt, nullptr, true, stream);
}

int mfvm_ns_advance(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w, int32_t integrator,
                    double cfl, int32_t nsteps, void* stream) {
  int rc = check_ns(g, p);
  if (!rc) rc = check_work(w, integrator, true);
  if (rc) return rc;
  if (!w->partials || !w->dts) return fail(MFVM_E_ARG, "ns_advance: NULL partials/dts");
  if (nsteps < 0) return fail(MFVM_E_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return MFVM_OK;
  if ((rc = mfvm_ns_dt(g, p, w->U, cfl, w->partials, w->dts, stream))) return rc;
  NSCtx x{g, p, w->xy};
  PhysOps ph{5, p->dtype, &p->bc[0][0], &x, ns_pre, ns_stage_op};
  for (int n = 0; n < nsteps; ++n) {
    if ((rc = step_impl(g, ph, w, integrator, w->dts + n, 0.0, w->partials, n == nsteps - 1, stream)))
      return rc;
    // dt of the next step from the new state (navier_stokes.py:512-585)
    if ((rc = hip_check(launch_dt_finalize(g, p, w->partials, cfl, w->dts + n + 1,
                                           static_cast<hipStream_t>(stream)),
                        "ns_advance")))
      return rc;
  }
  return MFVM_OK;
}

// ---- compensated FP32 Navier-Stokes ----
int mfvm_ns_comp_advance(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_comp_work* w,
                         int32_t integrator, double cfl, int32_t nsteps, void* stream) {
  int rc = check_ns(g, p);
  if (rc) return rc;
  if (p->dtype != MFVM_F32) return fail(MFVM_E_ARG, "ns_comp_advance: dtype must be f32");
  if (!w || !w->hi || !w->lo || !w->Sa || !w->Sb || !w->C || !w->dinc || !w->kacc || !w->partials || !w->dts)
    return fail(MFVM_E_ARG, "ns_comp_advance: NULL work buffer");
  if (integrator < MFVM_INT_EULER || integrator > MFVM_INT_RK4) return fail(MFVM_E_ARG, "unknown integrator");
  if (nsteps < 0) return fail(MFVM_E_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return MFVM_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int32_t* bc = &p->bc[0][0];
  if ((rc = hip_check(launch_cfl<float>(g, p, w->hi, w->lo, cfl, w->partials, w->dts, s), "ns_comp_advance")))
    return rc;
  struct S {
    void* in;
    void* next;
    int comp, acc;
    double c0, ac;
  };
  S seq[4];
  int ns = 0;
  switch (integrator) {
    case MFVM_INT_EULER:
      seq[0] = {w->hi, nullptr, MFVM_COMP_FINAL_SET, MFVM_ACC_NONE, 1.0, 0};
      ns = 1;
      break;
    case MFVM_INT_RK3:
      seq[0] = {w->hi, w->Sa, MFVM_COMP_SET, MFVM_ACC_SET, 0.5, 0};
      seq[1] = {w->Sa, w->Sb, MFVM_COMP_SET, MFVM_ACC_ADD, 1.0, 4.0};
      seq[2] = {w->Sb, nullptr, MFVM_COMP_FINAL_ACC6, MFVM_ACC_NONE, 0, 0};
      ns = 3;
      break;
    case MFVM_INT_SSPRK3:
      seq[0] = {w->hi, w->Sa, MFVM_COMP_SET, MFVM_ACC_NONE, 1.0, 0};
      seq[1] = {w->Sa, w->Sb, MFVM_COMP_SSP, MFVM_ACC_NONE, 0.25, 0};
      seq[2] = {w->Sb, nullptr, MFVM_COMP_FINAL_SSP, MFVM_ACC_NONE, 2.0 / 3.0, 0};
      ns = 3;
      break;
    default:
      seq[0] = {w->hi, w->Sa, MFVM_COMP_SET, MFVM_ACC_SET, 0.5, 0};
      seq[1] = {w->Sa, w->Sb, MFVM_COMP_SET, MFVM_ACC_ADD, 0.5, 2.0};
      seq[2] = {w->Sb, w->Sa, MFVM_COMP_SET, MFVM_ACC_ADD, 1.0, 2.0};
      seq[3] = {w->Sa, nullptr, MFVM_COMP_FINAL_ACC6, MFVM_ACC_NONE, 0, 0};
      ns = 4;
      break;
  }
  for (int n = 0; n < nsteps; ++n) {
    for (int q = 0; q < ns; ++q) {
      if ((rc = mfvm_apply_bcs(g, 5, bc, MFVM_F32, seq[q].in, seq[q].in, stream))) return rc;
      // the closure without P: the float32 stage forms P from U
      if ((rc = hip_check(launch_closure<float>(g, p, seq[q].in, w->C, s, false), "ns_comp_advance"))) return rc;
      const bool fin = q == ns - 1;
      mfvm_stage st = {};
      st.U = seq[q].in;
      st.C = w->C;
      st.U0 = w->hi;
      st.out = fin ? (integrator == MFVM_INT_EULER ? w->Sa : w->hi) : nullptr;
      st.kacc = w->kacc;
      st.update = MFVM_UPD_NONE;
      st.acc = seq[q].acc;
      st.c0 = seq[q].c0;
      st.acc_coef = seq[q].ac;
      st.dt_dev = w->dts + n;
      st.partials = fin ? w->partials : nullptr;
      st.lo = w->lo;
      st.dinc = w->dinc;
      st.snext = seq[q].next;
      st.comp = seq[q].comp;
      st.xy = w->xy;
      if ((rc = mfvm_ns_stage(g, p, &st, stream))) return rc;
    }
    if (integrator == MFVM_INT_EULER && (rc = mfvm_copy_interior(g, 5, MFVM_F32, w->Sa, w->hi, stream)))
      return rc;
    if ((rc = hip_check(launch_dt_finalize(g, p, w->partials, cfl, w->dts + n + 1, s), "ns_comp_advance")))
      return rc;
  }
  if (ns > 1 && (rc = mfvm_apply_bcs(g, 5, bc, MFVM_F32, seq[ns - 1].in, w->hi, stream))) return rc;
  return MFVM_OK;
}

// ---- batched ensemble ----
static int check_ens_members(const mfvm_ens_member* m, int32_t n) {
  if (!m || n < 1) return fail(MFVM_E_ARG, "ensemble: need at least one member");
  for (int i = 0; i < n; ++i) {
    int rc = check_grid(&m[i].grid);
    if (!rc) rc = check_bc(5, &m[i].bc[0][0]);
    if (rc) return rc;
    const int bx = m[i].tile[0] > 0 ? m[i].tile[0] : 64, by = m[i].tile[1] > 0 ? m[i].tile[1] : 4;
    if (bx < 2 || by < 1 || bx * by != 256 || bx + by > 256 || m[i].tile[2] < 0)
      return fail(MFVM_E_ARG, "ensemble: member tile must have x1 threads >= 2, x1*x2 threads == 256, "
                              "x1 + x2 threads <= 256, x3 planes >= 0");
  }
  return MFVM_OK;
}

int64_t mfvm_ens_member_tiles(const mfvm_ens_member* m) {
  if (check_ens_members(m, 1)) return -1;
  return ens_member_tiles(m);
}

int64_t mfvm_ens_table_bytes(const mfvm_ens_member* m, int32_t n, int32_t integrator) {
  if (check_ens_members(m, n)) return -1;
  if (ens_nstages(integrator) == 0) {
    fail(MFVM_E_UNSUP, "ensemble: integrator must be rk3, rk3_ssp or rk4");
    return -1;
  }
  return ens_table_bytes(m, n, integrator);
}

int mfvm_ens_advance(const mfvm_ens_member* m, int32_t n, const mfvm_ns_params* p, const mfvm_ens_opts* o,
                     int32_t integrator, int32_t compensated, double cfl, int32_t nsteps, void* table,
                     int64_t table_bytes, void* stream) {
  int rc = check_ens_members(m, n);
  if (rc) return rc;
  if (!p || !o) return fail(MFVM_E_ARG, "ensemble: NULL params or options");
  if (p->limiter < MFVM_LIM_NONE || p->limiter > MFVM_LIM_VANLEER)
    return fail(MFVM_E_UNSUP, "unsupported slope limiter / reconstructor");
  if (p->riemann != MFVM_RS_HLL && p->riemann != MFVM_RS_RUSANOV)
    return fail(MFVM_E_UNSUP, "unsupported Riemann solver");
  if (ens_nstages(integrator) == 0) return fail(MFVM_E_UNSUP, "ensemble: integrator must be rk3, rk3_ssp or rk4");
  if (compensated && p->dtype != MFVM_F32) return fail(MFVM_E_ARG, "ensemble: compensated runs need dtype f32");
  if (!compensated && p->dtype != MFVM_F64) return fail(MFVM_E_ARG, "ensemble: plain runs need dtype f64");
  if (o->order < MFVM_ENS_ORDER_MEMBER || o->order > MFVM_ENS_ORDER_INTERLEAVE || o->persist < 0 || o->graph < 0)
    return fail(MFVM_E_ARG, "ensemble: bad options");
  for (int i = 0; i < n; ++i) {
    const mfvm_ens_member& x = m[i];
    if (!x.U || !x.S1 || !x.S2 || !x.C || !x.kacc || !x.partials || !x.dts)
      return fail(MFVM_E_ARG, "ensemble: NULL member buffer");
    if (compensated && (!x.lo || !x.dinc)) return fail(MFVM_E_ARG, "ensemble: compensated member needs lo and dinc");
  }
  if (nsteps < 0) return fail(MFVM_E_ARG, "nsteps must be >= 0");
  const long long need = ens_table_bytes(m, n, integrator);
  if (!table || table_bytes < need) return fail(MFVM_E_ARG, "ensemble: table too small");
  if (nsteps == 0) return MFVM_OK;
  return hip_check(ens_advance(m, n, p, o, integrator, compensated != 0, cfl, nsteps, table,
                               static_cast<hipStream_t>(stream)),
                   "ens_advance");
}

// ---- learned flux-correction closure ----
static int check_cnn_dims(int32_t batch, int32_t n1, int32_t n2) {
  if (batch < 1 || n1 < 16 || n2 < 1 || n1 % 16 != 0)
    return fail(MFVM_E_ARG, "cnn: need batch >= 1, nx2 >= 1 and nx1 a positive multiple of 16");
  return MFVM_OK;
}

int mfvm_cnn_conv(int32_t mode, int32_t cin, int32_t cout, int32_t batch, int32_t n1, int32_t n2, const void* x,
                  const void* w, const float* bias, const void* mask, void* y, void* stream) {
  int rc = check_cnn_dims(batch, n1, n2);
  if (rc) return rc;
  if (mode < MFVM_CNN_FWD_RELU || mode > MFVM_CNN_BWD_DATA) return fail(MFVM_E_ARG, "cnn_conv: bad mode");
  if (!cnn_shape_ok(mode, cin, cout)) return fail(MFVM_E_UNSUP, "cnn_conv: unsupported channel counts");
  if (!x || !w || !y || (mode != MFVM_CNN_BWD_DATA && !bias) || (mode == MFVM_CNN_BWD_DATA && !mask))
    return fail(MFVM_E_ARG, "cnn_conv: NULL buffer");
  return hip_check(cnn_conv(mode, cin, cout, batch, n1, n2, x, w, bias, mask, y, static_cast<hipStream_t>(stream)),
                   "cnn_conv");
}

int mfvm_cnn_wgrad(int32_t cin, int32_t cout, int32_t batch, int32_t n1, int32_t n2, const void* x, const void* gy,
                   float* gw, float* gb, void* stream) {
  int rc = check_cnn_dims(batch, n1, n2);
  if (rc) return rc;
  if (!cnn_shape_ok(MFVM_CNN_FWD_RELU, cin, cout)) return fail(MFVM_E_UNSUP, "cnn_wgrad: unsupported channel counts");
  if (!x || !gy || !gw || !gb) return fail(MFVM_E_ARG, "cnn_wgrad: NULL buffer");
  return hip_check(cnn_wgrad(cin, cout, batch, n1, n2, x, gy, gw, gb, static_cast<hipStream_t>(stream)), "cnn_wgrad");
}

int mfvm_cnn_pack(int32_t cin, int32_t cout, const float* w, void* wf, void* wb, void* stream) {
  if (!cnn_shape_ok(MFVM_CNN_FWD_RELU, cin, cout)) return fail(MFVM_E_UNSUP, "cnn_pack: unsupported channel counts");
  if (!w || !wf || !wb) return fail(MFVM_E_ARG, "cnn_pack: NULL buffer");
  return hip_check(cnn_pack(cin, cout, w, wf, wb, static_cast<hipStream_t>(stream)), "cnn_pack");
}

int mfvm_cnn_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                  int32_t t, void* stream) {
  if (n < 1 || !w || !g || !m || !v || t < 1) return fail(MFVM_E_ARG, "cnn_adam: bad arguments");
  return hip_check(cnn_adam(n, w, g, m, v, lr, b1, b2, eps, t, static_cast<hipStream_t>(stream)), "cnn_adam");
}

int mfvm_cnn_features(const mfvm_grid* g, int32_t dtype, double gamma, int32_t batch, int64_t batch_stride,
                      const void* U, const float* shift, const float* scale, void* x, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (g->n[2] != 1) return fail(MFVM_E_UNSUP, "cnn_features: 2D grids only (nx3 == 1)");
  if (batch < 1 || !U || !shift || !scale || !x) return fail(MFVM_E_ARG, "cnn_features: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == MFVM_F64)
    return hip_check(cnn_features<double>(g, gamma, batch, batch_stride, U, shift, scale, x, s), "cnn_features");
  if (dtype == MFVM_F32)
    return hip_check(cnn_features<float>(g, gamma, batch, batch_stride, U, shift, scale, x, s), "cnn_features");
  return fail(MFVM_E_ARG, "dtype must be 0 (f64) or 1 (f32)");
}

int mfvm_cnn_loss(int32_t batch, int32_t n1, int32_t n2, double dx1, double dx2, const float* y, const float* target,
                  const float* out_scale, const float* wf, void* gy, float* loss, void* stream) {
  if (batch < 1 || n1 < 1 || n2 < 1 || !(dx1 > 0) || !(dx2 > 0)) return fail(MFVM_E_ARG, "cnn_loss: bad sizes");
  if (!y || !target || !out_scale || !wf || !gy || !loss) return fail(MFVM_E_ARG, "cnn_loss: NULL buffer");
  return hip_check(cnn_loss(batch, n1, n2, (float)(1.0 / dx1), (float)(1.0 / dx2), y, target, out_scale, wf, gy, loss,
                            static_cast<hipStream_t>(stream)),
                   "cnn_loss");
}

// One rk3_ssp step (time_integrators.py:143-159) with the learned correction
// in every stage, then the dt of the next step appended to w->dts.
static int closure_step(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w, const mfvm_cnn* net,
                        const mfvm_closure_work* cw, double cfl, double t_end, hipStream_t s) {
  struct Q {
    void* in;
    void* out;
    int update;
    double c0, c1, coef;
  };
  const Q seq[3] = {{w->U, w->S1, MFVM_UPD_EULER, 0.0, 0.0, 1.0},
                    {w->S1, w->S2, MFVM_UPD_SSP, 0.75, 0.25, 0.25},
                    {w->S2, w->U, MFVM_UPD_SSP, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  const int n1 = g->n[0], n2 = g->n[1];
  const int32_t* bc = &p->bc[0][0];
  // (the stage flags start at 0: the previous step's dt append, or the set-up, cleared them)
  hipError_t e = hipSuccess;
  // (pending: the selection of stage q-1 is still to be made, fused with
  // this stage's ghost fill; the last stage's, with the CFL partials)
  bool pending = false;
  for (int q = 0; q < 3 && e == hipSuccess; ++q) {
    if (!pending || !launch_select_bcs<double>(g, 5, bc, cw->T, seq[q].in, cw->flag + q - 1, cw->counts + q - 1, s, &e)) {
      if (pending) e = closure_select<double>(g, cw->T, seq[q].in, cw->flag + q - 1, cw->counts + q - 1, s);
      if (e == hipSuccess) e = launch_bcs<double>(g, 5, bc, seq[q].in, seq[q].in, s);
    }
    pending = false;
    if (e == hipSuccess) e = launch_closure<double>(g, p, seq[q].in, w->C, s);
    mfvm_stage st = {};
    st.U = seq[q].in;
    st.C = w->C;
    st.U0 = w->U;
    st.out = seq[q].out;
    st.kacc = w->kacc;
    st.update = seq[q].update;
    st.acc = MFVM_ACC_NONE;
    st.c0 = seq[q].c0;
    st.c1 = seq[q].c1;
    st.dt_dev = cw->dt_cur;
    st.xy = w->xy;  // split stages where the caller provides the scratch (same results)
    if (e == hipSuccess) e = launch_stage<double>(g, p, &st, s);
    if (!net) continue;
    if (!cw->layered) {
      // features, the three layers and the corrected state in one launch
      const void* wf[3] = {net->wf[0], net->wf[1], net->wf[2]};
      const float* b[3] = {net->b[0], net->b[1], net->b[2]};
      if (e == hipSuccess)
        e = closure_net_apply(g, p->gamma, seq[q].in, seq[q].out, net->in_shift, net->in_scale, wf, b, net->out_scale,
                              seq[q].coef, cw->dt_cur, cw->T, cw->flag + q, nullptr, s);
    } else {
      // features of the stage input and the first layer in one launch, then one per layer
      if (e == hipSuccess)
        e = cnn_conv_features(g, p->gamma, seq[q].in, net->in_shift, net->in_scale, net->wf[0], net->b[0], cw->a1, s);
      if (e == hipSuccess)
        e = cnn_conv(MFVM_CNN_FWD_RELU, 32, 32, 1, n1, n2, cw->a1, net->wf[1], net->b[1], nullptr, cw->a2, s);
      if (e == hipSuccess)
        e = cnn_conv(MFVM_CNN_FWD_LINEAR, 32, 8, 1, n1, n2, cw->a2, net->wf[2], net->b[2], nullptr, cw->y, s);
      if (e == hipSuccess)
        e = closure_apply<double>(g, p->gamma, seq[q].out, cw->y, net->out_scale, seq[q].coef, cw->dt_cur, cw->T,
                                  cw->flag + q, s);
    }
    pending = true;
  }
  if (e == hipSuccess)
    e = pending ? launch_select_cfl_partials(g, p, cw->T, w->U, cw->flag + 2, cw->counts + 2, w->partials, s)
                : launch_cfl_partials<double>(g, p, w->U, nullptr, w->partials, s);
  if (e == hipSuccess)
    e = launch_dt_append(g, p, w->partials, cfl, t_end, w->dts, cw->dt_cur, cw->t_cur, cw->step, cw->flag, 3, s);
  return hip_check(e, "closure_advance");
}

int mfvm_closure_advance(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w, const mfvm_cnn* net,
                         const mfvm_closure_work* cw, double cfl, double t_end, int32_t nsteps, int32_t graph,
                         void* stream) {
  int rc = check_ns(g, p);
  if (!rc) rc = check_work(w, MFVM_INT_SSPRK3, true);
  if (rc) return rc;
  if (p->dtype != MFVM_F64) return fail(MFVM_E_ARG, "closure_advance: dtype must be f64");
  if (g->n[2] != 1 || !g->uniform || g->n[0] % 16 != 0)
    return fail(MFVM_E_UNSUP, "closure_advance: needs a uniform Cartesian 2D grid with nx1 a multiple of 16");
  for (int v = 0; v < 5; ++v)
    for (int d = 0; d < 4; ++d)
      if (p->bc[v][d] != MFVM_BC_PERIODIC) return fail(MFVM_E_UNSUP, "closure_advance: x1 and x2 must be periodic");
  if (!w->partials || !w->dts) return fail(MFVM_E_ARG, "closure_advance: NULL partials/dts");
  if (!cw) return fail(MFVM_E_ARG, "closure_advance: NULL work");
  if (!cw->flag || !cw->counts || !cw->dt_cur || !cw->t_cur || !cw->step)
    return fail(MFVM_E_ARG, "closure_advance: NULL work buffer");
  if (net) {
    for (int l = 0; l < 3; ++l)
      if (!net->wf[l] || !net->b[l]) return fail(MFVM_E_ARG, "closure_advance: NULL network weights");
    if (!net->in_shift || !net->in_scale || !net->out_scale) return fail(MFVM_E_ARG, "closure_advance: NULL scales");
    if (!cw->T || (cw->layered && (!cw->a1 || !cw->a2 || !cw->y)))
      return fail(MFVM_E_ARG, "closure_advance: NULL work buffer");
  }
  if (t_end != t_end) return fail(MFVM_E_ARG, "closure_advance: t_end is NaN");
  if (nsteps < 0 || graph < 0) return fail(MFVM_E_ARG, "nsteps and graph must be >= 0");
  if (nsteps == 0) return MFVM_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = fill_int(cw->step, 1, 0, s);
  if (e == hipSuccess) e = fill_int(cw->counts, 3, 0, s);
  if (e == hipSuccess) e = hipMemsetAsync(cw->t_cur, 0, sizeof(double), s);  // not captured
  if (e == hipSuccess) e = launch_cfl_partials<double>(g, p, w->U, nullptr, w->partials, s);
  if (e == hipSuccess)
    e = launch_dt_append(g, p, w->partials, cfl, t_end, w->dts, cw->dt_cur, cw->t_cur, cw->step, cw->flag, 3, s);
  if ((rc = hip_check(e, "closure_advance"))) return rc;
  int done = 0;
  if (graph > 0) {
    const int chunk = graph < nsteps ? graph : nsteps;
    hipGraph_t gr = nullptr;
    hipGraphExec_t ex = nullptr;
    if ((rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "closure_advance: capture")))
      return rc;
    for (int c = 0; c < chunk && !rc; ++c) rc = closure_step(g, p, w, net, cw, cfl, t_end, s);
    e = hipStreamEndCapture(s, &gr);
    if (!rc) rc = hip_check(e, "closure_advance: end capture");
    if (!rc) rc = hip_check(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0), "closure_advance: instantiate");
    for (int r = 0; !rc && r < nsteps / chunk; ++r) {
      rc = hip_check(hipGraphLaunch(ex, s), "closure_advance: launch");
      done += chunk;
    }
    const hipError_t se = hipStreamSynchronize(s);
    if (ex) (void)hipGraphExecDestroy(ex);
    if (gr) (void)hipGraphDestroy(gr);
    if (rc) return rc;
    if ((rc = hip_check(se, "closure_advance: sync"))) return rc;
  }
  for (; done < nsteps; ++done)
    if ((rc = closure_step(g, p, w, net, cw, cfl, t_end, s))) return rc;
  // ghosts as the reference leaves them: the fill of the last stage input
  return mfvm_apply_bcs(g, 5, &p->bc[0][0], MFVM_F64, w->S2, w->U, stream);
}

// ---- heat ----
static int check_heat(const mfvm_grid* g, const mfvm_heat_params* p) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (!p) return fail(MFVM_E_ARG, "params descriptor is NULL");
  if (p->dtype != MFVM_F64 && p->dtype != MFVM_F32) return fail(MFVM_E_ARG, "dtype must be 0 (f64) or 1 (f32)");
  if (p->limiter < MFVM_LIM_NONE || p->limiter > MFVM_LIM_VANLEER)
    return fail(MFVM_E_UNSUP, "unsupported slope limiter / reconstructor");
  return check_bc(1, &p->bc[0][0]);
}
struct HeatCtx {
  const mfvm_grid* g;
  const mfvm_heat_params* p;
};
static int heat_stage_op(void* c, const void* in, const void*, const StageIO& io, double*, void* stream) {
  HeatCtx* x = static_cast<HeatCtx*>(c);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(x->p->dtype == MFVM_F64 ? heat_launch_stage<double>(x->g, x->p, in, io, s)
                                           : heat_launch_stage<float>(x->g, x->p, in, io, s),
                   "heat_stage");
}

int mfvm_heat_rhs(const mfvm_grid* g, const mfvm_heat_params* p, const void* U, void* rhs, void* stream) {
  int rc = check_heat(g, p);
  if (rc) return rc;
  if (!U || !rhs) return fail(MFVM_E_ARG, "heat_rhs: NULL buffer");
  HeatCtx x{g, p};
  StageIO io = {};
  io.kacc = rhs;
  io.update = MFVM_UPD_NONE;
  io.acc = MFVM_ACC_SET;
  return heat_stage_op(&x, U, nullptr, io, nullptr, stream);
}

int mfvm_heat_dt(const mfvm_grid* g, const mfvm_heat_params* p, double cfl, double* dt_out, void* stream) {
  int rc = check_heat(g, p);
  if (rc) return rc;
  if (!dt_out) return fail(MFVM_E_ARG, "heat_dt: NULL output");
  return hip_check(heat_launch_dt(g, p, cfl, dt_out, static_cast<hipStream_t>(stream)), "heat_dt");
}

int mfvm_heat_step(const mfvm_grid* g, const mfvm_heat_params* p, const mfvm_ns_work* w, int32_t integrator,
                   double dt, void* stream) {
  int rc = check_heat(g, p);
  if (!rc) rc = check_work(w, integrator, false);
  if (rc) return rc;
  HeatCtx x{g, p};
  PhysOps ph{1, p->dtype, &p->bc[0][0], &x, nullptr, heat_stage_op};
  return step_impl(g, ph, w, integrator, nullptr, dt, nullptr, true, stream);
}

int mfvm_heat_advance(const mfvm_grid* g, const mfvm_heat_params* p, const mfvm_ns_work* w, int32_t integrator,
                      double cfl, int32_t nsteps, void* stream) {
  int rc = check_heat(g, p);
  if (!rc) rc = check_work(w, integrator, false);
  if (rc) return rc;
  if (!w->dts || nsteps < 0) return fail(MFVM_E_ARG, "heat_advance: bad dts / nsteps");
  HeatCtx x{g, p};
  PhysOps ph{1, p->dtype, &p->bc[0][0], &x, nullptr, heat_stage_op};
  // heat dt depends on the parameters and the grid only (heat.py:338-349):
  // computed once, every step's entry holds that value
  if ((rc = mfvm_heat_dt(g, p, cfl, w->dts, stream))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((rc = hip_check(heat_launch_dt_fill(w->dts, nsteps, s), "heat_dt"))) return rc;
  if (integrator != MFVM_INT_EULER) {
    for (int n = 0; n < nsteps; ++n)
      if ((rc = step_impl(g, ph, w, integrator, w->dts + n, 0.0, nullptr, n == nsteps - 1, stream))) return rc;
    return MFVM_OK;
  }
  // Euler: the steps alternate between U and S1 (time_integrators.py:43-47
  // per step: ghosts of the input, then out = in + dt k) instead of copying
  // S1 back into U after each step.
  void* buf[2] = {w->U, w->S1};
  int cur = 0;
  for (int n = 0; n < nsteps; ++n) {
    void* in = buf[cur];
    void* out = buf[1 - cur];
    if ((rc = mfvm_apply_bcs(g, 1, ph.bc, ph.dtype, in, in, stream))) return rc;
    StageIO io = {};
    io.U0 = in;
    io.out = out;
    io.kacc = w->kacc;
    io.update = MFVM_UPD_EULER;
    io.acc = MFVM_ACC_NONE;
    io.dt_dev = w->dts + n;
    if ((rc = heat_stage_op(&x, in, nullptr, io, nullptr, stream))) return rc;
    cur = 1 - cur;
  }
  // The reference's U after the last step: the ghosts filled from the last
  // input, the interior updated.
  if (nsteps > 0 && cur == 1) {
    if ((rc = mfvm_copy_interior(g, 1, p->dtype, w->S1, w->U, stream))) return rc;
  } else if (nsteps > 0) {
    if ((rc = mfvm_apply_bcs(g, 1, ph.bc, ph.dtype, w->S1, w->U, stream))) return rc;
  }
  return MFVM_OK;
}

// ---- Maxwell ----
static int check_mx(const mfvm_grid* g, const mfvm_maxwell_params* p) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (!p) return fail(MFVM_E_ARG, "params descriptor is NULL");
  if (p->dtype != MFVM_F64 && p->dtype != MFVM_F32) return fail(MFVM_E_ARG, "dtype must be 0 (f64) or 1 (f32)");
  if (p->limiter < MFVM_LIM_NONE || p->limiter > MFVM_LIM_VANLEER)
    return fail(MFVM_E_UNSUP, "unsupported slope limiter / reconstructor");
  if (p->riemann != MFVM_RS_HLL && p->riemann != MFVM_RS_RUSANOV)
    return fail(MFVM_E_UNSUP, "unsupported Riemann solver");
  if (!(p->mu > 0.0) || !(p->ep > 0.0)) return fail(MFVM_E_ARG, "permeability and permittivity must be > 0");
  return check_bc(6, &p->bc[0][0]);
}
struct MXCtx {
  const mfvm_grid* g;
  const mfvm_maxwell_params* p;
  void* xy;  // split stage scratch (mfvm_ns_work.xy), NULL: fused stage
};
static int mx_pre(void* c, const void* in, void* R, void* stream) {
  MXCtx* x = static_cast<MXCtx*>(c);
  return mfvm_maxwell_residual(x->g, x->p, in, R, stream);
}
static int mx_stage_op(void* c, const void* in, const void* R, const StageIO& io, double*, void* stream) {
  MXCtx* x = static_cast<MXCtx*>(c);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(x->p->dtype == MFVM_F64 ? maxwell_launch_stage<double>(x->g, x->p, in, R, x->xy, io, s)
                                           : maxwell_launch_stage<float>(x->g, x->p, in, R, x->xy, io, s),
                   "maxwell_stage");
}

int mfvm_maxwell_residual(const mfvm_grid* g, const mfvm_maxwell_params* p, const void* U, void* R,
                          void* stream) {
  int rc = check_mx(g, p);
  if (rc) return rc;
  if (!U || !R) return fail(MFVM_E_ARG, "maxwell_residual: NULL buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return hip_check(p->dtype == MFVM_F64 ? maxwell_launch_res<double>(g, p, U, R, s)
                                        : maxwell_launch_res<float>(g, p, U, R, s),
                   "maxwell_residual");
}

int mfvm_maxwell_rhs(const mfvm_grid* g, const mfvm_maxwell_params* p, const void* U, void* R, void* rhs,
                     void* stream) {
  int rc = mfvm_maxwell_residual(g, p, U, R, stream);
  if (rc) return rc;
  if (!rhs) return fail(MFVM_E_ARG, "maxwell_rhs: NULL output");
  MXCtx x{g, p, nullptr};
  StageIO io = {};
  io.kacc = rhs;
  io.update = MFVM_UPD_NONE;
  io.acc = MFVM_ACC_SET;
  return mx_stage_op(&x, U, R, io, nullptr, stream);
}

int mfvm_maxwell_dt(const mfvm_grid* g, const mfvm_maxwell_params* p, double cfl, double* dt_out, void* stream) {
  int rc = check_mx(g, p);
  if (rc) return rc;
  if (!dt_out) return fail(MFVM_E_ARG, "maxwell_dt: NULL output");
  return hip_check(maxwell_launch_dt(g, p, cfl, dt_out, static_cast<hipStream_t>(stream)), "maxwell_dt");
}

int mfvm_maxwell_step(const mfvm_grid* g, const mfvm_maxwell_params* p, const mfvm_ns_work* w,
                      int32_t integrator, double dt, void* stream) {
  int rc = check_mx(g, p);
  if (!rc) rc = check_work(w, integrator, true);
  if (rc) return rc;
  MXCtx x{g, p, w->xy};
  PhysOps ph{6, p->dtype, &p->bc[0][0], &x, mx_pre, mx_stage_op};
  return step_impl(g, ph, w, integrator, nullptr, dt, nullptr, true, stream);
}

int mfvm_maxwell_advance(const mfvm_grid* g, const mfvm_maxwell_params* p, const mfvm_ns_work* w,
                         int32_t integrator, double cfl, int32_t nsteps, void* stream) {
  int rc = check_mx(g, p);
  if (!rc) rc = check_work(w, integrator, true);
  if (rc) return rc;
  if (!w->dts || nsteps < 0) return fail(MFVM_E_ARG, "maxwell_advance: bad dts / nsteps");
  MXCtx x{g, p, w->xy};
  PhysOps ph{6, p->dtype, &p->bc[0][0], &x, mx_pre, mx_stage_op};
  for (int n = 0; n < nsteps; ++n) {
    // Maxwell dt depends on the grid and parameters only (maxwell.py:579-598)
    if ((rc = mfvm_maxwell_dt(g, p, cfl, w->dts + n, stream))) return rc;
    if ((rc = step_impl(g, ph, w, integrator, w->dts + n, 0.0, nullptr, n == nsteps - 1, stream))) return rc;
  }
  if (nsteps > 0 && (rc = mfvm_maxwell_dt(g, p, cfl, w->dts + nsteps, stream))) return rc;
  return MFVM_OK;
}

// ---- diagnostics ------------------------------------------------------------------

static int check_diag_size(const mfvm_grid* g, int32_t nvar) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (nvar < 1 || nvar > 6) return fail(MFVM_E_ARG, "diagnostics: nvar must be 1..6");
  if ((long long)g->n[0] * g->n[1] * g->n[2] > 0x7fffffffLL)
    return fail(MFVM_E_UNSUP, "diagnostics: more than 2^31-1 interior cells");
  return MFVM_OK;
}

int64_t mfvm_diag_partials_size(const mfvm_grid* g, int32_t nvar) {
  if (check_diag_size(g, nvar)) return -1;
  return diag_nblocks(g) * MFVM_DIAG_NOUT(nvar);
}

int mfvm_diagnostics(const mfvm_grid* g, int32_t nvar, int32_t dtype, const void* U, const void* lo,
                     const void* ref, double gamma, double* partials, double* out, void* stream) {
  int rc = check_diag_size(g, nvar);
  if (rc) return rc;
  if (dtype != MFVM_F64 && dtype != MFVM_F32) return fail(MFVM_E_ARG, "diagnostics: dtype must be f64 or f32");
  if (!U || !partials || !out) return fail(MFVM_E_ARG, "diagnostics: NULL buffer");
  if (lo && dtype != MFVM_F32) return fail(MFVM_E_ARG, "diagnostics: a compensation array needs dtype f32");
  if (gamma > 0.0 && nvar != 5) return fail(MFVM_E_ARG, "diagnostics: the pressure monitor needs the 5 NS fields");
  return hip_check(diag_launch(g, nvar, dtype, U, lo, ref, gamma, partials, out, static_cast<hipStream_t>(stream)),
                   "diagnostics");
}

}  // extern "C"
