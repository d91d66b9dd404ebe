/*
 * mfvm.h -- C-ABI of the MI355X (gfx950) finite-volume right-hand-side and
 * time-step library (libmfvm.so).
 *
 * The reference (multiphysics-fvm) is Python/NumPy; its "plugin API" is the
 * setup dict of callables consumed by src/solver.py.  Each entry point below
 * replaces one of those callables (cited per function) and is bound from
 * Python with ctypes by the mfvm package.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Memory: the caller owns every buffer (device memory).  The library
 *    allocates nothing; scratch comes in through the arguments.
 *  - Layout: a ghosted state of nvar fields is one SoA buffer
 *    [nvar][nx3+2][nx2+2][nx1+2] (x1 fastest, one ghost layer per side).
 *    Interior-only buffers (right-hand sides, accumulators) are
 *    [nvar][nx3][nx2][nx1].  dtype 0 = float64, 1 = float32.
 *  - Streams: every call is enqueued on the given hipStream_t (NULL = the
 *    null stream) and returns without synchronising, except where noted.
 *  - Errors: 0 on success; a negative MFVM_E_* code otherwise, with a
 *    message in mfvm_last_error() (thread-local).
 */
#ifndef MFVM_H
#define MFVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
#define MFVM_OK 0
#define MFVM_E_ARG (-1)     /* invalid argument (the Python layer raises ValueError) */
#define MFVM_E_HIP (-2)     /* HIP runtime error (RuntimeError) */
#define MFVM_E_UNSUP (-3)   /* unsupported scheme combination (NotImplementedError) */

/* ---- enums ------------------------------------------------------------------ */
#define MFVM_CART 0
#define MFVM_CYL 1
#define MFVM_SPH 2

#define MFVM_LIM_NONE 0     /* reconstructors.piecewise_constant */
#define MFVM_LIM_MINMOD 1   /* reconstructors.muscl + slope_limiters.minmod */
#define MFVM_LIM_MC 2       /* ... + monotonized_central */
#define MFVM_LIM_VANLEER 3  /* ... + van_leer */

#define MFVM_RS_HLL 0       /* riemann_solvers.hll */
#define MFVM_RS_RUSANOV 1   /* riemann_solvers.rusanov */

#define MFVM_BC_NEUMANN 0   /* boundary_conditions.neumann */
#define MFVM_BC_PERIODIC 1  /* boundary_conditions.periodic */
#define MFVM_BC_REFLECTING 2 /* boundary_conditions.reflecting */

#define MFVM_INT_EULER 0    /* time_integrators.euler_explicit */
#define MFVM_INT_RK3 1      /* time_integrators.rk3 */
#define MFVM_INT_SSPRK3 2   /* time_integrators.rk3_ssp */
#define MFVM_INT_RK4 3      /* time_integrators.rk4 */

#define MFVM_F64 0
#define MFVM_F32 1

/* Metric slots.  Metric m at cell/face (i,j,k) is (a[i]*b[j])*c[k] with
 * a = fac + (3m)*fstride, b = fac + (3m+1)*fstride, c = fac + (3m+2)*fstride
 * (0-based interior or face indices).  This reproduces geometry.build_grid
 * (src/grid/geometry.py:80-171) value for value. */
#define MFVM_MET_V 0
#define MFVM_MET_S1 1
#define MFVM_MET_S2 2
#define MFVM_MET_S3 3
#define MFVM_MET_L1 4
#define MFVM_MET_L2 5
#define MFVM_MET_L3 6
#define MFVM_MET_H1 7
#define MFVM_MET_H2 8
#define MFVM_MET_H3 9
#define MFVM_MET_D1H 10   /* d/dx1 ln(h1 h2 h3) */
#define MFVM_MET_D1H2 11
#define MFVM_MET_D1H3 12
#define MFVM_MET_D2H 13
#define MFVM_MET_D2H1 14
#define MFVM_MET_D2H3 15
#define MFVM_MET_D3H 16
#define MFVM_MET_D3H1 17
#define MFVM_MET_D3H2 18
#define MFVM_NMET 19

/* ---- descriptors ------------------------------------------------------------ */
typedef struct mfvm_grid {
  int32_t n[3];                 /* interior cells nx1, nx2, nx3 */
  int32_t coord;                /* MFVM_CART / MFVM_CYL / MFVM_SPH */
  const double* fac;            /* device: 1D metric factors, 3*MFVM_NMET rows of fstride */
  int32_t fstride;              /* >= max(nx1, nx2, nx3) + 1 */
  int32_t uniform;              /* Cartesian with one L_d per direction: L, V, S constant */
  int32_t lpow2[3];             /* 1 if L_d is the same power of two in every cell */
  int32_t vpow2;                /* 1 if V is the same power of two in every cell */
  double lval[3];               /* that L_d (when lpow2[d] or uniform) */
  double vval;                  /* that V (when vpow2 or uniform) */
  double sval[3];               /* the face areas S_d (when uniform) */
  double smax;                  /* max over cells of 1/L1^2 + 1/L2^2 + 1/L3^2 */
  double lmin[3];               /* min over cells of L_d */
} mfvm_grid;

typedef struct mfvm_ns_params {  /* navier_stokes.py: params + space_schemes */
  int32_t dtype;
  int32_t limiter;              /* MFVM_LIM_* */
  int32_t riemann;              /* MFVM_RS_* */
  int32_t bc[5][6];             /* [var][left,right,bottom,top,front,back] */
  double gamma;                 /* heat_capacity_ratio */
  double mu;                    /* dynamic_viscosity */
  double g[3];                  /* gravitational_acceleration */
  int32_t tile[3];              /* stage tile (x1 threads, x2 rows, x3 planes); 0 = default */
} mfvm_ns_params;

/* Stage update modes (time_integrators.py).  k is the right-hand side. */
#define MFVM_UPD_NONE 0   /* no state update (compute_rhs) */
#define MFVM_UPD_EULER 1  /* out = U + dt*k                               (:47) */
#define MFVM_UPD_BASE 2   /* out = U0 + (c0*dt)*k                   (:91,97,203,209,215) */
#define MFVM_UPD_SSP 3    /* out = c0*U0 + c1*(U + dt*k)                 (:153,159) */
#define MFVM_UPD_ACC6 4   /* out = U0 + (dt/6)*(kacc + k)               (:103,221) */

#define MFVM_ACC_NONE 0
#define MFVM_ACC_SET 1    /* kacc = k */
#define MFVM_ACC_ADD 2    /* kacc = kacc + acc_coef*k */

/* Compensated FP32 stages (dtype f32 with a compensation word per value):
 * U0 = hi + lo, FP32 stage states hi + d with an FP32 increment d.
 *   SET:        d = (c0*dt)*k               SSP: d = c0*(d + dt*k)
 *   FINAL_SET:  U0 += (c0*dt)*k             FINAL_SSP: U0 += c0*(d + dt*k)
 *   FINAL_ACC6: U0 += (dt/6)*(kacc + k)
 * Non-final stages write the next stage state hi + d to snext (an FP32
 * value: lo, below half an ulp of hi, is left out of it); final stages add
 * the increment to hi + lo in FP64 and store the rounded sum and its FP32
 * remainder in (out, lo).  No reference counterpart (the reference is
 * FP64 only): this is the reduced-precision mode of the build. */
#define MFVM_COMP_NONE 0
#define MFVM_COMP_SET 1
#define MFVM_COMP_SSP 2
#define MFVM_COMP_FINAL_SET 3
#define MFVM_COMP_FINAL_SSP 4
#define MFVM_COMP_FINAL_ACC6 5

typedef struct mfvm_stage {
  const void* U;                /* stage input (ghosts filled): stencil reads */
  const void* C;                /* closure of U (mfvm_ns_closure) */
  const void* U0;               /* base state (pointwise reads) */
  void* out;                    /* stage output: interior written */
  void* kacc;                   /* interior accumulator / rhs output */
  int32_t update;               /* MFVM_UPD_* */
  int32_t acc;                  /* MFVM_ACC_* */
  double c0, c1, acc_coef;
  double dt;                    /* used when dt_dev is NULL */
  const double* dt_dev;         /* device scalar dt, or NULL */
  double* partials;             /* non-NULL: reduce CFL rates of out */
  void* lo;                     /* compensated stages: compensation of U0 (ghosted) */
  void* dinc;                   /* compensated stages: increment (interior) */
  void* snext;                  /* compensated stages: next stage state (ghosted) */
  int32_t comp;                 /* MFVM_COMP_* (requires dtype f32) */
  void* xy;                     /* optional (f64, or f32 compensated): interior 5-field
                                   scratch of the stage's dtype; the stage then runs as two
                                   kernels (x1/x2 faces, then x3 faces and the update); same
                                   results */
} mfvm_stage;

typedef struct mfvm_heat_params {  /* heat.py: params + space_schemes */
  int32_t dtype;
  int32_t limiter;              /* MFVM_LIM_* (reconstruction of the heat flux) */
  int32_t bc[1][6];
  double rho_cp;                /* mass_density * specific_heat_capacity */
  double conductivity;          /* thermal_conductivity */
  double source;                /* volumetric_heat_source (scalar form) */
  const void* source_field;     /* array form (heat.py:294-297 broadcasts an array q): the
                                   source of every interior cell, [nx3][nx2][nx1] in the
                                   state's dtype; NULL: the scalar `source` everywhere */
} mfvm_heat_params;

typedef struct mfvm_maxwell_params {  /* maxwell.py: params + space_schemes */
  int32_t dtype;
  int32_t limiter;
  int32_t riemann;
  int32_t bc[6][6];
  double mu;                    /* magnetic_permeability */
  double ep;                    /* electric_permittivity */
  double sigma;                 /* electrical_conductivity (scalar form) */
  double rho;                   /* charge_density (scalar form) */
  double chi_b, chi_d;          /* magnetic / electric cleaning coefficients */
  const void* sigma_field;      /* array forms (maxwell.py:71,475 broadcast arrays): sigma and */
  const void* rho_field;        /* rho of every interior cell, [nx3][nx2][nx1] in the state's
                                   dtype; NULL: the scalar value everywhere */
} mfvm_maxwell_params;

typedef struct mfvm_ns_work {   /* scratch for whole steps (all physics) */
  void* U;                      /* the state (ghosted), advanced in place */
  void* S1;                     /* ghosted stage states */
  void* S2;
  void* C;                      /* ghosted closure: NS 7 fields (ghost entries 0),
                                   Maxwell 2 residual fields, heat unused */
  void* kacc;                   /* interior, 5 fields */
  double* partials;             /* >= mfvm_ns_partials_size() doubles */
  double* dts;                  /* device: dt of step s written to dts[s] */
  void* xy;                     /* optional, interior scratch for split stages in the state
                                   dtype: NS f64 5 fields (see mfvm_stage.xy), Maxwell 6
                                   fields; NULL: fused stages */
} mfvm_ns_work;

/* ---- library ------------------------------------------------------------------ */
const char* mfvm_last_error(void);
int32_t mfvm_abi_version(void);

/* Kernel timing with HIP events recorded on the launch stream around every
 * kernel of the given kinds (measurement support; no reference counterpart).
 * mode 1: enable and reset; mode 0: disable; mode 2: wait for the recorded
 * events and return per-kind totals: ms[k] milliseconds and count[k]
 * launches for k = MFVM_TK_*.  Host-thread local, not for concurrent use. */
#define MFVM_TK_STAGE 0
#define MFVM_TK_CLOSURE 1
#define MFVM_TK_BC 2
#define MFVM_TK_OTHER 3
#define MFVM_TK_N 4
int mfvm_timing(int32_t mode, double* ms, int64_t* count);

/* Ghost fill of nvar fields: face ghosts by bc[v][side], then edge and corner
 * averages.  Reads the interior of src, writes the ghost shell of dst
 * (src == dst is the in-place reference behaviour).
 * Replaces src/solver.py:13-61 (apply_bcs) with
 * src/grid/boundary_conditions.py:13-91 (neumann, periodic, reflecting). */
int mfvm_apply_bcs(const mfvm_grid* g, int32_t nvar, const int32_t* bc, int32_t dtype,
                   const void* src, void* dst, void* stream);

/* Copy the interior of src into the interior of dst (ghosts untouched). */
int mfvm_copy_interior(const mfvm_grid* g, int32_t nvar, int32_t dtype,
                       const void* src, void* dst, void* stream);

/* ---- compressible Navier-Stokes (src/physics/navier_stokes.py) -------------- */

/* Closure C = {P, T11, T22, T33, T12, T13, T23} (7 ghosted fields; P on
 * interior and face ghosts, T on the interior only, ghost T left untouched
 * and expected to be 0).  Replaces navier_stokes.py:44-156 (compute_closure). */
int mfvm_ns_closure(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, void* C,
                    void* stream);

/* Right-hand side into an interior buffer; C is scratch for the closure.
 * Replaces navier_stokes.py:588-711 (compute_rhs). */
int mfvm_ns_rhs(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, void* C,
                void* rhs, void* stream);

/* CFL time step of U into dt_out (device scalar).  partials: scratch of
 * mfvm_ns_partials_size() doubles.  Replaces navier_stokes.py:512-585
 * (compute_dt). */
int mfvm_ns_dt(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, double cfl,
               double* partials, double* dt_out, void* stream);

/* Largest grid of mfvm_ns_dt_noalloc, in CFL partial slots (workgroups of the
 * partials kernel, mfvm_ns_partials_size() / 2): 65536 slots = 1 MiB of device
 * memory per GPU, e.g. 512^3 cells with the default tile needs 16384. */
#define MFVM_NS_DT_NOALLOC_MAX_BLOCKS 65536

/* mfvm_ns_dt without the partials argument: the CFL partials go to a fixed
 * device array inside the library (MFVM_NS_DT_NOALLOC_MAX_BLOCKS slots per
 * GPU, kept until the process ends), so the call allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, allocation-free harnesses).
 * Grids above the maximum fail with MFVM_E_ARG.  Calls that overlap on one
 * GPU (several streams or host threads at once) share the array and race;
 * serialise them. */
int mfvm_ns_dt_noalloc(const mfvm_grid* g, const mfvm_ns_params* p, const void* U, double cfl,
                       double* dt_out, void* stream);

/* One fused Runge-Kutta stage: flux divergence of st->U (closure st->C),
 * sources, and the update of st->update, optionally the CFL partials of the
 * result.  Replaces one stage block of src/numerics/time_integrators.py
 * (e.g. rk3_ssp :144-147) including the compute_rhs call. */
int mfvm_ns_stage(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_stage* st,
                  void* stream);

/* One integrator call with a given dt: ghost fill, closure and fused stage
 * per Runge-Kutta stage, state advanced in place in w->U (w->partials and
 * w->dts unused).  Replaces time_integrators.py:10-221 (euler_explicit, rk3,
 * rk3_ssp, rk4) called with navier_stokes.compute_rhs and solver.apply_bcs. */
int mfvm_ns_step(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w,
                 int32_t integrator, double dt, void* stream);

/* nsteps whole steps of the loop body of src/solver.py:137-151 (dt from the
 * current state, one integrator call), dt kept on the device: dts[s] is the
 * dt of step s and dts[nsteps] the dt of the following step (nsteps+1
 * entries). */
int mfvm_ns_advance(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w,
                    int32_t integrator, double cfl, int32_t nsteps, void* stream);

int64_t mfvm_ns_partials_size(const mfvm_grid* g, const mfvm_ns_params* p);

/* Compensated FP32 Navier-Stokes steps (FP32 reconstruction and fluxes,
 * FP64 final-stage accumulation and CFL).  No reference counterpart: the
 * reduced-precision form beside the FP64 drop-in. */
typedef struct mfvm_ns_comp_work {
  void* hi;                     /* ghosted f32 state (rounded values), advanced in place */
  void* lo;                     /* ghosted f32 compensation words */
  void* Sa;                     /* ghosted f32 stage states */
  void* Sb;
  void* C;                      /* ghosted f32 closure, 7 fields; ghost entries must be 0 */
  void* dinc;                   /* interior f32 increment */
  void* kacc;                   /* interior f32 accumulator */
  double* partials;             /* >= mfvm_ns_partials_size() doubles */
  double* dts;                  /* device, nsteps+1 entries */
  void* xy;                     /* optional (nx3 > 1): interior 5-field f32 scratch for
                                   split stages; NULL: fused stages */
} mfvm_ns_comp_work;

int mfvm_ns_comp_advance(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_comp_work* w,
                         int32_t integrator, double cfl, int32_t nsteps, void* stream);

/* ---- batched ensemble of Navier-Stokes instances ---------------------------
 * No counterpart in the reference, which advances one grid per process
 * (src/solver.py:88-151); each member here advances exactly as
 * mfvm_ns_advance / mfvm_ns_comp_advance would advance it alone (same
 * arithmetic, its own CFL dt every step), while every phase of a stage (ghost
 * fill, closure, the fused stage or the split x1/x2 and x3 passes) is one
 * launch over a work list of all members (their tiles; for the split x1/x2
 * pass, the row march's wave tasks), and the steps can be replayed from
 * captured hipGraphs. */

#define MFVM_ENS_ORDER_MEMBER 0      /* work list: member by member */
#define MFVM_ENS_ORDER_COST 1        /* costliest tiles first (x3 planes per tile, then size) */
#define MFVM_ENS_ORDER_INTERLEAVE 2  /* round-robin over the members */

typedef struct mfvm_ens_member {
  mfvm_grid grid;               /* sizes and metric factors of this member */
  int32_t bc[5][6];             /* ghost rule per field and side */
  int32_t tile[3];              /* stage tile x1, x2 threads (product 256) and x3 planes
                                   (fused stage and split x3 pass; the split x1/x2 pass is
                                   the row march); 0 = 64, 4, min(nx3, 32) */
  void* U;                      /* ghosted state, advanced in place (compensated: hi, f32) */
  void* S1;                     /* ghosted stage states (compensated: Sa, Sb) */
  void* S2;
  void* C;                      /* ghosted closure, 7 fields; ghost entries must be 0 */
  void* kacc;                   /* interior accumulator, 5 fields */
  void* lo;                     /* compensated only: ghosted f32 compensation words */
  void* dinc;                   /* compensated only: interior f32 increment */
  double* partials;             /* >= 2 * mfvm_ens_member_tiles() doubles */
  double* dts;                  /* device, nsteps+1 entries: dt of every step */
  void* xy;                     /* interior 5-field scratch in the stage dtype: the stage runs
                                   split (x1/x2 pass, then x3 pass) as in mfvm_stage.xy; NULL
                                   in any member: every member takes the fused stage */
} mfvm_ens_member;

typedef struct mfvm_ens_opts {
  int32_t order;                /* MFVM_ENS_ORDER_* */
  int32_t persist;              /* 0: one workgroup per tile; k > 0: k workgroups per CU,
                                   each looping over the tile work list (the row-march
                                   x1/x2 pass always takes one wave per task) */
  int32_t graph;                /* > 0: steps per captured hipGraph (stream must not be the
                                   null stream); 0: direct launches */
  int32_t remap;                /* XCD-aware work order */
} mfvm_ens_opts;

/* Stage tiles of one member (its partials need twice as many doubles). */
int64_t mfvm_ens_member_tiles(const mfvm_ens_member* m);
/* Bytes of the device table mfvm_ens_advance needs for these members. */
int64_t mfvm_ens_table_bytes(const mfvm_ens_member* m, int32_t n, int32_t integrator);
/* nsteps steps of every member.  p gives the shared scheme and physics
 * (dtype, limiter, riemann, gamma, mu, g; its bc and tile are not used).
 * compensated != 0: FP32 fluxes with FP64-compensated accumulation (dtype
 * f32).  table: caller-owned device buffer of mfvm_ens_table_bytes().
 * Euler is not supported (MFVM_E_UNSUP).  Synchronises the stream before
 * returning when graph > 0. */
int mfvm_ens_advance(const mfvm_ens_member* m, int32_t n, const mfvm_ns_params* p,
                     const mfvm_ens_opts* o, int32_t integrator, int32_t compensated,
                     double cfl, int32_t nsteps, void* table, int64_t table_bytes,
                     void* stream);

/* ---- learned flux-correction closure -----------------------------------------
 * No counterpart in the reference.  A 3-layer 3x3 CNN (4 -> 32 -> 32 -> 8
 * channels, ReLU after layers 1 and 2, periodic padding) maps the normalised
 * primitives (rho, v1, v2, p) of a 2D periodic Cartesian state to
 * corrections dF of the x1-face (channels 0-3) and x2-face (4-7) fluxes of
 * mass, momentum 1, momentum 2 and energy at the faces i-1/2 and j-1/2 of
 * each cell; the corrected right-hand side is k - div(dF), with
 * dF = y * out_scale per field.  Tensors are channels-last
 * [batch][nx2][nx1][C]; activations BF16, accumulation FP32 on MFMA; nx1 must
 * be a multiple of 16.  Weights: FP32 masters [cout][cin][3][3]; the
 * convolutions read the BF16 layouts written by mfvm_cnn_pack. */

#define MFVM_CNN_FWD_RELU 0    /* y = relu(conv(x) + b), BF16 */
#define MFVM_CNN_FWD_LINEAR 1  /* y = conv(x) + b, FP32 */
#define MFVM_CNN_BWD_DATA 2    /* y = conv_transpose(x) * (mask > 0), BF16 */

/* One convolution; (cin, cout) in {(4,32), (32,32), (32,8)} forward and
 * {(32,32), (8,32)} for BWD_DATA (the transposed layers: x has cin = the
 * layer's cout channels, w the backward layout, mask the layer input). */
int mfvm_cnn_conv(int32_t mode, int32_t cin, int32_t cout, int32_t batch, int32_t n1, int32_t n2,
                  const void* x, const void* w, const float* bias, const void* mask, void* y,
                  void* stream);
/* Accumulates weight / bias gradients of one layer into gw, gb (FP32). */
int mfvm_cnn_wgrad(int32_t cin, int32_t cout, int32_t batch, int32_t n1, int32_t n2, const void* x,
                   const void* gy, float* gw, float* gb, void* stream);
/* FP32 master weights -> BF16 forward layout wf [9][round16(cout)][cin] and
 * backward layout wb [9][round16(cin)][cout]. */
int mfvm_cnn_pack(int32_t cin, int32_t cout, const float* w, void* wf, void* wb, void* stream);
/* One Adam step (step number t >= 1) on n FP32 parameters. */
int mfvm_cnn_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1,
                  float b2, float eps, int32_t t, void* stream);
/* Network input from `batch` ghosted 2D states spaced batch_stride elements:
 * x = ((rho, v1, v2, p) - shift) * scale, BF16 channels-last. */
int mfvm_cnn_features(const mfvm_grid* g, int32_t dtype, double gamma, int32_t batch,
                      int64_t batch_stride, const void* U, const float* shift, const float* scale,
                      void* x, void* stream);
/* Training loss sum_f wf[f] (kc_f - target_f)^2 averaged over the cells,
 * added to *loss; gy = dloss/dy (BF16).  target: [batch][nx2][nx1][4] FP32. */
int mfvm_cnn_loss(int32_t batch, int32_t n1, int32_t n2, double dx1, double dx2, const float* y,
                  const float* target, const float* out_scale, const float* wf, void* gy,
                  float* loss, void* stream);

typedef struct mfvm_cnn {
  const void* wf[3];            /* BF16 forward weights (mfvm_cnn_pack) */
  const float* b[3];            /* FP32 biases: 32, 32, 8 */
  const float* in_shift;        /* 4 floats (device) */
  const float* in_scale;        /* 4 floats (device) */
  const float* out_scale;       /* 4 floats (device) */
} mfvm_cnn;

typedef struct mfvm_closure_work {
  void* x;                      /* unused (the first layer forms its input features from the
                                   state); kept for layout compatibility, may be NULL */
  void* a1;                     /* BF16 nx1*nx2*32 */
  void* a2;                     /* BF16 nx1*nx2*32 */
  float* y;                     /* FP32 nx1*nx2*8 */
  void* T;                      /* ghosted 5-field state scratch */
  int* flag;                    /* 3 ints: fallback flag of each stage of the current step */
  int* counts;                  /* 3 ints: stages (1st, 2nd, 3rd) that fell back, accumulated */
  double* dt_cur;               /* 1 double */
  double* t_cur;                /* 1 double: time reached (from 0 at the first call's start) */
  int* step;                    /* 1 int: dts written so far */
  int32_t layered;              /* 0: the network and the correction of a stage in one launch
                                   (a1, a2, y unused, may be NULL); 1: one launch per layer */
} mfvm_closure_work;

/* nsteps rk3_ssp steps (time_integrators.py:143-159 with
 * navier_stokes.compute_rhs) whose every stage adds the network's correction
 * -div(dF) to the right-hand side; a stage whose corrected state has rho <= 0
 * or p <= 0 in any cell keeps the plain result instead (counted in
 * cw->counts).  net == NULL: the plain scheme.  dt from the CFL rule on every
 * new state, clipped so that the time never passes t_end (pass INFINITY for
 * the reference's unclipped loop, src/solver.py:137-151); steps after t_end
 * have dt = 0 and leave the state unchanged.  w->dts gets nsteps+1 entries.
 * 2D (nx3 == 1), uniform Cartesian, periodic in x1 and x2, float64.
 * graph > 0: steps per captured hipGraph (stream must not be the null
 * stream; synchronises it before returning).  With an all-zero network and
 * t_end = INFINITY the result equals mfvm_ns_advance bit for bit. */
int mfvm_closure_advance(const mfvm_grid* g, const mfvm_ns_params* p, const mfvm_ns_work* w,
                         const mfvm_cnn* net, const mfvm_closure_work* cw, double cfl,
                         double t_end, int32_t nsteps, int32_t graph, void* stream);

/* ---- heat equation (src/physics/heat.py) -------------------------------------- */

/* Right-hand side into an interior buffer.  Replaces heat.py:354-473. */
int mfvm_heat_rhs(const mfvm_grid* g, const mfvm_heat_params* p, const void* U, void* rhs,
                  void* stream);
/* dt (device scalar).  Replaces heat.py:305-351 (compute_dt). */
int mfvm_heat_dt(const mfvm_grid* g, const mfvm_heat_params* p, double cfl, double* dt_out,
                 void* stream);
/* One integrator call (time_integrators.py with heat.compute_rhs). */
int mfvm_heat_step(const mfvm_grid* g, const mfvm_heat_params* p, const mfvm_ns_work* w,
                   int32_t integrator, double dt, void* stream);
/* nsteps steps of solver.py:137-151; dts needs nsteps+1 entries.  The heat
 * dt depends on the parameters and the grid only: it is computed once and
 * written to every entry.  Euler steps alternate w->U and w->S1; w->U ends as
 * the reference leaves U. */
int mfvm_heat_advance(const mfvm_grid* g, const mfvm_heat_params* p, const mfvm_ns_work* w,
                      int32_t integrator, double cfl, int32_t nsteps, void* stream);

/* ---- Maxwell (src/physics/maxwell.py) ------------------------------------------ */

/* Cleaning residuals R = {div B, div D - rho} (2 ghosted fields, face ghosts
 * edge-padded).  Replaces maxwell.py:491-512. */
int mfvm_maxwell_residual(const mfvm_grid* g, const mfvm_maxwell_params* p, const void* U,
                          void* R, void* stream);
/* Right-hand side into an interior buffer; R is scratch.  Replaces
 * maxwell.py:603-729 (compute_rhs). */
int mfvm_maxwell_rhs(const mfvm_grid* g, const mfvm_maxwell_params* p, const void* U, void* R,
                     void* rhs, void* stream);
/* dt (device scalar).  Replaces maxwell.py:533-600 (compute_dt). */
int mfvm_maxwell_dt(const mfvm_grid* g, const mfvm_maxwell_params* p, double cfl,
                    double* dt_out, void* stream);
/* One integrator call (time_integrators.py with maxwell.compute_rhs). */
int mfvm_maxwell_step(const mfvm_grid* g, const mfvm_maxwell_params* p, const mfvm_ns_work* w,
                      int32_t integrator, double dt, void* stream);
/* nsteps steps of solver.py:137-151; dts needs nsteps+1 entries. */
int mfvm_maxwell_advance(const mfvm_grid* g, const mfvm_maxwell_params* p,
                         const mfvm_ns_work* w, int32_t integrator, double cfl,
                         int32_t nsteps, void* stream);

/* ---- diagnostics ---------------------------------------------------------------
 * No counterpart in the reference, which has no run-time monitor: conserved
 * totals, extrema, a positivity monitor and errors against a second state,
 * reduced in float64 over the interior cells of one ghosted SoA state, for
 * any physics (nvar 1..6).  The value of a cell is double(U) (+ double(lo)
 * for a compensated FP32 state).  out (device, MFVM_DIAG_NOUT(nvar) doubles):
 *   out[q]             sum over cells of U_q * V (the conserved totals)
 *   out[nvar + q]      min U_q                 out[2*nvar + q]  max U_q
 *   out[3*nvar + q]    sum |U_q - ref_q| * V / sum V (volume-weighted L1)
 *   out[4*nvar + q]    max |U_q - ref_q|  (both 0 when ref is NULL)
 *   out[5*nvar]        sum V
 *   out[5*nvar + 1]    gamma > 0 (Navier-Stokes, nvar 5): min of the ideal-gas
 *                      pressure of navier_stokes.py:94-100; else 0
 *   out[5*nvar + 2]    gamma > 0: number of cells with !(rho > 0) or
 *                      !(P > 0) (NaN counts); else 0
 * ref: NULL or a float64 ghosted state of the same shape.  partials:
 * mfvm_diag_partials_size() doubles of scratch.  Min and max propagate NaN.
 * The reduction order is fixed, so equal inputs give equal outputs. */
#define MFVM_DIAG_NOUT(nvar) (5 * (nvar) + 3)
int64_t mfvm_diag_partials_size(const mfvm_grid* g, int32_t nvar);
int mfvm_diagnostics(const mfvm_grid* g, int32_t nvar, int32_t dtype, const void* U, const void* lo,
                     const void* ref, double gamma, double* partials, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MFVM_H */
