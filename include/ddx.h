/*
 * ddx.h -- C ABI of libddx.so, the MI355X (gfx950) pseudospectral step.
 *
 * Replaces, behind dedalus-1.0's own class interfaces, the numerical work of
 *   - FourierRepresentation.fwd_np / rev_np / fwd_fftw / rev_fftw and the
 *     rPlan FFTW wrapper       (dedalus/data_objects/representations.py:318-357,
 *                               dedalus/utils/fftw/_fftw.pyx:246-309)
 *   - the 2/3 dealias kernels  (dedalus/data_objects/dealias_cy_2d.pyx:13-39,
 *                               dealias_cy_3d.pyx:13-46, representations.py:384-408)
 *   - deriv / k2 / projection  (representations.py:419-440, physics.py:180-195,
 *                               physics.py:407-416, physics.py:588-599)
 *   - Physics.RHS              (physics.py:134-149, 527-586, 664-712, 770-819)
 *   - euler / etd1 / etd2rk1 / etd2rk2
 *                              (dedalus/time_stepping/forward_step_cy_2d.pyx:21-120,
 *                               forward_step_cy_3d.pyx:17-124)
 *   - RK2mid.do_advance / RK2trap.do_advance
 *                              (dedalus/time_stepping/time_step.py:275-309, 360-392)
 *
 * Conventions
 *   - Every call returns 0 on success or a negative DDX_ERR_* code; the
 *     message of the last failure on the calling thread is ddx_last_error().
 *   - x-space arrays are float64 (or float32 for DDX_F32 plans) in the
 *     reference layout: (y, x) in 2D, (z, y, x) in 3D, C order.
 *   - "kref" arrays are complex128 in the reference's transposed
 *     half-spectrum layout: (kx, ky) = (Nx/2+1, Ny) in 2D and
 *     (ky, kz, kx) = (Ny, Nz, Nx/2+1) in 3D (representations.py:82-93).
 *   - "kc" arrays are the library's compact k layout: only the modes kept by
 *     the 2/3 rule, shape (Ky, Kz, Kxp) with Kz = 1 in 2D, kx fastest and
 *     padded to Kxp (pad entries are zero).  ddx_plan_kshape() gives the
 *     sizes; ddx_pack / ddx_unpack convert to and from kref.
 *   - All device pointers are plain device addresses (ddx_alloc, or any
 *     other allocation on the same device).  The library never frees
 *     caller memory.  Calls are asynchronous on the plan's stream except
 *     ddx_h2d / ddx_d2h / ddx_sync, which return after the stream drains.
 *   - One host thread per plan.
 */
#ifndef DDX_H
#define DDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddx_plan ddx_plan;

enum {
  DDX_OK = 0,
  DDX_ERR_ARG = -1,         /* bad argument (shape, pointer, enum) */
  DDX_ERR_HIP = -2,         /* HIP runtime error */
  DDX_ERR_UNSUPPORTED = -3, /* valid request this build does not handle */
  DDX_ERR_NOMEM = -4,       /* device allocation failed */
  DDX_ERR_NODEV = -5        /* no HIP device, or ddx_init not called */
};

enum { DDX_F64 = 0, DDX_F32 = 1 };
enum { DDX_HYDRO = 0, DDX_BOUSSINESQ = 1, DDX_MHD = 2 };
enum { DDX_RK2MID = 0, DDX_RK2TRAP = 1 };
enum { DDX_EULER = 0, DDX_ETD1 = 1, DDX_ETD2RK1 = 2, DDX_ETD2RK2 = 3 };

/* Physics parameters (RHS.parameters of physics.py:462-465, 642-646,
 * 753-754).  Coefficients equal to 0 mean "no integrating factor"
 * (the reference's integrating_factor = None, physics.py:510-513). */
typedef struct ddx_params {
  int physics;          /* DDX_HYDRO / DDX_BOUSSINESQ / DDX_MHD */
  int viscosity_order;  /* 'viscosity_order' */
  double nu;            /* 'nu' */
  double kappa;         /* 'kappa' (Boussinesq) */
  double eta;           /* 'eta' (MHD) */
  double g, alpha_t, beta; /* Boussinesq 'g', 'alpha_t', 'beta' */
  int bdir;             /* 'boussinesq_direction': 0 = x, 1 = y, 2 = z */
  double rho0;          /* MHD 'rho0' */
} ddx_params;

/* ---- lifecycle ---------------------------------------------------------- */

/* Select the HIP device; must precede every other call. */
int ddx_init(int device);
/* Message of the last failure on this thread ("" if none). */
const char* ddx_last_error(void);
/* ABI version (bumped on any signature change). */
int ddx_version(void);

/* New plan for one grid: ndim 2 or 3, xshape in reference order ((y, x) or
 * (z, y, x)), box lengths in the same order, precision DDX_F64 / DDX_F32.
 * Replaces FourierRepresentation.set_fft / create_fftw_plans
 * (representations.py:290-316) and _setup_k (representations.py:204-233). */
int ddx_plan_create(int ndim, const int64_t* xshape, const double* length, int precision,
                    ddx_plan** plan);
/* Dealiasing rules of FourierRepresentation.set_dealiasing
 * (representations.py:359-382):
 *   DDX_DEALIAS_23         '2/3' and '2/3 cython': |k_i| < (2/3) kny_i on
 *                          every axis (representations.py:384-408);
 *   DDX_DEALIAS_SPHERICAL  '2/3 spherical': |k| < (2/3) min(kny)
 *                          (representations.py:410-416); the compact layout
 *                          is the cube |k_i| < (2/3) min(kny) that holds the
 *                          sphere, and the modes outside the sphere are
 *                          zeroed wherever the reference dealiases: in place
 *                          on the input of a backward transform and of a
 *                          right-hand side, and on the output of a forward
 *                          transform (before forcing is added);
 *   DDX_DEALIAS_NONE       'None' / None / 0: every mode but the Nyquist
 *                          planes (zero_nyquist, representations.py:442-455),
 *                          products not dealiased.  ddx_rhs and ddx_step
 *                          refuse such plans (DDX_ERR_UNSUPPORTED): their
 *                          rotational form of the nonlinear terms equals the
 *                          reference's convective form only for dealiased
 *                          products; the per-operation kernels serve them. */
enum { DDX_DEALIAS_23 = 0, DDX_DEALIAS_SPHERICAL = 1, DDX_DEALIAS_NONE = 2 };
/* ddx_plan_create with a dealiasing rule (ddx_plan_create uses DDX_DEALIAS_23). */
int ddx_plan_create_ex(int ndim, const int64_t* xshape, const double* length, int precision,
                       int dealias, ddx_plan** plan);
/* New 2D plan for nbatch independent members on one ny x nx grid (an
 * ensemble).  No reference counterpart: the reference advances one state
 * per process; this runs many coarse or ensemble members in one set of
 * launches.  Per member everything is as for ddx_plan_create(2, {ny, nx}).
 * Layouts: x space (nbatch, ny, nx); compact k (Ky, nbatch, Kxp), i.e.
 * the member takes the place of kz; reference k = nbatch stacked (kx, ky)
 * arrays.  Every call acts on all members at once; ddx_xreduce and
 * ddx_kreduce reduce over all members together. */
int ddx_plan_create_batched(int nbatch, int64_t ny, int64_t nx, const double* length,
                            int precision, ddx_plan** plan);
int ddx_plan_destroy(ddx_plan* plan);
/* kref[3] = reference k shape ((Ny, Nz, Nx/2+1) in 3D, (Nx/2+1, Ny, nbatch)
 * in 2D with nbatch = 1 for a plain plan), kc[3] = (Ky, Kz, Kxp) with Kz =
 * nbatch in 2D, cut[3] = kept |index| per axis (x, y, z); *nkc = Ky*Kz*Kxp. */
int ddx_plan_kshape(const ddx_plan* plan, int64_t* kref, int64_t* kc, int64_t* cut, int64_t* nkc);
/* Wavenumbers of axis (0 = x, 1 = y, 2 = z) in compact order (n entries). */
int ddx_plan_ktable(const ddx_plan* plan, int axis, double* out, int64_t n);
/* Run subsequent work of this plan on an existing HIP stream (NULL = the
 * plan's own stream). */
int ddx_plan_set_stream(ddx_plan* plan, void* stream);
/* ddx_step replays one captured HIP graph per step (on = 1, the default
 * unless DDX_NO_GRAPH=1) or launches its kernels one by one (on = 0).  The
 * graph is recaptured when the buffers, parameters or dt change and the new
 * step repeats (see ddx_step).  No reference counterpart: launch overhead
 * matters on small grids. */
int ddx_plan_set_graph(ddx_plan* plan, int on);
/* Kernel choices of a plan that do not change its results, for autotuning:
 * every choice is bit-identical, except that FP64 16-column tiles exchange
 * through a buffer of reals (other instruction sequences) and agree with
 * 8-column ones to rounding, and the radix variant (below) agrees to
 * rounding.  value -1 restores the built-in rule.
 *   DDX_TUNE_WIDE_TILES: 16-column tiles of the 256-point column passes
 *     (0 off, 1 on);
 *   DDX_TUNE_XCD_ORDER: workgroup order of the spectral column passes over
 *     the 8 XCDs (0 plain grid, 1 the fields of a (slab, tile) together,
 *     2 whole slabs together);
 *   DDX_TUNE_FUSED_YEPI: forward y pass inside the epilogue for 64-point
 *     y axes of 2D plans, batched ones included (0 off, 1 on);
 *   DDX_TUNE_RADIX: radix plan of the size-specialised row and column
 *     transforms (0 the built-in plans; 1 rows with more points per thread
 *     and fewer passes, e.g. 256 = 8.8.4 instead of 4.4.4.4, and columns
 *     with fewer points per thread and more passes, e.g. 256 = 8.8.4
 *     instead of 16.16).  Results agree with the built-in plans to rounding
 *     (other butterflies), not bit for bit; radix 1 runs the forward y pass
 *     and the epilogue as two launches.
 * The step graph is recaptured at the next ddx_step.  No reference
 * counterpart: the choices belong to this implementation. */
enum { DDX_TUNE_WIDE_TILES = 0, DDX_TUNE_XCD_ORDER = 1, DDX_TUNE_FUSED_YEPI = 2, DDX_TUNE_RADIX = 3 };
int ddx_plan_set_tuning(ddx_plan* plan, int knob, int value);

/* ---- memory --------------------------------------------------------------- */

int ddx_alloc(size_t bytes, void** ptr);
int ddx_free(void* ptr);
int ddx_h2d(ddx_plan* plan, void* dst, const void* src, size_t bytes);
int ddx_d2h(ddx_plan* plan, void* dst, const void* src, size_t bytes);
/* Zero device memory with a fill kernel (safe inside captured graphs). */
int ddx_zero(ddx_plan* plan, void* dst, size_t bytes);
/* Device-to-device copy with a streaming kernel (16-byte vector loads when
 * both pointers and the size are 16-byte aligned).  Backs the drop-in's
 * buffer-preserving assignment between fields (representations.py:144-166). */
int ddx_copy(ddx_plan* plan, void* dst, const void* src, size_t bytes);
int ddx_sync(ddx_plan* plan);

/* ---- layout ------------------------------------------------------------- */

/* kref -> kc; modes the 2/3 rule removes are dropped (the dealias of
 * representations.py:384-397 applied on entry). */
int ddx_pack(ddx_plan* plan, const void* kref, void* kc);
/* kc -> kref; removed modes are written as zero. */
int ddx_unpack(ddx_plan* plan, const void* kc, void* kref);

/* ---- transforms (representations.py:335-357) ----------------------------- */

/* x -> kc: r2c, divided by N = prod(xshape), dealiased (FourierRepresentation.forward). */
int ddx_fwd(ddx_plan* plan, const void* x, void* kc);
/* kc -> x: unnormalised c2r of the dealiased data (FourierRepresentation.backward);
 * the kx = 0 plane contributes its Hermitian part, as irfftn does. */
int ddx_bwd(ddx_plan* plan, const void* kc, void* x);

/* ddx_fwd / ddx_bwd of nf <= 16 fields in one call: each pass is one launch
 * over all the fields.  The results equal nf single calls bit for bit for
 * x extents 64..512 (size-specialised rows); other extents pair two fields
 * per complex row transform, equal to rounding.  The
 * reference transforms one component per call (representations.py:335-357);
 * this batches the components a caller needs together, for small grids where
 * launches, not bytes, set the time. */
int ddx_fwd_many(ddx_plan* plan, int nf, const void* const* x, void* const* kc);
int ddx_bwd_many(ddx_plan* plan, int nf, const void* const* kc, void* const* x);
/* Backward transforms of linear combinations of nsrc compact k arrays, the
 * derived fields the reference forms with deriv() and then transforms
 * (e.g. the vorticity of curlX, physics.py:374-405, representations.py:
 * 419-425 then 347-357), in one pass: output q (< nout) is
 *   bwd( sum_{t < nterm[q]} coef[2q+t] * M(dim[2q+t]) * kc[src[2q+t]] )
 * with M = identity (dim -1) or i k_x / i k_y / i k_z (dim 0 / 1 / 2) and
 * nterm[q] 1 or 2.  The combination is formed in the first column pass, as
 * the fused right-hand side forms its derived fields. */
int ddx_bwd_combine(ddx_plan* plan, int nsrc, const void* const* kc, int nout, const int* nterm,
                    const int* src, const int* dim, const double* coef, void* const* x);

/* ---- per-operation kernels ------------------------------------------------- */

/* The plan's dealiasing rule in place on a compact k array (the
 * reference's dealias(), representations.py:384-455): zeroes the modes
 * outside the sphere for DDX_DEALIAS_SPHERICAL; for the other rules the
 * compact layout already holds only the kept modes and nothing is done. */
int ddx_dealias(ddx_plan* plan, void* kc);
/* out = i k_dim * kc, dim 0 = x, 1 = y, 2 = z (representations.py:419-425). */
int ddx_deriv(ddx_plan* plan, const void* kc, void* out, int dim);
/* In-place solenoidal projection F <- F - k (k.F)/|k|^2 of ndim components
 * (physics.py:588-599 = fields.py:169-195). */
int ddx_project(ddx_plan* plan, void* const* comps);
/* One ETD / Euler sweep (forward_step_cy_{2,3}d.pyx): kind DDX_EULER,
 * DDX_ETD1, DDX_ETD2RK1, DDX_ETD2RK2.  The integrating factor coeff*|k|^(2*order)
 * is recomputed from the wavenumbers (coeff = 0 is the reference's
 * integrating_factor = None).  out may alias y.  F2 is ignored for EULER
 * and ETD1.  The 2D plans use the Taylor-series phi0 of forward_step_cy_2d.pyx:54-56,
 * 3D plans exp() as forward_step_cy_3d.pyx:54-56. */
int ddx_update(ddx_plan* plan, int kind, const void* y, void* out, const void* F1,
               const void* F2, double coeff, int order, double dt);
/* The same sweep with an integrating factor given per mode: intfactor is a
 * device array of FP64 values in the compact k layout (the component's
 * integrating_factor array, positive for dissipation), and Z = (-IF) dt as
 * the reference's sweeps form it from the negated array they are passed
 * (time_step.py:289-304, 373-388; forward_step_cy_3d.pyx:34-60).  For
 * integrating factors that are not coeff*|k|^(2*order). */
int ddx_update_if(ddx_plan* plan, int kind, const void* y, void* out, const void* F1,
                  const void* F2, const double* intfactor, double dt);

/* Linear combination on compact k arrays (the k-space operators of
 * physics.py:160-195, 266-307, 374-416 and fields.py:169-195):
 *   out (+)= sum_t coef[t] * M_t(k) * src[t],   t < nterm <= 4,
 * M = identity (op -1), i k_x / i k_y / i k_z (op 0/1/2), 1/|k|^2 with
 * |k|^2 = 0 replaced by 1 (op 3).  accumulate = 0 overwrites out. */
int ddx_kcombine(ddx_plan* plan, void* out, int nterm, const void* const* src,
                 const double* coef, const int* op, int accumulate);
/* Pointwise products on x-space arrays (XgradY / XcrossY / XdotY of
 * physics.py:197-372): out (+)= sum_t coef[t] * a[t] * b[t], b[t] may be NULL. */
int ddx_xcombine(ddx_plan* plan, void* out, int nterm, const void* const* a,
                 const void* const* b, const double* coef, int accumulate);
/* Reduction of an x-space array: op 0 = max(x^2) (fields.py:153-157),
 * 1 = sum(x), 2 = sum(x^2).  Synchronous; result in *out. */
int ddx_xreduce(ddx_plan* plan, const void* x, int op, double* out);
/* Hermitian-weighted sum over a compact k array (weight 1 on kx = 0, 2
 * elsewhere; volume_average.py:71-96): op 0 = sum |k0|^2, op 1 = sum
 * |i kx k1 - i ky k0|^2 (2D vorticity of the pair), op 2 = sum Re(k0).
 * Synchronous; result in *out. */
int ddx_kreduce(ddx_plan* plan, const void* k0, const void* k1, int op, double* out);
/* ddx_kreduce per member of a 2D plan (batched or not): out is a device
 * array of nbatch doubles (1 for a plain plan), written asynchronously on the
 * plan's stream, member b summed in a fixed order.  No reference counterpart:
 * the per-member diagnostics of an ensemble (volume_average.py:109-118,
 * 191-200 for each member) without a host round trip per member. */
int ddx_kreduce_members(ddx_plan* plan, const void* k0, const void* k1, int op, double* out);

/* ---- fused right-hand side and step ---------------------------------------- */

/* Physics.RHS on compact state -> compact deriv.  State component order:
 * HYDRO u[ndim]; BOUSSINESQ u[ndim], T; MHD u[ndim], B[ndim].  forcing is
 * NULL or an array of the same length whose non-NULL entries (compact k)
 * are added to the tendency before the projection (velocity) or after
 * stratification (T), as physics.py:569-572 and 710-712 do. */
int ddx_rhs(ddx_plan* plan, const ddx_params* params, const void* const* state,
            void* const* deriv, const void* const* forcing);
/* nsteps RK2mid / RK2trap steps of size dt on compact state in place.
 * stage, deriv1, deriv2 are caller buffers with the state's component
 * count (the reference integrator's data2 / deriv1 / deriv2,
 * time_step.py:262-264); stage is unused by RK2trap and may be NULL there.
 * After the call, as after the reference's last do_advance: deriv1 holds
 * F(y) of the last step's start, stage (RK2mid) the midpoint state, and
 * deriv2 the second right-hand side (F(stage) for RK2mid, F of the ETD1
 * state for RK2trap; time_step.py:297, 381).  deriv2 may be NULL: the
 * second right-hand side is then not stored, which saves one write of every
 * component per step.  forcing is as in ddx_rhs and
 * is added at every right-hand side of the nsteps steps: the reference's
 * set_velocity_forcing / set_thermal_forcing hook with a function that does
 * not depend on the state (physics.py:524-525, 569-572).
 * Launches: a call with nsteps > 1, or a single step repeating the previous
 * call's buffers, parameters and dt, replays one captured HIP graph per
 * step; any other single step (e.g. dt changing every step under a CFL
 * condition) launches its kernels directly. */
int ddx_step(ddx_plan* plan, int integrator, const ddx_params* params, void* const* state,
             void* const* stage, void* const* deriv1, void* const* deriv2,
             const void* const* forcing, int nsteps, double dt);

/* ddx_step with the forcing given in x space as well as (or instead of) in
 * k space: xforcing (NULL, or one entry per state component, NULL entries
 * allowed; velocity components and the Boussinesq T only) holds x-space
 * arrays of the plan's x layout, added to the right-hand side's x-space
 * products before their forward transform, so F = forward(f) (1/N,
 * dealiased) joins every right-hand side as the k-space forcing does.  The
 * reference's forcing hook (physics.py:524-525, 569-572) returns k-space
 * arrays; this saves the forward transform of a forcing computed in x space
 * (the learned closure's tendency), which otherwise runs once per step. */
int ddx_step_xforced(ddx_plan* plan, int integrator, const ddx_params* prm, void* const* state,
                     void* const* stage, void* const* deriv1, void* const* deriv2,
                     const void* const* forcing, const void* const* xforcing, int nsteps,
                     double dt);
/* One ddx_step_xforced step (2D grids) whose first right-hand side takes its
 * x-space inputs from the caller instead of transforming the state: xin[q]
 * (one per input of the physics' x-row products, in this order: u_x, u_y,
 * omega = d_x u_y - d_y u_x; then T, d_x T, d_y T (Boussinesq) or B_x, B_y,
 * J = d_x B_y - d_y B_x (MHD)) in the plan's x layout, as ddx_bwd_combine
 * writes them from the same state.  A caller that already holds these fields
 * (the learned closure computes them for its network) saves the backward
 * column pass of that right-hand side.  The results equal ddx_step_xforced's
 * to rounding (a backward transform pairs rows where the step pairs fields).
 * Plans where the size-specialised x-row kernel does not apply (x extents
 * other than 64, 128, 256, 512; spherical rule) ignore xin.  No counterpart
 * in the reference (its right-hand side always transforms the state,
 * physics.py:527-586). */
int ddx_step_xin(ddx_plan* plan, int integrator, const ddx_params* prm, void* const* state,
                 void* const* stage, void* const* deriv1, void* const* deriv2,
                 const void* const* forcing, const void* const* xforcing, const void* const* xin,
                 double dt);
/* ddx_step that allocates nothing: no hipMalloc, hipFree or other allocating
 * call, and no internal graph capture (the kernels are launched directly on
 * the plan's stream, so a caller may capture the call in a graph of its own).
 * No counterpart in the reference; it is for callers that cannot allocate:
 * code under HIP graph capture and tight serving loops.  The right-hand
 * side's workspace (and, when the ETD sweeps run unfused and deriv2 is
 * NULL, the second derivative) lives in a static device array of 64 MiB per
 * GPU, held from ddx_init until the process ends.  A grid and physics that
 * need more (e.g. 2D up to 512^2, 3D hydro up to 96^3 in FP64 fit) fail with
 * DDX_ERR_ARG and a message naming the size; use ddx_step for them.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two threads) race.  Arguments and results are those of ddx_step;
 * per-launch profiling is not recorded during the call. */
int ddx_step_noalloc(ddx_plan* plan, int integrator, const ddx_params* params, void* const* state,
                     void* const* stage, void* const* deriv1, void* const* deriv2,
                     const void* const* forcing, int nsteps, double dt);

/* ---- instrumentation (no reference counterpart; replaces the wall-clock
 *      Timer of dedalus/utils/timer.py:29-65 at kernel granularity) ------- */

/* Turn per-launch HIP-event timing on (1) or off (0); clears records. */
int ddx_prof_enable(ddx_plan* plan, int on);
/* Aggregate records by kernel class since the last enable/read: up to
 * max_cat classes, names as 32-byte NUL-terminated strings, launch counts,
 * summed event durations (ms) and summed algorithmic bytes (every valid
 * input element read once and output element written once). */
int ddx_prof_read(ddx_plan* plan, int max_cat, int* ncat, char* names, int64_t* launches,
                  double* ms_total, double* bytes_total);

#ifdef __cplusplus
}
#endif

#endif /* DDX_H */
