/* adfvm.h -- C ABI of libadfvm.so, the MI355X implementation of adFVM's
 * compressible finite-volume step and its discrete adjoint.
 *
 * The reference's operator boundary is the adpy `Function` objects it builds
 * and calls as `outputs = self.map(*inputs)`:
 *   primal step   adFVM/density.py:80-105 (built), adFVM/solver.py:312-323 (called)
 *   adjoint step  apps/adjoint.py:94-126  (built), apps/adjoint.py:268-291 (called)
 *   init/ghosts   adFVM/density.py:64-78, :140-143
 * plus the native helpers those Functions and the mesh rely on:
 *   cmesh.build / buildBeforeWrite   adFVM/cpp/cmesh.cpp:3-270, :427-432
 *   cmesh.computeSensitivity         adFVM/cpp/cmesh.cpp:309-340
 *   cmesh.computeEnergy              adFVM/cpp/cmesh.cpp:342-425
 *   external_init / external_exit    adFVM/cpp/external.cpp:21-51
 *
 * Conventions: plain pointers and sizes.  "device" pointers are HIP device
 * memory (for example a torch tensor's data_ptr()); "host" pointers are CPU
 * memory.  Floating-point state arrays use the context's precision (float
 * when precision == 4, double when 8) and structure-of-arrays layout:
 *   state  u[5][nInternalCells]   rows rho, rhoUx, rhoUy, rhoUz, rhoE
 * Streams are hipStream_t passed as void* (NULL = default stream).  Every
 * function returns ADFVM_OK or an error code; adfvm_last_error() /
 * adfvm_mesh_last_error() give the message.  Calls on one context are not
 * reentrant.
 */
#ifndef ADFVM_H
#define ADFVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADFVM_OK 0
#define ADFVM_ERR_INVALID 1
#define ADFVM_ERR_HIP 2
#define ADFVM_ERR_UNSUPPORTED 3

/* patch topology (adFVM/config.py:388-391) */
#define ADFVM_PATCH_OTHER 0
#define ADFVM_PATCH_CYCLIC 1
#define ADFVM_PATCH_PROCESSOR 2

/* ------------------------------------------------------------------ mesh --- */

typedef struct {
    int32_t nPoints, nFaces, nInternalFaces, nPatches;
    const double* points;          /* host, nPoints x 3 */
    const int32_t* faces;          /* host, nFaces x 4 quad point indices */
    const int32_t* owner;          /* host, nFaces */
    const int32_t* neighbour;      /* host, nInternalFaces */
    const int32_t* patchStartFace; /* host, nPatches (file order) */
    const int32_t* patchNFaces;
    const int32_t* patchType;      /* ADFVM_PATCH_* */
    const int32_t* patchNeighbour; /* cyclic partner patch index, else -1 */
} adfvm_mesh_input;

typedef struct {
    int32_t nInternalFaces, nFaces, nBoundaryFaces, nInternalCells, nGhostCells, nCells;
} adfvm_mesh_sizes_t;

/* caller-allocated host outputs; `cells` and `cellNeighboursMatOp` may be NULL */
typedef struct {
    int32_t* cellFaces;          /* nInternalCells x 6 */
    int32_t* cells;              /* nInternalCells x 8 */
    int32_t* neighbour;          /* nFaces, ghost-extended */
    int32_t* cellOwner;          /* nInternalCells x 6 */
    int32_t* cellNeighbours;     /* nInternalCells x 6, ghost indices for boundary faces */
    int32_t* cellNeighboursMatOp;/* nInternalCells x 6, -1 for boundary faces */
    double* normals;             /* nFaces x 3 */
    double* faceCentres;         /* nFaces x 3 */
    double* areas;               /* nFaces */
    double* cellCentres;         /* nCells x 3 (ghost centres included) */
    double* volumes;             /* nInternalCells */
    double* deltas;              /* nFaces */
    double* deltasUnit;          /* nFaces x 3 */
    double* weights;             /* nFaces */
    double* linearWeights;       /* nFaces x 2 */
    double* quadraticWeights;    /* nFaces x 2 x 3 */
    double* volumesL;            /* nFaces */
    double* volumesR;            /* nInternalFaces */
    int32_t* nLocalCells;        /* 1 */
} adfvm_mesh_output;

/* Sizes, adFVM/mesh.py:517-523 (Mesh.populateSizes). Host only. */
int adfvm_mesh_sizes(const adfvm_mesh_input* in, adfvm_mesh_sizes_t* out);

/* Index construction and FP64 geometry in one call; replaces
 * cmesh.buildBeforeWrite (adFVM/cpp/cmesh.cpp:3-63), cmesh.build
 * (cmesh.cpp:65-270) with its createGhostCells callback (adFVM/mesh.py:758-819)
 * and volumesL/R (mesh.py:108-109).  Host only; no GPU needed. */
int adfvm_mesh_build(const adfvm_mesh_input* in, adfvm_mesh_output* out);
const char* adfvm_mesh_last_error(void);

/* ---------------------------------------------------------------- solver --- */

/* per-field ghost rules, adFVM/BCs.py */
#define ADFVM_BC_CYCLIC 0        /* BCs.py:86-95 */
#define ADFVM_BC_ZERO_GRADIENT 1 /* BCs.py:97-104; also empty, inletOutlet (:135-137) */
#define ADFVM_BC_SYMMETRY 2      /* BCs.py:120-133; symmetryPlane, slip */
#define ADFVM_BC_FIXED_VALUE 3   /* BCs.py:111-118 */
#define ADFVM_BC_CALCULATED 4    /* BCs.py:74-84: ghost keeps its value */

/* patch flux treatment, adFVM/density.py:391-400 */
#define ADFVM_FLUX_COUPLED 0        /* _coupledFlux, cyclic */
#define ADFVM_FLUX_CHARACTERISTIC 1 /* _characteristicFlux */
#define ADFVM_FLUX_BOUNDARY 2       /* _boundaryFlux */

/* characteristic ghost rules, adFVM/BCs.py:140-184 */
#define ADFVM_CBC_NONE 0
#define ADFVM_CBC_UPT 1
#define ADFVM_CBC_TOTAL_PT 2

/* boundary-value arrays of a patch (the value pointers of adfvm_patch_desc) */
#define ADFVM_BCV_U 0          /* valueU */
#define ADFVM_BCV_T 1          /* valueT */
#define ADFVM_BCV_P 2          /* valuep */
#define ADFVM_BCV_TT 3         /* Tt */
#define ADFVM_BCV_PT 4         /* pt */
#define ADFVM_BCV_DIRECTION 5  /* direction */

typedef struct {
    int32_t startFace, nFaces;
    int32_t fluxKind;        /* ADFVM_FLUX_* */
    int32_t neighbourPatch;  /* cyclic partner index into the patch array, else -1 */
    int32_t bcU, bcT, bcp;   /* ADFVM_BC_* */
    int32_t cbc;             /* ADFVM_CBC_* (overrides bcU/bcT/bcp) */
    /* host FP64 value arrays, nullable: fixedValue/calculated/CBC_UPT values
     * (U nFaces x 3, T and p nFaces); CBC_TOTAL_PT Tt, pt (nFaces) and
     * direction (nFaces x 3, NULL = face normal, BCs.py:172-175) */
    const double* valueU;
    const double* valueT;
    const double* valuep;
    const double* Tt;
    const double* pt;
    const double* direction;
} adfvm_patch_desc;

#define ADFVM_RIEMANN_ROE 0  /* adFVM/riemann.py:21-70 */
#define ADFVM_RIEMANN_LF 1   /* adFVM/riemann.py:6-18 */
#define ADFVM_MU_CONSTANT 0
#define ADFVM_MU_SUTHERLAND 1 /* muA*T^1.5/(T+muTs), adFVM/density.py:27 */
#define ADFVM_RK_SSPRK3 0     /* adFVM/timestep.py:17-25 */
#define ADFVM_RK_EULER 1      /* adFVM/timestep.py:9-13 */
#define ADFVM_OBJ_NONE 0
#define ADFVM_OBJ_KINETIC_ENERGY 1 /* sum_c 0.5 |rhoU|^2/rho V_c */
#define ADFVM_OBJ_DRAG 2           /* templates/cylinder_test.py:9-37 on one patch */
#define ADFVM_OBJSTATE_STAGE1 0    /* objective on the first intermediate stage (density.py:412-413) */
#define ADFVM_OBJSTATE_START 1     /* objective on the step's start state u_n */

typedef struct {
    int32_t precision;             /* 4 (float kernels) or 8 (double) */
    double gamma, Cp, Pr;          /* adFVM/density.py:25-28 */
    int32_t muModel;
    double mu0, muA, muTs;
    int32_t riemann, boundaryRiemann;
    double small;                  /* config.SMALL, adFVM/config.py:170-178 */
    int32_t integrator;            /* ADFVM_RK_* */
    int32_t objective;             /* ADFVM_OBJ_* */
    int32_t objectivePatch;        /* patch index for ADFVM_OBJ_DRAG */
    int32_t limiter;               /* 0: interp.secondOrder as in the reference (unlimited);
                                      1: reconstructed values clamped between the two cell
                                      values (an addition; needed for Sod-type jumps) */
} adfvm_physics;

typedef struct {
    adfvm_mesh_sizes_t sizes;
    /* host arrays as produced by adfvm_mesh_build */
    const int32_t* owner;          /* nFaces */
    const int32_t* neighbour;      /* nFaces (ghost-extended) */
    const int32_t* cellFaces;      /* nInternalCells x 6 */
    const int32_t* cellOwner;      /* nInternalCells x 6 */
    const int32_t* cellNeighbours; /* nInternalCells x 6 */
    const double* areas;
    const double* normals;
    const double* weights;
    const double* deltas;
    const double* deltasUnit;
    const double* linearWeights;
    const double* quadraticWeights;
    const double* volumes;
    int32_t nPatches;
    const adfvm_patch_desc* patches;
} adfvm_solver_desc;

typedef struct adfvm_ctx adfvm_ctx;

/* Uploads the mesh to the current HIP device in structure-of-arrays form,
 * builds the cell->face slot tables with cyclic neighbours resolved to
 * internal cells, and allocates the step's work buffers.  Replaces
 * Function.initialize -> external_init (adFVM/cpp/external.cpp:21-42). */
int adfvm_ctx_create(const adfvm_solver_desc* desc, const adfvm_physics* phys, adfvm_ctx** out);
void adfvm_ctx_destroy(adfvm_ctx* ctx);
const char* adfvm_last_error(const adfvm_ctx* ctx);
int adfvm_ctx_precision(const adfvm_ctx* ctx);
/* SHA-256 (hex) of the sources the library was built from; the Python
 * loader refuses a library whose hash differs from the sources beside it.
 * No counterpart in the reference (adpy recompiles generated code). */
const char* adfvm_source_hash(void);

/* Cell segments: an ensemble of independent meshes concatenated into one
 * context, mesh k owning cells [cellStart[k], cellStart[k+1]).  The
 * objective is then evaluated per segment: obj_out of adfvm_primal_step and
 * adfvm_objective receive nSeg doubles, each summed exactly as a context
 * holding that mesh alone would sum it.  cellStart: host, nSeg+1 entries
 * from 0 to nInternalCells.  nSeg = 1 restores one objective.  No
 * counterpart in the reference (one mesh per process there); it serves the
 * batched ensemble of perturbed instances. */
int adfvm_set_segments(adfvm_ctx* ctx, int32_t nSeg, const int64_t* cellStart);

/* Replaces one boundary-value array of patch `patch` (index into the
 * creation-time patch array) on the device: slot ADFVM_BCV_*, values host
 * FP64, n = nFaces x components of that array (3 for U and direction, else
 * 1).  The array must have been given at creation.  Ordered on `stream`
 * (returns after the copy).  Replaces the reference's in-place update of a
 * patch's input array, patch.inputs[k][1][:] += value (BCs.py:45-57,
 * Solver.run's doPerturb solver.py:262-270): there those arrays are passed
 * with every Function call, here they stay on the device. */
int adfvm_set_bc_values(adfvm_ctx* ctx, int32_t patch, int32_t slot, const double* values, int64_t n,
                        void* stream);
int adfvm_segments(const adfvm_ctx* ctx);

/* The state a step's objective is evaluated on: ADFVM_OBJSTATE_STAGE1 (the
 * default) is the reference's hook, RCF.equation evaluating the case's
 * objective on the first intermediate RK stage (density.py:354-357,412-413);
 * ADFVM_OBJSTATE_START evaluates it on the step's start state u_n, so that
 * the time-integrated objective is a sum over the trajectory's states
 * u_0 .. u_{N-1}.  Applies to obj_out of adfvm_primal_step, to the
 * objective's adjoint in adfvm_adjoint_step and to the stage at which
 * adfvm_adjoint_step_cotangent adds obj_grad.  dtc stays on the first
 * intermediate stage.  No counterpart in the reference (which has the first
 * form only); the start-state form is the second of the two objective
 * definitions the 64^3 Taylor-Green adjoint reports. */
int adfvm_set_objective_state(adfvm_ctx* ctx, int32_t which);
int adfvm_objective_state(const adfvm_ctx* ctx);

/* One explicit time step in place: u <- step(u).  The reference's primal
 * Function (density.py:97-105: SSP-RK3 via timestep.py:27-55 of
 * RCF.equation, density.py:339-414).  source: device [5][nIC] or NULL.
 * dtc_out (device double, nullable): max_c sum_f (|Un|+a)A/V evaluated on the
 * first intermediate stage, density.py:405-413.  obj_out (device double,
 * nullable): the objective on the same stage (or on u_n, see
 * adfvm_set_objective_state), in FP64.  tape (device,
 * adfvm_tape_bytes() bytes, nullable): receives the step's intermediate
 * stage states and stage gradients for a later adfvm_adjoint_step. */
int adfvm_primal_step(adfvm_ctx* ctx, void* u, double dt, const void* source,
                      double* dtc_out, double* obj_out, void* tape, void* stream);

/* Size in bytes of one step's tape (2 stage states + 3 stage gradient
 * records for SSP-RK3).  Keeping tapes trades HBM capacity for the forward
 * recomputation the reference's reverse sweep does (apps/adjoint.py:211-217). */
int64_t adfvm_tape_bytes(const adfvm_ctx* ctx);

/* Reverse of one time step (the primal_grad Function, apps/adjoint.py:94-126,
 * called at :268-291 with dtca = 0).  u_n: device [5][nIC], the primal state
 * at the start of the step (unchanged).  ua: device [5][nIC]; in: adjoint of
 * the step's output, out: adjoint of its input, d(obja*J_n + <ua_in,u_{n+1}>)/du_n.
 * source_grad (device [5][nIC], nullable): the source-term gradient of the
 * same functional is ADDED to it (parameter 'source', apps/adjoint.py:104-106).
 * bc_grad (device [5][nBoundaryFaces], nullable): the gradient with respect to
 * the boundary-condition values is ADDED to it, row i / boundary face b
 * (= face - nInternalFaces) holding the value that sets ghost component i
 * (parameter ('BCs', field, patch, key), apps/adjoint.py:111-118):
 *   fixedValue U value -> rows 0-2, T value -> 3, p value -> 4;
 *   CBC_UPT U0 -> 0-2, T0 -> 3, p0 -> 4;
 *   CBC_TOTAL_PT direction -> 0-2 (when given), Tt -> 3, pt -> 4.
 * Faces of other rules are left unchanged.
 * tape (nullable): the tape the forward step from u_n filled; without it the
 * step's stages are recomputed from u_n. */
int adfvm_adjoint_step(adfvm_ctx* ctx, const void* u_n, double dt, const void* source,
                       void* ua, double obja, void* source_grad, void* bc_grad, const void* tape,
                       void* stream);

/* adfvm_adjoint_step for an objective the caller evaluates itself: obj_grad
 * (device [5][nIC]) is dJ_n/du at the step's objective stage (the first
 * intermediate state, density.py:412-413 -- the stage-1 state a tape holds
 * first, or u_n for a one-stage integrator or under ADFVM_OBJSTATE_START),
 * added with weight obja to the
 * adjoint there, besides the context's own objective (if any).  The
 * reference's case-file objective(fields, solver) is differentiated by adpy
 * as part of the primal Function (density.py:354-357, apps/adjoint.py:101);
 * here the caller differentiates its objective and chains the primitive
 * fields to the state with adfvm_primitive_ghosts_vjp. */
int adfvm_adjoint_step_cotangent(adfvm_ctx* ctx, const void* u_n, double dt, const void* source, void* ua,
                                 double obja, const void* obj_grad, void* source_grad, void* bc_grad,
                                 const void* tape, void* stream);

/* Mesh-geometry gradients (the parameter 'mesh', apps/adjoint.py:107-110:
 * the step differentiated with respect to Mesh.gradFields, mesh.py:28-32,
 * each an independent input of the reference's Function).  After
 * adfvm_set_mesh_grad(ctx, g) with g a device array of
 * adfvm_mesh_grad_size(ctx) scalars of the context precision, every adjoint
 * step (adfvm_adjoint_step, adfvm_adjoint_step_cotangent) ADDS
 * d(obja*J_n + <ua_in, u_{n+1}>)/d(mesh) to g, laid out as the reference's
 * arrays one after another, row-major:
 *   areas [nF], volumesL [nF], volumesR [nIF], weights [nF], deltas [nF],
 *   normals [nF][3], deltasUnit [nF][3], linearWeights [nF][2],
 *   quadraticWeights [nF][2][3], volumes [nIC]
 * (the built-in objectives' geometry included; an objective the caller
 * differentiates itself, adfvm_adjoint_step_cotangent's obj_grad, adds only
 * its state dependence).  g = NULL switches it off.  Not with cell
 * segments. */
int64_t adfvm_mesh_grad_size(const adfvm_ctx* ctx);
int adfvm_set_mesh_grad(adfvm_ctx* ctx, void* mesh_grad);

/* Adjoint-viscosity stabilisation of the adjoint state after an adjoint step
 * (the viscous map of apps/adjoint.py:127-140 with computeAdjointViscosity
 * and viscositySolver, postpro.py:491-719; the GPU heat-equation solve of
 * matop_cuda.cpp:108-235).  From the primal state u_n (device [5][nIC], the
 * step's input) it builds the symmetriser matrix of `type` per cell
 * (ADFVM_VISC_UNIFORM, _ABARBANEL, _TURKEL; postpro.py:545-627), takes its
 * largest eigenvalue (postpro.py:629-633), scales it to
 * eig * scaling / sqrt(sum eig^2 V / sum V) (postpro.py:635-652), and applies
 * `iterations` Jacobi sweeps of the implicit step
 * (1 + dt/V sum_f A DT_f/delta_f) phi - dt/V sum_f (A DT_f/delta_f) phi_nb = ua/V
 * over internal faces (the reference runs 1000), then ua = phi V in place.
 * m2norm_out (device [nIC], nullable): the scaled viscosity per cell. */
#define ADFVM_VISC_UNIFORM 0
#define ADFVM_VISC_ABARBANEL 1
#define ADFVM_VISC_TURKEL 2
int adfvm_adjoint_viscosity(adfvm_ctx* ctx, const void* u_n, void* ua, double dt, int32_t type, double scaling,
                            int32_t iterations, void* m2norm_out, void* stream);

/* Block-size autotuner for the six stencil kernels (gradient, face flux,
 * stage update, face adjoint, adjoint gather, gradient adjoint).  Times each
 * at 64, 128, 256 and 512 threads per block (reps timed launches each) on
 * the device state u and keeps the fastest for this context.  Results do not
 * depend on the block size: each thread's arithmetic is unchanged and the
 * only cross-thread combination in these kernels is a maximum.  Uses the
 * context's work buffers and synchronizes the stream, so it is not for use
 * inside graph capture.  No counterpart in the reference (adpy fixes its
 * launch shapes at code generation). */
int adfvm_autotune(adfvm_ctx* ctx, const void* u, int32_t reps, void* stream);
/* i-th tunable kernel (0 <= i < 6): its name, the block size in use and the
 * time per launch adfvm_autotune measured for it (-1 if not tuned). */
int adfvm_kernel_config(const adfvm_ctx* ctx, int32_t i, char* name, int32_t len, int32_t* block,
                        double* ms);
/* Sets one tunable kernel's block size (64, 128, 256 or 512) by name. */
int adfvm_set_kernel_block(adfvm_ctx* ctx, const char* name, int32_t block);
/* Streaming policy of the stencil kernels: mode 1 loads the rows a launch
 * uses once (face and slot geometry, slot tables, a cell's own state)
 * non-temporal, so that they leave L2 first; 0 loads them plainly, keeping
 * them cached for the next kernels; -1 (the default) chooses 1 when one RK
 * stage's working set exceeds 256 MB, the MI355X Infinity Cache.  Results
 * do not depend on it.  No counterpart in the reference (a cache policy). */
int adfvm_set_streaming(adfvm_ctx* ctx, int32_t mode);
/* The policy in use: 1 streaming, 0 not, negative for a null context. */
int adfvm_streaming(const adfvm_ctx* ctx);

/* Operator-level entry points (the hot path's pieces, used by tests). */

/* Gauss cell gradients of the primitive fields of u, adFVM/op.py:45-63 via
 * density.py:245-251.  grad: device [15][nIC], rows dU_i/dx_k (i-major,
 * 9 rows), dT/dx_k (3), dp/dx_k (3). */
int adfvm_gradients(adfvm_ctx* ctx, const void* u, void* grad, void* stream);

/* Right-hand side R(u) = (1/V) sum_f F.A (div of the face fluxes,
 * adFVM/op.py:12-29 from density.py:253-331,388-403).  R: device [5][nIC];
 * dtc_cell: device [nIC] or NULL. */
int adfvm_residual(adfvm_ctx* ctx, const void* u, void* R, void* dtc_cell, void* stream);

/* ubar = (dR/du)^T rbar.  rbar, ubar: device [5][nIC]. */
int adfvm_residual_vjp(adfvm_ctx* ctx, const void* u, const void* rbar, void* ubar, void* stream);

/* Objective on a state (after ghost update), FP64 result in obj_out (device). */
int adfvm_objective(adfvm_ctx* ctx, const void* u, double* obj_out, void* stream);

/* Primitive fields with ghost cells, the 'init' Function's boundary update
 * (density.py:64-78): prim device [5][nCells] rows Ux,Uy,Uz,T,p. */
int adfvm_primitive_ghosts(adfvm_ctx* ctx, const void* u, void* prim, void* stream);

/* Reverse of adfvm_primitive_ghosts: ubar (device [5][nIC], overwritten) =
 * (d prim/d u)^T prim_bar for prim_bar device [5][nCells] (ghost rows
 * included): ghost rows chain to their owner through the boundary rule, a
 * cyclic ghost row to the paired cell it copies.  bc_grad (device
 * [5][nBoundaryFaces], nullable): the ghost rows' dependence on the
 * boundary values is ADDED to it (rows as in adfvm_adjoint_step).  Lets a
 * caller's objective of the primitive fields (templates/cylinder_test.py:
 * 9-37, density.py:354-357) be chained to the conservative state. */
int adfvm_primitive_ghosts_vjp(adfvm_ctx* ctx, const void* u, const void* prim_bar, void* ubar, void* bc_grad,
                               void* stream);

/* Gauss gradient of an arbitrary cell field given on all nCells (ghost values
 * included, no boundary update), the reference unit-test operator
 * (tests/test_op.py:20-91 -> op.gradCell).  phi: device [k][nCells];
 * grad: device [k*3][nIC]. */
int adfvm_grad_cell(adfvm_ctx* ctx, const void* phi, int32_t k, void* grad, void* stream);

/* Second-order face reconstruction interp.secondOrder (adFVM/interp.py:20-28),
 * swap 0 (owner side) or 1: phi device [k][nCells], grad device [k*3][nCells],
 * out device [k][nFaces]. */
int adfvm_interp_second_order(adfvm_ctx* ctx, const void* phi, const void* grad, int32_t k,
                              int32_t swap, void* out, void* stream);

/* Reductions with FP64 accumulation in a fixed order.
 * sensitivity = sum_j a[j]*b[j] (cmesh.cpp:309-340); n scalars, device. */
int adfvm_dot(adfvm_ctx* ctx, const void* a, const void* b, int64_t n, double* out, void* stream);
/* adfvm_dot without a context and without allocating: the same kernels and
 * the same fixed summation order as adfvm_dot (bitwise equal to it for n up
 * to that context's partial count x 256), with the block partials in a
 * file-scope device array of 4 MiB per GPU.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, allocation-free serving loops).  precision is 4
 * or 8 (a, b float or double; NULL allowed when n = 0, giving 0);
 * n <= 2^27 (= 134217728) scalars, larger n is
 * rejected with ADFVM_ERR_INVALID.  The first call on a device looks up that
 * device's copy of the array (make it before a capture).  Calls on one GPU
 * share that array: calls that overlap in time on one GPU (different streams
 * or threads) race and must be serialised by the caller.  Errors:
 * adfvm_noalloc_last_error() (per host thread). */
int adfvm_dot_noalloc(int32_t precision, const void* a, const void* b, int64_t n, double* out, void* stream);
const char* adfvm_noalloc_last_error(void);
/* adjoint energy sum_c w^T A0 w V_c with w = ua/V (cmesh.cpp:342-425). */
int adfvm_adjoint_energy(adfvm_ctx* ctx, const void* ua, const void* u, double* out, void* stream);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * each hot-path kernel (replaces the reference's wall-clock prints,
 * adFVM/solver.py:309-346, at kernel granularity).  enable = 1 starts
 * recording; enable = 0 synchronises the stream and folds the recorded
 * launches into per-kernel totals, read with adfvm_profile_entry. */
int adfvm_profile(adfvm_ctx* ctx, int32_t enable, void* stream);
int adfvm_profile_count(const adfvm_ctx* ctx);
int adfvm_profile_entry(const adfvm_ctx* ctx, int32_t i, char* name, int32_t len, double* total_ms,
                        int64_t* launches);
int adfvm_profile_reset(adfvm_ctx* ctx);

/* Layout helpers: reference layout (rho nIC, rhoU nIC x 3 row-major,
 * rhoE nIC; device, context precision) <-> state u[5][nIC]. */
int adfvm_pack_state(adfvm_ctx* ctx, const void* rho, const void* rhoU, const void* rhoE,
                     void* u, void* stream);
int adfvm_unpack_state(adfvm_ctx* ctx, const void* u, void* rho, void* rhoU, void* rhoE,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ADFVM_H */
