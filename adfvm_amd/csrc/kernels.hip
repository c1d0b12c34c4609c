// libadfvm: HIP kernels for gfx950 (MI355X) and the C ABI of include/adfvm.h.
//
// Design (one GPU, one context per mesh):
//  * state and gradients are structure-of-arrays rows of nInternalCells
//    scalars in the context precision (float or double);
//  * every face->cell accumulation is a gather: each cell walks its six
//    cellFaces slots in order (owned faces in face order, then neighbour
//    faces, cmesh.cpp:9-19) and reads what the face kernel stored for them.
//    No atomics: results are bit-identical from run to run;
//  * cyclic ghosts are resolved at set-up to the paired internal cell; the
//    ghost of a non-cyclic boundary face is evaluated where it is used, from
//    its owner cell and the patch's values (a ghost array is filled only for
//    the ghost-value entry point and the adjoint viscosity);
//  * reductions (objective, dtc) write one partial per block and a single
//    block folds the partials in index order, in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "adfvm.h"
#include "fvm_math.hpp"
#include "adjoint_math.hpp"
#include "dual.hpp"

using namespace adfvm;

#define BLOCK 256

// Workgroups of a launch are dealt to the 8 XCDs round-robin (block b runs
// on XCD b % 8), and each XCD has its own L2.  The stencil kernels remap
// block b to a logical block so that every XCD works through one contiguous
// eighth of the cells (or faces): the neighbours a block gathers (RCM
// order: within a few thousand rows) were then recently touched on the
// same XCD and hit its L2.  The map is a bijection of [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
    return (int)(x * q + (x < r ? x : r) + k);
}

// The face kernels read both cells of a face; there a contiguous eighth per
// XCD measured slower (though it cut their HBM reads by a fifth), while
// chunks of 16 consecutive blocks per XCD, the chunks dealt round-robin,
// measured 1-2% faster than the plain order.  Bijection of [0, gridDim.x)
// (the last partial chunk keeps the plain order).
__device__ __forceinline__ int xcd_chunk_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x, C = 16;
    const unsigned full = nb / (8u * C) * (8u * C);
    if (b >= full) return (int)b;
    const unsigned chunk = b / (8u * C), r = b % (8u * C);
    return (int)(chunk * 8u * C + (r & 7u) * C + (r >> 3));
}

// ------------------------------------------------------------- tables ----
// slot flags: bit 0 = this cell owns the face; bits 1-2 = slot kind; bit 3
// (boundary slots) = the ghost's values are the owner's (zeroGradient U, and
// zeroGradient or symmetry T and p: BCs.py:97-104,120-137), so neither the
// ghost rule nor its adjoint needs the patch
enum { SLOT_INTERNAL = 0, SLOT_COUPLED = 1, SLOT_CHARACTERISTIC = 2, SLOT_BOUNDARY = 3 };
constexpr int SLOT_GHOST_COPY = 8;
__host__ __device__ __forceinline__ int slot_kind(int flags) { return (flags >> 1) & 3; }
// face kind bytes (device copy): FACE_* in bits 0-1, bit 2 for a boundary
// face whose ghost copies its owner (as the slot flag above)
constexpr int FACE_GHOST_COPY = 4;
__host__ __device__ __forceinline__ int face_kind(int k) { return k & 3; }

// the ghost of a non-characteristic patch is its owner's values
inline bool ghost_copies_owner(const adfvm_patch_desc& Pd) {
    auto scalarCopy = [](int bc) { return bc == ADFVM_BC_ZERO_GRADIENT || bc == ADFVM_BC_SYMMETRY; };
    return Pd.fluxKind == ADFVM_FLUX_BOUNDARY && Pd.cbc == ADFVM_CBC_NONE && Pd.bcU == ADFVM_BC_ZERO_GRADIENT &&
           scalarCopy(Pd.bcT) && scalarCopy(Pd.bcp);
}

// face geometry rows (each nFaces long)
enum { G_A = 0, G_NX, G_NY, G_NZ, G_W, G_DELTA, G_DX, G_DY, G_DZ, G_LW0, G_LW1,
       G_Q0X, G_Q0Y, G_Q0Z, G_Q1X, G_Q1Y, G_Q1Z, G_ROWS };

struct DPatch {
    int start, n, fluxKind, bcU, bcT, bcp, cbc, hasDir;
    long valU, valT, valp, valTt, valpt, valDir;   // offsets into the value buffer, -1 if absent
};

template <typename T> struct DMesh {
    int nIC, nIF, nF, nBF;
    const T* geo;           // [G_ROWS][nF]
    const T* vol;           // [nIC]
    const T* ngeo;          // [6][3][nIC]: per coupled cell slot, the paired face's owner's gradient
                            // stencil seen from this cell: coef * S*n / V_other (zero elsewhere)
    const T* sgeo;          // [6][4][nIC]: per cell slot, S*n signed outward for the cell (3) and
                            // the weight of the other side's value in the face value (op.py:47-58)
    const int* sFace;       // [6][nIC]
    const int* sOther;      // [6][nIC]: internal cell, or -1-b for boundary face b
    const unsigned char* sFlags;  // [6][nIC]
    const int* sPair;       // [6][nIC]: paired face of a coupled slot, else -1
    const int* owner;       // [nF]
    const int* fOther;      // [nF]: neighbour, paired internal cell (cyclic) or -1-b (ghost)
    const unsigned char* fKind;  // [nF]: FACE_INTERIOR / FACE_CHARACTERISTIC / FACE_BOUNDARY
    const int* posP;        // [nF]: record index (plane*nIC + cell) of the face's owner-side adjoint
    const int* posN;        // [nF]: record index of its neighbour-side adjoint, -1 if none
    int ownerSum;           // the face adjoint sums each cell's owned internal faces into plane 0
    const int* bPatch;      // [nBF]
    const DPatch* patches;
    const T* bcval;
};

// Loads and stores of data a launch touches once, in rows whose cache lines
// one instruction consumes whole (face geometry, slot tables, a thread's own
// state rows): non-temporal, so that they are first out of L2 and the cell
// records that neighbouring threads read again stay in it (grad 6%, face
// flux 5%, gradient adjoint 8% faster at TGV 128^3 FP32).  Not for records
// that straddle lines or that lanes share: there the hint re-fetches lines
// (adjoint gather 55% slower with non-temporal record loads).
// ST selects it per context (adfvm_ctx::stream): on meshes whose stage
// working set fits the 256 MB Infinity Cache the same rows are reused by the
// next kernels and stages, and plain loads keep them there (the 512 x 256
// cylinder ran 9% slower streamed).
template <bool ST, typename T> __device__ __forceinline__ T ld(const T* p) {
    if constexpr (ST) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool ST, typename T> __device__ __forceinline__ void st(T* p, T v) {
    if constexpr (ST) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool ST, typename T> __device__ __forceinline__ void load_geo(const DMesh<T>& m, int f, FaceGeo<T>& g) {
    const T* G = m.geo;
    const long s = m.nF;
    g.A = ld<ST>(G + G_A * s + f);
    g.n[0] = ld<ST>(G + G_NX * s + f); g.n[1] = ld<ST>(G + G_NY * s + f); g.n[2] = ld<ST>(G + G_NZ * s + f);
    g.w = T(0);           // the face weight enters gradients only (slot geometry), not the face flux
    g.delta = ld<ST>(G + G_DELTA * s + f);
    g.D[0] = ld<ST>(G + G_DX * s + f); g.D[1] = ld<ST>(G + G_DY * s + f); g.D[2] = ld<ST>(G + G_DZ * s + f);
    g.lw0 = ld<ST>(G + G_LW0 * s + f); g.lw1 = ld<ST>(G + G_LW1 * s + f);
    g.qw0[0] = ld<ST>(G + G_Q0X * s + f); g.qw0[1] = ld<ST>(G + G_Q0Y * s + f); g.qw0[2] = ld<ST>(G + G_Q0Z * s + f);
    g.qw1[0] = ld<ST>(G + G_Q1X * s + f); g.qw1[1] = ld<ST>(G + G_Q1Y * s + f); g.qw1[2] = ld<ST>(G + G_Q1Z * s + f);
}

// primitive values of a value source: internal cell (from the state) or ghost.
// Branch-free (one strided load from either source, then a per-element
// select) so that w stays in registers: writing the array in two branches
// made the compiler place it in scratch memory.
template <typename T>
__device__ __forceinline__ void load_prim(const Phys<T>& ph, const T* __restrict__ u, long nIC,
                                          const T* __restrict__ ghost, long nBF, int src, T* w) {
    const bool internal = src >= 0;
    const T* base = internal ? u + src : ghost + (-1 - src);
    const long stride = internal ? nIC : nBF;
    T q[5], p[5];
    for (int i = 0; i < 5; i++) q[i] = base[i * stride];
    primitive(ph, q, p);
    for (int i = 0; i < 5; i++) w[i] = internal ? p[i] : q[i];
}

// Internal work buffers are arrays of records so that a gathered record is a
// few vector loads instead of one cache line per field:
//   gradients        [nIC][GREC]   15 used (padded to 16-byte multiples)
//   face fluxes      [nF][FREC]    F*A (5), lam*A, unpadded: 24 B in FP32 (48 B in
//                                  FP64), accessed as pairs of scalars
//   face adjoints    [nPlanes][nIC][AREC] one plane per cell slot, then cyclic
//                                  extras: gradient adjoint (15), primitive adjoint (5)
//   cell adjoints    [nIC][AREC]   same layout as one side of a face record
constexpr int GREC = 16, FREC = 6, AREC = 20;

// N contiguous values from / to a 16-byte aligned address (N a multiple of
// 16/sizeof(T)), as 16-byte memory operations.  The copies go through
// __builtin_memcpy so that no object is read through a pointer of another
// type: the T values are never accessed as a vector type or vice versa.
template <int N, typename T> __device__ __forceinline__ void vload(const T* __restrict__ p, T* out) {
    static_assert((N * sizeof(T)) % 16 == 0, "vector width");
    const T* a = static_cast<const T*>(__builtin_assume_aligned(p, 16));
#pragma unroll
    for (int k = 0; k < (int)(N * sizeof(T) / 16); k++) __builtin_memcpy(out + k * (16 / sizeof(T)), a + k * (16 / sizeof(T)), 16);
}

template <int N, typename T> __device__ __forceinline__ void vstore(T* __restrict__ p, const T* in) {
    static_assert((N * sizeof(T)) % 16 == 0, "vector width");
    T* a = static_cast<T*>(__builtin_assume_aligned(p, 16));
#pragma unroll
    for (int k = 0; k < (int)(N * sizeof(T) / 16); k++) __builtin_memcpy(a + k * (16 / sizeof(T)), in + k * (16 / sizeof(T)), 16);
}

// Records of an even number of scalars aligned to a pair of them (the face
// flux records): pairs of scalars per access.
template <int N, typename T> __device__ __forceinline__ void pload(const T* __restrict__ p, T* out) {
    static_assert(N % 2 == 0, "pairs");
    const T* a = static_cast<const T*>(__builtin_assume_aligned(p, 2 * sizeof(T)));
#pragma unroll
    for (int k = 0; k < N / 2; k++) __builtin_memcpy(out + 2 * k, a + 2 * k, 2 * sizeof(T));
}
template <int N, typename T> __device__ __forceinline__ void pstore(T* __restrict__ p, const T* in) {
    static_assert(N % 2 == 0, "pairs");
    T* a = static_cast<T*>(__builtin_assume_aligned(p, 2 * sizeof(T)));
#pragma unroll
    for (int k = 0; k < N / 2; k++) __builtin_memcpy(a + 2 * k, in + 2 * k, 2 * sizeof(T));
}

// the 15 gradient entries of cell c (g has room for GREC values)
template <typename T>
__device__ __forceinline__ void load_grad(const T* __restrict__ grad, long nIC, int c, T* g) {
    (void)nIC;
    vload<GREC>(grad + (long)GREC * c, g);
}

// ------------------------------------------------------------- kernels ---

// Ghost primitive values of non-cyclic boundary face b from its owner's
// primitive values w (Solver.boundary + characteristicBoundary,
// solver.py:106-114, density.py:333-337, BCs.py): every rule reads only the
// owner, the face normal and the patch's values, so the kernels that need a
// ghost evaluate it where they use it.  ghost_rule takes the normal from
// nrm(n) (called only by the rules that use it) and computes in T, S or
// Dual<S>; ghost_value reads the normal from the face geometry.
template <typename T, typename S, typename NormalFn>
__device__ __forceinline__ void ghost_rule(const DMesh<S>& m, const Phys<T>& ph, int b, const T* w,
                                           NormalFn nrm, T* out) {
    const DPatch P = m.patches[m.bPatch[b]];
    const int f = m.nIF + b;
    const int k = f - P.start;
    const S* V = m.bcval;
    if (P.cbc == ADFVM_CBC_UPT) {                    // BCs.py:152-161
        for (int i = 0; i < 3; i++) out[i] = T(V[P.valU + 3 * k + i]);
        out[3] = T(V[P.valT + k]);
        out[4] = T(V[P.valp + k]);
    } else if (P.cbc == ADFVM_CBC_TOTAL_PT) {        // BCs.py:164-184
        T d[3];
        if (P.hasDir) { for (int i = 0; i < 3; i++) d[i] = T(V[P.valDir + 3 * k + i]); }
        else nrm(d);
        T Tt = T(V[P.valTt + k]), pt = T(V[P.valpt + k]);
        T Un = w[0] * d[0] + w[1] * d[1] + w[2] * d[2];
        for (int i = 0; i < 3; i++) out[i] = Un * d[i];
        out[3] = Tt - T(0.5) * Un * Un / ph.Cp;
        out[4] = pt * pow(w[3] / Tt, ph.gamma / (ph.gamma - T(1)));
    } else {
        // U
        if (P.bcU == ADFVM_BC_FIXED_VALUE || P.bcU == ADFVM_BC_CALCULATED) {
            for (int i = 0; i < 3; i++) out[i] = T(V[P.valU + 3 * k + i]);
        } else {
            for (int i = 0; i < 3; i++) out[i] = w[i];
            if (P.bcU == ADFVM_BC_SYMMETRY) {        // BCs.py:125-133
                T n[3];
                nrm(n);
                T un = out[0] * n[0] + out[1] * n[1] + out[2] * n[2];
                out[0] -= un * n[0]; out[1] -= un * n[1]; out[2] -= un * n[2];
            }
        }
        out[3] = (P.bcT == ADFVM_BC_FIXED_VALUE || P.bcT == ADFVM_BC_CALCULATED) ? T(V[P.valT + k]) : w[3];
        out[4] = (P.bcp == ADFVM_BC_FIXED_VALUE || P.bcp == ADFVM_BC_CALCULATED) ? T(V[P.valp + k]) : w[4];
    }
}

template <typename T>
__device__ __forceinline__ void ghost_value(const DMesh<T>& m, const Phys<T>& ph, int b, const T* w, T* out) {
    const int f = m.nIF + b;
    ghost_rule(m, ph, b, w, [&](T* n) {
        const long s = m.nF;
        n[0] = m.geo[G_NX * s + f]; n[1] = m.geo[G_NY * s + f]; n[2] = m.geo[G_NZ * s + f];
    }, out);
}

// The ghost array (ghost primitives [5][nBF]), for the consumers that read
// it as an array: the ghost-value entry point and the adjoint viscosity.
template <typename T>
__global__ void __launch_bounds__(BLOCK) ghost_kernel(DMesh<T> m, Phys<T> ph, const T* __restrict__ u,
                                                      T* __restrict__ ghost) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nBF) return;
    const DPatch P = m.patches[m.bPatch[b]];
    if (P.bcU == ADFVM_BC_CYCLIC && P.cbc == ADFVM_CBC_NONE) return;   // resolved by the slot tables
    const long nIC = m.nIC, nBF = m.nBF;
    const int c = m.owner[m.nIF + b];
    T q[5], w[5], out[5];
    for (int i = 0; i < 5; i++) q[i] = u[i * nIC + c];
    primitive(ph, q, w);
    ghost_value(m, ph, b, w, out);
    for (int i = 0; i < 5; i++) ghost[i * nBF + b] = out[i];
}

// Gauss cell gradients of U, T, p (op.gradCell, op.py:45-63).
template <typename T, int BS, bool ST, bool GH>
__global__ void __launch_bounds__(BS) grad_kernel(DMesh<T> m, Phys<T> ph, const T* __restrict__ u,
                                                     const T* __restrict__ ghost, T* __restrict__ grad) {
    int c = xcd_block() * BS + threadIdx.x;
    if (c >= m.nIC) return;
    const long nIC = m.nIC;
    const T V = m.vol[c];       // loaded first: read after the slot loop it was a last, unhidden wait
    T wc[5];
    load_prim(ph, u, nIC, ghost, m.nBF, c, wc);
    T g[15];
    for (int i = 0; i < 15; i++) g[i] = T(0);
    for (int j = 0; j < 6; j++) {
        const int o = ld<ST>(m.sOther + j * nIC + c);
        T wo[5];
        if constexpr (GH) {
            // a ghost slot's value from this cell (the face's owner): a copy
            // without reading the patch when the slot is flagged so
            const bool copy = ld<ST>(m.sFlags + j * nIC + c) & SLOT_GHOST_COPY;
            T wi[5], wg[5];
            load_prim(ph, u, nIC, u, nIC, o < 0 ? c : o, wi);
            for (int i = 0; i < 5; i++) wg[i] = wc[i];
            if (o < 0 && !copy) ghost_value(m, ph, -1 - o, wc, wg);
            for (int i = 0; i < 5; i++) wo[i] = o < 0 ? wg[i] : wi[i];
        } else {
            load_prim(ph, u, nIC, ghost, m.nBF, o, wo);
        }
        const T* sg = m.sgeo + (long)(4 * j) * nIC + c;
        const T SN[3] = {ld<ST>(sg), ld<ST>(sg + nIC), ld<ST>(sg + 2 * nIC)};
        const T w = ld<ST>(sg + 3 * nIC);
        for (int i = 0; i < 5; i++) {
            T phiF = wc[i] * (T(1) - w) + wo[i] * w;
            g[3 * i + 0] += phiF * SN[0];
            g[3 * i + 1] += phiF * SN[1];
            g[3 * i + 2] += phiF * SN[2];
        }
    }
    T out[GREC];
    for (int i = 0; i < 15; i++) out[i] = g[i] / V;
    out[15] = T(0);
    vstore<GREC>(grad + (long)GREC * c, out);
}

// gradient records -> the public [15][nIC] rows
template <typename T>
__global__ void __launch_bounds__(BLOCK) grad_rows_kernel(long nIC, const T* __restrict__ grec, T* __restrict__ rows) {
    long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nIC) return;
    T g[GREC];
    vload<GREC>(grec + GREC * c, g);
    for (int i = 0; i < 15; i++) rows[i * nIC + c] = g[i];
}

// Block reduction of one double per thread into out[blockIdx.x]; op 0 sum, 1 max.
template <int OP, int BS = BLOCK>
__device__ void block_reduce_store(double v, double* out) {
    __shared__ double sh[BS];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = sh[threadIdx.x], b = sh[threadIdx.x + s];
            sh[threadIdx.x] = OP == 0 ? a + b : (a > b ? a : b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Folds n partials in index order inside one block: out = op(partials).
template <int OP>
__global__ void __launch_bounds__(BLOCK) finalize_kernel(const double* __restrict__ partials, int n, double* out) {
    double v = OP == 0 ? 0.0 : -1.0e308;
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        double x = partials[i];
        v = OP == 0 ? v + x : (v > x ? v : x);
    }
    __shared__ double sh[BLOCK];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = sh[threadIdx.x], b = sh[threadIdx.x + s];
            sh[threadIdx.x] = OP == 0 ? a + b : (a > b ? a : b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Face fluxes, one face per thread (faces are in owner order, so the owner
// reads of neighbouring threads coincide and the geometry rows stream):
//   fbuf[i][f] = F_i(f)*A_f (i < 5),  fbuf[5][f] = lam_f*A_f
// Internal and coupled faces: RCF.flux (density.py:253-305) between the
// owner and the neighbour (for a cyclic face, the paired internal cell);
// characteristic / other boundary faces: the ghost as the right state.
template <typename T, int BS, bool ST, bool GH>
__global__ void __launch_bounds__(BS) face_flux_kernel(DMesh<T> m, Phys<T> ph, const T* __restrict__ u,
                                                          const T* __restrict__ grad, const T* __restrict__ ghost,
                                                          T* __restrict__ fbuf) {
    const int f = xcd_chunk_block() * BS + threadIdx.x;
    if (f >= m.nF) return;
    const long nIC = m.nIC, nBF = m.nBF, nF = m.nF;
    const int P = m.owner[f];
    const int o = ld<ST>(m.fOther + f);
    const int kindByte = ld<ST>(m.fKind + f);
    const int kind = face_kind(kindByte);
    FaceGeo<T> fg;
    load_geo<ST>(m, f, fg);
    T wP[5], gP[GREC], wN[5], gN[GREC], F[5], lam;
    load_prim(ph, u, nIC, ghost, nBF, P, wP);
    load_grad(grad, nIC, P, gP);
    if constexpr (GH) {
        // a boundary face's ghost from its owner (a copy without reading the
        // patch when the face is flagged so)
        T wi[5], wg[5];
        load_prim(ph, u, nIC, u, nIC, o < 0 ? P : o, wi);
        for (int i = 0; i < 5; i++) wg[i] = wP[i];
        if (o < 0 && !(kindByte & FACE_GHOST_COPY)) ghost_value(m, ph, -1 - o, wP, wg);
        for (int i = 0; i < 5; i++) wN[i] = o < 0 ? wg[i] : wi[i];
    } else {
        load_prim(ph, u, nIC, ghost, nBF, o, wN);
    }
    if (kind == FACE_INTERIOR) {
        load_grad(grad, nIC, o, gN);
        face_flux(ph, FACE_INTERIOR, wP, gP, wN, gN, fg, F, &lam);
    } else {
        face_flux(ph, kind, wP, gP, wN, gP, fg, F, &lam);
    }
    T rec[FREC];
    for (int i = 0; i < 5; i++) rec[i] = F[i] * fg.A;
    rec[5] = lam * fg.A;
    pstore<FREC>(fbuf + (long)FREC * f, rec);
    (void)nF;
}

// Residual and stage update, one cell per thread, gathering the six face
// fluxes in cellFaces slot order:
//   L_c = (1/V_c) sum_slots sign*(F A)_f     (op.div, op.py:12-20)
//   lam_c = (1/V_c) sum_slots (lam A)_f      (op.absDiv, op.py:22-29)
// MODE 0: uout = a0*u0 + ai*ui + (-beta*(L-S))*dt   (timestep.update, timestep.py:35-45)
//         with a0 = 1 - ai when u0 is given (every such stage of the tables)
// MODE 1: uout = L (the residual)
// The last SSP-RK3 stage passes the step's state as both u0 and uout (each
// thread reads u0[c] before writing uout[c] at the same index), so neither
// is declared __restrict__.
template <typename T, int MODE, int BS, bool ST>
__global__ void __launch_bounds__(BS) stage_kernel(DMesh<T> m, const T* __restrict__ fbuf,
                                                      const T* __restrict__ ui, const T* u0,
                                                      T ai, T nbeta, T dt, const T* __restrict__ S,
                                                      T* uout, T* __restrict__ dtcCell,
                                                      double* __restrict__ dtcPartial) {
    int c = xcd_block() * BS + threadIdx.x;
    const long nIC = m.nIC, nF = m.nF;
    double dtcv = 0.0;
    if (c < m.nIC) {
        // the cell's own rows first, so their loads overlap the face gathers
        T v[5], v0[5], sv[5];
        if (MODE == 0) {
            for (int i = 0; i < 5; i++) v[i] = ld<ST>(ui + i * nIC + c);
            for (int i = 0; i < 5; i++) v0[i] = u0 ? ld<ST>(u0 + i * nIC + c) : T(0);
            for (int i = 0; i < 5; i++) sv[i] = S ? ld<ST>(S + i * nIC + c) : T(0);
        }
        T L[5] = {0, 0, 0, 0, 0};
        T lamSum = 0;
        for (int j = 0; j < 6; j++) {
            const int f = ld<ST>(m.sFace + j * nIC + c);
            const bool own = ld<ST>(m.sFlags + j * nIC + c) & 1;
            T rec[FREC];
            pload<FREC>(fbuf + (long)FREC * f, rec);
            for (int i = 0; i < 5; i++) L[i] += own ? rec[i] : -rec[i];
            lamSum += rec[5];
        }
        (void)nF;
        const T V = m.vol[c];
        for (int i = 0; i < 5; i++) L[i] = L[i] / V;
        lamSum = lamSum / V;
        if (MODE == 1) {
            for (int i = 0; i < 5; i++) uout[i * nIC + c] = L[i];
        } else {
            for (int i = 0; i < 5; i++) {
                const T Li = S ? L[i] - sv[i] : L[i];
                if (u0) {
                    // a0 = 1 - ai: uout = u0 + (ai*(ui - u0) - beta*(L-S)*dt), the
                    // change formed at its own scale and added to u0 in one
                    // rounding.  (ai*ui + a0*u0 rounds ai*ui on a finer grid than
                    // the sum's; with ai = 2/3 that is biased by -1/9 ulp a step.)
                    const T du = ai * (v[i] - v0[i]) + (nbeta * Li) * dt;
                    uout[i * nIC + c] = v0[i] + du;
                } else {
                    uout[i * nIC + c] = ai * v[i] + (nbeta * Li) * dt;
                }
            }
        }
        if (dtcCell) dtcCell[c] = lamSum;
        dtcv = (double)lamSum;
    }
    if (dtcPartial) block_reduce_store<1, BS>(dtcv, dtcPartial);
}

// Objective partial sums (FP64): kinetic energy 0.5*rho*|U|^2*V per cell.
template <typename T>
__global__ void __launch_bounds__(BLOCK) ke_kernel(DMesh<T> m, const T* __restrict__ u, double* partial) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    const long nIC = m.nIC;
    double v = 0.0;
    if (c < m.nIC) {
        T rho = u[c];
        T Ux = u[nIC + c] / rho, Uy = u[2 * nIC + c] / rho, Uz = u[3 * nIC + c] / rho;
        T ke = T(0.5) * rho * (Ux * Ux + Uy * Uy + Uz * Uz) * m.vol[c];
        v = (double)ke;
    }
    block_reduce_store<0>(v, partial);
}

// Kinetic energy per cell segment (an ensemble of meshes in one context):
// block b covers cells segCell[k] + 256*(b - segBlock[k]) .. of segment
// k = blockSeg[b], so each segment's partials group its cells exactly as a
// context holding that segment alone would.
template <typename T>
__global__ void __launch_bounds__(BLOCK) ke_seg_kernel(DMesh<T> m, const T* __restrict__ u,
                                                       const long* __restrict__ segCell,
                                                       const int* __restrict__ segBlock,
                                                       const int* __restrict__ blockSeg, double* partial) {
    const int k = blockSeg[blockIdx.x];
    const long c = segCell[k] + (long)(blockIdx.x - segBlock[k]) * BLOCK + threadIdx.x;
    const long nIC = m.nIC;
    double v = 0.0;
    if (c < segCell[k + 1]) {
        T rho = u[c];
        T Ux = u[nIC + c] / rho, Uy = u[2 * nIC + c] / rho, Uz = u[3 * nIC + c] / rho;
        T ke = T(0.5) * rho * (Ux * Ux + Uy * Uy + Uz * Uz) * m.vol[c];
        v = (double)ke;
    }
    block_reduce_store<0>(v, partial);
}

// One block per segment: out[k] = sum of its partials, folded as
// finalize_kernel folds a whole context's.
__global__ void __launch_bounds__(BLOCK) finalize_seg_kernel(const double* __restrict__ partials,
                                                             const int* __restrict__ segBlock, double* out) {
    const int k = blockIdx.x;
    const int b0 = segBlock[k], n = segBlock[k + 1] - b0;
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) v = v + partials[b0 + i];
    __shared__ double sh[BLOCK];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = BLOCK / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
}

// Drag on one patch (templates/cylinder_test.py:9-19): sum (p0 nx - mu(T0)(U0x-Uix)/delta) A.
template <typename T>
__global__ void __launch_bounds__(BLOCK) drag_kernel(DMesh<T> m, Phys<T> ph, const T* __restrict__ u, int start,
                                                     int n, double* partial) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (k < n) {
        const int f = start + k;
        const int b = f - m.nIF;
        const long nIC = m.nIC, nF = m.nF;
        T wi[5], wg[5];
        load_prim(ph, u, nIC, u, nIC, m.owner[f], wi);
        ghost_value(m, ph, b, wi, wg);
        T U0 = wg[0], p0 = wg[4], T0 = wg[3];
        T nx = m.geo[G_NX * nF + f];
        T mungUx = viscosity(ph, T0) * (U0 - wi[0]) / m.geo[G_DELTA * nF + f];
        v = (double)((p0 * nx - mungUx) * m.geo[G_A * nF + f]);
    }
    block_reduce_store<0>(v, partial);
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) prim_ghost_kernel(DMesh<T> m, Phys<T> ph, const T* __restrict__ u,
                                                           const T* __restrict__ ghost, T* __restrict__ prim) {
    // prim [5][nC]: internal rows from the state, ghost rows from the ghost
    // array or (cyclic) the paired owner.
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const long nIC = m.nIC, nBF = m.nBF, nC = (long)m.nIC + m.nBF;
    if (i >= nC) return;
    T w[5];
    if (i < m.nIC) {
        load_prim(ph, u, nIC, ghost, nBF, i, w);
    } else {
        int b = i - m.nIC;
        const DPatch P = m.patches[m.bPatch[b]];
        if (P.bcU == ADFVM_BC_CYCLIC && P.cbc == ADFVM_CBC_NONE) {
            // cyclic ghost: the paired owner, as resolved in the owner's slot
            int f = m.nIF + b;
            int c = m.owner[f];
            int partner = -1;
            for (int j = 0; j < 6; j++)
                if (m.sFace[j * nIC + c] == f) partner = m.sOther[j * nIC + c];
            load_prim(ph, u, nIC, ghost, nBF, partner, w);
        } else {
            for (int r = 0; r < 5; r++) w[r] = ghost[r * nBF + b];
        }
    }
    for (int r = 0; r < 5; r++) prim[r * nC + i] = w[r];
}

// op.gradCell over a caller field given on all nCells (reference cellNeighbours,
// no boundary update), k components.
template <typename T>
__global__ void __launch_bounds__(BLOCK) grad_field_kernel(DMesh<T> m, const int* __restrict__ cellNb,
                                                           const T* __restrict__ phi, int k, long nC,
                                                           T* __restrict__ grad) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nIC) return;
    const long nIC = m.nIC, nF = m.nF;
    for (int comp = 0; comp < k; comp++) {
        T g0 = 0, g1 = 0, g2 = 0;
        T pc = phi[comp * nC + c];
        for (int j = 0; j < 6; j++) {
            const int f = m.sFace[j * nIC + c];
            const bool own = m.sFlags[j * nIC + c] & 1;
            const int nb = cellNb[j * nIC + c];
            T S = m.geo[G_A * nF + f];
            T N[3] = {m.geo[G_NX * nF + f], m.geo[G_NY * nF + f], m.geo[G_NZ * nF + f]};
            T w = m.geo[G_W * nF + f];
            if (!own) { N[0] = -N[0]; N[1] = -N[1]; N[2] = -N[2]; }
            else w = T(1) - w;
            T phiF = pc * (T(1) - w) + phi[comp * nC + nb] * w;
            g0 += phiF * (S * N[0]); g1 += phiF * (S * N[1]); g2 += phiF * (S * N[2]);
        }
        T V = m.vol[c];
        grad[(3 * comp + 0) * nIC + c] = g0 / V;
        grad[(3 * comp + 1) * nIC + c] = g1 / V;
        grad[(3 * comp + 2) * nIC + c] = g2 / V;
    }
}

// interp.secondOrder on every face: phi [k][nC], grad [3k][nC], out [k][nF].
template <typename T>
__global__ void __launch_bounds__(BLOCK) interp_kernel(DMesh<T> m, const int* __restrict__ nbFull,
                                                       const T* __restrict__ phi, const T* __restrict__ grad,
                                                       int k, long nC, int swap, T* __restrict__ out) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    const long nF = m.nF;
    int p = m.owner[f], n = nbFull[f];
    if (swap) { int t = p; p = n; n = t; }
    T lw = m.geo[(swap ? G_LW1 : G_LW0) * nF + f];
    T qw[3];
    for (int i = 0; i < 3; i++) qw[i] = m.geo[((swap ? G_Q1X : G_Q0X) + i) * nF + f];
    for (int comp = 0; comp < k; comp++) {
        T pc = phi[comp * nC + p], pd = phi[comp * nC + n];
        T g0 = grad[(3 * comp) * nC + p], g1 = grad[(3 * comp + 1) * nC + p], g2 = grad[(3 * comp + 2) * nC + p];
        out[comp * nF + f] = pc + (pd - pc) * lw + (qw[0] * g0 + qw[1] * g1 + qw[2] * g2);
    }
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) dot_kernel(const T* __restrict__ a, const T* __restrict__ b, long n,
                                                    double* partial) {
    double v = 0.0;
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK)
        v += (double)a[i] * (double)b[i];
    block_reduce_store<0>(v, partial);
}

// adjoint energy, cmesh.cpp:342-425 (per-cell entropy-variable quadratic form)
template <typename T>
__global__ void __launch_bounds__(BLOCK) energy_kernel(DMesh<T> m, Phys<T> ph, const T* __restrict__ ua,
                                                       const T* __restrict__ u, double* partial) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const long nIC = m.nIC;
    double e = 0.0;
    if (i < m.nIC) {
        double g = (double)ph.gamma;
        double v = m.vol[i];
        double rho = u[i], rE = u[4 * nIC + i];
        double u1 = u[nIC + i] / rho, u2 = u[2 * nIC + i] / rho, u3 = u[3 * nIC + i] / rho;
        double q2 = u1 * u1 + u2 * u2 + u3 * u3;
        double p = (rE - rho * q2 / 2) * (g - 1);
        double H = g * p / (rho * (g - 1)) + q2 / 2;
        double w1 = ua[i] / v, w2 = ua[nIC + i] / v, w3 = ua[2 * nIC + i] / v, w4 = ua[3 * nIC + i] / v,
               w5 = ua[4 * nIC + i] / v;
        double A11 = rho, A12 = rho * u1, A13 = rho * u2, A14 = rho * u3, A15 = rE;
        double A22 = rho * u1 * u1 + p, A23 = rho * u1 * u2, A24 = rho * u1 * u3, A25 = rho * H * u1;
        double A33 = rho * u2 * u2 + p, A34 = rho * u2 * u3, A35 = rho * H * u2;
        double A44 = rho * u3 * u3 + p, A45 = rho * H * u3;
        double A55 = rho * H * H - g * p * p / (rho * (g - 1));
        e = (A11 * w1 * w1 + A22 * w2 * w2 + A33 * w3 * w3 + A44 * w4 * w4 + A55 * w5 * w5 +
             2 * (A12 * w1 * w2 + A13 * w1 * w3 + A14 * w1 * w4 + A15 * w1 * w5 + A23 * w2 * w3 + A24 * w2 * w4 +
                  A25 * w2 * w5 + A34 * w3 * w4 + A35 * w3 * w5 + A45 * w4 * w5)) * v;
    }
    block_reduce_store<0>(e, partial);
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) pack_kernel(long n, const T* __restrict__ rho, const T* __restrict__ rhoU,
                                                     const T* __restrict__ rhoE, T* __restrict__ u) {
    long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= n) return;
    u[c] = rho[c];
    u[n + c] = rhoU[3 * c]; u[2 * n + c] = rhoU[3 * c + 1]; u[3 * n + c] = rhoU[3 * c + 2];
    u[4 * n + c] = rhoE[c];
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) unpack_kernel(long n, const T* __restrict__ u, T* __restrict__ rho,
                                                       T* __restrict__ rhoU, T* __restrict__ rhoE) {
    long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= n) return;
    rho[c] = u[c];
    rhoU[3 * c] = u[n + c]; rhoU[3 * c + 1] = u[2 * n + c]; rhoU[3 * c + 2] = u[3 * n + c];
    rhoE[c] = u[4 * n + c];
}

__global__ void fill_kernel(double* p, int n, double v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) copy_kernel(long n, const T* __restrict__ x, T* __restrict__ y) {
    long i = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// ----------------------------------------------------------- context -----

struct adfvm_ctx {
    int prec = 4;
    int nIC = 0, nIF = 0, nF = 0, nBF = 0, nC = 0;
    int nPatches = 0;
    bool hasGhostKernel = false;
    int objective = ADFVM_OBJ_NONE, objectivePatch = -1;
    int objState = ADFVM_OBJSTATE_STAGE1;   // the state the objective is evaluated on
    int integrator = ADFVM_RK_SSPRK3;
    Phys<double> phd;
    std::string err;
    std::vector<DPatch> hpatches;
    std::vector<int> hBPatch;  // boundary face -> patch (host copy)
    // device
    void* geo = nullptr; void* vol = nullptr; void* bcval = nullptr;
    int* sFace = nullptr; int* sOther = nullptr; unsigned char* sFlags = nullptr; int* sPair = nullptr;
    int* owner = nullptr; int* bPatch = nullptr; DPatch* patches = nullptr;
    int* cellNbFull = nullptr;   // reference cellNeighbours [6][nIC]
    int* nbFull = nullptr;       // ghost-extended neighbour [nF]
    void* ghost = nullptr;
    void* sgeo = nullptr;      // [6][4][nIC] cell-slot gradient geometry
    void* ngeo = nullptr;      // [6][3][nIC] cell-slot transposed-stencil coefficients
    void* visc = nullptr;      // adjoint viscosity: [6][nIC] slot coefficients, [nIC] diagonal, [nIC] eigenvalue
    void* u1 = nullptr; void* u2 = nullptr;
    void* grad0 = nullptr; void* grad1 = nullptr; void* grad2 = nullptr;
    void* adjAcc = nullptr;    // [5][nIC] adjoint of u0 accumulated over the stages
    void* adjG = nullptr;      // [nIC][AREC] cell adjoint records (gradient 15, primitive 5)
    double* partial = nullptr; int nPartial = 0;
    // block size of each tunable stencil kernel (adfvm_autotune) and its measured time
    int bs[6] = {256, 256, 256, 256, 256, 256};
    double tuneMs[6] = {-1, -1, -1, -1, -1, -1};
    // cell segments (adfvm_set_segments): objective per segment
    int nSeg = 1, nSegBlocks = 0;
    long* segCell = nullptr; int* segBlock = nullptr; int* blockSeg = nullptr;
    double* scal = nullptr;    // small device scalars
    int* fOther = nullptr; unsigned char* fKind = nullptr;
    int* posP = nullptr; int* posN = nullptr;
    int nPlanes = 6;           // adjoint record planes: one per cell slot, then cyclic extras
    // internal faces sorted by owner: per block size, the face ranges of the
    // owner-summing face adjoint blocks (no owner split between two ranges)
    // boundary faces with a ghost by face kind (0 characteristic, 1 boundary)
    // boundary faces with a ghost by face kind (characteristic, boundary):
    // per block of BLOCK threads its first boundary index and face count,
    // blocks never crossing a gap between runs of one kind
    int* kindBlk[3] = {nullptr, nullptr, nullptr};
    int nKindBlk[3] = {0, 0, 0};
    bool ownerSum = false;
    // the streaming policy (ld / st): non-temporal single-use rows when one
    // stage's working set exceeds the Infinity Cache (adfvm_set_streaming)
    bool stream = false;
    int* fvjpBlk[4] = {nullptr, nullptr, nullptr, nullptr};
    int fvjpNBlk[4] = {0, 0, 0, 0};
    void* meshGrad = nullptr;  // caller's mesh-geometry gradient array (adfvm_set_mesh_grad), or none
    void* fbuf = nullptr;      // [6][nF] face fluxes * area
    void* frec = nullptr;      // [nPlanes][nIC][AREC] face adjoint records, by cell slot
    // per-kernel event timing (adfvm_profile)
    bool prof = false;
    std::vector<hipEvent_t> evPool;
    struct Pending { const char* name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<std::string> profNames;
    std::vector<double> profMs;
    std::vector<long long> profCount;
};

namespace {

// Records a start/stop event pair around the kernel launched in its scope
// when the context's profiling is on (on the launch stream, so the timing is
// that kernel's execution on that stream).
struct KTimer {
    adfvm_ctx* c; const char* name; hipStream_t s; hipEvent_t a = nullptr, b = nullptr;
    static hipEvent_t take(adfvm_ctx* c) {
        if (!c->evPool.empty()) { hipEvent_t e = c->evPool.back(); c->evPool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    KTimer(adfvm_ctx* c_, const char* n, hipStream_t s_) : c(c_), name(n), s(s_) {
        if (!c->prof) return;
        a = take(c); b = take(c);
        if (a) (void)hipEventRecord(a, s);
    }
    ~KTimer() {
        if (!c->prof || !a || !b) return;
        (void)hipEventRecord(b, s);
        c->pending.push_back({name, a, b});
    }
};

int set_err(adfvm_ctx* ctx, int code, const std::string& m) {
    if (ctx) ctx->err = m;
    return code;
}

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_err(ctx, ADFVM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int grid_for(long n) { return (int)((n + BLOCK - 1) / BLOCK); }
int grid_bs(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Tunable stencil kernels and the block sizes they are built for.
enum { K_GRAD = 0, K_FLUX, K_STAGE, K_FVJP, K_GATHER, K_GVJP, K_TUNABLE };
const char* const kTunableName[K_TUNABLE] = {"grad_kernel", "face_flux_kernel", "stage_kernel",
                                              "face_vjp_kernel", "vjp_gather_kernel", "grad_vjp_kernel"};
const int kBlockChoices[4] = {64, 128, 256, 512};
int block_choice(int bs) { return bs == 64 ? 0 : bs == 128 ? 1 : bs == 512 ? 3 : 2; }

// Calls f(std::integral_constant<int, BS>) for the block size bs.
template <typename F> void with_block(int bs, F&& f) {
    switch (bs) {
        case 64: f(std::integral_constant<int, 64>()); break;
        case 128: f(std::integral_constant<int, 128>()); break;
        case 512: f(std::integral_constant<int, 512>()); break;
        default: f(std::integral_constant<int, 256>()); break;
    }
}

// Calls f(std::bool_constant<st>): the streaming policy of the kernels that
// take one (ld / st).
template <typename F> void with_stream(bool st, F&& f) {
    if (st) f(std::bool_constant<true>());
    else f(std::bool_constant<false>());
}

// Calls f(std::bool_constant<gh>): whether the mesh has faces with a ghost
// (the kernels that evaluate ghosts in place are built without that code
// for meshes that have none).
template <typename F> void with_ghosts(bool gh, F&& f) {
    if (gh) f(std::bool_constant<true>());
    else f(std::bool_constant<false>());
}

// The automatic streaming policy: on when one RK stage's working set
// exceeds the 256 MB Infinity Cache.  Per face: 17 geometry rows, the face
// tables (9 B) and a flux record (8 scalars); per cell: state, gradient
// record, slot geometry, volume, three stage states (61 scalars) and the
// slot tables (30 B).
bool stream_auto(const adfvm_ctx* c) {
    const double s = c->prec;
    const double bytes = (double)c->nF * (25 * s + 9) + (double)c->nIC * (61 * s + 30);
    return bytes > 256.0 * (1 << 20);
}

template <typename T> Phys<T> phys_cast(const Phys<double>& p) {
    Phys<T> o;
    o.gamma = (T)p.gamma; o.Cp = (T)p.Cp; o.Cv = (T)p.Cv; o.Pr = (T)p.Pr; o.invCv = (T)p.invCv;
    o.small = (T)p.small; o.mu0 = (T)p.mu0; o.muA = (T)p.muA; o.muTs = (T)p.muTs;
    o.muModel = p.muModel; o.riemann = p.riemann; o.boundaryRiemann = p.boundaryRiemann; o.limit = p.limit;
    return o;
}

template <typename T> DMesh<T> dmesh(const adfvm_ctx* c) {
    DMesh<T> m;
    m.nIC = c->nIC; m.nIF = c->nIF; m.nF = c->nF; m.nBF = c->nBF;
    m.geo = (const T*)c->geo; m.vol = (const T*)c->vol; m.sgeo = (const T*)c->sgeo; m.ngeo = (const T*)c->ngeo;
    m.sFace = c->sFace; m.sOther = c->sOther; m.sFlags = c->sFlags; m.sPair = c->sPair;
    m.owner = c->owner; m.bPatch = c->bPatch; m.patches = c->patches; m.bcval = (const T*)c->bcval;
    m.fOther = c->fOther; m.fKind = c->fKind;
    m.posP = c->posP; m.posN = c->posN;
    m.ownerSum = c->ownerSum ? 1 : 0;
    return m;
}

template <typename T> int upload(adfvm_ctx* ctx, void** dst, const std::vector<T>& v) {
    HIPCHK(ctx, hipMalloc(dst, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) HIPCHK(ctx, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return ADFVM_OK;
}

template <typename T> int create_typed(adfvm_ctx* ctx, const adfvm_solver_desc* d) {
    const long nF = ctx->nF, nIC = ctx->nIC, nBF = ctx->nBF;
    std::vector<T> geo((size_t)G_ROWS * nF);
    for (long f = 0; f < nF; f++) {
        geo[G_A * nF + f] = (T)d->areas[f];
        for (int i = 0; i < 3; i++) geo[(G_NX + i) * nF + f] = (T)d->normals[3 * f + i];
        geo[G_W * nF + f] = (T)d->weights[f];
        geo[G_DELTA * nF + f] = (T)d->deltas[f];
        for (int i = 0; i < 3; i++) geo[(G_DX + i) * nF + f] = (T)d->deltasUnit[3 * f + i];
        geo[G_LW0 * nF + f] = (T)d->linearWeights[2 * f];
        geo[G_LW1 * nF + f] = (T)d->linearWeights[2 * f + 1];
        for (int i = 0; i < 3; i++) {
            geo[(G_Q0X + i) * nF + f] = (T)d->quadraticWeights[6 * f + i];
            geo[(G_Q1X + i) * nF + f] = (T)d->quadraticWeights[6 * f + 3 + i];
        }
    }
    int rc;
    if ((rc = upload(ctx, &ctx->geo, geo)) != ADFVM_OK) return rc;
    std::vector<T> vol(nIC);
    for (long c = 0; c < nIC; c++) vol[c] = (T)d->volumes[c];
    if ((rc = upload(ctx, &ctx->vol, vol)) != ADFVM_OK) return rc;
    // Gauss-gradient geometry per cell slot, computed in FP64 and cast once:
    // the cell's outward S*n and the weight w' of the other side's value
    // (phi_f = phi_c (1 - w') + phi_other w'), streamed by the gradient
    // kernels instead of gathering face rows.
    std::vector<T> sg((size_t)24 * nIC);
    for (long c = 0; c < nIC; c++)
        for (int j = 0; j < 6; j++) {
            const long f = d->cellFaces[6 * c + j];
            const bool own = d->cellOwner[6 * c + j] != 0;
            const double sgn = own ? 1.0 : -1.0;
            const double w = d->weights[f];
            for (int k = 0; k < 3; k++) sg[(size_t)(4 * j + k) * nIC + c] = (T)(sgn * (d->areas[f] * d->normals[3 * f + k]));
            sg[(size_t)(4 * j + 3) * nIC + c] = (T)(own ? 1.0 - w : w);
        }
    if ((rc = upload(ctx, &ctx->sgeo, sg)) != ADFVM_OK) return rc;
    // The transposed gradient stencil for the gradient adjoint: cell c's
    // slot j face value enters the gradient of the other side o with weight
    // coef = (1 - w') and outward vector S*n_o, so d(grad_o)/d(w_c) through
    // that face is coef * S*n_o / V_o.  Across an internal face that is
    // -(1 - w') S*n_c / V_o (w' the other side's weight), which the kernel
    // forms from the slot geometry; across a cyclic face it is the paired
    // face fp seen from its owner o: S*n_fp and coef = 1 - w_fp, tabulated
    // here.
    std::vector<T> ng((size_t)18 * nIC, T(0));
    for (long c = 0; c < nIC; c++)
        for (int j = 0; j < 6; j++) {
            const long f = d->cellFaces[6 * c + j];
            if (f < ctx->nIF) continue;
            const adfvm_patch_desc& Pd = d->patches[ctx->hBPatch[f - ctx->nIF]];
            if (Pd.fluxKind != ADFVM_FLUX_COUPLED) continue;
            const long g = d->patches[Pd.neighbourPatch].startFace + (f - Pd.startFace);
            const long o = d->owner[g];
            const double coef = 1.0 - d->weights[g];
            for (int k = 0; k < 3; k++)
                ng[(size_t)(3 * j + k) * nIC + c] = (T)(coef * (d->areas[g] * d->normals[3 * g + k]) / d->volumes[o]);
        }
    if ((rc = upload(ctx, &ctx->ngeo, ng)) != ADFVM_OK) return rc;
    // boundary-condition values
    std::vector<T> vals;
    for (int p = 0; p < d->nPatches; p++) {
        const adfvm_patch_desc& P = d->patches[p];
        DPatch& D = ctx->hpatches[p];
        auto add = [&](const double* src, long cnt) -> long {
            if (!src) return -1;
            long off = (long)vals.size();
            for (long i = 0; i < cnt; i++) vals.push_back((T)src[i]);
            return off;
        };
        D.valU = add(P.valueU, 3L * P.nFaces);
        D.valT = add(P.valueT, P.nFaces);
        D.valp = add(P.valuep, P.nFaces);
        D.valTt = add(P.Tt, P.nFaces);
        D.valpt = add(P.pt, P.nFaces);
        D.valDir = add(P.direction, 3L * P.nFaces);
        D.hasDir = P.direction != nullptr;
    }
    if ((rc = upload(ctx, &ctx->bcval, vals)) != ADFVM_OK) return rc;
    // work buffers
    size_t st = 5 * nIC * sizeof(T), gr = (size_t)GREC * nIC * sizeof(T);
    HIPCHK(ctx, hipMalloc(&ctx->ghost, std::max<size_t>(1, 5 * nBF) * sizeof(T)));
    HIPCHK(ctx, hipMemset(ctx->ghost, 0, std::max<size_t>(1, 5 * nBF) * sizeof(T)));
    HIPCHK(ctx, hipMalloc(&ctx->u1, st));
    HIPCHK(ctx, hipMalloc(&ctx->u2, st));
    HIPCHK(ctx, hipMalloc(&ctx->grad0, gr));
    HIPCHK(ctx, hipMalloc(&ctx->grad1, gr));
    HIPCHK(ctx, hipMalloc(&ctx->grad2, gr));
    HIPCHK(ctx, hipMalloc(&ctx->adjAcc, st));
    HIPCHK(ctx, hipMalloc(&ctx->adjG, (size_t)AREC * nIC * sizeof(T)));
    HIPCHK(ctx, hipMalloc(&ctx->fbuf, (size_t)FREC * nF * sizeof(T)));
    HIPCHK(ctx, hipMalloc(&ctx->frec, (size_t)AREC * ctx->nPlanes * nIC * sizeof(T)));
    return ADFVM_OK;
}

template <typename T> T* P(void* p) { return (T*)p; }

// ghost refresh for state u (no-op for meshes with only cyclic patches)
template <typename T> void launch_ghosts(adfvm_ctx* c, const T* u, hipStream_t s) {
    if (!c->hasGhostKernel) return;
    KTimer kt(c, "ghost_kernel", s);
    ghost_kernel<T><<<grid_for(c->nBF), BLOCK, 0, s>>>(dmesh<T>(c), phys_cast<T>(c->phd), u, P<T>(c->ghost));
}

template <typename T> void launch_grad(adfvm_ctx* c, const T* u, T* grad, hipStream_t s) {
    KTimer kt(c, "grad_kernel", s);
    with_block(c->bs[K_GRAD], [&](auto B) {
        with_stream(c->stream, [&](auto S) {
            with_ghosts(c->hasGhostKernel, [&](auto G) {
                constexpr int BS = decltype(B)::value;
                constexpr bool ST = decltype(S)::value, GH = decltype(G)::value;
                grad_kernel<T, BS, ST, GH><<<grid_bs(c->nIC, BS), BS, 0, s>>>(dmesh<T>(c), phys_cast<T>(c->phd),
                                                                               u, P<T>(c->ghost), grad);
            });
        });
    });
}

template <typename T> void launch_face_flux(adfvm_ctx* c, const T* u, const T* grad, hipStream_t s) {
    KTimer kt(c, "face_flux_kernel", s);
    with_block(c->bs[K_FLUX], [&](auto B) {
        with_stream(c->stream, [&](auto S) {
            with_ghosts(c->hasGhostKernel, [&](auto G) {
                constexpr int BS = decltype(B)::value;
                constexpr bool ST = decltype(S)::value, GH = decltype(G)::value;
                face_flux_kernel<T, BS, ST, GH><<<grid_bs(c->nF, BS), BS, 0, s>>>(
                    dmesh<T>(c), phys_cast<T>(c->phd), u, grad, P<T>(c->ghost), P<T>(c->fbuf));
            });
        });
    });
}

// u_{i+1} = a0*u0 + ai*u_i - beta*(L-S)*dt.  Every stage of the tables
// below with a0 != 0 has a0 + ai = 1, and stage_kernel evaluates it as
// u0 + (ai*(u_i - u0) - beta*(L-S)*dt), so sum(V*u) is kept up to unbiased
// rounding.  (Taken as a0*u0 + ai*u_i with the coefficients rounded apart,
// SSP-RK3's 1/3 and 2/3 both round up in FP32, to a sum of 1 + 2^-25: the
// state would grow by that factor every step.)  The adjoint uses a0 formed
// as 1 - ai in the compute precision, the coefficient the primal applies.
struct Stage {
    double a0, ai, beta;
    template <typename T> T a0_as() const { return a0 == 0.0 ? T(0) : T(1) - (T)ai; }
};

template <typename T, int MODE>
void launch_stage(adfvm_ctx* c, const T* ui, const T* u0, const Stage& st, double dt, const T* S, T* out,
                  T* dtcCell, double* dtcPartial, hipStream_t s) {
    KTimer kt(c, MODE == 0 ? "stage_kernel" : "residual_kernel", s);
    with_block(c->bs[K_STAGE], [&](auto B) {
        with_stream(c->stream, [&](auto SP) {
            constexpr int BS = decltype(B)::value;
            constexpr bool ST = decltype(SP)::value;
            stage_kernel<T, MODE, BS, ST><<<grid_bs(c->nIC, BS), BS, 0, s>>>(
                dmesh<T>(c), P<T>(c->fbuf), ui, u0, (T)st.ai, (T)(-st.beta), (T)dt, S, out, dtcCell,
                dtcPartial);
        });
    });
}

// Fluxes of all faces at state ui (gradients grad) and the stage update or
// residual.  (A fused cell-centric variant that evaluates every face from
// both sides instead of storing face records measured 2.3x slower: the face
// flux is ALU-heavy.)
template <typename T, int MODE>
void launch_flux_and_stage(adfvm_ctx* c, const T* ui, const T* grad, const T* u0, const Stage& st, double dt,
                           const T* S, T* out, T* dtcCell, double* dtcPartial, hipStream_t s) {
    launch_face_flux<T>(c, ui, grad, s);
    launch_stage<T, MODE>(c, ui, u0, st, dt, S, out, dtcCell, dtcPartial, s);
}

template <typename T> void launch_objective(adfvm_ctx* c, const T* u, double* out, hipStream_t s) {
    KTimer kt(c, "objective", s);
    if (c->objective == ADFVM_OBJ_KINETIC_ENERGY && c->nSeg > 1) {
        ke_seg_kernel<T><<<c->nSegBlocks, BLOCK, 0, s>>>(dmesh<T>(c), u, c->segCell, c->segBlock, c->blockSeg,
                                                         c->partial);
        finalize_seg_kernel<<<c->nSeg, BLOCK, 0, s>>>(c->partial, c->segBlock, out);
    } else if (c->objective == ADFVM_OBJ_KINETIC_ENERGY) {
        int g = grid_for(c->nIC);
        ke_kernel<T><<<g, BLOCK, 0, s>>>(dmesh<T>(c), u, c->partial);
        finalize_kernel<0><<<1, BLOCK, 0, s>>>(c->partial, g, out);
    } else if (c->objective == ADFVM_OBJ_DRAG) {
        const DPatch& Pp = c->hpatches[c->objectivePatch];
        int g = grid_for(Pp.n);
        drag_kernel<T><<<g, BLOCK, 0, s>>>(dmesh<T>(c), phys_cast<T>(c->phd), u, Pp.start, Pp.n, c->partial);
        finalize_kernel<0><<<1, BLOCK, 0, s>>>(c->partial, g, out);
    } else {
        fill_kernel<<<grid_for(c->nSeg), BLOCK, 0, s>>>(out, c->nSeg, 0.0);   // one zero per segment
    }
}

void stages(int integrator, std::vector<Stage>& st) {
    st.clear();
    if (integrator == ADFVM_RK_EULER) { st.push_back({0, 1, 1}); return; }
    st.push_back({0.0, 1.0, 1.0});
    st.push_back({3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0});
    st.push_back({1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0});
}

// A step tape holds the forward step's intermediate stage states (all
// stages but the first, whose state is the step's input) and every stage's
// gradient records: [(ns-1) x 5 x nIC][ns x GREC x nIC] scalars.
// The stage whose state the step's objective is evaluated on: the first
// intermediate stage (density.py:412-413; the state a tape holds first), or
// the step's start state u_n when the caller asked for it
// (ADFVM_OBJSTATE_START; no counterpart in the reference).  dtc is taken on
// the first intermediate stage either way.
int objective_stage(const adfvm_ctx* c, int ns) {
    return (ns > 1 && c->objState == ADFVM_OBJSTATE_STAGE1) ? 1 : 0;
}

long tape_scalars(const adfvm_ctx* c, int ns) { return ((long)(ns - 1) * 5 + (long)ns * GREC) * c->nIC; }

template <typename T> void tape_slots(const adfvm_ctx* c, int ns, T* tape, T** states, T** grads) {
    for (int i = 0; i < ns - 1; i++) states[i] = tape + (long)i * 5 * c->nIC;
    T* g = tape + (long)(ns - 1) * 5 * c->nIC;
    for (int i = 0; i < ns; i++) grads[i] = g + (long)i * GREC * c->nIC;
}

template <typename T>
int primal_step_t(adfvm_ctx* c, T* u, double dt, const T* S, double* dtc_out, double* obj_out, T* tape,
                  hipStream_t s) {
    std::vector<Stage> st;
    stages(c->integrator, st);
    const int ns = (int)st.size();
    T* bufs[3] = {P<T>(c->u1), P<T>(c->u2), u};
    T* grads[3] = {P<T>(c->grad0), P<T>(c->grad1), P<T>(c->grad2)};
    if (tape) tape_slots(c, ns, tape, bufs, grads);
    const T* ui = u;
    const int objStage = objective_stage(c, ns);
    const int dtcStage = ns > 1 ? 1 : 0;    // density.py:412-413
    for (int i = 0; i < ns; i++) {
        // the last stage writes u in place (each cell reads only its own u0
        // entry); a single-stage scheme reads u's neighbours, so it goes
        // through u1 and is copied back
        T* out = (i == ns - 1 && ns > 1) ? u : bufs[i];
        if (i == objStage && obj_out) launch_objective<T>(c, ui, obj_out, s);
        launch_grad<T>(c, ui, grads[i], s);
        const bool wantDtc = (i == dtcStage) && dtc_out;
        const T* u0 = (i > 0 && st[i].a0 != 0.0) ? (const T*)u : nullptr;
        int g = grid_bs(c->nIC, c->bs[K_STAGE]);
        launch_flux_and_stage<T, 0>(c, ui, grads[i], u0, st[i], dt, S, out, nullptr, wantDtc ? c->partial : nullptr, s);
        if (wantDtc) finalize_kernel<1><<<1, BLOCK, 0, s>>>(c->partial, g, dtc_out);
        ui = out;
    }
    if (ns == 1) copy_kernel<T><<<grid_for(5L * c->nIC), BLOCK, 0, s>>>(5L * c->nIC, bufs[0], u);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

}  // namespace

// ------------------------------------------------------------- C ABI -----

extern "C" int adfvm_ctx_create(const adfvm_solver_desc* d, const adfvm_physics* ph, adfvm_ctx** out) {
    if (!d || !ph || !out) return ADFVM_ERR_INVALID;
    *out = nullptr;
    adfvm_ctx* c = new adfvm_ctx();
    auto bad = [&](const std::string& m) { c->err = m; return ADFVM_ERR_INVALID; };
    if (ph->precision != 4 && ph->precision != 8) { int r = bad("precision must be 4 or 8"); *out = c; return r; }
    c->prec = ph->precision;
    c->nIC = d->sizes.nInternalCells; c->nIF = d->sizes.nInternalFaces; c->nF = d->sizes.nFaces;
    c->nBF = d->sizes.nBoundaryFaces; c->nC = d->sizes.nCells;
    c->nPatches = d->nPatches;
    c->integrator = ph->integrator;
    c->objective = ph->objective;
    c->objectivePatch = ph->objectivePatch;
    c->phd.gamma = ph->gamma; c->phd.Cp = ph->Cp; c->phd.Cv = ph->Cp / ph->gamma; c->phd.Pr = ph->Pr;
    c->phd.invCv = 1.0 / c->phd.Cv; c->phd.small = ph->small;
    c->phd.mu0 = ph->mu0; c->phd.muA = ph->muA; c->phd.muTs = ph->muTs; c->phd.muModel = ph->muModel;
    c->phd.riemann = ph->riemann; c->phd.boundaryRiemann = ph->boundaryRiemann;
    c->phd.limit = ph->limiter;
    *out = c;
    if (c->nIC <= 0 || c->nF <= 0 || c->nBF < 0) return bad("bad mesh sizes");
    if (c->objective == ADFVM_OBJ_DRAG && (c->objectivePatch < 0 || c->objectivePatch >= c->nPatches))
        return bad("drag objective needs a valid patch index");
    const long nIC = c->nIC, nF = c->nF, nIF = c->nIF, nBF = c->nBF;

    // boundary face -> patch, patch table
    std::vector<int> bPatch(nBF, -1);
    c->hpatches.resize(d->nPatches);
    for (int p = 0; p < d->nPatches; p++) {
        const adfvm_patch_desc& Pd = d->patches[p];
        if (Pd.startFace < nIF || Pd.startFace + Pd.nFaces > nF) return bad("patch outside boundary faces");
        for (int k = 0; k < Pd.nFaces; k++) bPatch[Pd.startFace - nIF + k] = p;
        DPatch& D = c->hpatches[p];
        D.start = Pd.startFace; D.n = Pd.nFaces; D.fluxKind = Pd.fluxKind;
        D.bcU = Pd.bcU; D.bcT = Pd.bcT; D.bcp = Pd.bcp; D.cbc = Pd.cbc;
        const bool cyc = Pd.fluxKind == ADFVM_FLUX_COUPLED;
        if (cyc && (Pd.neighbourPatch < 0 || Pd.neighbourPatch >= d->nPatches ||
                    d->patches[Pd.neighbourPatch].nFaces != Pd.nFaces))
            return bad("coupled patch without a matching neighbour patch");
        if (cyc && (Pd.bcU != ADFVM_BC_CYCLIC || Pd.bcT != ADFVM_BC_CYCLIC || Pd.bcp != ADFVM_BC_CYCLIC))
            return bad("coupled patches need cyclic conditions on U, T and p");
        if (!cyc && (Pd.bcU == ADFVM_BC_CYCLIC || Pd.bcT == ADFVM_BC_CYCLIC || Pd.bcp == ADFVM_BC_CYCLIC))
            return bad("cyclic condition on a non-coupled patch");
        if (Pd.cbc == ADFVM_CBC_UPT && (!Pd.valueU || !Pd.valueT || !Pd.valuep))
            return bad("CBC_UPT needs U0, T0, p0 values");
        if (Pd.cbc == ADFVM_CBC_TOTAL_PT && (!Pd.Tt || !Pd.pt)) return bad("CBC_TOTAL_PT needs Tt and pt");
        auto needs = [](int bc) { return bc == ADFVM_BC_FIXED_VALUE || bc == ADFVM_BC_CALCULATED; };
        if (Pd.cbc == ADFVM_CBC_NONE && ((needs(Pd.bcU) && !Pd.valueU) || (needs(Pd.bcT) && !Pd.valueT) ||
                                         (needs(Pd.bcp) && !Pd.valuep)))
            return bad("fixedValue/calculated condition without values");
        if (!cyc && Pd.nFaces > 0) c->hasGhostKernel = true;
    }
    for (long b = 0; b < nBF; b++)
        if (bPatch[b] < 0) return bad("boundary face not covered by a patch");

    // slot tables
    std::vector<int> sFace(6 * nIC), sOther(6 * nIC), sPair(6 * nIC, -1);
    std::vector<unsigned char> sFlags(6 * nIC);
    std::vector<int> cellNb(6 * nIC);
    for (long cc = 0; cc < nIC; cc++) {
        for (int j = 0; j < 6; j++) {
            int f = d->cellFaces[6 * cc + j];
            if (f < 0 || f >= nF) return bad("cellFaces out of range");
            int own = d->cellOwner[6 * cc + j];
            int other, kind;
            if (f < nIF) {
                kind = SLOT_INTERNAL;
                other = own ? d->neighbour[f] : d->owner[f];
            } else {
                if (!own) return bad("boundary face in a neighbour slot");
                int b = f - nIF;
                const adfvm_patch_desc& Pd = d->patches[bPatch[b]];
                if (Pd.fluxKind == ADFVM_FLUX_COUPLED) {
                    kind = SLOT_COUPLED;
                    int nbStart = d->patches[Pd.neighbourPatch].startFace;
                    int pf = nbStart + (f - Pd.startFace);
                    other = d->owner[pf];
                    sPair[j * nIC + cc] = pf;
                } else {
                    kind = Pd.fluxKind == ADFVM_FLUX_CHARACTERISTIC ? SLOT_CHARACTERISTIC : SLOT_BOUNDARY;
                    other = -1 - b;
                }
            }
            sFace[j * nIC + cc] = f;
            sOther[j * nIC + cc] = other;
            const bool copy = kind == SLOT_BOUNDARY && ghost_copies_owner(d->patches[bPatch[f - nIF]]);
            sFlags[j * nIC + cc] = (unsigned char)((own ? 1 : 0) | (kind << 1) | (copy ? SLOT_GHOST_COPY : 0));
            cellNb[j * nIC + cc] = d->cellNeighbours[6 * cc + j];
        }
    }
    // face tables: the neighbour side's value source and the flux kind
    std::vector<int> fOther(nF);
    std::vector<unsigned char> fKind(nF);
    for (long f = 0; f < nF; f++) {
        if (f < nIF) { fOther[f] = d->neighbour[f]; fKind[f] = FACE_INTERIOR; continue; }
        const long b = f - nIF;
        const adfvm_patch_desc& Pd = d->patches[bPatch[b]];
        if (Pd.fluxKind == ADFVM_FLUX_COUPLED) {
            fOther[f] = d->owner[d->patches[Pd.neighbourPatch].startFace + (f - Pd.startFace)];
            fKind[f] = FACE_INTERIOR;
        } else {
            fOther[f] = (int)(-1 - b);
            fKind[f] = Pd.fluxKind == ADFVM_FLUX_CHARACTERISTIC ? FACE_CHARACTERISTIC : FACE_BOUNDARY;
        }
    }
    // Face adjoint records sit in planes of nIC records: plane j holds, for
    // every cell, the record of its slot j (its own side of an owned face,
    // the neighbour side of a face it neighbours); planes 6, 7, ... hold, in
    // slot order, the neighbour sides of the paired faces of a cell's cyclic
    // slots (on which the cell is the ghost).  A cell's records for one slot
    // are then at the same index of every plane, so neighbouring threads of
    // the gather read neighbouring records.
    std::vector<int> posP(nF, -1), posN(nF, -1);
    int maxExtra = 0;
    for (long cc = 0; cc < nIC; cc++) {
        int e = 0;
        for (int j = 0; j < 6; j++) {
            const int fl = sFlags[j * nIC + cc];
            const int f = sFace[j * nIC + cc];
            int& slot = (fl & 1) ? posP[f] : posN[f];
            if (slot != -1) return bad("face in two slots of the same side");
            slot = (int)(j * nIC + cc);
            if (slot_kind(fl) == SLOT_COUPLED) e++;
        }
        maxExtra = std::max(maxExtra, e);
    }
    if ((long)(6 + maxExtra) * nIC > 0x7fffffffL) return bad("mesh too large for 32-bit record indices");
    for (long cc = 0; cc < nIC; cc++) {
        int e = 0;
        for (int j = 0; j < 6; j++) {
            if (slot_kind(sFlags[j * nIC + cc]) != SLOT_COUPLED) continue;
            const int g = sPair[j * nIC + cc];
            if (posN[g] != -1) return bad("face with two neighbour-side slots");
            posN[g] = (int)((6 + e++) * nIC + cc);
        }
    }
    for (long f = 0; f < nF; f++) {
        if (posP[f] < 0) return bad("face without an owner slot");
        if (fKind[f] == FACE_INTERIOR && posN[f] < 0) return bad("face without a neighbour-side slot");
    }
    c->nPlanes = 6 + maxExtra;
    int rc;
    // the three groups: characteristic faces, boundary-kind faces whose
    // ghost reads its patch, boundary-kind faces whose ghost copies the owner
    std::vector<int> group(nBF, -1);
    for (long b = 0; b < nBF; b++) {
        if (fKind[nIF + b] == FACE_CHARACTERISTIC) group[b] = 0;
        else if (fKind[nIF + b] == FACE_BOUNDARY) group[b] = ghost_copies_owner(d->patches[bPatch[b]]) ? 2 : 1;
    }
    for (int k = 0; k < 3; k++) {
        std::vector<int> blk;                   // (first boundary index, count) per block
        for (long b = 0; b < nBF;) {
            if (group[b] != k) { b++; continue; }
            long e = b;
            while (e < nBF && group[e] == k) e++;
            for (long b0 = b; b0 < e; b0 += BLOCK) {
                blk.push_back((int)b0);
                blk.push_back((int)std::min<long>(BLOCK, e - b0));
            }
            b = e;
        }
        c->nKindBlk[k] = (int)(blk.size() / 2);
        if ((rc = upload(c, (void**)&c->kindBlk[k], blk)) != ADFVM_OK) return rc;
    }
    // The face adjoint sums a cell's owner sides in LDS when one owner's
    // internal faces are consecutive (owners non-decreasing over the internal
    // faces) and are the cell's first slots, in face order: the sum is then
    // the gather's own sum over those slots.  Per block size, the face ranges
    // end at owner changes.
    c->ownerSum = true;
    for (long f = 1; f < nIF && c->ownerSum; f++)
        if (d->owner[f - 1] > d->owner[f]) c->ownerSum = false;
    for (long cc = 0; cc < nIC && c->ownerSum; cc++) {
        int prev = -1;
        bool more = true;              // still in the leading run of owned internal slots
        for (int j = 0; j < 6; j++) {
            const int fl = sFlags[j * nIC + cc];
            const bool ownInt = (fl & 1) && slot_kind(fl) == SLOT_INTERNAL;
            if (ownInt && (!more || sFace[j * nIC + cc] <= prev)) { c->ownerSum = false; break; }
            if (ownInt) prev = sFace[j * nIC + cc];
            else more = false;
        }
    }
    for (int k = 0; k < 4 && c->ownerSum; k++) {
        const long bs = kBlockChoices[k];
        std::vector<int> starts(1, 0);
        for (long f = 0; f < nIF;) {
            long e = std::min(f + bs, nIF);
            while (e < nIF && e > f && d->owner[e - 1] == d->owner[e]) e--;
            if (e == f) return bad("cell with more internal faces than a face adjoint block");
            starts.push_back((int)e);
            f = e;
        }
        c->fvjpNBlk[k] = (int)starts.size() - 1;
        if ((rc = upload(c, (void**)&c->fvjpBlk[k], starts)) != ADFVM_OK) return rc;
    }
    if ((rc = upload(c, (void**)&c->posP, posP)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->posN, posN)) != ADFVM_OK) return rc;
    std::vector<int> owner(d->owner, d->owner + nF), nbFull(d->neighbour, d->neighbour + nF);
    if ((rc = upload(c, (void**)&c->fOther, fOther)) != ADFVM_OK) return rc;
    {
        std::vector<unsigned char> fKindDev(fKind);
        for (long b = 0; b < nBF; b++)
            if (ghost_copies_owner(d->patches[bPatch[b]])) fKindDev[nIF + b] |= FACE_GHOST_COPY;
        if ((rc = upload(c, (void**)&c->fKind, fKindDev)) != ADFVM_OK) return rc;
    }
    if ((rc = upload(c, (void**)&c->sFace, sFace)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->sOther, sOther)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->sFlags, sFlags)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->sPair, sPair)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->cellNbFull, cellNb)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->owner, owner)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->nbFull, nbFull)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->bPatch, bPatch)) != ADFVM_OK) return rc;
    c->hBPatch = bPatch;
    rc = c->prec == 4 ? create_typed<float>(c, d) : create_typed<double>(c, d);
    if (rc != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->patches, c->hpatches)) != ADFVM_OK) return rc;
    c->stream = stream_auto(c);
    c->nPartial = grid_bs(std::max<long>(nIC, nF), kBlockChoices[0]) + 1;   // the smallest block size
    HIPCHK(c, hipMalloc((void**)&c->partial, c->nPartial * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&c->scal, 16 * sizeof(double)));
    return ADFVM_OK;
}

extern "C" void adfvm_ctx_destroy(adfvm_ctx* c) {
    if (!c) return;
    void* ptrs[] = {c->geo, c->vol, c->bcval, c->sFace, c->sOther, c->sFlags, c->sPair, c->owner, c->bPatch,
                    c->patches, c->cellNbFull, c->nbFull, c->ghost, c->u1, c->u2, c->grad0, c->grad1,
                    c->grad2, c->adjAcc, c->adjG,
                    c->partial, c->scal, c->fOther, c->fKind, c->fbuf, c->frec,
                    c->segCell, c->segBlock, c->blockSeg, c->posP, c->posN, c->sgeo, c->ngeo, c->visc,
                    c->fvjpBlk[0], c->fvjpBlk[1], c->fvjpBlk[2], c->fvjpBlk[3], c->kindBlk[0],
                    c->kindBlk[1], c->kindBlk[2]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : c->evPool) (void)hipEventDestroy(e);
    delete c;
}

extern "C" const char* adfvm_last_error(const adfvm_ctx* c) { return c ? c->err.c_str() : "null context"; }

#ifndef ADFVM_SOURCE_HASH
#define ADFVM_SOURCE_HASH "unknown"
#endif
extern "C" const char* adfvm_source_hash(void) { return ADFVM_SOURCE_HASH; }
extern "C" int adfvm_ctx_precision(const adfvm_ctx* c) { return c ? c->prec : 0; }

namespace {
template <typename T> int set_bc_values_t(adfvm_ctx* c, long off, const double* v, long n, hipStream_t s) {
    std::vector<T> h(v, v + n);
    HIPCHK(c, hipMemcpyAsync((T*)c->bcval + off, h.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return ADFVM_OK;
}
}  // namespace

extern "C" int adfvm_set_bc_values(adfvm_ctx* c, int32_t patch, int32_t slot, const double* values, int64_t n,
                                   void* stream) {
    if (!c || !values) return ADFVM_ERR_INVALID;
    if (patch < 0 || patch >= c->nPatches) return set_err(c, ADFVM_ERR_INVALID, "patch index out of range");
    const DPatch& D = c->hpatches[patch];
    long off;
    int k = 1;
    switch (slot) {
        case ADFVM_BCV_U: off = D.valU; k = 3; break;
        case ADFVM_BCV_T: off = D.valT; break;
        case ADFVM_BCV_P: off = D.valp; break;
        case ADFVM_BCV_TT: off = D.valTt; break;
        case ADFVM_BCV_PT: off = D.valpt; break;
        case ADFVM_BCV_DIRECTION: off = D.valDir; k = 3; break;
        default: return set_err(c, ADFVM_ERR_INVALID, "unknown boundary-value slot");
    }
    if (off < 0) return set_err(c, ADFVM_ERR_INVALID, "this patch was created without that value array");
    if (n != (int64_t)k * D.n) return set_err(c, ADFVM_ERR_INVALID, "value count is not nFaces x components");
    hipStream_t s = (hipStream_t)stream;
    return c->prec == 4 ? set_bc_values_t<float>(c, off, values, n, s) : set_bc_values_t<double>(c, off, values, n, s);
}

extern "C" int adfvm_set_segments(adfvm_ctx* c, int32_t nSeg, const int64_t* cellStart) {
    if (!c || nSeg < 1 || !cellStart) return ADFVM_ERR_INVALID;
    if (cellStart[0] != 0 || cellStart[nSeg] != c->nIC)
        return set_err(c, ADFVM_ERR_INVALID, "segments must start at cell 0 and end at nInternalCells");
    for (int k = 0; k < nSeg; k++)
        if (cellStart[k + 1] <= cellStart[k]) return set_err(c, ADFVM_ERR_INVALID, "empty or decreasing segment");
    if (nSeg > 1 && c->objective == ADFVM_OBJ_DRAG)
        return set_err(c, ADFVM_ERR_UNSUPPORTED, "segments with the drag objective");
    std::vector<long> cells(cellStart, cellStart + nSeg + 1);
    std::vector<int> first(nSeg + 1), owner;
    first[0] = 0;
    for (int k = 0; k < nSeg; k++) {
        int nb = grid_for(cells[k + 1] - cells[k]);
        first[k + 1] = first[k] + nb;
        for (int b = 0; b < nb; b++) owner.push_back(k);
    }
    if (first[nSeg] > c->nPartial) return set_err(c, ADFVM_ERR_INVALID, "too many segment blocks");
    for (void* p : {(void*)c->segCell, (void*)c->segBlock, (void*)c->blockSeg})
        if (p) (void)hipFree(p);
    c->segCell = nullptr; c->segBlock = nullptr; c->blockSeg = nullptr;
    int rc;
    if ((rc = upload(c, (void**)&c->segCell, cells)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->segBlock, first)) != ADFVM_OK) return rc;
    if ((rc = upload(c, (void**)&c->blockSeg, owner)) != ADFVM_OK) return rc;
    c->nSeg = nSeg;
    c->nSegBlocks = first[nSeg];
    return ADFVM_OK;
}

extern "C" int adfvm_segments(const adfvm_ctx* c) { return c ? c->nSeg : 0; }

extern "C" int adfvm_set_objective_state(adfvm_ctx* c, int32_t which) {
    if (!c) return ADFVM_ERR_INVALID;
    if (which != ADFVM_OBJSTATE_STAGE1 && which != ADFVM_OBJSTATE_START)
        return set_err(c, ADFVM_ERR_INVALID, "objective state must be ADFVM_OBJSTATE_STAGE1 or ADFVM_OBJSTATE_START");
    c->objState = which;
    return ADFVM_OK;
}

extern "C" int adfvm_objective_state(const adfvm_ctx* c) { return c ? c->objState : -1; }

#define DISPATCH(c, CALL_F, CALL_D) ((c)->prec == 4 ? (CALL_F) : (CALL_D))

extern "C" int adfvm_primal_step(adfvm_ctx* c, void* u, double dt, const void* S, double* dtc_out,
                                 double* obj_out, void* tape, void* stream) {
    if (!c || !u) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return DISPATCH(c, primal_step_t<float>(c, (float*)u, dt, (const float*)S, dtc_out, obj_out, (float*)tape, s),
                    primal_step_t<double>(c, (double*)u, dt, (const double*)S, dtc_out, obj_out, (double*)tape, s));
}

namespace {
template <typename T> int gradients_t(adfvm_ctx* c, const T* u, T* grad, hipStream_t s) {
    launch_grad<T>(c, u, P<T>(c->grad0), s);
    grad_rows_kernel<T><<<grid_for(c->nIC), BLOCK, 0, s>>>(c->nIC, P<T>(c->grad0), grad);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}
template <typename T> int residual_t(adfvm_ctx* c, const T* u, T* R, T* dtcCell, hipStream_t s) {
    launch_grad<T>(c, u, P<T>(c->grad0), s);
    launch_flux_and_stage<T, 1>(c, u, P<T>(c->grad0), nullptr, Stage{0, 0, 0}, 0.0, nullptr, R, dtcCell, nullptr, s);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}
template <typename T> int objective_t(adfvm_ctx* c, const T* u, double* out, hipStream_t s) {
    launch_objective<T>(c, u, out, s);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}
template <typename T> int prim_t(adfvm_ctx* c, const T* u, T* prim, hipStream_t s) {
    launch_ghosts<T>(c, u, s);
    prim_ghost_kernel<T><<<grid_for((long)c->nIC + c->nBF), BLOCK, 0, s>>>(dmesh<T>(c), phys_cast<T>(c->phd), u,
                                                                           P<T>(c->ghost), prim);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}
}  // namespace

extern "C" int adfvm_gradients(adfvm_ctx* c, const void* u, void* grad, void* stream) {
    if (!c || !u || !grad) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return DISPATCH(c, gradients_t<float>(c, (const float*)u, (float*)grad, s),
                    gradients_t<double>(c, (const double*)u, (double*)grad, s));
}

extern "C" int adfvm_residual(adfvm_ctx* c, const void* u, void* R, void* dtcCell, void* stream) {
    if (!c || !u || !R) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return DISPATCH(c, residual_t<float>(c, (const float*)u, (float*)R, (float*)dtcCell, s),
                    residual_t<double>(c, (const double*)u, (double*)R, (double*)dtcCell, s));
}

extern "C" int adfvm_objective(adfvm_ctx* c, const void* u, double* out, void* stream) {
    if (!c || !u || !out) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return DISPATCH(c, objective_t<float>(c, (const float*)u, out, s), objective_t<double>(c, (const double*)u, out, s));
}

extern "C" int adfvm_primitive_ghosts(adfvm_ctx* c, const void* u, void* prim, void* stream) {
    if (!c || !u || !prim) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return DISPATCH(c, prim_t<float>(c, (const float*)u, (float*)prim, s), prim_t<double>(c, (const double*)u, (double*)prim, s));
}

extern "C" int adfvm_grad_cell(adfvm_ctx* c, const void* phi, int32_t k, void* grad, void* stream) {
    if (!c || !phi || !grad || k <= 0) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    long nC = (long)c->nIC + c->nBF;
    if (c->prec == 4)
        grad_field_kernel<float><<<grid_for(c->nIC), BLOCK, 0, s>>>(dmesh<float>(c), c->cellNbFull, (const float*)phi, k, nC, (float*)grad);
    else
        grad_field_kernel<double><<<grid_for(c->nIC), BLOCK, 0, s>>>(dmesh<double>(c), c->cellNbFull, (const double*)phi, k, nC, (double*)grad);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

extern "C" int adfvm_interp_second_order(adfvm_ctx* c, const void* phi, const void* grad, int32_t k, int32_t swap,
                                         void* out, void* stream) {
    if (!c || !phi || !grad || !out || k <= 0 || (swap != 0 && swap != 1)) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    long nC = (long)c->nIC + c->nBF;
    if (c->prec == 4)
        interp_kernel<float><<<grid_for(c->nF), BLOCK, 0, s>>>(dmesh<float>(c), c->nbFull, (const float*)phi, (const float*)grad, k, nC, swap, (float*)out);
    else
        interp_kernel<double><<<grid_for(c->nF), BLOCK, 0, s>>>(dmesh<double>(c), c->nbFull, (const double*)phi, (const double*)grad, k, nC, swap, (double*)out);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

extern "C" int adfvm_dot(adfvm_ctx* c, const void* a, const void* b, int64_t n, double* out, void* stream) {
    if (!c || !a || !b || !out || n < 0) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int g = std::min<long>(std::max<long>(grid_for(n), 1), c->nPartial);
    if (c->prec == 4) dot_kernel<float><<<g, BLOCK, 0, s>>>((const float*)a, (const float*)b, n, c->partial);
    else dot_kernel<double><<<g, BLOCK, 0, s>>>((const double*)a, (const double*)b, n, c->partial);
    finalize_kernel<0><<<1, BLOCK, 0, s>>>(c->partial, g, out);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

// ---- the allocation-free dot product ------------------------------------
// Block partials of adfvm_dot_noalloc live in this file-scope array, one
// copy per GPU, held until the process ends.  In host code its name is not a
// device address: the address each device holds it at is looked up once per
// device (hipGetSymbolAddress with that device current) and only that
// pointer is passed to kernels.
constexpr long kScratchDoubles = 1L << 19;        // 4 MiB: n <= 2^19 * 256 scalars
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices];

// The current device's copy of scratch (nullptr and a message on failure).
double* scratch_on_current_device(std::string& err) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        err = "no current HIP device (or device index >= 64)";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[dev] = (double*)ptr;
    }
    return g_scratch_ptr[dev];
}
thread_local std::string g_noalloc_err;
}  // namespace

extern "C" const char* adfvm_noalloc_last_error(void) { return g_noalloc_err.c_str(); }

extern "C" int adfvm_dot_noalloc(int32_t precision, const void* a, const void* b, int64_t n, double* out,
                                 void* stream) {
    g_noalloc_err.clear();
    if ((n > 0 && (!a || !b)) || !out || n < 0 || (precision != 4 && precision != 8)) {
        g_noalloc_err = "adfvm_dot_noalloc: null pointer, negative n or precision not 4/8";
        return ADFVM_ERR_INVALID;
    }
    const long g = std::max<long>(grid_for(n), 1);
    if (g > kScratchDoubles) {
        g_noalloc_err = "adfvm_dot_noalloc: n = " + std::to_string((long long)n) + " exceeds the maximum " +
                        std::to_string((long long)(kScratchDoubles * BLOCK)) + " scalars";
        return ADFVM_ERR_INVALID;
    }
    double* part = scratch_on_current_device(g_noalloc_err);
    if (!part) return ADFVM_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    // every one of the g partials is written by dot_kernel before
    // finalize_kernel reads it, so nothing needs clearing
    if (precision == 4) dot_kernel<float><<<(int)g, BLOCK, 0, s>>>((const float*)a, (const float*)b, n, part);
    else dot_kernel<double><<<(int)g, BLOCK, 0, s>>>((const double*)a, (const double*)b, n, part);
    finalize_kernel<0><<<1, BLOCK, 0, s>>>(part, (int)g, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_noalloc_err = std::string("adfvm_dot_noalloc: ") + hipGetErrorString(e);
        return ADFVM_ERR_HIP;
    }
    return ADFVM_OK;
}

extern "C" int adfvm_adjoint_energy(adfvm_ctx* c, const void* ua, const void* u, double* out, void* stream) {
    if (!c || !ua || !u || !out) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int g = grid_for(c->nIC);
    if (c->prec == 4) energy_kernel<float><<<g, BLOCK, 0, s>>>(dmesh<float>(c), phys_cast<float>(c->phd), (const float*)ua, (const float*)u, c->partial);
    else energy_kernel<double><<<g, BLOCK, 0, s>>>(dmesh<double>(c), phys_cast<double>(c->phd), (const double*)ua, (const double*)u, c->partial);
    finalize_kernel<0><<<1, BLOCK, 0, s>>>(c->partial, g, out);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

extern "C" int adfvm_pack_state(adfvm_ctx* c, const void* rho, const void* rhoU, const void* rhoE, void* u, void* stream) {
    if (!c || !rho || !rhoU || !rhoE || !u) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    if (c->prec == 4) pack_kernel<float><<<grid_for(c->nIC), BLOCK, 0, s>>>(c->nIC, (const float*)rho, (const float*)rhoU, (const float*)rhoE, (float*)u);
    else pack_kernel<double><<<grid_for(c->nIC), BLOCK, 0, s>>>(c->nIC, (const double*)rho, (const double*)rhoU, (const double*)rhoE, (double*)u);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

extern "C" int adfvm_unpack_state(adfvm_ctx* c, const void* u, void* rho, void* rhoU, void* rhoE, void* stream) {
    if (!c || !rho || !rhoU || !rhoE || !u) return ADFVM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    if (c->prec == 4) unpack_kernel<float><<<grid_for(c->nIC), BLOCK, 0, s>>>(c->nIC, (const float*)u, (float*)rho, (float*)rhoU, (float*)rhoE);
    else unpack_kernel<double><<<grid_for(c->nIC), BLOCK, 0, s>>>(c->nIC, (const double*)u, (double*)rho, (double*)rhoU, (double*)rhoE);
    HIPCHK(c, hipGetLastError());
    return ADFVM_OK;
}

#include "mesh_adjoint.inc"
#include "adjoint_kernels.inc"
#include "viscosity.inc"
