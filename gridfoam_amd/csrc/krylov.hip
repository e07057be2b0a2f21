// Krylov pressure/momentum solves on the sliced-ELL face-neighbour operator.
//
// Restates solvers/cg.py:75-174 (PCG), solvers/bicgstab.py:76-272
// (right-preconditioned BiCGSTAB with restarts), solvers/preconditioners.py:
// 76-96 (Jacobi) and solvers/base.py:140-162 (threshold max(atol, rtol*|r0|),
// strict <).
//
// Synchronisation design for the 8-XCD MI355X (no intra-kernel fences):
//  * a kernel that needs a global sum writes one partial per workgroup; the
//    NEXT kernel reduces those partials in every workgroup (same fixed
//    order everywhere, so every workgroup derives the same scalars);
//  * solver state (scalars, phase, counters) is double-buffered by launch
//    sequence number s: launch s reads state[(s-1)&1] and its workgroup 0
//    writes state[s&1], so no workgroup reads what another one writes;
//  * a component that has finished turns every later kernel into a no-op,
//    and the host polls the state every `check_every` iterations while the
//    next chunk is already queued.
// Vectors are component-major [nrhs][C] so one SpMV streams the matrix once
// for all components; values are repacked once per solve into slice-of-64
// ELL order (one slice = one wavefront, coalesced column/value loads).
#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>
#include <type_traits>
#include <cstdlib>
#include <cmath>

#include "gf_common.h"

namespace gf {
namespace {

enum Phase : int { PH_ACTIVE = 0, PH_RESTART = 1, PH_DONE = 3 };

struct KState {
    double thresh, norm0, final_res;
    double rz_old, alpha, beta;
    double rho_old, omega, rhat2;
    double scale;  // mixed precision: true residual at the last refinement
    double rel;    // mixed precision: this solve's refinement trigger (see k_mxf_refine)
    double rec;    // mixed precision: the FP32 recursive residual that triggered the refinement
    int iter, phase, converged, restarts;
    int fresh;  // PCG: 1 = first direction (p = z), 2 = residual just replaced, 0 = steady
    int pcur;   // PCG: which of the two p buffers holds the current direction
    int iter_ref;  // mixed precision: the iteration count at the last refinement
};

constexpr int kMaxRhs = 3;
constexpr int kNQ = 3;        // partial quantities per rhs
constexpr int RB = 1024;      // threads per workgroup in solver kernels
constexpr int NBMAX = 1024;   // cap on the vector-kernel workgroup count

// Workgroups used by the vector kernels: one per CU on the 256-CU MI355X by
// default (GRIDFOAM_KRYLOV_WG overrides it for tuning).
int solver_workgroups(int C) {
    static int cap = [] {
        const char *e = getenv("GRIDFOAM_KRYLOV_WG");
        int v = e ? atoi(e) : 256;
        return v < 1 ? 1 : (v > NBMAX ? NBMAX : v);
    }();
    const int need = (C + RB - 1) / RB;
    return need < cap ? need : cap;
}

constexpr int SB = 256;  // threads per workgroup in SpMV kernels (one row per thread)

struct Ops {
    const int32_t *sell_ptr, *sell_col;
    const int32_t *bnbr;  // brick neighbours [nbs*6] (solver row order), null without bricks
    const int32_t *ecol;  // explicit-block columns [n_expl*64]
    int nbrick;           // brick slices: solver rows [0, 64 nbrick) are bricks (0: none)
    int C;
    int nb;       // BiCGSTAB vector kernels: workgroups (<= NBMAX)
    int nbs;      // BiCGSTAB SpMV kernels: workgroups (ceil(C / SB))
    int tb;       // PCG SpMV-type kernels: rows (= threads) per workgroup, 256, 512 or 1024
    int gs;       // PCG SpMV-type kernels: workgroups (ceil(C / tb) rounded up to a multiple of 8)
    int gd;       // PCG direction kernel: workgroups (its waves carry bricks or sliced-ELL slices)
    int gu;       // PCG vector kernels: workgroups (a multiple of 8)
    int pstride;  // partials per (rhs, quantity): the largest workgroup count
    int norm_inf, max_iter, max_restart;
    double atol, rtol;
    double reliable;   // mixed PCG: refine once |r| < reliable * (last true residual)
    int reliable_run;  // mixed PCG: ... or once this many FP32 iterations ran since the last refinement
    KState *st;        // [2][kMaxRhs]
    double *partials;  // [2][kMaxRhs][kNQ][pstride]
    int *poll;         // last launch of a chunk: pinned host word <- (s << 1) | all components done
};

inline int round8(int n) { return (n + 7) / 8 * 8; }
inline int spmv_workgroups(int C) { return (C + SB - 1) / SB; }
// Rows per workgroup of the PCG SpMV-type kernels (GRIDFOAM_PCG_TILE overrides it for tuning).
int pcg_tile() {
    static int tb = [] {
        const char *e = getenv("GRIDFOAM_PCG_TILE");
        const int v = e ? atoi(e) : 256;
        return (v == 512 || v == 1024) ? v : 256;
    }();
    return tb;
}
inline int pcg_spmv_workgroups(int C) { return round8((C + pcg_tile() - 1) / pcg_tile()); }
// Direction kernel: one wave per kBricks bricks, then one wave per remaining slice of 64 rows.
int pcg_dir_workgroups(int C, int nbrick);
inline int pcg_vector_workgroups(int C) { return round8(solver_workgroups(C)); }
inline int partial_stride(int C) {
    int m = solver_workgroups(C);
    for (int v : {spmv_workgroups(C), round8(spmv_workgroups(C)), pcg_vector_workgroups(C)}) m = v > m ? v : m;
    return m;  // pcg_spmv_workgroups(C) <= round8(spmv_workgroups(C)) for every tile size
}

// Tile of workgroup b in a grid of G (a multiple of 8) workgroups.  The
// hardware hands workgroups to the 8 XCDs round-robin, so XCD k = b % 8 gets
// the contiguous tiles [k G/8, (k+1) G/8): each XCD's L2 then holds one
// Morton-contiguous part of the mesh, where nearly all gathered neighbours lie.
__device__ __forceinline__ int xcd_tile(int b, int G) { return (b & 7) * (G >> 3) + (b >> 3); }

// Contiguous rows of tile t when C rows are split into G tiles (slices of 64 kept whole).
__device__ __forceinline__ void chunk_range(int C, int G, int t, int &lo, int &hi) {
    const int per = ((C + G - 1) / G + 63) & ~63;
    lo = min(C, t * per);
    hi = min(C, lo + per);
}

// Workgroup barrier for LDS traffic only.  __syncthreads() also waits for
// every global load the wave has in flight; here loads issued before the
// barrier stay in flight while the workgroup exchanges values through LDS.
__device__ __forceinline__ void lds_sync() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left alone
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ double nrm_part(double v, bool inf) { return inf ? fabs(v) : v * v; }
__device__ __forceinline__ double nrm_fin(double v, bool inf) { return inf ? v : sqrt(v); }

// Row i of the slice-of-64 ELL operator.  The slice width W is uniform in a
// wavefront, so entries are processed in batches of 8 with wave-uniform
// guards: all column and value loads of a batch are issued before any
// gather, which turns W dependent load chains into ceil(W/8).
constexpr int kBatch = 8;

template <typename VT, typename AT, int NR>
__device__ __forceinline__ void sell_row_batched(const int32_t *__restrict__ sell_ptr,
                                                 const int32_t *__restrict__ sell_col, const VT *__restrict__ val,
                                                 AT dg, int i, const VT *__restrict__ x, int C, AT (&acc)[NR]) {
    const int s = i >> 6, lane = i & 63;
    const int base = sell_ptr[s];
    const int W = (sell_ptr[s + 1] - base) >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = dg * (AT)x[(size_t)r * C + i];
    for (int kb = 0; kb < W; kb += kBatch) {
        int col[kBatch];
        VT v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            col[u] = i;
            v[u] = (VT)0;
            if (kb + u < W) {
                const int e = base + (kb + u) * 64 + lane;
                col[u] = __builtin_nontemporal_load(sell_col + e);
                v[u] = __builtin_nontemporal_load(val + e);
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (kb + u < W) {
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] += (AT)v[u] * (AT)x[(size_t)r * C + col[u]];
            }
        }
    }
}

template <typename VT, int NR>
__device__ __forceinline__ void sell_row(const int32_t *__restrict__ sell_ptr, const int32_t *__restrict__ sell_col,
                                         const VT *__restrict__ val, VT dg, int i, const VT *__restrict__ x, int C,
                                         VT (&acc)[NR]) {
    sell_row_batched<VT, VT, NR>(sell_ptr, sell_col, val, dg, i, x, C, acc);
}

// Contiguous row range of workgroup b (chunks of RB rows).
__device__ __forceinline__ void row_range(int C, int nb, int b, int &lo, int &hi) {
    const int per = ((C + nb - 1) / nb + 63) & ~63;  // keep slices inside one workgroup
    lo = b * per;
    hi = min(C, lo + per);
}

// Rows of SpMV workgroup b: one row per thread.
__device__ __forceinline__ void spmv_range(int C, int b, int &lo, int &hi) {
    lo = b * SB;
    hi = min(C, lo + SB);
}

// ---- state plumbing -----------------------------------------------------
template <int NR>
__device__ __forceinline__ void load_state(const Ops &o, int s, KState *sh) {
    if (threadIdx.x < NR) sh[threadIdx.x] = o.st[((s - 1) & 1) * kMaxRhs + threadIdx.x];
    lds_sync();
}
// The last launch of a chunk also posts one word to pinned host memory for
// the host's poll: a single aligned 32-bit system-scope atomic store (no
// copy of it stays in the GPU's caches), read by the host only after an
// event recorded behind the launch has completed.
template <int NR>
__device__ __forceinline__ void post_poll(const Ops &o, int s, bool all_done) {
    if (o.poll && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(o.poll, (s << 1) | (all_done ? 1 : 0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <int NR>
__device__ __forceinline__ void store_state(const Ops &o, int s, const KState *sh) {
    lds_sync();
    if (blockIdx.x == 0 && threadIdx.x < NR) o.st[(s & 1) * kMaxRhs + threadIdx.x] = sh[threadIdx.x];
    if (o.poll) {
        bool done = true;
        for (int r = 0; r < NR; ++r) done &= sh[r].phase == PH_DONE;
        post_poll<NR>(o, s, done);
    }
}

// State of the previous launch read with scalar loads: it is uniform across
// the grid, lands in SGPRs of every wave, and needs no LDS broadcast (the
// scalar cache is invalidated at every kernel start).
typedef const __attribute__((address_space(4))) KState ConstKState;
__device__ __forceinline__ KState state_in(const Ops &o, int s, int r) {
    ConstKState *c = (ConstKState *)(o.st + ((s - 1) & 1) * kMaxRhs + r);
    KState v;
    v.thresh = c->thresh;
    v.norm0 = c->norm0;
    v.final_res = c->final_res;
    v.rz_old = c->rz_old;
    v.alpha = c->alpha;
    v.beta = c->beta;
    v.rho_old = c->rho_old;
    v.omega = c->omega;
    v.rhat2 = c->rhat2;
    v.scale = c->scale;
    v.rel = c->rel;
    v.rec = c->rec;
    v.iter = c->iter;
    v.phase = c->phase;
    v.converged = c->converged;
    v.restarts = c->restarts;
    v.fresh = c->fresh;
    v.pcur = c->pcur;
    v.iter_ref = c->iter_ref;
    return v;
}
// Workgroup 0 writes this launch's state (every wave holds the same copy).
template <int NR>
__device__ __forceinline__ void state_out(const Ops &o, int s, const KState (&S)[NR]) {
    if (blockIdx.x != 0) return;
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (threadIdx.x == r) o.st[(s & 1) * kMaxRhs + r] = S[r];
    if (o.poll) {
        bool done = true;
#pragma unroll
        for (int r = 0; r < NR; ++r) done &= S[r].phase == PH_DONE;
        post_poll<NR>(o, s, done);
    }
}

// The first KP shares of the previous launch's partials (indices
// threadIdx.x + k BS), loaded ahead of everything that waits on the solver
// state; zero beyond nb_prev.
template <int NR, int NQ, int KP, int BS>
struct Prefetched {
    double x[KP][NR][NQ];
};
template <int NR, int NQ, int KP, int BS>
__device__ __forceinline__ void prefetch_prev(const Ops &o, int s, int nb_prev, Prefetched<NR, NQ, KP, BS> &pf) {
    const double *P = o.partials + (size_t)((s - 1) & 1) * kMaxRhs * kNQ * o.pstride;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int j = threadIdx.x + k * BS;
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                pf.x[k][r][q] = j < nb_prev ? P[((size_t)r * kNQ + q) * o.pstride + j] : 0.0;
    }
}

// Reduce the previous launch's partials: tot[r][q] in every thread.  With
// `pf`, the first KP shares were prefetched (same summation order either way).
template <int NR, int NQ, int BS = RB, int KP = 1>
__device__ __forceinline__ void reduce_prev(const Ops &o, int s, int nb_prev, const bool (&is_max)[NQ],
                                            double (&tot)[NR][NQ],
                                            const Prefetched<NR, NQ, KP, BS> *pf = nullptr) {
    __shared__ double lds[NR * NQ][BS / kWave];
    __shared__ double res[NR * NQ];
    const double *P = o.partials + (size_t)((s - 1) & 1) * kMaxRhs * kNQ * o.pstride;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    double acc[NR][NQ] = {};
    if (pf) {
#pragma unroll
        for (int k = 0; k < KP; ++k)
#pragma unroll
            for (int r = 0; r < NR; ++r)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const double x = pf->x[k][r][q];
                    acc[r][q] = is_max[q] ? fmax(acc[r][q], x) : acc[r][q] + x;
                }
    }
    for (int j = threadIdx.x + (pf ? KP * BS : 0); j < nb_prev; j += BS) {  // all quantities' loads issued together
        double x[NR][NQ];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int q = 0; q < NQ; ++q) x[r][q] = P[((size_t)r * kNQ + q) * o.pstride + j];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[r][q] = is_max[q] ? fmax(acc[r][q], x[r][q]) : acc[r][q] + x[r][q];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double v = is_max[q] ? wave_max(acc[r][q]) : wave_sum(acc[r][q]);
            if (lane == 0) lds[r * NQ + q][wid] = v;
        }
    lds_sync();
    if (threadIdx.x < NR * NQ) {
        const int q = threadIdx.x % NQ;
        double a = lds[threadIdx.x][0];
        for (int k = 1; k < BS / kWave; ++k) a = is_max[q] ? fmax(a, lds[threadIdx.x][k]) : a + lds[threadIdx.x][k];
        res[threadIdx.x] = a;
    }
    lds_sync();
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < NQ; ++q) tot[r][q] = res[r * NQ + q];
}

// Write this workgroup's partials for the next launch.
template <int NR, int NQ, int BS = RB>
__device__ __forceinline__ void write_partials(const Ops &o, int s, double (&v)[NR][NQ], const bool (&is_max)[NQ]) {
    __shared__ double lds[NR * NQ][BS / kWave];
    double *P = o.partials + (size_t)(s & 1) * kMaxRhs * kNQ * o.pstride;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x = is_max[q] ? wave_max(v[r][q]) : wave_sum(v[r][q]);
            if (lane == 0) lds[r * NQ + q][wid] = x;
        }
    lds_sync();
    if (threadIdx.x < NR * NQ) {
        const int r = threadIdx.x / NQ, q = threadIdx.x % NQ;
        double a = lds[threadIdx.x][0];
        for (int k = 1; k < BS / kWave; ++k) a = is_max[q] ? fmax(a, lds[threadIdx.x][k]) : a + lds[threadIdx.x][k];
        P[((size_t)r * kNQ + q) * o.pstride + blockIdx.x] = a;
    }
}

template <int NR>
__device__ __forceinline__ bool any_phase(const KState *sh, int ph) {
    bool a = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) a |= sh[r].phase == ph;
    return a;
}

// ---------------------------------------------------------------- packing
template <typename VT>
__global__ void k_sell_pack(int64_t n, const int32_t *__restrict__ code, const double *__restrict__ upper,
                            const double *__restrict__ lower, int transpose, VT *__restrict__ val) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    int c = code[e];
    double v = 0.0;
    if (c != INT32_MIN) {
        bool up = (c >= 0) != (transpose != 0);
        int f = c >= 0 ? c : -c - 1;
        v = up ? upper[f] : lower[f];
    }
    val[e] = (VT)v;
}

template <typename VT>
__global__ void k_diag_prep(int C, const double *__restrict__ diag, const int32_t *__restrict__ perm, int precond,
                            VT *__restrict__ dg, VT *__restrict__ inv_d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    double d = diag[perm ? perm[i] : i];  // solver row i is cell perm[i]
    dg[i] = (VT)d;
    // preconditioners.py:83-90: 1/diag, or 1 where |diag| < eps(dtype)
    const double eps = sizeof(VT) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
    double inv = 1.0;
    if (precond) inv = fabs(d) < eps ? 1.0 : 1.0 / d;
    if (inv_d) inv_d[i] = (VT)inv;
}

// Solve entry: b and x ([C, nrhs]) into the workspace's component-major
// copies, in solver row order (row i is cell perm[i]; identity without perm).
__global__ void k_solve_in(int C, int nrhs, const double *__restrict__ b, const double *__restrict__ x,
                           const int32_t *__restrict__ perm, double *__restrict__ wb, double *__restrict__ wx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const int c = perm ? perm[i] : i;
    for (int r = 0; r < nrhs; ++r) {
        wb[(size_t)r * C + i] = b[(size_t)c * nrhs + r];
        wx[(size_t)r * C + i] = x[(size_t)c * nrhs + r];
    }
}

// Solve exit: the solution back to [C, nrhs], and the final solver state of
// `nr` components into coherent pinned host memory (system-scope stores).
__global__ void k_solve_out(int C, int nrhs, const double *__restrict__ wx, const int32_t *__restrict__ perm,
                            double *__restrict__ x, const KState *__restrict__ st, KState *host_state) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C) {
        const int c = perm ? perm[i] : i;
        for (int r = 0; r < nrhs; ++r) x[(size_t)c * nrhs + r] = wx[(size_t)r * C + i];
    }
    constexpr int kWords = (int)(sizeof(KState) / sizeof(uint32_t));
    if (blockIdx.x == 0 && threadIdx.x < nrhs * kWords) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(st);
        uint32_t *dst = reinterpret_cast<uint32_t *>(host_state);
        __hip_atomic_store(dst + threadIdx.x, src[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Steady PCG state with beta = 0 for the SpMV benchmark entry.
__global__ void k_spmv_bench_state(KState *st, int n) {
    int i = threadIdx.x;
    if (i < n) {
        KState S{};
        S.phase = PH_ACTIVE;
        S.fresh = 2;
        S.rz_old = 1.0;
        S.thresh = -1.0;
        S.scale = 0.0;
        st[i] = S;
    }
}

__global__ void k_zero_state(KState *st, int n) {
    int i = threadIdx.x;
    if (i < n) st[i] = KState{};
}

// ---------------------------------------------------------------- PCG
// Two launches per iteration, equal in exact arithmetic to cg.py:120-170:
//   k_cg_dir : test |r| left by the previous update (cg.py:150-158);
//              beta = r.z / r.z_old ; p = z + beta p_old written to the
//              other p buffer, and Ap.  Every gathered p_j is rebuilt from
//              z_j and p_old_j with the same fma its owner row uses, so Ap
//              is A times the stored p ; partial p.Ap.
//   k_cg_upd : alpha = r.z_old / p.Ap ; x += alpha p ; r -= alpha Ap ;
//              z = M^-1 r ; partials |r|, r.z.
// The two p buffers alternate by the iteration's position in its chunk (an
// even number of iterations): iteration j reads p_old from buffer j & 1 and
// writes p to the other, so every gather is issued before the solver state
// is read.  The state still records the buffer of the current direction
// (pcur); when the mixed solve stops a chunk after an odd number of
// iterations, the refinement moves the direction back to buffer 0.

template <typename VT>
__device__ __forceinline__ VT dir_fma(VT beta, VT p_old, VT z);
template <>
__device__ __forceinline__ double dir_fma<double>(double beta, double p_old, double z) {
    return __builtin_fma(beta, p_old, z);
}
template <>
__device__ __forceinline__ float dir_fma<float>(float beta, float p_old, float z) {
    return __builtin_fmaf(beta, p_old, z);
}

// Entries per row loaded ahead of the partial reduction.  Six cover every
// row of a uniform 3-D region; nine also cover slices that hold a 2:1
// refinement interface (a fifth of the slices of a wall-refined octree) and
// keep the FP32 kernel at 68 VGPRs (seven waves per SIMD).  The FP64 kernel
// keeps six: with nine it needs 90 VGPRs (five waves per SIMD) and is no
// faster.
template <typename VT>
constexpr int head_entries() {
    return sizeof(VT) == 4 ? 9 : 6;
}

// ---- brick slices of the solver row order ----------------------------------
// A brick is 64 leaves of one level filling an aligned 4x4x4 block; lane
// l = x + 4 y + 16 z holds the cell with in-brick coordinates (x, y, z), so
// its neighbours inside the brick are lanes l +- 1, +- 4, +- 16 and their
// values come from cross-lane permutes.  Each lane keeps the value of its
// +axis face (symmetric operator: the -axis value is the -neighbour's +axis
// value).  A side toward a brick of the same level reads that brick's mirror
// lane (l ^ 3, l ^ 12, l ^ 48) and its -axis values are stored per side (16
// per axis); a side toward anything else is an explicit block of 64 slots
// (16 boundary cells x 4 faces, slot = 4 q + k).
__device__ __forceinline__ int bcoord(int l, int a) { return (l >> (2 * a)) & 3; }
// position q of a boundary cell on an axis-a side: c_b + 4 c_c, b = (a+1) % 3, c = (a+2) % 3
__device__ __forceinline__ int bface(int l, int a) {
    return a == 0 ? l >> 2 : (a == 2 ? l & 15 : (l >> 4) + 4 * (l & 3));
}

typedef const __attribute__((address_space(4))) int32_t ConstI32;

// Bricks per wave in the direction kernel: each wave carries kBricks
// bricks, all of whose loads are issued before the solver state is read, so
// the dependent neighbour-brick gathers of several bricks overlap.
constexpr int kBricks = 2;
// Threads per workgroup of the fused mixed-precision launch: 1024 (one
// workgroup per CU), or 512 / 256 with as many more workgroups
// (GRIDFOAM_FUSED_TB overrides the default for tuning).  One brick per unit.
constexpr int kFusedBricks = 1;
// Bricks (or 64-row slices) whose operands fit in one XCD's 4 MB L2.
constexpr int kL2Bricks = 2048;
int fused_tb() {
    static int tb = [] {
        const char *e = getenv("GRIDFOAM_FUSED_TB");
        const int v = e ? atoi(e) : 1024;
        return (v == 256 || v == 512) ? v : 1024;
    }();
    return tb;
}

// Waves that carry bricks: whole workgroups only (bricks left over go
// through the sliced ELL with the other rows).
__host__ __device__ __forceinline__ int brick_waves(int nbrick, int tb) {
    const int per = tb / 64;
    return nbrick / (kBricks * per * 8) * per * 8;
}
// The grid is [brick workgroups | sliced-ELL workgroups], each part a
// multiple of 8 so that both spread evenly over the 8 XCDs (each XCD gets a
// contiguous range of either part).
__host__ __device__ __forceinline__ int brick_workgroups(int nbrick, int tb) {
    return brick_waves(nbrick, tb) / (tb / 64);  // whole workgroups, a multiple of 8 by construction
}
int pcg_dir_workgroups(int C, int nbrick) {
    const int nsl = (C + 63) / 64, nbw = brick_waves(nbrick, pcg_tile()), per = pcg_tile() / 64;
    return brick_workgroups(nbrick, pcg_tile()) + round8((nsl - nbw * kBricks + per - 1) / per);
}
// Workgroups of the persistent fused launch, a multiple of 8: at most one per
// CU (GRIDFOAM_FUSED_WG overrides the cap for tuning), and about two chunks
// of 16 units each (a 64-row slice counted as two units) per workgroup, so a
// small grid leaves CUs to solves that run concurrently on other streams.
int fused_workgroup_cap() {
    static int g = [] {
        const char *e = getenv("GRIDFOAM_FUSED_WG");
        int v = e ? atoi(e) : 0;
        if (v <= 0) {
            int dev = 0, n = 0;
            if (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
                v = n;
            else
                v = 256;
            v *= 1024 / fused_tb();
        }
        return round8(v);
    }();
    return g;
}
int fused_workgroups(int C, int nbrick) {
    const int nsl = (C + 63) / 64, nw = fused_tb() / 64;
    const long work = (long)nbrick + 2L * (nsl - nbrick);
    const long want = (work + 2L * nw - 1) / (2L * nw);
    const int cap = fused_workgroup_cap();
    return want >= cap ? cap : round8((int)(want > 8 ? want : 8));
}
// Partials per (rhs, quantity) of a solve: the largest workgroup count of any of its kernels.
int solve_pstride(int C, int nbrick) {
    int m = partial_stride(C);
    for (int v : {pcg_dir_workgroups(C, nbrick), fused_workgroups(C, nbrick)}) m = v > m ? v : m;
    return m;
}

// Everything a brick row needs that does not depend on the solver state.
template <typename VT>
struct BrickRow {
    double d;
    VT z, po, vp[3], vm[3], zx[3], px[3];
    int nb[6];
    int d0, d1;  // first two explicit sides (uniform), -1 none
    VT e0v, e0z, e0p, e1v, e1z, e1p;
};

template <typename VT>
__device__ __forceinline__ void brick_load(const Ops &o, int sl, int lane, const double *__restrict__ dg,
                                           const VT *__restrict__ z, const VT *__restrict__ pold,
                                           const VT *__restrict__ bv, const VT *__restrict__ bm,
                                           const VT *__restrict__ ev, BrickRow<VT> &B) {
    const int i = sl * 64 + lane;
    ConstI32 *nbp = (ConstI32 *)(o.bnbr + 6 * sl);
#pragma unroll
    for (int k = 0; k < 6; ++k) B.nb[k] = nbp[k];
    B.d = dg[i];
    B.z = z[i];
    B.po = pold[i];
    // every lane loads (its own row where it has no neighbour brick): no divergence
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        B.vp[a] = bv[((size_t)sl * 3 + a) * 64 + lane];
        const int c = bcoord(lane, a);
        const int nbr = c == 3 ? B.nb[2 * a + 1] : B.nb[2 * a];
        const bool hx = (c == 0 || c == 3) && nbr >= 0;
        const int e = hx ? nbr * 64 + (lane ^ (3 << (2 * a))) : i;
        B.zx[a] = z[e];
        B.px[a] = pold[e];
        const VT m = bm[((size_t)sl * 3 + a) * 16 + bface(lane, a)];
        B.vm[a] = (c == 0 && hx) ? m : (VT)0;
    }
    B.d0 = -1;
    B.d1 = -1;
    int b0 = 0, b1 = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (B.nb[k] <= -2) {
            if (B.d0 < 0) {
                B.d0 = k;
                b0 = -2 - B.nb[k];
            } else if (B.d1 < 0) {
                B.d1 = k;
                b1 = -2 - B.nb[k];
            }
        }
    }
    B.e0v = B.e0z = B.e0p = B.e1v = B.e1z = B.e1p = (VT)0;
    if (B.d0 >= 0) {
        const int e = b0 * 64 + lane;
        const int c = o.ecol[e];
        B.e0v = ev[e];
        B.e0z = z[c];
        B.e0p = pold[c];
    }
    if (B.d1 >= 0) {
        const int e = b1 * 64 + lane;
        const int c = o.ecol[e];
        B.e1v = ev[e];
        B.e1z = z[c];
        B.e1p = pold[c];
    }
}

// Ap of a brick row; p is this row's new direction value.  Every lane of the
// wave takes part in the permutes (a lane masked off would read as zero).
template <typename VT>
__device__ __forceinline__ double brick_apply(const Ops &o, int lane, bool start, VT be, VT p,
                                              const BrickRow<VT> &B, const VT *__restrict__ z,
                                              const VT *__restrict__ pold, const VT *__restrict__ ev) {
    double acc = B.d * (double)p;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int st = 1 << (2 * a), c = bcoord(lane, a);
        const VT pe = start ? B.zx[a] : dir_fma<VT>(be, B.px[a], B.zx[a]);
        const VT sp = __shfl(p, lane + st);
        const VT sm = __shfl(p, lane - st);
        const VT sv = __shfl(B.vp[a], lane - st);
        const VT qp = c == 3 ? pe : sp;  // a boundary lane takes its outside neighbour
        const VT qm = c == 0 ? pe : sm;  // (zero value where there is none)
        const VT vmn = c == 0 ? B.vm[a] : sv;
        acc += (double)B.vp[a] * (double)qp;
        acc += (double)vmn * (double)qm;
    }
    // explicit side dd: slot l = 4 q + k holds face k of boundary cell q; the
    // four slots of a cell are summed across lanes, then read by the cell's lane
    auto side = [&](int dd, VT vv, VT zz, VT pp) {
        const VT pe = start ? zz : dir_fma<VT>(be, pp, zz);
        double cl = (double)vv * (double)pe;
        cl += __shfl_xor(cl, 1);
        cl += __shfl_xor(cl, 2);
        const int a = dd >> 1;
        const double t = __shfl(cl, 4 * bface(lane, a));
        if (bcoord(lane, a) == ((dd & 1) ? 3 : 0)) acc += t;
    };
    if (B.d0 >= 0) side(B.d0, B.e0v, B.e0z, B.e0p);
    if (B.d1 >= 0) side(B.d1, B.e1v, B.e1z, B.e1p);
#pragma unroll
    for (int dd = 0; dd < 6; ++dd) {  // sides beyond the first two (rare): loaded here
        if (B.nb[dd] <= -2 && dd != B.d0 && dd != B.d1) {
            const int e = (-2 - B.nb[dd]) * 64 + lane;
            const int c = o.ecol[e];
            side(dd, ev[e], z[c], pold[c]);
        }
    }
    return acc;
}

// The state transition of the direction kernel (cg.py:150-170), applied by
// every wave to the same inputs: convergence test on the previous update's
// residual, beta, the direction buffer flip.  The FP64 PCG only: the mixed
// precision PCG runs in k_cg_fused (MX stays false; may_refine is unused).
template <int NR, bool MX>
__device__ __forceinline__ void dir_transition(const Ops &o, int may_refine, const double (&tot)[NR][2],
                                               KState (&S)[NR], bool (&act)[NR], bool (&start)[NR]) {
    static_assert(!MX, "the mixed-precision PCG runs in k_cg_fused");
    (void)may_refine;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        KState &St = S[r];
        start[r] = St.fresh == 1;
        if (St.phase == PH_ACTIVE) {
            bool go = true;
            if (St.fresh == 0) {
                const double nr = nrm_fin(tot[r][0], o.norm_inf);
                St.iter += 1;
                St.final_res = nr;
                if (nr < St.thresh) {
                    St.converged = 1;
                    St.phase = PH_DONE;
                    go = false;
                } else if (St.iter >= o.max_iter) {
                    St.phase = PH_DONE;
                    go = false;
                }
            }
            if (go) {
                if (St.fresh != 1) {
                    St.beta = tot[r][1] / St.rz_old;
                    St.rz_old = tot[r][1];
                } else {
                    St.beta = 0.0;
                }
                St.fresh = 0;
                St.pcur ^= 1;
            }
        }
        act[r] = St.phase == PH_ACTIVE;
    }
}

// Direction work of a wave that carries kBricks bricks (one rhs, symmetric
// operator).  Everything that does not depend on the solver state is loaded
// before the state is read and stays in flight during the reduction.
template <typename VT, bool MX, int TB>
__device__ __forceinline__ void dir_bricks(const Ops &o, int s, int nb_prev, int may_refine, int pb, int sl,
                                           const Prefetched<1, 2, 1, TB> &pre, const double *__restrict__ dg,
                                           const VT *__restrict__ z, VT *__restrict__ p0, VT *__restrict__ p1,
                                           VT *__restrict__ ap, const VT *__restrict__ bv,
                                           const VT *__restrict__ bm, const VT *__restrict__ ev) {
    const int lane = threadIdx.x & 63;
    const VT *pold = pb ? p1 : p0;
    VT *pnew = pb ? p0 : p1;
    BrickRow<VT> Bk[kBricks];
#pragma unroll
    for (int k = 0; k < kBricks; ++k) brick_load<VT>(o, sl + k, lane, dg, z, pold, bv, bm, ev, Bk[k]);
    asm volatile("" ::: "memory");  // keep the loads above issued before the state test
    KState S[1];
    S[0] = state_in(o, s, 0);
    if (S[0].phase != PH_ACTIVE) {
        state_out<1>(o, s, S);
        return;
    }
    const bool mx[2] = {(bool)o.norm_inf, false};
    double tot[1][2];
    reduce_prev<1, 2, TB, 1>(o, s, nb_prev, mx, tot, &pre);
    bool act[1], start[1];
    dir_transition<1, MX>(o, may_refine, tot, S, act, start);
    const VT be = (VT)S[0].beta;
    double pv[1][1] = {{0.0}};
#pragma unroll
    for (int k = 0; k < kBricks; ++k) {
        const VT p = start[0] ? Bk[k].z : dir_fma<VT>(be, Bk[k].po, Bk[k].z);
        const double acc = brick_apply<VT>(o, lane, start[0], be, p, Bk[k], z, pold, ev);
        if (act[0]) {
            const int ik = (sl + k) * 64 + lane;
            const VT a = (VT)acc;
            pnew[ik] = p;
            ap[ik] = a;
            pv[0][0] += (double)p * (double)a;
        }
    }
    const bool mx1[1] = {false};
    write_partials<1, 1, TB>(o, s, pv, mx1);
    state_out<1>(o, s, S);
}

// Direction work of a wave that carries one slice of the sliced ELL (slice
// sl0; past the last slice the wave only takes part in the reductions).
template <typename VT, int NR, bool MX, int TB>
__device__ __forceinline__ void dir_sell(const Ops &o, int s, int nb_prev, int may_refine, int pb, int sl0,
                                         const Prefetched<NR, 2, 1, TB> &pre, const VT *__restrict__ val,
                                         const double *__restrict__ dg, const VT *__restrict__ z,
                                         VT *__restrict__ p0, VT *__restrict__ p1, VT *__restrict__ ap) {
    constexpr int kHead = head_entries<VT>();
    const int lane = threadIdx.x & 63;
    const int nsl = (o.C + 63) >> 6;
    const int sl = sl0 < nsl ? sl0 : nsl - 1;
    const int i = sl0 * 64 + lane;
    const int ic = i < o.C ? i : o.C - 1;
    const VT *pold[NR];
    VT *pnew[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {  // direction of the previous iteration in buffer pb; the new one goes to the other
        pold[r] = (pb ? p1 : p0) + (size_t)r * o.C;
        pnew[r] = (pb ? p0 : p1) + (size_t)r * o.C;
    }
    const double d = dg[ic];
    VT zi[NR], poi[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        zi[r] = z[(size_t)r * o.C + ic];
        poi[r] = pold[r][ic];
    }
    const int base = o.sell_ptr[sl];
    const int W = sl0 < nsl ? (o.sell_ptr[sl + 1] - base) >> 6 : 0;
    int col[kHead];
    VT v[kHead], zj[kHead][NR], pj[kHead][NR];
#pragma unroll
    for (int u = 0; u < kHead; ++u) {
        col[u] = ic;
        v[u] = (VT)0;
        if (u < W) {
            const int e = base + u * 64 + lane;
            col[u] = __builtin_nontemporal_load(o.sell_col + e);
            v[u] = __builtin_nontemporal_load(val + e);
        }
    }
#pragma unroll
    for (int u = 0; u < kHead; ++u)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            zj[u][r] = z[(size_t)r * o.C + col[u]];
            pj[u][r] = pold[r][col[u]];
        }
    asm volatile("" ::: "memory");  // keep the loads above issued before the state test
    KState S[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        S[r] = state_in(o, s, r);
        any |= S[r].phase == PH_ACTIVE;
    }
    if (!any) {
        state_out<NR>(o, s, S);
        return;
    }
    const bool mx[2] = {(bool)o.norm_inf, false};
    double tot[NR][2];
    reduce_prev<NR, 2, TB, 1>(o, s, nb_prev, mx, tot, &pre);
    bool act[NR], start[NR];
    dir_transition<NR, MX>(o, may_refine, tot, S, act, start);
    VT be[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) be[r] = (VT)S[r].beta;
    double acc[NR];
    VT pi[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        pi[r] = start[r] ? zi[r] : dir_fma<VT>(be[r], poi[r], zi[r]);
        acc[r] = d * (double)pi[r];
    }
#pragma unroll
    for (int u = 0; u < kHead; ++u)
        if (u < W)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const VT q = start[r] ? zj[u][r] : dir_fma<VT>(be[r], pj[u][r], zj[u][r]);
                acc[r] += (double)v[u] * (double)q;
            }
    for (int kb = kHead; kb < W; kb += kBatch) {  // rows wider than the head batch
        int c2[kBatch];
        VT v2[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            c2[u] = ic;
            v2[u] = (VT)0;
            if (kb + u < W) {
                const int e = base + (kb + u) * 64 + lane;
                c2[u] = __builtin_nontemporal_load(o.sell_col + e);
                v2[u] = __builtin_nontemporal_load(val + e);
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (kb + u < W)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const VT zz = z[(size_t)r * o.C + c2[u]];
                    const VT q = start[r] ? zz : dir_fma<VT>(be[r], pold[r][c2[u]], zz);
                    acc[r] += (double)v2[u] * (double)q;
                }
    }
    double pv[NR][1] = {};
    if (i < o.C) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const VT a = (VT)acc[r];
            pnew[r][i] = pi[r];
            ap[(size_t)r * o.C + i] = a;
            pv[r][0] = (double)pi[r] * (double)a;
        }
    }
    const bool mx1[1] = {false};
    write_partials<NR, 1, TB>(o, s, pv, mx1);
    state_out<NR>(o, s, S);
}

// Convergence test on the previous update's residual, beta and p = z + beta
// p_old ; Ap ; partial p.Ap (cg.py:120-170 in two launches, see above).
// Work per wave: the first workgroups (a multiple of 8) carry kBricks
// bricks per wave (values bv: +axis faces, bm: -axis sides toward bricks,
// ev: explicit blocks); every wave of the following workgroups one slice of
// the sliced ELL, from slice nbw * kBricks on.  A workgroup is all-brick or
// all-ELL, so both paths meet the same barriers.
template <typename VT, int NR, bool MX, int TB>
__global__ __launch_bounds__(TB) void k_cg_dir(Ops o, int s, int nb_prev, int may_refine, int pb,
                                               const VT *__restrict__ val, const double *__restrict__ dg,
                                               const VT *__restrict__ z, VT *__restrict__ p0, VT *__restrict__ p1,
                                               VT *__restrict__ ap, const VT *__restrict__ bv,
                                               const VT *__restrict__ bm, const VT *__restrict__ ev) {
    Prefetched<NR, 2, 1, TB> pre;
    prefetch_prev(o, s, nb_prev, pre);
    const int nbw = brick_waves(o.nbrick, TB), gb = brick_workgroups(o.nbrick, TB);
    const int wid = threadIdx.x >> 6;
    if constexpr (NR == 1) {
        if ((int)blockIdx.x < gb) {
            const int wv = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gb) * (TB / 64) + wid);
            dir_bricks<VT, MX, TB>(o, s, nb_prev, may_refine, pb, wv * kBricks, pre, dg, z, p0, p1, ap, bv, bm, ev);
            return;
        }
    }
    const int ws = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x - gb, gridDim.x - gb) * (TB / 64) + wid);
    dir_sell<VT, NR, MX, TB>(o, s, nb_prev, may_refine, pb, nbw * kBricks + ws, pre, val, dg, z, p0, p1, ap);
}

// alpha = rz_old / p.Ap ; x += alpha p ; r -= alpha Ap ; z = M^-1 r ; partials |r|, r.z.
// The grid has at most one 1024-thread workgroup per CU, so the kernel may
// use up to 128 VGPRs: each thread's first rows (kHoist of them; all rows of
// a 1.1M-cell mesh) and the first five shares of the direction kernel's
// partials are loaded before the state is read.
template <typename VT, int NR>
__global__ __launch_bounds__(RB) void k_cg_upd(Ops o, int s, int pb, const VT *__restrict__ inv_d,
                                               VT *__restrict__ x_, VT *__restrict__ r_, VT *__restrict__ z_,
                                               const VT *__restrict__ p0, const VT *__restrict__ p1,
                                               const VT *__restrict__ ap_) {
    constexpr int kHoist = (sizeof(VT) == 4 ? 6 : 5) / NR > 0 ? (sizeof(VT) == 4 ? 6 : 5) / NR : 1;
    Prefetched<NR, 1, 5, RB> pre;
    prefetch_prev(o, s, o.gd, pre);
    int lo, hi;
    chunk_range(o.C, gridDim.x, xcd_tile(blockIdx.x, gridDim.x), lo, hi);
    const VT *p_[NR];  // this iteration's direction: buffer pb
#pragma unroll
    for (int r = 0; r < NR; ++r) p_[r] = (pb ? p1 : p0) + (size_t)r * o.C;
    VT hx[kHoist][NR], hr[kHoist][NR], hp[kHoist][NR], ha[kHoist][NR], hid[kHoist];
#pragma unroll
    for (int h = 0; h < kHoist; ++h) {
        const int i = lo + threadIdx.x + h * RB;
        const int ic = i < hi ? i : o.C - 1;
        hid[h] = inv_d[ic];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * o.C + ic;
            hx[h][r] = x_[k];
            hr[h][r] = r_[k];
            hp[h][r] = p_[r][ic];
            ha[h][r] = ap_[k];
        }
    }
    asm volatile("" ::: "memory");  // keep the loads above issued before the state test
    KState S[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        S[r] = state_in(o, s, r);
        any |= S[r].phase == PH_ACTIVE;
    }
    if (!any) {
        state_out<NR>(o, s, S);
        return;
    }
    const bool mx1[1] = {false};
    double pap[NR][1];
    reduce_prev<NR, 1, RB, 5>(o, s, o.gd, mx1, pap, &pre);
    double v[NR][2] = {};
    bool act[NR];
    VT a[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (S[r].phase == PH_ACTIVE) S[r].alpha = S[r].rz_old / pap[r][0];
        act[r] = S[r].phase == PH_ACTIVE;
        a[r] = (VT)S[r].alpha;
    }
    auto row = [&](int i, VT id, const VT (&xv)[NR], const VT (&rv)[NR], const VT (&pv)[NR], const VT (&av)[NR]) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * o.C + i;
            const VT xx = xv[r] + a[r] * pv[r];
            const VT rr = rv[r] - a[r] * av[r];
            const VT zz = id * rr;
            x_[k] = xx;
            r_[k] = rr;
            z_[k] = zz;
            v[r][0] = o.norm_inf ? fmax(v[r][0], fabs((double)rr)) : v[r][0] + (double)rr * (double)rr;
            v[r][1] += (double)rr * (double)zz;
        }
    };
#pragma unroll
    for (int h = 0; h < kHoist; ++h) {
        const int i = lo + threadIdx.x + h * RB;
        if (i < hi) row(i, hid[h], hx[h], hr[h], hp[h], ha[h]);
    }
    for (int i = lo + threadIdx.x + kHoist * RB; i < hi; i += RB) {  // chunks beyond kHoist rows per thread
        VT xv[NR], rv[NR], pv[NR], av[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * o.C + i;
            xv[r] = x_[k];
            rv[r] = r_[k];
            pv[r] = p_[r][i];
            av[r] = ap_[k];
        }
        row(i, inv_d[i], xv, rv, pv, av);
    }
    const bool mx[2] = {(bool)o.norm_inf, false};
    write_partials<NR, 2, RB>(o, s, v, mx);
    state_out<NR>(o, s, S);
}

// ---------------------------------------------------------------- mixed PCG: one launch per iteration
// The FP32 inner solve of the mixed-precision PCG runs one launch per
// iteration.  Launch k finishes the update of iteration k-1 and forms the
// next direction in the same pass:
//   alpha = g / d ; r' = r - alpha Ap ; x' = x + alpha p_old ; z = M^-1 r' ;
//   beta = (g - 2 alpha m + alpha^2 n) / g ; p = z + beta p_old ; Ap = A p
// with g = r.z, d = p.Ap, m = Ap.z, n = Ap.M^-1 Ap of the previous launch
// (its partials), so the new r'.z (the numerator of beta) is expanded from
// the previous launch's dot products instead of reduced after the update.
// Every gathered neighbour p_j is rebuilt from r_j, Ap_j, M^-1_j and p_old_j
// by the same function its owner row uses (bit-identical).  r, Ap and p are
// double-buffered by launch parity: a launch reads buffer pb and writes
// pb ^ 1, so no workgroup reads what another one writes in the same launch.
// The operator is the FP32 off-diagonals and the FP32-rounded FP64 row sum
// e = diag + sum_j a_ij: (A p)_i = e_i p_i + sum_j a_ij (p_j - p_i), summed
// in FP64, so the near-null constant mode of a Neumann pressure operator
// keeps its FP64 row sums.  Launches without an update (fresh != 0): the
// first direction p = z (fresh 1), and the direction rebuilt after an FP64
// refinement replaced r, p = z + (r.z / rz_old) p_old (fresh 2).
enum FQ : int { FQ_NR = 0, FQ_G = 1, FQ_D = 2, FQ_M = 3, FQ_N = 4, kFQ = 5 };

struct FusedT {
    bool act, upd, start;
    float al, be;
};

// State transition of a fused launch, applied by every wave to the same
// inputs.  A residual (of the previous update) below the threshold or a
// spent budget hands the solve to the FP64 refinement (which decides on the
// true residual); nothing is updated then.
__device__ __forceinline__ void fused_transition(const Ops &o, int pb, const double (&t)[1][kFQ], KState &S,
                                                 FusedT &T) {
    T.act = T.upd = T.start = false;
    T.al = T.be = 0.0f;
    if (S.phase != PH_ACTIVE) return;
    if (S.fresh == 0) {
        const double nr = nrm_fin(t[0][FQ_NR], o.norm_inf);
        if (nr < S.thresh || S.iter >= o.max_iter) {
            S.phase = PH_RESTART;
            S.rec = nr;
            return;
        }
        const double g = t[0][FQ_G], al = g / t[0][FQ_D];
        const double rz = g - 2.0 * al * t[0][FQ_M] + al * al * t[0][FQ_N];
        S.alpha = al;
        S.beta = rz / g;
        S.rz_old = g;
        S.iter += 1;
        T.upd = true;
    } else {
        T.start = S.fresh == 1;
        S.alpha = 0.0;
        S.beta = T.start ? 0.0 : t[0][FQ_G] / S.rz_old;
        // A true residual much larger than the recursive one it replaced
        // (beta > 4: twice its M^-1 norm) means the FP32 iteration had run
        // into rounding noise, and so has its direction: restart from p = z.
        if (S.beta > 4.0) {
            S.beta = 0.0;
            T.start = true;
        }
    }
    S.fresh = 0;
    S.pcur = pb ^ 1;
    T.act = true;
    T.al = (float)S.alpha;
    T.be = (float)S.beta;
}

// r' , z and p of one row from its previous r, Ap, p and M^-1 (owner rows
// and gathered neighbours alike).
__device__ __forceinline__ float fused_dir(const FusedT &T, float r, float ap, float id, float po, float &rn,
                                           float &z) {
    rn = T.upd ? __builtin_fmaf(-T.al, ap, r) : r;
    z = __fmul_rn(id, rn);
    return T.start ? z : __builtin_fmaf(T.be, po, z);
}

struct FusedVec {
    const float *e, *id;
    float *x, *r0, *r1, *a0, *a1, *p0, *p1;
};

// Everything a brick row of the fused launch needs that does not depend on
// the solver state: its own row, its +axis values, the -axis values toward
// neighbour bricks and the mirror lanes across brick sides (its own row
// where there is none).  Explicit sides are loaded when they are applied.
struct FBrick {
    float e, id, r, ap, x, po;
    float vp[3], vm[3];
    float xr[3], xa[3], xi[3], xp[3];
    int nb[6];
};

__device__ __forceinline__ void fbrick_load(const Ops &o, int sl, int lane, const FusedVec &V, const float *rin,
                                            const float *ain, const float *pin, const float *__restrict__ bv,
                                            const float *__restrict__ bm, const float *__restrict__ ev, FBrick &B) {
    const int i = sl * 64 + lane;
    ConstI32 *nbp = (ConstI32 *)(o.bnbr + 6 * sl);
#pragma unroll
    for (int k = 0; k < 6; ++k) B.nb[k] = nbp[k];
    B.e = V.e[i];
    B.id = V.id[i];
    B.r = rin[i];
    B.ap = ain[i];
    B.x = V.x[i];
    B.po = pin[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        B.vp[a] = bv[((size_t)sl * 3 + a) * 64 + lane];
        const int c = bcoord(lane, a);
        const int nbr = c == 3 ? B.nb[2 * a + 1] : B.nb[2 * a];
        const bool hx = (c == 0 || c == 3) && nbr >= 0;
        const int e = hx ? nbr * 64 + (lane ^ (3 << (2 * a))) : i;
        B.xr[a] = rin[e];
        B.xa[a] = ain[e];
        B.xi[a] = V.id[e];
        B.xp[a] = pin[e];
        const float m = bm[((size_t)sl * 3 + a) * 16 + bface(lane, a)];
        B.vm[a] = (c == 0 && hx) ? m : 0.0f;
    }
}

// (A p)_i of a brick row whose new direction value is p, in the difference
// form e p + sum_j a_j (p_j - p) summed in FP32 (a constant p gives exactly
// e p, so the near-null mode keeps its FP64 row sum).  Every lane of the
// wave takes part in the permutes (a lane masked off would read as zero).
__device__ __forceinline__ float fbrick_apply(const Ops &o, int lane, const FusedT &T, float p, const FBrick &B,
                                              const FusedVec &V, const float *rin, const float *ain,
                                              const float *pin, const float *__restrict__ ev) {
    float acc = __fmul_rn(B.e, p);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int st = 1 << (2 * a), c = bcoord(lane, a);
        float rn, z;
        const float pe = fused_dir(T, B.xr[a], B.xa[a], B.xi[a], B.xp[a], rn, z);
        const float sp = __shfl(p, lane + st);
        const float sm = __shfl(p, lane - st);
        const float sv = __shfl(B.vp[a], lane - st);
        const float qp = c == 3 ? pe : sp;  // a boundary lane takes its outside neighbour
        const float qm = c == 0 ? pe : sm;  // (zero value where there is none)
        const float vmn = c == 0 ? B.vm[a] : sv;
        acc = __builtin_fmaf(B.vp[a], qp - p, acc);
        acc = __builtin_fmaf(vmn, qm - p, acc);
    }
    // explicit side dd: slot l = 4 q + k holds face k of boundary cell q; the
    // four slots of a cell are summed across lanes (sum a (p_j - p_q)), then
    // read by the cell's lane
#pragma unroll
    for (int dd = 0; dd < 6; ++dd) {  // explicit sides (2:1 level changes): loaded here
        if (B.nb[dd] <= -2) {
            const int a = dd >> 1;
            const int e = (-2 - B.nb[dd]) * 64 + lane;
            const int c = o.ecol[e];
            float rn, z;
            const float pe = fused_dir(T, rin[c], ain[c], V.id[c], pin[c], rn, z);
            // the boundary cell of slot l is lane cl: p of that cell from its lane
            const int cellq = (lane >> 2);  // q of this slot
            const int cl = a == 0 ? ((dd & 1) ? 3 : 0) + 4 * (cellq & 3) + 16 * (cellq >> 2)
                         : a == 1 ? (cellq >> 2) + 4 * ((dd & 1) ? 3 : 0) + 16 * (cellq & 3)
                                  : (cellq & 3) + 4 * (cellq >> 2) + 16 * ((dd & 1) ? 3 : 0);
            const float pq = __shfl(p, cl);
            float t = __fmul_rn(ev[e], pe - pq);
            t += __shfl_xor(t, 1);
            t += __shfl_xor(t, 2);
            const float tl = __shfl(t, 4 * bface(lane, a));
            if (bcoord(lane, a) == ((dd & 1) ? 3 : 0)) acc += tl;
        }
    }
    return acc;
}

// Own-row outputs of a fused launch and its partials.
__device__ __forceinline__ void fused_out(const Ops &o, const FusedT &T, int i, float id, float x, float po,
                                          float rn, float z, float p, float a, const FusedVec &V, float *rout,
                                          float *aout, float *pout, double (&v)[1][kFQ]) {
    if (T.upd) V.x[i] = __builtin_fmaf(T.al, po, x);
    rout[i] = rn;
    pout[i] = p;
    aout[i] = a;
    const double rd = rn, ad = a;
    v[0][FQ_NR] = o.norm_inf ? fmax(v[0][FQ_NR], fabs(rd)) : v[0][FQ_NR] + rd * rd;
    v[0][FQ_G] += rd * (double)z;
    v[0][FQ_D] += (double)p * ad;
    v[0][FQ_M] += ad * (double)z;
    v[0][FQ_N] += ad * (double)__fmul_rn(id, a);
}

// ---- fast fixed-order reductions of the fused launch -----------------------
// DPP moves a value across lanes of a row of 16 without the LDS crossbar;
// 64-bit values move as two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_d(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ double red_op(double a, double b, bool mx) { return mx ? fmax(a, b) : a + b; }
// Sum (or max) over the 64 lanes of a wave: pairs, quads, halves and rows by
// symmetric permutes (every lane of a row then holds the same row total),
// then the four rows in a fixed order.  Every lane gets the same result.
__device__ __forceinline__ double wave_total(double v, bool mx) {
    v = red_op(v, dpp_d<0xB1>(v), mx);   // quad_perm [1,0,3,2]
    v = red_op(v, dpp_d<0x4E>(v), mx);   // quad_perm [2,3,0,1]
    v = red_op(v, dpp_d<0x141>(v), mx);  // row_half_mirror
    v = red_op(v, dpp_d<0x140>(v), mx);  // row_mirror
    return red_op(red_op(lane_d(v, 0), lane_d(v, 16), mx), red_op(lane_d(v, 32), lane_d(v, 48), mx), mx);
}

// The previous launch's kFQ partial sets reduced in every thread: the first
// share was prefetched (one set per thread), later shares are added in
// thread order, then wave totals, then the waves' totals in wave order.
template <int TB>
__device__ __forceinline__ void fused_reduce_prev(const Ops &o, int s, int nb_prev,
                                                  const Prefetched<1, kFQ, 1, TB> &pf, double (&tot)[1][kFQ]) {
    constexpr int NW = TB / 64;
    __shared__ double lds[kFQ][NW];
    const double *P = o.partials + (size_t)((s - 1) & 1) * kMaxRhs * kNQ * o.pstride;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double a[kFQ];
#pragma unroll
    for (int q = 0; q < kFQ; ++q) a[q] = pf.x[0][0][q];
    for (int j = threadIdx.x + TB; j < nb_prev; j += TB)
#pragma unroll
        for (int q = 0; q < kFQ; ++q) a[q] = red_op(a[q], P[(size_t)q * o.pstride + j], q == FQ_NR && o.norm_inf);
#pragma unroll
    for (int q = 0; q < kFQ; ++q) {
        const double t = wave_total(a[q], q == FQ_NR && o.norm_inf);
        if (lane == 0) lds[q][wid] = t;
    }
    lds_sync();
#pragma unroll
    for (int q = 0; q < kFQ; ++q) {
        double w[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = lds[q][k];
        double t = w[0];
#pragma unroll
        for (int k = 1; k < NW; ++k) t = red_op(t, w[k], q == FQ_NR && o.norm_inf);
        tot[0][q] = t;
    }
}

// This workgroup's kFQ partials: wave totals, then the waves' totals in wave order.
template <int TB>
__device__ __forceinline__ void fused_write_partials(const Ops &o, int s, const double (&v)[1][kFQ]) {
    constexpr int NW = TB / 64;
    __shared__ double lds[kFQ][NW];
    double *P = o.partials + (size_t)(s & 1) * kMaxRhs * kNQ * o.pstride;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kFQ; ++q) {
        const double t = wave_total(v[0][q], q == FQ_NR && o.norm_inf);
        if (lane == 0) lds[q][wid] = t;
    }
    lds_sync();
    if (threadIdx.x < kFQ) {
        const int q = threadIdx.x;
        double w[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = lds[q][k];
        double t = w[0];
#pragma unroll
        for (int k = 1; k < NW; ++k) t = red_op(t, w[k], q == FQ_NR && o.norm_inf);
        P[(size_t)q * o.pstride + blockIdx.x] = t;
    }
}

// One unit of KB bricks of a fused launch whose loads were issued into Bk.
template <int KB>
__device__ __forceinline__ void fused_unit(const Ops &o, int lane, const FusedT &T, int sl, const FBrick (&Bk)[KB],
                                           const FusedVec &V, const float *rin, const float *ain, const float *pin,
                                           float *rout, float *aout, float *pout, const float *__restrict__ ev,
                                           double (&v)[1][kFQ]) {
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        float rn, z;
        const float p = fused_dir(T, Bk[k].r, Bk[k].ap, Bk[k].id, Bk[k].po, rn, z);
        const float acc = fbrick_apply(o, lane, T, p, Bk[k], V, rin, ain, pin, ev);
        fused_out(o, T, (sl + k) * 64 + lane, Bk[k].id, Bk[k].x, Bk[k].po, rn, z, p, acc, V, rout, aout, pout, v);
    }
}

// One slice of 64 rows of the sliced ELL (rows past C take no part).
__device__ __forceinline__ void fused_slice(const Ops &o, int lane, const FusedT &T, int sl,
                                            const float *__restrict__ val, const FusedVec &V, const float *rin,
                                            const float *ain, const float *pin, float *rout, float *aout,
                                            float *pout, double (&v)[1][kFQ]) {
    constexpr int kHead = 6;
    const int i = sl * 64 + lane;
    const int ic = i < o.C ? i : o.C - 1;
    const float e = V.e[ic], id = V.id[ic], r = rin[ic], ap = ain[ic], x = V.x[ic], po = pin[ic];
    const int base = o.sell_ptr[sl];
    const int W = (o.sell_ptr[sl + 1] - base) >> 6;
    int col[kHead];
    float vv[kHead], jr[kHead], ja[kHead], ji[kHead], jp[kHead];
#pragma unroll
    for (int u = 0; u < kHead; ++u) {
        col[u] = ic;
        vv[u] = 0.0f;
        if (u < W) {
            const int k = base + u * 64 + lane;
            col[u] = __builtin_nontemporal_load(o.sell_col + k);
            vv[u] = __builtin_nontemporal_load(val + k);
        }
    }
#pragma unroll
    for (int u = 0; u < kHead; ++u) {
        jr[u] = rin[col[u]];
        ja[u] = ain[col[u]];
        ji[u] = V.id[col[u]];
        jp[u] = pin[col[u]];
    }
    float rn, z;
    const float p = fused_dir(T, r, ap, id, po, rn, z);
    float acc = __fmul_rn(e, p);  // difference form summed in FP32, as for brick rows
#pragma unroll
    for (int u = 0; u < kHead; ++u)
        if (u < W) {
            float rj, zj;
            const float q = fused_dir(T, jr[u], ja[u], ji[u], jp[u], rj, zj);
            acc = __builtin_fmaf(vv[u], q - p, acc);
        }
    for (int kb = kHead; kb < W; kb += kBatch) {  // rows wider than the head batch
        int c2[kBatch];
        float v2[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            c2[u] = ic;
            v2[u] = 0.0f;
            if (kb + u < W) {
                const int k = base + (kb + u) * 64 + lane;
                c2[u] = __builtin_nontemporal_load(o.sell_col + k);
                v2[u] = __builtin_nontemporal_load(val + k);
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (kb + u < W) {
                float rj, zj;
                const float q = fused_dir(T, rin[c2[u]], ain[c2[u]], V.id[c2[u]], pin[c2[u]], rj, zj);
                acc = __builtin_fmaf(v2[u], q - p, acc);
            }
    }
    if (i < o.C) fused_out(o, T, i, id, x, po, rn, z, p, acc, V, rout, aout, pout, v);
}

// A persistent launch of G workgroups (one per CU by default, a multiple of
// 8): the waves on one XCD share a contiguous eighth of the brick units (KB
// bricks each; bricks beyond the last whole unit go through the sliced ELL)
// and of the remaining sliced-ELL slices, interleaved wave by wave.  The loads of a wave's first unit are issued
// before the solver state is read and stay in flight during the reduction
// of the previous launch's G partial sets; partials are written once per
// workgroup, so the next launch reduces G sets.
template <int TB, int KB>
__global__ __launch_bounds__(TB) void k_cg_fused(Ops o, int s, int nb_prev, int pb, const float *__restrict__ val,
                                                 FusedVec V, const float *__restrict__ bv,
                                                 const float *__restrict__ bm, const float *__restrict__ ev) {
    constexpr int NW = TB / 64;
    Prefetched<1, kFQ, 1, TB> pre;
    prefetch_prev(o, s, nb_prev, pre);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // XCD x = blockIdx.x % 8 (the hardware's round-robin) owns a contiguous
    // eighth of the units and of the slices.  The units go in chunks of NW
    // consecutive ones to a workgroup, one per wave, so the waves of a CU
    // gather from each other's bricks through its L1.  When the eighth's
    // bricks fit in the XCD's 4 MB L2 (about 1.5 KB of operands per brick),
    // each workgroup takes a contiguous run of chunks; otherwise chunks go
    // round the XCD's workgroups, so all its waves sweep the range together
    // and the neighbour bricks they gather were loaded into the L2 moments
    // before.  Slices go round the XCD's workgroups first.
    const int G = gridDim.x, xcd = blockIdx.x & 7, PX = G >> 3, m = blockIdx.x >> 3;
    const int U = o.nbrick / KB, pu = (U + 7) >> 3;
    const int u0 = min(U, xcd * pu), u1 = min(U, u0 + pu);
    const int nch = (pu + NW - 1) / NW, cpw = (nch + PX - 1) / PX;
    const bool fits = pu * KB <= kL2Bricks;
    const int nsl = (o.C + 63) >> 6, e0 = U * KB, pe = (nsl - e0 + 7) >> 3;
    const int s0 = min(nsl, e0 + xcd * pe), s1 = min(nsl, s0 + pe);
    // j-th unit of this wave (u1 or beyond: none)
    auto unit_of = [&](int j) {
        if (fits && j >= cpw) return u1;
        const int c = fits ? m * cpw + j : m + PX * j;
        return u0 + c * NW + w;
    };
    const float *rin = pb ? V.r1 : V.r0, *ain = pb ? V.a1 : V.a0, *pin = pb ? V.p1 : V.p0;
    float *rout = pb ? V.r0 : V.r1, *aout = pb ? V.a0 : V.a1, *pout = pb ? V.p0 : V.p1;
    FBrick Bk[KB];
    int ju = 0, u = unit_of(0);
    if (u < u1) {
#pragma unroll
        for (int k = 0; k < KB; ++k) fbrick_load(o, u * KB + k, lane, V, rin, ain, pin, bv, bm, ev, Bk[k]);
    }
    asm volatile("" ::: "memory");  // keep the loads above issued before the state test
    KState S[1];
    S[0] = state_in(o, s, 0);
    if (S[0].phase != PH_ACTIVE) {
        state_out<1>(o, s, S);
        return;
    }
    double tot[1][kFQ];
    fused_reduce_prev<TB>(o, s, nb_prev, pre, tot);
    FusedT T;
    fused_transition(o, pb, tot, S[0], T);
    double v[1][kFQ] = {};
    if (T.act) {
        if (u < u1) {
            fused_unit<KB>(o, lane, T, u * KB, Bk, V, rin, ain, pin, rout, aout, pout, ev, v);
            for (u = unit_of(++ju); u < u1; u = unit_of(++ju)) {
#pragma unroll
                for (int k = 0; k < KB; ++k) fbrick_load(o, u * KB + k, lane, V, rin, ain, pin, bv, bm, ev, Bk[k]);
                fused_unit<KB>(o, lane, T, u * KB, Bk, V, rin, ain, pin, rout, aout, pout, ev, v);
            }
        }
        // slices are heavier than brick units: spread over the XCD's
        // workgroups first (slice j goes to workgroup j % (G / 8))
        for (int sl = s0 + m + PX * w; sl < s1; sl += PX * NW)
            fused_slice(o, lane, T, sl, val, V, rin, ain, pin, rout, aout, pout, v);
    }
    fused_write_partials<TB>(o, s, v);
    state_out<1>(o, s, S);
}

// The fused launch with the configured workgroup size.
void launch_fused(dim3 grid, hipStream_t st, const Ops &o, int s, int nb_prev, int pb, const float *val,
                  const FusedVec &V, const float *bv, const float *bm, const float *ev) {
    const int tb = fused_tb();
    if (tb == 256)
        hipLaunchKernelGGL((k_cg_fused<256, kFusedBricks>), grid, dim3(256), 0, st, o, s, nb_prev, pb, val, V, bv,
                           bm, ev);
    else if (tb == 512)
        hipLaunchKernelGGL((k_cg_fused<512, kFusedBricks>), grid, dim3(512), 0, st, o, s, nb_prev, pb, val, V, bv,
                           bm, ev);
    else
        hipLaunchKernelGGL((k_cg_fused<1024, kFusedBricks>), grid, dim3(1024), 0, st, o, s, nb_prev, pb, val, V,
                           bv, bm, ev);
}

// Refinement for the fused launches, part 1: the trigger (the last launch's
// |r| below reliable * the last true residual, or a stop the launches
// asked for), then x64 += x32 ; x32 = 0 ; the direction of the last update
// moves to buffer 0 (where the rebuilding launch reads p_old).
__global__ __launch_bounds__(RB) void k_mxf_accumulate(Ops o, int s, int nb_prev, float *__restrict__ x32,
                                                      double *__restrict__ x64, float *__restrict__ p0,
                                                      const float *__restrict__ p1) {
    KState S[1];
    S[0] = state_in(o, s, 0);
    bool due = S[0].phase == PH_RESTART;
    if (S[0].phase == PH_ACTIVE) {
        const bool mx[kFQ] = {(bool)o.norm_inf, false, false, false, false};
        double tot[1][kFQ];
        reduce_prev<1, kFQ, RB>(o, s, nb_prev, mx, tot);
        const double nr = nrm_fin(tot[0][FQ_NR], o.norm_inf);
        if (nr < S[0].rel * S[0].scale || S[0].iter - S[0].iter_ref >= o.reliable_run) {
            S[0].phase = PH_RESTART;
            S[0].rec = nr;
            due = true;
        }
    }
    if (due) {
        const bool mv = S[0].pcur == 0;  // p_old of the last update is in buffer pcur ^ 1
        int lo, hi;
        chunk_range(o.C, gridDim.x, xcd_tile(blockIdx.x, gridDim.x), lo, hi);
        for (int i = lo + threadIdx.x; i < hi; i += RB) {
            x64[i] += (double)x32[i];
            x32[i] = 0.0f;
            if (mv) p0[i] = p1[i];
        }
        S[0].pcur = 1;
    }
    state_out<1>(o, s, S);
}

// FP32 iterations after which the mixed PCG refines whatever its residual
// drop (GRIDFOAM_MIXED_RUN overrides the default for tuning).
int mixed_run() {
    static int v = [] {
        const char *e = getenv("GRIDFOAM_MIXED_RUN");
        const int r = e ? atoi(e) : 128;
        return r < 1 ? 1 : r;
    }();
    return v;
}

// Adaptive refinement trigger of the fused mixed-precision PCG: a solve
// starts refining once its FP32 recursive residual has dropped by
// `reliable` since the last true residual; a refinement whose true residual
// exceeds kDrift times the recursive one that triggered it multiplies that
// factor by kRelGrow, up to kRelMax (ill-conditioned systems, whose FP32
// recursion drifts early, end up refining after smaller drops).
constexpr double kDrift = 2.0, kRelGrow = 10.0, kRelMax = 0.5;

// Refinement, part 3 (one workgroup): decide on the true residual (first
// solve start or refinement), or pass the last fused launch's partials
// through.  Either way one reduced set of the kFQ quantities is written for
// the fused launch that follows.
__global__ __launch_bounds__(RB) void k_mxf_refine(Ops o, int s, int first, int nb_res, int nb_upd) {
    __shared__ KState sh[1];
    __shared__ double red[kFQ];
    load_state<1>(o, s, sh);
    const bool rep = first || sh[0].phase == PH_RESTART;
    const bool act = sh[0].phase == PH_ACTIVE;
    if (rep || act) {
        const bool mx[kFQ] = {(bool)o.norm_inf, false, false, false, false};
        double tot[1][kFQ];
        if (rep) {
            const bool mx2[2] = {(bool)o.norm_inf, false};
            double t2[1][2];
            reduce_prev<1, 2, RB>(o, s, nb_res, mx2, t2);
            tot[0][FQ_NR] = t2[0][0];
            tot[0][FQ_G] = t2[0][1];
            tot[0][FQ_D] = tot[0][FQ_M] = tot[0][FQ_N] = 0.0;
        } else {
            reduce_prev<1, kFQ, RB>(o, s, nb_upd, mx, tot);
        }
        if (threadIdx.x == 0) {
            for (int q = 0; q < kFQ; ++q) red[q] = tot[0][q];
            if (rep) {
                KState &S = sh[0];
                const double nr = nrm_fin(tot[0][FQ_NR], o.norm_inf);
                if (first) {
                    S.norm0 = nr;
                    S.thresh = fmax(o.atol, o.rtol * nr);
                    S.iter = 0;
                    S.restarts = 0;
                    S.converged = 0;
                    S.rz_old = tot[0][FQ_G];
                    S.pcur = 0;
                    S.rel = o.reliable;
                    S.rec = 0.0;
                } else if (S.rec > 0.0 && nr > kDrift * S.rec) {
                    // The FP32 recursion had drifted from the true residual:
                    // refine after smaller drops for the rest of this solve.
                    S.rel = fmin(S.rel * kRelGrow, kRelMax);
                }
                S.final_res = nr;
                if (nr < S.thresh) {
                    S.converged = 1;
                    S.phase = PH_DONE;
                } else if (S.iter >= o.max_iter) {
                    S.phase = PH_DONE;
                } else {
                    S.scale = nr;  // true residual at this refinement
                    S.iter_ref = S.iter;
                    S.fresh = first ? 1 : 2;
                    S.phase = PH_ACTIVE;
                }
            }
        }
        lds_sync();
        if (blockIdx.x == 0 && threadIdx.x < kFQ) {
            double *P = o.partials + (size_t)(s & 1) * kMaxRhs * kNQ * o.pstride;
            P[(size_t)threadIdx.x * o.pstride] = red[threadIdx.x];
        }
    }
    store_state<1>(o, s, sh);
}

// r = b - A x ; z = M^-1 r ; partials |r|, r.z  (cg.py:88-118)
template <int NR, int TB>
__global__ __launch_bounds__(TB) void k_cg_init(Ops o, int s, const double *__restrict__ val,
                                                const double *__restrict__ dg, const double *__restrict__ inv_d,
                                                const double *__restrict__ b, const double *__restrict__ x,
                                                double *__restrict__ r_, double *__restrict__ z_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    const int i = xcd_tile(blockIdx.x, gridDim.x) * TB + threadIdx.x;
    double v[NR][2] = {};
    if (i < o.C) {
        double ax[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, x, o.C, ax);
        const double id = inv_d[i];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * o.C + i;
            const double rr = b[k] - ax[r];
            const double zz = id * rr;
            r_[k] = rr;
            z_[k] = zz;
            v[r][0] = o.norm_inf ? fabs(rr) : rr * rr;
            v[r][1] = rr * zz;
        }
    }
    const bool mx[2] = {(bool)o.norm_inf, false};
    write_partials<NR, 2, TB>(o, s, v, mx);
    store_state<NR>(o, s, sh);
}

// Decide on |r0| (one workgroup): thresh, zero-iteration exit, rz_old.
template <int NR>
__global__ __launch_bounds__(RB) void k_cg_init_fin(Ops o, int s, int nb_prev) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    const bool mx[2] = {(bool)o.norm_inf, false};
    double tot[NR][2];
    reduce_prev<NR, 2, RB>(o, s, nb_prev, mx, tot);
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            KState &S = sh[r];
            const double n0 = nrm_fin(tot[r][0], o.norm_inf);
            S.norm0 = n0;
            S.final_res = n0;
            S.thresh = fmax(o.atol, o.rtol * n0);
            S.iter = 0;
            S.restarts = 0;
            S.rz_old = tot[r][1];
            S.converged = n0 < S.thresh;
            S.phase = S.converged ? PH_DONE : PH_ACTIVE;
            S.fresh = 1;
            S.pcur = 0;
        }
    }
    store_state<NR>(o, s, sh);
}

// ---------------------------------------------------------------- mixed precision
// FP32-storage PCG with FP64 residual replacement ("reliable updates").
// Off-diagonal values and every CG vector are stored in FP32, the diagonal
// as the FP32-rounded FP64 row sum (see the fused launch above).  FP32 x
// FP32 products are exact in FP64 and are accumulated in FP64, so the
// near-null (constant) mode of a reference-pinned Neumann pressure matrix
// does not suffer cancellation.  At the end of a chunk, when the recursive
// residual has dropped by `reliable` since the last refinement (or below
// the threshold, or the budget ran out), the FP32 correction is added to
// the FP64 solution, the FP64 residual b - A x replaces the recursive one
// and the next launch rebuilds the direction from it; convergence is
// decided on that true residual only.

// val32 = (float) val64 ; e = (float)(diag + sum_row val64): the FP64 row sum
__global__ void k_mxf_prep(int C, const int32_t *sell_ptr, const double *val64, const double *diag,
                           const int32_t *perm, float *val32, float *e) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const int sl = i >> 6, lane = i & 63;
    const int base = sell_ptr[sl];
    const int W = (sell_ptr[sl + 1] - base) >> 6;
    double sum = diag[perm ? perm[i] : i];
    for (int k = 0; k < W; ++k) {
        const int q = base + k * 64 + lane;
        const double v = val64[q];
        val32[q] = (float)v;
        sum += v;
    }
    e[i] = (float)sum;
}

// Refinement, part 2: r64 = b - A64 x64 ; r32 = r64 ; z = M^-1 r32 ;
// partials |r64| (norm type), r32.z.  `first` also starts the solve (x32 = 0).
template <int TB>
__global__ __launch_bounds__(TB) void k_mx_residual(Ops o, int s, int first, const double *__restrict__ val,
                                                    const double *__restrict__ dg, const float *__restrict__ inv_d,
                                                    const double *__restrict__ b, const double *__restrict__ x64,
                                                    float *__restrict__ r32, float *__restrict__ x32) {
    __shared__ KState sh[1];
    load_state<1>(o, s, sh);
    if (first || sh[0].phase == PH_RESTART) {
        const int i = xcd_tile(blockIdx.x, gridDim.x) * TB + threadIdx.x;
        double v[1][2] = {};
        if (i < o.C) {
            double ax[1];
            sell_row<double, 1>(o.sell_ptr, o.sell_col, val, dg[i], i, x64, o.C, ax);
            const double rr = b[i] - ax[0];
            const float r32v = (float)rr;
            const float zz = __fmul_rn(inv_d[i], r32v);
            r32[i] = r32v;
            if (first) x32[i] = 0.0f;
            v[0][0] = o.norm_inf ? fabs(rr) : rr * rr;
            v[0][1] = (double)r32v * (double)zz;
        }
        const bool mx[2] = {(bool)o.norm_inf, false};
        write_partials<1, 2, TB>(o, s, v, mx);
    }
    store_state<1>(o, s, sh);
}

// ---------------------------------------------------------------- BiCGSTAB (FP64)
__device__ __forceinline__ void bi_breakdown(KState &S, const Ops &o) {
    // bicgstab.py:238-246: restart while the budget lasts, else stop
    if (S.restarts >= o.max_restart || S.iter >= o.max_iter) {
        S.phase = PH_DONE;
    } else {
        S.restarts += 1;
        S.phase = PH_RESTART;
    }
}

// r = b - A x ; partial |r| ; r_hat = r ; p = r ; partial r.r  (bicgstab.py:88-118)
template <int NR>
__global__ __launch_bounds__(SB) void k_bi_init(Ops o, int s, const double *val, const double *dg, const double *b,
                                                const double *x, double *r_, double *rh_, double *p_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    int lo, hi;
    spmv_range(o.C, blockIdx.x, lo, hi);
    double v[NR][2] = {};
    for (int i = lo + threadIdx.x; i < hi; i += SB) {
        double ax[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, x, o.C, ax);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * o.C + i;
            const double rr = b[k] - ax[r];
            r_[k] = rr;
            rh_[k] = rr;
            p_[k] = rr;
            v[r][0] = o.norm_inf ? fmax(v[r][0], fabs(rr)) : v[r][0] + rr * rr;
            v[r][1] += rr * rr;
        }
    }
    const bool mx[2] = {(bool)o.norm_inf, false};
    write_partials<NR, 2, SB>(o, s, v, mx);
    store_state<NR>(o, s, sh);
}

template <int NR>
__global__ __launch_bounds__(RB) void k_bi_init_fin(Ops o, int s, int nb_prev) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    const bool mx[2] = {(bool)o.norm_inf, false};
    double tot[NR][2];
    reduce_prev<NR, 2>(o, s, nb_prev, mx, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r) {
            KState &S = sh[r];
            const double n0 = nrm_fin(tot[r][0], o.norm_inf);
            S.norm0 = n0;
            S.final_res = n0;
            S.thresh = fmax(o.atol, o.rtol * n0);
            S.iter = 0;
            S.restarts = 0;
            S.rho_old = tot[r][1];
            S.converged = n0 < S.thresh;
            S.phase = S.converged ? PH_DONE : PH_ACTIVE;
        }
    store_state<NR>(o, s, sh);
}

// restart (if pending): r_hat = r ; p = r ; partial r.r
template <int NR>
__global__ __launch_bounds__(RB) void k_bi_restart(Ops o, int s, const double *r_, double *rh_, double *p_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (any_phase<NR>(sh, PH_RESTART)) {
        int lo, hi;
        row_range(o.C, o.nb, blockIdx.x, lo, hi);
        double v[NR][1] = {};
        for (int i = lo + threadIdx.x; i < hi; i += RB)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (sh[r].phase != PH_RESTART) continue;
                const size_t k = (size_t)r * o.C + i;
                const double rr = r_[k];
                rh_[k] = rr;
                p_[k] = rr;
                v[r][0] += rr * rr;
            }
        const bool mx[1] = {false};
        write_partials<NR, 1>(o, s, v, mx);
    }
    store_state<NR>(o, s, sh);
}

// (finish restart: rho_old) ; y = M^-1 p
template <int NR>
__global__ __launch_bounds__(RB) void k_bi_y(Ops o, int s, const double *inv_d, const double *p_, double *y_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (any_phase<NR>(sh, PH_RESTART)) {
        const bool mx[1] = {false};
        double tot[NR][1];
        reduce_prev<NR, 1>(o, s, o.nb, mx, tot);
        if (threadIdx.x == 0)
            for (int r = 0; r < NR; ++r)
                if (sh[r].phase == PH_RESTART) {
                    sh[r].rho_old = tot[r][0];
                    sh[r].phase = PH_ACTIVE;
                }
        __syncthreads();
    }
    if (any_phase<NR>(sh, PH_ACTIVE)) {
        int lo, hi;
        row_range(o.C, o.nb, blockIdx.x, lo, hi);
        for (int i = lo + threadIdx.x; i < hi; i += RB) {
            const double id = inv_d[i];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (sh[r].phase != PH_ACTIVE) continue;
                const size_t k = (size_t)r * o.C + i;
                y_[k] = id * p_[k];
            }
        }
    }
    store_state<NR>(o, s, sh);
}

// v = A y ; partials r_hat.v, |r_hat|^2, |v|^2
template <int NR>
__global__ __launch_bounds__(SB) void k_bi_v(Ops o, int s, const double *val, const double *dg, const double *y_,
                                             const double *rh_, double *v_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (any_phase<NR>(sh, PH_ACTIVE)) {
        int lo, hi;
        spmv_range(o.C, blockIdx.x, lo, hi);
        double v[NR][3] = {};
        for (int i = lo + threadIdx.x; i < hi; i += SB) {
            double av[NR];
            sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, y_, o.C, av);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const size_t k = (size_t)r * o.C + i;
                v_[k] = av[r];
                const double h = rh_[k];
                v[r][0] += h * av[r];
                v[r][1] += h * h;
                v[r][2] += av[r] * av[r];
            }
        }
        const bool mx[3] = {false, false, false};
        write_partials<NR, 3, SB>(o, s, v, mx);
    }
    store_state<NR>(o, s, sh);
}

// iteration count ; breakdown test on r_hat.v ; alpha ; s = r - alpha v ; partial |s|
template <int NR>
__global__ __launch_bounds__(RB) void k_bi_s(Ops o, int s, const double *r_, const double *v_, double *s_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (!any_phase<NR>(sh, PH_ACTIVE)) {
        store_state<NR>(o, s, sh);
        return;
    }
    const bool mx3[3] = {false, false, false};
    double tot[NR][3];
    reduce_prev<NR, 3>(o, s, o.nbs, mx3, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r) {
            KState &S = sh[r];
            if (S.phase != PH_ACTIVE) continue;
            S.iter += 1;
            S.rhat2 = tot[r][1];
            const double den = tot[r][0];
            if (fabs(den) <= 1e-6 * sqrt(tot[r][1]) * sqrt(tot[r][2]))
                bi_breakdown(S, o);
            else
                S.alpha = S.rho_old / den;
        }
    __syncthreads();
    int lo, hi;
    row_range(o.C, o.nb, blockIdx.x, lo, hi);
    double v[NR][1] = {};
    for (int i = lo + threadIdx.x; i < hi; i += RB)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (sh[r].phase != PH_ACTIVE) continue;
            const size_t k = (size_t)r * o.C + i;
            const double ss = r_[k] - sh[r].alpha * v_[k];
            s_[k] = ss;
            v[r][0] = o.norm_inf ? fmax(v[r][0], fabs(ss)) : v[r][0] + ss * ss;
        }
    const bool mx[1] = {(bool)o.norm_inf};
    write_partials<NR, 1>(o, s, v, mx);
    store_state<NR>(o, s, sh);
}

// |s| < thresh ? x += alpha y, done : z = M^-1 s
template <int NR>
__global__ __launch_bounds__(RB) void k_bi_z(Ops o, int s, const double *inv_d, const double *y_, const double *s_,
                                             double *x_, double *z_) {
    __shared__ KState sh[NR];
    __shared__ int conv_s[NR];
    load_state<NR>(o, s, sh);
    if (!any_phase<NR>(sh, PH_ACTIVE)) {
        store_state<NR>(o, s, sh);
        return;
    }
    const bool mx1[1] = {(bool)o.norm_inf};
    double tot[NR][1];
    reduce_prev<NR, 1>(o, s, o.nb, mx1, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r) {
            conv_s[r] = 0;
            KState &S = sh[r];
            if (S.phase != PH_ACTIVE) continue;
            const double ns = nrm_fin(tot[r][0], o.norm_inf);
            S.final_res = ns;
            if (ns < S.thresh) {
                S.converged = 1;
                S.phase = PH_DONE;
                conv_s[r] = 1;
            }
        }
    __syncthreads();
    int lo, hi;
    row_range(o.C, o.nb, blockIdx.x, lo, hi);
    for (int i = lo + threadIdx.x; i < hi; i += RB) {
        const double id = inv_d[i];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * o.C + i;
            if (conv_s[r])
                x_[k] = x_[k] + sh[r].alpha * y_[k];
            else if (sh[r].phase == PH_ACTIVE)
                z_[k] = id * s_[k];
        }
    }
    store_state<NR>(o, s, sh);
}

// t = A z ; partials t.t, t.s
template <int NR>
__global__ __launch_bounds__(SB) void k_bi_t(Ops o, int s, const double *val, const double *dg, const double *z_,
                                             const double *s_, double *t_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (any_phase<NR>(sh, PH_ACTIVE)) {
        int lo, hi;
        spmv_range(o.C, blockIdx.x, lo, hi);
        double v[NR][2] = {};
        for (int i = lo + threadIdx.x; i < hi; i += SB) {
            double at[NR];
            sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, z_, o.C, at);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const size_t k = (size_t)r * o.C + i;
                t_[k] = at[r];
                v[r][0] += at[r] * at[r];
                v[r][1] += at[r] * s_[k];
            }
        }
        const bool mx[2] = {false, false};
        write_partials<NR, 2, SB>(o, s, v, mx);
    }
    store_state<NR>(o, s, sh);
}

// omega (+ breakdowns) ; x += alpha y + omega z ; r = s - omega t ; partials |r|, r_hat.r, |r|^2
template <int NR>
__global__ __launch_bounds__(RB) void k_bi_update(Ops o, int s, const double *y_, const double *z_, const double *s_,
                                                  const double *t_, const double *rh_, double *x_, double *r_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (!any_phase<NR>(sh, PH_ACTIVE)) {
        store_state<NR>(o, s, sh);
        return;
    }
    const bool mx2[2] = {false, false};
    double tot[NR][2];
    reduce_prev<NR, 2>(o, s, o.nbs, mx2, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r) {
            KState &S = sh[r];
            if (S.phase != PH_ACTIVE) continue;
            const double tt = tot[r][0];
            bool brk = fabs(tt) <= 1e-6 * sqrt(tt) * sqrt(tt);
            double om = 0.0;
            if (!brk) {
                om = tot[r][1] / tt;
                brk = fabs(om) <= 1e-6;
            }
            if (brk)
                bi_breakdown(S, o);
            else
                S.omega = om;
        }
    __syncthreads();
    int lo, hi;
    row_range(o.C, o.nb, blockIdx.x, lo, hi);
    double v[NR][3] = {};
    for (int i = lo + threadIdx.x; i < hi; i += RB)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (sh[r].phase != PH_ACTIVE) continue;
            const size_t k = (size_t)r * o.C + i;
            const double a = sh[r].alpha, w = sh[r].omega;
            const double xx = x_[k] + a * y_[k] + w * z_[k];
            const double rr = s_[k] - w * t_[k];
            x_[k] = xx;
            r_[k] = rr;
            v[r][0] = o.norm_inf ? fmax(v[r][0], fabs(rr)) : v[r][0] + rr * rr;
            v[r][1] += rh_[k] * rr;
            v[r][2] += rr * rr;
        }
    const bool mx[3] = {(bool)o.norm_inf, false, false};
    write_partials<NR, 3>(o, s, v, mx);
    store_state<NR>(o, s, sh);
}

// |r| check ; rho breakdown ; beta ; p = r + beta (p - omega v)
template <int NR>
__global__ __launch_bounds__(RB) void k_bi_p(Ops o, int s, const double *r_, const double *v_, double *p_) {
    __shared__ KState sh[NR];
    load_state<NR>(o, s, sh);
    if (!any_phase<NR>(sh, PH_ACTIVE)) {
        store_state<NR>(o, s, sh);
        return;
    }
    const bool mx3[3] = {(bool)o.norm_inf, false, false};
    double tot[NR][3];
    reduce_prev<NR, 3>(o, s, o.nb, mx3, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r) {
            KState &S = sh[r];
            if (S.phase != PH_ACTIVE) continue;
            const double nr = nrm_fin(tot[r][0], o.norm_inf);
            S.final_res = nr;
            if (nr < S.thresh) {
                S.converged = 1;
                S.phase = PH_DONE;
                continue;
            }
            const double rho_new = tot[r][1];
            if (fabs(rho_new) <= 1e-6 * sqrt(S.rhat2) * sqrt(tot[r][2])) {
                bi_breakdown(S, o);
                continue;
            }
            S.beta = (rho_new / S.rho_old) * (S.alpha / S.omega);
            S.rho_old = rho_new;
            if (S.iter >= o.max_iter) S.phase = PH_DONE;
        }
    __syncthreads();
    int lo, hi;
    row_range(o.C, o.nb, blockIdx.x, lo, hi);
    for (int i = lo + threadIdx.x; i < hi; i += RB)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (sh[r].phase != PH_ACTIVE) continue;
            const size_t k = (size_t)r * o.C + i;
            p_[k] = r_[k] + sh[r].beta * (p_[k] - sh[r].omega * v_[k]);
        }
    store_state<NR>(o, s, sh);
}

// ---------------------------------------------------------------- workspace
struct WS {
    char *base;
    size_t off;
    template <typename T>
    T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *p = reinterpret_cast<T *>(base + off);
        off += n * sizeof(T);
        return p;
    }
};

struct Layout {
    double *val64, *dg64, *inv64, *b, *x, *r, *z, *p, *p1, *ap, *y, *s, *t, *rh, *v;
    float *val32, *dg32, *inv32, *e32, *x32, *r32, *r32b, *p32, *p32b, *ap32, *ap32b;
    double *bv64, *bm64, *ev64;  // brick values (plain PCG), null without bricks
    float *bv32, *bm32, *ev32;   // brick values (mixed PCG)
    double *partials;
    KState *st;
    size_t bytes;
};

// Brick layout of the solve: the PCG of a symmetric operator with one rhs on
// a mesh whose solver layout was built.
struct BrickPlan {
    bool on = false;
    int nbs = 0, ne = 0;
};
BrickPlan brick_plan(const gf_mesh *m, int nrhs, const gf_krylov_cfg &cfg) {
    BrickPlan b;
    b.on = cfg.method == 0 && nrhs == 1 && cfg.symmetric && m->n_bricks > 0 && m->psell_ptr && m->row_perm;
    if (b.on) {
        b.nbs = m->n_bricks;
        b.ne = m->n_expl;
    }
    return b;
}

Layout layout(char *base, int64_t C, int64_t nnz, int nrhs, const gf_krylov_cfg &cfg, const BrickPlan &bp) {
    WS w{base, 0};
    Layout L{};
    const size_t n = (size_t)C * nrhs;
    L.val64 = w.take<double>(nnz);
    L.dg64 = w.take<double>(C);
    L.inv64 = w.take<double>(C);
    L.b = w.take<double>(n);
    L.x = w.take<double>(n);
    L.r = w.take<double>(n);
    if (cfg.method == 0 && cfg.mixed) {
        L.val32 = w.take<float>(nnz);
        L.dg32 = w.take<float>(C);
        L.inv32 = w.take<float>(C);
        L.e32 = w.take<float>(C);
        L.x32 = w.take<float>(n);
        L.r32 = w.take<float>(n);
        L.r32b = w.take<float>(n);
        L.p32 = w.take<float>(n);
        L.p32b = w.take<float>(n);
        L.ap32 = w.take<float>(n);
        L.ap32b = w.take<float>(n);
    } else {
        L.z = w.take<double>(n);
        L.p = w.take<double>(n);
        if (cfg.method == 0) L.p1 = w.take<double>(n);
        L.ap = w.take<double>(n);
        if (cfg.method == 1) {
            L.y = w.take<double>(n);
            L.s = w.take<double>(n);
            L.t = w.take<double>(n);
            L.rh = w.take<double>(n);
            L.v = L.ap;
        }
    }
    if (bp.on) {
        const size_t nv = (size_t)bp.nbs * 192, nm = (size_t)bp.nbs * 48, ne = (size_t)bp.ne * 64 + 1;
        if (cfg.mixed) {
            L.bv32 = w.take<float>(nv);
            L.bm32 = w.take<float>(nm);
            L.ev32 = w.take<float>(ne);
        } else {
            L.bv64 = w.take<double>(nv);
            L.bm64 = w.take<double>(nm);
            L.ev64 = w.take<double>(ne);
        }
    }
    L.partials = w.take<double>((size_t)2 * kMaxRhs * kNQ * solve_pstride((int)C, bp.on ? bp.nbs : 0));
    L.st = w.take<KState>(2 * kMaxRhs);
    L.bytes = w.off + 256;
    return L;
}

// Pinned host memory of this thread: the final solver state, and two poll
// words the last launch of each chunk writes directly (no copy on the stream).
struct HostStage {
    KState *state = nullptr;      // [kMaxRhs] host view
    KState *dev_state = nullptr;  // device view
    int *words = nullptr;         // [2] host view
    int *dev_words = nullptr;     // [2] device view
};
// One stage per thread that solves; a thread that ends returns its stage to
// a pool the next new thread takes it from, so short-lived worker threads do
// not accumulate pinned allocations (and no HIP call runs at thread exit).
struct StagePool {
    std::mutex mu;
    std::vector<HostStage> free;
};
// Never destroyed: threads that outlive static destruction (the autograd
// engine's device threads run the backward solves and end after it at
// process exit) still return their stage from ~StageHolder.
StagePool &stage_pool() {
    static StagePool *p = new StagePool;
    return *p;
}
struct StageHolder {
    HostStage hs;
    ~StageHolder() {
        if (!hs.state) return;
        StagePool &p = stage_pool();
        std::lock_guard<std::mutex> g(p.mu);
        p.free.push_back(hs);
    }
};
HostStage *host_stage() {
    static thread_local StageHolder holder;
    HostStage &hs = holder.hs;
    if (hs.state) return &hs;
    {
        StagePool &p = stage_pool();
        std::lock_guard<std::mutex> g(p.mu);
        if (!p.free.empty()) {
            hs = p.free.back();
            p.free.pop_back();
            return &hs;
        }
    }
    // coherent: the GPU's stores reach host memory uncached and snooped,
    // so the host never reads a stale cached copy
    const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
    HostStage n;
    if (hipHostMalloc((void **)&n.state, sizeof(KState) * kMaxRhs, fl) != hipSuccess) return nullptr;
    if (hipHostGetDevicePointer((void **)&n.dev_state, n.state, 0) != hipSuccess ||
        hipHostMalloc((void **)&n.words, 2 * sizeof(int), fl) != hipSuccess ||
        hipHostGetDevicePointer((void **)&n.dev_words, n.words, 0) != hipSuccess) {
        (void)hipHostFree(n.state);
        if (n.words) (void)hipHostFree(n.words);
        return nullptr;
    }
    hs = n;
    return &hs;
}

// Chunk k's last launch posts to word k & 1; the host queues chunk k + 1
// before it waits for chunk k (an event behind it) and reads that word.  The
// host clears the word to kNoPost before it queues the chunk (the chunk that
// used it last has completed), so a word that still reads kNoPost after the
// event means the post is missing.  (Replayed graphs carry the sequence
// numbers of their capture, so the posted value cannot identify the chunk.)
constexpr int kNoPost = -1;
struct Poller {
    hipEvent_t ev[2] = {nullptr, nullptr};
    HostStage *h = nullptr;
    int chunk = 0;
    int init() {
        h = host_stage();
        GF_REQUIRE(h != nullptr, "pinned staging allocation failed");
        GF_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        GF_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        return GF_OK;
    }
    // The last chunk's event may still be pending when a solve ends (the
    // loop waits for the chunk before it); the owner keeps the Poller alive
    // until the stream has been synchronised, and an early error exit waits
    // here, so no event is destroyed while the GPU may still complete it.
    ~Poller() {
        for (hipEvent_t e : ev)
            if (e) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
            }
    }
    // cleared device address of the word the next chunk posts to
    int *next_word() const {
        __atomic_store_n(h->words + (chunk & 1), kNoPost, __ATOMIC_RELEASE);
        return h->dev_words + (chunk & 1);
    }
    // after queueing a chunk: records its event, waits for the previous chunk
    // and reports whether it finished
    int poll(hipStream_t st, bool *done) {
        GF_CHECK_HIP(hipEventRecord(ev[chunk & 1], st));
        *done = false;
        if (chunk > 0) {
            const int k = (chunk - 1) & 1;
            GF_CHECK_HIP(hipEventSynchronize(ev[k]));
            const int w = __atomic_load_n(h->words + k, __ATOMIC_ACQUIRE);
            GF_REQUIRE(w != kNoPost, "solver chunk completed without posting its state");
            *done = (w & 1) != 0;
        }
        ++chunk;
        return GF_OK;
    }
};

// Statistics from the final state k_solve_out wrote (after the stream wait).
int finish_read(int nr, double *stats) {
    const KState *h = host_stage()->state;
    for (int r = 0; r < nr; ++r) {
        stats[r * 4 + 0] = h[r].norm0;
        stats[r * 4 + 1] = h[r].final_res;
        stats[r * 4 + 2] = h[r].iter;
        stats[r * 4 + 3] = h[r].converged;
    }
    return GF_OK;
}

// ---- launch-sequence graphs --------------------------------------------
// A chunk of `every` PCG iterations (plus the refinement trio in mixed
// precision) is captured once into a HIP graph and replayed: one host call
// per chunk instead of 2*every+3 launches.  Launches read only the parity of
// their sequence number, so a graph captured at sequence s0 replays any chunk
// that starts at s0 + 2k.  Graphs are cached per workspace (they embed its
// pointers) and dropped by gf_krylov_release.
// One mutex guards the cache AND every capture/instantiation: the host
// threads of a serving run solve concurrently, and capture plus
// instantiation are done by one thread at a time (a miss is captured while
// the lock is held, so no two threads capture the same chunk either).
// Replays (hipGraphLaunch) run outside the lock.
struct GraphCache {
    std::mutex mu;
    std::unordered_map<const void *, std::vector<std::pair<std::string, hipGraphExec_t>>> by_ws;
};
GraphCache &graph_cache() {
    static GraphCache *c = new GraphCache;  // never destroyed, as the stage pool
    return *c;
}
std::atomic<int> &graphs_flag() {
    static std::atomic<int> on{[] {
        const char *e = getenv("GRIDFOAM_KRYLOV_GRAPHS");
        return (e && atoi(e) == 0) ? 0 : 1;
    }()};
    return on;
}
bool graphs_enabled() { return graphs_flag().load() != 0; }
template <typename T>
void key_add(std::string &k, const T &v) {
    k.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

template <int NR, int TB>
void enqueue_cg_chunk(const Ops &o, const Layout &L, bool mixed, int every, int s, int first_prev, int *poll,
                      hipStream_t st) {
    const dim3 GS(o.gs), GD(o.gd), BS(TB), GU(o.gu), BU(RB), G1(1);
    Ops ol = o;  // the chunk's last launch posts the poll word
    ol.poll = poll;
    if (!mixed) {
        for (int j = 0; j < every; ++j) {
            const int nb_prev = j == 0 ? first_prev : o.gu;
            const int pb = j & 1;  // buffer of p_old; `every` is even
            hipLaunchKernelGGL((k_cg_dir<double, NR, false, TB>), GD, BS, 0, st, o, s++, nb_prev, 0, pb, L.val64,
                               L.dg64, L.z, L.p, L.p1, L.ap, L.bv64, L.bm64, L.ev64);
            hipLaunchKernelGGL((k_cg_upd<double, NR>), GU, BU, 0, st, j == every - 1 ? ol : o, s++, pb ^ 1, L.inv64,
                               L.x, L.r, L.z, L.p, L.p1, L.ap);
        }
        return;
    }
    if constexpr (NR == 1) {
        // one fused launch per iteration (buffers alternate by launch; `every`
        // is even), then the FP64 refinement (no-ops unless it is due)
        const FusedVec V{L.e32, L.inv32, L.x32, L.r32, L.r32b, L.ap32, L.ap32b, L.p32, L.p32b};
        for (int j = 0; j < every; ++j)
            launch_fused(GD, st, o, s++, j == 0 ? first_prev : o.gd, j & 1, L.val32, V, L.bv32, L.bm32, L.ev32);
        hipLaunchKernelGGL(k_mxf_accumulate, GU, BU, 0, st, o, s++, o.gd, L.x32, L.x, L.p32, L.p32b);
        hipLaunchKernelGGL((k_mx_residual<TB>), GS, BS, 0, st, o, s++, 0, L.val64, L.dg64, L.inv32, L.b, L.x, L.r32,
                           L.x32);
        hipLaunchKernelGGL(k_mxf_refine, G1, BU, 0, st, ol, s++, 0, o.gs, o.gd);
    }
}

// Enqueue one chunk starting at sequence s, through a cached graph when possible.
template <int NR, int TB>
int launch_cg_chunk(const void *ws, const Ops &o, const Layout &L, bool mixed, int every, int s, int first_prev,
                    int *poll, hipStream_t st) {
    if (!graphs_enabled()) {
        enqueue_cg_chunk<NR, TB>(o, L, mixed, every, s, first_prev, poll, st);
        GF_CHECK_LAUNCH();
        return GF_OK;
    }
    std::string key;
    key_add(key, o);
    key_add(key, L);
    for (int v : {NR, TB, (int)mixed, every, s & 1, first_prev}) key_add(key, v);
    key_add(key, poll);
    GraphCache &c = graph_cache();
    hipGraphExec_t exec = nullptr;
    std::unique_lock<std::mutex> lock(c.mu);
    for (auto &kv : c.by_ws[ws])
        if (kv.first == key) exec = kv.second;
    if (!exec) {
        // Capture on a side stream of this thread (the caller's stream may be
        // the legacy default stream, which cannot be captured); the graph is
        // then launched on the caller's stream.
        static thread_local hipStream_t cap = nullptr;
        if (!cap) GF_CHECK_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        GF_CHECK_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        enqueue_cg_chunk<NR, TB>(o, L, mixed, every, s, first_prev, poll, cap);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(cap, &graph);
        if (le != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            set_error(std::string("cg chunk capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
            return GF_EHIP;
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            set_error(std::string("cg chunk graph instantiation failed: ") + hipGetErrorString(ie));
            return GF_EHIP;
        }
        c.by_ws[ws].emplace_back(key, exec);
    }
    lock.unlock();
    GF_CHECK_HIP(hipGraphLaunch(exec, st));
    return GF_OK;
}

template <int NR, int TB>
int run_cg(const void *ws, Ops o, const Layout &L, const gf_krylov_cfg &cfg, hipStream_t st, Poller &P,
           int *s_last) {
    if (int rc = P.init()) return rc;
    // iterations per chunk: even, so every chunk starts with p_old in buffer 0
    const int every = cfg.check_every > 1 ? (cfg.check_every + 1) & ~1 : 2;
    int s = 1;  // launch sequence; state[0] holds the zeroed initial state
    const dim3 GS(o.gs), BS(TB), BU(RB), G1(1);
    // Partial sets the first k_cg_dir of a chunk reduces: the refinement's one
    // reduced set in mixed precision; the update's in FP64 (the very first
    // direction ignores them, so every chunk is the same launch sequence).
    const int first_prev = cfg.mixed ? 1 : o.gu;
    if (!cfg.mixed) {
        hipLaunchKernelGGL((k_cg_init<NR, TB>), GS, BS, 0, st, o, s++, L.val64, L.dg64, L.inv64, L.b, L.x, L.r, L.z);
        hipLaunchKernelGGL(k_cg_init_fin<NR>, G1, BU, 0, st, o, s++, o.gs);
    } else {
        hipLaunchKernelGGL((k_mx_residual<TB>), GS, BS, 0, st, o, s++, 1, L.val64, L.dg64, L.inv32, L.b, L.x, L.r32,
                           L.x32);
        hipLaunchKernelGGL(k_mxf_refine, G1, BU, 0, st, o, s++, 1, o.gs, o.gd);
    }
    GF_CHECK_LAUNCH();
    const int per_chunk = cfg.mixed ? every + 3 : 2 * every;
    int k = 1;
    const long cap = (long)cfg.max_iter + 4L * every;
    for (;;) {
        if (int rc = launch_cg_chunk<NR, TB>(ws, o, L, cfg.mixed != 0, every, s, first_prev, P.next_word(), st))
            return rc;
        s += per_chunk;
        k += every;
        bool done = false;
        if (int rc = P.poll(st, &done)) return rc;
        if (done) break;
        if (k > cap) {
            set_error("cg: solver state machine did not finish");
            return GF_EINVAL;
        }
    }
    *s_last = s - 1;
    return GF_OK;
}

template <int NR>
int run_bicgstab(Ops o, const Layout &L, const gf_krylov_cfg &cfg, hipStream_t st, Poller &P, int *s_last) {
    const int nb = o.nb;
    if (int rc = P.init()) return rc;
    const int every = cfg.check_every > 0 ? cfg.check_every : 4;
    int s = 1;
    const dim3 G(nb), B(RB), G1(1), GS(o.nbs), BS(SB);
    hipLaunchKernelGGL(k_bi_init<NR>, GS, BS, 0, st, o, s++, L.val64, L.dg64, L.b, L.x, L.r, L.rh, L.p);
    hipLaunchKernelGGL(k_bi_init_fin<NR>, G1, B, 0, st, o, s++, o.nbs);
    GF_CHECK_LAUNCH();
    int k = 0;
    Ops ol = o;  // the chunk's last launch posts the poll word
    for (;;) {
        ol.poll = P.next_word();
        for (int j = 0; j < every; ++j, ++k) {
            hipLaunchKernelGGL(k_bi_restart<NR>, G, B, 0, st, o, s++, L.r, L.rh, L.p);
            hipLaunchKernelGGL(k_bi_y<NR>, G, B, 0, st, o, s++, L.inv64, L.p, L.y);
            hipLaunchKernelGGL(k_bi_v<NR>, GS, BS, 0, st, o, s++, L.val64, L.dg64, L.y, L.rh, L.v);
            hipLaunchKernelGGL(k_bi_s<NR>, G, B, 0, st, o, s++, L.r, L.v, L.s);
            hipLaunchKernelGGL(k_bi_z<NR>, G, B, 0, st, o, s++, L.inv64, L.y, L.s, L.x, L.z);
            hipLaunchKernelGGL(k_bi_t<NR>, GS, BS, 0, st, o, s++, L.val64, L.dg64, L.z, L.s, L.t);
            hipLaunchKernelGGL(k_bi_update<NR>, G, B, 0, st, o, s++, L.y, L.z, L.s, L.t, L.rh, L.x, L.r);
            hipLaunchKernelGGL(k_bi_p<NR>, G, B, 0, st, j == every - 1 ? ol : o, s++, L.r, L.v, L.p);
        }
        GF_CHECK_LAUNCH();
        bool done = false;
        if (int rc = P.poll(st, &done)) return rc;
        if (done) break;
        if (k > cfg.max_iter + 4 * every + 8) {
            set_error("bicgstab: solver state machine did not finish");
            return GF_EINVAL;
        }
    }
    *s_last = s - 1;
    return GF_OK;
}

}  // namespace
}  // namespace gf

using namespace gf;

extern "C" int64_t gf_krylov_workspace_bytes(const gf_mesh *m, int32_t nrhs, const gf_krylov_cfg *cfg) {
    if (!m || !cfg || nrhs < 1 || nrhs > kMaxRhs) return -1;
    const BrickPlan bp = brick_plan(m, nrhs, *cfg);
    Layout L = layout(nullptr, m->n_cells, bp.on ? m->pnnz_pad : m->nnz_pad, nrhs, *cfg, bp);
    return (int64_t)L.bytes;
}

extern "C" int gf_krylov_solve(const gf_mesh *m, const gf_krylov_cfg *cfg, int32_t nrhs, int32_t transpose,
                               const double *diag, const double *upper, const double *lower, const double *b,
                               double *x, void *workspace, double *stats_out, gf_stream_t stream) {
    GF_REQUIRE(m && cfg && diag && upper && lower && b && x && workspace && stats_out, "null argument");
    GF_REQUIRE(nrhs == 1 || nrhs == 3, "nrhs must be 1 or 3");
    GF_REQUIRE(cfg->method == 0 || cfg->method == 1, "unknown Krylov method");
    GF_REQUIRE(!(cfg->method == 1 && cfg->mixed), "mixed precision is implemented for cg only");
    GF_REQUIRE(!(cfg->mixed && nrhs != 1), "mixed precision is implemented for one right-hand side");
    GF_REQUIRE(!cfg->mixed || (cfg->inner_rtol > 0.0 && cfg->inner_rtol < 1.0),
               "mixed precision needs 0 < inner_rtol < 1 (refinement trigger)");
    GF_REQUIRE(cfg->max_iter >= 1, "max_iter must be >= 1");
    GF_REQUIRE(m->n_cells >= 1, "empty grid");
    hipStream_t s = S(stream);
    const int C = m->n_cells;
    // PCG of a symmetric operator with one rhs runs on the solver row order
    // (bricks first); everything else on the grid's order
    const BrickPlan bp = brick_plan(m, nrhs, *cfg);
    const int32_t *perm = bp.on ? m->row_perm : nullptr;
    const int32_t *sptr = bp.on ? m->psell_ptr : m->sell_ptr;
    const int32_t *scol = bp.on ? m->psell_col : m->sell_col;
    const int32_t *scode = bp.on ? m->psell_code : m->sell_code;
    const int64_t nnz = bp.on ? m->pnnz_pad : m->nnz_pad;
    Layout L = layout(reinterpret_cast<char *>(workspace), C, nnz, nrhs, *cfg, bp);
    hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, scode, upper, lower,
                       transpose, L.val64);
    hipLaunchKernelGGL(k_diag_prep<double>, dim3(grid_for(C)), dim3(kBlock), 0, s, C, diag, perm, cfg->precond,
                       L.dg64, L.inv64);
    if (cfg->method == 0 && cfg->mixed) {
        hipLaunchKernelGGL(k_mxf_prep, dim3(grid_for(C)), dim3(kBlock), 0, s, C, sptr, L.val64, diag, perm, L.val32,
                           L.e32);
        hipLaunchKernelGGL(k_diag_prep<float>, dim3(grid_for(C)), dim3(kBlock), 0, s, C, diag, perm, cfg->precond,
                           L.dg32, L.inv32);
    }
    if (bp.on) {  // brick values, in the precision of the direction kernel
        const int64_t nv = (int64_t)bp.nbs * 192, nm = (int64_t)bp.nbs * 48, ne = (int64_t)bp.ne * 64;
        if (cfg->mixed) {
            hipLaunchKernelGGL(k_sell_pack<float>, dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, m->brick_vp, upper,
                               lower, transpose, L.bv32);
            hipLaunchKernelGGL(k_sell_pack<float>, dim3(grid_for(nm)), dim3(kBlock), 0, s, nm, m->brick_vm, upper,
                               lower, transpose, L.bm32);
            if (ne)
                hipLaunchKernelGGL(k_sell_pack<float>, dim3(grid_for(ne)), dim3(kBlock), 0, s, ne, m->expl_code,
                                   upper, lower, transpose, L.ev32);
        } else {
            hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, m->brick_vp, upper,
                               lower, transpose, L.bv64);
            hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(nm)), dim3(kBlock), 0, s, nm, m->brick_vm, upper,
                               lower, transpose, L.bm64);
            if (ne)
                hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(ne)), dim3(kBlock), 0, s, ne, m->expl_code,
                                   upper, lower, transpose, L.ev64);
        }
    }
    hipLaunchKernelGGL(k_zero_state, dim3(1), dim3(64), 0, s, L.st, 2 * kMaxRhs);
    hipLaunchKernelGGL(k_solve_in, dim3(grid_for(C)), dim3(kBlock), 0, s, C, nrhs, b, x, perm, L.b, L.x);
    GF_CHECK_LAUNCH();
    HostStage *hs = host_stage();
    GF_REQUIRE(hs != nullptr, "pinned staging allocation failed");
    Ops o{};
    o.sell_ptr = sptr;
    o.sell_col = scol;
    if (bp.on) {
        o.bnbr = m->brick_nbr;
        o.ecol = m->expl_col;
        o.nbrick = bp.nbs;
    }
    o.C = C;
    o.nb = solver_workgroups(C);
    o.nbs = spmv_workgroups(C);
    o.tb = pcg_tile();
    o.gs = pcg_spmv_workgroups(C);
    o.gd = (cfg->method == 0 && cfg->mixed) ? fused_workgroups(C, o.nbrick)
                                            : pcg_dir_workgroups(C, o.nbrick);
    o.gu = pcg_vector_workgroups(C);
    o.pstride = solve_pstride(C, o.nbrick);
    o.norm_inf = cfg->norm;
    o.max_iter = cfg->max_iter;
    o.max_restart = cfg->max_restart;
    o.atol = cfg->atol;
    o.rtol = cfg->rtol;
    o.reliable = cfg->inner_rtol;
    o.reliable_run = mixed_run();
    o.st = L.st;
    o.partials = L.partials;
    int rc;
    int s_last = 0;
    Poller P;  // outlives the final stream synchronisation below
    if (cfg->method == 0) {
        auto cg = [&](auto nr, auto tb) {
            return run_cg<decltype(nr)::value, decltype(tb)::value>(workspace, o, L, *cfg, s, P, &s_last);
        };
        using I1 = std::integral_constant<int, 1>;
        using I3 = std::integral_constant<int, 3>;
        using T256 = std::integral_constant<int, 256>;
        using T512 = std::integral_constant<int, 512>;
        using T1024 = std::integral_constant<int, 1024>;
        if (o.tb == 1024)
            rc = nrhs == 1 ? cg(I1{}, T1024{}) : cg(I3{}, T1024{});
        else if (o.tb == 512)
            rc = nrhs == 1 ? cg(I1{}, T512{}) : cg(I3{}, T512{});
        else
            rc = nrhs == 1 ? cg(I1{}, T256{}) : cg(I3{}, T256{});
    }
    else
        rc = nrhs == 1 ? run_bicgstab<1>(o, L, *cfg, s, P, &s_last) : run_bicgstab<3>(o, L, *cfg, s, P, &s_last);
    if (rc != GF_OK) return rc;
    hipLaunchKernelGGL(k_solve_out, dim3(grid_for(C)), dim3(kBlock), 0, s, C, nrhs, L.x, perm, x,
                       L.st + (s_last & 1) * kMaxRhs, hs->dev_state);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(s));  // the solution and the final state, in one wait
    return finish_read(nrhs, stats_out);
}

// Replay PCG chunks through cached graphs (1, the default) or launch them one
// by one (0).  Returns the previous setting.
extern "C" int gf_krylov_use_graphs(int32_t on) { return graphs_flag().exchange(on ? 1 : 0); }

// Drop the launch graphs cached for a workspace (call before freeing it).
extern "C" int gf_krylov_release(const void *workspace) {
    GraphCache &c = graph_cache();
    std::vector<std::pair<std::string, hipGraphExec_t>> v;
    {
        std::lock_guard<std::mutex> g(c.mu);
        auto it = c.by_ws.find(workspace);
        if (it == c.by_ws.end()) return GF_OK;
        v.swap(it->second);
        c.by_ws.erase(it);
    }
    for (auto &kv : v) GF_CHECK_HIP(hipGraphExecDestroy(kv.second));
    return GF_OK;
}

// ---- benchmark of the dominant PCG kernel on its own -----------------------
// FP64: the direction/SpMV kernel k_cg_dir in its steady state with beta =
// 0, so p = z = x and Ap = A x.  FP32: the fused mixed-precision launch
// k_cg_fused in its update state with alpha = 0, beta = 1, M^-1 = 1 and
// p_old = 0, so r' = r = x, x' = x, p = z = x and Ap = A x, while every
// operand of a solver launch is read and written.  Both run on the layout a
// symmetric one-rhs PCG solve uses (bricks first when the mesh has them),
// read the previous launch's partials and reduce them, as inside the solver.
namespace gf {
namespace {
template <typename VT>
__global__ void k_bench_in(int C, const double *__restrict__ x, const int32_t *__restrict__ perm,
                           VT *__restrict__ z, VT *__restrict__ p0, VT *__restrict__ p1) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    z[i] = (VT)x[perm ? perm[i] : i];
    p0[i] = (VT)0;
    p1[i] = (VT)0;
}
template <typename VT>
__global__ void k_bench_out(int C, const VT *__restrict__ ap, const int32_t *__restrict__ perm,
                            double *__restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    y[perm ? perm[i] : i] = (double)ap[i];
}
__global__ void k_fill_f32(int C, float v, float *__restrict__ a, float *__restrict__ b, float *__restrict__ c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    a[i] = v;
    b[i] = v;
    c[i] = v;
}
// Update state of the fused launch for the benchmark: every launch reads
// state[0] and one reduced partial set (g = 1, d = 1e300, m = n = 0, |r| = 1).
__global__ void k_fused_bench_state(KState *st, int n, double *partials, int pstride) {
    int i = threadIdx.x;
    if (i < n) {
        KState S{};
        S.phase = PH_ACTIVE;
        S.fresh = 0;
        S.rz_old = 1.0;
        S.thresh = -1.0;
        st[i] = S;
    }
    if (i < kFQ) {
        const double v[kFQ] = {1.0, 1.0, 1e300, 0.0, 0.0};
        partials[(size_t)i * pstride] = v[i];
    }
}

struct BenchLayout {
    Ops o;
    BrickPlan bp;
    const int32_t *perm, *scode;
    int64_t nnz;
    void *val, *bv, *bm, *ev, *z, *p0, *p1, *ap;
    double *dg, *val64;
    float *e32, *inv32, *x32, *r1, *a1;
    size_t bytes;
};
BenchLayout bench_layout(const gf_mesh *m, int fp32, char *base) {
    gf_krylov_cfg cfg{};
    cfg.method = 0;
    cfg.symmetric = 1;
    BenchLayout B{};
    B.bp = brick_plan(m, 1, cfg);
    B.perm = B.bp.on ? m->row_perm : nullptr;
    B.scode = B.bp.on ? m->psell_code : m->sell_code;
    B.nnz = B.bp.on ? m->pnnz_pad : m->nnz_pad;
    WS w{base, 0};
    const size_t vb = fp32 ? sizeof(float) : sizeof(double);
    const int C = m->n_cells;
    Ops &o = B.o;
    o.sell_ptr = B.bp.on ? m->psell_ptr : m->sell_ptr;
    o.sell_col = B.bp.on ? m->psell_col : m->sell_col;
    if (B.bp.on) {
        o.bnbr = m->brick_nbr;
        o.ecol = m->expl_col;
        o.nbrick = B.bp.nbs;
    }
    o.C = C;
    o.nb = solver_workgroups(C);
    o.nbs = spmv_workgroups(C);
    o.tb = pcg_tile();
    o.gs = pcg_spmv_workgroups(C);
    o.gd = fp32 ? fused_workgroups(C, o.nbrick) : pcg_dir_workgroups(C, o.nbrick);
    o.gu = pcg_vector_workgroups(C);
    o.pstride = solve_pstride(C, o.nbrick);
    o.max_iter = 1 << 30;
    o.partials = w.take<double>((size_t)2 * kMaxRhs * kNQ * o.pstride);
    o.st = w.take<KState>(2 * kMaxRhs);
    B.val = w.take<char>(vb * B.nnz);
    B.bv = w.take<char>(vb * ((size_t)B.bp.nbs * 192 + 1));
    B.bm = w.take<char>(vb * ((size_t)B.bp.nbs * 48 + 1));
    B.ev = w.take<char>(vb * ((size_t)B.bp.ne * 64 + 1));
    B.z = w.take<char>(vb * C);  // FP32: r of buffer 0
    B.p0 = w.take<char>(vb * C);
    B.p1 = w.take<char>(vb * C);
    B.ap = w.take<char>(vb * C);  // FP32: Ap of buffer 0
    B.dg = w.take<double>(C);
    if (fp32) {
        B.val64 = w.take<double>(B.nnz);
        B.e32 = w.take<float>(C);
        B.inv32 = w.take<float>(C);
        B.x32 = w.take<float>(C);
        B.r1 = w.take<float>(C);
        B.a1 = w.take<float>(C);
    }
    B.bytes = w.off + 256;
    return B;
}
template <typename VT>
void bench_pack(const gf_mesh *m, const BenchLayout &B, const double *diag, const double *upper,
                const double *lower, const double *x, hipStream_t s) {
    const int C = m->n_cells;
    hipLaunchKernelGGL(k_sell_pack<VT>, dim3(grid_for(B.nnz)), dim3(kBlock), 0, s, B.nnz, B.scode, upper, lower, 0,
                       (VT *)B.val);
    if (B.bp.on) {
        const int64_t nv = (int64_t)B.bp.nbs * 192, nm = (int64_t)B.bp.nbs * 48, ne = (int64_t)B.bp.ne * 64;
        hipLaunchKernelGGL(k_sell_pack<VT>, dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, m->brick_vp, upper, lower, 0,
                           (VT *)B.bv);
        hipLaunchKernelGGL(k_sell_pack<VT>, dim3(grid_for(nm)), dim3(kBlock), 0, s, nm, m->brick_vm, upper, lower, 0,
                           (VT *)B.bm);
        if (ne)
            hipLaunchKernelGGL(k_sell_pack<VT>, dim3(grid_for(ne)), dim3(kBlock), 0, s, ne, m->expl_code, upper,
                               lower, 0, (VT *)B.ev);
    }
    hipLaunchKernelGGL(k_diag_prep<double>, dim3(grid_for(C)), dim3(kBlock), 0, s, C, diag, B.perm, 0, B.dg,
                       (double *)nullptr);
    hipLaunchKernelGGL(k_bench_in<VT>, dim3(grid_for(C)), dim3(kBlock), 0, s, C, x, B.perm, (VT *)B.z, (VT *)B.p0,
                       (VT *)B.p1);
    if constexpr (sizeof(VT) == 4) {
        hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(B.nnz)), dim3(kBlock), 0, s, B.nnz, B.scode, upper,
                           lower, 0, B.val64);
        hipLaunchKernelGGL(k_mxf_prep, dim3(grid_for(C)), dim3(kBlock), 0, s, C, B.o.sell_ptr, B.val64, diag, B.perm,
                           (float *)B.val, B.e32);
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(C)), dim3(kBlock), 0, s, C, 1.0f, B.inv32, B.inv32, B.inv32);
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(C)), dim3(kBlock), 0, s, C, 0.0f, B.x32, (float *)B.ap, B.a1);
        hipLaunchKernelGGL(k_fused_bench_state, dim3(1), dim3(64), 0, s, B.o.st, 2 * kMaxRhs, B.o.partials,
                           B.o.pstride);
    } else {
        hipLaunchKernelGGL(k_spmv_bench_state, dim3(1), dim3(64), 0, s, B.o.st, 2 * kMaxRhs);
    }
}
template <typename VT>
void bench_run(const BenchLayout &B, int reps, hipStream_t s) {
    auto launch = [&](auto tbc) {
        constexpr int TB = decltype(tbc)::value;
        // every launch uses sequence number 1: it reads state[0] (never
        // written here) and writes state[1] and partials[1]
        for (int k = 0; k < reps; ++k) {
            if constexpr (sizeof(VT) == 4) {
                const FusedVec V{B.e32, B.inv32, B.x32, (float *)B.z, B.r1, (float *)B.ap, B.a1, (float *)B.p0,
                                 (float *)B.p1};
                launch_fused(dim3(B.o.gd), s, B.o, 1, 1, 0, (const float *)B.val, V, (const float *)B.bv,
                             (const float *)B.bm, (const float *)B.ev);
            } else {
                hipLaunchKernelGGL((k_cg_dir<double, 1, false, TB>), dim3(B.o.gd), dim3(TB), 0, s, B.o, 1, 0, 0, 0,
                                   (const double *)B.val, (const double *)B.dg, (const double *)B.z,
                                   (double *)B.p0, (double *)B.p1, (double *)B.ap, (const double *)B.bv,
                                   (const double *)B.bm, (const double *)B.ev);
            }
        }
    };
    if (B.o.tb == 1024)
        launch(std::integral_constant<int, 1024>{});
    else if (B.o.tb == 512)
        launch(std::integral_constant<int, 512>{});
    else
        launch(std::integral_constant<int, 256>{});
}
}  // namespace
}  // namespace gf

extern "C" int64_t gf_pcg_dir_bench_bytes(const gf_mesh *m, int32_t fp32) {
    if (!m || m->n_cells < 1) return -1;
    return (int64_t)bench_layout(m, fp32, nullptr).bytes;
}

extern "C" int gf_pcg_dir_bench(const gf_mesh *m, int32_t phase, int32_t fp32, int32_t reps, const double *diag,
                                const double *upper, const double *lower, const double *x, double *y, void *scratch,
                                gf_stream_t stream) {
    GF_REQUIRE(m && scratch, "null argument");
    GF_REQUIRE(m->n_cells >= 1, "empty grid");
    GF_REQUIRE(phase >= 0 && phase <= 2, "phase must be 0 (prepare), 1 (launch) or 2 (result)");
    hipStream_t s = S(stream);
    const BenchLayout B = bench_layout(m, fp32, reinterpret_cast<char *>(scratch));
    if (phase == 0) {
        GF_REQUIRE(diag && upper && lower && x, "null argument");
        if (fp32)
            bench_pack<float>(m, B, diag, upper, lower, x, s);
        else
            bench_pack<double>(m, B, diag, upper, lower, x, s);
    } else if (phase == 1) {
        GF_REQUIRE(reps >= 1, "reps must be >= 1");
        if (fp32)  // the fused mixed-precision launch: FP32 values and vectors, FP64 accumulation
            bench_run<float>(B, reps, s);
        else
            bench_run<double>(B, reps, s);
    } else {
        GF_REQUIRE(y, "null argument");
        if (fp32)  // Ap of the launch: buffer 1 (launches read buffer 0)
            hipLaunchKernelGGL(k_bench_out<float>, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, s, m->n_cells,
                               (const float *)B.a1, B.perm, y);
        else
            hipLaunchKernelGGL(k_bench_out<double>, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, s, m->n_cells,
                               (const double *)B.ap, B.perm, y);
    }
    GF_CHECK_LAUNCH();
    return GF_OK;
}
