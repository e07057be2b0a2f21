// Krylov pressure/momentum solves on the sliced-ELL face-neighbour operator.
//
// Restates solvers/cg.py:75-174 (PCG), solvers/bicgstab.py:76-272
// (right-preconditioned BiCGSTAB with restarts), solvers/preconditioners.py:
// 76-96 (Jacobi) and solvers/base.py:140-162 (threshold max(atol, rtol*|r0|),
// strict <).  The host never synchronises per iteration: scalars live in a
// device state block, every kernel checks the per-component phase and turns
// into a no-op once its component has finished, and the host polls the state
// every `check_every` iterations while the next chunk is already queued.
//
// Data layout: vectors are component-major [nrhs][C] so one SpMV streams the
// matrix once for all components; matrix values are repacked once per solve
// into slice-of-64 ELL order (one slice = one wavefront, coalesced loads).
#include <algorithm>
#include <cmath>
#include <vector>

#include "gf_common.h"

namespace gf {
namespace {

enum Phase : int { PH_ACTIVE = 0, PH_RESTART = 1, PH_FIN_S = 2, PH_DONE = 3 };

struct KState {
    double thresh, norm0, final_res;
    double rz_old, alpha, beta;
    double rho_old, omega, rhat2;
    double scale;       // mixed: norm of the FP64 residual the inner solve sees
    double inner_thresh;
    int iter, phase, converged, restarts;
    int total_iter, outer_phase, outer_converged, sweeps;
};

struct Ops {
    // matrix
    const int32_t *sell_ptr, *sell_col;
    int C;
    // control
    int nrhs, norm_inf, max_iter, max_restart, precond;
    double atol, rtol;
};

constexpr int kMaxRhs = 3;
constexpr int kNQ = 4;  // partial quantities per rhs per kernel

__device__ __forceinline__ double nrm_part(double v, bool inf) { return inf ? fabs(v) : v * v; }
__device__ __forceinline__ double nrm_fin(double v, bool inf) { return inf ? v : sqrt(v); }

template <typename VT, int NR>
__device__ __forceinline__ void sell_row(const int32_t *__restrict__ sell_ptr,
                                         const int32_t *__restrict__ sell_col,
                                         const VT *__restrict__ val, VT dg, int i,
                                         const VT *__restrict__ x, int C, VT (&acc)[NR]) {
    const int s = i >> 6, lane = i & 63;
    const int base = sell_ptr[s];
    const int W = (sell_ptr[s + 1] - base) >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = dg * x[(size_t)r * C + i];
    for (int k = 0; k < W; ++k) {
        const int e = base + k * 64 + lane;
        const int col = __builtin_nontemporal_load(sell_col + e);
        const VT v = __builtin_nontemporal_load(val + e);
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += v * x[(size_t)r * C + col];
    }
}

// Reduce kNQ partial quantities per rhs and hand them to thread 0 of the
// last block.  partials layout: [NR][kNQ][gridDim.x].
template <int NR>
__device__ __forceinline__ bool finish_partials(double (&v)[NR][kNQ], const bool (&is_max)[kNQ],
                                                double *partials, unsigned int *counter,
                                                double (&tot)[NR][kNQ]) {
    double flat[NR * kNQ];
    bool fm[NR * kNQ];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < kNQ; ++q) {
            flat[r * kNQ + q] = v[r][q];
            fm[r * kNQ + q] = is_max[q];
        }
    block_reduce<NR * kNQ, kBlock>(flat, fm);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int q = 0; q < kNQ; ++q) partials[((size_t)r * kNQ + q) * nb + blockIdx.x] = flat[r * kNQ + q];
    }
    if (!last_block_done(counter)) return false;
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < kNQ; ++q) {
            double a = is_max[q] ? 0.0 : 0.0;
            for (int b = threadIdx.x; b < nb; b += kBlock) {
                double pv = partials[((size_t)r * kNQ + q) * nb + b];
                a = is_max[q] ? (pv > a ? pv : a) : a + pv;
            }
            flat[r * kNQ + q] = a;
        }
    block_reduce<NR * kNQ, kBlock>(flat, fm);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < kNQ; ++q) tot[r][q] = flat[r * kNQ + q];
    return true;
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
__global__ void k_diag_prep(int C, const double *__restrict__ diag, int precond, VT *__restrict__ dg,
                            VT *__restrict__ inv_d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    double d = diag[i];
    dg[i] = (VT)d;
    // preconditioners.py:83-90: 1/diag, or 1 where |diag| < eps(dtype)
    const double eps = sizeof(VT) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
    double inv = 1.0;
    if (precond) inv = fabs(d) < eps ? 1.0 : 1.0 / d;
    inv_d[i] = (VT)inv;
}

// [C][nrhs] (row-major cell field) <-> [nrhs][C]
__global__ void k_to_soa(int C, int nrhs, const double *__restrict__ in, double *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    for (int r = 0; r < nrhs; ++r) out[(size_t)r * C + i] = in[(size_t)i * nrhs + r];
}
__global__ void k_from_soa(int C, int nrhs, const double *__restrict__ in, double *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    for (int r = 0; r < nrhs; ++r) out[(size_t)i * nrhs + r] = in[(size_t)r * C + i];
}

// ---------------------------------------------------------------- CG (FP64 plain path)
// r = b - A x ; z = M^-1 r ; p = z ; rz = r.z ; |r0|
template <int NR>
__global__ __launch_bounds__(kBlock) void k_cg_init(Ops o, const double *val, const double *dg,
                                                    const double *inv_d, const double *b,
                                                    const double *x, double *r_, double *z_,
                                                    double *p_, KState *st, double *partials,
                                                    unsigned int *counter) {
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        double ax[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, x, C, ax);
        const double id = inv_d[i];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * C + i;
            double rr = b[k] - ax[r];
            double zz = id * rr;
            r_[k] = rr;
            z_[k] = zz;
            p_[k] = zz;
            v[r][0] = nrm_part(rr, o.norm_inf);
            v[r][1] = rr * zz;
        }
    }
    const bool mx[kNQ] = {(bool)o.norm_inf, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            KState &s = st[r];
            double n0 = nrm_fin(tot[r][0], o.norm_inf);
            s.norm0 = n0;
            s.final_res = n0;
            s.thresh = fmax(o.atol, o.rtol * n0);
            s.iter = 0;
            s.restarts = 0;
            s.rz_old = tot[r][1];
            s.converged = n0 < s.thresh;
            s.phase = s.converged ? PH_DONE : PH_ACTIVE;
        }
    }
}

// Ap = A p ; alpha = rz_old / (p . Ap)
template <typename VT, int NR>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(Ops o, const VT *val, const VT *dg, const VT *p_,
                                                    VT *ap_, KState *st, double *partials,
                                                    unsigned int *counter) {
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) any |= st[r].phase == PH_ACTIVE;
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        VT ap[NR];
        sell_row<VT, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, p_, C, ap);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * C + i;
            ap_[k] = ap[r];
            v[r][0] = (double)p_[k] * (double)ap[r];
        }
    }
    const bool mx[kNQ] = {false, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r)
            if (st[r].phase == PH_ACTIVE) st[r].alpha = st[r].rz_old / tot[r][0];
}

// x += alpha p ; r -= alpha Ap ; |r| check ; z = M^-1 r ; beta
template <typename VT, int NR>
__global__ __launch_bounds__(kBlock) void k_cg_update(Ops o, const VT *inv_d, VT *x_, VT *r_, VT *z_,
                                                      const VT *p_, const VT *ap_, KState *st,
                                                      double *partials, unsigned int *counter,
                                                      int inner) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        const VT id = inv_d[i];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            const VT a = (VT)st[r].alpha;
            VT xx = x_[k] + a * p_[k];
            VT rr = r_[k] - a * ap_[k];
            VT zz = id * rr;
            x_[k] = xx;
            r_[k] = rr;
            z_[k] = zz;
            v[r][0] = nrm_part((double)rr, o.norm_inf);
            v[r][1] = (double)rr * (double)zz;
        }
    }
    const bool mx[kNQ] = {(bool)o.norm_inf, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            KState &s = st[r];
            if (s.phase != PH_ACTIVE) continue;
            s.iter += 1;
            if (inner) s.total_iter += 1;
            double nr = nrm_fin(tot[r][0], o.norm_inf);
            const double th = inner ? s.inner_thresh : s.thresh;
            if (!inner) s.final_res = nr;
            if (nr < th) {
                if (!inner) s.converged = 1;
                s.phase = PH_DONE;
            } else if (s.iter >= o.max_iter) {
                s.phase = PH_DONE;
            } else {
                s.beta = tot[r][1] / s.rz_old;
                s.rz_old = tot[r][1];
            }
        }
    }
}

// p = z + beta p
template <typename VT, int NR>
__global__ __launch_bounds__(kBlock) void k_cg_pupdate(int C, const VT *z_, VT *p_, const KState *st) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!act[r]) continue;
        const size_t k = (size_t)r * C + i;
        p_[k] = z_[k] + p_[k] * (VT)st[r].beta;
    }
}

// ---------------------------------------------------------------- mixed precision
// Outer FP64 residual: r64 = b - A64 x64 ; |r64| ; decide outer convergence.
// On the first sweep it also fixes thresh from |r0| (base.py:140-162).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_mx_residual(Ops o, const double *val, const double *dg,
                                                        const double *b, const double *x,
                                                        double *r64, KState *st, double *partials,
                                                        unsigned int *counter, int first) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = first || (st[r].outer_phase == PH_ACTIVE && st[r].phase == PH_DONE);
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        double ax[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, x, C, ax);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            double rr = b[k] - ax[r];
            r64[k] = rr;
            v[r][0] = nrm_part(rr, o.norm_inf);
            v[r][1] = rr * rr;
        }
    }
    const bool mx[kNQ] = {(bool)o.norm_inf, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            KState &s = st[r];
            double nr = nrm_fin(tot[r][0], o.norm_inf);
            if (first) {
                s.norm0 = nr;
                s.thresh = fmax(o.atol, o.rtol * nr);
                s.total_iter = 0;
                s.sweeps = 0;
                s.converged = 0;
            }
            s.final_res = nr;
            if (nr < s.thresh) {
                s.converged = 1;
                s.outer_phase = PH_DONE;
                s.phase = PH_DONE;
            } else if (s.total_iter >= o.max_iter) {
                s.outer_phase = PH_DONE;
                s.phase = PH_DONE;
            } else {
                s.outer_phase = PH_ACTIVE;
                s.scale = sqrt(tot[r][1]);  // L2 scaling keeps FP32 in range
                s.phase = PH_RESTART;       // inner init pending
            }
        }
    }
}

// Inner FP32 PCG init from the scaled FP64 residual (x32 = 0).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_mx_inner_init(Ops o, const float *inv_d, const double *r64,
                                                          float *x_, float *r_, float *z_, float *p_,
                                                          KState *st, double *partials,
                                                          unsigned int *counter, double inner_rtol) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_RESTART;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        const float id = inv_d[i];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            const double inv_s = 1.0 / st[r].scale;
            float rr = (float)(r64[k] * inv_s);
            float zz = id * rr;
            x_[k] = 0.0f;
            r_[k] = rr;
            z_[k] = zz;
            p_[k] = zz;
            v[r][0] = (double)rr * (double)rr;
            v[r][1] = (double)rr * (double)zz;
        }
    }
    const bool mx[kNQ] = {false, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            KState &s = st[r];
            s.rz_old = tot[r][1];
            s.inner_thresh = inner_rtol * sqrt(tot[r][0]);
            s.iter = 0;
            s.sweeps += 1;
            s.phase = PH_ACTIVE;
        }
    }
}

// x64 += scale * x32 after an inner sweep finished.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_mx_correct(int C, const float *x32, double *x64,
                                                       const KState *st) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].outer_phase == PH_ACTIVE && st[r].phase == PH_DONE;
        any |= act[r];
    }
    if (!any) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!act[r]) continue;
        const size_t k = (size_t)r * C + i;
        x64[k] += st[r].scale * (double)x32[k];
    }
}

// ---------------------------------------------------------------- BiCGSTAB (FP64)
// restart: r_hat = r ; p = r ; rho_old = r.r   (bicgstab.py:248-260)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_restart(Ops o, const double *r_, double *rh_, double *p_,
                                                       KState *st, double *partials,
                                                       unsigned int *counter, int force) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = force ? st[r].phase == PH_ACTIVE : st[r].phase == PH_RESTART;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            double rr = r_[k];
            rh_[k] = rr;
            p_[k] = rr;
            v[r][0] = rr * rr;
        }
    }
    const bool mx[kNQ] = {false, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r)
            if (act[r]) {
                st[r].rho_old = tot[r][0];
                st[r].phase = PH_ACTIVE;
            }
}

// y = M^-1 p ; v = A y ; den = r_hat.v ; breakdown test ; alpha
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_v(Ops o, const double *val, const double *dg,
                                                 const double *inv_d, const double *y_,
                                                 const double *rh_, double *v_, KState *st,
                                                 double *partials, unsigned int *counter) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        double av[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, y_, C, av);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * C + i;
            v_[k] = av[r];
            const double h = rh_[k];
            v[r][0] = h * av[r];
            v[r][1] = h * h;
            v[r][2] = av[r] * av[r];
        }
    }
    const bool mx[kNQ] = {false, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            KState &s = st[r];
            if (s.phase != PH_ACTIVE) continue;
            s.iter += 1;
            s.rhat2 = tot[r][1];
            const double den = tot[r][0];
            if (fabs(den) <= 1e-6 * sqrt(tot[r][1]) * sqrt(tot[r][2])) {
                if (s.restarts >= o.max_restart || s.iter >= o.max_iter) {
                    s.phase = PH_DONE;
                } else {
                    s.restarts += 1;
                    s.phase = PH_RESTART;
                }
            } else {
                s.alpha = s.rho_old / den;
            }
        }
    }
}

// y = M^-1 p  (elementwise; also z = M^-1 s with the same kernel)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_precond(int C, const double *inv_d, const double *in,
                                                       double *out, const KState *st, int want_phase) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == want_phase;
        any |= act[r];
    }
    if (!any) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C) return;
    const double id = inv_d[i];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!act[r]) continue;
        const size_t k = (size_t)r * C + i;
        out[k] = id * in[k];
    }
}

// s = r - alpha v ; |s| ; converged via s -> PH_FIN_S
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_s(Ops o, const double *r_, const double *v_, double *s_,
                                                 KState *st, double *partials,
                                                 unsigned int *counter) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            double ss = r_[k] - st[r].alpha * v_[k];
            s_[k] = ss;
            v[r][0] = nrm_part(ss, o.norm_inf);
        }
    }
    const bool mx[kNQ] = {(bool)o.norm_inf, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            KState &s = st[r];
            double ns = nrm_fin(tot[r][0], o.norm_inf);
            s.final_res = ns;
            if (ns < s.thresh) {
                s.converged = 1;
                s.phase = PH_FIN_S;
            }
        }
    }
}

// converged via s: x += alpha y, then finish
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_fin_s(int C, const double *y_, double *x_, KState *st,
                                                     unsigned int *counter) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_FIN_S;
        any |= act[r];
    }
    if (!any) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < C) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            x_[k] = x_[k] + st[r].alpha * y_[k];
        }
    }
    if (!last_block_done(counter)) return;
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; ++r)
            if (act[r]) st[r].phase = PH_DONE;
}

// t = A z ; omega = t.s / t.t with both breakdown tests
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_t(Ops o, const double *val, const double *dg,
                                                 const double *z_, const double *s_, double *t_,
                                                 KState *st, double *partials,
                                                 unsigned int *counter) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        double at[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, z_, C, at);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * C + i;
            t_[k] = at[r];
            v[r][0] = at[r] * at[r];
            v[r][1] = at[r] * s_[k];
        }
    }
    const bool mx[kNQ] = {false, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            KState &s = st[r];
            if (s.phase != PH_ACTIVE) continue;
            const double tt = tot[r][0];
            bool brk = fabs(tt) <= 1e-6 * sqrt(tt) * sqrt(tt);
            double om = 0.0;
            if (!brk) {
                om = tot[r][1] / tt;
                brk = fabs(om) <= 1e-6;
            }
            if (brk) {
                if (s.restarts >= o.max_restart || s.iter >= o.max_iter) {
                    s.phase = PH_DONE;
                } else {
                    s.restarts += 1;
                    s.phase = PH_RESTART;
                }
            } else {
                s.omega = om;
            }
        }
    }
}

// x += alpha y + omega z ; r = s - omega t ; |r| ; rho_new ; beta
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_update(Ops o, const double *y_, const double *z_,
                                                      const double *s_, const double *t_,
                                                      const double *rh_, double *x_, double *r_,
                                                      KState *st, double *partials,
                                                      unsigned int *counter) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            const size_t k = (size_t)r * C + i;
            const double a = st[r].alpha, w = st[r].omega;
            double xx = x_[k] + a * y_[k] + w * z_[k];
            double rr = s_[k] - w * t_[k];
            x_[k] = xx;
            r_[k] = rr;
            v[r][0] = nrm_part(rr, o.norm_inf);
            v[r][1] = rh_[k] * rr;
            v[r][2] = rr * rr;
        }
    }
    const bool mx[kNQ] = {(bool)o.norm_inf, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            if (!act[r]) continue;
            KState &s = st[r];
            double nr = nrm_fin(tot[r][0], o.norm_inf);
            s.final_res = nr;
            if (nr < s.thresh) {
                s.converged = 1;
                s.phase = PH_DONE;
                continue;
            }
            const double rho_new = tot[r][1];
            if (fabs(rho_new) <= 1e-6 * sqrt(s.rhat2) * sqrt(tot[r][2])) {
                if (s.restarts >= o.max_restart || s.iter >= o.max_iter) {
                    s.phase = PH_DONE;
                } else {
                    s.restarts += 1;
                    s.phase = PH_RESTART;
                }
                continue;
            }
            s.beta = (rho_new / s.rho_old) * (s.alpha / s.omega);
            s.rho_old = rho_new;
            if (s.iter >= o.max_iter) s.phase = PH_DONE;
        }
    }
}

// p = r + beta (p - omega v)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_p(int C, const double *r_, const double *v_, double *p_,
                                                 const KState *st) {
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        act[r] = st[r].phase == PH_ACTIVE;
        any |= act[r];
    }
    if (!any) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!act[r]) continue;
        const size_t k = (size_t)r * C + i;
        p_[k] = r_[k] + st[r].beta * (p_[k] - st[r].omega * v_[k]);
    }
}

// bicgstab init: r = b - A x ; |r0| ; thresh
template <int NR>
__global__ __launch_bounds__(kBlock) void k_bi_init(Ops o, const double *val, const double *dg,
                                                    const double *b, const double *x, double *r_,
                                                    KState *st, double *partials,
                                                    unsigned int *counter) {
    const int C = o.C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[NR][kNQ] = {};
    if (i < C) {
        double ax[NR];
        sell_row<double, NR>(o.sell_ptr, o.sell_col, val, dg[i], i, x, C, ax);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const size_t k = (size_t)r * C + i;
            double rr = b[k] - ax[r];
            r_[k] = rr;
            v[r][0] = nrm_part(rr, o.norm_inf);
        }
    }
    const bool mx[kNQ] = {(bool)o.norm_inf, false, false, false};
    double tot[NR][kNQ];
    if (!finish_partials<NR>(v, mx, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            KState &s = st[r];
            double n0 = nrm_fin(tot[r][0], o.norm_inf);
            s.norm0 = n0;
            s.final_res = n0;
            s.thresh = fmax(o.atol, o.rtol * n0);
            s.iter = 0;
            s.restarts = 0;
            s.converged = n0 < s.thresh;
            s.phase = s.converged ? PH_DONE : PH_ACTIVE;
        }
    }
}

// ---------------------------------------------------------------- plain SpMV entry
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_sell_spmv(const int32_t *sell_ptr, const int32_t *sell_col,
                                                      int C, const VT *val, const VT *dg, const VT *x,
                                                      VT *y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C) return;
    VT acc[1];
    sell_row<VT, 1>(sell_ptr, sell_col, val, dg[i], i, x, C, acc);
    y[i] = acc[0];
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
    double *val64, *dg64, *inv64, *b, *x, *r, *z, *p, *ap, *y, *s, *t, *rh, *v;
    float *val32, *dg32, *inv32, *x32, *r32, *z32, *p32, *ap32;
    double *partials;
    KState *st;
    unsigned int *counters;
    size_t bytes;
};

Layout layout(char *base, int64_t C, int64_t nnz, int nrhs, const gf_krylov_cfg &cfg) {
    WS w{base, 0};
    Layout L{};
    const size_t n = (size_t)C * nrhs;
    const int nb = grid_for(C);
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
        L.x32 = w.take<float>(n);
        L.r32 = w.take<float>(n);
        L.z32 = w.take<float>(n);
        L.p32 = w.take<float>(n);
        L.ap32 = w.take<float>(n);
    } else {
        L.z = w.take<double>(n);
        L.p = w.take<double>(n);
        L.ap = w.take<double>(n);
        if (cfg.method == 1) {
            L.y = w.take<double>(n);
            L.s = w.take<double>(n);
            L.t = w.take<double>(n);
            L.rh = w.take<double>(n);
            L.v = L.ap;
        }
    }
    L.partials = w.take<double>((size_t)kMaxRhs * kNQ * nb);
    L.st = w.take<KState>(kMaxRhs);
    L.counters = w.take<unsigned int>(32);
    L.bytes = w.off + 256;
    return L;
}

// Pinned staging for the polled device state (one small buffer per thread).
KState *host_stage() {
    static thread_local KState *buf = nullptr;
    if (!buf) {
        if (hipHostMalloc((void **)&buf, sizeof(KState) * kMaxRhs * 2, hipHostMallocDefault) != hipSuccess)
            buf = nullptr;
    }
    return buf;
}

template <int NR>
int run_cg(const Ops &o, const Layout &L, const gf_krylov_cfg &cfg, hipStream_t s, double *stats) {
    const int C = o.C, nb = grid_for(C);
    unsigned int *cnt = L.counters;
    KState *hst = host_stage();
    GF_REQUIRE(hst != nullptr, "pinned staging allocation failed");
    hipEvent_t ev[2];
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const int every = cfg.check_every > 0 ? cfg.check_every : 8;
    int rc = GF_OK;
    if (!cfg.mixed) {
        hipLaunchKernelGGL(k_cg_init<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.val64, L.dg64, L.inv64, L.b,
                           L.x, L.r, L.z, L.p, L.st, L.partials, cnt + 0);
    } else {
        hipLaunchKernelGGL(k_mx_residual<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.val64, L.dg64, L.b, L.x,
                           L.r, L.st, L.partials, cnt + 0, 1);
        hipLaunchKernelGGL(k_mx_inner_init<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.inv32, L.r, L.x32,
                           L.r32, L.z32, L.p32, L.st, L.partials, cnt + 1, cfg.inner_rtol);
    }
    GF_CHECK_LAUNCH();
    int launched = 0, chunk = 0;
    const int cap = cfg.max_iter + 8;  // each chunk is `every` iterations
    for (;;) {
        for (int k = 0; k < every; ++k) {
            if (!cfg.mixed) {
                hipLaunchKernelGGL((k_cg_spmv<double, NR>), dim3(nb), dim3(kBlock), 0, s, o, L.val64,
                                   L.dg64, L.p, L.ap, L.st, L.partials, cnt + 2);
                hipLaunchKernelGGL((k_cg_update<double, NR>), dim3(nb), dim3(kBlock), 0, s, o, L.inv64,
                                   L.x, L.r, L.z, L.p, L.ap, L.st, L.partials, cnt + 3, 0);
                hipLaunchKernelGGL((k_cg_pupdate<double, NR>), dim3(nb), dim3(kBlock), 0, s, C, L.z, L.p,
                                   L.st);
            } else {
                hipLaunchKernelGGL((k_cg_spmv<float, NR>), dim3(nb), dim3(kBlock), 0, s, o, L.val32,
                                   L.dg32, L.p32, L.ap32, L.st, L.partials, cnt + 2);
                hipLaunchKernelGGL((k_cg_update<float, NR>), dim3(nb), dim3(kBlock), 0, s, o, L.inv32,
                                   L.x32, L.r32, L.z32, L.p32, L.ap32, L.st, L.partials, cnt + 3, 1);
                hipLaunchKernelGGL((k_cg_pupdate<float, NR>), dim3(nb), dim3(kBlock), 0, s, C, L.z32,
                                   L.p32, L.st);
            }
        }
        if (cfg.mixed) {
            hipLaunchKernelGGL(k_mx_correct<NR>, dim3(nb), dim3(kBlock), 0, s, C, L.x32, L.x, L.st);
            hipLaunchKernelGGL(k_mx_residual<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.val64, L.dg64, L.b,
                               L.x, L.r, L.st, L.partials, cnt + 0, 0);
            hipLaunchKernelGGL(k_mx_inner_init<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.inv32, L.r,
                               L.x32, L.r32, L.z32, L.p32, L.st, L.partials, cnt + 1, cfg.inner_rtol);
        }
        GF_CHECK_LAUNCH();
        launched += every;
        KState *slot = hst + (chunk & 1) * kMaxRhs;
        GF_CHECK_HIP(hipMemcpyAsync(slot, L.st, sizeof(KState) * NR, hipMemcpyDeviceToHost, s));
        GF_CHECK_HIP(hipEventRecord(ev[chunk & 1], s));
        if (chunk > 0) {
            GF_CHECK_HIP(hipEventSynchronize(ev[(chunk - 1) & 1]));
            KState *prev = hst + ((chunk - 1) & 1) * kMaxRhs;
            bool all = true;
            for (int r = 0; r < NR; ++r)
                all &= cfg.mixed ? prev[r].outer_phase == PH_DONE : prev[r].phase == PH_DONE;
            if (all) break;
        }
        ++chunk;
        if (!cfg.mixed && launched > cap + 2 * every) {
            set_error("cg: state machine did not finish within max_iter");
            rc = GF_EINVAL;
            break;
        }
        if (cfg.mixed && launched > 64 * (cap + every)) {
            set_error("mixed cg: refinement did not finish");
            rc = GF_EINVAL;
            break;
        }
        if (chunk > 2 && every < 64) {
            // grow the chunk once the solve is clearly long
        }
    }
    GF_CHECK_HIP(hipStreamSynchronize(s));
    GF_CHECK_HIP(hipMemcpy(hst, L.st, sizeof(KState) * NR, hipMemcpyDeviceToHost));
    for (int r = 0; r < NR; ++r) {
        stats[r * 4 + 0] = hst[r].norm0;
        stats[r * 4 + 1] = hst[r].final_res;
        stats[r * 4 + 2] = cfg.mixed ? hst[r].total_iter : hst[r].iter;
        stats[r * 4 + 3] = hst[r].converged;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    return rc;
}

template <int NR>
int run_bicgstab(const Ops &o, const Layout &L, const gf_krylov_cfg &cfg, hipStream_t s, double *stats) {
    const int C = o.C, nb = grid_for(C);
    unsigned int *cnt = L.counters;
    KState *hst = host_stage();
    GF_REQUIRE(hst != nullptr, "pinned staging allocation failed");
    hipEvent_t ev[2];
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const int every = cfg.check_every > 0 ? cfg.check_every : 4;
    hipLaunchKernelGGL(k_bi_init<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.val64, L.dg64, L.b, L.x, L.r,
                       L.st, L.partials, cnt + 0);
    hipLaunchKernelGGL(k_bi_restart<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.r, L.rh, L.p, L.st,
                       L.partials, cnt + 1, 1);
    GF_CHECK_LAUNCH();
    int launched = 0, chunk = 0, rc = GF_OK;
    for (;;) {
        for (int k = 0; k < every; ++k) {
            hipLaunchKernelGGL(k_bi_restart<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.r, L.rh, L.p, L.st,
                               L.partials, cnt + 1, 0);
            hipLaunchKernelGGL(k_bi_precond<NR>, dim3(nb), dim3(kBlock), 0, s, C, L.inv64, L.p, L.y, L.st,
                               (int)PH_ACTIVE);
            hipLaunchKernelGGL(k_bi_v<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.val64, L.dg64, L.inv64, L.y,
                               L.rh, L.v, L.st, L.partials, cnt + 2);
            hipLaunchKernelGGL(k_bi_s<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.r, L.v, L.s, L.st,
                               L.partials, cnt + 3);
            hipLaunchKernelGGL(k_bi_fin_s<NR>, dim3(nb), dim3(kBlock), 0, s, C, L.y, L.x, L.st, cnt + 4);
            hipLaunchKernelGGL(k_bi_precond<NR>, dim3(nb), dim3(kBlock), 0, s, C, L.inv64, L.s, L.z, L.st,
                               (int)PH_ACTIVE);
            hipLaunchKernelGGL(k_bi_t<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.val64, L.dg64, L.z, L.s, L.t,
                               L.st, L.partials, cnt + 5);
            hipLaunchKernelGGL(k_bi_update<NR>, dim3(nb), dim3(kBlock), 0, s, o, L.y, L.z, L.s, L.t, L.rh,
                               L.x, L.r, L.st, L.partials, cnt + 6);
            hipLaunchKernelGGL(k_bi_p<NR>, dim3(nb), dim3(kBlock), 0, s, C, L.r, L.v, L.p, L.st);
        }
        GF_CHECK_LAUNCH();
        launched += every;
        KState *slot = hst + (chunk & 1) * kMaxRhs;
        GF_CHECK_HIP(hipMemcpyAsync(slot, L.st, sizeof(KState) * NR, hipMemcpyDeviceToHost, s));
        GF_CHECK_HIP(hipEventRecord(ev[chunk & 1], s));
        if (chunk > 0) {
            GF_CHECK_HIP(hipEventSynchronize(ev[(chunk - 1) & 1]));
            KState *prev = hst + ((chunk - 1) & 1) * kMaxRhs;
            bool all = true;
            for (int r = 0; r < NR; ++r) all &= prev[r].phase == PH_DONE;
            if (all) break;
        }
        ++chunk;
        if (launched > cfg.max_iter + 4 * every + 8) {
            set_error("bicgstab: state machine did not finish within max_iter");
            rc = GF_EINVAL;
            break;
        }
    }
    GF_CHECK_HIP(hipStreamSynchronize(s));
    GF_CHECK_HIP(hipMemcpy(hst, L.st, sizeof(KState) * NR, hipMemcpyDeviceToHost));
    for (int r = 0; r < NR; ++r) {
        stats[r * 4 + 0] = hst[r].norm0;
        stats[r * 4 + 1] = hst[r].final_res;
        stats[r * 4 + 2] = hst[r].iter;
        stats[r * 4 + 3] = hst[r].converged;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    return rc;
}

}  // namespace
}  // namespace gf

using namespace gf;

extern "C" int64_t gf_krylov_workspace_bytes(const gf_mesh *m, int32_t nrhs, const gf_krylov_cfg *cfg) {
    if (!m || !cfg || nrhs < 1 || nrhs > kMaxRhs) return -1;
    Layout L = layout(nullptr, m->n_cells, m->nnz_pad, nrhs, *cfg);
    return (int64_t)L.bytes;
}

extern "C" int gf_krylov_solve(const gf_mesh *m, const gf_krylov_cfg *cfg, int32_t nrhs, int32_t transpose,
                               const double *diag, const double *upper, const double *lower,
                               const double *b, double *x, void *workspace, double *stats_out,
                               gf_stream_t stream) {
    GF_REQUIRE(m && cfg && diag && upper && lower && b && x && workspace && stats_out, "null argument");
    GF_REQUIRE(nrhs == 1 || nrhs == 3, "nrhs must be 1 or 3");
    GF_REQUIRE(cfg->method == 0 || cfg->method == 1, "unknown Krylov method");
    GF_REQUIRE(!(cfg->method == 1 && cfg->mixed), "mixed precision is implemented for cg only");
    GF_REQUIRE(cfg->max_iter >= 1, "max_iter must be >= 1");
    hipStream_t s = S(stream);
    const int C = m->n_cells;
    Layout L = layout(reinterpret_cast<char *>(workspace), C, m->nnz_pad, nrhs, *cfg);
    const int64_t nnz = m->nnz_pad;
    hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, m->sell_code, upper,
                       lower, transpose, L.val64);
    hipLaunchKernelGGL(k_diag_prep<double>, dim3(grid_for(C)), dim3(kBlock), 0, s, C, diag, cfg->precond,
                       L.dg64, L.inv64);
    if (cfg->method == 0 && cfg->mixed) {
        hipLaunchKernelGGL(k_sell_pack<float>, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, m->sell_code,
                           upper, lower, transpose, L.val32);
        hipLaunchKernelGGL(k_diag_prep<float>, dim3(grid_for(C)), dim3(kBlock), 0, s, C, diag, cfg->precond,
                           L.dg32, L.inv32);
    }
    GF_CHECK_HIP(hipMemsetAsync(L.counters, 0, 32 * sizeof(unsigned int), s));
    GF_CHECK_HIP(hipMemsetAsync(L.st, 0, sizeof(KState) * kMaxRhs, s));
    if (nrhs == 1) {
        GF_CHECK_HIP(hipMemcpyAsync(L.b, b, sizeof(double) * C, hipMemcpyDeviceToDevice, s));
        GF_CHECK_HIP(hipMemcpyAsync(L.x, x, sizeof(double) * C, hipMemcpyDeviceToDevice, s));
    } else {
        hipLaunchKernelGGL(k_to_soa, dim3(grid_for(C)), dim3(kBlock), 0, s, C, nrhs, b, L.b);
        hipLaunchKernelGGL(k_to_soa, dim3(grid_for(C)), dim3(kBlock), 0, s, C, nrhs, x, L.x);
    }
    GF_CHECK_LAUNCH();
    Ops o{m->sell_ptr, m->sell_col, C, nrhs, cfg->norm, cfg->max_iter, cfg->max_restart, cfg->precond,
          cfg->atol, cfg->rtol};
    int rc;
    if (cfg->method == 0)
        rc = nrhs == 1 ? run_cg<1>(o, L, *cfg, s, stats_out) : run_cg<3>(o, L, *cfg, s, stats_out);
    else
        rc = nrhs == 1 ? run_bicgstab<1>(o, L, *cfg, s, stats_out) : run_bicgstab<3>(o, L, *cfg, s, stats_out);
    if (rc != GF_OK) return rc;
    if (nrhs == 1) {
        GF_CHECK_HIP(hipMemcpyAsync(x, L.x, sizeof(double) * C, hipMemcpyDeviceToDevice, s));
    } else {
        hipLaunchKernelGGL(k_from_soa, dim3(grid_for(C)), dim3(kBlock), 0, s, C, nrhs, L.x, x);
        GF_CHECK_LAUNCH();
    }
    return GF_OK;
}

extern "C" int gf_sell_pack(const gf_mesh *m, int32_t transpose, const double *upper, const double *lower,
                            void *val, int32_t fp32, gf_stream_t stream) {
    GF_REQUIRE(m && upper && lower && val, "null argument");
    const int64_t nnz = m->nnz_pad;
    if (fp32)
        hipLaunchKernelGGL(k_sell_pack<float>, dim3(grid_for(nnz)), dim3(kBlock), 0, S(stream), nnz, m->sell_code,
                           upper, lower, transpose, (float *)val);
    else
        hipLaunchKernelGGL(k_sell_pack<double>, dim3(grid_for(nnz)), dim3(kBlock), 0, S(stream), nnz,
                           m->sell_code, upper, lower, transpose, (double *)val);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_sell_spmv(const gf_mesh *m, const void *diag, const void *val, const void *x, void *y,
                            int32_t fp32, gf_stream_t stream) {
    GF_REQUIRE(m && diag && val && x && y, "null argument");
    const int C = m->n_cells;
    if (fp32)
        hipLaunchKernelGGL(k_sell_spmv<float>, dim3(grid_for(C)), dim3(kBlock), 0, S(stream), m->sell_ptr,
                           m->sell_col, C, (const float *)val, (const float *)diag, (const float *)x, (float *)y);
    else
        hipLaunchKernelGGL(k_sell_spmv<double>, dim3(grid_for(C)), dim3(kBlock), 0, S(stream), m->sell_ptr,
                           m->sell_col, C, (const double *)val, (const double *)diag, (const double *)x,
                           (double *)y);
    GF_CHECK_LAUNCH();
    return GF_OK;
}
