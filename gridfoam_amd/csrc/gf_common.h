// Internal helpers shared by the gfx950 kernels of libgridfoam_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/gridfoam_hip.h"

namespace gf {

constexpr int kWave = 64;     // CDNA wavefront width
constexpr int kBlock = 256;   // 4 waves per workgroup for streaming kernels

void set_error(const std::string &msg);

inline hipStream_t S(gf_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

#define GF_CHECK_HIP(expr)                                                      \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            ::gf::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return GF_EHIP;                                                     \
        }                                                                       \
    } while (0)

#define GF_CHECK_LAUNCH() GF_CHECK_HIP(hipGetLastError())

#define GF_REQUIRE(cond, msg)          \
    do {                               \
        if (!(cond)) {                 \
            ::gf::set_error(msg);      \
            return GF_EINVAL;          \
        }                              \
    } while (0)

// The boundary faces k of cell c (per-cell lists bnd_row / bnd_face; any
// boundary-face storage order).
#define GF_FOR_BND(m, c, k)                                                       \
    for (int k##_i = (m).bnd_row[c], k = 0;                                       \
         k##_i < (m).bnd_row[(c) + 1] && ((k = (m).bnd_face[k##_i]), true); ++k##_i)

// Coefficient of (ref_v - psi_P) in the boundary normal gradient
// (fv/boundary_ops.py:372-392): fraction / |d|; zero on pure Neumann faces
// and on snapped Dirichlet faces (an immersed side within the Gibou distance
// with fraction 1: its cell carries the prescribed value instead).
__device__ __forceinline__ double bnd_value_coef(const gf_mesh &m, int k, double fr) {
    if (fr == 0.0) return 0.0;
    if (m.bnd_snap && m.bnd_snap[k] && fr == 1.0) return 0.0;
    return fr / m.bnd_mag_d[k];
}

inline int grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return (int)(g > 0 ? g : 1);
}

// ---- wave64 / block reductions (DPP-backed shuffles, then LDS) ----------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        T o = __shfl_down(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Block-wide reduction of NQ values; result valid in thread 0. Fixed order
// (lane tree, then wave index), so the result is deterministic.
template <int NQ, int BLOCK>
__device__ __forceinline__ void block_reduce(double (&v)[NQ], const bool (&is_max)[NQ]) {
    __shared__ double lds[NQ][BLOCK / kWave];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = is_max[q] ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) lds[q][wid] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double a = lds[q][0];
            for (int k = 1; k < BLOCK / kWave; ++k)
                a = is_max[q] ? (lds[q][k] > a ? lds[q][k] : a) : a + lds[q][k];
            v[q] = a;
        }
    }
    __syncthreads();
}

// No "last workgroup done" reductions here: on the 8-XCD MI355X every
// device-scope fence writes back the issuing XCD's L2, which costs hundreds
// of microseconds per kernel.  Partials are reduced by the next launch.

// Deterministic reduction of partials[n] by one block (fixed strided order
// then fixed tree).  Result valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ double reduce_partials(const double *p, int n, bool is_max) {
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        double v = p[i];
        a = is_max ? (v > a ? v : a) : a + v;
    }
    double q[1] = {a};
    const bool m[1] = {is_max};
    block_reduce<1, BLOCK>(q, m);
    return q[0];
}

}  // namespace gf
