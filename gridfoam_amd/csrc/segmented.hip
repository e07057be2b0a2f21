// Segmented Jacobi-PCG: many independent pressure systems in one solve.
//
// The batched solve server concatenates several cases into one mesh whose
// operator is block diagonal (no face couples two cases).  The rows of case
// s are the contiguous segment [seg_row[s], seg_row[s+1]).  Every launch
// covers all rows of all cases, but each segment keeps its own CG scalars
// (alpha, beta, r.z), its own threshold max(atol, rtol |r0|) and its own
// stop, so each case follows exactly the iteration of the single-case solver
// (solvers/cg.py:75-174 with the Jacobi preconditioner of
// solvers/preconditioners.py:76-96 and the strict threshold of
// solvers/base.py:140-162); only the summation order of its dot products
// differs.  A finished segment turns into no-ops in every later launch.
//
// Work split: segment s gets W_s = min(ceil(rows / 1024), 256) workgroups of
// 1024 threads that stride over its rows (a table gives every workgroup its
// segment, its segment's first workgroup, W_s and row range), so a workgroup
// belongs to one segment and a segment's dot product has at most 256
// partials, one per thread of the reducing workgroup.  Few, wide workgroups
// keep the re-reduction of the partials in every workgroup cheap
// (sum_s W_s^2 loads per launch) while a segment still fills the chip.  One
// iteration is two launches, with no fence inside a kernel:
//  * direction kernel: every workgroup of s first reduces the partials |r|^2
//    and r.z of the previous update (fixed order, so all of them derive the
//    same scalars: iteration count, convergence, beta), then builds
//    p = z + beta p_old on the fly for its rows and their gathered columns,
//    y = A p, and writes its partial p.y;
//  * update kernel: every workgroup of s reduces the p.y partials to alpha,
//    updates x, r, z on its rows and writes its partials |r|^2, r.z.
// Segment state is double buffered between the two kernels (the direction
// kernel reads buffer 0 and its tile-0 workgroup writes buffer 1; the
// update kernel the other way round), so no workgroup reads what another
// workgroup of the same launch writes.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "gf_common.h"

namespace gf {
namespace {

constexpr int TS = 1024;  // threads per workgroup (16 waves: four per SIMD)
constexpr int WMAX = 256;  // workgroups per segment at most (one partial per thread in a reduction)
static_assert(WMAX <= TS, "a reducing workgroup loads one partial per thread");

struct SegState {
    double thresh, norm0, final_res, rz_old, alpha, beta;
    int iter, active, converged, fresh;
};

struct SegOps {
    int C;
    const int32_t *sell_ptr, *sell_col;
    const double *val, *dg, *inv_d;
    const int32_t *seg_row;  // [n_seg+1]
    const int4 *wg;          // [G] per workgroup: segment, first workgroup of it, its workgroup count, first row
    const int32_t *wg_end;   // [G] end row of the workgroup's segment
    double *x, *r, *z, *y;
    int G;                   // workgroups of the launch (the partials' stride)
    double *pd;              // [G] partial p.y
    double *pu;              // [2][G] partials |r|^2, r.z
    SegState *st0, *st1;     // [n_seg] each
    int32_t *active;         // [n_seg]
    double atol, rtol;
    int max_iter;
};

// Workgroup sum of NQ values, returned to every thread: wave sums by
// shuffles, then the four wave sums in wave order (a fixed order).
template <int NQ>
__device__ __forceinline__ void wg_sum(double (&v)[NQ]) {
    __shared__ double sh[NQ][TS / kWave];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double x = wave_sum(v[q]);
        if (lane == 0) sh[q][wid] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double a = sh[q][0];
#pragma unroll
        for (int k = 1; k < TS / kWave; ++k) a += sh[q][k];
        v[q] = a;
    }
    __syncthreads();
}

// The partials of a segment's W <= WMAX workgroups (one per thread),
// loaded at kernel start, before anything that depends on the state.
template <int NQ>
__device__ __forceinline__ void load_partials(const double *__restrict__ part, int stride, int w0, int W,
                                              double (&v)[NQ]) {
    const int j = threadIdx.x;
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = j < W ? part[(size_t)q * stride + w0 + j] : 0.0;
}

__device__ __forceinline__ double ell_row(const SegOps &o, int i, const double *__restrict__ v) {
    const int sl = i >> 6, lane = i & 63;
    const int base = o.sell_ptr[sl];
    const int W = (o.sell_ptr[sl + 1] - base) >> 6;
    double acc = 0.0;
    for (int k = 0; k < W; ++k) {
        const int e = base + k * 64 + lane;
        acc += o.val[e] * v[o.sell_col[e]];
    }
    return acc;
}

__global__ void k_seg_pack(int64_t n, const int32_t *__restrict__ code, const double *__restrict__ upper,
                           const double *__restrict__ lower, double *__restrict__ val) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int c = code[e];
    double v = 0.0;
    if (c != INT32_MIN) v = c >= 0 ? upper[c] : lower[-c - 1];
    val[e] = v;
}

__global__ void k_seg_diag(int C, const double *__restrict__ diag, int precond, double *__restrict__ dg,
                           double *__restrict__ inv_d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double d = diag[i];
    dg[i] = d;
    inv_d[i] = precond ? (fabs(d) < 2.220446049250313e-16 ? 1.0 : 1.0 / d) : 1.0;
}

// r = b - A x ; z = M^-1 r ; partials |r|^2, r.z
__global__ __launch_bounds__(TS) void k_seg_init(SegOps o, const double *__restrict__ b) {
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int t = w - d.y, W = d.z, r1 = o.wg_end[w];
    double v[2] = {0.0, 0.0};
    for (int i = d.w + t * TS + (int)threadIdx.x; i < r1; i += W * TS) {
        const double rr = b[i] - (o.dg[i] * o.x[i] + ell_row(o, i, o.x));
        const double zz = o.inv_d[i] * rr;
        o.r[i] = rr;
        o.z[i] = zz;
        v[0] += rr * rr;
        v[1] += rr * zz;
    }
    wg_sum<2>(v);
    if (threadIdx.x == 0) {
        o.pu[w] = v[0];
        o.pu[(size_t)o.G + w] = v[1];
    }
}

// phase 0: decide on |r0|; 1: iteration test and beta; 2: the test only.
// Then (phases 0, 1) p = z + beta p_old, y = A p, partial p.y.  The loads of
// a thread's first row, its columns and their z, p_old do not depend on the
// segment state; they are issued before the state and the partials are read.
constexpr int KPRE = 6;  // ELL entries prefetched per row (a uniform 3-D stencil has six)
__global__ __launch_bounds__(TS) void k_seg_dir(SegOps o, int phase, const double *__restrict__ p_old,
                                                double *__restrict__ p_new) {
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int s = d.x, w0 = d.y, t = w - w0, W = d.z, r1 = o.wg_end[w];
    const int i0 = d.w + t * TS + (int)threadIdx.x;
    const bool has0 = phase != 2 && i0 < r1;
    double q[2];
    load_partials<2>(o.pu, o.G, w0, W, q);
    double z0 = 0.0, po0 = 0.0, dg0 = 0.0, hv[KPRE], hz[KPRE], hp[KPRE];
    int Wl0 = 0, base0 = 0;
    if (has0) {
        const int sl = i0 >> 6, lane = i0 & 63;
        base0 = o.sell_ptr[sl] + lane;
        Wl0 = (o.sell_ptr[sl + 1] - o.sell_ptr[sl]) >> 6;
        z0 = o.z[i0];
        po0 = p_old[i0];
        dg0 = o.dg[i0];
#pragma unroll
        for (int k = 0; k < KPRE; ++k) {
            hv[k] = 0.0;
            hz[k] = 0.0;
            hp[k] = 0.0;
            if (k < Wl0) {
                const int e = base0 + k * 64;
                const int c = o.sell_col[e];
                hv[k] = o.val[e];
                hz[k] = o.z[c];
                hp[k] = p_old[c];
            }
        }
    }
    SegState S = o.st0[s];
    if (phase == 0 || S.active) {
        wg_sum<2>(q);
        if (phase == 0) {
            S.norm0 = sqrt(q[0]);
            S.final_res = S.norm0;
            S.thresh = fmax(o.atol, o.rtol * S.norm0);
            S.rz_old = q[1];
            S.alpha = S.beta = 0.0;
            S.iter = 0;
            S.converged = S.norm0 < S.thresh;
            S.active = !S.converged;
            S.fresh = 1;
        } else {
            const double nr = sqrt(q[0]);
            S.iter += 1;
            S.final_res = nr;
            if (nr < S.thresh) {
                S.converged = 1;
                S.active = 0;
            } else if (S.iter >= o.max_iter) {
                S.active = 0;
            } else {
                S.beta = q[1] / S.rz_old;
                S.rz_old = q[1];
                S.fresh = 0;
            }
        }
    }
    if (t == 0 && threadIdx.x == 0) {
        o.st1[s] = S;
        o.active[s] = S.active;
    }
    if (phase == 2 || !S.active) return;
    const bool fresh = S.fresh != 0;
    const double be = S.beta;
    double v[1] = {0.0};
    if (has0) {  // the prefetched first row
        const double pi = fresh ? z0 : z0 + be * po0;
        double acc = dg0 * pi;
#pragma unroll
        for (int k = 0; k < KPRE; ++k)
            if (k < Wl0) acc += hv[k] * (fresh ? hz[k] : hz[k] + be * hp[k]);
        for (int k = KPRE; k < Wl0; ++k) {
            const int e = base0 + k * 64;
            const int c = o.sell_col[e];
            acc += o.val[e] * (fresh ? o.z[c] : o.z[c] + be * p_old[c]);
        }
        p_new[i0] = pi;
        o.y[i0] = acc;
        v[0] += pi * acc;
    }
    for (int i = i0 + W * TS; i < r1; i += W * TS) {
        const double pi = fresh ? o.z[i] : o.z[i] + be * p_old[i];
        const int sl = i >> 6, lane = i & 63;
        const int base = o.sell_ptr[sl];
        const int Wl = (o.sell_ptr[sl + 1] - base) >> 6;
        double acc = o.dg[i] * pi;
        for (int k = 0; k < Wl; ++k) {
            const int e = base + k * 64 + lane;
            const int c = o.sell_col[e];
            const double pc = fresh ? o.z[c] : o.z[c] + be * p_old[c];
            acc += o.val[e] * pc;
        }
        p_new[i] = pi;
        o.y[i] = acc;
        v[0] += pi * acc;
    }
    wg_sum<1>(v);
    if (threadIdx.x == 0) o.pd[w] = v[0];
}

// alpha = r.z / p.y ; x += alpha p ; r -= alpha y ; z = M^-1 r ; partials
// |r|^2, r.z (the first row's operands are loaded before the state)
__global__ __launch_bounds__(TS) void k_seg_upd(SegOps o, const double *__restrict__ p) {
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int s = d.x, w0 = d.y, t = w - w0, W = d.z, r1 = o.wg_end[w];
    const int i0 = d.w + t * TS + (int)threadIdx.x;
    const bool has0 = i0 < r1;
    double q[1];
    load_partials<1>(o.pd, o.G, w0, W, q);
    double x0 = 0.0, r0 = 0.0, y0 = 0.0, pp0 = 0.0, id0 = 0.0;
    if (has0) {
        x0 = o.x[i0];
        r0 = o.r[i0];
        y0 = o.y[i0];
        pp0 = p[i0];
        id0 = o.inv_d[i0];
    }
    // the state is read field by field: a local copy of the whole record
    // would be placed in LDS (one per thread) by the compiler
    const SegState *in = o.st1 + s;
    if (!in->active) {
        if (t == 0 && threadIdx.x == 0) o.st0[s] = *in;
        return;
    }
    const double rz = in->rz_old;
    wg_sum<1>(q);
    const double a = rz / q[0];
    double v[2] = {0.0, 0.0};
    if (has0) {
        o.x[i0] = x0 + a * pp0;
        const double rr = r0 - a * y0;
        const double zz = id0 * rr;
        o.r[i0] = rr;
        o.z[i0] = zz;
        v[0] += rr * rr;
        v[1] += rr * zz;
    }
    for (int i = i0 + W * TS; i < r1; i += W * TS) {
        o.x[i] = o.x[i] + a * p[i];
        const double rr = o.r[i] - a * o.y[i];
        const double zz = o.inv_d[i] * rr;
        o.r[i] = rr;
        o.z[i] = zz;
        v[0] += rr * rr;
        v[1] += rr * zz;
    }
    wg_sum<2>(v);
    if (threadIdx.x == 0) {
        o.pu[w] = v[0];
        o.pu[(size_t)o.G + w] = v[1];
        if (t == 0) {
            o.st0[s] = *in;
            o.st0[s].alpha = a;
        }
    }
}

__global__ void k_seg_stats(int n_seg, const SegState *__restrict__ st, int32_t *__restrict__ iters,
                            double *__restrict__ res, uint8_t *__restrict__ conv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seg) return;
    iters[s] = st[s].iter;
    res[2 * s] = st[s].norm0;
    res[2 * s + 1] = st[s].final_res;
    conv[s] = (uint8_t)st[s].converged;
}

// Workspace carve-up, each piece 256-byte aligned.
struct Carve {
    size_t val, dg, inv_d, x, b, r, z, p0, p1, y, pd, pu, st0, st1, act, wg, wg_end, iters, res, conv, total;
};
inline size_t al(size_t n) { return (n + 255) / 256 * 256; }
inline int tiles_of(int rows) { return std::max(1, std::min((rows + TS - 1) / TS, WMAX)); }
Carve carve(int C, int n_seg, int64_t nnz, int G) {
    Carve c{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += al(bytes);
        return at;
    };
    c.val = take(sizeof(double) * (size_t)nnz);
    c.dg = take(sizeof(double) * C);
    c.inv_d = take(sizeof(double) * C);
    c.x = take(sizeof(double) * C);
    c.b = take(sizeof(double) * C);
    c.r = take(sizeof(double) * C);
    c.z = take(sizeof(double) * C);
    c.p0 = take(sizeof(double) * C);
    c.p1 = take(sizeof(double) * C);
    c.y = take(sizeof(double) * C);
    c.pd = take(sizeof(double) * G);
    c.pu = take(sizeof(double) * 2 * (size_t)G);
    c.st0 = take(sizeof(SegState) * n_seg);
    c.st1 = take(sizeof(SegState) * n_seg);
    c.act = take(sizeof(int32_t) * n_seg);
    c.wg = take(sizeof(int4) * G);
    c.wg_end = take(sizeof(int32_t) * G);
    c.iters = take(sizeof(int32_t) * n_seg);
    c.res = take(sizeof(double) * 2 * n_seg);
    c.conv = take(n_seg);
    c.total = o;
    return c;
}
// Upper bound of the workgroup count over any split of C rows into n_seg segments.
inline int max_workgroups(int C, int n_seg) { return (C + TS - 1) / TS + n_seg; }

}  // namespace
}  // namespace gf

using namespace gf;

extern "C" int gf_seg_pcg_workspace(int32_t n_cells, int32_t n_seg, int64_t nnz_pad, int64_t *bytes) {
    GF_REQUIRE(bytes && n_cells >= 0 && n_seg >= 1 && nnz_pad >= 0, "bad arguments");
    *bytes = (int64_t)carve(n_cells, n_seg, nnz_pad, max_workgroups(n_cells, n_seg)).total;
    return GF_OK;
}

extern "C" int gf_seg_pcg(const gf_mesh *m, int32_t n_seg, const int32_t *seg_row, const int32_t *seg_row_host,
                          const int32_t *row_seg, const double *diag, const double *upper, const double *lower,
                          const double *b, double *x, int32_t precond, double atol, double rtol, int32_t max_iter,
                          int32_t check_every, void *workspace, int64_t workspace_bytes, int32_t *iters_out,
                          double *res_out, uint8_t *conv_out, int32_t *n_active_host, gf_stream_t stream) {
    GF_REQUIRE(m && seg_row && seg_row_host && diag && upper && lower && b && x && workspace && iters_out &&
                   res_out && conv_out && n_active_host,
               "null argument");
    (void)row_seg;  // the workgroup tables carry the row -> segment map
    GF_REQUIRE(n_seg >= 1 && max_iter >= 0 && check_every >= 1, "bad solver parameters");
    const int C = m->n_cells;
    GF_REQUIRE(seg_row_host[0] == 0 && seg_row_host[n_seg] == C, "segments must cover the rows");
    std::vector<int4> wg_tab;
    std::vector<int32_t> wg_end;
    for (int s = 0; s < n_seg; ++s) {
        const int rows = seg_row_host[s + 1] - seg_row_host[s];
        GF_REQUIRE(rows >= 1, "every segment needs at least one row");
        const int w0 = (int)wg_tab.size(), W = tiles_of(rows);
        for (int k = 0; k < W; ++k) {
            wg_tab.push_back(make_int4(s, w0, W, seg_row_host[s]));
            wg_end.push_back(seg_row_host[s + 1]);
        }
    }
    const int G = (int)wg_tab.size();
    const Carve cv = carve(C, n_seg, m->nnz_pad, max_workgroups(C, n_seg));
    GF_REQUIRE(G <= max_workgroups(C, n_seg), "workgroup table overflow");
    GF_REQUIRE(workspace_bytes >= (int64_t)cv.total, "workspace too small (gf_seg_pcg_workspace)");
    char *w = static_cast<char *>(workspace);
    hipStream_t s = S(stream);
    SegOps o{};
    o.C = C;
    o.G = G;
    o.sell_ptr = m->sell_ptr;
    o.sell_col = m->sell_col;
    o.val = (double *)(w + cv.val);
    o.dg = (double *)(w + cv.dg);
    o.inv_d = (double *)(w + cv.inv_d);
    o.seg_row = seg_row;
    o.wg = (const int4 *)(w + cv.wg);
    o.wg_end = (const int32_t *)(w + cv.wg_end);
    o.x = (double *)(w + cv.x);
    o.r = (double *)(w + cv.r);
    o.z = (double *)(w + cv.z);
    o.y = (double *)(w + cv.y);
    o.pd = (double *)(w + cv.pd);
    o.pu = (double *)(w + cv.pu);
    o.st0 = (SegState *)(w + cv.st0);
    o.st1 = (SegState *)(w + cv.st1);
    o.active = (int32_t *)(w + cv.act);
    o.atol = atol;
    o.rtol = rtol;
    o.max_iter = max_iter;
    double *pb[2] = {(double *)(w + cv.p0), (double *)(w + cv.p1)};
    double *wb = (double *)(w + cv.b);
    GF_CHECK_HIP(hipMemcpyAsync((void *)o.wg, wg_tab.data(), sizeof(int4) * G, hipMemcpyHostToDevice, s));
    GF_CHECK_HIP(hipMemcpyAsync((void *)o.wg_end, wg_end.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, s));
    GF_CHECK_HIP(hipMemcpyAsync(o.x, x, sizeof(double) * C, hipMemcpyDeviceToDevice, s));
    GF_CHECK_HIP(hipMemcpyAsync(wb, b, sizeof(double) * C, hipMemcpyDeviceToDevice, s));
    if (m->nnz_pad > 0)
        hipLaunchKernelGGL(k_seg_pack, dim3(grid_for(m->nnz_pad)), dim3(kBlock), 0, s, (int64_t)m->nnz_pad,
                           m->sell_code, upper, lower, (double *)o.val);
    hipLaunchKernelGGL(k_seg_diag, dim3(grid_for(C)), dim3(kBlock), 0, s, C, diag, precond, (double *)o.dg,
                       (double *)o.inv_d);
    hipLaunchKernelGGL(k_seg_init, dim3(G), dim3(TS), 0, s, o, (const double *)wb);
    GF_CHECK_LAUNCH();
    // the host arrays above must outlive their copies
    GF_CHECK_HIP(hipStreamSynchronize(s));
    // chunks of check_every iterations; the active flags of chunk k are read
    // while chunk k+1 runs (finished segments make the extra work a no-op)
    int *host_flag = n_active_host;  // pinned host words: [2][n_seg] active flags
    for (int k = 0; k < 2 * n_seg; ++k) host_flag[k] = 1;
    hipEvent_t ev[2];
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int it = 0, pc = 0, chunk = 0, pending = -1, rc = GF_OK;
    while (it < max_iter) {
        const int k = std::min(check_every, max_iter - it);
        for (int j = 0; j < k; ++j) {
            hipLaunchKernelGGL(k_seg_dir, dim3(G), dim3(TS), 0, s, o, it + j == 0 ? 0 : 1, (const double *)pb[pc],
                               pb[pc ^ 1]);
            hipLaunchKernelGGL(k_seg_upd, dim3(G), dim3(TS), 0, s, o, (const double *)pb[pc ^ 1]);
            pc ^= 1;
        }
        it += k;
        const int slot = chunk++ & 1;
        if (hipMemcpyAsync(host_flag + slot * n_seg, o.active, sizeof(int32_t) * n_seg, hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipEventRecord(ev[slot], s) != hipSuccess) {
            set_error("gf_seg_pcg: could not post the active flags");
            rc = GF_EINVAL;
            break;
        }
        if (pending >= 0) {
            if (hipEventSynchronize(ev[pending]) != hipSuccess) {
                set_error("gf_seg_pcg: event wait failed");
                rc = GF_EINVAL;
                break;
            }
            int any = 0;
            for (int q = 0; q < n_seg; ++q) any |= host_flag[pending * n_seg + q];
            if (!any) break;
        }
        pending = slot;
    }
    if (rc == GF_OK) {
        // the test of the last update (phase 2), then the statistics
        hipLaunchKernelGGL(k_seg_dir, dim3(G), dim3(TS), 0, s, o, max_iter == 0 ? 0 : 2, (const double *)pb[pc],
                           pb[pc ^ 1]);
        int32_t *d_it = (int32_t *)(w + cv.iters);
        double *d_res = (double *)(w + cv.res);
        uint8_t *d_conv = (uint8_t *)(w + cv.conv);
        hipLaunchKernelGGL(k_seg_stats, dim3(grid_for(n_seg)), dim3(kBlock), 0, s, n_seg, o.st1, d_it, d_res, d_conv);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(x, o.x, sizeof(double) * C, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(iters_out, d_it, sizeof(int32_t) * n_seg, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(res_out, d_res, sizeof(double) * 2 * n_seg, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(conv_out, d_conv, n_seg, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("gf_seg_pcg: could not read the statistics");
            rc = GF_EINVAL;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

