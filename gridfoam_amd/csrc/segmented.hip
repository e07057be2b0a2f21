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
// shuffles, then the wave sums added in wave order (a fixed order) by the
// first NQ threads and broadcast through LDS.
template <int NQ>
__device__ __forceinline__ void wg_sum(double (&v)[NQ]) {
    __shared__ double sh[NQ][TS / kWave];
    __shared__ double res[NQ];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double x = wave_sum(v[q]);
        if (lane == 0) sh[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double a = sh[threadIdx.x][0];
        for (int k = 1; k < TS / kWave; ++k) a += sh[threadIdx.x][k];
        res[threadIdx.x] = a;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = res[q];
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


// ---------------------------------------------------------------------------
// Segmented right-preconditioned BiCGSTAB for the three velocity components
// (solvers/bicgstab.py:76-272 with solvers/preconditioners.py:76-96), one
// scalar set per (segment, component).  Vectors are component-major [3][C].
// One iteration is three launches, each reducing the previous launch's
// partials of its segment in every workgroup:
//  * V: the tests of the last update (|r| < thresh, rho breakdown), beta,
//    pending restarts (r_hat = r, p = r, rho = r.r); p_new = r + beta (p -
//    omega v) built on the fly for the rows and their gathered columns,
//    y = M^-1 p_new, v = A y; partials r_hat.v, |r_hat|^2, |v|^2;
//  * T: the iteration count, the alpha breakdown test, alpha; s = r - alpha
//    v and z = M^-1 s on the fly, t = A z; partials |s|^2, t.t, t.s.  t is
//    computed before |s| is known: when s meets the threshold it is unused;
//  * U: |s| < thresh ? x += alpha y : the omega breakdown tests, x += alpha y
//    + omega z, r = s - omega t; partials |r|^2, r_hat.r.
// p and v are double buffered (their old values are gathered while the new
// ones are written); state and partials alternate between two buffers by
// launch parity.  Norms are 2-norms.
namespace gf {
namespace {

constexpr int NC = 3;       // components
constexpr int BQ = 3 * NC;  // partial quantities per launch at most
enum BiPhase : int { BI_ACTIVE = 0, BI_RESTART = 1, BI_DONE = 3 };

struct BiState {
    double thresh, norm0, final_res, rho_old, alpha, omega, beta, rhat2, rr;
    int iter, phase, converged, restarts;
};

struct BiOps {
    int C, G;
    const int32_t *sell_ptr, *sell_col;
    const double *val, *dg, *inv_d;
    const int4 *wg;
    const int32_t *wg_end;
    double *x, *r, *rh, *y, *s, *z, *t;
    double *part;  // [2][BQ][G]
    BiState *st;   // [2][n_seg][NC]
    int32_t *active;
    int n_seg;
    double atol, rtol;
    int max_iter, max_restart;
};

__device__ __forceinline__ void bi_break(BiState &S, const BiOps &o) {
    // bicgstab.py:238-246: restart while the budget lasts, else stop
    if (S.restarts >= o.max_restart || S.iter >= o.max_iter) {
        S.phase = BI_DONE;
    } else {
        S.restarts += 1;
        S.phase = BI_RESTART;
    }
}

__device__ __forceinline__ const double *bi_part_in(const BiOps &o, int q) {
    return o.part + (size_t)((q - 1) & 1) * BQ * o.G;
}
__device__ __forceinline__ double *bi_part_out(const BiOps &o, int q) {
    return o.part + (size_t)(q & 1) * BQ * o.G;
}

// The segment's three states: read by thread 0 into LDS, written back by
// the segment's first workgroup.
__device__ __forceinline__ void bi_state_load(const BiOps &o, int q, int s, BiState *sh) {
    if (threadIdx.x < NC) sh[threadIdx.x] = o.st[((size_t)((q - 1) & 1) * o.n_seg + s) * NC + threadIdx.x];
}
__device__ __forceinline__ void bi_state_store(const BiOps &o, int q, int s, int t, const BiState *sh) {
    if (t == 0 && threadIdx.x < NC) o.st[((size_t)(q & 1) * o.n_seg + s) * NC + threadIdx.x] = sh[threadIdx.x];
}

template <int NQ>
__device__ __forceinline__ void bi_partials_out(const BiOps &o, int q, int w, const double (&v)[NQ]) {
    if (threadIdx.x == 0) {
        double *P = bi_part_out(o, q);
#pragma unroll
        for (int k = 0; k < NQ; ++k) P[(size_t)k * o.G + w] = v[k];
    }
}

// r = b - A x ; partials |r|^2 per component
__global__ __launch_bounds__(TS) void k_sbi_init(BiOps o, int q, const double *__restrict__ b) {
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int t = w - d.y, W = d.z, r1 = o.wg_end[w];
    double v[NC] = {};
    for (int i = d.w + t * TS + (int)threadIdx.x; i < r1; i += W * TS) {
        const int sl = i >> 6, lane = i & 63;
        const int base = o.sell_ptr[sl];
        const int Wl = (o.sell_ptr[sl + 1] - base) >> 6;
        double ax[NC];
        const double dgi = o.dg[i];
#pragma unroll
        for (int c = 0; c < NC; ++c) ax[c] = dgi * o.x[(size_t)c * o.C + i];
        for (int k = 0; k < Wl; ++k) {
            const int e = base + k * 64 + lane;
            const int col = o.sell_col[e];
            const double a = o.val[e];
#pragma unroll
            for (int c = 0; c < NC; ++c) ax[c] += a * o.x[(size_t)c * o.C + col];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double rr = b[(size_t)c * o.C + i] - ax[c];
            o.r[(size_t)c * o.C + i] = rr;
            v[c] += rr * rr;
        }
    }
    wg_sum<NC>(v);
    bi_partials_out<NC>(o, q, w, v);
}

// V (see above).  first: the launch after k_sbi_init (decides on |r0|).
// po, vo: the current p and v; pn, vn: the buffers for the new ones
__global__ __launch_bounds__(TS) void k_sbi_v(BiOps o, int q, int first, const double *__restrict__ po,
                                              const double *__restrict__ vo, double *__restrict__ pn,
                                              double *__restrict__ vn) {
    __shared__ BiState sh[NC];
    __shared__ int mode_s[NC];  // 0 idle, 1 continue (beta update), 2 restart (p = r, r_hat = r)
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int s = d.x, w0 = d.y, t = w - w0, W = d.z, r1 = o.wg_end[w];
    double tot[2 * NC];  // first: |r|^2 per component; else |r|^2, r_hat.r per component
    load_partials<2 * NC>(bi_part_in(o, q), o.G, w0, W, tot);
    bi_state_load(o, q, s, sh);
    __syncthreads();
    wg_sum<2 * NC>(tot);
    if (threadIdx.x == 0) {
        for (int c = 0; c < NC; ++c) {
            BiState &S = sh[c];
            int mode = 0;
            if (first) {
                const double rr = tot[c];
                const double n0 = sqrt(rr);
                S.norm0 = n0;
                S.final_res = n0;
                S.thresh = fmax(o.atol, o.rtol * n0);
                S.iter = 0;
                S.restarts = 0;
                S.alpha = S.omega = S.beta = S.rhat2 = 0.0;
                S.converged = n0 < S.thresh;
                S.rr = rr;
                S.rho_old = rr;
                S.phase = S.converged ? BI_DONE : BI_ACTIVE;
                mode = S.converged ? 0 : 2;
            } else if (S.phase == BI_ACTIVE) {  // r was updated by the last U launch
                const double rr = tot[2 * c], rhr = tot[2 * c + 1];
                const double nr = sqrt(rr);
                S.final_res = nr;
                S.rr = rr;
                if (nr < S.thresh) {
                    S.converged = 1;
                    S.phase = BI_DONE;
                } else if (fabs(rhr) <= 1e-6 * sqrt(S.rhat2) * nr) {
                    bi_break(S, o);
                    if (S.phase == BI_RESTART) {  // restart at once
                        S.rho_old = rr;
                        S.phase = BI_ACTIVE;
                        mode = 2;
                    }
                } else {
                    S.beta = (rhr / S.rho_old) * (S.alpha / S.omega);
                    S.rho_old = rhr;
                    if (S.iter >= o.max_iter)
                        S.phase = BI_DONE;
                    else
                        mode = 1;
                }
            } else if (S.phase == BI_RESTART) {  // a breakdown in T or U: r is unchanged
                S.rho_old = S.rr;
                S.phase = BI_ACTIVE;
                mode = 2;
            }
            mode_s[c] = mode;
        }
        if (t == 0) {
            int any = 0;
            for (int c = 0; c < NC; ++c) any |= sh[c].phase != BI_DONE;
            o.active[s] = any;
        }
    }
    __syncthreads();
    bi_state_store(o, q, s, t, sh);
    int mode[NC];
    double be[NC], om[NC];
    bool work = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        mode[c] = mode_s[c];
        be[c] = sh[c].beta;
        om[c] = sh[c].omega;
        work |= mode[c] != 0;
    }
    if (!work) return;
    double v[BQ] = {};
    for (int i = d.w + t * TS + (int)threadIdx.x; i < r1; i += W * TS) {
        const int sl = i >> 6, lane = i & 63;
        const int base = o.sell_ptr[sl];
        const int Wl = (o.sell_ptr[sl + 1] - base) >> 6;
        const double idi = o.inv_d[i], dgi = o.dg[i];
        double acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const size_t k = (size_t)c * o.C + i;
            double pi = 0.0;
            if (mode[c] == 1) pi = o.r[k] + be[c] * (po[k] - om[c] * vo[k]);
            else if (mode[c] == 2) pi = o.r[k];
            if (mode[c]) {
                pn[k] = pi;
                o.y[k] = idi * pi;
            }
            acc[c] = dgi * (idi * pi);
        }
        for (int kk = 0; kk < Wl; ++kk) {
            const int e = base + kk * 64 + lane;
            const int col = o.sell_col[e];
            const double a = o.val[e] * o.inv_d[col];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const size_t k = (size_t)c * o.C + col;
                if (mode[c] == 1) acc[c] += a * (o.r[k] + be[c] * (po[k] - om[c] * vo[k]));
                else if (mode[c] == 2) acc[c] += a * o.r[k];
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (!mode[c]) continue;
            const size_t k = (size_t)c * o.C + i;
            double h;
            if (mode[c] == 2) {
                h = o.r[k];
                o.rh[k] = h;
            } else {
                h = o.rh[k];
            }
            vn[k] = acc[c];
            v[3 * c] += h * acc[c];
            v[3 * c + 1] += h * h;
            v[3 * c + 2] += acc[c] * acc[c];
        }
    }
    wg_sum<BQ>(v);
    bi_partials_out<BQ>(o, q, w, v);
}

// T (see above); vv: the v written by the last V launch
__global__ __launch_bounds__(TS) void k_sbi_t(BiOps o, int q, const double *__restrict__ vv) {
    __shared__ BiState sh[NC];
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int s = d.x, w0 = d.y, t = w - w0, W = d.z, r1 = o.wg_end[w];
    double tot[BQ];
    load_partials<BQ>(bi_part_in(o, q), o.G, w0, W, tot);
    bi_state_load(o, q, s, sh);
    __syncthreads();
    bool any = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) any |= sh[c].phase == BI_ACTIVE;
    if (!any) {
        bi_state_store(o, q, s, t, sh);
        return;
    }
    wg_sum<BQ>(tot);
    if (threadIdx.x == 0)
        for (int c = 0; c < NC; ++c) {
            BiState &S = sh[c];
            if (S.phase != BI_ACTIVE) continue;
            S.iter += 1;
            S.rhat2 = tot[3 * c + 1];
            const double den = tot[3 * c];
            if (fabs(den) <= 1e-6 * sqrt(tot[3 * c + 1]) * sqrt(tot[3 * c + 2]))
                bi_break(S, o);
            else
                S.alpha = S.rho_old / den;
        }
    __syncthreads();
    bi_state_store(o, q, s, t, sh);
    bool act[NC];
    double al[NC];
    bool work = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        act[c] = sh[c].phase == BI_ACTIVE;
        al[c] = sh[c].alpha;
        work |= act[c];
    }
    if (!work) return;
    double v[BQ] = {};
    for (int i = d.w + t * TS + (int)threadIdx.x; i < r1; i += W * TS) {
        const int sl = i >> 6, lane = i & 63;
        const int base = o.sell_ptr[sl];
        const int Wl = (o.sell_ptr[sl + 1] - base) >> 6;
        const double idi = o.inv_d[i], dgi = o.dg[i];
        double acc[NC], si[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const size_t k = (size_t)c * o.C + i;
            si[c] = act[c] ? o.r[k] - al[c] * vv[k] : 0.0;
            acc[c] = dgi * (idi * si[c]);
        }
        for (int kk = 0; kk < Wl; ++kk) {
            const int e = base + kk * 64 + lane;
            const int col = o.sell_col[e];
            const double a = o.val[e] * o.inv_d[col];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (!act[c]) continue;
                const size_t k = (size_t)c * o.C + col;
                acc[c] += a * (o.r[k] - al[c] * vv[k]);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (!act[c]) continue;
            const size_t k = (size_t)c * o.C + i;
            o.s[k] = si[c];
            o.z[k] = idi * si[c];
            o.t[k] = acc[c];
            v[3 * c] += si[c] * si[c];
            v[3 * c + 1] += acc[c] * acc[c];
            v[3 * c + 2] += acc[c] * si[c];
        }
    }
    wg_sum<BQ>(v);
    bi_partials_out<BQ>(o, q, w, v);
}

// U (see above)
__global__ __launch_bounds__(TS) void k_sbi_u(BiOps o, int q) {
    __shared__ BiState sh[NC];
    __shared__ int mode_s[NC];  // 0 idle, 1 full update, 2 converged on s (x += alpha y)
    const int w = blockIdx.x;
    const int4 d = o.wg[w];
    const int s = d.x, w0 = d.y, t = w - w0, W = d.z, r1 = o.wg_end[w];
    double tot[BQ];
    load_partials<BQ>(bi_part_in(o, q), o.G, w0, W, tot);
    bi_state_load(o, q, s, sh);
    __syncthreads();
    bool any = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) any |= sh[c].phase == BI_ACTIVE;
    if (!any) {
        bi_state_store(o, q, s, t, sh);
        return;
    }
    wg_sum<BQ>(tot);
    if (threadIdx.x == 0)
        for (int c = 0; c < NC; ++c) {
            BiState &S = sh[c];
            int mode = 0;
            if (S.phase == BI_ACTIVE) {
                const double ns = sqrt(tot[3 * c]);
                S.final_res = ns;
                if (ns < S.thresh) {
                    S.converged = 1;
                    S.phase = BI_DONE;
                    mode = 2;
                } else {
                    const double tt = tot[3 * c + 1];
                    bool brk = fabs(tt) <= 1e-6 * sqrt(tt) * sqrt(tt);
                    double om = 0.0;
                    if (!brk) {
                        om = tot[3 * c + 2] / tt;
                        brk = fabs(om) <= 1e-6;
                    }
                    if (brk) {
                        bi_break(S, o);
                    } else {
                        S.omega = om;
                        mode = 1;
                    }
                }
            }
            mode_s[c] = mode;
        }
    __syncthreads();
    bi_state_store(o, q, s, t, sh);
    int mode[NC];
    double al[NC], om[NC];
    bool work = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        mode[c] = mode_s[c];
        al[c] = sh[c].alpha;
        om[c] = sh[c].omega;
        work |= mode[c] != 0;
    }
    if (!work) return;
    double v[2 * NC] = {};
    for (int i = d.w + t * TS + (int)threadIdx.x; i < r1; i += W * TS) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const size_t k = (size_t)c * o.C + i;
            if (mode[c] == 2) {
                o.x[k] = o.x[k] + al[c] * o.y[k];
            } else if (mode[c] == 1) {
                o.x[k] = o.x[k] + al[c] * o.y[k] + om[c] * o.z[k];
                const double rr = o.s[k] - om[c] * o.t[k];
                o.r[k] = rr;
                v[2 * c] += rr * rr;
                v[2 * c + 1] += o.rh[k] * rr;
            }
        }
    }
    wg_sum<2 * NC>(v);
    bi_partials_out<2 * NC>(o, q, w, v);
}

// [C][3] <-> [3][C]
__global__ void k_sbi_transpose(int C, const double *__restrict__ in, double *__restrict__ out, int to_cm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (to_cm)
            out[(size_t)c * C + i] = in[(size_t)i * NC + c];
        else
            out[(size_t)i * NC + c] = in[(size_t)c * C + i];
    }
}

__global__ void k_sbi_stats(int n, const BiState *__restrict__ st, int32_t *__restrict__ iters,
                            double *__restrict__ res, uint8_t *__restrict__ conv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    iters[k] = st[k].iter;
    res[2 * k] = st[k].norm0;
    res[2 * k + 1] = st[k].final_res;
    conv[k] = (uint8_t)st[k].converged;
}

struct BiCarve {
    size_t val, dg, inv_d, b, x, r, rh, y, s, z, t, p0, p1, v0, v1, part, st, act, wg, wg_end, iters, res, conv, total;
};
BiCarve bi_carve(int C, int n_seg, int64_t nnz, int G) {
    BiCarve c{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += al(bytes);
        return at;
    };
    const size_t vec = sizeof(double) * NC * (size_t)C;
    c.val = take(sizeof(double) * (size_t)nnz);
    c.dg = take(sizeof(double) * C);
    c.inv_d = take(sizeof(double) * C);
    c.b = take(vec);
    c.x = take(vec);
    c.r = take(vec);
    c.rh = take(vec);
    c.y = take(vec);
    c.s = take(vec);
    c.z = take(vec);
    c.t = take(vec);
    c.p0 = take(vec);
    c.p1 = take(vec);
    c.v0 = take(vec);
    c.v1 = take(vec);
    c.part = take(sizeof(double) * 2 * BQ * (size_t)G);
    c.st = take(sizeof(BiState) * 2 * NC * (size_t)n_seg);
    c.act = take(sizeof(int32_t) * n_seg);
    c.wg = take(sizeof(int4) * G);
    c.wg_end = take(sizeof(int32_t) * G);
    c.iters = take(sizeof(int32_t) * NC * n_seg);
    c.res = take(sizeof(double) * 2 * NC * n_seg);
    c.conv = take(NC * (size_t)n_seg);
    c.total = o;
    return c;
}

}  // namespace
}  // namespace gf

extern "C" int gf_seg_bicgstab_workspace(int32_t n_cells, int32_t n_seg, int64_t nnz_pad, int64_t *bytes) {
    GF_REQUIRE(bytes && n_cells >= 0 && n_seg >= 1 && nnz_pad >= 0, "bad arguments");
    *bytes = (int64_t)bi_carve(n_cells, n_seg, nnz_pad, max_workgroups(n_cells, n_seg)).total;
    return GF_OK;
}

extern "C" int gf_seg_bicgstab(const gf_mesh *m, int32_t n_seg, const int32_t *seg_row_host, const double *diag,
                               const double *upper, const double *lower, const double *b, double *x,
                               int32_t precond, double atol, double rtol, int32_t max_iter, int32_t max_restart,
                               int32_t check_every, void *workspace, int64_t workspace_bytes, int32_t *iters_out,
                               double *res_out, uint8_t *conv_out, int32_t *n_active_host, gf_stream_t stream) {
    GF_REQUIRE(m && seg_row_host && diag && upper && lower && b && x && workspace && iters_out && res_out &&
                   conv_out && n_active_host,
               "null argument");
    GF_REQUIRE(n_seg >= 1 && max_iter >= 0 && max_restart >= 0 && check_every >= 1, "bad solver parameters");
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
    GF_REQUIRE(G <= max_workgroups(C, n_seg), "workgroup table overflow");
    const BiCarve cv = bi_carve(C, n_seg, m->nnz_pad, max_workgroups(C, n_seg));
    GF_REQUIRE(workspace_bytes >= (int64_t)cv.total, "workspace too small (gf_seg_bicgstab_workspace)");
    char *w = static_cast<char *>(workspace);
    hipStream_t st = S(stream);
    BiOps o{};
    o.C = C;
    o.G = G;
    o.sell_ptr = m->sell_ptr;
    o.sell_col = m->sell_col;
    o.val = (double *)(w + cv.val);
    o.dg = (double *)(w + cv.dg);
    o.inv_d = (double *)(w + cv.inv_d);
    o.wg = (const int4 *)(w + cv.wg);
    o.wg_end = (const int32_t *)(w + cv.wg_end);
    o.x = (double *)(w + cv.x);
    o.r = (double *)(w + cv.r);
    o.rh = (double *)(w + cv.rh);
    o.y = (double *)(w + cv.y);
    o.s = (double *)(w + cv.s);
    o.z = (double *)(w + cv.z);
    o.t = (double *)(w + cv.t);
    double *pbuf[2] = {(double *)(w + cv.p0), (double *)(w + cv.p1)};
    double *vbuf[2] = {(double *)(w + cv.v0), (double *)(w + cv.v1)};
    o.part = (double *)(w + cv.part);
    o.st = (BiState *)(w + cv.st);
    o.active = (int32_t *)(w + cv.act);
    o.n_seg = n_seg;
    o.atol = atol;
    o.rtol = rtol;
    o.max_iter = max_iter;
    o.max_restart = max_restart;
    double *wb = (double *)(w + cv.b);
    GF_CHECK_HIP(hipMemcpyAsync((void *)o.wg, wg_tab.data(), sizeof(int4) * G, hipMemcpyHostToDevice, st));
    GF_CHECK_HIP(hipMemcpyAsync((void *)o.wg_end, wg_end.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, st));
    if (m->nnz_pad > 0)
        hipLaunchKernelGGL(k_seg_pack, dim3(grid_for(m->nnz_pad)), dim3(kBlock), 0, st, (int64_t)m->nnz_pad,
                           m->sell_code, upper, lower, (double *)o.val);
    hipLaunchKernelGGL(k_seg_diag, dim3(grid_for(C)), dim3(kBlock), 0, st, C, diag, precond, (double *)o.dg,
                       (double *)o.inv_d);
    hipLaunchKernelGGL(k_sbi_transpose, dim3(grid_for(C)), dim3(kBlock), 0, st, C, x, o.x, 1);
    hipLaunchKernelGGL(k_sbi_transpose, dim3(grid_for(C)), dim3(kBlock), 0, st, C, b, wb, 1);
    int q = 0;
    hipLaunchKernelGGL(k_sbi_init, dim3(G), dim3(TS), 0, st, o, q++, (const double *)wb);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(st));  // the host tables above must outlive their copies
    int *host_flag = n_active_host;  // pinned host words: [2][n_seg] active flags
    for (int k = 0; k < 2 * n_seg; ++k) host_flag[k] = 1;
    hipEvent_t ev[2];
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    GF_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int it = 0, pb = 0, chunk = 0, pending = -1, rc = GF_OK;
    while (it < max_iter) {
        const int k = std::min(check_every, max_iter - it);
        for (int j = 0; j < k; ++j) {
            hipLaunchKernelGGL(k_sbi_v, dim3(G), dim3(TS), 0, st, o, q++, it + j == 0 ? 1 : 0,
                               (const double *)pbuf[pb], (const double *)vbuf[pb], pbuf[pb ^ 1], vbuf[pb ^ 1]);
            pb ^= 1;  // the new p, v
            hipLaunchKernelGGL(k_sbi_t, dim3(G), dim3(TS), 0, st, o, q++, (const double *)vbuf[pb]);
            hipLaunchKernelGGL(k_sbi_u, dim3(G), dim3(TS), 0, st, o, q++);
        }
        it += k;
        const int slot = chunk++ & 1;
        if (hipMemcpyAsync(host_flag + slot * n_seg, o.active, sizeof(int32_t) * n_seg, hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipEventRecord(ev[slot], st) != hipSuccess) {
            set_error("gf_seg_bicgstab: could not post the active flags");
            rc = GF_EINVAL;
            break;
        }
        if (pending >= 0) {
            if (hipEventSynchronize(ev[pending]) != hipSuccess) {
                set_error("gf_seg_bicgstab: event wait failed");
                rc = GF_EINVAL;
                break;
            }
            int any = 0;
            for (int s = 0; s < n_seg; ++s) any |= host_flag[pending * n_seg + s];
            if (!any) break;
        }
        pending = slot;
    }
    if (rc == GF_OK) {
        // the tests of the last update (every component that is still active stops here)
        hipLaunchKernelGGL(k_sbi_v, dim3(G), dim3(TS), 0, st, o, q++, max_iter == 0 ? 1 : 0,
                           (const double *)pbuf[pb], (const double *)vbuf[pb], pbuf[pb ^ 1], vbuf[pb ^ 1]);
        const BiState *last = o.st + (size_t)((q - 1) & 1) * n_seg * NC;
        int32_t *d_it = (int32_t *)(w + cv.iters);
        double *d_res = (double *)(w + cv.res);
        uint8_t *d_conv = (uint8_t *)(w + cv.conv);
        hipLaunchKernelGGL(k_sbi_stats, dim3(grid_for(NC * n_seg)), dim3(kBlock), 0, st, NC * n_seg, last, d_it,
                           d_res, d_conv);
        hipLaunchKernelGGL(k_sbi_transpose, dim3(grid_for(C)), dim3(kBlock), 0, st, C, (const double *)o.x, x, 0);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(iters_out, d_it, sizeof(int32_t) * NC * n_seg, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(res_out, d_res, sizeof(double) * 2 * NC * n_seg, hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipMemcpyAsync(conv_out, d_conv, NC * n_seg, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("gf_seg_bicgstab: could not read the statistics");
            rc = GF_EINVAL;
        }
    }
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}
