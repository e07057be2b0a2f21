// Octree face-neighbour operator build on the GPU.
//
// Replaces the external mesh builder that gridfoam calls through
// FluxelManager.build_axis_projected_mesh (src/gridfoam/core/grid/
// factory.py:32-65; refinement semantics from src/gridfoam/meta/config.py:
// 90-145: refine leaves intersecting a box below its level, then 2:1
// balance, then n_leaf uniform splits).
//
// Design: leaves are kept in Morton order at all times, so no sort is ever
// needed.  Root blocks are emitted in Morton order by compacting the codes
// of the bounding power-of-two cube; every refinement/balance pass computes
// a split flag per leaf, scans the child counts and writes each split leaf's
// eight children (in Morton child order) at its scanned position, which
// keeps the array sorted.  Neighbour lookups are predecessor binary searches
// over the sorted keys.  The integer output is bit-identical to the CPU
// oracle (oracle/octree_oracle.c), which reaches the same leaves by a
// different route (unsorted list + qsort per pass).
#include <vector>

#include "gf_common.h"
#include "gf_shapes.h"

namespace gf {

__device__ __host__ __forceinline__ uint64_t spread3(uint64_t x) {
    x &= 0x1fffffULL;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}
__device__ __host__ __forceinline__ uint64_t compact3(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffffULL;
    return x;
}
__device__ __host__ __forceinline__ uint64_t morton3(uint64_t x, uint64_t y, uint64_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

// ------------------------------------------------------------------ scan
// Exclusive scan of int32 counts (deterministic; three-phase, recursive).
__global__ void k_scan_block(const int32_t *in, int32_t *out, int32_t *block_sums, int32_t n) {
    __shared__ int32_t buf[1024];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    int32_t v = i < n ? in[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int32_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = buf[threadIdx.x] - v;
    if (threadIdx.x == 1023) block_sums[blockIdx.x] = buf[1023];
}
__global__ void k_scan_add(int32_t *out, const int32_t *offs, int32_t n) {
    const int i = blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += offs[blockIdx.x];
}

// Scratch: caller gives `tmp` with room for scan_tmp_ints(n) ints.
static int64_t scan_tmp_ints(int64_t n) {
    int64_t t = 0;
    while (n > 1) {
        n = (n + 1023) / 1024;
        t += 2 * n + 2;
    }
    return t + 4;
}
static void scan_rec(const int32_t *in, int32_t *out, int32_t n, int32_t *tmp, hipStream_t s) {
    const int nb = (n + 1023) / 1024;
    int32_t *sums = tmp, *sums_scan = tmp + nb + 1;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(1024), 0, s, in, out, sums, n);
    if (nb > 1) {
        scan_rec(sums, sums_scan, nb, tmp + 2 * nb + 2, s);
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, s, out, sums_scan, n);
    }
}

// -------------------------------------------------------------- octree
struct Leaf {
    uint64_t key;
    int32_t i, j, k, level;
};

struct Params {
    double lower[3], upper[3];
    int32_t root[3];
    int32_t lmax;
    int32_t n_regions;
    const double *rmin, *rmax;
    const int32_t *rlevel;
    int32_t n_shapes, target_level;  // surface refinement of analytic bodies
    const gf_shape *shapes;
};

__device__ __forceinline__ uint64_t leaf_key(int32_t i, int32_t j, int32_t k, int level, int lmax) {
    const int s = lmax - level;
    return morton3((uint64_t)i << s, (uint64_t)j << s, (uint64_t)k << s);
}

__global__ void k_root_flags(int64_t ncode, int P, Params p, int32_t *flag) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncode) return;
    uint64_t x = compact3((uint64_t)c), y = compact3((uint64_t)c >> 1), z = compact3((uint64_t)c >> 2);
    flag[c] = (x < (uint64_t)p.root[0] && y < (uint64_t)p.root[1] && z < (uint64_t)p.root[2]) ? 1 : 0;
}
__global__ void k_root_emit(int64_t ncode, Params p, const int32_t *flag, const int32_t *pos, Leaf *out) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncode || !flag[c]) return;
    Leaf l;
    l.i = (int32_t)compact3((uint64_t)c);
    l.j = (int32_t)compact3((uint64_t)c >> 1);
    l.k = (int32_t)compact3((uint64_t)c >> 2);
    l.level = 0;
    l.key = leaf_key(l.i, l.j, l.k, 0, p.lmax);
    out[pos[c]] = l;
}

__device__ __forceinline__ bool intersects(const Params &p, const Leaf &l, int r) {
    const int32_t idx[3] = {l.i, l.j, l.k};
    for (int a = 0; a < 3; ++a) {
        double h = (p.upper[a] - p.lower[a]) / ldexp((double)p.root[a], l.level);
        double lo = fma((double)idx[a], h, p.lower[a]);
        double hi = fma((double)(idx[a] + 1), h, p.lower[a]);
        if (!(lo < p.rmax[3 * r + a] && hi > p.rmin[3 * r + a])) return false;
    }
    return true;
}

// A body's surface may pass through the leaf: |G(centre)| <= half the
// leaf's diagonal (G bounds the distance to the surface from below).  Plane
// (kind 0), box (1) and bluff bodies (3).
__device__ bool near_surface(const Params &p, const Leaf &l) {
#pragma clang fp contract(off)
    const int32_t idx[3] = {l.i, l.j, l.k};
    double c[3], h[3];
    for (int a = 0; a < 3; ++a) {
        h[a] = (p.upper[a] - p.lower[a]) / ldexp((double)p.root[a], l.level);
        c[a] = fma((double)idx[a] + 0.5, h[a], p.lower[a]);
    }
    const double hd = 0.5 * sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (int k = 0; k < p.n_shapes; ++k) {
        const gf_shape &b = p.shapes[k];
        double g;
        if (b.kind == 0)
            g = c[b.axis] - b.p[0];
        else if (b.kind == 1)
            g = box_g(b, c[0], c[1], c[2]);
        else
            g = bluff_g(b, c[0], c[1], c[2]);
        if (fabs(g) <= hd) return true;
    }
    return false;
}

__global__ void k_region_flags(int32_t n, Params p, const Leaf *cells, int32_t *flag, int32_t *count) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const Leaf l = cells[c];
    int f = l.level < p.target_level && near_surface(p, l);
    for (int r = 0; r < p.n_regions && !f; ++r)
        if (l.level < p.rlevel[r] && intersects(p, l, r)) {
            f = 1;
            break;
        }
    flag[c] = f;
    if (f) atomicAdd(count, 1);
}

__device__ __forceinline__ int64_t find_leaf(const Leaf *cells, int32_t n, const Params &p, int64_t x,
                                             int64_t y, int64_t z) {
    if (x < 0 || y < 0 || z < 0) return -1;
    if (x >= ((int64_t)p.root[0] << p.lmax) || y >= ((int64_t)p.root[1] << p.lmax) ||
        z >= ((int64_t)p.root[2] << p.lmax))
        return -1;
    const uint64_t key = morton3((uint64_t)x, (uint64_t)y, (uint64_t)z);
    int64_t lo = 0, hi = n - 1, ans = -1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (cells[mid].key <= key) {
            ans = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    return ans;
}

__device__ __forceinline__ void across(const Leaf &l, int lmax, int dir, int64_t (&pt)[3]) {
    const int s = lmax - l.level;
    const int64_t lo[3] = {(int64_t)l.i << s, (int64_t)l.j << s, (int64_t)l.k << s};
    const int a = dir / 2;
    for (int b = 0; b < 3; ++b) pt[b] = lo[b];
    pt[a] = (dir & 1) ? lo[a] + ((int64_t)1 << s) : lo[a] - 1;
}

// A leaf more than one level finer than its face neighbour forces that
// neighbour to split (plain stores of 1: every writer writes the same value).
__global__ void k_balance_flags(int32_t n, Params p, const Leaf *cells, int32_t *flag) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const Leaf l = cells[c];
    for (int dir = 0; dir < 6; ++dir) {
        int64_t pt[3];
        across(l, p.lmax, dir, pt);
        int64_t nb = find_leaf(cells, n, p, pt[0], pt[1], pt[2]);
        if (nb >= 0 && cells[nb].level < l.level - 1) flag[nb] = 1;
    }
}

__global__ void k_count_flags(int32_t n, const int32_t *flag, int32_t *cnt, int32_t *total) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    cnt[c] = flag[c] ? 8 : 1;
    if (flag[c]) atomicAdd(total, 1);
}

__global__ void k_split(int32_t n, int lmax, const Leaf *cells, const int32_t *flag, const int32_t *pos,
                        Leaf *out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const Leaf l = cells[c];
    const int32_t w = pos[c];
    if (!flag[c]) {
        out[w] = l;
        return;
    }
    for (int ch = 0; ch < 8; ++ch) {
        Leaf q;
        q.level = l.level + 1;
        q.i = 2 * l.i + (ch & 1);
        q.j = 2 * l.j + ((ch >> 1) & 1);
        q.k = 2 * l.k + ((ch >> 2) & 1);
        q.key = leaf_key(q.i, q.j, q.k, q.level, lmax);
        out[w + ch] = q;
    }
}

__device__ __forceinline__ void sort4(int32_t (&v)[4]) {
    for (int a = 1; a < 4; ++a) {
        int32_t x = v[a];
        int b = a - 1;
        while (b >= 0 && v[b] > x) {
            v[b + 1] = v[b];
            --b;
        }
        v[b + 1] = x;
    }
}

// mode 0: count faces per cell; mode 1: fill
__global__ void k_faces(int32_t n, Params p, const Leaf *cells, int mode, int32_t *fcnt, int32_t *bcnt,
                        const int32_t *fpos, const int32_t *bpos, int32_t *f_owner, int32_t *f_nbr,
                        int32_t *f_axis, int32_t *b_owner, int32_t *b_dir) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const Leaf l = cells[c];
    const int32_t idx[3] = {l.i, l.j, l.k};
    int32_t nf = 0, nb = 0;
    int32_t fw = mode ? fpos[c] : 0, bw = mode ? bpos[c] : 0;
    for (int a = 0; a < 3; ++a) {
        int64_t pt[3];
        across(l, p.lmax, 2 * a + 1, pt);
        int64_t nbr = find_leaf(cells, n, p, pt[0], pt[1], pt[2]);
        if (nbr < 0) continue;
        if (cells[nbr].level <= l.level) {
            if (mode) {
                f_owner[fw] = c;
                f_nbr[fw] = (int32_t)nbr;
                f_axis[fw] = a;
                ++fw;
            }
            ++nf;
        } else {
            int32_t q4[4];
            const int s = p.lmax - (l.level + 1);
            const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
            for (int q = 0; q < 4; ++q) {
                int64_t pq[3];
                pq[a] = (int64_t)(2 * (idx[a] + 1)) << s;
                pq[t1] = (int64_t)(2 * idx[t1] + (q & 1)) << s;
                pq[t2] = (int64_t)(2 * idx[t2] + (q >> 1)) << s;
                q4[q] = (int32_t)find_leaf(cells, n, p, pq[0], pq[1], pq[2]);
            }
            sort4(q4);
            for (int q = 0; q < 4; ++q) {
                if (mode) {
                    f_owner[fw] = c;
                    f_nbr[fw] = q4[q];
                    f_axis[fw] = a;
                    ++fw;
                }
                ++nf;
            }
        }
    }
    for (int dir = 0; dir < 6; ++dir) {
        const int a = dir / 2;
        const int64_t extent = (int64_t)p.root[a] << l.level;
        const bool on = (dir & 1) ? (idx[a] + 1 == extent) : (idx[a] == 0);
        if (!on) continue;
        if (mode) {
            b_owner[bw] = c;
            b_dir[bw] = dir;
            ++bw;
        }
        ++nb;
    }
    if (!mode) {
        fcnt[c] = nf;
        bcnt[c] = nb;
    }
}

__global__ void k_unpack(int32_t n, const Leaf *cells, int32_t *level, int32_t *ijk) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const Leaf l = cells[c];
    level[c] = l.level;
    ijk[3 * c + 0] = l.i;
    ijk[3 * c + 1] = l.j;
    ijk[3 * c + 2] = l.k;
}

}  // namespace gf

using namespace gf;

struct gf_octree {
    int64_t n_cells, n_faces, n_bnd;
    int32_t lmax;
    Leaf *cells;
    int32_t *f_owner, *f_nbr, *f_axis, *b_owner, *b_dir;
};

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes > 0 ? bytes : 16); }
    template <typename T>
    T *as() const {
        return reinterpret_cast<T *>(p);
    }
};

// One refinement/balance expansion: returns the number of split leaves.
int expand(Leaf *&cells, int64_t &n, const int32_t *flag, int32_t lmax, hipStream_t s, int64_t &n_split) {
    DevBuf cnt, pos, tmp, total;
    GF_CHECK_HIP(cnt.alloc(sizeof(int32_t) * n));
    GF_CHECK_HIP(pos.alloc(sizeof(int32_t) * (n + 1)));
    GF_CHECK_HIP(tmp.alloc(sizeof(int32_t) * scan_tmp_ints(n + 1)));
    GF_CHECK_HIP(total.alloc(sizeof(int32_t)));
    GF_CHECK_HIP(hipMemsetAsync(total.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_count_flags, dim3(grid_for(n)), dim3(kBlock), 0, s, (int32_t)n, flag,
                       cnt.as<int32_t>(), total.as<int32_t>());
    int32_t h_total = 0;
    GF_CHECK_HIP(hipMemcpyAsync(&h_total, total.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GF_CHECK_HIP(hipStreamSynchronize(s));
    n_split = h_total;
    if (h_total == 0) return GF_OK;
    scan_rec(cnt.as<int32_t>(), pos.as<int32_t>(), (int32_t)n, tmp.as<int32_t>(), s);
    const int64_t n_new = n + 7 * (int64_t)h_total;
    GF_REQUIRE(n_new < (int64_t)INT32_MAX / 8, "octree too large for int32 indexing");
    Leaf *out = nullptr;
    GF_CHECK_HIP(hipMalloc(&out, sizeof(Leaf) * n_new));
    hipLaunchKernelGGL(k_split, dim3(grid_for(n)), dim3(kBlock), 0, s, (int32_t)n, lmax, cells, flag,
                       pos.as<int32_t>(), out);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(s));
    GF_CHECK_HIP(hipFree(cells));
    cells = out;
    n = n_new;
    return GF_OK;
}

}  // namespace

extern "C" int gf_exclusive_scan_i32(const int32_t *in, int32_t *out, int32_t n, gf_stream_t stream) {
    GF_REQUIRE(in && out && n >= 0, "bad scan arguments");
    if (n == 0) return GF_OK;
    DevBuf tmp;
    GF_CHECK_HIP(tmp.alloc(sizeof(int32_t) * scan_tmp_ints(n)));
    scan_rec(in, out, n, tmp.as<int32_t>(), S(stream));
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(S(stream)));
    return GF_OK;
}

extern "C" int gf_octree_build(const double lower[3], const double upper[3], const int32_t root[3],
                               int32_t n_regions, const double *rmin, const double *rmax,
                               const int32_t *rlevel, int32_t n_leaf, gf_octree **out, int64_t counts_out[4],
                               gf_stream_t stream) {
    return gf_octree_build_surface(lower, upper, root, n_regions, rmin, rmax, rlevel, 0, nullptr, 0, n_leaf, out,
                                   counts_out, stream);
}

extern "C" int gf_octree_build_surface(const double lower[3], const double upper[3], const int32_t root[3],
                                       int32_t n_regions, const double *rmin, const double *rmax,
                                       const int32_t *rlevel, int32_t n_shapes, const gf_shape *shapes,
                                       int32_t target_level, int32_t n_leaf, gf_octree **out,
                                       int64_t counts_out[4], gf_stream_t stream) {
    GF_REQUIRE(lower && upper && root && out && counts_out, "null argument");
    GF_REQUIRE(n_regions >= 0 && n_leaf >= 0, "negative region/leaf count");
    GF_REQUIRE(n_shapes >= 0 && n_shapes <= 16 && target_level >= 0 && (shapes || n_shapes == 0),
               "bad surface refinement arguments (at most 16 bodies)");
    for (int k = 0; k < n_shapes; ++k) {
        GF_REQUIRE(shapes[k].kind == 0 || shapes[k].kind == 1 || shapes[k].kind == 3,
                   "surface refinement is defined for plane, box and bluff bodies");
        GF_REQUIRE(shapes[k].kind != 0 || (shapes[k].axis >= 0 && shapes[k].axis <= 2), "plane axis must be 0..2");
        GF_REQUIRE(shapes[k].kind != 3 || (shapes[k].axis >= 0 && shapes[k].axis <= 4), "at most 4 struts");
    }
    for (int a = 0; a < 3; ++a) GF_REQUIRE(root[a] >= 1 && upper[a] > lower[a], "bad domain/root");
    hipStream_t s = S(stream);
    Params p{};
    int32_t lmax = n_shapes > 0 ? target_level : 0;
    for (int r = 0; r < n_regions; ++r) lmax = rlevel[r] > lmax ? rlevel[r] : lmax;
    lmax += n_leaf;
    int maxroot = root[0] > root[1] ? root[0] : root[1];
    maxroot = maxroot > root[2] ? maxroot : root[2];
    int P = 1;
    while (P < maxroot) P <<= 1;
    GF_REQUIRE((int64_t)P << lmax <= (1LL << 21), "octree extent exceeds 21-bit Morton range");
    for (int a = 0; a < 3; ++a) {
        p.lower[a] = lower[a];
        p.upper[a] = upper[a];
        p.root[a] = root[a];
    }
    p.lmax = lmax;
    p.n_regions = n_regions;
    DevBuf d_rmin, d_rmax, d_rlev;
    GF_CHECK_HIP(d_rmin.alloc(sizeof(double) * 3 * (n_regions + 1)));
    GF_CHECK_HIP(d_rmax.alloc(sizeof(double) * 3 * (n_regions + 1)));
    GF_CHECK_HIP(d_rlev.alloc(sizeof(int32_t) * (n_regions + 1)));
    if (n_regions > 0) {
        GF_CHECK_HIP(hipMemcpyAsync(d_rmin.p, rmin, sizeof(double) * 3 * n_regions, hipMemcpyHostToDevice, s));
        GF_CHECK_HIP(hipMemcpyAsync(d_rmax.p, rmax, sizeof(double) * 3 * n_regions, hipMemcpyHostToDevice, s));
        GF_CHECK_HIP(hipMemcpyAsync(d_rlev.p, rlevel, sizeof(int32_t) * n_regions, hipMemcpyHostToDevice, s));
    }
    p.rmin = d_rmin.as<double>();
    p.rmax = d_rmax.as<double>();
    p.rlevel = d_rlev.as<int32_t>();
    DevBuf d_shapes;
    GF_CHECK_HIP(d_shapes.alloc(sizeof(gf_shape) * (n_shapes + 1)));
    if (n_shapes > 0)
        GF_CHECK_HIP(hipMemcpyAsync(d_shapes.p, shapes, sizeof(gf_shape) * n_shapes, hipMemcpyHostToDevice, s));
    p.n_shapes = n_shapes;
    p.target_level = n_shapes > 0 ? target_level : 0;
    p.shapes = d_shapes.as<gf_shape>();

    // roots in Morton order
    const int64_t ncode = (int64_t)P * P * P;
    GF_REQUIRE(ncode < (int64_t)INT32_MAX, "root grid too large");
    DevBuf flag0, pos0, tmp0;
    GF_CHECK_HIP(flag0.alloc(sizeof(int32_t) * ncode));
    GF_CHECK_HIP(pos0.alloc(sizeof(int32_t) * ncode));
    GF_CHECK_HIP(tmp0.alloc(sizeof(int32_t) * scan_tmp_ints(ncode)));
    hipLaunchKernelGGL(k_root_flags, dim3(grid_for(ncode)), dim3(kBlock), 0, s, ncode, P, p, flag0.as<int32_t>());
    scan_rec(flag0.as<int32_t>(), pos0.as<int32_t>(), (int32_t)ncode, tmp0.as<int32_t>(), s);
    int64_t n = (int64_t)root[0] * root[1] * root[2];
    Leaf *cells = nullptr;
    GF_CHECK_HIP(hipMalloc(&cells, sizeof(Leaf) * n));
    hipLaunchKernelGGL(k_root_emit, dim3(grid_for(ncode)), dim3(kBlock), 0, s, ncode, p, flag0.as<int32_t>(),
                       pos0.as<int32_t>(), cells);
    GF_CHECK_LAUNCH();

    auto fail = [&](int rc) {
        if (cells) (void)hipFree(cells);
        return rc;
    };
    // surface and region refinement until no leaf below the target level
    // meets a body's surface and none below a box level intersects its box
    // (each pass raises the finest level by one, so lmax + 1 passes suffice)
    for (int pass = 0;; ++pass) {
        if (pass > lmax + 1) {
            set_error("octree region refinement did not reach a fixed point");
            return fail(GF_EINVAL);
        }
        DevBuf flag, cnt;
        if (flag.alloc(sizeof(int32_t) * n) != hipSuccess || cnt.alloc(sizeof(int32_t)) != hipSuccess)
            return fail(GF_ENOMEM);
        (void)hipMemsetAsync(cnt.p, 0, sizeof(int32_t), s);
        hipLaunchKernelGGL(k_region_flags, dim3(grid_for(n)), dim3(kBlock), 0, s, (int32_t)n, p, cells,
                           flag.as<int32_t>(), cnt.as<int32_t>());
        int64_t n_split = 0;
        int rc = expand(cells, n, flag.as<int32_t>(), lmax, s, n_split);
        if (rc != GF_OK) return fail(rc);
        if (n_split == 0) break;
    }
    // 2:1 face balance to a fixed point (a leaf can be split at most lmax times)
    for (int pass = 0;; ++pass) {
        if (pass > lmax + 1) {
            set_error("octree 2:1 balance did not reach a fixed point");
            return fail(GF_EINVAL);
        }
        DevBuf flag;
        if (flag.alloc(sizeof(int32_t) * n) != hipSuccess) return fail(GF_ENOMEM);
        (void)hipMemsetAsync(flag.p, 0, sizeof(int32_t) * n, s);
        hipLaunchKernelGGL(k_balance_flags, dim3(grid_for(n)), dim3(kBlock), 0, s, (int32_t)n, p, cells,
                           flag.as<int32_t>());
        int64_t n_split = 0;
        int rc = expand(cells, n, flag.as<int32_t>(), lmax, s, n_split);
        if (rc != GF_OK) return fail(rc);
        if (n_split == 0) break;
    }
    // uniform leaf refinement
    for (int r = 0; r < n_leaf; ++r) {
        DevBuf flag;
        if (flag.alloc(sizeof(int32_t) * n) != hipSuccess) return fail(GF_ENOMEM);
        std::vector<int32_t> ones((size_t)n, 1);
        (void)hipMemcpyAsync(flag.p, ones.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s);
        int64_t n_split = 0;
        int rc = expand(cells, n, flag.as<int32_t>(), lmax, s, n_split);
        if (rc != GF_OK) return fail(rc);
    }
    // faces: count, scan, fill
    DevBuf fcnt, bcnt, fpos, bpos, tmp;
    if (fcnt.alloc(sizeof(int32_t) * (n + 1)) != hipSuccess || bcnt.alloc(sizeof(int32_t) * (n + 1)) != hipSuccess ||
        fpos.alloc(sizeof(int32_t) * (n + 1)) != hipSuccess || bpos.alloc(sizeof(int32_t) * (n + 1)) != hipSuccess ||
        tmp.alloc(sizeof(int32_t) * scan_tmp_ints(n + 1)) != hipSuccess)
        return fail(GF_ENOMEM);
    (void)hipMemsetAsync(fcnt.p, 0, sizeof(int32_t) * (n + 1), s);
    (void)hipMemsetAsync(bcnt.p, 0, sizeof(int32_t) * (n + 1), s);
    hipLaunchKernelGGL(k_faces, dim3(grid_for(n)), dim3(kBlock), 0, s, (int32_t)n, p, cells, 0,
                       fcnt.as<int32_t>(), bcnt.as<int32_t>(), nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr);
    scan_rec(fcnt.as<int32_t>(), fpos.as<int32_t>(), (int32_t)(n + 1), tmp.as<int32_t>(), s);
    scan_rec(bcnt.as<int32_t>(), bpos.as<int32_t>(), (int32_t)(n + 1), tmp.as<int32_t>(), s);
    int32_t nf = 0, nbf = 0;
    (void)hipMemcpyAsync(&nf, fpos.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(&nbf, bpos.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(GF_EHIP);
    gf_octree *t = new gf_octree{};
    t->n_cells = n;
    t->n_faces = nf;
    t->n_bnd = nbf;
    t->lmax = lmax;
    t->cells = cells;
    cells = nullptr;
    bool ok = hipMalloc(&t->f_owner, sizeof(int32_t) * (nf + 1)) == hipSuccess &&
              hipMalloc(&t->f_nbr, sizeof(int32_t) * (nf + 1)) == hipSuccess &&
              hipMalloc(&t->f_axis, sizeof(int32_t) * (nf + 1)) == hipSuccess &&
              hipMalloc(&t->b_owner, sizeof(int32_t) * (nbf + 1)) == hipSuccess &&
              hipMalloc(&t->b_dir, sizeof(int32_t) * (nbf + 1)) == hipSuccess;
    if (!ok) {
        gf_octree_free(t);
        set_error("octree face allocation failed");
        return GF_ENOMEM;
    }
    hipLaunchKernelGGL(k_faces, dim3(grid_for(n)), dim3(kBlock), 0, s, (int32_t)n, p, t->cells, 1, nullptr,
                       nullptr, fpos.as<int32_t>(), bpos.as<int32_t>(), t->f_owner, t->f_nbr, t->f_axis,
                       t->b_owner, t->b_dir);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(s));
    counts_out[0] = n;
    counts_out[1] = nf;
    counts_out[2] = nbf;
    counts_out[3] = lmax;
    *out = t;
    return GF_OK;
}

extern "C" int gf_octree_fetch(const gf_octree *t, int32_t *level, int32_t *ijk, int32_t *f_owner, int32_t *f_nbr,
                               int32_t *f_axis, int32_t *b_owner, int32_t *b_dir, gf_stream_t stream) {
    GF_REQUIRE(t && level && ijk, "null argument");
    hipStream_t s = S(stream);
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(t->n_cells)), dim3(kBlock), 0, s, (int32_t)t->n_cells, t->cells,
                       level, ijk);
    GF_CHECK_LAUNCH();
    if (t->n_faces) {
        GF_CHECK_HIP(hipMemcpyAsync(f_owner, t->f_owner, sizeof(int32_t) * t->n_faces, hipMemcpyDeviceToDevice, s));
        GF_CHECK_HIP(hipMemcpyAsync(f_nbr, t->f_nbr, sizeof(int32_t) * t->n_faces, hipMemcpyDeviceToDevice, s));
        GF_CHECK_HIP(hipMemcpyAsync(f_axis, t->f_axis, sizeof(int32_t) * t->n_faces, hipMemcpyDeviceToDevice, s));
    }
    if (t->n_bnd) {
        GF_CHECK_HIP(hipMemcpyAsync(b_owner, t->b_owner, sizeof(int32_t) * t->n_bnd, hipMemcpyDeviceToDevice, s));
        GF_CHECK_HIP(hipMemcpyAsync(b_dir, t->b_dir, sizeof(int32_t) * t->n_bnd, hipMemcpyDeviceToDevice, s));
    }
    GF_CHECK_HIP(hipStreamSynchronize(s));
    return GF_OK;
}

extern "C" void gf_octree_free(gf_octree *t) {
    if (!t) return;
    if (t->cells) (void)hipFree(t->cells);
    if (t->f_owner) (void)hipFree(t->f_owner);
    if (t->f_nbr) (void)hipFree(t->f_nbr);
    if (t->f_axis) (void)hipFree(t->f_axis);
    if (t->b_owner) (void)hipFree(t->b_owner);
    if (t->b_dir) (void)hipFree(t->b_dir);
    delete t;
}
