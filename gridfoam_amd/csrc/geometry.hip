// Grid geometry, face geometry and the cell-centred operator layout.
//
// Restates src/gridfoam/core/grid/axis_projected.py:130-161, :715-811
// (volumes, overlap-box face centres, finer-cell face areas, domain face
// centres/areas) and src/gridfoam/fv/kernels/face_geometry.py:107-262
// (linear weights, delta coefficients, hanging-face detection and offsets,
// geometric least-squares normal matrices).  On top of the reference's
// face lists this file builds the MI355X layout every hot kernel reads:
//   * own_row / nbr_row+nbr_face / bnd_row: per-cell CSR of incident faces,
//     so face-to-cell sums are race-free gathers instead of atomic scatters;
//   * a slice-of-64 ELL operator (one slice per wavefront, column-major
//     inside the slice) whose columns are the face neighbours of each row.
#include <cmath>

#include "gf_common.h"

namespace gf {
namespace {

#pragma clang fp contract(off)

__global__ void k_cell_geom(int C, double3 lo, double3 hi, int3 root, const int32_t *level, const int32_t *ijk,
                            double *centers, double *sizes, double *volumes) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double l2 = ldexp(1.0, level[c]);
    const double h0 = (hi.x - lo.x) / ((double)root.x * l2);
    const double h1 = (hi.y - lo.y) / ((double)root.y * l2);
    const double h2 = (hi.z - lo.z) / ((double)root.z * l2);
    sizes[3 * c + 0] = h0;
    sizes[3 * c + 1] = h1;
    sizes[3 * c + 2] = h2;
    centers[3 * c + 0] = lo.x + ((double)ijk[3 * c + 0] + 0.5) * h0;
    centers[3 * c + 1] = lo.y + ((double)ijk[3 * c + 1] + 0.5) * h1;
    centers[3 * c + 2] = lo.z + ((double)ijk[3 * c + 2] + 0.5) * h2;
    volumes[c] = h0 * h1 * h2;
}

__global__ void k_face_geom(int F, const int32_t *owner, const int32_t *nbr, const int32_t *axis,
                            const double *centers, const double *sizes, double *fc, double *sf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int o = owner[f], n = nbr[f], a = axis[f];
    double mn[3];
    for (int d = 0; d < 3; ++d) {
        const double co = centers[3 * o + d], cn = centers[3 * n + d];
        const double so = sizes[3 * o + d], sn = sizes[3 * n + d];
        const double bot = fmax(co - so / 2.0, cn - sn / 2.0);
        const double top = fmin(co + so / 2.0, cn + sn / 2.0);
        fc[3 * f + d] = (bot + top) / 2.0;
        mn[d] = fmin(so, sn);
    }
    const double area = (mn[0] * mn[1] * mn[2]) / mn[a];
    for (int d = 0; d < 3; ++d) sf[3 * f + d] = d == a ? area : 0.0;
}

__global__ void k_bnd_geom(int B, const int32_t *b_owner, const int32_t *b_dir, const double *centers,
                           const double *sizes, double *bc, double *bsf) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const int o = b_owner[k], dir = b_dir[k], a = dir / 2;
    const double sign = 2.0 * (double)(dir % 2) - 1.0;
    const double *c = centers + 3 * o;
    const double *s = sizes + 3 * o;
    for (int d = 0; d < 3; ++d) {
        const double nrm = d == a ? sign : 0.0;
        bc[3 * k + d] = c[d] + 0.5 * s[d] * nrm;
    }
    const double area = sign * (s[0] * s[1] * s[2]) / s[a];
    for (int d = 0; d < 3; ++d) bsf[3 * k + d] = d == a ? area : 0.0;
}

// face_geometry.py:116-148
__global__ void k_face_weights(const gf_mesh m, const double *fc, double *w, double *mag_sf, double *delta,
                               uint8_t *is_hang, double tol) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f], a = m.axis[f];
    const double *co = m.centers + 3 * o, *cn = m.centers + 3 * n, *cf = fc + 3 * f;
    const double dn = fabs(cn[a] - co[a]);
    const double ww = fabs(cn[a] - cf[a]) / dn;
    w[f] = ww;
    // |Sf| of an axis-aligned face (area recomputed as in k_face_geom)
    double area = 0.0;
    {
        const double *so = m.sizes + 3 * o, *sn = m.sizes + 3 * n;
        double mn0 = fmin(so[0], sn[0]), mn1 = fmin(so[1], sn[1]), mn2 = fmin(so[2], sn[2]);
        const double mna = a == 0 ? mn0 : (a == 1 ? mn1 : mn2);
        area = (mn0 * mn1 * mn2) / mna;
    }
    mag_sf[f] = sqrt(area * area);
    delta[f] = 1.0 / dn;
    double off2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double cen = ww * co[d] + (1.0 - ww) * cn[d];
        const double od = cf[d] - cen;
        off2 += od * od;
    }
    is_hang[f] = sqrt(off2) > tol ? 1 : 0;
}

__global__ void k_row_counts(const gf_mesh m, int32_t *own_cnt, int32_t *nbr_cnt, int32_t *bnd_cnt) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < m.n_faces) {
        atomicAdd(own_cnt + m.owner[f], 1);
        atomicAdd(nbr_cnt + m.neighbour[f], 1);
    }
    if (f < m.n_bnd) atomicAdd(bnd_cnt + m.bnd_owner[f], 1);
}

// Face-order contract of the per-cell face lists (one bit per violated rule):
//   1 an index out of range (cell, axis or direction) or owner == neighbour
//   2 internal faces not grouped by owner in ascending order
//   8 own_row does not bound each face's owner group, or a boundary face is
//     missing from its owner's bnd_row / bnd_face list
//  16 the neighbour does not lie on the +axis side of the owner
// Every face is checked by one thread; flags are OR-ed into one word.
enum : int { kBadIndex = 1, kOwnerOrder = 2, kRowMismatch = 8, kNbrSide = 16 };
__global__ void k_mesh_validate(const gf_mesh m, int32_t *flags) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    if (f < m.n_faces) {
        const int o = m.owner[f], n = m.neighbour[f], a = m.axis[f];
        if (o < 0 || o >= m.n_cells || n < 0 || n >= m.n_cells || o == n || a < 0 || a > 2) {
            bad |= kBadIndex;
        } else {
            if (f > 0 && m.owner[f - 1] > o) bad |= kOwnerOrder;
            if (m.own_row && (f < m.own_row[o] || f >= m.own_row[o + 1])) bad |= kRowMismatch;
            if (m.centers && !(m.centers[3 * n + a] > m.centers[3 * o + a])) bad |= kNbrSide;
        }
    }
    if (f < m.n_bnd) {
        const int o = m.bnd_owner[f], d = m.bnd_dir[f];
        if (o < 0 || o >= m.n_cells || d < 0 || d > 5) {
            bad |= kBadIndex;
        } else if (m.bnd_row && m.bnd_face) {
            bool listed = false;
            for (int k = m.bnd_row[o]; k < m.bnd_row[o + 1]; ++k) listed |= m.bnd_face[k] == f;
            if (!listed) bad |= kRowMismatch;
        }
    }
    if (bad) atomicOr(flags, bad);
}

// boundary faces grouped by owner; order fixed afterwards (ascending face index)
__global__ void k_bnd_scatter(const gf_mesh m, int32_t *cursor, int32_t *bnd_face) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd) return;
    const int c = m.bnd_owner[k];
    const int slot = atomicAdd(cursor + c, 1);
    bnd_face[m.bnd_row[c] + slot] = k;
}
__global__ void k_bnd_sort(const gf_mesh m, int32_t *bnd_face) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    const int b = m.bnd_row[c], e = m.bnd_row[c + 1];
    for (int i = b + 1; i < e; ++i) {
        int32_t x = bnd_face[i];
        int j = i - 1;
        while (j >= b && bnd_face[j] > x) {
            bnd_face[j + 1] = bnd_face[j];
            --j;
        }
        bnd_face[j + 1] = x;
    }
}

// faces seen from their neighbour, grouped by cell; order fixed afterwards
__global__ void k_nbr_scatter(const gf_mesh m, int32_t *cursor, int32_t *nbr_face) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int n = m.neighbour[f];
    const int slot = atomicAdd(cursor + n, 1);
    nbr_face[m.nbr_row[n] + slot] = f;
}
__global__ void k_nbr_sort(const gf_mesh m, int32_t *nbr_face) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    const int b = m.nbr_row[c], e = m.nbr_row[c + 1];
    for (int i = b + 1; i < e; ++i) {
        int32_t x = nbr_face[i];
        int j = i - 1;
        while (j >= b && nbr_face[j] > x) {
            nbr_face[j + 1] = nbr_face[j];
            --j;
        }
        nbr_face[j + 1] = x;
    }
}

__global__ void k_sell_widths(const gf_mesh m, int32_t *slice_len) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m.n_slices) return;
    int w = 0;
    for (int l = 0; l < 64; ++l) {
        const int c = s * 64 + l;
        if (c >= m.n_cells) break;
        const int len = (m.own_row[c + 1] - m.own_row[c]) + (m.nbr_row[c + 1] - m.nbr_row[c]);
        w = len > w ? len : w;
    }
    slice_len[s] = 64 * w;
}

constexpr int kMaxRow = 32;  // 2:1 octree: at most 6 faces x 4 neighbours

__global__ void k_sell_fill(const gf_mesh m, int32_t *sell_col, int32_t *sell_code) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = c >> 6, lane = c & 63;
    if (s >= m.n_slices) return;
    const int base = m.sell_ptr[s];
    const int W = (m.sell_ptr[s + 1] - base) >> 6;
    int32_t col[kMaxRow], code[kMaxRow];
    int n = 0;
    if (c < m.n_cells) {
        for (int f = m.own_row[c]; f < m.own_row[c + 1] && n < kMaxRow; ++f) {
            col[n] = m.neighbour[f];
            code[n] = f;
            ++n;
        }
        for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1] && n < kMaxRow; ++k) {
            const int f = m.nbr_face[k];
            col[n] = m.owner[f];
            code[n] = -f - 1;
            ++n;
        }
        for (int i = 1; i < n; ++i) {  // ascending column
            int32_t xc = col[i], xk = code[i];
            int j = i - 1;
            while (j >= 0 && col[j] > xc) {
                col[j + 1] = col[j];
                code[j + 1] = code[j];
                --j;
            }
            col[j + 1] = xc;
            code[j + 1] = xk;
        }
    }
    const int self = c < m.n_cells ? c : m.n_cells - 1;
    for (int k = 0; k < W; ++k) {
        const int e = base + k * 64 + lane;
        sell_col[e] = k < n ? col[k] : self;
        sell_code[e] = k < n ? code[k] : INT32_MIN;
    }
}

// Symmetric 3x3 pseudo-inverse.  Well-conditioned matrices (every interior
// and boundary cell of an octree) use the adjugate; otherwise cyclic Jacobi
// eigen-decomposition with the cutoff 3*eps*max|lambda| used by
// torch.linalg.pinv (face_geometry.py:262).
__device__ void pinv_sym3(const double (&a)[9], double (&out)[9]) {
    const double c00 = a[4] * a[8] - a[5] * a[7];
    const double c01 = a[5] * a[6] - a[3] * a[8];
    const double c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    const double scale = fmax(fmax(fabs(a[0]), fabs(a[4])), fabs(a[8]));
    if (scale > 0.0 && fabs(det) > 1e-10 * scale * scale * scale) {
        const double inv = 1.0 / det;
        out[0] = c00 * inv;
        out[1] = (a[2] * a[7] - a[1] * a[8]) * inv;
        out[2] = (a[1] * a[5] - a[2] * a[4]) * inv;
        out[3] = c01 * inv;
        out[4] = (a[0] * a[8] - a[2] * a[6]) * inv;
        out[5] = (a[2] * a[3] - a[0] * a[5]) * inv;
        out[6] = c02 * inv;
        out[7] = (a[1] * a[6] - a[0] * a[7]) * inv;
        out[8] = (a[0] * a[4] - a[1] * a[3]) * inv;
        return;
    }
    double A[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = 0.5 * (a[3 * i + j] + a[3 * j + i]);
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (fabs(A[p][q]) < 1e-300) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 3; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = cs * akp - sn * akq;
                    A[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = cs * apk - sn * aqk;
                    A[q][k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = cs * vkp - sn * vkq;
                    V[k][q] = sn * vkp + cs * vkq;
                }
            }
    }
    const double lmax = fmax(fmax(fabs(A[0][0]), fabs(A[1][1])), fabs(A[2][2]));
    const double cut = 3.0 * 2.220446049250313e-16 * lmax;
    double il[3];
    for (int i = 0; i < 3; ++i) il[i] = fabs(A[i][i]) > cut ? 1.0 / A[i][i] : 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += V[i][k] * il[k] * V[j][k];
            out[3 * i + j] = acc;
        }
}

__device__ __forceinline__ void add_outer(double (&a)[9], double d0, double d1, double d2) {
    const double d2s = d0 * d0 + d1 * d1 + d2 * d2;
    const double w2 = 1.0 / (d2s > 0.0 ? d2s : 1.0);
    const double d[3] = {d0, d1, d2};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[3 * i + j] += w2 * d[i] * d[j];
}

// face_geometry.py:245-262: sum w^2 d d^T over single-sided faces (both
// sides) and every domain face, then pinv.
__global__ void k_ata_inv(const gf_mesh m, double *out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double *cc = m.centers + 3 * c;
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const double *cn = m.centers + 3 * m.neighbour[f];
        add_outer(a, cn[0] - cc[0], cn[1] - cc[1], cn[2] - cc[2]);
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const double *co = m.centers + 3 * m.owner[m.nbr_face[k]];
        add_outer(a, cc[0] - co[0], cc[1] - co[1], cc[2] - co[2]);
    }
    GF_FOR_BND(m, c, k)
        add_outer(a, m.bnd_d[3 * k + 0], m.bnd_d[3 * k + 1], m.bnd_d[3 * k + 2]);
    double inv[9];
    pinv_sym3(a, inv);
    for (int i = 0; i < 9; ++i) out[9 * c + i] = inv[i];
}

// face_geometry.py:141-155
__global__ void k_hang_geom(const gf_mesh m, const double *fc, double *offset, double *move_o, double *move_n) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= m.n_hang) return;
    const int f = m.hang_face[h];
    const int o = m.owner[f], n = m.neighbour[f], a = m.axis[f];
    const double ww = m.w[f];
    for (int d = 0; d < 3; ++d) {
        const double co = m.centers[3 * o + d], cn = m.centers[3 * n + d], cf = fc[3 * f + d];
        offset[3 * h + d] = cf - (ww * co + (1.0 - ww) * cn);
        move_o[3 * h + d] = d == a ? 0.0 : cf - co;
        move_n[3 * h + d] = d == a ? 0.0 : cf - cn;
    }
}

// hang-cell CSR: entries h (cell is the owner of hanging face h) or -h-1
__global__ void k_hc_fill(const gf_mesh m, const int32_t *cell_to_hc, int32_t *cursor, int32_t *hc_code) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= m.n_hang) return;
    const int f = m.hang_face[h];
    const int ho = cell_to_hc[m.owner[f]], hn = cell_to_hc[m.neighbour[f]];
    int so = atomicAdd(cursor + ho, 1);
    hc_code[m.hc_row[ho] + so] = h;
    int sn = atomicAdd(cursor + hn, 1);
    hc_code[m.hc_row[hn] + sn] = -h - 1;
}
__global__ void k_hc_sort(const gf_mesh m, int32_t *hc_code) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_hang_cells) return;
    const int b = m.hc_row[i], e = m.hc_row[i + 1];
    for (int p = b + 1; p < e; ++p) {
        int32_t x = hc_code[p];
        int q = p - 1;
        while (q >= b && hc_code[q] > x) {
            hc_code[q + 1] = hc_code[q];
            --q;
        }
        hc_code[q + 1] = x;
    }
}

}  // namespace
}  // namespace gf

using namespace gf;

extern "C" int gf_grid_geometry(int32_t n_cells, int32_t n_faces, int32_t n_bnd, const double lower[3],
                                const double upper[3], const int32_t root[3], const int32_t *level,
                                const int32_t *ijk, const int32_t *owner, const int32_t *nbr, const int32_t *axis,
                                const int32_t *b_owner, const int32_t *b_dir, double *centers, double *sizes,
                                double *volumes, double *face_centers, double *sf, double *bnd_centers,
                                double *bnd_sf, gf_stream_t stream) {
    GF_REQUIRE(lower && upper && root && level && ijk && centers && sizes && volumes, "null argument");
    hipStream_t s = S(stream);
    double3 lo = make_double3(lower[0], lower[1], lower[2]);
    double3 hi = make_double3(upper[0], upper[1], upper[2]);
    int3 rt = make_int3(root[0], root[1], root[2]);
    hipLaunchKernelGGL(k_cell_geom, dim3(grid_for(n_cells)), dim3(kBlock), 0, s, n_cells, lo, hi, rt, level, ijk,
                       centers, sizes, volumes);
    if (n_faces > 0)
        hipLaunchKernelGGL(k_face_geom, dim3(grid_for(n_faces)), dim3(kBlock), 0, s, n_faces, owner, nbr, axis,
                           centers, sizes, face_centers, sf);
    if (n_bnd > 0)
        hipLaunchKernelGGL(k_bnd_geom, dim3(grid_for(n_bnd)), dim3(kBlock), 0, s, n_bnd, b_owner, b_dir, centers,
                           sizes, bnd_centers, bnd_sf);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_face_geometry(const gf_mesh *m, const double *face_centers, double *w, double *mag_sf,
                                double *delta, uint8_t *is_hang, double h_min, gf_stream_t stream) {
    GF_REQUIRE(m && face_centers && w && mag_sf && delta && is_hang, "null argument");
    if (m->n_faces == 0) return GF_OK;
    hipLaunchKernelGGL(k_face_weights, dim3(grid_for(m->n_faces)), dim3(kBlock), 0, S(stream), *m, face_centers,
                       w, mag_sf, delta, is_hang, 1.0e-8 * h_min);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_row_counts(const gf_mesh *m, int32_t *own_cnt, int32_t *nbr_cnt, int32_t *bnd_cnt,
                             gf_stream_t stream) {
    GF_REQUIRE(m && own_cnt && nbr_cnt && bnd_cnt, "null argument");
    const int n = m->n_faces > m->n_bnd ? m->n_faces : m->n_bnd;
    if (n == 0) return GF_OK;
    hipLaunchKernelGGL(k_row_counts, dim3(grid_for(n)), dim3(kBlock), 0, S(stream), *m, own_cnt, nbr_cnt,
                       bnd_cnt);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_mesh_validate(const gf_mesh *m, int32_t *flags_out, gf_stream_t stream) {
    GF_REQUIRE(m && flags_out, "null argument");
    GF_REQUIRE(m->n_cells >= 1 && m->n_faces >= 0 && m->n_bnd >= 0, "bad mesh sizes");
    GF_REQUIRE(m->owner && m->neighbour && m->axis && m->bnd_owner && m->bnd_dir, "mesh topology missing");
    hipStream_t s = S(stream);
    int32_t *flags = nullptr;
    GF_CHECK_HIP(hipMalloc(&flags, sizeof(int32_t)));
    int32_t h = 0;
    hipError_t e = hipMemsetAsync(flags, 0, sizeof(int32_t), s);
    const int n = m->n_faces > m->n_bnd ? m->n_faces : m->n_bnd;
    if (e == hipSuccess && n > 0) {
        hipLaunchKernelGGL(k_mesh_validate, dim3(grid_for(n)), dim3(kBlock), 0, s, *m, flags);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const hipError_t fe = hipFree(flags);
    if (e == hipSuccess) e = fe;
    if (e != hipSuccess) {
        set_error(std::string("gf_mesh_validate: ") + hipGetErrorString(e));
        return GF_EHIP;
    }
    *flags_out = h;
    if (h) {
        std::string msg = "gf_mesh violates the face-order contract:";
        if (h & kBadIndex) msg += " index out of range or owner == neighbour;";
        if (h & kOwnerOrder) msg += " internal faces not grouped by owner in ascending order;";
        if (h & kRowMismatch) msg += " own_row/bnd_row/bnd_face do not match the face owners;";
        if (h & kNbrSide) msg += " neighbour not on the +axis side of the owner;";
        set_error(msg);
        return GF_EINVAL;
    }
    return GF_OK;
}

extern "C" int gf_fill_nbr_lists(const gf_mesh *m, int32_t *nbr_face, gf_stream_t stream) {
    GF_REQUIRE(m && nbr_face, "null argument");
    hipStream_t s = S(stream);
    int32_t *cursor = nullptr;
    GF_CHECK_HIP(hipMalloc(&cursor, sizeof(int32_t) * (m->n_cells + 1)));
    GF_CHECK_HIP(hipMemsetAsync(cursor, 0, sizeof(int32_t) * (m->n_cells + 1), s));
    if (m->n_faces > 0)
        hipLaunchKernelGGL(k_nbr_scatter, dim3(grid_for(m->n_faces)), dim3(kBlock), 0, s, *m, cursor, nbr_face);
    hipLaunchKernelGGL(k_nbr_sort, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, s, *m, nbr_face);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(s));
    GF_CHECK_HIP(hipFree(cursor));
    return GF_OK;
}

extern "C" int gf_fill_bnd_lists(const gf_mesh *m, int32_t *bnd_face, gf_stream_t stream) {
    GF_REQUIRE(m && bnd_face, "null argument");
    hipStream_t s = S(stream);
    int32_t *cursor = nullptr;
    GF_CHECK_HIP(hipMalloc(&cursor, sizeof(int32_t) * (m->n_cells + 1)));
    GF_CHECK_HIP(hipMemsetAsync(cursor, 0, sizeof(int32_t) * (m->n_cells + 1), s));
    if (m->n_bnd > 0)
        hipLaunchKernelGGL(k_bnd_scatter, dim3(grid_for(m->n_bnd)), dim3(kBlock), 0, s, *m, cursor, bnd_face);
    hipLaunchKernelGGL(k_bnd_sort, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, s, *m, bnd_face);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(s));
    GF_CHECK_HIP(hipFree(cursor));
    return GF_OK;
}

extern "C" int gf_sell_widths(const gf_mesh *m, int32_t *slice_len, gf_stream_t stream) {
    GF_REQUIRE(m && slice_len, "null argument");
    hipLaunchKernelGGL(k_sell_widths, dim3(grid_for(m->n_slices)), dim3(kBlock), 0, S(stream), *m, slice_len);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_sell_fill(const gf_mesh *m, int32_t *sell_col, int32_t *sell_code, gf_stream_t stream) {
    GF_REQUIRE(m && sell_col && sell_code, "null argument");
    hipLaunchKernelGGL(k_sell_fill, dim3(grid_for((int64_t)m->n_slices * 64)), dim3(kBlock), 0, S(stream), *m,
                       sell_col, sell_code);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_lsq_ata_inv(const gf_mesh *m, double *ata_inv, gf_stream_t stream) {
    GF_REQUIRE(m && ata_inv, "null argument");
    hipLaunchKernelGGL(k_ata_inv, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, S(stream), *m, ata_inv);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_hang_geometry(const gf_mesh *m, const double *face_centers, double *offset, double *move_o,
                                double *move_n, gf_stream_t stream) {
    GF_REQUIRE(m && face_centers, "null argument");
    if (m->n_hang == 0) return GF_OK;
    hipLaunchKernelGGL(k_hang_geom, dim3(grid_for(m->n_hang)), dim3(kBlock), 0, S(stream), *m, face_centers,
                       offset, move_o, move_n);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_hang_cell_lists(const gf_mesh *m, const int32_t *cell_to_hc, int32_t *hc_code,
                                  gf_stream_t stream) {
    GF_REQUIRE(m && cell_to_hc && hc_code, "null argument");
    if (m->n_hang == 0) return GF_OK;
    hipStream_t s = S(stream);
    int32_t *cursor = nullptr;
    GF_CHECK_HIP(hipMalloc(&cursor, sizeof(int32_t) * (m->n_hang_cells + 1)));
    GF_CHECK_HIP(hipMemsetAsync(cursor, 0, sizeof(int32_t) * (m->n_hang_cells + 1), s));
    hipLaunchKernelGGL(k_hc_fill, dim3(grid_for(m->n_hang)), dim3(kBlock), 0, s, *m, cell_to_hc, cursor, hc_code);
    hipLaunchKernelGGL(k_hc_sort, dim3(grid_for(m->n_hang_cells)), dim3(kBlock), 0, s, *m, hc_code);
    GF_CHECK_LAUNCH();
    GF_CHECK_HIP(hipStreamSynchronize(s));
    GF_CHECK_HIP(hipFree(cursor));
    return GF_OK;
}
