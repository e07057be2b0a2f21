// Finite-volume operators of the pressure-velocity path (FP64).
//
// Every face-to-cell sum is a gather over the per-cell CSR of incident
// faces (own_row / nbr_row+nbr_face / bnd_row), so there are no atomic
// scatters and the results are run-to-run deterministic.  The reference
// builds the same sums with torch index_add_ (FP64 atomics on GPU).
// Partial sums are grouped as the reference groups them (owner-side sum,
// then neighbour-side sum, then boundary terms).
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "gf_common.h"

namespace gf {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

namespace {

// ---------------------------------------------------------------- boundary state
// boundary_ops.py:311-326 (covered faces) and :545-593 (zero-gradient fill)
__global__ void k_bc_state(const gf_mesh m, int nc, const double *psi, const double *frac, const double *rv,
                           const double *rg, const uint8_t *covered, double *psi_b) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd) return;
    const int c = m.bnd_owner[k];
    if (!covered[k]) {
        for (int q = 0; q < nc; ++q) psi_b[k * nc + q] = psi[c * nc + q];
        return;
    }
    const double f = frac[k], d = m.bnd_mag_d[k];
    for (int q = 0; q < nc; ++q) psi_b[k * nc + q] = f * rv[k * nc + q] + (1.0 - f) * (psi[c * nc + q] + d * rg[k * nc + q]);
}

// slip.py:79-101: remove the wall-normal component of the cell velocity
__global__ void k_bc_slip(const gf_mesh m, const double *psi, const uint8_t *mask, double *rv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd || !mask[k]) return;
    const int c = m.bnd_owner[k], a = m.bnd_dir[k] / 2;
    const double sgn = m.bnd_sign[k];
    double n[3] = {0.0, 0.0, 0.0};
    n[a] = sgn;
    const double *u = psi + 3 * c;
    const double un = (u[0] * n[0] + u[1] * n[1]) + u[2] * n[2];
    for (int d = 0; d < 3; ++d) rv[3 * k + d] = u[d] - un * n[d];
}

// ---------------------------------------------------------------- least squares
constexpr int kMaxNc = 3;

__device__ __forceinline__ void lsq_wd(double d0, double d1, double d2, double (&wd)[3]) {
    const double d2s = (d0 * d0 + d1 * d1) + d2 * d2;
    const double w2 = 1.0 / (d2s > 0.0 ? d2s : 1.0);
    wd[0] = w2 * d0;
    wd[1] = w2 * d1;
    wd[2] = w2 * d2;
}

template <int NC>
__global__ void k_lsq_grad(const gf_mesh m, const double *psi, const double *psi_b, const uint8_t *covered,
                           const int32_t *cell_list, int n_list, double *grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_list) return;
    const int c = cell_list ? cell_list[i] : i;
    const double *cc = m.centers + 3 * c;
    double ao[NC][3] = {}, an[NC][3] = {}, ab[NC][3] = {};
    double pc[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) pc[q] = psi[c * NC + q];
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const int n = m.neighbour[f];
        const double *cn = m.centers + 3 * n;
        double wd[3];
        lsq_wd(cn[0] - cc[0], cn[1] - cc[1], cn[2] - cc[2], wd);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double dp = psi[n * NC + q] - pc[q];
#pragma unroll
            for (int j = 0; j < 3; ++j) ao[q][j] += dp * wd[j];
        }
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int o = m.owner[m.nbr_face[k]];
        const double *co = m.centers + 3 * o;
        double wd[3];
        lsq_wd(cc[0] - co[0], cc[1] - co[1], cc[2] - co[2], wd);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double dp = pc[q] - psi[o * NC + q];
#pragma unroll
            for (int j = 0; j < 3; ++j) an[q][j] += dp * wd[j];
        }
    }
    GF_FOR_BND(m, c, k) {
        if (!covered[k]) continue;
        double wd[3];
        lsq_wd(m.bnd_d[3 * k + 0], m.bnd_d[3 * k + 1], m.bnd_d[3 * k + 2], wd);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double dp = psi_b[k * NC + q] - pc[q];
#pragma unroll
            for (int j = 0; j < 3; ++j) ab[q][j] += dp * wd[j];
        }
    }
    const double *inv = m.ata_inv + 9 * c;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        double atb[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) atb[j] = (ao[q][j] + an[q][j]) + ab[q][j];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            grad[(c * NC + q) * 3 + r] = (inv[3 * r + 0] * atb[0] + inv[3 * r + 1] * atb[1]) + inv[3 * r + 2] * atb[2];
    }
}

// Reverse mode of k_lsq_grad: with v_c = A_c^T gbar_c (A = ata_inv) on the
// cells the forward evaluated, face f = (o, n) adds wd_f . (v_o + v_n) to
// psi_bar_n and subtracts it from psi_bar_o (both cells used wd_f (psi_n -
// psi_o)); a covered boundary face k of cell c gets psi_b_bar = wd_k . v_c,
// and c loses the same.  Gathered per cell (no atomics).
template <int NC>
__global__ void k_lsq_v(const gf_mesh m, const double *gbar, const int32_t *cell_list, int n_list, double *v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_list) return;
    const int c = cell_list ? cell_list[i] : i;
    const double *inv = m.ata_inv + 9 * c;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const double *g = gbar + (c * NC + q) * 3;
#pragma unroll
        for (int j = 0; j < 3; ++j) v[(c * NC + q) * 3 + j] = (inv[j] * g[0] + inv[3 + j] * g[1]) + inv[6 + j] * g[2];
    }
}

template <int NC>
__global__ void k_lsq_vjp(const gf_mesh m, const double *v, const uint8_t *covered, double *g_psi, double *g_psi_b) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    const double *cc = m.centers + 3 * c;
    double acc[NC] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const int n = m.neighbour[f];
        const double *cn = m.centers + 3 * n;
        double wd[3];
        lsq_wd(cn[0] - cc[0], cn[1] - cc[1], cn[2] - cc[2], wd);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double *vc = v + (c * NC + q) * 3, *vn = v + (n * NC + q) * 3;
            acc[q] -= (wd[0] * (vc[0] + vn[0]) + wd[1] * (vc[1] + vn[1])) + wd[2] * (vc[2] + vn[2]);
        }
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int o = m.owner[m.nbr_face[k]];
        const double *co = m.centers + 3 * o;
        double wd[3];
        lsq_wd(cc[0] - co[0], cc[1] - co[1], cc[2] - co[2], wd);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double *vc = v + (c * NC + q) * 3, *vo = v + (o * NC + q) * 3;
            acc[q] += (wd[0] * (vo[0] + vc[0]) + wd[1] * (vo[1] + vc[1])) + wd[2] * (vo[2] + vc[2]);
        }
    }
    GF_FOR_BND(m, c, k) {
        double wd[3];
        lsq_wd(m.bnd_d[3 * k + 0], m.bnd_d[3 * k + 1], m.bnd_d[3 * k + 2], wd);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double *vc = v + (c * NC + q) * 3;
            const double t = covered[k] ? (wd[0] * vc[0] + wd[1] * vc[1]) + wd[2] * vc[2] : 0.0;
            acc[q] -= t;
            g_psi_b[k * NC + q] = t;
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) g_psi[c * NC + q] = acc[q];
}

// Reverse mode of k_bc_state (and of k_bc_slip on slip faces), per cell over
// its boundary faces: psi_b = f ref_v + (1 - f)(psi_P + |d| ref_g) where
// covered, psi_P elsewhere; slip ref_v = psi_P - (psi_P . n) n.  Adds to
// g_psi and to the reference-value gradients g_ref_v / g_ref_g (which may
// hold other consumers' shares); on slip faces the whole ref_v gradient is
// carried on to psi_P and its entry is left 0.
template <int NC>
__global__ void k_bc_state_vjp(const gf_mesh m, const double *frac, const uint8_t *covered, const uint8_t *slip,
                               const double *g_psi_b, double *g_psi, double *g_ref_v, double *g_ref_g) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double acc[NC] = {};
    GF_FOR_BND(m, c, k) {
        double rv[NC], rg[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            rv[q] = g_ref_v ? g_ref_v[k * NC + q] : 0.0;
            rg[q] = g_ref_g ? g_ref_g[k * NC + q] : 0.0;
        }
        if (!covered[k]) {
#pragma unroll
            for (int q = 0; q < NC; ++q) acc[q] += g_psi_b[k * NC + q];
        } else {
            const double f = frac[k], d = m.bnd_mag_d[k];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const double gb = g_psi_b[k * NC + q];
                acc[q] += (1.0 - f) * gb;
                rv[q] += f * gb;
                rg[q] += (1.0 - f) * d * gb;
            }
            if constexpr (NC == 3) {
                if (slip && slip[k]) {  // ref_v is the tangential part of psi_P
                    const int a = m.bnd_dir[k] / 2;
                    double n[3] = {0.0, 0.0, 0.0};
                    n[a] = m.bnd_sign[k];
                    const double rn = (rv[0] * n[0] + rv[1] * n[1]) + rv[2] * n[2];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        acc[q] += rv[q] - rn * n[q];
                        rv[q] = 0.0;
                    }
                }
            }
        }
        if (g_ref_v)
#pragma unroll
            for (int q = 0; q < NC; ++q) g_ref_v[k * NC + q] = rv[q];
        if (g_ref_g)
#pragma unroll
            for (int q = 0; q < NC; ++q) g_ref_g[k * NC + q] = rg[q];
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) g_psi[c * NC + q] += acc[q];
}

// ---------------------------------------------------------------- interpolation
template <int NC>
__global__ void k_interp_linear(const gf_mesh m, const double *psi, double *psi_f) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f];
    const double w = m.w[f];
#pragma unroll
    for (int q = 0; q < NC; ++q) psi_f[f * NC + q] = w * psi[o * NC + q] + (1.0 - w) * psi[n * NC + q];
}

template <int NC>
__device__ __forceinline__ double hang_interp_corr(const gf_mesh &m, const double *grad, int h, int q) {
    const int f = m.hang_face[h];
    const int o = m.owner[f], n = m.neighbour[f];
    const double w = m.w[f];
    const double *go = grad + (o * NC + q) * 3, *gn = grad + (n * NC + q) * 3, *off = m.hang_offset + 3 * h;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) acc += (w * go[j] + (1.0 - w) * gn[j]) * off[j];
    return acc;
}

template <int NC>
__device__ __forceinline__ double hang_sn_corr(const gf_mesh &m, const double *grad, int h, int q) {
    const int f = m.hang_face[h];
    const int o = m.owner[f], n = m.neighbour[f];
    const double *go = grad + (o * NC + q) * 3, *gn = grad + (n * NC + q) * 3;
    const double *mo = m.hang_move_o + 3 * h, *mn = m.hang_move_n + 3 * h;
    const double cn = (gn[0] * mn[0] + gn[1] * mn[1]) + gn[2] * mn[2];
    const double co = (go[0] * mo[0] + go[1] * mo[1]) + go[2] * mo[2];
    return m.delta[f] * (cn - co);
}

template <int NC>
__global__ void k_interp_hang(const gf_mesh m, const double *grad, double *psi_f) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= m.n_hang) return;
    const int f = m.hang_face[h];
#pragma unroll
    for (int q = 0; q < NC; ++q) psi_f[f * NC + q] += hang_interp_corr<NC>(m, grad, h, q);
}

template <int NC>
__global__ void k_sn_grad_hang(const gf_mesh m, const double *grad, double *corr) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= m.n_hang) return;
#pragma unroll
    for (int q = 0; q < NC; ++q) corr[h * NC + q] = hang_sn_corr<NC>(m, grad, h, q);
}

// Green-Gauss cell gradient (fv/kernels/gauss_gradient.py:15-73): the sum
// over a cell's faces of psi_f Sf (owner +, neighbour -, boundary +) over
// its volume, gathered per cell (no atomics).  Sf lies along the face axis.
template <int NC>
__global__ void k_gauss_grad(const gf_mesh m, const double *psi_s, const double *psi_b, double *grad) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double acc[NC][3] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const int a = m.axis[f];
        const double sf = m.mag_sf[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q][a] += psi_s[f * NC + q] * sf;
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
        const int a = m.axis[f];
        const double sf = m.mag_sf[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q][a] -= psi_s[f * NC + q] * sf;
    }
    GF_FOR_BND(m, c, k) {
        const int a = m.bnd_dir[k] / 2;
        const double sf = m.bnd_sign[k] * m.bnd_mag_sf[k];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q][a] += psi_b[k * NC + q] * sf;
    }
    const double V = m.volumes[c];
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) grad[(c * NC + q) * 3 + j] = acc[q][j] / V;
}

// Reverse mode of k_gauss_grad: face value f enters its owner's gradient
// component axis[f] with +|Sf|/V_O and its neighbour's with -|Sf|/V_N; a
// boundary value enters its cell's with sign |Sf|/V (overwrites).
template <int NC>
__global__ void k_gauss_grad_vjp_faces(const gf_mesh m, const double *g_grad, double *g_psi_s) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f], a = m.axis[f];
    const double sf = m.mag_sf[f], vo = m.volumes[o], vn = m.volumes[n];
#pragma unroll
    for (int q = 0; q < NC; ++q)
        g_psi_s[f * NC + q] = sf * (g_grad[(o * NC + q) * 3 + a] / vo - g_grad[(n * NC + q) * 3 + a] / vn);
}

template <int NC>
__global__ void k_gauss_grad_vjp_bnd(const gf_mesh m, const double *g_grad, double *g_psi_b) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd) return;
    const int c = m.bnd_owner[k], a = m.bnd_dir[k] / 2;
    const double sf = m.bnd_sign[k] * m.bnd_mag_sf[k], v = m.volumes[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) g_psi_b[k * NC + q] = sf * g_grad[(c * NC + q) * 3 + a] / v;
}

// fvc.reconstruct (fvc/reconstruct.py:27-107): U_P = (sum |Sf| n n^T)^-1
// sum n phi_f over all faces of the cell, both sides of an internal face
// with the face's own normal; diagonal for axis-aligned faces.
__global__ void k_reconstruct(const gf_mesh m, const double *phi_s, const double *phi_b, double *U) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double num[3] = {}, den[3] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const int a = m.axis[f];
        num[a] += phi_s[f];
        den[a] += m.mag_sf[f];
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
        const int a = m.axis[f];
        num[a] += phi_s[f];
        den[a] += m.mag_sf[f];
    }
    GF_FOR_BND(m, c, k) {
        const int a = m.bnd_dir[k] / 2;
        num[a] += m.bnd_sign[k] * phi_b[k];
        den[a] += m.bnd_mag_sf[k];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) U[3 * c + j] = num[j] / den[j];
}

// Reverse mode of k_reconstruct: per cell, s[c][a] = g_U[c][a] / den_c[a]
// (the same area sums as the forward); then every face value gathers s of
// the cells it enters (both sides of an internal face with +, a boundary
// face with its sign).
__global__ void k_reconstruct_vjp_cells(const gf_mesh m, const double *g_U, double *s) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double den[3] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) den[m.axis[f]] += m.mag_sf[f];
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
        den[m.axis[f]] += m.mag_sf[f];
    }
    GF_FOR_BND(m, c, k) den[m.bnd_dir[k] / 2] += m.bnd_mag_sf[k];
#pragma unroll
    for (int j = 0; j < 3; ++j) s[3 * c + j] = g_U[3 * c + j] / den[j];
}

__global__ void k_reconstruct_vjp_faces(const gf_mesh m, const double *s, double *g_phi_s, double *g_phi_b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.n_faces) {
        const int a = m.axis[i];
        g_phi_s[i] = s[3 * m.owner[i] + a] + s[3 * m.neighbour[i] + a];
    } else if (i < m.n_faces + m.n_bnd) {
        const int k = i - m.n_faces;
        g_phi_b[k] = m.bnd_sign[k] * s[3 * m.bnd_owner[k] + m.bnd_dir[k] / 2];
    }
}

// ---------------------------------------------------------------- divergence
template <int NC>
__global__ void k_fvc_div(const gf_mesh m, const double *phi_s, const double *phi_b, double *out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double so[NC] = {}, sn[NC] = {}, sb[NC] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f)
#pragma unroll
        for (int q = 0; q < NC; ++q) so[q] += phi_s[f * NC + q];
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
#pragma unroll
        for (int q = 0; q < NC; ++q) sn[q] += -phi_s[f * NC + q];
    }
    GF_FOR_BND(m, c, k)
#pragma unroll
        for (int q = 0; q < NC; ++q) sb[q] += phi_b[k * NC + q];
    const double V = m.volumes[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) out[c * NC + q] = ((so[q] + sn[q]) + sb[q]) / V;
}

// Reverse mode of k_fvc_div: a face flux enters its owner's sum with + and
// its neighbour's with -, each divided by that cell's volume.
template <int NC>
__global__ void k_fvc_div_vjp_faces(const gf_mesh m, const double *g_out, double *g_phi_s) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f];
    const double vo = m.volumes[o], vn = m.volumes[n];
#pragma unroll
    for (int q = 0; q < NC; ++q) g_phi_s[f * NC + q] = g_out[o * NC + q] / vo - g_out[n * NC + q] / vn;
}

template <int NC>
__global__ void k_fvc_div_vjp_bnd(const gf_mesh m, const double *g_out, double *g_phi_b) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd) return;
    const int c = m.bnd_owner[k];
    const double v = m.volumes[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) g_phi_b[k * NC + q] = g_out[c * NC + q] / v;
}

// Reverse mode of k_interp_linear: each cell gathers w g_f over the faces it
// owns and (1 - w) g_f over the faces it neighbours, added to g_psi.
template <int NC>
__global__ void k_interp_vjp_cells(const gf_mesh m, const double *g_f, double *g_psi) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double acc[NC] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const double w = m.w[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] += w * g_f[f * NC + q];
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
        const double w1 = 1.0 - m.w[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] += w1 * g_f[f * NC + q];
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) g_psi[c * NC + q] += acc[q];
}

// Reverse mode of k_interp_hang: the correction (w grad_O + (1 - w) grad_N)
// . offset of hanging face h sends w offset g_f to the owner's gradient and
// (1 - w) offset g_f to the neighbour's; one thread per hang cell gathers its
// hanging faces (hc_code h: owner side, -h-1: neighbour side) and adds.
template <int NC>
__global__ void k_interp_vjp_hang(const gf_mesh m, const double *g_f, double *g_grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_hang_cells) return;
    const int c = m.hang_cell[i];
    double acc[NC][3] = {};
    for (int k = m.hc_row[i]; k < m.hc_row[i + 1]; ++k) {
        const int code = m.hc_code[k];
        const int h = code >= 0 ? code : -code - 1;
        const int f = m.hang_face[h];
        const double s = code >= 0 ? m.w[f] : 1.0 - m.w[f];
        const double *off = m.hang_offset + 3 * h;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double gs = s * g_f[f * NC + q];
#pragma unroll
            for (int j = 0; j < 3; ++j) acc[q][j] += gs * off[j];
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) g_grad[(c * NC + q) * 3 + j] += acc[q][j];
}

// ---------------------------------------------------------------- assembly
// Face diffusivity (laplacian.py:38-73): linear interpolation, or the
// distance-weighted harmonic mean 1 / ((1-w)/gO + w/gN) evaluated scaled by
// max(gO, gN) (no overflow; a zero side blocks the face with a finite value).
__device__ __forceinline__ double gamma_face(const gf_assembly &a, const gf_mesh &m, int f) {
    if (!a.gamma_cell) return a.gamma;
    const double w = m.w[f];
    const double go = a.gamma_cell[m.owner[f]], gn = a.gamma_cell[m.neighbour[f]];
    if (!a.lap_harmonic) return w * go + (1.0 - w) * gn;
    const double scale = fmax(go, gn);
    const double safe = scale > 0.0 ? scale : 1.0;
    const double den = w * (go / safe) + (1.0 - w) * (gn / safe);
    return fmin(go, gn) / (den > 0.0 ? den : 1.0);
}

// TVD convection (schemes/div.py:145-350): the upwind matrix plus, per face,
// the deferred-correction source -phi psi(r) (psi_linear - psi_upwind).
// Scheme codes 2..7: vanLeer, minmod, superbee, monotonized central,
// limitedLinear, Koren.
__device__ __forceinline__ double tvd_limiter(int scheme, double r) {
    switch (scheme) {
        case 2: return (r + fabs(r)) / (1.0 + fabs(r));
        case 3: return fmax(fmin(1.0, r), 0.0);
        case 4: return fmax(fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0)), 0.0);
        case 5: return fmax(fmin(fmin(2.0 * r, 0.5 * r + 0.5), 2.0), 0.0);
        case 6: return fmax(fmin(2.0 * r, 1.0), 0.0);
        default: return fmax(fmin(fmin(2.0 * r, (2.0 + r) / 3.0), 2.0), 0.0);
    }
}

__device__ __forceinline__ double sgn(double v) { return (double)((v > 0.0) - (v < 0.0)); }

// phi psi(r) of face f, with its linear (uncorrected) and upwind values.
// r = 2 (grad_U . d) / (psi_N - psi_O) - 1 from the upwind cell's gradient
// (vectors: component sums, NVDVTVDV); |num| >= 1000 |den| is steep.
template <int NC>
__device__ __forceinline__ double tvd_weight(const gf_mesh &m, const gf_assembly &a, int f, double (&base)[NC],
                                             double (&up)[NC]) {
    const int o = m.owner[f], n = m.neighbour[f];
    const double ph = a.phi_s[f], w = m.w[f];
    const bool own = ph > 0.0;
    const int u = own ? o : n;
    double d[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) d[j] = m.centers[3 * n + j] - m.centers[3 * o + j];
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const double po = a.psi[o * NC + q], pn = a.psi[n * NC + q];
        base[q] = w * po + (1.0 - w) * pn;
        up[q] = own ? po : pn;
        const double *gu = a.grad + (u * NC + q) * 3;
        const double gd = (gu[0] * d[0] + gu[1] * d[1]) + gu[2] * d[2];
        const double dp = pn - po;
        if (NC == 1) {
            num = gd;
            den = dp;
        } else {
            num += dp * gd;
            den += dp * dp;
        }
    }
    const bool steep = fabs(num) >= 1000.0 * fabs(den);
    const double r = steep ? 2.0 * 1000.0 * sgn(num) * sgn(den) - 1.0 : 2.0 * num / den - 1.0;
    return ph * tvd_limiter(a.div_scheme, r);
}

template <int NC>
__global__ void k_asm_faces(const gf_mesh m, const gf_assembly a, double *upper, double *lower, double *ffc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    double up = 0.0, lo = 0.0;
    if (a.s_div != 0.0) {
        const double ph = a.phi_s[f], w = m.w[f];
        double ud, ld;
        if (a.div_scheme != 1) {
            ud = fmin(ph, 0.0);
            ld = -fmax(ph, 0.0);
        } else {
            ud = ph * (1.0 - w);
            ld = -ph * w;
        }
        up = a.s_div * ud;
        lo = a.s_div * ld;
    }
    if (a.s_lap != 0.0) {
        const double coeff = gamma_face(a, m, f) * m.mag_sf[f] * m.delta[f];
        up = up + a.s_lap * coeff;
        lo = lo + a.s_lap * coeff;
    }
    upper[f] = up;
    lower[f] = lo;
    if (ffc)
#pragma unroll
        for (int q = 0; q < NC; ++q) ffc[f * NC + q] = 0.0;
}

template <int NC>
__global__ void k_asm_cells(const gf_mesh m, const gf_assembly a, double *diag, double *source) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    const double V = m.volumes[c];
    double dg = 0.0;
    double src[NC] = {};
    if (a.s_ddt != 0.0) {
        const double vdt = V / a.dt;
        dg = a.s_ddt * (a.ddt_a * vdt);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            double hist = a.ddt_b * a.psi_old[c * NC + q];
            if (a.ddt_c != 0.0) hist = hist - a.ddt_c * a.psi_older[c * NC + q];
            src[q] = a.s_ddt * (vdt * hist);
        }
    }
    if (a.s_div != 0.0) {
        double so = 0.0, sn = 0.0, sb = 0.0;
        const bool tvd = a.div_scheme >= 2;
        double ts[NC] = {};
        for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
            const double ph = a.phi_s[f];
            so += a.div_scheme != 1 ? fmax(ph, 0.0) : ph * m.w[f];
            if (tvd) {
                double base[NC], up[NC];
                const double lw = tvd_weight<NC>(m, a, f, base, up);
#pragma unroll
                for (int q = 0; q < NC; ++q) ts[q] += -lw * (base[q] - up[q]);
            }
        }
        for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
            const int f = m.nbr_face[k];
            const double ph = a.phi_s[f];
            sn += a.div_scheme != 1 ? -fmin(ph, 0.0) : -ph * (1.0 - m.w[f]);
            if (tvd) {
                double base[NC], up[NC];
                const double lw = tvd_weight<NC>(m, a, f, base, up);
#pragma unroll
                for (int q = 0; q < NC; ++q) ts[q] -= -lw * (base[q] - up[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) src[q] = src[q] + a.s_div * ts[q];
        double bs[NC] = {};
        GF_FOR_BND(m, c, k) {
            if (!a.covered[k]) continue;
            const double fr = a.frac[k], flux = a.phi_b[k], dd = m.bnd_mag_d[k];
            sb += flux * (1.0 - fr);
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const double bsrc = fr * a.ref_v[k * NC + q] + (1.0 - fr) * dd * a.ref_g[k * NC + q];
                bs[q] += -(flux * bsrc);
            }
        }
        dg = dg + a.s_div * ((so + sn) + sb);
#pragma unroll
        for (int q = 0; q < NC; ++q) src[q] = src[q] + a.s_div * bs[q];
    }
    if (a.s_lap != 0.0) {
        double so = 0.0, sn = 0.0, sb = 0.0;
        for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) so += -(gamma_face(a, m, f) * m.mag_sf[f] * m.delta[f]);
        for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
            const int f = m.nbr_face[k];
            sn += -(gamma_face(a, m, f) * m.mag_sf[f] * m.delta[f]);
        }
        double bs[NC] = {};
        const double gb = a.gamma_cell ? a.gamma_cell[c] : a.gamma;
        GF_FOR_BND(m, c, k) {
            if (!a.covered[k]) continue;
            const double fr = a.frac[k];
            const double cond = gb * m.bnd_mag_sf[k];
            const double coef = bnd_value_coef(m, k, fr);
            const double cv = cond * coef, cg = cond * (1.0 - fr);
            sb += -cv;
#pragma unroll
            for (int q = 0; q < NC; ++q) bs[q] += -(cv * a.ref_v[k * NC + q] + cg * a.ref_g[k * NC + q]);
        }
        dg = dg + a.s_lap * ((so + sn) + sb);
#pragma unroll
        for (int q = 0; q < NC; ++q) src[q] = src[q] + a.s_lap * bs[q];
    }
    diag[c] = dg;
#pragma unroll
    for (int q = 0; q < NC; ++q) source[c * NC + q] = src[q];
}

// Hanging-face explicit sources, applied after k_asm_cells by the one thread
// that owns each hang cell (fvm/laplacian.py:163-177, schemes/div.py:126-137).
template <int NC>
__global__ void k_asm_hang(const gf_mesh m, const gf_assembly a, double *source, double *ffc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_hang_cells) return;
    const int c = m.hang_cell[i];
    const bool do_div = a.s_div != 0.0 && a.div_scheme >= 1;
    const bool do_lap = a.s_lap != 0.0 && a.lap_corrected;
    if (!do_div && !do_lap) return;
    double add[NC] = {};
    for (int k = m.hc_row[i]; k < m.hc_row[i + 1]; ++k) {
        const int code = m.hc_code[k];
        const bool own = code >= 0;
        const int h = own ? code : -code - 1;
        const int f = m.hang_face[h];
        const int o = m.owner[f], n = m.neighbour[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            if (do_div) {
                const double w = m.w[f];
                const double base = w * a.psi[o * NC + q] + (1.0 - w) * a.psi[n * NC + q];
                const double corrected = base + hang_interp_corr<NC>(m, a.grad, h, q);
                double fl = a.phi_s[f];
                if (a.div_scheme >= 2) {  // TVD: the limited flux
                    double b2[NC], u2[NC];
                    fl = tvd_weight<NC>(m, a, f, b2, u2);
                }
                const double sface = -fl * (corrected - base);
                add[q] += a.s_div * (own ? sface : -sface);
            }
            if (do_lap) {
                const double lsrc = (gamma_face(a, m, f) * m.mag_sf[f]) * hang_sn_corr<NC>(m, a.grad, h, q);
                add[q] += a.s_lap * (own ? -lsrc : lsrc);
                if (own && ffc) ffc[f * NC + q] = a.s_lap * lsrc;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) source[c * NC + q] += add[q];
}

// ---------------------------------------------------------------- assembly VJP
// Reverse mode of k_asm_faces / k_asm_cells / k_asm_hang: the gradients of a
// scalar with respect to (diag, upper, lower, source, ffc) are carried back to
// the diffusivity, the face flux, the old field levels and the boundary
// reference values, following what autograd computes through the reference's
// torch assembly (fvm/laplacian.py:38-201, fvm/div.py:15-73,
// schemes/div.py:36-142, fvm/ddt.py:7-40).  Each output element is written by
// exactly one thread (face outputs by the face, boundary outputs by the
// owning cell), so the result is deterministic.  The field values enter only
// through the hanging-face corrections, which are held fixed.

// d gamma_f / d (gamma_O, gamma_N).  Linear: (w, 1-w).  Harmonic
// gamma_f = gO gN / (w gO + (1-w) gN): ((1-w) gN^2, w gO^2) / D^2, evaluated
// scaled by max(gO, gN); where both sides are zero the tie of min() splits
// the gradient in halves, as torch.minimum's backward does.
__device__ __forceinline__ void gamma_face_partials(const gf_assembly &a, const gf_mesh &m, int f, double &dgo,
                                                    double &dgn) {
    const double w = m.w[f];
    if (!a.lap_harmonic) {
        dgo = w;
        dgn = 1.0 - w;
        return;
    }
    const double go = a.gamma_cell[m.owner[f]], gn = a.gamma_cell[m.neighbour[f]];
    const double scale = fmax(go, gn);
    if (!(scale > 0.0)) {
        dgo = 0.5;
        dgn = 0.5;
        return;
    }
    const double gos = go / scale, gns = gn / scale;
    const double den = w * gos + (1.0 - w) * gns;
    const double ro = gns / den, rn = gos / den;
    dgo = (1.0 - w) * (ro * ro);
    dgn = w * (rn * rn);
}

// Per face: d/d gamma_f into work[f] and d/d phi_s[f] (upwind follows
// torch.clamp's backward, which passes the gradient at phi = 0 to both the
// min(phi, 0) and the max(phi, 0) branches).
template <int NC>
__global__ void k_asmv_faces(const gf_mesh m, const gf_assembly a, const double *__restrict__ gd,
                             const double *__restrict__ gu, const double *__restrict__ gl,
                             const double *__restrict__ gs, double *work, double *g_phi_s) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f];
    if (g_phi_s) {
        double gp = 0.0;
        if (a.s_div != 0.0) {
            const double ph = a.phi_s[f];
            if (a.div_scheme != 1) {
                const double neg = ph <= 0.0 ? gu[f] - gd[n] : 0.0;
                const double pos = ph >= 0.0 ? gd[o] - gl[f] : 0.0;
                gp = neg + pos;
                if (a.div_scheme >= 2 && ph != 0.0) {  // d/dphi of -phi psi(r) (base - up), r held fixed
                    double base[NC], up[NC];
                    const double lim = tvd_weight<NC>(m, a, f, base, up) / ph;
#pragma unroll
                    for (int q = 0; q < NC; ++q) gp += -lim * (base[q] - up[q]) * (gs[o * NC + q] - gs[n * NC + q]);
                }
            } else {
                const double w = m.w[f];
                gp = ((1.0 - w) * gu[f] - w * gl[f]) + (w * gd[o] - (1.0 - w) * gd[n]);
            }
            gp *= a.s_div;
        }
        g_phi_s[f] = gp;
    }
    if (work) {
        double gg = 0.0;
        if (a.s_lap != 0.0) gg = a.s_lap * (m.mag_sf[f] * m.delta[f]) * ((gu[f] + gl[f]) - (gd[o] + gd[n]));
        work[f] = gg;
    }
}

// Hanging faces: the explicit corrections' share (one thread per hanging face,
// after k_asmv_faces).
template <int NC>
__global__ void k_asmv_hang(const gf_mesh m, const gf_assembly a, const double *__restrict__ gs,
                            const double *__restrict__ gffc, double *work, double *g_phi_s) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= m.n_hang) return;
    const int f = m.hang_face[h];
    const int o = m.owner[f], n = m.neighbour[f];
    if (g_phi_s && a.s_div != 0.0 && a.div_scheme >= 1) {
        const double w = m.w[f];
        double lim = 1.0;
        if (a.div_scheme >= 2) {  // TVD: the limiter, r held fixed
            double b2[NC], u2[NC];
            const double ph = a.phi_s[f];
            lim = ph != 0.0 ? tvd_weight<NC>(m, a, f, b2, u2) / ph : 0.0;
        }
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double base = w * a.psi[o * NC + q] + (1.0 - w) * a.psi[n * NC + q];
            const double corrected = base + hang_interp_corr<NC>(m, a.grad, h, q);
            acc += -lim * (corrected - base) * (gs[o * NC + q] - gs[n * NC + q]);
        }
        g_phi_s[f] += a.s_div * acc;
    }
    if (work && a.s_lap != 0.0 && a.lap_corrected) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double sc = hang_sn_corr<NC>(m, a.grad, h, q);
            const double gf = gffc ? gffc[f * NC + q] : 0.0;
            acc += sc * ((gs[n * NC + q] - gs[o * NC + q]) + gf);
        }
        work[f] += a.s_lap * m.mag_sf[f] * acc;
    }
}

// The hanging corrections are linear in the (hang-cell) least-squares
// gradient: gather d J / d grad per hang cell over its hanging faces
// (owner side: w, -mo; neighbour side: 1 - w, +mn).  TVD sources are not
// covered (their ratio depends on the gradient nonlinearly).
template <int NC>
__global__ void k_asmv_hang_grad(const gf_mesh m, const gf_assembly a, const double *__restrict__ gs,
                                 const double *__restrict__ gffc, double *g_grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_hang_cells) return;
    const int c = m.hang_cell[i];
    const bool do_div = a.s_div != 0.0 && a.div_scheme == 1;
    const bool do_lap = a.s_lap != 0.0 && a.lap_corrected;
    double acc[NC][3] = {};
    for (int k = m.hc_row[i]; k < m.hc_row[i + 1]; ++k) {
        const int code = m.hc_code[k];
        const bool own = code >= 0;
        const int h = own ? code : -code - 1;
        const int f = m.hang_face[h];
        const int o = m.owner[f], n = m.neighbour[f];
        const double w = m.w[f];
        const double *off = m.hang_offset + 3 * h, *mo = m.hang_move_o + 3 * h, *mn = m.hang_move_n + 3 * h;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            if (do_div) {
                const double fac = -a.s_div * a.phi_s[f] * (gs[o * NC + q] - gs[n * NC + q]) * (own ? w : 1.0 - w);
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[q][j] += fac * off[j];
            }
            if (do_lap) {
                const double gf = gffc ? gffc[f * NC + q] : 0.0;
                const double fac = a.s_lap * (gamma_face(a, m, f) * m.mag_sf[f]) * m.delta[f] *
                                   ((gs[n * NC + q] - gs[o * NC + q]) + gf);
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[q][j] += fac * (own ? -mo[j] : mn[j]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) g_grad[(c * NC + q) * 3 + j] = acc[q][j];
}

// Per cell: old levels, the diffusivity (cell values, or per-cell shares of a
// uniform diffusivity whose sum is its gradient: each face counted at its
// owner) and the cell's boundary faces (flux, reference value and gradient).
template <int NC>
__global__ void k_asmv_cells(const gf_mesh m, const gf_assembly a, const double *__restrict__ gd,
                             const double *__restrict__ gs, const double *__restrict__ work, double *g_gamma,
                             double *g_phi_b, double *g_psi_old, double *g_psi_older, double *g_ref_v,
                             double *g_ref_g) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    const double vdt = a.s_ddt != 0.0 ? m.volumes[c] / a.dt : 0.0;
    if (g_psi_old)
#pragma unroll
        for (int q = 0; q < NC; ++q) g_psi_old[c * NC + q] = a.s_ddt * (vdt * a.ddt_b) * gs[c * NC + q];
    if (g_psi_older)
#pragma unroll
        for (int q = 0; q < NC; ++q) g_psi_older[c * NC + q] = -(a.s_ddt * (vdt * a.ddt_c)) * gs[c * NC + q];
    const double gdc = gd[c];
    double gsc[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) gsc[q] = gs[c * NC + q];
    double ggam = 0.0;
    if (g_gamma && a.s_lap != 0.0) {
        for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
            double dgo = 1.0, dgn = 0.0;
            if (a.gamma_cell) gamma_face_partials(a, m, f, dgo, dgn);
            ggam += dgo * work[f];
        }
        if (a.gamma_cell)
            for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
                const int f = m.nbr_face[k];
                double dgo, dgn;
                gamma_face_partials(a, m, f, dgo, dgn);
                ggam += dgn * work[f];
            }
    }
    const double gb = a.gamma_cell ? a.gamma_cell[c] : a.gamma;
    GF_FOR_BND(m, c, k) {
        double gphib = 0.0, grv[NC] = {}, grg[NC] = {};
        if (a.covered[k]) {
            const double fr = a.frac[k], dd = m.bnd_mag_d[k];
            if (a.s_div != 0.0) {
                const double flux = a.phi_b[k];
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    s += gsc[q] * (fr * a.ref_v[k * NC + q] + (1.0 - fr) * dd * a.ref_g[k * NC + q]);
                    grv[q] += -(a.s_div * flux * fr) * gsc[q];
                    grg[q] += -(a.s_div * flux * (1.0 - fr) * dd) * gsc[q];
                }
                gphib = a.s_div * (gdc * (1.0 - fr) - s);
            }
            if (a.s_lap != 0.0) {
                const double coef = bnd_value_coef(m, k, fr);
                const double cv = gb * m.bnd_mag_sf[k] * coef, cg = gb * m.bnd_mag_sf[k] * (1.0 - fr);
                double s = -coef * gdc;
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    s -= (coef * a.ref_v[k * NC + q] + (1.0 - fr) * a.ref_g[k * NC + q]) * gsc[q];
                    grv[q] += -(a.s_lap * cv) * gsc[q];
                    grg[q] += -(a.s_lap * cg) * gsc[q];
                }
                ggam += a.s_lap * m.bnd_mag_sf[k] * s;
            }
        }
        if (g_phi_b) g_phi_b[k] = gphib;
        if (g_ref_v)
#pragma unroll
            for (int q = 0; q < NC; ++q) g_ref_v[k * NC + q] = grv[q];
        if (g_ref_g)
#pragma unroll
            for (int q = 0; q < NC; ++q) g_ref_g[k * NC + q] = grg[q];
    }
    if (g_gamma) g_gamma[c] = ggam;
}

// ---------------------------------------------------------------- LDU algebra
template <int NC>
__global__ void k_ldu_multiply(const gf_mesh m, const double *diag, const double *upper, const double *lower,
                               const double *x, double *y) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double so[NC] = {}, sn[NC] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const int n = m.neighbour[f];
        const double u = upper[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) so[q] += u * x[n * NC + q];
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
        const int o = m.owner[f];
        const double l = lower[f];
#pragma unroll
        for (int q = 0; q < NC; ++q) sn[q] += l * x[o * NC + q];
    }
    const double d = diag[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) y[c * NC + q] = (d * x[c * NC + q] + so[q]) + sn[q];
}

__global__ void k_rau_hbya(const gf_mesh m, const double *diag, const double *upper, const double *lower,
                           const double *source, const double *U, double *rAU, double *HbyA) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double so[3] = {}, sn[3] = {};
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) {
        const int n = m.neighbour[f];
        const double u = upper[f];
#pragma unroll
        for (int q = 0; q < 3; ++q) so[q] += u * U[n * 3 + q];
    }
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
        const int f = m.nbr_face[k];
        const int o = m.owner[f];
        const double l = lower[f];
#pragma unroll
        for (int q = 0; q < 3; ++q) sn[q] += l * U[o * 3 + q];
    }
    const double d = diag[c], V = m.volumes[c];
    const double r = 1.0 / (d / V);
    rAU[c] = r;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double x = U[c * 3 + q];
        const double ax = (d * x + so[q]) + sn[q];
        const double H = (source[c * 3 + q] - ax + d * x) / V;
        HbyA[c * 3 + q] = r * H;
    }
}

template <int NC>
__global__ void k_ldu_flux(const gf_mesh m, double sign, const double *upper, const double *lower,
                           const double *ffc, const double *psi, double *flux) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        double v = upper[f] * psi[n * NC + q] - lower[f] * psi[o * NC + q];
        if (ffc) v = v + ffc[f * NC + q];
        flux[f * NC + q] = sign * v;
    }
}

// ---------------------------------------------------------------- fluxes
__global__ void k_face_flux_int(const gf_mesh m, const double *Uf, double *phi_s) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    phi_s[f] = Uf[3 * f + m.axis[f]] * m.mag_sf[f];
}
__global__ void k_face_flux_bnd(const gf_mesh m, const double *Ub, const uint8_t *covered, const uint8_t *mask,
                                double *phi_b) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd) return;
    if (mask && !mask[k]) return;
    const int a = m.bnd_dir[k] / 2;
    const double sf = m.bnd_sign[k] * m.bnd_mag_sf[k];
    phi_b[k] = (covered && !covered[k]) ? 0.0 : Ub[3 * k + a] * sf;
}

__global__ void k_ddt_corr_add(const gf_mesh m, const double *U0, const double *U00, const double *phi0,
                               const double *phi00, const double *rAU, double b, double cc, double dt,
                               double *phi_s) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f], a = m.axis[f];
    const double w = m.w[f], sf = m.mag_sf[f];
    const double u0f = w * U0[3 * o + a] + (1.0 - w) * U0[3 * n + a];
    const double corr = phi0[f] - u0f * sf;
    double hist = b * corr;
    if (cc != 0.0) {
        const double u00f = w * U00[3 * o + a] + (1.0 - w) * U00[3 * n + a];
        hist = hist - cc * (phi00[f] - u00f * sf);
    }
    const double tiny = 2.2250738585072014e-308;
    const double coupling = 1.0 - fmin(fabs(corr) / (fabs(phi0[f]) + tiny), 1.0);
    const double rf = w * rAU[o] + (1.0 - w) * rAU[n];
    phi_s[f] = phi_s[f] + rf * (coupling * hist / dt);
}

// adjust_phi.py:44-127 -- one workgroup, fixed summation order
__global__ void k_adjust_phi(const gf_mesh m, const int8_t *kind, double *phi_b) {
    double v[3] = {0.0, 0.0, 0.0};  // mass_in, fixed_out, adjustable_out
    for (int k = threadIdx.x; k < m.n_bnd; k += blockDim.x) {
        const int8_t t = kind[k];
        if (!t) continue;
        const double ph = phi_b[k];
        if (ph < 0.0) v[0] -= ph;
        if (ph > 0.0) {
            if (t == 1) v[1] += ph;
            else v[2] += ph;
        }
    }
    const bool mx[3] = {false, false, false};
    block_reduce<3, 1024>(v, mx);
    __shared__ double corr;
    __shared__ int apply;
    if (threadIdx.x == 0) {
        apply = v[2] > 0.0;
        corr = apply ? (v[0] - v[1]) / v[2] : 1.0;
    }
    __syncthreads();
    if (!apply) return;
    for (int k = threadIdx.x; k < m.n_bnd; k += blockDim.x)
        if (kind[k] == 2 && phi_b[k] > 0.0) phi_b[k] = phi_b[k] * corr;
}

__global__ void k_correct_phi_int(const gf_mesh m, const double *phi_h, const double *upper, const double *lower,
                                  const double *ffc, const double *p, double *phi_s) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const int o = m.owner[f], n = m.neighbour[f];
    double v = upper[f] * p[n] - lower[f] * p[o];
    if (ffc) v = v + ffc[f];
    phi_s[f] = phi_h[f] - (-v);
}
__global__ void k_correct_phi_bnd(const gf_mesh m, const double *phi_hb, const double *p, const double *rAtU,
                                  const double *frac, const double *rv, const double *rg, const uint8_t *covered,
                                  double *phi_b) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n_bnd) return;
    double out = phi_hb[k];
    if (covered[k]) {
        const int c = m.bnd_owner[k];
        const double fr = frac[k];
        const double coef = bnd_value_coef(m, k, fr);
        const double sng = coef * (rv[k] - p[c]) + (1.0 - fr) * rg[k];
        out = out - rAtU[c] * m.bnd_mag_sf[k] * sng;
    }
    phi_b[k] = out;
}

__global__ void k_correct_velocity(int C, const double *HbyA, const double *rAtU, const double *gp, double *U) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double r = rAtU[c];
#pragma unroll
    for (int q = 0; q < 3; ++q) U[3 * c + q] = HbyA[3 * c + q] - r * gp[3 * c + q];
}

// ---- symmetric elimination of prescribed cell values (fvmatrix.py:325-394)
__global__ void k_fix_mark(int K, int nc, const int32_t *__restrict__ cells, const double *__restrict__ values,
                           uint8_t *__restrict__ fixed, double *__restrict__ presc) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int c = cells[k];
    fixed[c] = 1;
    for (int q = 0; q < nc; ++q) presc[(size_t)c * nc + q] = values[(size_t)k * nc + q];
}

// Sum of the diagonal by one workgroup in a fixed order (the sign picks the
// unit diagonal of rows that have none).
__global__ __launch_bounds__(kBlock) void k_diag_sum(int C, const double *__restrict__ diag, double *__restrict__ out) {
    double a = 0.0;
    for (int i = threadIdx.x; i < C; i += kBlock) a += diag[i];
    double v[1] = {a};
    const bool mx[1] = {false};
    block_reduce<1, kBlock>(v, mx);
    if (threadIdx.x == 0) out[0] = v[0];
}

__global__ void k_fix_faces(const gf_mesh m, const uint8_t *__restrict__ fixed, const double *__restrict__ upper,
                            const double *__restrict__ lower, double *__restrict__ out_upper,
                            double *__restrict__ out_lower) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.n_faces) return;
    const bool conn = fixed[m.owner[f]] | fixed[m.neighbour[f]];
    out_upper[f] = conn ? 0.0 : upper[f];
    out_lower[f] = conn ? 0.0 : lower[f];
}

// Fixed rows: constraint diagonal (the row's own, or +-1 when it is zero)
// and source = diag * value.  Other rows: prescribed neighbours move to the
// right-hand side.  Gathered per cell over its face lists (no atomics).
__global__ void k_fix_cells(const gf_mesh m, int nc, const uint8_t *__restrict__ fixed,
                            const double *__restrict__ presc, const double *__restrict__ diag,
                            const double *__restrict__ upper, const double *__restrict__ lower,
                            const double *__restrict__ source, const double *__restrict__ dsum,
                            double *__restrict__ out_diag, double *__restrict__ out_source) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    const double d = diag[c];
    if (fixed[c]) {
        const double cd = d != 0.0 ? d : (dsum[0] < 0.0 ? -1.0 : 1.0);
        out_diag[c] = cd;
        for (int q = 0; q < nc; ++q) out_source[(size_t)c * nc + q] = cd * presc[(size_t)c * nc + q];
        return;
    }
    out_diag[c] = d;
    for (int q = 0; q < nc; ++q) {
        double s = source[(size_t)c * nc + q];
        for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) s -= upper[f] * presc[(size_t)m.neighbour[f] * nc + q];
        for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
            const int f = m.nbr_face[k];
            s -= lower[f] * presc[(size_t)m.owner[f] * nc + q];
        }
        out_source[(size_t)c * nc + q] = s;
    }
}

// H1 = -(sum of the row's off-diagonals) / V, gathered per cell over its face
// lists in a fixed order (fvmatrix.py:250-264).
__global__ void k_ldu_h1(const gf_mesh m, const double *__restrict__ upper, const double *__restrict__ lower,
                         double *__restrict__ h1) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_cells) return;
    double a = 0.0;
    for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f) a -= upper[f];
    for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) a -= lower[m.nbr_face[k]];
    h1[c] = a / m.volumes[c];
}

// SIMPLEC coefficient rAtU = 1 / (1/rAU - H1), bounded by 0.1/rAU
// (pressure_correction.py:163-193).
__global__ void k_simplec_rautu(int C, const double *__restrict__ rAU, const double *__restrict__ h1, int bounded,
                                double *__restrict__ rAtU) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double ir = 1.0 / rAU[c];
    double inv = ir - h1[c];
    if (bounded) inv = fmax(inv, 0.1 * ir);
    rAtU[c] = 1.0 / inv;
}

__global__ void k_set_reference(double *diag, double *source, int cell, double value) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double rd = diag[cell];
    diag[cell] = diag[cell] + rd;
    source[cell] = source[cell] + rd * value;
}

__global__ void k_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = a * x[i] + b * y[i];
}

// residual.py:12-36 -- ||A psi - b|| / (||b|| + eps), deterministic two-level sum
template <int NC>
__global__ __launch_bounds__(kBlock) void k_init_res(const gf_mesh m, const double *diag, const double *upper,
                                                     const double *lower, const double *src, const double *psi,
                                                     double *partials) {
    int c = blockIdx.x * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (c < m.n_cells) {
        double so[NC] = {}, sn[NC] = {};
        for (int f = m.own_row[c]; f < m.own_row[c + 1]; ++f)
#pragma unroll
            for (int q = 0; q < NC; ++q) so[q] += upper[f] * psi[m.neighbour[f] * NC + q];
        for (int k = m.nbr_row[c]; k < m.nbr_row[c + 1]; ++k) {
            const int f = m.nbr_face[k];
#pragma unroll
            for (int q = 0; q < NC; ++q) sn[q] += lower[f] * psi[m.owner[f] * NC + q];
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double ax = (diag[c] * psi[c * NC + q] + so[q]) + sn[q];
            const double r = ax - src[c * NC + q];
            v[0] += r * r;
            v[1] += src[c * NC + q] * src[c * NC + q];
        }
    }
    const bool mx[2] = {false, false};
    block_reduce<2, kBlock>(v, mx);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v[0];
        partials[nb + blockIdx.x] = v[1];
    }
}

// second level in one workgroup (no cross-workgroup fences on multi-XCD parts)
__global__ __launch_bounds__(1024) void k_init_res_fin(const double *partials, int nb, double *out) {
    double a = reduce_partials<1024>(partials, nb, false);
    double b = reduce_partials<1024>(partials + nb, nb, false);
    if (threadIdx.x == 0) out[0] = sqrt(a) / (sqrt(b) + 2.220446049250313e-16);
}

// ldu_grads.py:11-53
template <int NC>
__global__ void k_ldu_grads(const gf_mesh m, const double *lam, const double *x, double *gd, double *gu,
                            double *gl) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.n_cells) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NC; ++q) s += lam[i * NC + q] * x[i * NC + q];
        gd[i] = -s;
    }
    if (i < m.n_faces) {
        const int o = m.owner[i], n = m.neighbour[i];
        double su = 0.0, sl = 0.0;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            su += lam[o * NC + q] * x[n * NC + q];
            sl += lam[n * NC + q] * x[o * NC + q];
        }
        gu[i] = -su;
        gl[i] = -sl;
    }
}

// simple.py:302-311: A_new = A/alpha ; source += ((1-alpha)/alpha) A_old U
__global__ void k_relax_momentum(int C, double alpha, double *diag, double *source, const double *U) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double a_old = diag[c];
    diag[c] = a_old / alpha;
    const double rc = ((1.0 - alpha) / alpha) * a_old;
#pragma unroll
    for (int q = 0; q < 3; ++q) source[3 * c + q] = source[3 * c + q] + rc * U[3 * c + q];
}

// piso.py:247-250 and pressure_correction.py:121-124: out = source - V x
__global__ void k_predictor_source(int C, int nc, const double *source, const double *V, const double *x, double *out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    for (int q = 0; q < nc; ++q) out[nc * c + q] = source[nc * c + q] - x[nc * c + q] * V[c];
}

// simple.py:386: p = p_old + alpha (p - p_old)
__global__ void k_relax_field(int64_t n, double alpha, const double *old, double *x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = old[i] + alpha * (x[i] - old[i]);
}

}  // namespace
}  // namespace gf

using namespace gf;

#define GF_NC_DISPATCH(nc, KERNEL, GRID, ...)                                                          \
    do {                                                                                               \
        if ((nc) == 1)                                                                                 \
            hipLaunchKernelGGL(KERNEL<1>, dim3(GRID), dim3(kBlock), 0, s, __VA_ARGS__);               \
        else if ((nc) == 3)                                                                            \
            hipLaunchKernelGGL(KERNEL<3>, dim3(GRID), dim3(kBlock), 0, s, __VA_ARGS__);               \
        else {                                                                                         \
            set_error("component count must be 1 or 3");                                               \
            return GF_EINVAL;                                                                          \
        }                                                                                              \
    } while (0)

extern "C" const char *gf_last_error(void) { return g_err.c_str(); }

extern "C" int gf_device_count(int *count) {
    GF_REQUIRE(count, "null argument");
    GF_CHECK_HIP(hipGetDeviceCount(count));
    return GF_OK;
}

extern "C" int gf_bc_state(const gf_mesh *m, int32_t nc, const double *psi, const double *frac, const double *ref_v,
                           const double *ref_g, const uint8_t *covered, double *psi_b, gf_stream_t stream) {
    GF_REQUIRE(m && psi && frac && ref_v && ref_g && covered && psi_b, "null argument");
    if (m->n_bnd == 0) return GF_OK;
    hipLaunchKernelGGL(k_bc_state, dim3(grid_for(m->n_bnd)), dim3(kBlock), 0, S(stream), *m, nc, psi, frac, ref_v,
                       ref_g, covered, psi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_bc_slip(const gf_mesh *m, const double *psi, const uint8_t *slip_mask, double *ref_v,
                          gf_stream_t stream) {
    GF_REQUIRE(m && psi && slip_mask && ref_v, "null argument");
    if (m->n_bnd == 0) return GF_OK;
    hipLaunchKernelGGL(k_bc_slip, dim3(grid_for(m->n_bnd)), dim3(kBlock), 0, S(stream), *m, psi, slip_mask, ref_v);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_lsq_grad(const gf_mesh *m, int32_t nc, const double *psi, const double *psi_b,
                           const uint8_t *covered, int32_t hang_only, double *grad, gf_stream_t stream) {
    GF_REQUIRE(m && psi && psi_b && covered && grad, "null argument");
    hipStream_t s = S(stream);
    const int32_t *list = hang_only ? m->hang_cell : nullptr;
    const int n = hang_only ? m->n_hang_cells : m->n_cells;
    if (n == 0) return GF_OK;
    GF_NC_DISPATCH(nc, k_lsq_grad, grid_for(n), *m, psi, psi_b, covered, list, n, grad);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_lsq_grad_vjp(const gf_mesh *m, int32_t nc, const double *g_grad, const uint8_t *covered,
                               int32_t hang_only, double *work, double *g_psi, double *g_psi_b, gf_stream_t stream) {
    GF_REQUIRE(m && g_grad && covered && work && g_psi && g_psi_b, "null argument");
    hipStream_t s = S(stream);
    const int32_t *list = hang_only ? m->hang_cell : nullptr;
    const int n = hang_only ? m->n_hang_cells : m->n_cells;
    if (n > 0) GF_NC_DISPATCH(nc, k_lsq_v, grid_for(n), *m, g_grad, list, n, work);  // work arrives zeroed
    GF_NC_DISPATCH(nc, k_lsq_vjp, grid_for(m->n_cells), *m, work, covered, g_psi, g_psi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_bc_state_vjp(const gf_mesh *m, int32_t nc, const double *frac, const uint8_t *covered,
                               const uint8_t *slip_mask, const double *g_psi_b, double *g_psi, double *g_ref_v,
                               double *g_ref_g, gf_stream_t stream) {
    GF_REQUIRE(m && frac && covered && g_psi_b && g_psi, "null argument");
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(nc, k_bc_state_vjp, grid_for(m->n_cells), *m, frac, covered, slip_mask, g_psi_b, g_psi, g_ref_v,
                   g_ref_g);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_interpolate(const gf_mesh *m, int32_t nc, const double *psi, const double *grad, double *psi_f,
                              gf_stream_t stream) {
    GF_REQUIRE(m && psi && psi_f, "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces == 0) return GF_OK;
    GF_NC_DISPATCH(nc, k_interp_linear, grid_for(m->n_faces), *m, psi, psi_f);
    if (m->n_hang > 0) {
        GF_REQUIRE(grad, "hanging faces need a gradient");
        GF_NC_DISPATCH(nc, k_interp_hang, grid_for(m->n_hang), *m, grad, psi_f);
    }
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_fvc_div(const gf_mesh *m, int32_t nc, const double *phi_s, const double *phi_b, double *out,
                          gf_stream_t stream) {
    GF_REQUIRE(m && phi_s && phi_b && out, "null argument");
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(nc, k_fvc_div, grid_for(m->n_cells), *m, phi_s, phi_b, out);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_reconstruct_vjp(const gf_mesh *m, const double *g_U, double *work, double *g_phi_s,
                                  double *g_phi_b, gf_stream_t stream) {
    GF_REQUIRE(m && g_U && work && g_phi_s && g_phi_b, "null argument");
    hipStream_t s = S(stream);
    if (m->n_cells == 0) return GF_OK;
    hipLaunchKernelGGL(k_reconstruct_vjp_cells, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, s, *m, g_U, work);
    if (m->n_faces + m->n_bnd > 0)
        hipLaunchKernelGGL(k_reconstruct_vjp_faces, dim3(grid_for(m->n_faces + m->n_bnd)), dim3(kBlock), 0, s, *m, work,
                           g_phi_s, g_phi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_fvc_div_vjp(const gf_mesh *m, int32_t nc, const double *g_out, double *g_phi_s, double *g_phi_b,
                              gf_stream_t stream) {
    GF_REQUIRE(m && g_out && g_phi_s && g_phi_b, "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces > 0) GF_NC_DISPATCH(nc, k_fvc_div_vjp_faces, grid_for(m->n_faces), *m, g_out, g_phi_s);
    if (m->n_bnd > 0) GF_NC_DISPATCH(nc, k_fvc_div_vjp_bnd, grid_for(m->n_bnd), *m, g_out, g_phi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_interpolate_vjp(const gf_mesh *m, int32_t nc, const double *g_psi_f, double *g_psi,
                                  double *g_grad, gf_stream_t stream) {
    GF_REQUIRE(m && g_psi_f && g_psi, "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces == 0) return GF_OK;
    GF_NC_DISPATCH(nc, k_interp_vjp_cells, grid_for(m->n_cells), *m, g_psi_f, g_psi);
    if (m->n_hang > 0) {
        GF_REQUIRE(g_grad, "hanging faces need a gradient cotangent");
        GF_NC_DISPATCH(nc, k_interp_vjp_hang, grid_for(m->n_hang_cells), *m, g_psi_f, g_grad);
    }
    GF_CHECK_LAUNCH();
    return GF_OK;
}

// Volume-weighted continuity errors of a face flux, FP64, one workgroup in a
// fixed order (post/diagnostics/continuity.py:12-49): out = {dt sum|div| V / sum V,
// dt sum div V / sum V}.
__global__ __launch_bounds__(1024) void k_continuity(int C, const double *__restrict__ div,
                                                     const double *__restrict__ vol, double dt,
                                                     double *__restrict__ out) {
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < C; i += 1024) {
        const double v = vol[i], d = div[i];
        a[0] += fabs(d) * v;
        a[1] += d * v;
        a[2] += v;
    }
    const bool mx[3] = {false, false, false};
    block_reduce<3, 1024>(a, mx);
    if (threadIdx.x == 0) {
        out[0] = dt * a[0] / a[2];
        out[1] = dt * a[1] / a[2];
    }
}

extern "C" int gf_continuity_error(const gf_mesh *m, const double *phi_s, const double *phi_b, double dt,
                                   double *div_scratch, double *out, gf_stream_t stream) {
    GF_REQUIRE(m && phi_s && phi_b && div_scratch && out, "null argument");
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(1, k_fvc_div, grid_for(m->n_cells), *m, phi_s, phi_b, div_scratch);
    hipLaunchKernelGGL(k_continuity, dim3(1), dim3(1024), 0, s, m->n_cells, div_scratch, m->volumes, dt, out);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_reconstruct(const gf_mesh *m, const double *phi_s, const double *phi_b, double *U,
                              gf_stream_t stream) {
    GF_REQUIRE(m && phi_s && phi_b && U, "null argument");
    hipLaunchKernelGGL(k_reconstruct, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, S(stream), *m, phi_s, phi_b, U);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_gauss_grad(const gf_mesh *m, int32_t nc, const double *psi_s, const double *psi_b, double *grad,
                             gf_stream_t stream) {
    GF_REQUIRE(m && psi_s && psi_b && grad, "null argument");
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(nc, k_gauss_grad, grid_for(m->n_cells), *m, psi_s, psi_b, grad);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_gauss_grad_vjp(const gf_mesh *m, int32_t nc, const double *g_grad, double *g_psi_s,
                                 double *g_psi_b, gf_stream_t stream) {
    GF_REQUIRE(m && g_grad && g_psi_s && g_psi_b, "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces > 0) GF_NC_DISPATCH(nc, k_gauss_grad_vjp_faces, grid_for(m->n_faces), *m, g_grad, g_psi_s);
    if (m->n_bnd > 0) GF_NC_DISPATCH(nc, k_gauss_grad_vjp_bnd, grid_for(m->n_bnd), *m, g_grad, g_psi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_sn_grad_hang(const gf_mesh *m, int32_t nc, const double *grad, double *corr, gf_stream_t stream) {
    GF_REQUIRE(m && grad && corr, "null argument");
    hipStream_t s = S(stream);
    if (m->n_hang == 0) return GF_OK;
    GF_NC_DISPATCH(nc, k_sn_grad_hang, grid_for(m->n_hang), *m, grad, corr);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_assemble(const gf_mesh *m, const gf_assembly *a, double *diag, double *upper, double *lower,
                           double *source, double *ffc, gf_stream_t stream) {
    GF_REQUIRE(m && a && diag && upper && lower && source, "null argument");
    GF_REQUIRE(a->s_ddt == 0.0 || (a->psi_old && a->dt > 0.0), "ddt needs old values and dt > 0");
    GF_REQUIRE(a->s_ddt == 0.0 || a->ddt_c == 0.0 || a->psi_older, "backward ddt needs older values");
    GF_REQUIRE(a->s_div == 0.0 || (a->phi_s && a->phi_b), "div needs face flux");
    GF_REQUIRE(a->covered && a->frac && a->ref_v && a->ref_g, "boundary state missing");
    GF_REQUIRE(a->div_scheme >= 0 && a->div_scheme <= 7, "unknown div scheme");
    const bool tvd = a->s_div != 0.0 && a->div_scheme >= 2;
    const bool need_grad =
        tvd || (m->n_hang > 0 && ((a->s_div != 0.0 && a->div_scheme == 1) || (a->s_lap != 0.0 && a->lap_corrected)));
    GF_REQUIRE(!need_grad || (a->grad && a->psi), "hanging faces and TVD schemes need psi and its gradient");
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(a->nc, k_asm_faces, grid_for(m->n_faces), *m, *a, upper, lower, ffc);
    GF_NC_DISPATCH(a->nc, k_asm_cells, grid_for(m->n_cells), *m, *a, diag, source);
    if (need_grad && m->n_hang_cells > 0)
        GF_NC_DISPATCH(a->nc, k_asm_hang, grid_for(m->n_hang_cells), *m, *a, source, ffc);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_assemble_vjp(const gf_mesh *m, const gf_assembly *a, const double *g_diag, const double *g_upper,
                               const double *g_lower, const double *g_source, const double *g_ffc, double *work,
                               double *g_gamma, double *g_phi_s, double *g_phi_b, double *g_psi_old,
                               double *g_psi_older, double *g_ref_v, double *g_ref_g, double *g_grad,
                               gf_stream_t stream) {
    GF_REQUIRE(m && a && g_diag && g_upper && g_lower && g_source, "null argument");
    GF_REQUIRE(a->covered && a->frac && a->ref_v && a->ref_g, "boundary state missing");
    GF_REQUIRE(!g_gamma || work, "the diffusivity gradient needs a work array of one double per face");
    GF_REQUIRE(a->s_div == 0.0 || (a->phi_s && a->phi_b), "div needs face flux");
    GF_REQUIRE(a->s_ddt == 0.0 || a->dt > 0.0, "ddt needs dt > 0");
    const bool hang_div = a->s_div != 0.0 && a->div_scheme >= 1 && g_phi_s;
    const bool hang_lap = a->s_lap != 0.0 && a->lap_corrected && g_gamma;
    const bool tvd = a->s_div != 0.0 && a->div_scheme >= 2 && g_phi_s;
    const bool need_grad = tvd || (m->n_hang > 0 && (hang_div || hang_lap));
    GF_REQUIRE(!need_grad || (a->grad && a->psi), "hanging faces and TVD schemes need psi and its gradient");
    hipStream_t s = S(stream);
    double *wk = g_gamma ? work : nullptr;
    if (m->n_faces > 0 && (wk || g_phi_s))
        GF_NC_DISPATCH(a->nc, k_asmv_faces, grid_for(m->n_faces), *m, *a, g_diag, g_upper, g_lower, g_source, wk,
                       g_phi_s);
    if (m->n_hang > 0 && (hang_div || hang_lap))
        GF_NC_DISPATCH(a->nc, k_asmv_hang, grid_for(m->n_hang), *m, *a, g_source, g_ffc, wk, g_phi_s);
    GF_NC_DISPATCH(a->nc, k_asmv_cells, grid_for(m->n_cells), *m, *a, g_diag, g_source, wk, g_gamma, g_phi_b,
                   g_psi_old, g_psi_older, g_ref_v, g_ref_g);
    if (g_grad && m->n_hang_cells > 0) {
        GF_REQUIRE(!(a->s_div != 0.0 && a->div_scheme >= 2), "the gradient cotangent of TVD sources is not provided");
        GF_NC_DISPATCH(a->nc, k_asmv_hang_grad, grid_for(m->n_hang_cells), *m, *a, g_source, g_ffc, g_grad);
    }
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_ldu_multiply(const gf_mesh *m, int32_t nc, const double *diag, const double *upper,
                               const double *lower, const double *x, double *y, gf_stream_t stream) {
    GF_REQUIRE(m && diag && upper && lower && x && y, "null argument");
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(nc, k_ldu_multiply, grid_for(m->n_cells), *m, diag, upper, lower, x, y);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_rau_hbya(const gf_mesh *m, const double *diag, const double *upper, const double *lower,
                           const double *source, const double *U, double *rAU, double *HbyA, gf_stream_t stream) {
    GF_REQUIRE(m && diag && upper && lower && source && U && rAU && HbyA, "null argument");
    hipLaunchKernelGGL(k_rau_hbya, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, S(stream), *m, diag, upper, lower,
                       source, U, rAU, HbyA);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_ldu_flux(const gf_mesh *m, int32_t nc, double sign, const double *upper, const double *lower,
                           const double *ffc, const double *psi, double *flux, gf_stream_t stream) {
    GF_REQUIRE(m && upper && lower && psi && flux, "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces == 0) return GF_OK;
    GF_NC_DISPATCH(nc, k_ldu_flux, grid_for(m->n_faces), *m, sign, upper, lower, ffc, psi, flux);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_face_flux(const gf_mesh *m, const double *Uf_s, const double *Uf_b, const uint8_t *covered,
                            double *phi_s, double *phi_b, gf_stream_t stream) {
    GF_REQUIRE(m && Uf_s && Uf_b && covered && phi_s && phi_b, "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces)
        hipLaunchKernelGGL(k_face_flux_int, dim3(grid_for(m->n_faces)), dim3(kBlock), 0, s, *m, Uf_s, phi_s);
    if (m->n_bnd)
        hipLaunchKernelGGL(k_face_flux_bnd, dim3(grid_for(m->n_bnd)), dim3(kBlock), 0, s, *m, Uf_b, covered,
                           (const uint8_t *)nullptr, phi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_ddt_corr_add(const gf_mesh *m, const double *U_old, const double *U_older, const double *phi_old,
                               const double *phi_older, const double *rAU, double b, double c, double dt,
                               double *phi_s, gf_stream_t stream) {
    GF_REQUIRE(m && U_old && phi_old && rAU && phi_s && dt > 0.0, "null argument");
    GF_REQUIRE(c == 0.0 || (U_older && phi_older), "backward ddtCorr needs older history");
    if (m->n_faces == 0) return GF_OK;
    hipLaunchKernelGGL(k_ddt_corr_add, dim3(grid_for(m->n_faces)), dim3(kBlock), 0, S(stream), *m, U_old, U_older,
                       phi_old, phi_older, rAU, b, c, dt, phi_s);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_constrain_bnd_flux(const gf_mesh *m, const double *psi_b, const uint8_t *mask, double *phi_b,
                                     gf_stream_t stream) {
    GF_REQUIRE(m && psi_b && mask && phi_b, "null argument");
    if (m->n_bnd == 0) return GF_OK;
    hipLaunchKernelGGL(k_face_flux_bnd, dim3(grid_for(m->n_bnd)), dim3(kBlock), 0, S(stream), *m, psi_b,
                       (const uint8_t *)nullptr, mask, phi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_adjust_phi(const gf_mesh *m, const int8_t *kind, double *phi_b, double *scratch,
                             gf_stream_t stream) {
    (void)scratch;
    GF_REQUIRE(m && kind && phi_b, "null argument");
    if (m->n_bnd == 0) return GF_OK;
    hipLaunchKernelGGL(k_adjust_phi, dim3(1), dim3(1024), 0, S(stream), *m, kind, phi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_correct_phi(const gf_mesh *m, const double *phi_hbya_s, const double *phi_hbya_b,
                              const double *upper, const double *lower, const double *ffc, const double *p,
                              const double *rAtU, const double *frac, const double *ref_v, const double *ref_g,
                              const uint8_t *covered, double *phi_s, double *phi_b, gf_stream_t stream) {
    GF_REQUIRE(m && phi_hbya_s && phi_hbya_b && upper && lower && p && rAtU && frac && ref_v && ref_g && covered &&
                   phi_s && phi_b,
               "null argument");
    hipStream_t s = S(stream);
    if (m->n_faces)
        hipLaunchKernelGGL(k_correct_phi_int, dim3(grid_for(m->n_faces)), dim3(kBlock), 0, s, *m, phi_hbya_s, upper,
                           lower, ffc, p, phi_s);
    if (m->n_bnd)
        hipLaunchKernelGGL(k_correct_phi_bnd, dim3(grid_for(m->n_bnd)), dim3(kBlock), 0, s, *m, phi_hbya_b, p, rAtU,
                           frac, ref_v, ref_g, covered, phi_b);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_correct_velocity(int32_t n_cells, const double *HbyA, const double *rAtU, const double *grad_p,
                                   double *U, gf_stream_t stream) {
    GF_REQUIRE(HbyA && rAtU && grad_p && U, "null argument");
    hipLaunchKernelGGL(k_correct_velocity, dim3(grid_for(n_cells)), dim3(kBlock), 0, S(stream), n_cells, HbyA, rAtU,
                       grad_p, U);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_fixed_values(const gf_mesh *m, int32_t nc, int32_t n_fixed, const int32_t *cells,
                               const double *values, const double *diag, const double *upper, const double *lower,
                               const double *source, uint8_t *fixed, double *presc, double *dsum, double *out_diag,
                               double *out_upper, double *out_lower, double *out_source, gf_stream_t stream) {
    GF_REQUIRE(m && cells && values && diag && upper && lower && source && fixed && presc && dsum && out_diag &&
                   out_upper && out_lower && out_source,
               "null argument");
    GF_REQUIRE(nc >= 1 && n_fixed >= 1, "bad fixed-value arguments");
    hipStream_t s = S(stream);
    const int C = m->n_cells;
    GF_CHECK_HIP(hipMemsetAsync(fixed, 0, (size_t)C, s));
    GF_CHECK_HIP(hipMemsetAsync(presc, 0, sizeof(double) * (size_t)C * nc, s));
    hipLaunchKernelGGL(k_fix_mark, dim3(grid_for(n_fixed)), dim3(kBlock), 0, s, n_fixed, nc, cells, values, fixed,
                       presc);
    hipLaunchKernelGGL(k_diag_sum, dim3(1), dim3(kBlock), 0, s, C, diag, dsum);
    if (m->n_faces)
        hipLaunchKernelGGL(k_fix_faces, dim3(grid_for(m->n_faces)), dim3(kBlock), 0, s, *m, fixed, upper, lower,
                           out_upper, out_lower);
    hipLaunchKernelGGL(k_fix_cells, dim3(grid_for(C)), dim3(kBlock), 0, s, *m, nc, fixed, presc, diag, upper, lower,
                       source, dsum, out_diag, out_source);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_ldu_h1(const gf_mesh *m, const double *upper, const double *lower, double *h1,
                         gf_stream_t stream) {
    GF_REQUIRE(m && upper && lower && h1, "null argument");
    hipLaunchKernelGGL(k_ldu_h1, dim3(grid_for(m->n_cells)), dim3(kBlock), 0, S(stream), *m, upper, lower, h1);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_simplec_rautu(int32_t n_cells, const double *rAU, const double *h1, int32_t bounded, double *rAtU,
                                gf_stream_t stream) {
    GF_REQUIRE(rAU && h1 && rAtU, "null argument");
    hipLaunchKernelGGL(k_simplec_rautu, dim3(grid_for(n_cells)), dim3(kBlock), 0, S(stream), n_cells, rAU, h1,
                       bounded, rAtU);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_set_reference(double *diag, double *source, int32_t cell, double value, gf_stream_t stream) {
    GF_REQUIRE(diag && source && cell >= 0, "bad reference arguments");
    hipLaunchKernelGGL(k_set_reference, dim3(1), dim3(64), 0, S(stream), diag, source, cell, value);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out,
                        gf_stream_t stream) {
    GF_REQUIRE(x && y && out, "null argument");
    if (n == 0) return GF_OK;
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, S(stream), n, a, x, b, y, out);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

// ---- allocation-free initial residual -------------------------------------
// The per-workgroup partials of k_init_res live in a file-scope device array
// (one copy per GPU, held until the process ends) instead of a caller
// buffer: 2 x 2^17 doubles (2 MiB), enough for 2^17 workgroups of 256 cells,
// i.e. up to 33,554,432 cells.  Every workgroup writes its two partials
// before k_init_res_fin reads exactly those, so nothing needs clearing; calls
// that overlap on one GPU share the array and race.
constexpr int kNoAllocBlocks = 1 << 17;
static __device__ __align__(256) double scratch[2 * kNoAllocBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device: looked up once per
// device with that device current (the array's name is only a host-side
// stand-in and must never be passed to a kernel).
int scratch_address(double **out) {
    int dev = 0;
    GF_CHECK_HIP(hipGetDevice(&dev));
    GF_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        GF_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GF_REQUIRE(p != nullptr, "scratch symbol has no device address");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return GF_OK;
}
}  // namespace

extern "C" int gf_initial_residual_noalloc(const gf_mesh *m, int32_t nc, const double *diag, const double *upper,
                                           const double *lower, const double *source, const double *psi,
                                           double *out, gf_stream_t stream) {
    GF_REQUIRE(m && diag && upper && lower && source && psi && out, "null argument");
    GF_REQUIRE(nc == 1 || nc == 3, "nc must be 1 or 3");
    GF_REQUIRE(m->n_cells >= 1, "empty grid");
    const int nb = grid_for(m->n_cells);
    GF_REQUIRE(nb <= kNoAllocBlocks,
               "gf_initial_residual_noalloc: n_cells exceeds the static scratch capacity (33,554,432 cells)");
    double *sp = nullptr;
    if (int rc = scratch_address(&sp)) return rc;
    hipStream_t s = S(stream);
    GF_NC_DISPATCH(nc, k_init_res, nb, *m, diag, upper, lower, source, psi, sp);
    hipLaunchKernelGGL(k_init_res_fin, dim3(1), dim3(1024), 0, s, sp, nb, out);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_initial_residual(const gf_mesh *m, int32_t nc, const double *diag, const double *upper,
                                   const double *lower, const double *source, const double *psi, double *scratch,
                                   double *out, gf_stream_t stream) {
    GF_REQUIRE(m && diag && upper && lower && source && psi && scratch && out, "null argument");
    hipStream_t s = S(stream);
    const int nb = grid_for(m->n_cells);
    GF_NC_DISPATCH(nc, k_init_res, nb, *m, diag, upper, lower, source, psi, scratch);
    hipLaunchKernelGGL(k_init_res_fin, dim3(1), dim3(1024), 0, s, scratch, nb, out);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_ldu_grads(const gf_mesh *m, int32_t nc, const double *lambda, const double *x, double *g_diag,
                            double *g_upper, double *g_lower, gf_stream_t stream) {
    GF_REQUIRE(m && lambda && x && g_diag && g_upper && g_lower, "null argument");
    hipStream_t s = S(stream);
    const int n = m->n_cells > m->n_faces ? m->n_cells : m->n_faces;
    GF_NC_DISPATCH(nc, k_ldu_grads, grid_for(n), *m, lambda, x, g_diag, g_upper, g_lower);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_relax_momentum(int32_t n_cells, double alpha, double *diag, double *source, const double *U,
                                 gf_stream_t stream) {
    GF_REQUIRE(diag && source && U && alpha > 0.0, "bad relaxation arguments");
    hipLaunchKernelGGL(k_relax_momentum, dim3(grid_for(n_cells)), dim3(kBlock), 0, S(stream), n_cells, alpha, diag,
                       source, U);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_predictor_source(int32_t n_cells, int32_t nc, const double *source, const double *volumes,
                                   const double *x, double *out, gf_stream_t stream) {
    GF_REQUIRE(source && volumes && x && out && (nc == 1 || nc == 3), "bad arguments");
    hipLaunchKernelGGL(k_predictor_source, dim3(grid_for(n_cells)), dim3(kBlock), 0, S(stream), n_cells, nc, source,
                       volumes, x, out);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_relax_field(int64_t n, double alpha, const double *old, double *x, gf_stream_t stream) {
    GF_REQUIRE(old && x, "null argument");
    if (n == 0) return GF_OK;
    hipLaunchKernelGGL(k_relax_field, dim3(grid_for(n)), dim3(kBlock), 0, S(stream), n, alpha, old, x);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

// Plain 16-byte-per-lane copy: the HBM reference stream that the benchmark
// times beside the solver kernels and uses to scale the HBM counters.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void k_stream_copy(const f32x4 *__restrict__ src, f32x4 *__restrict__ dst,
                                                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        dst[i] = __builtin_nontemporal_load(src + i);
}

extern "C" int gf_stream_copy(const void *src, void *dst, int64_t bytes, gf_stream_t stream) {
    GF_REQUIRE(src && dst, "null argument");
    GF_REQUIRE(bytes % 16 == 0 && ((uintptr_t)src | (uintptr_t)dst) % 16 == 0,
               "gf_stream_copy needs 16-byte aligned buffers and sizes");
    const int64_t n = bytes / 16;
    if (n == 0) return GF_OK;
    const int64_t g = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_stream_copy, dim3((int)(g < 65536 ? g : 65536)), dim3(kBlock), 0, S(stream),
                       (const f32x4 *)src, (f32x4 *)dst, n);
    GF_CHECK_LAUNCH();
    return GF_OK;
}

extern "C" int gf_abi_sizes(int64_t out[3]) {
    GF_REQUIRE(out, "null argument");
    out[0] = (int64_t)sizeof(gf_mesh);
    out[1] = (int64_t)sizeof(gf_assembly);
    out[2] = (int64_t)sizeof(gf_krylov_cfg);
    return GF_OK;
}
