// Axis-projected immersed-boundary cut of the internal faces by analytic
// bodies.
//
// Restates the payload fluxel's AP-IBM builder hands to the grid
// (src/gridfoam/core/grid/axis_projected.py:163-211 loads it;
// docs/source/contributor_guide/architecture/data_contracts.rst:181-231
// defines it): per internal face, whether the axis ray between the owner's
// and the neighbour's centre meets the body surface (endpoints included),
// the distances along the axis from each centre to the first hit on its side,
// the near-boundary flags of Gibou et al. (distance / dx <= dx / L, dx the
// cell's width along the axis, L the largest domain extent), the patch of
// the body hit on each side and the body's outward unit normal at each hit
// (it stands in for the surface-mesh face the reference anchors a side to,
// post/forces/sampling.py:140-166).  The ray runs through the face centre (its
// tangential coordinates), so a 2:1 face is cut where its finer side's
// centre line meets the body.
//
// Bodies: a plane normal to one axis (open surface, both sides fluid); an
// axis-aligned box; a prism along z over a 2-D Fourier outline
// r(theta) = r0 (1 + sum_k a_k cos k theta + b_k sin k theta) around (cx, cy);
// a bluff body (a box with rounded nose edges, a rear slant and struts,
// gf_shapes.h).  One thread per face; the Fourier outline and the bluff body
// are intersected by bisection on the sign of their implicit function, which
// changes once along a segment shorter than the body's features.
#include <cmath>

#include "gf_common.h"
#include "gf_shapes.h"

namespace gf {
namespace {

constexpr int kMaxShapes = 16;
struct Shapes {
    int n;
    gf_shape s[kMaxShapes];
};

// Signed outline function of a Fourier body: negative inside.
__device__ double fourier_g(const gf_shape &b, double x, double y) {
    const double dx = x - b.p[0], dy = y - b.p[1];
    const double rho = sqrt(dx * dx + dy * dy);
    const double th = atan2(dy, dx);
    double r = 1.0;
    for (int k = 1; k <= 8; ++k) r += b.p[2 + k] * cos(k * th) + b.p[10 + k] * sin(k * th);
    return rho - b.p[2] * r;
}

// Implicit function of a Fourier (kind 2) or bluff (kind 3) body at q.
__device__ double implicit_g(const gf_shape &b, const double q[3]) {
    return b.kind == 2 ? fourier_g(b, q[0], q[1]) : bluff_g(b, q[0], q[1], q[2]);
}

// Hit of the segment s in [s0, s1] (axis a, other coordinates those of P) with
// body b, seen from the s0 end (from_lo) or from the s1 end; false if none.
__device__ bool body_hit(const gf_shape &b, int a, const double P[3], double s0, double s1, bool from_lo,
                         double *hit) {
    if (b.kind == 0) {  // plane normal to b.axis at p[0]
        if (b.axis != a) return false;
        const double s = b.p[0];
        if (s < s0 || s > s1) return false;
        *hit = s;
        return true;
    }
    if (b.kind == 1) {  // box [p0..p2] x [p3..p5]
        for (int j = 0; j < 3; ++j)
            if (j != a && (P[j] < b.p[j] || P[j] > b.p[3 + j])) return false;
        const double lo = b.p[a], hi = b.p[3 + a];
        const bool in_lo = lo >= s0 && lo <= s1, in_hi = hi >= s0 && hi <= s1;
        if (!in_lo && !in_hi) return false;
        if (in_lo && in_hi) {
            *hit = from_lo ? lo : hi;
        } else {
            *hit = in_lo ? lo : hi;
        }
        return true;
    }
    if (b.kind == 2 || b.kind == 3) {  // Fourier prism along z, bluff body
        if (b.kind == 2 && a == 2) return false;
        double q0[3] = {P[0], P[1], P[2]}, q1[3] = {P[0], P[1], P[2]};
        q0[a] = s0;
        q1[a] = s1;
        double g0 = implicit_g(b, q0);
        const double g1 = implicit_g(b, q1);
        if (g0 * g1 > 0.0) return false;
        if (g0 == 0.0) {
            *hit = s0;
            return true;
        }
        if (g1 == 0.0) {
            *hit = s1;
            return true;
        }
        double lo = s0, hi = s1;
        for (int it = 0; it < 64; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (mid <= lo || mid >= hi) break;
            double q[3] = {P[0], P[1], P[2]};
            q[a] = mid;
            const double gm = implicit_g(b, q);
            if ((gm <= 0.0) == (g0 <= 0.0)) {
                lo = mid;
                g0 = gm;
            } else {
                hi = mid;
            }
        }
        *hit = 0.5 * (lo + hi);
        return true;
    }
    return false;
}

// Body-outward unit normal of body b at the hit point (P with coordinate a
// replaced by s); at_lo: a box hit on its lower face along the axis.
__device__ void body_normal(const gf_shape &b, int a, const double P[3], double s, bool at_lo, double n[3]) {
    n[0] = n[1] = n[2] = 0.0;
    if (b.kind == 0) {  // the plane's own +axis normal
        n[b.axis] = 1.0;
    } else if (b.kind == 1) {
        n[a] = at_lo ? -1.0 : 1.0;
    } else if (b.kind == 3) {
#pragma clang fp contract(off)
        // central differences of the implicit function (step 1e-7)
        const double e = 1e-7;
        double q[3] = {P[0], P[1], P[2]};
        q[a] = s;
        double gr[3];
        for (int j = 0; j < 3; ++j) {
            double qp[3] = {q[0], q[1], q[2]}, qm[3] = {q[0], q[1], q[2]};
            qp[j] = q[j] + e;
            qm[j] = q[j] - e;
            gr[j] = (bluff_g(b, qp[0], qp[1], qp[2]) - bluff_g(b, qm[0], qm[1], qm[2])) / (2.0 * e);
        }
        const double gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
        for (int j = 0; j < 3; ++j) n[j] = gr[j] / gn;
    } else {
        // grad(|P - c| - R r(theta)) in the plane of the outline
        double q[3] = {P[0], P[1], P[2]};
        q[a] = s;
        const double dx = q[0] - b.p[0], dy = q[1] - b.p[1];
        const double rho2 = dx * dx + dy * dy, rho = sqrt(rho2);
        const double th = atan2(dy, dx);
        double dr = 0.0;  // r'(theta)
        for (int k = 1; k <= 8; ++k) dr += k * (b.p[10 + k] * cos(k * th) - b.p[2 + k] * sin(k * th));
        const double Rdr = b.p[2] * dr;
        const double gx = dx / rho + Rdr * dy / rho2, gy = dy / rho - Rdr * dx / rho2;
        const double gn = sqrt(gx * gx + gy * gy);
        n[0] = gx / gn;
        n[1] = gy / gn;
    }
}

__global__ void k_ap_cut(int F, const int32_t *__restrict__ owner, const int32_t *__restrict__ neighbour,
                         const int32_t *__restrict__ axis, const double *__restrict__ centers,
                         const double *__restrict__ sizes, const double *__restrict__ face_centers, Shapes sh,
                         double L, uint8_t *is_imm, double *dist_o, double *dist_n, uint8_t *near_o,
                         uint8_t *near_n, int32_t *patch_o, int32_t *patch_n, double *nrm_o, double *nrm_n) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int o = owner[f], n = neighbour[f], a = axis[f];
    const double P[3] = {face_centers[3 * f + 0], face_centers[3 * f + 1], face_centers[3 * f + 2]};
    const double s0 = centers[3 * o + a], s1 = centers[3 * n + a];
    double best_o = INFINITY, best_n = INFINITY, hit_o = 0.0, hit_n = 0.0;
    int po = -1, pn = -1, ko = 0, kn = 0;
    for (int k = 0; k < sh.n; ++k) {
        double h;
        if (body_hit(sh.s[k], a, P, s0, s1, true, &h) && h - s0 < best_o) {
            best_o = h - s0;
            hit_o = h;
            po = sh.s[k].patch;
            ko = k;
        }
        if (body_hit(sh.s[k], a, P, s0, s1, false, &h) && s1 - h < best_n) {
            best_n = s1 - h;
            hit_n = h;
            pn = sh.s[k].patch;
            kn = k;
        }
    }
    const bool cut = po >= 0;
    is_imm[f] = cut;
    dist_o[f] = cut ? best_o : 0.0;
    dist_n[f] = cut ? best_n : 0.0;
    patch_o[f] = po;
    patch_n[f] = pn;
    const double dxo = sizes[3 * o + a], dxn = sizes[3 * n + a];
    near_o[f] = cut && best_o / dxo <= dxo / L;
    near_n[f] = cut && best_n / dxn <= dxn / L;
    double no[3] = {0.0, 0.0, 0.0}, nn[3] = {0.0, 0.0, 0.0};
    if (cut) {
        const gf_shape &bo = sh.s[ko], &bn = sh.s[kn];
        body_normal(bo, a, P, hit_o, bo.kind == 1 && hit_o == bo.p[a], no);
        body_normal(bn, a, P, hit_n, bn.kind == 1 && hit_n == bn.p[a], nn);
    }
    for (int j = 0; j < 3; ++j) {
        nrm_o[3 * f + j] = no[j];
        nrm_n[3 * f + j] = nn[j];
    }
}

}  // namespace
}  // namespace gf

using namespace gf;

extern "C" int gf_ap_cut(int32_t n_faces, const int32_t *owner, const int32_t *neighbour, const int32_t *axis,
                         const double *centers, const double *sizes, const double *face_centers, int32_t n_shapes,
                         const gf_shape *shapes, double ref_length, uint8_t *is_immersed, double *dist_owner,
                         double *dist_neighbour, uint8_t *near_owner, uint8_t *near_neighbour, int32_t *patch_owner,
                         int32_t *patch_neighbour, double *normal_owner, double *normal_neighbour,
                         gf_stream_t stream) {
    GF_REQUIRE(n_faces >= 0 && n_shapes >= 0 && n_shapes <= kMaxShapes, "bad sizes (at most 16 bodies)");
    GF_REQUIRE(ref_length > 0.0, "reference length must be positive");
    if (n_faces == 0) return GF_OK;
    GF_REQUIRE(owner && neighbour && axis && centers && sizes && face_centers && is_immersed && dist_owner &&
                   dist_neighbour && near_owner && near_neighbour && patch_owner && patch_neighbour &&
                   normal_owner && normal_neighbour && (shapes || n_shapes == 0),
               "null argument");
    Shapes sh{};
    sh.n = n_shapes;
    for (int k = 0; k < n_shapes; ++k) {
        GF_REQUIRE(shapes[k].kind >= 0 && shapes[k].kind <= 3, "unknown body kind");
        GF_REQUIRE(shapes[k].kind != 3 || (shapes[k].axis >= 0 && shapes[k].axis <= 4), "at most 4 struts");
        GF_REQUIRE(shapes[k].patch >= 0, "body patch id must be >= 0");
        GF_REQUIRE(shapes[k].kind != 0 || (shapes[k].axis >= 0 && shapes[k].axis <= 2), "plane axis must be 0..2");
        sh.s[k] = shapes[k];
    }
    hipLaunchKernelGGL(k_ap_cut, dim3(grid_for(n_faces)), dim3(kBlock), 0, S(stream), n_faces, owner, neighbour,
                       axis, centers, sizes, face_centers, sh, ref_length, is_immersed, dist_owner, dist_neighbour,
                       near_owner, near_neighbour, patch_owner, patch_neighbour, normal_owner, normal_neighbour);
    GF_CHECK_LAUNCH();
    return GF_OK;
}
