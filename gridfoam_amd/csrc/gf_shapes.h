// Implicit functions of the analytic immersed bodies (negative inside, zero on
// the surface), shared by the axis-projected cut (immersed.hip) and the
// octree's surface refinement (octree.hip).
//
// kind 1, axis-aligned box p[0..2] - p[3..5]:
//   G = max(lo - x, x - hi) over the three axes.
// kind 3, bluff body: the box p[0..2] - p[3..5] (x0 its nose), with
//   - the four nose edges rounded with radius p[6]: for x < x0 + r,
//     G >= max(|(a, b_y)|, |(a, b_z)|) - r, where a = x0 + r - x and b_y, b_z
//     are the distances past the rounding's inner lines;
//   - the top rear edge cut by a slant plane, with outward normal
//     (p[7], 0, p[8]) = (sin, 0, cos) of the slant angle, through the top
//     face p[9] ahead of the rear;
//   - n = `axis` vertical struts (at most 4) with centres (p[12+2k],
//     p[13+2k]) and radius p[10], running from z0 - p[11] up into the body.
// Every piece is a max / min of 1-Lipschitz functions, so |G(x)| bounds the
// distance from x to the surface from below: the octree's surface test
// relies on that.  Evaluated without contraction (no fused multiply-add), so
// the host restatement of the same operations gives the same bits.
#pragma once

#include <cmath>

#include "gf_common.h"

namespace gf {

__host__ __device__ inline double box_g(const gf_shape &b, double x, double y, double z) {
#pragma clang fp contract(off)
    const double gx = fmax(b.p[0] - x, x - b.p[3]);
    const double gy = fmax(b.p[1] - y, y - b.p[4]);
    const double gz = fmax(b.p[2] - z, z - b.p[5]);
    return fmax(fmax(gx, gy), gz);
}

__host__ __device__ inline double bluff_g(const gf_shape &b, double x, double y, double z) {
#pragma clang fp contract(off)
    const double x0 = b.p[0], y0 = b.p[1], z0 = b.p[2], x1 = b.p[3], y1 = b.p[4], z1 = b.p[5];
    const double r = b.p[6];
    double g = box_g(b, x, y, z);
    if (x < x0 + r) {
        const double a = (x0 + r) - x;
        const double by = fmax(fmax((y0 + r) - y, y - (y1 - r)), 0.0);
        const double bz = fmax(fmax((z0 + r) - z, z - (z1 - r)), 0.0);
        g = fmax(g, fmax(sqrt(a * a + by * by) - r, sqrt(a * a + bz * bz) - r));
    }
    g = fmax(g, b.p[7] * (x - (x1 - b.p[9])) + b.p[8] * (z - z1));
    for (int k = 0; k < b.axis && k < 4; ++k) {
        const double dx = x - b.p[12 + 2 * k], dy = y - b.p[13 + 2 * k];
        const double gs = fmax(sqrt(dx * dx + dy * dy) - b.p[10], fmax((z0 - b.p[11]) - z, z - (z0 + b.p[10])));
        g = fmin(g, gs);
    }
    return g;
}

}  // namespace gf
