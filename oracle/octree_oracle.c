/*
 * ORACLE -- test infrastructure only.  Never linked into the product.
 *
 * CPU restatement of the octree mesh build that gridfoam obtains from its
 * external mesh builder (fluxel-cfd >=0.2.0,<0.3.0, pyproject.toml:36), as it
 * is consumed by src/gridfoam/core/grid/axis_projected.py:97-161 and
 * configured by src/gridfoam/meta/config.py:90-145:
 *
 *   root block grid  -> refine every leaf that meets an analytic body's
 *   surface below target_level (meta/config.py:100-106) or intersects a
 *   refinement box below the box level -> 2:1 face balance -> n_leaf
 *   uniform splits.
 *
 * Surface test: |G(leaf centre)| <= half the leaf's diagonal, with G the
 * body's implicit function (plane: signed offset along its axis; box:
 * max(lo - x, x - hi) over the axes; bluff body: that box with its nose
 * edges rounded, a rear slant plane and vertical struts).  These are
 * restated from their definition, operation for operation, so that without
 * contraction both builds round alike.
 *
 * Output contract (axis_projected.py:97-124, :763-767, :781-782):
 *   internal faces (owner, neighbour, axis) with the neighbour on the +axis
 *   side of the owner; domain faces (owner, dir) with axis = dir / 2 and
 *   sign = 2 (dir % 2) - 1.
 *
 * Canonical order (fluxel's own order is not observable here, so the order
 * is defined by this build and the GPU builder must match it bit for bit):
 *   cells      : ascending Morton key of the cell's min corner expressed at
 *                the finest level (depth-first leaf order);
 *   int. faces : by owner cell, then axis 0..2, then neighbour cell index;
 *   dom. faces : by owner cell, then dir 0..5.
 *
 * Algorithm here is deliberately different from the GPU one: leaves are
 * kept as an unsorted list, re-sorted with qsort after every pass, and
 * neighbours are found by predecessor binary search over the sorted keys.
 *
 * Floating-point box tests use explicit fma() so the GPU builder, which
 * uses the same correctly-rounded fma, takes identical decisions.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t key; /* Morton key of min corner at level lmax */
    int32_t i, j, k;
    int32_t level;
} leaf_t;

typedef struct {
    int32_t kind, axis; /* kind 0 plane, 1 box, 3 bluff body; axis: plane axis / strut count */
    double p[20];
} shape_t;

typedef struct {
    double lower[3], upper[3];
    int32_t root[3];
    int32_t lmax;
    int32_t n_regions;
    double *rmin, *rmax;
    int32_t *rlevel;
    int32_t n_shapes, target_level;
    shape_t *shapes;
    int32_t n_leaf;
    /* results */
    leaf_t *cells;
    int64_t n_cells;
    int32_t *f_owner, *f_nbr, *f_axis;
    int64_t n_faces;
    int32_t *b_owner, *b_dir;
    int64_t n_bfaces;
} oct_t;

static uint64_t spread3(uint64_t v) {
    /* place the low 21 bits of v at every third bit */
    uint64_t x = v & 0x1fffffULL;
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((x >> b) & 1ULL) << (3 * b);
    return r;
}

static uint64_t morton(uint64_t x, uint64_t y, uint64_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

static void set_key(const oct_t *o, leaf_t *c) {
    int s = o->lmax - c->level;
    c->key = morton((uint64_t)c->i << s, (uint64_t)c->j << s, (uint64_t)c->k << s);
}

static int cmp_leaf(const void *a, const void *b) {
    uint64_t ka = ((const leaf_t *)a)->key, kb = ((const leaf_t *)b)->key;
    return (ka > kb) - (ka < kb);
}

static double cell_h(const oct_t *o, int axis, int level) {
    return (o->upper[axis] - o->lower[axis]) / ldexp((double)o->root[axis], level);
}

static int intersects_region(const oct_t *o, const leaf_t *c, int r) {
    const int32_t idx[3] = {c->i, c->j, c->k};
    for (int a = 0; a < 3; ++a) {
        double h = cell_h(o, a, c->level);
        double lo = fma((double)idx[a], h, o->lower[a]);
        double hi = fma((double)(idx[a] + 1), h, o->lower[a]);
        if (!(lo < o->rmax[3 * r + a] && hi > o->rmin[3 * r + a])) return 0;
    }
    return 1;
}

static double box_g(const shape_t *b, double x, double y, double z) {
    double gx = fmax(b->p[0] - x, x - b->p[3]);
    double gy = fmax(b->p[1] - y, y - b->p[4]);
    double gz = fmax(b->p[2] - z, z - b->p[5]);
    return fmax(fmax(gx, gy), gz);
}

/* Bluff body: box p[0..2] - p[3..5]; nose edges rounded by p[6]; slant
 * plane with normal (p[7], 0, p[8]) through the top p[9] ahead of the rear;
 * `axis` struts of radius p[10] at (p[12+2k], p[13+2k]) reaching p[11] below. */
static double bluff_g(const shape_t *b, double x, double y, double z) {
    double x0 = b->p[0], y0 = b->p[1], z0 = b->p[2], x1 = b->p[3], y1 = b->p[4], z1 = b->p[5];
    double r = b->p[6];
    double g = box_g(b, x, y, z);
    if (x < x0 + r) {
        double a = (x0 + r) - x;
        double by = fmax(fmax((y0 + r) - y, y - (y1 - r)), 0.0);
        double bz = fmax(fmax((z0 + r) - z, z - (z1 - r)), 0.0);
        g = fmax(g, fmax(sqrt(a * a + by * by) - r, sqrt(a * a + bz * bz) - r));
    }
    g = fmax(g, b->p[7] * (x - (x1 - b->p[9])) + b->p[8] * (z - z1));
    for (int k = 0; k < b->axis && k < 4; ++k) {
        double dx = x - b->p[12 + 2 * k], dy = y - b->p[13 + 2 * k];
        double gs = fmax(sqrt(dx * dx + dy * dy) - b->p[10], fmax((z0 - b->p[11]) - z, z - (z0 + b->p[10])));
        g = fmin(g, gs);
    }
    return g;
}

static int near_surface(const oct_t *o, const leaf_t *c) {
    const int32_t idx[3] = {c->i, c->j, c->k};
    double ctr[3], h[3];
    for (int a = 0; a < 3; ++a) {
        h[a] = cell_h(o, a, c->level);
        ctr[a] = fma((double)idx[a] + 0.5, h[a], o->lower[a]);
    }
    double hd = 0.5 * sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (int k = 0; k < o->n_shapes; ++k) {
        const shape_t *b = &o->shapes[k];
        double g;
        if (b->kind == 0)
            g = ctr[b->axis] - b->p[0];
        else if (b->kind == 1)
            g = box_g(b, ctr[0], ctr[1], ctr[2]);
        else
            g = bluff_g(b, ctr[0], ctr[1], ctr[2]);
        if (fabs(g) <= hd) return 1;
    }
    return 0;
}

/* Replace every flagged leaf by its 8 children. */
static void split_flagged(oct_t *o, const uint8_t *flag, int64_t n_flag) {
    int64_t n_new = o->n_cells + 7 * n_flag;
    leaf_t *out = (leaf_t *)malloc(sizeof(leaf_t) * (size_t)(n_new > 0 ? n_new : 1));
    int64_t w = 0;
    for (int64_t c = 0; c < o->n_cells; ++c) {
        leaf_t p = o->cells[c];
        if (!flag[c]) {
            out[w++] = p;
            continue;
        }
        for (int ch = 0; ch < 8; ++ch) {
            leaf_t q;
            q.level = p.level + 1;
            q.i = 2 * p.i + (ch & 1);
            q.j = 2 * p.j + ((ch >> 1) & 1);
            q.k = 2 * p.k + ((ch >> 2) & 1);
            set_key(o, &q);
            out[w++] = q;
        }
    }
    free(o->cells);
    o->cells = out;
    o->n_cells = n_new;
    qsort(o->cells, (size_t)o->n_cells, sizeof(leaf_t), cmp_leaf);
}

/* Index of the leaf containing the finest-level point (x, y, z), or -1. */
static int64_t find_leaf(const oct_t *o, int64_t x, int64_t y, int64_t z) {
    int64_t ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = (int64_t)o->root[a] << o->lmax;
    if (x < 0 || y < 0 || z < 0 || x >= ext[0] || y >= ext[1] || z >= ext[2]) return -1;
    uint64_t key = morton((uint64_t)x, (uint64_t)y, (uint64_t)z);
    int64_t lo = 0, hi = o->n_cells - 1, ans = -1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (o->cells[mid].key <= key) {
            ans = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    return ans;
}

/* Finest-level point just across face `dir` of cell c (min corner on the
 * tangential axes). */
static void across(const oct_t *o, const leaf_t *c, int dir, int64_t p[3]) {
    int s = o->lmax - c->level;
    int64_t lo[3] = {(int64_t)c->i << s, (int64_t)c->j << s, (int64_t)c->k << s};
    int a = dir / 2;
    for (int b = 0; b < 3; ++b) p[b] = lo[b];
    if (dir % 2) p[a] = lo[a] + ((int64_t)1 << s);
    else p[a] = lo[a] - 1;
}

static int64_t refine_regions(oct_t *o) {
    int64_t total = 0;
    for (;;) {
        uint8_t *flag = (uint8_t *)calloc((size_t)o->n_cells, 1);
        int64_t n_flag = 0;
        for (int64_t c = 0; c < o->n_cells; ++c) {
            if (o->cells[c].level < o->target_level && near_surface(o, &o->cells[c])) {
                flag[c] = 1;
                ++n_flag;
                continue;
            }
            for (int r = 0; r < o->n_regions; ++r) {
                if (o->cells[c].level < o->rlevel[r] && intersects_region(o, &o->cells[c], r)) {
                    flag[c] = 1;
                    ++n_flag;
                    break;
                }
            }
        }
        if (n_flag) split_flagged(o, flag, n_flag);
        free(flag);
        total += n_flag;
        if (!n_flag) return total;
    }
}

static int64_t balance(oct_t *o) {
    int64_t total = 0;
    for (;;) {
        uint8_t *flag = (uint8_t *)calloc((size_t)o->n_cells, 1);
        int64_t n_flag = 0;
        for (int64_t c = 0; c < o->n_cells; ++c) {
            const leaf_t *cc = &o->cells[c];
            for (int dir = 0; dir < 6; ++dir) {
                int64_t p[3];
                across(o, cc, dir, p);
                int64_t n = find_leaf(o, p[0], p[1], p[2]);
                if (n >= 0 && o->cells[n].level < cc->level - 1 && !flag[n]) {
                    flag[n] = 1;
                    ++n_flag;
                }
            }
        }
        if (n_flag) split_flagged(o, flag, n_flag);
        free(flag);
        total += n_flag;
        if (!n_flag) return total;
    }
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static void enumerate_faces(oct_t *o) {
    /* pass 0 counts, pass 1 fills */
    for (int pass = 0; pass < 2; ++pass) {
        int64_t nf = 0, nb = 0;
        for (int64_t c = 0; c < o->n_cells; ++c) {
            const leaf_t *cc = &o->cells[c];
            const int32_t idx[3] = {cc->i, cc->j, cc->k};
            for (int a = 0; a < 3; ++a) {
                int64_t p[3];
                across(o, cc, 2 * a + 1, p);
                int64_t n = find_leaf(o, p[0], p[1], p[2]);
                if (n < 0) continue;
                if (o->cells[n].level <= cc->level) {
                    if (pass) {
                        o->f_owner[nf] = (int32_t)c;
                        o->f_nbr[nf] = (int32_t)n;
                        o->f_axis[nf] = a;
                    }
                    ++nf;
                } else {
                    /* four finer leaves cover the +a face */
                    int32_t nb4[4];
                    int s = o->lmax - (cc->level + 1);
                    int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
                    for (int q = 0; q < 4; ++q) {
                        int64_t pt[3];
                        pt[a] = (int64_t)(2 * (idx[a] + 1)) << s;
                        pt[t1] = (int64_t)(2 * idx[t1] + (q & 1)) << s;
                        pt[t2] = (int64_t)(2 * idx[t2] + (q >> 1)) << s;
                        nb4[q] = (int32_t)find_leaf(o, pt[0], pt[1], pt[2]);
                    }
                    qsort(nb4, 4, sizeof(int32_t), cmp_i32);
                    for (int q = 0; q < 4; ++q) {
                        if (pass) {
                            o->f_owner[nf] = (int32_t)c;
                            o->f_nbr[nf] = nb4[q];
                            o->f_axis[nf] = a;
                        }
                        ++nf;
                    }
                }
            }
            for (int dir = 0; dir < 6; ++dir) {
                int a = dir / 2;
                int64_t extent = (int64_t)o->root[a] << cc->level;
                int on_bnd = (dir % 2) ? (idx[a] + 1 == extent) : (idx[a] == 0);
                if (!on_bnd) continue;
                if (pass) {
                    o->b_owner[nb] = (int32_t)c;
                    o->b_dir[nb] = dir;
                }
                ++nb;
            }
        }
        if (!pass) {
            o->n_faces = nf;
            o->n_bfaces = nb;
            o->f_owner = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nf + 1));
            o->f_nbr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nf + 1));
            o->f_axis = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nf + 1));
            o->b_owner = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nb + 1));
            o->b_dir = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nb + 1));
        }
    }
}

/* ---- public C API used by the tests through ctypes ---------------------- */

void *oct_oracle_build_surface(const double *lower, const double *upper, const int32_t *root,
                               int32_t n_regions, const double *rmin, const double *rmax,
                               const int32_t *rlevel, int32_t n_shapes, const int32_t *kind,
                               const int32_t *axis, const double *params, int32_t target_level,
                               int32_t n_leaf, int64_t *counts_out);

void *oct_oracle_build(const double *lower, const double *upper, const int32_t *root,
                       int32_t n_regions, const double *rmin, const double *rmax,
                       const int32_t *rlevel, int32_t n_leaf, int64_t *counts_out) {
    return oct_oracle_build_surface(lower, upper, root, n_regions, rmin, rmax, rlevel, 0, NULL, NULL, NULL, 0,
                                    n_leaf, counts_out);
}

/* kind/axis [n_shapes], params [n_shapes][20] */
void *oct_oracle_build_surface(const double *lower, const double *upper, const int32_t *root,
                               int32_t n_regions, const double *rmin, const double *rmax,
                               const int32_t *rlevel, int32_t n_shapes, const int32_t *kind,
                               const int32_t *axis, const double *params, int32_t target_level,
                               int32_t n_leaf, int64_t *counts_out) {
    oct_t *o = (oct_t *)calloc(1, sizeof(oct_t));
    o->n_shapes = n_shapes;
    o->target_level = n_shapes > 0 ? target_level : 0;
    o->shapes = (shape_t *)calloc((size_t)(n_shapes + 1), sizeof(shape_t));
    for (int k = 0; k < n_shapes; ++k) {
        o->shapes[k].kind = kind[k];
        o->shapes[k].axis = axis[k];
        memcpy(o->shapes[k].p, params + 20 * k, sizeof(double) * 20);
    }
    memcpy(o->lower, lower, sizeof(double) * 3);
    memcpy(o->upper, upper, sizeof(double) * 3);
    memcpy(o->root, root, sizeof(int32_t) * 3);
    o->n_regions = n_regions;
    o->rmin = (double *)malloc(sizeof(double) * 3 * (size_t)(n_regions + 1));
    o->rmax = (double *)malloc(sizeof(double) * 3 * (size_t)(n_regions + 1));
    o->rlevel = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_regions + 1));
    int32_t lmax = o->target_level;
    for (int r = 0; r < n_regions; ++r) {
        for (int a = 0; a < 3; ++a) {
            o->rmin[3 * r + a] = rmin[3 * r + a];
            o->rmax[3 * r + a] = rmax[3 * r + a];
        }
        o->rlevel[r] = rlevel[r];
        if (rlevel[r] > lmax) lmax = rlevel[r];
    }
    o->n_leaf = n_leaf;
    o->lmax = lmax + n_leaf;

    o->n_cells = (int64_t)root[0] * root[1] * root[2];
    o->cells = (leaf_t *)malloc(sizeof(leaf_t) * (size_t)o->n_cells);
    int64_t w = 0;
    for (int32_t k = 0; k < root[2]; ++k)
        for (int32_t j = 0; j < root[1]; ++j)
            for (int32_t i = 0; i < root[0]; ++i) {
                leaf_t c = {0, i, j, k, 0};
                set_key(o, &c);
                o->cells[w++] = c;
            }
    qsort(o->cells, (size_t)o->n_cells, sizeof(leaf_t), cmp_leaf);

    refine_regions(o);
    balance(o);
    for (int r = 0; r < n_leaf; ++r) {
        uint8_t *flag = (uint8_t *)malloc((size_t)o->n_cells);
        memset(flag, 1, (size_t)o->n_cells);
        split_flagged(o, flag, o->n_cells);
        free(flag);
    }
    enumerate_faces(o);
    counts_out[0] = o->n_cells;
    counts_out[1] = o->n_faces;
    counts_out[2] = o->n_bfaces;
    counts_out[3] = o->lmax;
    return o;
}

void oct_oracle_fetch(void *h, int32_t *level, int32_t *ijk, int32_t *f_owner, int32_t *f_nbr,
                      int32_t *f_axis, int32_t *b_owner, int32_t *b_dir) {
    oct_t *o = (oct_t *)h;
    for (int64_t c = 0; c < o->n_cells; ++c) {
        level[c] = o->cells[c].level;
        ijk[3 * c + 0] = o->cells[c].i;
        ijk[3 * c + 1] = o->cells[c].j;
        ijk[3 * c + 2] = o->cells[c].k;
    }
    memcpy(f_owner, o->f_owner, sizeof(int32_t) * (size_t)o->n_faces);
    memcpy(f_nbr, o->f_nbr, sizeof(int32_t) * (size_t)o->n_faces);
    memcpy(f_axis, o->f_axis, sizeof(int32_t) * (size_t)o->n_faces);
    memcpy(b_owner, o->b_owner, sizeof(int32_t) * (size_t)o->n_bfaces);
    memcpy(b_dir, o->b_dir, sizeof(int32_t) * (size_t)o->n_bfaces);
}

void oct_oracle_free(void *h) {
    oct_t *o = (oct_t *)h;
    if (!o) return;
    free(o->cells);
    free(o->f_owner);
    free(o->f_nbr);
    free(o->f_axis);
    free(o->b_owner);
    free(o->b_dir);
    free(o->rmin);
    free(o->rmax);
    free(o->rlevel);
    free(o->shapes);
    free(o);
}
