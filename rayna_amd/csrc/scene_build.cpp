// Host-side scene preparation: the engine's object graph (object/list.rs:59-78), its SAH trees
// (shared/generic_bvh.rs:45-290), depth-first primitive ranks, tight refit and the BVH4 device layout.
#include "scene_build.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

namespace rayna {

namespace {

Aabb64 box_new(const double a[3], const double b[3]) {  // Aabb::new (shared/aabb.rs:85-99)
    Aabb64 r;
    for (int i = 0; i < 3; ++i) {
        r.mn[i] = std::fmin(a[i], b[i]);
        r.mx[i] = std::fmax(a[i], b[i]);
    }
    return r;
}

Aabb64 origin_box() { return Aabb64{{0, 0, 0}, {0, 0, 0}}; }  // Aabb::default()

void grow(Aabb64 &acc, const Aabb64 &b) {  // Aabb::encompass (aabb.rs:56-61)
    for (int i = 0; i < 3; ++i) {
        acc.mn[i] = std::fmin(acc.mn[i], b.mn[i]);
        acc.mx[i] = std::fmax(acc.mx[i], b.mx[i]);
    }
}

double area(const Aabb64 &b) {  // aabb.rs:90
    const double sx = b.mx[0] - b.mn[0], sy = b.mx[1] - b.mn[1], sz = b.mx[2] - b.mn[2];
    return ((sx * sy) + (sy * sz) + (sz * sx)) * 2.0;
}

Aabb64 points_box(const double *p, int n) {  // Aabb::encompass_points (aabb.rs:70-79)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < 3; ++i) {
            mn[i] = std::fmin(mn[i], p[3 * k + i]);
            mx[i] = std::fmax(mx[i], p[3 * k + i]);
        }
    return box_new(mn, mx);
}

Aabb64 min_padded(const Aabb64 &b, double thresh) {  // Aabb::min_padded (aabb.rs:84-89)
    double dims[3], centre[3], mn[3], mx[3];
    for (int i = 0; i < 3; ++i) {
        dims[i] = b.mx[i] - b.mn[i];
        centre[i] = b.mn[i] + dims[i] / 2.0;
        dims[i] = std::fmax(dims[i], thresh);
    }
    for (int i = 0; i < 3; ++i) {
        mn[i] = centre[i] - dims[i] / 2.0;
        mx[i] = centre[i] + dims[i] / 2.0;
    }
    return box_new(mn, mx);
}

// ---- ObjectTransform (object/transform.rs) ----
// An affine map is row-major 3x4: [m00 m01 m02 t0; m10 m11 m12 t1; m20 m21 m22 t2], p -> m p + t.  Points and
// vectors are mapped in the order of glam's transform_point3 / transform_vector3: ((c0 x + c1 y) + c2 z) + w.

bool affine_is_identity(const double *m) {  // ObjectTransform::is_identity (transform == Transform3::IDENTITY)
    static const double I[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (int i = 0; i < 12; ++i)
        if (!(m[i] == I[i])) return false;
    return true;
}

// The inverse of an affine map: the linear part by its adjugate over the determinant, the translation as
// -(inverse linear part) t.  False when the map is singular or the inverse is not finite.
bool affine_inverse(const double *m, double *inv) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    const double A = e * i - f * h, B = f * g - d * i, Cc = d * h - e * g;
    const double det = (a * A + b * B) + c * Cc;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double l[9] = {A / det, (c * h - b * i) / det, (b * f - c * e) / det,
                         B / det, (a * i - c * g) / det, (c * d - a * f) / det,
                         Cc / det, (b * g - a * h) / det, (a * e - b * d) / det};
    for (int r = 0; r < 3; ++r) {
        inv[4 * r] = l[3 * r];
        inv[4 * r + 1] = l[3 * r + 1];
        inv[4 * r + 2] = l[3 * r + 2];
        inv[4 * r + 3] = -((l[3 * r] * m[3] + l[3 * r + 1] * m[7]) + l[3 * r + 2] * m[11]);
    }
    for (int k = 0; k < 12; ++k)
        if (!std::isfinite(inv[k])) return false;
    return true;
}

void affine_point(const double *m, const double *p, double *out) {
    for (int r = 0; r < 3; ++r) out[r] = ((m[4 * r] * p[0] + m[4 * r + 1] * p[1]) + m[4 * r + 2] * p[2]) + m[4 * r + 3];
}

// ObjectTransform::calculate_aabb: the box around the eight mapped corners of the object-space box.
Aabb64 affine_box(const double *m, const Aabb64 &b) {
    double pts[24];
    for (int k = 0; k < 8; ++k) {
        const double c[3] = {(k & 1) ? b.mx[0] : b.mn[0], (k & 2) ? b.mx[1] : b.mn[1], (k & 4) ? b.mx[2] : b.mn[2]};
        affine_point(m, c, pts + 3 * k);
    }
    return points_box(pts, 8);
}

float round_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

// ---------------------------------------------------------------------------------------------------------
// The engine's SAH builder.  Each recursion level works on a range [lo, hi) of three index arrays that hold
// the range's items sorted by (aabb.min[axis], item index) for axis x, y, z; a split is a stable partition of
// all three, so no level sorts again.  The candidate splits, costs (unions folded from the origin, as
// Aabb::encompass_iter does) and tie rules are those of calculate_optimal_split::<1, 2>.
class Builder {
   public:
    explicit Builder(const std::vector<Aabb64> &boxes) : box_(boxes) {}

    RefTree build() {
        const int n = (int)box_.size();
        if (n == 0) return t_;
        for (int a = 0; a < 3; ++a) {
            ord_[a].resize(n);
            for (int i = 0; i < n; ++i) ord_[a][i] = i;
            std::sort(ord_[a].begin(), ord_[a].end(), [&](int x, int y) {
                const double kx = box_[x].mn[a], ky = box_[y].mn[a];
                if (kx < ky) return true;
                if (ky < kx) return false;
                return x < y;
            });
        }
        mark_.assign(n, 0);
        tmp_.resize(n);
        suf_.resize(n + 1);
        t_.root = generate(0, n, -1);
        return t_;
    }

   private:
    const std::vector<Aabb64> &box_;
    RefTree t_;
    std::vector<int32_t> ord_[3];
    std::vector<uint8_t> mark_;
    std::vector<int32_t> tmp_;
    std::vector<double> suf_;

    struct Split {
        int axis, left;
    };

    // calculate_optimal_split::<1, 2> on a range of more than two items (generic_bvh.rs:173-290)
    Split best(int lo, int hi) {
        const int n = hi - lo;
        Split s{0, 1};
        bool have = false;
        double best_cost = 0.0;
        for (int a = 0; a < 3; ++a) {
            const int32_t *o = &ord_[a][lo];
            Aabb64 acc = origin_box();
            for (int i = n - 1; i >= 1; --i) {
                grow(acc, box_[o[i]]);
                suf_[i] = area(acc);
            }
            acc = origin_box();
            for (int p = 1; p <= n - 2; ++p) {
                grow(acc, box_[o[p - 1]]);
                const double cost = (double)p * area(acc) + (double)(n - p) * suf_[p];
                if (!have || cost < best_cost) {  // std::cmp::min_by keeps the earlier of equal costs
                    have = true;
                    best_cost = cost;
                    s = Split{a, p};
                }
            }
        }
        return s;
    }

    // split_objects (generic_bvh.rs:156-168): the first `left` items along `axis` go left.
    void partition(int lo, int hi, int axis, int left) {
        for (int i = lo; i < lo + left; ++i) mark_[ord_[axis][i]] = 1;
        for (int a = 0; a < 3; ++a) {
            if (a == axis) continue;
            int w = lo, r = 0;
            for (int i = lo; i < hi; ++i) {
                const int32_t id = ord_[a][i];
                if (mark_[id]) ord_[a][w++] = id;
                else tmp_[r++] = id;
            }
            std::memcpy(&ord_[a][w], tmp_.data(), sizeof(int32_t) * r);
        }
        for (int i = lo; i < lo + left; ++i) mark_[ord_[axis][i]] = 0;
    }

    int add_node(int item) {
        t_.nodes.push_back(RefTree::Node{item, 0, 0});
        return (int)t_.nodes.size() - 1;
    }

    void set_children(int node, const std::vector<int32_t> &ch) {
        t_.nodes[node].first_child = (int32_t)t_.children.size();
        t_.nodes[node].n_children = (int32_t)ch.size();
        t_.children.insert(t_.children.end(), ch.begin(), ch.end());
    }

    // generate_nodes_sah (generic_bvh.rs:97-154).  `order_axis` is the axis of the sort that produced this
    // range's order in the engine (-1: the input order, only at the root).
    int generate(int lo, int hi, int order_axis) {
        const int n = hi - lo;
        if (n == 1) return add_node(ord_[0][lo]);
        if (n <= 8) {
            const int node = add_node(-1);
            std::vector<int32_t> ids;
            if (order_axis >= 0) ids.assign(ord_[order_axis].begin() + lo, ord_[order_axis].begin() + hi);
            else {
                ids.assign(ord_[0].begin() + lo, ord_[0].begin() + hi);
                std::sort(ids.begin(), ids.end());
            }
            std::vector<int32_t> ch;
            for (int32_t id : ids) ch.push_back(add_node(id));
            set_children(node, ch);
            return node;
        }
        const int node = add_node(-1);
        const Split s = best(lo, hi);
        partition(lo, hi, s.axis, s.left);
        std::vector<int32_t> ch;
        const int halves[2][2] = {{lo, lo + s.left}, {lo + s.left, hi}};
        for (int h = 0; h < 2; ++h) {
            const int hl = halves[h][0], hh = halves[h][1];
            if (hh - hl > 2) {
                const Split s2 = best(hl, hh);
                partition(hl, hh, s2.axis, s2.left);
                ch.push_back(generate(hl, hl + s2.left, s2.axis));
                ch.push_back(generate(hl + s2.left, hh, s2.axis));
            } else {
                ch.push_back(generate(hl, hh, s.axis));
            }
        }
        set_children(node, ch);
        return node;
    }
};

// ---------------------------------------------------------------------------------------------------------
// The traversal tree of a triangle list on the device.  The closest hit does not depend on the tree: culling is
// conservative and ties go to the higher depth-first rank of the engine's tree, which every triangle carries.
// So the kernels traverse a tree built for them.  The engine's SAH grows every candidate box from the origin
// (Aabb::encompass_iter, aabb.rs:64-67), so for a mesh away from the origin its costs barely depend on the
// split, and it stops at clusters of up to eight items.  This one is a binned SAH over centroids (32 bins per
// axis, tight boxes, a node costing two triangle tests) with leaves of at most RAYNA_LEAF_MAX triangles; its
// binary tree becomes 4-wide nodes when it is emitted, by opening the child of largest surface area until a
// node has four.
#ifndef RAYNA_LEAF_MAX
#define RAYNA_LEAF_MAX 4
#endif
#ifndef RAYNA_SAH_NODE_COST
#define RAYNA_SAH_NODE_COST 2.0  // a node's traversal cost in triangle tests
#endif
class TraversalTree {
   public:
    struct Node {
        Aabb64 box;
        int32_t left = -1, right = -1;  // children; left < 0: a leaf of items order[first, first + count)
        int32_t first = 0, count = 0;
    };
    std::vector<Node> nodes;
    std::vector<int32_t> order;

    explicit TraversalTree(const std::vector<Aabb64> &boxes) : box_(boxes) {
        const int n = (int)boxes.size();
        order.resize(n);
        cen_.resize(n);
        for (int i = 0; i < n; ++i) {
            order[i] = i;
            for (int a = 0; a < 3; ++a) cen_[i][a] = 0.5 * (boxes[i].mn[a] + boxes[i].mx[a]);
        }
        if (n > 0) build(0, n);
    }

   private:
    static constexpr int BINS = 32;
    const std::vector<Aabb64> &box_;
    std::vector<std::array<double, 3>> cen_;

    static Aabb64 empty_box() { return Aabb64{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}}; }
    static void add(Aabb64 &a, const Aabb64 &b) {
        for (int k = 0; k < 3; ++k) {
            a.mn[k] = std::min(a.mn[k], b.mn[k]);
            a.mx[k] = std::max(a.mx[k], b.mx[k]);
        }
    }
    static double half_area(const Aabb64 &b) {
        const double x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
        return x * y + y * z + z * x;
    }

    int build(int lo, int hi) {
        const int id = (int)nodes.size();
        nodes.emplace_back();
        Aabb64 box = empty_box(), cb = empty_box();
        for (int i = lo; i < hi; ++i) {
            add(box, box_[order[i]]);
            const Aabb64 c{{cen_[order[i]][0], cen_[order[i]][1], cen_[order[i]][2]},
                           {cen_[order[i]][0], cen_[order[i]][1], cen_[order[i]][2]}};
            add(cb, c);
        }
        nodes[id].box = box;
        const int n = hi - lo;
        // best binned split
        int best_axis = -1, best_bin = 0;
        double best_cost = INFINITY;
        for (int a = 0; a < 3; ++a) {
            const double ext = cb.mx[a] - cb.mn[a];
            if (!(ext > 0)) continue;
            Aabb64 bb[BINS];
            int cnt[BINS] = {};
            for (int b = 0; b < BINS; ++b) bb[b] = empty_box();
            const double scale = BINS / ext;
            for (int i = lo; i < hi; ++i) {
                const int b = std::min(BINS - 1, (int)((cen_[order[i]][a] - cb.mn[a]) * scale));
                ++cnt[b];
                add(bb[b], box_[order[i]]);
            }
            double right_area[BINS];
            int right_cnt[BINS];
            Aabb64 acc = empty_box();
            int c = 0;
            for (int b = BINS - 1; b >= 1; --b) {
                add(acc, bb[b]);
                c += cnt[b];
                right_area[b] = half_area(acc);
                right_cnt[b] = c;
            }
            acc = empty_box();
            c = 0;
            for (int b = 1; b < BINS; ++b) {
                add(acc, bb[b - 1]);
                c += cnt[b - 1];
                if (c == 0 || right_cnt[b] == 0) continue;
                const double cost = half_area(acc) * c + right_area[b] * right_cnt[b];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_bin = b;
                }
            }
        }
        const double parent = half_area(box);
        const double leaf_cost = (double)n;
        const double split_cost = parent > 0 ? RAYNA_SAH_NODE_COST + best_cost / parent : INFINITY;
        if (n <= RAYNA_LEAF_MAX && !(split_cost < leaf_cost)) {
            nodes[id].first = lo;
            nodes[id].count = n;
            return id;
        }
        int mid;
        if (best_axis >= 0) {
            const double ext = cb.mx[best_axis] - cb.mn[best_axis], scale = BINS / ext;
            const int32_t *m = std::partition(order.data() + lo, order.data() + hi, [&](int32_t it) {
                return std::min(BINS - 1, (int)((cen_[it][best_axis] - cb.mn[best_axis]) * scale)) < best_bin;
            });
            mid = (int)(m - order.data());
        } else {
            mid = lo + n / 2;  // coincident centroids: halves in input order
        }
        const int l = build(lo, mid);
        const int r = build(mid, hi);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

// ---------------------------------------------------------------------------------------------------------

struct Packer {
    const rayna_scene_desc &d;
    PackedScene &P;
    std::vector<RefTree> blas;              // per mesh (empty tree if not a triangle list)
    std::vector<std::vector<Aabb64>> blas_tight;  // per mesh, per node
    RefTree tlas;
    std::vector<Aabb64> tlas_tight;
    std::vector<int32_t> bounded;           // tlas item -> object
    std::vector<Aabb64> obj_tight;
    std::vector<int32_t> obj_xf;            // object -> index into xf_fwd / xf_inv (12 doubles each), -1: identity
    std::vector<double> xf_fwd, xf_inv;
    int32_t rank = 0;

    Packer(const rayna_scene_desc &dd, PackedScene &pp) : d(dd), P(pp) {}

    Aabb64 tri_box(int tri) const { return points_box(&d.tri_vertices[9 * (size_t)tri], 3); }

    // The box of a primitive mesh: sphere.rs:27-36, triangle.rs:33, parallelogram.rs:25-37, axis_box.rs:29-37.
    Aabb64 prim_box(int mesh) const {
        const rayna_mesh &m = d.meshes[mesh];
        switch (m.kind) {
            case RAYNA_MESH_SPHERE: {
                const double *s = &d.spheres[4 * (size_t)m.index];
                const double a[3] = {s[0] - s[3], s[1] - s[3], s[2] - s[3]};
                const double b[3] = {s[0] + s[3], s[1] + s[3], s[2] + s[3]};
                return box_new(a, b);
            }
            case RAYNA_MESH_TRIANGLE: return tri_box(m.index);
            case RAYNA_MESH_PARALLELOGRAM: {
                const double *p = &d.parallelograms[9 * (size_t)m.index];
                double pts[12];
                for (int i = 0; i < 3; ++i) {
                    pts[i] = p[i];
                    pts[3 + i] = p[i] + p[3 + i];
                    pts[6 + i] = p[i] + p[6 + i];
                    pts[9 + i] = p[i] + p[3 + i] + p[6 + i];
                }
                return min_padded(points_box(pts, 4), 1e-6);
            }
            default: {  // RAYNA_MESH_AXIS_BOX
                const double *b = &d.boxes[6 * (size_t)m.index];
                return box_new(b, b + 3);
            }
        }
    }

    static bool is_nested(const RefTree &t, int n) { return t.nodes[n].item < 0; }

    static bool all_items(const RefTree &t, int n) {
        const RefTree::Node &nd = t.nodes[n];
        for (int i = 0; i < nd.n_children; ++i)
            if (t.nodes[t.children[nd.first_child + i]].item < 0) return false;
        return true;
    }

    static std::vector<Aabb64> refit(const RefTree &t, const std::function<Aabb64(int)> &item_box) {
        std::vector<Aabb64> tight(t.nodes.size());
        // children always have larger indices than their parent in RefTree, so a reverse sweep is post-order
        for (int n = (int)t.nodes.size() - 1; n >= 0; --n) {
            const RefTree::Node &nd = t.nodes[n];
            if (nd.item >= 0) tight[n] = item_box(nd.item);
            else {
                Aabb64 b = tight[t.children[nd.first_child]];
                for (int i = 1; i < nd.n_children; ++i) grow(b, tight[t.children[nd.first_child + i]]);
                tight[n] = b;
            }
        }
        return tight;
    }

    static void set_slot_box(GNode &g, int k, const Aabb64 &b) {
        g.lo_x[k] = round_down(b.mn[0]);
        g.lo_y[k] = round_down(b.mn[1]);
        g.lo_z[k] = round_down(b.mn[2]);
        g.hi_x[k] = round_up(b.mx[0]);
        g.hi_y[k] = round_up(b.mx[1]);
        g.hi_z[k] = round_up(b.mx[2]);
    }

    static GNode empty_node() {
        GNode g;
        for (int k = 0; k < 4; ++k) {
            g.lo_x[k] = g.lo_y[k] = g.lo_z[k] = INFINITY;
            g.hi_x[k] = g.hi_y[k] = g.hi_z[k] = -INFINITY;
            g.child[k] = 0;
            g.info[k] = INFO_EMPTY;
        }
        return g;
    }

    // One triangle of the scene into the device arrays, ranked next in depth-first order; returns its index.
    int32_t add_triangle(int tri, int32_t material, int32_t object, int32_t local) {
        const int32_t i = push_triangle(tri, material, rank++);
        P.rank_object.push_back(object);
        P.rank_local.push_back(local);
        return i;
    }

    // One triangle into the device arrays with a given rank.
    int32_t push_triangle(int tri, int32_t material, int32_t tri_rank) {
        const double *v = &d.tri_vertices[9 * (size_t)tri];
        const double *nn = &d.tri_normals[9 * (size_t)tri];
        GTri g;
        const float v0[3] = {(float)v[0], (float)v[1], (float)v[2]};
        const float v1[3] = {(float)v[3], (float)v[4], (float)v[5]};
        const float v2[3] = {(float)v[6], (float)v[7], (float)v[8]};
        g.v0x = v0[0];
        g.v0y = v0[1];
        g.v0z = v0[2];
        g.e1x = v1[0] - v0[0];
        g.e1y = v1[1] - v0[1];
        g.e1z = v1[2] - v0[2];
        g.e2x = v2[0] - v0[0];
        g.e2y = v2[1] - v0[1];
        g.e2z = v2[2] - v0[2];
        g.rank = tri_rank;
        g.material = material;
        // Can sum_i b_i n_i vanish for barycentrics b?  Not if every n_i leans towards the mean direction.
        double c[3] = {nn[0] + nn[3] + nn[6], nn[1] + nn[4] + nn[7], nn[2] + nn[5] + nn[8]};
        const double cl = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        int flags = 1;
        if (cl > 0) {
            flags = 0;
            for (int i = 0; i < 3; ++i) {
                const double dd = (c[0] * nn[3 * i] + c[1] * nn[3 * i + 1] + c[2] * nn[3 * i + 2]) / cl;
                if (!(dd > 1e-3)) flags = 1;
            }
        }
        g.flags = flags;
        P.tris.push_back(g);
        GTriNormals tn;
        tn.n0x = (float)nn[0];
        tn.n0y = (float)nn[1];
        tn.n0z = (float)nn[2];
        tn.n1x = (float)nn[3];
        tn.n1y = (float)nn[4];
        tn.n1z = (float)nn[5];
        tn.n2x = (float)nn[6];
        tn.n2y = (float)nn[7];
        tn.n2z = (float)nn[8];
        tn.pad0 = tn.pad1 = tn.pad2 = 0;
        P.tri_normals.push_back(tn);
        return (int32_t)P.tris.size() - 1;
    }

    int32_t add_para(int idx) {  // Planar::new (planar/mod.rs:79-90) in double, then rounded
        const double *p = &d.parallelograms[9 * (size_t)idx];
        const double u[3] = {p[3], p[4], p[5]}, v[3] = {p[6], p[7], p[8]};
        const double nr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double l2 = (nr[0] * nr[0] + nr[1] * nr[1]) + nr[2] * nr[2];
        const double inv = 1.0 / std::sqrt(l2);
        const double n[3] = {nr[0] * inv, nr[1] * inv, nr[2] * inv};
        const double dd = -((n[0] * p[0] + n[1] * p[1]) + n[2] * p[2]);
        GPara g;
        g.nx = (float)n[0];
        g.ny = (float)n[1];
        g.nz = (float)n[2];
        g.d = (float)dd;
        g.wx = (float)(nr[0] / l2);
        g.wy = (float)(nr[1] / l2);
        g.wz = (float)(nr[2] / l2);
        g.px = (float)p[0];
        g.py = (float)p[1];
        g.pz = (float)p[2];
        g.ux = (float)u[0];
        g.uy = (float)u[1];
        g.uz = (float)u[2];
        g.vx = (float)v[0];
        g.vy = (float)v[1];
        g.vz = (float)v[2];
        g.pad0 = g.pad1 = g.pad2 = g.pad3 = 0;
        P.paras.push_back(g);
        return (int32_t)P.paras.size() - 1;
    }

    int32_t add_box(int idx) {  // AxisBoxMesh::new (axis_box.rs:29-37) in double, then rounded
        const double *b = &d.boxes[6 * (size_t)idx];
        GBox g;
        double c[3], r[3], iv[3];
        for (int i = 0; i < 3; ++i) {
            const double mn = std::fmin(b[i], b[3 + i]), mx = std::fmax(b[i], b[3 + i]);
            c[i] = (mn + mx) / 2.0;
            r[i] = (mx - mn) / 2.0;
            iv[i] = 1.0 / r[i];
        }
        g.cx = (float)c[0];
        g.cy = (float)c[1];
        g.cz = (float)c[2];
        g.rx = (float)r[0];
        g.ry = (float)r[1];
        g.rz = (float)r[2];
        g.ix = (float)iv[0];
        g.iy = (float)iv[1];
        g.iz = (float)iv[2];
        g.pad0 = g.pad1 = g.pad2 = 0;
        P.boxes.push_back(g);
        return (int32_t)P.boxes.size() - 1;
    }

    // ---- flattening ----

    // One primitive mesh of a BVH mesh as a leaf item, ranked next in depth-first order.
    GItem sub_item(int sub, int32_t material, int32_t object, int32_t local) {
        const rayna_mesh &sm = d.meshes[sub];
        GItem it{ITEM_NONE, 0, rank, material};
        switch (sm.kind) {
            case RAYNA_MESH_SPHERE:
                it.kind = ITEM_SPHERE;
                it.index = sm.index;
                break;
            case RAYNA_MESH_TRIANGLE:  // add_triangle ranks it
                it.kind = ITEM_TRIANGLE;
                it.index = add_triangle(sm.index, material, object, local);
                return it;
            case RAYNA_MESH_PARALLELOGRAM:
                it.kind = ITEM_PARALLELOGRAM;
                it.index = add_para(sm.index);
                break;
            default:  // RAYNA_MESH_AXIS_BOX
                it.kind = ITEM_AXIS_BOX;
                it.index = add_box(sm.index);
                break;
        }
        ++rank;
        P.rank_object.push_back(object);
        P.rank_local.push_back(local);
        return it;
    }

    // Leaf cluster of a mesh BVH in child order: triangles of a triangle list, or items of a BVH mesh.
    void blas_leaf(GNode &g, int k, const rayna_mesh &m, int32_t material, int32_t object,
                   const std::vector<int32_t> &items) {
        if (m.kind == RAYNA_MESH_BVH) {
            const int32_t start = (int32_t)P.items.size();
            for (int32_t it : items) {
                const GItem gi = sub_item(m.index + it, material, object, it);
                P.items.push_back(gi);
            }
            g.child[k] = start;
            g.info[k] = (int32_t)items.size() | (LEAF_ITEMS << 4);
            return;
        }
        const int32_t start = (int32_t)P.tris.size();
        for (int32_t it : items) add_triangle(m.index + it, material, object, it);
        g.child[k] = start;
        g.info[k] = (int32_t)items.size() | (LEAF_TRIS << 4);
    }

    // The items of an engine tree in the order the engine's depth-first walk meets them (the rank order).
    static void engine_item_order(const RefTree &t, std::vector<int32_t> &out) {
        std::function<void(int, bool)> walk = [&](int node, bool wrap) {
            std::vector<int32_t> slots;
            if (wrap) slots.push_back(node);
            else {
                const RefTree::Node &nd = t.nodes[node];
                for (int i = 0; i < nd.n_children; ++i) slots.push_back(t.children[nd.first_child + i]);
            }
            for (int32_t ch : slots) {
                if (!is_nested(t, ch)) out.push_back(t.nodes[ch].item);
                else if (all_items(t, ch)) {
                    const RefTree::Node &nd = t.nodes[ch];
                    for (int i = 0; i < nd.n_children; ++i) out.push_back(t.nodes[t.children[nd.first_child + i]].item);
                } else walk(ch, false);
            }
        };
        if (t.root >= 0) walk(t.root, !is_nested(t, t.root) || all_items(t, t.root));
    }

    // A triangle list: ranks from the engine's tree, nodes from the traversal tree.
    int32_t emit_triangle_tree(int mesh, int32_t object, int32_t material) {
        const rayna_mesh &m = d.meshes[mesh];
        std::vector<int32_t> ranked;
        engine_item_order(blas[mesh], ranked);
        std::vector<int32_t> rank_of(m.count, -1);
        for (int32_t it : ranked)
            if (it >= 0 && it < m.count) rank_of[it] = 0;
        if ((int32_t)ranked.size() != m.count || std::count(rank_of.begin(), rank_of.end(), 0) != m.count)
            return -1;  // not one rank per triangle: the caller emits the engine's tree
        for (size_t k = 0; k < ranked.size(); ++k) {
            rank_of[ranked[k]] = rank + (int32_t)k;
            P.rank_object.push_back(object);
            P.rank_local.push_back(ranked[k]);
        }
        rank += (int32_t)ranked.size();
        std::vector<Aabb64> boxes(m.count);
        for (int32_t i = 0; i < m.count; ++i) boxes[i] = tri_box(m.index + i);
        const TraversalTree tt(boxes);
        auto leaf = [&](GNode &g, int k, const TraversalTree::Node &nd) {
            g.child[k] = (int32_t)P.tris.size();
            g.info[k] = nd.count | (LEAF_TRIS << 4);
            for (int32_t i = 0; i < nd.count; ++i) {
                const int32_t it = tt.order[nd.first + i];
                push_triangle(m.index + it, material, rank_of[it]);
            }
        };
        std::function<int32_t(int, bool)> emit = [&](int node, bool wrap) -> int32_t {
            const int32_t gi = (int32_t)P.nodes.size();
            P.nodes.push_back(empty_node());
            std::vector<int32_t> slots;
            if (wrap) slots.push_back(node);
            else {
                slots = {tt.nodes[node].left, tt.nodes[node].right};
                for (;;) {  // open the inner child of largest surface area while there is room
                    int open = -1;
                    double best = -1.0;
                    for (int k = 0; k < (int)slots.size(); ++k) {
                        const TraversalTree::Node &c = tt.nodes[slots[k]];
                        if (c.left < 0) continue;
                        const Aabb64 &b = c.box;
                        const double x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
                        const double a = x * y + y * z + z * x;
                        if (a > best) {
                            best = a;
                            open = k;
                        }
                    }
                    if (open < 0 || slots.size() >= 4) break;
                    const int32_t c = slots[open];
                    slots[open] = tt.nodes[c].left;
                    slots.insert(slots.begin() + open + 1, tt.nodes[c].right);
                }
            }
            for (int k = 0; k < (int)slots.size(); ++k) {
                const TraversalTree::Node &c = tt.nodes[slots[k]];
                set_slot_box(P.nodes[gi], k, c.box);
                if (c.left < 0) {
                    GNode g = P.nodes[gi];
                    leaf(g, k, c);
                    P.nodes[gi] = g;
                } else {
                    const int32_t sub = emit(slots[k], false);
                    P.nodes[gi].child[k] = sub;
                    P.nodes[gi].info[k] = INFO_INNER;
                }
            }
            return gi;
        };
        return emit(0, tt.nodes[0].left < 0);
    }

    int32_t emit_blas(int mesh, int32_t object, int32_t material) {
        if (d.meshes[mesh].kind == RAYNA_MESH_TRIANGLE_LIST && d.meshes[mesh].count > 0) {
            const int32_t root = emit_triangle_tree(mesh, object, material);
            if (root >= 0) return root;
        }
        const RefTree &t = blas[mesh];
        const std::vector<Aabb64> &tight = blas_tight[mesh];
        const rayna_mesh &m = d.meshes[mesh];
        std::function<int32_t(int, bool)> emit = [&](int node, bool wrap) -> int32_t {
            const int32_t gi = (int32_t)P.nodes.size();
            P.nodes.push_back(empty_node());
            std::vector<int32_t> slots;
            if (wrap) slots.push_back(node);
            else {
                const RefTree::Node &nd = t.nodes[node];
                for (int i = 0; i < nd.n_children; ++i) slots.push_back(t.children[nd.first_child + i]);
            }
            for (int k = 0; k < (int)slots.size(); ++k) {
                const int ch = slots[k];
                set_slot_box(P.nodes[gi], k, tight[ch]);
                if (!is_nested(t, ch)) {
                    GNode g = P.nodes[gi];
                    blas_leaf(g, k, m, material, object, {t.nodes[ch].item});
                    P.nodes[gi] = g;
                } else if (all_items(t, ch)) {
                    std::vector<int32_t> items;
                    const RefTree::Node &nd = t.nodes[ch];
                    for (int i = 0; i < nd.n_children; ++i) items.push_back(t.nodes[t.children[nd.first_child + i]].item);
                    GNode g = P.nodes[gi];
                    blas_leaf(g, k, m, material, object, items);
                    P.nodes[gi] = g;
                } else {
                    const int32_t sub = emit(ch, false);
                    P.nodes[gi].child[k] = sub;
                    P.nodes[gi].info[k] = INFO_INNER;
                }
            }
            return gi;
        };
        const bool wrap = !is_nested(t, t.root) || all_items(t, t.root);
        return emit(t.root, wrap);
    }

    // A transformed object's item: the primitive or volume item `it` wrapped with its maps.
    GItem wrap_transform(int32_t o, GItem it) {
        const int32_t x = obj_xf[o];
        if (x < 0) return it;
        GXform g;
        g.inner_kind = it.kind;
        g.inner_index = it.index;
        g.pad0 = g.pad1 = 0;
        for (int k = 0; k < 12; ++k) {
            g.fwd[k] = (float)xf_fwd[12 * (size_t)x + k];
            g.inv[k] = (float)xf_inv[12 * (size_t)x + k];
        }
        P.xforms.push_back(g);
        it.kind = ITEM_XFORM;
        it.index = (int32_t)P.xforms.size() - 1;
        return it;
    }

    // One object of the object BVH as a leaf item.
    GItem object_item(int32_t o, std::string &error, rayna_status &st) {
        const GItem it = object_item_local(o, error, st);
        return st == RAYNA_OK ? wrap_transform(o, it) : it;
    }

    GItem object_item_local(int32_t o, std::string &error, rayna_status &st) {
        const rayna_object &ob = d.objects[o];
        const rayna_mesh &m = d.meshes[ob.mesh];
        GItem it{ITEM_NONE, 0, 0, ob.material};
        if (ob.kind == RAYNA_OBJ_SIMPLE) {
            switch (m.kind) {
                case RAYNA_MESH_SPHERE:
                    it.kind = ITEM_SPHERE;
                    it.index = m.index;
                    break;
                case RAYNA_MESH_TRIANGLE:
                    it.kind = ITEM_TRIANGLE;
                    it.rank = rank;
                    it.index = add_triangle(m.index, ob.material, o, 0);
                    return it;
                case RAYNA_MESH_PARALLELOGRAM:
                    it.kind = ITEM_PARALLELOGRAM;
                    it.index = add_para(m.index);
                    break;
                case RAYNA_MESH_AXIS_BOX:
                    it.kind = ITEM_AXIS_BOX;
                    it.index = add_box(m.index);
                    break;
                case RAYNA_MESH_TRIANGLE_LIST:
                case RAYNA_MESH_BVH:
                    it.kind = ITEM_LINK;
                    it.rank = rank;
                    it.index = emit_blas(ob.mesh, o, ob.material);
                    return it;
                default:
                    st = RAYNA_UNSUPPORTED;
                    error = "unsupported bounded mesh kind";
                    return it;
            }
        } else {
            GVolume v;
            v.object = o;
            v.neg_inv_density = -1.0f / (float)ob.density;
            if (m.kind == RAYNA_MESH_SPHERE) {
                v.boundary_kind = ITEM_SPHERE;
                v.boundary_index = m.index;
            } else if (m.kind == RAYNA_MESH_AXIS_BOX) {
                v.boundary_kind = ITEM_AXIS_BOX;
                v.boundary_index = add_box(m.index);
            } else {
                st = RAYNA_UNSUPPORTED;
                error = "volumes are supported with sphere and axis-box boundaries";
                return it;
            }
            P.volumes.push_back(v);
            it.kind = ITEM_VOLUME;
            it.index = (int32_t)P.volumes.size() - 1;
        }
        it.rank = rank++;
        P.rank_object.push_back(o);
        P.rank_local.push_back(0);
        return it;
    }

    rayna_status emit_tlas(std::string &error) {
        rayna_status st = RAYNA_OK;
        const RefTree &t = tlas;
        std::function<int32_t(int, bool)> emit = [&](int node, bool wrap) -> int32_t {
            const int32_t gi = (int32_t)P.nodes.size();
            P.nodes.push_back(empty_node());
            std::vector<int32_t> slots;
            if (wrap) slots.push_back(node);
            else {
                const RefTree::Node &nd = t.nodes[node];
                for (int i = 0; i < nd.n_children; ++i) slots.push_back(t.children[nd.first_child + i]);
            }
            for (int k = 0; k < (int)slots.size() && st == RAYNA_OK; ++k) {
                const int ch = slots[k];
                set_slot_box(P.nodes[gi], k, tlas_tight[ch]);
                std::vector<int32_t> objs;
                if (!is_nested(t, ch)) objs.push_back(bounded[t.nodes[ch].item]);
                else if (all_items(t, ch)) {
                    const RefTree::Node &nd = t.nodes[ch];
                    for (int i = 0; i < nd.n_children; ++i)
                        objs.push_back(bounded[t.nodes[t.children[nd.first_child + i]].item]);
                }
                if (!objs.empty()) {
                    const int32_t start = (int32_t)P.items.size();
                    P.items.resize(P.items.size() + objs.size());
                    for (size_t i = 0; i < objs.size(); ++i) {
                        const GItem it = object_item(objs[i], error, st);
                        P.items[start + i] = it;
                    }
                    P.nodes[gi].child[k] = start;
                    P.nodes[gi].info[k] = (int32_t)objs.size() | (LEAF_ITEMS << 4);
                } else {
                    const int32_t sub = emit(ch, false);
                    P.nodes[gi].child[k] = sub;
                    P.nodes[gi].info[k] = INFO_INNER;
                }
            }
            return gi;
        };
        if (t.root >= 0) {
            const bool wrap = !is_nested(t, t.root) || all_items(t, t.root);
            P.root = emit(t.root, wrap);
        }
        return st;
    }

    // The deepest traversal stacks the flattened tree can need, for any ray and any culling.
    //
    // Depth-first (the packets' wave stack, and the bound `dfs` below): an inner node pushes its m non-empty
    // children (inner nodes and leaf ranges alike) and they are popped one by one, so the height above the point
    // where the node was popped peaks at max(m, m - 1 + the deepest child's need).  A leaf range pushes the root
    // of each mesh-BVH item it tests; those are processed after the range, in any order.  An instance (a
    // transformed mesh BVH) ends the range's pass: it pushes the rest of the range, three saved words of the
    // closest hit, a marker and its root, so its tree runs 4 entries above the rest of the range.
    //
    // Speculative (k_extend's lane stack): a lane that holds a leaf range keeps popping and expanding inner nodes
    // until it pops a second range (pushed back) or its stack empties, and only then tests the held range, whose
    // mesh-BVH links and instance entries go on top of whatever the speculation pushed meanwhile.  The stack
    // below them is the state of a depth-first traversal in which the held range pushes nothing, so it holds at
    // most `dfs` entries; the held range's own entries and their trees add at most its leaf need, and no second
    // range with such entries can be pending meanwhile (a lane holds one range; the ranges inside mesh BVHs and
    // instances are triangle or primitive ranges, which push nothing).  So k_extend needs at most `dfs` plus the
    // largest leaf need of a range holding a link or an instance.
    void stack_needs(int32_t *speculative, int32_t *depth_first) const {
        std::vector<int32_t> memo(P.nodes.size(), -1);
        std::function<int32_t(int32_t)> node_need;
        std::function<int32_t(int32_t, int32_t)> leaf_need = [&](int32_t start, int32_t count) -> int32_t {
            int32_t links = 0, deepest_link = 0, peak = 0;
            for (int32_t i = 0; i < count; ++i) {
                const GItem &it = P.items[start + i];
                if (it.kind == ITEM_LINK) {
                    ++links;
                    deepest_link = std::max(deepest_link, node_need(it.index));
                } else if (it.kind == ITEM_XFORM && P.xforms[it.index].inner_kind == ITEM_LINK) {
                    const int32_t rest = i + 1 < count ? 1 : 0;
                    peak = std::max(peak, links + rest + 5);
                    peak = std::max(peak, links + rest + 4 + node_need(P.xforms[it.index].inner_index));
                    if (rest) peak = std::max(peak, links + leaf_need(start + i + 1, count - i - 1));
                    break;
                }
            }
            if (links > 0) peak = std::max(peak, std::max(links, links - 1 + deepest_link));
            return peak;
        };
        node_need = [&](int32_t g) -> int32_t {
            if (memo[g] >= 0) return memo[g];
            int32_t m = 0, deepest = 0;
            const GNode &n = P.nodes[g];
            for (int k = 0; k < 4; ++k) {
                if (n.info[k] == INFO_EMPTY) continue;
                ++m;
                if (n.info[k] == INFO_INNER) deepest = std::max(deepest, node_need(n.child[k]));
                else if ((n.info[k] >> 4) == LEAF_ITEMS) deepest = std::max(deepest, leaf_need(n.child[k], n.info[k] & 0xf));
            }
            memo[g] = m == 0 ? 0 : std::max(m, m - 1 + deepest);
            return memo[g];
        };
        *speculative = *depth_first = 0;
        if (P.root < 0) return;
        const int32_t dfs = std::max<int32_t>(1, node_need(P.root));
        int32_t held = 0;  // the largest leaf need of a range with a link or an instance
        for (const GNode &n : P.nodes)
            for (int k = 0; k < 4; ++k) {
                if (n.info[k] == INFO_EMPTY || n.info[k] == INFO_INNER || (n.info[k] >> 4) != LEAF_ITEMS) continue;
                const int32_t start = n.child[k], count = n.info[k] & 0xf;
                bool pushes = false;
                for (int32_t i = 0; i < count; ++i) {
                    const GItem &it = P.items[start + i];
                    pushes = pushes || it.kind == ITEM_LINK ||
                             (it.kind == ITEM_XFORM && P.xforms[it.index].inner_kind == ITEM_LINK);
                }
                if (pushes) held = std::max(held, leaf_need(start, count));
            }
        *depth_first = dfs;
        *speculative = dfs + held;
    }
};

}  // namespace

RefTree build_engine_bvh(const std::vector<Aabb64> &boxes) { return Builder(boxes).build(); }

void hash_tree(const RefTree &t, uint64_t &h) {
    if (t.root < 0) return;
    auto mix = [&](int64_t v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (uint64_t)((v >> (8 * i)) & 0xff);
            h *= 1099511628211ull;
        }
    };
    std::vector<int32_t> stack{t.root};
    while (!stack.empty()) {
        const int32_t n = stack.back();
        stack.pop_back();
        const RefTree::Node &nd = t.nodes[n];
        mix(nd.item);
        mix(nd.n_children);
        for (int i = nd.n_children - 1; i >= 0; --i) stack.push_back(t.children[nd.first_child + i]);
    }
}

// The escape sphere of an SDF as the kernels evaluate it (float centres, radii and parameters): outside it the
// distance estimate exceeds epsilon, with margin for float rounding.
//   sphere:      |p - c| - r > eps when |p - c| > r + eps
//   smooth min:  each fold subtracts at most k/4 from min(d, d_i), so the estimate is at least
//                min_i (|p - c_i| - r_i) - (n - 1) k / 4, above eps when
//                |p - E| > max_i (|c_i - E| + r_i) + (n - 1) k / 4 + eps
//   Mandelbulb:  beyond the bail-out radius the iteration stops at once and the estimate is 0.5 ln(r) r, which
//                exceeds eps once r > bailout and 0.5 ln(r) r > 2 eps.
// Points along a ray that moves away from E only get farther from it, so a march past the sphere cannot hit.
static void sdf_escape_sphere(GSdf &g, const std::vector<GSphere> &spheres) {
    const double eps = g.epsilon;
    double E[3] = {g.cx, g.cy, g.cz}, R = INFINITY;
    if (!(eps > 0.0) || eps > 0.1) {
        R = INFINITY;
    } else if (g.kind == RAYNA_SDF_SPHERE) {
        const GSphere &sp = spheres[g.first_sphere];
        E[0] = sp.cx;
        E[1] = sp.cy;
        E[2] = sp.cz;
        R = (double)sp.radius + eps;
    } else if (g.kind == RAYNA_SDF_SMIN_SPHERES) {
        double m[3] = {0, 0, 0};
        for (int i = 0; i < g.n_spheres; ++i) {
            const GSphere &sp = spheres[g.first_sphere + i];
            m[0] += sp.cx;
            m[1] += sp.cy;
            m[2] += sp.cz;
        }
        for (int a = 0; a < 3; ++a) E[a] = (double)(float)(m[a] / g.n_spheres);
        double far = 0.0;
        for (int i = 0; i < g.n_spheres; ++i) {
            const GSphere &sp = spheres[g.first_sphere + i];
            const double dx = sp.cx - E[0], dy = sp.cy - E[1], dz = sp.cz - E[2];
            far = std::max(far, std::sqrt(dx * dx + dy * dy + dz * dz) + (double)sp.radius);
        }
        R = (g.k >= 0.0f) ? far + (g.n_spheres - 1) * (double)g.k / 4.0 + eps : INFINITY;
    } else {  // Mandelbulb: smallest r >= 1 with 0.5 ln(r) r > 2 eps, by bisection
        double lo = 1.0, hi = 2.0 + 4.0 * eps;
        for (int it = 0; it < 80; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (0.5 * std::log(mid) * mid > 2.0 * eps) hi = mid;
            else lo = mid;
        }
        R = std::max((double)g.bailout, hi);
    }
    const double Rm = R * (1.0 + 1e-4) + 1e-4;
    g.ex = (float)E[0];
    g.ey = (float)E[1];
    g.ez = (float)E[2];
    float Rf = (float)Rm;
    if ((double)Rf < Rm) Rf = std::nextafter(Rf, INFINITY);
    g.escape_r = std::isfinite(Rm) ? Rf : INFINITY;
}

void noise_permutation(uint32_t seed, int32_t out[512]) {
    uint64_t state = seed;
    auto next = [&]() {  // SplitMix64 (Steele, Lea and Flood 2014)
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    int32_t p[256];
    for (int i = 0; i < 256; ++i) p[i] = i;
    for (int i = 255; i > 0; --i) {
        const int j = (int)(next() % (uint64_t)(i + 1));
        std::swap(p[i], p[j]);
    }
    for (int i = 0; i < 512; ++i) out[i] = p[i & 255];
}

// Textures and images: validation and device records.  A texture's odd / even name earlier textures, so the
// texture graph has no cycles and a lookup ends within n_textures steps.
static rayna_status pack_textures(const rayna_scene_desc &d, PackedScene &P, std::string &error) {
    if (d.n_textures < 0 || d.n_images < 0 || (d.n_textures && !d.textures) || (d.n_images && !d.images)) {
        error = "bad texture or image array";
        return RAYNA_INVALID_ARGUMENT;
    }
    for (int i = 0; i < d.n_images; ++i) {
        const rayna_image &im = d.images[i];
        const uint64_t n = (uint64_t)im.width * im.height;
        if (im.width == 0 || im.height == 0 || !im.rgb || n > (1ull << 28) ||
            P.texels.size() + 3 * n > (size_t)INT32_MAX) {
            error = "bad image";
            return RAYNA_INVALID_ARGUMENT;
        }
        P.images.push_back(GImage{(int32_t)(P.texels.size() / 3), im.width, im.height, 0});
        P.texels.insert(P.texels.end(), im.rgb, im.rgb + 3 * n);
    }
    for (int i = 0; i < d.n_textures; ++i) {
        const rayna_texture &t = d.textures[i];
        GTexture g;
        std::memset(&g, 0, sizeof(g));
        g.kind = t.kind;
        g.odd = t.odd - 1;
        g.even = t.even - 1;
        g.image = t.image;
        g.r = t.colour[0];
        g.g = t.colour[1];
        g.b = t.colour[2];
        g.scale = (float)t.scale;
        g.ox = (float)t.offset[0];
        g.oy = (float)t.offset[1];
        g.oz = (float)t.offset[2];
        g.sx = (float)t.image_scale[0];
        g.sy = (float)t.image_scale[1];
        g.rgb = t.noise_rgb ? 1 : 0;
        switch (t.kind) {
            case RAYNA_TEX_SOLID: break;
            case RAYNA_TEX_WORLD_CHECKER:
            case RAYNA_TEX_UV_CHECKER:
                if (t.odd < 1 || t.odd > i || t.even < 1 || t.even > i) {
                    error = "a checker texture must name earlier textures";
                    return RAYNA_INVALID_ARGUMENT;
                }
                break;
            case RAYNA_TEX_IMAGE:
                if (t.image < 0 || t.image >= d.n_images) {
                    error = "texture image index out of range";
                    return RAYNA_INVALID_ARGUMENT;
                }
                break;
            case RAYNA_TEX_UV_NOISE:
            case RAYNA_TEX_WORLD_NOISE:
            case RAYNA_TEX_LOCAL_NOISE:
                for (int k = 0; k < (g.rgb ? 3 : 1); ++k) {
                    g.perm[k] = (int32_t)P.perms.size();
                    int32_t table[512];
                    noise_permutation(t.noise_seed[k], table);
                    P.perms.insert(P.perms.end(), table, table + 512);
                }
                break;
            default:
                error = "unknown texture kind";
                return RAYNA_INVALID_ARGUMENT;
        }
        P.textures.push_back(g);
    }
    if (d.skybox == RAYNA_SKY_HDRI && (d.sky_image < 0 || d.sky_image >= d.n_images)) {
        error = "sky image index out of range";
        return RAYNA_INVALID_ARGUMENT;
    }
    P.sky_image = d.skybox == RAYNA_SKY_HDRI ? d.sky_image : 0;
    return RAYNA_OK;
}

static bool finite_all(const double *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

rayna_status pack_scene(const rayna_scene_desc &d, PackedScene &P, std::string &error) {
    P = PackedScene();
    // ---- validation ----
    if (d.n_objects < 0 || d.n_meshes < 0 || d.n_materials < 0 || d.n_spheres < 0 || d.n_triangles < 0 ||
        d.n_parallelograms < 0 || d.n_boxes < 0 || d.n_sdfs < 0 || d.n_transforms < 0) {
        error = "negative array size in scene";
        return RAYNA_INVALID_ARGUMENT;
    }
    if ((d.n_objects && !d.objects) || (d.n_meshes && !d.meshes) || (d.n_materials && !d.materials) ||
        (d.n_spheres && !d.spheres) || (d.n_triangles && (!d.tri_vertices || !d.tri_normals)) ||
        (d.n_parallelograms && !d.parallelograms) || (d.n_boxes && !d.boxes) || (d.n_sdfs && !d.sdfs) ||
        (d.n_transforms && !d.transforms)) {
        error = "null array in scene";
        return RAYNA_INVALID_ARGUMENT;
    }
    if (d.skybox < RAYNA_SKY_NONE || d.skybox > RAYNA_SKY_HDRI) {
        error = "unknown skybox";
        return RAYNA_INVALID_ARGUMENT;
    }
    if (!finite_all(d.spheres, 4 * (size_t)d.n_spheres) || !finite_all(d.tri_vertices, 9 * (size_t)d.n_triangles) ||
        !finite_all(d.tri_normals, 9 * (size_t)d.n_triangles) ||
        !finite_all(d.parallelograms, 9 * (size_t)d.n_parallelograms) || !finite_all(d.boxes, 6 * (size_t)d.n_boxes) ||
        !finite_all(d.transforms, 12 * (size_t)d.n_transforms)) {
        error = "non-finite geometry in scene";
        return RAYNA_INVALID_ARGUMENT;
    }
    for (int i = 0; i < d.n_materials; ++i) {
        const rayna_material &m = d.materials[i];
        if (m.kind < RAYNA_MAT_LAMBERTIAN || m.kind > RAYNA_MAT_LIGHT) {
            error = "unknown material kind";
            return RAYNA_INVALID_ARGUMENT;
        }
        if (m.texture < 0 || m.texture > d.n_textures) {
            error = "material texture index out of range";
            return RAYNA_INVALID_ARGUMENT;
        }
        GMaterial g{m.kind, m.colour[0], m.colour[1], m.colour[2], (float)m.fuzz, (float)m.refractive_index,
                    (float)m.density, m.texture - 1};
        P.materials.push_back(g);
    }
    {
        const rayna_status ts = pack_textures(d, P, error);
        if (ts != RAYNA_OK) return ts;
    }
    for (int i = 0; i < d.n_meshes; ++i) {
        const rayna_mesh &m = d.meshes[i];
        int64_t limit = 0;
        switch (m.kind) {
            case RAYNA_MESH_SPHERE: limit = d.n_spheres; break;
            case RAYNA_MESH_TRIANGLE: limit = d.n_triangles; break;
            case RAYNA_MESH_TRIANGLE_LIST: limit = d.n_triangles; break;
            case RAYNA_MESH_PARALLELOGRAM: limit = d.n_parallelograms; break;
            case RAYNA_MESH_AXIS_BOX: limit = d.n_boxes; break;
            case RAYNA_MESH_SDF: limit = d.n_sdfs; break;
            case RAYNA_MESH_BVH: limit = d.n_meshes; break;
            default: error = "unknown mesh kind"; return RAYNA_INVALID_ARGUMENT;
        }
        const bool ranged = m.kind == RAYNA_MESH_TRIANGLE_LIST || m.kind == RAYNA_MESH_BVH;
        const int64_t count = ranged ? m.count : 1;
        if (m.index < 0 || count < 0 || (int64_t)m.index + count > limit || (!ranged && m.index >= limit)) {
            error = "mesh index out of range";
            return RAYNA_INVALID_ARGUMENT;
        }
    }
    for (int i = 0; i < d.n_meshes; ++i) {  // a BVH mesh holds primitive meshes only
        const rayna_mesh &m = d.meshes[i];
        if (m.kind != RAYNA_MESH_BVH) continue;
        for (int j = m.index; j < m.index + m.count; ++j) {
            const int k = d.meshes[j].kind;
            if (k != RAYNA_MESH_SPHERE && k != RAYNA_MESH_TRIANGLE && k != RAYNA_MESH_PARALLELOGRAM &&
                k != RAYNA_MESH_AXIS_BOX) {
                error = "a BVH mesh holds sphere, triangle, parallelogram and axis-box meshes only";
                return RAYNA_UNSUPPORTED;
            }
        }
    }
    for (int i = 0; i < d.n_objects; ++i) {
        const rayna_object &o = d.objects[i];
        if (o.kind != RAYNA_OBJ_SIMPLE && o.kind != RAYNA_OBJ_VOLUMETRIC) {
            error = "unknown object kind";
            return RAYNA_INVALID_ARGUMENT;
        }
        if (o.mesh < 0 || o.mesh >= d.n_meshes || o.material < 0 || o.material >= d.n_materials) {
            error = "object mesh or material index out of range";
            return RAYNA_INVALID_ARGUMENT;
        }
        if (o.transform < 0 || o.transform > d.n_transforms) {
            error = "object transform index out of range";
            return RAYNA_INVALID_ARGUMENT;
        }
    }
    for (int i = 0; i < d.n_sdfs; ++i) {
        const rayna_sdf &f = d.sdfs[i];
        if (f.kind < RAYNA_SDF_SPHERE || f.kind > RAYNA_SDF_MANDELBULB || f.max_iterations < 0) {
            error = "invalid SDF";
            return RAYNA_INVALID_ARGUMENT;
        }
        if (f.kind != RAYNA_SDF_MANDELBULB &&
            (f.first_sphere < 0 || f.n_spheres < 1 || (int64_t)f.first_sphere + f.n_spheres > d.n_spheres)) {
            error = "SDF sphere range out of range";
            return RAYNA_INVALID_ARGUMENT;
        }
        if (f.kind == RAYNA_SDF_MANDELBULB &&
            (f.power != std::floor(f.power) || f.power < 2 || f.power > 16 || f.bulb_iterations < 0)) {
            error = "Mandelbulb power must be an integer in [2, 16]";
            return RAYNA_UNSUPPORTED;
        }
    }
    for (int i = 0; i < d.n_spheres; ++i) {
        const double *s = &d.spheres[4 * (size_t)i];
        P.spheres.push_back(GSphere{(float)s[0], (float)s[1], (float)s[2], (float)s[3]});
    }
    P.skybox = d.skybox;

    Packer K(d, P);
    // ---- object transforms: identity matrices are the identity transform (transform.rs: is_identity) ----
    K.obj_xf.assign(d.n_objects, -1);
    for (int o = 0; o < d.n_objects; ++o) {
        const rayna_object &ob = d.objects[o];
        if (ob.transform == 0) continue;
        const double *m = &d.transforms[12 * (size_t)(ob.transform - 1)];
        if (affine_is_identity(m)) continue;
        if (d.meshes[ob.mesh].kind == RAYNA_MESH_SDF) {
            error = "object transforms are not supported on SDF meshes";
            return RAYNA_UNSUPPORTED;
        }
        double inv[12];
        if (!affine_inverse(m, inv)) {
            error = "object transform is not invertible";
            return RAYNA_INVALID_ARGUMENT;
        }
        K.obj_xf[o] = (int32_t)(K.xf_fwd.size() / 12);
        K.xf_fwd.insert(K.xf_fwd.end(), m, m + 12);
        K.xf_inv.insert(K.xf_inv.end(), inv, inv + 12);
    }
    // ---- object bounds (reference boxes carry the origin for lists: mesh/advanced/list.rs:44-49) ----
    std::vector<Aabb64> obj_ref(d.n_objects);
    K.obj_tight.resize(d.n_objects);
    std::vector<int32_t> unbounded;
    for (int o = 0; o < d.n_objects; ++o) {
        const rayna_mesh &m = d.meshes[d.objects[o].mesh];
        bool bnd = true;
        switch (m.kind) {
            case RAYNA_MESH_TRIANGLE_LIST:
            case RAYNA_MESH_BVH: {
                // MeshList (mesh/advanced/list.rs:44-49) folds its triangles' boxes from the origin; BvhMesh takes its root's
                // box (bvh.rs:116-129), the origin-folded union, or the one mesh's own box
                if (m.count <= 0) {
                    bnd = false;
                    break;
                }
                const bool list = m.kind == RAYNA_MESH_TRIANGLE_LIST;
                Aabb64 r = origin_box(), t = list ? K.tri_box(m.index) : K.prim_box(m.index);
                for (int i = 0; i < m.count; ++i) {
                    const Aabb64 tb = list ? K.tri_box(m.index + i) : K.prim_box(m.index + i);
                    grow(r, tb);
                    grow(t, tb);
                }
                obj_ref[o] = (!list && m.count == 1) ? t : r;
                K.obj_tight[o] = t;
                break;
            }
            case RAYNA_MESH_SDF: bnd = false; break;
            default: obj_ref[o] = K.obj_tight[o] = K.prim_box(d.objects[o].mesh); break;
        }
        if (bnd && K.obj_xf[o] >= 0) {  // ObjectTransform::calculate_aabb
            const double *m = &K.xf_fwd[12 * (size_t)K.obj_xf[o]];
            obj_ref[o] = affine_box(m, obj_ref[o]);
            K.obj_tight[o] = affine_box(m, K.obj_tight[o]);
        }
        if (bnd) K.bounded.push_back(o);
        else unbounded.push_back(o);
    }
    // ---- unbounded objects: SDFs (empty triangle lists never hit) ----
    for (size_t i = 0; i < unbounded.size(); ++i) {
        const rayna_object &ob = d.objects[unbounded[i]];
        const rayna_mesh &m = d.meshes[ob.mesh];
        if (m.kind == RAYNA_MESH_TRIANGLE_LIST || m.kind == RAYNA_MESH_BVH) continue;
        if (ob.kind != RAYNA_OBJ_SIMPLE) {
            error = "volumes with unbounded boundaries are not supported";
            return RAYNA_UNSUPPORTED;
        }
        const rayna_sdf &f = d.sdfs[m.index];
        GSdf g;
        std::memset(&g, 0, sizeof(g));
        g.kind = f.kind;
        g.first_sphere = f.first_sphere;
        g.n_spheres = f.n_spheres;
        g.max_iterations = f.max_iterations;
        g.bulb_iterations = f.bulb_iterations;
        g.power = (int32_t)f.power;
        g.material = ob.material;
        g.list_pos = (int32_t)i;
        g.epsilon = (float)f.epsilon;
        g.k = (float)f.k;
        g.bailout = (float)f.bailout;
        g.cx = (float)f.centre[0];
        g.cy = (float)f.centre[1];
        g.cz = (float)f.centre[2];
        sdf_escape_sphere(g, P.spheres);
        P.sdfs.push_back(g);
    }
    // ---- mesh BVHs, built in parallel ----
    K.blas.resize(d.n_meshes);
    K.blas_tight.resize(d.n_meshes);
    {
        std::vector<int> lists;
        for (int m = 0; m < d.n_meshes; ++m)
            if ((d.meshes[m].kind == RAYNA_MESH_TRIANGLE_LIST || d.meshes[m].kind == RAYNA_MESH_BVH) &&
                d.meshes[m].count > 0)
                lists.push_back(m);
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (;;) {
                const size_t j = next.fetch_add(1);
                if (j >= lists.size()) break;
                const rayna_mesh &me = d.meshes[lists[j]];
                std::vector<Aabb64> tb(me.count);
                for (int i = 0; i < me.count; ++i)
                    tb[i] = me.kind == RAYNA_MESH_BVH ? K.prim_box(me.index + i) : K.tri_box(me.index + i);
                K.blas[lists[j]] = build_engine_bvh(tb);
                K.blas_tight[lists[j]] = Packer::refit(K.blas[lists[j]], [&](int it) { return tb[it]; });
            }
        };
        unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
        nt = std::min<unsigned>(nt, (unsigned)std::max<size_t>(1, lists.size()));
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
        work();
        for (auto &t : pool) t.join();
    }
    // ---- object BVH ----
    {
        std::vector<Aabb64> ob(K.bounded.size());
        for (size_t i = 0; i < K.bounded.size(); ++i) ob[i] = obj_ref[K.bounded[i]];
        K.tlas = build_engine_bvh(ob);
        K.tlas_tight = Packer::refit(K.tlas, [&](int it) { return K.obj_tight[K.bounded[it]]; });
    }
    P.tree_hash = 1469598103934665603ull;
    hash_tree(K.tlas, P.tree_hash);
    for (int m = 0; m < d.n_meshes; ++m) hash_tree(K.blas[m], P.tree_hash);
    const rayna_status st = K.emit_tlas(error);
    if (st != RAYNA_OK) return st;
    K.stack_needs(&P.stack_need, &P.packet_need);
    return RAYNA_OK;
}

std::vector<GNode> device_nodes(const std::vector<GNode> &nodes) {
    std::vector<GNode> out = nodes;
    for (GNode &g : out)
        for (int k = 0; k < 4; ++k) g.info[k] = child_code(g.child[k], g.info[k]);
    return out;
}

bool merge_scenes(const std::vector<PackedScene> &in, PackedScene &out, std::vector<SceneSlot> &slots) {
    out = PackedScene{};
    slots.clear();
    for (const PackedScene &p : in) {
        const int32_t node_off = (int32_t)out.nodes.size(), tri_off = (int32_t)out.tris.size();
        const int32_t item_off = (int32_t)out.items.size(), sphere_off = (int32_t)out.spheres.size();
        const int32_t para_off = (int32_t)out.paras.size(), box_off = (int32_t)out.boxes.size();
        const int32_t vol_off = (int32_t)out.volumes.size(), sdf_off = (int32_t)out.sdfs.size();
        const int32_t mat_off = (int32_t)out.materials.size(), xf_off = (int32_t)out.xforms.size();
        const int32_t tex_off = (int32_t)out.textures.size(), img_off = (int32_t)out.images.size();
        const int32_t texel_off = (int32_t)(out.texels.size() / 3), perm_off = (int32_t)out.perms.size();
        for (GNode g : p.nodes) {
            for (int k = 0; k < 4; ++k) {
                if (g.info[k] == INFO_INNER) g.child[k] += node_off;
                else if (g.info[k] > 0) g.child[k] += (g.info[k] >> 4) == LEAF_TRIS ? tri_off : item_off;
            }
            out.nodes.push_back(g);
        }
        for (GTri t : p.tris) {
            t.material += mat_off;
            out.tris.push_back(t);
        }
        out.tri_normals.insert(out.tri_normals.end(), p.tri_normals.begin(), p.tri_normals.end());
        for (GItem it : p.items) {
            switch (it.kind) {
                case ITEM_SPHERE: it.index += sphere_off; break;
                case ITEM_TRIANGLE: it.index += tri_off; break;
                case ITEM_PARALLELOGRAM: it.index += para_off; break;
                case ITEM_AXIS_BOX: it.index += box_off; break;
                case ITEM_VOLUME: it.index += vol_off; break;
                case ITEM_LINK: it.index += node_off; break;
                case ITEM_XFORM: it.index += xf_off; break;
                default: break;
            }
            it.material += mat_off;
            out.items.push_back(it);
        }
        for (GXform x : p.xforms) {
            switch (x.inner_kind) {
                case ITEM_LINK: x.inner_index += node_off; break;
                case ITEM_SPHERE: x.inner_index += sphere_off; break;
                case ITEM_TRIANGLE: x.inner_index += tri_off; break;
                case ITEM_PARALLELOGRAM: x.inner_index += para_off; break;
                case ITEM_AXIS_BOX: x.inner_index += box_off; break;
                default: x.inner_index += vol_off; break;  // ITEM_VOLUME
            }
            out.xforms.push_back(x);
        }
        out.spheres.insert(out.spheres.end(), p.spheres.begin(), p.spheres.end());
        out.paras.insert(out.paras.end(), p.paras.begin(), p.paras.end());
        out.boxes.insert(out.boxes.end(), p.boxes.begin(), p.boxes.end());
        for (GVolume v : p.volumes) {
            v.boundary_index += v.boundary_kind == ITEM_SPHERE ? sphere_off : box_off;
            out.volumes.push_back(v);
        }
        for (GSdf f : p.sdfs) {
            f.first_sphere += sphere_off;
            f.material += mat_off;
            out.sdfs.push_back(f);
        }
        for (GMaterial m : p.materials) {
            if (m.texture >= 0) m.texture += tex_off;
            out.materials.push_back(m);
        }
        for (GTexture t : p.textures) {
            t.odd += tex_off;
            t.even += tex_off;
            t.image += img_off;
            for (int k = 0; k < 3; ++k) t.perm[k] += perm_off;
            out.textures.push_back(t);
        }
        for (GImage im : p.images) {
            im.texel += texel_off;
            out.images.push_back(im);
        }
        out.texels.insert(out.texels.end(), p.texels.begin(), p.texels.end());
        out.perms.insert(out.perms.end(), p.perms.begin(), p.perms.end());
        slots.push_back(SceneSlot{p.root >= 0 ? p.root + node_off : -1, sdf_off, (int32_t)p.sdfs.size(), p.skybox,
                                  p.stack_need, p.packet_need, p.sky_image + img_off});
        out.stack_need = std::max(out.stack_need, p.stack_need);
        out.packet_need = std::max(out.packet_need, p.packet_need);
    }
    // leaf starts have 27 bits in a traversal stack code; hit references keep 28 bits of index
    return out.tris.size() < (1u << 27) && out.items.size() < (1u << 27) && out.sdfs.size() < (1u << 28);
}

}  // namespace rayna
