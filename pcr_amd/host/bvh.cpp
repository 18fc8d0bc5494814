// bvh.cpp -- SAH BVH builder, node layout and triangle order bit-identical to
// the reference builder (code/Bvh/Bvh.cpp:122-325):
//   - 16 centroid bins on each axis, axes skipped when the centroid extent is
//     below 1e-6, cost 1 + 2 (SA_L N_L + SA_R N_R) / SA_parent, a candidate
//     replaces the best only when strictly cheaper;
//   - leaf when N <= 4, no split, cost >= 2 N, or an empty side;
//   - nodes emitted in preorder (left child = parent + 1), leaves (first, count);
//   - triangles partitioned in place with libstdc++'s bidirectional
//     std::partition (two-pointer swap order), which fixes the leaf order.
// Work per node is done on precomputed centroids and triangle bounds (the same
// float expressions the reference evaluates on every visit, so caching them
// changes no bit) and the partition permutes 4-byte indices.
#include <algorithm>
#include <cfloat>
#include <numeric>

#include "scene.h"

namespace pcr {

namespace {

struct Box {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box &o) {
        for (int a = 0; a < 3; a++) {
            if (o.mn[a] < mn[a]) mn[a] = o.mn[a];
            if (o.mx[a] > mx[a]) mx[a] = o.mx[a];
        }
    }
    void grow_point(const float p[3]) {
        for (int a = 0; a < 3; a++) {
            if (p[a] < mn[a]) mn[a] = p[a];
            if (p[a] > mx[a]) mx[a] = p[a];
        }
    }
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder {
    std::vector<std::array<float, 3>> cen;  // centroids
    std::vector<Box> tb;                    // triangle bounds
    std::vector<int32_t> idx;               // permutation being built
    std::vector<BvhNode> nodes;

    // Triangle bounds as the reference accumulates them: vertex by vertex,
    // v0 then v1 then v2, with strict < / > updates.  Merging these boxes with
    // min/max is exact and order independent, so a bin box built from cached
    // triangle boxes equals the reference's vertex-by-vertex box.
    void prepare(const std::vector<Triangle> &t) {
        size_t n = t.size();
        cen.resize(n);
        tb.resize(n);
        idx.resize(n);
        for (size_t i = 0; i < n; i++) {
            const Vec3 *v[3] = {&t[i].v0, &t[i].v1, &t[i].v2};
            for (int a = 0; a < 3; a++) {
                float c0 = a == 0 ? v[0]->x : a == 1 ? v[0]->y : v[0]->z;
                float c1 = a == 0 ? v[1]->x : a == 1 ? v[1]->y : v[1]->z;
                float c2 = a == 0 ? v[2]->x : a == 1 ? v[2]->y : v[2]->z;
                cen[i][a] = (c0 + c1 + c2) * (1.f / 3.f);
            }
            Box b;
            for (int k = 0; k < 3; k++) {
                float p[3] = {v[k]->x, v[k]->y, v[k]->z};
                b.grow_point(p);
            }
            tb[i] = b;
            idx[i] = (int32_t)i;
        }
    }

    int build(int b, int e) {
        int ni = (int)nodes.size();
        nodes.emplace_back();
        Box box;
        for (int i = b; i < e; i++) box.grow(tb[idx[i]]);
        for (int a = 0; a < 3; a++) {
            nodes[ni].bmin[a] = box.mn[a];
            nodes[ni].bmax[a] = box.mx[a];
        }
        int n = e - b;
        auto leaf = [&] {
            nodes[ni].left = b;
            nodes[ni].right = 0;
            nodes[ni].count = n;
            return ni;
        };
        if (n <= 4) return leaf();
        Box cb;
        for (int i = b; i < e; i++) cb.grow_point(cen[idx[i]].data());
        float parent = box.area();
        int bestAxis = -1, bestBin = -1;
        float bestCost = FLT_MAX;
        if (parent > 0.f) {
            for (int a = 0; a < 3; a++) {
                float lo = cb.mn[a], ext = cb.mx[a] - lo;
                if (ext < 1e-6f) continue;
                float scale = 16.f / ext;
                Box bins[16];
                int cnt[16] = {};
                for (int i = b; i < e; i++) {
                    int t = idx[i];
                    int k = std::clamp((int)((cen[t][a] - lo) * scale), 0, 15);
                    bins[k].grow(tb[t]);
                    cnt[k]++;
                }
                float lsa[15], rsa[15];
                int lc[15], rc[15];
                Box acc;
                int ac = 0;
                for (int k = 0; k < 15; k++) {
                    acc.grow(bins[k]);
                    ac += cnt[k];
                    lsa[k] = acc.area();
                    lc[k] = ac;
                }
                acc = Box{};
                ac = 0;
                for (int k = 15; k > 0; k--) {
                    acc.grow(bins[k]);
                    ac += cnt[k];
                    rsa[k - 1] = acc.area();
                    rc[k - 1] = ac;
                }
                for (int k = 0; k < 15; k++) {
                    if (lc[k] == 0 || rc[k] == 0) continue;
                    float cost = 1.f + 2.f * (lsa[k] * lc[k] + rsa[k] * rc[k]) / parent;
                    if (cost < bestCost) {
                        bestCost = cost;
                        bestAxis = a;
                        bestBin = k;
                    }
                }
            }
        }
        if (bestAxis < 0 || bestCost >= (float)n * 2.f) return leaf();
        float lo = cb.mn[bestAxis];
        float scale = 16.f / (cb.mx[bestAxis] - lo);
        const int ax = bestAxis, bin = bestBin;
        auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](int32_t t) {
            return std::clamp((int)((cen[t][ax] - lo) * scale), 0, 15) <= bin;
        });
        int mid = (int)(it - idx.begin());
        if (mid == b || mid == e) return leaf();
        int l = build(b, mid);
        int r = build(mid, e);
        nodes[ni].left = l;
        nodes[ni].right = r;
        nodes[ni].count = 0;
        return ni;
    }
};

}  // namespace

std::vector<BvhNode> build_bvh(std::vector<Triangle> &tris) {
    if (tris.empty()) return {};
    Builder B;
    B.prepare(tris);
    B.nodes.reserve(2 * tris.size());
    B.build(0, (int)tris.size());
    std::vector<Triangle> perm(tris.size());
    for (size_t i = 0; i < tris.size(); i++) perm[i] = tris[B.idx[i]];
    tris.swap(perm);
    return std::move(B.nodes);
}

int bvh_depth(const std::vector<BvhNode> &nodes) {
    if (nodes.empty()) return 0;
    int best = 0;
    std::vector<std::pair<int, int>> st{{0, 1}};
    while (!st.empty()) {
        auto [n, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        if (nodes[n].count == 0) {
            st.push_back({nodes[n].left, d + 1});
            st.push_back({nodes[n].right, d + 1});
        }
    }
    return best;
}

}  // namespace pcr
