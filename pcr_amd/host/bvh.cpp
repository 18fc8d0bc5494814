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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <numeric>
#include <unistd.h>

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

namespace {
// nodes and the leaf order (input index of each position) of a build
std::vector<BvhNode> build_perm(const std::vector<Triangle> &tris, std::vector<int32_t> &perm) {
    Builder B;
    B.prepare(tris);
    B.nodes.reserve(2 * tris.size());
    B.build(0, (int)tris.size());
    perm = std::move(B.idx);
    return std::move(B.nodes);
}
void apply_perm(std::vector<Triangle> &tris, const std::vector<int32_t> &perm) {
    std::vector<Triangle> out(tris.size());
    for (size_t i = 0; i < tris.size(); i++) out[i] = tris[perm[i]];
    tris.swap(out);
}
}  // namespace

std::vector<BvhNode> build_bvh(std::vector<Triangle> &tris) {
    if (tris.empty()) return {};
    std::vector<int32_t> perm;
    auto nodes = build_perm(tris, perm);
    apply_perm(tris, perm);
    return nodes;
}

// ------------------------------------------------------- content cache ----
// The build depends only on the triangles' vertex positions, in input order.
// Entries are keyed by two independent 64-bit hashes of those float bits and
// the count (FNV-1a, and a splitmix64 chain), and a builder version.  Up to
// kMemEntries builds stay in memory; with a cache directory set, builds are
// also stored as files <dir>/<key>.bvh and read back after checks of size,
// keys, node ranges and that the leaf order is a permutation.
namespace {
constexpr uint32_t kBvhCacheVersion = 1;  // bump when the builder's output changes
constexpr size_t kMemEntries = 8;
struct Key {
    uint64_t a, b;
    bool operator==(const Key &o) const { return a == o.a && b == o.b; }
};
struct Entry {
    Key key;
    uint32_t n;
    std::vector<BvhNode> nodes;
    std::vector<int32_t> perm;
};
struct Cache {
    std::mutex m;
    std::deque<Entry> mem;
    std::string dir;
    bool dir_init = false;
    BvhCacheStats stats{};
};
Cache &cache() {
    static Cache c;
    return c;
}
uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
Key key_of(const std::vector<Triangle> &tris) {
    uint64_t a = 0xCBF29CE484222325ull, b = mix64(tris.size() ^ ((uint64_t)kBvhCacheVersion << 40));
    auto word = [&](uint32_t w) {
        for (int k = 0; k < 4; k++) {
            a ^= (w >> (8 * k)) & 0xFFu;
            a *= 0x100000001B3ull;
        }
        b = mix64(b ^ w);
    };
    word((uint32_t)tris.size());
    word(kBvhCacheVersion);
    for (const auto &t : tris)
        for (const Vec3 *v : {&t.v0, &t.v1, &t.v2}) {
            uint32_t w[3];
            std::memcpy(w, &v->x, 4);
            std::memcpy(w + 1, &v->y, 4);
            std::memcpy(w + 2, &v->z, 4);
            for (uint32_t x : w) word(x);
        }
    return Key{a, b};
}
bool valid(const Entry &e) {
    const int nn = (int)e.nodes.size(), n = (int)e.n;
    if (e.perm.size() != e.n || nn < 1 || nn > 2 * n) return false;
    std::vector<char> seen(e.n, 0);
    for (int32_t p : e.perm) {
        if (p < 0 || p >= n || seen[p]) return false;
        seen[p] = 1;
    }
    for (const auto &d : e.nodes) {
        if (d.count > 0) {
            if (d.left < 0 || d.left + d.count > n) return false;
        } else if (d.left < 1 || d.left >= nn || d.right < 1 || d.right >= nn) {
            return false;
        }
    }
    return true;
}
std::string file_of(const std::string &dir, const Key &k) {
    char name[64];
    std::snprintf(name, sizeof(name), "%016llx%016llx.bvh", (unsigned long long)k.a, (unsigned long long)k.b);
    return dir + "/" + name;
}
constexpr char kMagic[8] = {'P', 'C', 'R', 'B', 'V', 'H', '0', '1'};
bool read_file(const std::string &path, const Key &k, uint32_t n, Entry &e) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    char magic[8];
    uint32_t ver = 0, nt = 0, nn = 0;
    Key fk{};
    f.read(magic, 8);
    f.read((char *)&ver, 4);
    f.read((char *)&nt, 4);
    f.read((char *)&nn, 4);
    f.read((char *)&fk.a, 8);
    f.read((char *)&fk.b, 8);
    if (!f || std::memcmp(magic, kMagic, 8) != 0 || ver != kBvhCacheVersion || nt != n || !(fk == k) ||
        nn < 1 || nn > 2 * n)
        return false;
    e.key = k;
    e.n = n;
    e.nodes.resize(nn);
    e.perm.resize(n);
    f.read((char *)e.nodes.data(), (std::streamsize)(nn * sizeof(BvhNode)));
    f.read((char *)e.perm.data(), (std::streamsize)(n * sizeof(int32_t)));
    if (!f || f.peek() != std::char_traits<char>::eof()) return false;
    return valid(e);
}
void write_file(const std::string &path, const Entry &e) {
    const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return;
        const uint32_t ver = kBvhCacheVersion, nn = (uint32_t)e.nodes.size();
        f.write(kMagic, 8);
        f.write((const char *)&ver, 4);
        f.write((const char *)&e.n, 4);
        f.write((const char *)&nn, 4);
        f.write((const char *)&e.key.a, 8);
        f.write((const char *)&e.key.b, 8);
        f.write((const char *)e.nodes.data(), (std::streamsize)(nn * sizeof(BvhNode)));
        f.write((const char *)e.perm.data(), (std::streamsize)(e.n * sizeof(int32_t)));
        if (!f) {
            f.close();
            std::remove(tmp.c_str());
            return;
        }
    }
    std::error_code ec;
    std::filesystem::rename(tmp, path, ec);  // atomic: readers see the old file or the whole new one
    if (ec) std::remove(tmp.c_str());
}
}  // namespace

void bvh_cache_set_dir(const char *dir) {
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.m);
    c.dir = dir ? dir : "";
    c.dir_init = true;
}
void bvh_cache_clear_memory() {
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.m);
    c.mem.clear();
}
BvhCacheStats bvh_cache_stats() {
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.m);
    return c.stats;
}

std::vector<BvhNode> build_bvh_cached(std::vector<Triangle> &tris) {
    if (tris.empty()) return {};
    Cache &c = cache();
    const Key k = key_of(tris);
    const uint32_t n = (uint32_t)tris.size();
    std::string dir;
    {
        std::lock_guard<std::mutex> lk(c.m);
        if (!c.dir_init) {
            const char *env = std::getenv("PCR_BVH_CACHE");
            c.dir = env ? env : "";
            c.dir_init = true;
        }
        dir = c.dir;
        for (const auto &e : c.mem)
            if (e.key == k && e.n == n) {
                c.stats.memory_hits++;
                apply_perm(tris, e.perm);
                return e.nodes;
            }
    }
    Entry e;
    bool from_disk = false;
    if (!dir.empty()) from_disk = read_file(file_of(dir, k), k, n, e);
    if (!from_disk) {
        e.key = k;
        e.n = n;
        e.nodes = build_perm(tris, e.perm);
        if (!dir.empty()) {
            std::error_code ec;
            std::filesystem::create_directories(dir, ec);
            write_file(file_of(dir, k), e);
        }
    }
    apply_perm(tris, e.perm);
    std::vector<BvhNode> nodes = e.nodes;
    std::lock_guard<std::mutex> lk(c.m);
    (from_disk ? c.stats.disk_hits : c.stats.builds)++;
    c.mem.push_back(std::move(e));
    if (c.mem.size() > kMemEntries) c.mem.pop_front();
    return nodes;
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
