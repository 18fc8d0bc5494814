// refdump.cpp -- writes known-answer vectors from the compiled reference.
//
// Built by oracle/ref.mk against the reference's own sources (never copied
// into this repository) and run only in the build container by
// tests/golden/make_golden.py.  Every subcommand reads/writes raw
// little-endian binary files that the Python script converts to .npz.
//
// Subcommands
//   numgen SEED N OUT          NumGen::Epsilon streams of thread 0 and thread 1
//   cie OUT                    yBarIntegral, CMFs and singleLambdaXYZ on a grid
//   fits IN OUT                fitSigmoidCoefficients for RGB triples
//   spd FILE OUT               SpdLoader::loadSpdFile resampled table
//   optics IN OUT              Optics::dielectricBounce / schlick / cauchy
//   bvh IN OUT                 Bvh::build over triangles (matIdx = input index)
//   mt19937 OUT                the 200-triangle set of tests/unit/bvh.cpp + its BVH
//   scene JSON RAYS OUT        loadSceneFromFile: fits, BVH, sceneIntersect on rays
//   icosphere SUB SEED R AMP CX CY CZ OUT   the project's procedural mesh (input geometry)
//   bvhrays TRIS RAYS OUT      Bvh::build + Bvh::intersect over raw triangles

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "Bvh/Bvh.h"
#include "Includes/CIE.h"
#include "Includes/NumGen.h"
#include "Includes/Optics.h"
#include "Includes/RGBToSpectrum.h"
#include "Includes/SpdLoader.h"
#include "Scenes/Scene.h"
#include "Scenes/SceneLoader.h"
#include "../../pcr_amd/host/procmesh.h"

namespace {

std::vector<char> read_all(const char *path) {
    std::vector<char> buf;
    FILE *f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(n);
    if (n && std::fread(buf.data(), 1, n, f) != (size_t)n) { std::perror(path); std::exit(2); }
    std::fclose(f);
    return buf;
}

struct Writer {
    FILE *f;
    explicit Writer(const char *path) : f(std::fopen(path, "wb")) {
        if (!f) { std::perror(path); std::exit(2); }
    }
    ~Writer() { std::fclose(f); }
    template <class T> void put(const T &v) { std::fwrite(&v, sizeof(T), 1, f); }
    void putv(const Vec3f &v) { put(v[0]); put(v[1]); put(v[2]); }
};

void node_out(Writer &w, const Bvh::Node &n) {
    for (int i = 0; i < 3; i++) w.put(n.boxMin[i]);
    for (int i = 0; i < 3; i++) w.put(n.boxMax[i]);
    w.put(n.leftOrFirst);
    w.put(n.rightChild);
    w.put(n.count);
}

int cmd_numgen(uint64_t seed, int n, const char *out) {
    NumGen::setSeed(seed);
    std::vector<float> a(n), b(n);
    for (int i = 0; i < n; i++) a[i] = NumGen::Epsilon();  // main thread: index 0
    std::thread t([&]() { for (int i = 0; i < n; i++) b[i] = NumGen::Epsilon(); });  // index 1
    t.join();
    Writer w(out);
    for (float v : a) w.put(v);
    for (float v : b) w.put(v);
    return 0;
}

int cmd_cie(const char *out) {
    Writer w(out);
    w.put(CIE::yBarIntegral());
    // grid: 400 .. 700 in 0.37 nm steps plus exact 5 nm samples
    std::vector<float> lam;
    for (int i = 0; i <= 810; i++) lam.push_back(400.f + 0.37f * (float)i);
    for (int i = 0; i < Spectrum::kSamples; i++) lam.push_back(Spectrum::lambdaAt(i));
    int n = (int)lam.size();
    w.put(n);
    for (float l : lam) {
        w.put(l);
        w.put(CIE::xBar(l));
        w.put(CIE::yBar(l));
        w.put(CIE::zBar(l));
        w.putv(CIE::singleLambdaXYZ(l, 1.f));
        w.putv(CIE::singleLambdaXYZ(l, 3.7f));
    }
    Spectrum one(1.f);
    w.putv(CIE::spectrumToXYZ(one));
    Vec3f probe[4] = {Vec3f(1, 1, 1), Vec3f(0.4f, 0.7f, 0.2f), Vec3f(0.63f, 0.06f, 0.05f), Vec3f(0.2f, 0.3f, 0.9f)};
    for (auto &p : probe) {
        w.putv(CIE::linearSRGBToXYZ(p));
        w.putv(CIE::xyzToLinearSRGB(p));
    }
    return 0;
}

int cmd_fits(const char *in, const char *out) {
    auto buf = read_all(in);
    int n = (int)(buf.size() / (3 * sizeof(float)));
    const float *rgb = (const float *)buf.data();
    Writer w(out);
    float lamGrid[7] = {400.f, 433.3f, 480.f, 550.f, 612.5f, 660.f, 700.f};
    for (int i = 0; i < n; i++) {
        auto f = RGBToSpectrum::fitSigmoidCoefficients(Vec3f(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]));
        w.put(f.c0);
        w.put(f.c1);
        w.put(f.c2);
        w.put(f.scale);
        for (float l : lamGrid) w.put(RGBToSpectrum::evalSigmoidFit(f, l));
    }
    return 0;
}

int cmd_spd(const char *file, const char *out) {
    Spectrum s;
    std::string err;
    if (!SpdLoader::loadSpdFile(file, s, &err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    Writer w(out);
    for (int i = 0; i < Spectrum::kSamples; i++) w.put(s[i]);
    Vec3f xyz = CIE::spectrumToXYZ(s);
    w.putv(xyz);
    // tabulated lookups at a few off-grid wavelengths
    float probe[6] = {399.f, 400.f, 427.5f, 551.3f, 699.99f, 700.f};
    for (float l : probe) w.put(s(l));
    return 0;
}

int cmd_optics(const char *in, const char *out) {
    // input record: dir3 n3 hit3 entering(f) ior u
    auto buf = read_all(in);
    int n = (int)(buf.size() / (12 * sizeof(float)));
    const float *r = (const float *)buf.data();
    Writer w(out);
    for (int i = 0; i < n; i++, r += 12) {
        Vec3f d(r[0], r[1], r[2]), N(r[3], r[4], r[5]), h(r[6], r[7], r[8]);
        auto b = Optics::dielectricBounce(d, N, h, r[9] != 0.f, r[10], r[11]);
        w.putv(b.dir);
        w.putv(b.origin);
        float cosI = -d.dot(N);
        w.put(Optics::schlickFresnel(std::fabs(cosI), 1.f, r[10]));
        w.put(Optics::cauchyIor(r[10], 0.42f, 400.f + 300.f * r[11]));
    }
    return 0;
}

void dump_bvh(Writer &w, const std::vector<Bvh::Node> &nodes, const std::vector<Triangle> &tris) {
    int nn = (int)nodes.size(), nt = (int)tris.size();
    w.put(nn);
    w.put(nt);
    for (const auto &n : nodes) node_out(w, n);
    for (const auto &t : tris) w.put(t.matIdx);
    auto st = Bvh::statsOf(nodes);
    w.put(st.depth);
    w.put(st.leafCount);
    w.put(st.maxLeafTris);
}

int cmd_bvh(const char *in, const char *out) {
    auto buf = read_all(in);
    int n = (int)(buf.size() / (9 * sizeof(float)));
    const float *v = (const float *)buf.data();
    std::vector<Triangle> tris;
    tris.reserve(n);
    for (int i = 0; i < n; i++, v += 9)
        tris.emplace_back(Vec3f(v[0], v[1], v[2]), Vec3f(v[3], v[4], v[5]), Vec3f(v[6], v[7], v[8]), i);
    auto nodes = Bvh::build(tris);
    Writer w(out);
    dump_bvh(w, nodes, tris);
    return 0;
}

int cmd_mt19937(const char *out) {
    // The triangle set of tests/unit/bvh.cpp randomSceneTriangles(200).
    std::mt19937 rng(0xC0FFEEu);
    std::uniform_real_distribution<float> d(-1.f, 1.f);
    std::vector<Triangle> tris;
    for (int i = 0; i < 200; i++) {
        Vec3f a(d(rng), d(rng), d(rng) - 5.f);
        Vec3f b(a[0] + 0.05f * d(rng), a[1] + 0.05f * d(rng), a[2] + 0.05f * d(rng));
        Vec3f c(a[0] + 0.05f * d(rng), a[1] + 0.05f * d(rng), a[2] + 0.05f * d(rng));
        tris.emplace_back(a, b, c, i);
    }
    Writer w(out);
    for (const auto &t : tris) { w.putv(t.v0); w.putv(t.v1); w.putv(t.v2); }
    auto copy = tris;
    auto nodes = Bvh::build(tris);
    dump_bvh(w, nodes, tris);
    // the 30 rays of test_bvh_matches_brute_force and the BVH answers
    std::mt19937 r2(0xBADD);
    std::uniform_real_distribution<float> d2(-1.f, 1.f);
    for (int i = 0; i < 30; i++) {
        Vec3f origin(d2(r2), d2(r2), 0.f);
        Vec3f dir(d2(r2) * 0.2f, d2(r2) * 0.2f, -1.f);
        dir = dir.normalize();
        Ray ray(dir, origin);
        Vec3f hit, N;
        int m = -1;
        float t = 0.f;
        bool got = Bvh::intersect(nodes, tris, ray, hit, N, m, t, 1e30f);
        w.putv(origin);
        w.putv(dir);
        w.put(got ? m : -1);
        w.put(got ? t : INFINITY);
    }
    return 0;
}

int cmd_scene(const char *json, const char *raysPath, const char *out) {
    Scenes::SceneData s = Scenes::loadSceneFromFile(json);
    Writer w(out);
    int nm = (int)s.materials.size();
    w.put(nm);
    for (const auto &m : s.materials) {
        w.put(m.albedoFit.c0); w.put(m.albedoFit.c1); w.put(m.albedoFit.c2); w.put(m.albedoFit.scale);
        w.put(m.emissiveFit.c0); w.put(m.emissiveFit.c1); w.put(m.emissiveFit.c2); w.put(m.emissiveFit.scale);
        w.putv(m.albedo);
        w.putv(m.emissive);
        for (float l : {400.f, 455.5f, 550.f, 631.25f, 700.f}) {
            w.put(m.albedoAt(l));
            w.put(m.emissiveAt(l));
        }
    }
    // triangles in leaf order, identified by their vertices (caller maps to JSON order)
    int nt = (int)s.triangles.size();
    w.put(nt);
    for (const auto &t : s.triangles) { w.putv(t.v0); w.putv(t.v1); w.putv(t.v2); }
    int nn = (int)s.triangleBvh.size();
    w.put(nn);
    for (const auto &n : s.triangleBvh) node_out(w, n);
    int nl = (int)s.areaLights.size();
    w.put(nl);
    float total = 0.f;
    for (const auto &L : s.areaLights) {
        w.put(L.totalArea);
        total += L.totalArea;
    }
    w.put(total);
    // Renderer::render plane order (Renderer.cpp:123-130) then the
    // sceneIntersect sequence (Renderer.cpp:891-949) through public APIs.
    std::vector<Plane> planes;
    for (const auto &L : s.areaLights)
        if (L.kind == Scenes::AreaLightKind::Plane) planes.push_back(L.plane);
    for (const auto &p : s.walls) planes.push_back(p);
    auto rbuf = read_all(raysPath);
    int nr = (int)(rbuf.size() / (6 * sizeof(float)));
    const float *rr = (const float *)rbuf.data();
    w.put(nr);
    for (int i = 0; i < nr; i++, rr += 6) {
        Vec3f d(rr[3], rr[4], rr[5]), o(rr[0], rr[1], rr[2]);
        // const, as in Renderer::sceneIntersect: a non-const Ray would make
        // `ray.origin + ...` resolve to the mutating Vec3f::operator+.
        const Ray ray(d, o);
        float closest = std::numeric_limits<float>::max();
        float t0 = 0.f;
        Vec3f hit, N;
        int mat = -1, kind = -1, prim = -1;
        for (size_t k = 0; k < s.spheres.size(); k++) {
            const auto &sp = s.spheres[k];
            if (!sp.intersect(ray, t0) || t0 >= closest) continue;
            closest = t0;
            hit = ray.origin + ray.dir * t0;
            N = (hit - sp.center).normalize();
            mat = sp.matIdx;
            kind = 0;
            prim = (int)k;
        }
        for (size_t k = 0; k < planes.size(); k++) {
            if (!planes[k].intersect(ray, hit, t0, closest) || t0 >= closest) continue;
            closest = t0;
            N = planes[k].N;
            mat = planes[k].matIdx;
            kind = 1;
            prim = (int)k;
        }
        if (!s.triangleBvh.empty()) {
            Vec3f th, tn;
            int tm;
            float tt;
            if (Bvh::intersect(s.triangleBvh, s.triangles, ray, th, tn, tm, tt, closest)) {
                closest = tt;
                hit = th;
                N = tn;
                mat = tm;
                kind = 2;
                // recover the leaf-order index of the winning triangle: the
                // closest triangle with this t and material
                prim = -1;
                for (int q = 0; q < nt; q++) {
                    Vec3f qh, qn;
                    float qt;
                    if (s.triangles[q].intersect(ray, qh, qn, qt, FLT_MAX) && qt == tt &&
                        s.triangles[q].matIdx == tm && qh[0] == th[0] && qh[1] == th[1] && qh[2] == th[2]) {
                        prim = q;
                        break;
                    }
                }
            }
        }
        bool ok = closest < std::numeric_limits<float>::max();
        w.put(ok ? closest : INFINITY);
        w.put(ok ? kind : -1);
        w.put(ok ? prim : -1);
        w.put(ok ? mat : -1);
        w.putv(ok ? hit : Vec3f());
        w.putv(ok ? N : Vec3f());
    }
    return 0;
}

int cmd_icosphere(char **a, const char *out) {
    pcr::IcoParams p;
    p.subdiv = std::atoi(a[0]);
    p.seed = (uint32_t)std::stoul(a[1]);
    p.radius = std::atof(a[2]);
    p.amplitude = std::atof(a[3]);
    p.center[0] = std::atof(a[4]);
    p.center[1] = std::atof(a[5]);
    p.center[2] = std::atof(a[6]);
    auto v = pcr::make_displaced_icosphere(p);
    Writer w(out);
    for (float x : v) w.put(x);
    return 0;
}

int cmd_bvhrays(const char *trisPath, const char *raysPath, const char *out) {
    auto buf = read_all(trisPath);
    int n = (int)(buf.size() / (9 * sizeof(float)));
    const float *v = (const float *)buf.data();
    std::vector<Triangle> tris;
    tris.reserve(n);
    for (int i = 0; i < n; i++, v += 9)
        tris.emplace_back(Vec3f(v[0], v[1], v[2]), Vec3f(v[3], v[4], v[5]), Vec3f(v[6], v[7], v[8]), i);
    auto nodes = Bvh::build(tris);
    Writer w(out);
    dump_bvh(w, nodes, tris);
    auto rbuf = read_all(raysPath);
    int nr = (int)(rbuf.size() / (6 * sizeof(float)));
    const float *rr = (const float *)rbuf.data();
    w.put(nr);
    for (int i = 0; i < nr; i++, rr += 6) {
        Vec3f d(rr[3], rr[4], rr[5]), o(rr[0], rr[1], rr[2]);
        Ray ray(d, o);
        Vec3f hit, N;
        int m = -1;
        float t = 0.f;
        bool got = Bvh::intersect(nodes, tris, ray, hit, N, m, t, std::numeric_limits<float>::max());
        w.put(got ? t : INFINITY);
        w.put(got ? m : -1);
        w.putv(got ? hit : Vec3f());
        w.putv(got ? N : Vec3f());
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    std::string c = argv[1];
    try {
        if (c == "numgen" && argc == 5) return cmd_numgen(std::stoull(argv[2]), std::atoi(argv[3]), argv[4]);
        if (c == "cie" && argc == 3) return cmd_cie(argv[2]);
        if (c == "fits" && argc == 4) return cmd_fits(argv[2], argv[3]);
        if (c == "spd" && argc == 4) return cmd_spd(argv[2], argv[3]);
        if (c == "optics" && argc == 4) return cmd_optics(argv[2], argv[3]);
        if (c == "bvh" && argc == 4) return cmd_bvh(argv[2], argv[3]);
        if (c == "mt19937" && argc == 3) return cmd_mt19937(argv[2]);
        if (c == "scene" && argc == 5) return cmd_scene(argv[2], argv[3], argv[4]);
        if (c == "icosphere" && argc == 10) return cmd_icosphere(argv + 2, argv[9]);
        if (c == "bvhrays" && argc == 5) return cmd_bvhrays(argv[2], argv[3], argv[4]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "refdump %s: %s\n", c.c_str(), e.what());
        return 1;
    }
    std::fprintf(stderr, "usage: refdump <numgen|cie|fits|spd|optics|bvh|mt19937|scene|icosphere|bvhrays> ...\n");
    return 2;
}
