// loader.cpp -- scene files in the reference's JSON schema 1.0.
//
// Ingestion rules follow code/Scenes/SceneLoader.cpp:
//   - "schema" "major.minor" with major 1 (:67-85); "name", "version" required;
//   - "camera" {position, fov} (:217-234);
//   - "materials": object iterated in sorted-key order, so material indices
//     are alphabetical (:89-215); albedo / emissive RGB, albedo_spd /
//     emissive_spd measured spectra (+ emissive_scale), RGB derived from an
//     SPD when the RGB value is absent; metallic, transparent, ior, cauchyB;
//   - "primitives": sphere, triangle (optional n0 n1 n2), plane; "light": true
//     on a plane puts it only in the light list, on a triangle it becomes a
//     one-triangle light and stays geometry (:394-472);
//   - after parsing: BVH over the triangles (permuting them), then the
//     spectral fits of every material (:504-513).
// Extension of this project: an "icosphere" primitive (procedural displaced
// icosphere, pcr_amd/host/procmesh.h) for the traversal-stress scene.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "json.h"
#include "procmesh.h"
#include "scene.h"

namespace pcr {

namespace fs = std::filesystem;
using json::Value;

namespace {

Vec3 vec3_of(const Value &v, const std::string &field) {
    if (!v.is_array() || v.arr.size() != 3) throw SceneError(field + ": expected [x, y, z] array of 3 numbers");
    for (const auto &e : v.arr)
        if (!e.is_number()) throw SceneError(field + ": expected numbers");
    return Vec3{v.arr[0].as_float(), v.arr[1].as_float(), v.arr[2].as_float()};
}

std::string req_string(const Value &p, const char *key, const std::string &field, const std::string &path) {
    const Value *v = p.find(key);
    if (!v || !v->is_string()) throw SceneError(path + ": " + field + " missing string \"" + key + "\"");
    return v->str;
}

std::string resolve_spd(const std::string &name, const std::vector<std::string> &dirs, const std::string &scenePath) {
    std::vector<fs::path> roots;
    for (const auto &d : dirs) roots.emplace_back(d);
    // next to the scene, beside its Scenes/ folder (the reference repository
    // layout), then $PWD/spd as the reference does (SpdLoader.cpp:93-98)
    roots.push_back(fs::path(scenePath).parent_path() / "spd");
    roots.push_back(fs::path(scenePath).parent_path().parent_path() / "spd");
    roots.push_back(fs::current_path() / "spd");
    std::error_code ec;
    for (const auto &r : roots) {
        fs::path c = r / (name + ".spd");
        if (fs::exists(c, ec)) return c.string();
    }
    if (fs::exists(name, ec)) return name;
    std::string msg = "SPD '" + name + "' not found in any of: ";
    for (const auto &r : roots) msg += r.string() + " ";
    throw SceneError(msg);
}

}  // namespace

Scene load_scene(const std::string &path, const std::vector<std::string> &spdDirs) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw SceneError("Cannot open scene file: " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    Value j;
    try {
        j = json::parse(ss.str());
    } catch (const json::ParseError &e) {
        throw SceneError(path + ":" + std::to_string(e.line) + ":" + std::to_string(e.col) +
                         ": JSON parse error: " + e.what());
    }
    if (!j.is_object()) throw SceneError(path + ": top level must be an object");
    const Value *sv = j.find("schema");
    if (!sv) throw SceneError(path + ": missing required \"schema\" field");
    if (!sv->is_string()) throw SceneError(path + ": \"schema\" must be a string like \"1.0\"");
    int major = 0, minor = 0;
    if (std::sscanf(sv->str.c_str(), "%d.%d", &major, &minor) != 2)
        throw SceneError(path + ": malformed schema \"" + sv->str + "\" (expected major.minor)");
    if (major != 1) throw SceneError(path + ": schema major version " + std::to_string(major) + " is not supported");

    Scene s;
    s.path = path;
    s.name = req_string(j, "name", "(root)", path);
    s.version = req_string(j, "version", "(root)", path);
    if (const Value *c = j.find("camera")) {
        if (!c->is_object()) throw SceneError(path + ": \"camera\" must be an object");
        if (const Value *p = c->find("position")) s.cam = vec3_of(*p, "camera.position");
        if (const Value *f = c->find("fov")) {
            if (!f->is_number()) throw SceneError(path + ": camera.fov must be a number");
            s.fov = f->as_float();
        }
    }

    const Value *mats = j.find("materials");
    if (!mats || !mats->is_object()) throw SceneError(path + ": missing or non-object \"materials\" map");
    std::unordered_map<std::string, int> idx;
    for (const auto &[name, m] : mats->obj) {  // std::map: sorted keys
        if (!m.is_object()) throw SceneError(path + ": material \"" + name + "\" must be an object");
        Material mat;
        const std::string f = "materials." + name;
        if (const Value *a = m.find("albedo")) mat.albedo = vec3_of(*a, f + ".albedo");
        if (const Value *e = m.find("emissive")) mat.emissive = vec3_of(*e, f + ".emissive");
        if (const Value *sp = m.find("albedo_spd")) {
            if (!sp->is_string()) throw SceneError(path + ": " + f + ".albedo_spd must be a string");
            std::string err;
            if (!load_spd_file(resolve_spd(sp->str, spdDirs, path), mat.tabAlbedo, &err))
                throw SceneError(path + ": " + f + ".albedo_spd: " + err);
            mat.useTabAlbedo = true;
            if (!m.contains("albedo")) {
                Vec3 rgb = xyz_to_linear_srgb(spectrum_to_xyz(mat.tabAlbedo));
                mat.albedo = Vec3{std::max(0.f, rgb.x), std::max(0.f, rgb.y), std::max(0.f, rgb.z)};
            }
        }
        if (const Value *sp = m.find("emissive_spd")) {
            if (!sp->is_string()) throw SceneError(path + ": " + f + ".emissive_spd must be a string");
            std::string err;
            if (!load_spd_file(resolve_spd(sp->str, spdDirs, path), mat.tabEmissive, &err))
                throw SceneError(path + ": " + f + ".emissive_spd: " + err);
            mat.useTabEmissive = true;
            if (const Value *k = m.find("emissive_scale")) {
                if (!k->is_number()) throw SceneError(path + ": " + f + ".emissive_scale must be a number");
                float sc = k->as_float();
                for (auto &x : mat.tabEmissive) x *= sc;
            }
            if (!m.contains("emissive")) {
                Vec3 rgb = xyz_to_linear_srgb(spectrum_to_xyz(mat.tabEmissive));
                mat.emissive = Vec3{std::max(0.f, rgb.x), std::max(0.f, rgb.y), std::max(0.f, rgb.z)};
            }
        }
        auto flag = [&](const char *key, bool &dst) {
            if (const Value *v = m.find(key)) {
                if (!v->is_bool()) throw SceneError(path + ": " + f + "." + key + " must be a boolean");
                dst = v->b;
            }
        };
        auto number = [&](const char *key, float &dst) {
            if (const Value *v = m.find(key)) {
                if (!v->is_number()) throw SceneError(path + ": " + f + "." + key + " must be a number");
                dst = v->as_float();
            }
        };
        flag("metallic", mat.metallic);
        flag("transparent", mat.transparent);
        number("ior", mat.ior);
        number("cauchyB", mat.cauchyB);
        if (mat.metallic && mat.transparent)
            throw SceneError(path + ": " + f + " cannot be both metallic and transparent");
        idx[name] = (int)s.materials.size();
        s.materials.push_back(mat);
        s.materialNames.push_back(name);
    }
    if (s.materials.empty()) throw SceneError(path + ": \"materials\" map is empty");
    auto mat_of = [&](const Value &p, const std::string &field) {
        std::string ref = req_string(p, "material", field, path);
        auto it = idx.find(ref);
        if (it == idx.end()) throw SceneError(path + ": " + field + ".material references unknown material \"" + ref + "\"");
        return it->second;
    };

    const Value *prims = j.find("primitives");
    if (!prims || !prims->is_array()) throw SceneError(path + ": missing or non-array \"primitives\"");
    auto add_light_tris = [&](std::vector<Triangle> tris) {
        Light L;
        L.kind = 1;
        float acc = 0.f;
        for (const auto &t : tris) {
            Vec3 e1{t.v1.x - t.v0.x, t.v1.y - t.v0.y, t.v1.z - t.v0.z};
            Vec3 e2{t.v2.x - t.v0.x, t.v2.y - t.v0.y, t.v2.z - t.v0.z};
            Vec3 c{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
            acc += 0.5f * std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);  // Scene.h:117-126
            L.cum.push_back(acc);
        }
        L.area = acc;
        L.tris = std::move(tris);
        s.lights.push_back(std::move(L));
    };
    for (size_t i = 0; i < prims->arr.size(); i++) {
        const Value &p = prims->arr[i];
        const std::string field = "primitives[" + std::to_string(i) + "]";
        if (!p.is_object()) throw SceneError(path + ": " + field + " must be an object");
        std::string type = req_string(p, "type", field, path);
        bool light = false;
        if (const Value *l = p.find("light")) light = l->is_bool() && l->b;
        if (type == "sphere") {
            if (light)
                throw SceneError(path + ": " + field +
                                 ": light flag is only valid on plane, triangle, or mesh primitives");
            const Value *c = p.find("center"), *r = p.find("radius");
            if (!c || !r || !r->is_number())
                throw SceneError(path + ": " + field + " sphere needs center [x,y,z] and numeric radius");
            Sphere sp;
            sp.center = vec3_of(*c, field + ".center");
            sp.radius = r->as_float();
            sp.mat = mat_of(p, field);
            s.spheres.push_back(sp);
        } else if (type == "triangle") {
            const Value *a = p.find("v0"), *b = p.find("v1"), *c = p.find("v2");
            if (!a || !b || !c) throw SceneError(path + ": " + field + " triangle needs v0, v1, v2 vertices");
            Triangle t;
            t.v0 = vec3_of(*a, field + ".v0");
            t.v1 = vec3_of(*b, field + ".v1");
            t.v2 = vec3_of(*c, field + ".v2");
            t.mat = mat_of(p, field);
            int nn = p.contains("n0") + p.contains("n1") + p.contains("n2");
            if (nn != 0 && nn != 3)
                throw SceneError(path + ": " + field + " triangle normals require all of n0, n1, n2 or none");
            if (nn == 3) {
                t.smooth = true;
                t.n0 = vec3_of(*p.find("n0"), field + ".n0");
                t.n1 = vec3_of(*p.find("n1"), field + ".n1");
                t.n2 = vec3_of(*p.find("n2"), field + ".n2");
            }
            t.orig = (int)s.triangles.size();
            if (light) add_light_tris({t});
            s.triangles.push_back(t);
        } else if (type == "plane") {
            const Value *o = p.find("origin"), *u = p.find("u"), *v = p.find("v");
            if (!o || !u || !v) throw SceneError(path + ": " + field + " plane needs origin, u, v");
            Plane pl;
            pl.o = vec3_of(*o, field + ".origin");
            pl.u = vec3_of(*u, field + ".u");
            pl.v = vec3_of(*v, field + ".v");
            pl.mat = mat_of(p, field);
            if (light) {
                Light L;
                L.kind = 0;
                L.plane = (int)s.lightPlanes.size();
                Vec3 c{pl.u.y * pl.v.z - pl.u.z * pl.v.y, pl.u.z * pl.v.x - pl.u.x * pl.v.z,
                       pl.u.x * pl.v.y - pl.u.y * pl.v.x};
                L.area = std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);  // Plane::getArea
                s.lightPlanes.push_back(pl);
                s.lights.push_back(L);
            } else {
                s.walls.push_back(pl);
            }
        } else if (type == "icosphere") {
            IcoParams ip;
            auto num = [&](const char *key, double &dst) {
                const Value *v = p.find(key);
                if (!v || !v->is_number()) throw SceneError(path + ": " + field + " icosphere needs numeric " + key);
                dst = v->num;
            };
            double sub = 0, seed = 0;
            num("subdiv", sub);
            num("seed", seed);
            num("radius", ip.radius);
            num("displacement", ip.amplitude);
            if (sub < 0 || sub > 8) throw SceneError(path + ": " + field + " icosphere subdiv must be 0..8");
            ip.subdiv = (int)sub;
            ip.seed = (uint32_t)seed;
            const Value *c = p.find("center");
            if (!c) throw SceneError(path + ": " + field + " icosphere needs center");
            Vec3 cc = vec3_of(*c, field + ".center");
            ip.center[0] = cc.x;
            ip.center[1] = cc.y;
            ip.center[2] = cc.z;
            int m = mat_of(p, field);
            std::vector<float> v = make_displaced_icosphere(ip);
            size_t first = s.triangles.size();
            for (size_t k = 0; k + 8 < v.size(); k += 9) {
                Triangle t;
                t.v0 = Vec3{v[k], v[k + 1], v[k + 2]};
                t.v1 = Vec3{v[k + 3], v[k + 4], v[k + 5]};
                t.v2 = Vec3{v[k + 6], v[k + 7], v[k + 8]};
                t.mat = m;
                t.orig = (int)s.triangles.size();
                s.triangles.push_back(t);
            }
            if (light) add_light_tris(std::vector<Triangle>(s.triangles.begin() + first, s.triangles.end()));
        } else if (type == "mesh") {
            throw SceneError(path + ": " + field + ": OBJ meshes are not supported by this loader");
        } else {
            throw SceneError(path + ": " + field + ": unknown primitive type \"" + type + "\"");
        }
    }
    if (s.lights.empty()) throw SceneError(path + ": at least one primitive must have \"light\": true; found none");
    s.bvh = build_bvh_cached(s.triangles);
    for (auto &m : s.materials) populate_spectra(m);
    return s;
}

void pack_scene(const Scene &s, PackedScene &p) {
    p = PackedScene{};
    for (const auto &m : s.materials) {
        p.matAlbedo.insert(p.matAlbedo.end(), {m.albedo.x, m.albedo.y, m.albedo.z});
        p.matEmissive.insert(p.matEmissive.end(), {m.emissive.x, m.emissive.y, m.emissive.z});
        p.matFitA.insert(p.matFitA.end(), {m.fitAlbedo.c0, m.fitAlbedo.c1, m.fitAlbedo.c2, m.fitAlbedo.scale});
        p.matFitE.insert(p.matFitE.end(), {m.fitEmissive.c0, m.fitEmissive.c1, m.fitEmissive.c2, m.fitEmissive.scale});
        p.matTabA.insert(p.matTabA.end(), m.tabAlbedo.begin(), m.tabAlbedo.end());
        p.matTabE.insert(p.matTabE.end(), m.tabEmissive.begin(), m.tabEmissive.end());
        p.matIor.push_back(m.ior);
        p.matCauchy.push_back(m.cauchyB);
        uint32_t f = 0;
        if (m.metallic) f |= PCRHIP_MAT_METALLIC;
        if (m.transparent) f |= PCRHIP_MAT_TRANSPARENT;
        if (m.useTabAlbedo) f |= PCRHIP_MAT_TAB_ALBEDO;
        if (m.useTabEmissive) f |= PCRHIP_MAT_TAB_EMISSIVE;
        p.matFlags.push_back(f);
    }
    for (const auto &sp : s.spheres) {
        p.spheres.insert(p.spheres.end(), {sp.center.x, sp.center.y, sp.center.z, sp.radius});
        p.sphereMat.push_back(sp.mat);
    }
    // intersection order: light planes, then walls (Renderer.cpp:123-130)
    auto put_plane = [](std::vector<float> &v, const Plane &q) {
        v.insert(v.end(), {q.o.x, q.o.y, q.o.z, q.u.x, q.u.y, q.u.z, q.v.x, q.v.y, q.v.z});
    };
    for (const auto &L : s.lights)
        if (L.kind == 0) {
            put_plane(p.planes, s.lightPlanes[L.plane]);
            p.planeMat.push_back(s.lightPlanes[L.plane].mat);
        }
    for (const auto &w : s.walls) {
        put_plane(p.planes, w);
        p.planeMat.push_back(w.mat);
    }
    bool anySmooth = false;
    for (const auto &t : s.triangles) {
        p.tris.insert(p.tris.end(), {t.v0.x, t.v0.y, t.v0.z, t.v1.x, t.v1.y, t.v1.z, t.v2.x, t.v2.y, t.v2.z});
        p.triNormals.insert(p.triNormals.end(), {t.n0.x, t.n0.y, t.n0.z, t.n1.x, t.n1.y, t.n1.z, t.n2.x, t.n2.y, t.n2.z});
        p.triSmooth.push_back(t.smooth ? 1 : 0);
        p.triMat.push_back(t.mat);
        p.triOrig.push_back(t.orig);
        anySmooth |= t.smooth;
    }
    for (const auto &n : s.bvh) {
        int32_t w[9];
        std::memcpy(w, n.bmin, 12);
        std::memcpy(w + 3, n.bmax, 12);
        w[6] = n.left;
        w[7] = n.right;
        w[8] = n.count;
        p.nodes.insert(p.nodes.end(), w, w + 9);
    }
    for (const auto &L : s.lights) {
        p.lightKind.push_back(L.kind);
        p.lightArea.push_back(L.area);
        if (L.kind == 0) {
            p.lightFirst.push_back((int32_t)p.lightPlaneMat.size());
            p.lightCount.push_back(1);
            put_plane(p.lightPlanes, s.lightPlanes[L.plane]);
            p.lightPlaneMat.push_back(s.lightPlanes[L.plane].mat);
        } else {
            p.lightFirst.push_back((int32_t)p.lightTriMat.size());
            p.lightCount.push_back((int32_t)L.tris.size());
            for (size_t k = 0; k < L.tris.size(); k++) {
                const Triangle &t = L.tris[k];
                p.lightTris.insert(p.lightTris.end(),
                                   {t.v0.x, t.v0.y, t.v0.z, t.v1.x, t.v1.y, t.v1.z, t.v2.x, t.v2.y, t.v2.z});
                p.lightTriMat.push_back(t.mat);
                p.lightTriCum.push_back(L.cum[k]);
            }
        }
    }
    pcrhip_scene_desc &d = p.desc;
    d.n_materials = (int)s.materials.size();
    d.mat_albedo = p.matAlbedo.data();
    d.mat_emissive = p.matEmissive.data();
    d.mat_fit_albedo = p.matFitA.data();
    d.mat_fit_emissive = p.matFitE.data();
    d.mat_tab_albedo = p.matTabA.data();
    d.mat_tab_emissive = p.matTabE.data();
    d.mat_flags = p.matFlags.data();
    d.mat_ior = p.matIor.data();
    d.mat_cauchy_b = p.matCauchy.data();
    d.n_spheres = (int)s.spheres.size();
    d.spheres = p.spheres.data();
    d.sphere_mat = p.sphereMat.data();
    d.n_planes = (int)p.planeMat.size();
    d.planes = p.planes.data();
    d.plane_mat = p.planeMat.data();
    d.n_triangles = (int)s.triangles.size();
    d.triangles = p.tris.data();
    d.tri_normals = anySmooth ? p.triNormals.data() : nullptr;
    d.tri_smooth = anySmooth ? p.triSmooth.data() : nullptr;
    d.tri_mat = p.triMat.data();
    d.tri_orig = p.triOrig.data();
    d.n_nodes = (int)s.bvh.size();
    d.nodes = p.nodes.data();
    d.n_lights = (int)s.lights.size();
    d.light_kind = p.lightKind.data();
    d.light_first = p.lightFirst.data();
    d.light_count = p.lightCount.data();
    d.light_area = p.lightArea.data();
    d.n_light_planes = (int)p.lightPlaneMat.size();
    d.light_planes = p.lightPlanes.data();
    d.light_plane_mat = p.lightPlaneMat.data();
    d.n_light_tris = (int)p.lightTriMat.size();
    d.light_tris = p.lightTris.data();
    d.light_tri_mat = p.lightTriMat.data();
    d.light_tri_cum = p.lightTriCum.data();
    d.cam_pos[0] = s.cam.x;
    d.cam_pos[1] = s.cam.y;
    d.cam_pos[2] = s.cam.z;
    d.cam_fov = s.fov;
}

}  // namespace pcr
