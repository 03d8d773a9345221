// Wavefront OBJ/MTL import into the scene arrays of the C ABI: the subset of
// AssimpAssetImporter::Import (import/assimp_assetimporter.cpp:20-146) and of the Assimp 3.x
// steps it requests (OBJ loader, aiProcess_Triangulate, aiProcess_JoinIdenticalVertices,
// aiProcess_GenNormals) that the reference's assets use. Host code only.
//
// What the reference builds from a file, and what this import returns:
//  * materials: Assimp's default material first (diffuse 0.6), then the MTL materials in the
//    order they are defined. A material named "light" or with a nonzero Ke becomes
//    Emissive(0.6 (60, 55, 40)) (assimp_assetimporter.cpp:74-85), every other one
//    SimpleMaterial(Lambert(Kd)).
//  * bundles: one per non-empty (object, material) run of faces, in file order: `o` starts an
//    object, and `usemtl` starts a new mesh inside it (Assimp's OBJ loader).
//  * triangles: a quad a b c d becomes (a b c), (a c d), started at its concave corner if it has
//    one (Assimp's TriangulateProcess); a convex polygon of n > 4 corners is fanned from its first
//    corner. A face without normals gets its triangle's normal normalize((p2 - p1) x (p3 - p1))
//    (aiProcess_GenNormals).
//  * vertices: per bundle, the distinct (position, normal) pairs of its faces in first-use order
//    (aiProcess_JoinIdenticalVertices), so a bundle is the reference Mesh's vertex list.
//  * lights: one AreaLight(shape, bundle) per triangle of every bundle whose material is emissive
//    (assimp_assetimporter.cpp:137-145), in bundle then triangle order.
// Numbers are read with strtof (correctly rounded); texture maps (map_Kd) are counted and not
// applied.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/banshee_hip.h"

namespace bsh {
bsh_status fail(bsh_status code, const std::string& msg);
}

namespace {

struct V3d
{
    float x, y, z;
};

struct ObjMaterial
{
    std::string name;
    float kd[3] = {0.6f, 0.6f, 0.6f};   // Assimp's default diffuse
    float ke[3] = {0.f, 0.f, 0.f};
    bool has_map = false;
};

struct FaceVert
{
    int v, n;   // 0-based; n = -1 when the face has no normal
};

struct ObjMesh
{
    int material = 0;
    std::vector<std::vector<FaceVert>> faces;
};

std::string dir_of(const std::string& path)
{
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

bool read_floats(std::istringstream& ss, float* out, int n)
{
    for (int k = 0; k < n; ++k)
    {
        std::string tok;
        if (!(ss >> tok)) return false;
        char* end = nullptr;
        out[k] = std::strtof(tok.c_str(), &end);
        if (end == tok.c_str()) return false;
    }
    return true;
}

// MTL: newmtl, Kd, Ke, map_Kd (other statements do not change what the reference builds)
bool load_mtl(const std::string& path, std::vector<ObjMaterial>& mats, std::map<std::string, int>& index,
              std::string& err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open material library " + path;
        return false;
    }
    std::string line;
    int cur = -1;
    while (std::getline(f, line))
    {
        std::istringstream ss(line);
        std::string key;
        if (!(ss >> key) || key[0] == '#') continue;
        if (key == "newmtl")
        {
            std::string name;
            ss >> name;
            ObjMaterial m;
            m.name = name;
            index[name] = (int)mats.size();
            cur = (int)mats.size();
            mats.push_back(m);
        }
        else if (cur >= 0 && key == "Kd")
        {
            if (!read_floats(ss, mats[cur].kd, 3)) { err = "bad Kd in " + path; return false; }
        }
        else if (cur >= 0 && key == "Ke")
        {
            if (!read_floats(ss, mats[cur].ke, 3)) { err = "bad Ke in " + path; return false; }
        }
        else if (cur >= 0 && key == "map_Kd")
        {
            mats[cur].has_map = true;
        }
    }
    return true;
}

// One OBJ index: 1-based, or negative = relative to the elements read so far.
bool resolve(const std::string& s, int count, int& out)
{
    if (s.empty()) return false;
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (*end != '\0' || v == 0) return false;
    out = v > 0 ? (int)v - 1 : count + (int)v;
    return out >= 0 && out < count;
}

struct Asset
{
    std::vector<float> verts, normals;
    std::vector<int32_t> tv, tn, tm;
    std::vector<bsh_bundle> bundles;
    std::vector<bsh_material> mats;
    std::vector<bsh_area_light> lights;
    size_t maps = 0;
};

V3d sub(V3d a, V3d b) { return V3d{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3d nrm(V3d a)
{
    float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 0.f ? V3d{a.x / l, a.y / l, a.z / l} : a;
}
float dot(V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3d cross(V3d a, V3d b) { return V3d{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

bool import(const std::string& path, Asset& A, std::string& err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<V3d> pos, nor;
    std::vector<ObjMaterial> mats(1);
    mats[0].name = "DefaultMaterial";
    std::map<std::string, int> mindex;
    mindex[mats[0].name] = 0;
    std::vector<ObjMesh> meshes;    // in creation order (objects in file order, usemtl runs inside)
    int cur_mat = 0;
    bool have_mesh = false;
    std::string line;
    int lineno = 0;
    auto new_mesh = [&]() {
        ObjMesh m;
        m.material = cur_mat;
        meshes.push_back(m);
        have_mesh = true;
    };
    while (std::getline(f, line))
    {
        ++lineno;
        std::istringstream ss(line);
        std::string key;
        if (!(ss >> key) || key[0] == '#') continue;
        if (key == "v" || key == "vn")
        {
            float p[3];
            if (!read_floats(ss, p, 3)) { err = path + ":" + std::to_string(lineno) + ": bad " + key; return false; }
            (key == "v" ? pos : nor).push_back(V3d{p[0], p[1], p[2]});
        }
        else if (key == "mtllib")
        {
            std::string lib;
            while (ss >> lib)
                if (!load_mtl(dir_of(path) + lib, mats, mindex, err)) return false;
        }
        else if (key == "o")
        {
            new_mesh();
        }
        else if (key == "usemtl")
        {
            std::string name;
            ss >> name;
            auto it = mindex.find(name);
            cur_mat = it == mindex.end() ? 0 : it->second;   // unknown names take the default material
            new_mesh();
        }
        else if (key == "f")
        {
            if (!have_mesh) new_mesh();
            std::vector<FaceVert> fv;
            std::string tok;
            while (ss >> tok)
            {
                FaceVert v{-1, -1};
                std::string a = tok, c;
                size_t s1 = tok.find('/');
                if (s1 != std::string::npos)
                {
                    a = tok.substr(0, s1);
                    size_t s2 = tok.find('/', s1 + 1);
                    if (s2 != std::string::npos) c = tok.substr(s2 + 1);
                }
                if (!resolve(a, (int)pos.size(), v.v) || (!c.empty() && !resolve(c, (int)nor.size(), v.n)))
                {
                    err = path + ":" + std::to_string(lineno) + ": bad face index " + tok;
                    return false;
                }
                fv.push_back(v);
            }
            if (fv.size() < 3) { err = path + ":" + std::to_string(lineno) + ": face with fewer than 3 corners"; return false; }
            meshes.back().faces.push_back(fv);
        }
    }
    // materials, as the reference's onmaterial_ callback receives them
    std::vector<bool> emissive(mats.size());
    for (size_t i = 0; i < mats.size(); ++i)
    {
        const ObjMaterial& m = mats[i];
        bsh_material b;
        std::memset(&b, 0, sizeof(b));
        const float ke2 = m.ke[0] * m.ke[0] + m.ke[1] * m.ke[1] + m.ke[2] * m.ke[2];
        emissive[i] = ke2 > 0.f || m.name == "light";
        if (emissive[i])
        {
            b.type = (float)BSH_EMISSIVE;
            b.color[0] = 0.6f * 60.f;
            b.color[1] = 0.6f * 55.f;
            b.color[2] = 0.6f * 40.f;
        }
        else
        {
            b.type = (float)BSH_LAMBERT;
            for (int k = 0; k < 3; ++k) b.color[k] = m.kd[k];
            A.maps += m.has_map ? 1 : 0;
        }
        A.mats.push_back(b);
    }
    // meshes -> bundles
    for (const ObjMesh& m : meshes)
    {
        if (m.faces.empty()) continue;
        bsh_bundle b;
        b.tri_begin = (int32_t)A.tm.size();
        b.vert_begin = b.norm_begin = (int32_t)(A.verts.size() / 3);
        std::map<std::tuple<float, float, float, float, float, float>, int> seen;
        auto vertex = [&](V3d p, V3d n) -> int32_t {
            auto key = std::make_tuple(p.x, p.y, p.z, n.x, n.y, n.z);
            auto it = seen.find(key);
            if (it != seen.end()) return it->second;
            int32_t id = (int32_t)(A.verts.size() / 3);
            A.verts.insert(A.verts.end(), {p.x, p.y, p.z});
            A.normals.insert(A.normals.end(), {n.x, n.y, n.z});
            seen[key] = id;
            return id;
        };
        for (const auto& fv : m.faces)
        {
            const size_t n = fv.size();
            size_t start = 0;
            if (n == 4)
            {
                // a quad has at most one concave corner; fan from it (TriangulateProcess)
                for (size_t i = 0; i < 4; ++i)
                {
                    V3d v = pos[fv[i].v];
                    V3d left = nrm(sub(pos[fv[(i + 3) % 4].v], v)), diag = nrm(sub(pos[fv[(i + 2) % 4].v], v)),
                        right = nrm(sub(pos[fv[(i + 1) % 4].v], v));
                    if (std::acos(dot(left, diag)) + std::acos(dot(right, diag)) > 3.14159265358979323846f)
                    {
                        start = i;
                        break;
                    }
                }
            }
            for (size_t t = 1; t + 1 < n; ++t)
            {
                const FaceVert c[3] = {fv[start], fv[(start + t) % n], fv[(start + t + 1) % n]};
                V3d p[3] = {pos[c[0].v], pos[c[1].v], pos[c[2].v]};
                V3d gn = nrm(cross(sub(p[1], p[0]), sub(p[2], p[0])));
                for (int k = 0; k < 3; ++k)
                {
                    const int32_t id = vertex(p[k], c[k].n >= 0 ? nor[c[k].n] : gn);
                    A.tv.push_back(id);
                    A.tn.push_back(id);
                }
                A.tm.push_back(m.material);
            }
        }
        b.tri_count = (int32_t)A.tm.size() - b.tri_begin;
        b.vert_count = b.norm_count = (int32_t)(A.verts.size() / 3) - b.vert_begin;
        const int32_t bi = (int32_t)A.bundles.size();
        A.bundles.push_back(b);
        if (emissive[m.material])
            for (int32_t s = 0; s < b.tri_count; ++s) A.lights.push_back(bsh_area_light{bi, s});
    }
    return true;
}

}  // namespace

extern "C" bsh_status bsh_import_obj(const char* path, bsh_asset_counts* counts, float* verts, float* normals,
                                     int32_t* tri_vidx, int32_t* tri_nidx, int32_t* tri_material,
                                     bsh_bundle* bundles, bsh_material* materials, bsh_area_light* lights)
{
    if (!path || !counts) return bsh::fail(BSH_ERR_INVALID, "bsh_import_obj: NULL path or counts");
    Asset A;
    std::string err;
    if (!import(path, A, err)) return bsh::fail(BSH_ERR_INVALID, "bsh_import_obj: " + err);
    bsh_asset_counts c;
    c.nverts = A.verts.size() / 3;
    c.nnormals = A.normals.size() / 3;
    c.ntris = A.tm.size();
    c.nbundles = A.bundles.size();
    c.nmaterials = A.mats.size();
    c.nlights = A.lights.size();
    c.ignored_maps = A.maps;
    const bool want = verts || normals || tri_vidx || tri_nidx || tri_material || bundles || materials || lights;
    if (want)
    {
        if (!(verts && normals && tri_vidx && tri_nidx && tri_material && bundles && materials && (lights || !c.nlights)))
            return bsh::fail(BSH_ERR_INVALID, "bsh_import_obj: pass every output array, or none to query the counts");
        if (counts->nverts < c.nverts || counts->nnormals < c.nnormals || counts->ntris < c.ntris ||
            counts->nbundles < c.nbundles || counts->nmaterials < c.nmaterials || counts->nlights < c.nlights)
            return bsh::fail(BSH_ERR_INVALID, "bsh_import_obj: output arrays smaller than the counts of the asset");
        std::memcpy(verts, A.verts.data(), A.verts.size() * sizeof(float));
        std::memcpy(normals, A.normals.data(), A.normals.size() * sizeof(float));
        std::memcpy(tri_vidx, A.tv.data(), A.tv.size() * sizeof(int32_t));
        std::memcpy(tri_nidx, A.tn.data(), A.tn.size() * sizeof(int32_t));
        std::memcpy(tri_material, A.tm.data(), A.tm.size() * sizeof(int32_t));
        std::memcpy(bundles, A.bundles.data(), A.bundles.size() * sizeof(bsh_bundle));
        std::memcpy(materials, A.mats.data(), A.mats.size() * sizeof(bsh_material));
        if (c.nlights) std::memcpy(lights, A.lights.data(), A.lights.size() * sizeof(bsh_area_light));
    }
    *counts = c;
    return BSH_OK;
}
