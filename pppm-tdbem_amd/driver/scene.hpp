// Scene layer of the demo driver: the reference's config layout, OBJ meshes, Manual and Audio
// objects and their concatenation, the rigid bodies' data files, on the host (C++17, no device
// code). The numerics that
// decide which mesh and which accelerations a step sees follow the reference's float
// operations:
//   SimpleJsonReader          src/utils/simple_json_reader.h:13-86
//   Mesh::loadOBJ, bbox       src/utils/objIO.cu:94-108,169-220
//   ManualObject, AudioObject src/utils/Object.h:73-323, src/utils/Object.cu:6-270
//   ObjectCollection          src/utils/ObjectCollection.h, ObjectCollection.cu:22-160
//   RigidBody data files      src/utils/RigidBody.cu:90-188 (motion, contact impulses, tet mesh,
//                             eigen pairs), FindAllSurfaces src/utils/Object.cu:72-142,
//                             MaterialParameters src/utils/material.h:34-96
// Mesh repair (fix_mesh: a pymesh container) is not run; meshes are used as given, and the
// selected-surface map that repair would have built is computed against the same mesh.
#pragma once

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace pppm_demo
{

struct F3
{
    float x, y, z;
};
struct F4
{
    float x, y, z, w;
};
struct I3
{
    int x, y, z;
};
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(F3 a) { return std::sqrt(dot(a, a)); }

// ------------------------------------------------------------------ config
// The reference's line-oriented "json-like" layout: '{', inputDir, outputDir, numeric
// key: value lines up to the line holding '[', then per object '{', name, type, '}' up to ']'.
struct SceneConfig
{
    std::map<std::string, std::string> dirs;
    std::map<std::string, float> nums;
    std::vector<std::map<std::string, std::string>> objects;

    float num(const std::string &k, float fallback = 0.0f) const
    {
        auto it = nums.find(k);
        return it == nums.end() ? fallback : it->second;
    }

    static std::string trim(const std::string &s)
    {
        auto junk = [](char c) { return c == ' ' || c == '\t' || c == '"' || c == '\r'; };
        size_t b = 0, e = s.size();
        while (b < e && junk(s[b]))
            b++;
        while (e > b && junk(s[e - 1]))
            e--;
        return s.substr(b, e - b);
    }
    static void key_value(const std::string &line, std::string &k, std::string &v)
    {
        const size_t colon = line.find(':');
        const size_t comma = std::min(line.find(','), line.size());
        if (colon == std::string::npos)
            throw std::runtime_error("config: expected 'key: value' in line: " + line);
        k = trim(line.substr(0, colon));
        v = trim(line.substr(colon + 1, comma > colon ? comma - colon - 1 : 0));
    }

    static SceneConfig load(const std::string &path)
    {
        std::ifstream in(path);
        if (!in.good())
            throw std::runtime_error("config: cannot open " + path);
        SceneConfig c;
        std::string line, k, v;
        std::getline(in, line);
        if (trim(line) != "{")
            throw std::runtime_error("config: first line must be '{'");
        for (int i = 0; i < 2; i++)
        {
            std::getline(in, line);
            key_value(line, k, v);
            c.dirs[k] = v;
        }
        while (std::getline(in, line))
        {
            if (line.find('[') != std::string::npos)
                break;
            if (trim(line).empty())
                continue;
            key_value(line, k, v);
            c.nums[k] = std::stof(v);
        }
        while (std::getline(in, line))
        {
            if (line.find(']') != std::string::npos)
                break;
            std::map<std::string, std::string> one;
            for (int i = 0; i < 2; i++)
            {
                std::getline(in, line);
                key_value(line, k, v);
                one[k] = v;
            }
            std::getline(in, line);  // closing brace of the entry
            c.objects.push_back(std::move(one));
        }
        return c;
    }
};

// ------------------------------------------------------------------ meshes
struct Mesh
{
    std::vector<F3> verts;
    std::vector<I3> tris;

    // Wavefront OBJ: 'v x y z' and triangular 'f a b c' (a, a/b, a/b/c or a//c; 1-based,
    // negative = relative), in file order.
    static Mesh load_obj(const std::string &path)
    {
        std::ifstream in(path);
        if (!in.good())
            throw std::runtime_error("mesh: cannot open " + path);
        Mesh m;
        std::string line;
        while (std::getline(in, line))
        {
            std::istringstream ss(line);
            std::string tag;
            if (!(ss >> tag))
                continue;
            if (tag == "v")
            {
                F3 p{};
                ss >> p.x >> p.y >> p.z;
                m.verts.push_back(p);
            }
            else if (tag == "f")
            {
                std::vector<int> ids;
                std::string tok;
                while (ss >> tok)
                {
                    int id = std::stoi(tok.substr(0, tok.find('/')));
                    ids.push_back(id > 0 ? id - 1 : (int)m.verts.size() + id);
                }
                if (ids.size() != 3)
                    throw std::runtime_error("mesh: only triangles are supported (" + path + ")");
                m.tris.push_back({ids[0], ids[1], ids[2]});
            }
        }
        return m;
    }
};

struct BBox
{
    F3 min{1e10f, 1e10f, 1e10f}, max{-1e10f, -1e10f, -1e10f};
    void add(F3 p)
    {
        min = {std::min(min.x, p.x), std::min(min.y, p.y), std::min(min.z, p.z)};
        max = {std::max(max.x, p.x), std::max(max.y, p.y), std::max(max.z, p.z)};
    }
    void add(const BBox &b)
    {
        add(b.min);
        add(b.max);
    }
    F3 center() const { return {(min.x + max.x) / 2, (min.y + max.y) / 2, (min.z + max.z) / 2}; }
    float length() const { return std::max(std::max(max.x - min.x, max.y - min.y), max.z - min.z); }
};

// rotate v by the unit quaternion (x, y, z, w) (Object.cu:6-11)
inline F3 rotate(F4 q, F3 v)
{
    const F3 u{q.x, q.y, q.z};
    const float s = q.w;
    return ((2.0f * dot(u, v)) * u + (s * s - dot(u, u)) * v) + (2.0f * s) * cross(u, v);
}

// point of triangle (a, b, c) closest to p (Voronoi-region tests; a vertex of the triangle is
// returned exactly, so the triangles that hold a mesh vertex are at distance 0 from it)
F3 closest_point(F3 p, F3 a, F3 b, F3 c);

// ------------------------------------------------------------------ objects
struct ObjectState
{
    std::vector<F3> verts;
    int anim = -1, acc = 0;
    // rigid bodies (RigidBodyState, RigidBody.h:47-56): clock, next impulse, oscillators
    float time = 0;
    int imp = 0;
    std::vector<float> q3, accs;
};

// motion.txt: "t tx ty tz qx qy qz qw" per line (Object::LoadMotion_, Object.cu:195-222)
inline void load_motion_file(const std::string &path, std::vector<float> &times, std::vector<F3> &trans,
                             std::vector<F4> &rots)
{
    std::ifstream in(path);
    if (!in.good())
        return;
    std::string line;
    while (std::getline(in, line))
    {
        std::istringstream ss(line);
        float t;
        F3 p;
        F4 q;
        if (!(ss >> t >> p.x >> p.y >> p.z >> q.x >> q.y >> q.z >> q.w))
            continue;
        times.push_back(t);
        trans.push_back(p);
        rots.push_back(q);
    }
}

// Rayleigh damping (alpha, beta) of the reference's materials (material.h:34-96)
inline void material_damping(const std::string &name, float &alpha, float &beta)
{
    static const std::map<std::string, std::pair<float, float>> table = {
        {"ceramic", {40.0f, 1e-7f}}, {"plastic", {30.0f, 1e-6f}}, {"steel", {5.0f, 3e-8f}},
        {"mdf", {35.0f, 5e-6f}},     {"wood", {60.0f, 2e-6f}}};
    auto it = table.find(name);
    alpha = it == table.end() ? 0.0f : it->second.first;
    beta = it == table.end() ? 0.0f : it->second.second;
}

// Boundary triangles of a tetrahedral mesh, oriented outwards (FindAllSurfaces,
// Object.cu:72-142): faces used by exactly one tetrahedron, in the iteration order of a hash
// map keyed by the face's vertex triple with the reference's hash ((x << 20) + (y << 10) + z),
// flipped when the tetrahedron's fourth vertex lies on the normal's side.
struct TetSurface
{
    std::vector<I3> tris;
    std::vector<F3> normals;
};
inline TetSurface find_all_surfaces(const std::vector<std::array<int, 4>> &tets, const std::vector<F3> &v)
{
    struct Key
    {
        int x, y, z;
        bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash
    {
        int operator()(const Key &k) const { return (k.x << 20) + (k.y << 10) + k.z; }
    };
    struct Info
    {
        int cnt = 0, tet = 0, except = 0;
    };
    std::unordered_map<Key, Info, KeyHash> faces;
    for (int i = 0; i < (int)tets.size(); i++)
        for (int i0 = 0; i0 < 4; i0++)
        {
            int k[3], n = 0;
            for (int j = 0; j < 4; j++)
                if (j != i0)
                    k[n++] = tets[i][j];
            Info &f = faces[Key{k[0], k[1], k[2]}];
            f.cnt++;
            f.tet = i;
            f.except = i0;
        }
    TetSurface out;
    for (auto &kv : faces)
    {
        if (kv.second.cnt != 1)
            continue;
        I3 t{kv.first.x, kv.first.y, kv.first.z};
        const F3 c = (v[t.x] + v[t.y] + v[t.z]) / 3;
        const F3 ex = v[tets[kv.second.tet][kv.second.except]] - c;
        const F3 e1 = v[t.y] - v[t.x], e2 = v[t.z] - v[t.x];
        const F3 cr = cross(e1, e2);
        F3 nrm = cr * (1.0f / std::sqrt(dot(cr, cr)));
        if (dot(nrm, ex) > 0)
        {
            std::swap(t.y, t.z);
            nrm = F3{-nrm.x, -nrm.y, -nrm.z};
        }
        out.tris.push_back(t);
        out.normals.push_back(nrm);
    }
    return out;
}

// A rigid body's data directory (RigidBody::load_data, RigidBody.cu:90-188): motion.txt,
// contact.txt ("t vertex nx ny nz magnitude"), vertices.txt ("x y z"), tets.txt (4 vertex
// indices, sorted per tetrahedron), eigenvalues.txt, eigenvectors.txt (3 nv rows of all
// eigenvalues' columns); modes are kept up to the last eigenvalue below max_frequency Hz.
struct Impulse
{
    float time;
    int vertex;
    F3 vec;
};
struct RigidBodyData
{
    std::vector<F3> verts;
    TetSurface surf;
    std::vector<float> times;
    std::vector<F3> trans;
    std::vector<F4> rots;
    std::vector<Impulse> impulses;
    std::vector<float> eigvals;  // M
    std::vector<float> eigvecs;  // 3 nv x M, row-major
};
inline RigidBodyData load_rigid_body(const std::string &dir, float max_frequency = 15000.0f)
{
    RigidBodyData d;
    load_motion_file(dir + "/motion.txt", d.times, d.trans, d.rots);
    auto open = [&](const std::string &f) {
        std::ifstream in(dir + "/" + f);
        if (!in.good())
            throw std::runtime_error("cannot read " + dir + "/" + f);
        return in;
    };
    std::string line;
    {
        std::ifstream in = open("contact.txt");
        while (std::getline(in, line))
        {
            if (line.empty())
                continue;
            std::istringstream ss(line);
            float t, vid, mag;
            F3 nrm;
            ss >> t >> vid >> nrm.x >> nrm.y >> nrm.z >> mag;
            d.impulses.push_back({t, (int)(vid + 0.5f), nrm * mag});  // F2I (macro.h:81)
        }
    }
    {
        std::ifstream in = open("vertices.txt");
        while (std::getline(in, line))
        {
            if (line.empty())
                continue;
            std::istringstream ss(line);
            F3 p;
            ss >> p.x >> p.y >> p.z;
            d.verts.push_back(p);
        }
    }
    std::vector<std::array<int, 4>> tets;
    {
        std::ifstream in = open("tets.txt");
        while (std::getline(in, line))
        {
            if (line.empty())
                continue;
            std::istringstream ss(line);
            float a[4];
            ss >> a[0] >> a[1] >> a[2] >> a[3];
            std::array<int, 4> t{(int)(a[0] + 0.5f), (int)(a[1] + 0.5f), (int)(a[2] + 0.5f), (int)(a[3] + 0.5f)};
            std::sort(t.begin(), t.end());
            for (int k : t)
                if (k < 0 || k >= (int)d.verts.size())
                    throw std::runtime_error(dir + "/tets.txt: vertex index out of range");
            tets.push_back(t);
        }
    }
    d.surf = find_all_surfaces(tets, d.verts);
    std::vector<float> all;
    int modes = 0;
    {
        std::ifstream in = open("eigenvalues.txt");
        while (std::getline(in, line))
        {
            if (line.empty())
                continue;
            std::istringstream ss(line);
            float val;
            ss >> val;
            all.push_back(val);
            if (std::sqrt(val) / (2 * M_PI) < max_frequency)
                modes = (int)all.size();
        }
    }
    d.eigvals.assign(all.begin(), all.begin() + modes);
    const size_t rows = 3 * d.verts.size();
    d.eigvecs.assign(rows * modes, 0.0f);
    {
        std::ifstream in = open("eigenvectors.txt");
        for (size_t i = 0; i < rows; i++)
            for (size_t j = 0; j < all.size(); j++)
            {
                float val = 0;
                in >> val;
                if ((int)j < modes)
                    d.eigvecs[i * modes + j] = val;
            }
    }
    return d;
}

class Object
{
  public:
    virtual ~Object() = default;
    std::string name;
    std::vector<F3> verts;
    std::vector<I3> tris;
    virtual bool update_until(float time) = 0;           // true if the mesh moved
    virtual void submit_accelerations(float *out) = 0;   // one value per triangle
    virtual BBox bbox() = 0;
    virtual void save(ObjectState &) {}
    virtual void load(const ObjectState &) {}
    virtual float time_step() const { return FLT_MAX; }
};

// no motion, no sound (Object.h:73-132)
class ManualObject : public Object
{
  public:
    explicit ManualObject(const std::string &dir)
    {
        Mesh m = Mesh::load_obj(dir + "/mesh.obj");
        verts = m.verts;
        tris = m.tris;
    }
    bool update_until(float) override { return false; }
    void submit_accelerations(float *out) override { std::fill(out, out + tris.size(), 0.0f); }
    BBox bbox() override
    {
        BBox b;
        b.min = b.max = verts.empty() ? F3{0, 0, 0} : verts[0];
        for (auto &v : verts)
            b.add(v);
        return b;
    }
};

// rigid motion from motion.txt, surface accelerations from accs.txt at 44.1 kHz on the
// triangles nearest to selected_vertices.txt (Object.h:134-323, Object.cu:182-270)
class AudioObject : public Object
{
  public:
    explicit AudioObject(const std::string &dir, float sample_rate = 44100.0f)
    {
        Mesh m = Mesh::load_obj(dir + "/mesh.obj");
        verts = m.verts;
        base = m.verts;
        tris = m.tris;
        sample_time = 1 / sample_rate;
        load_motion(dir + "/motion.txt");
        load_accs(dir + "/accs.txt");
        load_selected(dir + "/selected_vertices.txt");
    }
    float time_step() const override { return sample_time; }

    bool update_until(float time) override
    {
        const bool moved = animate_until(time);
        // AccelerationUpdateUntil_: the largest k with sample_time * k <= time, at least k + 1 - 1
        acc++;
        while (sample_time * acc <= time)
            acc++;
        acc--;
        return moved;
    }
    void submit_accelerations(float *out) override
    {
        if (acc >= (int)accs.size())
        {
            std::fill(out, out + tris.size(), 0.0f);
            return;
        }
        for (size_t i = 0; i < tris.size(); i++)
            out[i] = selected[i] > 0 ? accs[acc] : 0.0f;
    }
    BBox bbox() override
    {
        BBox b;
        if (times.empty())
        {
            b.min = b.max = verts.empty() ? F3{0, 0, 0} : verts[0];
            for (auto &v : verts)
                b.add(v);
            return b;
        }
        for (size_t f = 0; f < times.size(); f++)
            for (auto &v : base)
                b.add(rotate(rots[f], v) + trans[f]);
        return b;
    }
    void save(ObjectState &s) override
    {
        s.verts = verts;
        s.anim = anim;
        s.acc = acc;
    }
    void load(const ObjectState &s) override
    {
        verts = s.verts;
        anim = s.anim;
        acc = s.acc;
    }

    std::vector<float> times, accs;
    std::vector<F3> trans;
    std::vector<F4> rots;
    std::vector<int> selected;  // per triangle: number of selected vertices mapped to it
    int anim = -1, acc = 0;

  private:
    std::vector<F3> base;
    float sample_time = 0;

    bool animate_until(float time)
    {
        if (anim + 1 >= (int)times.size())
            return false;
        bool moved = false;
        while (anim + 1 < (int)times.size() && times[anim + 1] <= time)
        {
            anim++;
            moved = true;
        }
        if (moved)
            for (size_t i = 0; i < base.size(); i++)
                verts[i] = rotate(rots[anim], base[i]) + trans[anim];
        return moved;
    }
    void load_motion(const std::string &path) { load_motion_file(path, times, trans, rots); }
    // LoadAccs_ (Object.cu:228-247): a value counts only if the stream is still good after
    // reading it, so a last value not followed by whitespace is dropped
    void load_accs(const std::string &path)
    {
        std::ifstream in(path);
        if (!in.good())
            return;
        float a = 0;
        while (true)
        {
            in >> a;
            if (!in.good())
                break;
            accs.push_back(a);
        }
    }
    // LoadCover_ (Object.cu:249-268) keeps the value of the read that fails at the end of the
    // file (C++ stores 0 there), so vertex 0 is selected whenever the list ends in whitespace.
    // The selection goes to the triangle closest to each listed vertex, first one on ties
    // (FindNearestVertex, Object.cu:37-70).
    void load_selected(const std::string &path)
    {
        selected.assign(tris.size(), 0);
        std::ifstream in(path);
        if (accs.empty())
            return;
        if (!in.good())
        {
            std::fill(selected.begin(), selected.end(), 1);
            return;
        }
        std::vector<int> ids;
        int v = 0;
        while (true)
        {
            in >> v;
            ids.push_back(v);
            if (!in.good())
                break;
        }
        for (int id : ids)
        {
            if (id < 0 || id >= (int)verts.size())
                continue;
            const F3 p = verts[id];
            float best = FLT_MAX;
            int bf = -1;
            for (size_t f = 0; f < tris.size(); f++)
            {
                const F3 q = closest_point(p, verts[tris[f].x], verts[tris[f].y], verts[tris[f].z]);
                const float d = length(q - p);
                if (d < best)
                {
                    best = d;
                    bf = (int)f;
                }
            }
            if (bf >= 0)
                selected[bf]++;
        }
    }
};

// ------------------------------------------------------------------ collection
class ObjectCollection
{
  public:
    using ModalFactory = std::function<std::unique_ptr<Object>(const std::string &dir)>;
    ObjectCollection(const std::string &dir, const std::vector<std::pair<std::string, std::string>> &named,
                     const ModalFactory &modal = nullptr)
    {
        size_t nv = 0, nt = 0;
        for (auto &[name, type] : named)
        {
            std::unique_ptr<Object> o;
            if (type == "Manual")
                o = std::make_unique<ManualObject>(dir + "/" + name);
            else if (type == "Audio")
                o = std::make_unique<AudioObject>(dir + "/" + name);
            else if (type == "Modal" && modal)
                o = modal(dir + "/" + name);
            else if (type == "Modal")
                throw std::runtime_error("Modal objects run their oscillators on the GPU: not available in this mode");
            else
                throw std::runtime_error("unsupported sound type: " + type + " (Audio, Manual and Modal are supported)");
            o->name = name;
            vert_off.push_back(nv);
            tri_off.push_back(nt);
            nv += o->verts.size();
            nt += o->tris.size();
            objects.push_back(std::move(o));
        }
        verts.resize(nv);
        tris.resize(nt);
        accs.assign(nt, 0.0f);
        update_mesh();
    }
    BBox bbox()
    {
        BBox b;
        for (auto &o : objects)
            b.add(o->bbox());
        return b;
    }
    float time_step() const
    {
        float m = FLT_MAX;
        for (auto &o : objects)
            m = std::min(m, o->time_step());
        return m;
    }
    void update_mesh()
    {
        for (size_t k = 0; k < objects.size(); k++)
        {
            auto &o = *objects[k];
            std::copy(o.verts.begin(), o.verts.end(), verts.begin() + vert_off[k]);
            for (size_t i = 0; i < o.tris.size(); i++)
            {
                const I3 t = o.tris[i];
                const int off = (int)vert_off[k];
                tris[tri_off[k] + i] = {t.x + off, t.y + off, t.z + off};
            }
        }
    }
    void update_acc()
    {
        for (size_t k = 0; k < objects.size(); k++)
            objects[k]->submit_accelerations(accs.data() + tri_off[k]);
    }

    std::vector<std::unique_ptr<Object>> objects;
    std::vector<size_t> vert_off, tri_off;
    std::vector<F3> verts;
    std::vector<I3> tris;
    std::vector<float> accs;
};

}  // namespace pppm_demo
