// pppm_demo — the reference's demo driver (experiments/demo/main.cu:38-275) on the MI355X
// library: reads the same scene config, moves Audio objects along their motion, feeds their
// surface accelerations, steps the PPPM solver once per audio sample, applies the demo's
// chunked restart with a cross-fade, and writes result.txt in the same layout. Modal objects
// (RigidBody, src/utils/RigidBody.cu) run their oscillators through the library's modal source
// at the solver's frame rate, with the contact impulses as modal forces.
//
//   pppm_demo <scene.cfg>                    run; writes <outputDir>/result.txt
//   pppm_demo --setup <scene.cfg>            print the derived grid / timing parameters
//   pppm_demo --calls <scene.cfg> <out.bin>  record every solver call of a run without a GPU
//   pppm_demo --trace <scene.cfg> <out.bin>  run (writes result.txt) and record the calls too
//
// The call record (for checking the loop against another solver) is a sequence of records:
//   int32 op; op 0 set_mesh: int32 nv, nt, float32 verts[3 nv], int32 tris[3 nt]
//             op 1 update_mesh: float32 verts[3 nv]
//             op 2 step: float32 neumann[nt]
//             op 3 read: int32 result_index, float32 weight (result[i] = w * p, or += p if w < 0)
//             op 4 reset: int32 t (fdtd/neumann/dirichlet reset, then fdtd.t = t)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "pppm_hip.h"
#include "pppm_solver.hpp"
#include "scene.hpp"

namespace pppm_demo
{

// Closest point of triangle (a, b, c) to p by Voronoi regions (Ericson, "Real-Time Collision
// Detection", 2005, section 5.1.5): the vertex regions, then the edge regions, then the face,
// each decided from the dot products of the edges with p's offsets from the vertices. A point in
// a vertex's region gets that vertex itself, so a mesh vertex is at distance 0 from every
// triangle that holds it.
F3 closest_point(F3 p, F3 a, F3 b, F3 c)
{
    const F3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0)
        return a;
    const F3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3)
        return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0)
        return a + (d1 / (d1 - d3)) * ab;
    const F3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6)
        return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0)
        return a + (d2 / (d2 - d6)) * ac;
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0)
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    const float den = 1.0f / (va + vb + vc);
    return a + (vb * den) * ab + (vc * den) * ac;
}

// A rigid body with modal sound (RigidBody, src/utils/RigidBody.h:53-166): per audio sample one
// animation frame at most, the contact impulses active at the sample's time as modal forces
// (CalculateIIR_, RigidBody.cu:254-295, summed on the host as the reference does), and the
// oscillators and surface accelerations of that sample on the GPU (pppm_modal_run).
class ModalObject : public Object
{
  public:
    explicit ModalObject(const std::string &dir, const std::string &material = "plastic")
        : d(load_rigid_body(dir))
    {
        material_damping(material, alpha, beta);
        verts = d.verts;
        tris = d.surf.tris;
        if (tris.empty() || d.eigvals.empty())
            throw std::runtime_error(dir + ": a modal object needs a closed surface and at least one mode");
    }
    ~ModalObject() override
    {
        if (m)
            pppm_modal_destroy(m);
        if (accs_dev)
            (void)hipFree(accs_dev);
    }
    // RigidBody::set_sample_rate (RigidBody.h:64-70) / InitIIR_ (RigidBody.cu:190-205)
    void set_sample_rate(float rate)
    {
        timestep = 1.0f / rate;
        if (m)
            pppm_modal_destroy(m);
        m = nullptr;
        if (pppm_modal_create(&m, d.eigvecs.data(), (int)d.verts.size(), (int)d.eigvals.size(), d.eigvals.data(),
                              &tris[0].x, (int)tris.size(), &d.surf.normals[0].x, alpha, beta, rate,
                              nullptr) != PPPM_OK)
            throw std::runtime_error(name + ": pppm_modal_create failed");
        if (!accs_dev && hipMalloc((void **)&accs_dev, tris.size() * sizeof(float)) != hipSuccess)
            throw std::runtime_error(name + ": hipMalloc failed");
        if (hipMemset(accs_dev, 0, tris.size() * sizeof(float)) != hipSuccess)
            throw std::runtime_error(name + ": hipMemset failed");
        time = 0;
        anim = 0;
    }
    float time_step() const override { return timestep; }
    bool update_until(float t) override
    {
        while (time < t)
            audio_step();
        return moved;
    }
    void submit_accelerations(float *out) override
    {
        if (hipMemcpy(out, accs_dev, tris.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error(name + ": hipMemcpy failed");
    }
    // get_bbox (RigidBody.cu:310-338): the frames from the one before the first impulse on
    BBox bbox() override
    {
        BBox b;
        size_t f = 0;
        if (!d.impulses.empty() && d.times.size() > 1)
        {
            f = 1;
            while (f < d.times.size() && d.times[f] < d.impulses[0].time)
                f++;
            f--;
        }
        for (; f < d.times.size(); f++)
            for (auto &v : d.verts)
                b.add(rotate(d.rots[f], v) + d.trans[f]);
        if (d.times.empty())
            for (auto &v : verts)
                b.add(v);
        return b;
    }
    // move_to_first_impulse (RigidBody.cu:355-362)
    void move_to_first_impulse()
    {
        if (d.impulses.empty())
            return;
        while (time + timestep < d.impulses[0].time)
            audio_step();
    }
    void save(ObjectState &s) override
    {
        s.verts = verts;
        s.anim = anim;
        s.time = time;
        s.imp = imp;
        s.q3.resize(3 * d.eigvals.size());
        s.accs.resize(tris.size());
        if (pppm_modal_state(m, s.q3.data()) != PPPM_OK)
            throw std::runtime_error(name + ": pppm_modal_state failed");
        submit_accelerations(s.accs.data());
    }
    void load(const ObjectState &s) override
    {
        verts = s.verts;
        anim = s.anim;
        time = s.time;
        imp = s.imp;
        if (pppm_modal_set_state(m, s.q3.data()) != PPPM_OK ||
            hipMemcpy(accs_dev, s.accs.data(), tris.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error(name + ": restoring the modal state failed");
        // the active-impulse list is not part of the saved state (RigidBodyState, RigidBody.h:47-56)
    }
    float time = 0;

  private:
    RigidBodyData d;
    float alpha = 0, beta = 0, timestep = 0;
    pppm_modal *m = nullptr;
    float *accs_dev = nullptr;
    int anim = 0, imp = 0;
    bool moved = false;
    std::vector<Impulse> active;
    std::vector<float> force;

    // audio_step (RigidBody.cu:340-353) and CalculateIIR_'s impulse part (RigidBody.cu:257-281)
    void audio_step()
    {
        moved = false;
        if (anim < (int)d.times.size() && d.times[anim] <= time)
        {
            for (size_t i = 0; i < verts.size(); i++)
                verts[i] = rotate(d.rots[anim], d.verts[i]) + d.trans[anim];
            anim++;
            moved = true;
        }
        time += timestep;
        while (imp < (int)d.impulses.size() && time >= d.impulses[imp].time)
            active.push_back(d.impulses[imp++]);
        while (!active.empty() && time - active[0].time > 0.0005f)  // Impulse::dead, RigidBody.h:30
            active.erase(active.begin());
        const int M = (int)d.eigvals.size();
        force.assign(M, 0.0f);
        for (int i = 0; i < M; i++)
            for (auto &a : active)
            {
                const float amp = (float)std::sin(M_PI * (time - a.time) / 0.0005f);  // Impulse::amp
                const F3 I = a.vec * amp;
                const size_t r = 3 * (size_t)a.vertex;
                force[i] += d.eigvecs[r * M + i] * I.x + d.eigvecs[(r + 1) * M + i] * I.y +
                            d.eigvecs[(r + 2) * M + i] * I.z;
            }
        if (pppm_modal_run(m, force.data(), 1, accs_dev, 0) != PPPM_OK)
            throw std::runtime_error(name + ": pppm_modal_run failed");
    }
};

// grid and timing derived from the scene (experiments/demo/main.cu:94-147)
struct Setup
{
    int res = 0, frame_rate = 0, frame_num = 0, chunk = 0, back = 0, mute = 0;
    float grid_size = 0, dt = 0, reflect = 0, start_time = 0, max_time = 0;
    F3 min_pos{};
    I3 check{};
};

Setup derive(const SceneConfig &cfg, ObjectCollection &col)
{
    Setup s;
    const BBox box = col.bbox();
    const F3 centre = box.center();
    const float grid_length = box.length() * cfg.num("gridLengthFactor");
    s.reflect = cfg.num("reflectCoeff");
    s.res = (int)cfg.num("res");
    if (s.res < 8)
        throw std::runtime_error("config: res must be >= 8");
    s.grid_size = grid_length / s.res;
    s.min_pos = centre - F3{grid_length / 2, grid_length / 2, grid_length / 2};
    // 1.01 x the CFL rate at the reference's speed of sound (AIR_WAVE_SPEED, macro.h:82)
    s.frame_rate = (int)(1.01f / (s.grid_size / std::sqrt(3) / 340.3f));
    s.dt = 1.0f / s.frame_rate;
    // modal objects run at the frame rate and skip ahead to their first contact; the samples
    // before it are silent (main.cu:104-124)
    float curr = 0.0f;
    for (auto &o : col.objects)
        if (auto *rb = dynamic_cast<ModalObject *>(o.get()))
        {
            rb->set_sample_rate((float)s.frame_rate);
            rb->move_to_first_impulse();
            curr = std::max(curr, rb->time);
        }
    s.mute = (int)(curr * s.frame_rate);
    col.update_mesh();
    s.start_time = curr + cfg.num("startTime");
    s.max_time = cfg.num("maxTime") + s.start_time;
    s.frame_num = (int)((s.max_time - s.start_time) / s.dt);
    s.check = {(int)(cfg.num("checkCoordX") * s.res), (int)(cfg.num("checkCoordY") * s.res),
               (int)(cfg.num("checkCoordZ") * s.res)};
    s.chunk = (int)cfg.num("chunkSize");
    s.back = (int)cfg.num("backStepSize");
    return s;
}

// The solver side of the loop: the GPU library, or a recorder of the calls.
struct Sink
{
    virtual ~Sink() = default;
    virtual void set_mesh(const std::vector<F3> &v, const std::vector<I3> &t) = 0;
    virtual void update_mesh(const std::vector<F3> &v) = 0;
    virtual void step(const std::vector<float> &neumann) = 0;
    virtual float read(int idx, float w, I3 c) = 0;  // returns the pressure (NaN check)
    virtual void reset(int t) = 0;
    virtual int time_idx() = 0;
};

struct GpuSink : Sink
{
    pppm_amd::PPPMSolver solver;
    explicit GpuSink(const Setup &s)
        : solver(s.res, s.grid_size, s.dt, {s.min_pos.x, s.min_pos.y, s.min_pos.z}, 0, s.reflect, 340.3f)
    {
    }
    void set_mesh(const std::vector<F3> &v, const std::vector<I3> &t) override
    {
        solver.set_mesh(reinterpret_cast<const pppm_amd::float3_ *>(v.data()), (int)v.size(),
                        reinterpret_cast<const pppm_amd::int3_ *>(t.data()), (int)t.size());
    }
    // a mesh update is held back and issued with the next step as one call (update_mesh, then
    // update_grid_and_face: the same two operations without a host round trip between them)
    void update_mesh(const std::vector<F3> &v) override
    {
        pending = v;
        has_pending = true;
    }
    void step(const std::vector<float> &n) override
    {
        if (has_pending)
        {
            solver.update_mesh_and_step(reinterpret_cast<const pppm_amd::float3_ *>(pending.data()), n.data());
            has_pending = false;
        }
        else
            solver.update_grid_and_face(n);
    }
    std::vector<F3> pending;
    bool has_pending = false;
    float read(int, float, I3 c) override { return solver.read_cell(c.x, c.y, c.z); }
    void reset(int t) override
    {
        solver.reset();
        pppm_set_time_idx(solver.handle(), t);
    }
    int time_idx() override { return solver.time_idx(); }
};

struct RecordSink : Sink
{
    FILE *f;
    int t = -1;
    explicit RecordSink(const std::string &path) : f(std::fopen(path.c_str(), "wb"))
    {
        if (!f)
            throw std::runtime_error("cannot write " + path);
    }
    ~RecordSink() override { std::fclose(f); }
    void op(int o) { std::fwrite(&o, sizeof(int), 1, f); }
    void set_mesh(const std::vector<F3> &v, const std::vector<I3> &tr) override
    {
        op(0);
        const int nv = (int)v.size(), nt = (int)tr.size();
        std::fwrite(&nv, sizeof(int), 1, f);
        std::fwrite(&nt, sizeof(int), 1, f);
        std::fwrite(v.data(), sizeof(F3), v.size(), f);
        std::fwrite(tr.data(), sizeof(I3), tr.size(), f);
    }
    void update_mesh(const std::vector<F3> &v) override
    {
        op(1);
        std::fwrite(v.data(), sizeof(F3), v.size(), f);
    }
    void step(const std::vector<float> &n) override
    {
        op(2);
        std::fwrite(n.data(), sizeof(float), n.size(), f);
        t++;
    }
    float read(int idx, float w, I3) override
    {
        op(3);
        std::fwrite(&idx, sizeof(int), 1, f);
        std::fwrite(&w, sizeof(float), 1, f);
        return 0.0f;
    }
    void reset(int tt) override
    {
        op(4);
        std::fwrite(&tt, sizeof(int), 1, f);
        t = tt;
    }
    int time_idx() override { return t; }
};

// the GPU run with every call also recorded
struct TeeSink : Sink
{
    Sink &a, &b;
    TeeSink(Sink &gpu, Sink &rec) : a(gpu), b(rec) {}
    void set_mesh(const std::vector<F3> &v, const std::vector<I3> &t) override
    {
        a.set_mesh(v, t);
        b.set_mesh(v, t);
    }
    void update_mesh(const std::vector<F3> &v) override
    {
        a.update_mesh(v);
        b.update_mesh(v);
    }
    void step(const std::vector<float> &n) override
    {
        a.step(n);
        b.step(n);
    }
    float read(int idx, float w, I3 c) override
    {
        b.read(idx, w, c);
        return a.read(idx, w, c);
    }
    void reset(int t) override
    {
        a.reset(t);
        b.reset(t);
    }
    int time_idx() override { return a.time_idx(); }
};

// AudioAndManualTest's loop (experiments/demo/main.cu:166-272)
std::vector<float> run(const Setup &s, ObjectCollection &col, Sink &sink)
{
    std::vector<float> result((size_t)s.mute + s.frame_num + 2, 0.0f);
    const size_t mute = (size_t)s.mute;
    result[0] = (float)s.frame_rate;
    float curr = s.start_time;
    bool mesh_set = false;
    std::vector<float> silent(col.accs.size(), 0.0f);

    auto update_sound = [&](bool mute) {
        bool moved = false;
        for (auto &o : col.objects)
            moved = moved || o->update_until(curr + s.dt);  // short-circuits, as the reference
        curr += s.dt;
        col.update_mesh();
        col.update_acc();
        if (!mesh_set)
        {
            sink.set_mesh(col.verts, col.tris);
            mesh_set = true;
        }
        else if (moved)
            sink.update_mesh(col.verts);
        sink.step(mute ? silent : col.accs);
    };

    // CheckNaN looks at result[progress + 1] (main.cu:158-165), which is the sample just
    // written inside the restart blocks but the next, unwritten one on the plain path
    auto nan_at = [&](int progress) {
        const size_t k = mute + progress + 1;
        return k < result.size() && std::isnan(result[k]);
    };
    int progress = 0;
    const bool chunked = s.chunk > 0 && s.back > 0 && s.back <= s.chunk;
    while (progress <= s.frame_num)
    {
        const int i = progress;
        if (chunked && i % s.chunk == s.chunk - s.back)
        {
            const int end = std::min(i + s.back, s.frame_num + 1);
            const int saved_t = sink.time_idx();
            const float saved_curr = curr;
            std::vector<ObjectState> states(col.objects.size());
            for (size_t k = 0; k < col.objects.size(); k++)
                col.objects[k]->save(states[k]);
            // fade the running simulation out with silent input
            bool ok = true;
            for (int j = i; j < end && ok; j++)
            {
                update_sound(true);
                const float factor = (float)(j - i) / (float)(end - i);
                const float p = sink.read((int)mute + j + 1, 1 - factor, s.check);
                result[mute + j + 1] = p * (1 - factor);
                ok = !nan_at(progress);
            }
            if (!ok)
                break;
            // restart from silence at the same time and fade it in
            sink.reset(saved_t);
            curr = saved_curr;
            for (size_t k = 0; k < col.objects.size(); k++)
                col.objects[k]->load(states[k]);
            for (int j = i; j < end && ok; j++)
            {
                update_sound(false);
                result[mute + j + 1] += sink.read((int)mute + j + 1, -1.0f, s.check);
                ok = !nan_at(progress);
                if (ok)
                    progress++;
            }
            if (!ok)
                break;
            continue;
        }
        update_sound(false);
        result[mute + i + 1] = sink.read((int)mute + i + 1, 1.0f, s.check);
        progress++;
        if (nan_at(progress))
            break;
    }
    return result;
}

// write_to_txt (src/utils/array_writer.h:22-31): "%e" per line, no newline after the last
void write_result(const std::string &path, const std::vector<float> &r)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f)
        throw std::runtime_error("cannot write " + path);
    for (size_t i = 0; i + 1 < r.size(); i++)
        std::fprintf(f, "%e\n", r[i]);
    std::fprintf(f, "%e", r.back());
    std::fclose(f);
}

std::vector<std::pair<std::string, std::string>> object_list(const SceneConfig &cfg)
{
    std::vector<std::pair<std::string, std::string>> v;
    for (auto &o : cfg.objects)
    {
        auto n = o.find("name"), t = o.find("type");
        if (n == o.end() || t == o.end())
            throw std::runtime_error("config: each object needs name and type");
        if (n->second == "water")
            throw std::runtime_error("config: the water object is not supported");
        v.emplace_back(n->second, t->second);
    }
    return v;
}

}  // namespace pppm_demo

int main(int argc, char **argv)
{
    using namespace pppm_demo;
    try
    {
        std::string mode = "run", cfg_path, out;
        if (argc >= 3 && !std::strcmp(argv[1], "--setup"))
            mode = "setup", cfg_path = argv[2];
        else if (argc >= 4 && !std::strcmp(argv[1], "--calls"))
            mode = "calls", cfg_path = argv[2], out = argv[3];
        else if (argc >= 4 && !std::strcmp(argv[1], "--trace"))
            mode = "trace", cfg_path = argv[2], out = argv[3];
        else if (argc == 2)
            cfg_path = argv[1];
        else
        {
            std::fprintf(stderr, "usage: pppm_demo <scene.cfg> | --setup <scene.cfg> | --calls <scene.cfg> <out> | "
                                 "--trace <scene.cfg> <out>\n");
            return 2;
        }
        const SceneConfig cfg = SceneConfig::load(cfg_path);
        // Modal objects (material "plastic", main.cu:62-66) need the GPU: not in the recording mode
        ObjectCollection::ModalFactory modal;
        if (mode != "calls")
            modal = [](const std::string &dir) { return std::unique_ptr<Object>(new ModalObject(dir)); };
        ObjectCollection col(cfg.dirs.at("inputDir"), object_list(cfg), modal);
        const Setup s = derive(cfg, col);
        if (mode == "setup")
        {
            std::printf("res %d\ngrid_size %.9g\nmin_pos %.9g %.9g %.9g\nframe_rate %d\ndt %.9g\nframe_num %d\n"
                        "check %d %d %d\nchunk %d %d\nvertices %zu\ntriangles %zu\nmute %d\n",
                        s.res, s.grid_size, s.min_pos.x, s.min_pos.y, s.min_pos.z, s.frame_rate, s.dt, s.frame_num,
                        s.check.x, s.check.y, s.check.z, s.chunk, s.back, col.verts.size(), col.tris.size(), s.mute);
            return 0;
        }
        if (mode == "calls")
        {
            RecordSink rec(out);
            run(s, col, rec);
            return 0;
        }
        GpuSink gpu(s);
        std::vector<float> r;
        if (mode == "trace")
        {
            RecordSink rec(out);
            TeeSink tee(gpu, rec);
            r = run(s, col, tee);
        }
        else
            r = run(s, col, gpu);
        const std::string odir = cfg.dirs.at("outputDir");
        std::filesystem::create_directories(odir);
        write_result(odir + "/result.txt", r);
        std::printf("wrote %s/result.txt (%d frames at %d Hz)\n", odir.c_str(), s.frame_num, s.frame_rate);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "pppm_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}
