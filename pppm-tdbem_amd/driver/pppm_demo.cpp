// pppm_demo — the reference's demo driver (experiments/demo/main.cu:38-275) on the MI355X
// library: reads the same scene config, moves Audio objects along their motion, feeds their
// surface accelerations, steps the PPPM solver once per audio sample, applies the demo's
// chunked restart with a cross-fade, and writes result.txt in the same layout.
//
//   pppm_demo <scene.cfg>                    run; writes <outputDir>/result.txt
//   pppm_demo --setup <scene.cfg>            print the derived grid / timing parameters
//   pppm_demo --calls <scene.cfg> <out.bin>  record every solver call of a run without a GPU
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

#include "pppm_solver.hpp"
#include "scene.hpp"

namespace pppm_demo
{

namespace
{
struct P2
{
    float s, t;
};
P2 edge02(float a11, float b1)
{
    const int neg = b1 < 0, past = a11 + b1 <= 0;
    return {0.0f, (float)neg * ((float)past - (float)(1 - past) * (b1 / a11))};
}
P2 edge12(float a01, float a11, float b1, float f10, float f01)
{
    const float g0 = a01 + b1 - f10, g1 = a11 + b1 - f01;
    const int neg = g0 < 0, past = g1 <= 0;
    const float t = (float)neg * ((float)past + (float)(1 - past) * g0 / (g0 - g1));
    return {1 - t, t};
}
P2 inside(P2 q0, float g0, P2 q1, float g1)
{
    const float z = g0 / (g0 - g1), w = 1 - z;
    return {w * q0.s + z * q1.s, w * q0.t + z * q1.t};
}
}  // namespace

F3 closest_point(F3 p, F3 v0, F3 v1, F3 v2)
{
    const F3 d = p - v0, e0 = v1 - v0, e1 = v2 - v0;
    const float a00 = dot(e0, e0), a01 = dot(e0, e1), a11 = dot(e1, e1);
    const float b0 = -dot(d, e0), b1 = -dot(d, e1);
    const float f00 = b0, f10 = b0 + a00, f01 = b0 + a01;
    P2 q0{0, 0}, q1{0, 0}, r{0, 0};
    float g0, g1;
    if (f00 >= 0)
    {
        if (f01 >= 0)
            r = edge02(a11, b1);
        else
        {
            q0 = {0, f00 / (f00 - f01)};
            q1.s = f01 / (f01 - f10);
            q1.t = 1 - q1.s;
            const float dt1 = q1.t - q0.t;
            g0 = dt1 * (a11 * q0.t + b1);
            if (g0 >= 0)
                r = edge02(a11, b1);
            else
            {
                g1 = dt1 * (a01 * q1.s + a11 * q1.t + b1);
                r = g1 <= 0 ? edge12(a01, a11, b1, f10, f01) : inside(q0, g0, q1, g1);
            }
        }
    }
    else if (f01 <= 0)
    {
        if (f10 <= 0)
            r = edge12(a01, a11, b1, f10, f01);
        else
        {
            q0 = {f00 / (f00 - f10), 0};
            q1.s = f01 / (f01 - f10);
            q1.t = 1 - q1.s;
            g0 = q1.t * (a01 * q0.s + b1);
            if (g0 >= 0)
                r = q0;
            else
            {
                g1 = q1.t * (a01 * q1.s + a11 * q1.t + b1);
                r = g1 <= 0 ? edge12(a01, a11, b1, f10, f01) : inside(q0, g0, q1, g1);
            }
        }
    }
    else if (f10 <= 0)
    {
        q0 = {0, f00 / (f00 - f01)};
        q1.s = f01 / (f01 - f10);
        q1.t = 1 - q1.s;
        const float dt1 = q1.t - q0.t;
        g0 = dt1 * (a11 * q0.t + b1);
        if (g0 >= 0)
            r = edge02(a11, b1);
        else
        {
            g1 = dt1 * (a01 * q1.s + a11 * q1.t + b1);
            r = g1 <= 0 ? edge12(a01, a11, b1, f10, f01) : inside(q0, g0, q1, g1);
        }
    }
    else
    {
        q0 = {f00 / (f00 - f10), 0};
        q1 = {0, f00 / (f00 - f01)};
        g0 = q1.t * (a01 * q0.s + b1);
        if (g0 >= 0)
            r = q0;
        else
        {
            g1 = q1.t * (a11 * q1.t + b1);
            r = g1 <= 0 ? edge02(a11, b1) : inside(q0, g0, q1, g1);
        }
    }
    return (v0 + r.s * e0) + r.t * e1;
}

// grid and timing derived from the scene (experiments/demo/main.cu:94-147)
struct Setup
{
    int res = 0, frame_rate = 0, frame_num = 0, chunk = 0, back = 0;
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
    s.start_time = 0.0f + cfg.num("startTime");
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
    void update_mesh(const std::vector<F3> &v) override
    {
        solver.update_mesh(reinterpret_cast<const pppm_amd::float3_ *>(v.data()));
    }
    void step(const std::vector<float> &n) override { solver.update_grid_and_face(n); }
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

// AudioAndManualTest's loop (experiments/demo/main.cu:166-272)
std::vector<float> run(const Setup &s, ObjectCollection &col, Sink &sink)
{
    std::vector<float> result((size_t)s.frame_num + 2, 0.0f);
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
        const size_t k = (size_t)progress + 1;
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
                const float p = sink.read(j + 1, 1 - factor, s.check);
                result[(size_t)j + 1] = p * (1 - factor);
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
                result[(size_t)j + 1] += sink.read(j + 1, -1.0f, s.check);
                ok = !nan_at(progress);
                if (ok)
                    progress++;
            }
            if (!ok)
                break;
            continue;
        }
        update_sound(false);
        result[(size_t)i + 1] = sink.read(i + 1, 1.0f, s.check);
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
        else if (argc == 2)
            cfg_path = argv[1];
        else
        {
            std::fprintf(stderr, "usage: pppm_demo <scene.cfg> | --setup <scene.cfg> | --calls <scene.cfg> <out>\n");
            return 2;
        }
        const SceneConfig cfg = SceneConfig::load(cfg_path);
        ObjectCollection col(cfg.dirs.at("inputDir"), object_list(cfg));
        const Setup s = derive(cfg, col);
        if (mode == "setup")
        {
            std::printf("res %d\ngrid_size %.9g\nmin_pos %.9g %.9g %.9g\nframe_rate %d\ndt %.9g\nframe_num %d\n"
                        "check %d %d %d\nchunk %d %d\nvertices %zu\ntriangles %zu\n",
                        s.res, s.grid_size, s.min_pos.x, s.min_pos.y, s.min_pos.z, s.frame_rate, s.dt, s.frame_num,
                        s.check.x, s.check.y, s.check.z, s.chunk, s.back, col.verts.size(), col.tris.size());
            return 0;
        }
        if (mode == "calls")
        {
            RecordSink rec(out);
            run(s, col, rec);
            return 0;
        }
        GpuSink gpu(s);
        const std::vector<float> r = run(s, col, gpu);
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
