// A driver written against the C++ PPPMSolver interface (include/pppm_solver.hpp) the way
// the reference's drivers and tests use its PPPMSolver (experiments/demo/main.cu:142-193,
// tests/PPPMTest.cu): construct, set_mesh, per step update_grid_and_face, read a cell.
//
//   drop_in args                      argument errors surface as std::runtime_error (no GPU use)
//   drop_in run <in.bin> <out.bin>    runs the scene in <in.bin>, writes the listener series
//
// in.bin: int32 res, nv, nt, steps, lx, ly, lz; float32 dl, dt, c, minx, miny, minz;
//         float32 verts[nv*3]; int32 tris[nt*3]; float32 neumann[steps*nt]
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "pppm_solver.hpp"

using pppm_amd::PPPMSolver;

static int args_mode()
{
    int thrown = 0;
    try
    {
        PPPMSolver bad(4, 0.01f, 1e-5f);  // res < 8
    }
    catch (const std::runtime_error &)
    {
        thrown++;
    }
    try
    {
        PPPMSolver bad(32, -1.0f, 1e-5f);  // dl <= 0
    }
    catch (const std::runtime_error &)
    {
        thrown++;
    }
    std::printf("%s %d\n", thrown == 2 ? "ok" : "fail", thrown);
    return thrown == 2 ? 0 : 1;
}

static int run_mode(const char *in_path, const char *out_path)
{
    FILE *f = std::fopen(in_path, "rb");
    if (!f)
        return 2;
    int hdr[7];
    float fh[6];
    if (std::fread(hdr, sizeof(int), 7, f) != 7 || std::fread(fh, sizeof(float), 6, f) != 6)
        return 2;
    const int res = hdr[0], nv = hdr[1], nt = hdr[2], steps = hdr[3];
    std::vector<pppm_amd::float3_> verts(nv);
    std::vector<pppm_amd::int3_> tris(nt);
    std::vector<float> neumann((size_t)steps * nt);
    if (std::fread(verts.data(), sizeof(float), (size_t)nv * 3, f) != (size_t)nv * 3 ||
        std::fread(tris.data(), sizeof(int), (size_t)nt * 3, f) != (size_t)nt * 3 ||
        std::fread(neumann.data(), sizeof(float), neumann.size(), f) != neumann.size())
        return 2;
    std::fclose(f);

    PPPMSolver solver(res, fh[0], fh[1], {fh[3], fh[4], fh[5]}, 0, 0.83f, fh[2]);
    solver.set_mesh(verts, tris);
    std::vector<float> out(steps);
    std::vector<float> row(nt);
    for (int i = 0; i < steps; i++)
    {
        std::memcpy(row.data(), neumann.data() + (size_t)i * nt, nt * sizeof(float));
        solver.update_grid_and_face(row);
        out[i] = solver.read_cell(hdr[4], hdr[5], hdr[6]);
    }
    FILE *g = std::fopen(out_path, "wb");
    if (!g)
        return 2;
    std::fwrite(out.data(), sizeof(float), out.size(), g);
    std::fclose(g);
    std::printf("ran %d steps, t = %d\n", steps, solver.time_idx());
    return 0;
}

int main(int argc, char **argv)
{
    try
    {
        if (argc >= 2 && !std::strcmp(argv[1], "args"))
            return args_mode();
        if (argc >= 4 && !std::strcmp(argv[1], "run"))
            return run_mode(argv[2], argv[3]);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "usage: drop_in args | run <in.bin> <out.bin>\n");
    return 2;
}
