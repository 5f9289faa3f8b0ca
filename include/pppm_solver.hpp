// pppm_solver.hpp — header-only C++ host interface over the C ABI of libpppm_hip.so.
//
// Mirrors the reference's PPPMSolver (src/algorithm/pppm.h:36-163) so that code written
// against it keeps its shape: same method names, same argument meaning, same call order
// (set_mesh, update_mesh, update_grid_and_face / update_step, the solve_* phases), and
// failures surface as std::runtime_error like the reference's cuSynchronize
// (src/core/macro.h:119-127). Arrays are plain host or device pointers (or std::vector);
// the library copies what it is given.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "pppm_hip.h"

namespace pppm_amd
{

struct float3_
{
    float x, y, z;
};
struct int3_
{
    int x, y, z;
};

class PPPMSolver
{
  public:
    // PPPMSolver(res, dl, dt, min_pos, pml_width, reflect_coeff) (pppm.h:39-55); c is the
    // reference's AIR_WAVE_SPEED (macro.h:82) made a parameter.
    PPPMSolver(int res, float dl, float dt, float3_ min_pos = {0, 0, 0}, int pml_width = 0,
               float reflect_coeff = 0.83f, float c = 340.3f, unsigned flags = 0)
        : res_(res), dl_(dl), dt_(dt), min_(min_pos)
    {
        const float mn[3] = {min_pos.x, min_pos.y, min_pos.z};
        check(pppm_create(&h_, res, dl, dt, mn, pml_width, reflect_coeff, c, flags), "PPPMSolver");
    }
    ~PPPMSolver() { clear(); }
    PPPMSolver(const PPPMSolver &) = delete;
    PPPMSolver &operator=(const PPPMSolver &) = delete;

    // accessors (pppm.h:56-73)
    float dt() const { return dt_; }
    float dl() const { return dl_; }
    float grid_size() const { return dl_; }
    int res() const { return res_; }
    float3_ min_coord() const { return min_; }
    float3_ max_coord() const { return {min_.x + res_ * dl_, min_.y + res_ * dl_, min_.z + res_ * dl_}; }
    int time_idx() const
    {
        int t = 0;
        check(pppm_time_idx(h_, &t), "time_idx");
        return t;
    }

    // set_mesh / update_mesh (pppm.h:75-109). on_device: the pointers are device memory.
    void set_mesh(const float3_ *verts, int nv, const int3_ *tris, int nt, bool on_device = false)
    {
        check(pppm_set_mesh(h_, &verts->x, nv, &tris->x, nt, on_device), "set_mesh");
        nt_ = nt;
        mesh_set = true;
    }
    void set_mesh(const std::vector<float3_> &verts, const std::vector<int3_> &tris)
    {
        set_mesh(verts.data(), (int)verts.size(), tris.data(), (int)tris.size());
    }
    void update_mesh(const float3_ *verts, bool on_device = false)
    {
        check(pppm_update_mesh(h_, &verts->x, on_device), "update_mesh");
    }
    void update_mesh(const std::vector<float3_> &verts) { update_mesh(verts.data()); }

    // update_grid_and_face / update_step (pppm.h:137-152): one time step with the Neumann
    // input of every triangle (nt floats).
    void update_grid_and_face(const float *neumann, bool on_device = false)
    {
        check(pppm_step(h_, neumann, on_device), "update_grid_and_face");
    }
    void update_grid_and_face(const std::vector<float> &neumann) { update_grid_and_face(neumann.data()); }
    void update_step(const std::vector<float> &neumann) { update_grid_and_face(neumann); }

    // update_mesh then update_grid_and_face in one call: the demo's per-sample UpdateSound with
    // a moving mesh (experiments/demo/main.cu:170-205)
    void update_mesh_and_step(const float3_ *verts, const float *neumann, bool on_device = false)
    {
        check(pppm_update_mesh_and_step(h_, &verts->x, neumann, on_device), "update_mesh_and_step");
    }

    // the phases, as PPPMTest.cu and FaceCacheTest.cu call them (pppm.h:127-135)
    void fdtd_step() { check(pppm_fdtd_step(h_), "fdtd.step"); }
    void solve_fdtd_far() { check(pppm_solve_far(h_), "solve_fdtd_far"); }
    void set_neumann_condition(const float *neumann, bool on_device = false)
    {
        check(pppm_set_neumann(h_, neumann, on_device), "set_neumann_condition");
    }
    void update_dirichlet() { check(pppm_update_dirichlet(h_), "update_dirichlet"); }
    void solve_fdtd_near() { check(pppm_solve_near(h_), "solve_fdtd_near"); }
    void reset() { check(pppm_reset(h_), "reset"); }

    // grids[t + rel] as a host copy (x-major, z contiguous; src/core/array3D.h:54-87) or as
    // a borrowed device pointer
    const float *grid_device(int rel = 0) const
    {
        float *p = nullptr;
        check(pppm_grid_ptr(h_, rel, &p), "grid_ptr");
        return p;
    }
    // pressure at a few cells (the demo's per-step listener read, experiments/demo/main.cu:266)
    std::vector<float> read_cells(const std::vector<std::array<int, 3>> &xyz) const
    {
        std::vector<float> out(xyz.size());
        check(pppm_read_cells(h_, xyz.empty() ? nullptr : xyz[0].data(), (int)xyz.size(), out.data()),
              "read_cells");
        return out;
    }
    float read_cell(int x, int y, int z) const { return read_cells({{x, y, z}})[0]; }

    void set_stream(void *hip_stream) { check(pppm_set_stream(h_, hip_stream), "set_stream"); }

    // PPPMSolver::clear (pppm.h:111-125)
    void clear()
    {
        if (h_)
            pppm_destroy(h_);
        h_ = nullptr;
        mesh_set = false;
    }

    pppm_solver *handle() const { return h_; }
    bool mesh_set = false;

  private:
    void check(int rc, const char *what) const
    {
        if (rc != PPPM_OK)
        {
            const char *m = h_ ? pppm_last_error(h_) : "";
            throw std::runtime_error(std::string("pppm ") + what + " failed (" + std::to_string(rc) + "): " +
                                     (m ? m : ""));
        }
    }

    pppm_solver *h_ = nullptr;
    int res_;
    float dl_, dt_;
    float3_ min_;
    int nt_ = 0;
};

}  // namespace pppm_amd
