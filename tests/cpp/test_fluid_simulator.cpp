// test_fluid_simulator.cpp -- the C++ host mirror (FluidSimulator /
// D2Q9Node over the gfx950 library) against the CPU restatement of the
// reference shaders (oracle/wgsl_restate.c), frame by frame, through the
// same sequence of user edits the reference's UI produces: creation with
// the Poiseuille scene, frames, a click (obstacle), a drag (external force
// cells with a block_iter countdown), a viscosity change and a reset.
// Every comparison is bit-exact: populations of non-solid cells in both
// buffers, the exported macro field and the lattice info records.
//
// Needs a gfx950 device.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "d2q9_node.hpp"

extern "C" {
typedef struct {
    int32_t material;
    int32_t block_iter;
    float vx;
    float vy;
} ws_info;
void ws_init(int nx, int ny, int fluid_ty, ws_info* info, float* A, float* B, float* macro4);
void ws_run(int nx, int ny, float tau, ws_info* info, float* buf0, float* buf1, float* macro4, int nsteps,
            int start_swap);
}

using namespace simuverse;

namespace {

int failures = 0;

void expect(bool ok, const char* what)
{
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    std::fflush(stdout);
    if (!ok) ++failures;
}

struct Oracle {
    int nx, ny;
    float tau;
    std::vector<ws_info> info;
    std::vector<float> A, B, macro4;
    Oracle(int nx_, int ny_, float tau_, const std::vector<LatticeInfo>& init)
        : nx(nx_), ny(ny_), tau(tau_), info(init.size()), A((size_t)9 * nx_ * ny_), B(A.size()),
          macro4((size_t)4 * nx_ * ny_)
    {
        std::memcpy(info.data(), init.data(), init.size() * sizeof(ws_info));
    }
    void reset(int fluid_ty) { ws_init(nx, ny, fluid_ty, info.data(), A.data(), B.data(), macro4.data()); }
    void frames(int k) { ws_run(nx, ny, tau, info.data(), A.data(), B.data(), macro4.data(), 2 * k, 0); }
};

bool is_solid(int32_t m) { return m == 2 || m == 4; }

bool same_state(const D2Q9Node& node, const Oracle& o, bool with_macro)
{
    const size_t n = (size_t)o.nx * o.ny;
    std::vector<float> f(9 * n);
    std::vector<LatticeInfo> info(n);
    if (lbm_read_lattice_info(node.handle(), info.data(), 0, n) != LBM_OK) return false;
    for (size_t c = 0; c < n; ++c)
        if (std::memcmp(&info[c], &o.info[c], sizeof(ws_info)) != 0) {
            std::printf("  lattice info differs at cell %zu\n", c);
            return false;
        }
    for (int b = 0; b < 2; ++b) {
        node.read_populations(b, f.data());
        const std::vector<float>& ref = b ? o.B : o.A;
        for (size_t c = 0; c < n; ++c) {
            if (is_solid(o.info[c].material)) continue;
            for (int i = 0; i < 9; ++i)
                if (std::memcmp(&f[c + i * n], &ref[c + i * n], 4) != 0) {
                    std::printf("  buffer %d differs at cell %zu direction %d: %.9g vs %.9g\n", b, c, i,
                                f[c + i * n], ref[c + i * n]);
                    return false;
                }
        }
    }
    if (with_macro) {
        std::vector<float> rho(n), ux(n), uy(n);
        node.read_macro(rho.data(), ux.data(), uy.data());
        for (size_t c = 0; c < n; ++c)
            if (std::memcmp(&ux[c], &o.macro4[4 * c], 4) || std::memcmp(&uy[c], &o.macro4[4 * c + 1], 4) ||
                std::memcmp(&rho[c], &o.macro4[4 * c + 2], 4)) {
                std::printf("  macro differs at cell %zu\n", c);
                return false;
            }
    }
    return true;
}

}  // namespace

int main()
{
    // Reference conventions: canvas pixels / lattice_pixel_size cells.
    const uint32_t W = 512, H = 256, px = 2;
    SettingObj setting;
    setting.fluid_viscosity = 0.05f;
    setting.animation_type = FieldAnimationType::Poiseuille;
    std::printf("case: FluidSimulator %ux%u canvas, pixel %u\n", W, H, px);
    std::fflush(stdout);
    FluidSimulator sim(W, H, px, setting);
    D2Q9Node& node = sim.fluid_compute_node();
    const int nx = (int)node.width(), ny = (int)node.height();
    expect(nx == 256 && ny == 128, "lattice size = canvas / lattice_pixel_size");

    // Materials: walls, ghosts, inlet and outlet columns, three disks whose
    // cells satisfy |p - c| <= 28 for the centres of lattice.rs:37-39.
    {
        const auto& info = node.lattice_info_data();
        size_t disk = 0, brute = 0;
        const float cx[3] = {nx / 7.0f - 28.0f, nx / 5.0f, nx / 5.0f};
        const float cy[3] = {ny / 2.0f, ny / 4.0f, ny * 0.75f};
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 2; x < nx - 2; ++x) {
                bool in = false;
                for (int k = 0; k < 3; ++k)
                    in |= (x - cx[k]) * (x - cx[k]) + (y - cy[k]) * (y - cy[k]) <= 28.0f * 28.0f;
                brute += in;
                disk += info[(size_t)y * nx + x].material == 4;
            }
        expect(disk == brute && disk > 0, "obstacle cells = brute-force disk count");
        expect(info[0].material == 2 && info[(size_t)5 * nx].material == 7 && info[(size_t)5 * nx + 1].material == 3 &&
                   info[(size_t)5 * nx + 1].vx == 0.12f && info[(size_t)5 * nx + nx - 2].material == 5,
               "wall, ghost, inlet (vx 0.12) and outlet records");
    }

    Oracle o(nx, ny, 3.0f * setting.fluid_viscosity + 0.5f, node.lattice_info_data());
    o.reset(0);
    expect(same_state(node, o, false), "state after creation equals init.wgsl");

    sim.compute();
    o.frames(1);
    expect(same_state(node, o, true), "1 frame (2 steps) bit-exact");
    for (int k = 0; k < 24; ++k) sim.compute();
    o.frames(24);
    expect(same_state(node, o, true), "25 frames bit-exact");

    // A click adds a disk (centre cell (200, 65)): the band of rows goes up
    // from the host copy.
    sim.on_click(Vec2{400.0f, 130.0f});
    {
        const uint32_t y0 = 65 - 28, y1 = y0 + 56;
        for (size_t c = (size_t)y0 * nx; c < (size_t)y1 * nx; ++c)
            std::memcpy(&o.info[c], &node.lattice_info_data()[c], sizeof(ws_info));
    }
    for (int k = 0; k < 10; ++k) sim.compute();
    o.frames(10);
    expect(same_state(node, o, true), "obstacle click, 10 frames bit-exact");

    // A drag writes ExternalForce records (block_iter 90) along the path,
    // cells (95, 62) .. (135, 66), all fluid (a record written over a solid
    // cell would expose slots the reference's boundary pass leaves to the
    // order of its invocations).
    sim.touch_begin();
    sim.touch_move(Vec2{190.0f, 124.0f});
    sim.touch_move(Vec2{230.0f, 128.0f});
    sim.touch_move(Vec2{270.0f, 132.0f});
    {
        std::vector<LatticeInfo> dev((size_t)nx * ny);
        if (lbm_read_lattice_info(node.handle(), dev.data(), 0, dev.size()) != LBM_OK) return 2;
        size_t forced = 0;
        for (size_t c = 0; c < dev.size(); ++c)
            if (dev[c].material == 6 && dev[c].block_iter == 90) {
                ++forced;
                std::memcpy(&o.info[c], &dev[c], sizeof(ws_info));
            }
        expect(forced > 0, "drag wrote external-force cells");
    }
    for (int k = 0; k < 30; ++k) sim.compute();
    o.frames(30);
    expect(same_state(node, o, true), "drag, 30 frames (countdown running) bit-exact");
    for (int k = 0; k < 20; ++k) sim.compute();
    o.frames(20);
    {
        size_t left = 0;
        for (const ws_info& r : o.info) left += (r.material == 6);
        expect(left == 0, "every external-force cell expired after 90 steps");
    }
    expect(same_state(node, o, true), "50 frames after the drag bit-exact");

    // A viscosity change rewrites tau only.
    setting.fluid_viscosity = 0.02f;
    sim.update_uniforms(setting);
    o.tau = 3.0f * setting.fluid_viscosity + 0.5f;
    for (int k = 0; k < 10; ++k) sim.compute();
    o.frames(10);
    expect(same_state(node, o, true), "new viscosity, 10 frames bit-exact");

    // Reset: the Poiseuille scene comes back (the clicked disk goes away).
    sim.reset();
    {
        const auto init = init_lattice_material(nx, ny, FieldAnimationType::Poiseuille);
        std::memcpy(o.info.data(), init.data(), init.size() * sizeof(ws_info));
    }
    o.reset(0);
    expect(same_state(node, o, false), "reset equals a fresh init.wgsl");

    // Errors surface as LbmError, as the reference would panic.
    bool threw = false;
    try {
        FluidSimulator tiny(2, 2, 2, setting);
    } catch (const LbmError& e) {
        threw = e.code() == LBM_ERR_INVALID && std::strstr(e.what(), "lattice") != nullptr;
    }
    expect(threw, "a 1x1 lattice throws LbmError(LBM_ERR_INVALID)");

    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
