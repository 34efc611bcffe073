// d2q9_node.hpp -- C++ host side of the D2Q9 path over the C ABI (lbm.h),
// with the reference's names and argument meanings:
//
//   simuverse/src/fluid/lattice.rs:5-98      LatticeInfo, LatticeType,
//                                            init_lattice_material
//   simuverse/src/fluid/mod.rs:1-59          OBSTACLE_RADIUS, is_sd_sphere
//   simuverse/src/fluid/d2q9_node.rs:31-312  D2Q9Node (new, reset,
//                                            reset_lattice_info, add_obstacle,
//                                            add_external_force,
//                                            compute_by_pass)
//   simuverse/src/fluid/fluid_simulator.rs   FluidSimulator (on_click,
//                                            touch_begin, touch_move,
//                                            update_uniforms, reset, compute)
//
// The reference panics on failure; here every failure throws LbmError with
// the library's message.  Rendering (macro texture, particles, curl) is not
// part of this path: macro fields are read back with read_macro.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm.h"

namespace simuverse {

class LbmError : public std::runtime_error {
public:
    LbmError(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// lattice.rs:15-24
enum class LatticeType : int32_t {
    Bulk = 1,
    Boundary = 2,
    Inlet = 3,
    Obstacle = 4,
    Outlet = 5,
    ExternalForce = 6,
    Ghost = 7,
};

// lib.rs:119-127 (only Poiseuille and LidDrivenCavity change the fluid).
enum class FieldAnimationType : uint32_t {
    Basic = 0,
    JuliaSet,
    Spirl,
    BlackHole,
    Poiseuille,
    LidDrivenCavity,
    Custom,
};

using LatticeInfo = lbm_lattice_info;  // lattice.rs:5-11, 16 bytes

struct Vec2 {
    float x = 0.0f, y = 0.0f;
};

constexpr float OBSTACLE_RADIUS = 28.0f;  // mod.rs:1

// mod.rs:57-59: |p| <= r.
bool is_sd_sphere(Vec2 p, float r);

// lattice.rs:26-98: materials of an nx x ny lattice for an animation type.
std::vector<LatticeInfo> init_lattice_material(uint32_t nx, uint32_t ny, FieldAnimationType ty);

// ---- edit geometry (host only, no device calls) ---------------------------

// Cells [first, first + count) of a row-major lattice.
struct CellRange {
    uint32_t first = 0, count = 0;
};

// d2q9_node.rs:215-245: rows y - r .. y + r - 1 (r = OBSTACLE_RADIUS) of the
// host copy `info` (nx*ny records) get an Obstacle record where
// |(xx + 0.5, yy + 0.5) - (x + 0.5, y + 0.5)| <= r; returns the band the
// reference uploads (whole rows, cells outside the disk with their host
// value).  The band must lie inside the lattice (y >= r, y + r <= ny).
CellRange obstacle_band(std::vector<LatticeInfo>& info, uint32_t nx, uint32_t ny, uint32_t x, uint32_t y);

// d2q9_node.rs:263-300: the record a drag from pre_pos to pos (canvas
// pixels) writes (ExternalForce, block_iter 90, force 0.1 |d| / 20 capped at
// 0.12 along the drag) and the cells it writes it to, in order: ceil(|d| /
// (pixel - 1)) points spaced |d| / that count from pre_pos, rounded half away
// from zero, converted as Rust's saturating `as u32`, divided by pixel,
// skipped outside 1 <= x < nx - 2, 1 <= y < ny - 2.  pixel >= 2.
struct ForceStroke {
    LatticeInfo record{};
    std::vector<uint32_t> cells;
};
ForceStroke force_stroke(uint32_t nx, uint32_t ny, uint32_t pixel, Vec2 pos, Vec2 pre_pos);

// fluid_simulator.rs:137-152: the lattice cell a click at canvas pos adds an
// obstacle at, or false when the click is outside the allowed band.
bool click_cell(uint32_t nx, uint32_t ny, uint32_t pixel, Vec2 pos, uint32_t* x, uint32_t* y);

// fluid_simulator.rs:158-173: given the previous drag position, whether a
// move to pos draws a stroke (and so calls add_external_force); *next is
// the new previous position.
bool touch_move_draws(Vec2 pre_pos, Vec2 pos, Vec2* next);

// The settings D2Q9Node::new reads (d2q9_node.rs:31-59).
struct SettingObj {
    float fluid_viscosity = 0.02f;
    FieldAnimationType animation_type = FieldAnimationType::Poiseuille;
};

// One LBM domain in the reference's parity mode (bit-exact reference
// arithmetic), owned like the reference's D2Q9Node owns its buffers.
class D2Q9Node {
public:
    // canvas_size / lattice_pixel_size cells (d2q9_node.rs:38-45);
    // tau = 3 * fluid_viscosity + 0.5 (:50); fluid_ty 1 for a lid-driven
    // cavity, else 0 (:53-57).  Materials come from init_lattice_material and
    // the populations from the reset kernel (:203-213, :247-261).
    D2Q9Node(uint32_t canvas_width, uint32_t canvas_height, uint32_t lattice_pixel_size, const SettingObj& setting);
    ~D2Q9Node();
    D2Q9Node(const D2Q9Node&) = delete;
    D2Q9Node& operator=(const D2Q9Node&) = delete;

    uint32_t width() const { return nx_; }
    uint32_t height() const { return ny_; }
    uint32_t lattice_pixel_size() const { return pixel_; }
    const std::vector<LatticeInfo>& lattice_info_data() const { return info_; }
    lbm_handle* handle() const { return h_; }

    // d2q9_node.rs:211-213: the init shader on both population buffers.
    void reset();
    // d2q9_node.rs:247-261: Poiseuille materials rewritten, then reset.
    void reset_lattice_info();
    // d2q9_node.rs:215-245: a disk of OBSTACLE_RADIUS centred at
    // (x + 0.5, y + 0.5), written row band by row band.
    void add_obstacle(uint32_t x, uint32_t y);
    // d2q9_node.rs:263-300: ExternalForce cells (block_iter 90) along the
    // segment pre_pos -> pos (canvas pixels).
    void add_external_force(Vec2 pos, Vec2 pre_pos);
    // d2q9_node.rs:302-312: one step from buffer swap_index into the other.
    void compute_by_pass(int swap_index);
    // fluid_simulator.rs:175-193: a new tau = 3 * fluid_viscosity + 0.5.
    void update_uniforms(const SettingObj& setting);

    // Macro fields (rho, ux, uy per cell) as the collide pass computed them
    // in its last step; the caller sizes each array nx*ny.
    void read_macro(float* rho, float* ux, float* uy) const;
    // Populations of buffer `buf` (9*nx*ny floats, f[i*nx*ny + y*nx + x]).
    void read_populations(int buf, float* f) const;
    void write_populations(int buf, const float* f);

private:
    void write_info_rows(uint32_t first_cell, uint32_t count);

    lbm_handle* h_ = nullptr;
    uint32_t nx_ = 0, ny_ = 0, pixel_ = 1;
    FieldAnimationType ty_ = FieldAnimationType::Poiseuille;
    std::vector<LatticeInfo> info_;
};

// The LBM part of the reference's FluidSimulator: user edits and two steps
// per frame (fluid_simulator.rs:217-232; particle passes are not part of
// this path).
class FluidSimulator {
public:
    FluidSimulator(uint32_t canvas_width, uint32_t canvas_height, uint32_t lattice_pixel_size,
                   const SettingObj& setting);

    D2Q9Node& fluid_compute_node() { return node_; }

    void on_click(Vec2 pos);    // :137-152
    void touch_begin();         // :154-156
    void touch_move(Vec2 pos);  // :158-173
    void update_uniforms(const SettingObj& setting);
    void reset();               // :205-210
    void compute();             // :217-232: compute_by_pass(0), then (1)

private:
    D2Q9Node node_;
    Vec2 pre_pos_{};
};

}  // namespace simuverse
