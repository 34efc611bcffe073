// d2q9_node.cpp -- the reference's D2Q9Node / FluidSimulator host logic
// over the gfx950 library's C ABI.
#include "d2q9_node.hpp"

#include <cmath>

namespace simuverse {

namespace {

void check(int rc)
{
    if (rc != LBM_OK) throw LbmError(rc, std::string("lbm: ") + lbm_last_error());
}

float tau_of(const SettingObj& s) { return 3.0f * s.fluid_viscosity + 0.5f; }  // d2q9_node.rs:50

}  // namespace

bool is_sd_sphere(Vec2 p, float r) { return std::sqrt(p.x * p.x + p.y * p.y) <= r; }

std::vector<LatticeInfo> init_lattice_material(uint32_t nx, uint32_t ny, FieldAnimationType ty)
{
    std::vector<LatticeInfo> info;
    info.reserve((size_t)nx * ny);
    // Disk centres of the Poiseuille scene (lattice.rs:37-39), in binary32.
    const Vec2 s0{(float)nx / 7.0f - OBSTACLE_RADIUS, (float)ny / 2.0f};
    const Vec2 s1{(float)nx / 5.0f, (float)ny / 4.0f};
    const Vec2 s2{(float)nx / 5.0f, (float)ny * 0.75f};
    auto in_disk = [](Vec2 p, Vec2 c) { return is_sd_sphere(Vec2{p.x - c.x, p.y - c.y}, OBSTACLE_RADIUS); };
    for (uint32_t y = 0; y < ny; ++y) {
        for (uint32_t x = 0; x < nx; ++x) {
            int32_t material = (int32_t)LatticeType::Bulk;
            float vx = 0.0f;
            switch (ty) {
            case FieldAnimationType::Custom:
                if (x == 0 || x == nx - 1 || y == 0 || y == ny - 1) material = (int32_t)LatticeType::Boundary;
                break;
            case FieldAnimationType::LidDrivenCavity:
                if (x == 0 || x == nx - 1 || y == ny - 1) {
                    material = (int32_t)LatticeType::Boundary;
                } else if (y == 0) {
                    material = (int32_t)LatticeType::Ghost;
                } else if (y == 1) {
                    material = (int32_t)LatticeType::ExternalForce;
                    vx = 0.13f;
                }
                break;
            case FieldAnimationType::Poiseuille:
                if (y == 0 || y == ny - 1) {
                    material = (int32_t)LatticeType::Boundary;
                } else if (x == 0 || x == nx - 1) {
                    material = (int32_t)LatticeType::Ghost;
                } else if (x == 1) {
                    material = (int32_t)LatticeType::Inlet;
                    vx = 0.12f;
                } else if (x == nx - 2) {
                    material = (int32_t)LatticeType::Outlet;
                } else {
                    const Vec2 p{(float)x, (float)y};
                    if (in_disk(p, s0) || in_disk(p, s1) || in_disk(p, s2)) material = (int32_t)LatticeType::Obstacle;
                }
                break;
            default:
                break;
            }
            info.push_back(LatticeInfo{material, -1, vx, 0.0f});
        }
    }
    return info;
}

// ---------------------------------------------------------------------------
// Edit geometry
// ---------------------------------------------------------------------------

CellRange obstacle_band(std::vector<LatticeInfo>& info, uint32_t nx, uint32_t ny, uint32_t x, uint32_t y)
{
    const uint32_t r = (uint32_t)OBSTACLE_RADIUS;
    if (y < r || y - r + 2 * r > ny || info.size() != (size_t)nx * ny)
        throw LbmError(LBM_ERR_INVALID, "add_obstacle: the disk's row band leaves the lattice");
    const LatticeInfo obstacle{(int32_t)LatticeType::Obstacle, -1, 0.0f, 0.0f};
    const Vec2 center{(float)x + 0.5f, (float)y + 0.5f};
    const uint32_t min_y = y - r, max_y = min_y + 2 * r;
    for (uint32_t yy = min_y; yy < max_y; ++yy)
        for (uint32_t xx = 0; xx < nx; ++xx)
            if (is_sd_sphere(Vec2{(float)xx + 0.5f - center.x, (float)yy + 0.5f - center.y}, OBSTACLE_RADIUS))
                info[(size_t)yy * nx + xx] = obstacle;
    // The whole band goes up: cells outside the disk carry the host copy.
    return CellRange{nx * min_y, nx * (max_y - min_y)};
}

ForceStroke force_stroke(uint32_t nx, uint32_t ny, uint32_t pixel, Vec2 pos, Vec2 pre_pos)
{
    if (pixel < 2) throw LbmError(LBM_ERR_INVALID, "add_external_force: lattice_pixel_size must be >= 2");
    const float dx = pos.x - pre_pos.x, dy = pos.y - pre_pos.y;
    const float dis = std::sqrt(dx * dx + dy * dy);
    float force = 0.1f * (dis / 20.0f);
    if (force > 0.12f) force = 0.12f;
    const float ridian = std::atan2(dy, dx);
    ForceStroke out;
    out.record = LatticeInfo{(int32_t)LatticeType::ExternalForce, 90, force * std::cos(ridian),
                             force * std::sin(ridian)};
    const float c = std::ceil(dis / (float)(pixel - 1));
    const float step = dis / c;
    for (int i = 0; i < (int)c; ++i) {
        const float px = std::round(pre_pos.x + step * (float)i * std::cos(ridian));
        const float py = std::round(pre_pos.y + step * (float)i * std::sin(ridian));
        // `as u32` saturates: negatives and NaN to 0
        const uint32_t x = (px > 0.0f ? (px < 4294967296.0f ? (uint32_t)px : 0xFFFFFFFFu) : 0u) / pixel;
        const uint32_t y = (py > 0.0f ? (py < 4294967296.0f ? (uint32_t)py : 0xFFFFFFFFu) : 0u) / pixel;
        if (x < 1 || x >= nx - 2 || y < 1 || y >= ny - 2) continue;
        out.cells.push_back(nx * y + x);
    }
    return out;
}

bool click_cell(uint32_t nx, uint32_t ny, uint32_t pixel, Vec2 pos, uint32_t* x, uint32_t* y)
{
    if (pos.x <= 0.0f || pos.y <= 0.0f) return false;
    const uint32_t cx = (uint32_t)pos.x / pixel;
    const uint32_t cy = (uint32_t)pos.y / pixel;
    const uint32_t half = (uint32_t)OBSTACLE_RADIUS;
    if (cx < half || cx >= nx - (half + 2) || cy < half || cy >= ny - (half + 2)) return false;
    *x = cx;
    *y = cy;
    return true;
}

bool touch_move_draws(Vec2 pre_pos, Vec2 pos, Vec2* next)
{
    if (pos.x <= 0.0f || pos.y <= 0.0f) {
        *next = Vec2{};
        return false;
    }
    const float dx = pos.x - pre_pos.x, dy = pos.y - pre_pos.y;
    const float dis = std::sqrt(dx * dx + dy * dy);
    *next = pos;
    return !((pre_pos.x == 0.0f && pre_pos.y == 0.0f) || dis > 300.0f);
}

// ---------------------------------------------------------------------------
// D2Q9Node
// ---------------------------------------------------------------------------

D2Q9Node::D2Q9Node(uint32_t canvas_width, uint32_t canvas_height, uint32_t lattice_pixel_size,
                   const SettingObj& setting)
{
    if (lattice_pixel_size < 1) throw LbmError(LBM_ERR_INVALID, "lattice_pixel_size must be >= 1");
    pixel_ = lattice_pixel_size;
    nx_ = canvas_width / pixel_;
    ny_ = canvas_height / pixel_;
    ty_ = setting.animation_type;
    lbm_desc d{};
    d.nx = (int32_t)nx_;
    d.ny = (int32_t)ny_;
    d.tau = tau_of(setting);
    d.mode = LBM_MODE_PARITY_WGSL;
    d.storage = LBM_STORAGE_F32;
    d.fluid_ty = ty_ == FieldAnimationType::LidDrivenCavity ? 1 : 0;  // d2q9_node.rs:53-57
    check(lbm_create(&d, &h_));
    try {
        info_ = init_lattice_material(nx_, ny_, ty_);  // :106
        write_info_rows(0, nx_ * ny_);
        reset_lattice_info();  // :203
    } catch (...) {
        lbm_destroy(h_);
        h_ = nullptr;
        throw;
    }
}

D2Q9Node::~D2Q9Node()
{
    if (h_) lbm_destroy(h_);
}

void D2Q9Node::write_info_rows(uint32_t first_cell, uint32_t count)
{
    check(lbm_write_lattice_info(h_, info_.data() + first_cell, first_cell, count));
}

void D2Q9Node::reset() { check(lbm_reset(h_)); }

void D2Q9Node::reset_lattice_info()
{
    if (ty_ == FieldAnimationType::Poiseuille) {
        info_ = init_lattice_material(nx_, ny_, ty_);
        write_info_rows(0, nx_ * ny_);
    }
    reset();
}

void D2Q9Node::add_obstacle(uint32_t x, uint32_t y)
{
    const CellRange band = obstacle_band(info_, nx_, ny_, x, y);
    write_info_rows(band.first, band.count);
}

void D2Q9Node::add_external_force(Vec2 pos, Vec2 pre_pos)
{
    const ForceStroke stroke = force_stroke(nx_, ny_, pixel_, pos, pre_pos);
    for (uint32_t cell : stroke.cells) check(lbm_write_lattice_info(h_, &stroke.record, cell, 1));
}

void D2Q9Node::compute_by_pass(int swap_index) { check(lbm_step(h_, swap_index, 1)); }

void D2Q9Node::update_uniforms(const SettingObj& setting) { check(lbm_set_tau(h_, tau_of(setting))); }

void D2Q9Node::read_macro(float* rho, float* ux, float* uy) const { check(lbm_read_macro(h_, rho, ux, uy)); }

void D2Q9Node::read_populations(int buf, float* f) const { check(lbm_read_populations(h_, buf, f)); }

void D2Q9Node::write_populations(int buf, const float* f) { check(lbm_write_populations(h_, buf, f)); }

// ---------------------------------------------------------------------------
// FluidSimulator
// ---------------------------------------------------------------------------

FluidSimulator::FluidSimulator(uint32_t canvas_width, uint32_t canvas_height, uint32_t lattice_pixel_size,
                               const SettingObj& setting)
    : node_(canvas_width, canvas_height, lattice_pixel_size, setting)
{
}

void FluidSimulator::on_click(Vec2 pos)
{
    uint32_t x = 0, y = 0;
    if (click_cell(node_.width(), node_.height(), node_.lattice_pixel_size(), pos, &x, &y)) node_.add_obstacle(x, y);
}

void FluidSimulator::touch_begin() { pre_pos_ = Vec2{}; }

void FluidSimulator::touch_move(Vec2 pos)
{
    Vec2 next;
    if (touch_move_draws(pre_pos_, pos, &next)) node_.add_external_force(pos, pre_pos_);
    pre_pos_ = next;
}

void FluidSimulator::update_uniforms(const SettingObj& setting) { node_.update_uniforms(setting); }

void FluidSimulator::reset()
{
    node_.reset_lattice_info();
    pre_pos_ = Vec2{};
}

void FluidSimulator::compute()
{
    node_.compute_by_pass(0);
    node_.compute_by_pass(1);
}

}  // namespace simuverse
