// Runs the host layer's edit geometry (no device calls) on commands read
// from stdin, one per line, and prints one result line each:
//   obstacle nx ny x y              -> "band first count" then the obstacle cells
//   stroke nx ny pixel px py qx qy  -> "record material block_iter vx vy" then the cells
//   click nx ny pixel px py         -> "click x y" or "click none"
//   move ax ay bx by                -> "move draws nx ny"
// Floats are printed with 9 significant digits (exact for binary32).
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "d2q9_node.hpp"

using namespace simuverse;

int main()
{
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        in >> cmd;
        try {
            if (cmd == "obstacle") {
                uint32_t nx, ny, x, y;
                in >> nx >> ny >> x >> y;
                std::vector<LatticeInfo> info((size_t)nx * ny, LatticeInfo{1, -1, 0.0f, 0.0f});
                const CellRange band = obstacle_band(info, nx, ny, x, y);
                std::printf("band %u %u", band.first, band.count);
                for (size_t c = 0; c < info.size(); ++c)
                    if (info[c].material == (int32_t)LatticeType::Obstacle) std::printf(" %zu", c);
                std::printf("\n");
            } else if (cmd == "stroke") {
                uint32_t nx, ny, pixel;
                Vec2 p, q;
                in >> nx >> ny >> pixel >> p.x >> p.y >> q.x >> q.y;
                const ForceStroke s = force_stroke(nx, ny, pixel, p, q);
                std::printf("record %d %d %.9g %.9g", s.record.material, s.record.block_iter, s.record.vx,
                            s.record.vy);
                for (uint32_t c : s.cells) std::printf(" %u", c);
                std::printf("\n");
            } else if (cmd == "click") {
                uint32_t nx, ny, pixel, x = 0, y = 0;
                Vec2 p;
                in >> nx >> ny >> pixel >> p.x >> p.y;
                if (click_cell(nx, ny, pixel, p, &x, &y)) std::printf("click %u %u\n", x, y);
                else std::printf("click none\n");
            } else if (cmd == "move") {
                Vec2 a, b, next;
                in >> a.x >> a.y >> b.x >> b.y;
                const bool draws = touch_move_draws(a, b, &next);
                std::printf("move %d %.9g %.9g\n", draws ? 1 : 0, next.x, next.y);
            } else {
                std::printf("error unknown command\n");
            }
        } catch (const LbmError& e) {
            std::printf("error %d\n", e.code());
        }
        std::fflush(stdout);
    }
    return 0;
}
