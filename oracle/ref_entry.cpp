// Entry points into the reference's own C++ sources, compiled from where they
// lie in the reference tree (initialize.h, core/force_lj.cpp) into
// oracle/_ref/libtinymd_ref.so.  Used only by tests to pin the oracle.
// Nothing here restates or replaces reference code: the two wrappers below
// only adapt the reference's std::tuple / std::function interface to C.
// initialize.h relies on its includer for <functional> and <iostream>, as
// main.cpp:1-7 provides them.
#include <functional>
#include <iostream>
#include "initialize.h"

extern "C" __attribute__((visibility("default")))
double ref_random(int* seed) { return random(seed); }

extern "C" __attribute__((visibility("default")))
int ref_generate_rectangular_grid(double* aabb, double* rank_aabb, int half, double spacing,
                                  double* mass, double* pos, double* vel, int cap) {
    double velocity[3] = {0.0, 0.0, 0.0};
    auto within = [rank_aabb](double x, double y, double z) { return is_within_aabb(x, y, z, rank_aabb); };
    auto t = generate_rectangular_grid(aabb, half != 0, spacing, 1.0, velocity, false, within);
    int n = (int)std::get<0>(t).size();
    for (int i = 0; i < n && i < cap; ++i) {
        mass[i] = std::get<0>(t)[i];
        pos[3 * i + 0] = std::get<1>(t)[i].x; pos[3 * i + 1] = std::get<1>(t)[i].y; pos[3 * i + 2] = std::get<1>(t)[i].z;
        vel[3 * i + 0] = std::get<2>(t)[i].x; vel[3 * i + 1] = std::get<2>(t)[i].y; vel[3 * i + 2] = std::get<2>(t)[i].z;
    }
    return n;
}
