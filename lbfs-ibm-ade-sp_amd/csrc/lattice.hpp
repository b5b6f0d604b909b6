// Discrete velocity sets for the LBFS face reconstruction.  D3Q15 keeps the
// reference order and weights (Solver_gpu.cu:1407-1445, Solver.cpp:1398-1445);
// D3Q19 and D2Q9 are the same construction with their own tables.
#pragma once

namespace lbfs {

struct D3Q15 {
    static constexpr int Q = 15;
    static constexpr int id = 0;
    static constexpr int dims = 3;
    static constexpr int ex[15] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    static constexpr int ey[15] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1};
    static constexpr int ez[15] = {0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1};
    static constexpr float w(int k) { return k == 0 ? 2.0f / 9.0f : (k < 7 ? 1.0f / 9.0f : 1.0f / 72.0f); }
};

struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr int id = 1;
    static constexpr int dims = 3;
    static constexpr int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr float w(int k) { return k == 0 ? 1.0f / 3.0f : (k < 7 ? 1.0f / 18.0f : 1.0f / 36.0f); }
};

struct D2Q9 {
    static constexpr int Q = 9;
    static constexpr int id = 2;
    static constexpr int dims = 2;
    static constexpr int ex[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    static constexpr int ey[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    static constexpr int ez[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr float w(int k) { return k == 0 ? 4.0f / 9.0f : (k < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }
};

}  // namespace lbfs
