// Host-side membrane set-up (FP64): geodesic spheres, the Evans-Fung
// biconcave shape, spring connectivity and the reference state of every edge.
#pragma once

#include <string>
#include <vector>

#include "lbfs_abi.h"

namespace lbfs {

// frequency-n geodesic sphere (Class I icosahedral subdivision): 10 n^2 + 2
// nodes, 20 n^2 triangles, outward winding
void icosphere(int frequency, std::vector<double> &nodes, std::vector<int> &tri);

// Evans-Fung shape of diameter d0 (Fedosov 2010 form), symmetry axis z
void evans_fung(const std::vector<double> &unit, double d0, std::vector<double> &out);

// springs (a, b, c, d) with edge a-c shared by triangles (a,b,c) and (a,c,d),
// in the order lagrangian_object.cpp:322-490 produces them
int build_springs(int n_nodes, const std::vector<int> &tri, std::vector<int> &springs, std::string &err);

// Reference state of one membrane from get_contour_lengths
// (spring_network.cu:516-604), get_reference_curvatures (389-513) and the
// initial force call of Solver_gpu.cu:726-732 (initials = true).
struct MembraneSetup {
    int n_nodes = 0, n_tri = 0, n_edges = 0;
    std::vector<int> tri, springs;
    std::vector<double> X;                 // initial positions
    std::vector<double> contour_length, spring_constant, curvature_0;
    std::vector<double> pow_ratio;         // WLC-POW: k_p / k_s per edge (0 with WLC)
    std::vector<double> spring_area_0;     // 2 per edge
    std::vector<double> curvature, spring_area;  // state after the initial call (2 per edge for areas)
    double V0 = 0, A0 = 0, MAD0 = 0, V = 0, A = 0, MAD = 0;
};
int membrane_setup(const lbfs_membrane_desc &m, MembraneSetup &s, std::string &err);

}  // namespace lbfs
