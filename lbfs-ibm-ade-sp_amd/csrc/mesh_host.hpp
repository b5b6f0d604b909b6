// Host-side mesh set-up: generated hex boxes, FP64 geometry, least-squares
// weights and the tile-blocked device layout.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lbfs_abi.h"

namespace lbfs {

// Counter-based generator: SplitMix64 (Steele, Lea, Flood 2014).  Portable
// across compilers and libraries, unlike std::uniform_real_distribution.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct HexCounts {
    int n_points, n_faces, n_internal;
};
HexCounts hex_counts(int nx, int ny, int nz, const int periodic[3]);
void hex_generate(int nx, int ny, int nz, const double L[3], const double origin[3], const int periodic[3],
                  double jitter, uint64_t seed, double *points, int *fv, int *owner, int *neighbour,
                  double *shift, int *patch);

// FP64 geometry of a polyMesh-form hex mesh.
struct Geometry {
    int n_cells = 0, n_faces = 0, n_int = 0, n_bc = 0;
    std::vector<int> owner, nbr;          // nbr of boundary face f = n_cells + f - n_int
    std::vector<double> shift;            // 3 per face
    std::vector<double> area, normal, fc, fdt;
    std::vector<double> cc;               // 3 per cell and ghost
    std::vector<double> vol;
    std::vector<int> gcell, gface, grole; // 6 per cell
};
int compute_geometry(const lbfs_mesh_desc &m, double visc, double gamma, double fneq_min, bool closed, Geometry &g,
                     std::string &err);

// LS set-up: stored inverse (9 per cell) and RHS weights (18 per cell).
int compute_ls(const Geometry &g, int grad_type, bool det_compat, std::vector<double> &lhs,
               std::vector<double> &rhs, std::string &err);

// Logical boxes of the mesh: one for a generated box, several for an
// ensemble; empty when the mesh has no logical numbering.
struct HostBlock {
    int n[3];
    double origin[3];
    double spacing;
    int cell0;
    double force[3];  // body force of the box
};
int mesh_blocks(const lbfs_mesh_desc &m, const double body_force[3], std::vector<HostBlock> &blocks,
                std::string &err);
// body force acting on caller cell c: its box's, or `dflt` outside every box
const double *cell_body_force(const std::vector<HostBlock> &blocks, int c, const double *dflt);

// Tile-blocked device layout (all indices are internal, tile-ordered ids).
struct DeviceLayout {
    int n_cells = 0, n_bc = 0, n_tiles = 0, max_tile_faces = 0;
    int tile[3] = {0, 0, 0};      // cells per tile along x, y, z of the logical boxes
    std::vector<int> perm;        // internal id -> caller id
    std::vector<int> inv;         // caller id -> internal id
    std::vector<int> tile_cell_start, tile_face_start;
    std::vector<int> tile_bface_start;  // per tile: first boundary face (interior faces come first)
    // per cell (internal order)
    std::vector<int> stencil;     // 6 * n, SoA [k*n + i]; ghost entries encoded -1 - bc
    std::vector<float> gw;        // 18 * n, SoA [(k*3+c)*n + i], W_k = LHS^-1 RHS_k
    std::vector<float> inv_vol;   // n
    std::vector<float> static_force; // 3 * n: body force + pressure closure defect of the cell
    std::vector<double> closure;  // 3 * n, caller order: sum of n A over the faces that carry flux
    std::vector<int16_t> cell_slots; // 6 * n, AoS [i*6+k]: tile-local face slot, sign in bit 15 (1 = owner)
    // per tile-face (duplicated across tiles)
    // 48 B per face: the area-weighted normal (area = its length), the
    // centroid offsets, and the owner and neighbour ids as int bits
    std::vector<float> fA;        // 4 per face: normal * area xyz, face dt
    std::vector<float> fB;        // 4 per face: owner centroid - face centroid, owner id
    std::vector<float> fC;        // 4 per face: neighbour image centroid - face centroid, neighbour id
                                  // (ghost: n_cells + bc)
    // fused fluid pass: per tile the cells across its border that its faces
    // reach (their gradients are computed into the tile's LDS after its own),
    // and per tile-face the LDS gradient slots of owner and neighbour
    // (tile cells first, then that halo; 0xffff for a ghost)
    std::vector<int> tile_halo_start, halo_ids;
    std::vector<uint32_t> face_gslot;
    int max_grad_slots = 0;
    std::vector<int> face_ref;    // caller face id of each tile-face
    std::vector<int> face_first;  // caller face id -> one tile-face holding it
    // boundary faces
    std::vector<int> bc_owner;    // internal owner id per bc
};
int build_layout(const lbfs_mesh_desc &m, const Geometry &g, const std::vector<double> &lhs,
                 const std::vector<double> &rhs, double rho0, const double body_force[3], int dims, const int tile[3],
                 const std::vector<HostBlock> &blocks, DeviceLayout &d, std::string &err);

}  // namespace lbfs
