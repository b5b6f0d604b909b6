// Mesh geometry and least-squares set-up on the GPU (FP64): the same numbers
// as compute_geometry and compute_ls in mesh_host.cpp, bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mesh_host.hpp"

namespace lbfs {

// Fills g, lhs and rhs as compute_geometry followed by compute_ls do, with
// the face and cell passes on the GPU (device buffers allocated and freed
// here, work on stream s, synchronised before returning).  Error codes and
// messages as the host passes give them.
int gpu_geometry(const lbfs_mesh_desc &m, double visc, double gamma, double fneq_min, bool closed, int grad_type,
                 bool det_compat, hipStream_t s, Geometry &g, std::vector<double> &lhs, std::vector<double> &rhs,
                 std::string &err);

}  // namespace lbfs
