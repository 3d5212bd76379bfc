// Shared host helpers of libfieldgrid: the thread-local error text behind fg_last_error().
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fieldgrid_sim.h"

namespace fg {

// Records `msg` as this thread's last error and returns `status` (the reference returns
// Result<_, String>; the C-ABI returns a status and keeps the String here).
int set_error(int status, const std::string &msg);

}  // namespace fg

namespace fg {

// One E edge whose material average (sim.rs:375-410) reads a cell of a design box: the edge's
// eps is EPS0 * (fixed + sum_q w[q] * eps_r[cell[q]]) over the box cells it touches.
struct DesignEdge {
    uint64_t f;        // c*nn + id
    double sig;        // averaged conductivity (S/m)
    double fixed;      // weighted eps_r of the edge's cells outside the box (weights normalised)
    uint32_t cell[4];  // box cell (k fastest) of each averaged cell, UINT32_MAX outside the box
    double w[4];       // normalised area weights
};

// The design edges of box [lo, hi) (cells) in `sim`, built from `desc`; edges eps does not reach
// (PEC and outer faces: cb = 0; Debye edges) are left out. Returns "" or the reason the box is
// refused: it touches a port, lumped or resistive edge, or reaches into the CPML.
std::string design_map(const fg_sim &sim, const fg_sim_desc &desc, const uint32_t lo[3], const uint32_t hi[3],
                       std::vector<DesignEdge> &out);

}  // namespace fg
