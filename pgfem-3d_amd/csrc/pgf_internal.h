// Internal host-side helpers shared by the set-up code and the device handle.
#pragma once
#include <cstdint>
#include <vector>

#include "pgfem3d_amd.h"

namespace pgf {

// Node -> incident element lists of a mesh (volume and cohesive elements together).
struct Incidence {
  std::vector<int64_t> ptr;  // [n_nodes+1]
  std::vector<int64_t> elems;  // element ids; cohesive element c is stored as n_elem + c
};

Incidence build_incidence(const pgf_mesh &m);

// Greedy colouring of the volume elements: two elements sharing a node never share a
// colour.  Elements are visited in mesh order, each taking the lowest free colour.
// Returns the number of colours (or -1 if more than 256 are needed).
int color_elements(const pgf_mesh &m, std::vector<int32_t> &color);

// Per-element, per node-pair column offsets into the CSR rows (see pgf_create).
// off[e_first_pair + a*nne + w] = position, inside the row of node a's first free DOF, of
// node w's first free DOF column; -1 when either node has no free DOF.
int build_scatter_offsets(const pgf_mesh &m, const int64_t *id, int64_t ndof,
                          const int32_t *Ap, const int32_t *Ai,
                          std::vector<int64_t> &pair_ptr, std::vector<int16_t> &off);

// The same for the interface elements (pair_ptr over coel_ptr).
int build_coel_offsets(const pgf_mesh &m, const int64_t *id, const int32_t *Ap, const int32_t *Ai,
                       std::vector<int64_t> &pair_ptr, std::vector<int16_t> &off);

// Greedy colouring of the interface elements among themselves (they are launched after all
// volume colours, so they never race with volume elements).
int color_coels(const pgf_mesh &m, std::vector<int32_t> &color);

}  // namespace pgf
