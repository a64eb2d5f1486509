// Layout of the neighbour-list workspace (dpk_nlist_build), shared by the
// kernels that reuse its cell list (the cell-blocked force kernel).
#pragma once

#include <cstddef>

namespace dpk {

struct NlistWs {
  int* cell_count;  // [ncells]
  int* cell_start;  // [ncells + 1] exclusive prefix
  int* cell_atoms;  // [natoms] atoms sorted by cell
  int* atom_cell;   // [natoms] global cell id of each atom
  int* atom_slot;   // [natoms] position of each atom inside its cell
  int* cell_tmp;    // [natoms] binning scratch (arrival order inside each cell)
};

inline size_t nlist_align256(size_t b) { return (b + 255) & ~size_t(255); }

inline size_t nlist_ws_bytes(int natoms, int ncells) {
  return nlist_align256(sizeof(int) * (size_t)ncells) +
         nlist_align256(sizeof(int) * ((size_t)ncells + 1)) +
         4 * nlist_align256(sizeof(int) * (size_t)natoms);
}

inline NlistWs nlist_ws(void* workspace, int natoms, int ncells) {
  char* ws = static_cast<char*>(workspace);
  NlistWs w;
  w.cell_count = reinterpret_cast<int*>(ws);
  ws += nlist_align256(sizeof(int) * (size_t)ncells);
  w.cell_start = reinterpret_cast<int*>(ws);
  ws += nlist_align256(sizeof(int) * ((size_t)ncells + 1));
  w.cell_atoms = reinterpret_cast<int*>(ws);
  ws += nlist_align256(sizeof(int) * (size_t)natoms);
  w.atom_cell = reinterpret_cast<int*>(ws);
  ws += nlist_align256(sizeof(int) * (size_t)natoms);
  w.atom_slot = reinterpret_cast<int*>(ws);
  ws += nlist_align256(sizeof(int) * (size_t)natoms);
  w.cell_tmp = reinterpret_cast<int*>(ws);
  return w;
}

}  // namespace dpk
