// Host-side mesh preprocessing for the HIP step: neighbour codes, the
// internal cell order and the tile adjacency used to skip dry tiles.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/cuteflow_hip.h"

namespace cfprep {

// Per-cell neighbour record, in the internal (device) cell order:
//   c[k] >= 0  neighbour across face k (internal index)
//   c[k] <  0  -(class | id << 2), class 1 inlet, 2 outlet, 3 wall
//   w          bits 2k..2k+1: face of the neighbour that shares face k's edge
struct Nbr {
    int c[3];
    int w;
};

struct Prep {
    int nelt = 0;
    int tile = 0;
    std::vector<int> perm;          // internal index -> caller's index
    std::vector<int> inv;           // caller's index -> internal index
    std::vector<Nbr> nbr;           // internal order
    std::vector<int> conn;          // 3*nelt node ids (1-based), internal order
    std::vector<int> tile_off;      // CSR offsets, ntiles + 1
    std::vector<int> tile_adj;      // tiles holding a neighbour of the tile's cells
    std::vector<unsigned char> tile_always;  // tiles that always run (inflow faces)
    std::vector<int> io_cells;      // internal indices of cells with an inflow or outflow face, caller order
};

// Builds the internal layout.  keep_order: use the caller's cell order;
// otherwise cells are sorted along a Hilbert curve of their centroids so that
// a tile of `tile` consecutive cells is a compact patch of the mesh.
int prepare(const cf_mesh *mesh, const cf_params *params, int tile, bool keep_order, Prep &out);

// The same on the current GPU (sorts, per-cell kernels), identical results;
// CF_ERR_HIP / CF_ERR_NOMEM on device failures.
int prepare_device(const cf_mesh *mesh, const cf_params *params, int tile, bool keep_order, Prep &out);

}  // namespace cfprep
