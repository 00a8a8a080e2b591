// Device-side addressing of the ComputeUnit arrays.
//
// Cell index n = i*(yg*zg) + j*zg + k over the whole array including ghosts
// (tdlbcpp/src/ComputeUnit.hpp:454-463): F is an array of {x,y,z} per cell, Nu
// one value per cell, O one byte per cell, all indexed by n.
//
// The population arrays have their own cell index qn = qcell(i,j,k):
//   IJKL: Q[qn*18 + l], qn = n                          (tdlbcpp/src/Field.hpp:18-144)
//   LIJK: Q[l*qsize + qn], qn = i*qyz + j*qz + (k-1) for interior k: k-rows of
//         the zg1 interior cells back to back, the k-ghost cells of each row in
//         a side region (qside), so the lines the stream-collide kernels move
//         hold interior cells only.
// At the boundary (download, upload, checkpoint) Q is the reference's dense array.
//
// The F-occupancy bitmask has one bit per cell, rows of k padded to 64-bit
// words: word (i*yg + j)*mask_w + k/64, bit k%64.  A set bit means F of that
// cell may be non-zero; a clear bit means F is zero whatever the array holds.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "lattice.h"

namespace tdlb {

struct Grid {
    int32_t xg, yg, zg;     // sizes including ghosts
    int32_t xg1, yg1, zg1;  // last interior index (ghost = 1)
    uint32_t yz;            // yg*zg
    uint32_t mask_w;        // 64-bit words per k-row of the bitmask
    uint64_t size;          // xg*yg*zg
    uint32_t qz, qyz;       // population arrays: row pitch and i-plane pitch in cells
    uint32_t qoff;          // population index of cell (0,0,0) (modulo 2^32)
    uint32_t qside;         // first cell of the k-ghost side region, 0: none
    uint64_t qsize;         // cells per slot plane (LIJK)
};

// slot l of the population cell qn
template <int L>
__device__ __forceinline__ uint64_t qaddr(const Grid& g, uint32_t qn, int l) {
    return L == LayoutIJKL ? (uint64_t)qn * kQ + (uint64_t)l : (uint64_t)l * g.qsize + (uint64_t)qn;
}

__host__ __device__ __forceinline__ uint32_t cell_index(const Grid& g, int i, int j, int k) {
    return (uint32_t)i * g.yz + (uint32_t)j * (uint32_t)g.zg + (uint32_t)k;
}

// (i,j,k) of cell index n
__device__ __forceinline__ void decode_cell(const Grid& g, uint64_t n, int& i, int& j, int& k) {
    i = (int)(n / g.yz);
    const uint32_t r = (uint32_t)(n - (uint64_t)i * g.yz);
    j = (int)(r / (uint32_t)g.zg);
    k = (int)(r - (uint32_t)j * (uint32_t)g.zg);
}

// population cell of (i,j,k); with a side region the rows hold the interior
// k = 1..zg1 only and the two k-ghost cells of row (i,j) sit at qside + 2(i*yg + j)
__host__ __device__ __forceinline__ uint32_t qcell(const Grid& g, int i, int j, int k) {
    if (g.qside && (k == 0 || k > g.zg1))
        return g.qside + 2u * ((uint32_t)i * (uint32_t)g.yg + (uint32_t)j) + (k != 0 ? 1u : 0u);
    return g.qoff + (uint32_t)i * g.qyz + (uint32_t)j * g.qz + (uint32_t)k;
}

__device__ __forceinline__ uint64_t mask_word_index(const Grid& g, int i, int j, int k) {
    return ((uint64_t)i * (uint64_t)g.yg + (uint64_t)j) * g.mask_w + (uint64_t)(k >> 6);
}

__device__ __forceinline__ bool mask_test(const uint64_t* __restrict__ m, const Grid& g, int i, int j, int k) {
    return (m[mask_word_index(g, i, j, k)] >> (k & 63)) & 1ull;
}

__device__ __forceinline__ bool is_interior(const Grid& g, int i, int j, int k) {
    return i >= 1 && i <= g.xg1 && j >= 1 && j <= g.yg1 && k >= 1 && k <= g.zg1;
}

// Decode a flat block id into (first k of the block, j, i) over interior rows.
// Blocks are laid out k-chunk fastest, then j, then i.
struct RowBlock {
    int i, j, k0;
};
__device__ __forceinline__ RowBlock row_block(const Grid& g, uint32_t b, uint32_t nkb, uint32_t bk) {
    RowBlock r;
    uint32_t kb = b % nkb;
    uint32_t rest = b / nkb;
    r.j = (int)(rest % (uint32_t)g.yg1) + 1;
    r.i = (int)(rest / (uint32_t)g.yg1) + 1;
    r.k0 = (int)(kb * bk) + 1;
    return r;
}

}  // namespace tdlb
