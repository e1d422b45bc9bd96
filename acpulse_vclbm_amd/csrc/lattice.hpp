// D3Q19 velocity set in the reference's direction order (Control.cpp:58-123):
//   0:+x 1:+y 2:-x 3:-y 4:+z 5:-z 6:(+x,+z) 7:(+x,-z) 8:(+y,+z) 9:(+y,-z)
//   10:(-x,+z) 11:(-x,-z) 12:(-y,+z) 13:(-y,-z) 14:(-x,+y) 15:(+x,-y)
//   16:(+x,+y) 17:(-x,-y) 18:rest
// and weights 1/18 (axis), 1/36 (diagonal), 1/3 (rest) (Control.cpp:126-131).
#pragma once

#include <cstdint>

namespace acp {

constexpr int Q = 19;

constexpr int EX[Q] = {1, 0, -1, 0, 0, 0, 1, 1, 0, 0, -1, -1, 0, 0, -1, 1, 1, -1, 0};
constexpr int EY[Q] = {0, 1, 0, -1, 0, 0, 0, 0, 1, 1, 0, 0, -1, -1, 1, -1, 1, -1, 0};
constexpr int EZ[Q] = {0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0};

// opposite direction (FluidNode.h:45-62)
constexpr int OPP[Q] = {2, 3, 0, 1, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 15, 14, 17, 16, 18};

__host__ __device__ constexpr int ecomp(int q, int a) { return a == 0 ? EX[q] : (a == 1 ? EY[q] : EZ[q]); }

// weight as the reference writes it: 1./18., 1./36., 1./3.
__host__ __device__ constexpr double weight(int q) { return q < 6 ? 1. / 18. : (q < 18 ? 1. / 36. : 1. / 3.); }

// Push rule of Lattice::setConnectivity (Lattice.cpp:244-1392): a node on the
// boundary pushes link q only if q points inward in every dimension in which the
// node lies on the boundary (bi, bj, bk = +1 at the low face, -1 at the high face,
// 0 inside).  Faces push 5 links, edges 1, corners none.
__host__ __device__ inline bool boundary_pushes(int q, int bi, int bj, int bk) {
    return (bi == 0 || EX[q] == bi) && (bj == 0 || EY[q] == bj) && (bk == 0 || EZ[q] == bk);
}

// Whether the node at x index s of a grid row pushes link q, for the bulk
// kernels' row-wise stores.  (bj, bk) is the row's y/z boundary class (0 on
// grids wrapped in y and z).  s may lie one node beyond the row, where the
// source is the node across the x wrap: none without wrap (LODI on all faces),
// none across an open or a LODI x face (xopen / xlodi: bit 0 the face s = 0,
// bit 1 the face s = nx - 1).  Inside the row the push rule above applies, and
// on a grid wrapped in y and z a LODI x-face node pushes its inward x links.
template <bool WRAPPED_YZ>
__host__ __device__ inline bool row_source_pushes(int q, int s, int nx, int bj, int bk, int xopen, int xlodi) {
    if (s < 0 || s >= nx) {
        if (!WRAPPED_YZ) return false;
        return ((xopen | xlodi) & (s < 0 ? 2 : 1)) == 0;
    }
    if (WRAPPED_YZ) {
        const int bi = (s == 0 && (xlodi & 1)) ? 1 : ((s == nx - 1 && (xlodi & 2)) ? -1 : 0);
        return bi == 0 || EX[q] == bi;
    }
    const int bi = s == 0 ? 1 : (s == nx - 1 ? -1 : 0);
    return (bi | bj | bk) == 0 || boundary_pushes(q, bi, bj, bk);
}

}  // namespace acp
