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

}  // namespace acp
