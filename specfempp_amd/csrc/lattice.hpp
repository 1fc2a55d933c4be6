// Device helpers shared by the 3-D kernels: tiled lattice addressing,
// element-force slot order and XCD-aware workgroup remapping.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace sfx {

// The 8 XCDs of the device receive workgroups round-robin; remap so each XCD
// processes one contiguous range of work items (and its L2 sees their shared
// data).
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  constexpr int NXCD = 8;
  const int per = nblocks / NXCD;  // blocks with index < per * NXCD are remapped
  if (b >= per * NXCD) return b;
  return (b % NXCD) * per + b / NXCD;
}

// Device index of local point (i, j, k) (0..4) of element (ex, ey, ez).
__device__ __forceinline__ int tiled_index(const Lattice &lat, int ex, int ey, int ez, int i, int j, int k) {
  const int tx = ex + (i >> 2), ty = ey + (j >> 2), tz = ez + (k >> 2);
  return ((tz * lat.TY + ty) * lat.TX + tx) * 64 + (((k & 3) * 4 + (j & 3)) * 4 + (i & 3));
}

// Element-force block: 3 components of 128 floats (125 used), so every
// component starts 16-byte aligned.
constexpr int RE_STRIDE = 128;
constexpr int RE_ELEM = 3 * RE_STRIDE;

// Slot of local point (i, j, k) in an element's block of re.
__device__ __forceinline__ int re_slot(int i, int j, int k) {
  const int fi = i >> 2, fj = j >> 2, fk = k >> 2;  // 1 on the local index-4 faces
  switch (fi | (fj << 1) | (fk << 2)) {
    case 0: return (k * 4 + j) * 4 + i;
    case 1: return 64 + k * 4 + j;
    case 2: return 80 + k * 4 + i;
    case 4: return 96 + j * 4 + i;
    case 3: return 112 + k;
    case 5: return 116 + j;
    case 6: return 120 + i;
    default: break;
  }
  return 124;
}

// ---- block assembly (2 x 2 x 2 elements, 9 x 9 x 9 lattice points) ----
// Block (bx, by, bz) covers lattice points 8 bx .. 8 bx + 8 (and so on).  Its
// 386 surface points are shared with neighbour blocks; their partial sums are
// kept per "owner": the owner of lattice point (I, J, K) with one coordinate a
// multiple of 8 is block (I / 8, J / 8, K / 8) (an owner grid of
// (nbx + 1) (nby + 1) (nbz + 1) covers the upper mesh faces), and the sum a
// block contributes from offset d = dx + 2 dy + 4 dz below the owner goes to
// the owner's region d.  Per owner: 386 slots (and per component a stride of
// 388 floats) -- region 0 the owner's own 169 surface points, regions 1, 2, 4
// the 64 points of its low x, y, z faces, 3, 5, 6 the 8 points of its low
// edges, 7 its low corner.
constexpr int OWNER_SLOTS = 386;
constexpr int BS_STRIDE = 388;
constexpr int BS_ELEM = 3 * BS_STRIDE;

// Slot of owner-local point (li, lj, lk) (0..7; zero on the axes of d) in
// region d.
__host__ __device__ __forceinline__ int owner_slot(int d, int li, int lj, int lk) {
  switch (d) {
    case 0: return lk == 0 ? lj * 8 + li : 64 + (lk - 1) * 15 + (lj == 0 ? li : 7 + lj);
    case 1: return 169 + lk * 8 + lj;
    case 2: return 233 + lk * 8 + li;
    case 4: return 297 + lj * 8 + li;
    case 3: return 361 + lk;
    case 5: return 369 + lj;
    case 6: return 377 + li;
    default: return 385;
  }
}

// Owner index and slot of block-surface point (bi, bj, bk) (0..8) of block
// (bx, by, bz); false for a point inside the block.
__host__ __device__ __forceinline__ bool block_point_owner(int nbx, int nby, int bx, int by, int bz, int bi, int bj,
                                                           int bk, int &owner, int &slot) {
  const bool on = bi == 0 || bj == 0 || bk == 0 || bi == 8 || bj == 8 || bk == 8;
  if (!on) return false;
  const int dx = bi >> 3, dy = bj >> 3, dz = bk >> 3;
  owner = ((bz + dz) * (nby + 1) + by + dy) * (nbx + 1) + bx + dx;
  slot = owner_slot(dx | (dy << 1) | (dz << 2), bi & 7, bj & 7, bk & 7);
  return true;
}

// Block layout of the fields (SFX_SCATTER_BLOCK): the 343 points strictly
// inside block b are slots b * 343 + ((lk - 1) * 7 + (lj - 1)) * 7 + li - 1;
// then, from slot nint_pad (343 * blocks, rounded up to 64), every owner has
// 176 slots of which its 169 own surface points use owner_slot(0, ...).  The
// block kernel then reads and writes its interior as one contiguous range and
// the surface kernel touches only surface slots.
constexpr int OWNER_STRIDE = 176;
__host__ __device__ __forceinline__ int block_layout_index(int nbx, int nby, int nint_pad, int I, int J, int K) {
  const int li = I & 7, lj = J & 7, lk = K & 7;
  if (li && lj && lk) {
    const int b = ((K >> 3) * nby + (J >> 3)) * nbx + (I >> 3);
    return b * 343 + ((lk - 1) * 7 + (lj - 1)) * 7 + (li - 1);
  }
  const int o = ((K >> 3) * (nby + 1) + (J >> 3)) * (nbx + 1) + (I >> 3);
  return nint_pad + o * OWNER_STRIDE + owner_slot(0, li, lj, lk);
}

// Block-local coordinates (0..8) of the q-th of a block's 386 surface points:
// the bk = 0 and bk = 8 planes, then the ring of 32 points of each layer
// bk = 1..7.
__host__ __device__ __forceinline__ void block_surface_point(int q, int &bi, int &bj, int &bk) {
  if (q < 162) {
    bk = q < 81 ? 0 : 8;
    const int r = q < 81 ? q : q - 81;
    bj = r / 9;
    bi = r - bj * 9;
    return;
  }
  const int r0 = q - 162;
  bk = 1 + r0 / 32;
  const int r = r0 - (r0 / 32) * 32;
  if (r < 9) {
    bj = 0, bi = r;
  } else if (r < 18) {
    bj = 8, bi = r - 9;
  } else if (r < 25) {
    bi = 0, bj = r - 17;
  } else {
    bi = 8, bj = r - 24;
  }
}

}  // namespace sfx
