// tiles.hpp -- 8x8x8 tiles of the B = 8 voxel numbering (brick_index): a
// tile is 512 consecutive internal voxels (eight 4^3 bricks), tiles ordered
// y, x, z.  Stencil kernels stage U over a tile and a one-voxel shell
// ([-1, 8]^3) in LDS; these helpers map tile-local offsets to internal
// indices with the periodic wrap of D (3d/tval3_gpu_3d_kernels.cu:170-215).
#pragma once

#include <cstdint>

namespace tv3 {

constexpr int kTileBox = 10;   // shell edge: voxels -1 .. 8

__device__ __forceinline__ int box_slot(int a, int b, int c) {
    return ((c + 1) * kTileBox + (b + 1)) * kTileBox + (a + 1);
}

// Tile voxel q (internal order) -> offsets (a, b, c) = (y, x, z) in the tile.
__device__ __forceinline__ void tile_pos(int q, int &a, int &b, int &c) {
    const int sub = q >> 6, w = q & 63;
    a = ((sub & 1) << 2) | (w & 3);
    b = (((sub >> 1) & 1) << 2) | ((w >> 2) & 3);
    c = ((sub >> 2) << 2) | (w >> 4);
}

// A voxel at offsets (a, b, c) in [-1, 8]^3 from a tile's corner lies in the
// tile shifted by (a >> 3, b >> 3, c >> 3) in {-1, 0, 1}^3 (periodic wrap of
// the tile coordinates: every dimension is a multiple of 8) at the in-tile
// position of (a & 7, b & 7, c & 7); both parts are fixed per thread.
struct TileOff {
    int da, db, dc, off;
};

__device__ __forceinline__ TileOff tile_off(int a, int b, int c) {
    const int ai = a & 7, bi = b & 7, ci = c & 7;
    const int sub = ((ci >> 2) << 2) | ((bi >> 2) << 1) | (ai >> 2);
    return TileOff{a >> 3, b >> 3, c >> 3, (sub << 6) + ((ci & 3) << 4) + ((bi & 3) << 2) + (ai & 3)};
}

// internal index (from the volume's first voxel) of that voxel for the tile
// at tile coordinates (ty, tx, tz) of a volume of nby x nbx x nbz tiles
__device__ __forceinline__ int64_t tile_gidx(int nby, int nbx, int nbz, int ty, int tx, int tz,
                                             const TileOff &w) {
    int y = ty + w.da, x = tx + w.db, z = tz + w.dc;
    y = y < 0 ? y + nby : (y >= nby ? y - nby : y);
    x = x < 0 ? x + nbx : (x >= nbx ? x - nbx : x);
    z = z < 0 ? z + nbz : (z >= nbz ? z - nbz : z);
    return ((int64_t)((z * nbx + x) * nby + y) << 9) + w.off;
}

// box entry k of [-1, 8]^3 (y fastest) as a tile offset
__device__ __forceinline__ TileOff box_off(int k) {
    return tile_off(k % kTileBox - 1, (k / kTileBox) % kTileBox - 1, k / 100 - 1);
}

}  // namespace tv3
