// phx_push_common.hpp -- launch geometry and box predicates shared by the FP64 and the
// mixed-precision particle kernels (the two must agree on the work items the sort builds).
#pragma once

#include "phx_kernels.hpp"

#ifndef PHX_SEG_ROWS
#define PHX_SEG_ROWS 32 // most segment rows per workgroup of the segmented push (default)
#endif
#ifndef PHX_TILE_SPREAD
#define PHX_TILE_SPREAD 1
#endif
#ifndef PHX_PPT_3D
#define PHX_PPT_3D 16 // particles per thread of the linear 3D pushes: a chunk of 4096 is one brick
#endif
#ifndef PHX_SEG_WAVES
#define PHX_SEG_WAVES 4 // waves per SIMD the register allocation of the segmented push targets
#endif

namespace phx
{
namespace pc
{
constexpr int BLOCK    = 256;
constexpr int NW       = BLOCK / 64; // wavefronts per workgroup
#ifndef PHX_SEG_BLOCK
#define PHX_SEG_BLOCK 256
#endif
constexpr int SEG_BLOCK = PHX_SEG_BLOCK;    // work-items per workgroup of the segmented pushes
constexpr int SEG_NW    = SEG_BLOCK / 64;
constexpr int LANES    = 64;
constexpr unsigned int TILE_BYTES_MAX = 65536;
constexpr int TILE_SPREAD = PHX_TILE_SPREAD; // cells of drift since the sort in a segment tile

__device__ __forceinline__ bool inBox(int const* c, int const* lo, int const* hi, int D)
{
    for (int d = 0; d < D; ++d)
        if (c[d] < lo[d] || c[d] > hi[d])
            return false;
    return true;
}

// UpdaterSelectionBoxing::nonLevelGhostBox membership (ion_updater.hpp:128-163)
__device__ __forceinline__ bool inNLGB(int const* c, Boxes const& b, int D)
{
    for (int i = 0; i < b.nnlgb; ++i)
        if (inBox(c, b.nlgb_lo[i], b.nlgb_hi[i], D))
            return true;
    return false;
}

// timing record of a launch (KTime): workgroup 0 (dispatched first) stamps the start; every
// workgroup with work counts itself out at its end, and the last of the `working` ones stamps
// the end, accumulates the launch and re-arms the record. One relaxed atomic per workgroup and
// no fences (a device-scope fence writes back the L2: far too slow to pay per workgroup); the
// stamps are exact to the spread of the first workgroups' dispatch.
__device__ __forceinline__ void ktBegin(KTime* kt)
{
    if (kt && blockIdx.x == 0 && threadIdx.x == 0)
        atomicExch(&kt->start, static_cast<unsigned long long>(wall_clock64()));
}
__device__ __forceinline__ void ktEnd(KTime* kt, unsigned int working)
{
    if (!kt || threadIdx.x != 0)
        return;
    if (atomicAdd(&kt->done, 1u) != working - 1)
        return;
    unsigned long long const e = static_cast<unsigned long long>(wall_clock64());
    unsigned long long const s = atomicAdd(&kt->start, 0ull);
    atomicAdd(&kt->total, e - s);
    atomicAdd(&kt->count, 1u);
    atomicExch(&kt->done, 0u);
}

// XCD-aware order of a grid's workgroups: the hardware deals workgroup b to XCD b % 8, so that
// XCD x gets workgroups x, x + 8, ...; the work item of b is chosen so that each XCD takes one
// contiguous run of items (neighbouring segments share their field halos and moment nodes in that
// XCD's L2). A bijection of [0, n).
__device__ __forceinline__ unsigned int xcdItem(unsigned int b, unsigned int n)
{
    constexpr unsigned int XCD = 8;
    unsigned int const q = n / XCD, r = n % XCD;
    unsigned int const x = b % XCD, i = b / XCD;
    return x * q + (x < r ? x : r) + i;
}

// rows [r0, r1) of work item j of a segment of R rows: the segment's ceil(R / rowsPerItem)
// items share its rows evenly (at most rowsPerItem each)
__device__ __forceinline__ void itemRows(unsigned int j, unsigned int R, unsigned int rowsPerItem,
                                         unsigned int& r0, unsigned int& r1)
{
    unsigned int const k = (R + rowsPerItem - 1) / rowsPerItem;
    r0 = static_cast<unsigned int>(static_cast<unsigned long long>(j) * R / k);
    r1 = static_cast<unsigned int>(static_cast<unsigned long long>(j + 1) * R / k);
}

// local node index of the LDS tile origin of segment `seg` (the segment's first cell less the
// tile's low margin), from its patch's record; the overflow segment has no cell range and gets an
// unreachable origin
template<int D>
__device__ __forceinline__ void segTileOrigin(Geom const& G, SegGeom const& S, TileSpec const& tile,
                                              unsigned int seg, int* org)
{
    if (seg < static_cast<unsigned int>(S.nsegs))
    {
        // one patch: its record comes with the kernel arguments, no dependent table reads
        SegPatch const P = S.npatch == 1 ? S.first : S.patch[S.segPatch[seg]];
        int rem          = static_cast<int>(seg - P.segBase);
        int sidx[3];
#pragma unroll
        for (int d = D - 1; d >= 0; --d)
        {
            sidx[d] = rem % P.nseg[d];
            rem /= P.nseg[d];
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
            org[d] = P.lo[d] + sidx[d] * P.seg[d] - (G.lower[d] - G.g) - tile.lo[d];
    }
    else
    {
#pragma unroll
        for (int d = 0; d < D; ++d)
            org[d] = -0x20000000;
    }
}

// node margins of a segment tile: a particle's deposit spans nodes [-1, 2] of its cell (orders
// 2, 3) or [0, 1] (order 1); one cell of motion and TILE_SPREAD cells of drift since the sort
#ifndef PHX_TILE_ROW_PAD
#define PHX_TILE_ROW_PAD 16
#endif
// node pitches of a segment tile: in 2D the rows (fastest dimension) are padded to a multiple of
// `pad` nodes (PHX_TILE_ROW_PAD unless tuned), so that rows of the tile start on the same LDS
// bank and lanes of one wavefront whose stencils start on neighbouring rows do not collide in
// banks; 1D/3D unpadded
inline void tilePitch(TileSpec& T, int dim, int pad)
{
    for (int d = 0; d < 3; ++d)
        T.pitch[d] = T.ext[d];
    if (dim == 2 && pad > 1)
        T.pitch[1] = (T.ext[1] + pad - 1) / pad * pad;
}

inline int segMarginLo(int order) { return (order >= 2 ? 1 : 0) + 1 + TILE_SPREAD; }
inline int segMarginHi(int order) { return (order == 1 ? 1 : 2) + 1 + TILE_SPREAD; }

} // namespace pc
} // namespace phx
