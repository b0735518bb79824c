// phx_particles.hip -- particle kernels for gfx950: fused prePush + gather + Boris + postPush +
// moment deposit (IonUpdater::updatePopulations, ion_updater.hpp:88-295, with the Boris pusher
// boris.hpp:93-138 and the Interpolator gather/deposit interpolator.hpp:420-510), the counting
// sort of domain arrays into cell-segmented storage, stable compaction of particle arrays, and
// the periodic re-entry of leaving particles.
//
// Particles are structure-of-arrays in HBM (int32 iCell[D], f64 delta[D], f64 v[3], f64 weight,
// f64 charge), each component a contiguous stream, in one of two storage orders:
//  - linear: a workgroup owns a chunk of consecutive particles, lanes taking them with a stride
//    across the chunk (any order is correct; cell order is faster);
//  - cell-segmented (built by the sort): the 64 lanes of a wavefront read one row of a segment
//    (a box of up to 64 cells whose particles, in cell order, fill its rows column by column), so
//    that each lane walks a run of cell-ordered particles and the lanes of one row hold cells
//    spread over the segment. Reads stay coalesced and the lanes of one LDS deposit instruction
//    add into distinct words, except where a crowded cell spans two lanes.
// Either way a workgroup accumulates its moments in an LDS tile over its cells and flushes the
// tile with one HBM atomic per touched node and moment; the gather reads E/B through the L1/L2
// caches, leaving the LDS to the deposit. Particles whose stencil leaves the tile deposit
// straight into HBM.
#include "phx_push_common.hpp"

#include <algorithm>
#include <mutex>

#ifndef PHX_PPT_ND
#define PHX_PPT_ND 2 // particles per thread of the linear 2D push
#endif

namespace phx
{

namespace
{
using pc::BLOCK;
using pc::inBox;
using pc::inNLGB;
using pc::LANES;
using pc::NW;
using pc::TILE_BYTES_MAX;

// Per (dim, order): whether the deposit is accumulated in registers over a union window, the
// particles per thread of the linear push and the window. With accumulation, a thread's
// successive particles (mostly one cell) add into the window and scatter to LDS only when the
// cell changes. Only 1D keeps the window (5 moments x 4 nodes) in registers without losing
// occupancy; 2D/3D scatter every particle to the LDS tile.
template<int D, int O>
struct Cfg
{
    static constexpr bool ACC  = D == 1;
    static constexpr int PPT   = ACC ? 8 : (D == 3 ? PHX_PPT_3D : PHX_PPT_ND);
    static constexpr int CHUNK = BLOCK * PPT;
    static constexpr int W     = O == 1 ? 2 : 4;  // window width per direction
    static constexpr int WB    = O == 1 ? 0 : 1;  // window base = local cell - WB
    static constexpr int WN    = D == 1 ? W : (D == 2 ? W * W : W * W * W);
};

// linear storage: (thread, step) -> slot in the chunk; consecutive threads read consecutive slots
// (coalesced). The loader's arrays are stored chunk-interleaved (chunkInterleave on the host):
// slot k*BLOCK + t holds the particle (r*PPT + k) of the caller's order, r = lane*NW + wave, so
// the lanes of a wavefront own runs spread across the chunk.
__device__ __forceinline__ int chunkOffset(int t, int k) { return k * BLOCK + t; }

// LDS view of the tile: moment slab (m, copy)
struct TileView
{
    double* lds;
    int org[3];
    int ext[3];
    int pitch[3];
    int copies, stride;
    __device__ double* mom(int m, int copy) const { return lds + (m * copies + copy) * stride; }
    __device__ int node(int i, int j, int k) const { return (i * pitch[1] + j) * pitch[2] + k; }
    // the O+1 nodes from start lie in the tile along d
    template<int O>
    __device__ bool holds(int d, int start) const
    {
        return start >= org[d] && start + O < org[d] + ext[d];
    }
};

template<int D, int O>
struct DepositData
{
    int start[D];
    double w[D][O + 1];
    double val[5];
};

template<int D, int O>
__device__ __forceinline__ void primalWeights(Geom const& G, int const* cell, double const* delta,
                                              DepositData<D, O>& dd)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int icl     = cell[d] - (G.lower[d] - G.g);
        dd.start[d] = indexAndWeights<O>(CP, icl, delta[d], dd.w[d]);
    }
}

// contribution of one particle to node (a,b,c) of its stencil, in the reference's
// multiplication order (interpolator.hpp:290-361)
template<int D, int O>
__device__ __forceinline__ double nodeWeight(DepositData<D, O> const& dd, double dep, int a, int b,
                                             int c)
{
    if constexpr (D == 1)
        return dep * dd.w[0][a];
    else if constexpr (D == 2)
        return dep * dd.w[0][a] * dd.w[1][b];
    else
        return dep * dd.w[0][a] * dd.w[1][b] * dd.w[2][c];
}


// E and B at the particle (Interpolator::operator() for the fields, interpolator.hpp:420-456)
template<int D, int O>
__device__ __forceinline__ void gatherEB(Geom const& G, EMPtrs const& em, Stencil<D, O> const& s,
                                         double* E, double* B)
{
    E[0] = interp<D, O>(G, em.E[0], QEX, s);
    E[1] = interp<D, O>(G, em.E[1], QEY, s);
    E[2] = interp<D, O>(G, em.E[2], QEZ, s);
    B[0] = interp<D, O>(G, em.B[0], QBX, s);
    B[1] = interp<D, O>(G, em.B[1], QBY, s);
    B[2] = interp<D, O>(G, em.B[2], QBZ, s);
}

// scatter-add the 5 moments of one particle: into the lane's copy of the LDS tile when its
// whole stencil lies in the tile, else straight into the HBM moment arrays. RN: the population
// has one charge, the tile skips the charge density and its flush derives it from n
template<int D, int O, bool RN = false>
__device__ __forceinline__ void depositOne(DepositData<D, O> const& dd, Geom const& G,
                                           TileView const& T, int copy, double* const* mom)
{
    constexpr int S  = O + 1;
    constexpr int SY = D > 1 ? S : 1;
    constexpr int SZ = D > 2 ? S : 1;
    bool inTile      = true;
#pragma unroll
    for (int d = 0; d < D; ++d)
        inTile = inTile && T.template holds<O>(d, dd.start[d]);
    if (inTile)
    {
        double* slab[5];
#pragma unroll
        for (int m = 0; m < 5; ++m)
            slab[m] = T.mom(m, copy);
#pragma unroll
        for (int a = 0; a < S; ++a)
#pragma unroll
            for (int b = 0; b < SY; ++b)
#pragma unroll
                for (int c = 0; c < SZ; ++c)
                {
                    int const node = T.node(dd.start[0] + a - T.org[0],
                                            D > 1 ? dd.start[1] + b - T.org[1] : 0,
                                            D > 2 ? dd.start[2] + c - T.org[2] : 0);
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        if (!RN || m != 1)
                            atomicAdd(slab[m] + node, nodeWeight<D, O>(dd, dd.val[m], a, b, c));
                }
    }
    else
    {
#pragma unroll
        for (int a = 0; a < S; ++a)
#pragma unroll
            for (int b = 0; b < SY; ++b)
#pragma unroll
                for (int c = 0; c < SZ; ++c)
                {
                    int node = G.idx(QMOM, dd.start[0] + a, D > 1 ? dd.start[1] + b : 0,
                                     D > 2 ? dd.start[2] + c : 0);
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        atomicAdd(mom[m] + node, nodeWeight<D, O>(dd, dd.val[m], a, b, c));
                }
    }
}

// register accumulator of one thread: the 5 moments over the union window of its current cell
template<int D, int O>
struct Acc
{
    static constexpr int W  = Cfg<D, O>::W;
    static constexpr int WN = Cfg<D, O>::WN;
    double a[5][WN];
    int base[D]; // local node index of the window origin
    bool live;
};

template<int D, int O>
__device__ __forceinline__ void accZero(Acc<D, O>& A)
{
#pragma unroll
    for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int x = 0; x < Acc<D, O>::WN; ++x)
            A.a[m][x] = 0.;
    A.live = false;
}

// scatter the window: nodes inside the LDS tile by LDS atomics into the lane's copy, others
// straight to HBM; never-touched nodes hold exactly zero and are skipped
template<int D, int O>
__device__ __forceinline__ void accFlush(Acc<D, O> const& A, Geom const& G, TileView const& T,
                                         int copy, double* const* mom)
{
    constexpr int W = Acc<D, O>::W;
    if (!A.live)
        return;
#pragma unroll
    for (int x = 0; x < Acc<D, O>::WN; ++x)
    {
        int const a = D == 1 ? x : (D == 2 ? x / W : x / (W * W));
        int const b = D == 1 ? 0 : (D == 2 ? x % W : (x / W) % W);
        int const c = D == 3 ? x % W : 0;
        int const g[3] = {A.base[0] + a, D > 1 ? A.base[1] + b : 0, D > 2 ? A.base[2] + c : 0};
        bool inTile    = true;
#pragma unroll
        for (int d = 0; d < D; ++d)
            inTile = inTile && g[d] >= T.org[d] && g[d] < T.org[d] + T.ext[d];
        int const tnode = T.node(g[0] - T.org[0], D > 1 ? g[1] - T.org[1] : 0,
                                 D > 2 ? g[2] - T.org[2] : 0);
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            double const v = A.a[m][x];
            if (v != 0.)
            {
                if (inTile)
                    atomicAdd(T.mom(m, copy) + tnode, v);
                else
                    atomicAdd(mom[m] + G.idx(QMOM, g[0], g[1], g[2]), v);
            }
        }
    }
}

// add one particle's deposit (already weighted per the reference's products) to the window
template<int D, int O>
__device__ __forceinline__ void accAdd(Acc<D, O>& A, DepositData<D, O> const& dd)
{
    constexpr int W = Acc<D, O>::W;
    constexpr int S = O + 1;
    // weights placed in the window: offset of the stencil start from the window base
    double w[D][W];
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int const off = dd.start[d] - A.base[d];
#pragma unroll
        for (int a = 0; a < W; ++a)
        {
            double x = 0.;
#pragma unroll
            for (int e = 0; e < S; ++e)
                x = (a - off == e) ? dd.w[d][e] : x;
            w[d][a] = x;
        }
    }
#pragma unroll
    for (int x = 0; x < Acc<D, O>::WN; ++x)
    {
        int const a = D == 1 ? x : (D == 2 ? x / W : x / (W * W));
        int const b = D == 1 ? 0 : (D == 2 ? x % W : (x / W) % W);
        int const c = D == 3 ? x % W : 0;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            double v;
            if constexpr (D == 1)
                v = dd.val[m] * w[0][a];
            else if constexpr (D == 2)
                v = dd.val[m] * w[0][a] * w[1][b];
            else
                v = dd.val[m] * w[0][a] * w[1][b] * w[2][c];
            A.a[m][x] += v;
        }
    }
}

// the fields of one particle read from its slot (cell[0] is DEAD_CELL for an empty slot)
template<int D>
struct Part
{
    int cell[D];
    double delta[D], v[3], w, q;
};

template<int D>
__device__ __forceinline__ void loadPart(PView const& in, unsigned long long i, Part<D>& p)
{
    p.cell[0] = in.ic[0][i];
    if (p.cell[0] == DEAD_CELL)
        return;
#pragma unroll
    for (int d = 1; d < D; ++d)
        p.cell[d] = in.ic[d][i];
#pragma unroll
    for (int d = 0; d < D; ++d)
        p.delta[d] = in.de[d][i];
    p.v[0] = in.v[0][i];
    p.v[1] = in.v[1][i];
    p.v[2] = in.v[2][i];
    p.w    = in.w[i];
    p.q    = in.q[i];
}

// One particle of updatePopulations: prePush, gather, Boris, postPush (boris.hpp:93-138), the
// selection of ion_updater.hpp:169-295, the deposit and, in mode all, the write-back. Dead slots
// (lost particles, empty slots of segmented storage) are skipped, and dropped by a compaction.
template<int D, int O, bool RN = false>
__device__ __forceinline__ void pushLoaded(PushArgs const& A, PView const& out, bool outOfPlace,
                                           unsigned long long i, Part<D> const& P,
                                           TileView const& T, int copy, Acc<D, O>& acc)
{
    using C       = Cfg<D, O>;
    Geom const& G = A.G;
    if (P.cell[0] == DEAD_CELL)
    {
        if (A.pushOn && A.mode == MODE_ALL && A.status)
            A.status[i] = ST_DROP;
        return;
    }
    int cell[D];
    double delta[D], v[3];
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        cell[d]  = P.cell[d];
        delta[d] = P.delta[d];
    }
    v[0] = P.v[0];
    v[1] = P.v[1];
    v[2] = P.v[2];
    double const w = P.w, q = P.q;

    bool deposit = true, sel1 = true, inDom = true;
    uint8_t st   = ST_KEEP;
    if (A.pushOn)
    {
        double bd, bv;
        // prePushStep_ (boris.hpp:180-216)
        if (!advancePosition<D>(cell, delta, v, A.hdtdx, bd, bv))
        {
            raiseError(A.err, 3u, 1u, i, bd, bv, A.kind, A.pop);
            return;
        }
        // first selector: noop (domain) or inGhostBox (level ghosts)
        sel1 = A.kind == KIND_DOMAIN ? true : inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
        if (sel1)
        {
            Stencil<D, O> s;
            makeStencil<D, O>(G, cell, delta, s);
            bool safe = true;
#pragma unroll
            for (int d = 0; d < D; ++d)
            {
                int const icl = cell[d] - (G.lower[d] - G.g);
                safe          = safe && icl >= G.safeLo[d] && icl <= G.safeHi[d];
            }
            if (!safe
                && (!stencilInside<D, O>(G, s, QEX) || !stencilInside<D, O>(G, s, QBX)
                    || !stencilInside<D, O>(G, s, QEY) || !stencilInside<D, O>(G, s, QBY)
                    || !stencilInside<D, O>(G, s, QEZ) || !stencilInside<D, O>(G, s, QBZ)))
            {
                raiseError(A.err, 4u, 2u, i, delta[0], v[0], A.kind, A.pop);
                return;
            }
            double E[3], B[3];
            gatherEB<D, O>(G, A.em, s, E, B);
            borisAccelerate(v, E, B, q, A.dto2m);
            // postPushStep_ (boris.hpp:218-234)
            if (!advancePosition<D>(cell, delta, v, A.hdtdx, bd, bv))
            {
                raiseError(A.err, 3u, 2u, i, bd, bv, A.kind, A.pop);
                return;
            }
        }
        // what is deposited and what the array keeps (ion_updater.hpp:169-295)
        inDom = inBox(cell, A.box.dom_lo, A.box.dom_hi, D);
        if (A.kind == KIND_DOMAIN)
        {
            bool const nl = inNLGB(cell, A.box, D);
            deposit       = nl;
            st            = inDom ? ST_KEEP : (nl ? ST_LEAVE : ST_DROP);
        }
        else
        {
            deposit = sel1 && inDom;
            bool const layer = sel1 && !inDom && inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
            st = !sel1 ? ST_DROP : (inDom ? ST_EXPORT : (layer ? ST_KEEP : ST_DROP));
        }
    }

    if (deposit)
    {
        DepositData<D, O> dd;
        primalWeights<D, O>(G, cell, delta, dd);
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d)
            ok = ok && dd.start[d] >= 0 && dd.start[d] + O < G.shape[QMOM][d];
        if (!ok)
            raiseError(A.err, 4u, 3u, i, delta[0], v[0], A.kind, A.pop);
        else
        {
            dd.val[0] = 1. * w * A.coef;
            dd.val[1] = q * w * A.coef;
            dd.val[2] = v[0] * w * A.coef;
            dd.val[3] = v[1] * w * A.coef;
            dd.val[4] = v[2] * w * A.coef;
            if constexpr (C::ACC)
            {
                // same cell as the window -> accumulate; else flush and restart there
                bool same = acc.live;
#pragma unroll
                for (int d = 0; d < D; ++d)
                    same = same && acc.base[d] == cell[d] - (G.lower[d] - G.g) - C::WB;
                if (!same)
                {
                    accFlush<D, O>(acc, G, T, copy, A.mom);
                    accZero(acc);
#pragma unroll
                    for (int d = 0; d < D; ++d)
                        acc.base[d] = cell[d] - (G.lower[d] - G.g) - C::WB;
                    acc.live = true;
                }
                accAdd<D, O>(acc, dd);
            }
            else
                depositOne<D, O, RN>(dd, G, T, copy, A.mom);
        }
    }

    if (A.pushOn && A.mode == MODE_ALL)
    {
        if (A.periodicWrap && A.kind == KIND_DOMAIN)
        {
            if (st == ST_LEAVE)
            {
                // fillIonGhostParticles: the leaver re-enters through the periodic image
#pragma unroll
                for (int d = 0; d < D; ++d)
                {
                    if (cell[d] < A.box.dom_lo[d])
                        cell[d] += G.n[d];
                    else if (cell[d] > A.box.dom_hi[d])
                        cell[d] -= G.n[d];
                }
                st = ST_KEEP;
            }
            else if (st == ST_DROP)
            {
                cell[0] = DEAD_CELL;
                atomicAdd(A.dropped, 1u);
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            out.ic[d][i] = cell[d];
            out.de[d][i] = delta[d];
        }
        out.v[0][i] = v[0];
        out.v[1][i] = v[1];
        out.v[2][i] = v[2];
        if (outOfPlace)
        {
            out.w[i] = w;
            out.q[i] = q;
        }
        if (A.status)
            A.status[i] = st;
    }
}

template<int D, int O, bool RN = false>
__device__ __forceinline__ void pushParticle(PushArgs const& A, PView const& in, PView const& out,
                                             unsigned long long i, TileView const& T, int copy,
                                             Acc<D, O>& acc)
{
    Part<D> P;
    loadPart<D>(in, i, P);
    pushLoaded<D, O, RN>(A, out, out.w != in.w, i, P, T, copy, acc);
}

template<int TB = BLOCK>
__device__ __forceinline__ void zeroTile(double* lds, TileSpec const& S)
{
    for (int x = threadIdx.x; x < 5 * S.copies * S.stride; x += TB)
        lds[x] = 0.;
}

// flush the tile: the copies of each node summed, one HBM atomic per touched node and moment;
// RN: the charge density is the population's charge times n
template<int D, bool RN = false, int TB = BLOCK>
__device__ __forceinline__ void flushTile(Geom const& G, TileView const& T, int nodes,
                                          double* const* mom, double charge = 0.)
{
    int const e12 = T.pitch[1] * T.pitch[2];
    for (int x = threadIdx.x; x < nodes; x += TB)
    {
        int node = -1;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            double val = 0.;
            if (RN && m == 1)
            {
                for (int c = 0; c < T.copies; ++c)
                    val += T.mom(0, c)[x];
                val *= charge;
            }
            else
                for (int c = 0; c < T.copies; ++c)
                    val += T.mom(m, c)[x];
            if (val != 0.)
            {
                if (node < 0)
                {
                    int const ti = x / e12;
                    int const r  = x - ti * e12;
                    int const tj = r / T.pitch[2];
                    int const tk = r - tj * T.pitch[2];
                    node = G.idx(QMOM, T.org[0] + ti, D > 1 ? T.org[1] + tj : 0,
                                 D > 2 ? T.org[2] + tk : 0);
                }
                atomicAdd(mom[m] + node, val);
            }
        }
    }
}

__device__ __forceinline__ void tileView(TileView& T, double* lds, TileSpec const& S)
{
    T.lds    = lds;
    T.copies = S.copies;
    T.stride = S.stride;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T.ext[d]   = S.ext[d];
        T.pitch[d] = S.pitch[d];
        T.org[d] = 0;
    }
}

} // namespace

// ---------------------------------------------------------------------------------------------
// fused push + deposit over linear storage (and deposit-only for unpushed arrays)
// ---------------------------------------------------------------------------------------------
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_push(PushArgs A, PView in, PView out)
{
    using C             = Cfg<D, O>;
    constexpr int PPT   = C::PPT;
    constexpr int CHUNK = C::CHUNK;
    extern __shared__ double lds[];
    if (A.G.stopped())
        return;

    Geom const& G  = A.G;
    int const t    = threadIdx.x;
    int const copy = (t & 63) & (A.tile.copies - 1);
    unsigned long long const base = static_cast<unsigned long long>(blockIdx.x) * CHUNK;

    zeroTile(lds, A.tile);
    // the tile is placed around the cell of the chunk's first particle (chunks of a cell-ordered
    // array hold neighbouring cells), with host-chosen margins for the chunk's extent, one cell
    // of motion and the stencil; brick tiles (3D, brick-ordered arrays) around the brick of that
    // cell; a dead first particle leaves the tile out of reach
    TileView T;
    tileView(T, lds, A.tile);
    int const c0   = in.ic[0][base];
    bool const far = c0 == DEAD_CELL;
    int const bk   = A.tile.brick;
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int first = in.ic[d][base] - G.lower[d];
        if (bk > 0)
            first = (first >= 0 ? first / bk : -((bk - 1 - first) / bk)) * bk;
        T.org[d] = far ? -0x20000000 : first + G.g - A.tile.lo[d];
    }
    __syncthreads();

    Acc<D, O> acc;
    if constexpr (C::ACC)
        accZero(acc);
#pragma unroll 1
    for (int k = 0; k < PPT; ++k)
    {
        unsigned long long const i = base + chunkOffset(t, k);
        if (i < A.n)
            pushParticle<D, O>(A, in, out, i, T, copy, acc);
    }
    if constexpr (C::ACC)
        accFlush<D, O>(acc, G, T, copy, A.mom);
    __syncthreads();
    flushTile<D>(G, T, A.tile.nodes, A.mom);
}

// ---------------------------------------------------------------------------------------------
// fused push + deposit over cell-segmented storage
// ---------------------------------------------------------------------------------------------
// rows [rA, rB) of segment seg through the LDS tile: zero the tile, push the rows (wavefront w
// takes rows rA + w, rA + w + NW, ...), flush the tile. Every work-item of the workgroup calls it.
template<int D, int O, bool RN>
__device__ __forceinline__ void pushPiece(PushArgs const& A, PView const& in, double* lds,
                                          unsigned int seg, unsigned int rA, unsigned int rB)
{
    PView const& out = in; // segmented arrays are pushed in place
    using C = Cfg<D, O>;
    Geom const& G          = A.G;
    unsigned long long const off = A.st.off[seg];
    int const t    = threadIdx.x;
    int const lane = t & 63;
    int const copy = lane & (A.tile.copies - 1);

    zeroTile<pc::SEG_BLOCK>(lds, A.tile);
    // the tile covers the segment's cells with host-chosen margins (stencil, one cell of
    // motion, drift since the sort); the overflow segment has no cell range
    TileView T;
    tileView(T, lds, A.tile);
    pc::segTileOrigin<D>(G, A.sg, A.tile, seg, T.org);
    __syncthreads();

    Acc<D, O> acc;
    if constexpr (C::ACC)
        accZero(acc);
    for (unsigned int r = rA + (t >> 6); r < rB; r += pc::SEG_NW)
        pushParticle<D, O, RN>(A, in, out, off + static_cast<unsigned long long>(r) * LANES + lane, T,
                               copy, acc);
    if constexpr (C::ACC)
        accFlush<D, O>(acc, G, T, copy, A.mom);
    __syncthreads();
    flushTile<D, RN, pc::SEG_BLOCK>(G, T, A.tile.nodes, A.mom, A.charge);
}

// one workgroup per work item (a segment and its share of the segment's rows, at most
// sg.rowsPerItem)
template<int D, int O, bool RN>
__global__ __launch_bounds__(pc::SEG_BLOCK) __attribute__((amdgpu_waves_per_eu(PHX_SEG_WAVES)))
void k_push_seg(PushArgs A, PView in)
{
    extern __shared__ double lds[];
    unsigned int const item = blockIdx.x;
    unsigned int const nItems = *A.st.nItems;
    if (item >= nItems)
        return;
    pc::ktBegin(A.kt);
    if (!A.G.stopped())
    {
        unsigned int const seg = A.st.items[2 * item];
        unsigned int const R   = A.st.rows[seg];
        unsigned int r0, rEnd;
        pc::itemRows(A.st.items[2 * item + 1], R, static_cast<unsigned int>(A.sg.rowsPerItem), r0, rEnd);
        pushPiece<D, O, RN>(A, in, lds, seg, r0, rEnd);
    }
    if (A.kt)
    {
        __syncthreads();
        pc::ktEnd(A.kt, nItems);
    }
}

// ---------------------------------------------------------------------------------------------
// fused push + deposit over cell-segmented storage, one workgroup per segment with all its rows
// (the staged 3D path, segments of 4x4x4 cells): wavefront w takes rows w, w + 4, ..., so that
// each instruction's 64 lanes walk 64 cells of the segment; the tile covers the segment with the
// stencil and one cell of motion on each side; the overflow segment (number nsegs) deposits
// straight into HBM. RN: the population has one charge, the tile leaves the charge density out
// and its flush takes the charge times n
// ---------------------------------------------------------------------------------------------
template<int D, int O, bool RN>
__global__ __launch_bounds__(BLOCK) void k_push_segwg(PushArgs A, PView in)
{
    extern __shared__ double lds[];
    if (A.G.stopped())
        return;
    Geom const& G          = A.G;
    unsigned int const seg = pc::xcdItem(blockIdx.x, gridDim.x);
    int const t            = threadIdx.x;
    int const lane         = t & 63;
    unsigned long long const off = A.st.off[seg];
    unsigned int const R         = A.st.rows[seg];
    zeroTile(lds, A.tile);
    TileView T;
    tileView(T, lds, A.tile);
    pc::segTileOrigin<D>(G, A.sg, A.tile, seg, T.org);
    __syncthreads();
    Acc<D, O> acc;
    if constexpr (Cfg<D, O>::ACC)
        accZero(acc);
#pragma unroll 1
    for (unsigned int r = static_cast<unsigned int>(t >> 6); r < R; r += NW)
        pushParticle<D, O, RN>(A, in, in, off + static_cast<unsigned long long>(r) * LANES + lane, T, 0, acc);
    if constexpr (Cfg<D, O>::ACC)
        accFlush<D, O>(acc, G, T, 0, A.mom);
    __syncthreads();
    flushTile<D, RN>(G, T, A.tile.nodes, A.mom, A.charge);
}

// ---------------------------------------------------------------------------------------------
// stable compaction: append the particles whose status equals `want` to dst at dstOffset
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_count_status(uint8_t const* __restrict__ status,
                                                        unsigned long long n, uint8_t want,
                                                        unsigned int* __restrict__ blockCounts)
{
    __shared__ unsigned int s;
    if (threadIdx.x == 0)
        s = 0;
    __syncthreads();
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    bool const f               = i < n && status[i] == want;
    unsigned long long const m = __ballot(f);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(&s, static_cast<unsigned int>(__popcll(m)));
    __syncthreads();
    if (threadIdx.x == 0)
        blockCounts[blockIdx.x] = s;
}

// exclusive scan of the block counts in one workgroup; total written at blockCounts[nblocks]
__global__ __launch_bounds__(1024) void k_scan_counts(unsigned int* __restrict__ blockCounts,
                                                      unsigned int nblocks,
                                                      unsigned long long* __restrict__ total)
{
    __shared__ unsigned int part[1024];
    unsigned int const per = (nblocks + 1023) / 1024;
    unsigned int const b0  = threadIdx.x * per;
    unsigned int sum       = 0;
    for (unsigned int b = b0; b < b0 + per && b < nblocks; ++b)
        sum += blockCounts[b];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (unsigned int off = 1; off < 1024; off <<= 1)
    {
        unsigned int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned int run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (unsigned int b = b0; b < b0 + per && b < nblocks; ++b)
    {
        unsigned int c = blockCounts[b];
        blockCounts[b] = run;
        run += c;
    }
    if (threadIdx.x == 1023)
        *total = part[1023];
}

__global__ __launch_bounds__(BLOCK) void k_scatter_status(PView src, PView dst,
                                                          uint8_t const* __restrict__ status,
                                                          unsigned long long n, uint8_t want,
                                                          unsigned int const* __restrict__ offsets,
                                                          unsigned long long dstOffset, int dim,
                                                          Wrap wrap)
{
    __shared__ unsigned int waveBase[BLOCK / 64];
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    bool const f               = i < n && status[i] == want;
    unsigned long long const m = __ballot(f);
    int const lane             = threadIdx.x & 63;
    int const wave             = threadIdx.x >> 6;
    if (lane == 0)
        waveBase[wave] = static_cast<unsigned int>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned int run = 0;
        for (int w = 0; w < BLOCK / 64; ++w)
        {
            unsigned int c = waveBase[w];
            waveBase[w]    = run;
            run += c;
        }
    }
    __syncthreads();
    if (!f)
        return;
    unsigned long long const below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    unsigned long long const j
        = dstOffset + offsets[blockIdx.x] + waveBase[wave] + static_cast<unsigned int>(__popcll(below));
    for (int d = 0; d < dim; ++d)
    {
        int c = src.ic[d][i];
        if (wrap.on)
        {
            // periodic shift of a patch-ghost particle into the domain of the image
            if (c < wrap.lo[d])
                c += wrap.n[d];
            else if (c > wrap.hi[d])
                c -= wrap.n[d];
        }
        dst.ic[d][j] = c;
        dst.de[d][j] = src.de[d][i];
    }
    dst.v[0][j] = src.v[0][i];
    dst.v[1][j] = src.v[1][i];
    dst.v[2][j] = src.v[2][i];
    dst.w[j]    = src.w[i];
    if (src.q)
        dst.q[j] = src.q[i];
}
// ---------------------------------------------------------------------------------------------
// counting sort of a domain array into cell-segmented storage (dead particles are dropped)
// ---------------------------------------------------------------------------------------------
namespace
{
constexpr unsigned int DEAD_KEY = 0xffffffffu;

// key of a particle: the row-major index of its cell inside the patch owning the cell, offset by
// the patch's first key; S.ncells for a cell no patch owns (outside the level box or not covered),
// DEAD_KEY for a lost particle. p receives the patch, c the cell relative to the patch's box.
template<class V>
__device__ __forceinline__ unsigned int cellKey(SegGeom const& S, V const& v, unsigned long long i,
                                                int* c, int& p)
{
    if (v.ic[0][i] == DEAD_CELL)
        return DEAD_KEY;
    unsigned int lin = 0;
    for (int d = 0; d < S.dim; ++d)
    {
        int const r = v.ic[d][i] - S.lo[d];
        if (r < 0 || r >= S.n[d])
            return static_cast<unsigned int>(S.ncells);
        lin = lin * static_cast<unsigned int>(S.n[d]) + static_cast<unsigned int>(r);
    }
    p = S.owner[lin];
    if (p < 0)
        return static_cast<unsigned int>(S.ncells);
    SegPatch const& P = S.patch[p];
    unsigned int key  = 0;
    for (int d = 0; d < S.dim; ++d)
    {
        c[d] = v.ic[d][i] - P.lo[d];
        key  = key * static_cast<unsigned int>(P.n[d]) + static_cast<unsigned int>(c[d]);
    }
    return P.cellBase + key;
}

// segment and lane of cell c (relative to patch p's box)
__device__ __forceinline__ void segLane(SegGeom const& S, int p, int const* c, unsigned int& seg,
                                        int& lane)
{
    SegPatch const& P = S.patch[p];
    unsigned int s    = 0;
    int l             = 0;
    for (int d = 0; d < S.dim; ++d)
    {
        int const q = c[d] / P.seg[d];
        s           = s * static_cast<unsigned int>(P.nseg[d]) + static_cast<unsigned int>(q);
        l           = l * P.seg[d] + (c[d] - q * P.seg[d]);
    }
    seg  = P.segBase + s;
    lane = l;
}

// key of lane l's cell in segment s, or DEAD_KEY where the segment overhangs its patch's box
__device__ __forceinline__ unsigned int laneKey(SegGeom const& S, unsigned int s, int l)
{
    SegPatch const& P = S.patch[S.segPatch[s]];
    int sidx[3], sub[3];
    unsigned int rem = s - P.segBase;
    int lr           = l;
    for (int d = S.dim - 1; d >= 0; --d)
    {
        sidx[d] = static_cast<int>(rem % static_cast<unsigned int>(P.nseg[d]));
        rem /= static_cast<unsigned int>(P.nseg[d]);
        sub[d] = lr % P.seg[d];
        lr /= P.seg[d];
    }
    if (lr != 0)
        return DEAD_KEY; // lane beyond the segment's cells
    unsigned int key = 0;
    for (int d = 0; d < S.dim; ++d)
    {
        int const c = sidx[d] * P.seg[d] + sub[d];
        if (c >= P.n[d])
            return DEAD_KEY;
        key = key * static_cast<unsigned int>(P.n[d]) + static_cast<unsigned int>(c);
    }
    return P.cellBase + key;
}
} // namespace

__global__ __launch_bounds__(BLOCK) void k_seg_zero(unsigned int* __restrict__ p, unsigned int m)
{
    unsigned int const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < m)
        p[i] = 0u;
}

// particles per cell key, and the rank of each particle among those of its key (the order in
// which the atomics reached the count): the scatter places particle i at row rank[i] of its cell
template<class V>
__global__ __launch_bounds__(BLOCK) void k_seg_count(SegGeom S, V v, unsigned long long n,
                                                     unsigned int* __restrict__ counts,
                                                     unsigned int* __restrict__ rank)
{
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    int c[3], p = 0;
    unsigned int const key = cellKey(S, v, i, c, p);
    if (key != DEAD_KEY)
        rank[i] = atomicAdd(counts + key, 1u);
}

// rows of each segment, ceil(particles / 64), and its work items; the count of each cell becomes
// the cell's first position in the segment's cell order (its particles take the positions after,
// in rank order) and total[s] the segment's particle count. One wave per segment, lane l taking
// the segment's l-th cell; the overflow segment packs its particles 64 to a row.
__global__ __launch_bounds__(BLOCK) void k_seg_rows(SegGeom S, unsigned int* __restrict__ counts,
                                                    unsigned int* __restrict__ rows,
                                                    unsigned int* __restrict__ itemCount,
                                                    unsigned int* __restrict__ total)
{
    __shared__ unsigned int sc[BLOCK];
    unsigned int const s = blockIdx.x * (BLOCK / LANES) + threadIdx.x / LANES;
    int const lane       = static_cast<int>(threadIdx.x % LANES);
    unsigned int key     = DEAD_KEY;
    unsigned int cnt     = 0;
    if (s < static_cast<unsigned int>(S.nsegs))
    {
        key = laneKey(S, s, lane);
        cnt = key != DEAD_KEY ? counts[key] : 0u;
    }
    // inclusive scan of the counts over each wave's lanes
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = 1; o < LANES; o <<= 1)
    {
        unsigned int const a = lane >= o ? sc[threadIdx.x - o] : 0u;
        __syncthreads();
        sc[threadIdx.x] += a;
        __syncthreads();
    }
    unsigned int T = sc[(threadIdx.x & ~static_cast<unsigned int>(LANES - 1)) + LANES - 1];
    if (key != DEAD_KEY)
        counts[key] = sc[threadIdx.x] - cnt;
    if (s == static_cast<unsigned int>(S.nsegs))
        T = counts[S.ncells];
    if (lane == 0 && s <= static_cast<unsigned int>(S.nsegs))
    {
        unsigned int const R = (T + LANES - 1) / LANES;
        rows[s]              = R;
        unsigned int const k = static_cast<unsigned int>(S.rowsPerItem);
        itemCount[s]         = (R + k - 1) / k;
        total[s]             = T;
    }
}

// exclusive scans over the m = nsegs + 1 segments in one workgroup: first slots (64 per row;
// off[m] = all slots) and first work items (in place; *nItems = all items)
__global__ __launch_bounds__(1024) void k_seg_scan(unsigned int const* __restrict__ rows,
                                                   unsigned long long* __restrict__ off,
                                                   unsigned int* __restrict__ itemCount,
                                                   unsigned int* __restrict__ nItems, unsigned int m)
{
    __shared__ unsigned long long ps[1024];
    __shared__ unsigned int pi[1024];
    unsigned int const per = (m + 1023) / 1024;
    unsigned int const b0  = threadIdx.x * per;
    unsigned int const b1  = min(b0 + per, m);
    unsigned long long s   = 0;
    unsigned int it        = 0;
    for (unsigned int b = b0; b < b1; ++b)
    {
        s += static_cast<unsigned long long>(rows[b]) * LANES;
        it += itemCount[b];
    }
    ps[threadIdx.x] = s;
    pi[threadIdx.x] = it;
    __syncthreads();
    for (unsigned int o = 1; o < 1024; o <<= 1)
    {
        unsigned long long const as = threadIdx.x >= o ? ps[threadIdx.x - o] : 0ull;
        unsigned int const ai       = threadIdx.x >= o ? pi[threadIdx.x - o] : 0u;
        __syncthreads();
        ps[threadIdx.x] += as;
        pi[threadIdx.x] += ai;
        __syncthreads();
    }
    unsigned long long rs = threadIdx.x ? ps[threadIdx.x - 1] : 0ull;
    unsigned int ri       = threadIdx.x ? pi[threadIdx.x - 1] : 0u;
    for (unsigned int b = b0; b < b1; ++b)
    {
        unsigned long long const a = static_cast<unsigned long long>(rows[b]) * LANES;
        unsigned int const c       = itemCount[b];
        off[b]                     = rs;
        itemCount[b]               = ri;
        rs += a;
        ri += c;
    }
    if (threadIdx.x == 1023)
    {
        off[m]  = ps[1023];
        *nItems = pi[1023];
    }
}

__global__ __launch_bounds__(BLOCK) void k_seg_items(unsigned int const* __restrict__ rows,
                                                     unsigned int const* __restrict__ itemOff,
                                                     unsigned int* __restrict__ items, unsigned int m,
                                                     unsigned int rowsPerItem)
{
    unsigned int const s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= m)
        return;
    unsigned int const k = (rows[s] + rowsPerItem - 1) / rowsPerItem;
    for (unsigned int j = 0; j < k; ++j)
    {
        items[2 * (itemOff[s] + j)]     = s;
        items[2 * (itemOff[s] + j) + 1] = j; // rows of item j: itemRows
    }
}

constexpr unsigned int NO_PARTICLE = 0xFFFFFFFFu;

// the slot of source particle i, recorded as perm[slot] = i: position p of segment s (R rows) is
// row p % R, lane p / R, so that the lanes of a row are R positions apart in the segment's cell
// order; the overflow segment packs its particles in rank order
template<class V>
__global__ __launch_bounds__(BLOCK) void k_seg_perm(SegGeom S, V src, unsigned long long n,
                                                    unsigned int const* __restrict__ rank,
                                                    unsigned int const* __restrict__ first,
                                                    unsigned int const* __restrict__ rows,
                                                    unsigned long long const* __restrict__ off,
                                                    unsigned int* __restrict__ perm)
{
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    int c[3], pt = 0;
    unsigned int const key = cellKey(S, src, i, c, pt);
    if (key == DEAD_KEY)
        return;
    unsigned long long j;
    if (key < static_cast<unsigned int>(S.ncells))
    {
        unsigned int seg;
        int lane;
        segLane(S, pt, c, seg, lane);
        unsigned int const p = first[key] + rank[i];
        unsigned int const R = rows[seg];
        j = off[seg] + static_cast<unsigned long long>(p % R) * LANES + p / R;
    }
    else
        j = off[S.nsegs] + rank[i];
    perm[j] = static_cast<unsigned int>(i);
}

// the empty slots of every segment (one workgroup per segment): those past the segment's
// particle count in its position order (the overflow segment: in slot order)
__global__ __launch_bounds__(BLOCK) void k_seg_pad(SegGeom S, unsigned int const* __restrict__ rows,
                                                   unsigned long long const* __restrict__ off,
                                                   unsigned int const* __restrict__ total,
                                                   unsigned int* __restrict__ perm)
{
    unsigned int const s     = blockIdx.x;
    unsigned int const R     = rows[s];
    unsigned int const slots = R * LANES;
    unsigned int const T     = total[s];
    bool const packed        = s < static_cast<unsigned int>(S.nsegs);
    for (unsigned int x = threadIdx.x; x < slots; x += BLOCK)
    {
        unsigned int const p = packed ? (x % LANES) * R + x / LANES : x;
        if (p >= T)
            perm[off[s] + x] = NO_PARTICLE;
    }
}

// slot j of the sorted storage takes source particle perm[j] (empty slots: a dead particle);
// consecutive work-items write consecutive slots
template<class V>
__global__ __launch_bounds__(BLOCK) void k_seg_gather(V src, V dst, unsigned int const* __restrict__ perm,
                                                      unsigned long long slots, int dim)
{
    unsigned long long const j = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (j >= slots)
        return;
    unsigned int const i = perm[j];
    if (i == NO_PARTICLE)
    {
        dst.ic[0][j] = DEAD_CELL;
        return;
    }
    for (int d = 0; d < dim; ++d)
    {
        dst.ic[d][j] = src.ic[d][i];
        dst.de[d][j] = src.de[d][i];
    }
    dst.v[0][j] = src.v[0][i];
    dst.v[1][j] = src.v[1][i];
    dst.v[2][j] = src.v[2][i];
    dst.w[j]    = src.w[i];
    if (src.q)
        dst.q[j] = src.q[i];
}

// the first n slots of every array of src into dst (a sorted storage back into the array the
// captured steps address)
template<class V>
__global__ __launch_bounds__(BLOCK) void k_seg_copy(V src, V dst, unsigned long long n, int dim)
{
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    for (int d = 0; d < dim; ++d)
    {
        dst.ic[d][i] = src.ic[d][i];
        dst.de[d][i] = src.de[d][i];
    }
    dst.v[0][i] = src.v[0][i];
    dst.v[1][i] = src.v[1][i];
    dst.v[2][i] = src.v[2][i];
    dst.w[i]    = src.w[i];
    if (src.q)
        dst.q[i] = src.q[i];
}

template<class V>
static void segCountT(SegGeom const& sg, V const& src, unsigned long long n, unsigned int* work,
                      unsigned int* rank, SegTables const& tables, hipStream_t s)
{
    unsigned int const nc = static_cast<unsigned int>(sg.ncells) + 1;
    unsigned int const m  = static_cast<unsigned int>(sg.nsegs) + 1;
    unsigned int* counts    = work;
    unsigned int* itemCount = work + nc;
    unsigned int* total     = work + nc + m;
    constexpr unsigned int SEGS_PER_BLOCK = BLOCK / LANES;
    k_seg_zero<<<(nc + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(counts, nc);
    if (n)
        k_seg_count<V><<<static_cast<unsigned int>((n + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(sg, src, n,
                                                                                          counts, rank);
    k_seg_rows<<<(m + SEGS_PER_BLOCK - 1) / SEGS_PER_BLOCK, BLOCK, 0, s>>>(sg, counts, tables.rows,
                                                                           itemCount, total);
    k_seg_scan<<<1, 1024, 0, s>>>(tables.rows, tables.off, itemCount, tables.nItems, m);
}

template<class V>
static void segScatterT(SegGeom const& sg, V const& src, V const& dst, unsigned long long n,
                        unsigned long long slots, unsigned int* work, unsigned int const* rank,
                        unsigned int* perm, SegTables const& tables, int dim, hipStream_t s)
{
    unsigned int const nc = static_cast<unsigned int>(sg.ncells) + 1;
    unsigned int const m  = static_cast<unsigned int>(sg.nsegs) + 1;
    unsigned int* first     = work; // the counts, turned into first positions by k_seg_rows
    unsigned int* itemCount = work + nc;
    unsigned int* total     = work + nc + m;
    k_seg_items<<<(m + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(tables.rows, itemCount, tables.items, m,
                                                         static_cast<unsigned int>(sg.rowsPerItem));
    if (n)
        k_seg_perm<V><<<static_cast<unsigned int>((n + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(
            sg, src, n, rank, first, tables.rows, tables.off, perm);
    k_seg_pad<<<m, BLOCK, 0, s>>>(sg, tables.rows, tables.off, total, perm);
    if (slots)
        k_seg_gather<V><<<static_cast<unsigned int>((slots + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(
            src, dst, perm, slots, dim);
}

void launchSegCount(SegGeom const& sg, PView const& src, unsigned long long n, unsigned int* work,
                    unsigned int* rank, SegTables const& tables, hipStream_t s)
{
    segCountT(sg, src, n, work, rank, tables, s);
}
void launchSegCount(SegGeom const& sg, PView32 const& src, unsigned long long n, unsigned int* work,
                    unsigned int* rank, SegTables const& tables, hipStream_t s)
{
    segCountT(sg, src, n, work, rank, tables, s);
}
void launchSegScatter(SegGeom const& sg, PView const& src, PView const& dst, unsigned long long n,
                      unsigned long long slots, unsigned int* work, unsigned int const* rank,
                      unsigned int* perm, SegTables const& tables, int dim, hipStream_t s)
{
    segScatterT(sg, src, dst, n, slots, work, rank, perm, tables, dim, s);
}
void launchSegScatter(SegGeom const& sg, PView32 const& src, PView32 const& dst, unsigned long long n,
                      unsigned long long slots, unsigned int* work, unsigned int const* rank,
                      unsigned int* perm, SegTables const& tables, int dim, hipStream_t s)
{
    segScatterT(sg, src, dst, n, slots, work, rank, perm, tables, dim, s);
}

void launchSegCopy(PView const& src, PView const& dst, unsigned long long n, int dim, hipStream_t s)
{
    if (n)
        k_seg_copy<PView><<<static_cast<unsigned int>((n + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(src, dst, n,
                                                                                              dim);
}
void launchSegCopy(PView32 const& src, PView32 const& dst, unsigned long long n, int dim, hipStream_t s)
{
    if (n)
        k_seg_copy<PView32><<<static_cast<unsigned int>((n + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(src, dst,
                                                                                                n, dim);
}

SegPatch segPatchOf(int dim, int const* lo, int const* hi, int const* pref)
{
    // the preferred segment shape `pref` (per dimension; by default a run of 64 cells in 1D, a
    // 2x32 block in 2D, 32 cells along the fastest dimension so that a row's gather loads stay
    // in runs of consecutive nodes, and a 4x4x4 block in 3D: compact blocks keep the LDS tile with
    // its margins small), clipped to the box; the cells left over go to the slower dimensions
    SegPatch P{};
    int room = LANES;
    for (int d = 0; d < 3; ++d)
    {
        P.lo[d] = d < dim ? lo[d] : 0;
        P.n[d]  = d < dim ? hi[d] - lo[d] + 1 : 1;
        P.seg[d] = 1;
    }
    for (int d = dim - 1; d >= 0; --d)
    {
        int const want = d == 0 ? room : std::min(std::max(pref[d], 1), room);
        P.seg[d]       = std::max(1, std::min(want, P.n[d]));
        room /= P.seg[d];
    }
    for (int d = 0; d < 3; ++d)
        P.nseg[d] = (P.n[d] + P.seg[d] - 1) / P.seg[d];
    return P;
}

TileSpec segTileShape(int dim, int order, SegGeom const& sg, int copies)
{
    TileSpec T{};
    T.copies = copies;
    int const mlo = pc::segMarginLo(order);
    int const mhi = pc::segMarginHi(order);
    for (int d = 0; d < 3; ++d)
    {
        T.lo[d]  = d < dim ? mlo : 0;
        T.ext[d] = d < dim ? sg.maxSeg[d] + mlo + mhi : 1;
    }
    pc::tilePitch(T, dim, sg.rowPad);
    T.nodes  = T.ext[0] * T.pitch[1] * T.pitch[2];
    T.stride = (T.nodes + 15) / 16 * 16; // copies 16 words apart: neighbouring lanes' banks differ
    T.bytes  = static_cast<unsigned int>(8 * 5 * copies * T.stride);
    if (T.bytes > TILE_BYTES_MAX)
    {
        // no room: every deposit goes straight to HBM
        for (int d = 0; d < 3; ++d)
        {
            T.ext[d]   = 1;
            T.pitch[d] = 1;
        }
        T.lo[0]  = -0x10000000;
        T.nodes  = 1;
        T.stride = 16;
        T.bytes  = static_cast<unsigned int>(8 * 5 * copies * T.stride);
    }
    return T;
}


unsigned int segItemsCap(SegGeom const& sg, unsigned long long capacity)
{
    return static_cast<unsigned int>(capacity / (LANES * static_cast<unsigned long long>(sg.rowsPerItem)))
           + static_cast<unsigned int>(sg.nsegs)
           + 2u;
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
template<int D, int O>
static void launchPushT(PushArgs const& A, PView const& in, PView const& out, hipStream_t s)
{
    if (A.segOn)
    {
        // segmented arrays are pushed in place (out is in)
        if (A.rhoFromN)
            k_push_seg<D, O, true><<<A.st.itemsCap, pc::SEG_BLOCK, A.tile.bytes, s>>>(A, in);
        else
            k_push_seg<D, O, false><<<A.st.itemsCap, pc::SEG_BLOCK, A.tile.bytes, s>>>(A, in);
        return;
    }
    if (A.n == 0)
        return;
    constexpr int CH = Cfg<D, O>::CHUNK;
    unsigned int const grid = static_cast<unsigned int>((A.n + CH - 1) / CH);
    k_push<D, O><<<grid, BLOCK, A.tile.bytes, s>>>(A, in, out);
}

// particles per workgroup of the linear push kernel
template<int D, int O>
static int chunkOf()
{
    return Cfg<D, O>::CHUNK;
}

int chunkSize(int dim, int order)
{
    switch (dim * 10 + order)
    {
        case 11: return chunkOf<1, 1>();
        case 12: return chunkOf<1, 2>();
        case 13: return chunkOf<1, 3>();
        case 21: return chunkOf<2, 1>();
        case 22: return chunkOf<2, 2>();
        case 23: return chunkOf<2, 3>();
        case 31: return chunkOf<3, 1>();
        case 32: return chunkOf<3, 2>();
        case 33: return chunkOf<3, 3>();
    }
    return BLOCK;
}

int blockThreads() { return BLOCK; }

TileSpec tileShape(int dim, int order, double ppc, int copies)
{
    TileSpec T{};
    T.copies        = copies;
    int const slabs = 5 * copies;
    if (dim == 3)
    {
        // brick tile: the chunk's brick of BRICK^3 cells with the margins below (stencil from
        // cell - 1, one cell of motion) and above (stencil, one cell of motion)
        TileSpec B = T;
        B.brick    = BRICK;
        for (int d = 0; d < 3; ++d)
        {
            B.lo[d]    = 2;
            B.ext[d]   = BRICK + 2 + order + 2;
            B.pitch[d] = B.ext[d];
        }
        B.nodes  = B.ext[0] * B.ext[1] * B.ext[2];
        B.stride = (B.nodes + 15) / 16 * 16 + 1;
        B.bytes  = static_cast<unsigned int>(8 * slabs * B.stride);
        if (B.bytes <= TILE_BYTES_MAX)
            return B;
    }
    // nodes that fit the LDS budget with the bank padding of every slab
    int const cap   = static_cast<int>(TILE_BYTES_MAX / (8u * static_cast<unsigned>(slabs))) - 17;
    int const CHUNK = chunkSize(dim, order);
    // margins around the first particle's cell: the stencil (from cell - 1 for order >= 2) and
    // one cell of motion
    int const mlo = 2, mhi = order + 2;
    // a chunk of a cell-ordered array spans one cell across the slower dimensions and about
    // CHUNK/ppc cells along the fastest
    for (int d = 0; d < 3; ++d)
    {
        T.lo[d]  = d < dim ? mlo : 0;
        T.ext[d] = d < dim ? 1 + mlo + mhi : 1;
    }
    int const fast = dim - 1;
    double cells   = ppc > 0 ? CHUNK / ppc : CHUNK;
    int want       = static_cast<int>(cells) + 1 + mlo + mhi;
    int others     = 1;
    for (int d = 0; d < dim - 1; ++d)
        others *= T.ext[d];
    int const maxFast = cap / others;
    T.ext[fast]       = want < maxFast ? want : maxFast;
    for (int d = 0; d < 3; ++d)
        T.pitch[d] = T.ext[d];
    T.nodes           = T.ext[0] * T.ext[1] * T.ext[2];
    // copies one word past a multiple of 16: lanes of one cell adding into different copies of
    // a node hit different banks
    T.stride          = (T.nodes + 15) / 16 * 16 + 1;
    T.bytes           = static_cast<unsigned int>(8 * slabs * T.stride);
    return T;
}

template<int D, int O>
static void launchDepositT(PushArgs const& A, PView const& in, hipStream_t s)
{
    PushArgs B = A;
    B.pushOn   = 0;
    launchPushT<D, O>(B, in, in, s);
}

template<int D, int O>
static void launchPushOnT(PushArgs const& A, PView const& in, PView const& out, hipStream_t s)
{
    PushArgs B = A;
    B.pushOn   = 1;
    launchPushT<D, O>(B, in, out, s);
}

#define PHX_DISPATCH_DO(FN, ...)                                                                   \
    switch (A.G.dim * 10 + A.G.order)                                                              \
    {                                                                                              \
        case 11: FN<1, 1>(__VA_ARGS__); break;                                                     \
        case 12: FN<1, 2>(__VA_ARGS__); break;                                                     \
        case 13: FN<1, 3>(__VA_ARGS__); break;                                                     \
        case 21: FN<2, 1>(__VA_ARGS__); break;                                                     \
        case 22: FN<2, 2>(__VA_ARGS__); break;                                                     \
        case 23: FN<2, 3>(__VA_ARGS__); break;                                                     \
        case 31: FN<3, 1>(__VA_ARGS__); break;                                                     \
        case 32: FN<3, 2>(__VA_ARGS__); break;                                                     \
        case 33: FN<3, 3>(__VA_ARGS__); break;                                                     \
    }

void launchPush(PushArgs const& A, PView const& in, PView const& out, hipStream_t s)
{
    PHX_DISPATCH_DO(launchPushOnT, A, in, out, s)
}

void launchDeposit(PushArgs const& A, PView const& in, hipStream_t s)
{
    PHX_DISPATCH_DO(launchDepositT, A, in, s)
}

unsigned int compactBlocks(unsigned long long n)
{
    return static_cast<unsigned int>((n + BLOCK - 1) / BLOCK);
}

void launchCompact(PView const& src, PView const& dst, uint8_t const* status, unsigned long long n,
                   uint8_t want, unsigned int* blockCounts, unsigned long long* total,
                   unsigned long long dstOffset, int dim, Wrap const& wrap, hipStream_t s)
{
    if (n == 0)
        return;
    unsigned int nb = compactBlocks(n);
    k_count_status<<<nb, BLOCK, 0, s>>>(status, n, want, blockCounts);
    k_scan_counts<<<1, 1024, 0, s>>>(blockCounts, nb, total);
    k_scatter_status<<<nb, BLOCK, 0, s>>>(src, dst, status, n, want, blockCounts, dstOffset, dim,
                                          wrap);
}

// tile of the workgroup-per-segment push: the largest segment plus the stencil and one cell of
// motion below (2 nodes) and above (order + 2 nodes), `slabs` moment slabs, the pitches of
// dimensions 1 and 2 padded to a multiple of `pad` nodes
TileSpec segwgTileShape(int dim, int order, SegGeom const& sg, int slabs, int pad)
{
    TileSpec T{};
    T.copies = 1;
    for (int d = 0; d < 3; ++d)
    {
        T.lo[d]    = d < dim ? 2 : 0;
        T.ext[d]   = d < dim ? sg.maxSeg[d] + 2 + order + 2 : 1;
        T.pitch[d] = T.ext[d];
    }
    auto size = [&]() {
        T.nodes  = T.ext[0] * T.pitch[1] * T.pitch[2];
        T.stride = (T.nodes + 15) / 16 * 16;
        T.bytes  = static_cast<unsigned int>(8 * slabs * T.stride);
    };
    if (dim == 3 && pad > 1)
    {
        // pitches of dimensions 1 and 2 padded to a multiple of pad nodes (the padding nodes are
        // never deposited into: the flush skips them as zeros); dropped if the tile outgrows LDS
        for (int d = 1; d < 3; ++d)
            T.pitch[d] = (T.ext[d] + pad - 1) / pad * pad;
        size();
        if (T.bytes > TILE_BYTES_MAX)
            for (int d = 1; d < 3; ++d)
                T.pitch[d] = T.ext[d];
    }
    size();
    if (T.bytes > TILE_BYTES_MAX)
    {
        // no room: every deposit goes straight to HBM
        for (int d = 0; d < 3; ++d)
        {
            T.ext[d]   = 1;
            T.pitch[d] = 1;
        }
        T.lo[0]  = -0x10000000;
        T.nodes  = 1;
        T.stride = 16;
        T.bytes  = static_cast<unsigned int>(8 * slabs * T.stride);
    }
    return T;
}

void launchPushSegWG(PushArgs const& A, PView const& in, hipStream_t s)
{
    unsigned int const grid = static_cast<unsigned int>(A.sg.nsegs) + 1u;
    switch (A.G.dim * 10 + A.G.order)
    {
        case 31: A.rhoFromN ? k_push_segwg<3, 1, true><<<grid, BLOCK, A.tile.bytes, s>>>(A, in)
                            : k_push_segwg<3, 1, false><<<grid, BLOCK, A.tile.bytes, s>>>(A, in); break;
        case 32: A.rhoFromN ? k_push_segwg<3, 2, true><<<grid, BLOCK, A.tile.bytes, s>>>(A, in)
                            : k_push_segwg<3, 2, false><<<grid, BLOCK, A.tile.bytes, s>>>(A, in); break;
        case 33: A.rhoFromN ? k_push_segwg<3, 3, true><<<grid, BLOCK, A.tile.bytes, s>>>(A, in)
                            : k_push_segwg<3, 3, false><<<grid, BLOCK, A.tile.bytes, s>>>(A, in); break;
        default: break;
    }
}

// ---------------------------------------------------------------------------------------------
// drop-in Pusher::move and Interpolator gather over a whole particle array
// ---------------------------------------------------------------------------------------------
namespace
{
__device__ __forceinline__ bool inAny(int const* c, SelBoxes const& S, int D)
{
    for (int b = 0; b < S.n; ++b)
        if (inBox(c, S.lo[b], S.hi[b], D))
            return true;
    return false;
}
} // namespace

// BorisPusher::move (boris.hpp:93-138) in place, one thread per slot: prePush of every particle;
// gather, Boris and postPush of those the first selector keeps (S.n == 0: every particle); the
// fate of each slot for the partition that follows: 0 kept by both selectors, 1 by the first
// only, 2 by neither, 3 an empty slot
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_move(PushArgs A, PView v, SelBoxes first, SelBoxes second,
                                                uint8_t* __restrict__ status)
{
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= A.n)
        return;
    Geom const& G = A.G;
    Part<D> P;
    loadPart<D>(v, i, P);
    if (P.cell[0] == DEAD_CELL)
    {
        status[i] = 3;
        return;
    }
    int cell[D];
    double delta[D], vel[3] = {P.v[0], P.v[1], P.v[2]};
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        cell[d]  = P.cell[d];
        delta[d] = P.delta[d];
    }
    double bd, bv;
    status[i] = 2;
    if (!advancePosition<D>(cell, delta, vel, A.hdtdx, bd, bv))
    {
        raiseError(A.err, 3u, 1u, i, bd, bv, A.kind, A.pop);
        return;
    }
    bool const sel1 = first.n == 0 || inAny(cell, first, D);
    bool sel2       = false;
    if (sel1)
    {
        Stencil<D, O> s;
        makeStencil<D, O>(G, cell, delta, s);
        if (!stencilInside<D, O>(G, s, QEX) || !stencilInside<D, O>(G, s, QBX)
            || !stencilInside<D, O>(G, s, QEY) || !stencilInside<D, O>(G, s, QBY)
            || !stencilInside<D, O>(G, s, QEZ) || !stencilInside<D, O>(G, s, QBZ))
        {
            raiseError(A.err, 4u, 2u, i, delta[0], vel[0], A.kind, A.pop);
            return;
        }
        double E[3], B[3];
        gatherEB<D, O>(G, A.em, s, E, B);
        borisAccelerate(vel, E, B, P.q, A.dto2m);
        if (!advancePosition<D>(cell, delta, vel, A.hdtdx, bd, bv))
        {
            raiseError(A.err, 3u, 2u, i, bd, bv, A.kind, A.pop);
            return;
        }
        sel2 = second.n == 0 || inAny(cell, second, D);
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        v.ic[d][i] = cell[d];
        v.de[d][i] = delta[d];
    }
    v.v[0][i] = vel[0];
    v.v[1][i] = vel[1];
    v.v[2][i] = vel[2];
    status[i] = sel2 ? 0 : (sel1 ? 1 : 2);
}

// Interpolator::operator()(particle, em, layout) (interpolator.hpp:420-456): E and B at every
// live slot, written as [3 * slot + component]
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_gather(PushArgs A, PView v, double* __restrict__ Eo,
                                                  double* __restrict__ Bo)
{
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= A.n)
        return;
    Part<D> P;
    loadPart<D>(v, i, P);
    if (P.cell[0] == DEAD_CELL)
        return;
    Stencil<D, O> s;
    makeStencil<D, O>(A.G, P.cell, P.delta, s);
    if (!stencilInside<D, O>(A.G, s, QEX) || !stencilInside<D, O>(A.G, s, QBX)
        || !stencilInside<D, O>(A.G, s, QEY) || !stencilInside<D, O>(A.G, s, QBY)
        || !stencilInside<D, O>(A.G, s, QEZ) || !stencilInside<D, O>(A.G, s, QBZ))
    {
        raiseError(A.err, 4u, 2u, i, P.delta[0], P.v[0], A.kind, A.pop);
        return;
    }
    double E[3], B[3];
    gatherEB<D, O>(A.G, A.em, s, E, B);
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
        Eo[3 * i + c] = E[c];
        Bo[3 * i + c] = B[c];
    }
}

template<int D, int O>
static void launchMoveT(PushArgs const& A, PView const& v, SelBoxes const& f, SelBoxes const& s2,
                        uint8_t* status, hipStream_t s)
{
    if (A.n)
        k_move<D, O><<<compactBlocks(A.n), BLOCK, 0, s>>>(A, v, f, s2, status);
}

template<int D, int O>
static void launchGatherT(PushArgs const& A, PView const& v, double* E, double* B, hipStream_t s)
{
    if (A.n)
        k_gather<D, O><<<compactBlocks(A.n), BLOCK, 0, s>>>(A, v, E, B);
}

void launchMove(PushArgs const& A, PView const& v, SelBoxes const& first, SelBoxes const& second,
                uint8_t* status, hipStream_t s)
{
    PHX_DISPATCH_DO(launchMoveT, A, v, first, second, status, s)
}

void launchGather(PushArgs const& A, PView const& v, double* E, double* B, hipStream_t s)
{
    PHX_DISPATCH_DO(launchGatherT, A, v, E, B, s)
}

// ---------------------------------------------------------------------------------------------
// gather without allocation: the same E/B values as launchGather, compacted to the live slots in
// storage order, staged in a file-scope device array instead of a library allocation
// ---------------------------------------------------------------------------------------------
namespace
{
constexpr unsigned long long NA_BLOCKS = (NOALLOC_GATHER_MAX_SLOTS + BLOCK - 1) / BLOCK;
// layout: E [3 x max], B [3 x max] (doubles), the total (one 64-bit word), the live count of each
// workgroup (NA_BLOCKS + 1 words)
constexpr size_t NA_E_OFF     = 0;
constexpr size_t NA_B_OFF     = NA_E_OFF + 3 * 8 * NOALLOC_GATHER_MAX_SLOTS;
constexpr size_t NA_TOTAL_OFF = NA_B_OFF + 3 * 8 * NOALLOC_GATHER_MAX_SLOTS;
constexpr size_t NA_COUNT_OFF = NA_TOTAL_OFF + 256;
constexpr size_t NA_BYTES     = NA_COUNT_OFF + 4 * (NA_BLOCKS + 1);

// the scratch of phx_gather_noalloc (every GPU the library runs on holds one until the process
// ends); host code must use the address hipGetSymbolAddress returns for the current device, never
// this name
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// live slots of each workgroup's range
__global__ __launch_bounds__(BLOCK) void k_count_live(int32_t const* __restrict__ ic0, unsigned long long n,
                                                      unsigned int* __restrict__ blockCounts)
{
    __shared__ unsigned int s;
    if (threadIdx.x == 0)
        s = 0;
    __syncthreads();
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    bool const f               = i < n && ic0[i] != DEAD_CELL;
    unsigned long long const m = __ballot(f);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(&s, static_cast<unsigned int>(__popcll(m)));
    __syncthreads();
    if (threadIdx.x == 0)
        blockCounts[blockIdx.x] = s;
}

// E and B of every live slot at its rank among the live slots (offsets: exclusive scan of the
// workgroup counts)
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_gather_compact(PushArgs A, PView v,
                                                          unsigned int const* __restrict__ offsets,
                                                          double* __restrict__ Eo, double* __restrict__ Bo)
{
    __shared__ unsigned int waveBase[BLOCK / 64];
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    Part<D> P;
    P.cell[0] = DEAD_CELL;
    if (i < A.n)
        loadPart<D>(v, i, P);
    bool const f               = P.cell[0] != DEAD_CELL;
    unsigned long long const m = __ballot(f);
    int const lane             = threadIdx.x & 63;
    int const wave             = threadIdx.x >> 6;
    if (lane == 0)
        waveBase[wave] = static_cast<unsigned int>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned int run = 0;
        for (int w = 0; w < BLOCK / 64; ++w)
        {
            unsigned int const c = waveBase[w];
            waveBase[w]          = run;
            run += c;
        }
    }
    __syncthreads();
    if (!f)
        return;
    unsigned long long const below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    unsigned long long const j
        = offsets[blockIdx.x] + waveBase[wave] + static_cast<unsigned int>(__popcll(below));
    Stencil<D, O> s;
    makeStencil<D, O>(A.G, P.cell, P.delta, s);
    if (!stencilInside<D, O>(A.G, s, QEX) || !stencilInside<D, O>(A.G, s, QBX)
        || !stencilInside<D, O>(A.G, s, QEY) || !stencilInside<D, O>(A.G, s, QBY)
        || !stencilInside<D, O>(A.G, s, QEZ) || !stencilInside<D, O>(A.G, s, QBZ))
    {
        raiseError(A.err, 4u, 2u, i, P.delta[0], P.v[0], A.kind, A.pop);
        return;
    }
    double E[3], B[3];
    gatherEB<D, O>(A.G, A.em, s, E, B);
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
        Eo[3 * j + c] = E[c];
        Bo[3 * j + c] = B[c];
    }
}

template<int D, int O>
static void launchGatherCompactT(PushArgs const& A, PView const& v, unsigned int const* offsets,
                                 double* E, double* B, hipStream_t s)
{
    k_gather_compact<D, O><<<compactBlocks(A.n), BLOCK, 0, s>>>(A, v, offsets, E, B);
}

// device address of `scratch` per device, looked up once per device (with that device current)
std::mutex naMutex;
void* naAddr[64] = {};
} // namespace

hipError_t noAllocScratch(void** out)
{
    *out      = nullptr;
    int dev   = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= 64)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(naMutex);
    if (!naAddr[dev])
    {
        void* p = nullptr;
        e       = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        naAddr[dev] = p;
    }
    *out = naAddr[dev];
    return hipSuccess;
}

void launchGatherNoAlloc(PushArgs const& A, PView const& v, void* scratchPtr, hipStream_t s)
{
    if (A.n == 0 || A.n > NOALLOC_GATHER_MAX_SLOTS)
        return;
    unsigned char* base   = static_cast<unsigned char*>(scratchPtr);
    double* E             = reinterpret_cast<double*>(base + NA_E_OFF);
    double* B             = reinterpret_cast<double*>(base + NA_B_OFF);
    auto* total           = reinterpret_cast<unsigned long long*>(base + NA_TOTAL_OFF);
    auto* counts          = reinterpret_cast<unsigned int*>(base + NA_COUNT_OFF);
    unsigned int const nb = compactBlocks(A.n);
    k_count_live<<<nb, BLOCK, 0, s>>>(v.ic[0], A.n, counts);
    k_scan_counts<<<1, 1024, 0, s>>>(counts, nb, total);
    PHX_DISPATCH_DO(launchGatherCompactT, A, v, counts, E, B, s)
}

void noAllocGatherOutputs(void* scratchPtr, double const** E, double const** B,
                          unsigned long long const** total)
{
    unsigned char const* base = static_cast<unsigned char const*>(scratchPtr);
    *E     = reinterpret_cast<double const*>(base + NA_E_OFF);
    *B     = reinterpret_cast<double const*>(base + NA_B_OFF);
    *total = reinterpret_cast<unsigned long long const*>(base + NA_TOTAL_OFF);
}

} // namespace phx
