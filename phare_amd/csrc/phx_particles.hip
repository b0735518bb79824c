// phx_particles.hip -- particle kernels for gfx950: fused prePush + gather + Boris + postPush +
// moment deposit (IonUpdater::updatePopulations, ion_updater.hpp:88-295, with the Boris pusher
// boris.hpp:93-138 and the Interpolator deposit interpolator.hpp:468-510), stable compaction of
// particle arrays, and the periodic re-entry of leaving particles.
//
// Particles are structure-of-arrays in HBM (int32 iCell[D], f64 delta[D], f64 v[3], f64 weight,
// f64 charge), each component a contiguous stream. A workgroup owns a chunk of 1024 consecutive
// particles (kept close to cell order by the loader), accumulates their moments in an LDS tile
// spanning the chunk's deposit footprint, then flushes the tile to the moment arrays. Lanes take
// particles with a stride across the chunk so that lanes of one wavefront fall into different
// cells and the LDS accumulations of a wavefront rarely hit the same node.
#include "phx_kernels.hpp"

namespace phx
{

namespace
{
constexpr int BLOCK = 256;
constexpr int NW    = BLOCK / 64; // wavefronts per workgroup

// Per (dim, order): whether the deposit is accumulated in registers over a union window, the
// particles per thread and the window. With accumulation, each thread owns PPT consecutive
// particles of the caller's order (mostly one cell) and scatters to LDS only when its cell
// changes. Only 1D keeps the window (5 moments x 4 nodes) in registers without losing
// occupancy; 2D/3D scatter every particle to the LDS tile.
template<int D, int O>
struct Cfg
{
    static constexpr bool ACC  = D == 1;
    static constexpr int PPT   = ACC ? 8 : 2;
    static constexpr int CHUNK = BLOCK * PPT;
    static constexpr int W     = O == 1 ? 2 : 4;  // window width per direction
    static constexpr int WB    = O == 1 ? 0 : 1;  // window base = local cell - WB
    static constexpr int WN    = D == 1 ? W : (D == 2 ? W * W : W * W * W);
};

// LDS moment tile: 5 moments x TILE nodes (20 KB in 1D/2D, 40 KB in 3D)
template<int D>
struct Tile
{
    static constexpr int NODES = D == 3 ? 1024 : 512;
};

__device__ __forceinline__ bool inBox(int const* c, int const* lo, int const* hi, int D)
{
    for (int d = 0; d < D; ++d)
        if (c[d] < lo[d] || c[d] > hi[d])
            return false;
    return true;
}

__device__ __forceinline__ bool inNLGB(int const* c, Boxes const& b, int D)
{
    for (int i = 0; i < b.nnlgb; ++i)
        if (inBox(c, b.nlgb_lo[i], b.nlgb_hi[i], D))
            return true;
    return false;
}

// (thread, step) -> slot in the chunk: consecutive threads read consecutive slots (coalesced).
// Arrays are stored chunk-interleaved (chunkInterleave on the host): slot k*BLOCK + t holds the
// particle (r*PPT + k) of the caller's order, r = lane*NW + wave. Each thread thus owns PPT
// consecutive particles, and the lanes of a wavefront own runs spread across the chunk.
__device__ __forceinline__ int chunkOffset(int t, int k) { return k * BLOCK + t; }

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

// scatter-add the 5 moments of one particle: into the LDS tile when its whole stencil lies in
// the tile, else straight into the HBM moment arrays
template<int D, int O>
__device__ __forceinline__ void depositOne(DepositData<D, O> const& dd, Geom const& G, double* tile,
                                           int const* torg, int const* text, double* const* mom)
{
    constexpr int S  = O + 1;
    constexpr int SY = D > 1 ? S : 1;
    constexpr int SZ = D > 2 ? S : 1;
    constexpr int TN = Tile<D>::NODES;
    bool inTile      = true;
#pragma unroll
    for (int d = 0; d < D; ++d)
        inTile = inTile && dd.start[d] >= torg[d] && dd.start[d] + O < torg[d] + text[d];
    if (inTile)
    {
#pragma unroll
        for (int a = 0; a < S; ++a)
#pragma unroll
            for (int b = 0; b < SY; ++b)
#pragma unroll
                for (int c = 0; c < SZ; ++c)
                {
                    int i    = dd.start[0] + a - torg[0];
                    int j    = D > 1 ? dd.start[1] + b - torg[1] : 0;
                    int k    = D > 2 ? dd.start[2] + c - torg[2] : 0;
                    int node = (i * text[1] + j) * text[2] + k;
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        atomicAdd(tile + m * TN + node, nodeWeight<D, O>(dd, dd.val[m], a, b, c));
                }
    }
    else
    {
        for (int a = 0; a < S; ++a)
            for (int b = 0; b < SY; ++b)
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

// scatter the window: nodes inside the LDS tile by LDS atomics, others straight to HBM;
// never-touched nodes hold exactly zero and are skipped
template<int D, int O>
__device__ __forceinline__ void accFlush(Acc<D, O> const& A, Geom const& G, double* tile, int const* torg,
                                         int const* text, double* const* mom)
{
    constexpr int W  = Acc<D, O>::W;
    constexpr int TN = Tile<D>::NODES;
    if (!A.live)
        return;
#pragma unroll
    for (int x = 0; x < Acc<D, O>::WN; ++x)
    {
        int const a = D == 1 ? x : (D == 2 ? x / W : x / (W * W));
        int const b = D == 1 ? 0 : (D == 2 ? x % W : (x / W) % W);
        int const c = D == 3 ? x % W : 0;
        int gi = A.base[0] + a;
        int gj = D > 1 ? A.base[1] + b : 0;
        int gk = D > 2 ? A.base[2] + c : 0;
        bool inTile = gi >= torg[0] && gi < torg[0] + text[0];
        if (D > 1)
            inTile = inTile && gj >= torg[1] && gj < torg[1] + text[1];
        if (D > 2)
            inTile = inTile && gk >= torg[2] && gk < torg[2] + text[2];
        int const tnode = ((gi - torg[0]) * text[1] + (D > 1 ? gj - torg[1] : 0)) * text[2]
                          + (D > 2 ? gk - torg[2] : 0);
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            double const v = A.a[m][x];
            if (v != 0.)
            {
                if (inTile)
                    atomicAdd(tile + m * TN + tnode, v);
                else
                    atomicAdd(mom[m] + G.idx(QMOM, gi, gj, gk), v);
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

} // namespace

// ---------------------------------------------------------------------------------------------
// fused push + deposit (and deposit-only for unpushed arrays)
// ---------------------------------------------------------------------------------------------
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_push(PushArgs A, PView in, PView out)
{
    using C                = Cfg<D, O>;
    constexpr int TN       = Tile<D>::NODES;
    constexpr int PPT      = C::PPT;
    constexpr int CHUNK    = C::CHUNK;
    __shared__ double tile[5 * TN];
    __shared__ int org[3];

    Geom const& G = A.G;
    int const t   = threadIdx.x;
    unsigned long long const base = static_cast<unsigned long long>(blockIdx.x) * CHUNK;
    Acc<D, O> acc;
    if constexpr (C::ACC)
        accZero(acc);

    for (int x = t; x < 5 * TN; x += BLOCK)
        tile[x] = 0.;
    if (t == 0)
    {
        // the tile is anchored at the chunk's first particle (particles are kept close to
        // cell order), offset by the host-chosen margins
        int c0 = in.ic[0][base];
        for (int d = 0; d < 3; ++d)
            org[d] = d < D ? (c0 == DEAD_CELL ? -0x20000000
                                              : in.ic[d][base] - (G.lower[d] - G.g) - A.tileLo[d])
                           : 0;
    }
    __syncthreads();
    int torg[3] = {org[0], org[1], org[2]};

    for (int k = 0; k < PPT; ++k)
    {
        unsigned long long const i = base + chunkOffset(t, k);
        if (i >= A.n)
            continue;
        int cell[D];
        double delta[D], v[3];
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            cell[d]  = in.ic[d][i];
            delta[d] = in.de[d][i];
        }
        if (cell[0] == DEAD_CELL)
            continue;
        v[0] = in.v[0][i];
        v[1] = in.v[1][i];
        v[2] = in.v[2][i];
        double const w = in.w[i], q = in.q[i];

        bool deposit = true, sel1 = true, inDom = true;
        uint8_t st   = ST_KEEP;
        if (A.pushOn)
        {
            double bd, bv;
            // prePushStep_ (boris.hpp:180-216)
            if (!advancePosition<D>(cell, delta, v, A.hdtdx, bd, bv))
            {
                raiseError(A.err, 3u, 1u, i, bd, bv, A.kind, A.pop);
                continue;
            }
            // first selector: noop (domain) or inGhostBox (level ghosts)
            sel1 = A.kind == KIND_DOMAIN ? true : inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
            if (sel1)
            {
                Stencil<D, O> s;
                makeStencil<D, O>(G, cell, delta, s);
                if (!stencilInside<D, O>(G, s, QEX) || !stencilInside<D, O>(G, s, QBX)
                    || !stencilInside<D, O>(G, s, QEY) || !stencilInside<D, O>(G, s, QBY)
                    || !stencilInside<D, O>(G, s, QEZ) || !stencilInside<D, O>(G, s, QBZ))
                {
                    raiseError(A.err, 4u, 2u, i, delta[0], v[0], A.kind, A.pop);
                    continue;
                }
                double E[3], B[3];
                E[0] = interp<D, O>(G, A.em.E[0], QEX, s);
                E[1] = interp<D, O>(G, A.em.E[1], QEY, s);
                E[2] = interp<D, O>(G, A.em.E[2], QEZ, s);
                B[0] = interp<D, O>(G, A.em.B[0], QBX, s);
                B[1] = interp<D, O>(G, A.em.B[1], QBY, s);
                B[2] = interp<D, O>(G, A.em.B[2], QBZ, s);
                borisAccelerate(v, E, B, q, A.dto2m);
                // postPushStep_ (boris.hpp:218-234)
                if (!advancePosition<D>(cell, delta, v, A.hdtdx, bd, bv))
                {
                    raiseError(A.err, 3u, 2u, i, bd, bv, A.kind, A.pop);
                    continue;
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
                        accFlush<D, O>(acc, G, tile, torg, A.tileExt, A.mom);
                        accZero(acc);
#pragma unroll
                        for (int d = 0; d < D; ++d)
                            acc.base[d] = cell[d] - (G.lower[d] - G.g) - C::WB;
                        acc.live = true;
                    }
                    accAdd<D, O>(acc, dd);
                }
                else
                    depositOne<D, O>(dd, G, tile, torg, A.tileExt, A.mom);
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
            if (out.w != in.w)
            {
                out.w[i] = w;
                out.q[i] = q;
            }
            if (A.status)
                A.status[i] = st;
        }
    }
    if constexpr (C::ACC)
        accFlush<D, O>(acc, G, tile, torg, A.tileExt, A.mom);
    __syncthreads();

    // flush the tile: one HBM atomic per touched node and moment
    int const vol = A.tileExt[0] * A.tileExt[1] * A.tileExt[2];
    for (int x = t; x < vol; x += BLOCK)
    {
        int const e12 = A.tileExt[1] * A.tileExt[2];
        int const ti  = x / e12;
        int const r   = x - ti * e12;
        int const tj  = r / A.tileExt[2];
        int const tk  = r - tj * A.tileExt[2];
        int node      = -1;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            double const val = tile[m * TN + x];
            if (val != 0.)
            {
                if (node < 0)
                    node = G.idx(QMOM, torg[0] + ti, D > 1 ? torg[1] + tj : 0, D > 2 ? torg[2] + tk : 0);
                atomicAdd(A.mom[m] + node, val);
            }
        }
    }
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
    dst.q[j]    = src.q[i];
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
template<int D, int O>
static void launchPushT(PushArgs const& A, PView const& in, PView const& out, hipStream_t s)
{
    if (A.n == 0)
        return;
    PushArgs B = A;
    B.pushOn   = 1;
    constexpr int CH = Cfg<D, O>::CHUNK;
    unsigned int const grid = static_cast<unsigned int>((A.n + CH - 1) / CH);
    k_push<D, O><<<grid, BLOCK, 0, s>>>(B, in, out);
}

// LDS tile extents for a level: the chunk of CHUNK consecutive particles spans about
// CHUNK/ppc cells along the fastest dimension; margins cover the stencil and one cell of motion
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

void tileShape(int dim, int order, double ppc, int* lo, int* ext)
{
    int const cap   = dim == 3 ? Tile<3>::NODES : Tile<2>::NODES;
    int const CHUNK = chunkSize(dim, order);
    int const mlo = 2, mhi = order + 2;
    for (int d = 0; d < 3; ++d)
    {
        lo[d]  = d < dim ? mlo : 0;
        ext[d] = d < dim ? 1 + mlo + mhi : 1;
    }
    int fast     = dim - 1;
    double cells = ppc > 0 ? CHUNK / ppc : CHUNK;
    int want     = static_cast<int>(cells) + 2 + mlo + mhi;
    int others   = 1;
    for (int d = 0; d < dim - 1; ++d)
        others *= ext[d];
    int maxFast = cap / others;
    ext[fast]   = want < maxFast ? want : maxFast;
}

template<int D, int O>
static void launchDepositT(PushArgs const& A, PView const& in, hipStream_t s)
{
    if (A.n == 0)
        return;
    PushArgs B = A;
    B.pushOn   = 0;
    constexpr int CH = Cfg<D, O>::CHUNK;
    unsigned int const grid = static_cast<unsigned int>((A.n + CH - 1) / CH);
    k_push<D, O><<<grid, BLOCK, 0, s>>>(B, in, in);
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
    PHX_DISPATCH_DO(launchPushT, A, in, out, s)
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

} // namespace phx
