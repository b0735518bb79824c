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
constexpr int PPT   = 4; // particles per thread
constexpr int CHUNK = BLOCK * PPT;
constexpr int TILE_NODES = 1536; // LDS tile capacity in nodes (x5 moments x8 B = 60 KB)

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

// map of (thread, iteration) -> particle offset inside the chunk: lane l of wave w at step k
// takes l*16 + w*4 + k, so one wavefront spans the whole chunk
__device__ __forceinline__ int chunkOffset(int t, int k)
{
    int lane = t & 63, wave = t >> 6;
    return lane * 16 + wave * 4 + k;
}

template<int D, int O>
struct DepositData
{
    int start[D];
    double w[D][O + 1];
    double val[5];
    bool on;
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

// contribution of one particle to one node (a,b,c) of its stencil, in the reference's
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

template<int D, int O>
__device__ __forceinline__ void depositTo(DepositData<D, O> const& dd, double* const* dst,
                                          int const* origin, int const* ext, bool lds)
{
    constexpr int S  = O + 1;
    constexpr int SY = D > 1 ? S : 1;
    constexpr int SZ = D > 2 ? S : 1;
    for (int a = 0; a < S; ++a)
        for (int b = 0; b < SY; ++b)
            for (int c = 0; c < SZ; ++c)
            {
                int i = dd.start[0] + a - origin[0];
                int j = D > 1 ? dd.start[1] + b - origin[1] : 0;
                int k = D > 2 ? dd.start[2] + c - origin[2] : 0;
                int node = (i * ext[1] + j) * ext[2] + k;
#pragma unroll
                for (int m = 0; m < 5; ++m)
                {
                    double contrib = nodeWeight<D, O>(dd, dd.val[m], a, b, c);
                    if (lds)
                        atomicAdd(dst[0] + m * TILE_NODES + node, contrib);
                    else
                        atomicAdd(dst[m] + node, contrib);
                }
            }
}

} // namespace

// ---------------------------------------------------------------------------------------------
// fused push + deposit
// ---------------------------------------------------------------------------------------------
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_push(PushArgs A, PView in, PView out)
{
    constexpr int S = O + 1;
    __shared__ double tile[5 * TILE_NODES];
    __shared__ int bb[6];

    Geom const& G = A.G;
    int const t   = threadIdx.x;
    if (t < 3)
    {
        bb[t]     = 0x7fffffff;
        bb[3 + t] = -0x7fffffff;
    }
    __syncthreads();
    unsigned long long const base = static_cast<unsigned long long>(blockIdx.x) * CHUNK;

    DepositData<D, O> dd[PPT];

#pragma unroll
    for (int k = 0; k < PPT; ++k)
    {
        dd[k].on                   = false;
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

        double bd, bv;
        // prePushStep_ (boris.hpp:180-216)
        if (!advancePosition<D>(cell, delta, v, A.hdtdx, bd, bv))
        {
            raiseError(A.err, 3u, 1u, i, bd, bv, A.kind, A.pop);
            continue;
        }
        // first selector: noop (domain) or inGhostBox (level ghosts)
        bool const sel1 = A.kind == KIND_DOMAIN ? true : inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
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

        // selection of what is deposited and what the array keeps
        bool const inDom = inBox(cell, A.box.dom_lo, A.box.dom_hi, D);
        bool deposit;
        uint8_t st = ST_KEEP;
        if (A.kind == KIND_DOMAIN)
        {
            bool const nl = inNLGB(cell, A.box, D);
            deposit       = nl;
            st            = inDom ? ST_KEEP : (nl ? ST_LEAVE : ST_DROP);
        }
        else
        {
            deposit = sel1 && inDom;
            bool const layer
                = sel1 && !inDom && inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
            st = !sel1 ? ST_DROP : (inDom ? ST_EXPORT : (layer ? ST_KEEP : ST_DROP));
        }

        if (deposit)
        {
            primalWeights<D, O>(G, cell, delta, dd[k]);
            bool ok = true;
#pragma unroll
            for (int d = 0; d < D; ++d)
                ok = ok && dd[k].start[d] >= 0 && dd[k].start[d] + O < G.shape[QMOM][d];
            if (!ok)
            {
                raiseError(A.err, 4u, 3u, i, delta[0], v[0], A.kind, A.pop);
            }
            else
            {
                dd[k].val[0]    = 1. * w * A.coef;
                dd[k].val[1]    = q * w * A.coef;
                dd[k].val[2]    = v[0] * w * A.coef;
                dd[k].val[3]    = v[1] * w * A.coef;
                dd[k].val[4]    = v[2] * w * A.coef;
                dd[k].on        = true;
#pragma unroll
                for (int d = 0; d < D; ++d)
                {
                    atomicMin(&bb[d], dd[k].start[d]);
                    atomicMax(&bb[3 + d], dd[k].start[d] + O);
                }
            }
        }

        if (A.mode == MODE_ALL)
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
    __syncthreads();

    // deposit: LDS tile over the chunk's footprint when it fits, else straight to HBM
    int origin[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
    long vol = 1;
    bool any = bb[0] <= bb[3];
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        origin[d] = bb[d];
        ext[d]    = bb[3 + d] - bb[d] + 1;
        vol *= ext[d];
    }
    if (!any)
        return;
    bool const useLds = vol <= TILE_NODES;
    if (useLds)
    {
        for (int x = t; x < 5 * TILE_NODES; x += BLOCK)
            tile[x] = 0.;
        __syncthreads();
        double* tp[1] = {tile};
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            if (dd[k].on)
                depositTo<D, O>(dd[k], tp, origin, ext, true);
        __syncthreads();
        // flush: one HBM atomic per touched node and moment
        for (int x = t; x < vol; x += BLOCK)
        {
            int i = x / (ext[1] * ext[2]);
            int r = x - i * ext[1] * ext[2];
            int j = r / ext[2];
            int k = r - j * ext[2];
            int gi = (i + origin[0]);
            int gj = D > 1 ? (j + origin[1]) : 0;
            int gk = D > 2 ? (k + origin[2]) : 0;
            int node = G.idx(QMOM, gi, gj, gk);
#pragma unroll
            for (int m = 0; m < 5; ++m)
            {
                double val = tile[m * TILE_NODES + x];
                if (val != 0.)
                    atomicAdd(A.mom[m] + node, val);
            }
        }
    }
    else
    {
        int zero[3] = {0, 0, 0};
        int ext3[3] = {G.shape[QMOM][0], G.shape[QMOM][1], G.shape[QMOM][2]};
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            if (dd[k].on)
                depositTo<D, O>(dd[k], A.mom, zero, ext3, false);
    }
}

// deposit without push (patch-ghost particles in mode all, ion_updater.hpp:292-293)
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_deposit(PushArgs A, PView in)
{
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= A.n)
        return;
    int cell[D];
    double delta[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        cell[d]  = in.ic[d][i];
        delta[d] = in.de[d][i];
    }
    if (cell[0] == DEAD_CELL)
        return;
    DepositData<D, O> dd;
    primalWeights<D, O>(A.G, cell, delta, dd);
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (dd.start[d] < 0 || dd.start[d] + O >= A.G.shape[QMOM][d])
        {
            raiseError(A.err, 4u, 3u, i, delta[0], 0., A.kind, A.pop);
            return;
        }
    double const w = in.w[i], q = in.q[i];
    dd.val[0]      = 1. * w * A.coef;
    dd.val[1]      = q * w * A.coef;
    dd.val[2]      = in.v[0][i] * w * A.coef;
    dd.val[3]      = in.v[1][i] * w * A.coef;
    dd.val[4]      = in.v[2][i] * w * A.coef;
    int zero[3]    = {0, 0, 0};
    int ext3[3]    = {A.G.shape[QMOM][0], A.G.shape[QMOM][1], A.G.shape[QMOM][2]};
    depositTo<D, O>(dd, A.mom, zero, ext3, false);
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
    unsigned int const grid = static_cast<unsigned int>((A.n + CHUNK - 1) / CHUNK);
    k_push<D, O><<<grid, BLOCK, 0, s>>>(A, in, out);
}

template<int D, int O>
static void launchDepositT(PushArgs const& A, PView const& in, hipStream_t s)
{
    if (A.n == 0)
        return;
    unsigned int const grid = static_cast<unsigned int>((A.n + BLOCK - 1) / BLOCK);
    k_deposit<D, O><<<grid, BLOCK, 0, s>>>(A, in);
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
