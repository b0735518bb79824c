// phx_amr.hip -- coarse-to-fine kernels of the AMR level coupling for gfx950: the particle split
// (ParticlesRefineOperator::refine_, src/amr/data/particles/refine/particles_data_split.hpp:157-260,
// with the split patterns of src/amr/data/particles/refine/split_{1,2,3}d.hpp dispatched as in
// splitter.hpp:53-104) and the linear time interpolation of fields
// (src/amr/data/field/time_interpolate/field_linear_time_interpolate.hpp).
//
// The split runs one thread per coarse particle: the particle on the fine grid (toFineGrid,
// particles_data_split.hpp:30-45), the candidates whose fine cell lies in a split box (a
// destination box grown by ceil((order + 1) / 2) cells), and each of its refined particles that
// lands in a destination box is appended to the destination array. The appends are ordered:
// coarse particle order, then pattern order, through a count pass, a scan of the workgroup
// counts and a write pass with the same in-workgroup prefix.
#include "phx_kernels.hpp"

namespace phx
{
namespace
{
constexpr int BLOCK = 256;

__device__ __forceinline__ bool inAnyBox(int const* c, SplitSpec const& S, int D, bool grown)
{
    for (int b = 0; b < S.nBoxes; ++b)
    {
        bool in = true;
        for (int d = 0; d < D; ++d)
        {
            int const lo = S.lo[b][d] - (grown ? S.grow : 0);
            int const hi = S.hi[b][d] + (grown ? S.grow : 0);
            in           = in && c[d] >= lo && c[d] <= hi;
        }
        if (in)
            return true;
    }
    return false;
}

// the coarse particle i on the fine grid (toFineGrid): cell * 2 + int(2 delta), the fraction as
// delta; false for an empty slot or a particle whose fine cell is in no split box
template<int D>
__device__ __forceinline__ bool fineParent(PView const& src, unsigned long long i, SplitSpec const& S,
                                           int* cell, double* delta)
{
    if (src.ic[0][i] == DEAD_CELL)
        return false;
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        double const fd = src.de[d][i] * static_cast<double>(S.ratio);
        int const fi    = static_cast<int>(fd);
        cell[d]         = src.ic[d][i] * S.ratio + fi;
        delta[d]        = fd - fi;
    }
    return inAnyBox(cell, S, D, true);
}

// refined particle k of a parent on the fine grid (PatternDispatcher::dispatch, splitter.hpp:
// 71-103): delta + the pattern's (single precision) offset, its floor moved into the cell
template<int D>
__device__ __forceinline__ void refined(SplitSpec const& S, int k, int const* pc, double const* pd,
                                        int* cell, double* delta)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        double x            = pd[d] + static_cast<double>(S.delta[k][d]);
        double const integra = floor(x);
        x -= integra;
        delta[d] = x;
        cell[d]  = pc[d] + static_cast<int>(integra);
    }
}

template<int D>
__device__ __forceinline__ unsigned int keptCount(PView const& src, unsigned long long i,
                                                  unsigned long long n, SplitSpec const& S)
{
    if (i >= n)
        return 0u;
    int pc[D];
    double pd[D];
    if (!fineParent<D>(src, i, S, pc, pd))
        return 0u;
    unsigned int k = 0;
    for (int r = 0; r < S.nRefined; ++r)
    {
        int c[D];
        double dl[D];
        refined<D>(S, r, pc, pd, c, dl);
        k += inAnyBox(c, S, D, false) ? 1u : 0u;
    }
    return k;
}

template<int D>
__global__ __launch_bounds__(BLOCK) void k_split_count(PView src, unsigned long long n, SplitSpec S,
                                                       unsigned int* __restrict__ blockCounts)
{
    __shared__ unsigned int s;
    if (threadIdx.x == 0)
        s = 0;
    __syncthreads();
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    unsigned int const k       = keptCount<D>(src, i, n, S);
    if (k)
        atomicAdd(&s, k);
    __syncthreads();
    if (threadIdx.x == 0)
        blockCounts[blockIdx.x] = s;
}

// exclusive scan of the workgroup counts in one workgroup of 1024; the total at *total
__global__ __launch_bounds__(1024) void k_split_scan(unsigned int* __restrict__ blockCounts,
                                                     unsigned int nblocks,
                                                     unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long part[1024];
    unsigned int const per = (nblocks + 1023) / 1024;
    unsigned int const b0  = threadIdx.x * per;
    unsigned long long sum = 0;
    for (unsigned int b = b0; b < b0 + per && b < nblocks; ++b)
        sum += blockCounts[b];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (unsigned int off = 1; off < 1024; off <<= 1)
    {
        unsigned long long const v = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = threadIdx.x ? part[threadIdx.x - 1] : 0ull;
    for (unsigned int b = b0; b < b0 + per && b < nblocks; ++b)
    {
        unsigned int const c = blockCounts[b];
        blockCounts[b]        = static_cast<unsigned int>(run);
        run += c;
    }
    if (threadIdx.x == 1023)
        *total = part[1023];
}

// the refined particles, at dstOffset + the workgroup's offset + the thread's in-workgroup prefix
template<int D>
__global__ __launch_bounds__(BLOCK) void k_split_write(PView src, unsigned long long n, SplitSpec S,
                                                       unsigned int const* __restrict__ offsets,
                                                       PView dst, unsigned long long dstOffset,
                                                       double weightFactor)
{
    __shared__ unsigned int scan[BLOCK];
    unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x;
    unsigned int const k       = keptCount<D>(src, i, n, S);
    scan[threadIdx.x]          = k;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1)
    {
        unsigned int const v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0u;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (k == 0)
        return;
    unsigned long long j = dstOffset + offsets[blockIdx.x] + scan[threadIdx.x] - k;
    int pc[D];
    double pd[D];
    fineParent<D>(src, i, S, pc, pd);
    double const w = src.w[i], q = src.q[i];
    double const vx = src.v[0][i], vy = src.v[1][i], vz = src.v[2][i];
    for (int r = 0; r < S.nRefined; ++r)
    {
        int c[D];
        double dl[D];
        refined<D>(S, r, pc, pd, c, dl);
        if (!inAnyBox(c, S, D, false))
            continue;
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            dst.ic[d][j] = c[d];
            dst.de[d][j] = dl[d];
        }
        dst.v[0][j] = vx;
        dst.v[1][j] = vy;
        dst.v[2][j] = vz;
        // particle.weight * weight * power[dimension - 1] (splitter.hpp:91)
        dst.w[j] = w * static_cast<double>(S.weight[r]) * weightFactor;
        dst.q[j] = q;
        ++j;
    }
}

// FieldLinearTimeInterpolate (field_linear_time_interpolate.hpp): dst = (1 - alpha) old + alpha
// new on the nodes of the given local box
__global__ __launch_bounds__(BLOCK) void k_time_interp(double const* __restrict__ a,
                                                       double const* __restrict__ b,
                                                       double* __restrict__ out, int s1, int s2,
                                                       int lo0, int lo1, int lo2, int n0, int n1,
                                                       int n2, double alpha)
{
    long const t = static_cast<long>(blockIdx.x) * BLOCK + threadIdx.x;
    long const total = static_cast<long>(n0) * n1 * n2;
    if (t >= total)
        return;
    int const k = static_cast<int>(t % n2);
    int const j = static_cast<int>((t / n2) % n1);
    int const i = static_cast<int>(t / (static_cast<long>(n1) * n2));
    long const x = (static_cast<long>(lo0 + i) * s1 + (lo1 + j)) * s2 + (lo2 + k);
    out[x]       = (1. - alpha) * a[x] + alpha * b[x];
}
// the coarsening of one coarse node from the fine nodes under it (ratio 2): MomentsCoarsener
// (src/amr/data/field/coarsening/moments_coarsener.hpp:39-75) injects the fine node 2c;
// ElectricFieldCoarsener (electric_field_coarsener.hpp:38-140) averages the two fine nodes 2c,
// 2c + 1 along the component's dual direction and injects along its primal ones. Coarse nodes
// (i, j, k) run over the local box cLo .. cLo + n - 1 of the coarse array; fOff is the fine local
// index of coarse local index 0 (2 x (coarse AMR origin) - fine AMR origin)
__global__ __launch_bounds__(BLOCK) void k_coarsen(double const* __restrict__ fine, int fs1, int fs2,
                                                   double* __restrict__ coarse, int cs1, int cs2,
                                                   int clo0, int clo1, int clo2, int n0, int n1,
                                                   int n2, int fo0, int fo1, int fo2, int dualDir,
                                                   int dim)
{
    long const t     = static_cast<long>(blockIdx.x) * BLOCK + threadIdx.x;
    long const total = static_cast<long>(n0) * n1 * n2;
    if (t >= total)
        return;
    int const c2 = clo2 + static_cast<int>(t % n2);
    int const c1 = clo1 + static_cast<int>((t / n2) % n1);
    int const c0 = clo0 + static_cast<int>(t / (static_cast<long>(n1) * n2));
    int const f0 = 2 * c0 + fo0;
    int const f1 = dim > 1 ? 2 * c1 + fo1 : 0;
    int const f2 = dim > 2 ? 2 * c2 + fo2 : 0;
    long const fi = (static_cast<long>(f0) * fs1 + f1) * fs2 + f2;
    double v      = fine[fi];
    if (dualDir >= 0)
    {
        long const step = dualDir == 0 ? static_cast<long>(fs1) * fs2 : (dualDir == 1 ? fs2 : 1);
        v               = 0.5 * (fine[fi] + fine[fi + step]);
    }
    coarse[(static_cast<long>(c0) * cs1 + c1) * cs2 + c2] = v;
}

// a += b * coef over n values (PlusEqualsProduct, src/core/utilities/types.hpp)
__global__ __launch_bounds__(BLOCK) void k_axpy(double* __restrict__ a, double const* __restrict__ b,
                                                double coef, long n)
{
    long const t = static_cast<long>(blockIdx.x) * BLOCK + threadIdx.x;
    if (t < n)
        a[t] += b[t] * coef;
}

// ---- coarse-to-fine field refinement (ratio 2) -------------------------------------------------
// floor(i / 2) for either sign: the coarse cell or node under fine index i (toCoarseIndex,
// src/amr/resources_manager/amr_utils.hpp:129-135)
__device__ __forceinline__ int halfFloor(int i) { return (i - (i < 0 ? 1 : 0)) / 2; }
__device__ __forceinline__ bool isOdd(int i) { return (i & 1) != 0; }

struct Grid3
{
    double const* p;
    int s1, s2;
    __device__ __forceinline__ double operator()(int i, int j, int k) const
    {
        return p[(static_cast<long>(i) * s1 + j) * s2 + k];
    }
};

// linear interpolation weights along one direction for fine index f (FieldRefineIndexesAndWeights
// with LinearWeighter, field_linear_refine.hpp:60-116, linear_weighter.cpp:6-53): the coarse node
// to start from and the weights of it and of its right neighbour. A primal fine node on a coarse
// node takes it whole (the neighbour with weight 0), one between two coarse nodes their mean; a
// dual fine node lies a quarter of a coarse cell from one coarse dual node and three quarters from
// the other.
__device__ __forceinline__ int linearStart(int f, bool primal, double& w0, double& w1)
{
    bool const odd = isOdd(f);
    if (primal)
    {
        w0 = odd ? 0.5 : 1.0;
        w1 = odd ? 0.5 : 0.0;
        return halfFloor(f);
    }
    w0 = odd ? 0.75 : 0.25;
    w1 = odd ? 0.25 : 0.75;
    return halfFloor(f - 1);
}

// DefaultFieldRefiner / FieldMomentsRefiner (field_refiner.hpp:58-161, field_moments_refiner.hpp:
// 56-160): the tensor-product linear interpolation, summed innermost direction first, every weight
// applied (a zero weight still multiplies its node)
template<int D>
__device__ __forceinline__ double linearValue(RefineArgs const& A, Grid3 const& C, int const* f)
{
    int s[3]        = {0, 0, 0};
    double w[3][2]  = {{1.0, 0.0}, {1.0, 0.0}, {1.0, 0.0}};
#pragma unroll
    for (int d = 0; d < D; ++d)
        s[d] = linearStart(f[d], A.primal[d] != 0, w[d][0], w[d][1]) - A.cOrg[d];
    double v = 0.;
    if constexpr (D == 1)
    {
        for (int a = 0; a < 2; ++a)
            v += C(s[0] + a, 0, 0) * w[0][a];
    }
    else if constexpr (D == 2)
    {
        for (int a = 0; a < 2; ++a)
        {
            double y = 0.;
            for (int b = 0; b < 2; ++b)
                y += C(s[0] + a, s[1] + b, 0) * w[1][b];
            v += y * w[0][a];
        }
    }
    else
    {
        for (int a = 0; a < 2; ++a)
        {
            double y = 0.;
            for (int b = 0; b < 2; ++b)
            {
                double z = 0.;
                for (int c = 0; c < 2; ++c)
                    z += C(s[0] + a, s[1] + b, s[2] + c) * w[2][c];
                y += z * w[1][b];
            }
            v += y * w[0][a];
        }
    }
    return v;
}

// ElectricFieldRefiner (electric_field_refiner.hpp:40-345): along the component's dual direction
// both fine edges take the coarse edge they halve; along its primal directions a fine edge on a
// coarse face takes the coarse value, one between two coarse faces their mean (the mean of four
// when between faces in both primal directions). 1D copies the coarse value for every component.
template<int D>
__device__ __forceinline__ double electricValue(RefineArgs const& A, Grid3 const& C, int const* f)
{
    int c[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d)
        c[d] = halfFloor(f[d]) - A.cOrg[d];
    if constexpr (D == 1)
        return C(c[0], 0, 0);
    int pd[2] = {0, 0};
    int np    = 0;
    for (int d = 0; d < D; ++d)
        if (A.primal[d] && np < 2)
            pd[np++] = d;
    // the coarse value shifted by one node along the first (da) / second (db) primal direction
    auto at = [&](int da, int db) {
        int i[3] = {c[0], c[1], c[2]};
        i[pd[0]] += da;
        i[pd[1]] += db;
        return C(i[0], i[1], i[2]);
    };
    if (np == 1)
        return isOdd(f[pd[0]]) ? 0.5 * (at(0, 0) + at(1, 0)) : at(0, 0);
    bool const oa = isOdd(f[pd[0]]);
    bool const ob = isOdd(f[pd[1]]);
    if (!oa && !ob)
        return at(0, 0);
    if (!oa)
        return 0.5 * (at(0, 0) + at(0, 1));
    if (!ob)
        return 0.5 * (at(0, 0) + at(1, 0));
    if (D == 3 && pd[0] == 1) // Ex in 3D: the pairs along y summed, then halved, z-pair by z-pair
        return 0.25 * (at(0, 0) + at(1, 0)) + 0.25 * (at(0, 1) + at(1, 1));
    return 0.25 * (at(0, 0) + at(1, 0) + at(0, 1) + at(1, 1));
}

template<int D>
__global__ __launch_bounds__(BLOCK) void k_refine_field(RefineArgs A)
{
    long const t     = static_cast<long>(blockIdx.x) * BLOCK + threadIdx.x;
    long const total = static_cast<long>(A.n[0]) * A.n[1] * A.n[2];
    if (t >= total)
        return;
    int f[3];
    f[2] = A.lo[2] + static_cast<int>(t % A.n[2]);
    f[1] = A.lo[1] + static_cast<int>((t / A.n[2]) % A.n[1]);
    f[0] = A.lo[0] + static_cast<int>(t / (static_cast<long>(A.n[1]) * A.n[2]));
    long const fi = (static_cast<long>(f[0] - A.fOrg[0]) * A.fs1 + (f[1] - A.fOrg[1])) * A.fs2
                    + (f[2] - A.fOrg[2]);
    bool const nanOnly = A.op != REFINE_LINEAR_OVERWRITE && A.op != REFINE_MAGNETIC_INIT;
    if (nanOnly && !isnan(A.fine[fi]))
        return;
    Grid3 const C{A.coarse, A.cs1, A.cs2};
    double v;
    if (A.op == REFINE_LINEAR || A.op == REFINE_LINEAR_OVERWRITE)
        v = linearValue<D>(A, C, f);
    else if (A.op == REFINE_ELECTRIC)
        v = electricValue<D>(A, C, f);
    else
    {
        // MagneticFieldRefiner / MagneticFieldInitRefiner (magnetic_field_refiner.hpp:50-194):
        // the fine faces on a coarse face take its value; the new faces between two coarse faces
        // (odd index along the component's primal direction) are left to the patch strategy
        int c[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            if (A.primal[d] && isOdd(f[d]))
                return;
            c[d] = halfFloor(f[d]) - A.cOrg[d];
        }
        v = C(c[0], c[1], c[2]);
    }
    A.fine[fi] = v;
}

// index along one direction of a B value around a new face (MagneticRefinePatchStrategy's
// p_plus / p_minus along the value's primal direction, d_plus / d_minus along its dual ones,
// magnetic_refine_patch_strategy.hpp:372-376): i - o on the minus side, i + 2 - o (primal) or
// i + 1 - o (dual) on the plus side
__device__ __forceinline__ int faceSide(int i, int o, bool primal, int plus)
{
    return plus ? i + (primal ? 2 : 1) - o : i - o;
}

// MagneticRefinePatchStrategy::postprocessRefine (magnetic_refine_patch_strategy.hpp:67-360): the
// new fine faces of component n (odd AMR index along n), the divergence-free reconstruction of
// Toth & Roe (2002) from the shared faces: the mean of the two faces along n, plus the cross terms
// of the other components (2D: one group of 4; 3D: two first-order groups of 8 and two third-order
// groups of 8 weighted by the mesh sizes)
template<int D>
__global__ __launch_bounds__(BLOCK) void k_mag_new_faces(NewFacesArgs A, int n, int e0, int e1, int e2)
{
    long const t     = static_cast<long>(blockIdx.x) * BLOCK + threadIdx.x;
    long const total = static_cast<long>(e0) * e1 * e2;
    if (t >= total)
        return;
    int f[3];
    f[2] = A.lo[2] + static_cast<int>(t % e2);
    f[1] = A.lo[1] + static_cast<int>((t / e2) % e1);
    f[0] = A.lo[0] + static_cast<int>(t / (static_cast<long>(e1) * e2));
    if (!isOdd(f[n]))
        return;
    int loc[3] = {0, 0, 0}, off[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        loc[d] = f[d] - A.org[d];
        off[d] = (d == n || isOdd(f[d])) ? 1 : 0;
    }
    auto at = [&](int comp, int i, int j, int k) {
        return A.b[comp][(static_cast<long>(i) * A.s1[comp] + j) * A.s2[comp] + k];
    };
    // component comp on the side (sx, sy, sz) of the face (0 minus, 1 plus per direction)
    auto val = [&](int comp, int sx, int sy, int sz) {
        int const s[3] = {sx, sy, sz};
        int i[3]       = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d)
            i[d] = faceSide(loc[d], off[d], d == comp, s[d]);
        return at(comp, i[0], i[1], i[2]);
    };
    int im[3] = {loc[0], loc[1], loc[2]}, ip[3] = {loc[0], loc[1], loc[2]};
    im[n] -= 1;
    ip[n] += 1;
    // at the edge of the arrays: no neighbour along the normal -> left as it is; cross terms that
    // would read beyond the arrays -> the mean of the two neighbours only
    if (im[n] < 0 || ip[n] >= A.shape[n][n])
        return;
    bool cross = true;
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
        if (c == n || (D == 2 && c == 2) || D == 1)
            continue;
#pragma unroll
        for (int d = 0; d < D; ++d)
            cross = cross && loc[d] - off[d] >= 0 && loc[d] + (d == c ? 2 : 1) - off[d] < A.shape[c][d];
    }
    double v = 0.5 * (at(n, im[0], im[1], im[2]) + at(n, ip[0], ip[1], ip[2]));
    if (!cross)
    {
        A.b[n][(static_cast<long>(loc[0]) * A.s1[n] + loc[1]) * A.s2[n] + loc[2]] = v;
        return;
    }
    if constexpr (D == 2)
    {
        double s;
        if (n == 0) // Bx from By, x-side outer
            s = val(1, 0, 0, 0) - val(1, 0, 1, 0) - val(1, 1, 0, 0) + val(1, 1, 1, 0);
        else // By from Bx, y-side outer
            s = val(0, 0, 0, 0) - val(0, 1, 0, 0) - val(0, 0, 1, 0) + val(0, 1, 1, 0);
        v = v + 0.25 * s;
    }
    else if constexpr (D == 3)
    {
        // first-order group of component comp: z-side slowest, then x, then y; a value counts
        // positive where its sides along n and along comp agree
        auto first = [&](int comp) {
            double s = val(comp, 0, 0, 0);
            for (int k = 1; k < 8; ++k)
            {
                int const sd[3] = {(k >> 1) & 1, k & 1, (k >> 2) & 1};
                double const x  = val(comp, sd[0], sd[1], sd[2]);
                s               = sd[n] == sd[comp] ? s + x : s - x;
            }
            return s;
        };
        // third-order group: all plus, one minus (x, y, z), two minus (y z, x z, x y), all minus;
        // sign (-1)^(number of minus sides)
        auto third = [&](int comp) {
            constexpr int ORD[8][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0},
                                       {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};
            double s = val(comp, 1, 1, 1);
            for (int k = 1; k < 8; ++k)
            {
                double const x  = val(comp, ORD[k][0], ORD[k][1], ORD[k][2]);
                int const minus = 3 - ORD[k][0] - ORD[k][1] - ORD[k][2];
                s               = (minus & 1) ? s - x : s + x;
            }
            return s;
        };
        // weight of the third-order group of comp: +-1/8 Dt^2 / (Dn^2 + Dt^2), t the direction
        // that is neither n nor comp, minus where the face is on the even side along t
        auto weight = [&](int comp) {
            int const tt = 3 - n - comp;
            double const sgn = off[tt] ? 1.0 : -1.0;
            return 0.125 * sgn * A.dx[tt] * A.dx[tt] / (A.dx[n] * A.dx[n] + A.dx[tt] * A.dx[tt]);
        };
        int const lo = n == 0 ? 1 : 0;
        int const hi = n == 2 ? 1 : 2;
        int const c1 = (n + 1) % 3, c2 = (n + 2) % 3;
        v = v + 0.125 * first(lo) + 0.125 * first(hi) + weight(c1) * third(c1) + weight(c2) * third(c2);
    }
    A.b[n][(static_cast<long>(loc[0]) * A.s1[n] + loc[1]) * A.s2[n] + loc[2]] = v;
}
} // namespace

void launchCoarsen(double const* fine, int const* fshape, double* coarse, int const* cshape,
                   int const* clo, int const* n, int const* fineOff, int dualDir, int dim, hipStream_t s)
{
    long const total = static_cast<long>(n[0]) * n[1] * n[2];
    if (total <= 0)
        return;
    unsigned int const nb = static_cast<unsigned int>((total + BLOCK - 1) / BLOCK);
    k_coarsen<<<nb, BLOCK, 0, s>>>(fine, fshape[1], fshape[2], coarse, cshape[1], cshape[2], clo[0],
                                   clo[1], clo[2], n[0], n[1], n[2], fineOff[0], fineOff[1], fineOff[2],
                                   dualDir, dim);
}

void launchRefineField(RefineArgs const& A, hipStream_t s)
{
    long const total = static_cast<long>(A.n[0]) * A.n[1] * A.n[2];
    if (total <= 0)
        return;
    unsigned int const nb = static_cast<unsigned int>((total + BLOCK - 1) / BLOCK);
    switch (A.dim)
    {
        case 1: k_refine_field<1><<<nb, BLOCK, 0, s>>>(A); break;
        case 2: k_refine_field<2><<<nb, BLOCK, 0, s>>>(A); break;
        case 3: k_refine_field<3><<<nb, BLOCK, 0, s>>>(A); break;
    }
}

void launchMagneticNewFaces(NewFacesArgs const& A, hipStream_t s)
{
    for (int n = 0; n < A.dim; ++n)
    {
        int e[3] = {1, 1, 1};
        for (int d = 0; d < A.dim; ++d)
            e[d] = A.hi[d] - A.lo[d] + 1 + (d == n ? 1 : 0);
        long const total = static_cast<long>(e[0]) * e[1] * e[2];
        if (total <= 0)
            continue;
        unsigned int const nb = static_cast<unsigned int>((total + BLOCK - 1) / BLOCK);
        switch (A.dim)
        {
            case 1: k_mag_new_faces<1><<<nb, BLOCK, 0, s>>>(A, n, e[0], e[1], e[2]); break;
            case 2: k_mag_new_faces<2><<<nb, BLOCK, 0, s>>>(A, n, e[0], e[1], e[2]); break;
            case 3: k_mag_new_faces<3><<<nb, BLOCK, 0, s>>>(A, n, e[0], e[1], e[2]); break;
        }
    }
}

void launchAxpy(double* a, double const* b, double coef, long n, hipStream_t s)
{
    if (n <= 0)
        return;
    k_axpy<<<static_cast<unsigned int>((n + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(a, b, coef, n);
}

// The split patterns (src/amr/data/particles/refine/split_{1,2,3}d.hpp): refined particle
// offsets from the parent (in fine cells) and weights (fractions of the parent's weight before the
// factor ratio^dim), in the order PatternDispatcher places them. Patterns: black = the parent's
// position; pink = the 2 D offsets +-delta along each axis (1D: +d, -d; 2D: (0,-d), (-d,0),
// (+d,0), (0,+d); 3D: +d along x, z, y, then -d along x, z, y); purple (2D) = the 4 corners
// (-d,-d), (-d,+d), (+d,-d), (+d,+d). The default refined particle numbers of PHARE (2 in 1D,
// 4 in 2D, 6 in 3D) and every 1D variant.
bool splitPattern(int dim, int order, int nbr, SplitSpec& S)
{
    S.nRefined = 0;
    auto put = [&S](float w, float dx, float dy, float dz) {
        S.weight[S.nRefined]   = w;
        S.delta[S.nRefined][0] = dx;
        S.delta[S.nRefined][1] = dy;
        S.delta[S.nRefined][2] = dz;
        ++S.nRefined;
    };
    auto black = [&](float w) { put(w, 0.f, 0.f, 0.f); };
    auto pink1 = [&](float w, float d) {
        put(w, d, 0.f, 0.f);
        put(w, -d, 0.f, 0.f);
    };
    auto pink2 = [&](float w, float d) {
        put(w, 0.f, -d, 0.f);
        put(w, -d, 0.f, 0.f);
        put(w, d, 0.f, 0.f);
        put(w, 0.f, d, 0.f);
    };
    auto purple2 = [&](float w, float d) {
        put(w, -d, -d, 0.f);
        put(w, -d, d, 0.f);
        put(w, d, -d, 0.f);
        put(w, d, d, 0.f);
    };
    auto pink3 = [&](float w, float d) {
        for (int i = 0; i < 2; ++i)
        {
            float const m = d * (i % 2 ? -1.f : 1.f);
            put(w, m, 0.f, 0.f);
            put(w, 0.f, 0.f, m);
            put(w, 0.f, m, 0.f);
        }
    };
    switch (dim * 100 + order * 10 + nbr)
    {
        case 112: pink1(0.5f, 0.551569f); break;
        case 113:
            black(0.5f);
            pink1(0.25f, 1.f);
            break;
        case 122: pink1(0.5f, 0.663959f); break;
        case 123:
            black(0.468137f);
            pink1(0.265931f, 1.112033f);
            break;
        case 124:
            pink1(0.375f, 0.5f);
            pink1(0.125f, 1.5f);
            break;
        case 132: pink1(0.5f, 0.752399f); break;
        case 133:
            black(0.473943f);
            pink1(0.263028f, 1.275922f);
            break;
        case 134:
            pink1(0.364766f, 0.542949f);
            pink1(0.135234f, 1.664886f);
            break;
        case 135:
            black(0.375f);
            pink1(0.25f, 1.f);
            pink1(0.0625f, 2.f);
            break;
        case 214: purple2(0.25f, 0.571783f); break;
        case 224: pink2(0.25f, 0.683734f); break;
        case 234: purple2(0.25f, 0.776459f); break;
        case 316: pink3(.166666f, .966431f); break;
        case 326: pink3(.166666f, 1.149658f); break;
        case 336: pink3(.166666f, 1.312622f); break;
        default: return false;
    }
    return true;
}

unsigned long long launchSplitCount(PView const& src, unsigned long long n, SplitSpec const& S, int dim,
                                    unsigned int* blockCounts, unsigned long long* total, hipStream_t s)
{
    unsigned int const nb = static_cast<unsigned int>((n + BLOCK - 1) / BLOCK);
    if (nb == 0)
        return 0;
    switch (dim)
    {
        case 1: k_split_count<1><<<nb, BLOCK, 0, s>>>(src, n, S, blockCounts); break;
        case 2: k_split_count<2><<<nb, BLOCK, 0, s>>>(src, n, S, blockCounts); break;
        case 3: k_split_count<3><<<nb, BLOCK, 0, s>>>(src, n, S, blockCounts); break;
    }
    k_split_scan<<<1, 1024, 0, s>>>(blockCounts, nb, total);
    return nb;
}

void launchSplitWrite(PView const& src, unsigned long long n, SplitSpec const& S, int dim,
                      unsigned int const* offsets, PView const& dst, unsigned long long dstOffset,
                      hipStream_t s)
{
    unsigned int const nb = static_cast<unsigned int>((n + BLOCK - 1) / BLOCK);
    if (nb == 0)
        return;
    double const factor = static_cast<double>(dim == 1 ? S.ratio : (dim == 2 ? S.ratio * S.ratio
                                                                              : S.ratio * S.ratio * S.ratio));
    switch (dim)
    {
        case 1: k_split_write<1><<<nb, BLOCK, 0, s>>>(src, n, S, offsets, dst, dstOffset, factor); break;
        case 2: k_split_write<2><<<nb, BLOCK, 0, s>>>(src, n, S, offsets, dst, dstOffset, factor); break;
        case 3: k_split_write<3><<<nb, BLOCK, 0, s>>>(src, n, S, offsets, dst, dstOffset, factor); break;
    }
}

void launchTimeInterp(double const* a, double const* b, double* out, int const* shape, int const* lo,
                      int const* n, double alpha, hipStream_t s)
{
    long const total = static_cast<long>(n[0]) * n[1] * n[2];
    if (total <= 0)
        return;
    unsigned int const nb = static_cast<unsigned int>((total + BLOCK - 1) / BLOCK);
    k_time_interp<<<nb, BLOCK, 0, s>>>(a, b, out, shape[1], shape[2], lo[0], lo[1], lo[2], n[0], n[1],
                                       n[2], alpha);
}

} // namespace phx
