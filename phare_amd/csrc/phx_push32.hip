// phx_push32.hip -- the mixed-precision particle family for gfx950: FP32 delta, velocity and
// weight (int32 iCell, the charge per population), pushed and gathered in FP32 from the FP64
// fields; the FP32 deposit contributions are summed in an FP64 LDS tile and flushed into the FP64
// moment arrays. The field solve, the moments it reads and the energy and charge reductions stay
// FP64.
//
// Same method as the FP64 family (IonUpdater::updatePopulations, ion_updater.hpp:88-295; Boris,
// boris.hpp:93-300; Interpolator gather and deposit, interpolator.hpp:54-125,379-510), with
// these changes for FP32:
//  - B-spline weights are evaluated at u = position - stencil start, formed from the small integer
//    (start shift) and delta, instead of from the absolute local position: an FP32 position of a
//    few hundred cells keeps only ~15 bits of delta;
//  - a new delta that rounds to 1 (a tiny negative step) moves the particle to the next cell with
//    delta 0, the same position;
//  - the tile keeps n, Fx, Fy, Fz (the charge density is charge x n at the flush, the charge being
//    the population's) and sums the FP32 contributions in FP64: on gfx950 an LDS FP64 atomic add
//    is several times faster than an FP32 one (ds_add_f64 / ds_add_f32), and FP64 sums of FP32
//    terms are more accurate than compensated FP32 sums (a TwoSum-compensated FP32 tile, whose
//    adds must return the old value, measured slower still);
//  - E and B are gathered from an FP32 copy of the averaged set (written beside it by the average
//    kernel) and interpolated in FP32; in 1D and 2D a component's stencil values are loaded
//    together before the sums, so that their loads are in flight at once.
#include "phx_push_common.hpp"

namespace phx
{
namespace
{
using pc::BLOCK;
using pc::inBox;
using pc::inNLGB;
using pc::LANES;
using pc::NW;

// B-spline weights at u = np - start (u in [0, O+1)); the products of interpolator.hpp:54-125
// written in u: order 2 uses 1 - u for (start + 1 - np), order 3 uses -u for (start - np)
template<int O>
__device__ __forceinline__ void bspline32(float u, float* w)
{
    if constexpr (O == 1)
    {
        w[1] = u;
        w[0] = 1.f - u;
    }
    else if constexpr (O == 2)
    {
        float const d  = 1.f - u;
        float const c1 = .5f + d;
        float const c3 = .5f - d;
        w[0]           = .5f * c1 * c1;
        w[1]           = .75f - d * d;
        w[2]           = .5f * c3 * c3;
    }
    else
    {
        float const index = -u;
        float const c1    = 1.f + .5f * index;
        float const c2    = index + 1.f;
        float const c3    = index + 2.f;
        float const c4    = 1.f - .5f * (index + 3.f);
        float const c2s   = c2 * c2;
        float const c3s   = c3 * c3;
        w[0]              = (4.f / 3.f) * c1 * c1 * c1;
        w[1]              = (2.f / 3.f) - c2s - .5f * c2s * c2;
        w[2]              = (2.f / 3.f) - c3s + .5f * c3s * c3;
        w[3]              = (4.f / 3.f) * c4 * c4 * c4;
    }
}

// start index and weights of one direction (interpolator.hpp:379-406): np - start is the start
// shift plus delta (less one half for the dual centering)
template<int O>
__device__ __forceinline__ int indexAndWeights32(int c, int icl, float delta, float* w)
{
    int const sh = startShift<O>(c, static_cast<double>(delta));
    float u      = static_cast<float>(sh) + delta;
    if (c == CD)
        u -= .5f;
    bspline32<O>(u, w);
    return icl - sh;
}

template<int D, int O>
struct Stencil32
{
    int sp[D], sd[D];
    float wp[D][O + 1], wd[D][O + 1];
};

template<int D, int O>
__device__ __forceinline__ void makeStencil32(Geom const& G, int const* cell, float const* delta,
                                              Stencil32<D, O>& s)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int const icl = cell[d] - (G.lower[d] - G.g);
        s.sd[d]       = indexAndWeights32<O>(CD, icl, delta[d], s.wd[d]);
        s.sp[d]       = indexAndWeights32<O>(CP, icl, delta[d], s.wp[d]);
    }
}

template<int D, int O>
__device__ __forceinline__ bool stencilInside32(Geom const& G, Stencil32<D, O> const& s, int q)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int const st = centering(q, d) == CP ? s.sp[d] : s.sd[d];
        if (st < 0 || st + O >= G.shape[q][d])
            return false;
    }
    return true;
}

// FP32 copy of a field, FP32 weights and sums, in MeshToParticle's order of sums
template<int D, int O>
__device__ __forceinline__ float interp32(Geom const& G, float const* __restrict__ f, int q,
                                          Stencil32<D, O> const& s)
{
    constexpr int S = O + 1;
    int st[D];
    float const* w[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        bool const p = centering(q, d) == CP;
        st[d]        = p ? s.sp[d] : s.sd[d];
        w[d]         = p ? s.wp[d] : s.wd[d];
    }
    int const s1 = G.shape[q][1], s2 = G.shape[q][2];
    float r      = 0.f;
    if constexpr (D == 1)
    {
        float vals[S];
#pragma unroll
        for (int i = 0; i < S; ++i)
            vals[i] = f[st[0] + i];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < S; ++i)
            r = __builtin_fmaf(vals[i], w[0][i], r);
    }
    else if constexpr (D == 2)
    {
        float vals[S][S];
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
            for (int j = 0; j < S; ++j)
                vals[i][j] = f[(st[0] + i) * s1 + st[1] + j];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < S; ++i)
        {
            float yi = 0.f;
#pragma unroll
            for (int j = 0; j < S; ++j)
                yi = __builtin_fmaf(vals[i][j], w[1][j], yi);
            r = __builtin_fmaf(yi, w[0][i], r);
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < S; ++i)
        {
            float yi = 0.f;
#pragma unroll
            for (int j = 0; j < S; ++j)
            {
                float const* pencil = f + ((st[0] + i) * s1 + st[1] + j) * s2 + st[2];
                float zi            = 0.f;
#pragma unroll
                for (int k = 0; k < S; ++k)
                    zi = __builtin_fmaf(pencil[k], w[2][k], zi);
                yi = __builtin_fmaf(zi, w[1][j], yi);
            }
            r = __builtin_fmaf(yi, w[0][i], r);
        }
    }
    return r;
}

// Boris rotation (boris.hpp:240-300) in FP32
__device__ __forceinline__ void borisAccelerate32(float* v, float const* E, float const* B,
                                                  float coef1)
{
    float const vx1  = v[0] + coef1 * E[0];
    float const vy1  = v[1] + coef1 * E[1];
    float const vz1  = v[2] + coef1 * E[2];
    float const rx   = coef1 * B[0];
    float const ry   = coef1 * B[1];
    float const rz   = coef1 * B[2];
    float const rx2  = rx * rx;
    float const ry2  = ry * ry;
    float const rz2  = rz * rz;
    float const rxry = rx * ry;
    float const rxrz = rx * rz;
    float const ryrz = ry * rz;
    float const invDet = 1.f / (1.f + rx2 + ry2 + rz2);
    float const mxx  = 1.f + rx2 - ry2 - rz2;
    float const mxy  = 2.f * (rxry + rz);
    float const mxz  = 2.f * (rxrz - ry);
    float const myx  = 2.f * (rxry - rz);
    float const myy  = 1.f + ry2 - rx2 - rz2;
    float const myz  = 2.f * (ryrz + rx);
    float const mzx  = 2.f * (rxrz + ry);
    float const mzy  = 2.f * (ryrz - rx);
    float const mzz  = 1.f + rz2 - rx2 - ry2;
    float const vx2  = (mxx * vx1 + mxy * vy1 + mxz * vz1) * invDet;
    float const vy2  = (myx * vx1 + myy * vy1 + myz * vz1) * invDet;
    float const vz2  = (mzx * vx1 + mzy * vy1 + mzz * vz1) * invDet;
    v[0]             = vx2 + coef1 * E[0];
    v[1]             = vy2 + coef1 * E[1];
    v[2]             = vz2 + coef1 * E[2];
}

// half position step (boris.hpp:156-172) in FP32; false when |delta| > 2
template<int D>
__device__ __forceinline__ bool advancePosition32(int* cell, float* delta, float const* v,
                                                  float const* hdtdx, float& badDelta, float& badV)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        float const dl = delta[d] + hdtdx[d] * v[d];
        if (fabsf(dl) > 2.f)
        {
            badDelta = dl;
            badV     = v[d];
            return false;
        }
        int ic   = static_cast<int>(floorf(dl));
        float nd = dl - static_cast<float>(ic);
        if (nd >= 1.f)
        {
            // dl a tiny negative number: dl + 1 rounds to 1; the same position is cell + 1, 0
            nd -= 1.f;
            ++ic;
        }
        delta[d] = nd;
        cell[d] += ic;
    }
    return true;
}

// LDS tile of the mixed-precision push: FP64 slab m of n, Fx, Fy, Fz
struct Tile32
{
    double* lds;
    int org[3], ext[3], pitch[3];
    int stride;
    __device__ double* mom(int m) const { return lds + m * stride; }
    __device__ int node(int i, int j, int k) const { return (i * pitch[1] + j) * pitch[2] + k; }
    template<int O>
    __device__ bool holds(int d, int start) const
    {
        return start >= org[d] && start + O < org[d] + ext[d];
    }
};

template<int D, int O>
struct Dep32
{
    int start[D];
    float w[D][O + 1];
    float val[4]; // n, Fx, Fy, Fz contributions (weight x coef, v x weight x coef)
};

template<int D, int O>
__device__ __forceinline__ float nodeW32(Dep32<D, O> const& dd, float dep, int a, int b, int c)
{
    if constexpr (D == 1)
        return dep * dd.w[0][a];
    else if constexpr (D == 2)
        return dep * dd.w[0][a] * dd.w[1][b];
    else
        return dep * dd.w[0][a] * dd.w[1][b] * dd.w[2][c];
}

// one particle's deposit: into the tile when its stencil lies there, else straight into the FP64
// moment arrays (charge density from the population's charge)
template<int D, int O>
__device__ __forceinline__ void deposit32(Dep32<D, O> const& dd, Geom const& G, Tile32 const& T,
                                          double* const* mom, double charge)
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
                    for (int m = 0; m < 4; ++m)
                        atomicAdd(T.mom(m) + node,
                                  static_cast<double>(nodeW32<D, O>(dd, dd.val[m], a, b, c)));
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
                    int const node = G.idx(QMOM, dd.start[0] + a, D > 1 ? dd.start[1] + b : 0,
                                           D > 2 ? dd.start[2] + c : 0);
                    double const nw = static_cast<double>(nodeW32<D, O>(dd, dd.val[0], a, b, c));
                    atomicAdd(mom[0] + node, nw);
                    atomicAdd(mom[1] + node, charge * nw);
#pragma unroll
                    for (int m = 1; m < 4; ++m)
                        atomicAdd(mom[m + 1] + node,
                                  static_cast<double>(nodeW32<D, O>(dd, dd.val[m], a, b, c)));
                }
    }
}

// register window of one lane (WIN variant of the mixed-precision push): the n, Fx, Fy, Fz
// contributions of the lane's successive particles of one cell, summed in FP32 over the union of
// their stencils (nodes cell - WB .. cell - WB + W - 1 per dimension), flushed into the FP64 LDS
// tile (or HBM) when the cell changes and at the end. A lane walks a run of cell-ordered
// particles, so one flush replaces the LDS adds of several particles.
template<int D, int O>
struct Win32
{
    static constexpr int W  = O == 1 ? 2 : 4;
    static constexpr int WB = O == 1 ? 0 : 1;
    static constexpr int WN = D == 1 ? W : (D == 2 ? W * W : W * W * W);
    float a[4][WN];
    int base[D]; // local node index of the window origin
    bool live;
};

template<int D, int O>
__device__ __forceinline__ void winZero(Win32<D, O>& A)
{
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int x = 0; x < Win32<D, O>::WN; ++x)
            A.a[m][x] = 0.f;
    A.live = false;
}

template<int D, int O>
__device__ __forceinline__ void winFlush(Win32<D, O> const& A, Geom const& G, Tile32 const& T,
                                         double* const* mom, double charge)
{
    constexpr int W = Win32<D, O>::W;
    if (!A.live)
        return;
#pragma unroll
    for (int x = 0; x < Win32<D, O>::WN; ++x)
    {
        int const a  = D == 1 ? x : (D == 2 ? x / W : x / (W * W));
        int const b  = D == 1 ? 0 : (D == 2 ? x % W : (x / W) % W);
        int const c  = D == 3 ? x % W : 0;
        int const g[3] = {A.base[0] + a, D > 1 ? A.base[1] + b : 0, D > 2 ? A.base[2] + c : 0};
        bool inTile    = true;
#pragma unroll
        for (int d = 0; d < D; ++d)
            inTile = inTile && g[d] >= T.org[d] && g[d] < T.org[d] + T.ext[d];
        if (inTile)
        {
            int const tnode = T.node(g[0] - T.org[0], D > 1 ? g[1] - T.org[1] : 0,
                                     D > 2 ? g[2] - T.org[2] : 0);
#pragma unroll
            for (int m = 0; m < 4; ++m)
                if (A.a[m][x] != 0.f)
                    atomicAdd(T.mom(m) + tnode, static_cast<double>(A.a[m][x]));
        }
        else
        {
            int const node = G.idx(QMOM, g[0], g[1], g[2]);
            if (A.a[0][x] != 0.f)
            {
                atomicAdd(mom[0] + node, static_cast<double>(A.a[0][x]));
                atomicAdd(mom[1] + node, charge * static_cast<double>(A.a[0][x]));
            }
#pragma unroll
            for (int m = 1; m < 4; ++m)
                if (A.a[m][x] != 0.f)
                    atomicAdd(mom[m + 1] + node, static_cast<double>(A.a[m][x]));
        }
    }
}

// one particle's deposit into the lane's window (re-based, after a flush, when its cell differs
// from the window's); the node products in nodeW32's order
template<int D, int O>
__device__ __forceinline__ void winAdd(Win32<D, O>& A, Dep32<D, O> const& dd, int const* icl,
                                       Geom const& G, Tile32 const& T, double* const* mom,
                                       double charge)
{
    constexpr int W = Win32<D, O>::W;
    constexpr int S = O + 1;
    bool same       = A.live;
#pragma unroll
    for (int d = 0; d < D; ++d)
        same = same && A.base[d] == icl[d] - Win32<D, O>::WB;
    if (!same)
    {
        winFlush<D, O>(A, G, T, mom, charge);
        winZero(A);
#pragma unroll
        for (int d = 0; d < D; ++d)
            A.base[d] = icl[d] - Win32<D, O>::WB;
        A.live = true;
    }
    float w[D][W];
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int const off = dd.start[d] - A.base[d];
#pragma unroll
        for (int a = 0; a < W; ++a)
        {
            float x = 0.f;
#pragma unroll
            for (int e = 0; e < S; ++e)
                x = (a - off == e) ? dd.w[d][e] : x;
            w[d][a] = x;
        }
    }
#pragma unroll
    for (int x = 0; x < Win32<D, O>::WN; ++x)
    {
        int const a = D == 1 ? x : (D == 2 ? x / W : x / (W * W));
        int const b = D == 1 ? 0 : (D == 2 ? x % W : (x / W) % W);
        int const c = D == 3 ? x % W : 0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            float v;
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

// the window variant exists where its accumulators fit the registers: 1D, 2D, and 3D at order 1
template<int D, int O>
constexpr bool winFits()
{
    return D < 3 || O == 1;
}

// per-launch constants in FP32
struct Push32
{
    float hdtdx[3];
    float coef1; // charge * dt / (2 m)
    float coef;
};

// one slot of segmented storage: prePush, gather, Boris, postPush, selection, deposit and the
// in-place write-back (the FP64 family's pushLoaded in FP32)
template<int D, int O, bool WIN>
__device__ __forceinline__ void pushSlot32(PushArgs const& A, Push32 const& K, PView32 const& v,
                                           unsigned long long i, Tile32 const& T, Win32<D, O>& win)
{
    Geom const& G = A.G;
    int cell[D];
    cell[0] = v.ic[0][i];
    if (cell[0] == DEAD_CELL)
    {
        if (A.pushOn && A.mode == MODE_ALL && A.status)
            A.status[i] = ST_DROP;
        return;
    }
    float delta[D], vel[3];
#pragma unroll
    for (int d = 1; d < D; ++d)
        cell[d] = v.ic[d][i];
#pragma unroll
    for (int d = 0; d < D; ++d)
        delta[d] = v.de[d][i];
    vel[0]        = v.v[0][i];
    vel[1]        = v.v[1][i];
    vel[2]        = v.v[2][i];
    float const w = v.w[i];

    bool deposit = true, sel1 = true;
    uint8_t st   = ST_KEEP;
    if (A.pushOn)
    {
        float bd, bv;
        if (!advancePosition32<D>(cell, delta, vel, K.hdtdx, bd, bv))
        {
            raiseError(A.err, 3u, 1u, i, bd, bv, A.kind, A.pop);
            return;
        }
        sel1 = A.kind == KIND_DOMAIN ? true : inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
        if (sel1)
        {
            Stencil32<D, O> s;
            makeStencil32<D, O>(G, cell, delta, s);
            bool safe = true;
#pragma unroll
            for (int d = 0; d < D; ++d)
            {
                int const icl = cell[d] - (G.lower[d] - G.g);
                safe          = safe && icl >= G.safeLo[d] && icl <= G.safeHi[d];
            }
            if (!safe
                && (!stencilInside32<D, O>(G, s, QEX) || !stencilInside32<D, O>(G, s, QBX)
                    || !stencilInside32<D, O>(G, s, QEY) || !stencilInside32<D, O>(G, s, QBY)
                    || !stencilInside32<D, O>(G, s, QEZ) || !stencilInside32<D, O>(G, s, QBZ)))
            {
                raiseError(A.err, 4u, 2u, i, delta[0], vel[0], A.kind, A.pop);
                return;
            }
            float E[3], B[3];
            E[0] = interp32<D, O>(G, A.em32[0], QEX, s);
            E[1] = interp32<D, O>(G, A.em32[1], QEY, s);
            E[2] = interp32<D, O>(G, A.em32[2], QEZ, s);
            B[0] = interp32<D, O>(G, A.em32[3], QBX, s);
            B[1] = interp32<D, O>(G, A.em32[4], QBY, s);
            B[2] = interp32<D, O>(G, A.em32[5], QBZ, s);
            borisAccelerate32(vel, E, B, K.coef1);
            if (!advancePosition32<D>(cell, delta, vel, K.hdtdx, bd, bv))
            {
                raiseError(A.err, 3u, 2u, i, bd, bv, A.kind, A.pop);
                return;
            }
        }
        bool const inDom = inBox(cell, A.box.dom_lo, A.box.dom_hi, D);
        if (A.kind == KIND_DOMAIN)
        {
            bool const nl = inNLGB(cell, A.box, D);
            deposit       = nl;
            st            = inDom ? ST_KEEP : (nl ? ST_LEAVE : ST_DROP);
        }
        else
        {
            deposit          = sel1 && inDom;
            bool const layer = sel1 && !inDom && inBox(cell, A.box.gh_lo, A.box.gh_hi, D);
            st = !sel1 ? ST_DROP : (inDom ? ST_EXPORT : (layer ? ST_KEEP : ST_DROP));
        }
    }

    if (deposit)
    {
        Dep32<D, O> dd;
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            int const icl = cell[d] - (G.lower[d] - G.g);
            dd.start[d]   = indexAndWeights32<O>(CP, icl, delta[d], dd.w[d]);
            ok            = ok && dd.start[d] >= 0 && dd.start[d] + O < G.shape[QMOM][d];
        }
        if (!ok)
            raiseError(A.err, 4u, 3u, i, delta[0], vel[0], A.kind, A.pop);
        else
        {
            float const wc = w * K.coef;
            dd.val[0]      = wc;
            dd.val[1]      = vel[0] * wc;
            dd.val[2]      = vel[1] * wc;
            dd.val[3]      = vel[2] * wc;
            if constexpr (WIN)
            {
                int icl[D];
#pragma unroll
                for (int d = 0; d < D; ++d)
                    icl[d] = cell[d] - (G.lower[d] - G.g);
                winAdd<D, O>(win, dd, icl, G, T, A.mom, A.charge);
            }
            else
                deposit32<D, O>(dd, G, T, A.mom, A.charge);
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
            v.ic[d][i] = cell[d];
            v.de[d][i] = delta[d];
        }
        v.v[0][i] = vel[0];
        v.v[1][i] = vel[1];
        v.v[2][i] = vel[2];
        if (A.status)
            A.status[i] = st;
    }
}

} // namespace

// ---------------------------------------------------------------------------------------------
// fused push + deposit over cell-segmented FP32 storage, in place
// ---------------------------------------------------------------------------------------------
#ifndef PHX_SEG_WAVES32
#define PHX_SEG_WAVES32 4 // waves per SIMD the register allocation of the mixed-precision push targets
#endif
#ifndef PHX_SEG_WAVES32_WIN
#define PHX_SEG_WAVES32_WIN 3 // the same for the register-window variant (its window needs room)
#endif
namespace
{
// rows [rA, rB) of segment seg through the FP64 LDS tile: zero the tile, push the rows (wavefront
// w takes rows rA + w, rA + w + NW, ...), flush the tile. Every work-item of the workgroup calls it.
template<int D, int TB>
__device__ __forceinline__ void flushTile32(PushArgs const& A, Tile32 const& T);

template<int D, int O, bool WIN>
__device__ __forceinline__ void pushPiece32(PushArgs const& A, PView32 const& v, double* lds32,
                                            unsigned int seg, unsigned int rA, unsigned int rB)
{
    Geom const& G          = A.G;
    unsigned long long const off = A.st.off[seg];
    int const t    = threadIdx.x;
    int const lane = t & 63;

    int const slabs = 4 * A.tile.stride;
    for (int x = t; x < slabs; x += pc::SEG_BLOCK)
        lds32[x] = 0.;
    Tile32 T;
    T.lds    = lds32;
    T.stride = A.tile.stride;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T.ext[d]   = A.tile.ext[d];
        T.pitch[d] = A.tile.pitch[d];
        T.org[d] = 0;
    }
    pc::segTileOrigin<D>(G, A.sg, A.tile, seg, T.org);
    Push32 K;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        K.hdtdx[d] = static_cast<float>(A.hdtdx[d]);
    K.coef1 = static_cast<float>(A.charge * A.dto2m);
    K.coef  = static_cast<float>(A.coef);
    __syncthreads();

    Win32<D, O> win;
    if constexpr (WIN)
        winZero(win);
    for (unsigned int r = rA + (t >> 6); r < rB; r += pc::SEG_NW)
        pushSlot32<D, O, WIN>(A, K, v, off + static_cast<unsigned long long>(r) * LANES + lane, T, win);
    if constexpr (WIN)
        winFlush<D, O>(win, G, T, A.mom, A.charge);
    __syncthreads();
    flushTile32<D, pc::SEG_BLOCK>(A, T);
}

template<int D, int TB>
__device__ __forceinline__ void flushTile32(PushArgs const& A, Tile32 const& T)
{
    Geom const& G = A.G;
    int const t   = threadIdx.x;
    // flush: one HBM atomic per touched node and moment
    int const e12 = T.pitch[1] * T.pitch[2];
    for (int x = t; x < A.tile.nodes; x += TB)
    {
        int node = -1;
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            double const val = T.mom(m)[x];
            if (val != 0.)
            {
                if (node < 0)
                {
                    int const ti = x / e12;
                    int const rr = x - ti * e12;
                    int const tj = rr / T.pitch[2];
                    int const tk = rr - tj * T.pitch[2];
                    node = G.idx(QMOM, T.org[0] + ti, D > 1 ? T.org[1] + tj : 0,
                                 D > 2 ? T.org[2] + tk : 0);
                }
                if (m == 0)
                {
                    atomicAdd(A.mom[0] + node, val);
                    atomicAdd(A.mom[1] + node, A.charge * val);
                }
                else
                    atomicAdd(A.mom[m + 1] + node, val);
            }
        }
    }
}
} // namespace

// one workgroup per work item (a segment and its share of the segment's rows, at most
// sg.rowsPerItem)
template<int D, int O, bool WIN>
__global__ __launch_bounds__(pc::SEG_BLOCK)
__attribute__((amdgpu_waves_per_eu(WIN ? PHX_SEG_WAVES32_WIN : PHX_SEG_WAVES32)))
void k_push_seg32(PushArgs A, PView32 v)
{
    extern __shared__ double lds32[];
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
        pushPiece32<D, O, WIN>(A, v, lds32, seg, r0, rEnd);
    }
    if (A.kt)
    {
        __syncthreads();
        pc::ktEnd(A.kt, nItems);
    }
}

template<int D, int O>
static void launchPush32T(PushArgs const& A, PView32 const& v, hipStream_t s)
{
    if constexpr (winFits<D, O>())
        if (A.depositWindow)
        {
            k_push_seg32<D, O, true><<<A.st.itemsCap, pc::SEG_BLOCK, A.tile.bytes, s>>>(A, v);
            return;
        }
    k_push_seg32<D, O, false><<<A.st.itemsCap, pc::SEG_BLOCK, A.tile.bytes, s>>>(A, v);
}

// ---------------------------------------------------------------------------------------------
// fused push + deposit over brick-ordered linear FP32 storage (3D, the staged 3D path), in place:
// chunks of 4096 particles, 16 per thread, the tile over the brick of the chunk's first particle
// (the storage of the FP64 family's linear 3D push, in FP32)
// ---------------------------------------------------------------------------------------------
namespace
{
constexpr int LIN32_PPT = PHX_PPT_3D; // the FP64 linear 3D push's, so that both share the brick sort's interleave

template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_push32_lin(PushArgs A, PView32 v)
{
    extern __shared__ double lds32[];
    if (A.G.stopped())
        return;
    Geom const& G = A.G;
    int const t   = threadIdx.x;
    unsigned long long const base = static_cast<unsigned long long>(blockIdx.x) * (BLOCK * LIN32_PPT);
    for (int x = t; x < 4 * A.tile.stride; x += BLOCK)
        lds32[x] = 0.;
    Tile32 T;
    T.lds    = lds32;
    T.stride = A.tile.stride;
    int const bk = A.tile.brick;
    bool const far = v.ic[0][base] == DEAD_CELL;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T.ext[d]   = A.tile.ext[d];
        T.pitch[d] = A.tile.pitch[d];
        T.org[d]   = 0;
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int first = v.ic[d][base] - G.lower[d];
        first     = (first >= 0 ? first / bk : -((bk - 1 - first) / bk)) * bk;
        T.org[d]  = far ? -0x20000000 : first + G.g - A.tile.lo[d];
    }
    Push32 K;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        K.hdtdx[d] = static_cast<float>(A.hdtdx[d]);
    K.coef1 = static_cast<float>(A.charge * A.dto2m);
    K.coef  = static_cast<float>(A.coef);
    __syncthreads();
    Win32<D, O> win;
#pragma unroll 1
    for (int k = 0; k < LIN32_PPT; ++k)
    {
        unsigned long long const i = base + static_cast<unsigned long long>(k) * BLOCK + t;
        if (i < A.n)
            pushSlot32<D, O, false>(A, K, v, i, T, win);
    }
    __syncthreads();
    flushTile32<D, BLOCK>(A, T);
}
} // namespace

// the same over cell-segmented FP32 storage, one workgroup per segment with all its rows (the
// staged 3D path's layout after its sorts; the FP64 family's k_push_segwg in FP32)
template<int D, int O>
__global__ __launch_bounds__(BLOCK) void k_push32_segwg(PushArgs A, PView32 v)
{
    extern __shared__ double lds32[];
    if (A.G.stopped())
        return;
    Geom const& G          = A.G;
    unsigned int const seg = pc::xcdItem(blockIdx.x, gridDim.x);
    int const t            = threadIdx.x;
    int const lane         = t & 63;
    unsigned long long const off = A.st.off[seg];
    unsigned int const R         = A.st.rows[seg];
    for (int x = t; x < 4 * A.tile.stride; x += BLOCK)
        lds32[x] = 0.;
    Tile32 T;
    T.lds    = lds32;
    T.stride = A.tile.stride;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T.ext[d]   = A.tile.ext[d];
        T.pitch[d] = A.tile.pitch[d];
        T.org[d]   = 0;
    }
    pc::segTileOrigin<D>(G, A.sg, A.tile, seg, T.org);
    Push32 K;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        K.hdtdx[d] = static_cast<float>(A.hdtdx[d]);
    K.coef1 = static_cast<float>(A.charge * A.dto2m);
    K.coef  = static_cast<float>(A.coef);
    __syncthreads();
    Win32<D, O> win;
#pragma unroll 1
    for (unsigned int r = static_cast<unsigned int>(t >> 6); r < R; r += NW)
        pushSlot32<D, O, false>(A, K, v, off + static_cast<unsigned long long>(r) * LANES + lane, T, win);
    __syncthreads();
    flushTile32<D, BLOCK>(A, T);
}

void launchPush32SegWG(PushArgs const& A, PView32 const& v, hipStream_t s)
{
    unsigned int const grid = static_cast<unsigned int>(A.sg.nsegs) + 1u;
    switch (A.G.dim * 10 + A.G.order)
    {
        case 31: k_push32_segwg<3, 1><<<grid, BLOCK, A.tile.bytes, s>>>(A, v); break;
        case 32: k_push32_segwg<3, 2><<<grid, BLOCK, A.tile.bytes, s>>>(A, v); break;
        case 33: k_push32_segwg<3, 3><<<grid, BLOCK, A.tile.bytes, s>>>(A, v); break;
        default: break;
    }
}

TileSpec brickTileShape32(int order)
{
    TileSpec T{};
    T.copies = 1;
    T.brick  = BRICK;
    for (int d = 0; d < 3; ++d)
    {
        T.lo[d]    = 2;
        T.ext[d]   = BRICK + 2 + order + 2;
        T.pitch[d] = T.ext[d];
    }
    T.nodes  = T.ext[0] * T.ext[1] * T.ext[2];
    T.stride = (T.nodes + 15) / 16 * 16;
    T.bytes  = static_cast<unsigned int>(8 * 4 * T.stride);
    return T;
}

void launchPush32Linear(PushArgs const& A, PView32 const& v, hipStream_t s)
{
    if (A.n == 0 || A.G.dim != 3)
        return;
    unsigned int const grid = static_cast<unsigned int>((A.n + BLOCK * LIN32_PPT - 1) / (BLOCK * LIN32_PPT));
    switch (A.G.order)
    {
        case 1: k_push32_lin<3, 1><<<grid, BLOCK, A.tile.bytes, s>>>(A, v); break;
        case 2: k_push32_lin<3, 2><<<grid, BLOCK, A.tile.bytes, s>>>(A, v); break;
        case 3: k_push32_lin<3, 3><<<grid, BLOCK, A.tile.bytes, s>>>(A, v); break;
    }
}

void launchPush32(PushArgs const& A, PView32 const& v, hipStream_t s)
{
    switch (A.G.dim * 10 + A.G.order)
    {
        case 11: launchPush32T<1, 1>(A, v, s); break;
        case 12: launchPush32T<1, 2>(A, v, s); break;
        case 13: launchPush32T<1, 3>(A, v, s); break;
        case 21: launchPush32T<2, 1>(A, v, s); break;
        case 22: launchPush32T<2, 2>(A, v, s); break;
        case 23: launchPush32T<2, 3>(A, v, s); break;
        case 31: launchPush32T<3, 1>(A, v, s); break;
        case 32: launchPush32T<3, 2>(A, v, s); break;
        case 33: launchPush32T<3, 3>(A, v, s); break;
    }
}

TileSpec segTileShape32(int dim, int order, SegGeom const& sg)
{
    TileSpec T{};
    T.copies      = 1;
    int const mlo = pc::segMarginLo(order);
    int const mhi = pc::segMarginHi(order);
    for (int d = 0; d < 3; ++d)
    {
        T.lo[d]  = d < dim ? mlo : 0;
        T.ext[d] = d < dim ? sg.maxSeg[d] + mlo + mhi : 1;
    }
    pc::tilePitch(T, dim, sg.rowPad);
    T.nodes  = T.ext[0] * T.pitch[1] * T.pitch[2];
    T.stride = (T.nodes + 15) / 16 * 16;
    T.bytes  = static_cast<unsigned int>(8 * 4 * T.stride);
    if (T.bytes > pc::TILE_BYTES_MAX)
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
        T.bytes  = static_cast<unsigned int>(8 * 4 * T.stride);
    }
    return T;
}

} // namespace phx
