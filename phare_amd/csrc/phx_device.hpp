// phx_device.hpp -- device building blocks for gfx950: Yee index space, B-spline weights,
// field gather and the Boris step. Arithmetic follows the reference's operation order and is
// compiled without FMA contraction, so an FP64 push reproduces the reference bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace phx
{

// centering: 0 primal, 1 dual
enum : int
{
    CP = 0,
    CD = 1
};

// quantity ids: 0..2 B, 3..5 E, 6..8 J, 9 all-primal (moments, Pe, Vi)
enum : int
{
    QBX = 0,
    QBY,
    QBZ,
    QEX,
    QEY,
    QEZ,
    QJX,
    QJY,
    QJZ,
    QMOM,
    NQ
};

// Yee hybrid centering table (src/core/data/grid/impl/yee/gridlayout_hybrid_yee.hpp:51-62)
__host__ __device__ constexpr int centering(int q, int dir)
{
    switch (q)
    {
        case QBX: return dir == 0 ? CP : CD;
        case QBY: return dir == 1 ? CP : CD;
        case QBZ: return dir == 2 ? CP : CD;
        case QEX:
        case QJX: return dir == 0 ? CD : CP;
        case QEY:
        case QJY: return dir == 1 ? CD : CP;
        case QEZ:
        case QJZ: return dir == 2 ? CD : CP;
        default: return CP;
    }
}

// Patch geometry in the GridLayout's local index space (gridlayout.hpp:1349-1440):
// physical nodes start at g for both centerings; primal has n+1, dual n physical nodes.
struct Geom
{
    int dim, order;
    int n[3];     // physical cells (1 for unused dimensions)
    int lower[3]; // AMR lower cell
    int g, pg;    // field / particle ghost widths ({2,4,4}, {1,2,2})
    double dx[3];
    double inv[3]; // 1./dx
    int shape[NQ][3];
    // local cells from which every E/B gather stencil lies inside its array, whatever delta
    int safeLo[3], safeHi[3];
    // the context's first-error word (DevError::code): once a push has raised an error, every
    // kernel given this geometry returns at entry, so that the state stays as the failing stage
    // left it (the reference throws out of SolverPPC::moveIons_, solver_ppc.hpp:558-563)
    unsigned int const* stop;

    __host__ __device__ bool stopped() const { return stop && *stop != 0u; }

    __host__ __device__ int size(int q) const { return shape[q][0] * shape[q][1] * shape[q][2]; }
    __host__ __device__ int idx(int q, int i, int j, int k) const
    {
        return (i * shape[q][1] + j) * shape[q][2] + k;
    }
    __host__ __device__ int physEnd(int q, int d) const { return g + n[d] - centering(q, d); }
    __host__ __device__ int ghostEnd(int q, int d) const { return physEnd(q, d) + g; }
};

// ---------------------------------------------------------------------------------------------
// B-spline weights (interpolator.hpp:54-125) and start shift (interpolator.hpp:519-549)
// ---------------------------------------------------------------------------------------------
template<int O>
__device__ __forceinline__ int startShift(int c, double delta)
{
    if constexpr (O == 1)
        return c == CP ? 0 : (delta < .5 ? 1 : 0);
    else if constexpr (O == 2)
        return c == CP ? (delta < .5 ? 1 : 0) : 1;
    else
        return c == CP ? 1 : (delta < .5 ? 2 : 1);
}

template<int O>
__device__ __forceinline__ void bsplineWeights(double np, int start, double* w)
{
    if constexpr (O == 1)
    {
        w[1] = np - static_cast<double>(start);
        w[0] = 1. - w[1];
    }
    else if constexpr (O == 2)
    {
        double delta = static_cast<double>(start + 1) - np;
        double c1    = 0.5 + delta;
        double c3    = 0.5 - delta;
        w[0]         = 0.5 * c1 * c1;
        w[1]         = 0.75 - delta * delta;
        w[2]         = 0.5 * c3 * c3;
    }
    else
    {
        double const f43 = 4. / 3.;
        double const f23 = 2. / 3.;
        double index     = static_cast<double>(start) - np;
        double c1        = 1. + 0.5 * index;
        double c2        = index + 1;
        double c3        = index + 2;
        double c4        = 1. - 0.5 * (index + 3);
        double c2s       = c2 * c2;
        double c2c       = c2s * c2;
        double c3s       = c3 * c3;
        double c3c       = c3s * c3;
        w[0]             = f43 * c1 * c1 * c1;
        w[1]             = f23 - c2s - 0.5 * c2c;
        w[2]             = f23 - c3s + 0.5 * c3c;
        w[3]             = f43 * c4 * c4 * c4;
    }
}

// start index and weights of one direction for a centering (interpolator.hpp:379-406)
template<int O>
__device__ __forceinline__ int indexAndWeights(int c, int iclLocal, double delta, double* w)
{
    int start = iclLocal - startShift<O>(c, delta);
    double np = static_cast<double>(iclLocal) + delta;
    if (c == CD)
        np -= .5;
    bsplineWeights<O>(np, start, w);
    return start;
}

// ---------------------------------------------------------------------------------------------
// Gather (interpolator.hpp:420-456, MeshToParticle :152-264)
// ---------------------------------------------------------------------------------------------
template<int D, int O>
struct Stencil
{
    int sp[D], sd[D];          // primal / dual start index per direction
    double wp[D][O + 1];       // primal weights
    double wd[D][O + 1];       // dual weights
};

template<int D, int O>
__device__ __forceinline__ void makeStencil(Geom const& G, int const* cell, double const* delta,
                                            Stencil<D, O>& s)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int icl  = cell[d] - (G.lower[d] - G.g);
        s.sd[d]  = indexAndWeights<O>(CD, icl, delta[d], s.wd[d]);
        s.sp[d]  = indexAndWeights<O>(CP, icl, delta[d], s.wp[d]);
    }
}

// range of start shifts of a centering over delta in [0,1) (the shift table of startShift)
template<int O>
__host__ __device__ constexpr int minShift(int c)
{
    return O == 1 ? 0 : (O == 2 ? (c == CP ? 0 : 1) : 1);
}
template<int O>
__host__ __device__ constexpr int maxShift(int c)
{
    return O == 1 ? (c == CP ? 0 : 1) : (O == 2 ? 1 : (c == CP ? 1 : 2));
}

// true when every node the stencil touches for quantity q is inside its array
template<int D, int O>
__device__ __forceinline__ bool stencilInside(Geom const& G, Stencil<D, O> const& s, int q)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        int st = centering(q, d) == CP ? s.sp[d] : s.sd[d];
        if (st < 0 || st + O >= G.shape[q][d])
            return false;
    }
    return true;
}

// start index and weights of quantity q's stencil
template<int D, int O>
__device__ __forceinline__ void stencilOf(int q, Stencil<D, O> const& s, int* st, double const** w)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        bool p = centering(q, d) == CP;
        st[d]  = p ? s.sp[d] : s.sd[d];
        w[d]   = p ? s.wp[d] : s.wd[d];
    }
}

// interpolation of a C-order array with row strides s1 (dimension 1) and s2 (dimension 2) from
// the stencil starting at st, in the reference's order of sums (MeshToParticle)
template<int D, int O>
__device__ __forceinline__ double interpStrided(double const* __restrict__ f, int const* st,
                                                double const* const* w, int s1, int s2)
{
    constexpr int S = O + 1;
    if constexpr (D == 1)
    {
        double r = 0.;
#pragma unroll
        for (int i = 0; i < S; ++i)
            r += f[st[0] + i] * w[0][i];
        return r;
    }
    else if constexpr (D == 2)
    {
        double r = 0.;
#pragma unroll
        for (int i = 0; i < S; ++i)
        {
            double yi         = 0.;
            double const* row = f + (st[0] + i) * s1 + st[1];
#pragma unroll
            for (int j = 0; j < S; ++j)
                yi += row[j] * w[1][j];
            r += yi * w[0][i];
        }
        return r;
    }
    else
    {
        double r = 0.;
#pragma unroll
        for (int i = 0; i < S; ++i)
        {
            double yi = 0.;
#pragma unroll
            for (int j = 0; j < S; ++j)
            {
                double zi           = 0.;
                double const* pencil = f + ((st[0] + i) * s1 + st[1] + j) * s2 + st[2];
#pragma unroll
                for (int k = 0; k < S; ++k)
                    zi += pencil[k] * w[2][k];
                yi += zi * w[1][j];
            }
            r += yi * w[0][i];
        }
        return r;
    }
}

template<int D, int O>
__device__ __forceinline__ double interp(Geom const& G, double const* __restrict__ f, int q,
                                         Stencil<D, O> const& s)
{
    int st[D];
    double const* w[D];
    stencilOf<D, O>(q, s, st, w);
    return interpStrided<D, O>(f, st, w, G.shape[q][1], G.shape[q][2]);
}

struct EMPtrs
{
    double const* E[3];
    double const* B[3];
};

// ---------------------------------------------------------------------------------------------
// Boris rotation (boris.hpp:240-300)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void borisAccelerate(double* v, double const* E, double const* B,
                                                double charge, double dto2m)
{
    double const coef1 = charge * dto2m;
    double vx1         = v[0] + coef1 * E[0];
    double vy1         = v[1] + coef1 * E[1];
    double vz1         = v[2] + coef1 * E[2];
    double const rx    = coef1 * B[0];
    double const ry    = coef1 * B[1];
    double const rz    = coef1 * B[2];
    double const rx2   = rx * rx;
    double const ry2   = ry * ry;
    double const rz2   = rz * rz;
    double const rxry  = rx * ry;
    double const rxrz  = rx * rz;
    double const ryrz  = ry * rz;
    double const invDet = 1. / (1. + rx2 + ry2 + rz2);
    double const mxx   = 1. + rx2 - ry2 - rz2;
    double const mxy   = 2. * (rxry + rz);
    double const mxz   = 2. * (rxrz - ry);
    double const myx   = 2. * (rxry - rz);
    double const myy   = 1. + ry2 - rx2 - rz2;
    double const myz   = 2. * (ryrz + rx);
    double const mzx   = 2. * (rxrz + ry);
    double const mzy   = 2. * (ryrz - rx);
    double const mzz   = 1. + rz2 - rx2 - ry2;
    double const vx2   = (mxx * vx1 + mxy * vy1 + mxz * vz1) * invDet;
    double const vy2   = (myx * vx1 + myy * vy1 + myz * vz1) * invDet;
    double const vz2   = (mzx * vx1 + mzy * vy1 + mzz * vz1) * invDet;
    v[0]               = vx2 + coef1 * E[0];
    v[1]               = vy2 + coef1 * E[1];
    v[2]               = vz2 + coef1 * E[2];
}

// half position step (boris.hpp:156-172); false when |delta| > 2 (MoveTwoCellException)
template<int D>
__device__ __forceinline__ bool advancePosition(int* cell, double* delta, double const* v,
                                                double const* hdtdx, double& badDelta, double& badV)
{
#pragma unroll
    for (int d = 0; d < D; ++d)
    {
        double const dl = delta[d] + hdtdx[d] * v[d];
        if (fabs(dl) > 2)
        {
            badDelta = dl;
            badV     = v[d];
            return false;
        }
        int const ic = static_cast<int>(floor(dl));
        delta[d]     = dl - ic;
        cell[d] += ic;
    }
    return true;
}

// first-error-wins record of a device error (particle, values, stage)
struct DevError
{
    unsigned int code; // 0 none, else a phx_status
    unsigned int stage;
    unsigned long long particle;
    double delta, v;
    int kind, pop;
};

__device__ __forceinline__ void raiseError(DevError* e, unsigned code, unsigned stage,
                                           unsigned long long i, double delta, double v, int kind,
                                           int pop)
{
    if (atomicCAS(&e->code, 0u, code) == 0u)
    {
        e->stage    = stage;
        e->particle = i;
        e->delta    = delta;
        e->v        = v;
        e->kind     = kind;
        e->pop      = pop;
        __threadfence();
    }
}

} // namespace phx
