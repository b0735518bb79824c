// phx_fields.hip -- Yee-grid kernels for gfx950: Faraday (faraday.hpp:27-95), Ampere
// (ampere.hpp:26-94), the electron closure (electrons.hpp:102-128, 202-212), Ohm
// (ohm.hpp:48-269), the predictor average (algorithm.hpp:68-76), the ion totals
// (ions.hpp:75-144) and the same-level fills of one periodic patch (ghost copy, border sum of
// population moments, border max of the totals; hybrid_hybrid_messenger_strategy.hpp:376-497).
//
// One thread per node, FP64, arrays in the reference's C-order layout with ghost nodes. The
// stencils read neighbours straight from HBM through the L1/L2 caches.
#include "phx_kernels.hpp"

#include <algorithm>

namespace phx
{

namespace
{
constexpr int FB = 128;

// the patch geometry with its dimension known at compile time: the helpers below take either
// and fold every per-direction loop and projection when it is a GeomD
template<int DIM>
struct GeomD : Geom
{
    static constexpr int dim = DIM;
    __host__ __device__ explicit GeomD(Geom const& g)
        : Geom(g)
    {
    }
};

template<class GT>
__device__ __forceinline__ bool nodeOf(GT const& G, int q, long t, int& i, int& j, int& k)
{
    if (G.stopped())
        return false;
    int const s1 = G.shape[q][1], s2 = G.shape[q][2];
    long const plane = static_cast<long>(s1) * s2;
    i                = static_cast<int>(t / plane);
    long r           = t - static_cast<long>(i) * plane;
    j                = static_cast<int>(r / s2);
    k                = static_cast<int>(r - static_cast<long>(j) * s2);
    return i < G.shape[q][0];
}

__device__ __forceinline__ int coord(int d, int i, int j, int k) { return d == 0 ? i : (d == 1 ? j : k); }

// inside [lo, hi] of quantity q in the directions < dim
template<class GT>
__device__ __forceinline__ bool inRange(GT const& G, int i, int j, int k, int const* lo, int const* hi)
{
    for (int d = 0; d < G.dim; ++d)
    {
        int c = coord(d, i, j, k);
        if (c < lo[d] || c > hi[d])
            return false;
    }
    return true;
}

// GridLayout::deriv (gridlayout.hpp:571-632): operand centering decides next/prev; f(i, j, k)
// gives the operand at a node
template<class GT, class F>
__device__ __forceinline__ double derivFn(GT const& G, F const& f, int q, int dir, int i, int j,
                                          int k)
{
    int const c = centering(q, dir);
    int in = i, jn = j, kn = k, ip = i, jp = j, kp = k;
    int const up = c == CP ? 1 : 0;
    int const dn = c == CP ? 0 : -1;
    if (dir == 0)
    {
        in = i + up;
        ip = i + dn;
    }
    else if (dir == 1)
    {
        jn = j + up;
        jp = j + dn;
    }
    else
    {
        kn = k + up;
        kp = k + dn;
    }
    return G.inv[dir] * (f(in, jn, kn) - f(ip, jp, kp));
}

template<class GT>
__device__ __forceinline__ double deriv(GT const& G, double const* f, int q, int dir, int i, int j,
                                        int k)
{
    return derivFn(G, [&](int a, int b, int cc) { return f[G.idx(q, a, b, cc)]; }, q, dir, i, j, k);
}

// GridLayout::laplacian (gridlayout.hpp:638-700)
template<class GT>
__device__ __forceinline__ double laplacian(GT const& G, double const* f, int q, int i, int j, int k)
{
    double res    = 0.;
    double const h = f[G.idx(q, i, j, k)];
    for (int d = 0; d < G.dim; ++d)
    {
        double nx = f[G.idx(q, i + (d == 0), j + (d == 1), k + (d == 2))];
        double px = f[G.idx(q, i - (d == 0), j - (d == 1), k - (d == 2))];
        double lap = G.inv[d] * G.inv[d] * (nx - 2.0 * h + px);
        res        = d == 0 ? lap : res + lap;
    }
    return res;
}

// GridLayout::project with the Yee tables (gridlayout_hybrid_yee.hpp:351-446): per direction
// none (identity), P2D (primal->dual: offsets 0,+1) or D2P (dual->primal: -1,0), coefficients
// 0.5 each, tensor order X outer, Z inner, accumulated from 0.
enum : int
{
    PN  = 0,
    P2D = 1,
    D2P = 2
};
template<class GT, class F>
__device__ __forceinline__ double projectFn(GT const& G, F const& f, int i, int j, int k, int mx,
                                            int my, int mz)
{
    int m[3] = {mx, my, mz};
    int off[3][2];
    double cf[3][2];
    int np[3];
    for (int d = 0; d < 3; ++d)
    {
        int mode = d < G.dim ? m[d] : PN;
        if (mode == PN)
        {
            np[d]     = 1;
            off[d][0] = 0;
            cf[d][0]  = 1.0;
        }
        else
        {
            np[d]     = 2;
            int base  = mode == P2D ? 1 : 0;
            off[d][0] = base - 1;
            off[d][1] = base;
            cf[d][0]  = 0.5;
            cf[d][1]  = 0.5;
        }
    }
    double r = 0.;
    for (int a = 0; a < np[0]; ++a)
        for (int b = 0; b < np[1]; ++b)
            for (int c = 0; c < np[2]; ++c)
                r += cf[0][a] * cf[1][b] * cf[2][c] * f(i + off[0][a], j + off[1][b], k + off[2][c]);
    return r;
}

template<class GT>
__device__ __forceinline__ double project(GT const& G, double const* f, int q, int i, int j, int k,
                                          int mx, int my, int mz)
{
    return projectFn(G, [&](int a, int b, int cc) { return f[G.idx(q, a, b, cc)]; }, i, j, k, mx,
                     my, mz);
}

template<class GT>
__device__ __forceinline__ void physBox(GT const& G, int q, int* lo, int* hi)
{
    for (int d = 0; d < 3; ++d)
    {
        lo[d] = d < G.dim ? G.g : 0;
        hi[d] = d < G.dim ? G.physEnd(q, d) : 0;
    }
}

// ---- Faraday ------------------------------------------------------------------------------
template<int DIM>
__global__ __launch_bounds__(FB) void k_faraday(Geom G0, CVecPtr B, CVecPtr E, VecPtr Bn, double dt)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    // iterate over the primal-in-every-direction shape (the largest), each component on its box
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const D = G.dim;
    int lo[3], hi[3];
    physBox(G, QBX, lo, hi);
    if (inRange(G, i, j, k, lo, hi))
    {
        int const n = G.idx(QBX, i, j, k);
        double r;
        if (D == 1)
            r = B.c[0][n];
        else if (D == 2)
            r = B.c[0][n] - dt * deriv(G, E.c[2], QEZ, 1, i, j, k);
        else
            r = B.c[0][n] - dt * deriv(G, E.c[2], QEZ, 1, i, j, k) + dt * deriv(G, E.c[1], QEY, 2, i, j, k);
        Bn.c[0][n] = r;
    }
    physBox(G, QBY, lo, hi);
    if (inRange(G, i, j, k, lo, hi))
    {
        int const n = G.idx(QBY, i, j, k);
        double r;
        if (D < 3)
            r = B.c[1][n] + dt * deriv(G, E.c[2], QEZ, 0, i, j, k);
        else
            r = B.c[1][n] - dt * deriv(G, E.c[0], QEX, 2, i, j, k) + dt * deriv(G, E.c[2], QEZ, 0, i, j, k);
        Bn.c[1][n] = r;
    }
    physBox(G, QBZ, lo, hi);
    if (inRange(G, i, j, k, lo, hi))
    {
        int const n = G.idx(QBZ, i, j, k);
        double r;
        if (D == 1)
            r = B.c[2][n] - dt * deriv(G, E.c[1], QEY, 0, i, j, k);
        else
            r = B.c[2][n] - dt * deriv(G, E.c[1], QEY, 0, i, j, k) + dt * deriv(G, E.c[0], QEX, 1, i, j, k);
        Bn.c[2][n] = r;
    }
}

// ---- Ampere (ghost box shrunk by one) -------------------------------------------------------
template<int DIM>
__global__ __launch_bounds__(FB) void k_ampere(Geom G0, CVecPtr B, VecPtr J)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const D = G.dim;
    for (int c = 0; c < 3; ++c)
    {
        int q = QJX + c;
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d)
        {
            lo[d] = d < D ? 1 : 0;
            hi[d] = d < D ? G.ghostEnd(q, d) - 1 : 0;
        }
        if (!inRange(G, i, j, k, lo, hi))
            continue;
        double r;
        if (c == 0)
        {
            if (D == 1)
                r = 0.0;
            else if (D == 2)
                r = deriv(G, B.c[2], QBZ, 1, i, j, k);
            else
                r = deriv(G, B.c[2], QBZ, 1, i, j, k) - deriv(G, B.c[1], QBY, 2, i, j, k);
        }
        else if (c == 1)
        {
            if (D < 3)
                r = -deriv(G, B.c[2], QBZ, 0, i, j, k);
            else
                r = deriv(G, B.c[0], QBX, 2, i, j, k) - deriv(G, B.c[2], QBZ, 0, i, j, k);
        }
        else
        {
            if (D == 1)
                r = deriv(G, B.c[1], QBY, 0, i, j, k);
            else
                r = deriv(G, B.c[1], QBY, 0, i, j, k) - deriv(G, B.c[0], QBX, 1, i, j, k);
        }
        J.c[c][G.idx(q, i, j, k)] = r;
    }
}

// ---- electrons: Ve on the physical primal box, Pe everywhere ---------------------------------
template<int DIM>
__global__ __launch_bounds__(FB) void k_electrons(Geom G0, double Te, CVecPtr J, CVecPtr Vi,
                                                   double const* Ne, VecPtr Ve, double* Pe)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    // every load before the first store (a store that may alias a later load would serialize them)
    int const n     = G.idx(QMOM, i, j, k);
    double const ne = Ne[n];
    int lo[3], hi[3];
    physBox(G, QMOM, lo, hi);
    if (!inRange(G, i, j, k, lo, hi))
    {
        Pe[n] = ne * Te;
        return;
    }
    // JxToMoments / JyToMoments / JzToMoments: dual->primal along the component's direction
    double const jx = project(G, J.c[0], QJX, i, j, k, D2P, PN, PN);
    double const jy = project(G, J.c[1], QJY, i, j, k, PN, D2P, PN);
    double const jz = project(G, J.c[2], QJZ, i, j, k, PN, PN, D2P);
    double const vx = Vi.c[0][n], vy = Vi.c[1][n], vz = Vi.c[2][n];
    Pe[n]           = ne * Te;
    Ve.c[0][n]      = vx - jx / ne;
    Ve.c[1][n]      = vy - jy / ne;
    Ve.c[2][n]      = vz - jz / ne;
}

// ---- Ohm --------------------------------------------------------------------------------------
// component c of E at a node of its physical box (Ohm::operator(), ohm.hpp:48-269): ideal
// -(Ve x B), pressure -grad(Pe)/n, resistive eta J, hyper-resistive (constant or spatial).
// ve(d, i, j, k) and pe(i, j, k) give the electron fields at primal nodes.
template<class GT, class VeF, class PeF>
__device__ __forceinline__ double ohmAt(GT const& G, int c, int i, int j, int k, double eta,
                                        double nu, int hyper, double lvlCoef, double const* nn,
                                        VeF const& ve, PeF const& pe, CVecPtr B, CVecPtr J)
{
    int const D = G.dim;
    auto veOf   = [&](int d) { return [&, d](int a, int b, int cc) { return ve(d, a, b, cc); }; };
    // moments -> E_c: primal->dual along c
    int const mx = c == 0 ? P2D : PN, my = c == 1 ? P2D : PN, mz = c == 2 ? P2D : PN;
    double ideal, press = 0., bxE, byE, bzE;
    if (c == 0)
    {
        double const vy = projectFn(G, veOf(1), i, j, k, mx, my, mz);
        double const vz = projectFn(G, veOf(2), i, j, k, mx, my, mz);
        byE             = project(G, B.c[1], QBY, i, j, k, PN, PN, D2P); // ByToEx
        bzE             = project(G, B.c[2], QBZ, i, j, k, PN, D2P, PN); // BzToEx
        ideal           = -vy * bzE + vz * byE;
    }
    else if (c == 1)
    {
        double const vx = projectFn(G, veOf(0), i, j, k, mx, my, mz);
        double const vz = projectFn(G, veOf(2), i, j, k, mx, my, mz);
        bxE             = project(G, B.c[0], QBX, i, j, k, PN, PN, D2P); // BxToEy
        bzE             = project(G, B.c[2], QBZ, i, j, k, D2P, PN, PN); // BzToEy
        ideal           = -vz * bxE + vx * bzE;
    }
    else
    {
        double const vx = projectFn(G, veOf(0), i, j, k, mx, my, mz);
        double const vy = projectFn(G, veOf(1), i, j, k, mx, my, mz);
        bxE             = project(G, B.c[0], QBX, i, j, k, PN, D2P, PN); // BxToEz
        byE             = project(G, B.c[1], QBY, i, j, k, D2P, PN, PN); // ByToEz
        ideal           = -vx * byE + vy * bxE;
    }
    if (c < D)
    {
        double const nOnE = project(G, nn, QMOM, i, j, k, mx, my, mz);
        double const grad = derivFn(G, pe, QMOM, c, i, j, k);
        press             = -grad / nOnE;
    }
    int const n       = G.idx(QEX + c, i, j, k);
    double const res  = eta * (0. + 1.0 * J.c[c][n]);
    double const lapJ = laplacian(G, J.c[c], QJX + c, i, j, k);
    double hyp;
    if (hyper == 0)
        hyp = -nu * lapJ;
    else
    {
        // BxToE_c, ByToE_c, BzToE_c and momentsToE_c (gridlayout_hybrid_yee.hpp:450-745)
        double bx, by, bz;
        if (c == 0)
        {
            bx = project(G, B.c[0], QBX, i, j, k, P2D, D2P, D2P);
            by = project(G, B.c[1], QBY, i, j, k, PN, PN, D2P);
            bz = project(G, B.c[2], QBZ, i, j, k, PN, D2P, PN);
        }
        else if (c == 1)
        {
            bx = project(G, B.c[0], QBX, i, j, k, PN, PN, D2P);
            by = project(G, B.c[1], QBY, i, j, k, D2P, P2D, D2P);
            bz = project(G, B.c[2], QBZ, i, j, k, D2P, PN, PN);
        }
        else
        {
            bx = project(G, B.c[0], QBX, i, j, k, PN, D2P, PN);
            by = project(G, B.c[1], QBY, i, j, k, D2P, PN, PN);
            bz = project(G, B.c[2], QBZ, i, j, k, D2P, D2P, P2D);
        }
        double const nOnE = project(G, nn, QMOM, i, j, k, mx, my, mz);
        double const b    = sqrt(bx * bx + by * by + bz * bz);
        hyp               = -nu * (b / (nOnE + 0.1) + 1) * lvlCoef * lapJ;
    }
    return ideal + press + res + hyp;
}

// node (i0, j0, k0) of E_c: false outside its array (or, without fill, outside its physical
// box); else (i, j, k) is where Ohm is evaluated: the node itself, or for a ghost node its
// periodic source
template<class GT>
__device__ __forceinline__ bool ohmNode(GT const& G, int c, int fill, int i0, int j0, int k0, int& i,
                                        int& j, int& k)
{
    int const q = QEX + c;
    i = i0;
    j = j0;
    k = k0;
    if (fill)
    {
        // ghost layers too: a ghost node gets Ohm evaluated at its periodic source
        if (i0 >= G.shape[q][0] || j0 >= G.shape[q][1] || k0 >= G.shape[q][2])
            return false;
        int s[3] = {i0, j0, k0};
        for (int d = 0; d < G.dim; ++d)
        {
            if (s[d] < G.g)
                s[d] += G.n[d];
            else if (s[d] > G.physEnd(q, d))
                s[d] -= G.n[d];
        }
        i = s[0];
        j = s[1];
        k = s[2];
        return true;
    }
    int lo[3], hi[3];
    physBox(G, q, lo, hi);
    return inRange(G, i, j, k, lo, hi);
}

template<int DIM>
__global__ __launch_bounds__(FB) void k_ohm(Geom G0, double eta, double nu, int hyper, double lvlCoef,
                                            double const* nn, CVecPtr Ve, double const* Pe, CVecPtr B,
                                            CVecPtr J, VecPtr E, int fill)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i0, j0, k0;
    if (!nodeOf(G, QMOM, t, i0, j0, k0))
        return;
    int const c = static_cast<int>(blockIdx.y); // one component per grid row
    int i, j, k;
    if (!ohmNode(G, c, fill, i0, j0, k0, i, j, k))
        return;
    auto ve = [&](int d, int a, int b, int cc) { return Ve.c[d][G.idx(QMOM, a, b, cc)]; };
    auto pe = [&](int a, int b, int cc) { return Pe[G.idx(QMOM, a, b, cc)]; };
    E.c[c][G.idx(QEX + c, i0, j0, k0)] = ohmAt(G, c, i, j, k, eta, nu, hyper, lvlCoef, nn, ve, pe, B, J);
}

struct Mirror
{
    float* c[6];
};

// Ohm of a periodic level's predictor and SolverPPC::average_ (solver_ppc.hpp:483-510) in one
// launch: Ohm is evaluated at every E node (a ghost node at its periodic source) and stored as
// Epred on the physical nodes; Eavg = (E + Epred) / 2 at every node from its source node and
// Bavg = (B + Bpred) / 2, as k_ohm followed by k_average_fill computes them. M32: the averaged
// set is also written in FP32 (the mixed-precision push gathers from it). Every load precedes
// the first store.
template<int DIM, bool M32>
__global__ __launch_bounds__(FB) void k_ohm_avg(Geom G0, double eta, double nu, int hyper,
                                                double lvlCoef, double const* nn, CVecPtr Ve,
                                                double const* Pe, CVecPtr Bp, CVecPtr J, VecPtr Ep,
                                                CVecPtr E, CVecPtr B, VecPtr Ea, VecPtr Ba,
                                                Mirror m32)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i0, j0, k0;
    if (!nodeOf(G, QMOM, t, i0, j0, k0))
        return;
    int const c = static_cast<int>(blockIdx.y); // one component per grid row
    int i, j, k;
    bool const onE = ohmNode(G, c, 1, i0, j0, k0, i, j, k);
    bool const onB = i0 < G.shape[QBX + c][0] && j0 < G.shape[QBX + c][1] && k0 < G.shape[QBX + c][2];
    double ev = 0., eo = 0., bs = 0.;
    if (onE)
    {
        auto ve = [&](int d, int a, int b, int cc) { return Ve.c[d][G.idx(QMOM, a, b, cc)]; };
        auto pe = [&](int a, int b, int cc) { return Pe[G.idx(QMOM, a, b, cc)]; };
        ev = ohmAt(G, c, i, j, k, eta, nu, hyper, lvlCoef, nn, ve, pe, Bp, J);
        eo = E.c[c][G.idx(QEX + c, i, j, k)];
    }
    int const nb = onB ? G.idx(QBX + c, i0, j0, k0) : 0;
    if (onB)
        bs = B.c[c][nb] + Bp.c[c][nb];
    if (onB)
    {
        double const v = bs * .5;
        Ba.c[c][nb]    = v;
        if constexpr (M32)
            m32.c[3 + c][nb] = static_cast<float>(v);
    }
    if (onE)
    {
        int const n = G.idx(QEX + c, i0, j0, k0);
        if (i == i0 && j == j0 && k == k0)
            Ep.c[c][n] = ev;
        double const v = (eo + ev) * .5;
        Ea.c[c][n]     = v;
        if constexpr (M32)
            m32.c[c][n] = static_cast<float>(v);
    }
}

// ---- elementwise ------------------------------------------------------------------------------
__global__ __launch_bounds__(FB) void k_average(double const* a, double const* b, double* o, long n)
{
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    if (t < n)
        o[t] = (a[t] + b[t]) * .5;
}

__global__ __launch_bounds__(FB) void k_zero(double* p, long n)
{
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    if (t < n)
        p[t] = 0.;
}

__global__ __launch_bounds__(FB) void k_copy(double const* s, double* d, long n)
{
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    if (t < n)
        d[t] = s[t];
}

__global__ __launch_bounds__(FB) void k_ion_totals(PopMoments P, double* rho, double* md, VecPtr Vi,
                                                   long n)
{
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    if (t >= n)
        return;
    double r = 0., m = 0., vx = 0., vy = 0., vz = 0.;
    for (int p = 0; p < P.npop; ++p)
        r = r + P.rho[p][t];
    for (int p = 0; p < P.npop; ++p)
        m = m + P.n[p][t] * P.mass[p];
    for (int p = 0; p < P.npop; ++p)
    {
        vx = vx + P.F[p][0][t] * P.mass[p];
        vy = vy + P.F[p][1][t] * P.mass[p];
        vz = vz + P.F[p][2][t] * P.mass[p];
    }
    rho[t]     = r;
    md[t]      = m;
    Vi.c[0][t] = vx / m;
    Vi.c[1][t] = vy / m;
    Vi.c[2][t] = vz / m;
}

// ---- periodic same-level fills ----------------------------------------------------------------
// pure ghost nodes take the value of the periodic interior node (interior not overwritten)
template<int DIM>
__global__ __launch_bounds__(FB) void k_fill_ghosts(Geom G0, int q, double* f)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, q, t, i, j, k))
        return;
    int src[3] = {i, j, k};
    bool ghost = false;
    for (int d = 0; d < G.dim; ++d)
    {
        int c = src[d];
        if (c < G.g)
        {
            src[d] = c + G.n[d];
            ghost  = true;
        }
        else if (c > G.physEnd(q, d))
        {
            src[d] = c - G.n[d];
            ghost  = true;
        }
    }
    if (ghost)
        f[G.idx(q, i, j, k)] = f[G.idx(q, src[0], src[1], src[2])];
}

// setNaNsOnFieldGhosts of a single patch (hybrid_hybrid_messenger_strategy.hpp:959-994): every
// node outside the quantity's physical nodes becomes NaN (what the NaN-only refine operators of a
// level-ghost fill then complete)
template<int DIM>
__global__ __launch_bounds__(FB) void k_nan_ghosts(Geom G0, int q, double* f)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, q, t, i, j, k))
        return;
    int const c[3] = {i, j, k};
    bool ghost     = false;
    for (int d = 0; d < G.dim; ++d)
        ghost = ghost || c[d] < G.g || c[d] > G.physEnd(q, d);
    if (ghost)
        f[G.idx(q, i, j, k)] = __builtin_nan("");
}

// the nodes of the all-primal ghost box (S nodes per direction) that a periodic image overlaps:
// the m = S - N layers at each end of every direction, all S when the two ends meet (K nodes);
// enumerated as the nodes near along direction 0, then those inside along 0 and near along 1,
// then inside along 0 and 1 and near along 2
struct Shell
{
    int S[3], m[3], K[3];
    long count;
};

Shell shellOf(Geom const& G)
{
    Shell H{};
    for (int d = 0; d < 3; ++d)
    {
        H.S[d] = d < G.dim ? G.shape[QMOM][d] : 1;
        H.m[d] = d < G.dim ? H.S[d] - G.n[d] : 0;
        H.K[d] = d < G.dim ? (2 * H.m[d] < H.S[d] ? 2 * H.m[d] : H.S[d]) : 0;
    }
    long const in0 = H.S[0] - H.K[0], in1 = H.S[1] - H.K[1];
    H.count = static_cast<long>(H.K[0]) * H.S[1] * H.S[2];
    if (G.dim > 1)
        H.count += in0 * H.K[1] * H.S[2];
    if (G.dim > 2)
        H.count += in0 * in1 * H.K[2];
    return H;
}

// coordinate of the u-th near node / v-th inner node along d
__device__ __forceinline__ int shellNear(Shell const& H, int d, int u)
{
    return H.K[d] == H.S[d] || u < H.m[d] ? u : H.S[d] - H.K[d] + u;
}

template<int DIM>
__device__ __forceinline__ bool shellNode(Shell const& H, long t, int* c)
{
    if (t >= H.count)
        return false;
    long const r0 = static_cast<long>(H.K[0]) * H.S[1] * H.S[2];
    c[1] = c[2] = 0;
    if (t < r0)
    {
        long const pl = static_cast<long>(H.S[1]) * H.S[2];
        c[0]          = shellNear(H, 0, static_cast<int>(t / pl));
        long const r  = t % pl;
        c[1]          = static_cast<int>(r / H.S[2]);
        c[2]          = static_cast<int>(r % H.S[2]);
        return true;
    }
    t -= r0;
    int const in0 = H.S[0] - H.K[0];
    long const r1 = DIM > 1 ? static_cast<long>(in0) * H.K[1] * H.S[2] : 0;
    if (t < r1)
    {
        long const pl = static_cast<long>(H.K[1]) * H.S[2];
        c[0]          = H.m[0] + static_cast<int>(t / pl);
        long const r  = t % pl;
        c[1]          = shellNear(H, 1, static_cast<int>(r / H.S[2]));
        c[2]          = static_cast<int>(r % H.S[2]);
        return true;
    }
    t -= r1;
    int const in1 = H.S[1] - H.K[1];
    long const pl = static_cast<long>(in1) * H.K[2];
    c[0]          = H.m[0] + static_cast<int>(t / pl);
    long const r  = t % pl;
    c[1]          = H.m[1] + static_cast<int>(r / H.K[2]);
    c[2]          = shellNear(H, 2, static_cast<int>(r % H.K[2]));
    return true;
}

// image k of {-1,0,1}^dim (lexicographic, zero excluded) overlaps the all-primal ghost box on
// [max(0, k*N), min(end, end + k*N)]; node n receives orig[n - k*N]. Launched over the shell of
// nodes some image overlaps: the others keep their value (out already holds it, the callers
// copy out into orig first)
template<int DIM, bool MAX>
__global__ __launch_bounds__(FB) void k_border(Geom G0, Shell H, double const* orig, double* out)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int c[3];
    if (!shellNode<DIM>(H, t, c))
        return;
    int const i = c[0], j = c[1], k = c[2];
    int total = 1;
    for (int d = 0; d < G.dim; ++d)
        total *= 3;
    int const n = G.idx(QMOM, i, j, k);
    double v    = orig[n];
    for (int im = 0; im < total; ++im)
    {
        int kk[3] = {0, 0, 0};
        int r     = im;
        bool zero = true;
        for (int d = G.dim - 1; d >= 0; --d)
        {
            kk[d] = (r % 3) - 1;
            r /= 3;
            if (kk[d])
                zero = false;
        }
        if (zero)
            continue;
        bool in = true;
        int s[3] = {i, j, k};
        for (int d = 0; d < G.dim; ++d)
        {
            int const end = G.ghostEnd(QMOM, d);
            int const sh  = kk[d] * G.n[d];
            int const lo  = sh > 0 ? sh : 0;
            int const hi  = end + sh < end ? end + sh : end;
            if (c[d] < lo || c[d] > hi)
            {
                in = false;
                break;
            }
            s[d] = c[d] - sh;
        }
        if (!in)
            continue;
        double const o = orig[G.idx(QMOM, s[0], s[1], s[2])];
        if (MAX)
            v = v < o ? o : v;
        else
            v += o;
    }
    out[n] = v;
}

// ---- fused kernels of the periodic fast path ------------------------------------------------
// A pure ghost node of a periodic patch holds the value of its periodic source node; computing
// the stencil directly at the source gives that value bit for bit, so the producer kernels
// below write their ghost layers themselves instead of a separate fill.
template<class GT>
__device__ __forceinline__ bool wrapNode(GT const& G, int q, int* c)
{
    bool ghost = false;
    for (int d = 0; d < G.dim; ++d)
    {
        if (c[d] < G.g)
        {
            c[d] += G.n[d];
            ghost = true;
        }
        else if (c[d] > G.physEnd(q, d))
        {
            c[d] -= G.n[d];
            ghost = true;
        }
    }
    return ghost;
}

template<class GT>
__device__ __forceinline__ bool inShape(GT const& G, int q, int i, int j, int k)
{
    return i < G.shape[q][0] && j < G.shape[q][1] && k < G.shape[q][2];
}

template<class GT>
__device__ __forceinline__ double faradayAt(GT const& G, CVecPtr B, CVecPtr E, double dt, int c,
                                            int i, int j, int k)
{
    int const D = G.dim;
    int const n = G.idx(QBX + c, i, j, k);
    if (c == 0)
    {
        if (D == 1)
            return B.c[0][n];
        if (D == 2)
            return B.c[0][n] - dt * deriv(G, E.c[2], QEZ, 1, i, j, k);
        return B.c[0][n] - dt * deriv(G, E.c[2], QEZ, 1, i, j, k) + dt * deriv(G, E.c[1], QEY, 2, i, j, k);
    }
    if (c == 1)
    {
        if (D < 3)
            return B.c[1][n] + dt * deriv(G, E.c[2], QEZ, 0, i, j, k);
        return B.c[1][n] - dt * deriv(G, E.c[0], QEX, 2, i, j, k) + dt * deriv(G, E.c[2], QEZ, 0, i, j, k);
    }
    if (D == 1)
        return B.c[2][n] - dt * deriv(G, E.c[1], QEY, 0, i, j, k);
    return B.c[2][n] - dt * deriv(G, E.c[1], QEY, 0, i, j, k) + dt * deriv(G, E.c[0], QEX, 1, i, j, k);
}

// Faraday into a separate output, ghost layers included (predictor: B, E -> Bpred)
template<int DIM>
__global__ __launch_bounds__(FB) void k_faraday_fill(Geom G0, CVecPtr B, CVecPtr E, VecPtr Bn, double dt)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const c = static_cast<int>(blockIdx.y); // one component per grid row
    {
        if (!inShape(G, QBX + c, i, j, k))
            return;
        int s[3] = {i, j, k};
        wrapNode(G, QBX + c, s);
        Bn.c[c][G.idx(QBX + c, i, j, k)] = faradayAt(G, B, E, dt, c, s[0], s[1], s[2]);
    }
}

// component c of curl B at a J node; bAt(d, i, j, k) gives B component d at a node
template<class GT, class BF>
__device__ __forceinline__ double ampereAtFn(GT const& G, BF const& bAt, int c, int i, int j, int k)
{
    int const D = G.dim;
    auto comp   = [&](int d) { return [&, d](int a, int b, int cc) { return bAt(d, a, b, cc); }; };
    if (c == 0)
    {
        if (D == 1)
            return 0.0;
        if (D == 2)
            return derivFn(G, comp(2), QBZ, 1, i, j, k);
        return derivFn(G, comp(2), QBZ, 1, i, j, k) - derivFn(G, comp(1), QBY, 2, i, j, k);
    }
    if (c == 1)
    {
        if (D < 3)
            return -derivFn(G, comp(2), QBZ, 0, i, j, k);
        return derivFn(G, comp(0), QBX, 2, i, j, k) - derivFn(G, comp(2), QBZ, 0, i, j, k);
    }
    if (D == 1)
        return derivFn(G, comp(1), QBY, 0, i, j, k);
    return derivFn(G, comp(1), QBY, 0, i, j, k) - derivFn(G, comp(0), QBX, 1, i, j, k);
}

template<class GT>
__device__ __forceinline__ double ampereAt(GT const& G, CVecPtr B, int c, int i, int j, int k)
{
    return ampereAtFn(G, [&](int d, int a, int b, int cc) { return B.c[d][G.idx(QBX + d, a, b, cc)]; },
                      c, i, j, k);
}

// Ampere + fillCurrentGhosts
template<int DIM>
__global__ __launch_bounds__(FB) void k_ampere_fill(Geom G0, CVecPtr B, VecPtr J)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const c = static_cast<int>(blockIdx.y); // one component per grid row
    {
        if (!inShape(G, QJX + c, i, j, k))
            return;
        int s[3] = {i, j, k};
        wrapNode(G, QJX + c, s);
        J.c[c][G.idx(QJX + c, i, j, k)] = ampereAt(G, B, c, s[0], s[1], s[2]);
    }
}

// ghost fill of three components in one launch
template<int DIM>
__global__ __launch_bounds__(FB) void k_fill3(Geom G0, int q0, VecPtr F)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const c = static_cast<int>(blockIdx.y); // one component per grid row
    {
        int const q = q0 + c;
        if (!inShape(G, q, i, j, k))
            return;
        int s[3] = {i, j, k};
        if (wrapNode(G, q, s))
            F.c[c][G.idx(q, i, j, k)] = F.c[c][G.idx(q, s[0], s[1], s[2])];
    }
}

// average of E and B sets (algorithm.hpp:68-76) + fillElectricGhosts of Eavg
template<int DIM, bool M32>
__global__ __launch_bounds__(FB) void k_average_fill(Geom G0, CVecPtr E, CVecPtr Ep, CVecPtr B,
                                                     CVecPtr Bp, VecPtr Ea, VecPtr Ba, Mirror m32)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const c = static_cast<int>(blockIdx.y); // one component per grid row
    {
        if (inShape(G, QEX + c, i, j, k))
        {
            int s[3] = {i, j, k};
            wrapNode(G, QEX + c, s);
            int const src  = G.idx(QEX + c, s[0], s[1], s[2]);
            int const n    = G.idx(QEX + c, i, j, k);
            double const v = (E.c[c][src] + Ep.c[c][src]) * .5;
            Ea.c[c][n]     = v;
            if constexpr (M32)
                m32.c[c][n] = static_cast<float>(v);
        }
        if (inShape(G, QBX + c, i, j, k))
        {
            int const n    = G.idx(QBX + c, i, j, k);
            double const v = (B.c[c][n] + Bp.c[c][n]) * .5;
            Ba.c[c][n]     = v;
            if constexpr (M32)
                m32.c[3 + c][n] = static_cast<float>(v);
        }
    }
}

struct MomSet
{
    double const* raw[4][5];
    double* mom[4][5];
    double* zero[4][5]; // the other raw set, cleared for the next deposit
    double mass[4];
    int npop;
};

// periodic images of a node of the all-primal grid: image m (lexicographic over shifts
// {-1,0,1}^D less the zero shift) overlaps the node when the node lies in the array shifted by
// that many patch widths; its source is then at the linear offset -off[m]
template<int D>
struct Images
{
    static constexpr int N = D == 1 ? 2 : (D == 2 ? 8 : 26);
    bool valid[N];
    int off[N];
};

template<class GT>
__device__ __forceinline__ void imagesOf(GT const& G, int const* c, Images<GT::dim>& I)
{
    constexpr int D     = GT::dim;
    constexpr int total = D == 1 ? 3 : (D == 2 ? 9 : 27);
    int m               = 0;
#pragma unroll
    for (int im = 0; im < total; ++im)
    {
        int kk[3] = {0, 0, 0};
        int r     = im;
#pragma unroll
        for (int d = D - 1; d >= 0; --d)
        {
            kk[d] = (r % 3) - 1;
            r /= 3;
        }
        if (kk[0] == 0 && kk[1] == 0 && kk[2] == 0)
            continue;
        bool in   = true;
        int shift = 0;
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            int const end = G.ghostEnd(QMOM, d);
            int const sh  = kk[d] * G.n[d];
            int const lo  = sh > 0 ? sh : 0;
            int const hi  = end + sh < end ? end + sh : end;
            in            = in && c[d] >= lo && c[d] <= hi;
            int const stride = d == 0 ? G.shape[QMOM][1] * G.shape[QMOM][2]
                                      : (d == 1 ? G.shape[QMOM][2] : 1);
            shift += sh * stride;
        }
        I.valid[m] = in;
        I.off[m]   = shift;
        ++m;
    }
}

// fillFluxBorders + fillDensityBorders (border sum of every population moment) + updateIons
// (totals into tmp) ; the max fill of the totals follows in k_max5. A node's sources are itself,
// then its overlapping periodic images in lexicographic order. Loads are unconditional (an
// image that does not overlap reads the node itself and is not added) and precede the
// population's stores, so that they are in flight together.
template<class GT>
__device__ __forceinline__ void foldTotalsAt(GT const& G, MomSet const& M, int i, int j, int k,
                                             double* trho, double* tmd, VecPtr tvi)
{
    constexpr int DIM = GT::dim;
    int const c[3] = {i, j, k};
    int const n    = G.idx(QMOM, i, j, k);
    Images<DIM> I;
    imagesOf(G, c, I);
    int src[Images<DIM>::N];
#pragma unroll
    for (int x = 0; x < Images<DIM>::N; ++x)
        src[x] = I.valid[x] ? n - I.off[x] : n;
    double r = 0., m = 0., vx = 0., vy = 0., vz = 0.;
    for (int p = 0; p < M.npop; ++p)
    {
        double f[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            double const* raw = M.raw[p][q];
            double v          = raw[n];
#pragma unroll
            for (int x = 0; x < Images<DIM>::N; ++x)
            {
                double const o = raw[src[x]];
                v              = I.valid[x] ? v + o : v;
            }
            f[q] = v;
        }
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            M.mom[p][q][n]  = f[q];
            M.zero[p][q][n] = 0.;
        }
        r  = r + f[1];
        m  = m + f[0] * M.mass[p];
        vx = vx + f[2] * M.mass[p];
        vy = vy + f[3] * M.mass[p];
        vz = vz + f[4] * M.mass[p];
    }
    trho[n]     = r;
    tmd[n]      = m;
    tvi.c[0][n] = vx / m;
    tvi.c[1][n] = vy / m;
    tvi.c[2][n] = vz / m;
}

template<int DIM>
__global__ __launch_bounds__(FB) void k_fold_totals(Geom G0, MomSet M, double* trho, double* tmd,
                                                    VecPtr tvi)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    foldTotalsAt(G, M, i, j, k, trho, tmd, tvi);
}

// the fold of a moveIons and the Faraday that follows it in a captured step, in one launch (the
// two are independent): grid rows 0-2 compute component c of Faraday, as k_faraday_fill does
// (INPLACE false: into Bn with its ghost layers) or as k_faraday does (INPLACE true: Bn is B, the
// physical nodes only), row 3 computes the fold as k_fold_totals does
template<int DIM, bool INPLACE>
__global__ __launch_bounds__(FB) void k_fold_faraday(Geom G0, MomSet M, double* trho, double* tmd,
                                                     VecPtr tvi, CVecPtr B, CVecPtr E, VecPtr Bn,
                                                     double dt)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const c = static_cast<int>(blockIdx.y);
    if (c == 3)
    {
        foldTotalsAt(G, M, i, j, k, trho, tmd, tvi);
        return;
    }
    if (!inShape(G, QBX + c, i, j, k))
        return;
    int s[3] = {i, j, k};
    if constexpr (INPLACE)
    {
        int lo[3], hi[3];
        physBox(G, QBX + c, lo, hi);
        if (!inRange(G, i, j, k, lo, hi))
            return;
    }
    else
        wrapNode(G, QBX + c, s);
    double const val = faradayAt(G, B, E, dt, c, s[0], s[1], s[2]);
    Bn.c[c][G.idx(QBX + c, i, j, k)] = val;
    if constexpr (INPLACE)
    {
        // fillMagneticGhosts: every ghost node whose periodic source is this node (the nodes
        // q = (i, j, k) + sum_d k_d n_d e_d, k_d in {-1, 0, 1}, that wrapNode maps back here)
        // takes the new value. Faraday reads B only at its own node, so no thread reads a
        // ghost value this launch writes.
        constexpr int NI = DIM == 1 ? 3 : (DIM == 2 ? 9 : 27);
        for (int m = 0; m < NI; ++m)
        {
            int q[3]  = {i, j, k};
            int mm    = m;
            bool self = true;
            for (int d = 0; d < DIM; ++d)
            {
                int const kd = mm % 3 - 1;
                mm /= 3;
                q[d] += kd * G.n[d];
                self = self && kd == 0;
            }
            if (self || q[0] < 0 || q[1] < 0 || q[2] < 0 || !inShape(G, QBX + c, q[0], q[1], q[2]))
                continue;
            int w[3] = {q[0], q[1], q[2]};
            if (!wrapNode(G, QBX + c, w) || w[0] != i || w[1] != j || w[2] != k)
                continue;
            Bn.c[c][G.idx(QBX + c, q[0], q[1], q[2])] = val;
        }
    }
}

struct Five
{
    double const* src[5];
    double* dst[5];
};

// the five totals at primal node (i, j, k) (index n): the max over its periodic images of the
// values in F.src. A non-overlapping image reads the node itself, which leaves the max as it is,
// so every load is unconditional and its value always used (the loads are issued together rather
// than each behind its own branch).
template<int DIM>
__device__ __forceinline__ void maxOverImages(GeomD<DIM> const& G, Five const& F, int i, int j, int k,
                                              int n, double* v)
{
    int const c[3] = {i, j, k};
    Images<DIM> I;
    imagesOf(G, c, I);
#pragma unroll
    for (int q = 0; q < 5; ++q)
    {
        v[q] = F.src[q][n];
#pragma unroll
        for (int x = 0; x < Images<DIM>::N; ++x)
        {
            double const o = F.src[q][I.valid[x] ? n - I.off[x] : n];
            v[q]           = v[q] < o ? o : v[q];
        }
    }
}

// fillIonBorders: max over the periodic images (snapshot in src)
template<int DIM>
__global__ __launch_bounds__(FB) void k_max5(Geom G0, Five F)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    int const n = G.idx(QMOM, i, j, k);
    double v[5];
    maxOverImages(G, F, i, j, k, n, v); // every load before the stores
#pragma unroll
    for (int q = 0; q < 5; ++q)
        F.dst[q][n] = v[q];
}

// Ampere with J's ghost layers and the electron closure in one launch (periodic level). A thread
// stores J at its node (every component whose shape holds the node) and runs the closure at the
// same primal node; the J values the closure projects are recomputed from B, each one the value
// k_ampere_fill stores at that node: curl B at the node's periodic source. MAX: the closure takes
// the totals from the border max of F.src over the node's periodic images (stored in F.dst), as
// k_max5 takes it, else from Ne, Vi as k_electrons does.
template<int DIM, bool MAX>
__global__ __launch_bounds__(FB) void k_ampere_electrons(Geom G0, double Te, CVecPtr B, VecPtr J,
                                                         Five F, CVecPtr Vi, double const* Ne,
                                                         VecPtr Ve, double* Pe)
{
    GeomD<DIM> const G(G0);
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    int i, j, k;
    if (!nodeOf(G, QMOM, t, i, j, k))
        return;
    auto bAt = [&](int d, int a, int b, int cc) { return B.c[d][G.idx(QBX + d, a, b, cc)]; };
    auto jAt = [&](int c, int a, int b, int cc) {
        int s[3] = {a, b, cc};
        wrapNode(G, QJX + c, s);
        return ampereAtFn(G, bAt, c, s[0], s[1], s[2]);
    };
    // every load before the first store
    bool onJ[3];
    double jv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
        onJ[c] = inShape(G, QJX + c, i, j, k);
        jv[c]  = onJ[c] ? jAt(c, i, j, k) : 0.;
    }
    int const n = G.idx(QMOM, i, j, k);
    double v[5];
    if (MAX)
        maxOverImages(G, F, i, j, k, n, v);
    else
    {
        v[0] = Ne[n];
        v[1] = 0.;
        v[2] = Vi.c[0][n];
        v[3] = Vi.c[1][n];
        v[4] = Vi.c[2][n];
    }
    int lo[3], hi[3];
    physBox(G, QMOM, lo, hi);
    bool const phys = inRange(G, i, j, k, lo, hi);
    double jx = 0., jy = 0., jz = 0.;
    if (phys)
    {
        // JxToMoments / JyToMoments / JzToMoments: dual->primal along the component's direction
        jx = projectFn(G, [&](int a, int b, int cc) { return jAt(0, a, b, cc); }, i, j, k, D2P, PN, PN);
        jy = projectFn(G, [&](int a, int b, int cc) { return jAt(1, a, b, cc); }, i, j, k, PN, D2P, PN);
        jz = projectFn(G, [&](int a, int b, int cc) { return jAt(2, a, b, cc); }, i, j, k, PN, PN, D2P);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (onJ[c])
            J.c[c][G.idx(QJX + c, i, j, k)] = jv[c];
    if (MAX)
    {
#pragma unroll
        for (int q = 0; q < 5; ++q)
            F.dst[q][n] = v[q];
    }
    Pe[n] = v[0] * Te;
    if (phys)
    {
        Ve.c[0][n] = v[2] - jx / v[0];
        Ve.c[1][n] = v[3] - jy / v[0];
        Ve.c[2][n] = v[4] - jz / v[0];
    }
}

unsigned int gridFor(long n) { return static_cast<unsigned int>((n + FB - 1) / FB); }

template<bool MAX>
void ampereElectronsT(Geom const& G, double Te, CVecPtr B, VecPtr J, Five const& F, CVecPtr Vi,
                      double const* Ne, VecPtr Ve, double* Pe, hipStream_t s)
{
    unsigned int const g = gridFor(G.size(QMOM));
    switch (G.dim)
    {
        case 1: k_ampere_electrons<1, MAX><<<g, FB, 0, s>>>(G, Te, B, J, F, Vi, Ne, Ve, Pe); break;
        case 2: k_ampere_electrons<2, MAX><<<g, FB, 0, s>>>(G, Te, B, J, F, Vi, Ne, Ve, Pe); break;
        case 3: k_ampere_electrons<3, MAX><<<g, FB, 0, s>>>(G, Te, B, J, F, Vi, Ne, Ve, Pe); break;
    }
}

} // namespace

void launchAmpereElectrons(Geom const& G, double Te, CVecPtr B, VecPtr J, double const* const* src,
                           double* const* dst, CVecPtr Vi, double const* Ne, VecPtr Ve, double* Pe,
                           bool borderMax, hipStream_t s)
{
    Five F{};
    if (borderMax)
        for (int q = 0; q < 5; ++q)
        {
            F.src[q] = src[q];
            F.dst[q] = dst[q];
        }
    if (borderMax)
        ampereElectronsT<true>(G, Te, B, J, F, Vi, Ne, Ve, Pe, s);
    else
        ampereElectronsT<false>(G, Te, B, J, F, Vi, Ne, Ve, Pe, s);
}

void launchFaraday(Geom const& G, CVecPtr B, CVecPtr E, VecPtr Bnew, double dt, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_faraday<1><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, B, E, Bnew, dt); break;
        case 2: k_faraday<2><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, B, E, Bnew, dt); break;
        case 3: k_faraday<3><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, B, E, Bnew, dt); break;
    }
}
void launchAmpere(Geom const& G, CVecPtr B, VecPtr J, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_ampere<1><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, B, J); break;
        case 2: k_ampere<2><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, B, J); break;
        case 3: k_ampere<3><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, B, J); break;
    }
}
void launchElectrons(Geom const& G, double Te, CVecPtr J, CVecPtr Vi, double const* Ne, VecPtr Ve,
                     double* Pe, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_electrons<1><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, Te, J, Vi, Ne, Ve, Pe); break;
        case 2: k_electrons<2><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, Te, J, Vi, Ne, Ve, Pe); break;
        case 3: k_electrons<3><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, Te, J, Vi, Ne, Ve, Pe); break;
    }
}
void launchOhm(Geom const& G, double eta, double nu, int hyper, int level, double const* n,
               CVecPtr Ve, double const* Pe, CVecPtr B, CVecPtr J, VecPtr E, hipStream_t s, int fill)
{
    double lvlCoef = 1.;
    for (int l = 0; l < level; ++l)
        lvlCoef *= 4.;
    lvlCoef = 1. / lvlCoef;
    switch (G.dim)
    {
        case 1: k_ohm<1><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, B, J, E, fill); break;
        case 2: k_ohm<2><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, B, J, E, fill); break;
        case 3: k_ohm<3><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, B, J, E, fill); break;
    }
}
void launchOhmAverage(Geom const& G, double eta, double nu, int hyper, int level, double const* n,
                      CVecPtr Ve, double const* Pe, CVecPtr Bp, CVecPtr J, VecPtr Ep, CVecPtr E,
                      CVecPtr B, VecPtr Ea, VecPtr Ba, hipStream_t s, float* const* mirror)
{
    double lvlCoef = 1.;
    for (int l = 0; l < level; ++l)
        lvlCoef *= 4.;
    lvlCoef = 1. / lvlCoef;
    Mirror m{};
    for (int k = 0; k < 6 && mirror; ++k)
        m.c[k] = mirror[k];
    dim3 const grid(gridFor(G.size(QMOM)), 3);
    switch (G.dim * 2 + (mirror ? 1 : 0))
    {
        case 2: k_ohm_avg<1, false><<<grid, FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, Bp, J, Ep, E, B, Ea, Ba, m); break;
        case 3: k_ohm_avg<1, true><<<grid, FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, Bp, J, Ep, E, B, Ea, Ba, m); break;
        case 4: k_ohm_avg<2, false><<<grid, FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, Bp, J, Ep, E, B, Ea, Ba, m); break;
        case 5: k_ohm_avg<2, true><<<grid, FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, Bp, J, Ep, E, B, Ea, Ba, m); break;
        case 6: k_ohm_avg<3, false><<<grid, FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, Bp, J, Ep, E, B, Ea, Ba, m); break;
        case 7: k_ohm_avg<3, true><<<grid, FB, 0, s>>>(G, eta, nu, hyper, lvlCoef, n, Ve, Pe, Bp, J, Ep, E, B, Ea, Ba, m); break;
    }
}

void launchAverage(double const* a, double const* b, double* out, long count, hipStream_t s)
{
    k_average<<<gridFor(count), FB, 0, s>>>(a, b, out, count);
}
void launchZero(double* p, long count, hipStream_t s) { k_zero<<<gridFor(count), FB, 0, s>>>(p, count); }
void launchCopy(double const* src, double* dst, long count, hipStream_t s)
{
    k_copy<<<gridFor(count), FB, 0, s>>>(src, dst, count);
}
void launchNanGhosts(Geom const& G, int q, double* f, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_nan_ghosts<1><<<gridFor(G.size(q)), FB, 0, s>>>(G, q, f); break;
        case 2: k_nan_ghosts<2><<<gridFor(G.size(q)), FB, 0, s>>>(G, q, f); break;
        case 3: k_nan_ghosts<3><<<gridFor(G.size(q)), FB, 0, s>>>(G, q, f); break;
    }
}
void launchFillGhosts(Geom const& G, int q, double* f, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_fill_ghosts<1><<<gridFor(G.size(q)), FB, 0, s>>>(G, q, f); break;
        case 2: k_fill_ghosts<2><<<gridFor(G.size(q)), FB, 0, s>>>(G, q, f); break;
        case 3: k_fill_ghosts<3><<<gridFor(G.size(q)), FB, 0, s>>>(G, q, f); break;
    }
}
void launchBorderSum(Geom const& G, double const* orig, double* out, hipStream_t s)
{
    Shell const H = shellOf(G);
    if (H.count == 0)
        return;
    switch (G.dim)
    {
        case 1: k_border<1, false><<<gridFor(H.count), FB, 0, s>>>(G, H, orig, out); break;
        case 2: k_border<2, false><<<gridFor(H.count), FB, 0, s>>>(G, H, orig, out); break;
        case 3: k_border<3, false><<<gridFor(H.count), FB, 0, s>>>(G, H, orig, out); break;
    }
}
void launchBorderMax(Geom const& G, double const* orig, double* inout, hipStream_t s)
{
    Shell const H = shellOf(G);
    if (H.count == 0)
        return;
    switch (G.dim)
    {
        case 1: k_border<1, true><<<gridFor(H.count), FB, 0, s>>>(G, H, orig, inout); break;
        case 2: k_border<2, true><<<gridFor(H.count), FB, 0, s>>>(G, H, orig, inout); break;
        case 3: k_border<3, true><<<gridFor(H.count), FB, 0, s>>>(G, H, orig, inout); break;
    }
}
void launchIonTotals(PopMoments const& P, double* rho, double* massDensity, VecPtr Vi, long count,
                     hipStream_t s)
{
    k_ion_totals<<<gridFor(count), FB, 0, s>>>(P, rho, massDensity, Vi, count);
}

void launchFaradayFill(Geom const& G, CVecPtr B, CVecPtr E, VecPtr Bnew, double dt, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_faraday_fill<1><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, B, E, Bnew, dt); break;
        case 2: k_faraday_fill<2><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, B, E, Bnew, dt); break;
        case 3: k_faraday_fill<3><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, B, E, Bnew, dt); break;
    }
}
void launchAmpereFill(Geom const& G, CVecPtr B, VecPtr J, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_ampere_fill<1><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, B, J); break;
        case 2: k_ampere_fill<2><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, B, J); break;
        case 3: k_ampere_fill<3><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, B, J); break;
    }
}
void launchFill3(Geom const& G, int q0, VecPtr F, hipStream_t s)
{
    switch (G.dim)
    {
        case 1: k_fill3<1><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, q0, F); break;
        case 2: k_fill3<2><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, q0, F); break;
        case 3: k_fill3<3><<<dim3(gridFor(G.size(QMOM)), 3), FB, 0, s>>>(G, q0, F); break;
    }
}
void launchAverageFill(Geom const& G, CVecPtr E, CVecPtr Ep, CVecPtr B, CVecPtr Bp, VecPtr Ea,
                       VecPtr Ba, hipStream_t s, float* const* mirror)
{
    Mirror m{};
    for (int k = 0; k < 6 && mirror; ++k)
        m.c[k] = mirror[k];
    dim3 const grid(gridFor(G.size(QMOM)), 3);
    switch (G.dim * 2 + (mirror ? 1 : 0))
    {
        case 2: k_average_fill<1, false><<<grid, FB, 0, s>>>(G, E, Ep, B, Bp, Ea, Ba, m); break;
        case 3: k_average_fill<1, true><<<grid, FB, 0, s>>>(G, E, Ep, B, Bp, Ea, Ba, m); break;
        case 4: k_average_fill<2, false><<<grid, FB, 0, s>>>(G, E, Ep, B, Bp, Ea, Ba, m); break;
        case 5: k_average_fill<2, true><<<grid, FB, 0, s>>>(G, E, Ep, B, Bp, Ea, Ba, m); break;
        case 6: k_average_fill<3, false><<<grid, FB, 0, s>>>(G, E, Ep, B, Bp, Ea, Ba, m); break;
        case 7: k_average_fill<3, true><<<grid, FB, 0, s>>>(G, E, Ep, B, Bp, Ea, Ba, m); break;
    }
}

// FP32 copy of the 6 E/B components (one grid row per component)
__global__ __launch_bounds__(FB) void k_em_to_f32(EMPtrs src, Mirror dst, int n0, int n1, int n2,
                                                  int n3, int n4, int n5)
{
    int const c  = static_cast<int>(blockIdx.y);
    int const nc = c == 0 ? n0 : (c == 1 ? n1 : (c == 2 ? n2 : (c == 3 ? n3 : (c == 4 ? n4 : n5))));
    long const t = static_cast<long>(blockIdx.x) * FB + threadIdx.x;
    if (t >= nc)
        return;
    double const* s = c < 3 ? src.E[c] : src.B[c - 3];
    dst.c[c][t]     = static_cast<float>(s[t]);
}

void launchEmToF32(Geom const& G, double const* const* src, float* const* dst, hipStream_t s)
{
    EMPtrs e{};
    Mirror m{};
    int n[6];
    for (int k = 0; k < 6; ++k)
    {
        if (k < 3)
            e.E[k] = src[k];
        else
            e.B[k - 3] = src[k];
        m.c[k] = dst[k];
        n[k]   = G.size(k < 3 ? QEX + k : QBX + k - 3);
    }
    int const mx = *std::max_element(n, n + 6);
    k_em_to_f32<<<dim3(gridFor(mx), 6), FB, 0, s>>>(e, m, n[0], n[1], n[2], n[3], n[4], n[5]);
}
void launchFoldTotals(Geom const& G, MomSetArgs const& A, double* trho, double* tmd, VecPtr tvi,
                      hipStream_t s)
{
    MomSet M{};
    M.npop = A.npop;
    for (int p = 0; p < A.npop; ++p)
    {
        M.mass[p] = A.mass[p];
        for (int q = 0; q < 5; ++q)
        {
            M.raw[p][q]  = A.raw[p][q];
            M.mom[p][q]  = A.mom[p][q];
            M.zero[p][q] = A.zero[p][q];
        }
    }
    switch (G.dim)
    {
        case 1: k_fold_totals<1><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, M, trho, tmd, tvi); break;
        case 2: k_fold_totals<2><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, M, trho, tmd, tvi); break;
        case 3: k_fold_totals<3><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, M, trho, tmd, tvi); break;
    }
}

void launchFoldFaraday(Geom const& G, MomSetArgs const& A, double* trho, double* tmd, VecPtr tvi,
                       CVecPtr B, CVecPtr E, VecPtr Bn, double dt, bool inPlace, hipStream_t s)
{
    MomSet M{};
    M.npop = A.npop;
    for (int p = 0; p < A.npop; ++p)
    {
        M.mass[p] = A.mass[p];
        for (int q = 0; q < 5; ++q)
        {
            M.raw[p][q]  = A.raw[p][q];
            M.mom[p][q]  = A.mom[p][q];
            M.zero[p][q] = A.zero[p][q];
        }
    }
    dim3 const grid(gridFor(G.size(QMOM)), 4);
    switch (G.dim * 2 + (inPlace ? 1 : 0))
    {
        case 2: k_fold_faraday<1, false><<<grid, FB, 0, s>>>(G, M, trho, tmd, tvi, B, E, Bn, dt); break;
        case 3: k_fold_faraday<1, true><<<grid, FB, 0, s>>>(G, M, trho, tmd, tvi, B, E, Bn, dt); break;
        case 4: k_fold_faraday<2, false><<<grid, FB, 0, s>>>(G, M, trho, tmd, tvi, B, E, Bn, dt); break;
        case 5: k_fold_faraday<2, true><<<grid, FB, 0, s>>>(G, M, trho, tmd, tvi, B, E, Bn, dt); break;
        case 6: k_fold_faraday<3, false><<<grid, FB, 0, s>>>(G, M, trho, tmd, tvi, B, E, Bn, dt); break;
        case 7: k_fold_faraday<3, true><<<grid, FB, 0, s>>>(G, M, trho, tmd, tvi, B, E, Bn, dt); break;
    }
}

void launchMax5(Geom const& G, double const* const* src, double* const* dst, hipStream_t s)
{
    Five F{};
    for (int q = 0; q < 5; ++q)
    {
        F.src[q] = src[q];
        F.dst[q] = dst[q];
    }
    switch (G.dim)
    {
        case 1: k_max5<1><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, F); break;
        case 2: k_max5<2><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, F); break;
        case 3: k_max5<3><<<gridFor(G.size(QMOM)), FB, 0, s>>>(G, F); break;
    }
}

} // namespace phx
