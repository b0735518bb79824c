// refcpu.hpp -- CPU ORACLE (test infrastructure only).
//
// A clean-room FP64 restatement of PHARE's hybrid-PIC ion update and field solve, used ONLY
// as the checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
// Nothing in the product (phare_amd/) includes, links or calls this code.
//
// Every function cites the reference file:line it restates (paths relative to the PHARE
// source tree). Arithmetic is written in the reference's operation order so that, built
// without FMA contraction (see Makefile), a single call reproduces the reference bit for bit.
//
// Parity pins (see tests/): B-spline weights vs scipy BSpline; gather/deposit/Faraday/Ampere/
// gridlayout/ParticleArray-partition vs the reference headers compiled into oracle/_ref;
// deposit vs the reference's hand-tuned particle collections; Boris vs a closed-form
// gyration and an ODE trajectory; Ohm and Electrons vs polynomial fields for which the
// discrete operators are exact. The single-patch periodic fills restate SAMRAI semantics
// (parity unpinned at the SAMRAI boundary except through invariants: conservation of charge,
// equality of periodic images, translation of a uniform plasma).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace refcpu
{

// ------------------------------------------------------------------------------------------
// centering (gridlayout_hybrid_yee.hpp:51-96)
// ------------------------------------------------------------------------------------------
enum Cent : int { primal = 0, dual = 1 };

// quantity ids: B, E, J components and one all-primal id for moments/Pe/Vi
enum Qty : int { qBx = 0, qBy, qBz, qEx, qEy, qEz, qJx, qJy, qJz, qMom };

inline Cent centering(Qty q, int dir)
{
    // gridlayout_hybrid_yee.hpp:51-62: Bx pdd, By dpd, Bz ddp, Ex dpp, Ey pdp, Ez ppd (J as E)
    switch (q)
    {
        case qBx: return dir == 0 ? primal : dual;
        case qBy: return dir == 1 ? primal : dual;
        case qBz: return dir == 2 ? primal : dual;
        case qEx:
        case qJx: return dir == 0 ? dual : primal;
        case qEy:
        case qJy: return dir == 1 ? dual : primal;
        case qEz:
        case qJz: return dir == 2 ? dual : primal;
        default: return primal;
    }
}

// hybrid_options.hpp:24-25
inline int field_ghost_width(int order) { return std::array{2, 4, 4}[order - 1]; }
inline int particle_ghost_width(int order) { return std::array{1, 2, 2}[order - 1]; }

// ------------------------------------------------------------------------------------------
// Box (core/utilities/box/box.hpp: row-major iteration, last index fastest :186-190; isIn :265)
// ------------------------------------------------------------------------------------------
template<int D>
struct Box
{
    std::array<int, D> lo{}, hi{};

    bool contains(std::array<int, D> const& c) const
    {
        for (int d = 0; d < D; ++d)
            if (c[d] < lo[d] || c[d] > hi[d])
                return false;
        return true;
    }
    std::size_t size() const
    {
        std::size_t s = 1;
        for (int d = 0; d < D; ++d)
            s *= static_cast<std::size_t>(hi[d] - lo[d] + 1);
        return s;
    }
    Box grown(int n) const
    {
        Box b = *this;
        for (int d = 0; d < D; ++d)
        {
            b.lo[d] -= n;
            b.hi[d] += n;
        }
        return b;
    }
    std::size_t lin(std::array<int, D> const& c) const
    {
        std::size_t l = 0;
        for (int d = 0; d < D; ++d)
            l = l * static_cast<std::size_t>(hi[d] - lo[d] + 1) + static_cast<std::size_t>(c[d] - lo[d]);
        return l;
    }
    template<typename Fn>
    void for_each(Fn&& fn) const
    {
        std::array<int, D> c = lo;
        if (size() == 0)
            return;
        while (true)
        {
            fn(c);
            int d = D - 1;
            while (d >= 0)
            {
                if (++c[d] <= hi[d])
                    break;
                c[d] = lo[d];
                --d;
            }
            if (d < 0)
                break;
        }
    }
};

// ------------------------------------------------------------------------------------------
// GridLayout index space (gridlayout.hpp:1349-1440 tables, :746-763 AMRToLocal,
// :550-563 + :1319-1337 next/prev index, :571-632 deriv, :638-700 laplacian)
// ------------------------------------------------------------------------------------------
template<int D>
struct Layout
{
    std::array<int, D> ncells{};
    std::array<double, D> dx{};
    std::array<double, D> inv{}; // 1./dx  (gridlayout.hpp:136)
    std::array<int, D> lower{};  // AMR lower cell
    int order = 1;
    int g     = 2; // field ghost width
    int pg    = 1; // particle ghost width
    int level = 0;

    Layout() = default;
    Layout(std::array<int, D> n, std::array<double, D> h, std::array<int, D> lo, int o, int lvl = 0)
        : ncells{n}
        , dx{h}
        , lower{lo}
        , order{o}
        , g{field_ghost_width(o)}
        , pg{particle_ghost_width(o)}
        , level{lvl}
    {
        for (int d = 0; d < D; ++d)
            inv[d] = 1. / dx[d];
    }

    int physStart() const { return g; }
    int physEnd(Cent c, int d) const { return g + ncells[d] - (c == dual ? 1 : 0); }
    int ghostEnd(Cent c, int d) const { return physEnd(c, d) + g; }
    int nodes(Cent c, int d) const { return ghostEnd(c, d) + 1; }
    int amrToLocal(int amr, int d) const { return amr - (lower[d] - g); }

    Box<D> amrBox() const
    {
        Box<D> b;
        for (int d = 0; d < D; ++d)
        {
            b.lo[d] = lower[d];
            b.hi[d] = lower[d] + ncells[d] - 1;
        }
        return b;
    }
    std::array<int, 3> shape(Qty q) const
    {
        std::array<int, 3> s{1, 1, 1};
        for (int d = 0; d < D; ++d)
            s[d] = nodes(centering(q, d), d);
        return s;
    }
};

// next/prev index of the opposite centering (gridlayout.hpp:1294-1320)
inline int nextIdx(Cent operandCentering, int i) { return operandCentering == primal ? i + 1 : i; }
inline int prevIdx(Cent operandCentering, int i) { return operandCentering == primal ? i : i - 1; }

// ------------------------------------------------------------------------------------------
// Field: C-order array with ghosts (ndarray_vector.hpp:34-50)
// ------------------------------------------------------------------------------------------
struct Field
{
    Qty q = qMom;
    std::array<int, 3> shape{1, 1, 1};
    std::vector<double> v;

    Field() = default;
    template<int D>
    Field(Layout<D> const& L, Qty qq)
        : q{qq}
        , shape{L.shape(qq)}
        , v(static_cast<std::size_t>(shape[0]) * shape[1] * shape[2], 0.)
    {
    }
    std::size_t idx(int i, int j = 0, int k = 0) const
    {
        return (static_cast<std::size_t>(i) * shape[1] + j) * shape[2] + k;
    }
    template<std::size_t D>
    double& at(std::array<int, D> const& c)
    {
        if constexpr (D == 1)
            return v[idx(c[0])];
        else if constexpr (D == 2)
            return v[idx(c[0], c[1])];
        else
            return v[idx(c[0], c[1], c[2])];
    }
    template<std::size_t D>
    double at(std::array<int, D> const& c) const
    {
        if constexpr (D == 1)
            return v[idx(c[0])];
        else if constexpr (D == 2)
            return v[idx(c[0], c[1])];
        else
            return v[idx(c[0], c[1], c[2])];
    }
    void zero() { std::fill(v.begin(), v.end(), 0.); }
};

struct VecField
{
    std::array<Field, 3> c;
    VecField() = default;
    template<int D>
    VecField(Layout<D> const& L, Qty q0)
        : c{Field{L, Qty(q0)}, Field{L, Qty(q0 == qMom ? qMom : q0 + 1)},
            Field{L, Qty(q0 == qMom ? qMom : q0 + 2)}}
    {
    }
    Field& operator[](int i) { return c[i]; }
    Field const& operator[](int i) const { return c[i]; }
    void zero()
    {
        for (auto& f : c)
            f.zero();
    }
};

struct EM
{
    VecField E, B;
};

// physical / ghost / shrunk-ghost loops (gridlayout.hpp:1198-1283)
template<int D, typename Fn>
void evalOnRange(std::array<int, D> const& lo, std::array<int, D> const& hi, Fn&& fn)
{
    Box<D> b;
    b.lo = lo;
    b.hi = hi;
    b.for_each(fn);
}
template<int D, typename Fn>
void evalOnBox(Layout<D> const& L, Qty q, Fn&& fn)
{
    std::array<int, D> lo, hi;
    for (int d = 0; d < D; ++d)
    {
        lo[d] = L.physStart();
        hi[d] = L.physEnd(centering(q, d), d);
    }
    evalOnRange<D>(lo, hi, fn);
}
template<int D, typename Fn>
void evalOnShrunkGhostBox(Layout<D> const& L, Qty q, int shrink, Fn&& fn)
{
    std::array<int, D> lo, hi;
    for (int d = 0; d < D; ++d)
    {
        lo[d] = 0 + shrink;
        hi[d] = L.ghostEnd(centering(q, d), d) - shrink;
    }
    evalOnRange<D>(lo, hi, fn);
}

// deriv (gridlayout.hpp:571-632): inverseMeshSize * (next - prev)
template<int D>
double deriv(Layout<D> const& L, Field const& f, int dir, std::array<int, D> idx)
{
    Cent c                 = centering(f.q, dir);
    std::array<int, D> nxt = idx, prv = idx;
    nxt[dir]               = nextIdx(c, idx[dir]);
    prv[dir]               = prevIdx(c, idx[dir]);
    return L.inv[dir] * (f.at(nxt) - f.at(prv));
}

// laplacian (gridlayout.hpp:638-700)
template<int D>
double laplacian(Layout<D> const& L, Field const& f, std::array<int, D> idx)
{
    double res = 0.;
    for (int d = 0; d < D; ++d)
    {
        auto p = idx, n = idx;
        p[d] -= 1;
        n[d] += 1;
        double lap = L.inv[d] * L.inv[d] * (f.at(n) - 2.0 * f.at(idx) + f.at(p));
        res        = (d == 0) ? lap : res + lap;
    }
    return res;
}

// GridLayout::project (gridlayout.hpp:784-795) with Yee directionalInterp (order 2) and
// tensorProduct (gridlayout_hybrid_yee.hpp:351-446): a list of (offset, coef) summed from 0.
struct WP
{
    std::array<int, 3> off;
    double coef;
};
// one direction: DualToPrimal -> offsets {-1,0}, PrimalToDual -> {0,+1}; dir >= D -> identity
template<int D>
std::vector<WP> dirInterp(int dir, bool primalToDual)
{
    if (dir >= D)
        return {WP{{0, 0, 0}, 1.0}};
    int base = primalToDual ? 1 : 0;
    WP a{{0, 0, 0}, 0.5}, b{{0, 0, 0}, 0.5};
    a.off[dir] = base - 1;
    b.off[dir] = base;
    return {a, b};
}
inline std::vector<WP> tensor(std::vector<WP> const& s1, std::vector<WP> const& s2)
{
    std::vector<WP> r;
    for (auto const& p1 : s1)
        for (auto const& p2 : s2)
        {
            WP w;
            for (int d = 0; d < 3; ++d)
                w.off[d] = p1.off[d] + p2.off[d];
            w.coef = p1.coef * p2.coef;
            r.push_back(w);
        }
    return r;
}
inline std::vector<WP> tensor(std::vector<WP> const& s1, std::vector<WP> const& s2,
                              std::vector<WP> const& s3)
{
    std::vector<WP> r;
    for (auto const& p1 : s1)
        for (auto const& p2 : s2)
            for (auto const& p3 : s3)
            {
                WP w;
                for (int d = 0; d < 3; ++d)
                    w.off[d] = p1.off[d] + p2.off[d] + p3.off[d];
                w.coef = p1.coef * p2.coef * p3.coef;
                r.push_back(w);
            }
    return r;
}

template<int D>
double project(std::vector<WP> const& wps, Field const& f, std::array<int, D> idx)
{
    double result = 0.;
    for (auto const& wp : wps)
    {
        auto i = idx;
        for (int d = 0; d < D; ++d)
            i[d] += wp.off[d];
        result += wp.coef * f.at(i);
    }
    return result;
}

// projection tables (gridlayout_hybrid_yee.hpp:450-745)
template<int D>
struct Proj
{
    enum
    {
        X = 0,
        Y = 1,
        Z = 2
    };
    static std::vector<WP> momentsToE(int c) { return dirInterp<D>(c, true); }
    static std::vector<WP> JToMoments(int c) { return dirInterp<D>(c, false); }
    static std::vector<WP> BxToEx()
    {
        return tensor(dirInterp<D>(X, true), dirInterp<D>(Y, false), dirInterp<D>(Z, false));
    }
    static std::vector<WP> ByToEx() { return dirInterp<D>(Z, false); }
    static std::vector<WP> BzToEx() { return dirInterp<D>(Y, false); }
    static std::vector<WP> BxToEy() { return dirInterp<D>(Z, false); }
    static std::vector<WP> ByToEy()
    {
        return tensor(dirInterp<D>(X, false), dirInterp<D>(Y, true), dirInterp<D>(Z, false));
    }
    static std::vector<WP> BzToEy() { return dirInterp<D>(X, false); }
    static std::vector<WP> BxToEz() { return dirInterp<D>(Y, false); }
    static std::vector<WP> ByToEz() { return dirInterp<D>(X, false); }
    static std::vector<WP> BzToEz()
    {
        return tensor(dirInterp<D>(X, false), dirInterp<D>(Y, false), dirInterp<D>(Z, true));
    }
};

// ------------------------------------------------------------------------------------------
// B-spline weights (interpolator.hpp:54-125) and start shift (interpolator.hpp:519-549)
// ------------------------------------------------------------------------------------------
template<int O>
inline void computeWeight(double normalizedPos, int startIndex, double* w)
{
    if constexpr (O == 1)
    {
        w[1] = normalizedPos - static_cast<double>(startIndex);
        w[0] = 1. - w[1];
    }
    else if constexpr (O == 2)
    {
        auto index   = startIndex + 1;
        auto delta   = static_cast<double>(index) - normalizedPos;
        double coef1 = 0.5 + delta;
        double coef2 = delta;
        double coef3 = 0.5 - delta;
        w[0]         = 0.5 * coef1 * coef1;
        w[1]         = 0.75 - coef2 * coef2;
        w[2]         = 0.5 * coef3 * coef3;
    }
    else
    {
        constexpr double _4_over_3 = 4. / 3.;
        constexpr double _2_over_3 = 2. / 3.;
        auto index                 = static_cast<double>(startIndex) - normalizedPos;
        double coef1               = 1. + 0.5 * index;
        double coef2               = index + 1;
        double coef3               = index + 2;
        double coef4               = 1. - 0.5 * (index + 3);
        double coef2_sq            = coef2 * coef2;
        double coef2_cub           = coef2_sq * coef2;
        double coef3_sq            = coef3 * coef3;
        double coef3_cub           = coef3_sq * coef3;
        w[0]                       = _4_over_3 * coef1 * coef1 * coef1;
        w[1]                       = _2_over_3 - coef2_sq - 0.5 * coef2_cub;
        w[2]                       = _2_over_3 - coef3_sq + 0.5 * coef3_cub;
        w[3]                       = _4_over_3 * coef4 * coef4 * coef4;
    }
}

template<int O>
inline int startLeftShift(Cent c, double delta)
{
    if constexpr (O == 1)
        return c == primal ? 0 : (delta < .5 ? 1 : 0);
    else if constexpr (O == 2)
        return c == primal ? (delta < .5 ? 1 : 0) : 1;
    else
        return c == primal ? 1 : (delta < .5 ? 2 : 1);
}

// indexAndWeights_ (interpolator.hpp:379-406)
template<int D, int O>
struct IW
{
    std::array<int, D> start;
    std::array<std::array<double, O + 1>, D> w;
};
template<int D, int O>
inline IW<D, O> indexAndWeights(Layout<D> const& L, Cent c, std::array<int, D> const& iCellAMR,
                                std::array<double, D> const& delta)
{
    IW<D, O> r;
    for (int d = 0; d < D; ++d)
    {
        std::uint32_t icl = static_cast<std::uint32_t>(L.amrToLocal(iCellAMR[d], d));
        r.start[d]        = static_cast<int>(icl - static_cast<std::uint32_t>(startLeftShift<O>(c, delta[d])));
        double normalizedPos = icl + delta[d];
        if (c == dual)
            normalizedPos -= .5;
        computeWeight<O>(normalizedPos, r.start[d], r.w[d].data());
    }
    return r;
}

// ------------------------------------------------------------------------------------------
// Particles (particle.hpp:37-61) and ParticleArray + CellMap (particle_array.hpp:21-238,
// cellmap.hpp:25-220,347-470, indexer.hpp:24-80)
// ------------------------------------------------------------------------------------------
template<int D>
struct Particle
{
    double weight = 0;
    double charge = 0;
    std::array<int, D> iCell{};
    std::array<double, D> delta{};
    std::array<double, 3> v{};
};

template<int D>
struct CellMap
{
    Box<D> box;
    std::vector<std::vector<std::size_t>> lists;

    explicit CellMap(Box<D> b = {})
        : box{b}
        , lists(b.size())
    {
    }
    std::vector<std::size_t>* list(std::array<int, D> const& c)
    {
        if (!box.contains(c))
            return nullptr; // outside the map: the reference indexes out of range (UB); skip
        return &lists[box.lin(c)];
    }
    void add(std::array<int, D> const& c, std::size_t i)
    {
        if (auto l = list(c))
            l->push_back(i);
    }
    static void remove(std::vector<std::size_t>& l, std::size_t i)
    {
        auto it = std::find(l.begin(), l.end(), i);
        if (it != l.end())
            l.erase(it);
    }
    static void updateIndex(std::vector<std::size_t>& l, std::size_t o, std::size_t n)
    {
        auto it = std::find(l.begin(), l.end(), o);
        if (it != l.end())
            *it = n;
    }
};

template<int D>
struct ParticleArray
{
    std::vector<Particle<D>> p;
    CellMap<D> map;

    explicit ParticleArray(Box<D> b = {})
        : map{b}
    {
    }
    std::size_t size() const { return p.size(); }
    void clear()
    {
        p.clear();
        for (auto& l : map.lists)
            l.clear();
    }
    void push_back(Particle<D> const& q)
    {
        p.push_back(q);
        map.add(q.iCell, p.size() - 1);
    }
    // particle_array.hpp:183-190, cellmap.hpp:391-404
    void change_icell(std::array<int, D> const& nc, std::size_t i)
    {
        auto old    = p[i].iCell;
        p[i].iCell = nc;
        if (map.box.size() > 1)
        {
            if (auto ol = map.list(old))
                CellMap<D>::remove(*ol, i);
            map.add(nc, i);
        }
    }
    // cellmap.hpp:409-452 ; returns the pivot of [b, e)
    template<typename Pred>
    std::size_t partition(std::size_t b, std::size_t e, Pred&& pred)
    {
        std::size_t toSwap = e - 1;
        std::size_t pivot  = e;
        map.box.for_each([&](std::array<int, D> const& cell) {
            if (pred(cell))
                return;
            auto& items = map.lists[map.box.lin(cell)];
            for (std::size_t k = 0; k < items.size(); ++k)
            {
                std::size_t cur = items[k];
                if (cur >= b && cur < e)
                {
                    --pivot;
                    if (cur < toSwap)
                    {
                        while (!pred(p[toSwap].iCell) && toSwap > cur)
                            --toSwap;
                        if (toSwap > cur)
                        {
                            CellMap<D>::updateIndex(items, cur, toSwap);
                            if (auto l = map.list(p[toSwap].iCell))
                                CellMap<D>::updateIndex(*l, toSwap, cur);
                            std::swap(p[cur], p[toSwap]);
                            --toSwap;
                        }
                    }
                }
            }
        });
        return pivot;
    }
    // cellmap.hpp:457-470 (erase of a tail range)
    void erase_from(std::size_t from)
    {
        for (std::size_t i = from; i < p.size(); ++i)
            if (auto l = map.list(p[i].iCell))
                CellMap<D>::remove(*l, i);
        p.erase(p.begin() + static_cast<std::ptrdiff_t>(from), p.end());
    }
    // cellmap.hpp:347-360 export_if
    template<typename Pred>
    void export_if(ParticleArray& dest, Pred&& pred) const
    {
        map.box.for_each([&](std::array<int, D> const& cell) {
            if (!pred(cell))
                return;
            for (auto i : map.lists[map.box.lin(cell)])
                dest.push_back(p[i]);
        });
    }
};

// ------------------------------------------------------------------------------------------
// Interpolator: gather (interpolator.hpp:420-456, MeshToParticle :152-264) and deposit
// (interpolator.hpp:468-510, ParticleToMesh :277-363)
// ------------------------------------------------------------------------------------------
template<int D, int O>
double meshToParticle(Field const& f, IW<D, O> const& iwd, IW<D, O> const& iwp)
{
    // per-direction centering of the field selects the dual or primal index/weights
    std::array<int, D> st;
    std::array<double const*, D> w;
    for (int d = 0; d < D; ++d)
    {
        bool isPrimal = centering(f.q, d) == primal;
        st[d]         = isPrimal ? iwp.start[d] : iwd.start[d];
        w[d]          = isPrimal ? iwp.w[d].data() : iwd.w[d].data();
    }
    constexpr int S = O + 1;
    if constexpr (D == 1)
    {
        double r = 0.;
        for (int ik = 0; ik < S; ++ik)
            r += f.v[f.idx(st[0] + ik)] * w[0][ik];
        return r;
    }
    else if constexpr (D == 2)
    {
        double r = 0.;
        for (int ix = 0; ix < S; ++ix)
        {
            double Y = 0.;
            for (int iy = 0; iy < S; ++iy)
                Y += f.v[f.idx(st[0] + ix, st[1] + iy)] * w[1][iy];
            r += Y * w[0][ix];
        }
        return r;
    }
    else
    {
        double r = 0.;
        for (int ix = 0; ix < S; ++ix)
        {
            double Y = 0.;
            for (int iy = 0; iy < S; ++iy)
            {
                double Z = 0.;
                for (int iz = 0; iz < S; ++iz)
                    Z += f.v[f.idx(st[0] + ix, st[1] + iy, st[2] + iz)] * w[2][iz];
                Y += Z * w[1][iy];
            }
            r += Y * w[0][ix];
        }
        return r;
    }
}

template<int D, int O>
void gather(Layout<D> const& L, EM const& em, Particle<D> const& part, std::array<double, 3>& pE,
            std::array<double, 3>& pB)
{
    auto iwd = indexAndWeights<D, O>(L, dual, part.iCell, part.delta);
    auto iwp = indexAndWeights<D, O>(L, primal, part.iCell, part.delta);
    for (int c = 0; c < 3; ++c)
        pE[c] = meshToParticle<D, O>(em.E[c], iwd, iwp);
    for (int c = 0; c < 3; ++c)
        pB[c] = meshToParticle<D, O>(em.B[c], iwd, iwp);
}

template<int D, int O>
void particleToMesh(Field& f, double deposit, IW<D, O> const& iw)
{
    constexpr int S = O + 1;
    if constexpr (D == 1)
    {
        for (int ik = 0; ik < S; ++ik)
            f.v[f.idx(iw.start[0] + ik)] += deposit * iw.w[0][ik];
    }
    else if constexpr (D == 2)
    {
        for (int ix = 0; ix < S; ++ix)
            for (int iy = 0; iy < S; ++iy)
                f.v[f.idx(iw.start[0] + ix, iw.start[1] + iy)] += deposit * iw.w[0][ix] * iw.w[1][iy];
    }
    else
    {
        for (int ix = 0; ix < S; ++ix)
            for (int iy = 0; iy < S; ++iy)
                for (int iz = 0; iz < S; ++iz)
                    f.v[f.idx(iw.start[0] + ix, iw.start[1] + iy, iw.start[2] + iz)]
                        += deposit * iw.w[0][ix] * iw.w[1][iy] * iw.w[2][iz];
    }
}

template<int D, int O>
void depositOne(Layout<D> const& L, Particle<D> const& part, Field& n, Field& rho, VecField& F,
                double coef)
{
    auto iw = indexAndWeights<D, O>(L, primal, part.iCell, part.delta);
    particleToMesh<D, O>(n, 1. * part.weight * coef, iw);
    particleToMesh<D, O>(rho, part.charge * part.weight * coef, iw);
    particleToMesh<D, O>(F[0], part.v[0] * part.weight * coef, iw);
    particleToMesh<D, O>(F[1], part.v[1] * part.weight * coef, iw);
    particleToMesh<D, O>(F[2], part.v[2] * part.weight * coef, iw);
}

template<int D, int O>
void deposit(Layout<D> const& L, ParticleArray<D> const& arr, std::size_t b, std::size_t e,
             Field& n, Field& rho, VecField& F, double coef = 1.)
{
    for (std::size_t i = b; i < e; ++i)
        depositOne<D, O>(L, arr.p[i], n, rho, F, coef);
}

// ------------------------------------------------------------------------------------------
// Boris pusher (boris.hpp:93-300)
// ------------------------------------------------------------------------------------------
struct PushError
{
    bool set = false;
    double delta = 0, vel = 0;
    int stage = 0; // 1 = prePush, 2 = postPush
};

using CellPred1 = std::function<bool(std::array<int, 1> const&)>;

template<int D>
struct Selector
{
    // empty -> noop (no partition); otherwise a cell predicate used by CellMap::partition
    std::function<bool(std::array<int, D> const&)> pred;
};

template<int D, int O>
struct BorisPusher
{
    std::array<double, D> halfDtOverDl{};
    double dt = 0;

    void setMeshAndTimeStep(std::array<double, D> const& ms, double ts)
    {
        for (int d = 0; d < D; ++d)
            halfDtOverDl[d] = 0.5 * ts / ms[d];
        dt = ts;
    }

    // boris.hpp:156-172
    bool advancePosition(Particle<D> const& in, Particle<D>& out, std::array<int, D>& newCell,
                         PushError& err, int stage) const
    {
        for (int d = 0; d < D; ++d)
        {
            double const delta = in.delta[d] + static_cast<double>(halfDtOverDl[d] * in.v[d]);
            if (std::abs(delta) > 2)
            {
                err = PushError{true, delta, in.v[d], stage};
                return false;
            }
            auto const iCell = static_cast<int>(std::floor(delta));
            out.delta[d]     = delta - iCell;
            newCell[d]       = iCell + in.iCell[d];
        }
        return true;
    }

    // boris.hpp:240-300
    static void accelerate(Particle<D>& part, std::array<double, 3> const& pE,
                           std::array<double, 3> const& pB, double dto2m)
    {
        double const coef1 = part.charge * dto2m;
        double velx1       = part.v[0] + coef1 * pE[0];
        double vely1       = part.v[1] + coef1 * pE[1];
        double velz1       = part.v[2] + coef1 * pE[2];
        double const rx    = coef1 * pB[0];
        double const ry    = coef1 * pB[1];
        double const rz    = coef1 * pB[2];
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
        double const velx2 = (mxx * velx1 + mxy * vely1 + mxz * velz1) * invDet;
        double const vely2 = (myx * velx1 + myy * vely1 + myz * velz1) * invDet;
        double const velz2 = (mzx * velx1 + mzy * vely1 + mzz * velz1) * invDet;
        velx1              = velx2 + coef1 * pE[0];
        vely1              = vely2 + coef1 * pE[1];
        velz1              = velz2 + coef1 * pE[2];
        part.v[0]          = velx1;
        part.v[1]          = vely1;
        part.v[2]          = velz1;
    }

    // boris.hpp:93-138 with rangeIn == rangeOut (in place, as the IonUpdater calls it).
    // Returns the end of the range selected by `second` (begin is b).
    // firstEnd (optional) receives the end of the range selected by `first`.
    std::size_t move(ParticleArray<D>& arr, std::size_t b, std::size_t e, EM const& em, double mass,
                     Layout<D> const& L, Selector<D> const& first, Selector<D> const& second,
                     PushError& err, std::size_t* firstEnd = nullptr) const
    {
        // prePushStep_ (boris.hpp:180-216)
        for (std::size_t i = b; i < e; ++i)
        {
            std::array<int, D> nc{};
            if (!advancePosition(arr.p[i], arr.p[i], nc, err, 1))
                return b;
            if (nc != arr.p[i].iCell)
                arr.change_icell(nc, i);
        }
        std::size_t e1 = first.pred ? arr.partition(b, e, first.pred) : e;
        if (firstEnd)
            *firstEnd = e1;
        double const dto2m = 0.5 * dt / mass;
        for (std::size_t i = b; i < e1; ++i)
        {
            auto& part = arr.p[i];
            std::array<double, 3> pE, pB;
            gather<D, O>(L, em, part, pE, pB);
            accelerate(part, pE, pB, dto2m);
            std::array<int, D> nc{};
            if (!advancePosition(part, part, nc, err, 2))
                return b;
            if (nc != part.iCell)
                arr.change_icell(nc, i);
        }
        return second.pred ? arr.partition(b, e1, second.pred) : e1;
    }
};

// ------------------------------------------------------------------------------------------
// Ions (ions.hpp:75-144, ion_population.hpp, particle_pack.hpp)
// ------------------------------------------------------------------------------------------
template<int D>
struct Population
{
    double mass = 1.;
    ParticleArray<D> domain, patchGhost, levelGhost, levelGhostOld, levelGhostNew;
    Field n, rho;
    VecField F;
};

template<int D>
struct Ions
{
    std::vector<Population<D>> pops;
    Field rho, massDensity;
    VecField Vi;

    // ions.hpp:75-90
    void computeChargeDensity()
    {
        rho.zero();
        for (auto const& pop : pops)
            for (std::size_t i = 0; i < rho.v.size(); ++i)
                rho.v[i] = rho.v[i] + pop.rho.v[i];
    }
    // ions.hpp:92-108
    void computeMassDensity()
    {
        massDensity.zero();
        for (auto const& pop : pops)
            for (std::size_t i = 0; i < massDensity.v.size(); ++i)
                massDensity.v[i] = massDensity.v[i] + pop.n.v[i] * pop.mass;
    }
    // ions.hpp:111-144
    void computeBulkVelocity()
    {
        computeMassDensity();
        Vi.zero();
        for (auto const& pop : pops)
            for (int c = 0; c < 3; ++c)
                for (std::size_t i = 0; i < Vi[c].v.size(); ++i)
                    Vi[c].v[i] = Vi[c].v[i] + pop.F[c].v[i] * pop.mass;
        for (int c = 0; c < 3; ++c)
            for (std::size_t i = 0; i < Vi[c].v.size(); ++i)
                Vi[c].v[i] = Vi[c].v[i] / massDensity.v[i];
    }
};

// ------------------------------------------------------------------------------------------
// IonUpdater (ion_updater.hpp:22-295) with UpdaterSelectionBoxing (:119-163)
// ------------------------------------------------------------------------------------------
enum class UpdaterMode { domain_only = 1, all = 2 };

template<int D>
struct Boxing
{
    Layout<D> layout;
    std::vector<Box<D>> nonLevelGhostBox;
    Box<D> domainBox;
    Box<D> ghostBox;

    Boxing(Layout<D> const& L, std::vector<Box<D>> nlgb)
        : layout{L}
        , nonLevelGhostBox{std::move(nlgb)}
        , domainBox{L.amrBox()}
        , ghostBox{L.amrBox().grown(L.pg)}
    {
    }
    Selector<D> noop() const { return {}; }
    Selector<D> inDomainBox() const
    {
        auto b = domainBox;
        return {[b](auto const& c) { return b.contains(c); }};
    }
    Selector<D> inGhostBox() const
    {
        auto b = ghostBox;
        return {[b](auto const& c) { return b.contains(c); }};
    }
    Selector<D> inNonLevelGhostBox() const
    {
        auto bs = nonLevelGhostBox;
        return {[bs](auto const& c) {
            for (auto const& b : bs)
                if (b.contains(c))
                    return true;
            return false;
        }};
    }
    Selector<D> inGhostLayer() const
    {
        auto g = ghostBox;
        auto d = domainBox;
        return {[g, d](auto const& c) { return g.contains(c) && !d.contains(c); }};
    }
};

template<int D, int O>
struct IonUpdater
{
    BorisPusher<D, O> pusher;
    ParticleArray<D> tmp;

    // ion_updater.hpp:88-107 (+ resetMoments moments.hpp:15-23)
    PushError updatePopulations(Ions<D>& ions, EM const& em, Boxing<D> const& boxing, double dt,
                                UpdaterMode mode)
    {
        for (auto& pop : ions.pops)
        {
            pop.n.zero();
            pop.rho.zero();
            pop.F.zero();
        }
        pusher.setMeshAndTimeStep(boxing.layout.dx, dt);
        PushError err;
        if (mode == UpdaterMode::domain_only)
            domain_(ions, em, boxing, err);
        else
            all_(ions, em, boxing, err);
        return err;
    }

    // ion_updater.hpp:111-116
    void updateIons(Ions<D>& ions)
    {
        ions.computeChargeDensity();
        ions.computeBulkVelocity();
    }

private:
    // ion_updater.hpp:169-219
    void domain_(Ions<D>& ions, EM const& em, Boxing<D> const& bx, PushError& err)
    {
        auto const& L = bx.layout;
        for (auto& pop : ions.pops)
        {
            tmp = pop.domain;
            auto e = pusher.move(tmp, 0, tmp.size(), em, pop.mass, L, bx.noop(),
                                 bx.inNonLevelGhostBox(), err);
            if (err.set)
                return;
            deposit<D, O>(L, tmp, 0, e, pop.n, pop.rho, pop.F);
            if (pop.levelGhost.size())
            {
                tmp = pop.levelGhost;
                auto e2 = pusher.move(tmp, 0, tmp.size(), em, pop.mass, L, bx.inGhostBox(),
                                      bx.inDomainBox(), err);
                if (err.set)
                    return;
                deposit<D, O>(L, tmp, 0, e2, pop.n, pop.rho, pop.F);
            }
        }
    }

    // ion_updater.hpp:226-295
    void all_(Ions<D>& ions, EM const& em, Boxing<D> const& bx, PushError& err)
    {
        auto const& L = bx.layout;
        for (auto& pop : ions.pops)
        {
            auto& dom     = pop.domain;
            auto inDomain = pusher.move(dom, 0, dom.size(), em, pop.mass, L, bx.noop(),
                                        bx.inDomainBox(), err);
            if (err.set)
                return;
            // now_ghosts = [inDomain, size); not_level_ghosts = inNonLevelGhostBox(now_ghosts)
            auto nlg = dom.partition(inDomain, dom.size(), bx.inNonLevelGhostBox().pred);
            for (std::size_t i = inDomain; i < nlg; ++i)
                pop.patchGhost.push_back(dom.p[i]);
            dom.erase_from(inDomain);

            if (pop.levelGhost.size())
            {
                auto& lg = pop.levelGhost;
                auto e   = pusher.move(lg, 0, lg.size(), em, pop.mass, L, bx.inGhostBox(),
                                       bx.inGhostLayer(), err);
                if (err.set)
                    return;
                auto dbox = bx.domainBox;
                lg.export_if(dom, [&](auto const& c) { return dbox.contains(c); });
                lg.erase_from(e);
            }
            deposit<D, O>(L, dom, 0, dom.size(), pop.n, pop.rho, pop.F);
            deposit<D, O>(L, pop.patchGhost, 0, pop.patchGhost.size(), pop.n, pop.rho, pop.F);
        }
    }
};

// ------------------------------------------------------------------------------------------
// Field solve: Faraday (faraday.hpp:27-95), Ampere (ampere.hpp:26-94), Ohm (ohm.hpp:48-269),
// Electrons (electrons.hpp:102-128, 202-212), average (algorithm.hpp:68-76)
// ------------------------------------------------------------------------------------------
template<int D>
void faraday(Layout<D> const& L, VecField const& B, VecField const& E, VecField& Bnew, double dt)
{
    auto const &Ex = E[0], &Ey = E[1], &Ez = E[2];
    evalOnBox<D>(L, qBx, [&](std::array<int, D> const& i) {
        if constexpr (D == 1)
            Bnew[0].at(i) = B[0].at(i);
        if constexpr (D == 2)
            Bnew[0].at(i) = B[0].at(i) - dt * deriv<D>(L, Ez, 1, i);
        if constexpr (D == 3)
            Bnew[0].at(i) = B[0].at(i) - dt * deriv<D>(L, Ez, 1, i) + dt * deriv<D>(L, Ey, 2, i);
    });
    evalOnBox<D>(L, qBy, [&](std::array<int, D> const& i) {
        if constexpr (D == 1 || D == 2)
            Bnew[1].at(i) = B[1].at(i) + dt * deriv<D>(L, Ez, 0, i);
        if constexpr (D == 3)
            Bnew[1].at(i) = B[1].at(i) - dt * deriv<D>(L, Ex, 2, i) + dt * deriv<D>(L, Ez, 0, i);
    });
    evalOnBox<D>(L, qBz, [&](std::array<int, D> const& i) {
        if constexpr (D == 1)
            Bnew[2].at(i) = B[2].at(i) - dt * deriv<D>(L, Ey, 0, i);
        else
            Bnew[2].at(i) = B[2].at(i) - dt * deriv<D>(L, Ey, 0, i) + dt * deriv<D>(L, Ex, 1, i);
    });
}

template<int D>
void ampere(Layout<D> const& L, VecField const& B, VecField& J)
{
    auto const &Bx = B[0], &By = B[1], &Bz = B[2];
    evalOnShrunkGhostBox<D>(L, qJx, 1, [&](std::array<int, D> const& i) {
        if constexpr (D == 1)
            J[0].at(i) = 0.0;
        if constexpr (D == 2)
            J[0].at(i) = deriv<D>(L, Bz, 1, i);
        if constexpr (D == 3)
            J[0].at(i) = deriv<D>(L, Bz, 1, i) - deriv<D>(L, By, 2, i);
    });
    evalOnShrunkGhostBox<D>(L, qJy, 1, [&](std::array<int, D> const& i) {
        if constexpr (D == 1 || D == 2)
            J[1].at(i) = -deriv<D>(L, Bz, 0, i);
        if constexpr (D == 3)
            J[1].at(i) = deriv<D>(L, Bx, 2, i) - deriv<D>(L, Bz, 0, i);
    });
    evalOnShrunkGhostBox<D>(L, qJz, 1, [&](std::array<int, D> const& i) {
        if constexpr (D == 1)
            J[2].at(i) = deriv<D>(L, By, 0, i);
        else
            J[2].at(i) = deriv<D>(L, By, 0, i) - deriv<D>(L, Bx, 1, i);
    });
}

// StandardHybridElectronFluxComputer::computeBulkVelocity + isothermal closure
template<int D>
void electrons(Layout<D> const& L, double Te, VecField const& J, VecField const& Vi, Field const& Ne,
               VecField& Ve, Field& Pe)
{
    using P = Proj<D>;
    auto pjx = P::JToMoments(0), pjy = P::JToMoments(1), pjz = P::JToMoments(2);
    evalOnBox<D>(L, qMom, [&](std::array<int, D> const& i) {
        auto const JxOnVx = project<D>(pjx, J[0], i);
        auto const JyOnVy = project<D>(pjy, J[1], i);
        auto const JzOnVz = project<D>(pjz, J[2], i);
        Ve[0].at(i)       = Vi[0].at(i) - JxOnVx / Ne.at(i);
        Ve[1].at(i)       = Vi[1].at(i) - JyOnVy / Ne.at(i);
        Ve[2].at(i)       = Vi[2].at(i) - JzOnVz / Ne.at(i);
    });
    for (std::size_t k = 0; k < Pe.v.size(); ++k)
        Pe.v[k] = Ne.v[k] * Te;
}

enum class HyperMode { constant = 0, spatial = 1 };

template<int D>
void ohm(Layout<D> const& L, double eta, double nu, HyperMode mode, Field const& n, VecField const& Ve,
         Field const& Pe, VecField const& B, VecField const& J, VecField& E)
{
    using P              = Proj<D>;
    double const lvlCoef = 1. / std::pow(4, L.level);
    auto hyper = [&](int comp, std::array<int, D> const& i, std::vector<WP> const& bx,
                     std::vector<WP> const& by, std::vector<WP> const& bz,
                     std::vector<WP> const& nP) {
        if (mode == HyperMode::constant)
            return -nu * laplacian<D>(L, J[comp], i);
        auto const BxOnE = project<D>(bx, B[0], i);
        auto const ByOnE = project<D>(by, B[1], i);
        auto const BzOnE = project<D>(bz, B[2], i);
        auto const nOnE  = project<D>(nP, n, i);
        auto b           = std::sqrt(BxOnE * BxOnE + ByOnE * ByOnE + BzOnE * BzOnE);
        return -nu * (b / (nOnE + 0.1) + 1) * lvlCoef * laplacian<D>(L, J[comp], i);
    };
    auto const mEx = P::momentsToE(0), mEy = P::momentsToE(1), mEz = P::momentsToE(2);
    auto const bxEx = P::BxToEx(), byEx = P::ByToEx(), bzEx = P::BzToEx();
    auto const bxEy = P::BxToEy(), byEy = P::ByToEy(), bzEy = P::BzToEy();
    auto const bxEz = P::BxToEz(), byEz = P::ByToEz(), bzEz = P::BzToEz();
    std::vector<WP> const ident{WP{{0, 0, 0}, 1.0}};

    evalOnBox<D>(L, qEx, [&](std::array<int, D> const& i) {
        auto const vyOnEx = project<D>(mEx, Ve[1], i);
        auto const vzOnEx = project<D>(mEx, Ve[2], i);
        auto const byOnEx = project<D>(byEx, B[1], i);
        auto const bzOnEx = project<D>(bzEx, B[2], i);
        double ideal      = -vyOnEx * bzOnEx + vzOnEx * byOnEx;
        auto const nOnEx  = project<D>(mEx, n, i);
        auto gradP        = deriv<D>(L, Pe, 0, i);
        double press      = -gradP / nOnEx;
        double resist     = eta * project<D>(ident, J[0], i);
        E[0].at(i)        = ideal + press + resist + hyper(0, i, bxEx, byEx, bzEx, mEx);
    });
    evalOnBox<D>(L, qEy, [&](std::array<int, D> const& i) {
        auto const vxOnEy = project<D>(mEy, Ve[0], i);
        auto const vzOnEy = project<D>(mEy, Ve[2], i);
        auto const bxOnEy = project<D>(bxEy, B[0], i);
        auto const bzOnEy = project<D>(bzEy, B[2], i);
        double ideal      = -vzOnEy * bxOnEy + vxOnEy * bzOnEy;
        double press      = 0.;
        if constexpr (D >= 2)
        {
            auto const nOnEy = project<D>(mEy, n, i);
            auto gradP       = deriv<D>(L, Pe, 1, i);
            press            = -gradP / nOnEy;
        }
        double resist = eta * project<D>(ident, J[1], i);
        E[1].at(i)    = ideal + press + resist + hyper(1, i, bxEy, byEy, bzEy, mEy);
    });
    evalOnBox<D>(L, qEz, [&](std::array<int, D> const& i) {
        auto const vxOnEz = project<D>(mEz, Ve[0], i);
        auto const vyOnEz = project<D>(mEz, Ve[1], i);
        auto const bxOnEz = project<D>(bxEz, B[0], i);
        auto const byOnEz = project<D>(byEz, B[1], i);
        double ideal      = -vxOnEz * byOnEz + vyOnEz * bxOnEz;
        double press      = 0.;
        if constexpr (D >= 3)
        {
            auto const nOnEz = project<D>(mEz, n, i);
            auto gradP       = deriv<D>(L, Pe, 2, i);
            press            = -gradP / nOnEz;
        }
        double resist = eta * project<D>(ident, J[2], i);
        E[2].at(i)    = ideal + press + resist + hyper(2, i, bxEz, byEz, bzEz, mEz);
    });
}

inline void average(Field const& f1, Field const& f2, Field& avg)
{
    for (std::size_t i = 0; i < f1.v.size(); ++i)
        avg.v[i] = (f1.v[i] + f2.v[i]) * .5;
}

// ------------------------------------------------------------------------------------------
// Same-level fills for ONE periodic patch covering the whole domain (SAMRAI semantics as used
// by hybrid_hybrid_messenger_strategy.hpp:376-548; fill patterns
// field_variable_fill_pattern.hpp:27-125 (ghost copy, interior not overwritten) and :240-290
// (border sum/max over the intersection of ghost field boxes of periodic images);
// ParticleDomainFromGhostFillPattern for patch-ghost particles).
// ------------------------------------------------------------------------------------------
template<int D>
struct PeriodicFills
{
    Layout<D> L;

    // iterate the 3^D - 1 periodic images in lexicographic order of k in {-1,0,1}^D
    template<typename Fn>
    void images(Fn&& fn) const
    {
        int total = 1;
        for (int d = 0; d < D; ++d)
            total *= 3;
        for (int t = 0; t < total; ++t)
        {
            std::array<int, D> k;
            int r      = t;
            bool allz  = true;
            for (int d = D - 1; d >= 0; --d)
            {
                k[d] = (r % 3) - 1;
                r /= 3;
                if (k[d] != 0)
                    allz = false;
            }
            if (!allz)
                fn(k);
        }
    }

    // all-primal ghost box overlap between this patch and image k, in local node indices:
    // dst node i receives src node i - k*N.
    bool overlap(std::array<int, D> const& k, std::array<int, D>& lo, std::array<int, D>& hi) const
    {
        for (int d = 0; d < D; ++d)
        {
            int N   = L.ncells[d];
            int end = L.ghostEnd(primal, d);
            lo[d]   = std::max(0, k[d] * N);
            hi[d]   = std::min(end, end + k[d] * N);
            if (lo[d] > hi[d])
                return false;
        }
        return true;
    }

    // PatchFieldBorderSum: sum = copy(f); sum[overlap] += f[src]; f = sum
    void borderSum(Field& f) const
    {
        std::vector<double> sum = f.v;
        images([&](std::array<int, D> const& k) {
            std::array<int, D> lo, hi;
            if (!overlap(k, lo, hi))
                return;
            evalOnRange<D>(lo, hi, [&](std::array<int, D> const& i) {
                auto s = i;
                for (int d = 0; d < D; ++d)
                    s[d] -= k[d] * L.ncells[d];
                Field const& cf = f;
                sum[f.idx(i[0], D > 1 ? i[1] : 0, D > 2 ? i[2] : 0)] += cf.at(s);
            });
        });
        f.v = sum;
    }

    // PatchFieldBorderMax: f[overlap] = max(f[overlap], f[src])
    void borderMax(Field& f) const
    {
        std::vector<double> orig = f.v;
        images([&](std::array<int, D> const& k) {
            std::array<int, D> lo, hi;
            if (!overlap(k, lo, hi))
                return;
            evalOnRange<D>(lo, hi, [&](std::array<int, D> const& i) {
                auto s = i;
                for (int d = 0; d < D; ++d)
                    s[d] -= k[d] * L.ncells[d];
                auto& dst = f.at(i);
                dst       = std::max(dst, orig[f.idx(s[0], D > 1 ? s[1] : 0, D > 2 ? s[2] : 0)]);
            });
        });
    }

    // ghost fill of a Yee component: pure ghost nodes take the periodic interior value
    void fillGhosts(Field& f) const
    {
        std::array<int, D> lo, hi;
        for (int d = 0; d < D; ++d)
        {
            lo[d] = 0;
            hi[d] = L.ghostEnd(centering(f.q, d), d);
        }
        evalOnRange<D>(lo, hi, [&](std::array<int, D> const& i) {
            bool ghost = false;
            auto s     = i;
            for (int d = 0; d < D; ++d)
            {
                int a = L.physStart(), b = L.physEnd(centering(f.q, d), d);
                int N = L.ncells[d];
                if (i[d] < a)
                {
                    s[d] = i[d] + N;
                    ghost = true;
                }
                else if (i[d] > b)
                {
                    s[d] = i[d] - N;
                    ghost = true;
                }
            }
            if (ghost)
                f.at(i) = static_cast<Field const&>(f).at(s);
        });
    }
    void fillGhosts(VecField& v) const
    {
        for (int c = 0; c < 3; ++c)
            fillGhosts(v[c]);
    }

    // fillIonGhostParticles: patch-ghost particles of the images enter the domain with the
    // periodic shift, then the patch-ghost arrays are emptied
    void ghostParticlesToDomain(Population<D>& pop) const
    {
        auto dom = L.amrBox();
        images([&](std::array<int, D> const& k) {
            // destination region = domain ∩ shift(source ghost box, k*N)
            for (auto const& part : pop.patchGhost.p)
            {
                auto q = part;
                for (int d = 0; d < D; ++d)
                    q.iCell[d] += k[d] * L.ncells[d];
                if (dom.contains(q.iCell))
                    pop.domain.push_back(q);
            }
        });
        pop.patchGhost.clear();
    }
};

// ------------------------------------------------------------------------------------------
// Maxwellian loader (maxwellian_particle_initializer.hpp:137-198, .cpp:12-22;
// ParticleDeltaDistribution particle.hpp:18-27). Per-cell profiles are given as arrays in the
// AMR row-major cell order of layout.indices(AMRBox).
// ------------------------------------------------------------------------------------------
template<int D>
void loadMaxwellian(Layout<D> const& L, double const* n, std::array<double const*, 3> V,
                    std::array<double const*, 3> Vth, double charge, std::uint32_t ppc,
                    std::uint64_t seed, double cutoff, ParticleArray<D>& out)
{
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> deltaDistrib{0, 1. - std::numeric_limits<double>::epsilon()};
    std::size_t flat = 0;
    L.amrBox().for_each([&](std::array<int, D> const& cell) {
        std::size_t c = flat++;
        if (n[c] < cutoff)
            return;
        double const cellWeight = n[c] / ppc;
        for (std::uint32_t ip = 0; ip < ppc; ++ip)
        {
            std::array<double, 3> vel;
            {
                std::normal_distribution<> mx(V[0][c], Vth[0][c]);
                std::normal_distribution<> my(V[1][c], Vth[1][c]);
                std::normal_distribution<> mz(V[2][c], Vth[2][c]);
                vel[0] = mx(gen);
                vel[1] = my(gen);
                vel[2] = mz(gen);
            }
            Particle<D> p;
            p.weight = cellWeight;
            p.charge = charge;
            p.iCell  = cell;
            for (int d = 0; d < D; ++d)
                p.delta[d] = deltaDistrib(gen);
            p.v = vel;
            out.push_back(p);
        }
    });
}

} // namespace refcpu
