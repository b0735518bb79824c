// ref_harness.cpp -- CPU ORACLE support (test infrastructure only).
//
// Compiles the reference's OWN header-only numerics (Interpolator, GridLayout + Yee tables,
// Faraday, Ampere, ParticleArray/CellMap) from the PHARE source tree into oracle/_ref/ and
// exposes them through plain C entry points so tests can check the restatement in
// refcpu.hpp against the real thing. Built by oracle/Makefile only when the PHARE sources are
// present (-I <phare>/src); no reference source is copied into this repository.
// Headers that need the un-vendored cppdict (boris.hpp, ohm.hpp, ion_updater.hpp,
// electrons.hpp, ions.hpp) are NOT compiled: they are unbuildable here.
#include "core/data/field/field.hpp"
#include "core/data/grid/gridlayout.hpp"
#include "core/data/grid/gridlayoutimplyee.hpp"
#include "core/data/particles/particle.hpp"
#include "core/data/particles/particle_array.hpp"
#include "core/data/vecfield/vecfield.hpp"
#include "core/models/options/hybrid_options.hpp"
#include "core/numerics/ampere/ampere.hpp"
#include "core/numerics/faraday/faraday.hpp"
#include "core/numerics/interpolator/interpolator.hpp"
#include "core/utilities/types.hpp"
#include "amr/data/particles/refine/split.hpp"

// defined in the reference's maxwellian_particle_initializer.cpp, compiled beside this harness
// (its header includes the un-vendored cppdict, so the declaration is restated here)
namespace PHARE::core
{
void maxwellianVelocity(std::array<double, 3> const& V, std::array<double, 3> const& Vth,
                        std::mt19937_64& generator, std::array<double, 3>& partVelocity);
}

#include <array>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace
{
using namespace PHARE::core;
using PHARE::HybridFieldOptions;
using PHARE::HybridOptions;

// the (dimension, interp_order) pair the reference's HybridFieldOptions template expects
struct HarnessOpts
{
    std::size_t dimension;
    std::size_t interp_order;
};

template<std::size_t D, std::size_t O>
struct T
{
    static constexpr HarnessOpts opts{D, O};
    static constexpr auto fopts = HybridFieldOptions<opts>{};
    static constexpr auto hopts = HybridOptions<fopts>{};
    using GL                    = GridLayout<hopts>;
    using Fld                   = Field<D, HybridQuantity::Scalar>;
    using VF                    = VecField<Fld, HybridQuantity>;

    static GL layout(int const* n, double const* dx, int const* lo, int level = 0)
    {
        std::array<double, D> h;
        std::array<std::uint32_t, D> nc;
        Point<double, D> origin;
        Box<int, D> amr;
        for (std::size_t d = 0; d < D; ++d)
        {
            h[d]         = dx[d];
            nc[d]        = static_cast<std::uint32_t>(n[d]);
            origin[d]    = 0.;
            amr.lower[d] = lo[d];
            amr.upper[d] = lo[d] + n[d] - 1;
        }
        return GL{h, nc, origin, amr, level};
    }

    static Fld field(GL const& L, HybridQuantity::Scalar q, double* data, std::string name)
    {
        auto s = L.allocSize(q);
        return Fld{name, q, data, s};
    }

    // a VecField view over three caller buffers
    struct VFV
    {
        VF vf;
        std::array<Fld, 3> comps;
        VFV(GL const& L, std::string name, HybridQuantity::Vector vq, double* const* data)
            : vf{name, vq}
            , comps{field(L, HybridQuantity::componentsQuantities(vq)[0], data[0], name + "_x"),
                    field(L, HybridQuantity::componentsQuantities(vq)[1], data[1], name + "_y"),
                    field(L, HybridQuantity::componentsQuantities(vq)[2], data[2], name + "_z")}
        {
            vf.setBuffer(&comps);
        }
    };

    // the Interpolator reads Em.E() / Em.B(): VecField members called for their components
    struct EMV
    {
        VFV e, b;
        VF const& E;
        VF const& B;
        EMV(VFV&& ee, VFV&& bb)
            : e{std::move(ee)}
            , b{std::move(bb)}
            , E{e.vf}
            , B{b.vf}
        {
        }
    };

    static Particle<D> particle(int i, int const* ic, double const* de, double const* v,
                                double const* w, double const* q)
    {
        Particle<D> p;
        p.weight = w ? w[i] : 1.;
        p.charge = q ? q[i] : 1.;
        for (std::size_t d = 0; d < D; ++d)
        {
            p.iCell[d] = ic[i * D + d];
            p.delta[d] = de[i * D + d];
        }
        for (int c = 0; c < 3; ++c)
            p.v[c] = v ? v[i * 3 + c] : 0.;
        return p;
    }
};

using Q = HybridQuantity::Scalar;

Q qty(int q)
{
    // ids of the oracle: 0..2 B, 3..5 E, 6..8 J, 9 moments (rho)
    static constexpr Q t[] = {Q::Bx, Q::By, Q::Bz, Q::Ex, Q::Ey, Q::Ez, Q::Jx, Q::Jy, Q::Jz, Q::rho};
    return t[q];
}

template<std::size_t D, std::size_t O>
int layout_info(int const* n, int q, int* out)
{
    double dx[3]  = {1, 1, 1};
    int lo[3]     = {0, 0, 0};
    auto L        = T<D, O>::layout(n, dx, lo);
    auto s        = L.allocSize(qty(q));
    auto cent     = L.centering(qty(q));
    for (std::size_t d = 0; d < D; ++d)
    {
        out[d * 4 + 0] = static_cast<int>(s[d]);
        out[d * 4 + 1] = static_cast<int>(L.physicalStartIndex(qty(q), static_cast<Direction>(d)));
        out[d * 4 + 2] = static_cast<int>(L.physicalEndIndex(qty(q), static_cast<Direction>(d)));
        out[d * 4 + 3] = static_cast<int>(L.ghostEndIndex(qty(q), static_cast<Direction>(d)));
        (void)cent;
    }
    return 0;
}

template<std::size_t D, std::size_t O>
int gather(int const* n, double const* dx, int const* lo, double* const* E, double* const* B,
           long np, int const* ic, double const* de, double* outE, double* outB)
{
    using TT = T<D, O>;
    auto L   = TT::layout(n, dx, lo);
    typename TT::EMV em{typename TT::VFV{L, "E", HybridQuantity::Vector::E, E},
                        typename TT::VFV{L, "B", HybridQuantity::Vector::B, B}};
    Interpolator<D, O> interp;
    for (long i = 0; i < np; ++i)
    {
        auto p          = TT::particle(static_cast<int>(i), ic, de, nullptr, nullptr, nullptr);
        auto const [e, b] = interp(p, em, L);
        for (int c = 0; c < 3; ++c)
        {
            outE[i * 3 + c] = e[c];
            outB[i * 3 + c] = b[c];
        }
    }
    return 0;
}

template<std::size_t D, std::size_t O>
int depositF(int const* n, double const* dx, int const* lo, long np, int const* ic, double const* de,
             double const* v, double const* w, double const* q, double coef, double* dn,
             double* drho, double* const* dF)
{
    using TT = T<D, O>;
    auto L   = TT::layout(n, dx, lo);
    auto fn  = TT::field(L, Q::rho, dn, "n");
    auto fr  = TT::field(L, Q::rho, drho, "rho");
    typename TT::VFV F{L, "F", HybridQuantity::Vector::V, dF};
    std::vector<Particle<D>> parts;
    for (long i = 0; i < np; ++i)
        parts.push_back(TT::particle(static_cast<int>(i), ic, de, v, w, q));
    Interpolator<D, O> interp;
    interp(parts, fn, fr, F.vf, L, coef);
    return 0;
}

template<std::size_t D, std::size_t O>
int faradayF(int const* n, double const* dx, int const* lo, double* const* B, double* const* E,
             double* const* Bnew, double dt)
{
    using TT = T<D, O>;
    auto L   = TT::layout(n, dx, lo);
    typename TT::VFV b{L, "B", HybridQuantity::Vector::B, B};
    typename TT::VFV e{L, "E", HybridQuantity::Vector::E, E};
    typename TT::VFV bn{L, "Bn", HybridQuantity::Vector::B, Bnew};
    Faraday<typename TT::GL> far{L};
    far(b.vf, e.vf, bn.vf, dt);
    return 0;
}

template<std::size_t D, std::size_t O>
int ampereF(int const* n, double const* dx, int const* lo, double* const* B, double* const* J)
{
    using TT = T<D, O>;
    auto L   = TT::layout(n, dx, lo);
    typename TT::VFV b{L, "B", HybridQuantity::Vector::B, B};
    typename TT::VFV j{L, "J", HybridQuantity::Vector::J, J};
    Ampere<typename TT::GL> amp{L};
    amp(b.vf, j.vf);
    return 0;
}

// laplacian of a field of quantity q, written on the physical box of q
template<std::size_t D, std::size_t O>
int laplacianF(int const* n, double const* dx, int const* lo, int q, double* src, double* out)
{
    using TT = T<D, O>;
    auto L   = TT::layout(n, dx, lo);
    auto f   = TT::field(L, qty(q), src, "f");
    auto o   = TT::field(L, qty(q), out, "f");
    L.evalOnBox(o, [&](auto const&... ijk) {
        o(ijk...) = L.laplacian(f, {static_cast<std::uint32_t>(ijk)...});
    });
    return 0;
}

// GridLayout::project with one of the Yee tables, evaluated on the physical box of `dstq`
template<std::size_t D, std::size_t O>
int projectF(int const* n, double const* dx, int const* lo, int which, int srcq, int dstq,
             double* src, double* out)
{
    using TT = T<D, O>;
    using GL = typename TT::GL;
    auto L   = TT::layout(n, dx, lo);
    auto f   = TT::field(L, qty(srcq), src, "f");
    auto o   = TT::field(L, qty(dstq), out, "o");
    L.evalOnBox(o, [&](auto const&... ijk) {
        MeshIndex<D> idx{static_cast<std::uint32_t>(ijk)...};
        double r = 0.;
        switch (which)
        {
            case 0: r = GL::template project<GL::momentsToEx>(f, idx); break;
            case 1: r = GL::template project<GL::momentsToEy>(f, idx); break;
            case 2: r = GL::template project<GL::momentsToEz>(f, idx); break;
            case 3: r = GL::template project<GL::BxToEx>(f, idx); break;
            case 4: r = GL::template project<GL::ByToEx>(f, idx); break;
            case 5: r = GL::template project<GL::BzToEx>(f, idx); break;
            case 6: r = GL::template project<GL::BxToEy>(f, idx); break;
            case 7: r = GL::template project<GL::ByToEy>(f, idx); break;
            case 8: r = GL::template project<GL::BzToEy>(f, idx); break;
            case 9: r = GL::template project<GL::BxToEz>(f, idx); break;
            case 10: r = GL::template project<GL::ByToEz>(f, idx); break;
            case 11: r = GL::template project<GL::BzToEz>(f, idx); break;
            case 12: r = GL::template project<GL::JxToMoments>(f, idx); break;
            case 13: r = GL::template project<GL::JyToMoments>(f, idx); break;
            case 14: r = GL::template project<GL::JzToMoments>(f, idx); break;
        }
        o(ijk...) = r;
    });
    return 0;
}

// ParticleArray: push_back all, apply change_icell, partition by a box predicate; report the
// resulting order (as original positions) and the pivot
template<std::size_t D>
long partitionF(int const* mapLo, int const* mapHi, long np, int const* ic0, int const* ic1,
                int const* predLo, int const* predHi, long* order)
{
    Box<int, D> mapBox, pred;
    for (std::size_t d = 0; d < D; ++d)
    {
        mapBox.lower[d] = mapLo[d];
        mapBox.upper[d] = mapHi[d];
        pred.lower[d]   = predLo[d];
        pred.upper[d]   = predHi[d];
    }
    ParticleArray<D> arr{mapBox};
    for (long i = 0; i < np; ++i)
    {
        Particle<D> p;
        p.weight = static_cast<double>(i);
        for (std::size_t d = 0; d < D; ++d)
            p.iCell[d] = ic0[i * D + d];
        arr.push_back(p);
    }
    for (long i = 0; i < np; ++i)
    {
        std::array<int, D> nc;
        for (std::size_t d = 0; d < D; ++d)
            nc[d] = ic1[i * D + d];
        if (nc != arr[i].iCell)
            arr.change_icell(nc, static_cast<std::size_t>(i));
    }
    auto range = makeIndexRange(arr);
    auto sel   = arr.partition(range, [&](auto const& cell) { return isIn(Point{cell}, pred); });
    for (long i = 0; i < np; ++i)
        order[i] = static_cast<long>(arr[i].weight);
    return static_cast<long>(sel.iend());
}

// the Maxwellian loader's draws (MaxwellianParticleInitializer::loadParticles,
// maxwellian_particle_initializer.hpp:137-198) from the reference's own pieces: the cell order of
// GridLayout::indices(AMRBox) (gridlayout.hpp:219), maxwellianVelocity (compiled from
// maxwellian_particle_initializer.cpp:12-22), ParticleDeltaDistribution<double> (particle.hpp:18-27)
// and an mt19937_64 seeded as getRNG seeds it (maxwellian_particle_initializer.hpp:71-80). The
// loop is restated (the class itself includes the un-vendored cppdict); profiles are per cell in
// that order, the magnetic basis is not used (cartesian). Writes at most cap particles.
template<std::size_t D>
long loadMaxwellianF(int const* n, int const* lo, double const* dens, double const* const* V,
                     double const* const* Vth, double charge, unsigned ppc, unsigned long long seed,
                     double cutoff, long cap, int* ic, double* de, double* v, double* w, double* q)
{
    double dx[3] = {1., 1., 1.};
    auto const L = T<D, 1>::layout(n, dx, lo);
    auto const cells = L.indices(L.AMRBox());
    std::mt19937_64 gen(static_cast<std::size_t>(seed));
    ParticleDeltaDistribution<double> deltaDistrib;
    long np = 0;
    for (std::size_t flat = 0; flat < cells.size(); ++flat)
    {
        if (dens[flat] < cutoff)
            continue;
        double const cellWeight = dens[flat] / ppc;
        auto const cell = for_N_make_array<D>([&](auto i) { return std::get<i>(cells[flat]); });
        for (unsigned ip = 0; ip < ppc; ++ip)
        {
            std::array<double, 3> vel;
            maxwellianVelocity({V[0][flat], V[1][flat], V[2][flat]},
                               {Vth[0][flat], Vth[1][flat], Vth[2][flat]}, gen, vel);
            auto const delta = for_N_make_array<D>([&](auto) { return deltaDistrib(gen); });
            if (np < cap)
            {
                for (std::size_t d = 0; d < D; ++d)
                {
                    ic[np * D + d] = cell[d];
                    de[np * D + d] = delta[d];
                }
                for (int c = 0; c < 3; ++c)
                    v[np * 3 + c] = vel[c];
                w[np] = cellWeight;
                q[np] = charge;
            }
            ++np;
        }
    }
    return np;
}

template<template<std::size_t, std::size_t> class F>
struct Dispatch;

#define PHX_REF_DISPATCH(dim, order, fn, ...)                                                      \
    [&]() -> int {                                                                                 \
        switch (dim * 10 + order)                                                                  \
        {                                                                                          \
            case 11: return fn<1, 1>(__VA_ARGS__);                                                 \
            case 12: return fn<1, 2>(__VA_ARGS__);                                                 \
            case 13: return fn<1, 3>(__VA_ARGS__);                                                 \
            case 21: return fn<2, 1>(__VA_ARGS__);                                                 \
            case 22: return fn<2, 2>(__VA_ARGS__);                                                 \
            case 23: return fn<2, 3>(__VA_ARGS__);                                                 \
            case 31: return fn<3, 1>(__VA_ARGS__);                                                 \
            case 32: return fn<3, 2>(__VA_ARGS__);                                                 \
            case 33: return fn<3, 3>(__VA_ARGS__);                                                 \
        }                                                                                          \
        return -1;                                                                                 \
    }()

} // namespace

extern "C" {

int ref_layout_info(int dim, int order, int const* n, int q, int* out)
{
    return PHX_REF_DISPATCH(dim, order, layout_info, n, q, out);
}

void ref_weights(int order, int dual, int icell_local, double delta, int* start, double* w)
{
    auto one = [&](auto W, int shift) {
        *start    = icell_local - shift;
        double np = static_cast<double>(static_cast<std::uint32_t>(icell_local)) + delta;
        if (dual)
            np -= .5;
        std::array<double, nbrPointsSupport(decltype(W)::interp_order)> ww;
        W.computeWeight(np, *start, ww);
        for (std::size_t i = 0; i < ww.size(); ++i)
            w[i] = ww[i];
    };
    if (order == 1)
        one(Weighter<1>{}, dual ? Interpolator<1, 1>::computeStartLeftShift<QtyCentering, QtyCentering::dual>(delta)
                                : Interpolator<1, 1>::computeStartLeftShift<QtyCentering, QtyCentering::primal>(delta));
    else if (order == 2)
        one(Weighter<2>{}, dual ? Interpolator<1, 2>::computeStartLeftShift<QtyCentering, QtyCentering::dual>(delta)
                                : Interpolator<1, 2>::computeStartLeftShift<QtyCentering, QtyCentering::primal>(delta));
    else
        one(Weighter<3>{}, dual ? Interpolator<1, 3>::computeStartLeftShift<QtyCentering, QtyCentering::dual>(delta)
                                : Interpolator<1, 3>::computeStartLeftShift<QtyCentering, QtyCentering::primal>(delta));
}

int ref_gather(int dim, int order, int const* n, double const* dx, int const* lo, double* const* E,
               double* const* B, long np, int const* ic, double const* de, double* outE, double* outB)
{
    return PHX_REF_DISPATCH(dim, order, gather, n, dx, lo, E, B, np, ic, de, outE, outB);
}

int ref_deposit(int dim, int order, int const* n, double const* dx, int const* lo, long np,
                int const* ic, double const* de, double const* v, double const* w, double const* q,
                double coef, double* dn, double* drho, double* const* dF)
{
    return PHX_REF_DISPATCH(dim, order, depositF, n, dx, lo, np, ic, de, v, w, q, coef, dn, drho, dF);
}

int ref_faraday(int dim, int order, int const* n, double const* dx, int const* lo, double* const* B,
                double* const* E, double* const* Bnew, double dt)
{
    return PHX_REF_DISPATCH(dim, order, faradayF, n, dx, lo, B, E, Bnew, dt);
}

int ref_ampere(int dim, int order, int const* n, double const* dx, int const* lo, double* const* B,
               double* const* J)
{
    return PHX_REF_DISPATCH(dim, order, ampereF, n, dx, lo, B, J);
}

int ref_laplacian(int dim, int order, int const* n, double const* dx, int const* lo, int q,
                  double* src, double* out)
{
    return PHX_REF_DISPATCH(dim, order, laplacianF, n, dx, lo, q, src, out);
}

int ref_project(int dim, int order, int const* n, double const* dx, int const* lo, int which,
                int srcq, int dstq, double* src, double* out)
{
    return PHX_REF_DISPATCH(dim, order, projectF, n, dx, lo, which, srcq, dstq, src, out);
}

long ref_load_maxwellian(int dim, int const* n, int const* lo, double const* dens,
                         double const* const* V, double const* const* Vth, double charge, unsigned ppc,
                         unsigned long long seed, double cutoff, long cap, int* ic, double* de,
                         double* v, double* w, double* q)
{
    if (dim == 1)
        return loadMaxwellianF<1>(n, lo, dens, V, Vth, charge, ppc, seed, cutoff, cap, ic, de, v, w, q);
    if (dim == 2)
        return loadMaxwellianF<2>(n, lo, dens, V, Vth, charge, ppc, seed, cutoff, cap, ic, de, v, w, q);
    if (dim == 3)
        return loadMaxwellianF<3>(n, lo, dens, V, Vth, charge, ppc, seed, cutoff, cap, ic, de, v, w, q);
    return -1;
}

long ref_partition(int dim, int const* mapLo, int const* mapHi, long np, int const* ic0,
                   int const* ic1, int const* predLo, int const* predHi, long* order)
{
    if (dim == 1)
        return partitionF<1>(mapLo, mapHi, np, ic0, ic1, predLo, predHi, order);
    if (dim == 2)
        return partitionF<2>(mapLo, mapHi, np, ic0, ic1, predLo, predHi, order);
    if (dim == 3)
        return partitionF<3>(mapLo, mapHi, np, ic0, ic1, predLo, predHi, order);
    return -1;
}

// the reference's Splitter<dim, order, nbr> applied to one particle already on the fine grid
// (the input of PatternDispatcher::operator()): its nbr refined particles, in dispatch order
long ref_split_one(int dim, int order, int nbr, int const* ic, double const* de, double const* v,
                   double w, double q, int* oic, double* ode, double* ov, double* ow, double* oq)
{
    long r = -1;
    auto run = [&](auto dimc, auto splitter) {
        constexpr std::size_t D = decltype(dimc)::value;
        constexpr auto N        = decltype(splitter)::nbRefinedPart;
        Particle<D> p{};
        p.weight = w;
        p.charge = q;
        for (std::size_t d = 0; d < D; ++d)
        {
            p.iCell[d] = ic[d];
            p.delta[d] = de[d];
        }
        for (std::size_t d = 0; d < 3; ++d)
            p.v[d] = v[d];
        std::array<Particle<D>, N> out{};
        splitter(p, out);
        for (std::size_t k = 0; k < N; ++k)
        {
            for (std::size_t d = 0; d < D; ++d)
            {
                oic[k * D + d] = out[k].iCell[d];
                ode[k * D + d] = out[k].delta[d];
            }
            for (std::size_t d = 0; d < 3; ++d)
                ov[k * 3 + d] = out[k].v[d];
            ow[k] = out[k].weight;
            oq[k] = out[k].charge;
        }
        r = static_cast<long>(N);
    };
    using PHARE::amr::Splitter;
    using PHARE::core::DimConst;
    using PHARE::core::InterpConst;
    using PHARE::core::RefinedParticlesConst;
    using D1 = std::integral_constant<std::size_t, 1>;
    using D2 = std::integral_constant<std::size_t, 2>;
    using D3 = std::integral_constant<std::size_t, 3>;
    switch (dim * 100 + order * 10 + nbr)
    {
        case 112: run(D1{}, Splitter<DimConst<1>, InterpConst<1>, RefinedParticlesConst<2>>{}); break;
        case 113: run(D1{}, Splitter<DimConst<1>, InterpConst<1>, RefinedParticlesConst<3>>{}); break;
        case 122: run(D1{}, Splitter<DimConst<1>, InterpConst<2>, RefinedParticlesConst<2>>{}); break;
        case 123: run(D1{}, Splitter<DimConst<1>, InterpConst<2>, RefinedParticlesConst<3>>{}); break;
        case 124: run(D1{}, Splitter<DimConst<1>, InterpConst<2>, RefinedParticlesConst<4>>{}); break;
        case 132: run(D1{}, Splitter<DimConst<1>, InterpConst<3>, RefinedParticlesConst<2>>{}); break;
        case 133: run(D1{}, Splitter<DimConst<1>, InterpConst<3>, RefinedParticlesConst<3>>{}); break;
        case 134: run(D1{}, Splitter<DimConst<1>, InterpConst<3>, RefinedParticlesConst<4>>{}); break;
        case 135: run(D1{}, Splitter<DimConst<1>, InterpConst<3>, RefinedParticlesConst<5>>{}); break;
        case 214: run(D2{}, Splitter<DimConst<2>, InterpConst<1>, RefinedParticlesConst<4>>{}); break;
        case 224: run(D2{}, Splitter<DimConst<2>, InterpConst<2>, RefinedParticlesConst<4>>{}); break;
        case 234: run(D2{}, Splitter<DimConst<2>, InterpConst<3>, RefinedParticlesConst<4>>{}); break;
        case 316: run(D3{}, Splitter<DimConst<3>, InterpConst<1>, RefinedParticlesConst<6>>{}); break;
        case 326: run(D3{}, Splitter<DimConst<3>, InterpConst<2>, RefinedParticlesConst<6>>{}); break;
        case 336: run(D3{}, Splitter<DimConst<3>, InterpConst<3>, RefinedParticlesConst<6>>{}); break;
        default: break;
    }
    return r;
}

// the reference's Splitter::maxCellDistanceFromSplit of (dim, order) (splitter.hpp:116-120)
int ref_split_max_cell_distance(int order)
{
    using PHARE::amr::ASplitter;
    using PHARE::core::DimConst;
    using PHARE::core::InterpConst;
    using PHARE::core::RefinedParticlesConst;
    switch (order)
    {
        case 1: return ASplitter<DimConst<1>, InterpConst<1>, RefinedParticlesConst<2>>::maxCellDistanceFromSplit();
        case 2: return ASplitter<DimConst<1>, InterpConst<2>, RefinedParticlesConst<2>>::maxCellDistanceFromSplit();
        case 3: return ASplitter<DimConst<1>, InterpConst<3>, RefinedParticlesConst<2>>::maxCellDistanceFromSplit();
    }
    return -1;
}

} // extern "C"
