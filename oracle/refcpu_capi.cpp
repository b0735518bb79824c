// refcpu_capi.cpp -- CPU ORACLE (test infrastructure only): plain C entry points over the
// restatement in refcpu.hpp / refcpu_solver.hpp, loaded with ctypes by tests/ and by the
// cpu_baseline leg of bench.py. The product never loads this library.
//
// Field ids and particle-array kinds use the same numbering as include/phx.h so that tests
// can address both sides with one id; they are redefined here on purpose (no shared header).
#include "refcpu_amr.hpp"
#include "refcpu_solver.hpp"

#include <memory>

namespace
{
using namespace refcpu;

enum FieldId
{
    F_E = 0,
    F_B = 3,
    F_EPRED = 6,
    F_BPRED = 9,
    F_EAVG = 12,
    F_BAVG = 15,
    F_J = 18,
    F_VE = 21,
    F_PE = 24,
    F_RHO = 25,
    F_MASSD = 26,
    F_VI = 27,
    F_POP0 = 32, // + 8*pop + {0:n, 1:rho, 2..4:F}
};

struct Base
{
    virtual ~Base()                                                                      = default;
    virtual Field* field(int fid)                                                        = 0;
    virtual long count(int pop, int kind)                                                = 0;
    virtual int setParticles(int pop, int kind, long n, int const*, double const*, double const*,
                             double const*, double const*)                               = 0;
    virtual int getParticles(int pop, int kind, int*, double*, double*, double*, double*) = 0;
    virtual int loadMaxwellian(int pop, double const* n, std::array<double const*, 3> V,
                               std::array<double const*, 3> Vth, double charge, unsigned ppc,
                               unsigned long long seed, double cutoff)                   = 0;
    virtual int initialize()                                                             = 0;
    virtual int updatePopulations(int emset, double dt, int mode)                        = 0;
    virtual int updateIons()                                                             = 0;
    virtual int fillMomentBorders()                                                      = 0;
    virtual int fillIonBorders()                                                         = 0;
    virtual int moveIons(double dt, int mode)                                            = 0;
    virtual int advance(double dt, int nsteps)                                           = 0;
    virtual int faraday(int bset, int eset, int bnewset, double dt)                      = 0;
    virtual int ampere(int bset)                                                         = 0;
    virtual int electrons()                                                              = 0;
    virtual int ohm(int bset, int eset)                                                  = 0;
    virtual int average()                                                                = 0;
    virtual int fillGhosts(int fid)                                                      = 0;
    virtual int gather(int emset, int pop, int kind, double* E, double* B)               = 0;
    virtual int pushError(double* out)                                                   = 0;
    virtual int push(int emset, int pop, int kind, double dt, int nf, int const* flo, int const* fhi,
                     int ns, int const* slo, int const* shi, long* out)                  = 0;
    virtual int deposit(int pop, int kind, double coef)                                  = 0;
    virtual int diagnostics(double* out)                                                 = 0;
};

template<int D, int O>
struct Impl : Base
{
    Solver<D, O> s;

    Impl(Layout<D> const& L, std::vector<double> const& masses)
        : s{L, masses}
    {
    }

    EM& emset(int i) { return i == 0 ? s.em : (i == 1 ? s.emPred : s.emAvg); }

    Field* field(int fid) override
    {
        if (fid >= F_POP0)
        {
            int p = (fid - F_POP0) / 8, k = (fid - F_POP0) % 8;
            if (p >= static_cast<int>(s.ions.pops.size()) || k > 4)
                return nullptr;
            auto& pop = s.ions.pops[p];
            return k == 0 ? &pop.n : (k == 1 ? &pop.rho : &pop.F[k - 2]);
        }
        if (fid < 18)
        {
            auto& e = emset(fid / 6);
            int r   = fid % 6;
            return r < 3 ? &e.E[r] : &e.B[r - 3];
        }
        if (fid < 21)
            return &s.J[fid - 18];
        if (fid < 24)
            return &s.Ve[fid - 21];
        if (fid == F_PE)
            return &s.Pe;
        if (fid == F_RHO)
            return &s.ions.rho;
        if (fid == F_MASSD)
            return &s.ions.massDensity;
        if (fid >= F_VI && fid < F_VI + 3)
            return &s.ions.Vi[fid - F_VI];
        return nullptr;
    }

    ParticleArray<D>* arr(int pop, int kind)
    {
        if (pop < 0 || pop >= static_cast<int>(s.ions.pops.size()))
            return nullptr;
        auto& P = s.ions.pops[pop];
        switch (kind)
        {
            case 0: return &P.domain;
            case 1: return &P.patchGhost;
            case 2: return &P.levelGhost;
            case 3: return &P.levelGhostOld;
            case 4: return &P.levelGhostNew;
            default: return nullptr;
        }
    }

    long count(int pop, int kind) override
    {
        auto a = arr(pop, kind);
        return a ? static_cast<long>(a->size()) : -1;
    }

    int setParticles(int pop, int kind, long n, int const* ic, double const* de, double const* v,
                     double const* w, double const* q) override
    {
        auto a = arr(pop, kind);
        if (!a)
            return 1;
        a->clear();
        for (long i = 0; i < n; ++i)
        {
            Particle<D> p;
            p.weight = w[i];
            p.charge = q[i];
            for (int d = 0; d < D; ++d)
            {
                p.iCell[d] = ic[i * D + d];
                p.delta[d] = de[i * D + d];
            }
            for (int c = 0; c < 3; ++c)
                p.v[c] = v[i * 3 + c];
            a->push_back(p);
        }
        return 0;
    }

    int getParticles(int pop, int kind, int* ic, double* de, double* v, double* w,
                     double* q) override
    {
        auto a = arr(pop, kind);
        if (!a)
            return 1;
        for (std::size_t i = 0; i < a->size(); ++i)
        {
            auto const& p = a->p[i];
            w[i]          = p.weight;
            q[i]          = p.charge;
            for (int d = 0; d < D; ++d)
            {
                ic[i * D + d] = p.iCell[d];
                de[i * D + d] = p.delta[d];
            }
            for (int c = 0; c < 3; ++c)
                v[i * 3 + c] = p.v[c];
        }
        return 0;
    }

    int loadMaxwellian(int pop, double const* n, std::array<double const*, 3> V,
                       std::array<double const*, 3> Vth, double charge, unsigned ppc,
                       unsigned long long seed, double cutoff) override
    {
        auto a = arr(pop, 0);
        if (!a)
            return 1;
        refcpu::loadMaxwellian<D>(s.L, n, V, Vth, charge, ppc, seed, cutoff, *a);
        return 0;
    }

    int initialize() override
    {
        s.initialize();
        return 0;
    }
    int updatePopulations(int e, double dt, int mode) override
    {
        s.lastError = s.updater.updatePopulations(s.ions, emset(e), s.boxing(), dt,
                                                  mode == 1 ? UpdaterMode::domain_only
                                                            : UpdaterMode::all);
        return s.lastError.set ? 2 : 0;
    }
    int updateIons() override
    {
        s.updater.updateIons(s.ions);
        return 0;
    }
    int fillMomentBorders() override
    {
        s.fillMomentBorders();
        return 0;
    }
    int fillIonBorders() override
    {
        s.fillIonBorders();
        return 0;
    }
    int moveIons(double dt, int mode) override
    {
        return s.moveIons(dt, mode == 1 ? UpdaterMode::domain_only : UpdaterMode::all) ? 0 : 2;
    }
    int advance(double dt, int nsteps) override
    {
        for (int i = 0; i < nsteps; ++i)
            if (!s.advance(dt))
                return 2;
        return 0;
    }
    int faraday(int b, int e, int bn, double dt) override
    {
        refcpu::faraday<D>(s.L, emset(b).B, emset(e).E, emset(bn).B, dt);
        return 0;
    }
    int ampere(int b) override
    {
        refcpu::ampere<D>(s.L, emset(b).B, s.J);
        return 0;
    }
    int electrons() override
    {
        s.updateElectrons();
        return 0;
    }
    int ohm(int b, int e) override
    {
        s.runOhm(emset(b).B, emset(e).E);
        return 0;
    }
    int average() override
    {
        s.averageStage();
        return 0;
    }
    int fillGhosts(int fid) override
    {
        auto f = field(fid);
        if (!f)
            return 1;
        s.fills.fillGhosts(*f);
        return 0;
    }
    int gather(int e, int pop, int kind, double* E, double* B) override
    {
        auto a = arr(pop, kind);
        if (!a)
            return 1;
        for (std::size_t i = 0; i < a->size(); ++i)
        {
            std::array<double, 3> pE, pB;
            refcpu::gather<D, O>(s.L, emset(e), a->p[i], pE, pB);
            for (int c = 0; c < 3; ++c)
            {
                E[i * 3 + c] = pE[c];
                B[i * 3 + c] = pB[c];
            }
        }
        return 0;
    }
    // BorisPusher::move (boris.hpp:93-138) in place, with selectors given as box lists (an empty
    // list is the noop selector); out = {end of the second selection, end of the first}
    int push(int e, int pop, int kind, double dt, int nf, int const* flo, int const* fhi, int ns,
             int const* slo, int const* shi, long* out) override
    {
        auto a = arr(pop, kind);
        if (!a)
            return 1;
        auto sel = [](int nb, int const* lo, int const* hi) {
            Selector<D> S;
            if (nb <= 0)
                return S;
            std::vector<Box<D>> bx(static_cast<std::size_t>(nb));
            for (int b = 0; b < nb; ++b)
                for (int d = 0; d < D; ++d)
                {
                    bx[b].lo[d] = lo[3 * b + d];
                    bx[b].hi[d] = hi[3 * b + d];
                }
            S.pred = [bx](std::array<int, D> const& c) {
                for (auto const& x : bx)
                    if (x.contains(c))
                        return true;
                return false;
            };
            return S;
        };
        BorisPusher<D, O> P;
        P.setMeshAndTimeStep(s.L.dx, dt);
        PushError err;
        std::size_t e1 = 0;
        std::size_t const e2 = P.move(*a, 0, a->size(), emset(e), s.ions.pops[pop].mass, s.L,
                                      sel(nf, flo, fhi), sel(ns, slo, shi), err, &e1);
        s.lastError = err;
        out[0] = static_cast<long>(e2);
        out[1] = static_cast<long>(e1);
        return err.set ? 2 : 0;
    }
    // Interpolator::operator() deposit (interpolator.hpp:468-510) of a whole array, added to the
    // population's moments
    int deposit(int pop, int kind, double coef) override
    {
        auto a = arr(pop, kind);
        if (!a)
            return 1;
        auto& P = s.ions.pops[pop];
        refcpu::deposit<D, O>(s.L, *a, 0, a->size(), P.n, P.rho, P.F, coef);
        return 0;
    }
    int pushError(double* out) override
    {
        out[0] = s.lastError.set ? 1. : 0.;
        out[1] = s.lastError.delta;
        out[2] = s.lastError.vel;
        out[3] = s.lastError.stage;
        return 0;
    }
    // [kinetic, magnetic, total charge, particle count]; one period of nodes per direction
    int diagnostics(double* out) override
    {
        double vol = 1.;
        for (int d = 0; d < D; ++d)
            vol *= s.L.dx[d];
        double kin = 0., charge = 0., count = 0.;
        for (auto const& pop : s.ions.pops)
        {
            double k = 0.;
            for (auto const& p : pop.domain.p)
            {
                k += p.weight * (p.v[0] * p.v[0] + p.v[1] * p.v[1] + p.v[2] * p.v[2]);
                charge += p.charge * p.weight;
            }
            kin += 0.5 * pop.mass * k;
            count += static_cast<double>(pop.domain.size());
        }
        double mag = 0.;
        for (int c = 0; c < 3; ++c)
        {
            auto const& f = s.em.B[c];
            std::array<int, D> lo, hi;
            for (int d = 0; d < D; ++d)
            {
                lo[d] = s.L.physStart();
                hi[d] = s.L.physStart() + s.L.ncells[d] - 1;
            }
            evalOnRange<D>(lo, hi, [&](std::array<int, D> const& i) {
                double b = f.at(i);
                mag += b * b;
            });
        }
        out[0] = kin * vol;
        out[1] = 0.5 * mag * vol;
        out[2] = charge * vol;
        out[3] = count;
        return 0;
    }
};

template<int D>
Base* make2(int order, Layout<D> const& L, std::vector<double> const& m)
{
    switch (order)
    {
        case 1: return new Impl<D, 1>(L, m);
        case 2: return new Impl<D, 2>(L, m);
        case 3: return new Impl<D, 3>(L, m);
    }
    return nullptr;
}

template<int D>
Base* make1(int order, int const* n, double const* dx, int const* lo, int level,
            std::vector<double> const& m)
{
    std::array<int, D> nn, ll;
    std::array<double, D> hh;
    for (int d = 0; d < D; ++d)
    {
        nn[d] = n[d];
        ll[d] = lo[d];
        hh[d] = dx[d];
    }
    return make2<D>(order, Layout<D>{nn, hh, ll, order, level}, m);
}

Base* B_(void* h) { return static_cast<Base*>(h); }

} // namespace

extern "C" {

void* ora_create(int dim, int order, int const* ncells, double const* dx, int const* lower, int level,
                 int npop, double const* masses, double Te, double eta, double nu, int hyper_mode)
{
    if (order < 1 || order > 3 || npop < 1)
        return nullptr;
    std::vector<double> m(masses, masses + npop);
    Base* b = nullptr;
    if (dim == 1)
        b = make1<1>(order, ncells, dx, lower, level, m);
    else if (dim == 2)
        b = make1<2>(order, ncells, dx, lower, level, m);
    else if (dim == 3)
        b = make1<3>(order, ncells, dx, lower, level, m);
    if (!b)
        return nullptr;
    // solver parameters
    auto set = [&](auto* impl) {
        impl->s.Te    = Te;
        impl->s.eta   = eta;
        impl->s.nu    = nu;
        impl->s.hyper = hyper_mode == 0 ? HyperMode::constant : HyperMode::spatial;
    };
#define PHX_ORA_SET(DD, OO)                                                                        \
    if (dim == DD && order == OO)                                                                  \
        set(static_cast<Impl<DD, OO>*>(b));
    PHX_ORA_SET(1, 1)
    PHX_ORA_SET(1, 2)
    PHX_ORA_SET(1, 3)
    PHX_ORA_SET(2, 1)
    PHX_ORA_SET(2, 2)
    PHX_ORA_SET(2, 3)
    PHX_ORA_SET(3, 1)
    PHX_ORA_SET(3, 2)
    PHX_ORA_SET(3, 3)
#undef PHX_ORA_SET
    return b;
}

void ora_destroy(void* h) { delete B_(h); }

int ora_field_shape(void* h, int fid, int* shape)
{
    auto f = B_(h)->field(fid);
    if (!f)
        return 1;
    for (int i = 0; i < 3; ++i)
        shape[i] = f->shape[i];
    return 0;
}
int ora_set_field(void* h, int fid, double const* data)
{
    auto f = B_(h)->field(fid);
    if (!f)
        return 1;
    std::copy(data, data + f->v.size(), f->v.begin());
    return 0;
}
int ora_get_field(void* h, int fid, double* data)
{
    auto f = B_(h)->field(fid);
    if (!f)
        return 1;
    std::copy(f->v.begin(), f->v.end(), data);
    return 0;
}
long ora_num_particles(void* h, int pop, int kind) { return B_(h)->count(pop, kind); }
int ora_set_particles(void* h, int pop, int kind, long n, int const* ic, double const* de,
                      double const* v, double const* w, double const* q)
{
    return B_(h)->setParticles(pop, kind, n, ic, de, v, w, q);
}
int ora_get_particles(void* h, int pop, int kind, int* ic, double* de, double* v, double* w, double* q)
{
    return B_(h)->getParticles(pop, kind, ic, de, v, w, q);
}
int ora_load_maxwellian(void* h, int pop, double const* n, double const* vx, double const* vy,
                        double const* vz, double const* tx, double const* ty, double const* tz,
                        double charge, unsigned ppc, unsigned long long seed, double cutoff)
{
    return B_(h)->loadMaxwellian(pop, n, {vx, vy, vz}, {tx, ty, tz}, charge, ppc, seed, cutoff);
}
int ora_initialize(void* h) { return B_(h)->initialize(); }
int ora_update_populations(void* h, int emset, double dt, int mode)
{
    return B_(h)->updatePopulations(emset, dt, mode);
}
int ora_update_ions(void* h) { return B_(h)->updateIons(); }
int ora_fill_moment_borders(void* h) { return B_(h)->fillMomentBorders(); }
int ora_fill_ion_borders(void* h) { return B_(h)->fillIonBorders(); }
int ora_move_ions(void* h, double dt, int mode) { return B_(h)->moveIons(dt, mode); }
int ora_advance(void* h, double dt, int nsteps) { return B_(h)->advance(dt, nsteps); }
int ora_faraday(void* h, int bset, int eset, int bnewset, double dt)
{
    return B_(h)->faraday(bset, eset, bnewset, dt);
}
int ora_ampere(void* h, int bset) { return B_(h)->ampere(bset); }
int ora_electrons(void* h) { return B_(h)->electrons(); }
int ora_ohm(void* h, int bset, int eset) { return B_(h)->ohm(bset, eset); }
int ora_average(void* h) { return B_(h)->average(); }
int ora_fill_ghosts(void* h, int fid) { return B_(h)->fillGhosts(fid); }
int ora_gather(void* h, int emset, int pop, int kind, double* E, double* B)
{
    return B_(h)->gather(emset, pop, kind, E, B);
}
int ora_push_error(void* h, double* out) { return B_(h)->pushError(out); }
int ora_push(void* h, int emset, int pop, int kind, double dt, int nf, int const* flo, int const* fhi,
             int ns, int const* slo, int const* shi, long* out)
{
    return B_(h)->push(emset, pop, kind, dt, nf, flo, fhi, ns, slo, shi, out);
}
int ora_deposit(void* h, int pop, int kind, double coef) { return B_(h)->deposit(pop, kind, coef); }
int ora_diagnostics(void* h, double* out) { return B_(h)->diagnostics(out); }

// B-spline weights of one direction, as the Interpolator computes them for a local cell index
void ora_weights(int order, int dual, int icell_local, double delta, int* start, double* w)
{
    Cent c    = dual ? refcpu::dual : refcpu::primal;
    auto icl  = static_cast<std::uint32_t>(icell_local);
    int shift = order == 1 ? startLeftShift<1>(c, delta)
                           : (order == 2 ? startLeftShift<2>(c, delta) : startLeftShift<3>(c, delta));
    *start    = static_cast<int>(icl - static_cast<std::uint32_t>(shift));
    double np = icl + delta;
    if (dual)
        np -= .5;
    if (order == 1)
        computeWeight<1>(np, *start, w);
    else if (order == 2)
        computeWeight<2>(np, *start, w);
    else
        computeWeight<3>(np, *start, w);
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// Stand-alone operators with the same signatures as oracle/_ref's ref_* entry points, so the
// restatement can be compared against the reference headers call for call.
// ---------------------------------------------------------------------------------------------
namespace
{
template<int D>
Layout<D> mkLayout(int order, int const* n, double const* dx, int const* lo)
{
    std::array<int, D> nn, ll;
    std::array<double, D> hh;
    for (int d = 0; d < D; ++d)
    {
        nn[d] = n[d];
        ll[d] = lo[d];
        hh[d] = dx[d];
    }
    return Layout<D>{nn, hh, ll, order};
}

// quantity ids: 0..2 B, 3..5 E, 6..8 J, 9 moments
Qty oqty(int q) { return q == 9 ? qMom : Qty(q); }

template<int D>
Field wrap(Layout<D> const& L, int q, double const* data)
{
    Field f{L, oqty(q)};
    std::copy(data, data + f.v.size(), f.v.begin());
    return f;
}

template<int D, int O>
int s_deposit(int const* n, double const* dx, int const* lo, long np, int const* ic,
              double const* de, double const* v, double const* w, double const* q, double coef,
              double* dn, double* drho, double* const* dF)
{
    auto L   = mkLayout<D>(O, n, dx, lo);
    Field fn = wrap<D>(L, 9, dn), fr = wrap<D>(L, 9, drho);
    VecField F{L, qMom};
    for (int c = 0; c < 3; ++c)
        F[c] = wrap<D>(L, 9, dF[c]);
    for (long i = 0; i < np; ++i)
    {
        Particle<D> p;
        p.weight = w[i];
        p.charge = q[i];
        for (int d = 0; d < D; ++d)
        {
            p.iCell[d] = ic[i * D + d];
            p.delta[d] = de[i * D + d];
        }
        for (int c = 0; c < 3; ++c)
            p.v[c] = v[i * 3 + c];
        depositOne<D, O>(L, p, fn, fr, F, coef);
    }
    std::copy(fn.v.begin(), fn.v.end(), dn);
    std::copy(fr.v.begin(), fr.v.end(), drho);
    for (int c = 0; c < 3; ++c)
        std::copy(F[c].v.begin(), F[c].v.end(), dF[c]);
    return 0;
}

template<int D, int O>
int s_laplacian(int const* n, double const* dx, int const* lo, int q, double* src, double* out)
{
    auto L  = mkLayout<D>(O, n, dx, lo);
    Field f = wrap<D>(L, q, src), o = wrap<D>(L, q, out);
    evalOnBox<D>(L, oqty(q), [&](std::array<int, D> const& i) { o.at(i) = laplacian<D>(L, f, i); });
    std::copy(o.v.begin(), o.v.end(), out);
    return 0;
}

template<int D, int O>
int s_project(int const* n, double const* dx, int const* lo, int which, int srcq, int dstq,
              double* src, double* out)
{
    using P = Proj<D>;
    auto L  = mkLayout<D>(O, n, dx, lo);
    Field f = wrap<D>(L, srcq, src), o = wrap<D>(L, dstq, out);
    std::vector<WP> t;
    switch (which)
    {
        case 0: t = P::momentsToE(0); break;
        case 1: t = P::momentsToE(1); break;
        case 2: t = P::momentsToE(2); break;
        case 3: t = P::BxToEx(); break;
        case 4: t = P::ByToEx(); break;
        case 5: t = P::BzToEx(); break;
        case 6: t = P::BxToEy(); break;
        case 7: t = P::ByToEy(); break;
        case 8: t = P::BzToEy(); break;
        case 9: t = P::BxToEz(); break;
        case 10: t = P::ByToEz(); break;
        case 11: t = P::BzToEz(); break;
        case 12: t = P::JToMoments(0); break;
        case 13: t = P::JToMoments(1); break;
        case 14: t = P::JToMoments(2); break;
    }
    evalOnBox<D>(L, oqty(dstq), [&](std::array<int, D> const& i) { o.at(i) = project<D>(t, f, i); });
    std::copy(o.v.begin(), o.v.end(), out);
    return 0;
}

template<int D, int O>
int s_faraday(int const* n, double const* dx, int const* lo, double* const* B, double* const* E,
              double* const* Bnew, double dt)
{
    auto L = mkLayout<D>(O, n, dx, lo);
    VecField b{L, qBx}, e{L, qEx}, bn{L, qBx};
    for (int c = 0; c < 3; ++c)
    {
        b[c]  = wrap<D>(L, c, B[c]);
        e[c]  = wrap<D>(L, 3 + c, E[c]);
        bn[c] = wrap<D>(L, c, Bnew[c]);
    }
    refcpu::faraday<D>(L, b, e, bn, dt);
    for (int c = 0; c < 3; ++c)
        std::copy(bn[c].v.begin(), bn[c].v.end(), Bnew[c]);
    return 0;
}

template<int D, int O>
int s_ampere(int const* n, double const* dx, int const* lo, double* const* B, double* const* J)
{
    auto L = mkLayout<D>(O, n, dx, lo);
    VecField b{L, qBx}, j{L, qJx};
    for (int c = 0; c < 3; ++c)
    {
        b[c] = wrap<D>(L, c, B[c]);
        j[c] = wrap<D>(L, 6 + c, J[c]);
    }
    refcpu::ampere<D>(L, b, j);
    for (int c = 0; c < 3; ++c)
        std::copy(j[c].v.begin(), j[c].v.end(), J[c]);
    return 0;
}

template<int D>
long s_partition(int const* mapLo, int const* mapHi, long np, int const* ic0, int const* ic1,
                 int const* predLo, int const* predHi, long* order)
{
    Box<D> mb, pb;
    for (int d = 0; d < D; ++d)
    {
        mb.lo[d] = mapLo[d];
        mb.hi[d] = mapHi[d];
        pb.lo[d] = predLo[d];
        pb.hi[d] = predHi[d];
    }
    ParticleArray<D> arr{mb};
    for (long i = 0; i < np; ++i)
    {
        Particle<D> p;
        p.weight = static_cast<double>(i);
        for (int d = 0; d < D; ++d)
            p.iCell[d] = ic0[i * D + d];
        arr.push_back(p);
    }
    for (long i = 0; i < np; ++i)
    {
        std::array<int, D> nc{};
        for (int d = 0; d < D; ++d)
            nc[d] = ic1[i * D + d];
        if (nc != arr.p[i].iCell)
            arr.change_icell(nc, static_cast<std::size_t>(i));
    }
    auto piv = arr.partition(0, arr.size(), [&](auto const& c) { return pb.contains(c); });
    for (long i = 0; i < np; ++i)
        order[i] = static_cast<long>(arr.p[i].weight);
    return static_cast<long>(piv);
}

#define PHX_ORA_DISPATCH(dim, order, fn, ...)                                                      \
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

int ora_layout_shape(int dim, int order, int const* n, int q, int* shape3)
{
    std::array<int, 3> s{1, 1, 1};
    double dx[3] = {1, 1, 1};
    int lo[3]    = {0, 0, 0};
    if (dim == 1)
        s = mkLayout<1>(order, n, dx, lo).shape(oqty(q));
    else if (dim == 2)
        s = mkLayout<2>(order, n, dx, lo).shape(oqty(q));
    else
        s = mkLayout<3>(order, n, dx, lo).shape(oqty(q));
    for (int i = 0; i < 3; ++i)
        shape3[i] = s[i];
    return 0;
}

int ora_s_deposit(int dim, int order, int const* n, double const* dx, int const* lo, long np,
                  int const* ic, double const* de, double const* v, double const* w, double const* q,
                  double coef, double* dn, double* drho, double* const* dF)
{
    return PHX_ORA_DISPATCH(dim, order, s_deposit, n, dx, lo, np, ic, de, v, w, q, coef, dn, drho, dF);
}
int ora_s_laplacian(int dim, int order, int const* n, double const* dx, int const* lo, int q,
                    double* src, double* out)
{
    return PHX_ORA_DISPATCH(dim, order, s_laplacian, n, dx, lo, q, src, out);
}
int ora_s_project(int dim, int order, int const* n, double const* dx, int const* lo, int which,
                  int srcq, int dstq, double* src, double* out)
{
    return PHX_ORA_DISPATCH(dim, order, s_project, n, dx, lo, which, srcq, dstq, src, out);
}
int ora_s_faraday(int dim, int order, int const* n, double const* dx, int const* lo,
                  double* const* B, double* const* E, double* const* Bnew, double dt)
{
    return PHX_ORA_DISPATCH(dim, order, s_faraday, n, dx, lo, B, E, Bnew, dt);
}
int ora_s_ampere(int dim, int order, int const* n, double const* dx, int const* lo, double* const* B,
                 double* const* J)
{
    return PHX_ORA_DISPATCH(dim, order, s_ampere, n, dx, lo, B, J);
}
long ora_partition(int dim, int const* mapLo, int const* mapHi, long np, int const* ic0,
                   int const* ic1, int const* predLo, int const* predHi, long* order)
{
    if (dim == 1)
        return s_partition<1>(mapLo, mapHi, np, ic0, ic1, predLo, predHi, order);
    if (dim == 2)
        return s_partition<2>(mapLo, mapHi, np, ic0, ic1, predLo, predHi, order);
    return s_partition<3>(mapLo, mapHi, np, ic0, ic1, predLo, predHi, order);
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// Single-particle Boris trajectory in uniform fields, as the reference's pusher test drives it
// (tests/core/numerics/pusher/test_pusher.cpp:180-243: constant-field interpolator, positions
// recorded as (iCell + delta) * dx before every move).
// ---------------------------------------------------------------------------------------------
namespace
{
template<int D>
void s_trajectory(long nsteps, double dt, double const* dx, double mass, double charge,
                  double const* E, double const* B, int const* ic0, double const* de0, double const* v0,
                  double* out)
{
    BorisPusher<D, 1> pusher;
    std::array<double, D> ms;
    for (int d = 0; d < D; ++d)
        ms[d] = dx[d];
    pusher.setMeshAndTimeStep(ms, dt);
    Particle<D> in;
    in.weight = 1.;
    in.charge = charge;
    for (int d = 0; d < D; ++d)
    {
        in.iCell[d] = ic0[d];
        in.delta[d] = de0[d];
    }
    for (int c = 0; c < 3; ++c)
        in.v[c] = v0[c];
    Particle<D> outp = in;
    std::array<double, 3> pE{E[0], E[1], E[2]}, pB{B[0], B[1], B[2]};
    PushError err;
    double const dto2m = 0.5 * dt / mass;
    for (long s = 0; s < nsteps; ++s)
    {
        for (int d = 0; d < D; ++d)
            out[s * D + d] = (outp.iCell[d] + outp.delta[d]) * dx[d];
        std::array<int, D> nc{};
        outp.charge = in.charge;
        outp.weight = in.weight;
        outp.v      = in.v;
        pusher.advancePosition(in, outp, nc, err, 1);
        outp.iCell = nc;
        BorisPusher<D, 1>::accelerate(outp, pE, pB, dto2m);
        pusher.advancePosition(outp, outp, nc, err, 2);
        outp.iCell = nc;
        in         = outp;
    }
}
} // namespace

extern "C" int ora_boris_trajectory(int dim, long nsteps, double dt, double const* dx, double mass,
                                    double charge, double const* E, double const* B, int const* ic0,
                                    double const* de0, double const* v0, double* out)
{
    if (dim == 1)
        s_trajectory<1>(nsteps, dt, dx, mass, charge, E, B, ic0, de0, v0, out);
    else if (dim == 2)
        s_trajectory<2>(nsteps, dt, dx, mass, charge, E, B, ic0, de0, v0, out);
    else
        s_trajectory<3>(nsteps, dt, dx, mass, charge, E, B, ic0, de0, v0, out);
    return 0;
}

// the coarse-to-fine particle split (refcpu_amr.hpp): coarse particles (SoA, n) -> refined
// particles kept in the destination boxes, written to the out arrays (capacity cap); returns the
// count of refined particles, or -1 when the out arrays are too small or the splitter unknown
extern "C" long ora_split(int dim, int order, int nbr, long n, int const* ic, double const* de,
                          double const* v, double const* w, double const* q, int nb, int const* lo,
                          int const* hi, long cap, int* oic, double* ode, double* ov, double* ow,
                          double* oq)
{
    if (refcpu::splitTable(dim, order, nbr).w.empty())
        return -1;
    std::vector<refcpu::SplitPart> in(static_cast<size_t>(n)), out;
    for (long i = 0; i < n; ++i)
    {
        auto& p = in[static_cast<size_t>(i)];
        for (int d = 0; d < 3; ++d)
        {
            p.ic[d] = d < dim ? ic[i * dim + d] : 0;
            p.de[d] = d < dim ? de[i * dim + d] : 0.;
            p.v[d]  = v[i * 3 + d];
        }
        p.w = w[i];
        p.q = q[i];
    }
    refcpu::splitParticles(dim, order, nbr, in, nb, lo, hi, out);
    if (static_cast<long>(out.size()) > cap)
        return -1;
    for (size_t k = 0; k < out.size(); ++k)
    {
        for (int d = 0; d < dim; ++d)
        {
            oic[k * dim + d] = out[k].ic[d];
            ode[k * dim + d] = out[k].de[d];
        }
        for (int d = 0; d < 3; ++d)
            ov[k * 3 + d] = out[k].v[d];
        ow[k] = out[k].w;
        oq[k] = out[k].q;
    }
    return static_cast<long>(out.size());
}

// the coarse-to-fine field refinement (refcpu_amr.hpp): fine and coarse arrays in local indices
// with their shapes (x, y, z; unused directions 1) and AMR origins (AMR index of local index 0);
// lo/hi: the fine AMR box of the quantity's nodes; primal: its centering per direction
extern "C" int ora_refine_field(int dim, int op, int const* primal, double* fine, int const* fshape,
                                int const* forg, double const* coarse, int const* cshape,
                                int const* corg, int const* lo, int const* hi)
{
    refcpu::AmrField F{fine, {fshape[0], fshape[1], fshape[2]}, {forg[0], forg[1], forg[2]}};
    refcpu::AmrField C{const_cast<double*>(coarse), {cshape[0], cshape[1], cshape[2]},
                       {corg[0], corg[1], corg[2]}};
    refcpu::refineField(dim, op, primal, F, C, lo, hi);
    return 0;
}

// MagneticRefinePatchStrategy::postprocessRefine over the fine cell box lo..hi: the three B
// components (shapes sx, sy, sz; one AMR origin) updated in place
extern "C" int ora_magnetic_new_faces(int dim, double* bx, int const* sx, double* by, int const* sy,
                                      double* bz, int const* sz, int const* org, double const* dx,
                                      int const* lo, int const* hi)
{
    refcpu::AmrField X{bx, {sx[0], sx[1], sx[2]}, {org[0], org[1], org[2]}};
    refcpu::AmrField Y{by, {sy[0], sy[1], sy[2]}, {org[0], org[1], org[2]}};
    refcpu::AmrField Z{bz, {sz[0], sz[1], sz[2]}, {org[0], org[1], org[2]}};
    refcpu::magneticNewFaces(dim, X, Y, Z, dx, lo, hi);
    return 0;
}
