// refcpu_solver.hpp -- CPU ORACLE (test infrastructure only): SolverPPC on one periodic patch.
//
// Restates the predictor-corrector stage order of src/amr/solvers/solver_ppc.hpp:315-598 and
// the root-level initialisation of src/amr/level_initializer/hybrid_level_initializer.hpp:
// 26-188 for a level made of ONE patch that covers a periodic domain, with the same-level
// messenger fills of hybrid_hybrid_messenger_strategy.hpp:376-548 restated in PeriodicFills.
// Used only as the checker and as the timed CPU baseline; never by the product.
#pragma once

#include "refcpu.hpp"

namespace refcpu
{

template<int D, int O>
struct Solver
{
    Layout<D> L;
    double Te  = 0.1;
    double eta = 0., nu = 1e-4;
    HyperMode hyper = HyperMode::constant;

    EM em, emPred, emAvg;
    VecField J, Ve;
    Field Pe;
    Ions<D> ions;
    IonUpdater<D, O> updater;
    PeriodicFills<D> fills;
    PushError lastError;

    Solver(Layout<D> const& layout, std::vector<double> const& masses)
        : L{layout}
        , fills{layout}
    {
        for (auto* e : {&em, &emPred, &emAvg})
        {
            e->E = VecField{L, qEx};
            e->B = VecField{L, qBx};
        }
        J  = VecField{L, qJx};
        Ve = VecField{L, qMom};
        Pe = Field{L, qMom};
        ions.rho         = Field{L, qMom};
        ions.massDensity = Field{L, qMom};
        ions.Vi          = VecField{L, qMom};
        auto mapBox      = L.amrBox().grown(L.pg + 1);
        for (double m : masses)
        {
            Population<D> p;
            p.mass          = m;
            p.domain        = ParticleArray<D>{mapBox};
            p.patchGhost    = ParticleArray<D>{mapBox};
            p.levelGhost    = ParticleArray<D>{mapBox};
            p.levelGhostOld = ParticleArray<D>{mapBox};
            p.levelGhostNew = ParticleArray<D>{mapBox};
            p.n             = Field{L, qMom};
            p.rho           = Field{L, qMom};
            p.F             = VecField{L, qMom};
            ions.pops.push_back(std::move(p));
        }
        updater.tmp = ParticleArray<D>{mapBox};
    }

    // a single periodic patch: the non-level-ghost boxes are the domain plus the ghost layer
    // shared with its own periodic images (amr_utils.hpp:233-251)
    Boxing<D> boxing() const { return Boxing<D>{L, {L.amrBox(), L.amrBox().grown(L.pg)}}; }

    void fillMomentBorders()
    {
        for (auto& pop : ions.pops)
        {
            for (int c = 0; c < 3; ++c)
                fills.borderSum(pop.F[c]); // fillFluxBorders
            fills.borderSum(pop.n);        // fillDensityBorders
            fills.borderSum(pop.rho);
        }
    }
    void fillIonBorders()
    {
        fills.borderMax(ions.rho);
        fills.borderMax(ions.massDensity);
        for (int c = 0; c < 3; ++c)
            fills.borderMax(ions.Vi[c]);
    }

    void updateElectrons() { electrons<D>(L, Te, J, ions.Vi, ions.rho, Ve, Pe); }
    void runOhm(VecField const& B, VecField& E)
    {
        ohm<D>(L, eta, nu, hyper, ions.rho, Ve, Pe, B, J, E);
    }

    // hybrid_level_initializer.hpp:103-171 (root level)
    void initialize()
    {
        fills.fillGhosts(em.B);
        for (auto& pop : ions.pops)
        {
            pop.n.zero();
            pop.rho.zero();
            pop.F.zero();
            deposit<D, O>(L, pop.domain, 0, pop.domain.size(), pop.n, pop.rho, pop.F);
        }
        fillMomentBorders();
        updater.updateIons(ions);
        fillIonBorders();
        ampere<D>(L, em.B, J);
        fills.fillGhosts(J);
        updateElectrons();
        runOhm(em.B, em.E);
        fills.fillGhosts(em.E);
    }

    // solver_ppc.hpp:537-598
    bool moveIons(double dt, UpdaterMode mode)
    {
        lastError = updater.updatePopulations(ions, emAvg, boxing(), dt, mode);
        if (lastError.set)
            return false;
        fillMomentBorders();
        if (mode != UpdaterMode::domain_only)
            for (auto& pop : ions.pops)
                fills.ghostParticlesToDomain(pop);
        updater.updateIons(ions);
        fillIonBorders();
        return true;
    }

    // solver_ppc.hpp:483-510
    void averageStage()
    {
        for (int c = 0; c < 3; ++c)
        {
            average(em.B[c], emPred.B[c], emAvg.B[c]);
            average(em.E[c], emPred.E[c], emAvg.E[c]);
        }
        fills.fillGhosts(emAvg.E);
    }

    // solver_ppc.hpp:346-428 (predictor1 uses E, predictor2 uses Eavg)
    void predictor(double dt, VecField const& Euse)
    {
        faraday<D>(L, em.B, Euse, emPred.B, dt);
        fills.fillGhosts(emPred.B);
        ampere<D>(L, emPred.B, J);
        fills.fillGhosts(J);
        updateElectrons();
        runOhm(emPred.B, emPred.E);
    }

    // solver_ppc.hpp:433-479
    void corrector(double dt)
    {
        faraday<D>(L, em.B, emAvg.E, em.B, dt);
        fills.fillGhosts(em.B);
        ampere<D>(L, em.B, J);
        fills.fillGhosts(J);
        updateElectrons();
        runOhm(em.B, em.E);
        fills.fillGhosts(em.E);
    }

    // solver_ppc.hpp:315-341
    bool advance(double dt)
    {
        predictor(dt, em.E);
        averageStage();
        if (!moveIons(dt, UpdaterMode::domain_only))
            return false;
        predictor(dt, emAvg.E);
        averageStage();
        if (!moveIons(dt, UpdaterMode::all))
            return false;
        corrector(dt);
        return true;
    }
};

} // namespace refcpu
