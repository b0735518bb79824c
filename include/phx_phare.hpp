/* phx_phare.hpp -- header-only C++ adapter between PHARE's own types and the C ABI of phx.h.
 *
 * What a PHARE build includes to hand its patch data to the MI355X path and take it back:
 *  - particle arrays: any ParticleArray-like type whose vector() is a std::vector of
 *    Particle<dim> records (src/core/data/particles/particle_array.hpp:228-229,
 *    particle.hpp:37-61) goes through the AoS entry points unchanged;
 *  - fields: any Field-like type with data() and shape() in GridLayout::allocSize order
 *    (src/core/data/field/field.hpp, src/core/data/grid/grid.hpp) is copied with a shape check;
 *  - IonUpdaterPHX: IonUpdater's interface (updatePopulations, updateIons;
 *    src/core/numerics/ion_updater/ion_updater.hpp:58-63) over an Ions-like range of populations
 *    (domainParticles(), patchGhostParticles(), levelGhostParticles(), particleDensity(),
 *    chargeDensity(), flux()), as SolverPPC::moveIons_ calls it per patch
 *    (src/amr/solvers/solver_ppc.hpp:553-556).
 * Errors: a two-cell move becomes the exception type the caller names (PHARE passes its
 * DictionaryException-based error, boris.hpp:120-134), every other failure std::runtime_error
 * with phx_last_error's message.
 */
#ifndef PHX_PHARE_HPP
#define PHX_PHARE_HPP

#include "phx.h"

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace phx_phare
{
struct TwoCellMove : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

template<typename TwoCellError = TwoCellMove>
inline void check(phx_ctx* c, phx_status s, char const* what)
{
    if (s == PHX_OK)
        return;
    std::string msg = std::string(what) + ": " + (c ? phx_last_error(c) : "no context");
    if (s == PHX_ERR_MOVE_TWO_CELLS)
        throw TwoCellError(msg);
    throw std::runtime_error(msg);
}

// the reference's Particle<dim> record: 56/64/80 B for dim 1/2/3, the layout phx_set_particles_aos
// reads (weight, charge, iCell[dim], delta[dim], v[3])
template<typename ParticleArray>
inline void uploadParticles(phx_ctx* c, int pop, int kind, ParticleArray const& arr)
{
    auto const& v = arr.vector();
    check(c, phx_set_particles_aos(c, pop, kind, v.size(), v.data()), "phx_set_particles_aos");
}

// the array's content replaced by the context's particles (push_back keeps its cell map)
template<typename ParticleArray>
inline void downloadParticles(phx_ctx* c, int pop, int kind, ParticleArray& arr)
{
    std::size_t n = 0;
    check(c, phx_particle_count(c, pop, kind, &n), "phx_particle_count");
    using Particle_t = typename ParticleArray::value_type;
    std::vector<Particle_t> tmp(n);
    if (n)
        check(c, phx_get_particles_aos(c, pop, kind, tmp.data()), "phx_get_particles_aos");
    arr.clear();
    arr.reserve(n);
    for (auto& p : tmp)
        arr.push_back(p);
}

template<typename Field>
inline void checkShape(phx_ctx* c, int fid, Field const& f)
{
    int shape[3] = {1, 1, 1};
    check(c, phx_field_shape(c, fid, shape), "phx_field_shape");
    auto const s = f.shape();
    for (std::size_t d = 0; d < s.size(); ++d)
        if (static_cast<int>(s[d]) != shape[d])
            throw std::runtime_error("phx_phare: field " + std::to_string(fid)
                                     + " differs in shape from the context's layout");
}

template<typename Field>
inline void uploadField(phx_ctx* c, int fid, Field const& f)
{
    checkShape(c, fid, f);
    check(c, phx_set_field(c, fid, f.data()), "phx_set_field");
}

template<typename Field>
inline void downloadField(phx_ctx* c, int fid, Field& f)
{
    checkShape(c, fid, f);
    check(c, phx_get_field(c, fid, f.data()), "phx_get_field");
}

// IonUpdater's interface over one patch's context: the caller creates the context for the patch
// (phx_create with its boxes), this class moves the data in and out around each call
template<typename TwoCellError = TwoCellMove>
class IonUpdaterPHX
{
public:
    explicit IonUpdaterPHX(phx_ctx* ctx)
        : ctx_{ctx}
    {
    }

    // IonUpdater::updatePopulations(ions, em, boxing, dt, mode) (ion_updater.hpp:88-107): E and
    // B into the averaged set, the particle arrays of every population in, the update, then the
    // moments and (mode all) the particle arrays back
    template<typename Ions, typename Electromag>
    void updatePopulations(Ions& ions, Electromag const& em, double dt, bool modeAll)
    {
        for (int c = 0; c < 3; ++c)
        {
            uploadField(ctx_, PHX_F_EAVG_X + c, em.E[c]);
            uploadField(ctx_, PHX_F_EAVG_X + 3 + c, em.B[c]);
        }
        int p = 0;
        for (auto& pop : ions)
        {
            uploadParticles(ctx_, p, PHX_DOMAIN, pop.domainParticles());
            uploadParticles(ctx_, p, PHX_PATCH_GHOST, pop.patchGhostParticles());
            uploadParticles(ctx_, p, PHX_LEVEL_GHOST, pop.levelGhostParticles());
            ++p;
        }
        check<TwoCellError>(ctx_,
                            phx_update_populations(ctx_, PHX_EM_AVG, dt,
                                                   modeAll ? PHX_MODE_ALL : PHX_MODE_DOMAIN_ONLY),
                            "Updater::updatePopulations");
        p = 0;
        for (auto& pop : ions)
        {
            downloadField(ctx_, PHX_F_POP(p, 0), pop.particleDensity());
            downloadField(ctx_, PHX_F_POP(p, 1), pop.chargeDensity());
            for (int c = 0; c < 3; ++c)
                downloadField(ctx_, PHX_F_POP(p, 2 + c), pop.flux()[c]);
            if (modeAll)
            {
                downloadParticles(ctx_, p, PHX_DOMAIN, pop.domainParticles());
                downloadParticles(ctx_, p, PHX_PATCH_GHOST, pop.patchGhostParticles());
                downloadParticles(ctx_, p, PHX_LEVEL_GHOST, pop.levelGhostParticles());
            }
            ++p;
        }
    }

    // IonUpdater::updateIons (ion_updater.hpp:111-116): the totals from the population moments
    // already in the context
    template<typename Field, typename VecField>
    void updateIons(Field& chargeDensity, Field& massDensity, VecField& bulkVelocity)
    {
        check(ctx_, phx_update_ions(ctx_), "phx_update_ions");
        downloadField(ctx_, PHX_F_RHO, chargeDensity);
        downloadField(ctx_, PHX_F_MASSD, massDensity);
        for (int c = 0; c < 3; ++c)
            downloadField(ctx_, PHX_F_VIX + c, bulkVelocity[c]);
    }

private:
    phx_ctx* ctx_;
};

} // namespace phx_phare

#endif /* PHX_PHARE_HPP */
