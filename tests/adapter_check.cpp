// Compile-and-link check of include/phx_phare.hpp against the reference's own particle and field
// types (src/core/data/particles/particle_array.hpp, src/core/data/field/field.hpp): the adapter
// templates are instantiated with ParticleArray<2> and Field<2, HybridQuantity::Scalar> and the
// program links against libphx.so. Built by tests/test_adapter_compile.py; not run (it needs a GPU).
#include "core/data/field/field.hpp"
#include "core/data/particles/particle_array.hpp"
#include "core/models/quantities/hybrid_quantities.hpp"

#include "phx_phare.hpp"

#include <vector>

namespace
{
using PHARE::core::Field;
using PHARE::core::HybridQuantity;
using PHARE::core::ParticleArray;
using Field2 = Field<2, HybridQuantity::Scalar>;

struct Pop
{
    ParticleArray<2> dom{PHARE::core::Box<int, 2>{{0, 0}, {9, 9}}};
    ParticleArray<2> pg{PHARE::core::Box<int, 2>{{0, 0}, {9, 9}}};
    ParticleArray<2> lg{PHARE::core::Box<int, 2>{{0, 0}, {9, 9}}};
    std::vector<double> buf = std::vector<double>(5 * 17 * 17);
    Field2 n{"n", HybridQuantity::Scalar::rho, buf.data(), {17, 17}};
    Field2 rho{"rho", HybridQuantity::Scalar::rho, buf.data() + 289, {17, 17}};
    std::array<Field2, 3> F{Field2{"Fx", HybridQuantity::Scalar::Vx, buf.data() + 578, {17, 17}},
                            Field2{"Fy", HybridQuantity::Scalar::Vy, buf.data() + 867, {17, 17}},
                            Field2{"Fz", HybridQuantity::Scalar::Vz, buf.data() + 1156, {17, 17}}};
    auto& domainParticles() { return dom; }
    auto& patchGhostParticles() { return pg; }
    auto& levelGhostParticles() { return lg; }
    auto& particleDensity() { return n; }
    auto& chargeDensity() { return rho; }
    auto& flux() { return F; }
};

struct EM
{
    std::vector<double> buf = std::vector<double>(6 * 17 * 17);
    std::array<Field2, 3> E{Field2{"Ex", HybridQuantity::Scalar::Ex, buf.data(), {16, 17}},
                            Field2{"Ey", HybridQuantity::Scalar::Ey, buf.data() + 289, {17, 16}},
                            Field2{"Ez", HybridQuantity::Scalar::Ez, buf.data() + 578, {17, 17}}};
    std::array<Field2, 3> B{Field2{"Bx", HybridQuantity::Scalar::Bx, buf.data() + 867, {17, 16}},
                            Field2{"By", HybridQuantity::Scalar::By, buf.data() + 1156, {16, 17}},
                            Field2{"Bz", HybridQuantity::Scalar::Bz, buf.data() + 1445, {16, 16}}};
};
} // namespace

int main()
{
    phx_ctx* c = nullptr;
    std::vector<Pop> ions(1);
    EM em;
    phx_phare::uploadParticles(c, 0, PHX_DOMAIN, ions[0].dom);
    phx_phare::downloadParticles(c, 0, PHX_DOMAIN, ions[0].dom);
    phx_phare::uploadField(c, PHX_F_RHO, ions[0].rho);
    phx_phare::downloadField(c, PHX_F_RHO, ions[0].rho);
    phx_phare::IonUpdaterPHX<> updater{c};
    updater.updatePopulations(ions, em, 0.01, true);
    updater.updateIons(ions[0].n, ions[0].rho, ions[0].F);
    return 0;
}
