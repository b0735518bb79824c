// refcpu_amr.hpp -- CPU ORACLE (test infrastructure only): the coarse-to-fine particle split of
// PHARE restated plainly, with the split patterns as data:
//  - toFineGrid (src/amr/data/particles/refine/particles_data_split.hpp:30-45);
//  - ParticlesRefineOperator::refine_ (same file, :157-260): for each destination box, each
//    coarse domain particle whose fine cell lies in the box grown by maxCellDistanceFromSplit
//    (splitter.hpp:116-120) is split and its refined particles inside the box are kept;
//  - PatternDispatcher::dispatch (splitter.hpp:71-103) with the patterns and constants of
//    split_1d.hpp (every 1D splitter), split_2d.hpp (4 refined particles) and split_3d.hpp (6);
// and the coarse-to-fine field refinement (ratio 2) of src/amr/data/field/refine/:
//  - DefaultFieldRefiner / FieldMomentsRefiner (field_refiner.hpp, field_moments_refiner.hpp) with
//    FieldRefineIndexesAndWeights (field_linear_refine.hpp) and LinearWeighter (linear_weighter.cpp);
//  - ElectricFieldRefiner (electric_field_refiner.hpp);
//  - MagneticFieldRefiner / MagneticFieldInitRefiner (magnetic_field_refiner.hpp,
//    magnetic_field_init_refiner.hpp);
//  - MagneticRefinePatchStrategy::postprocessRefine (magnetic_refine_patch_strategy.hpp).
// Used only by tests/ as the checker of the HIP kernels; never by the product.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace refcpu
{
struct SplitPart
{
    int ic[3];
    double de[3], v[3], w, q;
};

struct SplitTable
{
    std::vector<float> w;
    std::vector<std::array<float, 3>> d;
};

// the pattern list of one splitter; empty when not restated here
inline SplitTable splitTable(int dim, int order, int nbr)
{
    SplitTable T;
    auto put   = [&](float w, float x, float y, float z) {
        T.w.push_back(w);
        T.d.push_back({x, y, z});
    };
    auto black = [&](float w) { put(w, 0, 0, 0); };
    // PinkPattern<1>: {+d}, {-d}
    auto pink1 = [&](float w, float d) {
        put(w, d, 0, 0);
        put(w, -d, 0, 0);
    };
    // PinkPattern<2>: {0,-d}, {-d,0}, {+d,0}, {0,+d}
    auto pink2 = [&](float w, float d) {
        put(w, 0, -d, 0);
        put(w, -d, 0, 0);
        put(w, d, 0, 0);
        put(w, 0, d, 0);
    };
    // PurplePattern<2>: the four corners
    auto purple2 = [&](float w, float d) {
        put(w, -d, -d, 0);
        put(w, -d, d, 0);
        put(w, d, -d, 0);
        put(w, d, d, 0);
    };
    // PinkPattern<3>: (m,0,0), (0,0,m), (0,m,0) for m = +d then -d
    auto pink3 = [&](float w, float d) {
        for (float m : {d, -d})
        {
            put(w, m, 0, 0);
            put(w, 0, 0, m);
            put(w, 0, m, 0);
        }
    };
    switch (dim * 100 + order * 10 + nbr)
    {
        case 112: pink1(0.5f, 0.551569f); break;
        case 113: black(0.5f); pink1(0.25f, 1.f); break;
        case 122: pink1(0.5f, 0.663959f); break;
        case 123: black(0.468137f); pink1(0.265931f, 1.112033f); break;
        case 124: pink1(0.375f, 0.5f); pink1(0.125f, 1.5f); break;
        case 132: pink1(0.5f, 0.752399f); break;
        case 133: black(0.473943f); pink1(0.263028f, 1.275922f); break;
        case 134: pink1(0.364766f, 0.542949f); pink1(0.135234f, 1.664886f); break;
        case 135: black(0.375f); pink1(0.25f, 1.f); pink1(0.0625f, 2.f); break;
        case 214: purple2(0.25f, 0.571783f); break;
        case 224: pink2(0.25f, 0.683734f); break;
        case 234: purple2(0.25f, 0.776459f); break;
        case 316: pink3(.166666f, .966431f); break;
        case 326: pink3(.166666f, 1.149658f); break;
        case 336: pink3(.166666f, 1.312622f); break;
        default: break;
    }
    return T;
}

inline SplitPart toFineGrid(SplitPart p, int dim)
{
    for (int d = 0; d < dim; ++d)
    {
        double fineDelta = p.de[d] * 2;
        int fineDeltaInt = static_cast<int>(fineDelta);
        p.ic[d]          = p.ic[d] * 2 + fineDeltaInt;
        p.de[d]          = fineDelta - fineDeltaInt;
    }
    return p;
}

inline bool inSplitBox(int const* ic, int const* lo, int const* hi, int dim, int grow)
{
    for (int d = 0; d < dim; ++d)
        if (ic[d] < lo[d] - grow || ic[d] > hi[d] + grow)
            return false;
    return true;
}

// refine_ over the destination boxes (boxes: nb x (lo[3], hi[3])); appends to out
inline void splitParticles(int dim, int order, int nbr, std::vector<SplitPart> const& coarse,
                           int nb, int const* lo, int const* hi, std::vector<SplitPart>& out)
{
    SplitTable const T = splitTable(dim, order, nbr);
    int const grow     = static_cast<int>(std::ceil((order + 1) * 0.5));
    double const power = dim == 1 ? 2. : (dim == 2 ? 4. : 8.);
    for (int b = 0; b < nb; ++b)
    {
        int const* blo = lo + 3 * b;
        int const* bhi = hi + 3 * b;
        for (auto const& cp : coarse)
        {
            SplitPart const fp = toFineGrid(cp, dim);
            if (!inSplitBox(fp.ic, blo, bhi, dim, grow))
                continue;
            for (size_t r = 0; r < T.w.size(); ++r)
            {
                SplitPart q = fp;
                q.w         = fp.w * static_cast<double>(T.w[r]) * power;
                for (int d = 0; d < dim; ++d)
                {
                    q.de[d] += static_cast<double>(T.d[r][d]);
                    double const integra = std::floor(q.de[d]);
                    q.de[d] -= integra;
                    q.ic[d] += static_cast<int>(integra);
                }
                if (inSplitBox(q.ic, blo, bhi, dim, 0))
                    out.push_back(q);
            }
        }
    }
}
} // namespace refcpu

namespace refcpu
{
// a field array addressed by AMR index: row-major over (x, y, z), org = the AMR index of local
// index 0 per direction (unused directions: extent 1, org 0)
struct AmrField
{
    double* p;
    int shape[3];
    int org[3];
    double& operator()(int i, int j = 0, int k = 0) const
    {
        return p[(static_cast<long>(i - org[0]) * shape[1] + (j - org[1])) * shape[2] + (k - org[2])];
    }
};

// toCoarseIndex (src/amr/resources_manager/amr_utils.hpp:129-135)
inline int toCoarse(int i) { return i >= 0 ? i / 2 : i / 2 + i % 2; }

// coarseStartIndex and the weight table (field_linear_refine.hpp:60-116, linear_weighter.cpp):
// primal: start floor(f / 2), weights {1, 0} (f even) or {1/2, 1/2}; dual: start
// floor((f + 1/2) / 2 - 1/2), weights {1/4, 3/4} (f even) or {3/4, 1/4}; the weight row is |f| % 2
inline int linearStart(int f, bool primal)
{
    double const shift = primal ? 0.0 : 0.5;
    return static_cast<int>(std::floor(static_cast<double>(f + shift) / 2 - shift));
}
inline std::array<double, 2> linearWeights(int f, bool primal)
{
    int const row = std::abs(f) % 2;
    // distance of the fine node from the left coarse node, in coarse cells
    double const dist = primal ? (row == 0 ? 0.0 : 0.5) : (row == 0 ? 0.75 : 0.25);
    return {1. - dist, dist};
}

enum
{
    ORA_REFINE_LINEAR           = 0,
    ORA_REFINE_LINEAR_OVERWRITE = 1,
    ORA_REFINE_ELECTRIC         = 2,
    ORA_REFINE_MAGNETIC         = 3,
    ORA_REFINE_MAGNETIC_INIT    = 4
};

// one fine node of the linear refiners (field_refiner.hpp:58-161)
inline double linearNode(int dim, AmrField const& c, int const* f, int const* primal)
{
    int s[3] = {0, 0, 0};
    std::array<double, 2> w[3];
    for (int d = 0; d < dim; ++d)
    {
        s[d] = linearStart(f[d], primal[d] != 0);
        w[d] = linearWeights(f[d], primal[d] != 0);
    }
    double value = 0.;
    if (dim == 1)
    {
        for (int ix = 0; ix < 2; ++ix)
            value += c(s[0] + ix) * w[0][ix];
    }
    else if (dim == 2)
    {
        for (int ix = 0; ix < 2; ++ix)
        {
            double yi = 0.;
            for (int iy = 0; iy < 2; ++iy)
                yi += c(s[0] + ix, s[1] + iy) * w[1][iy];
            value += yi * w[0][ix];
        }
    }
    else
    {
        for (int ix = 0; ix < 2; ++ix)
        {
            double yi = 0.;
            for (int iy = 0; iy < 2; ++iy)
            {
                double zi = 0.;
                for (int iz = 0; iz < 2; ++iz)
                    zi += c(s[0] + ix, s[1] + iy, s[2] + iz) * w[2][iz];
                yi += zi * w[1][iy];
            }
            value += yi * w[0][ix];
        }
    }
    return value;
}

// one fine edge of ElectricFieldRefiner (electric_field_refiner.hpp:40-345); P = primal flags
inline double electricNode(int dim, AmrField const& c, int const* f, int const* P)
{
    int const x = toCoarse(f[0]);
    int const y = dim > 1 ? toCoarse(f[1]) : 0;
    int const z = dim > 2 ? toCoarse(f[2]) : 0;
    auto even   = [&](int d) { return f[d] % 2 == 0; };
    if (dim == 1)
        return c(x);
    if (dim == 2)
    {
        if (!P[0] && P[1]) // Ex
            return even(1) ? c(x, y) : 0.5 * (c(x, y) + c(x, y + 1));
        if (P[0] && !P[1]) // Ey
            return even(0) ? c(x, y) : 0.5 * (c(x, y) + c(x + 1, y));
        // Ez
        if (even(0) && even(1))
            return c(x, y);
        if (even(0))
            return 0.5 * (c(x, y) + c(x, y + 1));
        if (even(1))
            return 0.5 * (c(x, y) + c(x + 1, y));
        return 0.25 * (c(x, y) + c(x + 1, y) + c(x, y + 1) + c(x + 1, y + 1));
    }
    if (!P[0]) // Ex
    {
        if (even(1) && even(2))
            return c(x, y, z);
        if (even(1))
            return 0.5 * (c(x, y, z) + c(x, y, z + 1));
        if (even(2))
            return 0.5 * (c(x, y, z) + c(x, y + 1, z));
        return 0.25 * (c(x, y, z) + c(x, y + 1, z)) + 0.25 * (c(x, y, z + 1) + c(x, y + 1, z + 1));
    }
    if (!P[1]) // Ey
    {
        if (even(0) && even(2))
            return c(x, y, z);
        if (even(0))
            return 0.5 * (c(x, y, z) + c(x, y, z + 1));
        if (even(2))
            return 0.5 * (c(x, y, z) + c(x + 1, y, z));
        return 0.25 * (c(x, y, z) + c(x + 1, y, z) + c(x, y, z + 1) + c(x + 1, y, z + 1));
    }
    // Ez
    if (even(0) && even(1))
        return c(x, y, z);
    if (even(0))
        return 0.5 * (c(x, y, z) + c(x, y + 1, z));
    if (even(1))
        return 0.5 * (c(x, y, z) + c(x + 1, y, z));
    return 0.25 * (c(x, y, z) + c(x + 1, y, z) + c(x, y + 1, z) + c(x + 1, y + 1, z));
}

// FieldRefineOperator::refine over one fine box (inclusive, fine AMR indices of the quantity)
inline void refineField(int dim, int op, int const* primal, AmrField const& fine, AmrField const& coarse,
                        int const* lo, int const* hi)
{
    int L[3] = {0, 0, 0}, H[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d)
    {
        L[d] = lo[d];
        H[d] = hi[d];
    }
    for (int i = L[0]; i <= H[0]; ++i)
        for (int j = L[1]; j <= H[1]; ++j)
            for (int k = L[2]; k <= H[2]; ++k)
            {
                int const f[3] = {i, j, k};
                double& dst    = fine(i, j, k);
                bool const nanOnly = op != ORA_REFINE_LINEAR_OVERWRITE && op != ORA_REFINE_MAGNETIC_INIT;
                if (op == ORA_REFINE_LINEAR || op == ORA_REFINE_LINEAR_OVERWRITE)
                {
                    double const v = linearNode(dim, coarse, f, primal);
                    if (!nanOnly || std::isnan(dst))
                        dst = v;
                }
                else if (op == ORA_REFINE_ELECTRIC)
                {
                    if (std::isnan(dst))
                        dst = electricNode(dim, coarse, f, primal);
                }
                else
                {
                    // magnetic: a component primal along d (its normal) is copied on the fine
                    // faces with an even index along d; the all-dual ones (By, Bz in 1D, Bz in 2D)
                    // everywhere
                    bool shared = true;
                    for (int d = 0; d < dim; ++d)
                        if (primal[d] && f[d] % 2 != 0)
                            shared = false;
                    if (shared && (!nanOnly || std::isnan(dst)))
                        dst = coarse(toCoarse(i), dim > 1 ? toCoarse(j) : 0, dim > 2 ? toCoarse(k) : 0);
                }
            }
}

// MagneticRefinePatchStrategy::postprocessRefine over a fine cell box (magnetic_refine_patch_
// strategy.hpp:67-360), restated per component as there: Bx (1D, 2D, 3D), By (2D, 3D), Bz (3D)
inline void magneticNewFaces(int dim, AmrField const& bx, AmrField const& by, AmrField const& bz,
                             double const* meshSize, int const* lo, int const* hi)
{
    auto pp = [](int i, int o) { return i + 2 - o; };
    auto pm = [](int i, int o) { return i - o; };
    auto dp = [](int i, int o) { return i + 1 - o; };
    auto dm = [](int i, int o) { return i - o; };
    auto odd = [](int i) { return i % 2 != 0; };
    // ORA_PLANT_*: mistakes planted by the tests to show that their checks catch them (the
    // defaults are the reference's values and leave every result bit-identical)
#ifndef ORA_PLANT_IJK
#define ORA_PLANT_IJK -1., 1.
#endif
#ifndef ORA_PLANT_CROSS
#define ORA_PLANT_CROSS 1.
#endif
    double const ijk[2] = {ORA_PLANT_IJK};
    double const cross  = ORA_PLANT_CROSS;
    double const Dx = meshSize[0], Dy = dim > 1 ? meshSize[1] : 0., Dz = dim > 2 ? meshSize[2] : 0.;
    // the field box of a component: the cell box with one more node along its primal direction
    auto box = [&](int comp, int d, int& a, int& b) {
        a = d < dim ? lo[d] : 0;
        b = d < dim ? hi[d] + (d == comp ? 1 : 0) : 0;
    };
    for (int comp = 0; comp < dim; ++comp)
    {
        int a0, b0, a1, b1, a2, b2;
        box(comp, 0, a0, b0);
        box(comp, 1, a1, b1);
        box(comp, 2, a2, b2);
        for (int ix = a0; ix <= b0; ++ix)
            for (int iy = a1; iy <= b1; ++iy)
                for (int iz = a2; iz <= b2; ++iz)
                {
                    int const idx[3] = {ix, iy, iz};
                    if (!odd(idx[comp]))
                        continue;
                    // at the edge of the arrays (this build's rule; the reference refines into
                    // a scratch region one stencil wider): without both neighbours along the
                    // normal the face is left, with cross terms reading beyond the arrays it takes
                    // the neighbours' mean
                    AmrField const& own = comp == 0 ? bx : (comp == 1 ? by : bz);
                    auto inF = [&](AmrField const& F, int i, int j, int k) {
                        int const a[3] = {i, j, k};
                        for (int d = 0; d < 3; ++d)
                            if (a[d] - F.org[d] < 0 || a[d] - F.org[d] >= F.shape[d])
                                return false;
                        return true;
                    };
                    int lo3[3] = {ix, iy, iz}, hi3[3] = {ix, iy, iz};
                    lo3[comp] -= 1;
                    hi3[comp] += 1;
                    if (!inF(own, lo3[0], lo3[1], lo3[2]) || !inF(own, hi3[0], hi3[1], hi3[2]))
                        continue;
                    bool crossIn = true;
                    for (int c = 0; c < dim && dim > 1; ++c)
                    {
                        if (c == comp)
                            continue;
                        AmrField const& F = c == 0 ? bx : (c == 1 ? by : bz);
                        for (int sx = 0; sx < 2; ++sx)
                            for (int sy = 0; sy < (dim > 1 ? 2 : 1); ++sy)
                                for (int sz = 0; sz < (dim > 2 ? 2 : 1); ++sz)
                                {
                                    int const sd[3] = {sx, sy, sz};
                                    int r[3]        = {0, 0, 0};
                                    for (int d = 0; d < dim; ++d)
                                    {
                                        int const o = d == comp ? 1 : (odd(idx[d]) ? 1 : 0);
                                        r[d] = sd[d] ? (d == c ? pp(idx[d], o) : dp(idx[d], o))
                                                     : (d == c ? pm(idx[d], o) : dm(idx[d], o));
                                    }
                                    crossIn = crossIn && inF(F, r[0], r[1], r[2]);
                                }
                    }
                    if (dim > 1 && !crossIn)
                    {
                        own(ix, iy, iz) = 0.5 * (own(lo3[0], lo3[1], lo3[2]) + own(hi3[0], hi3[1], hi3[2]));
                        continue;
                    }
                    if (dim == 1)
                    {
                        bx(ix) = 0.5 * (bx(ix - 1) + bx(ix + 1));
                        continue;
                    }
                    if (dim == 2)
                    {
                        if (comp == 0)
                        {
                            int const xo = 1, yo = odd(iy) ? 1 : 0;
                            bx(ix, iy) = 0.5 * (bx(ix - 1, iy) + bx(ix + 1, iy))
                                         + cross * 0.25
                                               * (by(dm(ix, xo), pm(iy, yo)) - by(dm(ix, xo), pp(iy, yo))
                                                  - by(dp(ix, xo), pm(iy, yo)) + by(dp(ix, xo), pp(iy, yo)));
                        }
                        else
                        {
                            int const xo = odd(ix) ? 1 : 0, yo = 1;
                            by(ix, iy) = 0.5 * (by(ix, iy - 1) + by(ix, iy + 1))
                                         + cross * 0.25
                                               * (bx(pm(ix, xo), dm(iy, yo)) - bx(pp(ix, xo), dm(iy, yo))
                                                  - bx(pm(ix, xo), dp(iy, yo)) + bx(pp(ix, xo), dp(iy, yo)));
                        }
                        continue;
                    }
                    int const xo = comp == 0 ? 1 : (odd(ix) ? 1 : 0);
                    int const yo = comp == 1 ? 1 : (odd(iy) ? 1 : 0);
                    int const zo = comp == 2 ? 1 : (odd(iz) ? 1 : 0);
                    // the value of a component at signed offsets: s = -1 (minus side) or +1 (plus);
                    // p-type index along the component's own direction, d-type along the others
                    auto V = [&](AmrField const& F, int c, int sx, int sy, int sz) {
                        int const i = c == 0 ? (sx > 0 ? pp(ix, xo) : pm(ix, xo)) : (sx > 0 ? dp(ix, xo) : dm(ix, xo));
                        int const j = c == 1 ? (sy > 0 ? pp(iy, yo) : pm(iy, yo)) : (sy > 0 ? dp(iy, yo) : dm(iy, yo));
                        int const k = c == 2 ? (sz > 0 ? pp(iz, zo) : pm(iz, zo)) : (sz > 0 ? dp(iz, zo) : dm(iz, zo));
                        return F(i, j, k);
                    };
                    if (comp == 0)
                    {
                        bx(ix, iy, iz)
                            = 0.5 * (bx(ix - 1, iy, iz) + bx(ix + 1, iy, iz))
                              + cross * 0.125
                                    * (V(by, 1, -1, -1, -1) - V(by, 1, -1, +1, -1) - V(by, 1, +1, -1, -1)
                                       + V(by, 1, +1, +1, -1) + V(by, 1, -1, -1, +1) - V(by, 1, -1, +1, +1)
                                       - V(by, 1, +1, -1, +1) + V(by, 1, +1, +1, +1))
                              + cross * 0.125
                                    * (V(bz, 2, -1, -1, -1) + V(bz, 2, -1, +1, -1) - V(bz, 2, +1, -1, -1)
                                       - V(bz, 2, +1, +1, -1) - V(bz, 2, -1, -1, +1) - V(bz, 2, -1, +1, +1)
                                       + V(bz, 2, +1, -1, +1) + V(bz, 2, +1, +1, +1))
                              + (0.125 * ijk[zo] * Dz * Dz / (Dx * Dx + Dz * Dz))
                                    * (V(by, 1, +1, +1, +1) - V(by, 1, -1, +1, +1) - V(by, 1, +1, -1, +1)
                                       - V(by, 1, +1, +1, -1) + V(by, 1, +1, -1, -1) + V(by, 1, -1, +1, -1)
                                       + V(by, 1, -1, -1, +1) - V(by, 1, -1, -1, -1))
                              + (0.125 * ijk[yo] * Dy * Dy / (Dx * Dx + Dy * Dy))
                                    * (V(bz, 2, +1, +1, +1) - V(bz, 2, -1, +1, +1) - V(bz, 2, +1, -1, +1)
                                       - V(bz, 2, +1, +1, -1) + V(bz, 2, +1, -1, -1) + V(bz, 2, -1, +1, -1)
                                       + V(bz, 2, -1, -1, +1) - V(bz, 2, -1, -1, -1));
                    }
                    else if (comp == 1)
                    {
                        by(ix, iy, iz)
                            = 0.5 * (by(ix, iy - 1, iz) + by(ix, iy + 1, iz))
                              + cross * 0.125
                                    * (V(bx, 0, -1, -1, -1) - V(bx, 0, -1, +1, -1) - V(bx, 0, +1, -1, -1)
                                       + V(bx, 0, +1, +1, -1) + V(bx, 0, -1, -1, +1) - V(bx, 0, -1, +1, +1)
                                       - V(bx, 0, +1, -1, +1) + V(bx, 0, +1, +1, +1))
                              + cross * 0.125
                                    * (V(bz, 2, -1, -1, -1) - V(bz, 2, -1, +1, -1) + V(bz, 2, +1, -1, -1)
                                       - V(bz, 2, +1, +1, -1) - V(bz, 2, -1, -1, +1) + V(bz, 2, -1, +1, +1)
                                       - V(bz, 2, +1, -1, +1) + V(bz, 2, +1, +1, +1))
                              + (0.125 * ijk[xo] * Dx * Dx / (Dx * Dx + Dy * Dy))
                                    * (V(bz, 2, +1, +1, +1) - V(bz, 2, -1, +1, +1) - V(bz, 2, +1, -1, +1)
                                       - V(bz, 2, +1, +1, -1) + V(bz, 2, +1, -1, -1) + V(bz, 2, -1, +1, -1)
                                       + V(bz, 2, -1, -1, +1) - V(bz, 2, -1, -1, -1))
                              + (0.125 * ijk[zo] * Dz * Dz / (Dy * Dy + Dz * Dz))
                                    * (V(bx, 0, +1, +1, +1) - V(bx, 0, -1, +1, +1) - V(bx, 0, +1, -1, +1)
                                       - V(bx, 0, +1, +1, -1) + V(bx, 0, +1, -1, -1) + V(bx, 0, -1, +1, -1)
                                       + V(bx, 0, -1, -1, +1) - V(bx, 0, -1, -1, -1));
                    }
                    else
                    {
                        bz(ix, iy, iz)
                            = 0.5 * (bz(ix, iy, iz - 1) + bz(ix, iy, iz + 1))
                              + cross * 0.125
                                    * (V(bx, 0, -1, -1, -1) + V(bx, 0, -1, +1, -1) - V(bx, 0, +1, -1, -1)
                                       - V(bx, 0, +1, +1, -1) - V(bx, 0, -1, -1, +1) - V(bx, 0, -1, +1, +1)
                                       + V(bx, 0, +1, -1, +1) + V(bx, 0, +1, +1, +1))
                              + cross * 0.125
                                    * (V(by, 1, -1, -1, -1) - V(by, 1, -1, +1, -1) + V(by, 1, +1, -1, -1)
                                       - V(by, 1, +1, +1, -1) - V(by, 1, -1, -1, +1) + V(by, 1, -1, +1, +1)
                                       - V(by, 1, +1, -1, +1) + V(by, 1, +1, +1, +1))
                              + (0.125 * ijk[yo] * Dy * Dy / (Dy * Dy + Dz * Dz))
                                    * (V(bx, 0, +1, +1, +1) - V(bx, 0, -1, +1, +1) - V(bx, 0, +1, -1, +1)
                                       - V(bx, 0, +1, +1, -1) + V(bx, 0, +1, -1, -1) + V(bx, 0, -1, +1, -1)
                                       + V(bx, 0, -1, -1, +1) - V(bx, 0, -1, -1, -1))
                              + (0.125 * ijk[xo] * Dx * Dx / (Dx * Dx + Dz * Dz))
                                    * (V(by, 1, +1, +1, +1) - V(by, 1, -1, +1, +1) - V(by, 1, +1, -1, +1)
                                       - V(by, 1, +1, +1, -1) + V(by, 1, +1, -1, -1) + V(by, 1, -1, +1, -1)
                                       + V(by, 1, -1, -1, +1) - V(by, 1, -1, -1, -1));
                    }
                }
    }
}
} // namespace refcpu
