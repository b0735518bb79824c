// refcpu_amr.hpp -- CPU ORACLE (test infrastructure only): the coarse-to-fine particle split of
// PHARE restated plainly, with the split patterns as data:
//  - toFineGrid (src/amr/data/particles/refine/particles_data_split.hpp:30-45);
//  - ParticlesRefineOperator::refine_ (same file, :157-260): for each destination box, each
//    coarse domain particle whose fine cell lies in the box grown by maxCellDistanceFromSplit
//    (splitter.hpp:116-120) is split and its refined particles inside the box are kept;
//  - PatternDispatcher::dispatch (splitter.hpp:71-103) with the patterns and constants of
//    split_1d.hpp (every 1D splitter), split_2d.hpp (4 refined particles) and split_3d.hpp (6).
// Used only by tests/ as the checker of the HIP split; never by the product.
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
