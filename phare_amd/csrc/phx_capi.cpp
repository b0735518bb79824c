// phx_capi.cpp -- host side of the C ABI declared in include/phx.h.
//
// A context holds one patch level resident in HBM: the Yee fields of SolverPPC (E/B state,
// predicted and averaged sets, J, electron Ve/Pe), the ion totals and per-population moments,
// and the particle arrays of each population in structure-of-arrays form. Every kernel runs on
// the context's own HIP stream; SolverPPC::advanceLevel is captured once into a hipGraph and
// replayed per step. Device memory is allocated at creation (and when a particle array must
// grow); the hot path allocates nothing.
#include "../../include/phx.h"
#include "phx_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <vector>

// LDS tile policy of the push kernel: moment copies per dimension (lanes of a wavefront add into
// different copies)
#ifndef PHX_TILE_COPIES_1D
#define PHX_TILE_COPIES_1D 1
#endif
#ifndef PHX_TILE_COPIES_2D
#define PHX_TILE_COPIES_2D 8
#endif
#ifndef PHX_TILE_COPIES_3D
#define PHX_TILE_COPIES_3D 1
#endif
// moment copies of the segmented push (its lanes own distinct cells)
#ifndef PHX_SEG_COPIES
#define PHX_SEG_COPIES 1
#endif
// default launch shapes of the segmented push (phx_tuning)
#ifndef PHX_SEG_ROWS
#define PHX_SEG_ROWS 32
#endif
#ifndef PHX_TILE_ROW_PAD
#define PHX_TILE_ROW_PAD 16
#endif

using namespace phx;

static_assert(NOALLOC_GATHER_MAX_SLOTS == PHX_GATHER_NOALLOC_MAX_SLOTS,
              "the scratch array of phx_gather_noalloc is sized for the limit phx.h states");

namespace
{

// particle arrays of one kind: int32 cells, and delta, velocity, weight (and charge) as FP64
// (real = 8) or, in the mixed-precision family, FP32 without a charge array (real = 4)
struct Store
{
    size_t n = 0, cap = 0;
    bool seg = false; // cell-segmented by the sort (the population's segment tables describe it)
    int real = 8;
    int32_t* ic[3] = {nullptr, nullptr, nullptr};
    void* de[3]    = {nullptr, nullptr, nullptr};
    void* v[3]     = {nullptr, nullptr, nullptr};
    void* w        = nullptr;
    void* q        = nullptr;

    template<class R>
    PViewT<R> viewT() const
    {
        PViewT<R> p;
        for (int d = 0; d < 3; ++d)
        {
            p.ic[d] = ic[d];
            p.de[d] = static_cast<R*>(de[d]);
            p.v[d]  = static_cast<R*>(v[d]);
        }
        p.w = static_cast<R*>(w);
        p.q = static_cast<R*>(q);
        return p;
    }
    PView view() const { return viewT<double>(); }
    PView32 view32() const { return viewT<float>(); }
};

struct Pop
{
    double mass = 1.;
    double charge = 0.; // the particles' charge (mixed-precision family: one per population)
    bool hasCharge = false;
    bool uniformCharge = true; // every particle uploaded so far has charge `charge`
    Store arr[5]; // domain, patch ghost, level ghost, level ghost old, level ghost new
    double* mom[5] = {};
    SegTables seg{}; // tables of the segmented domain array (items sized for itemsAlloc items)
    unsigned int itemsAlloc = 0;
};

} // namespace

struct phx_ctx
{
    phx_config cfg{};
    phx_patch patch{};
    Geom G{};
    Boxes boxes{};
    hipStream_t stream = nullptr;

    std::vector<void*> allocs;
    double* em[3][6] = {};    // E/B of each set, one block per set (em[set][0] is its base)
    float* em32[6]   = {};    // mixed-precision family: FP32 copy of the averaged set (gather)
    double* J[3]     = {};
    double* Bold[3]    = {}; // SolverPPC::Bold_ (prepareStep)
    double* fluxSum[3] = {}; // SolverPPC::fluxSumE_ (E-centered)
    double* Ve[3]    = {};
    double* Pe = nullptr, *rho = nullptr, *massD = nullptr, *tmp = nullptr;
    double* Vi[3] = {};
    Pop pops[PHX_MAX_POP];
    Store scratch;
    // fused periodic path: two raw deposit sets (one filled while the other is cleared) and
    // the totals before their border max
    double* raw[2][PHX_MAX_POP][5] = {};
    int rawCur                     = 0;
    double* tot5[5]                = {};

    DevError* err           = nullptr;
    unsigned int* dropped   = nullptr;
    uint8_t* status         = nullptr;
    size_t statusCap        = 0;
    unsigned int* blockCnt  = nullptr;
    size_t blockCap         = 0;
    unsigned long long* tot = nullptr;

    // cell sort of the domain arrays: a second store per population (the sort writes there and
    // the two swap), the key counts, and the interval in PPC steps of phx_advance
    Store sortBuf[PHX_MAX_POP];
    SegGeom sg{};
    // the level's patches: AMR domain boxes (host) and their segment records
    std::vector<std::array<int, 6>> patchBox;
    std::vector<SegPatch> segPatch;
    unsigned int* segWork = nullptr; // per-cell counts and per-segment work items of a sort
    unsigned int* sortRank = nullptr; // rank of each particle in its cell
    size_t rankCap         = 0;
    unsigned int* sortPerm = nullptr; // source particle of each sorted slot
    size_t permCap         = 0;
    int sortInterval      = 30;
    int stepsSinceSort    = 0;
    // launch shapes of the segmented push (phx_set_tuning)
    phx_tuning tuning{};
    bool tileFallbackNoted = false; // the note that a segment tile outgrew LDS was printed

    // captured PPC steps, keyed on dt, the raw deposit set and the particle storage they address
    struct GraphEntry
    {
        hipGraphExec_t exec = nullptr;
        double dt           = 0.;
        int rawCur          = -1;
        std::vector<unsigned long long> key;
    };
    GraphEntry graphs[2];
    // device address of the scratch array of phx_gather_noalloc on the context's device
    void* noAllocScratch = nullptr;
    // timing of the pushes of the staged 3D path: a HIP event pair around each push launch
    // (eager launches), kept until phx_set_push_timing resets them
    struct EventPair
    {
        hipEvent_t a = nullptr, b = nullptr;
        int mode = 0, pop = 0;
    };
    std::vector<EventPair> events;
    size_t eventsUsed = 0;
    int graphNext = 0;
    // in-kernel timing of the pushes of the fast path, per (mode, population)
    KTime* kt   = nullptr;
    bool timing = false;
    std::string lastError;
};

namespace
{

phx_status fail(phx_ctx* c, phx_status s, std::string const& msg)
{
    if (c)
        c->lastError = msg;
    return s;
}

#define PHX_HIP(ctx, call)                                                                         \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, PHX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define PHX_TRY(expr)                                                                              \
    do                                                                                             \
    {                                                                                              \
        phx_status s_ = (expr);                                                                    \
        if (s_ != PHX_OK)                                                                          \
            return s_;                                                                             \
    } while (0)

phx_status devAlloc(phx_ctx* c, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0)
        bytes = 8;
    PHX_HIP(c, hipMalloc(p, bytes));
    c->allocs.push_back(*p);
    return PHX_OK;
}

void devFree(phx_ctx* c, void* p)
{
    if (!p)
        return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it != c->allocs.end())
        c->allocs.erase(it);
    (void)hipFree(p);
}

bool mixed(phx_ctx const* c) { return c->cfg.precision == PHX_PRECISION_FP32; }

phx_status storeAlloc(phx_ctx* c, Store& s, size_t cap)
{
    int const D = c->cfg.dim;
    s.real      = mixed(c) ? 4 : 8;
    size_t const r = static_cast<size_t>(s.real);
    for (int d = 0; d < 3; ++d)
    {
        PHX_TRY(devAlloc(c, (void**)&s.ic[d], d < D ? cap * 4 : 8));
        PHX_TRY(devAlloc(c, &s.de[d], d < D ? cap * r : 8));
        PHX_TRY(devAlloc(c, &s.v[d], cap * r));
    }
    PHX_TRY(devAlloc(c, &s.w, cap * r));
    if (s.real == 8)
        PHX_TRY(devAlloc(c, &s.q, cap * r));
    s.cap = cap;
    return PHX_OK;
}

void storeFree(phx_ctx* c, Store& s)
{
    for (int d = 0; d < 3; ++d)
    {
        devFree(c, s.ic[d]);
        devFree(c, s.de[d]);
        devFree(c, s.v[d]);
        s.ic[d] = nullptr;
        s.de[d] = nullptr;
        s.v[d]  = nullptr;
    }
    devFree(c, s.w);
    devFree(c, s.q);
    s.w = s.q = nullptr;
    s.cap = s.n = 0;
}

// grow a store to hold at least `need` particles, keeping its content
phx_status storeReserve(phx_ctx* c, Store& s, size_t need)
{
    if (need <= s.cap)
        return PHX_OK;
    size_t cap = std::max(need, s.cap + s.cap / 2 + 1024);
    Store ns;
    PHX_TRY(storeAlloc(c, ns, cap));
    int const D    = c->cfg.dim;
    size_t const r = static_cast<size_t>(s.real);
    if (s.n)
    {
        for (int d = 0; d < D; ++d)
        {
            PHX_HIP(c, hipMemcpyAsync(ns.ic[d], s.ic[d], s.n * 4, hipMemcpyDeviceToDevice, c->stream));
            PHX_HIP(c, hipMemcpyAsync(ns.de[d], s.de[d], s.n * r, hipMemcpyDeviceToDevice, c->stream));
        }
        for (int d = 0; d < 3; ++d)
            PHX_HIP(c, hipMemcpyAsync(ns.v[d], s.v[d], s.n * r, hipMemcpyDeviceToDevice, c->stream));
        PHX_HIP(c, hipMemcpyAsync(ns.w, s.w, s.n * r, hipMemcpyDeviceToDevice, c->stream));
        if (s.q)
            PHX_HIP(c, hipMemcpyAsync(ns.q, s.q, s.n * r, hipMemcpyDeviceToDevice, c->stream));
        PHX_HIP(c, hipStreamSynchronize(c->stream));
    }
    ns.n = s.n;
    storeFree(c, s);
    s = ns;
    return PHX_OK;
}

phx_status scratchReserve(phx_ctx* c, size_t n)
{
    PHX_TRY(storeReserve(c, c->scratch, n));
    if (n > c->statusCap)
    {
        devFree(c, c->status);
        size_t cap = std::max(n, c->statusCap * 2);
        PHX_TRY(devAlloc(c, (void**)&c->status, cap));
        c->statusCap = cap;
    }
    size_t nb = compactBlocks(n) + 1;
    if (nb > c->blockCap)
    {
        devFree(c, c->blockCnt);
        PHX_TRY(devAlloc(c, (void**)&c->blockCnt, nb * 4 * 2));
        c->blockCap = nb * 2;
    }
    return PHX_OK;
}

// field id -> (pointer, quantity)
bool fieldOf(phx_ctx const* c, int fid, double*& p, int& q)
{
    if (fid >= PHX_F_POP0 && fid < PHX_F_POP0 + 8 * PHX_MAX_POP)
    {
        int pop = (fid - PHX_F_POP0) / 8, k = (fid - PHX_F_POP0) % 8;
        if (pop >= c->cfg.n_pop || k > 4)
            return false;
        p = c->pops[pop].mom[k];
        q = QMOM;
        return true;
    }
    if (fid >= 0 && fid < 18)
    {
        int set = fid / 6, r = fid % 6;
        p       = c->em[set][r];
        q       = r < 3 ? QEX + r : QBX + (r - 3);
        return true;
    }
    if (fid >= PHX_F_JX && fid <= PHX_F_JZ)
    {
        p = c->J[fid - PHX_F_JX];
        q = QJX + (fid - PHX_F_JX);
        return true;
    }
    if (fid >= PHX_F_BOLD_X && fid <= PHX_F_BOLD_X + 2)
    {
        p = c->Bold[fid - PHX_F_BOLD_X];
        q = QBX + (fid - PHX_F_BOLD_X);
        return true;
    }
    if (fid >= PHX_F_FLUXSUM_X && fid <= PHX_F_FLUXSUM_X + 2)
    {
        p = c->fluxSum[fid - PHX_F_FLUXSUM_X];
        q = QEX + (fid - PHX_F_FLUXSUM_X);
        return true;
    }
    q = QMOM;
    if (fid >= PHX_F_VEX && fid <= PHX_F_VEZ)
        p = c->Ve[fid - PHX_F_VEX];
    else if (fid == PHX_F_PE)
        p = c->Pe;
    else if (fid == PHX_F_RHO)
        p = c->rho;
    else if (fid == PHX_F_MASSD)
        p = c->massD;
    else if (fid >= PHX_F_VIX && fid <= PHX_F_VIZ)
        p = c->Vi[fid - PHX_F_VIX];
    else
        return false;
    return true;
}

CVecPtr cvec(double* const* p) { return CVecPtr{{p[0], p[1], p[2]}}; }
VecPtr vec(double* const* p) { return VecPtr{{p[0], p[1], p[2]}}; }

phx_status checkLaunch(phx_ctx* c, char const* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(c, PHX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return PHX_OK;
}

// read the device error record; raise the reference's error semantics
phx_status checkDeviceError(phx_ctx* c)
{
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    DevError e{};
    PHX_HIP(c, hipMemcpy(&e, c->err, sizeof(e), hipMemcpyDeviceToHost));
    if (e.code == 0)
        return PHX_OK;
    DevError z{};
    PHX_HIP(c, hipMemcpy(c->err, &z, sizeof(z), hipMemcpyHostToDevice));
    std::ostringstream os;
    if (e.code == PHX_ERR_MOVE_TWO_CELLS)
        os << "Updater::updatePopulations: " << (e.stage == 1 ? "PrePush" : "PostPush")
           << " Particle moved 2 cells with delta/vel: " << e.delta << "/" << e.v << " (pop " << e.pop
           << ", array " << e.kind << ", particle " << e.particle << ")";
    else
        os << "Updater::updatePopulations: particle " << e.particle << " (pop " << e.pop << ", array "
           << e.kind << ") has an interpolation stencil outside the ghost box at stage " << e.stage;
    return fail(c, static_cast<phx_status>(e.code), os.str());
}

// arguments of the push kernel; `segOk` lets a cell-segmented domain array use the segmented
// kernel (callers that need a fate for every slot, the compacting path, pass false)
// a segment tile that outgrew the LDS budget (an extreme segment shape from phx_set_tuning):
// the launch deposits every particle straight into the moment arrays, correct but slow; said
// once per context on stderr
void noteTileFallback(phx_ctx* c, TileSpec const& T)
{
    if (T.lo[0] != -0x10000000 || c->tileFallbackNoted)
        return;
    c->tileFallbackNoted = true;
    std::fprintf(stderr,
                 "phx: the LDS tile of the segment push does not fit for segments of %dx%dx%d "
                 "cells; its deposits go straight to the moment arrays (slower)\n",
                 c->tuning.segment_shape[0], c->tuning.segment_shape[1], c->tuning.segment_shape[2]);
}

PushArgs pushArgs(phx_ctx* c, int pop, int kind, int mode, int emSet, double dt, bool segOk)
{
    PushArgs A{};
    A.G = c->G;
    for (int i = 0; i < 3; ++i)
    {
        A.em.E[i] = c->em[emSet][i];
        A.em.B[i] = c->em[emSet][3 + i];
    }
    for (int k = 0; k < 6; ++k)
        A.em32[k] = c->em32[k];
    for (int m = 0; m < 5; ++m)
        A.mom[m] = c->pops[pop].mom[m];
    for (int d = 0; d < 3; ++d)
        A.hdtdx[d] = d < c->cfg.dim ? 0.5 * dt / c->G.dx[d] : 0.;
    A.dto2m        = 0.5 * dt / c->pops[pop].mass;
    A.coef         = 1.;
    A.box          = c->boxes;
    A.kind         = kind;
    A.mode         = mode;
    A.periodicWrap = 0;
    A.pop          = pop;
    A.n            = c->pops[pop].arr[kind].n;
    A.err          = c->err;
    A.status       = nullptr;
    A.dropped      = c->dropped;
    A.pushOn       = 1;
    A.charge       = c->pops[pop].charge;
    // tile shape from the density of particles per cell of this array
    double ppc = static_cast<double>(A.n);
    for (int d = 0; d < c->cfg.dim; ++d)
        ppc /= c->G.n[d];
    int const copies = c->tuning.tile_copies;
    A.tile = tileShape(c->cfg.dim, c->cfg.interp_order, ppc, copies);
    Pop const& P   = c->pops[pop];
    Store const& s = P.arr[kind];
    // (3D: the linear-storage kernel on any storage order; the segmented push stays off in 3D with
    // the captured fused step)
    if (segOk && c->cfg.dim != 3 && kind == KIND_DOMAIN && s.seg
        && segItemsCap(c->sg, s.cap) <= P.itemsAlloc)
    {
        A.segOn       = 1;
        A.sg          = c->sg;
        A.st          = P.seg;
        A.st.itemsCap = segItemsCap(c->sg, s.cap);
        A.tile        = mixed(c) ? segTileShape32(c->cfg.dim, c->cfg.interp_order, c->sg)
                                 : segTileShape(c->cfg.dim, c->cfg.interp_order, c->sg, PHX_SEG_COPIES);
        A.rhoFromN    = !mixed(c) && P.hasCharge && P.uniformCharge ? 1 : 0;
        A.depositWindow = c->tuning.deposit_window;
        noteTileFallback(c, A.tile);
    }
    return A;
}

long count(phx_ctx const* c, int q) { return c->G.size(q); }

void zeroMoments(phx_ctx* c)
{
    for (int p = 0; p < c->cfg.n_pop; ++p)
        for (int m = 0; m < 5; ++m)
            launchZero(c->pops[p].mom[m], count(c, QMOM), c->stream);
}

// fillFluxBorders + fillDensityBorders: sum = copy(moment); sum += periodic images; copy back
phx_status fillMomentBorders(phx_ctx* c)
{
    if (!c->patch.periodic)
        return PHX_OK;
    for (int p = 0; p < c->cfg.n_pop; ++p)
        for (int m = 0; m < 5; ++m)
        {
            launchCopy(c->pops[p].mom[m], c->tmp, count(c, QMOM), c->stream);
            launchBorderSum(c->G, c->tmp, c->pops[p].mom[m], c->stream);
        }
    return checkLaunch(c, "fillMomentBorders");
}

phx_status fillIonBorders(phx_ctx* c)
{
    if (!c->patch.periodic)
        return PHX_OK;
    double* f[5] = {c->rho, c->massD, c->Vi[0], c->Vi[1], c->Vi[2]};
    for (auto* x : f)
    {
        launchCopy(x, c->tmp, count(c, QMOM), c->stream);
        launchBorderMax(c->G, c->tmp, x, c->stream);
    }
    return checkLaunch(c, "fillIonBorders");
}

void fillGhosts(phx_ctx* c, double* const* comps, int q0)
{
    if (!c->patch.periodic)
        return;
    for (int k = 0; k < 3; ++k)
        launchFillGhosts(c->G, q0 + k, comps[k], c->stream);
}

void updateIons(phx_ctx* c)
{
    PopMoments P{};
    P.npop = c->cfg.n_pop;
    for (int p = 0; p < P.npop; ++p)
    {
        P.n[p]   = c->pops[p].mom[0];
        P.rho[p] = c->pops[p].mom[1];
        for (int k = 0; k < 3; ++k)
            P.F[p][k] = c->pops[p].mom[2 + k];
        P.mass[p] = c->pops[p].mass;
    }
    launchIonTotals(P, c->rho, c->massD, vec(c->Vi), count(c, QMOM), c->stream);
}

void electrons(phx_ctx* c)
{
    launchElectrons(c->G, c->cfg.Te, cvec(c->J), cvec(c->Vi), c->rho, vec(c->Ve), c->Pe, c->stream);
}

void ohm(phx_ctx* c, int bset, int eset)
{
    launchOhm(c->G, c->cfg.eta, c->cfg.nu, c->cfg.hyper_mode, c->patch.level, c->rho, cvec(c->Ve),
              c->Pe, cvec(c->em[bset] + 3), cvec(c->J), vec(c->em[eset]), c->stream);
}

// the fast path of a periodic level: the `all` push wraps leavers into the domain in place
bool fastPath(phx_ctx* c)
{
    if (!c->patch.periodic)
        return false;
    for (int p = 0; p < c->cfg.n_pop; ++p)
        if (c->pops[p].arr[PHX_PATCH_GHOST].n || c->pops[p].arr[PHX_LEVEL_GHOST].n)
            return false;
    return true;
}

// enqueue the particle stage of SolverPPC::moveIons_ on the fast path (no host sync): the push
// deposits into the current raw set, then one kernel border-sums every population moment,
// clears the other raw set and computes the totals, and one kernel applies their border max
// a Faraday to run in the fold's launch: B of set bset, E of eset -> B of bnset (inPlace: the
// corrector's update of the state B on its physical nodes)
struct FaradaySpec
{
    int bset, eset, bnset;
    double dt;
    bool inPlace;
};

// borderMax false: the border max of the totals is left to the next electron closure
// (launchAmpereElectrons), which a captured step runs before anything reads the totals; far: the
// Faraday that follows in the step, launched with the fold
phx_status enqueueMoveIonsFast(phx_ctx* c, double dt, int mode, bool borderMax = true,
                               FaradaySpec const* far = nullptr)
{
    int const cur = c->rawCur;
    MomSetArgs M{};
    M.npop = c->cfg.n_pop;
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        PushArgs A = pushArgs(c, p, KIND_DOMAIN, mode, PHX_EM_AVG, dt, true);
        for (int m = 0; m < 5; ++m)
            A.mom[m] = c->raw[cur][p][m];
        A.periodicWrap = 1;
        A.kt           = c->timing ? c->kt + (mode == PHX_MODE_ALL ? PHX_MAX_POP : 0) + p : nullptr;
        Store const& dom = c->pops[p].arr[PHX_DOMAIN];
        if (mixed(c))
        {
            if (!A.segOn)
                return fail(c, PHX_ERR_STATE, "mixed-precision push on unsorted storage");
            launchPush32(A, dom.view32(), c->stream);
        }
        else
            launchPush(A, dom.view(), dom.view(), c->stream);
        M.mass[p] = c->pops[p].mass;
        for (int m = 0; m < 5; ++m)
        {
            M.raw[p][m]  = c->raw[cur][p][m];
            M.mom[p][m]  = c->pops[p].mom[m];
            M.zero[p][m] = c->raw[1 - cur][p][m];
        }
    }
    PHX_TRY(checkLaunch(c, "push"));
    if (far)
        launchFoldFaraday(c->G, M, c->tot5[0], c->tot5[1], vec(c->tot5 + 2), cvec(c->em[far->bset] + 3),
                          cvec(c->em[far->eset]), vec(c->em[far->bnset] + 3), far->dt, far->inPlace,
                          c->stream);
    else
        launchFoldTotals(c->G, M, c->tot5[0], c->tot5[1], vec(c->tot5 + 2), c->stream);
    if (borderMax)
    {
        double const* src[5] = {c->tot5[0], c->tot5[1], c->tot5[2], c->tot5[3], c->tot5[4]};
        double* dst[5]       = {c->rho, c->massD, c->Vi[0], c->Vi[1], c->Vi[2]};
        launchMax5(c->G, src, dst, c->stream);
    }
    c->rawCur = 1 - cur;
    return checkLaunch(c, "moment fills");
}

// SolverPPC predictor1_/predictor2_ (solver_ppc.hpp:346-428): Faraday writes Bpred with its
// ghost layers, Ampere writes J with its ghost layers
void enqueuePredictor(phx_ctx* c, double dt, int eset)
{
    if (!c->patch.periodic)
    {
        // a patch of a wider level: its ghost layers come from the caller's messenger
        launchFaraday(c->G, cvec(c->em[PHX_EM] + 3), cvec(c->em[eset]), vec(c->em[PHX_EM_PRED] + 3), dt,
                      c->stream);
        launchAmpere(c->G, cvec(c->em[PHX_EM_PRED] + 3), vec(c->J), c->stream);
        electrons(c);
        ohm(c, PHX_EM_PRED, PHX_EM_PRED);
        return;
    }
    launchFaradayFill(c->G, cvec(c->em[PHX_EM] + 3), cvec(c->em[eset]), vec(c->em[PHX_EM_PRED] + 3), dt,
                      c->stream);
    launchAmpereFill(c->G, cvec(c->em[PHX_EM_PRED] + 3), vec(c->J), c->stream);
    electrons(c);
    ohm(c, PHX_EM_PRED, PHX_EM_PRED);
}

// SolverPPC::average_ (solver_ppc.hpp:483-510) with the Eavg ghost fill
void enqueueAverage(phx_ctx* c)
{
    if (!c->patch.periodic)
    {
        for (int k = 0; k < 6; ++k)
            launchAverage(c->em[PHX_EM][k], c->em[PHX_EM_PRED][k], c->em[PHX_EM_AVG][k],
                          count(c, k < 3 ? QEX + k : QBX + k - 3), c->stream);
        return;
    }
    // the mixed-precision family's push gathers from an FP32 copy of the averaged set
    launchAverageFill(c->G, cvec(c->em[PHX_EM]), cvec(c->em[PHX_EM_PRED]), cvec(c->em[PHX_EM] + 3),
                      cvec(c->em[PHX_EM_PRED] + 3), vec(c->em[PHX_EM_AVG]), vec(c->em[PHX_EM_AVG] + 3),
                      c->stream, mixed(c) ? c->em32 : nullptr);
}

// SolverPPC::corrector_ (solver_ppc.hpp:433-479): in-place Faraday, then the B ghost fill;
// Ampere and Ohm write their ghost layers themselves
void enqueueCorrector(phx_ctx* c, double dt)
{
    launchFaraday(c->G, cvec(c->em[PHX_EM] + 3), cvec(c->em[PHX_EM_AVG]), vec(c->em[PHX_EM] + 3), dt,
                  c->stream);
    if (!c->patch.periodic)
    {
        launchAmpere(c->G, cvec(c->em[PHX_EM] + 3), vec(c->J), c->stream);
        electrons(c);
        ohm(c, PHX_EM, PHX_EM);
        return;
    }
    launchFill3(c->G, QBX, vec(c->em[PHX_EM] + 3), c->stream);
    launchAmpereFill(c->G, cvec(c->em[PHX_EM] + 3), vec(c->J), c->stream);
    electrons(c);
    launchOhm(c->G, c->cfg.eta, c->cfg.nu, c->cfg.hyper_mode, c->patch.level, c->rho, cvec(c->Ve),
              c->Pe, cvec(c->em[PHX_EM] + 3), cvec(c->J), vec(c->em[PHX_EM]), c->stream, 1);
}

// Ampere on B (em set bset, ghost layers filled) with J's ghost layers and the electron closure,
// in one launch
void ampereElectronsAfterFold(phx_ctx* c, int bset, bool borderMax)
{
    double const* src[5] = {c->tot5[0], c->tot5[1], c->tot5[2], c->tot5[3], c->tot5[4]};
    double* dst[5]       = {c->rho, c->massD, c->Vi[0], c->Vi[1], c->Vi[2]};
    launchAmpereElectrons(c->G, c->cfg.Te, cvec(c->em[bset] + 3), vec(c->J), src, dst, cvec(c->Vi),
                          c->rho, vec(c->Ve), c->Pe, borderMax, c->stream);
}

// predictor + average of a periodic level: Faraday with its ghost layers (unless the previous
// fold's launch ran it), Ampere with its ghost layers and the electron closure (with the pending
// border max) in one launch, then Ohm and the average in one kernel
void enqueuePredictorAverageFast(phx_ctx* c, double dt, int eset, bool borderMax, bool faraday)
{
    if (faraday)
        launchFaradayFill(c->G, cvec(c->em[PHX_EM] + 3), cvec(c->em[eset]), vec(c->em[PHX_EM_PRED] + 3),
                          dt, c->stream);
    ampereElectronsAfterFold(c, PHX_EM_PRED, borderMax);
    launchOhmAverage(c->G, c->cfg.eta, c->cfg.nu, c->cfg.hyper_mode, c->patch.level, c->rho,
                     cvec(c->Ve), c->Pe, cvec(c->em[PHX_EM_PRED] + 3), cvec(c->J),
                     vec(c->em[PHX_EM_PRED]), cvec(c->em[PHX_EM]), cvec(c->em[PHX_EM] + 3),
                     vec(c->em[PHX_EM_AVG]), vec(c->em[PHX_EM_AVG] + 3), c->stream,
                     mixed(c) ? c->em32 : nullptr);
}

// SolverPPC::corrector_ of a periodic level after the in-place Faraday (run with the last fold,
// which also fills B's ghost layers): Ampere and the electron closure with the pending border max
// in one launch, then Ohm with its ghost layers
void enqueueCorrectorFast(phx_ctx* c)
{
    ampereElectronsAfterFold(c, PHX_EM, true);
    launchOhm(c->G, c->cfg.eta, c->cfg.nu, c->cfg.hyper_mode, c->patch.level, c->rho, cvec(c->Ve),
              c->Pe, cvec(c->em[PHX_EM] + 3), cvec(c->J), vec(c->em[PHX_EM]), c->stream, 1);
}

// one PPC step of a periodic level, as captured: each moveIons leaves the border max of the
// totals to the electron closure that follows it (the first reader of the totals), and its fold
// shares a launch with the next Faraday (predictor 2's, then the corrector's in-place one)
phx_status enqueueAdvanceFast(phx_ctx* c, double dt)
{
    enqueuePredictorAverageFast(c, dt, PHX_EM, false, true);
    FaradaySpec const pred2{PHX_EM, PHX_EM_AVG, PHX_EM_PRED, dt, false};
    PHX_TRY(enqueueMoveIonsFast(c, dt, PHX_MODE_DOMAIN_ONLY, false, &pred2));
    enqueuePredictorAverageFast(c, dt, PHX_EM_AVG, true, false);
    FaradaySpec const corr{PHX_EM, PHX_EM_AVG, PHX_EM, dt, true};
    PHX_TRY(enqueueMoveIonsFast(c, dt, PHX_MODE_ALL, false, &corr));
    enqueueCorrectorFast(c);
    return checkLaunch(c, "advance");
}

// A periodic level in 3D runs SolverPPC::advanceLevel staged from the kernels of the drop-in
// entry points (linear-storage push, Faraday/Ampere/Ohm on the physical nodes followed by the
// periodic ghost fill, the border sum and max of the moments, the ion totals), eagerly on the
// context stream: the captured fused step (segmented push, fused field kernels) stays off in 3D,
// one GPU run of it having faulted without a cause shown.
bool staged3d(phx_ctx const* c) { return c->cfg.dim == 3 && c->patch.periodic; }

// the push (or, pushOn 0, the deposit) of population p's domain array on the staged 3D path:
// cell-segmented storage (after a sort) with one workgroup per segment, otherwise the linear
// brick-tiled kernels; FP64 or mixed precision
void enqueuePushStaged3d(phx_ctx* c, int p, PushArgs& A)
{
    Pop const& P     = c->pops[p];
    Store const& dom = P.arr[PHX_DOMAIN];
    A.segOn          = 0;
    if (dom.seg)
    {
        A.sg   = c->sg;
        A.st   = P.seg;
        A.tile = segwgTileShape(c->cfg.dim, c->cfg.interp_order, c->sg, mixed(c) ? 4 : 5,
                                c->tuning.tile_row_pad);
        noteTileFallback(c, A.tile);
        A.rhoFromN = !mixed(c) && P.hasCharge && P.uniformCharge ? 1 : 0;
        if (mixed(c))
            launchPush32SegWG(A, dom.view32(), c->stream);
        else
            launchPushSegWG(A, dom.view(), c->stream);
        return;
    }
    if (mixed(c))
    {
        A.tile = brickTileShape32(c->cfg.interp_order);
        launchPush32Linear(A, dom.view32(), c->stream);
    }
    else if (A.pushOn)
        launchPush(A, dom.view(), dom.view(), c->stream);
    else
        launchDeposit(A, dom.view(), c->stream);
}

// SolverPPC::moveIons_ (solver_ppc.hpp:537-598) of a 3D periodic level: the push deposits into
// the population moments and wraps leavers back through the periodic image in place
// (fillIonGhostParticles), then fillFluxBorders/fillDensityBorders, updateIons, fillIonBorders
phx_status moveIonsStaged3d(phx_ctx* c, double dt, int mode)
{
    zeroMoments(c);
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        PushArgs A     = pushArgs(c, p, KIND_DOMAIN, mode, PHX_EM_AVG, dt, false);
        A.periodicWrap = 1;
        Store const& dom = c->pops[p].arr[PHX_DOMAIN];
        phx_ctx::EventPair* ev = nullptr;
        if (c->timing)
        {
            if (c->eventsUsed == c->events.size())
            {
                phx_ctx::EventPair e;
                PHX_HIP(c, hipEventCreate(&e.a));
                PHX_HIP(c, hipEventCreate(&e.b));
                c->events.push_back(e);
            }
            ev       = &c->events[c->eventsUsed++];
            ev->mode = mode;
            ev->pop  = p;
            PHX_HIP(c, hipEventRecord(ev->a, c->stream));
        }
        enqueuePushStaged3d(c, p, A);
        if (ev)
            PHX_HIP(c, hipEventRecord(ev->b, c->stream));
    }
    PHX_TRY(checkLaunch(c, "push"));
    PHX_TRY(fillMomentBorders(c));
    updateIons(c);
    return fillIonBorders(c);
}

// predictor1_/predictor2_ (solver_ppc.hpp:346-428) of a 3D periodic level: Faraday, then the B
// ghost fill; Ampere, then the J ghost fill; the electron closure; Ohm on the physical nodes (the
// average that follows takes Eavg's ghost nodes from their periodic sources)
void predictorStaged3d(phx_ctx* c, double dt, int eset)
{
    launchFaraday(c->G, cvec(c->em[PHX_EM] + 3), cvec(c->em[eset]), vec(c->em[PHX_EM_PRED] + 3), dt,
                  c->stream);
    fillGhosts(c, c->em[PHX_EM_PRED] + 3, QBX);
    launchAmpere(c->G, cvec(c->em[PHX_EM_PRED] + 3), vec(c->J), c->stream);
    fillGhosts(c, c->J, QJX);
    electrons(c);
    ohm(c, PHX_EM_PRED, PHX_EM_PRED);
}

// corrector_ (solver_ppc.hpp:433-479) of a 3D periodic level, each stage followed by its fill
void correctorStaged3d(phx_ctx* c, double dt)
{
    launchFaraday(c->G, cvec(c->em[PHX_EM] + 3), cvec(c->em[PHX_EM_AVG]), vec(c->em[PHX_EM] + 3), dt,
                  c->stream);
    fillGhosts(c, c->em[PHX_EM] + 3, QBX);
    launchAmpere(c->G, cvec(c->em[PHX_EM] + 3), vec(c->J), c->stream);
    fillGhosts(c, c->J, QJX);
    electrons(c);
    ohm(c, PHX_EM, PHX_EM);
    fillGhosts(c, c->em[PHX_EM], QEX);
}

phx_status advanceStaged3d(phx_ctx* c, double dt)
{
    predictorStaged3d(c, dt, PHX_EM);
    enqueueAverage(c);
    PHX_TRY(moveIonsStaged3d(c, dt, PHX_MODE_DOMAIN_ONLY));
    predictorStaged3d(c, dt, PHX_EM_AVG);
    enqueueAverage(c);
    PHX_TRY(moveIonsStaged3d(c, dt, PHX_MODE_ALL));
    correctorStaged3d(c, dt);
    return checkLaunch(c, "advance");
}

phx_status validPop(phx_ctx* c, int pop, int kind)
{
    if (!c)
        return PHX_ERR_ARG;
    if (pop < 0 || pop >= c->cfg.n_pop || kind < 0 || kind > PHX_LEVEL_GHOST_NEW)
        return fail(c, PHX_ERR_ARG, "bad population or particle array kind");
    return PHX_OK;
}

// storage order inside each chunk of the push kernel: slot k*BLOCK + t (t = wave*64 + lane)
// holds the particle (r*PPT + k) of the caller's order with r = lane*NW + wave, so a thread
// owns PPT consecutive particles and the lanes of a wavefront own runs spread across the chunk;
// a last partial chunk keeps the caller's order
std::vector<size_t> chunkInterleave(size_t n, int dim, int order)
{
    size_t const C   = static_cast<size_t>(chunkSize(dim, order));
    size_t const BT  = static_cast<size_t>(blockThreads());
    size_t const PPT = C / BT, NWv = BT / 64;
    std::vector<size_t> ord(n);
    for (size_t c0 = 0; c0 < n; c0 += C)
    {
        size_t const m = std::min(C, n - c0);
        for (size_t s = 0; s < m; ++s)
        {
            if (m < C)
            {
                ord[c0 + s] = c0 + s;
                continue;
            }
            size_t const k = s / BT, t = s % BT;
            size_t const wave = t / 64, lane = t % 64;
            size_t const r    = lane * NWv + wave;
            ord[c0 + s]       = c0 + r * PPT + k;
        }
    }
    return ord;
}

// 3D particles (cells ic[3 n], AMR) in brick order: by brick of BRICK^3 cells on the grid of
// bricks from the patch's lower cell (row-major), then by cell inside the brick (row-major), the
// caller's order kept among particles of one cell (a counting sort). Particles beyond the bricks
// that cover the patch and its ghost layers go last, in the caller's order. With 64 particles per
// cell every chunk of the linear push is one brick, and at each step the 64 lanes of a wavefront
// hold particles of 64 different cells.
std::vector<size_t> brickOrder(Geom const& G, int32_t const* ic, size_t n)
{
    int nb[3];
    for (int d = 0; d < 3; ++d)
        nb[d] = (G.n[d] + BRICK - 1) / BRICK + 2; // one brick of ghost cells on each side
    size_t const bricks = static_cast<size_t>(nb[0]) * nb[1] * nb[2];
    size_t const keys   = bricks * BRICK * BRICK * BRICK + 1; // the last key: beyond the bricks
    auto key = [&](size_t i) -> size_t {
        size_t b = 0, cib = 0;
        for (int d = 0; d < 3; ++d)
        {
            int const r  = ic[i * 3 + d] - G.lower[d];
            int const bd = (r >= 0 ? r / BRICK : -((BRICK - 1 - r) / BRICK)) + 1;
            if (bd < 0 || bd >= nb[d])
                return keys - 1;
            b   = b * nb[d] + static_cast<size_t>(bd);
            cib = cib * BRICK + static_cast<size_t>(r - (bd - 1) * BRICK);
        }
        return b * BRICK * BRICK * BRICK + cib;
    };
    std::vector<size_t> count(keys + 1, 0);
    std::vector<uint32_t> k(n);
    for (size_t i = 0; i < n; ++i)
    {
        size_t const kk = key(i);
        k[i]            = static_cast<uint32_t>(kk);
        ++count[kk + 1];
    }
    for (size_t j = 1; j <= keys; ++j)
        count[j] += count[j - 1];
    std::vector<size_t> ord(n);
    for (size_t i = 0; i < n; ++i)
        ord[count[k[i]]++] = i;
    return ord;
}

// FP64 host values into a device array of the store's precision
phx_status putReals(phx_ctx* c, Store const& s, void* dst, size_t offset, std::vector<double> const& src)
{
    size_t const n = src.size();
    if (s.real == 8)
    {
        PHX_HIP(c, hipMemcpy(static_cast<double*>(dst) + offset, src.data(), n * 8, hipMemcpyHostToDevice));
        return PHX_OK;
    }
    std::vector<float> f(n);
    for (size_t i = 0; i < n; ++i)
        f[i] = static_cast<float>(src[i]);
    PHX_HIP(c, hipMemcpy(static_cast<float*>(dst) + offset, f.data(), n * 4, hipMemcpyHostToDevice));
    return PHX_OK;
}

phx_status getReals(phx_ctx* c, Store const& s, void const* src, std::vector<double>& dst)
{
    size_t const n = dst.size();
    if (s.real == 8)
    {
        PHX_HIP(c, hipMemcpy(dst.data(), src, n * 8, hipMemcpyDeviceToHost));
        return PHX_OK;
    }
    std::vector<float> f(n);
    PHX_HIP(c, hipMemcpy(f.data(), src, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i)
        dst[i] = static_cast<double>(f[i]);
    return PHX_OK;
}

phx_status upload(phx_ctx* c, Pop& P, Store& s, size_t n, int32_t const* ic, double const* de,
                  double const* v, double const* w, double const* q, size_t offset)
{
    int const D = c->cfg.dim;
    if (s.real == 8 && n)
    {
        // FP64 family: note whether the population keeps a single charge (the segmented push
        // then derives the charge density from n)
        if (!P.hasCharge)
        {
            P.charge        = q[0];
            P.hasCharge     = true;
            P.uniformCharge = true;
        }
        for (size_t i = 0; i < n && P.uniformCharge; ++i)
            P.uniformCharge = q[i] == P.charge;
    }
    if (s.real == 4 && n)
    {
        // mixed precision: one charge per population, kept on the host
        double const q0 = P.hasCharge && offset > 0 ? P.charge : q[0];
        for (size_t i = 0; i < n; ++i)
            if (q[i] != q0)
                return fail(c, PHX_ERR_ARG, "the mixed-precision family keeps one charge per population");
        P.charge    = q0;
        P.hasCharge = true;
    }
    PHX_TRY(storeReserve(c, s, offset + n));
    s.seg = false;
    // the copies below run on the null stream, which does not wait for the context's
    // non-blocking stream: nothing queued there may still use the array
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    // appended particles start a fresh chunk alignment only when the array was empty; in 3D the
    // particles of an empty array are put in brick order first (the linear push's brick tiles)
    std::vector<size_t> ord = offset == 0 ? chunkInterleave(n, D, c->cfg.interp_order)
                                          : chunkInterleave(0, D, c->cfg.interp_order);
    if (offset == 0 && D == 3 && n)
    {
        std::vector<size_t> const bo = brickOrder(c->G, ic, n);
        for (size_t i = 0; i < n; ++i)
            ord[i] = bo[ord[i]];
    }
    std::vector<size_t> identity;
    if (offset != 0)
    {
        identity.resize(n);
        for (size_t i = 0; i < n; ++i)
            identity[i] = i;
    }
    std::vector<size_t> const& o = offset == 0 ? ord : identity;
    std::vector<int32_t> ib(n);
    std::vector<double> db(n);
    for (int d = 0; d < D; ++d)
    {
        for (size_t i = 0; i < n; ++i)
            ib[i] = ic[o[i] * D + d];
        PHX_HIP(c, hipMemcpy(s.ic[d] + offset, ib.data(), n * 4, hipMemcpyHostToDevice));
        for (size_t i = 0; i < n; ++i)
        {
            db[i] = de[o[i] * D + d];
            // an FP32 delta must stay below 1: the largest float under 1 for a delta that rounds up
            if (s.real == 4 && static_cast<float>(db[i]) >= 1.f)
                db[i] = static_cast<double>(std::nextafter(1.f, 0.f));
        }
        PHX_TRY(putReals(c, s, s.de[d], offset, db));
    }
    for (int d = 0; d < 3; ++d)
    {
        for (size_t i = 0; i < n; ++i)
            db[i] = v[o[i] * 3 + d];
        PHX_TRY(putReals(c, s, s.v[d], offset, db));
    }
    for (size_t i = 0; i < n; ++i)
        db[i] = w[o[i]];
    PHX_TRY(putReals(c, s, s.w, offset, db));
    if (s.q)
    {
        for (size_t i = 0; i < n; ++i)
            db[i] = q[o[i]];
        PHX_TRY(putReals(c, s, s.q, offset, db));
    }
    s.n = offset + n;
    return PHX_OK;
}

// live (not lost) particles of a store, downloaded to host SoA buffers (FP64)
phx_status download(phx_ctx* c, Pop const& P, Store const& s, std::vector<int32_t>& ic,
                    std::vector<double>& de, std::vector<double>& v, std::vector<double>& w,
                    std::vector<double>& q)
{
    int const D  = c->cfg.dim;
    size_t const n = s.n;
    std::vector<int32_t> icd(n * D);
    std::vector<double> ded(n * D), vd(n * 3), wd(n), qd(n, P.charge);
    std::vector<int32_t> ib(n);
    std::vector<double> db(n);
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    for (int d = 0; d < D; ++d)
    {
        PHX_HIP(c, hipMemcpy(ib.data(), s.ic[d], n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i)
            icd[i * D + d] = ib[i];
        PHX_TRY(getReals(c, s, s.de[d], db));
        for (size_t i = 0; i < n; ++i)
            ded[i * D + d] = db[i];
    }
    for (int d = 0; d < 3; ++d)
    {
        PHX_TRY(getReals(c, s, s.v[d], db));
        for (size_t i = 0; i < n; ++i)
            vd[i * 3 + d] = db[i];
    }
    PHX_TRY(getReals(c, s, s.w, wd));
    if (s.q)
        PHX_TRY(getReals(c, s, s.q, qd));
    ic.clear();
    de.clear();
    v.clear();
    w.clear();
    q.clear();
    for (size_t i = 0; i < n; ++i)
    {
        if (icd[i * D] == DEAD_CELL)
            continue;
        for (int d = 0; d < D; ++d)
        {
            ic.push_back(icd[i * D + d]);
            de.push_back(ded[i * D + d]);
        }
        for (int d = 0; d < 3; ++d)
            v.push_back(vd[i * 3 + d]);
        w.push_back(wd[i]);
        q.push_back(qd[i]);
    }
    return PHX_OK;
}

size_t recordSize(int dim) { return dim == 1 ? 56 : (dim == 2 ? 64 : 80); }

} // namespace

// =============================================================================================
// C ABI
// =============================================================================================
namespace
{
// counting sort of every population's domain array into cell-segmented storage: scattered into
// the population's sort buffer, then copied back, so that the domain array keeps its address and
// the captured step its graph. The two swap when the array is linear (the first sort after an
// upload) or too small for the sorted slots, and always on the staged 3D path (no captured step).
// Lost particles are dropped.
phx_status sortDomain(phx_ctx* c)
{
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        Pop& P     = c->pops[p];
        Store& dom = P.arr[PHX_DOMAIN];
        Store& buf = c->sortBuf[p];
        if (dom.n >= 0xFFFFFFFFull)
            return fail(c, PHX_ERR_ARG, "particle sort: more than 2^32 - 2 particles in one array");
        if (c->rankCap < dom.n)
        {
            devFree(c, c->sortRank);
            c->sortRank = nullptr;
            c->rankCap  = 0;
            size_t const want = dom.n + dom.n / 2 + 4096;
            PHX_TRY(devAlloc(c, (void**)&c->sortRank, want * 4));
            c->rankCap = want;
        }
        if (mixed(c))
            launchSegCount(c->sg, dom.view32(), dom.n, c->segWork, c->sortRank, P.seg, c->stream);
        else
            launchSegCount(c->sg, dom.view(), dom.n, c->segWork, c->sortRank, P.seg, c->stream);
        PHX_TRY(checkLaunch(c, "particle sort (count)"));
        unsigned long long slots = 0;
        PHX_HIP(c, hipStreamSynchronize(c->stream));
        PHX_HIP(c, hipMemcpy(&slots, P.seg.off + c->sg.nsegs + 1, 8, hipMemcpyDeviceToHost));
        if (buf.cap < slots || buf.cap < dom.cap)
        {
            // room for segments that gain particles between sorts (each rounds its particle
            // count up to whole rows), so that later sorts reuse it
            size_t const want = slots + slots / 2 + 4096;
            storeFree(c, buf);
            PHX_TRY(storeAlloc(c, buf, std::max(want, dom.cap)));
        }
        if (c->permCap < buf.cap)
        {
            devFree(c, c->sortPerm);
            c->sortPerm = nullptr;
            c->permCap  = 0;
            PHX_TRY(devAlloc(c, (void**)&c->sortPerm, buf.cap * 4));
            c->permCap = buf.cap;
        }
        // work items for the larger of the two storages the sort alternates between
        unsigned int const need = segItemsCap(c->sg, std::max(buf.cap, dom.cap));
        if (need > P.itemsAlloc)
        {
            devFree(c, P.seg.items);
            P.seg.items = nullptr;
            PHX_TRY(devAlloc(c, (void**)&P.seg.items, static_cast<size_t>(need) * 8));
            P.itemsAlloc = need;
        }
        if (mixed(c))
            launchSegScatter(c->sg, dom.view32(), buf.view32(), dom.n, slots, c->segWork, c->sortRank,
                             c->sortPerm, P.seg, c->cfg.dim, c->stream);
        else
            launchSegScatter(c->sg, dom.view(), buf.view(), dom.n, slots, c->segWork, c->sortRank,
                             c->sortPerm, P.seg, c->cfg.dim, c->stream);
        PHX_TRY(checkLaunch(c, "particle sort (scatter)"));
        // a linear array (an upload, sized for it) moves into the buffer with its headroom; the
        // staged 3D path captures no step, so its two storages always swap
        if (dom.seg && slots <= dom.cap && !staged3d(c))
        {
            if (mixed(c))
                launchSegCopy(buf.view32(), dom.view32(), slots, c->cfg.dim, c->stream);
            else
                launchSegCopy(buf.view(), dom.view(), slots, c->cfg.dim, c->stream);
            PHX_TRY(checkLaunch(c, "particle sort (copy)"));
            dom.n   = slots;
            dom.seg = true;
            continue;
        }
        buf.n   = slots;
        buf.seg = true;
        std::swap(dom, buf);
        buf.seg = false;
        buf.n   = 0;
        if (buf.cap < dom.cap)
        {
            // the two storages the sorts alternate between keep one capacity: no reallocation
            // (and no new captured step) at later sorts
            PHX_HIP(c, hipStreamSynchronize(c->stream));
            storeFree(c, buf);
            PHX_TRY(storeAlloc(c, buf, dom.cap));
        }
    }
    c->stepsSinceSort = 0;
    return PHX_OK;
}

// the mixed-precision push runs on segmented storage only: sort what an upload left linear
phx_status sortIfLinear(phx_ctx* c)
{
    if (!mixed(c))
        return PHX_OK;
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        Pop const& P = c->pops[p];
        Store const& s = P.arr[PHX_DOMAIN];
        if (!s.seg || segItemsCap(c->sg, s.cap) > P.itemsAlloc)
            return sortDomain(c);
    }
    return PHX_OK;
}

// the particle storage a captured step addresses: pointers and counts of every array
std::vector<unsigned long long> storageKey(phx_ctx const* c)
{
    std::vector<unsigned long long> k;
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        // the segment tables a segmented push reads (reallocated when the storage grows)
        k.push_back(reinterpret_cast<unsigned long long>(c->pops[p].seg.items));
        // the population's charge, which the push takes as an argument when it is uniform
        double const q = c->pops[p].charge;
        unsigned long long qb;
        std::memcpy(&qb, &q, 8);
        k.push_back(qb);
        k.push_back(c->pops[p].uniformCharge ? 1ull : 0ull);
        k.push_back(c->timing ? 1ull : 0ull);
        for (Store const& s : c->pops[p].arr)
        {
            // segmented storage: the kernels read their extent from the segment tables, so a
            // sort that changes the slot count keeps the captured step
            k.push_back(reinterpret_cast<unsigned long long>(s.ic[0]));
            k.push_back(reinterpret_cast<unsigned long long>(s.w));
            k.push_back(s.cap);
            k.push_back(s.seg ? ~0ull : s.n);
        }
    }
    return k;
}

// the captured PPC step for dt on the current storage (captured on first use; two are kept, so
// that the two storages a periodic sort alternates between each keep theirs)
phx_status stepGraph(phx_ctx* c, double dt, phx_ctx::GraphEntry** out)
{
    std::vector<unsigned long long> const key = storageKey(c);
    for (auto& g : c->graphs)
        if (g.exec && g.dt == dt && g.rawCur == c->rawCur && g.key == key)
        {
            *out = &g;
            return PHX_OK;
        }
    phx_ctx::GraphEntry& g = c->graphs[c->graphNext];
    c->graphNext           = (c->graphNext + 1) % 2;
    if (g.exec)
    {
        PHX_HIP(c, hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    hipGraph_t graph = nullptr;
    int const cur0   = c->rawCur;
    PHX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    phx_status s = enqueueAdvanceFast(c, dt);
    c->rawCur    = cur0; // a step uses both raw sets and ends on the one it started with
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (s != PHX_OK)
    {
        if (graph)
            (void)hipGraphDestroy(graph);
        return s;
    }
    if (e != hipSuccess)
        return fail(c, PHX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
    {
        g.exec = nullptr;
        return fail(c, PHX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    g.dt     = dt;
    g.rawCur = c->rawCur;
    g.key    = key;
    *out     = &g;
    return PHX_OK;
}
} // namespace

namespace
{
// the level's patch table: each patch cut into segments (numbered patch after patch), the owner
// map of the level box, the segment -> patch table; uploaded once, read by the sort and the push
phx_status buildPatchTable(phx_ctx* c, int const* plo, int const* phi, int np)
{
    int const D = c->cfg.dim;
    SegGeom& S  = c->sg;
    devFree(c, const_cast<SegPatch*>(S.patch));
    devFree(c, const_cast<int*>(S.owner));
    devFree(c, const_cast<unsigned int*>(S.segPatch));
    S             = SegGeom{};
    S.dim         = D;
    S.npatch      = np;
    S.rowsPerItem = c->tuning.rows_per_item;
    S.rowPad      = c->tuning.tile_row_pad;
    size_t boxCells = 1;
    for (int d = 0; d < 3; ++d)
    {
        S.lo[d] = d < D ? c->boxes.dom_lo[d] : 0;
        S.n[d]  = d < D ? c->boxes.dom_hi[d] - c->boxes.dom_lo[d] + 1 : 1;
        boxCells *= static_cast<size_t>(S.n[d]);
    }
    std::vector<int> owner(boxCells, -1);
    std::vector<unsigned int> segOf;
    c->patchBox.clear();
    c->segPatch.clear();
    unsigned int segBase = 0, cellBase = 0;
    for (int p = 0; p < np; ++p)
    {
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d)
        {
            lo[d] = plo[3 * p + d];
            hi[d] = phi[3 * p + d];
            if (hi[d] < lo[d] || lo[d] < S.lo[d] || hi[d] >= S.lo[d] + S.n[d])
                return fail(c, PHX_ERR_ARG, "patch " + std::to_string(p) + " is empty or leaves the level box");
        }
        SegPatch P = segPatchOf(D, lo, hi, c->tuning.segment_shape);
        P.segBase  = segBase;
        P.cellBase = cellBase;
        unsigned int cells = 1, segs = 1;
        for (int d = 0; d < D; ++d)
        {
            cells *= static_cast<unsigned int>(P.n[d]);
            segs *= static_cast<unsigned int>(P.nseg[d]);
            S.maxSeg[d] = std::max(S.maxSeg[d], P.seg[d]);
        }
        // owner map: every cell of the patch, row-major over the level box
        for (unsigned int k = 0; k < cells; ++k)
        {
            unsigned int r = k;
            size_t lin = 0;
            int cc[3] = {0, 0, 0};
            for (int d = D - 1; d >= 0; --d)
            {
                cc[d] = static_cast<int>(r % static_cast<unsigned int>(P.n[d]));
                r /= static_cast<unsigned int>(P.n[d]);
            }
            for (int d = 0; d < D; ++d)
                lin = lin * static_cast<size_t>(S.n[d]) + static_cast<size_t>(lo[d] + cc[d] - S.lo[d]);
            if (owner[lin] >= 0)
                return fail(c, PHX_ERR_ARG, "patches " + std::to_string(owner[lin]) + " and "
                                                + std::to_string(p) + " overlap");
            owner[lin] = p;
        }
        segOf.insert(segOf.end(), segs, static_cast<unsigned int>(p));
        segBase += segs;
        cellBase += cells;
        c->segPatch.push_back(P);
        c->patchBox.push_back({lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
    }
    for (int d = D; d < 3; ++d)
        S.maxSeg[d] = 1;
    S.nsegs  = static_cast<int>(segBase);
    S.ncells = static_cast<int>(cellBase);
    SegPatch* dP = nullptr;
    int* dO      = nullptr;
    unsigned int* dS = nullptr;
    PHX_TRY(devAlloc(c, (void**)&dP, c->segPatch.size() * sizeof(SegPatch)));
    PHX_TRY(devAlloc(c, (void**)&dO, owner.size() * sizeof(int)));
    PHX_TRY(devAlloc(c, (void**)&dS, std::max<size_t>(1, segOf.size()) * sizeof(unsigned int)));
    PHX_HIP(c, hipMemcpy(dP, c->segPatch.data(), c->segPatch.size() * sizeof(SegPatch), hipMemcpyHostToDevice));
    PHX_HIP(c, hipMemcpy(dO, owner.data(), owner.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!segOf.empty())
        PHX_HIP(c, hipMemcpy(dS, segOf.data(), segOf.size() * sizeof(unsigned int), hipMemcpyHostToDevice));
    S.patch    = dP;
    S.owner    = dO;
    S.segPatch = dS;
    S.first    = c->segPatch.front();
    return PHX_OK;
}

// the sort's work arrays and every population's segment tables, sized for the patch table
phx_status allocSegTables(phx_ctx* c)
{
    size_t const nsegs = static_cast<size_t>(c->sg.nsegs);
    devFree(c, c->segWork);
    c->segWork = nullptr;
    PHX_TRY(devAlloc(c, (void**)&c->segWork, (static_cast<size_t>(c->sg.ncells) + 1 + 2 * (nsegs + 1)) * 4));
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        SegTables& T = c->pops[p].seg;
        devFree(c, T.rows);
        devFree(c, T.off);
        devFree(c, T.nItems);
        T.rows   = nullptr;
        T.off    = nullptr;
        T.nItems = nullptr;
        PHX_TRY(devAlloc(c, (void**)&T.rows, (nsegs + 1) * 4));
        PHX_TRY(devAlloc(c, (void**)&T.off, (nsegs + 2) * 8));
        PHX_TRY(devAlloc(c, (void**)&T.nItems, 64));
    }
    return PHX_OK;
}

// the launch shapes of this build before any tuning: the autotuner's winners on the 2D Harris
// workload (segments of 2x32 cells, at most 32 rows per work item, tile rows padded to 16 nodes,
// a sort every 30 PPC steps); runs of 64 cells in 1D, 4x4x4 blocks in 3D (the staged 3D path's
// one-workgroup-per-segment push gives each lane of a wavefront one cell of the block)
phx_tuning defaultTuning(int dim, int order)
{
    phx_tuning t{};
    static int const pref[3][3] = {{64, 1, 1}, {2, 32, 1}, {4, 4, 4}};
    for (int d = 0; d < 3; ++d)
        t.segment_shape[d] = pref[dim - 1][d];
    t.rows_per_item = PHX_SEG_ROWS;
    // 3D: the segment tile's pitches padded to a multiple of 4 nodes at orders 1 and 3 (every
    // particle's stencil starts at a fixed node of its cell, so the 64 lanes of one LDS atomic
    // fall on distinct bank pairs); unpadded at order 2 (starts shift with the delta at random,
    // and the larger tile would halve the workgroups per CU)
    t.tile_row_pad  = dim == 3 ? (order == 2 ? 1 : 4) : PHX_TILE_ROW_PAD;
    t.sort_interval = 30;
    t.tile_copies   = dim == 1 ? PHX_TILE_COPIES_1D : (dim == 2 ? PHX_TILE_COPIES_2D : PHX_TILE_COPIES_3D);
    t.deposit_window = 0;
    return t;
}

// a level of n_patches patches (AMR domain boxes) inside the box of `level` (its mesh, level
// number, periodicity and selection boxes); one patch covering the level box is a patch level
phx_status createLevel(phx_config const* cfg, phx_patch const* patch, int const* plo, int const* phi,
                       int np, size_t particle_capacity, phx_ctx** out)
{
    if (!cfg || !patch || !out || np < 1)
        return PHX_ERR_ARG;
    *out = nullptr;
    if (cfg->dim < 1 || cfg->dim > 3 || cfg->interp_order < 1 || cfg->interp_order > 3
        || cfg->n_pop < 1 || cfg->n_pop > PHX_MAX_POP
        || (cfg->precision != PHX_PRECISION_FP64 && cfg->precision != PHX_PRECISION_FP32))
        return PHX_ERR_ARG;
    // the mixed-precision family runs the resident path of a periodic level
    if (cfg->precision == PHX_PRECISION_FP32 && !patch->periodic)
        return PHX_ERR_ARG;
    if (patch->n_nonlevel_ghost_boxes < 0 || patch->n_nonlevel_ghost_boxes > MAX_NLGB)
        return PHX_ERR_ARG;
    auto* c  = new phx_ctx;
    c->cfg   = *cfg;
    c->patch = *patch;
    int const D = cfg->dim, O = cfg->interp_order;
    Geom& G = c->G;
    G.dim   = D;
    G.order = O;
    G.g     = O == 1 ? 2 : 4;
    G.pg    = O == 1 ? 1 : 2;
    for (int d = 0; d < 3; ++d)
    {
        G.n[d]     = d < D ? patch->upper[d] - patch->lower[d] + 1 : 1;
        G.lower[d] = d < D ? patch->lower[d] : 0;
        G.dx[d]    = d < D ? patch->dx[d] : 1.;
        G.inv[d]   = 1. / G.dx[d];
        if (d < D && patch->periodic && G.n[d] <= 2 * G.g)
        {
            delete c;
            return PHX_ERR_ARG; // a periodic patch must be wider than both ghost layers
        }
    }
    for (int q = 0; q < NQ; ++q)
        for (int d = 0; d < 3; ++d)
            G.shape[q][d] = d < D ? G.n[d] + 1 + 2 * G.g - centering(q, d) : 1;
    // local cells whose every E/B stencil fits its array: start = cell - shift >= 0 and
    // start + order < shape for all shifts the centering allows
    for (int d = 0; d < 3; ++d)
    {
        G.safeLo[d] = 0;
        G.safeHi[d] = 0x3fffffff;
        for (int q = QBX; q <= QEZ && d < D; ++q)
        {
            int const cq = centering(q, d);
            int const mx = O == 1 ? maxShift<1>(cq) : (O == 2 ? maxShift<2>(cq) : maxShift<3>(cq));
            int const mn = O == 1 ? minShift<1>(cq) : (O == 2 ? minShift<2>(cq) : minShift<3>(cq));
            G.safeLo[d]  = std::max(G.safeLo[d], mx);
            G.safeHi[d]  = std::min(G.safeHi[d], G.shape[q][d] - 1 - O + mn);
        }
    }

    Boxes& B = c->boxes;
    for (int d = 0; d < 3; ++d)
    {
        B.dom_lo[d] = d < D ? patch->lower[d] : 0;
        B.dom_hi[d] = d < D ? patch->upper[d] : 0;
        B.gh_lo[d]  = d < D ? patch->lower[d] - G.pg : 0;
        B.gh_hi[d]  = d < D ? patch->upper[d] + G.pg : 0;
    }
    if (patch->n_nonlevel_ghost_boxes == 0)
    {
        B.nnlgb = 2;
        for (int d = 0; d < 3; ++d)
        {
            B.nlgb_lo[0][d] = B.dom_lo[d];
            B.nlgb_hi[0][d] = B.dom_hi[d];
            B.nlgb_lo[1][d] = B.gh_lo[d];
            B.nlgb_hi[1][d] = B.gh_hi[d];
        }
    }
    else
    {
        B.nnlgb = patch->n_nonlevel_ghost_boxes;
        for (int b = 0; b < B.nnlgb; ++b)
            for (int d = 0; d < 3; ++d)
            {
                B.nlgb_lo[b][d] = d < D ? patch->nonlevel_ghost_lower[b][d] : 0;
                B.nlgb_hi[b][d] = d < D ? patch->nonlevel_ghost_upper[b][d] : 0;
            }
    }

    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
    {
        delete c;
        return PHX_ERR_HIP;
    }
    // the device address of the no-allocation gather's scratch array on this device (looked up
    // once per device, here with the context's device current)
    e = noAllocScratch(&c->noAllocScratch);
    if (e != hipSuccess)
    {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PHX_ERR_HIP;
    }
    auto alloc = [&](double** p, int q) { return devAlloc(c, (void**)p, static_cast<size_t>(G.size(q)) * 8); };
    phx_status s = PHX_OK;
    // each E/B set is one block (Ex, Ey, Ez, Bx, By, Bz, each starting on a 256-byte boundary)
    size_t emWords = 0;
    size_t emOff[6];
    for (int k = 0; k < 6; ++k)
    {
        emOff[k] = emWords;
        emWords += (static_cast<size_t>(G.size(k < 3 ? QEX + k : QBX + k - 3)) + 31) / 32 * 32;
    }
    for (int set = 0; set < 3 && s == PHX_OK; ++set)
    {
        double* block = nullptr;
        s             = devAlloc(c, (void**)&block, emWords * 8);
        for (int k = 0; k < 6 && s == PHX_OK; ++k)
            c->em[set][k] = block + emOff[k];
    }
    if (s == PHX_OK && cfg->precision == PHX_PRECISION_FP32)
    {
        float* block = nullptr;
        s            = devAlloc(c, (void**)&block, emWords * 4);
        for (int k = 0; k < 6 && s == PHX_OK; ++k)
            c->em32[k] = block + emOff[k];
    }
    for (int k = 0; k < 3 && s == PHX_OK; ++k)
    {
        s = alloc(&c->Bold[k], QBX + k);
        if (s == PHX_OK)
            s = alloc(&c->fluxSum[k], QEX + k);
        if (s == PHX_OK)
            s = alloc(&c->J[k], QJX + k);
        if (s == PHX_OK)
            s = alloc(&c->Ve[k], QMOM);
        if (s == PHX_OK)
            s = alloc(&c->Vi[k], QMOM);
    }
    if (s == PHX_OK)
        s = alloc(&c->Pe, QMOM);
    if (s == PHX_OK)
        s = alloc(&c->rho, QMOM);
    if (s == PHX_OK)
        s = alloc(&c->massD, QMOM);
    if (s == PHX_OK)
        s = alloc(&c->tmp, QMOM);
    for (int k = 0; k < 5 && s == PHX_OK; ++k)
        s = alloc(&c->tot5[k], QMOM);
    for (int p = 0; p < cfg->n_pop && s == PHX_OK; ++p)
        for (int r = 0; r < 2 && s == PHX_OK; ++r)
            for (int m = 0; m < 5 && s == PHX_OK; ++m)
                s = alloc(&c->raw[r][p][m], QMOM);
    for (int p = 0; p < cfg->n_pop && s == PHX_OK; ++p)
    {
        c->pops[p].mass = cfg->mass[p];
        for (int m = 0; m < 5 && s == PHX_OK; ++m)
            s = alloc(&c->pops[p].mom[m], QMOM);
        for (int k = 0; k < 5 && s == PHX_OK; ++k)
            s = storeAlloc(c, c->pops[p].arr[k], k == 0 ? std::max<size_t>(particle_capacity, 1024) : 1024);
    }
    c->tuning = defaultTuning(D, cfg->interp_order);
    c->sortInterval = c->tuning.sort_interval;
    if (s == PHX_OK)
        s = buildPatchTable(c, plo, phi, np);
    if (s == PHX_OK)
        s = allocSegTables(c);
    if (s == PHX_OK)
        s = devAlloc(c, (void**)&c->err, sizeof(DevError));
    if (s == PHX_OK)
        G.stop = &c->err->code;
    if (s == PHX_OK)
        s = devAlloc(c, (void**)&c->dropped, 64);
    if (s == PHX_OK)
        s = devAlloc(c, (void**)&c->tot, 64);
    if (s != PHX_OK)
    {
        phx_destroy(c);
        return s;
    }
    // every field starts at zero
    for (int set = 0; set < 3; ++set)
        for (int k = 0; k < 6; ++k)
            launchZero(c->em[set][k], G.size(k < 3 ? QEX + k : QBX + k - 3), c->stream);
    for (int k = 0; k < 3; ++k)
    {
        launchZero(c->Bold[k], G.size(QBX + k), c->stream);
        launchZero(c->fluxSum[k], G.size(QEX + k), c->stream);
    }
    for (int k = 0; k < 3; ++k)
    {
        launchZero(c->J[k], G.size(QJX + k), c->stream);
        launchZero(c->Ve[k], G.size(QMOM), c->stream);
        launchZero(c->Vi[k], G.size(QMOM), c->stream);
    }
    for (double* p : {c->Pe, c->rho, c->massD, c->tmp})
        launchZero(p, G.size(QMOM), c->stream);
    for (int k = 0; k < 5; ++k)
        launchZero(c->tot5[k], G.size(QMOM), c->stream);
    for (int p = 0; p < cfg->n_pop; ++p)
        for (int r = 0; r < 2; ++r)
            for (int m = 0; m < 5; ++m)
                launchZero(c->raw[r][p][m], G.size(QMOM), c->stream);
    zeroMoments(c);
    DevError z{};
    (void)hipMemcpy(c->err, &z, sizeof(z), hipMemcpyHostToDevice);
    unsigned int zz[16] = {};
    (void)hipMemcpy(c->dropped, zz, 64, hipMemcpyHostToDevice);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
    {
        phx_destroy(c);
        return PHX_ERR_HIP;
    }
    *out = c;
    return PHX_OK;
}
} // namespace

extern "C" {

phx_status phx_create(phx_config const* cfg, phx_patch const* patch, size_t particle_capacity,
                      phx_ctx** out)
{
    if (!patch)
        return PHX_ERR_ARG;
    return createLevel(cfg, patch, patch->lower, patch->upper, 1, particle_capacity, out);
}

phx_status phx_create_level(phx_config const* cfg, phx_patch const* level, int n_patches,
                            int const* patch_lower, int const* patch_upper, size_t particle_capacity,
                            phx_ctx** out)
{
    if (!level || !patch_lower || !patch_upper || n_patches < 1)
        return PHX_ERR_ARG;
    // the borders of a level that does not cover a periodic domain are fed by the coarser level
    if (n_patches > 1 && !level->periodic)
        return PHX_ERR_ARG;
    return createLevel(cfg, level, patch_lower, patch_upper, n_patches, particle_capacity, out);
}

void phx_destroy(phx_ctx* c)
{
    if (!c)
        return;
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs)
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
    for (auto& e : c->events)
    {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (void* p : c->allocs)
        (void)hipFree(p);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

char const* phx_last_error(phx_ctx const* c) { return c ? c->lastError.c_str() : "null context"; }

phx_status phx_synchronize(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    return PHX_OK;
}

void* phx_stream(phx_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

phx_status phx_field_shape(phx_ctx const* c, int fid, int shape[3])
{
    double* p;
    int q;
    if (!c || !shape || !fieldOf(c, fid, p, q))
        return PHX_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        shape[d] = c->G.shape[q][d];
    return PHX_OK;
}

// ---- the patches of a level -------------------------------------------------------------------
namespace
{
phx_status validPatch(phx_ctx const* c, int p)
{
    if (!c || p < 0 || p >= static_cast<int>(c->patchBox.size()))
        return PHX_ERR_ARG;
    return PHX_OK;
}

// shape of quantity q on patch p (GridLayout::allocSize of the patch's layout) and the offset of
// the patch's array origin inside the level's array (both include g ghost nodes)
void patchShape(phx_ctx const* c, int p, int q, int* shape, int* off)
{
    for (int d = 0; d < 3; ++d)
    {
        bool const on = d < c->cfg.dim;
        int const n   = on ? c->patchBox[p][3 + d] - c->patchBox[p][d] + 1 : 1;
        shape[d]      = on ? n + 1 + 2 * c->G.g - centering(q, d) : 1;
        off[d]        = on ? c->patchBox[p][d] - c->boxes.dom_lo[d] : 0;
    }
}
} // namespace

phx_status phx_patch_box(phx_ctx const* c, int p, int lower[3], int upper[3])
{
    PHX_TRY(validPatch(c, p));
    if (!lower || !upper)
        return PHX_ERR_ARG;
    for (int d = 0; d < 3; ++d)
    {
        lower[d] = c->patchBox[p][d];
        upper[d] = c->patchBox[p][3 + d];
    }
    return PHX_OK;
}

phx_status phx_patch_nonlevel_ghost_boxes(phx_ctx const* c, int p, int* n, int lower[][3],
                                          int upper[][3])
{
    PHX_TRY(validPatch(c, p));
    if (!n || !lower || !upper)
        return PHX_ERR_ARG;
    int const D  = c->cfg.dim;
    int const pg = c->G.pg;
    auto const& me = c->patchBox[p];
    int cnt = 0;
    auto put = [&](int const* lo, int const* hi) {
        if (cnt >= MAX_NLGB)
            return false;
        for (int d = 0; d < 3; ++d)
        {
            lower[cnt][d] = d < D ? lo[d] : 0;
            upper[cnt][d] = d < D ? hi[d] : 0;
        }
        ++cnt;
        return true;
    };
    put(&me[0], &me[3]);
    // neighbours (periodic images of the level included) intersected with the particle ghost box
    int nshift = 1;
    for (int d = 0; d < D; ++d)
        nshift *= c->patch.periodic ? 3 : 1;
    for (int q = 0; q < static_cast<int>(c->patchBox.size()); ++q)
        for (int sh = 0; sh < nshift; ++sh)
        {
            int shift[3] = {0, 0, 0}, r = sh;
            bool self = q == p;
            for (int d = 0; d < D && c->patch.periodic; ++d)
            {
                shift[d] = (r % 3 - 1) * (c->boxes.dom_hi[d] - c->boxes.dom_lo[d] + 1);
                r /= 3;
                self = self && shift[d] == 0;
            }
            if (self)
                continue;
            int lo[3], hi[3];
            bool empty = false;
            for (int d = 0; d < D; ++d)
            {
                lo[d] = std::max(me[d] - pg, c->patchBox[q][d] + shift[d]);
                hi[d] = std::min(me[3 + d] + pg, c->patchBox[q][3 + d] + shift[d]);
                empty = empty || hi[d] < lo[d];
            }
            if (!empty && !put(lo, hi))
                return fail(const_cast<phx_ctx*>(c), PHX_ERR_CAPACITY, "more than 27 selection boxes");
        }
    *n = cnt;
    return PHX_OK;
}

phx_status phx_patch_field_shape(phx_ctx const* c, int p, int fid, int shape[3])
{
    PHX_TRY(validPatch(c, p));
    double* f;
    int q, off[3];
    if (!shape || !fieldOf(c, fid, f, q))
        return PHX_ERR_ARG;
    patchShape(c, p, q, shape, off);
    return PHX_OK;
}

phx_status phx_set_patch_field(phx_ctx* c, int p, int fid, double const* host)
{
    PHX_TRY(validPatch(c, p));
    double* f;
    int q, shape[3], off[3];
    if (!host || !fieldOf(c, fid, f, q))
        return fail(c, PHX_ERR_ARG, "bad field id");
    patchShape(c, p, q, shape, off);
    std::vector<double> lv(static_cast<size_t>(c->G.size(q)));
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    PHX_HIP(c, hipMemcpy(lv.data(), f, lv.size() * 8, hipMemcpyDeviceToHost));
    // the patch's physical nodes (GridLayout physicalStartIndex..physicalEndIndex per direction)
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d)
    {
        bool const on = d < c->cfg.dim;
        lo[d] = on ? c->G.g : 0;
        hi[d] = on ? shape[d] - 1 - c->G.g : 0;
    }
    for (int i = lo[0]; i <= hi[0]; ++i)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int k = lo[2]; k <= hi[2]; ++k)
                lv[static_cast<size_t>(c->G.idx(q, i + off[0], j + off[1], k + off[2]))]
                    = host[(static_cast<size_t>(i) * shape[1] + j) * shape[2] + k];
    PHX_HIP(c, hipMemcpy(f, lv.data(), lv.size() * 8, hipMemcpyHostToDevice));
    return PHX_OK;
}

phx_status phx_get_patch_field(phx_ctx* c, int p, int fid, double* host)
{
    PHX_TRY(validPatch(c, p));
    double* f;
    int q, shape[3], off[3];
    if (!host || !fieldOf(c, fid, f, q))
        return fail(c, PHX_ERR_ARG, "bad field id");
    patchShape(c, p, q, shape, off);
    std::vector<double> lv(static_cast<size_t>(c->G.size(q)));
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    PHX_HIP(c, hipMemcpy(lv.data(), f, lv.size() * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < shape[0]; ++i)
        for (int j = 0; j < shape[1]; ++j)
            for (int k = 0; k < shape[2]; ++k)
                host[(static_cast<size_t>(i) * shape[1] + j) * shape[2] + k]
                    = lv[static_cast<size_t>(c->G.idx(q, i + off[0], j + off[1], k + off[2]))];
    return PHX_OK;
}

phx_status phx_get_patch_particles(phx_ctx* c, int p, int pop, size_t* n, int32_t* icell,
                                   double* delta, double* v, double* weight, double* charge)
{
    PHX_TRY(validPatch(c, p));
    PHX_TRY(validPop(c, pop, PHX_DOMAIN));
    if (!n)
        return PHX_ERR_ARG;
    int const D = c->cfg.dim;
    std::vector<int32_t> ic;
    std::vector<double> de, vv, w, q;
    PHX_TRY(download(c, c->pops[pop], c->pops[pop].arr[PHX_DOMAIN], ic, de, vv, w, q));
    auto const& b = c->patchBox[p];
    size_t m = 0;
    for (size_t i = 0; i < w.size(); ++i)
    {
        bool in = true;
        for (int d = 0; d < D; ++d)
            in = in && ic[i * D + d] >= b[d] && ic[i * D + d] <= b[3 + d];
        if (!in)
            continue;
        if (icell)
        {
            for (int d = 0; d < D; ++d)
            {
                icell[m * D + d] = ic[i * D + d];
                delta[m * D + d] = de[i * D + d];
            }
            for (int k = 0; k < 3; ++k)
                v[m * 3 + k] = vv[i * 3 + k];
            weight[m] = w[i];
            charge[m] = q[i];
        }
        ++m;
    }
    *n = m;
    return PHX_OK;
}

phx_status phx_add_particles(phx_ctx* c, int pop, int kind, size_t n, int32_t const* icell,
                             double const* delta, double const* v, double const* weight,
                             double const* charge)
{
    PHX_TRY(validPop(c, pop, kind));
    if (n && (!icell || !delta || !v || !weight || !charge))
        return fail(c, PHX_ERR_ARG, "null particle buffer");
    if (n && kind != PHX_DOMAIN && mixed(c))
        return fail(c, PHX_ERR_STATE, "the mixed-precision family holds domain particles only");
    Store& s = c->pops[pop].arr[kind];
    if (s.seg)
        return fail(c, PHX_ERR_STATE, "phx_add_particles: the array is cell-segmented (sorted); append "
                                      "before phx_initialize or phx_sort_particles");
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    return upload(c, c->pops[pop], s, n, icell, delta, v, weight, charge, s.n);
}

phx_status phx_set_field(phx_ctx* c, int fid, double const* host)
{
    double* p;
    int q;
    if (!c || !host || !fieldOf(c, fid, p, q))
        return fail(c, PHX_ERR_ARG, "bad field id");
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    PHX_HIP(c, hipMemcpy(p, host, static_cast<size_t>(c->G.size(q)) * 8, hipMemcpyHostToDevice));
    return PHX_OK;
}

phx_status phx_get_field(phx_ctx* c, int fid, double* host)
{
    double* p;
    int q;
    if (!c || !host || !fieldOf(c, fid, p, q))
        return fail(c, PHX_ERR_ARG, "bad field id");
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    PHX_HIP(c, hipMemcpy(host, p, static_cast<size_t>(c->G.size(q)) * 8, hipMemcpyDeviceToHost));
    return PHX_OK;
}

phx_status phx_set_particles(phx_ctx* c, int pop, int kind, size_t n, int32_t const* icell,
                             double const* delta, double const* v, double const* weight,
                             double const* charge)
{
    PHX_TRY(validPop(c, pop, kind));
    if (n && (!icell || !delta || !v || !weight || !charge))
        return fail(c, PHX_ERR_ARG, "null particle buffer");
    if (n && kind != PHX_DOMAIN && mixed(c))
        return fail(c, PHX_ERR_STATE, "the mixed-precision family holds domain particles only");
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    return upload(c, c->pops[pop], c->pops[pop].arr[kind], n, icell, delta, v, weight, charge, 0);
}

phx_status phx_particle_count(phx_ctx* c, int pop, int kind, size_t* n)
{
    PHX_TRY(validPop(c, pop, kind));
    if (!n)
        return PHX_ERR_ARG;
    Store const& s = c->pops[pop].arr[kind];
    if (kind == PHX_DOMAIN)
    {
        // lost particles and the empty slots of segmented storage hold dead entries
        unsigned int dropped = 0;
        PHX_HIP(c, hipStreamSynchronize(c->stream));
        PHX_HIP(c, hipMemcpy(&dropped, c->dropped, 4, hipMemcpyDeviceToHost));
        if (dropped || s.seg)
        {
            std::vector<int32_t> ic;
            std::vector<double> de, v, w, q;
            PHX_TRY(download(c, c->pops[pop], s, ic, de, v, w, q));
            *n = w.size();
            return PHX_OK;
        }
    }
    *n = s.n;
    return PHX_OK;
}

phx_status phx_get_particles(phx_ctx* c, int pop, int kind, int32_t* icell, double* delta, double* v,
                             double* weight, double* charge)
{
    PHX_TRY(validPop(c, pop, kind));
    std::vector<int32_t> ic;
    std::vector<double> de, vv, w, q;
    PHX_TRY(download(c, c->pops[pop], c->pops[pop].arr[kind], ic, de, vv, w, q));
    std::copy(ic.begin(), ic.end(), icell);
    std::copy(de.begin(), de.end(), delta);
    std::copy(vv.begin(), vv.end(), v);
    std::copy(w.begin(), w.end(), weight);
    std::copy(q.begin(), q.end(), charge);
    return PHX_OK;
}

// Particle<dim> records: double weight, charge; int iCell[dim]; double delta[dim]; double v[3]
phx_status phx_set_particles_aos(phx_ctx* c, int pop, int kind, size_t n, void const* records)
{
    PHX_TRY(validPop(c, pop, kind));
    int const D       = c->cfg.dim;
    size_t const rs   = recordSize(D);
    size_t const oIc  = 16;
    size_t const oDe  = D == 1 ? 24 : (D == 2 ? 24 : 32);
    size_t const oV   = oDe + 8 * D;
    auto const* bytes = static_cast<unsigned char const*>(records);
    std::vector<int32_t> ic(n * D);
    std::vector<double> de(n * D), v(n * 3), w(n), q(n);
    for (size_t i = 0; i < n; ++i)
    {
        unsigned char const* r = bytes + i * rs;
        std::memcpy(&w[i], r, 8);
        std::memcpy(&q[i], r + 8, 8);
        std::memcpy(&ic[i * D], r + oIc, 4 * D);
        std::memcpy(&de[i * D], r + oDe, 8 * D);
        std::memcpy(&v[i * 3], r + oV, 24);
    }
    return phx_set_particles(c, pop, kind, n, ic.data(), de.data(), v.data(), w.data(), q.data());
}

phx_status phx_get_particles_aos(phx_ctx* c, int pop, int kind, void* records)
{
    PHX_TRY(validPop(c, pop, kind));
    int const D     = c->cfg.dim;
    size_t const rs = recordSize(D);
    size_t const oDe = D == 1 ? 24 : (D == 2 ? 24 : 32);
    size_t const oV  = oDe + 8 * D;
    std::vector<int32_t> ic;
    std::vector<double> de, v, w, q;
    PHX_TRY(download(c, c->pops[pop], c->pops[pop].arr[kind], ic, de, v, w, q));
    auto* bytes = static_cast<unsigned char*>(records);
    for (size_t i = 0; i < w.size(); ++i)
    {
        unsigned char* r = bytes + i * rs;
        std::memset(r, 0, rs);
        std::memcpy(r, &w[i], 8);
        std::memcpy(r + 8, &q[i], 8);
        std::memcpy(r + 16, &ic[i * D], 4 * D);
        std::memcpy(r + oDe, &de[i * D], 8 * D);
        std::memcpy(r + oV, &v[i * 3], 24);
    }
    return PHX_OK;
}

phx_status phx_load_maxwellian(phx_ctx* c, int pop, double const* n, double const* vx,
                               double const* vy, double const* vz, double const* tx,
                               double const* ty, double const* tz, double charge, uint32_t ppc,
                               uint64_t seed, double cutoff)
{
    PHX_TRY(validPop(c, pop, PHX_DOMAIN));
    int const D = c->cfg.dim;
    // MaxwellianParticleInitializer::loadParticles: cells in row-major AMR order, per particle
    // three normal draws (fresh distributions each) then dim uniform deltas in [0, 1-eps)
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> deltaDistrib{0, 1. - std::numeric_limits<double>::epsilon()};
    std::vector<int32_t> ic;
    std::vector<double> de, v, w, q;
    int cell[3] = {c->patch.lower[0], c->patch.lower[1], c->patch.lower[2]};
    size_t ncell = 1;
    for (int d = 0; d < D; ++d)
        ncell *= static_cast<size_t>(c->G.n[d]);
    for (size_t flat = 0; flat < ncell; ++flat)
    {
        if (!(n[flat] < cutoff))
        {
            double const cw = n[flat] / ppc;
            for (uint32_t ip = 0; ip < ppc; ++ip)
            {
                std::normal_distribution<> mx(vx[flat], tx[flat]);
                std::normal_distribution<> my(vy[flat], ty[flat]);
                std::normal_distribution<> mz(vz[flat], tz[flat]);
                double const a = mx(gen);
                double const b = my(gen);
                double const cc = mz(gen);
                w.push_back(cw);
                q.push_back(charge);
                for (int d = 0; d < D; ++d)
                    ic.push_back(cell[d]);
                for (int d = 0; d < D; ++d)
                    de.push_back(deltaDistrib(gen));
                v.push_back(a);
                v.push_back(b);
                v.push_back(cc);
            }
        }
        // next cell, last dimension fastest
        for (int d = D - 1; d >= 0; --d)
        {
            if (++cell[d] <= c->patch.upper[d])
                break;
            cell[d] = c->patch.lower[d];
        }
    }
    Store& s = c->pops[pop].arr[PHX_DOMAIN];
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    return upload(c, c->pops[pop], s, w.size(), ic.data(), de.data(), v.data(), w.data(), q.data(), s.n);
}

phx_status phx_update_populations(phx_ctx* c, int em_set, double dt, int mode)
{
    if (!c)
        return PHX_ERR_ARG;
    if (em_set < 0 || em_set > 2 || (mode != PHX_MODE_DOMAIN_ONLY && mode != PHX_MODE_ALL))
        return fail(c, PHX_ERR_ARG, "bad em set or mode");
    if (mixed(c))
        return fail(c, PHX_ERR_STATE,
                    "the mixed-precision family runs the resident level path (phx_move_ions, phx_advance)");
    zeroMoments(c);
    int const D = c->cfg.dim;
    Wrap noWrap{};
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        Pop& P = c->pops[p];
        Store& dom = P.arr[PHX_DOMAIN];
        Store& pg  = P.arr[PHX_PATCH_GHOST];
        Store& lg  = P.arr[PHX_LEVEL_GHOST];
        if (mode == PHX_MODE_DOMAIN_ONLY)
        {
            // pushed copies: nothing written back
            PushArgs A = pushArgs(c, p, KIND_DOMAIN, mode, em_set, dt, true);
            launchPush(A, dom.view(), dom.view(), c->stream);
            if (lg.n)
            {
                PushArgs L = pushArgs(c, p, KIND_LEVEL_GHOST, mode, em_set, dt, false);
                launchPush(L, lg.view(), lg.view(), c->stream);
            }
            PHX_TRY(checkLaunch(c, "push"));
            continue;
        }
        // mode all: existing patch ghosts are deposited unpushed
        if (pg.n)
        {
            PushArgs G = pushArgs(c, p, KIND_PATCH_GHOST, mode, em_set, dt, false);
            launchDeposit(G, pg.view(), c->stream);
        }
        size_t const nd = dom.n, nl = lg.n;
        PHX_TRY(scratchReserve(c, std::max(nd, nl)));
        // worst-case room: every domain particle may leave, every level ghost may enter
        PHX_TRY(storeReserve(c, pg, pg.n + nd));
        // domain: push in place with a fate per particle, then split keep / leave
        PushArgs A = pushArgs(c, p, KIND_DOMAIN, mode, em_set, dt, false);
        A.status   = c->status;
        launchPush(A, dom.view(), dom.view(), c->stream);
        PHX_TRY(checkLaunch(c, "push"));
        PHX_TRY(checkDeviceError(c));
        // leavers appended to patchGhost (updateAndDepositAll_ :248-254)
        unsigned long long got = 0;
        launchCompact(dom.view(), pg.view(), c->status, nd, ST_LEAVE, c->blockCnt, c->tot, pg.n, D,
                      noWrap, c->stream);
        PHX_HIP(c, hipStreamSynchronize(c->stream));
        PHX_HIP(c, hipMemcpy(&got, c->tot, 8, hipMemcpyDeviceToHost));
        pg.n += got;
        // kept domain particles -> scratch, then swap (erase of the leavers, :273)
        launchCompact(dom.view(), c->scratch.view(), c->status, nd, ST_KEEP, c->blockCnt, c->tot, 0,
                      D, noWrap, c->stream);
        PHX_HIP(c, hipStreamSynchronize(c->stream));
        PHX_HIP(c, hipMemcpy(&got, c->tot, 8, hipMemcpyDeviceToHost));
        std::swap(dom, c->scratch);
        dom.n          = got;
        dom.seg        = false;
        c->scratch.n   = 0;
        c->scratch.seg = false;
        if (nl)
        {
            // level ghosts (:275-288): pushed with inGhostBox / inGhostLayer selection; those that
            // entered the domain were deposited by the push and join the domain array
            PHX_TRY(storeReserve(c, dom, dom.n + nl));
            PHX_TRY(scratchReserve(c, nl));
            PushArgs L = pushArgs(c, p, KIND_LEVEL_GHOST, mode, em_set, dt, false);
            L.status   = c->status;
            launchPush(L, lg.view(), lg.view(), c->stream);
            PHX_TRY(checkDeviceError(c));
            launchCompact(lg.view(), dom.view(), c->status, nl, ST_EXPORT, c->blockCnt, c->tot, dom.n,
                          D, noWrap, c->stream);
            PHX_HIP(c, hipStreamSynchronize(c->stream));
            PHX_HIP(c, hipMemcpy(&got, c->tot, 8, hipMemcpyDeviceToHost));
            dom.n += got;
            launchCompact(lg.view(), c->scratch.view(), c->status, nl, ST_KEEP, c->blockCnt, c->tot, 0,
                          D, noWrap, c->stream);
            PHX_HIP(c, hipStreamSynchronize(c->stream));
            PHX_HIP(c, hipMemcpy(&got, c->tot, 8, hipMemcpyDeviceToHost));
            std::swap(lg, c->scratch);
            lg.n         = got;
            c->scratch.n = 0;
        }
    }
    return checkDeviceError(c);
}

phx_status phx_push(phx_ctx* c, int pop, int kind, int em_set, double dt, int n_first,
                    int const* first_lower, int const* first_upper, int n_second,
                    int const* second_lower, int const* second_upper, size_t counts[2])
{
    PHX_TRY(validPop(c, pop, kind));
    if (!counts || em_set < 0 || em_set > 2 || n_first < 0 || n_first > MAX_NLGB || n_second < 0
        || n_second > MAX_NLGB || (n_first && (!first_lower || !first_upper))
        || (n_second && (!second_lower || !second_upper)))
        return fail(c, PHX_ERR_ARG, "phx_push: bad em set, selector box list or counts pointer");
    if (mixed(c))
        return fail(c, PHX_ERR_STATE, "phx_push: the mixed-precision family runs the resident level path");
    int const D = c->cfg.dim;
    auto boxes = [D](int nb, int const* lo, int const* hi) {
        SelBoxes S{};
        S.n = nb;
        for (int b = 0; b < nb; ++b)
            for (int d = 0; d < 3; ++d)
            {
                S.lo[b][d] = d < D ? lo[3 * b + d] : 0;
                S.hi[b][d] = d < D ? hi[3 * b + d] : 0;
            }
        return S;
    };
    Store& s       = c->pops[pop].arr[kind];
    size_t const n = s.n;
    counts[0] = counts[1] = 0;
    if (n == 0)
        return PHX_OK;
    PHX_TRY(scratchReserve(c, n));
    PushArgs A = pushArgs(c, pop, kind, PHX_MODE_ALL, em_set, dt, false);
    launchMove(A, s.view(), boxes(n_first, first_lower, first_upper),
               boxes(n_second, second_lower, second_upper), c->status, c->stream);
    PHX_TRY(checkLaunch(c, "phx_push"));
    PHX_TRY(checkDeviceError(c));
    // the partitions of the two selectors: [both | first only | neither]; empty slots dropped
    Wrap noWrap{};
    unsigned long long got[3] = {0, 0, 0}, off = 0;
    for (uint8_t want = 0; want < 3; ++want)
    {
        launchCompact(s.view(), c->scratch.view(), c->status, n, want, c->blockCnt, c->tot, off, D, noWrap,
                      c->stream);
        PHX_HIP(c, hipStreamSynchronize(c->stream));
        PHX_HIP(c, hipMemcpy(&got[want], c->tot, 8, hipMemcpyDeviceToHost));
        off += got[want];
    }
    std::swap(s, c->scratch);
    s.n            = off;
    s.seg          = false;
    c->scratch.n   = 0;
    c->scratch.seg = false;
    counts[0]      = got[0];
    counts[1]      = got[0] + got[1];
    return PHX_OK;
}

phx_status phx_gather(phx_ctx* c, int em_set, int pop, int kind, double* E, double* B)
{
    PHX_TRY(validPop(c, pop, kind));
    if (!E || !B || em_set < 0 || em_set > 2)
        return fail(c, PHX_ERR_ARG, "phx_gather: bad em set or output buffer");
    if (mixed(c))
        return fail(c, PHX_ERR_STATE, "phx_gather: the mixed-precision family runs the resident level path");
    Store const& s = c->pops[pop].arr[kind];
    size_t const n = s.n;
    if (n == 0)
        return PHX_OK;
    double* dev = nullptr;
    PHX_TRY(devAlloc(c, (void**)&dev, n * 6 * 8));
    PushArgs A = pushArgs(c, pop, kind, PHX_MODE_ALL, em_set, 0., false);
    launchGather(A, s.view(), dev, dev + 3 * n, c->stream);
    phx_status st = checkLaunch(c, "phx_gather");
    if (st == PHX_OK)
        st = checkDeviceError(c);
    std::vector<double> h(n * 6);
    std::vector<int32_t> ic0(n);
    if (st == PHX_OK && hipMemcpy(h.data(), dev, n * 6 * 8, hipMemcpyDeviceToHost) != hipSuccess)
        st = fail(c, PHX_ERR_HIP, "phx_gather: copy of the gathered fields");
    if (st == PHX_OK && hipMemcpy(ic0.data(), s.ic[0], n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        st = fail(c, PHX_ERR_HIP, "phx_gather: copy of the cells");
    devFree(c, dev);
    if (st != PHX_OK)
        return st;
    // live slots in storage order, as phx_get_particles returns them
    size_t o = 0;
    for (size_t i = 0; i < n; ++i)
    {
        if (ic0[i] == DEAD_CELL)
            continue;
        for (int k = 0; k < 3; ++k)
        {
            E[3 * o + k] = h[3 * i + k];
            B[3 * o + k] = h[3 * n + 3 * i + k];
        }
        ++o;
    }
    return PHX_OK;
}

phx_status phx_gather_noalloc(phx_ctx* c, int em_set, int pop, int kind, double* E, double* B)
{
    PHX_TRY(validPop(c, pop, kind));
    if (!E || !B || em_set < 0 || em_set > 2)
        return fail(c, PHX_ERR_ARG, "phx_gather_noalloc: bad em set or output buffer");
    if (mixed(c))
        return fail(c, PHX_ERR_STATE,
                    "phx_gather_noalloc: the mixed-precision family runs the resident level path");
    Store const& s = c->pops[pop].arr[kind];
    size_t const n = s.n;
    if (n > NOALLOC_GATHER_MAX_SLOTS)
        return fail(c, PHX_ERR_CAPACITY,
                    "phx_gather_noalloc: the array holds more slots than the scratch array allows ("
                        + std::to_string(n) + " > " + std::to_string(NOALLOC_GATHER_MAX_SLOTS) + ")");
    if (n == 0)
        return PHX_OK;
    if (!c->noAllocScratch)
        return fail(c, PHX_ERR_STATE, "phx_gather_noalloc: no scratch address for this device");
    PushArgs A = pushArgs(c, pop, kind, PHX_MODE_ALL, em_set, 0., false);
    launchGatherNoAlloc(A, s.view(), c->noAllocScratch, c->stream);
    PHX_TRY(checkLaunch(c, "phx_gather_noalloc"));
    PHX_TRY(checkDeviceError(c));
    double const *dE = nullptr, *dB = nullptr;
    unsigned long long const* dTotal = nullptr;
    noAllocGatherOutputs(c->noAllocScratch, &dE, &dB, &dTotal);
    unsigned long long live = 0;
    PHX_HIP(c, hipMemcpy(&live, dTotal, 8, hipMemcpyDeviceToHost));
    if (live > n)
        return fail(c, PHX_ERR_STATE, "phx_gather_noalloc: inconsistent live slot count");
    if (live)
    {
        PHX_HIP(c, hipMemcpy(E, dE, live * 3 * 8, hipMemcpyDeviceToHost));
        PHX_HIP(c, hipMemcpy(B, dB, live * 3 * 8, hipMemcpyDeviceToHost));
    }
    return PHX_OK;
}

phx_status phx_deposit(phx_ctx* c, int pop, int kind, double coef)
{
    PHX_TRY(validPop(c, pop, kind));
    if (mixed(c))
        return fail(c, PHX_ERR_STATE, "phx_deposit: the mixed-precision family runs the resident level path");
    PushArgs A = pushArgs(c, pop, kind, PHX_MODE_ALL, PHX_EM, 0., true);
    A.coef     = coef;
    launchDeposit(A, c->pops[pop].arr[kind].view(), c->stream);
    PHX_TRY(checkLaunch(c, "phx_deposit"));
    return checkDeviceError(c);
}

phx_status phx_refine_particles(phx_ctx* fine, phx_ctx const* coarse, int pop, int dest_kind,
                                int n_boxes, int const* box_lower, int const* box_upper,
                                int refined_particle_nbr)
{
    if (!fine || !coarse || fine == coarse)
        return fine ? fail(fine, PHX_ERR_ARG, "phx_refine_particles: two different contexts") : PHX_ERR_ARG;
    PHX_TRY(validPop(fine, pop, dest_kind));
    int const D = fine->cfg.dim, O = fine->cfg.interp_order;
    if (coarse->cfg.dim != D || coarse->cfg.interp_order != O || pop >= coarse->cfg.n_pop)
        return fail(fine, PHX_ERR_ARG, "phx_refine_particles: coarse and fine levels differ in "
                                       "dimension, order or populations");
    if (mixed(fine) || mixed(coarse))
        return fail(fine, PHX_ERR_STATE, "phx_refine_particles: the split runs on FP64 levels");
    if (dest_kind == PHX_PATCH_GHOST)
        return fail(fine, PHX_ERR_ARG, "phx_refine_particles: the split fills the domain or the "
                                       "level ghost arrays");
    if (n_boxes < 1 || n_boxes > MAX_SPLIT_BOXES || !box_lower || !box_upper)
        return fail(fine, PHX_ERR_ARG, "phx_refine_particles: 1.." + std::to_string(MAX_SPLIT_BOXES)
                                           + " destination boxes");
    SplitSpec S{};
    int const nbr = refined_particle_nbr ? refined_particle_nbr : (D == 1 ? 2 : (D == 2 ? 4 : 6));
    if (!splitPattern(D, O, nbr, S))
        return fail(fine, PHX_ERR_ARG, "phx_refine_particles: no split pattern for dim "
                                           + std::to_string(D) + ", order " + std::to_string(O)
                                           + ", " + std::to_string(nbr) + " refined particles");
    S.nBoxes = n_boxes;
    S.ratio  = 2;
    // Splitter::maxCellDistanceFromSplit (splitter.hpp:116-120): ceil((order + 1) / 2)
    S.grow = (O + 2) / 2;
    for (int b = 0; b < n_boxes; ++b)
        for (int d = 0; d < 3; ++d)
        {
            S.lo[b][d] = d < D ? box_lower[3 * b + d] : 0;
            S.hi[b][d] = d < D ? box_upper[3 * b + d] : 0;
        }
    Store const& src = coarse->pops[pop].arr[PHX_DOMAIN];
    Store& dst       = fine->pops[pop].arr[dest_kind];
    if (src.n == 0)
        return PHX_OK;
    // the coarse level's arrays are read on the fine level's stream
    PHX_HIP(fine, hipStreamSynchronize(coarse->stream));
    size_t const nb = (src.n + 255) / 256 + 1;
    if (nb > fine->blockCap)
    {
        devFree(fine, fine->blockCnt);
        fine->blockCnt = nullptr;
        fine->blockCap = 0;
        PHX_TRY(devAlloc(fine, (void**)&fine->blockCnt, nb * 2 * 4));
        fine->blockCap = nb * 2;
    }
    launchSplitCount(src.view(), src.n, S, D, fine->blockCnt, fine->tot, fine->stream);
    PHX_TRY(checkLaunch(fine, "phx_refine_particles (count)"));
    unsigned long long got = 0;
    PHX_HIP(fine, hipStreamSynchronize(fine->stream));
    PHX_HIP(fine, hipMemcpy(&got, fine->tot, 8, hipMemcpyDeviceToHost));
    if (got == 0)
        return PHX_OK;
    PHX_TRY(storeReserve(fine, dst, dst.n + got));
    launchSplitWrite(src.view(), src.n, S, D, fine->blockCnt, dst.view(), dst.n, fine->stream);
    PHX_TRY(checkLaunch(fine, "phx_refine_particles (write)"));
    dst.n += got;
    if (dest_kind == PHX_DOMAIN)
        dst.seg = false;
    return PHX_OK;
}

phx_status phx_level_ghost_last_step(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        Pop& P = c->pops[p];
        // std::swap(levelGhostParticlesNew, levelGhostParticlesOld); New.clear(); levelGhost = Old
        std::swap(P.arr[PHX_LEVEL_GHOST_NEW], P.arr[PHX_LEVEL_GHOST_OLD]);
        P.arr[PHX_LEVEL_GHOST_NEW].n = 0;
        Store const& old = P.arr[PHX_LEVEL_GHOST_OLD];
        Store& lg        = P.arr[PHX_LEVEL_GHOST];
        lg.n             = 0;
        PHX_TRY(storeReserve(c, lg, old.n));
        if (old.n)
        {
            launchSegCopy(old.view(), lg.view(), old.n, c->cfg.dim, c->stream);
            PHX_TRY(checkLaunch(c, "phx_level_ghost_last_step"));
        }
        lg.n   = old.n;
        lg.seg = false;
    }
    return PHX_OK;
}

phx_status phx_time_interpolate(phx_ctx* dst, int dst_field, phx_ctx const* old_ctx, int old_field,
                                phx_ctx const* new_ctx, int new_field, int const lower[3],
                                int const upper[3], double alpha)
{
    if (!dst || !old_ctx || !new_ctx || !lower || !upper)
        return dst ? fail(dst, PHX_ERR_ARG, "phx_time_interpolate: null argument") : PHX_ERR_ARG;
    double *pd, *po, *pn;
    int qd, qo, qn;
    if (!fieldOf(dst, dst_field, pd, qd) || !fieldOf(old_ctx, old_field, po, qo)
        || !fieldOf(new_ctx, new_field, pn, qn))
        return fail(dst, PHX_ERR_ARG, "phx_time_interpolate: bad field id");
    int shape[3], lo[3], n[3];
    for (int d = 0; d < 3; ++d)
    {
        shape[d] = dst->G.shape[qd][d];
        if (old_ctx->G.shape[qo][d] != shape[d] || new_ctx->G.shape[qn][d] != shape[d])
            return fail(dst, PHX_ERR_ARG, "phx_time_interpolate: the three fields differ in shape");
        lo[d] = d < dst->cfg.dim ? lower[d] : 0;
        int const hi = d < dst->cfg.dim ? upper[d] : 0;
        if (lo[d] < 0 || hi >= shape[d] || hi < lo[d])
            return fail(dst, PHX_ERR_ARG, "phx_time_interpolate: the box leaves the field");
        n[d] = hi - lo[d] + 1;
    }
    if (old_ctx != dst)
        PHX_HIP(dst, hipStreamSynchronize(old_ctx->stream));
    if (new_ctx != dst && new_ctx != old_ctx)
        PHX_HIP(dst, hipStreamSynchronize(new_ctx->stream));
    launchTimeInterp(po, pn, pd, shape, lo, n, alpha, dst->stream);
    return checkLaunch(dst, "phx_time_interpolate");
}

phx_status phx_prepare_step(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    for (int k = 0; k < 3; ++k)
        launchCopy(c->em[PHX_EM][3 + k], c->Bold[k], count(c, QBX + k), c->stream);
    return checkLaunch(c, "phx_prepare_step");
}

phx_status phx_reset_flux_sum(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    for (int k = 0; k < 3; ++k)
        launchZero(c->fluxSum[k], count(c, QEX + k), c->stream);
    return checkLaunch(c, "phx_reset_flux_sum");
}

phx_status phx_accumulate_flux_sum(phx_ctx* c, double coef)
{
    if (!c)
        return PHX_ERR_ARG;
    for (int k = 0; k < 3; ++k)
        launchAxpy(c->fluxSum[k], c->em[PHX_EM_AVG][k], coef, count(c, QEX + k), c->stream);
    return checkLaunch(c, "phx_accumulate_flux_sum");
}

phx_status phx_reflux(phx_ctx* c, double dt)
{
    if (!c)
        return PHX_ERR_ARG;
    launchFaraday(c->G, cvec(c->Bold), cvec(c->em[PHX_EM_AVG]), vec(c->em[PHX_EM] + 3), dt, c->stream);
    fillGhosts(c, c->em[PHX_EM] + 3, QBX);
    return checkLaunch(c, "phx_reflux");
}

phx_status phx_coarsen(phx_ctx* coarse, int coarse_field, phx_ctx const* fine, int fine_field, int op,
                       int const lower[3], int const upper[3])
{
    if (!coarse || !fine || !lower || !upper)
        return coarse ? fail(coarse, PHX_ERR_ARG, "phx_coarsen: null argument") : PHX_ERR_ARG;
    if (op != PHX_COARSEN_MOMENTS && op != PHX_COARSEN_ELECTRIC)
        return fail(coarse, PHX_ERR_ARG, "phx_coarsen: bad operator");
    double *pc, *pf;
    int qc, qf;
    if (!fieldOf(coarse, coarse_field, pc, qc) || !fieldOf(fine, fine_field, pf, qf) || qc != qf)
        return fail(coarse, PHX_ERR_ARG, "phx_coarsen: bad field ids (one quantity on both levels)");
    int const D = coarse->cfg.dim;
    if (fine->cfg.dim != D)
        return fail(coarse, PHX_ERR_ARG, "phx_coarsen: the levels differ in dimension");
    int dualDir = -1;
    for (int d = 0; d < D; ++d)
        if (centering(qc, d) == CD)
        {
            if (op == PHX_COARSEN_MOMENTS || dualDir >= 0)
                return fail(coarse, PHX_ERR_ARG,
                            op == PHX_COARSEN_MOMENTS ? "phx_coarsen: the moments operator takes "
                                                        "all-primal quantities"
                                                      : "phx_coarsen: the electric operator takes "
                                                        "E-centered quantities");
            dualDir = d;
        }
    int cshape[3], fshape[3], clo[3], n[3], fo[3];
    for (int d = 0; d < 3; ++d)
    {
        cshape[d] = coarse->G.shape[qc][d];
        fshape[d] = fine->G.shape[qf][d];
        if (d >= D)
        {
            clo[d] = 0;
            n[d]   = 1;
            fo[d]  = 0;
            continue;
        }
        int const cOrg = coarse->G.lower[d] - coarse->G.g; // AMR index of local index 0
        int const fOrg = fine->G.lower[d] - fine->G.g;
        clo[d]         = lower[d] - cOrg;
        int const chi  = upper[d] - cOrg;
        n[d]           = chi - clo[d] + 1;
        fo[d]          = 2 * cOrg - fOrg;
        int const fLast = 2 * chi + fo[d] + (dualDir == d ? 1 : 0);
        if (clo[d] < 0 || chi >= cshape[d] || n[d] < 1 || 2 * clo[d] + fo[d] < 0 || fLast >= fshape[d])
            return fail(coarse, PHX_ERR_ARG, "phx_coarsen: the box leaves the coarse or the fine field");
    }
    PHX_HIP(coarse, hipStreamSynchronize(fine->stream));
    launchCoarsen(pf, fshape, pc, cshape, clo, n, fo, dualDir, D, coarse->stream);
    return checkLaunch(coarse, "phx_coarsen");
}

namespace
{
int floorHalf(int i) { return (i - (i < 0 ? 1 : 0)) / 2; }
} // namespace

phx_status phx_refine_field(phx_ctx* fine, int fine_field, phx_ctx const* coarse, int coarse_field,
                            int op, int const lower[3], int const upper[3])
{
    if (!fine || !coarse || !lower || !upper)
        return fine ? fail(fine, PHX_ERR_ARG, "phx_refine_field: null argument") : PHX_ERR_ARG;
    if (op < PHX_REFINE_LINEAR || op > PHX_REFINE_MAGNETIC_INIT)
        return fail(fine, PHX_ERR_ARG, "phx_refine_field: bad operator");
    double *pf, *pc;
    int qf, qc;
    if (!fieldOf(fine, fine_field, pf, qf) || !fieldOf(coarse, coarse_field, pc, qc) || qf != qc)
        return fail(fine, PHX_ERR_ARG, "phx_refine_field: bad field ids (one quantity on both levels)");
    int const D = fine->cfg.dim;
    if (coarse->cfg.dim != D)
        return fail(fine, PHX_ERR_ARG, "phx_refine_field: the levels differ in dimension");
    bool const isE = qf >= QEX && qf <= QJZ;
    bool const isB = qf >= QBX && qf <= QBZ;
    if (op == PHX_REFINE_ELECTRIC && !isE)
        return fail(fine, PHX_ERR_ARG, "phx_refine_field: the electric operator takes E-centered "
                                       "quantities (E, J)");
    if ((op == PHX_REFINE_MAGNETIC || op == PHX_REFINE_MAGNETIC_INIT) && !isB)
        return fail(fine, PHX_ERR_ARG, "phx_refine_field: the magnetic operators take B");
    RefineArgs A{};
    A.fine   = pf;
    A.coarse = pc;
    A.fs1    = fine->G.shape[qf][1];
    A.fs2    = fine->G.shape[qf][2];
    A.cs1    = coarse->G.shape[qc][1];
    A.cs2    = coarse->G.shape[qc][2];
    A.op     = op;
    A.dim    = D;
    for (int d = 0; d < 3; ++d)
    {
        A.primal[d] = centering(qf, d) == CP ? 1 : 0;
        if (d >= D)
        {
            A.fOrg[d] = A.cOrg[d] = A.lo[d] = 0;
            A.n[d]                         = 1;
            continue;
        }
        A.fOrg[d] = fine->G.lower[d] - fine->G.g;
        A.cOrg[d] = coarse->G.lower[d] - coarse->G.g;
        A.lo[d]   = lower[d];
        A.n[d]    = upper[d] - lower[d] + 1;
        if (A.n[d] < 1 || lower[d] - A.fOrg[d] < 0 || upper[d] - A.fOrg[d] >= fine->G.shape[qf][d])
            return fail(fine, PHX_ERR_ARG, "phx_refine_field: the box leaves the fine field");
        // the coarse nodes the box reads
        bool const p = A.primal[d] != 0;
        int cmin, cmax;
        if (op == PHX_REFINE_LINEAR || op == PHX_REFINE_LINEAR_OVERWRITE)
        {
            cmin = p ? floorHalf(lower[d]) : floorHalf(lower[d] - 1);
            cmax = (p ? floorHalf(upper[d]) : floorHalf(upper[d] - 1)) + 1;
        }
        else
        {
            cmin = floorHalf(lower[d]);
            cmax = op == PHX_REFINE_ELECTRIC && p && D >= 2 ? floorHalf(upper[d] + 1)
                                                           : floorHalf(upper[d]);
        }
        if (cmin - A.cOrg[d] < 0 || cmax - A.cOrg[d] >= coarse->G.shape[qc][d])
            return fail(fine, PHX_ERR_ARG, "phx_refine_field: the box reads beyond the coarse field");
    }
    if (coarse != fine)
        PHX_HIP(fine, hipStreamSynchronize(coarse->stream));
    launchRefineField(A, fine->stream);
    return checkLaunch(fine, "phx_refine_field");
}

phx_status phx_set_ghosts_nan(phx_ctx* ctx, int field_id)
{
    if (!ctx)
        return PHX_ERR_ARG;
    double* p;
    int q;
    if (!fieldOf(ctx, field_id, p, q))
        return fail(ctx, PHX_ERR_ARG, "phx_set_ghosts_nan: bad field id");
    launchNanGhosts(ctx->G, q, p, ctx->stream);
    return checkLaunch(ctx, "phx_set_ghosts_nan");
}

phx_status phx_refine_magnetic_new_faces(phx_ctx* fine, int b_field_x, int const cell_lower[3],
                                         int const cell_upper[3])
{
    if (!fine || !cell_lower || !cell_upper)
        return fine ? fail(fine, PHX_ERR_ARG, "phx_refine_magnetic_new_faces: null argument")
                    : PHX_ERR_ARG;
    NewFacesArgs A{};
    for (int k = 0; k < 3; ++k)
    {
        int q;
        if (!fieldOf(fine, b_field_x + k, A.b[k], q) || q != QBX + k)
            return fail(fine, PHX_ERR_ARG,
                        "phx_refine_magnetic_new_faces: b_field_x must name the x component of a "
                        "magnetic field (B of a set, or Bold)");
        A.s1[k] = fine->G.shape[QBX + k][1];
        A.s2[k] = fine->G.shape[QBX + k][2];
        for (int d = 0; d < 3; ++d)
            A.shape[k][d] = fine->G.shape[QBX + k][d];
    }
    int const D = fine->cfg.dim;
    A.dim       = D;
    for (int d = 0; d < 3; ++d)
    {
        A.dx[d] = fine->G.dx[d];
        if (d >= D)
        {
            A.org[d] = A.lo[d] = A.hi[d] = 0;
            continue;
        }
        A.org[d] = fine->G.lower[d] - fine->G.g;
        A.lo[d]  = cell_lower[d];
        A.hi[d]  = cell_upper[d];
        if (A.hi[d] < A.lo[d])
            return fail(fine, PHX_ERR_ARG, "phx_refine_magnetic_new_faces: empty box");
    }
    // the box's cells within the arrays (faces whose stencil leaves them take the edge rule)
    for (int d = 0; d < D; ++d)
        if (A.lo[d] - A.org[d] < 0 || A.hi[d] - A.org[d] >= fine->G.shape[QBX + (d + 1) % 3][d])
            return fail(fine, PHX_ERR_ARG, "phx_refine_magnetic_new_faces: the box leaves the B arrays");
    launchMagneticNewFaces(A, fine->stream);
    return checkLaunch(fine, "phx_refine_magnetic_new_faces");
}

phx_status phx_update_ions(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    updateIons(c);
    return checkLaunch(c, "updateIons");
}

phx_status phx_fill_moment_borders(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    return fillMomentBorders(c);
}

phx_status phx_fill_ion_borders(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    return fillIonBorders(c);
}

phx_status phx_fill_ion_ghost_particles(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    int const D = c->cfg.dim;
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        Store& pg  = c->pops[p].arr[PHX_PATCH_GHOST];
        Store& dom = c->pops[p].arr[PHX_DOMAIN];
        if (c->patch.periodic && pg.n)
        {
            PHX_TRY(storeReserve(c, dom, dom.n + pg.n));
            PHX_TRY(scratchReserve(c, pg.n));
            PHX_HIP(c, hipMemsetAsync(c->status, ST_KEEP, pg.n, c->stream));
            Wrap wrap{};
            wrap.on = 1;
            for (int d = 0; d < 3; ++d)
            {
                wrap.n[d]  = c->G.n[d];
                wrap.lo[d] = c->boxes.dom_lo[d];
                wrap.hi[d] = c->boxes.dom_hi[d];
            }
            launchCompact(pg.view(), dom.view(), c->status, pg.n, ST_KEEP, c->blockCnt, c->tot, dom.n,
                          D, wrap, c->stream);
            dom.n += pg.n;
            dom.seg = false;
        }
        pg.n = 0;
    }
    return checkLaunch(c, "fillIonGhostParticles");
}

phx_status phx_fill_ghosts(phx_ctx* c, int fid)
{
    double* p;
    int q;
    if (!c || !fieldOf(c, fid, p, q))
        return fail(c, PHX_ERR_ARG, "bad field id");
    if (!c->patch.periodic)
        return PHX_OK;
    for (int k = 0; k < 3; ++k)
    {
        double* pk;
        int qk;
        if (!fieldOf(c, fid + k, pk, qk))
            return fail(c, PHX_ERR_ARG, "bad vector field id");
        launchFillGhosts(c->G, qk, pk, c->stream);
    }
    return checkLaunch(c, "fillGhosts");
}

phx_status phx_faraday(phx_ctx* c, int bset, int eset, int bnewset, double dt)
{
    if (!c || bset < 0 || bset > 2 || eset < 0 || eset > 2 || bnewset < 0 || bnewset > 2)
        return fail(c, PHX_ERR_ARG, "bad field set");
    launchFaraday(c->G, cvec(c->em[bset] + 3), cvec(c->em[eset]), vec(c->em[bnewset] + 3), dt, c->stream);
    return checkLaunch(c, "faraday");
}

phx_status phx_ampere(phx_ctx* c, int bset)
{
    if (!c || bset < 0 || bset > 2)
        return fail(c, PHX_ERR_ARG, "bad field set");
    launchAmpere(c->G, cvec(c->em[bset] + 3), vec(c->J), c->stream);
    return checkLaunch(c, "ampere");
}

phx_status phx_electrons_update(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    electrons(c);
    return checkLaunch(c, "electrons");
}

phx_status phx_ohm(phx_ctx* c, int bset, int eset)
{
    if (!c || bset < 0 || bset > 2 || eset < 0 || eset > 2)
        return fail(c, PHX_ERR_ARG, "bad field set");
    ohm(c, bset, eset);
    return checkLaunch(c, "ohm");
}

phx_status phx_average(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    enqueueAverage(c);
    return checkLaunch(c, "average");
}

phx_status phx_move_ions(phx_ctx* c, double dt, int mode)
{
    if (!c)
        return PHX_ERR_ARG;
    if (mode != PHX_MODE_DOMAIN_ONLY && mode != PHX_MODE_ALL)
        return fail(c, PHX_ERR_ARG, "bad mode");
    if (fastPath(c) && staged3d(c))
    {
        // the averaged set may have been set or computed outside the average kernel
        if (mixed(c))
            launchEmToF32(c->G, c->em[PHX_EM_AVG], c->em32, c->stream);
        PHX_TRY(moveIonsStaged3d(c, dt, mode));
        return checkDeviceError(c);
    }
    if (fastPath(c))
    {
        // the averaged set may have been set or computed outside the average kernel
        PHX_TRY(sortIfLinear(c));
        if (mixed(c))
            launchEmToF32(c->G, c->em[PHX_EM_AVG], c->em32, c->stream);
        PHX_TRY(enqueueMoveIonsFast(c, dt, mode));
        return checkDeviceError(c);
    }
    PHX_TRY(phx_update_populations(c, PHX_EM_AVG, dt, mode));
    PHX_TRY(fillMomentBorders(c));
    if (mode == PHX_MODE_ALL)
        PHX_TRY(phx_fill_ion_ghost_particles(c));
    updateIons(c);
    PHX_TRY(fillIonBorders(c));
    return checkDeviceError(c);
}


phx_status phx_sort_particles(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    return sortDomain(c);
}

phx_status phx_set_sort_interval(phx_ctx* c, int steps)
{
    if (!c || steps < 0)
        return PHX_ERR_ARG;
    c->sortInterval          = steps;
    c->tuning.sort_interval = steps;
    return PHX_OK;
}

phx_status phx_get_tuning(phx_ctx const* c, phx_tuning* out)
{
    if (!c || !out)
        return PHX_ERR_ARG;
    *out = c->tuning;
    return PHX_OK;
}

phx_status phx_set_tuning(phx_ctx* c, phx_tuning const* t)
{
    if (!c || !t)
        return PHX_ERR_ARG;
    int const D       = c->cfg.dim;
    phx_tuning next   = *t;
    phx_tuning const d = defaultTuning(D, c->cfg.interp_order);
    int cells         = 1;
    for (int k = 0; k < 3; ++k)
    {
        if (next.segment_shape[k] == 0 || k >= D)
            next.segment_shape[k] = k < D ? d.segment_shape[k] : 1;
        if (next.segment_shape[k] < 1 || next.segment_shape[k] > 64)
            return fail(c, PHX_ERR_ARG, "phx_set_tuning: segment extents are 1..64 cells");
        cells *= next.segment_shape[k];
    }
    if (cells > 64)
        return fail(c, PHX_ERR_ARG, "phx_set_tuning: a segment holds at most 64 cells");
    if (next.rows_per_item < 1 || next.rows_per_item > 4096 || next.tile_row_pad < 1
        || next.tile_row_pad > 64 || next.sort_interval < 0
        || (next.tile_copies != 1 && next.tile_copies != 2 && next.tile_copies != 4
            && next.tile_copies != 8)
        || (next.deposit_window != 0 && next.deposit_window != 1))
        return fail(c, PHX_ERR_ARG, "phx_set_tuning: rows per item 1..4096, row pad 1..64, "
                                    "sort interval >= 0, tile copies 1, 2, 4 or 8, deposit "
                                    "window 0 or 1");
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    // the captured steps address the segment tables: drop them
    for (auto& g : c->graphs)
    {
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
        g = phx_ctx::GraphEntry{};
    }
    c->tuning       = next;
    c->sortInterval = next.sort_interval;
    std::vector<int> plo, phi;
    for (auto const& b : c->patchBox)
    {
        plo.insert(plo.end(), {b[0], b[1], b[2]});
        phi.insert(phi.end(), {b[3], b[4], b[5]});
    }
    PHX_TRY(buildPatchTable(c, plo.data(), phi.data(), static_cast<int>(c->patchBox.size())));
    PHX_TRY(allocSegTables(c));
    bool segmented = false;
    for (int p = 0; p < c->cfg.n_pop; ++p)
        segmented = segmented || c->pops[p].arr[PHX_DOMAIN].seg;
    if (segmented)
        PHX_TRY(sortDomain(c));
    return PHX_OK;
}

phx_status phx_initialize(phx_ctx* c)
{
    if (!c)
        return PHX_ERR_ARG;
    fillGhosts(c, c->em[PHX_EM] + 3, QBX);
    PHX_TRY(sortDomain(c));
    zeroMoments(c);
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        PushArgs A = pushArgs(c, p, KIND_DOMAIN, PHX_MODE_ALL, PHX_EM, 0., true);
        if (staged3d(c))
        {
            A.pushOn = 0;
            enqueuePushStaged3d(c, p, A);
        }
        else if (mixed(c))
        {
            A.pushOn = 0;
            launchPush32(A, c->pops[p].arr[PHX_DOMAIN].view32(), c->stream);
        }
        else
            launchDeposit(A, c->pops[p].arr[PHX_DOMAIN].view(), c->stream);
    }
    PHX_TRY(checkLaunch(c, "initial deposit"));
    PHX_TRY(fillMomentBorders(c));
    updateIons(c);
    PHX_TRY(fillIonBorders(c));
    launchAmpere(c->G, cvec(c->em[PHX_EM] + 3), vec(c->J), c->stream);
    fillGhosts(c, c->J, QJX);
    electrons(c);
    ohm(c, PHX_EM, PHX_EM);
    fillGhosts(c, c->em[PHX_EM], QEX);
    PHX_TRY(checkLaunch(c, "initialize"));
    return checkDeviceError(c);
}

phx_status phx_advance(phx_ctx* c, double dt, int nsteps)
{
    if (!c || nsteps < 0)
        return PHX_ERR_ARG;
    if (!fastPath(c))
    {
        // general level: the staged sequence of SolverPPC::advanceLevel
        for (int s = 0; s < nsteps; ++s)
        {
            enqueuePredictor(c, dt, PHX_EM);
            enqueueAverage(c);
            PHX_TRY(phx_move_ions(c, dt, PHX_MODE_DOMAIN_ONLY));
            enqueuePredictor(c, dt, PHX_EM_AVG);
            enqueueAverage(c);
            PHX_TRY(phx_move_ions(c, dt, PHX_MODE_ALL));
            enqueueCorrector(c, dt);
            if (c->sortInterval > 0 && ++c->stepsSinceSort >= c->sortInterval)
                PHX_TRY(sortDomain(c));
        }
        return checkDeviceError(c);
    }
    if (staged3d(c))
    {
        for (int s = 0; s < nsteps; ++s)
        {
            PHX_TRY(advanceStaged3d(c, dt));
            if (c->sortInterval > 0 && ++c->stepsSinceSort >= c->sortInterval)
                PHX_TRY(sortDomain(c));
        }
        return checkDeviceError(c);
    }
    PHX_TRY(sortIfLinear(c));
    for (int s = 0; s < nsteps; ++s)
    {
        phx_ctx::GraphEntry* g = nullptr;
        PHX_TRY(stepGraph(c, dt, &g));
        PHX_HIP(c, hipGraphLaunch(g->exec, c->stream));
        if (c->sortInterval > 0 && ++c->stepsSinceSort >= c->sortInterval)
            PHX_TRY(sortDomain(c));
    }
    // a two-cell move or a stencil out of the ghost box inside the steps is reported here (the
    // steps run asynchronously; the first error is kept with its particle and stage)
    return checkDeviceError(c);
}

phx_status phx_diagnostics(phx_ctx* c, double out[4])
{
    if (!c || !out)
        return PHX_ERR_ARG;
    PHX_TRY(checkDeviceError(c));
    int const D = c->cfg.dim;
    double vol  = 1.;
    for (int d = 0; d < D; ++d)
        vol *= c->G.dx[d];
    double kin = 0., charge = 0., cnt = 0.;
    for (int p = 0; p < c->cfg.n_pop; ++p)
    {
        std::vector<int32_t> ic;
        std::vector<double> de, v, w, q;
        PHX_TRY(download(c, c->pops[p], c->pops[p].arr[PHX_DOMAIN], ic, de, v, w, q));
        double k = 0.;
        for (size_t i = 0; i < w.size(); ++i)
        {
            k += w[i] * (v[i * 3] * v[i * 3] + v[i * 3 + 1] * v[i * 3 + 1] + v[i * 3 + 2] * v[i * 3 + 2]);
            charge += q[i] * w[i];
        }
        kin += 0.5 * c->pops[p].mass * k;
        cnt += static_cast<double>(w.size());
    }
    double mag = 0.;
    for (int k = 0; k < 3; ++k)
    {
        int q = QBX + k;
        std::vector<double> f(static_cast<size_t>(c->G.size(q)));
        PHX_HIP(c, hipMemcpy(f.data(), c->em[PHX_EM][3 + k], f.size() * 8, hipMemcpyDeviceToHost));
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d)
        {
            lo[d] = c->G.g;
            hi[d] = c->G.g + c->G.n[d] - 1;
        }
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int kk = lo[2]; kk <= hi[2]; ++kk)
                {
                    double b = f[static_cast<size_t>(c->G.idx(q, i, j, kk))];
                    mag += b * b;
                }
    }
    out[0] = kin * vol;
    out[1] = 0.5 * mag * vol;
    out[2] = charge * vol;
    out[3] = cnt;
    return PHX_OK;
}

phx_status phx_set_push_timing(phx_ctx* c, int on)
{
    if (!c)
        return PHX_ERR_ARG;
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    if (on && !c->kt)
        PHX_TRY(devAlloc(c, (void**)&c->kt, sizeof(KTime) * 2 * PHX_MAX_POP));
    if (on)
    {
        KTime z[2 * PHX_MAX_POP];
        for (auto& k : z)
            k = KTime{~0ull, 0ull, 0ull, 0u, 0u};
        PHX_HIP(c, hipMemcpy(c->kt, z, sizeof(z), hipMemcpyHostToDevice));
    }
    c->timing     = on != 0;
    c->eventsUsed = 0;
    return PHX_OK;
}

phx_status phx_push_timing(phx_ctx* c, int pop, int mode, double* avg_ms, int* launches)
{
    if (!c || !avg_ms || !launches || pop < 0 || pop >= c->cfg.n_pop
        || (mode != PHX_MODE_DOMAIN_ONLY && mode != PHX_MODE_ALL))
        return PHX_ERR_ARG;
    if (!c->kt)
        return fail(c, PHX_ERR_STATE, "push timing was never enabled");
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    if (staged3d(c))
    {
        // the event pairs of the staged 3D path's pushes
        double total = 0.;
        int n        = 0;
        for (size_t i = 0; i < c->eventsUsed; ++i)
        {
            auto const& e = c->events[i];
            if (e.mode != mode || e.pop != pop)
                continue;
            float ms = 0.f;
            PHX_HIP(c, hipEventElapsedTime(&ms, e.a, e.b));
            total += ms;
            ++n;
        }
        *launches = n;
        *avg_ms   = n ? total / n : 0.;
        return PHX_OK;
    }
    KTime k{};
    PHX_HIP(c, hipMemcpy(&k, c->kt + (mode == PHX_MODE_ALL ? PHX_MAX_POP : 0) + pop, sizeof(k),
                         hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    PHX_HIP(c, hipGetDevice(&dev));
    PHX_HIP(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    *launches = static_cast<int>(k.count);
    *avg_ms   = k.count && khz ? static_cast<double>(k.total) / k.count / khz : 0.;
    return PHX_OK;
}

phx_status phx_profile_push(phx_ctx* c, double dt, int mode, int reps, double* avg_ms)
{
    if (!c || !avg_ms || reps < 1 || (mode != PHX_MODE_DOMAIN_ONLY && mode != PHX_MODE_ALL))
        return PHX_ERR_ARG;
    if (!fastPath(c))
        return fail(c, PHX_ERR_STATE, "phx_profile_push needs a periodic level without ghost arrays");
    if (staged3d(c))
        return fail(c, PHX_ERR_STATE,
                    "phx_profile_push: a 3D level runs the staged path (its pushes are timed by "
                    "phx_set_push_timing)");
    // the push kernel alone, captured in a graph and replayed back to back between two events
    // on the context stream (the mean over the replays is the kernel time; per-launch event
    // pairs around direct launches read high on this runtime)
    PHX_TRY(sortIfLinear(c));
    PushArgs A     = pushArgs(c, 0, KIND_DOMAIN, mode, PHX_EM_AVG, dt, true);
    A.periodicWrap = 1;
    Store const& dom = c->pops[0].arr[PHX_DOMAIN];
    if (mixed(c) && !A.segOn)
        return fail(c, PHX_ERR_STATE, "mixed-precision push on unsorted storage");
    if (mixed(c))
        launchEmToF32(c->G, c->em[PHX_EM_AVG], c->em32, c->stream);
    PHX_HIP(c, hipStreamSynchronize(c->stream));
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    PHX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (mixed(c))
        launchPush32(A, dom.view32(), c->stream);
    else
        launchPush(A, dom.view(), dom.view(), c->stream);
    PHX_HIP(c, hipStreamEndCapture(c->stream, &g));
    PHX_HIP(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    hipEvent_t a = nullptr, b = nullptr;
    PHX_HIP(c, hipEventCreate(&a));
    PHX_HIP(c, hipEventCreate(&b));
    PHX_HIP(c, hipGraphLaunch(ge, c->stream)); // warm-up
    PHX_HIP(c, hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; ++r)
        PHX_HIP(c, hipGraphLaunch(ge, c->stream));
    PHX_HIP(c, hipEventRecord(b, c->stream));
    PHX_HIP(c, hipEventSynchronize(b));
    float ms = 0.f;
    PHX_HIP(c, hipEventElapsedTime(&ms, a, b));
    double total = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipGraphExecDestroy(ge);
    *avg_ms = total / reps;
    return checkDeviceError(c);
}

} // extern "C"
