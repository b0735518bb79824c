// phx_kernels.hpp -- kernel argument blocks and launchers shared by the HIP translation units.
#pragma once

#include "phx_device.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace phx
{

constexpr int DEAD_CELL = -0x40000000; // iCell[0] of a particle lost beyond the ghost box
constexpr int MAX_NLGB  = 27;

enum : int
{
    KIND_DOMAIN = 0,
    KIND_PATCH_GHOST = 1,
    KIND_LEVEL_GHOST = 2
};
enum : int
{
    MODE_DOMAIN_ONLY = 1,
    MODE_ALL         = 2
};
// per-particle fate after a push in mode all
enum : uint8_t
{
    ST_KEEP   = 0, // stays in its array
    ST_LEAVE  = 1, // domain particle now in the patch ghost layer -> patchGhost
    ST_DROP   = 2, // outside every box that keeps it
    ST_EXPORT = 3  // level-ghost particle that entered the domain -> domain
};

// structure-of-arrays view of a particle array: R = double for the plain FP64 family (per-particle
// charge), float for the mixed-precision family (FP32 delta, velocity and weight; the charge is
// the population's, q is null)
template<class R>
struct PViewT
{
    int32_t* ic[3];
    R* de[3];
    R* v[3];
    R* w;
    R* q;
};
using PView   = PViewT<double>;
using PView32 = PViewT<float>;

struct Boxes
{
    int dom_lo[3], dom_hi[3]; // AMR domain box
    int gh_lo[3], gh_hi[3];   // domain grown by the particle ghost width
    int nnlgb;
    int nlgb_lo[MAX_NLGB][3], nlgb_hi[MAX_NLGB][3];
};

// LDS tile of one workgroup: a box of nodes placed over the chunk's cells, holding `copies`
// copies of the 5 moment accumulators. Every slab (one moment of one copy) is `stride` doubles
// long. The linear push's tile takes stride = nodes rounded up to 16, plus 1, so that the copies
// of a node sit in different LDS banks; the segmented push's single-copy tile takes nodes rounded
// up to 16.
struct TileSpec
{
    int lo[3];   // margin below the lowest cell of the chunk (nodes)
    int ext[3];  // extents
    int pitch[3]; // node (i, j, k) sits at (i * pitch[1] + j) * pitch[2] + k (pitch[d] >= ext[d])
    int nodes;  // ext[0]*pitch[1]*pitch[2]
    int copies; // moment copies (power of two): lane l accumulates into copy l % copies
    int stride;
    unsigned int bytes; // dynamic LDS of the launch
    int brick;          // > 0: the tile covers the brick (cube of `brick` cells, on the grid of
                        // bricks from the patch's lower cell) of the chunk's first particle
};
// cells per side of the bricks of the linear 3D push: brick-ordered arrays put the particles of
// one brick (BRICK^3 = 64 cells, one per lane of a wavefront) in one chunk
constexpr int BRICK = 4;

// Cell-segmented storage of a level's domain array: every patch of the level (a device patch
// table, patches of uneven sizes) is cut into segments, boxes of at most 64 cells (shape `seg`),
// numbered patch after patch; a segment holding P particles has R = ceil(P / 64) rows of 64 slots.
// Its particles, in the order of their cells (row-major inside the segment), take the slots
// column by column: position p is slot p / R of row p % R. The slots past P hold dead particles.
// One more segment at the end holds, unordered, the particles outside every patch. A particle
// belongs to the patch whose domain box holds its cell (the owner map), so a sort also migrates
// the particles that crossed into a neighbouring patch.
struct SegPatch
{
    int lo[3];             // lowest AMR cell of the patch's domain box
    int n[3];              // cells per dimension
    int seg[3];            // segment shape in cells
    int nseg[3];           // segments per dimension
    unsigned int segBase;  // number of the patch's first segment
    unsigned int cellBase; // key of the patch's first cell (keys are row-major inside a patch)
};
struct SegGeom
{
    int dim;
    int npatch;
    int nsegs;     // segments of all patches (the overflow segment is number nsegs)
    int ncells;    // cells of all patches
    int lo[3];     // lowest AMR cell of the level box (the box of the owner map)
    int n[3];      // level box cells per dimension
    int maxSeg[3]; // largest segment extent over the patches (the launch's LDS tile)
    int rowsPerItem; // most segment rows per work item of the segmented push
    int rowPad;      // 2D LDS tile rows padded to a multiple of this many nodes
    SegPatch first;               // patch 0 (a one-patch level reads it from the kernel arguments)
    SegPatch const* patch;        // [npatch] (device)
    int const* owner;             // [level box cells, row-major] patch of each cell, -1: none
    unsigned int const* segPatch; // [nsegs] patch of each segment
};
// tables of the current segmented storage, rebuilt by each sort
struct SegTables
{
    unsigned int* rows;       // [nsegs + 1] rows per segment
    unsigned long long* off;  // [nsegs + 2] first slot per segment; off[nsegs + 1] = all slots
    unsigned int* items;      // [2 * itemsCap] (segment, first row) of each work item
    unsigned int* nItems;     // number of work items
    unsigned int itemsCap;    // work items the storage's capacity allows (grid of the push)
};

// in-kernel timing of one push launch site (a node of the captured step): the earliest
// workgroup start and the latest workgroup end of a launch, in wall-clock ticks; the launch's last
// workgroup adds end - start to the total and re-arms the record for the next launch
struct KTime
{
    unsigned long long start; // ~0 between launches
    unsigned long long end;
    unsigned long long total;
    unsigned int done;  // workgroups of the current launch that have finished
    unsigned int count; // launches in total
};

struct PushArgs
{
    Geom G;
    EMPtrs em;
    double* mom[5]; // n, rho, Fx, Fy, Fz of the population
    double hdtdx[3];
    double dto2m;
    double coef;
    Boxes box;
    int kind, mode, periodicWrap, pop;
    int pushOn; // 0: deposit only (unpushed arrays, initial moments)
    TileSpec tile;
    unsigned long long n;
    DevError* err;
    uint8_t* status;
    unsigned int* dropped;
    int segOn; // the array is cell-segmented (sg, st valid)
    SegGeom sg;
    SegTables st;
    float const* em32[6]; // mixed-precision family: FP32 copy of the E/B set (Ex..Bz)
    KTime* kt;            // timing record of this launch site (null: not timed)
    double charge; // the population's charge (mixed-precision family; FP64 family with rhoFromN)
    int rhoFromN;  // FP64 segmented push: every particle has charge `charge`, so the LDS tile
                   // leaves the charge density out and its flush takes charge x n
    int depositWindow; // mixed-precision segmented push: sum each lane's successive particles of
                       // one cell in a register window before the LDS tile (1D, 2D, 3D order 1)
};

// box list of a particle selector (UpdaterSelectionBoxing, ion_updater.hpp:119-163): a cell is
// selected when it lies in one of the n boxes (inclusive AMR cells)
struct SelBoxes
{
    int n;
    int lo[MAX_NLGB][3], hi[MAX_NLGB][3];
};

struct Wrap
{
    int on;
    int n[3], lo[3], hi[3];
};

void launchPush(PushArgs const& A, PView const& in, PView const& out, hipStream_t s);
void launchDeposit(PushArgs const& A, PView const& in, hipStream_t s);
// tile of a level whose arrays hold `ppc` particles per cell on average
TileSpec tileShape(int dim, int order, double ppc, int copies);
int chunkSize(int dim, int order);
int blockThreads();
unsigned int compactBlocks(unsigned long long n);

// segmented storage: segment shape of a patch box (segBase/cellBase left to the caller), LDS
// tile of the segmented push, the most work items a storage of `capacity` slots can need
SegPatch segPatchOf(int dim, int const* lo, int const* hi, int const* pref);
TileSpec segTileShape(int dim, int order, SegGeom const& sg, int copies);
unsigned int segItemsCap(SegGeom const& sg, unsigned long long capacity);
// first half of a sort of src (n slots) into cell-segmented storage: counts per cell, the rank
// of every particle in its cell (rank: n words) and the segment tables up to the total slot count
// (tables.off[nsegs + 1]); `work` holds ncells + 1 + 2 (nsegs + 1) words
void launchSegCount(SegGeom const& sg, PView const& src, unsigned long long n, unsigned int* work,
                    unsigned int* rank, SegTables const& tables, hipStream_t s);
// second half, once dst can hold the slot count: work items, the source particle of every slot
// (perm: slots words) with the empty ones marked, then dst filled slot by slot
void launchSegScatter(SegGeom const& sg, PView const& src, PView const& dst, unsigned long long n,
                      unsigned long long slots, unsigned int* work, unsigned int const* rank,
                      unsigned int* perm, SegTables const& tables, int dim, hipStream_t s);
// mixed-precision family: the same sort over FP32 arrays
void launchSegCount(SegGeom const& sg, PView32 const& src, unsigned long long n, unsigned int* work,
                    unsigned int* rank, SegTables const& tables, hipStream_t s);
void launchSegScatter(SegGeom const& sg, PView32 const& src, PView32 const& dst, unsigned long long n,
                      unsigned long long slots, unsigned int* work, unsigned int const* rank,
                      unsigned int* perm, SegTables const& tables, int dim, hipStream_t s);
// copy of the first n slots of every array (dim position components)
void launchSegCopy(PView const& src, PView const& dst, unsigned long long n, int dim, hipStream_t s);
void launchSegCopy(PView32 const& src, PView32 const& dst, unsigned long long n, int dim, hipStream_t s);
// mixed-precision push + gather + deposit over cell-segmented FP32 storage, in place (A.pushOn 0:
// deposit only); the LDS tile of segTileShape32
void launchPush32(PushArgs const& A, PView32 const& v, hipStream_t s);
TileSpec segTileShape32(int dim, int order, SegGeom const& sg);
// the same over brick-ordered linear FP32 storage (3D), in place; the LDS tile of brickTileShape32
void launchPush32Linear(PushArgs const& A, PView32 const& v, hipStream_t s);
TileSpec brickTileShape32(int order);
// push + gather + deposit over cell-segmented storage with one workgroup per segment and all its
// rows, in place (3D; the staged 3D path), FP64 and mixed precision; grid nsegs + 1; the LDS tile
// of segwgTileShape (5 slabs FP64, 4 mixed precision)
void launchPushSegWG(PushArgs const& A, PView const& in, hipStream_t s);
void launchPush32SegWG(PushArgs const& A, PView32 const& v, hipStream_t s);
TileSpec segwgTileShape(int dim, int order, SegGeom const& sg, int slabs, int pad);
// BorisPusher::move in place over A.n slots with box-list selectors; status[i] gets the slot's
// fate for the partition (0 both selectors, 1 first only, 2 neither, 3 empty slot)
void launchMove(PushArgs const& A, PView const& v, SelBoxes const& first, SelBoxes const& second,
                uint8_t* status, hipStream_t s);
// E and B interpolated at every live slot of v (A.n slots): [3 * slot + component]
void launchGather(PushArgs const& A, PView const& v, double* E, double* B, hipStream_t s);
// gather without allocation (phx_gather_noalloc): at most NOALLOC_GATHER_MAX_SLOTS slots of an
// array; its staging lives in a file-scope device array, `scratch` of phx_particles.hip, shared
// by every call on a GPU (calls that overlap on one GPU race)
constexpr unsigned long long NOALLOC_GATHER_MAX_SLOTS = 1ull << 20;
// device address of that array on the current device (looked up once per device)
hipError_t noAllocScratch(void** out);
// E/B of the live slots of v (A.n <= NOALLOC_GATHER_MAX_SLOTS), compacted in storage order into the
// array at scratchPtr; noAllocGatherOutputs gives where the E, B values and their count lie
void launchGatherNoAlloc(PushArgs const& A, PView const& v, void* scratchPtr, hipStream_t s);
void noAllocGatherOutputs(void* scratchPtr, double const** E, double const** B,
                          unsigned long long const** total);
void launchCompact(PView const& src, PView const& dst, uint8_t const* status, unsigned long long n,
                   uint8_t want, unsigned int* blockCounts, unsigned long long* total,
                   unsigned long long dstOffset, int dim, Wrap const& wrap, hipStream_t s);

// ---- coarse-to-fine coupling (phx_amr.hip) ------------------------------------------------------
constexpr int MAX_SPLIT_BOXES = 32;
constexpr int MAX_REFINED     = 27;
// one split: destination boxes (fine AMR cells, inclusive, disjoint), the split-box growth
// ceil((order + 1) / 2), the refinement ratio and the pattern of refined particles (weight and
// offset of each, in single precision as the reference's split tables hold them)
struct SplitSpec
{
    int nBoxes;
    int lo[MAX_SPLIT_BOXES][3], hi[MAX_SPLIT_BOXES][3];
    int grow;
    int ratio;
    int nRefined;
    float weight[MAX_REFINED];
    float delta[MAX_REFINED][3];
};
// the split pattern of (dim, order, refined particles) into S (nRefined, weight, delta); false
// when the combination is not supported
bool splitPattern(int dim, int order, int nbr, SplitSpec& S);
// count pass: the refined particles of src (n slots) kept in the destination boxes, per
// workgroup, scanned in place (blockCounts: one word per 256 slots), the total at *total; returns
// the workgroup count
unsigned long long launchSplitCount(PView const& src, unsigned long long n, SplitSpec const& S, int dim,
                                    unsigned int* blockCounts, unsigned long long* total, hipStream_t s);
// write pass: the refined particles appended to dst at dstOffset in coarse-particle order
void launchSplitWrite(PView const& src, unsigned long long n, SplitSpec const& S, int dim,
                      unsigned int const* offsets, PView const& dst, unsigned long long dstOffset,
                      hipStream_t s);
// out = (1 - alpha) a + alpha b over the local box lo .. lo + n - 1 of arrays of the given shape
void launchTimeInterp(double const* a, double const* b, double* out, int const* shape, int const* lo,
                      int const* n, double alpha, hipStream_t s);

// coarse node values from the fine level (ratio 2) over the coarse local box clo .. clo + n - 1:
// dualDir < 0 injects the fine node 2c (MomentsCoarsener; the primal directions of E), dualDir = d
// averages the fine nodes 2c and 2c + 1 along d (ElectricFieldCoarsener); fineOff: the fine local
// index of coarse local index 0 per direction
void launchCoarsen(double const* fine, int const* fshape, double* coarse, int const* cshape,
                   int const* clo, int const* n, int const* fineOff, int dualDir, int dim, hipStream_t s);
// a += b * coef over n values
void launchAxpy(double* a, double const* b, double coef, long n, hipStream_t s);

// coarse-to-fine refinement of one field (ratio 2) over the fine AMR box lo .. lo + n - 1 of the
// quantity's own nodes. Arrays row-major (x slowest); org: the AMR index of local index 0 per
// direction; primal[d]: the quantity's centering per direction
enum RefineOp
{
    REFINE_LINEAR           = 0, // DefaultFieldRefiner: linear weights, NaN nodes only
    REFINE_LINEAR_OVERWRITE = 1, // FieldMomentsRefiner: linear weights, every node
    REFINE_ELECTRIC         = 2, // ElectricFieldRefiner (NaN nodes only)
    REFINE_MAGNETIC         = 3, // MagneticFieldRefiner: shared faces, NaN nodes only
    REFINE_MAGNETIC_INIT    = 4  // MagneticFieldInitRefiner: shared faces, every node
};
struct RefineArgs
{
    double* fine;
    double const* coarse;
    int fs1, fs2, cs1, cs2; // strides of the two slower directions' successors (shape[1], shape[2])
    int fOrg[3], cOrg[3];
    int lo[3], n[3];
    int primal[3];
    int op, dim;
};
void launchRefineField(RefineArgs const& A, hipStream_t s);
// the new fine faces of B inside the fine AMR cell box lo .. hi (the faces with an odd index along
// their normal), from the shared faces already refined (MagneticRefinePatchStrategy); at the edge
// of the arrays, a face whose cross terms would read beyond them takes the mean of its two
// neighbours along the normal, and one without both neighbours is left as it is
struct NewFacesArgs
{
    double* b[3];
    int s1[3], s2[3]; // shape[1], shape[2] of each component's array
    int shape[3][3];  // each component's array shape (the reads' bounds)
    int org[3];
    int lo[3], hi[3];
    double dx[3];
    int dim;
};
void launchMagneticNewFaces(NewFacesArgs const& A, hipStream_t s);

// ---- fields ----------------------------------------------------------------------------------
struct VecPtr
{
    double* c[3];
};
struct CVecPtr
{
    double const* c[3];
};

void launchFaraday(Geom const& G, CVecPtr B, CVecPtr E, VecPtr Bnew, double dt, hipStream_t s);
void launchAmpere(Geom const& G, CVecPtr B, VecPtr J, hipStream_t s);
void launchElectrons(Geom const& G, double Te, CVecPtr J, CVecPtr Vi, double const* Ne, VecPtr Ve,
                     double* Pe, hipStream_t s);
void launchOhm(Geom const& G, double eta, double nu, int hyper, int level, double const* n,
               CVecPtr Ve, double const* Pe, CVecPtr B, CVecPtr J, VecPtr E, hipStream_t s,
               int fill = 0);
void launchAverage(double const* a, double const* b, double* out, long count, hipStream_t s);
// predictor Ohm (Ep on the physical nodes from Bp, J) and the average (Ea, Ba from E, Ep, B, Bp)
// of a periodic level in one launch; mirror as launchAverageFill
void launchOhmAverage(Geom const& G, double eta, double nu, int hyper, int level, double const* n,
                      CVecPtr Ve, double const* Pe, CVecPtr Bp, CVecPtr J, VecPtr Ep, CVecPtr E,
                      CVecPtr B, VecPtr Ea, VecPtr Ba, hipStream_t s, float* const* mirror = nullptr);
void launchFillGhosts(Geom const& G, int q, double* f, hipStream_t s);
// every node of f (quantity q) outside the physical nodes set to NaN
void launchNanGhosts(Geom const& G, int q, double* f, hipStream_t s);
// out = orig combined (sum / max) with its periodic images on the nodes they overlap; out must
// already hold orig (the nodes no image overlaps are not written)
void launchBorderSum(Geom const& G, double const* orig, double* out, hipStream_t s);
void launchBorderMax(Geom const& G, double const* orig, double* inout, hipStream_t s);
void launchZero(double* p, long count, hipStream_t s);
void launchCopy(double const* src, double* dst, long count, hipStream_t s);
struct PopMoments
{
    double const* n[4];
    double const* rho[4];
    double const* F[4][3];
    double mass[4];
    int npop;
};
void launchIonTotals(PopMoments const& P, double* rho, double* massDensity, VecPtr Vi, long count,
                     hipStream_t s);

// fused kernels of the periodic fast path
struct MomSetArgs
{
    double const* raw[4][5];
    double* mom[4][5];
    double* zero[4][5];
    double mass[4];
    int npop;
};
void launchFaradayFill(Geom const& G, CVecPtr B, CVecPtr E, VecPtr Bnew, double dt, hipStream_t s);
void launchAmpereFill(Geom const& G, CVecPtr B, VecPtr J, hipStream_t s);
void launchFill3(Geom const& G, int q0, VecPtr F, hipStream_t s);
// mirror (nullable): 6 FP32 arrays (Ex..Bz) that also receive the averaged set, for the
// mixed-precision push's gather
void launchAverageFill(Geom const& G, CVecPtr E, CVecPtr Ep, CVecPtr B, CVecPtr Bp, VecPtr Ea,
                       VecPtr Ba, hipStream_t s, float* const* mirror = nullptr);
// FP32 copy of an E/B set (6 components, Ex..Bz)
void launchEmToF32(Geom const& G, double const* const* src, float* const* dst, hipStream_t s);
void launchFoldTotals(Geom const& G, MomSetArgs const& A, double* trho, double* tmd, VecPtr tvi,
                      hipStream_t s);
void launchMax5(Geom const& G, double const* const* src, double* const* dst, hipStream_t s);
// launchFoldTotals and the Faraday that follows it in a captured step, in one launch: B, E -> Bn
// with Bn's ghost layers, or (inPlace: Bn is B) on B's physical nodes, each node's new value also
// written to its periodic images in the ghost layers
void launchFoldFaraday(Geom const& G, MomSetArgs const& A, double* trho, double* tmd, VecPtr tvi,
                       CVecPtr B, CVecPtr E, VecPtr Bn, double dt, bool inPlace, hipStream_t s);
// Ampere with J's ghost layers and the electron closure in one launch (periodic level; B with
// its ghost layers): borderMax: the closure takes the totals from the border max of src (stored
// in dst: rho, massDensity, Vi), else from Ne, Vi
void launchAmpereElectrons(Geom const& G, double Te, CVecPtr B, VecPtr J, double const* const* src,
                           double* const* dst, CVecPtr Vi, double const* Ne, VecPtr Ve, double* Pe,
                           bool borderMax, hipStream_t s);

} // namespace phx
