/* phx.h -- C ABI of the MI355X (gfx950) hybrid-PIC ion update and field solve.
 *
 * One context = one AMR patch level held resident in HBM (round-one scope: a level made of one
 * patch; a periodic level is the patch covering the whole domain). Plain pointers and sizes only;
 * no torch or C++ types cross this boundary. All entry points return a phx_status; on failure
 * phx_last_error(ctx) describes the error (for a two-cell move: the particle, delta, velocity
 * and stage, as the reference's DictionaryException does).
 *
 * Each entry point names the reference interface it replaces (paths relative to the PHARE tree).
 * Data formats are the reference's:
 *  - fields: C-order FP64 arrays with the GridLayout's ghost nodes, sizes of
 *    GridLayout::allocSize (src/core/data/grid/gridlayout.hpp:849-864): per direction
 *    nbrCells + 1 + 2*ghosts for primal, one less for dual; ghosts = {2,4,4}[order-1];
 *  - particles: Particle<dim> (src/core/data/particles/particle.hpp:37-61), either as the
 *    reference's AoS records (56/64/80 B for dim 1/2/3) or as SoA arrays; iCell is in AMR index
 *    space, delta in [0,1).
 * Thread safety: a context is used by one host thread at a time (the reference is
 * single-threaded per rank); different contexts are independent.
 */
#ifndef PHX_H
#define PHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phx_ctx phx_ctx;

typedef enum {
    PHX_OK                 = 0,
    PHX_ERR_ARG            = 1, /* bad argument (id, size, dimension) */
    PHX_ERR_HIP            = 2, /* HIP runtime error (message in phx_last_error) */
    PHX_ERR_MOVE_TWO_CELLS = 3, /* Boris moved a particle more than 2 cells in a half step
                                   (boris.hpp:164-165, rethrown as "Updater::updatePopulations") */
    PHX_ERR_OUT_OF_GHOST   = 4, /* a particle's interpolation stencil left the ghost box */
    PHX_ERR_CAPACITY       = 5, /* particle capacity exceeded */
    PHX_ERR_STATE          = 6  /* call not valid in the current state */
} phx_status;

/* arithmetic of the particle path */
enum { PHX_PRECISION_FP64 = 0, /* plain: the reference method in FP64 */
       PHX_PRECISION_FP32 = 1  /* mixed: FP32 delta, velocity and weight (one charge per
                                  population), FP64 fields and moments; resident path of a
                                  periodic level only (phx_update_populations: PHX_ERR_STATE) */ };

/* UpdaterMode (src/core/numerics/ion_updater/ion_updater.hpp:22) */
enum { PHX_MODE_DOMAIN_ONLY = 1, PHX_MODE_ALL = 2 };

/* HyperMode (src/core/numerics/ohm/ohm.hpp:15) */
enum { PHX_HYPER_CONSTANT = 0, PHX_HYPER_SPATIAL = 1 };

/* particle arrays of a population: ParticlesPack (src/core/data/ions/particle_pack.hpp:20-114):
   domain, patch ghost, level ghost, and the level ghost particles split from the coarser level
   at the start (old) and the end (new) of its step (IonPopulation::levelGhostParticlesOld/New) */
enum { PHX_DOMAIN = 0, PHX_PATCH_GHOST = 1, PHX_LEVEL_GHOST = 2, PHX_LEVEL_GHOST_OLD = 3,
       PHX_LEVEL_GHOST_NEW = 4 };

/* electromagnetic field sets of SolverPPC (solver_ppc.hpp:52-53): state, predicted, averaged */
enum { PHX_EM = 0, PHX_EM_PRED = 1, PHX_EM_AVG = 2 };

/* field ids: E/B of each set, J, electron Ve/Pe, ion totals, per-population moments */
enum {
    PHX_F_EX = 0, PHX_F_EY, PHX_F_EZ, PHX_F_BX, PHX_F_BY, PHX_F_BZ,   /* state E, B */
    PHX_F_EPRED_X = 6,  /* .. 11: predicted E, B */
    PHX_F_EAVG_X  = 12, /* .. 17: averaged E, B */
    PHX_F_JX      = 18, PHX_F_JY, PHX_F_JZ,
    PHX_F_VEX     = 21, PHX_F_VEY, PHX_F_VEZ,
    PHX_F_PE      = 24,
    PHX_F_RHO     = 25, /* ion charge density (Ions::chargeDensity) */
    PHX_F_MASSD   = 26, /* ion mass density */
    PHX_F_VIX     = 27, PHX_F_VIY, PHX_F_VIZ,
    PHX_F_POP0    = 32  /* + 8*pop + {0: particleDensity, 1: chargeDensity, 2..4: flux x,y,z} */
};
#define PHX_F_POP(pop, k) (PHX_F_POP0 + 8 * (pop) + (k))

#define PHX_MAX_POP 4
/* level synchronisation fields of SolverPPC (solver_ppc.hpp:242-290): Bold (B at the start of the
   step, prepareStep) and fluxSumE (the time-integrated averaged E a finer level hands to the
   coarser one's reflux), 3 components each */
enum { PHX_F_BOLD_X = 64, PHX_F_FLUXSUM_X = 67 };

typedef struct {
    int dim;          /* 1, 2, 3 */
    int interp_order; /* 1, 2, 3 */
    int precision;    /* PHX_PRECISION_* */
    int n_pop;
    double mass[PHX_MAX_POP];  /* IonPopulation::mass */
    double eta;                /* OhmInfo::eta (resistivity) */
    double nu;                 /* OhmInfo::nu (hyper-resistivity) */
    int hyper_mode;            /* PHX_HYPER_* */
    double Te;                 /* isothermal electron temperature (electrons.hpp:202-212) */
} phx_config;

typedef struct {
    int lower[3];     /* AMR box of the patch, cells, inclusive */
    int upper[3];
    double dx[3];     /* GridLayout::meshSize */
    int level;        /* GridLayout::levelNumber */
    int periodic;     /* 1: the patch covers a periodic domain (same-level fills wrap onto itself) */
    /* UpdaterSelectionBoxing::nonLevelGhostBox (ion_updater.hpp:128; amr_utils.hpp:233-251);
       n_nonlevel_ghost_boxes == 0 means {domain box, domain box grown by the particle ghost width} */
    int n_nonlevel_ghost_boxes;
    int nonlevel_ghost_lower[27][3];
    int nonlevel_ghost_upper[27][3];
} phx_patch;

/* ---- context ------------------------------------------------------------------------------ */
/* replaces constructing IonUpdater{dict} (ion_updater.hpp:53-56) + SolverPPC{dict}
   (solver_ppc.hpp:69-75) + allocating the patch data the ResourcesManager would bind */
phx_status phx_create(const phx_config* cfg, const phx_patch* patch, size_t particle_capacity,
                      phx_ctx** out);
void phx_destroy(phx_ctx* ctx);
const char* phx_last_error(const phx_ctx* ctx);
phx_status phx_synchronize(phx_ctx* ctx);

/* A level of many patches of uneven sizes inside the level box given by `level` (its mesh,
   level number and periodicity; its lower/upper the level's AMR box): the batched form of the
   per-patch loops of SolverPPC (`for (auto& patch : rm.enumerate(level, ...))`,
   solver_ppc.hpp:553-556) and of the same-level messenger fills between the patches
   (hybrid_hybrid_messenger_strategy.hpp:376-548). patch_lower/patch_upper hold n_patches x 3
   AMR cells (inclusive) of non-overlapping patch domain boxes inside the level box. The fields,
   moments and particles of all patches live in one level-wide set of arrays in HBM; every stage
   runs as one launch over all patches, the particle stage from a device patch table (each patch
   cut into segments of at most 64 cells). A multi-patch level must be periodic (its borders are
   not fed by a coarser level here). phx_create is the level of one patch covering the box. */
phx_status phx_create_level(const phx_config* cfg, const phx_patch* level, int n_patches,
                            const int* patch_lower, const int* patch_upper, size_t particle_capacity,
                            phx_ctx** out);
/* AMR domain box of patch p */
phx_status phx_patch_box(const phx_ctx* ctx, int p, int lower[3], int upper[3]);
/* makeNonLevelGhostBoxFor (src/amr/resources_manager/amr_utils.hpp:233-251): the patch's
   domain box, then its particle ghost box intersected with each same-level neighbour's domain
   (periodic images of the level included); at most 27 boxes */
phx_status phx_patch_nonlevel_ghost_boxes(const phx_ctx* ctx, int p, int* n, int lower[][3],
                                          int upper[][3]);
/* a field of patch p in the patch's own GridLayout (allocSize of the patch, gridlayout.hpp:849-864):
   set writes the patch's physical nodes; get returns every node, ghost nodes holding what the
   same-level fills leave there (neighbour or periodic values, border sums of the moments) */
phx_status phx_patch_field_shape(const phx_ctx* ctx, int p, int field_id, int shape[3]);
phx_status phx_set_patch_field(phx_ctx* ctx, int p, int field_id, const double* host);
phx_status phx_get_patch_field(phx_ctx* ctx, int p, int field_id, double* host);
/* the domain particles of patch p (those whose cell lies in its box); with icell null only the
   count *n is returned */
phx_status phx_get_patch_particles(phx_ctx* ctx, int p, int pop, size_t* n, int32_t* icell,
                                   double* delta, double* v, double* weight, double* charge);

/* ---- data movement (host <-> HBM) ---------------------------------------------------------- */
/* shape of a field id as GridLayout::allocSize gives it (unused dimensions are 1) */
phx_status phx_field_shape(const phx_ctx* ctx, int field_id, int shape[3]);
phx_status phx_set_field(phx_ctx* ctx, int field_id, const double* host);
phx_status phx_get_field(phx_ctx* ctx, int field_id, double* host);

/* SoA particle arrays: icell[n*dim], delta[n*dim], v[n*3], weight[n], charge[n] */
phx_status phx_set_particles(phx_ctx* ctx, int pop, int kind, size_t n, const int32_t* icell,
                             const double* delta, const double* v, const double* weight,
                             const double* charge);
/* appends to the array (unsorted storage only: before phx_initialize / phx_sort_particles) */
phx_status phx_add_particles(phx_ctx* ctx, int pop, int kind, size_t n, const int32_t* icell,
                             const double* delta, const double* v, const double* weight,
                             const double* charge);
phx_status phx_particle_count(phx_ctx* ctx, int pop, int kind, size_t* n);
phx_status phx_get_particles(phx_ctx* ctx, int pop, int kind, int32_t* icell, double* delta,
                             double* v, double* weight, double* charge);
/* reference AoS records Particle<dim> (particle.hpp:37-61), e.g. ParticleArray::vector().data() */
phx_status phx_set_particles_aos(phx_ctx* ctx, int pop, int kind, size_t n, const void* records);
phx_status phx_get_particles_aos(phx_ctx* ctx, int pop, int kind, void* records);

/* MaxwellianParticleInitializer::loadParticles (maxwellian_particle_initializer.hpp:137-198):
   per-cell profiles in the row-major AMR cell order of layout.indices(AMRBox); the particles
   are appended to the population's domain array */
phx_status phx_load_maxwellian(phx_ctx* ctx, int pop, const double* density, const double* vx,
                               const double* vy, const double* vz, const double* vthx,
                               const double* vthy, const double* vthz, double charge,
                               uint32_t ppc, uint64_t seed, double density_cut_off);

/* ---- hot path --------------------------------------------------------------------------- */
/* IonUpdater::updatePopulations(ions, em, boxing, dt, mode) (ion_updater.hpp:58-60, 88-107):
   resets the population moments, pushes (Boris: pusher/boris.hpp:93-138) with the selection
   semantics of updateAndDepositDomain_ / updateAndDepositAll_ (:169-295) and deposits
   (Interpolator::operator() interpolator.hpp:468-510). em_set selects the E/B set. */
phx_status phx_update_populations(phx_ctx* ctx, int em_set, double dt, int mode);
/* BorisPusher::move(rangeIn, rangeOut, em, mass, interpolator, layout, firstSelector,
   secondSelector) (src/core/numerics/pusher/boris.hpp:93-138; Pusher interface pusher.hpp:28-35,
   factory name "modified_boris" pusher_factory.hpp:18-29), in place (rangeIn == rangeOut, as
   IonUpdater::updateAndDepositAll_ calls it) over array `kind` of population `pop`, with the
   population's mass, dt and the context's mesh (setMeshAndTimeStep, boris.hpp:143-148) and E/B of
   em_set. Each selector is a list of AMR boxes (n x 3 lower, n x 3 upper, inclusive cells): a
   particle is selected when its cell lies in one of them; n = 0 is the noop selector. Every
   particle is prePushed; those the first selector keeps are gathered, accelerated and
   postPushed. The array is left partitioned: [kept by both selectors | by the first only | by
   neither]; counts[0] = end of the second selection (the returned ParticleRange), counts[1] = end
   of the first. A move of more than 2 cells returns PHX_ERR_MOVE_TWO_CELLS (the reference's
   DictionaryException). */
phx_status phx_push(phx_ctx* ctx, int pop, int kind, int em_set, double dt, int n_first,
                    const int* first_lower, const int* first_upper, int n_second,
                    const int* second_lower, const int* second_upper, size_t counts[2]);
/* Interpolator::operator()(particle, em, layout) (src/core/numerics/interpolator/
   interpolator.hpp:420-456) over a whole array: E[3 n], B[3 n] at each particle, in the order
   phx_get_particles returns the array */
phx_status phx_gather(phx_ctx* ctx, int em_set, int pop, int kind, double* E, double* B);
/* No counterpart in the reference: phx_gather without any device allocation, for callers that
   cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc). Same arguments and outputs as phx_gather; the values are staged in a fixed device
   array of the library (one per GPU, ~48 MiB, held until the process ends) instead of a
   per-call allocation. At most PHX_GATHER_NOALLOC_MAX_SLOTS slots (live plus empty) in the array,
   else PHX_ERR_CAPACITY. Every call on a GPU shares that GPU's array: calls that overlap on one
   GPU (several host threads, several contexts) race and must be serialised by the caller. */
#define PHX_GATHER_NOALLOC_MAX_SLOTS (1u << 20)
phx_status phx_gather_noalloc(phx_ctx* ctx, int em_set, int pop, int kind, double* E, double* B);
/* Interpolator::operator()(range, n, rho, F, layout, coef) (interpolator.hpp:468-510): the
   particles of array `kind` added into the population's density, charge density and flux with
   weight coef (the moments are not reset; phx_update_populations resets them) */
phx_status phx_deposit(phx_ctx* ctx, int pop, int kind, double coef);

/* ---- coarse-to-fine coupling of two levels (both contexts on the same GPU) ------------------ */
/* ParticlesRefineOperator::refine_ (src/amr/data/particles/refine/particles_data_split.hpp:157-260)
   with the Splitter of (dim, interp order, refined_particle_nbr)
   (src/amr/data/particles/refine/split_{1,2,3}d.hpp; 0 = PHARE's default: 2 in 1D, 4 in 2D, 6 in
   3D; every 1D variant is available): the domain particles of population `pop` of the coarse level
   are put on the fine grid (refinement ratio 2), those whose fine cell lies within
   ceil((order + 1) / 2) cells of a destination box are split, and the refined particles inside a
   destination box are appended to array dest_kind of the fine level (PHX_DOMAIN: interior
   refinement; PHX_LEVEL_GHOST / _OLD / _NEW: the coarse-to-fine boundary). box_lower/box_upper:
   n_boxes x 3 fine AMR cells, inclusive, disjoint (at most 32). FP64 levels only. */
phx_status phx_refine_particles(phx_ctx* fine, const phx_ctx* coarse, int pop, int dest_kind,
                                int n_boxes, const int* box_lower, const int* box_upper,
                                int refined_particle_nbr);
/* HybridHybridMessengerStrategy::lastStep (hybrid_hybrid_messenger_strategy.hpp:610-633): per
   population, levelGhostNew and levelGhostOld swap, levelGhostNew is emptied and levelGhost
   becomes a copy of levelGhostOld. The moments of fillIonPopMomentGhosts (:507-548) are
   phx_deposit of PHX_LEVEL_GHOST_OLD with coef 1 - alpha and of PHX_LEVEL_GHOST_NEW with alpha. */
phx_status phx_level_ghost_last_step(phx_ctx* ctx);
/* linear_time_interpolate (src/amr/data/field/time_interpolate/field_linear_time_interpolate.hpp:
   20-32): dst = (1 - alpha) old + alpha new on the nodes lower..upper (inclusive, local node
   indices of the field arrays, which must have one shape); the contexts may be the same one */
phx_status phx_time_interpolate(phx_ctx* dst, int dst_field, const phx_ctx* old_ctx, int old_field,
                                const phx_ctx* new_ctx, int new_field, const int lower[3],
                                const int upper[3], double alpha);
/* ---- level synchronisation (fine to coarse) ------------------------------------------------- */
/* SolverPPC::prepareStep (solver_ppc.hpp:240-257): Bold = B */
phx_status phx_prepare_step(phx_ctx* ctx);
/* SolverPPC::resetFluxSum / accumulateFluxSum (solver_ppc.hpp:260-290): fluxSumE = 0;
   fluxSumE += Eavg * coef (the multiphysics integrator's coef is 1 / ratio^2,
   multiphysics_integrator.hpp:541-548) */
phx_status phx_reset_flux_sum(phx_ctx* ctx);
phx_status phx_accumulate_flux_sum(phx_ctx* ctx, double coef);
/* SolverPPC::reflux (solver_ppc.hpp:293-308): B = Faraday(Bold, Eavg, dt) on the physical nodes,
   then the magnetic ghost fill (periodic levels; others are the caller's messenger). The coarser
   level's Eavg (its "refluxElectric") holds the finer level's coarsened fluxSumE where the finer
   level covers it (phx_coarsen with PHX_COARSEN_ELECTRIC from PHX_F_FLUXSUM_X). */
phx_status phx_reflux(phx_ctx* ctx, double dt);
/* the coarsen operators of HybridHybridMessengerStrategy's synchronisation and reflux
   (hybrid_hybrid_messenger_strategy.hpp:106-110, 890-905, 194-198), ratio 2: the coarse nodes
   lower..upper (inclusive, coarse AMR node indices) of coarse_field take the values of fine_field
   (the same quantity) of the finer level: PHX_COARSEN_MOMENTS injects the fine node 2c
   (MomentsCoarsener, src/amr/data/field/coarsening/moments_coarsener.hpp:39-75; all-primal
   quantities: ion charge density, bulk velocity, population moments), PHX_COARSEN_ELECTRIC
   averages the two fine nodes 2c, 2c + 1 along the component's dual direction
   (ElectricFieldCoarsener, electric_field_coarsener.hpp:38-140; E and fluxSumE) */
enum { PHX_COARSEN_MOMENTS = 0, PHX_COARSEN_ELECTRIC = 1 };
phx_status phx_coarsen(phx_ctx* coarse, int coarse_field, const phx_ctx* fine, int fine_field, int op,
                       const int lower[3], const int upper[3]);
/* the refine operators of HybridHybridMessengerStrategy (hybrid_hybrid_messenger_strategy.hpp:
   87-93, 1130-1136; FieldRefineOperator::refine, src/amr/data/field/refine/
   field_refine_operator.hpp:75-108), ratio 2: the fine nodes lower..upper (inclusive, fine AMR
   indices of the quantity's own nodes, the reference's intersection box) of fine_field take values
   from coarse_field (the same quantity) of the coarser level.
   PHX_REFINE_LINEAR: DefaultFieldRefiner (field_refiner.hpp:31-166), linear interpolation
   (LinearWeighter, linear_weighter.cpp), only where the fine node holds NaN (the moment ghosts);
   PHX_REFINE_LINEAR_OVERWRITE: FieldMomentsRefiner (field_moments_refiner.hpp), the same on every
   node; PHX_REFINE_ELECTRIC: ElectricFieldRefiner (electric_field_refiner.hpp), E and J, NaN nodes
   only; PHX_REFINE_MAGNETIC: MagneticFieldRefiner (magnetic_field_refiner.hpp), the fine faces on
   coarse faces take the coarse face value, NaN nodes only; PHX_REFINE_MAGNETIC_INIT:
   MagneticFieldInitRefiner (magnetic_field_init_refiner.hpp), the same on every node. The new
   faces between coarse faces are phx_refine_magnetic_new_faces' work. */
enum { PHX_REFINE_LINEAR = 0, PHX_REFINE_LINEAR_OVERWRITE = 1, PHX_REFINE_ELECTRIC = 2,
       PHX_REFINE_MAGNETIC = 3, PHX_REFINE_MAGNETIC_INIT = 4 };
phx_status phx_refine_field(phx_ctx* fine, int fine_field, const phx_ctx* coarse, int coarse_field,
                            int op, const int lower[3], const int upper[3]);
/* setNaNsOnFieldGhosts (hybrid_hybrid_messenger_strategy.hpp:959-994) for a one-patch context:
   every node of the field outside its physical nodes becomes NaN, so that the NaN-only refine
   operators of a level-ghost fill (phx_refine_field) complete exactly those nodes */
phx_status phx_set_ghosts_nan(phx_ctx* ctx, int field_id);
/* MagneticRefinePatchStrategy::postprocessRefine (src/amr/data/field/refine/
   magnetic_refine_patch_strategy.hpp:67-125): in the fine AMR cell box cell_lower..cell_upper the
   new faces of the three components of the B field whose x component is b_field_x (PHX_F_B of a
   set, or PHX_F_BOLD_X) are set from the shared faces (Toth & Roe 2002, divergence-free). The
   box's cells must lie in the arrays. At their edge (no counterpart in the reference, which
   refines into a scratch region one stencil wider), a face whose cross terms would read beyond the
   arrays takes the mean of its two neighbours along the normal, and a face without both neighbours
   is left unchanged. */
phx_status phx_refine_magnetic_new_faces(phx_ctx* fine, int b_field_x, const int cell_lower[3],
                                         const int cell_upper[3]);
/* IonUpdater::updateIons (ion_updater.hpp:111-116) */
phx_status phx_update_ions(phx_ctx* ctx);
/* same-level fills of HybridHybridMessengerStrategy (hybrid_hybrid_messenger_strategy.hpp):
   fillFluxBorders + fillDensityBorders (:419-483), fillIonBorders (:486-497),
   fillIonGhostParticles (:404-415), field ghost fills (:376-401) */
phx_status phx_fill_moment_borders(phx_ctx* ctx);
phx_status phx_fill_ion_borders(phx_ctx* ctx);
phx_status phx_fill_ion_ghost_particles(phx_ctx* ctx);
phx_status phx_fill_ghosts(phx_ctx* ctx, int field_id_x); /* the 3 components from _x */

/* Faraday<GL>{layout}(B, E, Bnew, dt) (faraday.hpp:27-48) with E/B sets */
phx_status phx_faraday(phx_ctx* ctx, int b_set, int e_set, int bnew_set, double dt);
/* Ampere<GL>{layout}(B, J) (ampere.hpp:26-47) */
phx_status phx_ampere(phx_ctx* ctx, int b_set);
/* Electrons::update (electrons.hpp:304-314): Ve and Pe from J, Vi and the charge density */
phx_status phx_electrons_update(phx_ctx* ctx);
/* Ohm<GL>{info, layout}(n, Ve, Pe, B, J, Enew) (ohm.hpp:48-64) */
phx_status phx_ohm(phx_ctx* ctx, int b_set, int e_out_set);
/* SolverPPC::average_ (solver_ppc.hpp:483-510) including the Eavg ghost fill */
phx_status phx_average(phx_ctx* ctx);

/* SolverPPC::moveIons_ (solver_ppc.hpp:537-598): updatePopulations + fills + updateIons */
phx_status phx_move_ions(phx_ctx* ctx, double dt, int mode);
/* HybridLevelInitializer::initialize root branch (hybrid_level_initializer.hpp:103-171):
   moments from the domain particles, then J by Ampere and E by Ohm */
phx_status phx_initialize(phx_ctx* ctx);
/* SolverPPC::advanceLevel (solver_ppc.hpp:315-341), nsteps times; the whole step is a
   captured hipGraph replayed per step (periodic level only) */
phx_status phx_advance(phx_ctx* ctx, double dt, int nsteps);

/* Storage order (no counterpart in the reference, whose ParticleArray keeps insertion order and
   a CellMap): reorders every population's domain particles by cell, row-major over the AMR box,
   the order the push kernel's LDS deposit tile relies on. phx_initialize sorts once and
   phx_advance every `steps` PPC steps (0: never; default 30). Results do not depend on the order
   beyond the summation order of the deposit. */
phx_status phx_sort_particles(phx_ctx* ctx);
phx_status phx_set_sort_interval(phx_ctx* ctx, int steps);

/* FP64 diagnostics: [kinetic energy, magnetic energy, total charge, particle count] */
phx_status phx_diagnostics(phx_ctx* ctx, double out[4]);

/* the HIP stream the context launches on (a hipStream_t), for timing with HIP events */
void* phx_stream(phx_ctx* ctx);

/* ---- tuning (no counterpart in the reference: launch shapes of this build) ------------------- */
/* The shapes the segmented push and the sort use; the defaults are the winners of the project's
   autotuner (phare_amd/autotune.py) on the 2D Harris workload. phx_set_tuning rebuilds the
   segment tables of the level's patches, re-sorts the domain arrays that are cell-segmented and
   drops the captured steps (the next phx_advance captures again). */
typedef struct {
    int segment_shape[3]; /* preferred segment shape in cells per dimension (the product is at
                             most 64; clipped to each patch, the cells left over go to the slower
                             dimensions); 0 keeps the default of the dimension */
    int rows_per_item;    /* most segment rows per workgroup of the segmented push (1..4096) */
    int tile_row_pad;     /* LDS tile pitches padded to a multiple of this many nodes (1: none):
                             2D the rows; 3D (segment push) the pitches of dimensions 1 and 2 */
    int sort_interval;    /* PPC steps of phx_advance between two sorts (0: never) */
    int tile_copies;      /* moment copies of the LDS tile of the linear-storage push (1, 2, 4, 8;
                             lane l adds into copy l % copies) */
    int deposit_window;   /* mixed-precision push: 1 sums each lane's successive particles of one
                             cell in registers before the LDS tile (1D, 2D and 3D order 1), 0 adds
                             every particle into the tile */
} phx_tuning;
phx_status phx_get_tuning(const phx_ctx* ctx, phx_tuning* out);
phx_status phx_set_tuning(phx_ctx* ctx, const phx_tuning* tuning);
/* ---- measurement (no counterpart in the reference: a benchmark hook) ------------------------ */
/* captures only the fused push+gather+deposit kernel of a periodic level (population 0, E/B of
   the averaged set) in a hipGraph, replays it `reps` times back to back between two HIP events
   on the context stream, and returns the mean kernel time in milliseconds. Mode PHX_MODE_ALL
   advances the particles in place. */
phx_status phx_profile_push(phx_ctx* ctx, double dt, int mode, int reps, double* avg_ms);
/* in-kernel timing of the fused pushes inside the steps phx_advance / phx_move_ions run on a
   periodic level: each launch's span from its first workgroup's start to its last workgroup's
   end (device wall clock), summed per population and mode. on = 1 enables and resets the sums
   (a change of the setting re-captures the step once); phx_push_timing returns the mean launch
   time in milliseconds and the number of launches summed. */
phx_status phx_set_push_timing(phx_ctx* ctx, int on);
phx_status phx_push_timing(phx_ctx* ctx, int pop, int mode, double* avg_ms, int* launches);

#ifdef __cplusplus
}
#endif

#endif /* PHX_H */
