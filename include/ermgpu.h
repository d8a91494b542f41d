/*
 * ermgpu.h -- C-ABI of the MI355X (gfx950) solute-solvent energy pipeline.
 *
 * Drop-in for ERmod's per-frame energy path: the module procedures of
 * engproc.F90 / realcal.F90 / recpcal.F90 that engconst drives
 * (engproc.F90:301-448).  A Fortran host binds every entry point through
 * ISO_C_BINDING (see ermod_amd/fortran/ermgpu.F90); Python binds it through
 * ctypes (ermod_amd/api.py).  No torch or HIP types appear here: plain
 * pointers, sizes and int32 status codes.
 *
 * Conventions
 *  - all indices are 0-based; mol_begin has nmol+1 entries;
 *  - coordinates are xyz-interleaved doubles (site i at pos[3*i .. 3*i+2]);
 *  - 3x3 matrices are column-major (m[3*j+i] = M(i+1,j+1)), the Fortran
 *    layout of engmain's `cell` (columns are the box vectors);
 *  - uvengy arrays hold slvmax+1 doubles: [0] is the solute self energy,
 *    [k+1] the energy with the k-th physical molecule (tagpt order,
 *    engproc.F90:336-347);
 *  - host arrays are caller-owned and copied; device memory is owned by the
 *    context; calls on one context are serialised on its HIP stream and must
 *    come from one thread at a time;
 *  - every call returns ERM_OK or an error code; nothing aborts the process.
 *    erm_last_error() returns the message of the last failing call.
 */
#ifndef ERMGPU_H
#define ERMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1-9 are the hot path's halt_with_error / warning codes
 * (mpiproc.F90:203-264) returned instead of aborting. */
enum erm_status {
  ERM_OK = 0,
  ERM_E_ENG_MIN = 1,     /* energy at or below the lowest edge   engproc.F90:992-995   */
  ERM_E_ENG_SFT = 2,     /* soln: above the soft region           engproc.F90:1000-1005 */
  ERM_E_MBIN = 3,        /* overflow bin below 1e3 (warning)      engproc.F90:1010-1015 */
  ERM_E_ENG_ECD = 4,     /* inconsistent &hist                    engproc.F90:185       */
  ERM_E_RCP_CNS = 5,     /* molecule without spline slot          recpcal.F90:333       */
  ERM_E_NONCUBOID = 6,   /* non-cuboid cell                       realcal.F90:859-864   */
  ERM_E_ENG_BUG = 7,     /* internal inconsistency                eng_bug               */
  ERM_E_ENG_CNS = 8,     /* solvent species <= 0                  engproc.F90:812       */
  ERM_E_ENG_SIZ = 9,     /* too many bins                         engproc.F90:194       */
  ERM_E_ARG = 10,        /* invalid argument or call order */
  ERM_E_UNSUPPORTED = 11,/* option outside this implementation */
  ERM_E_DEVICE = 12,     /* HIP runtime error */
  ERM_E_EOF = 13,        /* end of the trajectory (not an error) */
  ERM_E_INS_LIMIT = 14   /* no reference-structure insertion accepted in 4096 attempts */
};

enum erm_precision { ERM_FP64 = 0, ERM_FP32 = 1 };

/* &ene_param values the hot path reads (engmain.F90:331-343), after
 * setparam's defaulting (setconf.F90:144-453). */
typedef struct erm_params {
  int32_t slttype;        /* 1 soln, 2 refs rigid, 3 refs flexible */
  int32_t estype;         /* 1 NVT, 2 NPT (volcorrect weights) */
  int32_t ljswitch;       /* 0 CHARMM pot, 1 GMX pot, 2 CHARMM force, 3 GMX force */
  int32_t splodr;         /* PME B-spline order, 2..12 */
  int32_t ms[3];          /* PME grid ms1max..ms3max; factors 2,3,5 only, <= 256 */
  int32_t corrcal;        /* accumulate the correlation matrix */
  int32_t selfcal;        /* accumulate the self-energy distribution */
  int32_t wgtslf;         /* weight by the self energy */
  int32_t force_calculation; /* continue past the mbin warning */
  double elecut, lwljcut, upljcut, screen, temp, block_threshold;
  int32_t precision;      /* ERM_FP64 or ERM_FP32 pair/PME arithmetic */
  int32_t max_batch;      /* insertions per device batch; 0 picks a default */
  int32_t device;         /* HIP device ordinal */
  int32_t pme_path;       /* reciprocal-space kernels: 0 auto, 1 fused slab with
                             LDS FFTs, 2 separate plane and gather kernels,
                             3 register-column FFTs + fused x/gather slab */
  int32_t cltype;         /* electrostatics, the reference's numbering
                             (engmain.F90:314): ERM_EL_COULOMB bare 1/r
                             (realcal_bare), ERM_EL_PME, ERM_EL_PPPM; Ewald (1)
                             is refused as by setconf.F90:329-331 */
  int32_t boxshp;         /* ERM_SYS_PERIODIC, or ERM_SYS_NONPERIODIC (bare
                             Coulomb, estype 1 and slttype 1 only,
                             setconf.F90:368-372, insertion.F90:155); the
                             cell is then not read.  Zero in both fields
                             means bare Coulomb without a box */
} erm_params;

enum { ERM_EL_COULOMB = 0, ERM_EL_PME = 2, ERM_EL_PPPM = 3 };
/* insertion modes, the reference's numbering (engmain.F90:320-327) */
enum { ERM_INSORG_ORIGIN = 0, ERM_INSORG_NOCHANGE = 1, ERM_INSORG_AGGCEN = 2, ERM_INSORG_REFSTR = 3 };
enum {
  ERM_INSPOS_RANDOM = 0, ERM_INSPOS_NOCHANGE = 1, ERM_INSPOS_SPHERE = 2, ERM_INSPOS_SLAB_GENERIC = 3,
  ERM_INSPOS_SLAB_SYMMETRIC = 4, ERM_INSPOS_RMSD = 5, ERM_INSPOS_GAUSS = 6
};
enum { ERM_INSROT_RANDOM = 0, ERM_INSROT_NOCHANGE = 1 };
enum { ERM_SYS_NONPERIODIC = 0, ERM_SYS_PERIODIC = 1 };

/* &hist namelist of one species (engproc.F90:66-67) */
typedef struct erm_hist {
  double ecdmin, ecfmns, ecdcen, eccore, ecdmax;
  double eclbin, ecfbin, ec0bin, finfac;
  int32_t pecore;
} erm_hist;

typedef struct erm_ctx erm_ctx;

/* ---- energy coordinate ------------------------------------------------ */
/* Bin edges of one species from its &hist; replaces the edge construction of
 * enginit (engproc.F90:124-237).  self_energy != 0 returns the pesoft edges
 * of the solute self-energy coordinate. Host-only. */
int32_t erm_build_edges(const erm_hist *h, int32_t self_energy, double *ercrd, int32_t cap,
                        int32_t *pemax, int32_t *pesoft);

/* hash of FP64 bit patterns; replaces hash_double_ (hash_real.c:5-15) used by
 * get_solute_hash (engproc.F90:1176-1182). Host-only. */
uint64_t erm_hash_double(const double *v, int32_t n);

/* ---- context ------------------------------------------------------------ */
/* Creates a context on params->device; replaces recpcal_init's setup
 * (recpcal.F90:42-85) and the parameter part of enginit. */
int32_t erm_create(const erm_params *params, erm_ctx **out);
void erm_destroy(erm_ctx *ctx);
const char *erm_last_error(const erm_ctx *ctx);

/* Topology: the engmain arrays setparam fills (setconf.F90:706-834):
 * scaled charges, 0-based LJ types, ntype*ntype LJ tables (epsilon and
 * sigma^2, symmetric), sluvid, solvent species (uvspec, engproc.F90:96-103),
 * molecular charges and site masses.  bfcoord (3*nsite of the test particle)
 * is required for erm_insert_random, else may be NULL. */
int32_t erm_set_topology(erm_ctx *ctx, int32_t natom, int32_t nmol, const int32_t *mol_begin,
                         const double *charge, const int32_t *ljtype, int32_t ntype,
                         const double *lj_eps, const double *lj_sig2, const int32_t *sluvid,
                         const int32_t *uvspec, const double *mol_charge, const double *sitemass,
                         const double *bfcoord);

/* Energy coordinate: enginit's output (engproc.F90:227-247); uvcrd holds
 * sum(uvmax) edges, escrd esmax edges (NULL when selfcal = 0). */
int32_t erm_set_bins(erm_ctx *ctx, int32_t numslv, const int32_t *uvmax, const int32_t *uvsoft,
                     const double *uvcrd, int32_t esmax, const double *escrd);

/* Zeroes the accumulators of a division; replaces engclear
 * (engproc.F90:276-288). */
int32_t erm_engclear(erm_ctx *ctx);

/* Placement of device-generated test-particle poses; replaces the
 * insorigin / insposition / insorient / lwreg / upreg / hostspec handling of
 * instslt (set_solute_origin, set_shift_com, apply_orientation,
 * com_aggregate; insertion.F90:103-366, 539-566) with the pairings and
 * bounds setconf.F90:399-432 enforces.  Supported: ORIGIN + RANDOM (the
 * default), AGGCEN + SPHERE / SLAB_GENERIC / SLAB_SYMMETRIC, each with
 * RANDOM or NOCHANGE orientation; hostspec lists nhost solvent species
 * (uvspec numbering) whose molecules define the aggregate centre, computed
 * per frame; REFSTR + RMSD / GAUSS after erm_set_refstructure.  Applies
 * to erm_insert_random and erm_solute_gradient's random poses. */
int32_t erm_set_insertion(erm_ctx *ctx, int32_t insorigin, int32_t insposition, int32_t insorient,
                          double lwreg, double upreg, int32_t nhost, const int32_t *hostspec);
/* Random insertions of a flexible test particle (slttype 3): confs holds n
 * body configurations (n * nsite * 3, as getsolute reads them from SltConf,
 * insertion.F90:384-480), weights their SltWght values (NULL: 1); body b is
 * placed as insertion first_index + b of the stream seed under the
 * insertion modes (not REFSTR), then as erm_insert_random. */
int32_t erm_insert_flexible(erm_ctx *ctx, int32_t n, const double *confs, const double *weights, uint64_t seed,
                            uint64_t first_index, double *uvengy);
/* The reference structure of the REFSTR modes as load_refstructure
 * (insertion.F90:623-723) leaves it: nhost atoms (0-based indices) of the
 * refspec species with their reference coordinates (3 x nhost) and fitting
 * weights, the test particle's reference coordinates (3 x nsite) and
 * weights.  With insorigin REFSTR and insposition RMSD or GAUSS (rigid test
 * particle), every frame runs reffit (insertion.F90:598-622) on the host
 * (Horn's quaternion fit, bestfit.F90:92-241) and every insertion draws
 * attempts until the RMSD to the best-fit reference solute lies in
 * [lwreg, upreg] (check_solute_configuration, insertion.F90:569-595);
 * GAUSS divides the insertion's weight by the mixture density at each
 * attempt, as set_shift_com does.  4096 attempts without acceptance return
 * ERM_E_INS_LIMIT. */
int32_t erm_set_refstructure(erm_ctx *ctx, int32_t nhost, const int32_t *host_atoms, const double *host_crd,
                             const double *host_weight, const double *slt_crd, const double *slt_weight);

/* The intermediate-state method of extensions/Intermediate/engproc.F90
 * (solution systems): zerodns[ermax] != 0 marks the bins whose density is
 * at most 1e-30 in every reference-solvent engref file (lines 286-315);
 * NULL switches it off.  Each configuration whose instantaneous histogram
 * touches none of them adds its weight to soln_intm and norm_intm, its
 * histogram to edintm and its per-species energy sums to uvintm; every
 * configuration adds to soln_total (lines 996-1009).  erm_engclear clears
 * edintm, uvintm and norm_intm, erm_engstore divides edintm by norm_intm. */
int32_t erm_set_intermediate(erm_ctx *ctx, const int32_t *zerodns);
/* edintm[ermax], uvintm[numslv+1] ([0] unused), scal = {norm_intm,
 * soln_intm, soln_total}; any pointer may be NULL. */
int32_t erm_fetch_intermediate(erm_ctx *ctx, double *edintm, double *uvintm, double *scal);

/* ---- frame ingestion ----------------------------------------------------
 * DCD trajectories as trajectory-formats/dcd.f90:39-105 reads them (Fortran
 * unformatted records, 'CORD' header, optional unit-cell record, FP32 x, y, z
 * records), the PermIndex permutation and the SysWght frame weights
 * (getconf_parallel / read_weight, setconf.F90:860-996). */
typedef struct erm_dcd erm_dcd;
/* Host-only.  natom_file: the atom-count record; has_cell: icntrl(11). */
int32_t erm_dcd_open(const char *path, erm_dcd **out, int32_t *natom_file, int32_t *has_cell);
/* Host-only.  Next frame: xyz = x[natom], y[natom], z[natom] (the first natom
 * values of each record); with periodic != 0 the cell record fills the
 * column-major cell's (1,1), (1,2), (2,2), (1,3), (2,3), (3,3) as dcd.f90
 * does (the rest 0).  ERM_E_EOF at the end of the file. */
int32_t erm_dcd_read(erm_dcd *h, int32_t natom, int32_t periodic, float *xyz, double *cell);
void erm_dcd_close(erm_dcd *h);
/* Host-only.  PermIndex ("i p" lines, 1-based; setconf.F90:866-905) into the
 * 0-based perm[natom_traj]; identity and *found = 0 when the file does not
 * exist; ERM_E_ARG for entries out of range or targets not used once. */
int32_t erm_read_permutation(const char *path, int32_t natom_traj, int32_t *perm, int32_t *found);

typedef struct erm_traj erm_traj;
/* Frame source for a context.  The file holds the natom_traj atoms of the
 * solvent and solute molecules (they come first; test-particle slots
 * follow); perm (0-based, NULL for none) sends trajectory atom i to atom
 * perm[i]; weight_path names a SysWght file (NULL: weight 1); every frame
 * used is the last of skpcnf frames read.  Close it before the context. */
int32_t erm_traj_open(erm_ctx *ctx, const char *dcd_path, int32_t natom_traj, const int32_t *perm,
                      const char *weight_path, int32_t skpcnf, erm_traj **out);
/* getconf_parallel + erm_set_frame for the next frame: the FP32 coordinates
 * go from pinned memory to the device asynchronously and are scattered
 * there through the permutation into FP64; the following frame is read by
 * a host thread meanwhile.  cell (9) and weight are optional outputs.
 * ERM_E_EOF after the last frame. */
int32_t erm_traj_next(erm_traj *t, double *cell, double *weight);
void erm_traj_close(erm_traj *t);

/* One frame: cell inverse and volume (get_inverted_cell, engproc.F90:898-919),
 * Green's function when the cell changed (recpcal_spline_greenfunc,
 * recpcal.F90:111-152, or recpcal_pppm_greenfunc, recpcal.F90:205-325),
 * solvent B-spline weights (recpcal_prepare_solvent, recpcal.F90:328-340),
 * coordinate normalisation (realcal_prepare, realcal.F90:137-146) and the
 * cell list (set_block_info / blockify / sort_block, realcal.F90:373-525),
 * built once per frame on the device.  sitepos: natom*3 host doubles
 * (test-particle slots ignored); cell may be NULL without a box. */
int32_t erm_set_frame(erm_ctx *ctx, const double *sitepos, const double *cell,
                      double stat_weight_system);
/* Same, with sitepos already in device memory (natom*3 doubles).  The copy is
 * ordered on the context's stream only: the caller makes sure the
 * coordinates are written (synchronises the stream that produced them)
 * before the call. */
int32_t erm_set_frame_device(erm_ctx *ctx, const double *d_sitepos, const double *cell,
                             double stat_weight_system);

/* get_uv_energy (engproc.F90:662-756) for a batch of n solute
 * configurations without touching the accumulators.  soln mode: solute is
 * the molecule index and solute_pos may be NULL (coordinates of the frame).
 * refs mode: solute is the test particle and solute_pos holds n*nsite*3
 * doubles.  uvengy: n*(slvmax+1) doubles. */
int32_t erm_get_uv_energy(erm_ctx *ctx, int32_t solute, int32_t n, const double *solute_pos,
                          double *uvengy);

/* The pieces of get_uv_energy for one configuration, as the reference's
 * module procedures compute them:
 *   realcal_proc  (realcal.F90:58-135)  -> ereal[slvmax]
 *   realcal_self  (realcal.F90:284-328) -> *self_real
 *   recpcal_prepare_solute + recpcal_energy (recpcal.F90:342-502)
 *                                       -> erecp[slvmax], *self_recp  */
int32_t erm_realcal_proc(erm_ctx *ctx, int32_t solute, const double *solute_pos, double *ereal);
int32_t erm_realcal_self(erm_ctx *ctx, int32_t solute, const double *solute_pos, double *self_real);
int32_t erm_recpcal(erm_ctx *ctx, int32_t solute, const double *solute_pos, double *erecp,
                    double *self_recp);

/* The hot loop of engconst (engproc.F90:388-392): get_uv_energy followed by
 * update_histogram (engproc.F90:758-856) for every configuration.
 * soln mode: n solute molecule indices in `solutes` (sltlist entries);
 * refs mode: n test-particle configurations in solute_pos (n*nsite*3), with
 * optional per-configuration weights (stat_weight_solute, NULL = 1).
 * uvengy (optional) receives n*(slvmax+1) energies. */
int32_t erm_engconst(erm_ctx *ctx, int32_t n, const int32_t *solutes, const double *solute_pos,
                     const double *weights, double *uvengy);

/* Test-particle insertion with device-generated random poses
 * (instslt INSPOS_RANDOM + INSROT_RANDOM, insertion.F90:44-366): insertion
 * `first_index + i` draws its position and quaternion from a counter-based
 * generator keyed by `seed`, then get_uv_energy + update_histogram run on
 * the device.  uvengy optional as above. */
int32_t erm_insert_random(erm_ctx *ctx, int32_t n, uint64_t seed, uint64_t first_index,
                          double *uvengy);

/* erm_insert_random on nctx contexts at once (no counterpart in the
 * reference, which runs one solute species per run): each context's batches
 * are enqueued on its own stream, so solutes of different sizes run
 * concurrently, and one wait per context follows.  n, seed, first_index
 * hold one value per context.  Returns the first error; each context's
 * message is in its own erm_last_error. */
int32_t erm_insert_random_multi(erm_ctx *const *ctxs, int32_t nctx, const int32_t *n, const uint64_t *seed,
                                const uint64_t *first_index);

/* ---- several test-particle species on one solvent ----------------------
 * No counterpart in the reference, which runs one solute species per run
 * (engproc.F90:89-91).  A group binds nctx (<= 16) contexts that describe the
 * same solvent (physical molecules, charges, LJ types, species), grid and
 * pair parameters, each with its own rigid test particle (slttype 2,
 * bfcoord, random positions and orientations), bins and accumulators.
 * erm_group_insert_random(g, n, seed, first_index) leaves in every member
 * the accumulators erm_insert_random(ctx[i], n[i], seed[i], first_index[i])
 * leaves: the same poses and energies, the counts bit for bit.  It runs as
 * one sequence on member 0's stream: batches mix the members'
 * configurations in runs (largest solutes first); the y and x transforms
 * and the gather of a batch are one launch each over every run; the z
 * columns and the real-space pairs are one persistent launch each whose
 * workgroups pull (configuration, block) items of every run, or one launch
 * per run (erm_group_set_tuning); the solvent's frame data and the batch
 * buffers are member 0's, and so is the PME tuning.  After
 * erm_group_set_frame the other members hold the frame's cell but not its
 * solvent data: erm_set_frame them before using one alone.  Needs the
 * register PME path (grid lengths 16-64, 96 in FP32; spline order 4 or 6). */
typedef struct erm_group erm_group;
int32_t erm_group_create(erm_ctx *const *ctxs, int32_t nctx, erm_group **out);
void erm_group_destroy(erm_group *g);
const char *erm_group_last_error(const erm_group *g);
int32_t erm_group_set_frame(erm_group *g, const double *sitepos, const double *cell, double stat_weight_system);
int32_t erm_group_set_frame_device(erm_group *g, const double *d_sitepos, const double *cell,
                                   double stat_weight_system);
int32_t erm_group_insert_random(erm_group *g, const int32_t *n, const uint64_t *seed, const uint64_t *first_index);
/* max_batch: configurations per batch over all members, up to 65536 (0
 * keeps it; 1024 by default); persistent: a bit mask of the stages that run
 * as one persistent grid over every run (workgroups pull items through an
 * atomic counter), bit 0 the z columns, bit 1 the pairs; a stage whose bit
 * is clear runs one launch per run; 2 by default, -1 keeps it. */
int32_t erm_group_set_tuning(erm_group *g, int32_t max_batch, int32_t persistent);
/* on = 1: the first erm_group_insert_random of a shape runs eagerly and
 * captures its batches' launches in a HIP graph; later calls of the same
 * shape (counts, batch size, persistence, the members' parameters and
 * buffers, the slab shape of the frame) copy the generator arguments and
 * launch the graph; any change captures again.  Stage times of a replayed
 * call hold the total only.  0 (the default): every call runs eagerly. */
int32_t erm_group_set_graph(erm_group *g, int32_t on);
/* stage times of the last erm_group_insert_random, as erm_stage_times */
int32_t erm_group_stage_times(const erm_group *g, double *ms10);

/* update_histogram alone, for energies computed elsewhere (n configurations,
 * uvengy n*(slvmax+1); solutes as in erm_engconst, NULL in refs mode). */
int32_t erm_update_histogram(erm_ctx *ctx, int32_t n, const int32_t *solutes,
                             const double *weights, const double *uvengy);

/* erm_update_histogram without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as erm_update_histogram; its scratch (the
 * energies as given and sorted, per-configuration histograms, weights and
 * per-species sums) lives in a 64 MiB device array of the library, one per
 * GPU, instead of context buffers.  n must fit that array:
 * n * (16 (slvmax + 1) + 8 (ermax + 35) + 8) bytes plus alignment, e.g. about
 * 950 configurations for 4000 molecules and 600 bins; larger n returns
 * ERM_E_ARG with the limit in the message.  Calls on one GPU share the array:
 * calls that overlap on one GPU (several host threads) race.  The context
 * must already hold its frame and bins (erm_set_frame, erm_set_bins). */
int32_t erm_update_histogram_static(erm_ctx *ctx, int32_t n, const int32_t *solutes,
                                    const double *weights, const double *uvengy);

/* engstore's normalisation (engproc.F90:539-553) on the device accumulators
 * (one process: no reduction step). */
int32_t erm_engstore(erm_ctx *ctx);

/* Copies the accumulators to the host.  Any pointer may be NULL.
 * edens[ermax]; ecorr[ermax*ermax] column-major as ecorr(iduvp, iduv);
 * eself[esmax]; minuv/maxuv[numslv+1]; slnuv[numslv+1] (index 0 unused);
 * scalars[4] = {engnorm, engsmpl, avslf, voffset}. */
int32_t erm_fetch(erm_ctx *ctx, double *edens, double *ecorr, double *eself, double *minuv,
                  double *maxuv, double *slnuv, double *scalars);

/* sizes: slvmax (physical molecules), ermax (total bins) */
int32_t erm_sizes(const erm_ctx *ctx, int32_t *slvmax, int32_t *ermax);

/* Device time of the last engconst / insert_random call, split by stage
 * (milliseconds, HIP events on the context's stream, summed over batches):
 * [0] poses and solute splines, [1] PME z columns (spread, z FFT, Green's
 * function), [2] PME inverse y transform, [3] PME x transform and gather
 * into per-site sums, [4] per-molecule sums, [5] real-space pairs, [6] self
 * terms, [7] binning, weights and edens, [8] correlation matrix, [9] total.
 * A call replayed from a graph reports the total only. */
int32_t erm_stage_times(const erm_ctx *ctx, double *ms10);

/* ---- batching and tuning ------------------------------------------------- */
/* No counterpart in the reference (its engconst runs one insertion at a
 * time).  max_batch: configurations per device batch (0 keeps the current);
 * pme_path as in erm_params; use_graph = 1 replays erm_insert_random calls
 * without uvengy from a captured HIP graph (the first call of a given shape
 * runs eagerly and captures; a change of frame geometry, batch shape or
 * buffers re-captures); slab_planes (1..16) caps the z planes of one fused
 * PME slab and group_planes (1..16) the planes the register path transforms
 * per group, slab_threads and column_threads (256 or 512) the workgroup
 * sizes of the register path's x slab and z columns (0 keeps any of them);
 * fold_y = 1 moves the inverse y transform into 512-thread x slabs of a
 * grid with K2 = K1, 0 keeps it in its own kernel, -1 keeps the setting. */
int32_t erm_set_tuning(erm_ctx *ctx, int32_t max_batch, int32_t pme_path, int32_t use_graph,
                       int32_t slab_planes, int32_t group_planes, int32_t slab_threads,
                       int32_t column_threads, int32_t fold_y);
/* out10 = {max_batch, pme_path, use_graph, slab threads (256 / 512), slab
 * planes in use, group planes in use, slab cap, group cap, z-column
 * threads (256 / 512), y transform inside the x slab (0 / 1)} */
int32_t erm_get_tuning(const erm_ctx *ctx, int32_t *out10);
/* Real-space kernel (no counterpart in the reference): 0 (default) walks the
 * candidates in tiles of 64 solvent sites and evaluates the pairs inside the
 * cutoff from a compacted per-wave list (realcal.F90:585-784 blocks the
 * same loop for the CPU); 1 gives each candidate molecule to one thread.
 * erm_autotune picks the faster.  erm_get_pair_path returns the setting. */
int32_t erm_set_pair_path(erm_ctx *ctx, int32_t path);
int32_t erm_get_pair_path(const erm_ctx *ctx);
/* Times n_trial random insertions (seed) on the current frame for batch
 * sizes 128..1024, every applicable PME path and slab / group shapes, the
 * real-space kernel, then graph replay against eager launches, and keeps
 * the fastest; accumulators
 * are untouched.  best_ms (optional) receives the chosen time. */
int32_t erm_autotune(erm_ctx *ctx, int32_t n_trial, uint64_t seed, double *best_ms);

/* ---- differentiation ---------------------------------------------------- */
/* Gradient of the solute-solvent energy u = sum_{k>=1} uvengy(k) (the sum
 * get_uv_energy, engproc.F90:662-756, leaves in uvengy(1:slvmax)) with
 * respect to the solute's per-site scaled charge q_i, LJ epsilon_i and
 * sigma_i.  The reference has no gradient; this is the adjoint of its energy:
 * real space over the pairs of realcal.F90:527-784, reciprocal space as the
 * solvent's PME potential at each solute site (recpcal.F90:342-502 is a
 * symmetric bilinear form), residual_ene (engproc.F90:859-869).
 * Configurations as in erm_engconst (soln: `solutes`, single solute type;
 * refs: solute_pos, or NULL for device random poses from seed/first_index as
 * in erm_insert_random).  q/eps/sigma (ns each, NULL = the topology's
 * values) override the solute's parameters; solvent LJ parameters are the
 * diagonals of the tables, combined by cmbrule (0 arithmetic sigma,
 * 1 geometric; epsilon geometric, setconf.F90:810-820).  Adds into
 * acc[2 + 3*ns]: {sum w, sum w u, sum w du/dq[ns], sum w du/deps[ns],
 * sum w du/dsigma[ns]} with w = stat_weight_system * weights[i] (NULL = 1).
 * Needs wgtslf = 0 and estype = 1 (weights independent of the parameters)
 * and register-FFT grid lengths (16, 24, 32, 48, 64; 96 in FP32). */
int32_t erm_solute_gradient(erm_ctx *ctx, int32_t n, const int32_t *solutes, const double *solute_pos,
                            uint64_t seed, uint64_t first_index, const double *weights, int32_t cmbrule,
                            const double *q, const double *eps, const double *sigma, double *acc);

#ifdef __cplusplus
}
#endif
#endif
