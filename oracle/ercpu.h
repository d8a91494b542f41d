/*
 * ercpu.h -- the oracle: an FP64 CPU restatement of ERmod's per-frame
 * solute-solvent energy pipeline (engproc / realcal / reciprocal / spline /
 * ptinsrt / quaternion), written from the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  It is linked by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, where it is the
 * checker and the timed CPU baseline.  The product library (libermgpu.so)
 * never links, loads or calls it.
 *
 * Conventions: 0-based indices; coordinates are AoS xyz (site i at
 * pos[3*i..3*i+2]); 3x3 matrices are column-major (m[3*j+i] = M(i+1,j+1)),
 * which is the reference's Fortran layout of `cell` / `invcl`.
 */
#ifndef ERCPU_H
#define ERCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: the reference's halt_with_error codes (mpiproc.F90:203-248)
 * that the hot path can raise; the values equal the product's ERM_E_* codes. */
enum {
  ORC_OK = 0,
  ORC_E_ENG_MIN = 1,     /* energy at or below the lowest edge (engproc.F90:992-995) */
  ORC_E_ENG_SFT = 2,     /* soln: energy above the soft region (engproc.F90:1000-1005) */
  ORC_E_MBIN = 3,        /* overflow bin below 1e3 (engproc.F90:1010-1015) */
  ORC_E_ENG_ECD = 4,     /* inconsistent bin parameters (engproc.F90:185) */
  ORC_E_RCP_CNS = 5,     /* molecule without a spline slot (recpcal.F90:333) */
  ORC_E_NONCUBOID = 6,   /* non-cuboid cell (realcal.F90:859-864) */
  ORC_E_ENG_BUG = 7,     /* internal inconsistency (eng_bug) */
  ORC_E_ENG_CNS = 8,     /* solvent with species <= 0 (engproc.F90:812) */
  ORC_E_ENG_SIZ = 9,     /* too many bins (engproc.F90:194) */
  ORC_E_ARG = 10,        /* invalid argument */
  ORC_E_UNSUPPORTED = 11 /* option outside the restated path */
};

/* &hist namelist (engproc.F90:66-67) for one species */
typedef struct orc_hist {
  double ecdmin, ecfmns, ecdcen, eccore, ecdmax;
  double eclbin, ecfbin, ec0bin, finfac;
  int32_t pecore;
} orc_hist;

/* Parameters and topology of the system: the engmain globals the hot path
 * reads (engmain.F90:252-301), as setparam leaves them (charges already
 * multiplied by 18.22261721, LJ tables already combined). */
typedef struct orc_system {
  int32_t slttype;      /* 1 soln, 2 refs rigid, 3 refs flexible */
  int32_t estype;       /* 1 NVT, 2 NPT */
  int32_t ljswitch;     /* 0 CHARMM pot, 1 GMX pot, 2 CHARMM force, 3 GMX force */
  int32_t splodr;       /* B-spline order p */
  int32_t ms[3];        /* PME grid */
  int32_t corrcal, selfcal, wgtslf;
  int32_t force_calculation;
  double elecut, lwljcut, upljcut, screen, temp, block_threshold;
  int32_t natom, nmol, ntype;
  const int32_t *mol_begin;   /* nmol+1 */
  const double *charge;       /* natom, scaled */
  const int32_t *ljtype;      /* natom, 0-based */
  const double *lj_eps;       /* ntype*ntype */
  const double *lj_sig2;      /* ntype*ntype */
  const int32_t *sluvid;      /* nmol: 0 solvent, 1 solute, 2/3 test particle */
  const int32_t *uvspec;      /* nmol: species 1..numslv for solvent, 0 for solute */
  const double *mol_charge;   /* nmol */
  /* energy coordinate (enginit output) */
  int32_t numslv;
  const int32_t *uvmax;       /* numslv */
  const int32_t *uvsoft;      /* numslv */
  const double *uvcrd;        /* ermax = sum(uvmax) */
  int32_t esmax;
  const double *escrd;        /* esmax (may be NULL when selfcal = 0) */
  /* electrostatics and boundary (engmain.F90:306,314) */
  int32_t cltype;             /* 0 bare Coulomb, 2 PME, 3 PPPM (1 Ewald: refused) */
  int32_t boxshp;             /* 0 non-periodic (bare Coulomb only), 1 periodic */
} orc_system;

typedef struct orc_ctx orc_ctx;

/* --- spline.F90 ------------------------------------------------------- */
void orc_spline_coeff(int32_t order, double *coeff /* order+1 */);
double orc_spline_value(const double *coeff, int32_t order, double x);

/* --- enginit edge construction (engproc.F90:124-237) -------------------- */
/* writes pemax edges into ercrd[0..pemax-1]; self_energy != 0 treats the
 * species as the solute self-energy (pecore forced to 0).  Returns status. */
int32_t orc_build_edges(const orc_hist *h, int32_t self_energy, double *ercrd,
                        int32_t cap, int32_t *pemax, int32_t *pesoft);
int32_t orc_binsearch(const double *coord, int32_t n, double v);
void orc_repval(const double *crd, int32_t idmax, int32_t idsoft, int32_t idpt,
                double *out);

/* --- hash_real.c:5-15 ---------------------------------------------------- */
uint64_t orc_hash_double(const double *v, int32_t n);

/* --- cell (engproc.F90:889-919) ----------------------------------------- */
void orc_inverted_cell(const double *cell, double *invcl, double *volume);

/* --- FFT with the reference's FFT-library conventions ------------------- */
/* r2c: real (K1,K2,K3) x-fastest -> complex (K1/2+1,K2,K3), forward e^{-i},
 * unnormalised.  c2r: the inverse with e^{+i}, unnormalised. Interleaved
 * complex (re, im). */
int32_t orc_fft_r2c_3d(const int32_t *K, const double *in, double *out);
int32_t orc_fft_c2r_3d(const int32_t *K, const double *in, double *out);

/* --- recpcal.F90 pieces ------------------------------------------------- */
void orc_spline_axis(int32_t K, int32_t p, double *splfc);
void orc_greenfunc(const int32_t *K, int32_t p, const double *invcl, double volume,
                   double screen, const double *splfc1, const double *splfc2,
                   const double *splfc3, double *engfac /* (K1/2+1)*K2*K3 */);
/* PPPM (recpcal.F90:153-325): the denominator coefficients gf_b (p of
 * them, calc_gfb_pppm) and the optimal influence function on the
 * (K1/2+1, K2, K3) half spectrum.  Returns ORC_E_ARG for an odd order or a
 * negative alias range. */
void orc_gfb_pppm(int32_t p, double *gf_b);
int32_t orc_pppm_greenfunc(const int32_t *K, int32_t p, const double *cell, const double *invcl,
                           double volume, double screen, double *engfac);
void orc_spline_site(const double *x, const double *invcl, const int32_t *K,
                     int32_t p, const double *coeff, double *w /* p*3 */,
                     int32_t *g /* 3 */);

/* --- test-particle pose (insertion.F90:103-366, bestfit.F90:49-203) ------ */
void orc_place_solute(int32_t n, const double *bfcoord, const double *mass,
                      const double *cell, const double *r /* 3 in [0,1) */,
                      const double *q /* 4, unit */, double *out);
/* counter-based uniform draws and the random pose of insertion `index`,
 * identical bit for bit to the device generator. */
double orc_rng_uniform(uint64_t seed, uint64_t index, uint32_t draw);
void orc_random_pose(uint64_t seed, uint64_t index, double *r, double *q);

/* insertion modes (engmain.F90:320-327): insorigin 0 ORIGIN / 2 AGGCEN,
 * insposition 0 RANDOM / 2 SPHERE / 3 SLAB_GENERIC / 4 SLAB_SYMMETRIC,
 * insorient 0 RANDOM / 1 NOCHANGE; lwreg, upreg the region bounds */
typedef struct orc_insertion {
  int32_t insorigin, insposition, insorient;
  double lwreg, upreg;
} orc_insertion;
/* com_aggregate (insertion.F90:539-566): centre of mass of the sites of the
 * molecules with hostmol[m] != 0, sequential sums in molecule order */
void orc_aggregate_center(int32_t nmol, const int32_t *mol_begin, const int32_t *hostmol,
                          const double *sitemass, const double *pos, double *center);
/* instslt's placement (insertion.F90:103-366) of insertion `index` of the
 * generator stream `seed` under the given modes: the draws of
 * orc_random_pose for the random position and the orientation, a second
 * stream (seed ^ 0xD1B54A32D192ED03) for the sphere and slab draws.
 * aggcen: the aggregate centre (AGGCEN only). */
void orc_insertion_pose(int32_t n, const double *bfcoord, const double *mass, const double *cell,
                        const orc_insertion *ins, const double *aggcen, uint64_t seed, uint64_t index,
                        double *out);

/* bestfit.F90: fit_a_rotate_b (214-241; Horn's quaternion from the top
 * eigenvector of the 4x4 key matrix, here by Jacobi rotations instead of
 * dsyev) and rmsd_nofit (252-268); coordinates 3 x n */
void orc_fit_a_rotate_b(int32_t na, const double *refa, const double *a, const double *w, int32_t nb,
                        const double *b, double *bout);
double orc_rmsd_nofit(int32_t n, const double *a, const double *b, const double *w);
/* A reference-structure insertion (insorigin REFSTR, insposition 5 RMSD or
 * 6 GAUSS) from the frame's fitted solute and best-fit reference solute:
 * attempts of position + orientation until rmsd_nofit lies in [lwreg,
 * upreg].  gconst = (L1/2)/upreg, (L2/2)/upreg, (L3/2)/upreg, z0, z1 (GAUSS).
 * *weight: the GAUSS weight (1 for RMSD).  Returns 14 after 4096 attempts. */
int32_t orc_refstr_pose(int32_t n, const double *mass, const double *cell, const orc_insertion *ins,
                        const double *fitted, const double *best, const double *w, const double *gconst,
                        uint64_t seed, uint64_t index, double *out, double *weight);

/* --- context: the per-frame pipeline ---------------------------------- */
orc_ctx *orc_create(const orc_system *sys, int32_t *status);
void orc_destroy(orc_ctx *c);
int32_t orc_slvmax(const orc_ctx *c);
/* frame: sitepos natom*3 (test-particle slots ignored), cell column-major */
int32_t orc_set_frame(orc_ctx *c, const double *sitepos, const double *cell,
                      double stat_weight_system);
/* get_uv_energy (engproc.F90:662-756) for solute molecule `solute` (0-based).
 * solute_pos: coordinates of its sites (NULL: take them from the frame).
 * uvengy: slvmax+1 values, [0] = self energy. */
int32_t orc_uv_energy(orc_ctx *c, int32_t solute, const double *solute_pos,
                      double *uvengy);
/* the pieces of get_uv_energy, for stage-wise checks */
int32_t orc_realcal_proc(orc_ctx *c, int32_t solute, const double *solute_pos,
                         double *ereal /* slvmax */);
int32_t orc_realcal_brute(orc_ctx *c, int32_t solute, const double *solute_pos,
                          double *ereal /* slvmax */);
/* realcal_bare (realcal.F90:155-281): bare Coulomb + LJ of the solute with
 * every physical molecule (cutoff elecut under the minimum image when
 * periodic, none otherwise) */
int32_t orc_realcal_bare(orc_ctx *c, int32_t solute, const double *solute_pos,
                         double *ereal /* slvmax */);
int32_t orc_realcal_self(orc_ctx *c, int32_t solute, const double *solute_pos,
                         double *pairep);
int32_t orc_recpcal(orc_ctx *c, int32_t solute, const double *solute_pos,
                    double *erecp /* slvmax */, double *self_energy,
                    double *phi /* K1*K2*K3 or NULL */);
/* update_histogram (engproc.F90:758-856); insdst (ermax) optional out */
int32_t orc_update_histogram(orc_ctx *c, int32_t solute, double stat_weight_solute,
                             const double *uvengy, int32_t *insdst);
void orc_engclear(orc_ctx *c);
/* engstore normalisation (engproc.F90:475-566) for a single process */
int32_t orc_engstore(orc_ctx *c);
/* accumulator access (pointers into the context) */
typedef struct orc_acc {
  double *edens, *ecorr, *eself, *minuv, *maxuv, *slnuv;
  double engnorm, engsmpl, avslf, voffset;
  int32_t ermax, esmax, numslv, voffset_initialized;
} orc_acc;
void orc_get_acc(orc_ctx *c, orc_acc *out);
/* The intermediate-state method (extensions/Intermediate/engproc.F90,
 * solution systems): zerodns[ermax] marks the bins empty in the reference
 * solvent (NULL switches it off).  Accumulates edintm, uvintm (numslv+1,
 * [0] unused), norm_intm (per division) and soln_intm, soln_total (over
 * all divisions); scal = {norm_intm, soln_intm, soln_total}.  engstore
 * divides edintm by norm_intm. */
int32_t orc_set_intermediate(orc_ctx *c, const int32_t *zerodns);
int32_t orc_get_intermediate(const orc_ctx *c, double *edintm, double *uvintm, double *scal);
/* per-species sum of solute-solvent energy of the last update (flceng) */
const double *orc_flceng(const orc_ctx *c);

/* Solute-solvent energy u = sum_{k>=1} uvengy(k) of one configuration and
 * its gradient with respect to the solute's per-site scaled charges q,
 * LJ epsilon and sigma (NULL: the topology's values; solvent LJ parameters
 * are the table diagonals, combined by cmbrule 0 arithmetic / 1 geometric).
 * out: [u, du/dq (ns), du/deps (ns), du/dsigma (ns)]. Brute-force real
 * space; reciprocal space through the solvent's PME potential. */
int32_t orc_solute_gradient(orc_ctx *c, int32_t solute, const double *solute_pos,
                            int32_t cmbrule, const double *q, const double *eps,
                            const double *sigma, double *out);


/* CPU baseline (not part of the reference's interface): n random insertions
 * of rigid test particles, insertion i of solute type i mod ntypes, pose from
 * the counter-based generator (seed, first_index + i) as orc_random_pose +
 * orc_place_solute, then uv_energy + update_histogram on
 * ctxs[thread * ntypes + type] (each context holds the frame already).
 * Stops early once `seconds` have passed; done_per_type[ntypes] counts the
 * insertions completed. */
int32_t orc_bench_insertions(orc_ctx *const *ctxs, int32_t nthreads, int32_t ntypes,
                             const int32_t *solute_mol, const int32_t *nsite,
                             const double *const *bfcoord, const double *const *mass,
                             const double *cell, int64_t n, uint64_t seed, uint64_t first_index,
                             double seconds, int64_t *done_per_type);

#ifdef __cplusplus
}
#endif
#endif
