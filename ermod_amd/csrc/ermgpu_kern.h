// ermgpu_kern.h -- kernels of the solute-solvent energy pipeline (gfx950).
//
// Data layout in HBM (per context):
//   frame, sorted by cell of each physical molecule's first site:
//     mol_*[j]   molecule j of the sorted order: tagpt index k, first sorted
//                site, site count, molecular charge, species
//     s*[t]      sorted sites, SoA: wrapped x/y/z (f64), charge (T), compact
//                LJ type, PME grid base (packed 3 x 10 bits) and the 3 x p
//                B-spline weights (T)
//     cell_start[c]  first sorted molecule of cell c (x fastest)
//   batch of B solute configurations b:
//     spos[b][s][3] (f64), solute spline data, Hk[b][z][m2][m1] (complex T,
//     half spectrum after the z-convolution), phi[b][z][y][x] (T),
//     uv[b][0..slvmax] (f64, sorted molecule order), hist[b][ermax] (f64)
#pragma once

#include "ermgpu_dev.h"
#include "ermgpu_erfc.h"

namespace erm {

struct Params {
  // grid / PME
  int K1, K2, K3, H, p;
  FftAxis ax1, ax2, ax3;
  double coeff[MAXP + 1];
  double L[3], halfL[3], invcl[9], volume, screen;
  // real space
  int ljswitch;
  double lw2, up2, el2, cut2, rcmax, lwljcut, upljcut;
  double lw3, up3, lw6, up6, repA, repB, repC, attA, attB, attC;
  int nb[3];          // molecule cell grid
  double hcell[3];
  // sizes
  int slvmax, ns, nvt, ermax, numslv, esmax;
  int slttype, wgtslf, estype, selfcal, force;
  double temp, stat_weight_system, total_charge, solute_charge;
  // electrostatics (0 bare Coulomb, 2 PME, 3 PPPM) and boundary (0 none)
  int cltype, boxshp;
  double elecut;
  // test-particle placement: insorigin, insposition, insorient and the
  // region bounds lwreg / upreg (engmain.F90:320-327)
  int ins_org, ins_pos, ins_rot;
  double lwreg, upreg;
  // intermediate-state accumulation on (solution systems)
  int intm;
  // reference-structure insertion (GAUSS): (L/2)/upreg per axis and the
  // normalisations z0, z1 of the uniform and Gaussian parts
  double gl[3], gz0, gz1;
};

// --------------------------------------------------------------------------
// per-frame preparation
// --------------------------------------------------------------------------
__device__ __forceinline__ double wrap1(double x, double L) { return x - L * floor(x * (1.0 / L)); }

__device__ __forceinline__ int cell_index(const Params &P, const double *w) {
  int c[3];
  for (int d = 0; d < 3; ++d) {
    int t = (int)floor(w[d] * (1.0 / P.L[d]) * (double)P.nb[d]);
    t %= P.nb[d];
    if (t < 0) t += P.nb[d];
    c[d] = t;
  }
  return c[0] + P.nb[0] * (c[1] + P.nb[1] * c[2]);
}

// cell of each physical molecule (by its first site) and the largest
// distance of a site from its molecule's first site
static __global__ void k_mol_cell(Params P, const double *__restrict__ pos, const int *__restrict__ tagpt,
                           const int *__restrict__ mol_begin, int *cellid, int *count,
                           unsigned long long *rm_ord) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.slvmax) return;
  int m = tagpt[k];
  int a0 = mol_begin[m], a1 = mol_begin[m + 1];
  double w[3];
  for (int d = 0; d < 3; ++d) w[d] = wrap1(pos[3 * a0 + d], P.L[d]);
  int c = P.boxshp ? cell_index(P, w) : 0;
  cellid[k] = c;
  atomicAdd(&count[c], 1);
  double rm = 0.0;
  for (int a = a0 + 1; a < a1; ++a) {
    double s2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      double x = pos[3 * a + d] - pos[3 * a0 + d];
      if (P.boxshp) x -= P.L[d] * rint(x / P.L[d]);
      s2 += x * x;
    }
    rm = fmax(rm, sqrt(s2));
  }
  atomicMax(rm_ord, dbl_to_ord(rm));
}

// Exclusive scan of in[0, n) into out[0, n], out[n] the total (one
// workgroup).  copy: a second copy of out[0, n) (the fill cursors of a
// counting sort), or null; planes: planes[z] = out[z KB] for z <= K3, or
// null.
static __global__ __launch_bounds__(1024) void k_scan(const int *__restrict__ in, int *out, int n, int *copy, int *planes,
                                               int K3, int KB) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int b0 = t * chunk;
  int s = 0;
  for (int i = 0; i < chunk; ++i)
    if (b0 + i < n) s += in[b0 + i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : part[t - 1];
  for (int i = 0; i < chunk; ++i)
    if (b0 + i < n) {
      int v = in[b0 + i];
      out[b0 + i] = run;
      if (copy) copy[b0 + i] = run;
      run += v;
    }
  if (t == 1023) out[n] = part[1023];
  if (planes) {
    __syncthreads();
    for (int z = t; z <= K3; z += blockDim.x) planes[z] = out[(size_t)z * KB];
  }
}

// the per-frame counters zeroed in one launch: cell counts, z-bucket
// counts (nz may be 0) and the molecule-extent maximum
static __global__ void k_frame_clear(int *count, int nc, int *zcount, int nz, unsigned long long *rm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) count[i] = 0;
  if (i < nz) zcount[i] = 0;
  if (i == 0) *rm = 0ULL;
}

static __global__ void k_mol_scatter(Params P, const int *__restrict__ cellid, int *fill,
                              const int *__restrict__ tagpt, const int *__restrict__ mol_begin,
                              int *mol_k, int *mol_ns, int *jof) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.slvmax) return;
  int j = atomicAdd(&fill[cellid[k]], 1);
  int m = tagpt[k];
  mol_k[j] = k;
  jof[k] = j;
  mol_ns[j] = mol_begin[m + 1] - mol_begin[m];
}

// sorted site arrays: wrapped coordinates (realcal.F90:789-885, cuboid),
// charges, compact LJ types and B-spline weights (calc_spline_molecule,
// recpcal.F90:408-442)
template <typename T>
__global__ void k_site_fill(Params P, const double *__restrict__ pos, const int *__restrict__ tagpt,
                            const int *__restrict__ mol_begin, const double *__restrict__ charge,
                            const int *__restrict__ cvt, const double *__restrict__ mol_charge,
                            const int *__restrict__ uvspec, const int *__restrict__ mol_k,
                            const int *__restrict__ mol_first, double *sx, double *sy, double *sz,
                            T *sq, int *sct, int *sg, T *sw, double *mol_q, int *mol_spec,
                            int *mol_m, double *mol_d1) {
#pragma clang fp contract(off)
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.slvmax) return;
  int m = tagpt[mol_k[j]];
  mol_q[j] = mol_charge[m];
  mol_spec[j] = uvspec[m];
  mol_m[j] = m;
  {
    // the molecule's reach from its first site (minimum images in a
    // periodic box, the pair kernels' convention): no site of the molecule
    // lies farther than mol_d1 from its first
    const double *x0 = pos + 3 * mol_begin[m];
    double d1 = 0.0;
    for (int a = mol_begin[m] + 1; a < mol_begin[m + 1]; ++a) {
      double dd = 0.0;
      for (int e = 0; e < 3; ++e) {
        double d = pos[3 * a + e] - x0[e];
        if (P.boxshp) d = d - P.L[e] * rint(d / P.L[e]);
        dd = dd + d * d;
      }
      d1 = fmax(d1, sqrt(dd));
    }
    mol_d1[j] = d1;
  }
  int t = mol_first[j];
  const int Ks[3] = {P.K1, P.K2, P.K3};
  for (int a = mol_begin[m]; a < mol_begin[m + 1]; ++a, ++t) {
    const double *x = pos + 3 * a;
    // raw coordinates without a box (realcal_prepare, realcal.F90:137-146)
    sx[t] = P.boxshp ? wrap1(x[0], P.L[0]) : x[0];
    sy[t] = P.boxshp ? wrap1(x[1], P.L[1]) : x[1];
    sz[t] = P.boxshp ? wrap1(x[2], P.L[2]) : x[2];
    sq[t] = (T)charge[a];
    sct[t] = cvt[a];
    if (P.cltype == 0) continue;  // no reciprocal space
    int g[3];
    for (int d = 0; d < 3; ++d) {
      double f = P.invcl[d] * x[0] + P.invcl[3 + d] * x[1] + P.invcl[6 + d] * x[2];
      f = f - floor(f);
      double factor = f * (double)Ks[d];
      // f - floor(f) may round to 1.0: factor == K gives the weights of
      // base K, i.e. of base 0 (indices are taken mod K); NaN maps to 0
      int rci = (factor >= 0.0 && factor <= (double)Ks[d]) ? (int)factor : 0;
      for (int s = 0; s < P.p; ++s)
        sw[((size_t)t * 3 + d) * P.p + s] = (T)spline_value_dev(P.coeff, P.p, factor - (double)(rci - s));
      g[d] = rci % Ks[d];
    }
    sg[t] = g[0] | (g[1] << 10) | (g[2] << 20);
  }
}

// --------------------------------------------------------------------------
// solute configurations
// --------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double rng_uniform(uint64_t seed, uint64_t index, uint32_t draw) {
  uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ULL);
  uint64_t ctr = index * 256ULL + (uint64_t)draw;
  uint64_t z = mix64(key ^ mix64(ctr + 0x632BE59BD9B4E019ULL));
  return (double)(z >> 11) * 0x1.0p-53;
}

static __device__ void com_dev(int n, const double *pts, const double *mass, double *c) {
#pragma clang fp contract(off)
  double msum = 0.0;
  for (int j = 0; j < n; ++j) msum = msum + mass[j];
  for (int i = 0; i < 3; ++i) {
    double d = 0.0;
    for (int j = 0; j < n; ++j) d = d + pts[3 * j + i] * mass[j];
    c[i] = d / msum;
  }
}

static __device__ void qprod_dev(const double *a, const double *b, double *p) {
#pragma clang fp contract(off)
  double dot = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  p[0] = a[0] * b[0] - dot;
  double cr[3] = {a[2] * b[3] - a[3] * b[2], a[3] * b[1] - a[1] * b[3], a[1] * b[2] - a[2] * b[1]};
  for (int i = 0; i < 3; ++i) p[1 + i] = cr[i] + a[0] * b[1 + i] + b[0] * a[1 + i];
}

// a trajectory frame (FP32 SoA x[n], y[n], z[n]) into the FP64 AoS
// coordinates through the permutation (setconf.F90:983-987): trajectory atom
// i becomes atom perm[i]; the conversion to FP64 is exact
static __global__ void k_frame_scatter(int n, const float *__restrict__ xyz, const int *__restrict__ perm, double *pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = perm ? perm[i] : i;
  pos[3 * a] = (double)xyz[i];
  pos[3 * a + 1] = (double)xyz[n + i];
  pos[3 * a + 2] = (double)xyz[2 * n + i];
}

// coordinates of the listed atoms (the reference-structure host atoms)
static __global__ void k_gather_atoms(int n, const int *__restrict__ idx, const double *__restrict__ pos, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) out[3 * i + d] = pos[3 * idx[i] + d];
}

// com_aggregate (insertion.F90:539-566): centre of mass of every site of
// the host molecules, summed in molecule order by a single thread so that the
// rounding is that of the sequential sums (center_of_mass, bestfit.F90:158-172)
static __global__ void k_aggcen(int nmol, const int *__restrict__ mol_begin, const int *__restrict__ hostmol,
                         const double *__restrict__ mass, const double *__restrict__ pos, double *c) {
#pragma clang fp contract(off)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double msum = 0.0;
  for (int m = 0; m < nmol; ++m)
    if (hostmol[m])
      for (int a = mol_begin[m]; a < mol_begin[m + 1]; ++a) msum = msum + mass[a];
  for (int i = 0; i < 3; ++i) {
    double d = 0.0;
    for (int m = 0; m < nmol; ++m)
      if (hostmol[m])
        for (int a = mol_begin[m]; a < mol_begin[m + 1]; ++a) d = d + pos[3 * a + i] * mass[a];
    c[i] = d / msum;
  }
}

// second generator stream, for the position draws of the sphere and slab
// modes (the random-position draws 0-2 and the orientation draws 3 + 4a + i
// stay in the first)
constexpr uint64_t POS_STREAM = 0xD1B54A32D192ED03ULL;

// test-particle pose (instslt, insertion.F90:43-101) placed with the
// reference's operation sequence: set_solute_origin (ORIGIN / NOCHANGE /
// AGGCEN, insertion.F90:103-131), set_shift_com (RANDOM / SPHERE / SLAB
// generic and symmetric, insertion.F90:133-199) and apply_orientation
// (RANDOM / NOCHANGE, insertion.F90:325-366; bestfit.F90:49-100, 174-203).
// args = {seed, first index of the call}; insertion b of this launch is
// number args[1] + offset + b of the generator stream (device-side arguments,
// so a captured graph replays with new seeds).  aggcen: the aggregate centre
// of the frame (AGGCEN only).  bf_stride 0: the rigid body bf for every
// insertion; 3 * ns: body b of a flexible solute (getsolute's SltConf
// configurations, insertion.F90:384-480).
// One 64-lane workgroup per insertion: the lanes update the sites in
// parallel and every lane forms the same scalars (draws, centres of mass in
// site order) from the positions, kept in LDS when 4 ns doubles fit
// (POSE_LDS_SITES), else in spos itself.
constexpr int POSE_LDS_SITES = 2048;
// IN_LDS: the sites and masses are staged in dynamic LDS (ns <=
// POSE_LDS_SITES), else the sites are updated in place in the output; a
// compile-time choice keeps each pointer in one address space
template <bool IN_LDS>
static __global__ __launch_bounds__(64) void k_pose(Params P, int n, const uint64_t *__restrict__ args, uint64_t offset,
                                             const double *__restrict__ bf, int bf_stride,
                                             const double *__restrict__ mass, const double *__restrict__ cellm,
                                             const double *__restrict__ aggcen, double *spos) {
#pragma clang fp contract(off)
  extern __shared__ double pose_lds[];
  const int b = blockIdx.x;
  if (b >= n) return;
  const int ns = P.ns, ln = threadIdx.x, nl = blockDim.x;
  const uint64_t seed = args[0], sp = seed ^ POS_STREAM;
  const uint64_t index = args[1] + offset + (uint64_t)b;
  double *out = spos + (size_t)b * ns * 3;
  const double *body = bf + (size_t)b * bf_stride;  // one body per insertion for a flexible solute
  constexpr bool in_lds = IN_LDS;
  double *x = in_lds ? pose_lds : out;
  const double *ms = in_lds ? pose_lds + 3 * ns : mass;
  for (int i = ln; i < 3 * ns; i += nl) x[i] = body[i];
  if constexpr (in_lds)
    for (int j = ln; j < ns; j += nl) pose_lds[3 * ns + j] = mass[j];
  __syncthreads();
  double c[3];
  if (P.ins_org != 1) {
    // COM to the origin or to the aggregate centre (set_solute_com)
    const double o[3] = {P.ins_org == 2 ? aggcen[0] : 0.0, P.ins_org == 2 ? aggcen[1] : 0.0,
                         P.ins_org == 2 ? aggcen[2] : 0.0};
    com_dev(ns, x, ms, c);
    __syncthreads();
    for (int j = ln; j < ns; j += nl)
      for (int i = 0; i < 3; ++i) x[3 * j + i] = (x[3 * j + i] - c[i]) + o[i];
    __syncthreads();
  }
  if (P.ins_pos == 0) {
    // fully random position in the cell: COM = cell . r
    double r[3], com[3];
    for (int i = 0; i < 3; ++i) r[i] = rng_uniform(seed, index, (uint32_t)i);
    for (int i = 0; i < 3; ++i) com[i] = cellm[i] * r[0] + cellm[3 + i] * r[1] + cellm[6 + i] * r[2];
    com_dev(ns, x, ms, c);
    __syncthreads();
    for (int j = ln; j < ns; j += nl)
      for (int i = 0; i < 3; ++i) x[3 * j + i] = (x[3 * j + i] - c[i]) + com[i];
    __syncthreads();
  } else if (P.ins_pos == 2) {
    // shell lwreg <= |shift| <= upreg, uniform in volume, direction by
    // rejection from the cube
    const double lw3 = P.lwreg * P.lwreg * P.lwreg, up3 = P.upreg * P.upreg * P.upreg;
    double t = rng_uniform(sp, index, 0);
    t = pow(t * (up3 - lw3) + lw3, 1.0 / 3.0);
    double r[3] = {0.0, 0.0, 1.0}, nrm = 1.0;
    for (int a = 0; a < 84; ++a) {
      double u[3];
      for (int i = 0; i < 3; ++i) u[i] = rng_uniform(sp, index, (uint32_t)(1 + 3 * a + i)) * 2.0 - 1.0;
      double s2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      if (0.0 < s2 && s2 <= 1.0) {
        for (int i = 0; i < 3; ++i) r[i] = u[i];
        nrm = s2;
        break;
      }
    }
    nrm = sqrt(nrm);
    for (int j = ln; j < ns; j += nl)
      for (int i = 0; i < 3; ++i) x[3 * j + i] = x[3 * j + i] + t * r[i] / nrm;
    __syncthreads();
  } else if (P.ins_pos == 3 || P.ins_pos == 4) {
    // slab: random in-plane position, lwreg <= z offset <= upreg (either
    // sign with equal probability for the symmetric slab)
    double r[3];
    r[0] = rng_uniform(sp, index, 0);
    r[1] = rng_uniform(sp, index, 1);
    double t = rng_uniform(sp, index, 2);
    t = t * (P.upreg - P.lwreg) + P.lwreg;
    if (P.ins_pos == 4) {
      double dir = rng_uniform(sp, index, 3);
      if (dir > 0.5) t = -t;
    }
    const double len3 = sqrt(cellm[6] * cellm[6] + cellm[7] * cellm[7] + cellm[8] * cellm[8]);
    r[2] = t / len3;
    double com[3];
    for (int i = 0; i < 3; ++i) com[i] = cellm[i] * r[0] + cellm[3 + i] * r[1] + cellm[6 + i] * r[2];
    for (int j = ln; j < ns; j += nl)
      for (int i = 0; i < 3; ++i) x[3 * j + i] = x[3 * j + i] + com[i];
    __syncthreads();
  }
  if (P.ins_rot == 0) {
    // uniform rotation: quaternion by rejection from the 4-cube
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < 63; ++a) {
      double t[4];
      for (int i = 0; i < 4; ++i) t[i] = rng_uniform(seed, index, (uint32_t)(3 + 4 * a + i)) * 2.0 - 1.0;
      double s = t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3];
      if (s < 1.0) {
        double nrm = sqrt(s);
        for (int i = 0; i < 4; ++i) q[i] = t[i] / nrm;
        break;
      }
    }
    com_dev(ns, x, ms, c);
    __syncthreads();
    double cj[4] = {q[0], -q[1], -q[2], -q[3]};
    for (int j = ln; j < ns; j += nl) {
      double q2[4] = {0.0, x[3 * j] - c[0], x[3 * j + 1] - c[1], x[3 * j + 2] - c[2]};
      double t[4], r2[4];
      qprod_dev(q, q2, t);
      qprod_dev(t, cj, r2);
      for (int i = 0; i < 3; ++i) x[3 * j + i] = r2[1 + i] + c[i];
    }
    __syncthreads();
  }
  if (in_lds)
    for (int i = ln; i < 3 * ns; i += nl) out[i] = x[i];
}

template <typename T>
__global__ __launch_bounds__(128) void k_solute_prep(Params P, const double *__restrict__ spos,
                                                     const double *__restrict__ qsolute,
                                                     const typename C2<T>::type *__restrict__ tw1,
                                                     const typename C2<T>::type *__restrict__ tw2,
                                                     const typename C2<T>::type *__restrict__ tw3,
                                                     double *sref, int *sg3, T *sw3, T *sqv,
                                                     typename C2<T>::type *w1h,
                                                     typename C2<T>::type *w2h,
                                                     typename C2<T>::type *w3h) {
#pragma clang fp contract(off)
  using V = typename C2<T>::type;
  const int b = blockIdx.x, ns = P.ns, p = P.p;
  const double *x = spos + (size_t)b * ns * 3;
  __shared__ double red[128 * 3];
  __shared__ int gsh[3 * 64];
  __shared__ double wsh[3 * MAXP * 64];
  // reference point
  double s0 = 0, s1 = 0, s2 = 0;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    s0 += x[3 * s];
    s1 += x[3 * s + 1];
    s2 += x[3 * s + 2];
  }
  red[threadIdx.x * 3] = s0;
  red[threadIdx.x * 3 + 1] = s1;
  red[threadIdx.x * 3 + 2] = s2;
  __syncthreads();
  __shared__ double cpt[4];
  if (threadIdx.x == 0) {
    double a = 0, bb = 0, c = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      a += red[3 * i];
      bb += red[3 * i + 1];
      c += red[3 * i + 2];
    }
    cpt[0] = a / ns;
    cpt[1] = bb / ns;
    cpt[2] = c / ns;
  }
  __syncthreads();
  double rmax = 0.0;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    double d0 = x[3 * s] - cpt[0], d1 = x[3 * s + 1] - cpt[1], d2 = x[3 * s + 2] - cpt[2];
    rmax = fmax(rmax, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
  }
  red[threadIdx.x] = rmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) r = fmax(r, red[i]);
    sref[4 * b + 0] = cpt[0];
    sref[4 * b + 1] = cpt[1];
    sref[4 * b + 2] = cpt[2];
    sref[4 * b + 3] = r;
  }
  if (P.cltype == 0) return;  // bare Coulomb: no reciprocal space
  // splines, in chunks of 64 sites through LDS
  const int Ks[3] = {P.K1, P.K2, P.K3};
  for (int s0c = 0; s0c < ns; s0c += 64) {
    int nc = min(64, ns - s0c);
    __syncthreads();
    for (int t = threadIdx.x; t < nc * 3; t += blockDim.x) {
      int s = s0c + t / 3, d = t % 3;
      const double *xs = x + 3 * s;
      double f = P.invcl[d] * xs[0] + P.invcl[3 + d] * xs[1] + P.invcl[6 + d] * xs[2];
      f = f - floor(f);
      double factor = f * (double)Ks[d];
      int rci = (factor >= 0.0 && factor <= (double)Ks[d]) ? (int)factor : 0;
      for (int k = 0; k < p; ++k)
        wsh[((t / 3) * 3 + d) * MAXP + k] = spline_value_dev(P.coeff, p, factor - (double)(rci - k));
      gsh[(t / 3) * 3 + d] = rci % Ks[d];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nc; t += blockDim.x) {
      int s = s0c + t;
      sg3[(size_t)b * ns + s] = gsh[t * 3 + 2];
      sqv[(size_t)b * ns + s] = (T)qsolute[s];
      for (int k = 0; k < p; ++k) sw3[((size_t)b * ns + s) * p + k] = (T)wsh[(t * 3 + 2) * MAXP + k];
    }
    // spectra along the three axes (axis 3 only when w3h is requested)
    int nm = P.H + P.K2 + (w3h ? P.K3 : 0);
    for (int t = threadIdx.x; t < nc * nm; t += blockDim.x) {
      int sl = t / nm, mm = t - sl * nm;
      int d = (mm < P.H) ? 0 : (mm < P.H + P.K2) ? 1 : 2;
      int m = (d == 0) ? mm : (d == 1) ? mm - P.H : mm - P.H - P.K2;
      int K = Ks[d];
      const V *tw = (d == 0) ? tw1 : (d == 1) ? tw2 : tw3;
      int g = gsh[sl * 3 + d];
      V acc;
      acc.x = 0;
      acc.y = 0;
      // twiddle index m (g - k) mod K, stepped by -m from m g mod K (m, g
      // < K <= 512, so m g fits an int)
      int ti = (m * g) % K;
      for (int k = 0; k < p; ++k) {
        V w = tw[ti];
        ti -= m;
        ti += (ti < 0) ? K : 0;
        T wt = (T)wsh[(sl * 3 + d) * MAXP + k];
        acc.x += wt * w.x;
        acc.y += wt * w.y;
      }
      int s = s0c + sl;
      if (d == 0)
        w1h[((size_t)b * ns + s) * P.H + m] = acc;
      else if (d == 1)
        w2h[((size_t)b * ns + s) * P.K2 + m] = acc;
      else
        w3h[((size_t)b * ns + s) * P.K3 + m] = acc;
    }
  }
}

// --------------------------------------------------------------------------
// PME: z columns.  For every half-spectrum column (m1, m2):
//   G(z)  = sum_s q_s what1_s(m1) what2_s(m2) w3_s(z)   (sparse in z)
//   Qhat  = FFT_z(G)                     (the r2c transform of the spread grid)
//   E_self += weight(m1) * engfac * |Qhat|^2             (recpcal.F90:385-394)
//   Hk(z) = IFFT_z(engfac * Qhat)                         (recpcal.F90:398,402)
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHR) void k_pme_zcol(Params P, int ncolwg, const T *__restrict__ engfac,
                                                   const int *__restrict__ sg3, const T *__restrict__ sw3,
                                                   const T *__restrict__ sqv,
                                                   const typename C2<T>::type *__restrict__ w1h,
                                                   const typename C2<T>::type *__restrict__ w2h,
                                                   const typename C2<T>::type *__restrict__ tw3,
                                                   typename C2<T>::type *Hk, double *selfp) {
  using V = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *buf0 = reinterpret_cast<V *>(lds_raw);
  V *buf1 = buf0 + ncolwg * P.K3;
  V *Acol = buf1 + ncolwg * P.K3;                       // [ZSCH][ncolwg]
  T *wz = reinterpret_cast<T *>(Acol + ZSCH * ncolwg);  // [ZSCH][p]
  int *gz = reinterpret_cast<int *>(wz + ZSCH * P.p);   // [ZSCH]
  __shared__ double red[NTHR];
  const int b = blockIdx.y, ns = P.ns, p = P.p, K3 = P.K3, H = P.H;
  const int ncol_tot = H * P.K2;
  const int c0 = blockIdx.x * ncolwg;
  const int nc = min(ncolwg, ncol_tot - c0);
  for (int t = threadIdx.x; t < ncolwg * K3; t += blockDim.x) {
    buf0[t].x = 0;
    buf0[t].y = 0;
  }
  // the column of (m1, m2) along z: sum_s q_s what1_s(m1) what2_s(m2) w3_s(z),
  // sites in chunks of ZSCH through LDS, one thread per (column, z)
  for (int s0 = 0; s0 < ns; s0 += ZSCH) {
    const int nsc = min(ZSCH, ns - s0);
    __syncthreads();
    for (int t = threadIdx.x; t < nsc * ncolwg; t += blockDim.x) {
      int sl = t / ncolwg, cl = t - sl * ncolwg;
      int c = c0 + cl;
      V a;
      a.x = 0;
      a.y = 0;
      if (cl < nc) {
        size_t bs = (size_t)b * ns + s0 + sl;
        a = cmul(w1h[bs * H + c % H], w2h[bs * P.K2 + c / H]);
        T q = sqv[bs];
        a.x *= q;
        a.y *= q;
      }
      Acol[t] = a;
    }
    for (int t = threadIdx.x; t < nsc; t += blockDim.x) {
      size_t bs = (size_t)b * ns + s0 + t;
      gz[t] = sg3[bs];
      for (int k = 0; k < p; ++k) wz[t * p + k] = sw3[bs * p + k];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ncolwg * K3; t += blockDim.x) {
      int cl = t / K3, z = t - cl * K3;
      V acc = buf0[t];
      for (int sl = 0; sl < nsc; ++sl) {
        int k = gz[sl] - z;
        k += (k < 0) ? K3 : 0;
        if (k < p) {
          V a = Acol[sl * ncolwg + cl];
          T w = wz[sl * p + k];
          acc.x += a.x * w;
          acc.y += a.y * w;
        }
      }
      buf0[t] = acc;
    }
  }
  __syncthreads();
  V *res = lds_fft(buf0, buf1, ncolwg, K3, 1, P.ax3, -1, tw3);
  // Green's function and self energy
  double self = 0.0;
  for (int t = threadIdx.x; t < nc * K3; t += blockDim.x) {
    int cl = t / K3, m3 = t - cl * K3;
    int c = c0 + cl;
    int m1 = c % H, m2 = c / H;
    T g = engfac[((size_t)m3 * P.K2 + m2) * H + m1];
    V v = res[cl * K3 + m3];
    double wgt = (m1 == 0 || (P.K1 % 2 == 0 && m1 == H - 1)) ? 0.5 : 1.0;
    self += wgt * (double)g * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    v.x *= g;
    v.y *= g;
    res[cl * K3 + m3] = v;
  }
  red[threadIdx.x] = self;
  __syncthreads();
  V *other = (res == buf0) ? buf1 : buf0;
  V *out = lds_fft(res, other, ncolwg, K3, 1, P.ax3, +1, tw3);
  for (int s = NTHR / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) selfp[(size_t)b * gridDim.x + blockIdx.x] = red[0];
  // Hk[b][z][m2][m1]: consecutive columns are contiguous for fixed z
  for (int t = threadIdx.x; t < nc * K3; t += blockDim.x) {
    int z = t / nc, cl = t - z * nc;
    Hk[((size_t)b * K3 + z) * ncol_tot + c0 + cl] = out[cl * K3 + z];
  }
}

// --------------------------------------------------------------------------
// PME: one z plane per workgroup.  y inverse transform of the K2 x H block,
// then the x c2r transform, two real rows per complex FFT
// (Z = X_y + i X_{y+1} with Hermitian extension), into phi[b][z][y][x].
// --------------------------------------------------------------------------
// The inverse transform of one z plane in LDS: y inverse FFT of the K2 x H
// block src[m2][m1], then the x c2r transform two real rows per complex FFT
// (Z = X_y + i X_{y+1} with Hermitian extension).  Returns R with
// phi(y, x) = (y & 1) ? R[(y >> 1) K1 + x].y : R[(y >> 1) K1 + x].x.
template <typename T>
__device__ typename C2<T>::type *plane_c2r(const Params &P, const typename C2<T>::type *__restrict__ src,
                                           typename C2<T>::type *A, typename C2<T>::type *A2,
                                           const typename C2<T>::type *__restrict__ tw1,
                                           const typename C2<T>::type *__restrict__ tw2) {
  using V = typename C2<T>::type;
  const int K1 = P.K1, K2 = P.K2, H = P.H;
  const int blk = K2 * H;
  for (int t = threadIdx.x; t < blk; t += blockDim.x) A[t] = src[t];
  __syncthreads();
  V *Y = lds_fft(A, A2, H, 1, H, P.ax2, +1, tw2);
  V *Z = (Y == A) ? A2 : A;
  const int npair = (K2 + 1) / 2;
  for (int t = threadIdx.x; t < npair * K1; t += blockDim.x) {
    int pr = t / K1, x = t - pr * K1;
    int y0 = 2 * pr, y1 = y0 + 1;
    V a, c;
    if (x < H) {
      a = Y[y0 * H + x];
      c = (y1 < K2) ? Y[y1 * H + x] : V{0, 0};
    } else {
      V aa = Y[y0 * H + (K1 - x)];
      V cc = (y1 < K2) ? Y[y1 * H + (K1 - x)] : V{0, 0};
      a.x = aa.x;
      a.y = -aa.y;
      c.x = cc.x;
      c.y = -cc.y;
    }
    if (x == 0 || (K1 % 2 == 0 && x == K1 / 2)) {
      a.y = 0;
      c.y = 0;
    }
    V zz;
    zz.x = a.x - c.y;
    zz.y = a.y + c.x;
    Z[pr * K1 + x] = zz;
  }
  __syncthreads();
  V *Zo = (Z == A) ? A2 : A;
  return lds_fft(Z, Zo, npair, K1, 1, P.ax1, +1, tw1);
}

// --------------------------------------------------------------------------
// PME: one z plane per workgroup into phi[b][z][y][x] (used when the fused
// slab kernel below does not fit).
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHR) void k_pme_plane(Params P, const typename C2<T>::type *__restrict__ Hk,
                                                    const typename C2<T>::type *__restrict__ tw1,
                                                    const typename C2<T>::type *__restrict__ tw2,
                                                    T *phi) {
  using V = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int K1 = P.K1, K2 = P.K2, H = P.H;
  V *A = reinterpret_cast<V *>(lds_raw);
  V *A2 = A + K2 * H;
  const int z = blockIdx.x, b = blockIdx.y;
  V *R = plane_c2r<T>(P, Hk + ((size_t)b * P.K3 + z) * K2 * H, A, A2, tw1, tw2);
  T *dst = phi + ((size_t)b * P.K3 + z) * (size_t)K2 * K1;
  for (int t = threadIdx.x; t < K2 * K1; t += blockDim.x) {
    int y = t / K1, x = t - y * K1;
    V v = R[(y >> 1) * K1 + x];
    dst[t] = (y & 1) ? v.y : v.x;
  }
}

// --------------------------------------------------------------------------
// PME, fused inverse transform and gather (recpcal.F90:402 and 451-502).
// Workgroup (slab, b) owns the solvent sites whose grid base g3 lies in
// [zb, zb + S); it computes the planes zb - p + 1 .. zb + S - 1 one at a time
// in LDS (plane_c2r) and, after each plane zz, adds that plane's share
//   q w3[g3 - zz] sum_{c2,c1} w2[c2] w1[c1] phi_zz(g2 - c2, g1 - c1)
// to every owned site whose support contains zz.  Thread t owns the slab
// sites t, t + blockDim, ... (sites are sorted by g3, so a wave's sites are
// active together); the per-site sums live in registers (FP64) and are
// written once to e_site[b][t_site].  phi never leaves the CU.
// --------------------------------------------------------------------------
template <typename T, int MAXR>
__global__ __launch_bounds__(NTHR) void k_pme_slab(Params P, int S, int nsites,
                                                   const typename C2<T>::type *__restrict__ Hk,
                                                   const typename C2<T>::type *__restrict__ tw1,
                                                   const typename C2<T>::type *__restrict__ tw2,
                                                   const int *__restrict__ z_start, const int *__restrict__ zg,
                                                   const T *__restrict__ zq, const T *__restrict__ zw,
                                                   const int *__restrict__ zsite, double *e_site) {
  using V = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int K1 = P.K1, K2 = P.K2, K3 = P.K3, H = P.H, p = P.p;
  V *A = reinterpret_cast<V *>(lds_raw);
  V *A2 = A + K2 * H;
  const int b = blockIdx.y;
  const int zb = blockIdx.x * S, ze = min(zb + S, K3);
  const int s_beg = z_start[zb], s_end = z_start[ze];
  double acc[MAXR];
  int gz[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    int idx = s_beg + r * (int)blockDim.x + (int)threadIdx.x;
    acc[r] = 0.0;
    gz[r] = (idx < s_end) ? ((zg[idx] >> 20) & 1023) : -(1 << 20);
  }
  const size_t np = (size_t)nsites;
  for (int zz = zb - p + 1; zz < ze; ++zz) {
    const int z = ((zz % K3) + K3) % K3;
    V *R = plane_c2r<T>(P, Hk + ((size_t)b * K3 + z) * K2 * H, A, A2, tw1, tw2);
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int c3 = gz[r] - zz;
      if (c3 >= 0 && c3 < p) {
        const int idx = s_beg + r * (int)blockDim.x + (int)threadIdx.x;
        const int gp = zg[idx];
        const int g1 = gp & 1023, g2 = (gp >> 10) & 1023;
        const T *wz = zw + (size_t)idx * 3 * p;  // the site's weights [3][p]
        const T f1 = zq[idx] * wz[2 * p + c3];
        T sum = 0;
        for (int c2 = 0; c2 < p; ++c2) {
          int y = g2 - c2;
          y += (y < 0) ? K2 : 0;
          const V *row = R + (y >> 1) * K1;
          const bool odd = (y & 1);
          T rs = 0;
          for (int c1 = 0; c1 < p; ++c1) {
            int x = g1 - c1;
            x += (x < 0) ? K1 : 0;
            V v = row[x];
            rs += wz[c1] * (odd ? v.y : v.x);
          }
          sum += wz[p + c2] * rs;
        }
        acc[r] += (double)(f1 * sum);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    int idx = s_beg + r * (int)blockDim.x + (int)threadIdx.x;
    if (idx < s_end) e_site[(size_t)b * np + idx] = acc[r];
  }
}

// --------------------------------------------------------------------------
// Register-column PME path.  Three kernels, each thread owning one column in
// registers (compile-time lengths, no LDS in the transforms):
//  k_pme_zcol_reg  column (m1, m2): forward z transform of the solute's
//                  separable spectrum, sum_s q_s what1_s(m1) what2_s(m2)
//                  what3_s(m3) (a direct sum over sites: the r2c transform of
//                  the spread grid without the grid), Green's function and
//                  self energy (recpcal.F90:377-398), inverse z FFT -> Hk
//  k_pme_ycol_reg  column (z, m1): y FFT in place in Hk (inverse by default)
//  k_pme_xslab     slab of z planes: x c2r of row pairs into an LDS ring of
//                  planes, then the gather into per-site FP64 sums
// --------------------------------------------------------------------------
constexpr int ZREG_SCH = 16;  // solute sites per LDS chunk of k_pme_zcol_reg

// dynamic LDS of k_pme_zcol_reg with XT threads: ZREG_SCH sites x
// (H + K2 + K3) spectra + charges + the self-energy reduction
template <typename T> __host__ __device__ inline size_t zcol_reg_lds(int H, int K2, int K3, int XT = NTHR) {
  return (size_t)ZREG_SCH * (H + K2 + K3) * 2 * sizeof(T) + ZREG_SCH * sizeof(T) + XT * sizeof(double);
}

template <typename T, int K3, int XT = NTHR>
__global__ __launch_bounds__(XT) void k_pme_zcol_reg(Params P, const T *__restrict__ engfac,
                                                       const T *__restrict__ sqv,
                                                       const typename C2<T>::type *__restrict__ w1h,
                                                       const typename C2<T>::type *__restrict__ w2h,
                                                       const typename C2<T>::type *__restrict__ w3h,
                                                       const typename C2<T>::type *__restrict__ tw3,
                                                       typename C2<T>::type *Hk, double *selfp) {
  using V = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int b = blockIdx.y, ns = P.ns, H = P.H, K2 = P.K2;
  // staged spectra per site: x and y, and z in FP64 (FP32 reads the z
  // spectrum through the scalar cache)
  const int nsp = H + K2 + (sizeof(T) == 8 ? K3 : 0);
  V *l1 = reinterpret_cast<V *>(lds_raw);  // [ZREG_SCH][H]
  V *l2 = l1 + ZREG_SCH * H;               // [ZREG_SCH][K2]
  V *l3 = l2 + ZREG_SCH * K2;              // [ZREG_SCH][K3]
  T *lq = reinterpret_cast<T *>(l3 + ZREG_SCH * K3);
  double *red = reinterpret_cast<double *>(lds_raw + zcol_reg_lds<T>(H, K2, K3, XT) - XT * sizeof(double));
  const int ncol = H * K2;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = c < ncol;
  const int m1 = act ? c % H : 0, m2 = act ? c / H : 0;
  V X[K3];
#pragma unroll
  for (int m3 = 0; m3 < K3; ++m3) {
    X[m3].x = 0;
    X[m3].y = 0;
  }
  for (int s0 = 0; s0 < ns; s0 += ZREG_SCH) {
    const int nc = min(ZREG_SCH, ns - s0);
    __syncthreads();
    for (int t = threadIdx.x; t < nc * nsp; t += blockDim.x) {
      const int sl = t / nsp, m = t - sl * nsp;
      const size_t bs = (size_t)b * ns + s0 + sl;
      if (m < H)
        l1[sl * H + m] = w1h[bs * H + m];
      else if (m < H + K2)
        l2[sl * K2 + (m - H)] = w2h[bs * K2 + (m - H)];
      else
        l3[sl * K3 + (m - H - K2)] = w3h[bs * K3 + (m - H - K2)];
    }
    for (int t = threadIdx.x; t < nc; t += blockDim.x) lq[t] = sqv[(size_t)b * ns + s0 + t];
    __syncthreads();
    for (int sl = 0; sl < nc; ++sl) {
      V a = cmul(l1[sl * H + m1], l2[sl * K2 + m2]);
      const T q = lq[sl];
      a.x *= q;
      a.y *= q;
      // the site's z spectrum is the same for every lane: FP32 reads it
      // through the scalar cache into SGPRs (an LDS broadcast costs a vector
      // read per four FMAs), FP64 from LDS (its 2 K3 values would not stay
      // in SGPRs)
      const V *w3 = (sizeof(T) == 4) ? w3h + ((size_t)b * ns + s0 + sl) * K3 : l3 + sl * K3;
#pragma unroll
      for (int m3 = 0; m3 < K3; ++m3) {
        const V w = w3[m3];
        X[m3].x = fma(a.x, w.x, X[m3].x);
        X[m3].x = fma(-a.y, w.y, X[m3].x);
        X[m3].y = fma(a.x, w.y, X[m3].y);
        X[m3].y = fma(a.y, w.x, X[m3].y);
      }
    }
  }
  double self = 0.0;
  const double wgt = (m1 == 0 || (P.K1 % 2 == 0 && m1 == H - 1)) ? 0.5 : 1.0;
#pragma unroll
  for (int m3 = 0; m3 < K3; ++m3) {
    const T g = act ? engfac[((size_t)m3 * K2 + m2) * H + m1] : (T)0;
    self += wgt * (double)g * ((double)X[m3].x * (double)X[m3].x + (double)X[m3].y * (double)X[m3].y);
    X[m3].x *= g;
    X[m3].y *= g;
  }
  red[threadIdx.x] = self;
  reg_fft<K3, +1>(X, tw3);
  __syncthreads();
  for (int st = XT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) selfp[(size_t)b * gridDim.x + blockIdx.x] = red[0];
  if (act) {
#pragma unroll
    for (int z = 0; z < K3; ++z) Hk[((size_t)b * K3 + z) * ncol + c] = X[z];
  }
}

// the value of the other lane of an even/odd lane pair (DPP quad_perm
// [1, 0, 3, 2]; every lane of the wave must be active)
__device__ __forceinline__ float pair_swap(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_swap(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// One value of the radix-2 step of a length-2M transform held by a lane
// pair (lane h owns y[n + h M] at index n): s y + y' (y' the other lane's
// value; s = 1 on lane 0, -1 on lane 1), times w on lane 1 (hs = h blends
// w with 1).
template <typename T, typename V> __device__ __forceinline__ void pair_step(V &y, V w, T sg, T hs) {
  V t;
  t.x = fma(sg, y.x, pair_swap(y.x));
  t.y = fma(sg, y.y, pair_swap(y.y));
  const T wr = fma(hs, w.x - (T)1, (T)1), wi = hs * w.y;
  y.x = t.x * wr - t.y * wi;
  y.y = t.x * wi + t.y * wr;
}
// (a scheduling barrier after every fourth value keeps the exchanged copies
// from all being live at once)
template <typename T, int M, int SIGN, typename V, int... J>
__device__ __forceinline__ void pair_steps(V (&Y)[M], T sg, T hs, std::integer_sequence<int, J...>) {
  ((pair_step<T>(Y[J], root_v<V, SIGN, J, 2 * M>(), sg, hs),
    (J % 4 == 3) ? __builtin_amdgcn_sched_barrier(0) : (void)0),
   ...);
}
// every value times w^J (lane 1) or 1 (lane 0), w = e^{SIGN 2 pi i / 2M}
template <typename T, int M, int SIGN, typename V, int... J>
__device__ __forceinline__ void pair_twiddle(V (&Y)[M], T hs, std::integer_sequence<int, J...>) {
  auto one = [&](V &y, V w) {
    const T wr = fma(hs, w.x - (T)1, (T)1), wi = hs * w.y;
    const T yr = y.x * wr - y.y * wi;
    y.y = y.x * wi + y.y * wr;
    y.x = yr;
  };
  (one(Y[J], root_v<V, SIGN, J, 2 * M>()), ...);
}

// Y += a w[k] at the PP planes (G - k) mod K3 of a site at plane G, on
// the lane pair that splits the column (lane h owns planes [h K3/2,
// (h+1) K3/2)); hm[o] is 1 on the lane that owns half o, else 0.  When the
// PP planes lie in one half (all but the buckets next to a half's edge)
// one multiply masks a for every plane; else each weight is masked.
template <int K3, int G, typename T, typename V, int... Kk>
__device__ __forceinline__ void zspread_half(V (&Y)[K3 / 2], const V a, const T (&w)[sizeof...(Kk)],
                                             const T (&hm)[2], std::integer_sequence<int, Kk...>) {
  constexpr int HK = K3 / 2;
  constexpr int h0 = G / HK;
  constexpr bool same = ((((G - Kk + K3) % K3) / HK == h0) && ...);
  if constexpr (same) {
    V am;
    am.x = a.x * hm[h0];
    am.y = a.y * hm[h0];
    auto one = [&](V &y, T wk) {
      y.x = fma(am.x, wk, y.x);
      y.y = fma(am.y, wk, y.y);
    };
    (one(Y[((G - Kk + K3) % K3) % HK], w[Kk]), ...);
  } else {
    auto one = [&](V &y, T hmo, T wk) {
      const T ww = hmo * wk;
      y.x = fma(a.x, ww, y.x);
      y.y = fma(a.y, ww, y.y);
    };
    (one(Y[((G - Kk + K3) % K3) % HK], hm[((G - Kk + K3) % K3) / HK], w[Kk]), ...);
  }
}

// k_pme_zcol_fft stages its sites in chunks of zfft_sch<T>, sorted by z
// plane (128 in FP32: the 150-site solutes take two chunks, and three
// 256-thread workgroups still fit a CU's LDS)
template <typename T> constexpr int zfft_sch = sizeof(T) == 4 ? 128 : 64;
constexpr int ZFFT_KMAX = 96;  // largest z grid length of the register kernels

// lanes per z column in k_pme_zcol_fft: one in FP32 up to K3 = 64 (the
// whole column, 2 K3 floats, in one lane's registers), else a lane pair
__host__ __device__ constexpr int zfft_lanes(int tbytes, int K3) { return (tbytes == 4 && K3 <= 64) ? 1 : 2; }

// m2 rows a workgroup of XT threads (XT / L columns, m1 fastest) touches
__host__ __device__ inline int zcol_fft_rows(int H, int K2, int XT, int L = 2) {
  const int r = (XT / L - 1) / H + 2;
  return r < K2 ? r : K2;
}

// dynamic LDS of k_pme_zcol_fft with XT threads and L lanes per column:
// zfft_sch<T> sites x (H + the block's m2 rows) spectra, charge-scaled z
// weights, planes and ranks, the bucket starts, and the self-energy
// reduction
template <typename T>
__host__ __device__ inline size_t zcol_fft_lds(int H, int K2, int XT = NTHR, int L = 2) {
  const size_t spec = (size_t)zfft_sch<T> * (H + zcol_fft_rows(H, K2, XT, L)) * 2 * sizeof(T);
  const size_t small = (size_t)zfft_sch<T> * (sizeof(T) * MAXP + 2 * sizeof(int)) + (ZFFT_KMAX + 1) * sizeof(int);
  return ((spec + small + 15) & ~(size_t)15) + XT * sizeof(double);
}

// Site chunk of the z-column kernels in LDS: ranks by (plane, index), the
// bucket starts per plane, the sorted spectra what1 (all m1) and what2 (the
// block's nm2 rows from m2lo) and the charge-scaled z weights.  All threads
// call it; the caller's barrier precedes the next chunk's call.
template <typename T, int PP>
__device__ __forceinline__ void zcol_stage_chunk(int s0, int nc, int K3, int H, int K2, int m2lo, int nm2,
                                                 const T *__restrict__ sqv,
                                                 const typename C2<T>::type *__restrict__ w1h,
                                                 const typename C2<T>::type *__restrict__ w2h,
                                                 const int *__restrict__ sg3, const T *__restrict__ sw3,
                                                 typename C2<T>::type *l1, typename C2<T>::type *l2, T *lw,
                                                 int *lg, int *lr, int *lst) {
  __syncthreads();  // the previous chunk is consumed
  for (int t = threadIdx.x; t < nc; t += blockDim.x) lg[t] = sg3[(size_t)s0 + t];
  __syncthreads();
  // rank by (plane, index); bucket starts = sites on lower planes (a plane
  // outside [0, K3) falls in no bucket)
  for (int t = threadIdx.x; t < nc; t += blockDim.x) {
    const int g = lg[t];
    int r = 0;
    for (int u = 0; u < nc; ++u) {
      const int gu = lg[u];
      r += (gu < g) || (gu == g && u < t);
    }
    lr[t] = r;
  }
  for (int t = threadIdx.x; t <= K3; t += blockDim.x) {
    int r = 0;
    for (int u = 0; u < nc; ++u) r += lg[u] < t;
    lst[t] = r;
  }
  __syncthreads();
  const int nsp = H + nm2;
  for (int t = threadIdx.x; t < nc * nsp; t += blockDim.x) {
    const int sl = t / nsp, m = t - sl * nsp;
    const size_t bs = (size_t)s0 + sl;
    const int r = lr[sl];
    if (m < H)
      l1[r * H + m] = w1h[bs * H + m];
    else
      l2[r * nm2 + (m - H)] = w2h[bs * K2 + m2lo + (m - H)];
  }
  for (int t = threadIdx.x; t < nc * PP; t += blockDim.x) {
    const int sl = t / PP, k = t - sl * PP;
    lw[lr[sl] * PP + k] = sqv[(size_t)s0 + sl] * sw3[(size_t)s0 * PP + t];
  }
  __syncthreads();
}

// One lane per column (m1, m2) (FP32, K3 <= 64): the same spread, Green's
// function and transforms as the lane-pair form below, with the whole column
// in one lane's registers in plane order.  Every site adds to its PP planes
// directly (no lane masks), the forward and inverse z FFTs are one K3-point
// register FFT each, and a wave's stores of one plane are one contiguous
// run of 64 columns.
template <typename T, int K3, int PP, int XT>
__device__ __forceinline__ void zcol_fft_body1(const Params &P, int ns, int blk, const T *__restrict__ engfac,
                                               const T *__restrict__ sqv,
                                               const typename C2<T>::type *__restrict__ w1h,
                                               const typename C2<T>::type *__restrict__ w2h,
                                               const int *__restrict__ sg3, const T *__restrict__ sw3,
                                               typename C2<T>::type *Hk, double *self_out) {
  using V = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int H = P.H, K2 = P.K2;
  const int rows = zcol_fft_rows(H, K2, XT, 1);
  V *l1 = reinterpret_cast<V *>(lds_raw);
  V *l2 = l1 + zfft_sch<T> * H;
  T *lw = reinterpret_cast<T *>(l2 + zfft_sch<T> * rows);
  int *lg = reinterpret_cast<int *>(lw + zfft_sch<T> * MAXP);
  int *lr = lg + zfft_sch<T>;
  int *lst = lr + zfft_sch<T>;
  double *red = reinterpret_cast<double *>(lds_raw + zcol_fft_lds<T>(H, K2, XT, 1) - XT * sizeof(double));
  const int ncol = H * K2;
  const int c = blk * XT + (int)threadIdx.x;
  const bool act = c < ncol;
  const int m1 = act ? c % H : 0;
  const int m2lo = (blk * XT) / H, m2hi = (min(ncol, (blk + 1) * XT) - 1) / H;
  const int nm2 = m2hi - m2lo + 1;
  const int m2 = act ? c / H - m2lo : 0;
  V Y[K3];
#pragma unroll
  for (int z = 0; z < K3; ++z) {
    Y[z].x = 0;
    Y[z].y = 0;
  }
  const int lane = threadIdx.x & 63;
  int lsa = 0, lsb = 0;
  auto bucket = [&](auto gc) {
    constexpr int G = decltype(gc)::value;
    const int b0 = __builtin_amdgcn_readlane(G < 64 ? lsa : lsb, G & 63);
    const int b1 = __builtin_amdgcn_readlane(G + 1 < 64 ? lsa : lsb, (G + 1) & 63);
    for (int i = b0; i < b1; ++i) {
      const V a = cmul(l1[i * H + m1], l2[i * nm2 + m2]);
      T w[PP];
#pragma unroll
      for (int k = 0; k < PP; ++k) w[k] = lw[i * PP + k];  // one broadcast read
#pragma unroll
      for (int k = 0; k < PP; ++k) {
        V &y = Y[(G - k + K3) % K3];
        y.x = fma(a.x, w[k], y.x);
        y.y = fma(a.y, w[k], y.y);
      }
    }
  };
  for (int s0 = 0; s0 < ns; s0 += zfft_sch<T>) {
    const int nc = min(zfft_sch<T>, ns - s0);
    zcol_stage_chunk<T, PP>(s0, nc, K3, H, K2, m2lo, nm2, sqv, w1h, w2h, sg3, sw3, l1, l2, lw, lg, lr, lst);
    lsa = lst[lane];
    if constexpr (K3 >= 64) lsb = lst[64 + lane < ZFFT_KMAX + 1 ? 64 + lane : ZFFT_KMAX];
    [&]<int... Gs>(std::integer_sequence<int, Gs...>) {
      (bucket(std::integral_constant<int, Gs>{}), ...);
    }(std::make_integer_sequence<int, K3>{});
  }
  reg_fft<K3, -1>(Y);
  double self = 0.0;
  const double wgt = (m1 == 0 || (P.K1 % 2 == 0 && m1 == H - 1)) ? 0.5 : 1.0;
  // engfac[m3][m2][m1]: a uniform row per m3 plus the lane's offset c
  const unsigned goff = act ? (unsigned)c : 0u;
  const T gmask = act ? (T)1 : (T)0;
#pragma unroll
  for (int m = 0; m < K3; ++m) {
    const T g = gmask * engfac[(size_t)m * ncol + goff];
    self += wgt * (double)g * ((double)Y[m].x * (double)Y[m].x + (double)Y[m].y * (double)Y[m].y);
    Y[m].x *= g;
    Y[m].y *= g;
  }
  red[threadIdx.x] = self;
  reg_fft<K3, +1>(Y);
  __syncthreads();
  for (int st = XT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *self_out = red[0];
  if (act) {
#pragma unroll
    for (int n = 0; n < K3; ++n) Hk[(size_t)n * ncol + c] = Y[n];
  }
}

// Column (m1, m2) of the same product as k_pme_zcol_reg, with the z axis
// spread and transformed instead of summed spectrally, on a lane pair that
// holds half of the column each:
//  - every site adds q_s what1_s(m1) what2_s(m2) w3_s,k at its PP planes
//    (g3 - k) mod K3 (the reference's spread along z, recpcal.F90:451-470).
//    The sites of a chunk are sorted by plane g3 in LDS; one loop per plane
//    (unrolled over the K3 planes, each over its own sites) updates the
//    PP registers of that plane at compile-time indices, so the column
//    stays in place in registers;
//  - forward z FFT (the r2c transform's last axis): one radix-2 step across
//    the pair, then a K3/2 FFT per lane (lane h: frequencies 2m + h);
//  - Green's function and self energy (recpcal.F90:377-398);
//  - inverse z FFT: a K3/2 FFT per lane, then the radix-2 step across the
//    pair (lane h: planes h K3/2 + n) -> Hk.
// Column block blk of one configuration (ns sites; the site arrays start at
// the configuration's first site, Hk at its spectrum); the block's
// self-energy sum goes to *self_out.  All threads of the workgroup call it.
template <typename T, int K3, int PP, int XT>
__device__ __forceinline__ void zcol_fft_body(const Params &P, int ns, int blk, const T *__restrict__ engfac,
                                              const T *__restrict__ sqv,
                                              const typename C2<T>::type *__restrict__ w1h,
                                              const typename C2<T>::type *__restrict__ w2h,
                                              const int *__restrict__ sg3, const T *__restrict__ sw3,
                                              typename C2<T>::type *Hk, double *self_out) {
  static_assert(K3 <= ZFFT_KMAX, "z grid length");
  if constexpr (zfft_lanes(sizeof(T), K3) == 1) {
    zcol_fft_body1<T, K3, PP, XT>(P, ns, blk, engfac, sqv, w1h, w2h, sg3, sw3, Hk, self_out);
    return;
  }
  using V = typename C2<T>::type;
  constexpr int HK = K3 / 2;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int H = P.H, K2 = P.K2;
  const int rows = zcol_fft_rows(H, K2, XT);
  V *l1 = reinterpret_cast<V *>(lds_raw);           // [zfft_sch<T>][H], sorted order
  V *l2 = l1 + zfft_sch<T> * H;                        // [zfft_sch<T>][nm2] the block's m2 rows
  T *lw = reinterpret_cast<T *>(l2 + zfft_sch<T> * rows);  // [zfft_sch<T>][PP] z weights x charge
  int *lg = reinterpret_cast<int *>(lw + zfft_sch<T> * MAXP);  // [zfft_sch<T>] planes, chunk order
  int *lr = lg + zfft_sch<T>;                          // [zfft_sch<T>] rank of each site by plane
  int *lst = lr + zfft_sch<T>;                         // [K3 + 1] first sorted site of each plane
  double *red = reinterpret_cast<double *>(lds_raw + zcol_fft_lds<T>(H, K2, XT) - XT * sizeof(double));
  const int ncol = H * K2;
  const int h = threadIdx.x & 1;
  const int c = (blk * XT + (int)threadIdx.x) >> 1;
  const bool act = c < ncol;
  const int m1 = act ? c % H : 0;
  const int m2lo = (blk * (XT / 2)) / H, m2hi = (min(ncol, (blk + 1) * (XT / 2)) - 1) / H;
  const int nm2 = m2hi - m2lo + 1;
  const int m2 = act ? c / H - m2lo : 0;
  const T sg = h ? (T)-1 : (T)1, hs = (T)h;
  const T hm[2] = {(T)(1 - h), (T)h};
  V Y[HK];
#pragma unroll
  for (int z = 0; z < HK; ++z) {
    Y[z].x = 0;
    Y[z].y = 0;
  }
  // the bucket starts of the current chunk, one per lane: lane l holds
  // lst[l] in lsa and lst[64 + l] in lsb (lst spans ZFFT_KMAX + 1 words)
  const int lane = threadIdx.x & 63;
  int lsa = 0, lsb = 0;
  // the sites of plane G (bucket [lst[G], lst[G + 1]) of the sorted chunk)
  // add to the registers of planes G, G - 1, ..., G - PP + 1; the bounds come
  // from the lanes' copies (no LDS round trip per plane)
  auto bucket = [&](auto gc) {
    constexpr int G = decltype(gc)::value;
    const int b0 = __builtin_amdgcn_readlane(G < 64 ? lsa : lsb, G & 63);
    const int b1 = __builtin_amdgcn_readlane(G + 1 < 64 ? lsa : lsb, (G + 1) & 63);
    for (int i = b0; i < b1; ++i) {
      const V a = cmul(l1[i * H + m1], l2[i * nm2 + m2]);
      T w[PP];
#pragma unroll
      for (int k = 0; k < PP; ++k) w[k] = lw[i * PP + k];  // one broadcast read
      zspread_half<K3, G, T>(Y, a, w, hm, std::make_integer_sequence<int, PP>{});
    }
  };
  for (int s0 = 0; s0 < ns; s0 += zfft_sch<T>) {
    const int nc = min(zfft_sch<T>, ns - s0);
    zcol_stage_chunk<T, PP>(s0, nc, K3, H, K2, m2lo, nm2, sqv, w1h, w2h, sg3, sw3, l1, l2, lw, lg, lr, lst);
    lsa = lst[lane];
    if constexpr (K3 >= 64) lsb = lst[64 + lane < ZFFT_KMAX + 1 ? 64 + lane : ZFFT_KMAX];
    [&]<int... Gs>(std::integer_sequence<int, Gs...>) {
      (bucket(std::integral_constant<int, Gs>{}), ...);
    }(std::make_integer_sequence<int, K3>{});
  }
  // forward: radix-2 step across the pair, K3/2 FFT per lane
  pair_steps<T, HK, -1>(Y, sg, hs, std::make_integer_sequence<int, HK>{});
  reg_fft<HK, -1>(Y);
  double self = 0.0;
  const double wgt = (m1 == 0 || (P.K1 % 2 == 0 && m1 == H - 1)) ? 0.5 : 1.0;
  // engfac[m3][m2][m1] at m3 = 2m + h: a uniform row per m plus the lane's
  // 32-bit offset (scalar base + vector offset addressing)
  const unsigned goff = act ? (unsigned)(h * ncol + c) : 0u;
  const T gmask = act ? (T)1 : (T)0;
#pragma unroll
  for (int m = 0; m < HK; ++m) {
    const T *grow = engfac + (size_t)(2 * m) * ncol;
    const T g = gmask * grow[goff];
    self += wgt * (double)g * ((double)Y[m].x * (double)Y[m].x + (double)Y[m].y * (double)Y[m].y);
    Y[m].x *= g;
    Y[m].y *= g;
  }
  red[threadIdx.x] = self;
  // inverse: K3/2 FFT per lane, lane 1 times e^{+2 pi i n / K3}, radix-2
  // step across the pair
  reg_fft<HK, +1>(Y);
  pair_twiddle<T, HK, +1>(Y, hs, std::make_integer_sequence<int, HK>{});
#pragma unroll
  for (int n = 0; n < HK; ++n) {
    Y[n].x = fma(sg, Y[n].x, pair_swap(Y[n].x));
    Y[n].y = fma(sg, Y[n].y, pair_swap(Y[n].y));
    if (n % 4 == 3) __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  for (int st = XT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *self_out = red[0];
  if (act) {
    const unsigned hoff = (unsigned)(h * HK * ncol + c);
#pragma unroll
    for (int n = 0; n < HK; ++n) {
      V *hrow = Hk + (size_t)n * ncol;
      hrow[hoff] = Y[n];
    }
  }
}

template <typename T, int K3, int PP, int XT = NTHR>
__global__ __launch_bounds__(XT) void k_pme_zcol_fft(Params P, const T *__restrict__ engfac,
                                                       const T *__restrict__ sqv,
                                                       const typename C2<T>::type *__restrict__ w1h,
                                                       const typename C2<T>::type *__restrict__ w2h,
                                                       const int *__restrict__ sg3, const T *__restrict__ sw3,
                                                       typename C2<T>::type *Hk, double *selfp) {
  const int b = blockIdx.y, ns = P.ns;
  const size_t so = (size_t)b * ns, ncol = (size_t)P.H * P.K2;
  zcol_fft_body<T, K3, PP, XT>(P, ns, blockIdx.x, engfac, sqv + so, w1h + so * P.H, w2h + so * P.K2, sg3 + so,
                               sw3 + so * PP, Hk + (size_t)b * K3 * ncol, selfp + (size_t)b * gridDim.x + blockIdx.x);
}

// One run of configurations of one solute type within a batch of several
// (the heterogeneous insertion batch): configurations [cfg0, cfg0 + ncfg)
// of the batch, ns sites each from site site0 of the batch's site arrays;
// the work items of the persistent kernels the run owns start at zitem0
// (z columns: ncfg x column blocks) and pitem0 (pairs: ncfg x nsplit); the
// type's charges and LJ rows in T and in FP64; member: the host's index of
// the type.
struct GRun {
  int cfg0, ncfg, ns, site0;
  int zitem0, pitem0, nsplit, member;
  const void *q_t, *lj_t;
  const double *q64, *lj64;
};

// Work items of the persistent kernels: a workgroup takes ITEM_CHUNK
// consecutive items per atomic (the same for every thread: the barrier
// after the read keeps the slot until all have it) until the counter
// passes nitem.
constexpr int ITEM_CHUNK = 4;
__device__ __forceinline__ int next_items(int *counter) {
  __shared__ int item_sh;
  __syncthreads();
  if (threadIdx.x == 0) item_sh = atomicAdd(counter, ITEM_CHUNK);
  __syncthreads();
  return item_sh;
}

// z columns of a heterogeneous batch as one persistent launch: workgroups
// pull (configuration, column block) items, the runs in the order given
// (largest solutes first), until nitem are taken
template <typename T, int K3, int PP, int XT = NTHR>
__global__ __launch_bounds__(XT) void k_pme_zcol_fft_runs(Params P, const T *__restrict__ engfac,
                                                            const T *__restrict__ sqv,
                                                            const typename C2<T>::type *__restrict__ w1h,
                                                            const typename C2<T>::type *__restrict__ w2h,
                                                            const int *__restrict__ sg3, const T *__restrict__ sw3,
                                                            typename C2<T>::type *Hk, double *selfp,
                                                            const GRun *__restrict__ runs, int nrun, int nblk,
                                                            int nitem, int *counter) {
  const size_t ncol = (size_t)P.H * P.K2;
  for (int base = next_items(counter); base < nitem; base = next_items(counter))
  for (int it = base; it < min(nitem, base + ITEM_CHUNK); ++it) {
    int r = 0;
    while (r + 1 < nrun && runs[r + 1].zitem0 <= it) ++r;
    const int ns = runs[r].ns, rel = it - runs[r].zitem0;
    const int bl = rel / nblk, blk = rel - bl * nblk, b = runs[r].cfg0 + bl;
    const size_t so = (size_t)runs[r].site0 + (size_t)bl * ns;
    zcol_fft_body<T, K3, PP, XT>(P, ns, blk, engfac, sqv + so, w1h + so * P.H, w2h + so * P.K2, sg3 + so,
                                 sw3 + so * PP, Hk + (size_t)b * K3 * ncol, selfp + (size_t)b * nblk + blk);
  }
}

// Inverse y transform of column (z, m1) in place on a lane pair: lane h
// loads m2 in [h K2/2, (h+1) K2/2), the radix-2 step across the pair (DPP)
// and a K2/2 FFT per lane leave y = 2k + h in lane h (half the registers of
// one lane per column, so more waves hide the loads)
template <typename T, int K2>
__global__ __launch_bounds__(NTHR) void k_pme_ycol_pair(Params P, typename C2<T>::type *Hk) {
  using V = typename C2<T>::type;
  constexpr int HK = K2 / 2;
  const int b = blockIdx.y, H = P.H, K3 = P.K3;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = t & 1, c = t >> 1;  // c = (z, m1)
  const bool act = c < K3 * H;
  const int cc = act ? c : 0;  // idle lanes of the last workgroup read column 0 and store nothing
  const int z = cc / H, m1 = cc - z * H;
  V *col = Hk + ((size_t)b * K3 + z) * K2 * H + m1;
  V Y[HK];
#pragma unroll
  for (int j = 0; j < HK; ++j) Y[j] = col[(size_t)(h * HK + j) * H];
  const T sg = h ? (T)-1 : (T)1, hs = (T)h;
  pair_steps<T, HK, +1>(Y, sg, hs, std::make_integer_sequence<int, HK>{});
  reg_fft<HK, +1>(Y);
  if (act) {
#pragma unroll
    for (int k = 0; k < HK; ++k) col[(size_t)(2 * k + h) * H] = Y[k];
  }
}

template <typename T, int K2, int SIGN = +1>
__global__ __launch_bounds__(NTHR) void k_pme_ycol_reg(Params P, const typename C2<T>::type *__restrict__ tw2,
                                                       typename C2<T>::type *Hk) {
  using V = typename C2<T>::type;
  const int b = blockIdx.y, H = P.H, K3 = P.K3;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;  // (z, m1)
  if (c >= K3 * H) return;
  const int z = c / H, m1 = c - z * H;
  V *col = Hk + ((size_t)b * K3 + z) * K2 * H + m1;
  V X[K2];
#pragma unroll
  for (int m2 = 0; m2 < K2; ++m2) X[m2] = col[(size_t)m2 * H];
  reg_fft<K2, SIGN>(X, tw2);
#pragma unroll
  for (int y = 0; y < K2; ++y) col[(size_t)y * H] = X[y];
}

// Slab of z planes [zb, zb + S) of configuration b.  The planes a site
// touches, [g3 - p + 1, g3], are processed Q at a time: the Q complex planes
// Hk[b][z][y][m1] are copied into LDS with coalesced loads; thread (q, y)
// holds row y of plane q in registers and inverts it through a half-length
// complex FFT (c2r_half), then writes the real row in place over its own
// complex row as (x[2m], x[2m+1]) pairs with a row stride of K1 + 2 values
// (consecutive lanes on distinct LDS banks; each real row fills exactly its
// complex row's bytes), and repeats x[0], x[1] in the row's two spare values
// K1, K1 + 1.  Each thread then adds the planes' contributions to its MAXR
// owned sites (FP64 sums).
// gather of one group of nq real planes (first plane zz0, rows of K1 + 2)
// into the per-site sums acc[] of the sites a thread owns.  PP: the spline
// order at compile time (weights, row offsets and the p^3 products in
// registers, loads independent); PP = 0 takes it from P.p.
template <typename T, int K1, int MAXR, int PP, int XT>
__device__ __forceinline__ void xslab_gather(const Params &P, const T *real, int zz0, int nq, int s_beg,
                                             const int *gz, const int *__restrict__ zg, const T *__restrict__ zq,
                                             const T *__restrict__ zw, double *acc) {
  constexpr int K1P = K1 + 2;
  const int K2 = P.K2, p = PP ? PP : P.p;
  const size_t plane = (size_t)K2 * (K1 / 2 + 1) * 2;  // reals per plane
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    const int g3 = gz[r];
    const int clo = max(0, g3 - (zz0 + nq - 1)), chi = min(p - 1, g3 - zz0);
    if (clo > chi) continue;
    int idx = s_beg + r * XT + (int)threadIdx.x;
    // formed here, not hoisted out of the caller's plane-group loop (which
    // would hold the load addresses of every site in registers across the
    // transforms)
    asm volatile("" : "+v"(idx));
    const int gp = zg[idx];
    const int g1 = gp & 1023, g2 = (gp >> 10) & 1023;
    const T qv = zq[idx];
    const T *wz = zw + (size_t)idx * 3 * p;  // the site's weights [3][p]
    if constexpr (PP > 0) {
      T w1[PP], w2[PP];
      int xo[PP], yo[PP];
#pragma unroll
      for (int k = 0; k < PP; ++k) {
        w1[k] = wz[k];
        w2[k] = wz[PP + k];
        int x = g1 - k, y = g2 - k;
        xo[k] = x + ((x < 0) ? K1 : 0);
        yo[k] = (y + ((y < 0) ? K2 : 0)) * K1P;
      }
      for (int c3 = clo; c3 <= chi; ++c3) {
        const T *ph = real + (size_t)(g3 - c3 - zz0) * plane;
        T sum = 0;
#pragma unroll
        for (int c2 = 0; c2 < PP; ++c2) {
          const T *row = ph + yo[c2];
          T rs = 0;
#pragma unroll
          for (int c1 = 0; c1 < PP; ++c1) rs += w1[c1] * row[xo[c1]];
          sum += w2[c2] * rs;
        }
        acc[r] += (double)(qv * wz[2 * PP + c3] * sum);
      }
    } else {
      T w1[MAXP], w2[MAXP];
      for (int k = 0; k < p; ++k) {
        w1[k] = wz[k];
        w2[k] = wz[p + k];
      }
      for (int c3 = clo; c3 <= chi; ++c3) {
        const T *ph = real + (size_t)(g3 - c3 - zz0) * plane;
        T sum = 0;
        for (int c2 = 0; c2 < p; ++c2) {
          int y = g2 - c2;
          y += (y < 0) ? K2 : 0;
          const T *row = ph + (size_t)y * K1P;
          T rs = 0;
          for (int c1 = 0; c1 < p; ++c1) {
            int xx = g1 - c1;
            xx += (xx < 0) ? K1 : 0;
            rs += w1[c1] * row[xx];
          }
          sum += w2[c2] * rs;
        }
        acc[r] += (double)(qv * wz[2 * p + c3] * sum);
      }
    }
  }
}

// The gather for spline order 4.  The weights of the owned sites that touch
// the group are loaded four sites at a time (each site's twelve weights are
// one 16-byte-aligned record: three 16-byte loads in FP32), so their loads
// are in flight together (one wait per four sites rather than one per
// site).  The four x values g1 - 3 .. g1 of a row are two pairs of adjacent
// values, (xa, xa + 1) and (xb, xb + 1), each one two-value LDS read: with
// x[0], x[1] repeated at K1, K1 + 1 every wrap of the stencil over the row's
// end is such a pair (g1 = 0: (K1-3, K1-2), (K1-1, K1); g1 = 1: (K1-2, K1-1),
// (0, 1); g1 = 2: (K1-1, K1), (1, 2)).  Products and sums are formed in the
// order of xslab_gather, so both give the same values.
template <typename T, int K1, int MAXR, int XT>
__device__ __forceinline__ void xslab_gather4(const Params &P, const T *real, int zz0, int nq, int s_beg,
                                              size_t np, const int *gz, const int *__restrict__ zg,
                                              const T *__restrict__ zq, const T *__restrict__ zw, double *acc) {
  constexpr int K1P = K1 + 2;
  const int K2 = P.K2;
  const ptrdiff_t plane = (ptrdiff_t)K2 * (K1 / 2 + 1) * 2;  // reals per plane
  constexpr int RB = MAXR < 4 ? MAXR : 4;  // sites whose loads fly together
#pragma unroll
  for (int r0 = 0; r0 < MAXR; r0 += RB) {
    int gp[RB];
    T qv[RB], w[RB][12];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      // site r meets the group's planes [zz0, zz0 + nq) when
      // g3 - 3 <= zz0 + nq - 1 and g3 >= zz0 (none for the padding, g3 < 0)
      const int g3 = gz[r0 + i];
      const bool need = g3 >= zz0 && g3 <= zz0 + nq + 2;
      gp[i] = 0;
      qv[i] = 0;
#pragma unroll
      for (int k = 0; k < 12; ++k) w[i][k] = 0;
      if (__builtin_amdgcn_ballot_w64(need)) {  // uniform: some lane's site r
        int idx = min(s_beg + (r0 + i) * XT + (int)threadIdx.x, (int)np - 1);
        asm volatile("" : "+v"(idx));  // not hoisted out of the plane-group loop
        gp[i] = zg[idx];
        qv[i] = zq[idx];
        const T *wz = static_cast<const T *>(__builtin_assume_aligned(zw + (size_t)idx * 12, 16));
#pragma unroll
        for (int k = 0; k < 12; ++k) w[i][k] = wz[k];
      }
    }
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int r = r0 + i;
      const int g3 = gz[r];
      const int clo = max(0, g3 - (zz0 + nq - 1)), chi = min(3, g3 - zz0);
      if (clo > chi) continue;
      const int g1 = gp[i] & 1023, g2 = (gp[i] >> 10) & 1023;
      const int xa = g1 - 3 + ((g1 < 3) ? K1 : 0);
      const int xb = g1 - 1 + ((g1 < 1) ? K1 : 0);
      int ra[4], rb[4];  // the row's two pairs, offsets within a plane
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int y = g2 - k;
        const int ro = (y + ((y < 0) ? K2 : 0)) * K1P;
        ra[k] = ro + xa;
        rb[k] = ro + xb;
      }
      const T *ph0 = real + (ptrdiff_t)(g3 - zz0) * plane;  // plane of c3 = 0
#pragma unroll
      for (int c3 = 0; c3 < 4; ++c3) {
        if (c3 < clo || c3 > chi) continue;
        const T *ph = ph0 - (ptrdiff_t)c3 * plane;
        T sum = 0;
#pragma unroll
        for (int c2 = 0; c2 < 4; ++c2) {
          const T *a = ph + ra[c2], *bb = ph + rb[c2];
          const T a0 = a[0], a1 = a[1], b0 = bb[0], b1 = bb[1];
          T rs = 0;
          rs += w[i][0] * b1;
          rs += w[i][1] * b0;
          rs += w[i][2] * a1;
          rs += w[i][3] * a0;
          sum += w[i][4 + c2] * rs;
        }
        acc[r] += (double)(qv[i] * w[i][8 + c3] * sum);
      }
    }
  }
}

// Planes [zz0, zz0 + nq) (periodic in K3) of configuration b's spectrum
// Hk[b][z][m2][m1] into LDS and transformed there: (with YF the inverse y
// transform per (plane, m1) column, then) the inverse x transform of every
// row, each real row written over its own complex row with x[0], x[1]
// repeated at K1, K1 + 1.  Returns the real planes [nq][K2][K1 + 2]; every
// thread of the workgroup takes part, and the last barrier has been passed.
template <typename T, int K1, bool YF>
__device__ __forceinline__ T *xslab_stage(const typename C2<T>::type *__restrict__ Hk,
                                          typename C2<T>::type *stage, int b, int zz0, int nq, int K2,
                                          int K3) {
  using V = typename C2<T>::type;
  constexpr int H = K1 / 2 + 1;
  constexpr int K1P = K1 + 2;
  const int cplane = K2 * H;
  // the nq complex planes straight into LDS: 16-byte global-to-LDS loads,
  // one 1 KiB run per wave instruction, all in flight until the wait
  {
    const int ln = threadIdx.x & 63, nch = cplane * (int)sizeof(V) / 16;
    for (int q = 0; q < nq; ++q) {
      const int z = (((zz0 + q) % K3) + K3) % K3;
      const uint4 *src = reinterpret_cast<const uint4 *>(Hk + ((size_t)b * K3 + z) * cplane);
      uint4 *dst = reinterpret_cast<uint4 *>(stage + (size_t)q * cplane);
      for (int e0 = (int)(threadIdx.x & ~63u); e0 < nch; e0 += blockDim.x)
        if (e0 + ln < nch)
          __builtin_amdgcn_global_load_lds(src + e0 + ln, (__attribute__((address_space(3))) void *)(dst + e0), 16,
                                           0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
  if constexpr (YF) {
    // inverse y transform of column (q, m1), in place (consecutive lanes
    // take consecutive m1: distinct banks)
    for (int t = threadIdx.x; t < nq * H; t += blockDim.x) {
      const int q = t / H, m1 = t - q * H;
      V *col = stage + (size_t)q * cplane + m1;
      V X[K1];
#pragma unroll
      for (int m2 = 0; m2 < K1; ++m2) X[m2] = col[(size_t)m2 * H];
      reg_fft<K1, +1>(X);
#pragma unroll
      for (int y = 0; y < K1; ++y) col[(size_t)y * H] = X[y];
    }
    __syncthreads();
  }
  // the real row (q, y) occupies exactly the bytes of the complex row
  // (q, y) (K1P = 2H), which only this thread reads: no barrier between
  // the transform and its write, any number of rows per thread
  T *real = reinterpret_cast<T *>(stage);
  for (int t = threadIdx.x; t < nq * K2; t += blockDim.x) {
    const int q = t / K2, y = t - q * K2;
    V Z[K1 / 2];
    c2r_half<K1>(stage + (size_t)q * cplane + (size_t)y * H, Z);
    V *d = reinterpret_cast<V *>(real + (size_t)q * cplane * 2 + (size_t)y * K1P);
#pragma unroll
    for (int m = 0; m < K1 / 2; ++m) d[m] = Z[m];
    d[K1 / 2] = Z[0];  // x[0], x[1] again at K1, K1 + 1 (the gather's wrap)
  }
  __syncthreads();
  return real;
}

// YF: the planes arrive before the inverse y transform, which is done here
// per (plane, m1) column in LDS (K2 = K1), so k_pme_ycol_reg is skipped
template <typename T, int K1, int MAXR, int XT = NTHR, bool YF = false>
__global__ __launch_bounds__(XT) void k_pme_xslab(Params P, int S, int Q, int nsites,
                                                    const typename C2<T>::type *__restrict__ Hk,
                                                    const typename C2<T>::type *__restrict__ tw1,
                                                    const int *__restrict__ z_start, const int *__restrict__ zg,
                                                    const T *__restrict__ zq, const T *__restrict__ zw,
                                                    const int *__restrict__ zsite, double *e_site) {
  using V = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int K2 = P.K2, K3 = P.K3, p = P.p;
  V *stage = reinterpret_cast<V *>(lds_raw);  // [Q][K2][H] complex, then [Q][K2][K1 + 2] real in place
  // XCD-aware order: workgroups go to the 8 XCDs round-robin by linear
  // index, so XCD x is given the x-th contiguous share of the (configuration,
  // slab) items; consecutive slabs of a configuration then run on one XCD
  // and find their p - 1 shared halo planes in its L2
  const int nwg = gridDim.x * gridDim.y;
  int L = blockIdx.y * gridDim.x + blockIdx.x;
  if (nwg % 8 == 0) L = (L % 8) * (nwg / 8) + L / 8;
  const int b = L / gridDim.x;
  const int zb = (L - b * gridDim.x) * S, ze = min(zb + S, K3);
  const int s_beg = z_start[zb], s_end = z_start[ze];
  double acc[MAXR];
  int gz[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    int idx = s_beg + r * (int)blockDim.x + (int)threadIdx.x;
    acc[r] = 0.0;
    gz[r] = (idx < s_end) ? ((zg[idx] >> 20) & 1023) : -(1 << 20);
  }
  const size_t np = (size_t)nsites;
  for (int zz0 = zb - p + 1; zz0 < ze; zz0 += Q) {
    const int nq = min(Q, ze - zz0);
    T *real = xslab_stage<T, K1, YF>(Hk, stage, b, zz0, nq, K2, K3);
    // gather: owned sites whose support g3 - c3 meets planes [zz0, zz0 + nq);
    // the common spline orders get fully unrolled register code
    if (p == 4)
      xslab_gather4<T, K1, MAXR, XT>(P, real, zz0, nq, s_beg, np, gz, zg, zq, zw, acc);
    else if (p == 6)
      xslab_gather<T, K1, MAXR, 6, XT>(P, real, zz0, nq, s_beg, gz, zg, zq, zw, acc);
    else
      xslab_gather<T, K1, MAXR, 0, XT>(P, real, zz0, nq, s_beg, gz, zg, zq, zw, acc);
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    int idx = s_beg + r * (int)blockDim.x + (int)threadIdx.x;
    if (idx < s_end) e_site[(size_t)b * np + idx] = acc[r];
  }
}

// per-molecule reciprocal energy from the per-site sums (stored in the
// z-sorted site order; zpos maps a cell-sorted site to it), plus the
// neutralising background (residual_ene, engproc.F90:859-869):
// uv[b][1 + j] = -pi Q_s Q_j / (alpha^2 V) + sum over the sites of j, in
// site order.  A thread takes molecule j for MOLSUM_CFG consecutive
// configurations and its sites four at a time: the four zpos reads, then
// the per-site sums of every configuration, all in flight together (one
// molecule per thread and configuration was a chain of dependent loads).
constexpr int MOLSUM_CFG = 4;
static __global__ void k_mol_sum(Params P, int nsites, int ncfg, const double *__restrict__ e_site,
                                 const int *__restrict__ zpos, const int *__restrict__ mol_first,
                                 const int *__restrict__ mol_ns, const double *__restrict__ mol_q,
                                 const int *__restrict__ slt_j, double *uv) {
  // XCD-aware order (workgroups go to the 8 XCDs round-robin by linear
  // index): the workgroups of a configuration group run on one XCD, whose
  // L2 then holds the group's per-site sums for all of them
  const int nwg = gridDim.x * gridDim.y;
  int L = blockIdx.y * gridDim.x + blockIdx.x;
  if (nwg % 8 == 0) L = (L % 8) * (nwg / 8) + L / 8;
  const int bg = L / gridDim.x;
  const int j = (L - bg * gridDim.x) * blockDim.x + threadIdx.x;
  if (j >= P.slvmax) return;
  const int b0 = bg * MOLSUM_CFG;
  const int nb = min(MOLSUM_CFG, ncfg - b0);
  const int t0 = mol_first[j], t1 = t0 + mol_ns[j];
  const double *es = e_site + (size_t)b0 * nsites;
  double e[MOLSUM_CFG];
#pragma unroll
  for (int i = 0; i < MOLSUM_CFG; ++i) e[i] = 0.0;
  for (int t = t0; t < t1; t += 4) {
    int pz[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) pz[k] = zpos[min(t + k, t1 - 1)];  // a molecule has a site
    double v[MOLSUM_CFG][4];
#pragma unroll
    for (int i = 0; i < MOLSUM_CFG; ++i) {
      const double *ei = es + (size_t)min(i, nb - 1) * nsites;
#pragma unroll
      for (int k = 0; k < 4; ++k) v[i][k] = ei[pz[k]];
    }
#pragma unroll
    for (int i = 0; i < MOLSUM_CFG; ++i)
#pragma unroll
      for (int k = 0; k < 4; ++k) e[i] = (t + k < t1) ? e[i] + v[i][k] : e[i];
  }
  const double epcl = 3.1415926535897932 * P.solute_charge * mol_q[j] / P.screen / P.screen / P.volume;
#pragma unroll
  for (int i = 0; i < MOLSUM_CFG; ++i) {
    if (i >= nb) break;
    double *u = uv + (size_t)(b0 + i) * (P.slvmax + 1) + 1;
    u[j] = (j == slt_j[b0 + i]) ? 0.0 : (0.0 - epcl) + e[i];
  }
}

// per-frame sort of the solvent sites by grid base g3 (z-sorted copies of
// base, charge and B-spline weights, the latter one record [3][p] per site,
// [nsites][3][p]).  Within a plane (KB = ZCLASS) the sites are dealt over
// the 32 LDS bank classes (rowp g2 + g1) mod 32 of their gather reads (rowp
// = K1 + 2 values per real row; a plane is a multiple of 32 values on the
// grids used, so the class does not depend on the plane or on the stencil
// offsets): round r takes the r-th site of every class that has one, the
// classes in the order 0, 2, ..., 30, 1, 3, ..., 31.  Any 32 consecutive
// sites of full rounds then read 32 distinct banks, and any 16 read distinct
// bank pairs, so a half-wave's gather reads do not conflict.  KB = 1: by
// plane only.
constexpr int ZBUCKET_MAX = 4096;
constexpr int ZCLASS = 32;
// bucket of a site: plane g3, then the dealing position of its class
__device__ __forceinline__ int zbucket(int g, int KB, int rowp) {
  const int g3 = (g >> 20) & 1023;
  if (KB == 1) return g3;
  const int c = (rowp * ((g >> 10) & 1023) + (g & 1023)) & (ZCLASS - 1);
  return g3 * ZCLASS + ((c & 1) ? ZCLASS / 2 + (c >> 1) : (c >> 1));
}

// counts per bucket in LDS, one global add per bucket and workgroup
static __global__ __launch_bounds__(256) void k_zcount(int nsites, int nbk, int KB, int rowp,
                                                       const int *__restrict__ sg, int *zcount) {
  __shared__ int cnt[ZBUCKET_MAX];
  for (int z = threadIdx.x; z < nbk; z += blockDim.x) cnt[z] = 0;
  __syncthreads();
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nsites) atomicAdd(&cnt[zbucket(sg[t], KB, rowp)], 1);
  __syncthreads();
  for (int z = threadIdx.x; z < nbk; z += blockDim.x)
    if (cnt[z]) atomicAdd(&zcount[z], cnt[z]);
}


// slot of every site: its rank within its bucket (a rank within the
// workgroup's sites of that bucket (LDS) plus the workgroup's base, one
// global add per bucket and workgroup), then with class buckets its place in
// the plane's dealing order: plane start + sum over the classes of
// min(count, rank) + the classes before its own with more than rank sites
template <typename T>
__global__ __launch_bounds__(256) void k_zscatter(int nsites, int nbk, int KB, int rowp, int p,
                                                  const int *__restrict__ sg, const T *__restrict__ sq,
                                                  const T *__restrict__ sw, const int *__restrict__ zcount,
                                                  const int *__restrict__ zfine, int *zfill, int *zsite, int *zpos,
                                                  int *zg, T *zq, T *zw) {
  __shared__ int cnt[ZBUCKET_MAX], base[ZBUCKET_MAX];
  for (int z = threadIdx.x; z < nbk; z += blockDim.x) cnt[z] = 0;
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = (t < nsites) ? zbucket(sg[t], KB, rowp) : 0;
  const int rank = (t < nsites) ? atomicAdd(&cnt[g], 1) : 0;
  __syncthreads();
  for (int z = threadIdx.x; z < nbk; z += blockDim.x) base[z] = cnt[z] ? atomicAdd(&zfill[z], cnt[z]) : 0;
  __syncthreads();
  if (t >= nsites) return;
  int zt = base[g] + rank;
  if (KB > 1) {
    const int rho = zt - zfine[g];  // rank within the bucket
    const int g0 = g - g % KB;      // the plane's first bucket
    zt = zfine[g0];
    for (int q = 0; q < KB; ++q) {
      const int n = zcount[g0 + q];
      zt += min(n, rho) + ((g0 + q < g && n > rho) ? 1 : 0);
    }
  }
  zsite[zt] = t;
  zpos[t] = zt;
  zg[zt] = sg[t];
  zq[zt] = sq[t];
  for (int k = 0; k < 3 * p; ++k) zw[(size_t)zt * 3 * p + k] = sw[(size_t)t * 3 * p + k];
}

// --------------------------------------------------------------------------
// PME gather (recpcal_energy, recpcal.F90:451-502) plus the neutralising
// background (residual_ene, engproc.F90:859-869); writes the sorted-order
// energies uv[b][1 + j] = -pi Q_s Q_j / (alpha^2 V) + E_recip.  Per-row partial
// sums in T, per-molecule accumulation in FP64.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHR) void k_gather(Params P, const T *__restrict__ phi,
                                                 const int *__restrict__ mol_first,
                                                 const int *__restrict__ mol_ns,
                                                 const double *__restrict__ mol_q,
                                                 const int *__restrict__ sg, const T *__restrict__ sw,
                                                 const T *__restrict__ sq, const int *__restrict__ slt_j,
                                                 double *uv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (j >= P.slvmax) return;
  double *u = uv + (size_t)b * (P.slvmax + 1) + 1;
  if (j == slt_j[b]) {
    u[j] = 0.0;
    return;
  }
  const int K1 = P.K1, K2 = P.K2, K3 = P.K3, p = P.p;
  const T *ph = phi + (size_t)b * K3 * K2 * K1;
  double e = 0.0;
  const int t0 = mol_first[j], nsm = mol_ns[j];
  for (int t = t0; t < t0 + nsm; ++t) {
    const int gp = sg[t];
    const int g1 = gp & 1023, g2 = (gp >> 10) & 1023, g3 = (gp >> 20) & 1023;
    const T *w = sw + (size_t)t * 3 * p;
    const T chr = sq[t];
    for (int c3 = 0; c3 < p; ++c3) {
      T f1 = chr * w[2 * p + c3];
      int r3 = g3 - c3;
      r3 += (r3 < 0) ? K3 : 0;
      for (int c2 = 0; c2 < p; ++c2) {
        T f2 = f1 * w[p + c2];
        int r2 = g2 - c2;
        r2 += (r2 < 0) ? K2 : 0;
        const T *row = ph + ((size_t)r3 * K2 + r2) * K1;
        T rs = 0;
        if (g1 >= p - 1) {
          for (int c1 = 0; c1 < p; ++c1) rs += w[c1] * row[g1 - c1];
        } else {
          for (int c1 = 0; c1 < p; ++c1) {
            int r1 = g1 - c1;
            r1 += (r1 < 0) ? K1 : 0;
            rs += w[c1] * row[r1];
          }
        }
        e += (double)(f2 * rs);
      }
    }
  }
  double epcl = 3.1415926535897932 * P.solute_charge * mol_q[j] / P.screen / P.screen / P.volume;
  u[j] = (0.0 - epcl) + e;
}

// --------------------------------------------------------------------------
// Real space (realcal_proc, realcal.F90:58-135 / get_pair_energy_block
// 585-784).  One workgroup slice per configuration: the molecules whose
// first site lies within rc + R_solute + R_molecule of the configuration's
// centre are found from the cell list by pillars (y, z) with an x range per
// pillar; each candidate molecule is owned by one thread, which sums every
// solute-site x molecule-site pair under the minimum image and adds the
// result to uv[b][1 + j].
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lj_low_t(const Params &P, T eps, T sig2, T dist) {
  T invr2 = sig2 / dist;
  T invr6 = invr2 * invr2 * invr2;
  if (P.ljswitch <= 1) return (T)4 * eps * invr6 * (invr6 - (T)1);
  T sg6 = sig2 * sig2 * sig2;
  if (P.ljswitch == 2) {
    T vdwa = invr6 * invr6 - sg6 * sg6 / (T)(P.lw6 * P.up6);
    T vdwb = invr6 - sg6 / (T)(P.lw3 * P.up3);
    return (T)4 * eps * (vdwa - vdwb);
  }
  T vdwa = invr6 * invr6 - sg6 * sg6 * (T)P.repC;
  T vdwb = invr6 - sg6 * (T)P.attC;
  return (T)4 * eps * (vdwa - vdwb);
}

template <typename T>
__device__ __forceinline__ T lj_switch_t(const Params &P, T eps, T sig2, T dist) {
  T invr2 = sig2 / dist;
  T invr6 = invr2 * invr2 * invr2;
  if (P.ljswitch == 0) {
    T up2 = (T)P.up2, lw2 = (T)P.lw2;
    T d = dist - up2, w = up2 - lw2;
    return (T)4 * eps * invr6 * (invr6 - (T)1) * ((T)2 * dist + up2 - (T)3 * lw2) * (d * d) / (w * w * w);
  }
  if (P.ljswitch == 1) {
    T sw = (sqrt(dist) - (T)P.lwljcut) / (T)(P.upljcut - P.lwljcut);
    T sw3 = sw * sw * sw;
    return (T)4 * eps * invr6 * (invr6 - (T)1) * ((T)1 - (T)10 * sw3 + (T)15 * (sw3 * sw) - (T)6 * (sw3 * sw * sw));
  }
  T sg6 = sig2 * sig2 * sig2;
  if (P.ljswitch == 2) {
    T invr3 = sqrt(invr6);
    T sg3 = sqrt(sg6);
    T a = invr6 - sg6 / (T)P.up6, c = invr3 - sg3 / (T)P.up3;
    T vdwa = (T)(P.up6 / (P.up6 - P.lw6)) * (a * a);
    T vdwb = (T)(P.up3 / (P.up3 - P.lw3)) * (c * c);
    return (T)4 * eps * (vdwa - vdwb);
  }
  T sw = sqrt(dist) - (T)P.lwljcut;
  T sw3 = sw * sw * sw, sw4 = sw3 * sw;
  T vdwa = invr6 * invr6 - sg6 * sg6 * ((T)P.repA * sw3 + (T)P.repB * sw4 + (T)P.repC);
  T vdwb = invr6 - sg6 * ((T)P.attA * sw3 + (T)P.attB * sw4 + (T)P.attC);
  return (T)4 * eps * (vdwa - vdwb);
}

template <typename T> __device__ __forceinline__ T coul_t(T qq, T r, T alpha);
template <> __device__ __forceinline__ float coul_t<float>(float qq, float r, float alpha) {
  return qq * erfcf(alpha * r) / r;
}
// FP64 mirrors the reference's (1 - erf(alpha r)) / r
template <> __device__ __forceinline__ double coul_t<double>(double qq, double r, double alpha) {
  return qq * (1.0 - erf(alpha * r)) / r;
}

// LJ of one pair as realcal_bare writes it (realcal.F90:213-266): zero
// beyond upljcut, the potential switches applied as a factor beyond
// lwljcut, the force switches branching on the distance rst
template <typename T>
__device__ __forceinline__ T lj_bare_t(const Params &P, T eps, T sig2, T dis2, T rst) {
  if (rst > (T)P.upljcut) return (T)0;
  if (P.ljswitch >= 2)
    return (rst <= (T)P.lwljcut) ? lj_low_t<T>(P, eps, sig2, dis2) : lj_switch_t<T>(P, eps, sig2, dis2);
  T invr2 = sig2 / dis2;
  T invr6 = invr2 * invr2 * invr2;
  T eplj = (T)4 * eps * invr6 * (invr6 - (T)1);
  if (rst > (T)P.lwljcut) {
    T swth;
    if (P.ljswitch == 0) {
      T up2 = (T)P.up2, lw2 = (T)P.lw2;
      T d = dis2 - up2, w = up2 - lw2;
      swth = ((T)2 * dis2 + up2 - (T)3 * lw2) * (d * d) / (w * w * w);
    } else {
      T sw = (rst - (T)P.lwljcut) / (T)(P.upljcut - P.lwljcut);
      T sw3 = sw * sw * sw;
      swth = (T)1 - (T)10 * sw3 + (T)15 * (sw3 * sw) - (T)6 * (sw3 * sw * sw);
    }
    eplj = swth * eplj;
  }
  return eplj;
}

// inclusive prefix sum of a[0..n) in LDS, n <= 8 * blockDim.x; all threads call
static __device__ void block_inclusive_scan(int *a, int n) {
  __shared__ int part[NTHR];
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int b0 = threadIdx.x * per;
  int s = 0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < n) s += a[b0 + i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    int v = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = (threadIdx.x == 0) ? 0 : part[threadIdx.x - 1];
  for (int i = 0; i < per; ++i)
    if (b0 + i < n) {
      run += a[b0 + i];
      a[b0 + i] = run;
    }
  __syncthreads();
}

// Candidate molecules of a configuration with (wrapped) centre Cw and reach
// R: pillars (y, z) of the molecule cell grid whose distance from Cw is at
// most R, each with the x range of cells within reach, as up to two
// contiguous ranges of cell-sorted molecules.  Every pillar owns two fixed
// segment slots (empty when unused), so all workgroups of one configuration
// flatten the candidates in the same order.  seg_beg[i] is the first
// molecule of segment i and seg_pre[i] the number of candidates before it;
// returns the segment count nsg (seg_pre[nsg] = total).  All threads call.
static __device__ int pillar_segments(const Params &P, const double Cw[3], double R, const int *__restrict__ cell_start,
                               int *seg_beg, int *seg_pre) {
  const double R2 = R * R;
  int lo[3], cnt[3];
  for (int d = 1; d < 3; ++d) {
    int n = (int)ceil(R / P.hcell[d]) + 1;
    if (2 * n + 1 >= P.nb[d]) {
      lo[d] = 0;
      cnt[d] = P.nb[d];
    } else {
      lo[d] = (int)floor(Cw[d] / P.hcell[d]) - n;
      cnt[d] = 2 * n + 1;
    }
  }
  const int nsg = 2 * cnt[1] * cnt[2];
  for (int t = threadIdx.x; t < nsg; t += blockDim.x) {
    seg_beg[t] = 0;
    seg_pre[t + 1] = 0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt[1] * cnt[2]; t += blockDim.x) {
    int oy = t % cnt[1], oz = t / cnt[1];
    int yc = ((lo[1] + oy) % P.nb[1] + P.nb[1]) % P.nb[1];
    int zc = ((lo[2] + oz) % P.nb[2] + P.nb[2]) % P.nb[2];
    double dyz[2];
    const int cc[2] = {yc, zc};
    for (int e = 0; e < 2; ++e) {
      int d = e + 1;
      double ctr = ((double)cc[e] + 0.5) * P.hcell[d];
      double dd = fabs(Cw[d] - ctr);
      dd = fmin(dd, P.L[d] - dd);
      dyz[e] = fmax(0.0, dd - 0.5 * P.hcell[d]);
    }
    double l2 = dyz[0] * dyz[0] + dyz[1] * dyz[1];
    if (l2 > R2) continue;
    double w = sqrt(R2 - l2);
    int xa = (int)floor((Cw[0] - w) / P.hcell[0]);
    int xb = (int)floor((Cw[0] + w) / P.hcell[0]);
    int rowb = P.nb[0] * (yc + P.nb[1] * zc);
    int n0 = P.nb[0];
    int a0[2], a1[2], npg = 0;
    if (xb - xa + 1 >= n0) {
      a0[0] = 0;
      a1[0] = n0;
      npg = 1;
    } else {
      int xw = ((xa % n0) + n0) % n0;
      int len = xb - xa + 1;
      if (xw + len <= n0) {
        a0[0] = xw;
        a1[0] = xw + len;
        npg = 1;
      } else {
        a0[0] = xw;
        a1[0] = n0;
        a0[1] = 0;
        a1[1] = xw + len - n0;
        npg = 2;
      }
    }
    for (int g = 0; g < npg; ++g) {
      int mb = cell_start[rowb + a0[g]], me = cell_start[rowb + a1[g]];
      seg_beg[2 * t + g] = mb;
      seg_pre[2 * t + g + 1] = me - mb;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) seg_pre[0] = 0;
  block_inclusive_scan(seg_pre + 1, nsg);
  return nsg;
}

// FP32 mode evaluates a pair in FP64 (from the FP64 displacement and
// parameters) when its energy is large enough for FP32 rounding to matter at
// 1e-4 kcal/mol: a repulsive LJ overlap (r < 2^(1/6) sigma, where r^-12
// amplifies the rounding of r) or a Coulomb term |q q| / r above COUL_HOT
// kcal/mol (first-shell pairs of charged sites).  Such pairs are few.
constexpr float COUL_HOT = 24.0f;
// erfc(x) for 0 <= x <= ERFC_XMAX in FP32: exp(-x^2) times the fitted
// polynomial R(u), u = x / ERFC_XMAX (ermgpu_erfc.h, from erfc_fit.py);
// largest absolute error about 1e-7 over the range (one FP32 unit of erfc
// near 0), in 15 operations
__device__ __forceinline__ float erfc_fit(float x) {
  const float u = x * (1.0f / ERFC_XMAX);
  float r = ERFC_R[ERFC_DEG];
#pragma unroll
  for (int k = ERFC_DEG - 1; k >= 0; --k) r = __builtin_fmaf(r, u, ERFC_R[k]);
  // exp(-x^2) = 2^(-x^2 log2 e); the argument stays above -29.3, no
  // denormal or range handling needed
  const float e = __builtin_amdgcn_exp2f(-(x * x) * 1.44269504088896341f);
  return e * r;
}

template <typename T> __device__ __forceinline__ bool hot_pair(T eps, T sig2, T qq, T r2) {
  return (eps > (T)0 && sig2 > (T)1.2599211 * r2) || (qq * qq > (T)(COUL_HOT * COUL_HOT) * r2);
}

// BARE: bare Coulomb q q / r below elecut and the LJ of realcal_bare
// (realcal.F90:155-281) in place of the screened real-space pair
template <typename T, bool BARE = false>
__global__ __launch_bounds__(NTHR) void k_pair(Params P, const double *__restrict__ spos,
                                               const double *__restrict__ sref,
                                               const T *__restrict__ slt_q, const T *__restrict__ slt_lj,
                                               const int *__restrict__ cell_start,
                                               const int *__restrict__ mol_first,
                                               const int *__restrict__ mol_ns, const double *__restrict__ sx,
                                               const double *__restrict__ sy, const double *__restrict__ sz,
                                               const T *__restrict__ sq, const int *__restrict__ sct,
                                               const int *__restrict__ slt_j, const double *__restrict__ rm_ord_p,
                                               const double *__restrict__ slt_lj64,
                                               const double *__restrict__ slt_q64, double *uv) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int ns = P.ns, nvt = P.nvt;
  T *ux = reinterpret_cast<T *>(lds_raw);
  T *uy = ux + ns;
  T *uz = uy + ns;
  T *uq = uz + ns;
  T *ulj = uq + ns;                       // [ns][nvt][2] (eps, sig2)
  T *ulo = ulj + 2 * ns * nvt;            // [ns][3] low parts of the offsets
  int *seg_beg = reinterpret_cast<int *>(ulo + 3 * ns);
  int *seg_pre = seg_beg + MAXSEG;        // MAXSEG + 1 prefix sums
  const int b = blockIdx.x;
  const double C0 = sref[4 * b], C1 = sref[4 * b + 1], C2v = sref[4 * b + 2];
  const double Rs = sref[4 * b + 3];
  const double *xs = spos + (size_t)b * ns * 3;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    ux[s] = (T)(xs[3 * s] - C0);
    uy[s] = (T)(xs[3 * s + 1] - C1);
    uz[s] = (T)(xs[3 * s + 2] - C2v);
    uq[s] = slt_q[s];
    // low parts of the solute offsets (zero in FP64)
    ulo[3 * s] = (T)((xs[3 * s] - C0) - (double)ux[s]);
    ulo[3 * s + 1] = (T)((xs[3 * s + 1] - C1) - (double)uy[s]);
    ulo[3 * s + 2] = (T)((xs[3 * s + 2] - C2v) - (double)uz[s]);
  }
  for (int t = threadIdx.x; t < 2 * ns * nvt; t += blockDim.x) ulj[t] = slt_lj[t];
  // pillars within reach of the centre
  const double Rm = ord_to_dbl(*(const unsigned long long *)rm_ord_p);
  const double R = P.rcmax + Rs + Rm + 1e-6 * (1.0 + P.rcmax);
  const double Cw[3] = {wrap1(C0, P.L[0]), wrap1(C1, P.L[1]), wrap1(C2v, P.L[2])};
  const int nsg = pillar_segments(P, Cw, R, cell_start, seg_beg, seg_pre);
  const int total = seg_pre[nsg];
  const int sj = slt_j[b];
  const T L0 = (T)P.L[0], L1 = (T)P.L[1], L2 = (T)P.L[2];
  const T h0 = (T)P.halfL[0], h1 = (T)P.halfL[1], h2 = (T)P.halfL[2];
  const T lw2 = (T)P.lw2, up2 = (T)P.up2, el2 = (T)P.el2, cut2 = (T)P.cut2, alpha = (T)P.screen;
  double *u = uv + (size_t)b * (P.slvmax + 1) + 1;
  for (int idx = blockIdx.y * blockDim.x + threadIdx.x; idx < total; idx += gridDim.y * blockDim.x) {
    int lo_s = 0, hi_s = nsg;  // last segment with seg_pre <= idx
    while (hi_s - lo_s > 1) {
      int mid = (lo_s + hi_s) >> 1;
      if (seg_pre[mid] <= idx)
        lo_s = mid;
      else
        hi_s = mid;
    }
    const int j = seg_beg[lo_s] + (idx - seg_pre[lo_s]);
    if (j == sj) continue;
    double e = 0.0;
    const int t0 = mol_first[j], nsm = mol_ns[j];
    for (int t = t0; t < t0 + nsm; ++t) {
      double dx = sx[t] - C0, dy = sy[t] - C1, dz = sz[t] - C2v;
      dx -= P.L[0] * rint(dx / P.L[0]);
      dy -= P.L[1] * rint(dy / P.L[1]);
      dz -= P.L[2] * rint(dz / P.L[2]);
      // each offset as hi + lo in T: (hi_v - hi_u) is exact for close pairs
      // (Sterbenz), so FP32 distances keep ~1e-7 relative accuracy
      const T vx = (T)dx, vy = (T)dy, vz = (T)dz;
      const T lx = (T)(dx - (double)vx), ly = (T)(dy - (double)vy), lz = (T)(dz - (double)vz);
      const T qv = sq[t];
      const int cv = sct[t];
      for (int s = 0; s < ns; ++s) {
        T ax = fabs(((vx - ux[s]) + lx) - ulo[3 * s]);
        T ay = fabs(((vy - uy[s]) + ly) - ulo[3 * s + 1]);
        T az = fabs(((vz - uz[s]) + lz) - ulo[3 * s + 2]);
        ax = (ax > h0) ? L0 - ax : ax;
        ay = (ay > h1) ? L1 - ay : ay;
        az = (az > h2) ? L2 - az : az;
        T r2 = ax * ax + ay * ay + az * az;
        if (r2 <= cut2) {
          T ep = 0;
          T eps = ulj[(s * nvt + cv) * 2], sig2 = ulj[(s * nvt + cv) * 2 + 1];
          if constexpr (sizeof(T) == 4) {
            // a repulsive overlap or a large Coulomb term: the pair in FP64
            // from the FP64 displacement and parameters
            if (hot_pair<T>(eps, sig2, uq[s] * qv, r2)) {
              const double ox = (double)ux[s] + (double)ulo[3 * s];
              const double oy = (double)uy[s] + (double)ulo[3 * s + 1];
              const double oz = (double)uz[s] + (double)ulo[3 * s + 2];
              double bx = fabs(dx - ox), by = fabs(dy - oy), bz = fabs(dz - oz);
              bx = (bx > P.halfL[0]) ? P.L[0] - bx : bx;
              by = (by > P.halfL[1]) ? P.L[1] - by : by;
              bz = (bz > P.halfL[2]) ? P.L[2] - bz : bz;
              const double r2d = bx * bx + by * by + bz * bz;
              const double e64 = slt_lj64[(s * nvt + cv) * 2], s64 = slt_lj64[(s * nvt + cv) * 2 + 1];
              double epd = 0.0;
              if constexpr (BARE) {
                const double rd = sqrt(r2d);
                epd = lj_bare_t<double>(P, e64, s64, r2d, rd);
                if (rd < P.elecut) epd += slt_q64[s] * (double)qv / rd;
              } else {
                if (r2d <= P.lw2)
                  epd = lj_low_t<double>(P, e64, s64, r2d);
                else if (r2d <= P.up2)
                  epd = lj_switch_t<double>(P, e64, s64, r2d);
                if (r2d <= P.el2) epd += coul_t<double>(slt_q64[s] * (double)qv, sqrt(r2d), P.screen);
              }
              e += epd;
              continue;
            }
          }
          if constexpr (BARE) {
            const T rst = sqrt(r2);
            ep = lj_bare_t<T>(P, eps, sig2, r2, rst);
            if (rst < (T)P.elecut) ep += uq[s] * qv / rst;
            e += (double)ep;
            continue;
          }
          if (eps > (T)0) {
            if (r2 <= lw2)
              ep += lj_low_t<T>(P, eps, sig2, r2);
            else if (r2 <= up2)
              ep += lj_switch_t<T>(P, eps, sig2, r2);
          }
          if (r2 <= el2) ep += coul_t<T>(uq[s] * qv, sqrt(r2), alpha);
          e += (double)ep;
        }
      }
    }
    u[j] += e;
  }
}

// --------------------------------------------------------------------------
// Real space in tiles with compacted pair lists (get_pair_energy_block,
// realcal.F90:585-784, rebuilt for 64-wide waves).  A workgroup takes one
// configuration; its waves take chunks of 64 candidate molecules (the
// pillar candidates of k_pair, chunk c to wave c mod #waves), one molecule
// per lane.  The chunk's sites are walked in tiles of 64 (one site per lane,
// SoA loads, the FP64 displacement from the configuration centre and its
// wrap formed once per site).  For every group of PT_GRP solute sites the
// lanes test their site against each solute site and append the pairs
// inside the cutoff to a wave-local queue (ballot + mbcnt, no atomics); the
// queued pairs are then evaluated with every lane busy, and each site adds
// its pairs' energies in solute-site order.  Per-site sums are reduced per
// molecule in site order by the lane that owns the molecule, so every sum
// is deterministic.  A per-pair minimum image is only formed when the reach
// rc + R_solute is at least half the box (otherwise the site's own wrap is
// already the minimum image of every pair within the cutoff).
// --------------------------------------------------------------------------
constexpr int PT_GRP = 8;                 // solute sites per compaction group
constexpr int PT_QCAP = 64 * PT_GRP;      // queued pairs per wave (at most)
constexpr float PT_FAR = 1e18f;           // sentinel displacement (A): (2 PT_FAR)^2 < FLT_MAX

template <typename T> struct V4T;
template <> struct V4T<float> { using type = float4; };
template <> struct V4T<double> { using type = double4; };

// LDS of k_pair_tile.  Head: the solute's sites as (x, y, z, q), their low
// parts and their test rows (-2x, -2y, -2z, threshold) (T4 each, ns rounded
// up to whole groups of PT_GRP: the slots past ns hold sentinels that fail
// every test), the LJ rows
// [ns][nvt] (eps, sig2), the pillar segments
// (segcap = 2 nb1 nb2 slots).  Per wave: the tile's sites (x, y, z, q) and
// (low parts, LJ type), molecule offsets [65] and first sites [64], the queue
// [PT_QCAP + 64] (int) and its energies [PT_QCAP + 64] (f64; the last 64 are
// the lanes' spare slots, energy 0), the per-site sums [64] (f64), the slots
// of the queued pairs evaluated in FP64 [PT_QCAP] (int).
template <typename T> __host__ __device__ inline size_t pair_tile_wave_lds() {
  return ((size_t)2 * 64 * 4 * sizeof(T) + 65 * 4 + 64 * 4 + (size_t)PT_QCAP * (4 + 8 + 4) + 64 * (4 + 8) + 64 * 8 +
          16 + 15) &
         ~(size_t)15;
}
template <typename T> __host__ __device__ inline size_t pair_tile_head(int ns, int nvt, int segcap) {
  const size_t nsp = (size_t)(ns + PT_GRP - 1) / PT_GRP * PT_GRP;
  const size_t a = sizeof(T) * ((size_t)12 * nsp + (size_t)2 * ns * nvt) + sizeof(int) * (2 * (size_t)segcap + 2) + 64;
  return (a + 15) & ~(size_t)15;
}
template <typename T> __host__ __device__ inline size_t pair_tile_lds(int ns, int nvt, int segcap, int nwave) {
  return pair_tile_head<T>(ns, nvt, segcap) + (size_t)nwave * pair_tile_wave_lds<T>();
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Configuration b (ns sites at xs), split `split` of nsplit workgroups;
// all threads of the workgroup call it
template <typename T, bool BARE>
__device__ __forceinline__ void pair_tile_body(const Params &P, int ns, int b, int split, int nsplit,
                                               const double *__restrict__ xs, const double *__restrict__ sref,
                                               const T *__restrict__ slt_q, const T *__restrict__ slt_lj,
                                               const int *__restrict__ cell_start,
                                               const int *__restrict__ mol_first,
                                               const int *__restrict__ mol_ns, const double *__restrict__ sx,
                                               const double *__restrict__ sy, const double *__restrict__ sz,
                                               const T *__restrict__ sq, const int *__restrict__ sct,
                                               const int *__restrict__ slt_j,
                                               const double *__restrict__ rm_ord_p,
                                               const double *__restrict__ slt_lj64,
                                               const double *__restrict__ slt_q64,
                                               const double *__restrict__ mol_d1, double *uv) {
  constexpr bool F32 = sizeof(T) == 4;
  using T4 = typename V4T<T>::type;
  using T2 = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int nvt = P.nvt;
  const int nsp = (ns + PT_GRP - 1) / PT_GRP * PT_GRP;
  T4 *us4 = reinterpret_cast<T4 *>(lds_raw);  // [nsp] (x, y, z, q) offsets from the centre
  T4 *ul4 = us4 + nsp;                        // [nsp] low parts (x, y, z, 0)
  T4 *ut4 = ul4 + nsp;                        // [nsp] test rows (-2x, -2y, -2z, threshold)
  T2 *ulj = reinterpret_cast<T2 *>(ut4 + nsp); // [ns][nvt] (eps, sig2)
  const int segcap = 2 * P.nb[1] * P.nb[2];
  int *seg_beg = reinterpret_cast<int *>(ulj + ns * nvt);
  int *seg_pre = seg_beg + segcap;
  const size_t base = pair_tile_head<T>(ns, nvt, segcap);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned char *wl = lds_raw + base + (size_t)wv * pair_tile_wave_lds<T>();
  T4 *tu4 = reinterpret_cast<T4 *>(wl);   // [64] (x, y, z, q) of the tile's sites
  T4 *tl4 = tu4 + 64;                     // [64] (low x, low y, low z, LJ type bits)
  int *woff = reinterpret_cast<int *>(tl4 + 64);  // [65] first chunk site of molecule lane
  int *wfirst = woff + 65;                // [64] first sorted site of molecule lane
  int *wq = wfirst + 64;                  // [PT_QCAP + 64] queued pairs: solute site << 6 | tile lane
  // (a byte offset from wl, rounded down to 8 over the spare slot of wq, so
  // the pointers stay LDS pointers for the compiler)
  double *weq = reinterpret_cast<double *>(
      wl + (((size_t)128 * sizeof(T4) + (size_t)(65 + 64 + PT_QCAP + 64 + 1) * sizeof(int)) & ~(size_t)7));
  double *wes = weq + PT_QCAP + 64;       // [64] per-site sums of a tile
  int *whq = reinterpret_cast<int *>(wes + 64);  // [PT_QCAP] queue slots of the pairs evaluated in FP64

  const double C0 = sref[4 * b], C1 = sref[4 * b + 1], C2v = sref[4 * b + 2];
  const double Rs = sref[4 * b + 3];
  // sites farther than rc + Rs from the centre meet no solute site within
  // the cutoff; the minimum image of a pair is the site's own wrap unless
  // the reach rc + Rs reaches half the box
  const double reach = P.rcmax + Rs + 1e-6 * (1.0 + P.rcmax), reach2 = reach * reach;
  const bool img = reach >= 0.5 * fmin(P.L[0], fmin(P.L[1], P.L[2]));
  // Without per-pair images a tile site t (|t| <= reach) and a solute site s
  // are tested as |t|^2 - 2 t.s <= cut2 - |s|^2 (three FMAs).  The bound on
  // its rounding, a few units in the last place of (reach + Rs)^2 (about
  // 4e-7 relative in FP32), sits well inside the margin added to cut2; the
  // exact distance of every queued pair decides.
  const double xr = reach + Rs;
  const double cut2x = F32 ? P.cut2 + fmax(P.cut2 * 1e-5, 4e-6 * xr * xr) : P.cut2 + 1e-12 * xr * xr;
  for (int s = threadIdx.x; s < nsp; s += blockDim.x) {
    T4 u, l, t;
    l.x = l.y = l.z = l.w = 0;
    // padding: never below the threshold (|t|^2 >= 0 > -1)
    t.x = t.y = t.z = 0;
    t.w = (T)-1;
    if (s < ns) {
      const double d0 = xs[3 * s] - C0, d1 = xs[3 * s + 1] - C1, d2 = xs[3 * s + 2] - C2v;
      u.x = (T)d0;
      u.y = (T)d1;
      u.z = (T)d2;
      u.w = slt_q[s];
      l.x = (T)(d0 - (double)u.x);
      l.y = (T)(d1 - (double)u.y);
      l.z = (T)(d2 - (double)u.z);
      t.x = (T)-2 * u.x;
      t.y = (T)-2 * u.y;
      t.z = (T)-2 * u.z;
      t.w = (T)(cut2x - ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z));
    } else {
      // sentinel: x displacement -PT_FAR, so its distance to any tile site
      // (at most PT_FAR away, the inactive lanes' sentinel) squares to at
      // least PT_FAR^2, far outside every cutoff and finite in FP32
      u.x = -(T)PT_FAR;
      u.y = u.z = u.w = 0;
    }
    us4[s] = u;
    ul4[s] = l;
    ut4[s] = t;
  }
  for (int t = threadIdx.x; t < ns * nvt; t += blockDim.x) {
    T2 v;
    v.x = slt_lj[2 * t];
    v.y = slt_lj[2 * t + 1];
    ulj[t] = v;
  }
  const double Rm = ord_to_dbl(*(const unsigned long long *)rm_ord_p);
  const double R = P.rcmax + Rs + Rm + 1e-6 * (1.0 + P.rcmax);
  const double Cw[3] = {wrap1(C0, P.L[0]), wrap1(C1, P.L[1]), wrap1(C2v, P.L[2])};
  const int nsg = pillar_segments(P, Cw, R, cell_start, seg_beg, seg_pre);
  const int total = seg_pre[nsg];
  const int sj = slt_j[b];
  const double iL0 = 1.0 / P.L[0], iL1 = 1.0 / P.L[1], iL2 = 1.0 / P.L[2];
  const T L0 = (T)P.L[0], L1 = (T)P.L[1], L2 = (T)P.L[2];
  const T h0 = (T)P.halfL[0], h1 = (T)P.halfL[1], h2 = (T)P.halfL[2];
  const T lw2 = (T)P.lw2, up2 = (T)P.up2, el2 = (T)P.el2, cut2 = (T)P.cut2, alpha = (T)P.screen;
  const T w3 = (up2 - lw2) * (up2 - lw2) * (up2 - lw2), iw3 = (T)1 / w3;  // switch denominator
  // with per-pair images the queue test is the plain distance on the high
  // parts (FP32, with a margin); the exact test on the full displacement
  // decides
  const T cut2q = F32 ? (T)(P.cut2 * (1.0 + 1e-5)) : cut2;
  // FP32 Coulomb: the fitted erfc where alpha r stays within its range
  const bool erfc_in_fit = P.screen * sqrt(P.el2) <= (double)ERFC_XMAX;
  double *u = uv + (size_t)b * (P.slvmax + 1) + 1;
  // the lanes' spare energy slots hold 0 throughout (only 0 is written
  // there); every queue slot holds a valid entry (solute site 0), so the
  // evaluation reads and writes a full round of slots without branches
  weq[PT_QCAP + lane] = 0.0;
  for (int k = lane; k < PT_QCAP + 64; k += 64) wq[k] = lane;
  const int nchunk = (total + 63) >> 6;
  for (int c = split * nw + wv; c < nchunk; c += nsplit * nw) {
    // molecule of this lane: flattened candidate c*64 + lane
    const int idx = c * 64 + lane;
    int j = -1, nsm = 0, first = 0;
    if (idx < total) {
      int lo_s = 0, hi_s = nsg;  // last segment with seg_pre <= idx
      while (hi_s - lo_s > 1) {
        const int mid = (lo_s + hi_s) >> 1;
        if (seg_pre[mid] <= idx)
          lo_s = mid;
        else
          hi_s = mid;
      }
      j = seg_beg[lo_s] + (idx - seg_pre[lo_s]);
      if (j != sj) {
        first = mol_first[j];
        // a molecule whose first site lies farther than reach + mol_d1 from
        // the centre has no site within reach (triangle inequality): it
        // takes no tile lanes
        double dx = sx[first] - C0, dy = sy[first] - C1, dz = sz[first] - C2v;
        dx -= P.L[0] * rint(dx * iL0);
        dy -= P.L[1] * rint(dy * iL1);
        dz -= P.L[2] * rint(dz * iL2);
        const double kr = reach + mol_d1[j] + 1e-6 * (1.0 + reach);
        if (img || dx * dx + dy * dy + dz * dz <= kr * kr) nsm = mol_ns[j];
      }
    }
    // exclusive scan of the site counts over the wave
    int off = nsm;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(off, d, 64);
      if (lane >= d) off += v;
    }
    const int tot_sites = __shfl(off, 63, 64);
    off -= nsm;
    woff[lane] = off;
    if (lane == 63) woff[64] = tot_sites;
    wfirst[lane] = first;
    wave_lds_sync();
    double acc = 0.0;  // energy of molecule `lane`
    // the site data of a tile are loaded one tile ahead: the loads of tile
    // t0 + 64 are in flight while tile t0's groups run
    double px = 0.0, py = 0.0, pz = 0.0;
    T pq = 0;
    int pc = 0;
    bool pv = false;
    auto fetch = [&](int t0n) {
      const int cs = t0n + lane;
      pv = cs < tot_sites;
      if (pv) {
        int a = 0, e = 64;  // molecule k of chunk site cs: last k with woff[k] <= cs
        while (e - a > 1) {
          const int mid = (a + e) >> 1;
          if (woff[mid] <= cs)
            a = mid;
          else
            e = mid;
        }
        const int t = wfirst[a] + (cs - woff[a]);
        px = sx[t];
        py = sy[t];
        pz = sz[t];
        pq = sq[t];
        pc = sct[t];
      }
    };
    fetch(0);
    for (int t0 = 0; t0 < tot_sites; t0 += 64) {
      bool act = pv;
      T4 tu, tl;
      tu.x = tu.y = tu.z = tu.w = 0;
      tl.x = tl.y = tl.z = tl.w = 0;
      if (act) {
        double dx = px - C0, dy = py - C1, dz = pz - C2v;
        dx -= P.L[0] * rint(dx * iL0);
        dy -= P.L[1] * rint(dy * iL1);
        dz -= P.L[2] * rint(dz * iL2);
        act = (dx * dx + dy * dy + dz * dz) <= reach2;
        tu.x = (T)dx;
        tu.y = (T)dy;
        tu.z = (T)dz;
        tu.w = pq;
        if constexpr (F32) {
          tl.x = (T)(dx - (double)tu.x);
          tl.y = (T)(dy - (double)tu.y);
          tl.z = (T)(dz - (double)tu.z);
          tl.w = __int_as_float(pc);
        } else {
          tl.w = __longlong_as_double((long long)pc);
        }
      }
      // a lane without an active site takes a sentinel displacement, so the
      // distance compare alone decides every pair below
      if (!act) tu.x = (T)PT_FAR;
      tu4[lane] = tu;
      tl4[lane] = tl;
      if (t0 + 64 < tot_sites) fetch(t0 + 64);
      const T t2 = tu.x * tu.x + tu.y * tu.y + tu.z * tu.z;  // PT_FAR^2 for an inactive lane
      double e = 0.0;  // this site's energy, in solute-site order
      for (int s0 = 0; s0 < ns; s0 += PT_GRP) {
        int pend[PT_GRP];
        int qn = 0;
        // appends pair (s0 + g, lane) to the queue where `in` holds; inactive
        // lanes and the padding sites past ns fail every test through their
        // sentinels, so a bare compare is the ballot's mask
        auto enqueue = [&](int g, bool in) {
          const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
          const int pos = qn + lanes_below(m);
          // a lane without the pair writes its own spare slot, whose energy
          // stays 0: no branch here or in the sums (a select on the ballot
          // mask, kept from becoming a branch around pos)
          int slot;
          asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(slot) : "v"(PT_QCAP + lane), "v"(pos), "s"(m));
          pend[g] = slot;
          wq[slot] = ((s0 + g) << 6) | lane;
          qn += (int)__popcll(m);
        };
        // the group's solute rows, all read before the queue writes below
        // (one wait for the eight broadcast reads)
        T4 usg[PT_GRP];
        if (!img) {
#pragma unroll
          for (int g = 0; g < PT_GRP; ++g) usg[g] = ut4[s0 + g];
#pragma unroll
          for (int g = 0; g < PT_GRP; ++g) {
            const T4 ut = usg[g];
            T v = fma(tu.x, ut.x, t2);
            v = fma(tu.y, ut.y, v);
            v = fma(tu.z, ut.z, v);
            enqueue(g, v <= ut.w);
          }
        } else {
#pragma unroll
          for (int g = 0; g < PT_GRP; ++g) usg[g] = us4[s0 + g];
#pragma unroll
          for (int g = 0; g < PT_GRP; ++g) {
            const T4 us = usg[g];
            T ax = fabs(tu.x - us.x), ay = fabs(tu.y - us.y), az = fabs(tu.z - us.z);
            ax = (ax > h0) ? L0 - ax : ax;
            ay = (ay > h1) ? L1 - ay : ay;
            az = (az > h2) ? L2 - az : az;
            enqueue(g, (ax * ax + ay * ay + az * az) <= cut2q);
          }
        }
        wave_lds_sync();
        // the queued pairs, every lane busy; in FP32 the pairs that need FP64
        // (hot_pair) go to a second queue evaluated afterwards, so the FP64
        // code runs on full waves too
        int nh = 0;
        for (int q0 = 0; q0 < qn; q0 += 64) {
          // a round of 64 slots: the lanes past qn read a stale (valid)
          // entry and write energy 0 (slots below PT_QCAP + 64, so a spare
          // slot keeps its 0)
          const int qi = q0 + lane;
          const bool valid = qi < qn;
          const int ent = wq[qi];
          const int i = ent & 63, s = ent >> 6;
          const T4 v = tu4[i], vl = tl4[i], us = us4[s];
          T ax, ay, az;
          if constexpr (F32) {
            const T4 ls = ul4[s];
            ax = ((v.x - us.x) + vl.x) - ls.x;
            ay = ((v.y - us.y) + vl.y) - ls.y;
            az = ((v.z - us.z) + vl.z) - ls.z;
          } else {
            ax = v.x - us.x;
            ay = v.y - us.y;
            az = v.z - us.z;
          }
          if (img) {  // the distance only squares the components otherwise
            ax = fabs(ax);
            ay = fabs(ay);
            az = fabs(az);
            ax = (ax > h0) ? L0 - ax : ax;
            ay = (ay > h1) ? L1 - ay : ay;
            az = (az > h2) ? L2 - az : az;
          }
          const T r2 = ax * ax + ay * ay + az * az;
          const T qsv = v.w;
          int cvi;
          if constexpr (F32)
            cvi = __float_as_int(vl.w);
          else
            cvi = (int)__double_as_longlong(vl.w);
          const T2 lj = ulj[__umul24(s, nvt) + cvi];  // s < 4096, nvt < 12800
          const T eps = lj.x, sig2 = lj.y;
          const bool inr = valid && r2 <= cut2;
          bool hot = false;
          if constexpr (F32) hot = inr & hot_pair<T>(eps, sig2, us.w * qsv, r2);  // both evaluated
          const unsigned long long mh = __ballot(hot);
          if (hot) whq[nh + lanes_below(mh)] = qi;
          nh += (int)__popcll(mh);
          T ep = 0;
          if constexpr (F32 && !BARE) {
            if (P.ljswitch == 0) {
              // every lane computes, one select keeps the pairs inside the
              // cutoff: one hardware reciprocal square root (1 ulp) per pair
              // in place of the divisions and the rounded square root (the
              // pairs it reaches are not hot, so their terms are below
              // 24 kcal/mol (Coulomb) and eps (LJ)); the CHARMM potential
              // switch as the low-range term, scaled by the switch beyond
              // lwljcut, zero beyond upljcut or for eps = 0 (the expressions
              // of lj_low_t and lj_switch_t)
              const T rs = (T)__builtin_amdgcn_rsqf((float)r2);
              const T invr2 = sig2 * (rs * rs);
              const T invr6 = invr2 * invr2 * invr2;
              const T elo = (T)4 * eps * invr6 * (invr6 - (T)1);
              const T d = r2 - up2;
              const T sw = ((T)2 * r2 + up2 - (T)3 * lw2) * (d * d);
              const T esw = elo * sw * iw3;
              T elj = (r2 <= lw2) ? elo : esw;
              elj = (r2 <= up2 && eps > (T)0) ? elj : (T)0;
              const float xe = alpha * (r2 * rs);
              const T ec = us.w * qsv * (erfc_in_fit ? erfc_fit(xe) : erfcf(xe)) * rs;
              const T et = elj + ((r2 <= el2) ? ec : (T)0);
              ep = (inr && !hot) ? et : (T)0;
            } else if (inr && !hot) {
              const T rs = (T)__builtin_amdgcn_rsqf((float)r2);
              if (eps > (T)0) {
                if (r2 <= lw2)
                  ep += lj_low_t<T>(P, eps, sig2, r2);
                else if (r2 <= up2)
                  ep += lj_switch_t<T>(P, eps, sig2, r2);
              }
              if (r2 <= el2) {
                const float xe = alpha * (r2 * rs);
                ep += us.w * qsv * (erfc_in_fit ? erfc_fit(xe) : erfcf(xe)) * rs;
              }
            }
          } else if (inr && !hot) {
            if constexpr (BARE) {
              const T rst = sqrt(r2);
              ep = lj_bare_t<T>(P, eps, sig2, r2, rst);
              if (rst < (T)P.elecut) ep += us.w * qsv / rst;
            } else {
              if (P.ljswitch == 0) {
                const T invr2 = sig2 / r2;
                const T invr6 = invr2 * invr2 * invr2;
                const T elo = (T)4 * eps * invr6 * (invr6 - (T)1);
                const T d = r2 - up2;
                const T sw = ((T)2 * r2 + up2 - (T)3 * lw2) * (d * d);
                const T esw = elo * sw / (w3);
                T elj = (r2 <= lw2) ? elo : esw;
                elj = (r2 <= up2 && eps > (T)0) ? elj : (T)0;
                ep = elj;
              } else if (eps > (T)0) {
                if (r2 <= lw2)
                  ep += lj_low_t<T>(P, eps, sig2, r2);
                else if (r2 <= up2)
                  ep += lj_switch_t<T>(P, eps, sig2, r2);
              }
              if (r2 <= el2) ep += coul_t<T>(us.w * qsv, sqrt(r2), alpha);
            }
          }
          weq[qi] = (double)ep;  // a hot pair's slot is rewritten in FP64 below
        }
        if constexpr (F32) {
          if (nh > 0) {
            wave_lds_sync();
            for (int h0 = 0; h0 < nh; h0 += 64) {
              const int hi = h0 + lane;
              if (hi < nh) {
                const int qi = whq[hi];
                const int ent = wq[qi];
                const int i = ent & 63, s = ent >> 6;
                const T4 v = tu4[i], vl = tl4[i], us = us4[s], ls = ul4[s];
                const int cvi = __float_as_int(vl.w);
                const T qsv = v.w;
                double bx = ((double)v.x + (double)vl.x) - ((double)us.x + (double)ls.x);
                double by = ((double)v.y + (double)vl.y) - ((double)us.y + (double)ls.y);
                double bz = ((double)v.z + (double)vl.z) - ((double)us.z + (double)ls.z);
                bx = fabs(bx);
                by = fabs(by);
                bz = fabs(bz);
                if (img) {
                  bx = (bx > P.halfL[0]) ? P.L[0] - bx : bx;
                  by = (by > P.halfL[1]) ? P.L[1] - by : by;
                  bz = (bz > P.halfL[2]) ? P.L[2] - bz : bz;
                }
                const double r2d = bx * bx + by * by + bz * bz;
                const double e64 = slt_lj64[(s * nvt + cvi) * 2], s64 = slt_lj64[(s * nvt + cvi) * 2 + 1];
                double epd = 0.0;
                if constexpr (BARE) {
                  const double rd = sqrt(r2d);
                  epd = lj_bare_t<double>(P, e64, s64, r2d, rd);
                  if (rd < P.elecut) epd += slt_q64[s] * (double)qsv / rd;
                } else {
                  if (r2d <= P.lw2)
                    epd = lj_low_t<double>(P, e64, s64, r2d);
                  else if (r2d <= P.up2)
                    epd = lj_switch_t<double>(P, e64, s64, r2d);
                  if (r2d <= P.el2) epd += coul_t<double>(slt_q64[s] * (double)qsv, sqrt(r2d), P.screen);
                }
                weq[qi] = epd;
              }
            }
          }
        }
        wave_lds_sync();
#pragma unroll
        for (int g = 0; g < PT_GRP; ++g)
          e += weq[pend[g]];
        wave_lds_sync();
      }
      wes[lane] = act ? e : 0.0;
      wave_lds_sync();
      // molecule `lane` adds its sites of this tile in site order
      if (nsm > 0) {
        const int a = max(off, t0), z = min(off + nsm, t0 + 64);
        for (int q = a; q < z; ++q) acc += wes[q - t0];
      }
      wave_lds_sync();
    }
    if (nsm > 0) u[j] += acc;
  }
}

template <typename T, bool BARE = false>
__global__ __launch_bounds__(NTHR) void k_pair_tile(Params P, const double *__restrict__ spos,
                                                    const double *__restrict__ sref,
                                                    const T *__restrict__ slt_q, const T *__restrict__ slt_lj,
                                                    const int *__restrict__ cell_start,
                                                    const int *__restrict__ mol_first,
                                                    const int *__restrict__ mol_ns, const double *__restrict__ sx,
                                                    const double *__restrict__ sy, const double *__restrict__ sz,
                                                    const T *__restrict__ sq, const int *__restrict__ sct,
                                                    const int *__restrict__ slt_j,
                                                    const double *__restrict__ rm_ord_p,
                                                    const double *__restrict__ slt_lj64,
                                                    const double *__restrict__ slt_q64,
                                                    const double *__restrict__ mol_d1, double *uv) {
  const int b = blockIdx.x;
  pair_tile_body<T, BARE>(P, P.ns, b, blockIdx.y, gridDim.y, spos + (size_t)b * P.ns * 3, sref, slt_q, slt_lj,
                          cell_start, mol_first, mol_ns, sx, sy, sz, sq, sct, slt_j, rm_ord_p, slt_lj64, slt_q64,
                          mol_d1, uv);
}

// real space of a heterogeneous batch as one persistent launch:
// workgroups pull (configuration, split) items, the runs in the order given
// (largest solutes first), each with its solute type's tables
template <typename T, bool BARE = false>
__global__ __launch_bounds__(NTHR) void k_pair_tile_runs(Params P, const double *__restrict__ spos,
                                                         const double *__restrict__ sref,
                                                         const int *__restrict__ cell_start,
                                                         const int *__restrict__ mol_first,
                                                         const int *__restrict__ mol_ns,
                                                         const double *__restrict__ sx,
                                                         const double *__restrict__ sy,
                                                         const double *__restrict__ sz, const T *__restrict__ sq,
                                                         const int *__restrict__ sct, const int *__restrict__ slt_j,
                                                         const double *__restrict__ rm_ord_p,
                                                         const GRun *__restrict__ runs, int nrun, int nitem,
                                                         int *counter, const double *__restrict__ mol_d1,
                                                         double *uv) {
  for (int base = next_items(counter); base < nitem; base = next_items(counter))
  for (int it = base; it < min(nitem, base + ITEM_CHUNK); ++it) {
    __syncthreads();  // the previous item's LDS is free
    int r = 0;
    while (r + 1 < nrun && runs[r + 1].pitem0 <= it) ++r;
    const int ns = runs[r].ns, nsplit = runs[r].nsplit, rel = it - runs[r].pitem0;
    const int bl = rel / nsplit, split = rel - bl * nsplit, b = runs[r].cfg0 + bl;
    const double *xs = spos + ((size_t)runs[r].site0 + (size_t)bl * ns) * 3;
    pair_tile_body<T, BARE>(P, ns, b, split, nsplit, xs, sref, static_cast<const T *>(runs[r].q_t),
                            static_cast<const T *>(runs[r].lj_t), cell_start, mol_first, mol_ns, sx, sy, sz, sq, sct,
                            slt_j, rm_ord_p, runs[r].lj64, runs[r].q64, mol_d1, uv);
  }
}

// realcal_bare without a box (realcal.F90:155-281 with boxshp = 0): the
// solute with every site of every physical molecule, raw coordinates, no
// Coulomb cutoff, LJ up to upljcut.  Grid (configuration, molecule blocks),
// one molecule per thread.  Offsets as hi + lo pairs and the FP64 overlap
// recomputation as in k_pair.
template <typename T>
__global__ __launch_bounds__(NTHR) void k_pair_open(Params P, const double *__restrict__ spos,
                                                    const double *__restrict__ sref,
                                                    const T *__restrict__ slt_q, const T *__restrict__ slt_lj,
                                                    const int *__restrict__ mol_first,
                                                    const int *__restrict__ mol_ns, const double *__restrict__ sx,
                                                    const double *__restrict__ sy, const double *__restrict__ sz,
                                                    const T *__restrict__ sq, const int *__restrict__ sct,
                                                    const int *__restrict__ slt_j,
                                                    const double *__restrict__ slt_lj64,
                                                    const double *__restrict__ slt_q64, double *uv) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int ns = P.ns, nvt = P.nvt;
  T *ux = reinterpret_cast<T *>(lds_raw);
  T *uy = ux + ns;
  T *uz = uy + ns;
  T *uq = uz + ns;
  T *ulj = uq + ns;             // [ns][nvt][2] (eps, sig2)
  T *ulo = ulj + 2 * ns * nvt;  // [ns][3] low parts of the offsets
  const int b = blockIdx.x;
  const double C0 = sref[4 * b], C1 = sref[4 * b + 1], C2v = sref[4 * b + 2];
  const double *xs = spos + (size_t)b * ns * 3;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    ux[s] = (T)(xs[3 * s] - C0);
    uy[s] = (T)(xs[3 * s + 1] - C1);
    uz[s] = (T)(xs[3 * s + 2] - C2v);
    uq[s] = slt_q[s];
    ulo[3 * s] = (T)((xs[3 * s] - C0) - (double)ux[s]);
    ulo[3 * s + 1] = (T)((xs[3 * s + 1] - C1) - (double)uy[s]);
    ulo[3 * s + 2] = (T)((xs[3 * s + 2] - C2v) - (double)uz[s]);
  }
  for (int t = threadIdx.x; t < 2 * ns * nvt; t += blockDim.x) ulj[t] = slt_lj[t];
  __syncthreads();
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= P.slvmax || j == slt_j[b]) return;
  double e = 0.0;
  const int t0 = mol_first[j], nsm = mol_ns[j];
  for (int t = t0; t < t0 + nsm; ++t) {
    const double dx = sx[t] - C0, dy = sy[t] - C1, dz = sz[t] - C2v;
    const T vx = (T)dx, vy = (T)dy, vz = (T)dz;
    const T lx = (T)(dx - (double)vx), ly = (T)(dy - (double)vy), lz = (T)(dz - (double)vz);
    const T qv = sq[t];
    const int cv = sct[t];
    for (int s = 0; s < ns; ++s) {
      const T ax = ((vx - ux[s]) + lx) - ulo[3 * s];
      const T ay = ((vy - uy[s]) + ly) - ulo[3 * s + 1];
      const T az = ((vz - uz[s]) + lz) - ulo[3 * s + 2];
      const T r2 = ax * ax + ay * ay + az * az;
      const T eps = ulj[(s * nvt + cv) * 2], sig2 = ulj[(s * nvt + cv) * 2 + 1];
      if constexpr (sizeof(T) == 4) {
        if (hot_pair<T>(eps, sig2, uq[s] * qv, r2)) {
          const double bx = dx - ((double)ux[s] + (double)ulo[3 * s]);
          const double by = dy - ((double)uy[s] + (double)ulo[3 * s + 1]);
          const double bz = dz - ((double)uz[s] + (double)ulo[3 * s + 2]);
          const double r2d = bx * bx + by * by + bz * bz, rd = sqrt(r2d);
          e += lj_bare_t<double>(P, slt_lj64[(s * nvt + cv) * 2], slt_lj64[(s * nvt + cv) * 2 + 1], r2d, rd) +
               slt_q64[s] * (double)qv / rd;
          continue;
        }
      }
      const T rst = sqrt(r2);
      e += (double)(lj_bare_t<T>(P, eps, sig2, r2, rst) + uq[s] * qv / rst);
    }
  }
  uv[(size_t)b * (P.slvmax + 1) + 1 + j] += e;
}

// realcal_self (realcal.F90:284-328) of configuration b, FP64
static __global__ __launch_bounds__(NTHR) void k_selfreal(Params P, const double *__restrict__ spos, int b,
                                                   const double *__restrict__ qsolute, double *out) {
  __shared__ double red[NTHR];
  const int ns = P.ns;
  const double *x = spos + (size_t)b * ns * 3;
  double e = 0.0;
  const double a = P.screen;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    double qi = qsolute[i];
    e += -qi * qi * a / sqrt(3.1415926535897932);
  }
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    for (int jj = i + 1; jj < ns; ++jj) {
      double d2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        double dd = wrap1(x[3 * jj + d], P.L[d]) - wrap1(x[3 * i + d], P.L[d]);
        dd = P.halfL[d] - fabs(P.halfL[d] - fabs(dd));
        d2 += dd * dd;
      }
      double r = sqrt(d2);
      e += -qsolute[i] * qsolute[jj] * erf(a * r) / r;
    }
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = NTHR / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// uvengy(0) = self_recp + realcal_self + residual_self (engproc.F90:701-732)
static __global__ void k_u0(Params P, int n, int ngrp, const double *__restrict__ selfp,
                     const double *__restrict__ selfreal, int per_item, double *uv) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double s = 0.0;
  for (int g = 0; g < ngrp; ++g) s += selfp[(size_t)b * ngrp + g];
  double pairep = per_item ? selfreal[b] : selfreal[0];
  double epcl = 3.1415926535897932 * P.solute_charge * P.solute_charge / P.screen / P.screen / P.volume;
  epcl = epcl / 2.0;
  double res = 0.0 - epcl;
  uv[(size_t)b * (P.slvmax + 1)] = s + pairep + res;
}

// --------------------------------------------------------------------------
// update_histogram (engproc.F90:758-856): bins of every molecule of
// configuration b (getiduv / binsearch, engproc.F90:927-1020), the
// instantaneous histogram, per-species sums and extrema.
// --------------------------------------------------------------------------
__device__ __forceinline__ void set_error(int *status, int code, int b, int j, double v) {
  if (atomicCAS(&status[0], 0, code) == 0) {
    status[1] = b;
    status[2] = j;
    *reinterpret_cast<double *>(&status[4]) = v;
  }
}

// generator streams of attempt a of a reference-structure insertion
__device__ __forceinline__ uint64_t att_seed(uint64_t seed, int a, uint64_t salt) {
  return seed ^ mix64((uint64_t)a * 0x9E3779B97F4A7C15ULL + salt);
}
constexpr uint64_t SALT_POS = 0x243F6A8885A308D3ULL, SALT_ROT = 0x13198A2E03707344ULL, SALT_MIX = 0xA4093822299F31D0ULL;
constexpr int REF_ATTEMPTS = 4096;

// Test-particle pose against a reference structure (insorigin REFSTR,
// insposition RMSD or GAUSS; instslt, insertion.F90:84-99, 198-230, 236-288,
// 569-595): from the solute fitted onto the frame's reference (ref[0..3ns)),
// a random shift within upreg + 3 * (largest COM distance) (RMSD) or from the
// uniform / Gaussian mixture (GAUSS, which divides the weight by the mixture
// density at every attempt), the orientation, and acceptance when the RMSD
// to the best-fit reference solute (ref[3ns..6ns), weights refw) lies in
// [lwreg, upreg]; else another attempt.  wsol[b] receives the weight (GAUSS).
static __global__ void k_pose_refstr(Params P, int n, const uint64_t *__restrict__ args, uint64_t offset,
                              const double *__restrict__ mass, const double *__restrict__ cellm,
                              const double *__restrict__ ref, const double *__restrict__ refw, double *spos,
                              double *wsol, int *status) {
#pragma clang fp contract(off)
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int ns = P.ns;
  const uint64_t seed = args[0];
  const uint64_t index = args[1] + offset + (uint64_t)b;
  double *out = spos + (size_t)b * ns * 3;
  const double *fitted = ref, *best = ref + 3 * ns;
  double weight = 1.0, c[3];
  bool ok = false;
  for (int a = 0; a < REF_ATTEMPTS && !ok; ++a) {
    const uint64_t sp = att_seed(seed, a, SALT_POS), sr = att_seed(seed, a, SALT_ROT);
    for (int i = 0; i < 3 * ns; ++i) out[i] = fitted[i];
    double com[3];
    if (P.ins_pos == 5) {
      com_dev(ns, out, mass, c);
      double maxdis = 0.0;
      for (int j = 0; j < ns; ++j) {
        double d0 = out[3 * j] - c[0], d1 = out[3 * j + 1] - c[1], d2 = out[3 * j + 2] - c[2];
        double dst = d0 * d0 + d1 * d1 + d2 * d2;
        if (dst > maxdis) maxdis = dst;
      }
      maxdis = sqrt(maxdis);
      const double movmax = P.upreg + 3.0 * maxdis;
      for (int i = 0; i < 3; ++i) com[i] = movmax * (2.0 * rng_uniform(sp, index, (uint32_t)i) - 1.0);
    } else {
      // uniform_gauss_mixture: uniform or Gaussian with probability 1/2 each
      double sc[3];
      const bool uni = rng_uniform(att_seed(seed, a, SALT_MIX), index, 0) < 0.5;
      uint32_t d = 0;
      for (int i = 0; i < 3; ++i) {
        if (uni) {
          sc[i] = (2.0 * rng_uniform(sp, index, d++) - 1.0) * P.gl[i];
        } else {
          double r = 0.0;
          for (int t = 0; t < 40; ++t) {
            double r1 = 1.0 - rng_uniform(sp, index, d++);
            double r2 = rng_uniform(sp, index, d++);
            r = sqrt(-2.0 * log(r1)) * cos(2.0 * 3.1415926535897932 * r2) / sqrt(2.0);
            if (fabs(r) < P.gl[i]) break;
          }
          sc[i] = r;
        }
      }
      double sq = sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2];
      weight = weight / (0.5 * 1.0 / P.gz0 + (1.0 - 0.5) * exp(-sq) / P.gz1);
      double u[3];
      for (int i = 0; i < 3; ++i) u[i] = sc[i] / (P.gl[i] * 2.0);
      for (int i = 0; i < 3; ++i) com[i] = cellm[i] * u[0] + cellm[3 + i] * u[1] + cellm[6 + i] * u[2];
    }
    for (int j = 0; j < ns; ++j)
      for (int i = 0; i < 3; ++i) out[3 * j + i] = out[3 * j + i] + com[i];
    if (P.ins_rot == 0) {
      double q[4] = {1.0, 0.0, 0.0, 0.0};
      for (int t = 0; t < 63; ++t) {
        double v[4];
        for (int i = 0; i < 4; ++i) v[i] = rng_uniform(sr, index, (uint32_t)(4 * t + i)) * 2.0 - 1.0;
        double s2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
        if (s2 < 1.0) {
          double nrm = sqrt(s2);
          for (int i = 0; i < 4; ++i) q[i] = v[i] / nrm;
          break;
        }
      }
      com_dev(ns, out, mass, c);
      double cj[4] = {q[0], -q[1], -q[2], -q[3]};
      for (int j = 0; j < ns; ++j) {
        double q2[4] = {0.0, out[3 * j] - c[0], out[3 * j + 1] - c[1], out[3 * j + 2] - c[2]};
        double t4[4], r2[4];
        qprod_dev(q, q2, t4);
        qprod_dev(t4, cj, r2);
        for (int i = 0; i < 3; ++i) out[3 * j + i] = r2[1 + i] + c[i];
      }
    }
    // rmsd_nofit (bestfit.F90:252-268) against the best-fit reference solute
    double wsum = 0.0, disp = 0.0;
    for (int j = 0; j < ns; ++j) wsum = wsum + refw[j];
    for (int j = 0; j < ns; ++j) {
      double d0 = out[3 * j] - best[3 * j], d1 = out[3 * j + 1] - best[3 * j + 1], d2 = out[3 * j + 2] - best[3 * j + 2];
      disp = disp + refw[j] * (d0 * d0 + d1 * d1 + d2 * d2);
    }
    const double rmsd = sqrt(disp / wsum);
    ok = !(P.lwreg > rmsd || rmsd > P.upreg);
  }
  if (!ok) set_error(status, ERM_E_INS_LIMIT, (int)index, -1, 0.0);
  if (P.ins_pos == 6) wsol[b] = weight;
}

static __global__ __launch_bounds__(NTHR) void k_bin(Params P, const double *__restrict__ uv,
                                              const double *__restrict__ edges,
                                              const int *__restrict__ spec_off, const int *__restrict__ uvmax,
                                              const int *__restrict__ uvsoft, const int *__restrict__ mol_spec,
                                              const int *__restrict__ slt_j, double *hist, double *flceng,
                                              unsigned long long *mnmx, int *status, int first) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *e = reinterpret_cast<double *>(lds_raw);
  // per-thread, per-species energy sums [numslv][blockDim.x] (each thread
  // adds only to its own slots; they are summed in thread order below, so
  // flceng does not depend on the order the threads run in)
  double *part = e + P.ermax;
  int *h = reinterpret_cast<int *>(part + (size_t)P.numslv * blockDim.x);
  __shared__ unsigned long long mn[MAXSPEC], mx[MAXSPEC];
  const int b = blockIdx.x;
  for (int t = threadIdx.x; t < P.ermax; t += blockDim.x) {
    e[t] = edges[t];
    h[t] = 0;
  }
  for (int t = 0; t < P.numslv; ++t) part[(size_t)t * blockDim.x + threadIdx.x] = 0.0;
  if (threadIdx.x < MAXSPEC) {
    mn[threadIdx.x] = ~0ULL;
    mx[threadIdx.x] = 0ULL;
  }
  __syncthreads();
  const double *u = uv + (size_t)b * (P.slvmax + 1) + 1;
  const int sj = slt_j[b];
  // runs of equal species and equal bins along a thread's molecules are
  // summed in registers and added to LDS once per run (the far molecules of
  // a configuration mostly share a species and a bin near zero)
  int rs = -1, rb = -1, rc = 0;
  double rsum = 0.0;
  unsigned long long rmn = ~0ULL, rmx = 0ULL;
  auto flush_spec = [&]() {
    if (rs >= 0) {
      part[(size_t)rs * blockDim.x + threadIdx.x] += rsum;
      atomicMin(&mn[rs], rmn);
      atomicMax(&mx[rs], rmx);
    }
  };
  // a thread's molecules j = tid, tid + blockDim.x, ... in order, their
  // energies and species loaded four at a time (all in flight together)
  constexpr int JB = 4;
  for (int j0 = threadIdx.x; j0 < P.slvmax; j0 += JB * (int)blockDim.x) {
    double vb[JB];
    int sb[JB];
#pragma unroll
    for (int k = 0; k < JB; ++k) {
      const int jk = min(j0 + k * (int)blockDim.x, P.slvmax - 1);
      vb[k] = u[jk];
      sb[k] = mol_spec[jk];
    }
#pragma unroll
    for (int k = 0; k < JB; ++k) {
      const int j = j0 + k * (int)blockDim.x;
      if (j >= P.slvmax) break;
      if (j == sj) continue;
      int pti = sb[k];
      if (pti <= 0 || pti > P.numslv) {
        set_error(status, 8, first + b, j, 0.0);
        continue;
      }
      double v = vb[k];
      int off = spec_off[pti - 1], n = uvmax[pti - 1];
      int lo = 0, hi = n;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (e[off + mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo <= 0) {
        set_error(status, 1, first + b, j, v);
        continue;
      }
      if (P.slttype == 1 && lo > uvsoft[pti - 1]) {
        set_error(status, 2, first + b, j, v);
        continue;
      }
      if (lo == n && v < 1.0e3 && !P.force) set_error(status, 3, first + b, j, v);
      const int bin = off + lo - 1;
      if (bin != rb) {
        if (rc) atomicAdd(&h[rb], rc);
        rb = bin;
        rc = 0;
      }
      ++rc;
      if (pti - 1 != rs) {
        flush_spec();
        rs = pti - 1;
        rsum = 0.0;
        rmn = ~0ULL;
        rmx = 0ULL;
      }
      rsum += v;
      const unsigned long long o = dbl_to_ord(v);
      rmn = (o < rmn) ? o : rmn;
      rmx = (o > rmx) ? o : rmx;
    }
  }
  if (rc) atomicAdd(&h[rb], rc);
  flush_spec();
  __syncthreads();
  for (int t = threadIdx.x; t < P.ermax; t += blockDim.x) hist[(size_t)b * P.ermax + t] = (double)h[t];
  if (threadIdx.x < P.numslv) {
    double sum = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) sum += part[(size_t)threadIdx.x * blockDim.x + t];
    flceng[(size_t)b * MAXSPEC + threadIdx.x] = sum;
    atomicMin(&mnmx[2 * (threadIdx.x + 1)], mn[threadIdx.x]);
    atomicMax(&mnmx[2 * (threadIdx.x + 1) + 1], mx[threadIdx.x]);
  }
}

// Scalar part of update_histogram over the configurations of a batch
// (engproc.F90:776-808): each thread forms the weight of one configuration
// (wgtslf, volcorrect) and its self-energy bin; thread 0 then adds them in
// configuration order, as the reference's sequential loop does.
// scal: [0] engnorm [1] engsmpl [2] avslf [3] voffset [4] voffset set
static __global__ __launch_bounds__(NTHR) void k_scalars(Params P, int n, const double *__restrict__ uv,
                                                  const double *__restrict__ wsol,
                                                  const double *__restrict__ escrd, double *scal, double *eself,
                                                  unsigned long long *mnmx, double *wout, int *status, int first) {
  __shared__ double wsh[NTHR], twsh[NTHR];
  __shared__ int ish[NTHR];
  __shared__ double voff;
  __shared__ unsigned long long rmn[NTHR], rmx[NTHR];
  unsigned long long lmn = ~0ULL, lmx = 0ULL;  // this thread's extrema (ordered)
  if (threadIdx.x == 0) {
    if (P.wgtslf && scal[4] == 0.0 && n > 0) {
      scal[3] = uv[0];
      scal[4] = 1.0;
    }
    voff = scal[3];
  }
  __syncthreads();
  for (int b0 = 0; b0 < n; b0 += NTHR) {
    const int b = b0 + threadIdx.x;
    if (b < n) {
      double u0 = uv[(size_t)b * (P.slvmax + 1)];
      double engnmfc = 1.0;
      if (P.wgtslf) {
        double f = u0 - voff;
        engnmfc = (P.slttype == 1) ? exp(f / P.temp) : exp(-f / P.temp);
      }
      double tw = P.stat_weight_system * (wsol ? wsol[b] : 1.0);
      if (P.estype == 2) {
        tw = tw * P.volume;
        if (P.cltype != 0) {
          double factor = 3.1415926535897932 * (P.total_charge * P.total_charge) / (P.screen * P.screen) /
                          P.volume / 2.0;
          tw = tw * exp(factor / P.temp);
        }
      }
      engnmfc = engnmfc * tw;
      int bin = -1;
      if (P.selfcal) {
        int lo = 0, hi = P.esmax;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (escrd[mid] < u0)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (lo <= 0) {
          set_error(status, 1, first + b, -1, u0);
        } else {
          if (lo == P.esmax && u0 < 1.0e3 && !P.force) set_error(status, 3, first + b, -1, u0);
          bin = lo - 1;
        }
      }
      wsh[threadIdx.x] = engnmfc;
      twsh[threadIdx.x] = tw;
      ish[threadIdx.x] = bin;
      wout[b] = engnmfc;
      const unsigned long long o = dbl_to_ord(u0);
      lmn = (o < lmn) ? o : lmn;
      lmx = (o > lmx) ? o : lmx;
    }
    __syncthreads();
    const int m = min(NTHR, n - b0);
    if (threadIdx.x == 0) {
      double en = scal[0], es = scal[1], av = scal[2];
      for (int i = 0; i < m; ++i) {
        en = en + wsh[i];
        es = es + 1.0;
        av = av + twsh[i];
      }
      scal[0] = en;
      scal[1] = es;
      scal[2] = av;
    }
    // self-energy bins, one thread per bin, each adding its configurations
    // in configuration order (the same sums as the sequential loop)
    if (P.selfcal)
      for (int e = threadIdx.x; e < P.esmax; e += NTHR) {
        double v = eself[e];
        bool hit = false;
        for (int i = 0; i < m; ++i)
          if (ish[i] == e) {
            v = v + wsh[i];
            hit = true;
          }
        if (hit) eself[e] = v;
      }
    __syncthreads();
  }
  rmn[threadIdx.x] = lmn;
  rmx[threadIdx.x] = lmx;
  __syncthreads();
  for (int st = NTHR / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      rmn[threadIdx.x] = (rmn[threadIdx.x + st] < rmn[threadIdx.x]) ? rmn[threadIdx.x + st] : rmn[threadIdx.x];
      rmx[threadIdx.x] = (rmx[threadIdx.x + st] > rmx[threadIdx.x]) ? rmx[threadIdx.x + st] : rmx[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && n > 0) {
    mnmx[0] = (rmn[0] < mnmx[0]) ? rmn[0] : mnmx[0];
    mnmx[1] = (rmx[0] > mnmx[1]) ? rmx[0] : mnmx[1];
  }
}

// edens(i) += engnmfc * insdst(i) (engproc.F90:836-840).  A workgroup takes
// 32 bins; its 8 waves' halves sum 8 consecutive chunks of the
// configurations (in order within a chunk, empty bins skipped as the
// reference skips them), then the chunk sums are added in chunk order.
static __global__ __launch_bounds__(256) void k_edens(int n, int ermax, const double *__restrict__ hist,
                                               const double *__restrict__ w, double *edens) {
  __shared__ double part[8][32];
  const int lb = threadIdx.x & 31, ch = threadIdx.x >> 5;
  const int i = blockIdx.x * 32 + lb;
  const int per = (n + 7) / 8, b0 = min(n, ch * per), b1 = min(n, b0 + per);
  double e = 0.0;
  if (i < ermax) {
    for (int bb = b0; bb < b1; bb += 16) {
      double h[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) h[u] = (bb + u < b1) ? hist[(size_t)(bb + u) * ermax + i] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (h[u] != 0.0) e = e + w[bb + u] * h[u];
    }
  }
  part[ch][lb] = e;
  __syncthreads();
  if (ch == 0 && i < ermax) {
    double t = edens[i];
    for (int k = 0; k < 8; ++k) t = t + part[k][lb];
    edens[i] = t;
  }
}

// intermediate state (extensions/Intermediate/engproc.F90:996-1009): the
// weight of configuration b if its instantaneous histogram touches no bin
// that is empty in the reference solvent, else 0; one workgroup per b
static __global__ __launch_bounds__(256) void k_intm_weight(int ermax, const double *__restrict__ hist,
                                                     const int *__restrict__ zerodns, const double *__restrict__ w,
                                                     double *wi) {
  __shared__ int hit;
  const int b = blockIdx.x;
  if (threadIdx.x == 0) hit = 0;
  __syncthreads();
  const double *h = hist + (size_t)b * ermax;
  int my = 0;
  for (int i = threadIdx.x; i < ermax; i += blockDim.x) my |= (h[i] > 0.0 && zerodns[i]) ? 1 : 0;
  if (my) atomicOr(&hit, 1);
  __syncthreads();
  if (threadIdx.x == 0) wi[b] = hit ? 0.0 : w[b];
}

// soln_total += engnmfc; soln_intm and norm_intm += the intermediate weight
// (adding the 0 of an excluded configuration leaves the sum unchanged), in
// configuration order; sc = {norm_intm, soln_intm, soln_total}
static __global__ void k_intm_scal(int n, const double *__restrict__ w, const double *__restrict__ wi, double *sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double a = sc[0], i = sc[1], t = sc[2];
  for (int b = 0; b < n; ++b) {
    t = t + w[b];
    i = i + wi[b];
    a = a + wi[b];
  }
  sc[0] = a;
  sc[1] = i;
  sc[2] = t;
}

// slnuv(:) += flceng(:, cntdst) * engnmfc (engproc.F90:832-834)
static __global__ void k_slnuv(int n, int numslv, const double *__restrict__ flceng, const double *__restrict__ w,
                        double *slnuv) {
  int s = threadIdx.x;
  if (s >= numslv) return;
  double v = slnuv[s + 1];
  for (int b = 0; b < n; ++b) v = v + flceng[(size_t)b * MAXSPEC + s] * w[b];
  slnuv[s + 1] = v;
}

// --------------------------------------------------------------------------
// Correlation matrix on FP64 MFMA: ecorr(p, i) += sum_b (w_b h_b[i]) h_b[p]
// (engproc.F90:841-853), one 16 x 16 tile per wave, v_mfma_f64_16x16x4f64
// over the batch in steps of 4 configurations.  ecorr column-major
// (element (p, i) at p + i * ermax).  The matrix is symmetric: the
// workgroups take the tiles on and above the diagonal (p tile <= i tile),
// and an off-diagonal tile also writes its transpose.
// --------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

static __global__ __launch_bounds__(256) void k_corr(int n, int ermax, const double *__restrict__ hist,
                                              const double *__restrict__ w, double *ecorr) {
  __shared__ dbl4 part[3][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int tile = blockIdx.x;
  // tile = ti (ti + 1) / 2 + tp with tp <= ti
  int ti = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > tile) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
  const int tp = tile - ti * (ti + 1) / 2;  // tile rows p, columns i
  const int p0 = tp * 16, i0 = ti * 16;
  const int r = lane & 15, kq = lane >> 4;
  // the four waves of the workgroup take four consecutive quarters of the
  // configurations (a quarter of the dependent MFMA chain each); the
  // quarter sums are added in order at the end.  D layout of the f64
  // 16x16x4 MFMA: register v of a lane holds row (lane >> 4) + 4 v, column
  // lane & 15
  const int nq4 = (((n + 3) / 4 + 3) / 4) * 4;  // configurations per quarter, a multiple of 4
  const int bq0 = min(n, q * nq4), bq1 = min(n, bq0 + nq4);
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  // U steps of 4 configurations per round: all 2U operand loads issued
  // before the U dependent MFMAs
  constexpr int U = 8;
  const bool pin = p0 + r < ermax, iin = i0 + r < ermax;
  int b0 = bq0;
  for (; b0 + 4 * U <= bq1; b0 += 4 * U) {
    double a[U], bb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + 4 * u + kq;
      a[u] = pin ? hist[(size_t)b * ermax + p0 + r] : 0.0;
      bb[u] = iin ? w[b] * hist[(size_t)b * ermax + i0 + r] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bb[u], acc, 0, 0, 0);
  }
  for (; b0 < bq1; b0 += 4) {
    int b = b0 + kq;
    double a = 0.0, bb = 0.0;
    if (b < bq1) {
      if (pin) a = hist[(size_t)b * ermax + p0 + r];
      if (iin) bb = w[b] * hist[(size_t)b * ermax + i0 + r];
    }
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
  }
  if (q > 0) part[q - 1][lane] = acc;
  __syncthreads();
  if (q != 0) return;
  for (int k = 0; k < 3; ++k) {
    const dbl4 o = part[k][lane];
    for (int v = 0; v < 4; ++v) acc[v] = acc[v] + o[v];
  }
  for (int v = 0; v < 4; ++v) {
    int pr = p0 + kq + 4 * v, ic = i0 + r;
    if (pr < ermax && ic < ermax) {
      const double e = ecorr[(size_t)ic * ermax + pr] + acc[v];
      ecorr[(size_t)ic * ermax + pr] = e;
      if (tp != ti) ecorr[(size_t)pr * ermax + ic] = e;
    }
  }
}

// uvengy in tagpt order from the sorted order
// the inverse of k_unsort for energies given in tagpt order, and each
// configuration's solute as a sorted index (kof: tagpt index of a molecule,
// -1 for none; jof: sorted index of a tagpt index; solutes NULL: none)
static __global__ void k_sort_uv(int n, int slvmax, const int *__restrict__ mol_k, const int *__restrict__ jof,
                                 const int *__restrict__ kof, const int *__restrict__ solutes,
                                 const double *__restrict__ in, double *uv, int *slt_j) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (b >= n) return;
  const double *u = in + (size_t)b * (slvmax + 1);
  double *o = uv + (size_t)b * (slvmax + 1);
  if (j == 0) {
    o[0] = u[0];
    int sj = -1;
    if (solutes) {
      const int kk = kof[solutes[b]];
      sj = (kk >= 0) ? jof[kk] : -1;
    }
    slt_j[b] = sj;
  }
  if (j < slvmax) o[1 + j] = u[1 + mol_k[j]];
}

static __global__ void k_unsort(int n, int slvmax, const int *__restrict__ mol_k, const double *__restrict__ uv,
                         double *out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (b >= n) return;
  const double *u = uv + (size_t)b * (slvmax + 1);
  double *o = out + (size_t)b * (slvmax + 1);
  if (j == 0) o[0] = u[0];
  if (j < slvmax) o[1 + mol_k[j]] = u[1 + j];
}

}  // namespace erm
