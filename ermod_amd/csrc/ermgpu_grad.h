// ermgpu_grad.h -- adjoint kernels: gradient of the solute-solvent energy
// u = sum_{k>=1} uvengy(k) with respect to the solute's per-site charge q_i,
// LJ epsilon_i and sigma_i.
//
//  real space   per (configuration, solute site): sums over the candidate
//               solvent sites of the pair kernel
//                 dU/dq_i      += q_v erfc(alpha r) / r
//                 U_LJ,i       += rep - att          (dU/deps_i = U_LJ,i / 2 eps_i)
//                 dU/dsigma_i  += (6 rep - 3 att) f_iv
//               where every LJ variant of realcal.F90:708-776 is rep - att
//               with rep proportional to sigma^12 and att to sigma^6, and
//               f_iv = d(sigma_iv^2)/d(sigma_i) / sigma_iv^2.
//  reciprocal   the PME energy is the symmetric bilinear form
//               q_slt^T W^T Phi W q_slv (recpcal.F90:342-502), so dU/dq_i is
//               the solvent's PME potential at site i: once per frame the
//               solvent charges are spread, transformed (register FFTs),
//               multiplied by the Green's function and transformed back; per
//               configuration each solute site interpolates that grid.
//  residual     residual_ene (engproc.F90:859-869): -pi Q_slt Q_solv / (alpha^2 V).
#pragma once

#include "ermgpu_kern.h"

namespace erm {

// rep and att of one LJ pair (U = rep - att) for the four switches of
// realcal.F90:708-776; rep scales with sigma^12, att with sigma^6
template <typename T>
__device__ __forceinline__ void lj_parts_t(const Params &P, T eps, T sig2, T dist, T &rep, T &att) {
  T invr2 = sig2 / dist;
  T invr6 = invr2 * invr2 * invr2;
  T sg6 = sig2 * sig2 * sig2;
  const bool low = dist <= (T)P.lw2;
  if (P.ljswitch <= 1) {
    T sw = (T)1;
    if (!low) {
      if (P.ljswitch == 0) {
        T up2 = (T)P.up2, lw2 = (T)P.lw2;
        T d = dist - up2, w = up2 - lw2;
        sw = ((T)2 * dist + up2 - (T)3 * lw2) * (d * d) / (w * w * w);
      } else {
        T x = (sqrt(dist) - (T)P.lwljcut) / (T)(P.upljcut - P.lwljcut);
        T x3 = x * x * x;
        sw = (T)1 - (T)10 * x3 + (T)15 * (x3 * x) - (T)6 * (x3 * x * x);
      }
    }
    rep = (T)4 * eps * invr6 * invr6 * sw;
    att = (T)4 * eps * invr6 * sw;
    return;
  }
  if (P.ljswitch == 2) {
    if (low) {
      rep = (T)4 * eps * (invr6 * invr6 - sg6 * sg6 / (T)(P.lw6 * P.up6));
      att = (T)4 * eps * (invr6 - sg6 / (T)(P.lw3 * P.up3));
    } else {
      T invr3 = sqrt(invr6), sg3 = sqrt(sg6);
      T a = invr6 - sg6 / (T)P.up6, c = invr3 - sg3 / (T)P.up3;
      rep = (T)4 * eps * (T)(P.up6 / (P.up6 - P.lw6)) * (a * a);
      att = (T)4 * eps * (T)(P.up3 / (P.up3 - P.lw3)) * (c * c);
    }
    return;
  }
  T a = (T)P.repC, b = (T)P.attC;
  if (!low) {
    T sw = sqrt(dist) - (T)P.lwljcut;
    T sw3 = sw * sw * sw, sw4 = sw3 * sw;
    a = (T)P.repA * sw3 + (T)P.repB * sw4 + (T)P.repC;
    b = (T)P.attA * sw3 + (T)P.attB * sw4 + (T)P.attC;
  }
  rep = (T)4 * eps * (invr6 * invr6 - sg6 * sg6 * a);
  att = (T)4 * eps * (invr6 - sg6 * b);
}

template <typename T> __device__ __forceinline__ T erfc_over_r(T r, T alpha);
template <> __device__ __forceinline__ float erfc_over_r<float>(float r, float alpha) { return erfcf(alpha * r) / r; }
template <> __device__ __forceinline__ double erfc_over_r<double>(double r, double alpha) {
  return (1.0 - erf(alpha * r)) / r;
}

// Real-space adjoint.  Grid (n, nsplit); every workgroup of configuration b
// builds the pair kernel's candidate segments, then its threads take items
// (solute site s, slice): a slice walks every (nsplit * nsl)-th candidate
// molecule.  Per-site sums go through LDS in slice order and are written as
// part[(b * nsplit + y)][4][ns] = {u, dU/dq, U_LJ, dU/dsigma}: deterministic.
// LDS: 7 ns T (solute offsets and charges) + 3 ns nvt T (eps_iv, sigma_iv^2,
// f_iv) + segments + max(NTHR, ns) * 4 doubles.
template <typename T>
__global__ __launch_bounds__(NTHR) void k_pair_grad(Params P, const double *__restrict__ spos,
                                                    const double *__restrict__ sref, const T *__restrict__ gq,
                                                    const T *__restrict__ gtab, const int *__restrict__ cell_start,
                                                    const int *__restrict__ mol_first,
                                                    const int *__restrict__ mol_ns, const double *__restrict__ sx,
                                                    const double *__restrict__ sy, const double *__restrict__ sz,
                                                    const T *__restrict__ sq, const int *__restrict__ sct,
                                                    const int *__restrict__ slt_j,
                                                    const double *__restrict__ rm_ord_p, double *part) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int ns = P.ns, nvt = P.nvt;
  const int nsl = max(1, (int)blockDim.x / ns);
  double *red = reinterpret_cast<double *>(lds_raw);  // [nsl][4][ns]
  const int nred = 4 * nsl * ns;
  T *ux = reinterpret_cast<T *>(red + nred);
  T *uy = ux + ns;
  T *uz = uy + ns;
  T *uq = uz + ns;
  T *ulo = uq + ns;                  // [ns][3]
  T *tab = ulo + 3 * ns;             // [ns][nvt][3]
  int *seg_beg = reinterpret_cast<int *>(tab + 3 * ns * nvt);
  int *seg_pre = seg_beg + MAXSEG;
  const int b = blockIdx.x;
  const double C0 = sref[4 * b], C1 = sref[4 * b + 1], C2v = sref[4 * b + 2];
  const double Rs = sref[4 * b + 3];
  const double *xs = spos + (size_t)b * ns * 3;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    ux[s] = (T)(xs[3 * s] - C0);
    uy[s] = (T)(xs[3 * s + 1] - C1);
    uz[s] = (T)(xs[3 * s + 2] - C2v);
    uq[s] = gq[s];
    ulo[3 * s] = (T)((xs[3 * s] - C0) - (double)ux[s]);
    ulo[3 * s + 1] = (T)((xs[3 * s + 1] - C1) - (double)uy[s]);
    ulo[3 * s + 2] = (T)((xs[3 * s + 2] - C2v) - (double)uz[s]);
  }
  for (int t = threadIdx.x; t < 3 * ns * nvt; t += blockDim.x) tab[t] = gtab[t];
  for (int t = threadIdx.x; t < nred; t += blockDim.x) red[t] = 0.0;
  const double Rm = ord_to_dbl(*(const unsigned long long *)rm_ord_p);
  const double R = P.rcmax + Rs + Rm + 1e-6 * (1.0 + P.rcmax);
  const double Cw[3] = {wrap1(C0, P.L[0]), wrap1(C1, P.L[1]), wrap1(C2v, P.L[2])};
  const int nsg = pillar_segments(P, Cw, R, cell_start, seg_beg, seg_pre);
  const int total = seg_pre[nsg];
  const int sj = slt_j[b];
  const T L0 = (T)P.L[0], L1 = (T)P.L[1], L2 = (T)P.L[2];
  const T h0 = (T)P.halfL[0], h1 = (T)P.halfL[1], h2 = (T)P.halfL[2];
  const T up2 = (T)P.up2, el2 = (T)P.el2, cut2 = (T)P.cut2, alpha = (T)P.screen;
  const int nslices = gridDim.y * nsl;
  for (int it = threadIdx.x; it < ns * nsl; it += blockDim.x) {
    const int s = it % ns, sl = it / ns;
    const T usx = ux[s], usy = uy[s], usz = uz[s], qs = uq[s];
    const T lsx = ulo[3 * s], lsy = ulo[3 * s + 1], lsz = ulo[3 * s + 2];
    double a_u = 0.0, a_q = 0.0, a_lj = 0.0, a_sg = 0.0;
    for (int idx = blockIdx.y * nsl + sl; idx < total; idx += nslices) {
      int lo_s = 0, hi_s = nsg;
      while (hi_s - lo_s > 1) {
        int mid = (lo_s + hi_s) >> 1;
        if (seg_pre[mid] <= idx)
          lo_s = mid;
        else
          hi_s = mid;
      }
      const int j = seg_beg[lo_s] + (idx - seg_pre[lo_s]);
      if (j == sj) continue;
      T e_u = 0, e_q = 0, e_lj = 0, e_sg = 0;
      const int t0 = mol_first[j], nsm = mol_ns[j];
      for (int t = t0; t < t0 + nsm; ++t) {
        double dx = sx[t] - C0, dy = sy[t] - C1, dz = sz[t] - C2v;
        dx -= P.L[0] * rint(dx / P.L[0]);
        dy -= P.L[1] * rint(dy / P.L[1]);
        dz -= P.L[2] * rint(dz / P.L[2]);
        const T vx = (T)dx, vy = (T)dy, vz = (T)dz;
        const T lx = (T)(dx - (double)vx), ly = (T)(dy - (double)vy), lz = (T)(dz - (double)vz);
        T ax = fabs(((vx - usx) + lx) - lsx);
        T ay = fabs(((vy - usy) + ly) - lsy);
        T az = fabs(((vz - usz) + lz) - lsz);
        ax = (ax > h0) ? L0 - ax : ax;
        ay = (ay > h1) ? L1 - ay : ay;
        az = (az > h2) ? L2 - az : az;
        const T r2 = ax * ax + ay * ay + az * az;
        if (r2 <= cut2) {
          const T *tb = tab + (s * nvt + sct[t]) * 3;
          if (tb[0] > (T)0 && r2 <= up2) {
            T rep, att;
            lj_parts_t<T>(P, tb[0], tb[1], r2, rep, att);
            e_u += rep - att;
            e_lj += rep - att;
            e_sg += ((T)6 * rep - (T)3 * att) * tb[2];
          }
          if (r2 <= el2) {
            T c = sq[t] * erfc_over_r<T>(sqrt(r2), alpha);
            e_u += qs * c;
            e_q += c;
          }
        }
      }
      a_u += (double)e_u;
      a_q += (double)e_q;
      a_lj += (double)e_lj;
      a_sg += (double)e_sg;
    }
    red[(sl * 4 + 0) * ns + s] = a_u;
    red[(sl * 4 + 1) * ns + s] = a_q;
    red[(sl * 4 + 2) * ns + s] = a_lj;
    red[(sl * 4 + 3) * ns + s] = a_sg;
  }
  __syncthreads();
  double *out = part + ((size_t)b * gridDim.y + blockIdx.y) * 4 * ns;
  for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) {
    double v = 0.0;
    for (int sl = 0; sl < nsl; ++sl) v += red[sl * 4 * ns + t];
    out[t] = v;
  }
}

// ---- per-frame solvent potential grid -----------------------------------
template <typename T> __global__ void k_fill(T *p, size_t n, T v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// spreads the sites of every sorted physical molecule except `excl` onto the
// real grid grid[z][y][x] (recpcal.F90:357-373 with the solvent's charges)
template <typename T>
__global__ void k_spread_slv(Params P, int excl, const int *__restrict__ mol_first, const int *__restrict__ mol_ns,
                             const int *__restrict__ sg, const T *__restrict__ sw, const T *__restrict__ sq,
                             T *grid) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.slvmax || j == excl) return;
  const int p = P.p, K1 = P.K1, K2 = P.K2, K3 = P.K3;
  for (int t = mol_first[j]; t < mol_first[j] + mol_ns[j]; ++t) {
    const int gp = sg[t];
    const int g1 = gp & 1023, g2 = (gp >> 10) & 1023, g3 = (gp >> 20) & 1023;
    const T *w = sw + (size_t)t * 3 * p;
    const T q = sq[t];
    for (int c3 = 0; c3 < p; ++c3) {
      int z = g3 - c3;
      z += (z < 0) ? K3 : 0;
      const T f3 = q * w[2 * p + c3];
      for (int c2 = 0; c2 < p; ++c2) {
        int y = g2 - c2;
        y += (y < 0) ? K2 : 0;
        const T f2 = f3 * w[p + c2];
        T *row = grid + ((size_t)z * K2 + y) * K1;
        for (int c1 = 0; c1 < p; ++c1) {
          int x = g1 - c1;
          x += (x < 0) ? K1 : 0;
          atomicAdd(row + x, f2 * w[c1]);
        }
      }
    }
  }
}

// forward x transform of real rows, two rows per complex FFT: thread per
// (z, row pair); Hk[z][y][m1], m1 < K1/2 + 1
template <typename T, int K1>
__global__ __launch_bounds__(NTHR) void k_xrow_r2c(Params P, const T *__restrict__ grid,
                                                   const typename C2<T>::type *__restrict__ tw1,
                                                   typename C2<T>::type *Hk) {
  using V = typename C2<T>::type;
  constexpr int H = K1 / 2 + 1;
  const int K2 = P.K2, K3 = P.K3, npair = (K2 + 1) / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K3 * npair) return;
  const int z = t / npair, y0 = 2 * (t - z * npair), y1 = y0 + 1;
  const bool has1 = y1 < K2;
  const T *r0 = grid + ((size_t)z * K2 + y0) * K1;
  V Z[K1];
#pragma unroll
  for (int x = 0; x < K1; ++x) {
    Z[x].x = r0[x];
    Z[x].y = has1 ? r0[K1 + x] : (T)0;
  }
  reg_fft<K1, -1>(Z, tw1);
  V *o0 = Hk + ((size_t)z * K2 + y0) * H;
#pragma unroll
  for (int m = 0; m < H; ++m) {
    const V a = Z[m], w = Z[(K1 - m) % K1];
    V A, Cc;
    A.x = (T)0.5 * (a.x + w.x);
    A.y = (T)0.5 * (a.y - w.y);
    Cc.x = (T)0.5 * (a.y + w.y);
    Cc.y = (T)-0.5 * (a.x - w.x);
    o0[m] = A;
    if (has1) o0[H + m] = Cc;
  }
}

// z columns of the half spectrum: forward FFT, Green's function, inverse FFT
template <typename T, int K3>
__global__ __launch_bounds__(NTHR) void k_zcol_green(Params P, const T *__restrict__ engfac,
                                                     const typename C2<T>::type *__restrict__ tw3,
                                                     typename C2<T>::type *Hk) {
  using V = typename C2<T>::type;
  const int ncol = P.H * P.K2;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  V X[K3];
#pragma unroll
  for (int z = 0; z < K3; ++z) X[z] = Hk[(size_t)z * ncol + c];
  reg_fft<K3, -1>(X, tw3);
#pragma unroll
  for (int z = 0; z < K3; ++z) {
    const T g = engfac[(size_t)z * ncol + c];
    X[z].x *= g;
    X[z].y *= g;
  }
  reg_fft<K3, +1>(X, tw3);
#pragma unroll
  for (int z = 0; z < K3; ++z) Hk[(size_t)z * ncol + c] = X[z];
}

// inverse x transform of Hermitian row pairs into real rows grid[z][y][x]
template <typename T, int K1>
__global__ __launch_bounds__(NTHR) void k_xrow_c2r(Params P, const typename C2<T>::type *__restrict__ Hk,
                                                   const typename C2<T>::type *__restrict__ tw1, T *grid) {
  using V = typename C2<T>::type;
  constexpr int H = K1 / 2 + 1;
  const int K2 = P.K2, K3 = P.K3, npair = (K2 + 1) / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K3 * npair) return;
  const int z = t / npair, y0 = 2 * (t - z * npair), y1 = y0 + 1;
  const bool has1 = y1 < K2;
  const V *r0 = Hk + ((size_t)z * K2 + y0) * H;
  V Z[K1];
#pragma unroll
  for (int x = 0; x < H; ++x) {
    V a = r0[x];
    V cc = has1 ? r0[H + x] : V{0, 0};
    if (x == 0 || x == K1 / 2) {
      a.y = 0;
      cc.y = 0;
    }
    Z[x].x = a.x - cc.y;
    Z[x].y = a.y + cc.x;
    if (x > 0 && x < K1 - x) {
      Z[K1 - x].x = a.x + cc.y;
      Z[K1 - x].y = -a.y + cc.x;
    }
  }
  reg_fft<K1, +1>(Z, tw1);
  T *d0 = grid + ((size_t)z * K2 + y0) * K1;
#pragma unroll
  for (int x = 0; x < K1; ++x) d0[x] = Z[x].x;
  if (has1) {
#pragma unroll
    for (int x = 0; x < K1; ++x) d0[K1 + x] = Z[x].y;
  }
}

// solvent potential at every solute site of every configuration
// (calc_spline_molecule, recpcal.F90:408-442, then the gather of
// recpcal.F90:451-502 against the solvent grid); thread per (b, s)
template <typename T>
__global__ void k_grad_recip(Params P, int n, const double *__restrict__ spos, const T *__restrict__ grid,
                             double *phi) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * P.ns) return;
  const double *x = spos + (size_t)i * 3;
  const int Ks[3] = {P.K1, P.K2, P.K3};
  const int p = P.p;
  double w[3][MAXP];
  int g[3];
  for (int d = 0; d < 3; ++d) {
    double f = P.invcl[d] * x[0] + P.invcl[3 + d] * x[1] + P.invcl[6 + d] * x[2];
    f = f - floor(f);
    double factor = f * (double)Ks[d];
    int rci = (factor >= 0.0 && factor <= (double)Ks[d]) ? (int)factor : 0;
    for (int k = 0; k < MAXP; ++k)
      w[d][k] = (k < p) ? spline_value_dev(P.coeff, p, factor - (double)(rci - k)) : 0.0;
    g[d] = rci % Ks[d];
  }
  double acc = 0.0;
  for (int c3 = 0; c3 < p; ++c3) {
    int z = g[2] - c3;
    z += (z < 0) ? P.K3 : 0;
    for (int c2 = 0; c2 < p; ++c2) {
      int y = g[1] - c2;
      y += (y < 0) ? P.K2 : 0;
      const T *row = grid + ((size_t)z * P.K2 + y) * P.K1;
      double rs = 0.0;
      for (int c1 = 0; c1 < p; ++c1) {
        int xx = g[0] - c1;
        xx += (xx < 0) ? P.K1 : 0;
        rs += w[0][c1] * (double)row[xx];
      }
      acc += w[2][c3] * w[1][c2] * rs;
    }
  }
  phi[i] = acc;
}

// per configuration: u_b = sum_s (u_real + q_s phi_s) + Q_slt * rf
static __global__ void k_grad_ub(Params P, int n, int nsplit, const double *__restrict__ part,
                          const double *__restrict__ phi, const double *__restrict__ qd, double rf, double qslt,
                          double *ub) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int ns = P.ns;
  double u = 0.0;
  for (int y = 0; y < nsplit; ++y) {
    const double *pp = part + ((size_t)b * nsplit + y) * 4 * ns;
    for (int s = 0; s < ns; ++s) u += pp[s];
  }
  for (int s = 0; s < ns; ++s) u += qd[s] * phi[(size_t)b * ns + s];
  ub[b] = u + qslt * rf;
}

// weighted batch sums: out = [sum w, sum w u, sum w du/dq (ns),
// sum w du/deps (ns), sum w du/dsigma (ns)]; thread per output, b in order
static __global__ void k_grad_reduce(Params P, int n, int nsplit, const double *__restrict__ part,
                              const double *__restrict__ phi, const double *__restrict__ ub,
                              const double *__restrict__ w, const double *__restrict__ epsd, double rf,
                              double *out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  const int ns = P.ns;
  if (o >= 2 + 3 * ns) return;
  double acc = 0.0;
  if (o == 0) {
    for (int b = 0; b < n; ++b) acc += w[b];
  } else if (o == 1) {
    for (int b = 0; b < n; ++b) acc += w[b] * ub[b];
  } else {
    const int c = (o - 2) / ns, s = (o - 2) - c * ns;  // 0 q, 1 eps, 2 sigma
    const int comp = (c == 0) ? 1 : (c == 1) ? 2 : 3;
    for (int b = 0; b < n; ++b) {
      double v = 0.0;
      for (int y = 0; y < nsplit; ++y) v += part[(((size_t)b * nsplit + y) * 4 + comp) * ns + s];
      if (c == 0) v += phi[(size_t)b * ns + s] + rf;
      if (c == 1) v = (epsd[s] > 0.0) ? v / (2.0 * epsd[s]) : 0.0;
      acc += w[b] * v;
    }
  }
  out[o] = acc;
}

}  // namespace erm
