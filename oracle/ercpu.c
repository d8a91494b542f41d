/*
 * ercpu.c -- the oracle: FP64 CPU restatement of ERmod's per-frame
 * solute-solvent energy pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: linked by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker; never by the product.
 *
 * Each function names the reference lines it restates.  Floating-point
 * operations follow the reference's source order (left-to-right Fortran
 * evaluation, sequential array sums); the file is compiled with
 * -ffp-contract=off so that no multiply-add is fused behind our back.
 *
 * Pinning (tests/test_oracle_*.py): spline values against the reference's
 * own spline_test.F90 known answers and against spline.F90 compiled from the
 * reference (oracle/_ref); the hash against hash_real.c compiled from the
 * reference; edges against their closed form; the binary search against
 * numpy.searchsorted; the FFT against numpy.fft; the Green's function against
 * an independent closed form; the real-space kernel against a brute-force
 * double loop; reciprocal energies against a direct Ewald k-sum; the composed
 * energy against the NaCl Madelung constant and against its independence of
 * the Ewald screening parameter; histograms and the correlation matrix
 * against brute-force counting.  Parity with the reference's exact FP64 bits
 * (its -ffast-math build, its FFT library, its Fortran runtime erf/exp) is
 * not pinnable without running the reference and is not claimed.
 */
#include "ercpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* engmain.F90:203 -- PI = 3.1415926535897932 promoted to real(8) */
static const double PI_REF = 3.1415926535897932;

/* ====================================================================== */
/* spline.F90:26-59                                                       */
/* ====================================================================== */
void orc_spline_coeff(int32_t order, double *coeff) {
  for (int i = 0; i <= order; ++i) {
    double f = 1.0;
    for (int k = 1; k <= i; ++k) f = f * (double)k;
    for (int k = 1; k <= order - i; ++k) f = f * (double)k;
    f = (double)order / f;
    if (i % 2 == 1) f = -f;
    coeff[i] = f;
  }
}

/* integer power as repeated squaring, the lowering of Fortran x**n */
static double ipow(double x, int n) {
  double r = 1.0;
  int first = 1;
  double b = x;
  while (n > 0) {
    if (n & 1) {
      r = first ? b : r * b;
      first = 0;
    }
    n >>= 1;
    if (n) b = b * b;
  }
  return first ? 1.0 : r;
}

double orc_spline_value(const double *coeff, int32_t order, double x) {
  double f = 0.0;
  if (x > 0.0 && x < (double)order) {
    int k = (int)x;
    for (int i = 0; i <= k; ++i) f = f + coeff[i] * ipow(x - (double)i, order - 1);
  }
  return f;
}

/* ====================================================================== */
/* enginit edge construction, engproc.F90:163-218                         */
/* ====================================================================== */
static long fortran_nint(double x) { return (long)(x >= 0 ? floor(x + 0.5) : -floor(-x + 0.5)); }

int32_t orc_build_edges(const orc_hist *h, int32_t self_energy, double *ercrd_out,
                        int32_t cap, int32_t *pemax_out, int32_t *pesoft_out) {
  enum { RGLMAX = 5, LARGE = 10000 };
  double ecdmin = h->ecdmin, ecfmns = h->ecfmns, ecdcen = h->ecdcen;
  double eccore = h->eccore, ecdmax = h->ecdmax;
  double eclbin = h->eclbin, ecfbin = h->ecfbin, ec0bin = h->ec0bin;
  double finfac = h->finfac;
  int pecore = h->pecore;
  double ectmvl = finfac * ecfbin;
  ecdmin = ecdmin - ectmvl;
  ecfmns = ecfmns - ectmvl;
  double ecmns0 = ecdcen - ectmvl;
  double ecpls0 = 2.0 * ecdcen - ecmns0;
  double ecfpls = 2.0 * ecdcen - ecfmns;
  eccore = eccore + ecfpls - ecdcen;
  double cdrgvl[RGLMAX + 2] = {ecdmin, ecfmns, ecmns0, ecpls0, ecfpls, eccore, ecdmax};
  int uprgcd[RGLMAX + 2];
  uprgcd[0] = 1;
  for (int regn = 1; regn <= RGLMAX; ++regn) {
    double factor = (regn == 1 || regn == 5) ? eclbin : (regn == 2 || regn == 4) ? ecfbin : ec0bin;
    double incre = cdrgvl[regn] - cdrgvl[regn - 1];
    if (incre < factor) return ORC_E_ENG_ECD;
    long iduv = fortran_nint(incre / factor);
    uprgcd[regn] = uprgcd[regn - 1] + (int)iduv;
  }
  int pesoft = uprgcd[RGLMAX] - uprgcd[0];
  int pemax = pesoft + pecore;
  uprgcd[RGLMAX + 1] = pemax;
  if (pemax > LARGE) return ORC_E_ENG_SIZ;
  /* 1-based work array large enough for index pesoft+1 */
  int top = (pemax > uprgcd[RGLMAX] ? pemax : uprgcd[RGLMAX]) + 1;
  double *e = (double *)calloc((size_t)top + 1, sizeof(double));
  if (!e) return ORC_E_ARG;
  for (int r = 0; r <= RGLMAX + 1; ++r) e[uprgcd[r]] = cdrgvl[r];
  int nreg = (pecore == 0) ? RGLMAX : RGLMAX + 1;
  for (int regn = 1; regn <= nreg; ++regn) {
    int minrg = uprgcd[regn - 1], maxrg = uprgcd[regn];
    double incre = 0.0;
    if (regn <= RGLMAX) incre = e[maxrg] - e[minrg];
    if (regn == RGLMAX + 1) incre = log(e[maxrg] / e[minrg]);
    double factor = incre / (double)(maxrg - minrg);
    for (int iduv = minrg; iduv <= maxrg - 1; ++iduv) {
      incre = factor * (double)(iduv - minrg);
      if (regn <= RGLMAX) e[iduv] = e[minrg] + incre;
      if (regn == RGLMAX + 1) e[iduv] = e[minrg] * exp(incre);
    }
  }
  int n = self_energy ? pesoft : pemax;
  if (n > cap) {
    free(e);
    return ORC_E_ARG;
  }
  for (int i = 0; i < n; ++i) ercrd_out[i] = e[i + 1];
  free(e);
  *pemax_out = n;
  *pesoft_out = pesoft;
  return ORC_OK;
}

/* engproc.F90:927-957: count of edges strictly below v */
int32_t orc_binsearch(const double *coord, int32_t n, double v) {
  if (v < coord[0]) return 0;
  if (v > coord[n - 1]) return n;
  int rmin = 1, rmax = n + 1;
  while (rmax - rmin > 1) {
    int rmid = (rmin + rmax - 1) / 2;
    if (v > coord[rmid - 1])
      rmin = rmid + 1;
    else
      rmax = rmid + 1;
  }
  return rmin - 1;
}

/* engproc.F90:1185-1223, 'intn' branch for one species (idpt 1-based) */
void orc_repval(const double *crd, int32_t idmax, int32_t idsoft, int32_t idpt, double *out) {
  if (idpt <= idsoft)
    *out = (crd[idpt - 1] + crd[idpt]) / 2.0;
  else if (idpt < idmax)
    *out = sqrt(crd[idpt - 1] * crd[idpt]);
  else
    *out = crd[idpt - 1];
}

/* ====================================================================== */
/* hash_real.c:5-15                                                       */
/* ====================================================================== */
uint64_t orc_hash_double(const double *v, int32_t n) {
  uint64_t x = 0, b;
  for (int32_t i = 0; i < n; ++i) {
    memcpy(&b, &v[i], sizeof b);
    x = ((x << 7) | (x >> 57)) ^ b;
  }
  return x;
}

/* ====================================================================== */
/* engproc.F90:889-919                                                    */
/* ====================================================================== */
#define CM(m, i, j) ((m)[3 * ((j)-1) + ((i)-1)]) /* 1-based column-major */
void orc_inverted_cell(const double *cell, double *invcl, double *volume) {
  double v = CM(cell, 1, 1) * CM(cell, 2, 2) * CM(cell, 3, 3) + CM(cell, 1, 2) * CM(cell, 2, 3) * CM(cell, 3, 1) +
             CM(cell, 1, 3) * CM(cell, 2, 1) * CM(cell, 3, 2) - CM(cell, 1, 3) * CM(cell, 2, 2) * CM(cell, 3, 1) -
             CM(cell, 1, 2) * CM(cell, 2, 1) * CM(cell, 3, 3) - CM(cell, 1, 1) * CM(cell, 2, 3) * CM(cell, 3, 2);
  CM(invcl, 1, 1) = CM(cell, 2, 2) * CM(cell, 3, 3) - CM(cell, 2, 3) * CM(cell, 3, 2);
  CM(invcl, 1, 2) = CM(cell, 1, 3) * CM(cell, 3, 2) - CM(cell, 1, 2) * CM(cell, 3, 3);
  CM(invcl, 1, 3) = CM(cell, 1, 2) * CM(cell, 2, 3) - CM(cell, 1, 3) * CM(cell, 2, 2);
  CM(invcl, 2, 1) = CM(cell, 2, 3) * CM(cell, 3, 1) - CM(cell, 2, 1) * CM(cell, 3, 3);
  CM(invcl, 2, 2) = CM(cell, 1, 1) * CM(cell, 3, 3) - CM(cell, 1, 3) * CM(cell, 3, 1);
  CM(invcl, 2, 3) = CM(cell, 1, 3) * CM(cell, 2, 1) - CM(cell, 1, 1) * CM(cell, 2, 3);
  CM(invcl, 3, 1) = CM(cell, 2, 1) * CM(cell, 3, 2) - CM(cell, 2, 2) * CM(cell, 3, 1);
  CM(invcl, 3, 2) = CM(cell, 1, 2) * CM(cell, 3, 1) - CM(cell, 1, 1) * CM(cell, 3, 2);
  CM(invcl, 3, 3) = CM(cell, 1, 1) * CM(cell, 2, 2) - CM(cell, 1, 2) * CM(cell, 2, 1);
  for (int i = 0; i < 9; ++i) invcl[i] = invcl[i] / v;
  *volume = v;
}

/* ====================================================================== */
/* FFT: mixed-radix (2,3,5,7) Stockham autosort on interleaved complex.   */
/* Library-convention wrapper for the reference's fft_iface.F90 calls.    */
/* ====================================================================== */
typedef struct {
  int n, nf, fac[40];
  double *tw; /* tw[2j], tw[2j+1] = cos, sin of -2 pi j / n */
} fftplan;

static int fft_plan(int n, fftplan *p) {
  p->n = n;
  p->nf = 0;
  p->tw = NULL;
  int m = n;
  static const int radices[] = {4, 2, 3, 5, 7};
  for (int ri = 0; ri < 5; ++ri) {
    int r = radices[ri];
    while (m % r == 0) {
      p->fac[p->nf++] = r;
      m /= r;
    }
  }
  if (m != 1) return 0;
  p->tw = (double *)malloc(sizeof(double) * 2 * n);
  for (int j = 0; j < n; ++j) {
    double ang = -2.0 * M_PI * (double)j / (double)n;
    p->tw[2 * j] = cos(ang);
    p->tw[2 * j + 1] = sin(ang);
  }
  return 1;
}
static void fft_unplan(fftplan *p) { free(p->tw); }

/* one 1-D transform of length n on x (interleaved complex, stride in
 * complex units); sign -1 forward, +1 backward; work: 4n doubles.
 * Stockham autosort: stage with radix R and span p reads a[i + q m]
 * (m = n/R), applies the twiddle w^(q k) with k = i mod p and w the
 * (p R)-th root of unity, and writes the R-point DFT to b[(i-k)R + k + q p]. */
static void fft1d(const fftplan *pl, double *x, long stride, int sign, double *work) {
  int n = pl->n;
  const double *tw = pl->tw;
  double *a = work, *b = work + 2 * n;
  for (int i = 0; i < n; ++i) {
    a[2 * i] = x[2 * stride * i];
    a[2 * i + 1] = x[2 * stride * i + 1];
  }
  double sg = (double)(-sign); /* tw holds e^{-i}; conjugate for sign +1 */
  int p = 1;
  for (int f = 0; f < pl->nf; ++f) {
    int R = pl->fac[f];
    int m = n / R;
    int tstep = n / (p * R);
    for (int i = 0; i < m; ++i) {
      int k = i % p;
      double xr[8], xi[8];
      for (int q = 0; q < R; ++q) {
        double re = a[2 * (i + q * m)], im = a[2 * (i + q * m) + 1];
        if (q && k) {
          int t = (q * k * tstep) % n;
          double c = tw[2 * t], s = sg * tw[2 * t + 1];
          double u = re * c - im * s;
          im = re * s + im * c;
          re = u;
        }
        xr[q] = re;
        xi[q] = im;
      }
      int j = (i - k) * R + k;
      if (R == 2) {
        b[2 * j] = xr[0] + xr[1];
        b[2 * j + 1] = xi[0] + xi[1];
        b[2 * (j + p)] = xr[0] - xr[1];
        b[2 * (j + p) + 1] = xi[0] - xi[1];
      } else if (R == 4) {
        /* y_q = sum_s x_s e^{sign 2 pi i s q / 4}; e^{sign i pi/2} = sign*i */
        double ar = xr[0] + xr[2], ai = xi[0] + xi[2];
        double br = xr[0] - xr[2], bi = xi[0] - xi[2];
        double cr = xr[1] + xr[3], ci = xi[1] + xi[3];
        double dr = xr[1] - xr[3], di = xi[1] - xi[3];
        double s = (double)sign;
        b[2 * j] = ar + cr;
        b[2 * j + 1] = ai + ci;
        b[2 * (j + 2 * p)] = ar - cr;
        b[2 * (j + 2 * p) + 1] = ai - ci;
        /* q=1: b + sign*i*d ; q=3: b - sign*i*d */
        b[2 * (j + p)] = br - s * di;
        b[2 * (j + p) + 1] = bi + s * dr;
        b[2 * (j + 3 * p)] = br + s * di;
        b[2 * (j + 3 * p) + 1] = bi - s * dr;
      } else {
        int rstep = n / R;
        for (int q = 0; q < R; ++q) {
          double sr = 0.0, si = 0.0;
          for (int s2 = 0; s2 < R; ++s2) {
            int t = ((s2 * q) % R) * rstep;
            double c = tw[2 * t], s = sg * tw[2 * t + 1];
            sr += xr[s2] * c - xi[s2] * s;
            si += xr[s2] * s + xi[s2] * c;
          }
          b[2 * (j + q * p)] = sr;
          b[2 * (j + q * p) + 1] = si;
        }
      }
    }
    double *t = a;
    a = b;
    b = t;
    p *= R;
  }
  for (int i = 0; i < n; ++i) {
    x[2 * stride * i] = a[2 * i];
    x[2 * stride * i + 1] = a[2 * i + 1];
  }
}

int32_t orc_fft_r2c_3d(const int32_t *K, const double *in, double *out) {
  int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
  fftplan p1 = {0}, p2 = {0}, p3 = {0};
  if (!fft_plan(K1, &p1) || !fft_plan(K2, &p2) || !fft_plan(K3, &p3)) {
    fft_unplan(&p1);
    fft_unplan(&p2);
    fft_unplan(&p3);
    return ORC_E_UNSUPPORTED;
  }
  int nmax = K1 > K2 ? K1 : K2;
  nmax = nmax > K3 ? nmax : K3;
  double *work = (double *)malloc(sizeof(double) * 4 * nmax);
  double *row = (double *)malloc(sizeof(double) * 2 * K1);
  /* x: real rows -> half spectrum, two rows per complex transform:
   * Z = FFT(a + i b), A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i */
  long nrow = (long)K2 * K3;
  for (long rw = 0; rw < nrow; rw += 2) {
    const double *sa = in + rw * K1;
    const double *sb = (rw + 1 < nrow) ? in + (rw + 1) * K1 : NULL;
    for (int x = 0; x < K1; ++x) {
      row[2 * x] = sa[x];
      row[2 * x + 1] = sb ? sb[x] : 0.0;
    }
    fft1d(&p1, row, 1, -1, work);
    double *da = out + 2 * rw * H;
    double *db = sb ? out + 2 * (rw + 1) * H : NULL;
    for (int k = 0; k < H; ++k) {
      int nk = (K1 - k) % K1;
      double zr = row[2 * k], zi = row[2 * k + 1], cr = row[2 * nk], ci = -row[2 * nk + 1];
      da[2 * k] = 0.5 * (zr + cr);
      da[2 * k + 1] = 0.5 * (zi + ci);
      if (db) {
        db[2 * k] = 0.5 * (zi - ci);
        db[2 * k + 1] = -0.5 * (zr - cr);
      }
    }
  }
  /* y */
  for (int z = 0; z < K3; ++z)
    for (int x = 0; x < H; ++x) fft1d(&p2, out + 2 * ((long)z * K2 * H + x), H, -1, work);
  /* z */
  for (int y = 0; y < K2; ++y)
    for (int x = 0; x < H; ++x) fft1d(&p3, out + 2 * ((long)y * H + x), (long)K2 * H, -1, work);
  free(work);
  free(row);
  fft_unplan(&p1);
  fft_unplan(&p2);
  fft_unplan(&p3);
  return ORC_OK;
}

int32_t orc_fft_c2r_3d(const int32_t *K, const double *in, double *out) {
  int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
  fftplan p1 = {0}, p2 = {0}, p3 = {0};
  if (!fft_plan(K1, &p1) || !fft_plan(K2, &p2) || !fft_plan(K3, &p3)) {
    fft_unplan(&p1);
    fft_unplan(&p2);
    fft_unplan(&p3);
    return ORC_E_UNSUPPORTED;
  }
  int nmax = K1 > K2 ? K1 : K2;
  nmax = nmax > K3 ? nmax : K3;
  double *work = (double *)malloc(sizeof(double) * 4 * nmax);
  double *row = (double *)malloc(sizeof(double) * 2 * K1);
  long nc = (long)H * K2 * K3;
  double *t = (double *)malloc(sizeof(double) * 2 * nc);
  memcpy(t, in, sizeof(double) * 2 * nc);
  for (int y = 0; y < K2; ++y)
    for (int x = 0; x < H; ++x) fft1d(&p3, t + 2 * ((long)y * H + x), (long)K2 * H, +1, work);
  for (int z = 0; z < K3; ++z)
    for (int x = 0; x < H; ++x) fft1d(&p2, t + 2 * ((long)z * K2 * H + x), H, +1, work);
  /* x: hermitian half rows -> real rows, two rows per complex transform
   * (Z = A + i B with Hermitian extension; x = Re, y = Im of IFFT(Z)); the
   * imaginary parts of the self-conjugate entries (x = 0, K1/2) are ignored */
  long nrow = (long)K2 * K3;
  for (long rw = 0; rw < nrow; rw += 2) {
    const double *sa = t + 2 * rw * H;
    const double *sb = (rw + 1 < nrow) ? t + 2 * (rw + 1) * H : NULL;
    for (int x = 0; x < K1; ++x) {
      int hx = x < H ? x : K1 - x;
      double sg = x < H ? 1.0 : -1.0;
      double ar = sa[2 * hx], ai = sg * sa[2 * hx + 1];
      double br = sb ? sb[2 * hx] : 0.0, bi = sb ? sg * sb[2 * hx + 1] : 0.0;
      if (x == 0 || (K1 % 2 == 0 && x == K1 / 2)) {
        ai = 0.0;
        bi = 0.0;
      }
      row[2 * x] = ar - bi;
      row[2 * x + 1] = ai + br;
    }
    fft1d(&p1, row, 1, +1, work);
    double *da = out + rw * K1;
    double *db = sb ? out + (rw + 1) * K1 : NULL;
    for (int x = 0; x < K1; ++x) {
      da[x] = row[2 * x];
      if (db) db[x] = row[2 * x + 1];
    }
  }
  free(t);
  free(work);
  free(row);
  fft_unplan(&p1);
  fft_unplan(&p2);
  fft_unplan(&p3);
  return ORC_OK;
}

/* ====================================================================== */
/* recpcal.F90:87-152, 408-442                                            */
/* ====================================================================== */
void orc_spline_axis(int32_t K, int32_t p, double *splfc) {
  double coeff[32];
  orc_spline_coeff(p, coeff);
  for (int rci = 0; rci < K; ++rci) {
    double re = 0.0, im = 0.0;
    for (int spi = 0; spi <= p - 2; ++spi) {
      double chr = orc_spline_value(coeff, p, (double)(spi + 1));
      double rtp2 = 2.0 * PI_REF * (double)(spi * rci) / (double)K;
      re = re + chr * cos(rtp2);
      im = im + chr * sin(rtp2);
    }
    /* real(rcpi * conjg(rcpi)) */
    splfc[rci] = re * re - im * (-im);
  }
}

void orc_greenfunc(const int32_t *K, int32_t p, const double *invcl, double volume, double screen,
                   const double *splfc1, const double *splfc2, const double *splfc3, double *engfac) {
  int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
  for (int rc3 = 0; rc3 < K3; ++rc3)
    for (int rc2 = 0; rc2 < K2; ++rc2)
      for (int rc1 = 0; rc1 < H; ++rc1) {
        double factor = 0.0;
        long idx = ((long)rc3 * K2 + rc2) * H + rc1;
        if (rc1 == 0 && rc2 == 0 && rc3 == 0) {
          engfac[idx] = 0.0; /* left unset by the reference, zeroed after */
          continue;
        }
        int rci[3] = {rc1, rc2, rc3};
        double inm[3];
        int skip = 0;
        for (int m = 0; m < 3; ++m) {
          if ((p % 2 == 1) && (2 * abs(rci[m]) == K[m])) {
            skip = 1;
            break;
          }
          inm[m] = (rci[m] <= K[m] / 2) ? (double)rci[m] : (double)(rci[m] - K[m]);
        }
        if (!skip) {
          double xst[3];
          for (int m = 0; m < 3; ++m)
            xst[m] = invcl[3 * m + 0] * inm[0] + invcl[3 * m + 1] * inm[1] + invcl[3 * m + 2] * inm[2];
          double rtp2 = xst[0] * xst[0] + xst[1] * xst[1] + xst[2] * xst[2];
          double chr = (PI_REF * PI_REF) * rtp2 / (screen * screen);
          factor = exp(-chr) / rtp2 / PI_REF / volume;
          rtp2 = splfc1[rc1] * splfc2[rc2] * splfc3[rc3];
          factor = factor / rtp2;
        }
        engfac[idx] = factor;
      }
  engfac[0] = 0.0;
}

/* calc_gfb_pppm (recpcal.F90:168-186): the coefficients of
 * sum_l gf_b(l) x^(l-1), x = sin^2(k h / 2), the alias sum of the squared
 * charge-assignment function; the inner loop's index is 0 after it ends */
void orc_gfb_pppm(int32_t p, double *gf_b) {
  for (int i = 1; i < p; ++i) gf_b[i] = 0.0;
  gf_b[0] = 1.0;
  for (int m = 1; m <= p - 1; ++m) {
    int l;
    for (l = m; l >= 1; --l)
      gf_b[l] = 4.0 * (gf_b[l] * (double)(l - m) * ((double)(l - m) - 0.5) -
                       gf_b[l - 1] * (double)(l - m - 1) * (double)(l - m - 1));
    gf_b[0] = 4.0 * (gf_b[0] * (double)(l - m) * ((double)(l - m) - 0.5));
  }
  double fact = 1.0;
  for (int i = 1; i <= 2 * p - 1; ++i) fact *= (double)i;
  double gaminv = 1.0 / fact;
  for (int m = 0; m < p; ++m) gf_b[m] = gf_b[m] * gaminv;
}

/* gf_denom (recpcal.F90:189-202) */
static double gf_denom(const double *gf_b, int p, double x, double y, double z) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int l = p - 1; l >= 0; --l) {
    s0 = gf_b[l] + s0 * x;
    s1 = gf_b[l] + s1 * y;
    s2 = gf_b[l] + s2 * z;
  }
  double pr = s0 * s1 * s2;
  return pr * pr;
}

/* recpcal_pppm_greenfunc (recpcal.F90:205-325).  The z alias offset is
 * mz * K2, as the reference writes it (recpcal.F90:293); it equals mz * K3
 * for K2 = K3. */
int32_t orc_pppm_greenfunc(const int32_t *K, int32_t p, const double *cell, const double *invcl, double volume,
                           double screen, double *engfac) {
  int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
  if (p % 2 == 1) return ORC_E_ARG;
  double gf_b[32];
  orc_gfb_pppm(p, gf_b);
  const double EPS_HOC = 0.0000001;
  double m_max[3];
  for (int d = 0; d < 3; ++d) m_max[d] = screen / (PI_REF * (double)K[d]) * pow(-log(EPS_HOC), 0.25);
  int mmax[3];
  for (int j = 0; j < 3; ++j)
    mmax[j] = (int)(cell[3 * j] * m_max[0] + cell[3 * j + 1] * m_max[1] + cell[3 * j + 2] * m_max[2]);
  if (mmax[0] < 0 || mmax[1] < 0 || mmax[2] < 0) return ORC_E_ARG;
  for (int rc3 = 0; rc3 < K3; ++rc3) {
    double inm3 = (rc3 <= K3 / 2) ? (double)rc3 : (double)(rc3 - K3);
    double s3 = sin(PI_REF * inm3 / (double)K3);
    s3 = s3 * s3;
    for (int rc2 = 0; rc2 < K2; ++rc2) {
      double inm2 = (rc2 <= K2 / 2) ? (double)rc2 : (double)(rc2 - K2);
      double s2 = sin(PI_REF * inm2 / (double)K2);
      s2 = s2 * s2;
      for (int rc1 = 0; rc1 < H; ++rc1) {
        long idx = ((long)rc3 * K2 + rc2) * H + rc1;
        if (rc1 == 0 && rc2 == 0 && rc3 == 0) {
          engfac[idx] = 0.0;
          continue;
        }
        double inm1 = (rc1 <= K1 / 2) ? (double)rc1 : (double)(rc1 - K1);
        double s1 = sin(PI_REF * inm1 / (double)K1);
        s1 = s1 * s1;
        const double inm[3] = {inm1, inm2, inm3};
        double k[3];
        for (int i = 0; i < 3; ++i)
          k[i] = 2.0 * PI_REF * (invcl[3 * i] * inm[0] + invcl[3 * i + 1] * inm[1] + invcl[3 * i + 2] * inm[2]);
        double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        double factor = 0.0;
        if (k2 > 0.0) {
          double sum = 0.0;
          for (int mx = -mmax[0]; mx <= mmax[0]; ++mx) {
            double t1 = PI_REF * (inm1 / (double)K1 + (double)mx);
            double u1 = (t1 == 0.0) ? 1.0 : ipow(sin(t1) / t1, 2 * p);
            for (int my = -mmax[1]; my <= mmax[1]; ++my) {
              double t2 = PI_REF * (inm2 / (double)K2 + (double)my);
              double u2 = (t2 == 0.0) ? 1.0 : ipow(sin(t2) / t2, 2 * p);
              for (int mz = -mmax[2]; mz <= mmax[2]; ++mz) {
                double t3 = PI_REF * (inm3 / (double)K3 + (double)mz);
                double u3 = (t3 == 0.0) ? 1.0 : ipow(sin(t3) / t3, 2 * p);
                const double mN[3] = {(double)(mx * K1), (double)(my * K2), (double)(mz * K2)};
                double km[3], gam = 1.0;
                for (int i = 0; i < 3; ++i) {
                  km[i] = k[i] + 2.0 * PI_REF * (invcl[3 * i] * mN[0] + invcl[3 * i + 1] * mN[1] +
                                                 invcl[3 * i + 2] * mN[2]);
                }
                double g[3];
                for (int i = 0; i < 3; ++i) g[i] = exp(-0.25 * ipow(km[i] / screen, 2));
                gam = g[0] * g[1] * g[2];
                double kkm = k[0] * km[0] + k[1] * km[1] + k[2] * km[2];
                double kmkm = km[0] * km[0] + km[1] * km[1] + km[2] * km[2];
                sum = sum + kkm * gam * (u1 * u2 * u3) / kmkm;
              }
            }
          }
          factor = sum * 4.0 * PI_REF / (k2 * gf_denom(gf_b, p, s1, s2, s3)) / volume;
        }
        engfac[idx] = factor;
      }
    }
  }
  engfac[0] = 0.0;
  return ORC_OK;
}

void orc_spline_site(const double *x, const double *invcl, const int32_t *K, int32_t p, const double *coeff,
                     double *w, int32_t *g) {
  double inm[3];
  for (int k = 0; k < 3; ++k) {
    /* dot_product(invcl(k,:), xst(:)) */
    double f = invcl[k] * x[0] + invcl[3 + k] * x[1] + invcl[6 + k] * x[2];
    f = f - floor(f);
    inm[k] = f;
  }
  for (int m = 0; m < 3; ++m) {
    double factor = inm[m] * (double)K[m];
    int rci = (int)factor;
    for (int spi = 0; spi < p; ++spi) {
      double rtp2 = factor - (double)(rci - spi);
      w[m * p + spi] = orc_spline_value(coeff, p, rtp2);
    }
    g[m] = rci;
  }
}

/* ====================================================================== */
/* pose: insertion.F90:103-366, bestfit.F90:49-100,158-203                */
/* ====================================================================== */
static void com_of(int n, const double *pts, const double *mass, double *c) {
  double msum = 0.0;
  for (int j = 0; j < n; ++j) msum = msum + mass[j];
  for (int i = 0; i < 3; ++i) {
    double d = 0.0;
    for (int j = 0; j < n; ++j) d = d + pts[3 * j + i] * mass[j];
    c[i] = d / msum;
  }
}

static void qprod(const double *a, const double *b, double *p) {
  double dot = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  p[0] = a[0] * b[0] - dot;
  double cr[3] = {a[2] * b[3] - a[3] * b[2], a[3] * b[1] - a[1] * b[3], a[1] * b[2] - a[2] * b[1]};
  for (int i = 0; i < 3; ++i) p[1 + i] = cr[i] + a[0] * b[1 + i] + b[0] * a[1 + i];
}

static void qrotate(const double *v, const double *q, double *res) {
  double q2[4] = {0.0, v[0], v[1], v[2]};
  double cj[4] = {q[0], -q[1], -q[2], -q[3]};
  double t[4], r2[4];
  qprod(q, q2, t);
  qprod(t, cj, r2);
  res[0] = r2[1];
  res[1] = r2[2];
  res[2] = r2[3];
}

void orc_place_solute(int32_t n, const double *bfcoord, const double *mass, const double *cell, const double *r,
                      const double *q, double *out) {
  double c[3];
  memcpy(out, bfcoord, sizeof(double) * 3 * n);
  /* set_solute_origin, INSORG_ORIGIN: COM to (0,0,0) */
  com_of(n, out, mass, c);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < 3; ++i) out[3 * j + i] = (out[3 * j + i] - c[i]) + 0.0;
  /* set_shift_com, INSPOS_RANDOM: COM = cell . r */
  double com[3];
  for (int i = 0; i < 3; ++i) com[i] = cell[i] * r[0] + cell[3 + i] * r[1] + cell[6 + i] * r[2];
  com_of(n, out, mass, c);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < 3; ++i) out[3 * j + i] = (out[3 * j + i] - c[i]) + com[i];
  /* apply_orientation, INSROT_RANDOM */
  com_of(n, out, mass, c);
  for (int j = 0; j < n; ++j) {
    double v[3] = {out[3 * j] - c[0], out[3 * j + 1] - c[1], out[3 * j + 2] - c[2]};
    double t[3];
    qrotate(v, q, t);
    for (int i = 0; i < 3; ++i) out[3 * j + i] = t[i] + c[i];
  }
}

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double orc_rng_uniform(uint64_t seed, uint64_t index, uint32_t draw) {
  uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ULL);
  uint64_t ctr = index * 256ULL + (uint64_t)draw;
  uint64_t z = mix64(key ^ mix64(ctr + 0x632BE59BD9B4E019ULL));
  return (double)(z >> 11) * 0x1.0p-53;
}

/* insertion.F90:153-163 (position) then 340-356 (quaternion by rejection) */
void orc_random_pose(uint64_t seed, uint64_t index, double *r, double *q) {
  for (int i = 0; i < 3; ++i) r[i] = orc_rng_uniform(seed, index, (uint32_t)i);
  for (int a = 0; a < 63; ++a) {
    double t[4];
    for (int i = 0; i < 4; ++i) t[i] = orc_rng_uniform(seed, index, (uint32_t)(3 + 4 * a + i)) * 2.0 - 1.0;
    double s = t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3];
    if (s < 1.0) {
      double nrm = sqrt(s);
      for (int i = 0; i < 4; ++i) q[i] = t[i] / nrm;
      return;
    }
  }
  q[0] = 1.0;
  q[1] = q[2] = q[3] = 0.0;
}

void orc_aggregate_center(int32_t nmol, const int32_t *mol_begin, const int32_t *hostmol, const double *sitemass,
                          const double *pos, double *center) {
  double msum = 0.0;
  for (int m = 0; m < nmol; ++m)
    if (hostmol[m])
      for (int a = mol_begin[m]; a < mol_begin[m + 1]; ++a) msum = msum + sitemass[a];
  for (int i = 0; i < 3; ++i) {
    double d = 0.0;
    for (int m = 0; m < nmol; ++m)
      if (hostmol[m])
        for (int a = mol_begin[m]; a < mol_begin[m + 1]; ++a) d = d + pos[3 * a + i] * sitemass[a];
    center[i] = d / msum;
  }
}

void orc_insertion_pose(int32_t n, const double *bfcoord, const double *mass, const double *cell,
                        const orc_insertion *ins, const double *aggcen, uint64_t seed, uint64_t index,
                        double *out) {
  const uint64_t sp = seed ^ 0xD1B54A32D192ED03ULL;
  double c[3];
  memcpy(out, bfcoord, sizeof(double) * 3 * n);
  /* set_solute_origin: ORIGIN or AGGCEN move the COM there (set_solute_com) */
  if (ins->insorigin != 1) {
    double o[3] = {0.0, 0.0, 0.0};
    if (ins->insorigin == 2)
      for (int i = 0; i < 3; ++i) o[i] = aggcen[i];
    com_of(n, out, mass, c);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < 3; ++i) out[3 * j + i] = (out[3 * j + i] - c[i]) + o[i];
  }
  /* set_shift_com */
  if (ins->insposition == 0) {
    double r[3], com[3];
    for (int i = 0; i < 3; ++i) r[i] = orc_rng_uniform(seed, index, (uint32_t)i);
    for (int i = 0; i < 3; ++i) com[i] = cell[i] * r[0] + cell[3 + i] * r[1] + cell[6 + i] * r[2];
    com_of(n, out, mass, c);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < 3; ++i) out[3 * j + i] = (out[3 * j + i] - c[i]) + com[i];
  } else if (ins->insposition == 2) {
    /* INSPOS_SPHERE (insertion.F90:165-180) */
    double t = orc_rng_uniform(sp, index, 0);
    t = pow(t * (ipow(ins->upreg, 3) - ipow(ins->lwreg, 3)) + ipow(ins->lwreg, 3), 1.0 / 3.0);
    double r[3] = {0.0, 0.0, 1.0}, norm = 1.0;
    for (int a = 0; a < 84; ++a) {
      double u[3];
      for (int i = 0; i < 3; ++i) u[i] = orc_rng_uniform(sp, index, (uint32_t)(1 + 3 * a + i)) * 2.0 - 1.0;
      double s2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      if (0.0 < s2 && s2 <= 1.0) {
        memcpy(r, u, sizeof r);
        norm = s2;
        break;
      }
    }
    norm = sqrt(norm);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < 3; ++i) out[3 * j + i] = out[3 * j + i] + t * r[i] / norm;
  } else if (ins->insposition == 3 || ins->insposition == 4) {
    /* INSPOS_SLAB_GENERIC / SYMMETRIC (insertion.F90:181-198) */
    double r[3], com[3];
    r[0] = orc_rng_uniform(sp, index, 0);
    r[1] = orc_rng_uniform(sp, index, 1);
    double t = orc_rng_uniform(sp, index, 2);
    t = t * (ins->upreg - ins->lwreg) + ins->lwreg;
    if (ins->insposition == 4) {
      double dir = orc_rng_uniform(sp, index, 3);
      if (dir > 0.5) t = -t;
    }
    double len3 = sqrt(cell[6] * cell[6] + cell[7] * cell[7] + cell[8] * cell[8]);
    r[2] = t / len3;
    for (int i = 0; i < 3; ++i) com[i] = cell[i] * r[0] + cell[3 + i] * r[1] + cell[6 + i] * r[2];
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < 3; ++i) out[3 * j + i] = out[3 * j + i] + com[i];
  }
  /* apply_orientation */
  if (ins->insorient == 0) {
    double r[3], q[4];
    orc_random_pose(seed, index, r, q);
    com_of(n, out, mass, c);
    for (int j = 0; j < n; ++j) {
      double v[3] = {out[3 * j] - c[0], out[3 * j + 1] - c[1], out[3 * j + 2] - c[2]};
      double t[3];
      qrotate(v, q, t);
      for (int i = 0; i < 3; ++i) out[3 * j + i] = t[i] + c[i];
    }
  }
}

/* ---- bestfit.F90: Horn's quaternion superposition ----------------------- */
/* eigenvector of the largest eigenvalue of a symmetric 4x4 matrix by cyclic
 * Jacobi rotations (the reference calls LAPACK dsyev) */
static void sym4_top_eigvec(const double Ain[4][4], double v[4]) {
  double A[4][4], V[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      A[i][j] = Ain[i][j];
      V[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) off += A[p][q] * A[p][q];
    if (off == 0.0) break;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) {
        if (A[p][q] == 0.0) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < 4; ++k) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = cs * akp - sn * akq;
          A[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 4; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = cs * apk - sn * aqk;
          A[q][k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < 4; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = cs * vkp - sn * vkq;
          V[k][q] = sn * vkp + cs * vkq;
        }
      }
  }
  int best = 0;
  for (int i = 1; i < 4; ++i)
    if (A[i][i] > A[best][best]) best = i;
  for (int i = 0; i < 4; ++i) v[i] = V[i][best];
}

/* find_rotation_quaternion (bestfit.F90:92-146) */
static void fit_quaternion(int n, const double *ref, const double *mov, const double *w, double q[4]) {
  double ip[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc = acc + ref[3 * k + i] * mov[3 * k + j] * w[k];
      ip[j][i] = acc;
    }
  double M[4][4];
  M[0][0] = ip[0][0] + ip[1][1] + ip[2][2];
  M[1][1] = ip[0][0] - ip[1][1] - ip[2][2];
  M[2][2] = -ip[0][0] + ip[1][1] - ip[2][2];
  M[3][3] = -ip[0][0] - ip[1][1] + ip[2][2];
  M[0][1] = ip[1][2] - ip[2][1];
  M[0][2] = ip[2][0] - ip[0][2];
  M[0][3] = ip[0][1] - ip[1][0];
  M[1][2] = ip[0][1] + ip[1][0];
  M[2][3] = ip[1][2] + ip[2][1];
  M[1][3] = ip[2][0] + ip[0][2];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) M[i][j] = M[j][i];
  sym4_top_eigvec(M, q);
}

void orc_fit_a_rotate_b(int32_t na, const double *refa, const double *a, const double *w, int32_t nb, const double *b,
                        double *bout) {
  double *wr = (double *)malloc(sizeof(double) * 3 * (na ? na : 1));
  double *wa = (double *)malloc(sizeof(double) * 3 * (na ? na : 1));
  memcpy(wr, refa, sizeof(double) * 3 * na);
  memcpy(wa, a, sizeof(double) * 3 * na);
  double cr[3], ca[3];
  com_of(na, wr, w, cr);
  for (int k = 0; k < na; ++k)
    for (int i = 0; i < 3; ++i) wr[3 * k + i] = wr[3 * k + i] - cr[i];
  com_of(na, wa, w, ca);
  for (int k = 0; k < na; ++k)
    for (int i = 0; i < 3; ++i) wa[3 * k + i] = wa[3 * k + i] - ca[i];
  double q[4];
  fit_quaternion(na, wr, wa, w, q);
  for (int k = 0; k < nb; ++k) {
    double d[3] = {b[3 * k] - ca[0], b[3 * k + 1] - ca[1], b[3 * k + 2] - ca[2]}, r[3];
    qrotate(d, q, r);
    for (int i = 0; i < 3; ++i) bout[3 * k + i] = r[i] + cr[i];
  }
  free(wr);
  free(wa);
}

double orc_rmsd_nofit(int32_t n, const double *a, const double *b, const double *w) {
  double ws = 0.0, disp = 0.0;
  for (int j = 0; j < n; ++j) ws = ws + w[j];
  for (int j = 0; j < n; ++j) {
    double d0 = b[3 * j] - a[3 * j], d1 = b[3 * j + 1] - a[3 * j + 1], d2 = b[3 * j + 2] - a[3 * j + 2];
    disp = disp + w[j] * (d0 * d0 + d1 * d1 + d2 * d2);
  }
  return sqrt(disp / ws);
}

static uint64_t att_seed(uint64_t seed, int a, uint64_t salt) {
  return seed ^ mix64((uint64_t)a * 0x9E3779B97F4A7C15ULL + salt);
}

int32_t orc_refstr_pose(int32_t n, const double *mass, const double *cell, const orc_insertion *ins,
                        const double *fitted, const double *best, const double *w, const double *gconst,
                        uint64_t seed, uint64_t index, double *out, double *weight) {
  const double PI_F = 3.1415926535897932;
  double c[3], wt = 1.0;
  for (int a = 0; a < 4096; ++a) {
    uint64_t sp = att_seed(seed, a, 0x243F6A8885A308D3ULL), sr = att_seed(seed, a, 0x13198A2E03707344ULL);
    memcpy(out, fitted, sizeof(double) * 3 * n);
    double com[3];
    if (ins->insposition == 5) {
      /* INSPOS_RMSD (insertion.F90:199-217) */
      com_of(n, out, mass, c);
      double maxdis = 0.0;
      for (int j = 0; j < n; ++j) {
        double d0 = out[3 * j] - c[0], d1 = out[3 * j + 1] - c[1], d2 = out[3 * j + 2] - c[2];
        double dst = d0 * d0 + d1 * d1 + d2 * d2;
        if (dst > maxdis) maxdis = dst;
      }
      maxdis = sqrt(maxdis);
      double movmax = ins->upreg + 3.0 * maxdis;
      for (int i = 0; i < 3; ++i) com[i] = movmax * (2.0 * orc_rng_uniform(sp, index, (uint32_t)i) - 1.0);
    } else {
      /* INSPOS_GAUSS: uniform_gauss_mixture (insertion.F90:236-288) */
      double sc[3];
      int uni = orc_rng_uniform(att_seed(seed, a, 0xA4093822299F31D0ULL), index, 0) < 0.5;
      uint32_t d = 0;
      for (int i = 0; i < 3; ++i) {
        if (uni) {
          sc[i] = (2.0 * orc_rng_uniform(sp, index, d++) - 1.0) * gconst[i];
        } else {
          double r = 0.0;
          for (int t = 0; t < 40; ++t) {
            double r1 = 1.0 - orc_rng_uniform(sp, index, d++);
            double r2 = orc_rng_uniform(sp, index, d++);
            r = sqrt(-2.0 * log(r1)) * cos(2.0 * PI_F * r2) / sqrt(2.0);
            if (fabs(r) < gconst[i]) break;
          }
          sc[i] = r;
        }
      }
      double sq = sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2];
      wt = wt / (0.5 * 1.0 / gconst[3] + (1.0 - 0.5) * exp(-sq) / gconst[4]);
      double u[3];
      for (int i = 0; i < 3; ++i) u[i] = sc[i] / (gconst[i] * 2.0);
      for (int i = 0; i < 3; ++i) com[i] = cell[i] * u[0] + cell[3 + i] * u[1] + cell[6 + i] * u[2];
    }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < 3; ++i) out[3 * j + i] = out[3 * j + i] + com[i];
    if (ins->insorient == 0) {
      double q[4] = {1.0, 0.0, 0.0, 0.0};
      for (int t = 0; t < 63; ++t) {
        double v[4];
        for (int i = 0; i < 4; ++i) v[i] = orc_rng_uniform(sr, index, (uint32_t)(4 * t + i)) * 2.0 - 1.0;
        double s2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
        if (s2 < 1.0) {
          double nrm = sqrt(s2);
          for (int i = 0; i < 4; ++i) q[i] = v[i] / nrm;
          break;
        }
      }
      com_of(n, out, mass, c);
      for (int j = 0; j < n; ++j) {
        double v[3] = {out[3 * j] - c[0], out[3 * j + 1] - c[1], out[3 * j + 2] - c[2]}, t[3];
        qrotate(v, q, t);
        for (int i = 0; i < 3; ++i) out[3 * j + i] = t[i] + c[i];
      }
    }
    double rmsd = orc_rmsd_nofit(n, best, out, w);
    if (!(ins->lwreg > rmsd || rmsd > ins->upreg)) {
      *weight = wt;
      return ORC_OK;
    }
  }
  *weight = wt;
  return 14;
}

/* ====================================================================== */
/* context                                                                */
/* ====================================================================== */
struct orc_ctx {
  orc_system s;
  /* owned copies of the system arrays */
  int32_t *mol_begin, *ljtype, *sluvid, *uvspec, *uvmax, *uvsoft;
  double *charge, *lj_eps, *lj_sig2, *mol_charge, *uvcrd, *escrd;
  int32_t ermax;
  /* tagpt (engproc.F90:336-347) */
  int32_t slvmax, *tagpt, *kof; /* kof[mol] = index into tagpt or -1 */
  /* PME */
  int32_t p, K[3], H;
  double coeff[33];
  double *splfc[3], *engfac, *cnv, *rcp;
  int32_t *slvtag, nslot;
  double *splslv;
  int32_t *grdslv;
  int32_t have_cell;
  double prevcl[9];
  /* frame */
  double cell[9], invcl[9], volume, L[3], stat_weight_system;
  double *pos, *posn;
  /* accumulators */
  double *edens, *ecorr, *eself, *minuv, *maxuv, *slnuv, *flceng;
  double engnorm, engsmpl, avslf, voffset;
  int32_t voffset_init, warn_bin_first;
  int32_t *insdst;
  /* intermediate state (extensions/Intermediate/engproc.F90:34-39) */
  int32_t *zerodns;
  double *edintm, *uvintm, norm_intm, soln_intm, soln_total;
  /* realcal_self cache (engproc.F90:712-730) */
  uint64_t solute_hash;
  double usreal;
  /* scratch */
  double *spos, *sposn;
};

static int32_t *dupi(const int32_t *src, long n) {
  int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (src && n > 0) memcpy(d, src, sizeof(int32_t) * (size_t)n);
  return d;
}
static double *dupd(const double *src, long n) {
  double *d = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  if (src && n > 0) memcpy(d, src, sizeof(double) * (size_t)n);
  return d;
}
#define DUPI(dst, src, n) ((dst) = dupi((src), (long)(n)))
#define DUPD(dst, src, n) ((dst) = dupd((src), (long)(n)))

int32_t orc_slvmax(const orc_ctx *c) { return c->slvmax; }
const double *orc_flceng(const orc_ctx *c) { return c->flceng; }

orc_ctx *orc_create(const orc_system *sys, int32_t *status) {
  *status = ORC_OK;
  if (sys->splodr < 2 || sys->splodr > 32 || sys->ms[0] < 2 || sys->ms[1] < 2 || sys->ms[2] < 2) {
    *status = ORC_E_ARG;
    return NULL;
  }
  /* cltype / boxshp consistency (setconf.F90:326-339, 368-372) */
  if ((sys->cltype != 0 && sys->cltype != 2 && sys->cltype != 3) || (sys->boxshp != 0 && sys->boxshp != 1) ||
      (sys->boxshp == 0 && (sys->cltype != 0 || sys->estype == 2)) || (sys->cltype == 3 && sys->splodr % 2 == 1)) {
    *status = ORC_E_ARG;
    return NULL;
  }
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->s = *sys;
  int na = sys->natom, nm = sys->nmol, nt = sys->ntype;
  DUPI(c->mol_begin, sys->mol_begin, nm + 1);
  DUPI(c->ljtype, sys->ljtype, na);
  DUPI(c->sluvid, sys->sluvid, nm);
  DUPI(c->uvspec, sys->uvspec, nm);
  DUPI(c->uvmax, sys->uvmax, sys->numslv);
  DUPI(c->uvsoft, sys->uvsoft, sys->numslv);
  DUPD(c->charge, sys->charge, na);
  DUPD(c->lj_eps, sys->lj_eps, nt * nt);
  DUPD(c->lj_sig2, sys->lj_sig2, nt * nt);
  DUPD(c->mol_charge, sys->mol_charge, nm);
  c->ermax = 0;
  for (int i = 0; i < sys->numslv; ++i) c->ermax += sys->uvmax[i];
  DUPD(c->uvcrd, sys->uvcrd, c->ermax);
  DUPD(c->escrd, sys->selfcal ? sys->escrd : NULL, sys->selfcal ? sys->esmax : 0);
  c->s.mol_begin = c->mol_begin;
  c->s.ljtype = c->ljtype;
  c->s.sluvid = c->sluvid;
  c->s.uvspec = c->uvspec;
  c->s.uvmax = c->uvmax;
  c->s.uvsoft = c->uvsoft;
  c->s.charge = c->charge;
  c->s.lj_eps = c->lj_eps;
  c->s.lj_sig2 = c->lj_sig2;
  c->s.mol_charge = c->mol_charge;
  c->s.uvcrd = c->uvcrd;
  c->s.escrd = c->escrd;
  /* tagpt: physical molecules (solvent or solute) */
  c->tagpt = (int32_t *)malloc(sizeof(int32_t) * nm);
  c->kof = (int32_t *)malloc(sizeof(int32_t) * nm);
  c->slvmax = 0;
  for (int i = 0; i < nm; ++i) {
    c->kof[i] = -1;
    if (sys->sluvid[i] == 0 || sys->sluvid[i] == 1) {
      c->kof[i] = c->slvmax;
      c->tagpt[c->slvmax++] = i;
    }
  }
  /* recpcal_init (recpcal.F90:42-85) */
  c->p = sys->splodr;
  for (int m = 0; m < 3; ++m) c->K[m] = sys->ms[m];
  c->H = c->K[0] / 2 + 1;
  orc_spline_coeff(c->p, c->coeff);
  c->slvtag = (int32_t *)malloc(sizeof(int32_t) * nm);
  c->nslot = 0;
  for (int i = 0; i < nm; ++i) c->slvtag[i] = -1;
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    c->slvtag[m] = c->nslot;
    c->nslot += c->mol_begin[m + 1] - c->mol_begin[m];
  }
  c->splslv = (double *)malloc(sizeof(double) * 3 * c->p * (c->nslot ? c->nslot : 1));
  c->grdslv = (int32_t *)malloc(sizeof(int32_t) * 3 * (c->nslot ? c->nslot : 1));
  for (int m = 0; m < 3; ++m) {
    c->splfc[m] = (double *)malloc(sizeof(double) * c->K[m]);
    orc_spline_axis(c->K[m], c->p, c->splfc[m]);
  }
  long nr = (long)c->K[0] * c->K[1] * c->K[2];
  long nh = (long)c->H * c->K[1] * c->K[2];
  c->engfac = (double *)malloc(sizeof(double) * nh);
  c->cnv = (double *)malloc(sizeof(double) * nr);
  c->rcp = (double *)malloc(sizeof(double) * 2 * nh);
  c->pos = (double *)malloc(sizeof(double) * 3 * (na ? na : 1));
  c->posn = (double *)malloc(sizeof(double) * 3 * (na ? na : 1));
  int maxsite = 1;
  for (int i = 0; i < nm; ++i) {
    int ns = c->mol_begin[i + 1] - c->mol_begin[i];
    if (ns > maxsite) maxsite = ns;
  }
  c->spos = (double *)malloc(sizeof(double) * 3 * maxsite);
  c->sposn = (double *)malloc(sizeof(double) * 3 * maxsite);
  /* accumulators */
  int ns1 = sys->numslv + 1;
  c->edens = (double *)calloc(c->ermax ? c->ermax : 1, sizeof(double));
  c->ecorr = sys->corrcal ? (double *)calloc((size_t)c->ermax * c->ermax, sizeof(double)) : NULL;
  c->eself = sys->selfcal ? (double *)calloc(sys->esmax ? sys->esmax : 1, sizeof(double)) : NULL;
  c->minuv = (double *)malloc(sizeof(double) * ns1);
  c->maxuv = (double *)malloc(sizeof(double) * ns1);
  c->slnuv = (double *)calloc(ns1, sizeof(double));
  c->flceng = (double *)calloc(ns1, sizeof(double));
  c->insdst = (int32_t *)calloc(c->ermax ? c->ermax : 1, sizeof(int32_t));
  for (int i = 0; i < ns1; ++i) {
    c->minuv[i] = 1.0e50;
    c->maxuv[i] = -1.0e50;
  }
  c->voffset = -1.0e50;
  c->warn_bin_first = 1;
  orc_engclear(c);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->mol_begin, c->ljtype, c->sluvid, c->uvspec, c->uvmax, c->uvsoft, c->charge, c->lj_eps,
                  c->lj_sig2, c->mol_charge, c->uvcrd, c->escrd, c->tagpt, c->kof, c->slvtag, c->splslv,
                  c->grdslv, c->splfc[0], c->splfc[1], c->splfc[2], c->engfac, c->cnv, c->rcp, c->pos,
                  c->posn, c->spos, c->sposn, c->edens, c->ecorr, c->eself, c->minuv, c->maxuv, c->slnuv,
                  c->flceng, c->insdst, c->zerodns, c->edintm, c->uvintm};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  free(c);
}

/* engproc.F90:276-288 */
void orc_engclear(orc_ctx *c) {
  memset(c->edens, 0, sizeof(double) * c->ermax);
  if (c->ecorr) memset(c->ecorr, 0, sizeof(double) * (size_t)c->ermax * c->ermax);
  if (c->eself) memset(c->eself, 0, sizeof(double) * c->s.esmax);
  memset(c->slnuv, 0, sizeof(double) * (c->s.numslv + 1));
  c->avslf = 0.0;
  c->engnorm = 0.0;
  c->engsmpl = 0.0;
  /* soln_intm and soln_total run over all divisions */
  if (c->zerodns) {
    memset(c->edintm, 0, sizeof(double) * c->ermax);
    memset(c->uvintm, 0, sizeof(double) * (c->s.numslv + 1));
    c->norm_intm = 0.0;
  }
}

static double wrap1(double x, double L) { return x - L * floor(x * (1.0 / L)); }

/* per-frame part of engconst (engproc.F90:352-383) plus the cuboid
 * normalisation of realcal_prepare (realcal.F90:137-146, 789-885) */
int32_t orc_set_frame(orc_ctx *c, const double *sitepos, const double *cell, double stat_weight_system) {
  c->stat_weight_system = stat_weight_system;
  if (c->s.boxshp == 0) {
    /* non-periodic: raw coordinates, no cell (realcal.F90:137-146) */
    memcpy(c->pos, sitepos, sizeof(double) * 3 * c->s.natom);
    memcpy(c->posn, sitepos, sizeof(double) * 3 * c->s.natom);
    return ORC_OK;
  }
  for (int i = 1; i <= 3; ++i)
    for (int j = 1; j <= 3; ++j)
      if (i != j && fabs(CM(cell, i, j)) > 1e-8) return ORC_E_NONCUBOID;
  memcpy(c->cell, cell, sizeof(double) * 9);
  orc_inverted_cell(c->cell, c->invcl, &c->volume);
  for (int i = 0; i < 3; ++i) c->L[i] = fabs(CM(cell, i + 1, i + 1));
  int changed = !c->have_cell;
  for (int i = 0; i < 9 && !changed; ++i) changed = (c->prevcl[i] != cell[i]);
  if (changed && c->s.cltype == 2)
    orc_greenfunc(c->K, c->p, c->invcl, c->volume, c->s.screen, c->splfc[0], c->splfc[1], c->splfc[2], c->engfac);
  if (changed && c->s.cltype == 3) {
    int st = orc_pppm_greenfunc(c->K, c->p, c->cell, c->invcl, c->volume, c->s.screen, c->engfac);
    if (st) return st;
  }
  memcpy(c->prevcl, cell, sizeof(double) * 9);
  c->have_cell = 1;
  memcpy(c->pos, sitepos, sizeof(double) * 3 * c->s.natom);
  for (int a = 0; a < c->s.natom; ++a)
    for (int i = 0; i < 3; ++i) c->posn[3 * a + i] = wrap1(c->pos[3 * a + i], c->L[i]);
  if (c->s.cltype == 0) return ORC_OK;
  /* recpcal_prepare_solvent for every physical molecule */
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    int slot = c->slvtag[m];
    if (slot < 0) return ORC_E_RCP_CNS;
    for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a, ++slot)
      orc_spline_site(c->pos + 3 * a, c->invcl, c->K, c->p, c->coeff, c->splslv + (long)3 * c->p * slot,
                      c->grdslv + 3 * slot);
  }
  return ORC_OK;
}

/* solute coordinates: caller-supplied (test particle) or from the frame */
static int solute_coords(orc_ctx *c, int solute, const double *solute_pos, int *nsite) {
  if (solute < 0 || solute >= c->s.nmol) return ORC_E_ARG;
  int b = c->mol_begin[solute], e = c->mol_begin[solute + 1];
  *nsite = e - b;
  for (int j = 0; j < *nsite; ++j)
    for (int i = 0; i < 3; ++i) {
      double x = solute_pos ? solute_pos[3 * j + i] : c->pos[3 * (b + j) + i];
      c->spos[3 * j + i] = x;
      c->sposn[3 * j + i] = c->s.boxshp ? wrap1(x, c->L[i]) : x;
    }
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* real-space kernel: realcal.F90:58-135, 373-784                          */
/* ---------------------------------------------------------------------- */
typedef struct {
  double lw2, up2, el2, lw3, up3, lw6, up6, repA, repB, repC, attA, attB, attC;
} ljconst;

/* realcal.F90:888-904 */
static void gmx_force_params(int pw, double lw, double up, double *A, double *B, double *C) {
  double d = up - lw;
  *A = -(double)pw * ((double)(pw + 4) * up - (double)(pw + 1) * lw) / (ipow(up, pw + 2) * ipow(d, 2)) / 3.0;
  *B = (double)pw * ((double)(pw + 3) * up - (double)(pw + 1) * lw) / (ipow(up, pw + 2) * ipow(d, 3)) / 4.0;
  *C = 1.0 / ipow(up, pw) - *A * ipow(d, 3) - *B * ipow(d, 4);
}

static void make_ljconst(const orc_system *s, ljconst *k) {
  k->lw2 = ipow(s->lwljcut, 2);
  k->up2 = ipow(s->upljcut, 2);
  k->el2 = ipow(s->elecut, 2);
  k->lw3 = ipow(s->lwljcut, 3);
  k->up3 = ipow(s->upljcut, 3);
  k->lw6 = k->lw3 * k->lw3;
  k->up6 = k->up3 * k->up3;
  gmx_force_params(12, s->lwljcut, s->upljcut, &k->repA, &k->repB, &k->repC);
  gmx_force_params(6, s->lwljcut, s->upljcut, &k->attA, &k->attB, &k->attC);
}

/* LJ for a pair inside the lower cutoff (realcal.F90:708-734) */
static double lj_low(const orc_system *s, const ljconst *k, double eps, double sig2, double dist) {
  double invr2 = sig2 / dist;
  double invr6 = invr2 * invr2 * invr2;
  switch (s->ljswitch) {
    case 0:
    case 1:
      return 4.0 * eps * invr6 * (invr6 - 1.0);
    case 2: {
      double sg6 = sig2 * sig2 * sig2;
      double vdwa = invr6 * invr6 - sg6 * sg6 / (k->lw6 * k->up6);
      double vdwb = invr6 - sg6 / (k->lw3 * k->up3);
      return 4.0 * eps * (vdwa - vdwb);
    }
    default: {
      double sg6 = sig2 * sig2 * sig2;
      double vdwa = invr6 * invr6 - sg6 * sg6 * k->repC;
      double vdwb = invr6 - sg6 * k->attC;
      return 4.0 * eps * (vdwa - vdwb);
    }
  }
}

/* LJ in the switching region (realcal.F90:737-776) */
static double lj_switch(const orc_system *s, const ljconst *k, double eps, double sig2, double dist) {
  double invr2 = sig2 / dist;
  double invr6 = invr2 * invr2 * invr2;
  switch (s->ljswitch) {
    case 0:
      return 4.0 * eps * invr6 * (invr6 - 1.0) * (2.0 * dist + k->up2 - 3.0 * k->lw2) *
             ipow(dist - k->up2, 2) / ipow(k->up2 - k->lw2, 3);
    case 1: {
      double sw = (sqrt(dist) - s->lwljcut) / (s->upljcut - s->lwljcut);
      return 4.0 * eps * invr6 * (invr6 - 1.0) *
             (1.0 - 10.0 * ipow(sw, 3) + 15.0 * ipow(sw, 4) - 6.0 * ipow(sw, 5));
    }
    case 2: {
      double invr3 = sqrt(invr6);
      double sg6 = sig2 * sig2 * sig2;
      double sg3 = sqrt(sg6);
      double vdwa = k->up6 / (k->up6 - k->lw6) * ipow(invr6 - sg6 / k->up6, 2);
      double vdwb = k->up3 / (k->up3 - k->lw3) * ipow(invr3 - sg3 / k->up3, 2);
      return 4.0 * eps * (vdwa - vdwb);
    }
    default: {
      double sg6 = sig2 * sig2 * sig2;
      double sw = sqrt(dist) - s->lwljcut;
      double vdwa = invr6 * invr6 - sg6 * sg6 * (k->repA * ipow(sw, 3) + k->repB * ipow(sw, 4) + k->repC);
      double vdwb = invr6 - sg6 * (k->attA * ipow(sw, 3) + k->attB * ipow(sw, 4) + k->attC);
      return 4.0 * eps * (vdwa - vdwb);
    }
  }
}

static double minimg2(const double *a, const double *b, const double *half) {
  double d2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    double d = b[i] - a[i];
    d = half[i] - fabs(half[i] - fabs(d));
    d2 = d2 + d * d;
  }
  return d2;
}

typedef struct {
  int n;
  double *eps, *sig2, *dist;
  int *bel;
} plist;

typedef struct {
  int bs[3];
  double unit[3];
  int npill, *pill_y, *pill_z, *pill_xlen;
  /* sorted solute and solvent sites */
  int nu, nv;
  int *u_atom, *u_psum; /* psum: ncell+1 */
  double *u_pos;
  int *v_atom, *v_bel, *v_psum;
  double *v_pos;
  plist low, sw, el;
  double *el_q;
} cellgrid;

/* set_block_info (realcal.F90:373-437) */
static void cells_setup(orc_ctx *c, cellgrid *g) {
  const orc_system *s = &c->s;
  int bmax = 0;
  for (int i = 0; i < 3; ++i) {
    g->bs[i] = (int)ceil(c->L[i] / s->block_threshold);
    g->unit[i] = c->L[i] / (double)g->bs[i];
    if (g->bs[i] > bmax) bmax = g->bs[i];
  }
  double *gd = (double *)malloc(sizeof(double) * 3 * bmax);
  double *bd = (double *)malloc(sizeof(double) * 3 * bmax);
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < g->bs[j]; ++i) {
      int a = i, b = ((-i) % g->bs[j] + g->bs[j]) % g->bs[j];
      gd[j * bmax + i] = (double)(a < b ? a : b) * g->unit[j];
    }
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < g->bs[j]; ++i) {
      int n = g->bs[j];
      double v0 = gd[j * bmax + i], v1 = gd[j * bmax + (i - 1 + n) % n], v2 = gd[j * bmax + (i + 1) % n];
      double m = v0 < v1 ? v0 : v1;
      bd[j * bmax + i] = m < v2 ? m : v2;
    }
  double cut = s->upljcut > s->elecut ? s->upljcut : s->elecut;
  double cut2 = ipow(cut, 2);
  g->npill = 0;
  g->pill_y = (int *)malloc(sizeof(int) * g->bs[1] * g->bs[2]);
  g->pill_z = (int *)malloc(sizeof(int) * g->bs[1] * g->bs[2]);
  g->pill_xlen = (int *)malloc(sizeof(int) * g->bs[1] * g->bs[2]);
  for (int k = 0; k < g->bs[2]; ++k)
    for (int j = 0; j < g->bs[1]; ++j) {
      double l = ipow(bd[1 * bmax + j], 2) + ipow(bd[2 * bmax + k], 2);
      if (l < cut2) {
        g->pill_y[g->npill] = j;
        g->pill_z[g->npill] = k;
        g->pill_xlen[g->npill] = (int)ceil(sqrt(cut2 - l) / g->unit[0]);
        g->npill++;
      }
    }
  free(gd);
  free(bd);
}

static int cell_of(const cellgrid *g, const double *x, const double *L) {
  int b[3];
  for (int i = 0; i < 3; ++i) {
    int t = (int)floor(x[i] * (1.0 / L[i]) * (double)g->bs[i]);
    t %= g->bs[i];
    if (t < 0) t += g->bs[i];
    b[i] = t;
  }
  return b[0] + g->bs[0] * (b[1] + g->bs[1] * b[2]);
}

/* blockify + sort_block (realcal.F90:439-525): stable counting sort */
static void cells_sort(const cellgrid *g, int n, const double *pos, const double *L, int *order, int *psum) {
  int nc = g->bs[0] * g->bs[1] * g->bs[2];
  int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i <= nc; ++i) psum[i] = 0;
  for (int i = 0; i < n; ++i) {
    cell[i] = cell_of(g, pos + 3 * i, L);
    psum[cell[i] + 1]++;
  }
  for (int i = 0; i < nc; ++i) psum[i + 1] += psum[i];
  int *fill = (int *)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
  for (int i = 0; i < nc; ++i) fill[i] = psum[i];
  for (int i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
  free(fill);
  free(cell);
}

static void plist_alloc(plist *l, int cap) {
  l->eps = (double *)malloc(sizeof(double) * cap);
  l->sig2 = (double *)malloc(sizeof(double) * cap);
  l->dist = (double *)malloc(sizeof(double) * cap);
  l->bel = (int *)malloc(sizeof(int) * cap);
  l->n = 0;
}
static void plist_free(plist *l) {
  free(l->eps);
  free(l->sig2);
  free(l->dist);
  free(l->bel);
}

/* get_pair_energy_block (realcal.F90:585-784) over solvent cells [vb, ve) */
static void pair_block(orc_ctx *c, cellgrid *g, const ljconst *k, int upos, int vb, int ve, double *eng) {
  const orc_system *s = &c->s;
  double half[3] = {0.5 * c->L[0], 0.5 * c->L[1], 0.5 * c->L[2]};
  g->low.n = g->sw.n = g->el.n = 0;
  int nt = s->ntype;
  for (int ui = g->u_psum[upos]; ui < g->u_psum[upos + 1]; ++ui) {
    int ua = g->u_atom[ui];
    int tu = c->ljtype[ua];
    const double *cu = g->u_pos + 3 * ui;
    for (int vi = g->v_psum[vb]; vi < g->v_psum[ve]; ++vi) {
      int va = g->v_atom[vi];
      int tv = c->ljtype[va];
      double dist = minimg2(cu, g->v_pos + 3 * vi, half);
      double eps = c->lj_eps[tv * nt + tu];
      if (eps > 0) {
        if (dist <= k->lw2) {
          int n = g->low.n++;
          g->low.eps[n] = eps;
          g->low.sig2[n] = c->lj_sig2[tv * nt + tu];
          g->low.dist[n] = dist;
          g->low.bel[n] = g->v_bel[vi];
        } else if (dist <= k->up2) {
          int n = g->sw.n++;
          g->sw.eps[n] = eps;
          g->sw.sig2[n] = c->lj_sig2[tv * nt + tu];
          g->sw.dist[n] = dist;
          g->sw.bel[n] = g->v_bel[vi];
        }
      }
      if (dist <= k->el2) {
        int n = g->el.n++;
        g->el_q[n] = c->charge[ua] * c->charge[va];
        g->el.dist[n] = dist;
        g->el.bel[n] = g->v_bel[vi];
      }
    }
  }
  for (int i = 0; i < g->low.n; ++i) eng[g->low.bel[i]] += lj_low(s, k, g->low.eps[i], g->low.sig2[i], g->low.dist[i]);
  for (int i = 0; i < g->sw.n; ++i) eng[g->sw.bel[i]] += lj_switch(s, k, g->sw.eps[i], g->sw.sig2[i], g->sw.dist[i]);
  for (int i = 0; i < g->el.n; ++i) {
    double r = sqrt(g->el.dist[i]);
    eng[g->el.bel[i]] += g->el_q[i] * (1.0 - erf(s->screen * r)) / r;
  }
}

int32_t orc_realcal_proc(orc_ctx *c, int32_t solute, const double *solute_pos, double *ereal) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  const orc_system *s = &c->s;
  cellgrid g;
  memset(&g, 0, sizeof g);
  cells_setup(c, &g);
  int nc = g.bs[0] * g.bs[1] * g.bs[2];
  /* solvent atoms in tagpt order, excluding the solute (set_solv_atoms) */
  int nv = 0;
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    if (m == solute) continue;
    nv += c->mol_begin[m + 1] - c->mol_begin[m];
  }
  int *vat = (int *)malloc(sizeof(int) * (nv ? nv : 1));
  int *vbel = (int *)malloc(sizeof(int) * (nv ? nv : 1));
  double *vpos = (double *)malloc(sizeof(double) * 3 * (nv ? nv : 1));
  int n = 0;
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    if (m == solute) continue;
    for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a) {
      vat[n] = a;
      vbel[n] = k;
      memcpy(vpos + 3 * n, c->posn + 3 * a, sizeof(double) * 3);
      n++;
    }
  }
  int *vord = (int *)malloc(sizeof(int) * (nv ? nv : 1));
  int *uord = (int *)malloc(sizeof(int) * ns);
  g.v_psum = (int *)malloc(sizeof(int) * (nc + 1));
  g.u_psum = (int *)malloc(sizeof(int) * (nc + 1));
  cells_sort(&g, nv, vpos, c->L, vord, g.v_psum);
  cells_sort(&g, ns, c->sposn, c->L, uord, g.u_psum);
  g.v_atom = (int *)malloc(sizeof(int) * (nv ? nv : 1));
  g.v_bel = (int *)malloc(sizeof(int) * (nv ? nv : 1));
  g.v_pos = (double *)malloc(sizeof(double) * 3 * (nv ? nv : 1));
  for (int i = 0; i < nv; ++i) {
    g.v_atom[i] = vat[vord[i]];
    g.v_bel[i] = vbel[vord[i]];
    memcpy(g.v_pos + 3 * i, vpos + 3 * vord[i], sizeof(double) * 3);
  }
  g.u_atom = (int *)malloc(sizeof(int) * ns);
  g.u_pos = (double *)malloc(sizeof(double) * 3 * ns);
  int b = c->mol_begin[solute];
  for (int i = 0; i < ns; ++i) {
    g.u_atom[i] = b + uord[i];
    memcpy(g.u_pos + 3 * i, c->sposn + 3 * uord[i], sizeof(double) * 3);
  }
  /* list capacity: max solute cell x max solvent pillar (realcal.F90:100-109) */
  int maxu = 0, maxv = 0;
  for (int i = 0; i < nc; ++i) {
    int d = g.u_psum[i + 1] - g.u_psum[i];
    if (d > maxu) maxu = d;
  }
  for (int z = 0; z < g.bs[2]; ++z)
    for (int y = 0; y < g.bs[1]; ++y) {
      int base = g.bs[0] * (y + g.bs[1] * z);
      int d = g.v_psum[base + g.bs[0]] - g.v_psum[base];
      if (d > maxv) maxv = d;
    }
  int cap = maxu * maxv;
  if (cap < 1) cap = 1;
  plist_alloc(&g.low, cap);
  plist_alloc(&g.sw, cap);
  plist_alloc(&g.el, cap);
  g.el_q = (double *)malloc(sizeof(double) * cap);
  ljconst k;
  make_ljconst(s, &k);
  for (int i = 0; i < c->slvmax; ++i) ereal[i] = 0.0;
  /* get_pair_energy (realcal.F90:527-582) */
  for (int u3 = 0; u3 < g.bs[2]; ++u3)
    for (int u2 = 0; u2 < g.bs[1]; ++u2)
      for (int u1 = 0; u1 < g.bs[0]; ++u1) {
        int upos = u1 + g.bs[0] * (u2 + g.bs[1] * u3);
        if (g.u_psum[upos + 1] == g.u_psum[upos]) continue;
        for (int i = 0; i < g.npill; ++i) {
          int v2 = (u2 + g.pill_y[i]) % g.bs[1];
          int v3 = (u3 + g.pill_z[i]) % g.bs[2];
          int vbase = g.bs[0] * (v2 + g.bs[1] * v3);
          int xlen = g.pill_xlen[i];
          if (2 * xlen + 1 >= g.bs[0]) {
            pair_block(c, &g, &k, upos, vbase, vbase + g.bs[0], ereal);
          } else {
            int vbeg = vbase + u1 - xlen, vend = vbase + u1 + xlen + 1, vline = vbase + g.bs[0];
            if (vbeg < vbase) {
              pair_block(c, &g, &k, upos, vbeg + g.bs[0], vline, ereal);
              pair_block(c, &g, &k, upos, vbase, vend, ereal);
            } else if (vend > vline) {
              pair_block(c, &g, &k, upos, vbeg, vline, ereal);
              pair_block(c, &g, &k, upos, vbase, vend - g.bs[0], ereal);
            } else {
              pair_block(c, &g, &k, upos, vbeg, vend, ereal);
            }
          }
        }
      }
  plist_free(&g.low);
  plist_free(&g.sw);
  plist_free(&g.el);
  free(g.el_q);
  free(vat);
  free(vbel);
  free(vpos);
  free(vord);
  free(uord);
  free(g.v_psum);
  free(g.u_psum);
  free(g.v_atom);
  free(g.v_bel);
  free(g.v_pos);
  free(g.u_atom);
  free(g.u_pos);
  free(g.pill_y);
  free(g.pill_z);
  free(g.pill_xlen);
  return ORC_OK;
}

/* The same sum as orc_realcal_proc by an O(N_s N_v) minimum-image double
 * loop with no cell list: an independent check of the cell traversal. */
int32_t orc_realcal_brute(orc_ctx *c, int32_t solute, const double *solute_pos, double *ereal) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  const orc_system *s = &c->s;
  ljconst k;
  make_ljconst(s, &k);
  double half[3] = {0.5 * c->L[0], 0.5 * c->L[1], 0.5 * c->L[2]};
  int b = c->mol_begin[solute], nt = s->ntype;
  for (int kk = 0; kk < c->slvmax; ++kk) {
    int m = c->tagpt[kk];
    double e = 0.0;
    if (m != solute) {
      for (int j = 0; j < ns; ++j)
        for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a) {
          double dist = minimg2(c->sposn + 3 * j, c->posn + 3 * a, half);
          int tu = c->ljtype[b + j], tv = c->ljtype[a];
          double eps = c->lj_eps[tv * nt + tu], sig2 = c->lj_sig2[tv * nt + tu];
          if (eps > 0) {
            if (dist <= k.lw2)
              e += lj_low(s, &k, eps, sig2, dist);
            else if (dist <= k.up2)
              e += lj_switch(s, &k, eps, sig2, dist);
          }
          if (dist <= k.el2) {
            double r = sqrt(dist);
            e += c->charge[b + j] * c->charge[a] * (1.0 - erf(s->screen * r)) / r;
          }
        }
    }
    ereal[kk] = e;
  }
  return ORC_OK;
}

/* LJ of one pair as realcal_bare writes it (realcal.F90:213-266): cut
 * beyond upljcut, switched beyond lwljcut, comparisons on the distance */
static double lj_bare(const orc_system *s, const ljconst *k, double eps, double sig2, double dis2, double rst) {
  if (rst > s->upljcut) return 0.0;
  double invr2 = sig2 / dis2;
  double invr6 = invr2 * invr2 * invr2;
  switch (s->ljswitch) {
    case 0:
    case 1: {
      double eplj = 4.0 * eps * invr6 * (invr6 - 1.0);
      if (rst > s->lwljcut) {
        double swth;
        if (s->ljswitch == 0) {
          double lw2 = ipow(s->lwljcut, 2), up2 = ipow(s->upljcut, 2);
          swth = (2.0 * dis2 + up2 - 3.0 * lw2) * ipow(dis2 - up2, 2) / ipow(up2 - lw2, 3);
        } else {
          double swfac = (rst - s->lwljcut) / (s->upljcut - s->lwljcut);
          swth = 1.0 - 10.0 * ipow(swfac, 3) + 15.0 * ipow(swfac, 4) - 6.0 * ipow(swfac, 5);
        }
        eplj = swth * eplj;
      }
      return eplj;
    }
    case 2: {
      double sg6 = sig2 * sig2 * sig2, vdwa, vdwb;
      if (rst <= s->lwljcut) {
        vdwa = invr6 * invr6 - sg6 * sg6 / (k->lw6 * k->up6);
        vdwb = invr6 - sg6 / (k->lw3 * k->up3);
      } else {
        double invr3 = sqrt(invr6), sg3 = sqrt(sg6);
        vdwa = k->up6 / (k->up6 - k->lw6) * ipow(invr6 - sg6 / k->up6, 2);
        vdwb = k->up3 / (k->up3 - k->lw3) * ipow(invr3 - sg3 / k->up3, 2);
      }
      return 4.0 * eps * (vdwa - vdwb);
    }
    default: {
      double sg6 = sig2 * sig2 * sig2, vdwa, vdwb;
      if (rst <= s->lwljcut) {
        vdwa = invr6 * invr6 - sg6 * sg6 * k->repC;
        vdwb = invr6 - sg6 * k->attC;
      } else {
        double swfac = rst - s->lwljcut;
        vdwa = invr6 * invr6 - sg6 * sg6 * (k->repA * ipow(swfac, 3) + k->repB * ipow(swfac, 4) + k->repC);
        vdwb = invr6 - sg6 * (k->attA * ipow(swfac, 3) + k->attB * ipow(swfac, 4) + k->attC);
      }
      return 4.0 * eps * (vdwa - vdwb);
    }
  }
}

/* realcal_bare (realcal.F90:155-281) for every physical molecule but the
 * solute: sites of both molecules in the normalised cell, cuboid minimum
 * image and Coulomb cutoff elecut when periodic; raw coordinates and no
 * Coulomb cutoff otherwise */
int32_t orc_realcal_bare(orc_ctx *c, int32_t solute, const double *solute_pos, double *ereal) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  const orc_system *s = &c->s;
  ljconst k;
  make_ljconst(s, &k);
  const double reelcut = s->boxshp ? s->elecut : 1.0e50;
  const double half[3] = {0.5 * c->L[0], 0.5 * c->L[1], 0.5 * c->L[2]};
  int b = c->mol_begin[solute], nt = s->ntype;
  for (int kk = 0; kk < c->slvmax; ++kk) {
    int m = c->tagpt[kk];
    double pairep = 0.0;
    if (m != solute) {
      for (int j = 0; j < ns; ++j)
        for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a) {
          double dis2 = 0.0;
          for (int i = 0; i < 3; ++i) {
            double x = c->sposn[3 * j + i] - c->posn[3 * a + i];
            if (s->boxshp) x = half[i] - fabs(half[i] - fabs(x));
            dis2 = dis2 + x * x;
          }
          double rst = sqrt(dis2);
          int tu = c->ljtype[b + j], tv = c->ljtype[a];
          double eplj = lj_bare(s, &k, c->lj_eps[tv * nt + tu], c->lj_sig2[tv * nt + tu], dis2, rst);
          double epcl = (rst >= reelcut) ? 0.0 : c->charge[b + j] * c->charge[a] / rst;
          pairep = pairep + eplj + epcl;
        }
    }
    ereal[kk] = pairep;
  }
  return ORC_OK;
}

/* realcal_self (realcal.F90:284-328) */
int32_t orc_realcal_self(orc_ctx *c, int32_t solute, const double *solute_pos, double *pairep) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  double half[3] = {0.5 * c->L[0], 0.5 * c->L[1], 0.5 * c->L[2]};
  int b = c->mol_begin[solute];
  double a = c->s.screen, e = 0.0;
  for (int i = 0; i < ns; ++i) {
    double qi = c->charge[b + i];
    e = e + (-qi * qi * a / sqrt(PI_REF));
    for (int j = i + 1; j < ns; ++j) {
      double dist = minimg2(c->sposn + 3 * i, c->sposn + 3 * j, half);
      double r = sqrt(dist);
      e = e + (-qi * c->charge[b + j] * erf(a * r) / r);
    }
  }
  *pairep = e;
  return ORC_OK;
}

/* recpcal_prepare_solute (recpcal.F90:342-406) and recpcal_energy for
 * every physical molecule (recpcal.F90:451-502) */
int32_t orc_recpcal(orc_ctx *c, int32_t solute, const double *solute_pos, double *erecp, double *self_energy,
                    double *phi) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  int p = c->p, K1 = c->K[0], K2 = c->K[1], K3 = c->K[2], H = c->H;
  long nr = (long)K1 * K2 * K3;
  memset(c->cnv, 0, sizeof(double) * nr);
  int b = c->mol_begin[solute];
  double w[3 * 32];
  int32_t g[3];
  for (int sid = 0; sid < ns; ++sid) {
    orc_spline_site(c->spos + 3 * sid, c->invcl, c->K, p, c->coeff, w, g);
    double chr = c->charge[b + sid];
    for (int c3 = 0; c3 < p; ++c3)
      for (int c2 = 0; c2 < p; ++c2)
        for (int c1 = 0; c1 < p; ++c1) {
          int r1 = ((g[0] - c1) % K1 + K1) % K1;
          int r2 = ((g[1] - c2) % K2 + K2) % K2;
          int r3 = ((g[2] - c3) % K3 + K3) % K3;
          double f = chr * w[0 * p + c1] * w[1 * p + c2] * w[2 * p + c3];
          c->cnv[((long)r3 * K2 + r2) * K1 + r1] += f;
        }
  }
  orc_fft_r2c_3d(c->K, c->cnv, c->rcp);
  /* self energy on the half spectrum with symmetry weights (385-394) */
  double s_mid = 0.0, s_0 = 0.0, s_last = 0.0;
  int last = H - 1;
  int even = (K1 % 2 == 0);
  for (int r3 = 0; r3 < K3; ++r3)
    for (int r2 = 0; r2 < K2; ++r2)
      for (int r1 = 1; r1 <= (even ? last - 1 : last); ++r1) {
        long i = ((long)r3 * K2 + r2) * H + r1;
        double re = c->rcp[2 * i], im = c->rcp[2 * i + 1];
        s_mid = s_mid + c->engfac[i] * (re * re - im * (-im));
      }
  for (int r3 = 0; r3 < K3; ++r3)
    for (int r2 = 0; r2 < K2; ++r2) {
      long i = ((long)r3 * K2 + r2) * H;
      double re = c->rcp[2 * i], im = c->rcp[2 * i + 1];
      s_0 = s_0 + c->engfac[i] * (re * re - im * (-im));
      if (even) {
        long j = i + last;
        re = c->rcp[2 * j];
        im = c->rcp[2 * j + 1];
        s_last = s_last + c->engfac[j] * (re * re - im * (-im));
      }
    }
  *self_energy = even ? s_mid + 0.5 * s_0 + 0.5 * s_last : s_mid + 0.5 * s_0;
  long nh = (long)H * K2 * K3;
  for (long i = 0; i < nh; ++i) {
    c->rcp[2 * i] = c->engfac[i] * c->rcp[2 * i];
    c->rcp[2 * i + 1] = c->engfac[i] * c->rcp[2 * i + 1];
  }
  orc_fft_c2r_3d(c->K, c->rcp, c->cnv);
  if (phi) memcpy(phi, c->cnv, sizeof(double) * nr);
  /* gather */
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    double e = 0.0;
    if (m != solute) {
      int svi = c->slvtag[m];
      if (svi < 0) return ORC_E_RCP_CNS;
      int nsm = c->mol_begin[m + 1] - c->mol_begin[m];
      for (int sid = 0; sid < nsm; ++sid) {
        int slot = svi + sid;
        const double *sw = c->splslv + (long)3 * p * slot;
        const int32_t *gg = c->grdslv + 3 * slot;
        double chr = c->charge[c->mol_begin[m] + sid];
        for (int c3 = 0; c3 < p; ++c3) {
          double f1 = chr * sw[2 * p + c3];
          int r3 = ((gg[2] - c3) % K3 + K3) % K3;
          for (int c2 = 0; c2 < p; ++c2) {
            double f2 = f1 * sw[1 * p + c2];
            int r2 = ((gg[1] - c2) % K2 + K2) % K2;
            for (int c1 = 0; c1 < p; ++c1) {
              double f3 = f2 * sw[0 * p + c1];
              int r1 = ((gg[0] - c1) % K1 + K1) % K1;
              e = e + f3 * c->cnv[((long)r3 * K2 + r2) * K1 + r1];
            }
          }
        }
      }
    }
    erecp[k] = e;
  }
  return ORC_OK;
}

/* residual_ene (engproc.F90:859-869) */
static double residual(const orc_ctx *c, int i, int j) {
  double epcl = PI_REF * c->mol_charge[i] * c->mol_charge[j] / c->s.screen / c->s.screen / c->volume;
  if (i == j) epcl = epcl / 2.0;
  return -epcl;
}

/* get_uv_energy (engproc.F90:662-756): the PME / PPPM branch, or with bare
 * Coulomb realcal_bare per molecule and a zero self energy (realcal_self
 * and residual_ene return 0, realcal.F90:294, engproc.F90:865) */
int32_t orc_uv_energy(orc_ctx *c, int32_t solute, const double *solute_pos, double *uvengy) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  int n = c->slvmax;
  if (c->s.cltype == 0) {
    uvengy[0] = 0.0;
    return orc_realcal_bare(c, solute, solute_pos, uvengy + 1);
  }
  double *ereal = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *erecp = (double *)malloc(sizeof(double) * (n ? n : 1));
  double self_recp;
  st = orc_recpcal(c, solute, solute_pos, erecp, &self_recp, NULL);
  if (!st) st = orc_realcal_proc(c, solute, solute_pos, ereal);
  if (st) {
    free(ereal);
    free(erecp);
    return st;
  }
  uvengy[0] = self_recp;
  for (int k = 0; k < n; ++k) uvengy[k + 1] = ereal[k];
  /* realcal_self is reused while the solute hash is unchanged, and always
   * after the first call for a rigid test particle (engproc.F90:712-730) */
  double pairep;
  uint64_t h = orc_hash_double(c->spos, 3 * ns);
  if (h == c->solute_hash || (c->s.slttype == 2 && c->solute_hash != 0)) {
    pairep = c->usreal;
  } else {
    orc_realcal_self(c, solute, solute_pos, &pairep);
    c->usreal = pairep;
  }
  c->solute_hash = h;
  double res = residual(c, solute, solute);
  uvengy[0] = uvengy[0] + pairep + res;
  for (int k = 0; k < n; ++k) {
    int m = c->tagpt[k];
    if (m == solute) continue;
    double pe = 0.0 + residual(c, solute, m);
    pe = pe + erecp[k];
    uvengy[k + 1] = uvengy[k + 1] + pe;
  }
  free(ereal);
  free(erecp);
  return ORC_OK;
}

/* getiduv (engproc.F90:960-1020); returns status, writes the 0-based bin */
static int getiduv(orc_ctx *c, int pti, double v, int *iduv) {
  int idmin = 0, idnum, idpti;
  if (pti > 0) {
    for (int i = 0; i < pti - 1; ++i) idmin += c->uvmax[i];
    idnum = c->uvmax[pti - 1];
    idpti = orc_binsearch(c->uvcrd + idmin, idnum, v);
  } else if (pti == 0) {
    idnum = c->s.esmax;
    idpti = orc_binsearch(c->escrd, idnum, v);
  } else {
    return ORC_E_ENG_BUG;
  }
  if (idpti <= 0) return ORC_E_ENG_MIN;
  if (idpti > idnum) return ORC_E_ENG_BUG;
  if (c->s.slttype == 1 && pti > 0 && idpti > c->uvsoft[pti - 1]) return ORC_E_ENG_SFT;
  if (idpti == idnum && v < 1.0e3 && c->warn_bin_first) {
    c->warn_bin_first = 0;
    if (!c->s.force_calculation) return ORC_E_MBIN;
  }
  *iduv = idmin + idpti - 1;
  return ORC_OK;
}

/* volcorrect (engproc.F90:871-887) */
static double volcorrect(const orc_ctx *c, double weight) {
  weight = weight * c->volume;
  if (c->s.cltype == 0) return weight;
  double tq = 0.0;
  for (int i = 0; i < c->s.nmol; ++i)
    if (c->sluvid[i] == 0 || c->sluvid[i] == 1) tq = tq + c->mol_charge[i];
  double factor = PI_REF * ipow(tq, 2) / ipow(c->s.screen, 2) / c->volume / 2.0;
  return weight * exp(factor / c->s.temp);
}

/* update_histogram (engproc.F90:758-856) */
int32_t orc_update_histogram(orc_ctx *c, int32_t solute, double stat_weight_solute, const double *uvengy,
                             int32_t *insdst) {
  const orc_system *s = &c->s;
  double engnmfc;
  if (s->wgtslf == 0) {
    engnmfc = 1.0;
  } else {
    if (!c->voffset_init) {
      c->voffset = uvengy[0];
      c->voffset_init = 1;
    }
    double f = uvengy[0] - c->voffset;
    engnmfc = (s->slttype == 1) ? exp(f / s->temp) : exp(-f / s->temp);
  }
  double tw = c->stat_weight_system * stat_weight_solute;
  if (s->estype == 2) tw = volcorrect(c, tw);
  engnmfc = engnmfc * tw;
  c->engnorm = c->engnorm + engnmfc;
  c->engsmpl = c->engsmpl + 1.0;
  c->avslf = c->avslf + tw;
  int iduv;
  if (s->selfcal) {
    int st = getiduv(c, 0, uvengy[0], &iduv);
    if (st) return st;
    c->eself[iduv] = c->eself[iduv] + engnmfc;
  }
  if (uvengy[0] < c->minuv[0]) c->minuv[0] = uvengy[0];
  if (uvengy[0] > c->maxuv[0]) c->maxuv[0] = uvengy[0];
  memset(c->insdst, 0, sizeof(int32_t) * c->ermax);
  for (int i = 0; i <= s->numslv; ++i) c->flceng[i] = 0.0;
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    if (m == solute) continue;
    int pti = c->uvspec[m];
    if (pti <= 0) return ORC_E_ENG_CNS;
    double e = uvengy[k + 1];
    int st = getiduv(c, pti, e, &iduv);
    if (st) return st;
    c->insdst[iduv] += 1;
    c->flceng[pti] = c->flceng[pti] + e;
    if (e < c->minuv[pti]) c->minuv[pti] = e;
    if (e > c->maxuv[pti]) c->maxuv[pti] = e;
  }
  if (s->slttype == 1)
    for (int i = 1; i <= s->numslv; ++i) c->slnuv[i] = c->slnuv[i] + c->flceng[i] * engnmfc;
  for (int i = 0; i < c->ermax; ++i) {
    int k = c->insdst[i];
    if (k == 0) continue;
    c->edens[i] = c->edens[i] + engnmfc * (double)k;
  }
  if (s->corrcal) {
    for (int i = 0; i < c->ermax; ++i) {
      int k = c->insdst[i];
      if (k == 0) continue;
      for (int j = 0; j < c->ermax; ++j) {
        int q = c->insdst[j];
        if (q == 0) continue;
        /* ecorr(iduvp, iduv), column-major: element (j, i) at j + i*ermax */
        c->ecorr[(long)i * c->ermax + j] = c->ecorr[(long)i * c->ermax + j] + engnmfc * (double)k * (double)q;
      }
    }
  }
  /* the intermediate state: configurations with no molecule in a bin that
   * is empty in the reference solvent (extensions/Intermediate/
   * engproc.F90:996-1009) */
  if (c->zerodns) {
    c->soln_total = c->soln_total + engnmfc;
    int hit = 0;
    for (int i = 0; i < c->ermax; ++i)
      if (c->insdst[i] > 0 && c->zerodns[i]) hit = 1;
    if (!hit) {
      c->soln_intm = c->soln_intm + engnmfc;
      c->norm_intm = c->norm_intm + engnmfc;
      for (int i = 0; i < c->ermax; ++i)
        if (c->insdst[i] > 0) c->edintm[i] = c->edintm[i] + engnmfc * (double)c->insdst[i];
      for (int i = 1; i <= s->numslv; ++i) c->uvintm[i] = c->uvintm[i] + c->flceng[i] * engnmfc;
    }
  }
  if (insdst) memcpy(insdst, c->insdst, sizeof(int32_t) * c->ermax);
  return ORC_OK;
}

int32_t orc_set_intermediate(orc_ctx *c, const int32_t *zerodns) {
  if (c->s.slttype != 1) return ORC_E_ARG;
  free(c->zerodns);
  free(c->edintm);
  free(c->uvintm);
  c->zerodns = NULL;
  c->edintm = c->uvintm = NULL;
  if (!zerodns) return ORC_OK;
  c->zerodns = dupi(zerodns, c->ermax);
  c->edintm = (double *)calloc(c->ermax ? c->ermax : 1, sizeof(double));
  c->uvintm = (double *)calloc(c->s.numslv + 1, sizeof(double));
  c->norm_intm = c->soln_intm = c->soln_total = 0.0;
  return ORC_OK;
}

int32_t orc_get_intermediate(const orc_ctx *c, double *edintm, double *uvintm, double *scal) {
  if (!c->zerodns) return ORC_E_ARG;
  if (edintm) memcpy(edintm, c->edintm, sizeof(double) * c->ermax);
  if (uvintm) memcpy(uvintm, c->uvintm, sizeof(double) * (c->s.numslv + 1));
  if (scal) {
    scal[0] = c->norm_intm;
    scal[1] = c->soln_intm;
    scal[2] = c->soln_total;
  }
  return ORC_OK;
}

/* engstore normalisation for one process (engproc.F90:539-553) */
int32_t orc_engstore(orc_ctx *c) {
  for (int i = 0; i < c->ermax; ++i) c->edens[i] = c->edens[i] / c->engnorm;
  if (c->ecorr)
    for (long i = 0; i < (long)c->ermax * c->ermax; ++i) c->ecorr[i] = c->ecorr[i] / c->engnorm;
  if (c->eself)
    for (int i = 0; i < c->s.esmax; ++i) c->eself[i] = c->eself[i] / c->engnorm;
  c->avslf = c->avslf / c->engnorm;
  if (c->zerodns)
    for (int i = 0; i < c->ermax; ++i) c->edintm[i] = c->edintm[i] / c->norm_intm;
  return ORC_OK;
}

void orc_get_acc(orc_ctx *c, orc_acc *o) {
  o->edens = c->edens;
  o->ecorr = c->ecorr;
  o->eself = c->eself;
  o->minuv = c->minuv;
  o->maxuv = c->maxuv;
  o->slnuv = c->slnuv;
  o->engnorm = c->engnorm;
  o->engsmpl = c->engsmpl;
  o->avslf = c->avslf;
  o->voffset = c->voffset;
  o->ermax = c->ermax;
  o->esmax = c->s.esmax;
  o->numslv = c->s.numslv;
  o->voffset_initialized = c->voffset_init;
}

/* ---------------------------------------------------------------------- */
/* Gradient of the solute-solvent energy u = sum_{k>=1} uvengy(k) with     */
/* respect to the solute's per-site charge q_i, LJ epsilon_i and sigma_i. */
/* The reference has no gradient; this is the analytic derivative of the  */
/* energy that get_uv_energy (engproc.F90:662-756) computes:              */
/*  - real space (realcal.F90:676-779): pairs under the minimum image,    */
/*    every LJ variant written as rep - att with rep proportional to      */
/*    sigma^12 and att to sigma^6, so dU/d(sigma^2) = (6 rep - 3 att) /   */
/*    sigma^2; epsilon_iv = sqrt(eps_i eps_v) (setconf.F90:820), sigma_iv */
/*    arithmetic (cmbrule 0) or geometric (1) (setconf.F90:810-817);      */
/*  - reciprocal space: the PME energy is the symmetric bilinear form     */
/*    q_slt^T W^T Phi W q_slv (recpcal.F90:342-502), so dE/dq_i is the    */
/*    solvent's PME potential interpolated at site i;                      */
/*  - residual_ene (engproc.F90:859-869): -pi Q_slt Q_j / (alpha^2 V).     */
/* Solvent LJ parameters are the diagonals of the type tables.            */
/* ---------------------------------------------------------------------- */
static void lj_parts(const orc_system *s, const ljconst *k, double eps, double sig2, double dist, double *rep,
                     double *att) {
  double invr2 = sig2 / dist;
  double invr6 = invr2 * invr2 * invr2;
  double sg6 = sig2 * sig2 * sig2;
  int low = dist <= k->lw2;
  switch (s->ljswitch) {
    case 0:
    case 1: {
      double sw = 1.0;
      if (!low) {
        if (s->ljswitch == 0) {
          sw = (2.0 * dist + k->up2 - 3.0 * k->lw2) * ipow(dist - k->up2, 2) / ipow(k->up2 - k->lw2, 3);
        } else {
          double x = (sqrt(dist) - s->lwljcut) / (s->upljcut - s->lwljcut);
          sw = 1.0 - 10.0 * ipow(x, 3) + 15.0 * ipow(x, 4) - 6.0 * ipow(x, 5);
        }
      }
      *rep = 4.0 * eps * invr6 * invr6 * sw;
      *att = 4.0 * eps * invr6 * sw;
      return;
    }
    case 2:
      if (low) {
        *rep = 4.0 * eps * (invr6 * invr6 - sg6 * sg6 / (k->lw6 * k->up6));
        *att = 4.0 * eps * (invr6 - sg6 / (k->lw3 * k->up3));
      } else {
        double invr3 = sqrt(invr6), sg3 = sqrt(sg6);
        *rep = 4.0 * eps * k->up6 / (k->up6 - k->lw6) * ipow(invr6 - sg6 / k->up6, 2);
        *att = 4.0 * eps * k->up3 / (k->up3 - k->lw3) * ipow(invr3 - sg3 / k->up3, 2);
      }
      return;
    default: {
      double a = k->repC, b = k->attC;
      if (!low) {
        double sw = sqrt(dist) - s->lwljcut;
        a = k->repA * ipow(sw, 3) + k->repB * ipow(sw, 4) + k->repC;
        b = k->attA * ipow(sw, 3) + k->attB * ipow(sw, 4) + k->attC;
      }
      *rep = 4.0 * eps * (invr6 * invr6 - sg6 * sg6 * a);
      *att = 4.0 * eps * (invr6 - sg6 * b);
      return;
    }
  }
}

int32_t orc_solute_gradient(orc_ctx *c, int32_t solute, const double *solute_pos, int32_t cmbrule,
                            const double *q_in, const double *eps_in, const double *sig_in, double *out) {
  int ns, st = solute_coords(c, solute, solute_pos, &ns);
  if (st) return st;
  if (cmbrule != 0 && cmbrule != 1) return ORC_E_ARG;
  if (c->s.cltype != 2 && c->s.cltype != 3) return ORC_E_UNSUPPORTED;
  const orc_system *s = &c->s;
  const int nt = s->ntype, b0 = c->mol_begin[solute];
  double *q = (double *)malloc(sizeof(double) * 3 * ns), *ep = q + ns, *sg = q + 2 * ns;
  for (int i = 0; i < ns; ++i) {
    int t = c->ljtype[b0 + i];
    q[i] = q_in ? q_in[i] : c->charge[b0 + i];
    ep[i] = eps_in ? eps_in[i] : c->lj_eps[t * nt + t];
    sg[i] = sig_in ? sig_in[i] : sqrt(c->lj_sig2[t * nt + t]);
  }
  double *gq = out + 1, *ge = out + 1 + ns, *gs = out + 1 + 2 * ns;
  double u = 0.0;
  for (int i = 0; i < 3 * ns; ++i) out[1 + i] = 0.0;
  double *elj = (double *)calloc((size_t)ns, sizeof(double));
  /* real space */
  ljconst k;
  make_ljconst(s, &k);
  double half[3] = {0.5 * c->L[0], 0.5 * c->L[1], 0.5 * c->L[2]};
  double qsolv = 0.0;
  for (int kk = 0; kk < c->slvmax; ++kk) {
    int m = c->tagpt[kk];
    if (m == solute) continue;
    qsolv += c->mol_charge[m];
    for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a) {
      int tv = c->ljtype[a];
      double ev = c->lj_eps[tv * nt + tv], sv = sqrt(c->lj_sig2[tv * nt + tv]);
      for (int i = 0; i < ns; ++i) {
        double dist = minimg2(c->sposn + 3 * i, c->posn + 3 * a, half);
        double eiv = sqrt(ep[i] * ev);
        double s2 = cmbrule ? sg[i] * sv : ipow(0.5 * (sg[i] + sv), 2);
        if (eiv > 0 && dist <= k.up2) {
          double rep, att;
          lj_parts(s, &k, eiv, s2, dist, &rep, &att);
          double f = cmbrule ? 1.0 / sg[i] : 2.0 / (sg[i] + sv);
          u += rep - att;
          elj[i] += rep - att;
          gs[i] += (6.0 * rep - 3.0 * att) * f;
        }
        if (dist <= k.el2) {
          double r = sqrt(dist);
          double phi = c->charge[a] * (1.0 - erf(s->screen * r)) / r;
          u += q[i] * phi;
          gq[i] += phi;
        }
      }
    }
  }
  for (int i = 0; i < ns; ++i) ge[i] = ep[i] > 0 ? elj[i] / (2.0 * ep[i]) : 0.0;
  /* reciprocal space: solvent potential on the grid, gathered at the sites */
  int p = c->p, K1 = c->K[0], K2 = c->K[1], K3 = c->K[2];
  long nr = (long)K1 * K2 * K3, nh = (long)c->H * K2 * K3;
  memset(c->cnv, 0, sizeof(double) * nr);
  for (int kk = 0; kk < c->slvmax; ++kk) {
    int m = c->tagpt[kk];
    if (m == solute) continue;
    int svi = c->slvtag[m];
    if (svi < 0) {
      free(q);
      free(elj);
      return ORC_E_RCP_CNS;
    }
    for (int sid = 0; sid < c->mol_begin[m + 1] - c->mol_begin[m]; ++sid) {
      const double *sw = c->splslv + (long)3 * p * (svi + sid);
      const int32_t *gg = c->grdslv + 3 * (svi + sid);
      double chr = c->charge[c->mol_begin[m] + sid];
      for (int c3 = 0; c3 < p; ++c3)
        for (int c2 = 0; c2 < p; ++c2)
          for (int c1 = 0; c1 < p; ++c1) {
            int r1 = ((gg[0] - c1) % K1 + K1) % K1;
            int r2 = ((gg[1] - c2) % K2 + K2) % K2;
            int r3 = ((gg[2] - c3) % K3 + K3) % K3;
            c->cnv[((long)r3 * K2 + r2) * K1 + r1] += chr * sw[c1] * sw[p + c2] * sw[2 * p + c3];
          }
    }
  }
  orc_fft_r2c_3d(c->K, c->cnv, c->rcp);
  for (long i = 0; i < nh; ++i) {
    c->rcp[2 * i] *= c->engfac[i];
    c->rcp[2 * i + 1] *= c->engfac[i];
  }
  orc_fft_c2r_3d(c->K, c->rcp, c->cnv);
  double w[3 * 32];
  int32_t g[3];
  for (int i = 0; i < ns; ++i) {
    orc_spline_site(c->spos + 3 * i, c->invcl, c->K, p, c->coeff, w, g);
    double phi = 0.0;
    for (int c3 = 0; c3 < p; ++c3)
      for (int c2 = 0; c2 < p; ++c2)
        for (int c1 = 0; c1 < p; ++c1) {
          int r1 = ((g[0] - c1) % K1 + K1) % K1;
          int r2 = ((g[1] - c2) % K2 + K2) % K2;
          int r3 = ((g[2] - c3) % K3 + K3) % K3;
          phi += w[c1] * w[p + c2] * w[2 * p + c3] * c->cnv[((long)r3 * K2 + r2) * K1 + r1];
        }
    u += q[i] * phi;
    gq[i] += phi;
  }
  /* residual */
  double qslt = 0.0;
  for (int i = 0; i < ns; ++i) qslt += q[i];
  double rf = -PI_REF * qsolv / s->screen / s->screen / c->volume;
  u += qslt * rf;
  for (int i = 0; i < ns; ++i) gq[i] += rf;
  out[0] = u;
  free(q);
  free(elj);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: engconst's test-particle loop (engproc.F90:388-392: pose,  */
/* get_uv_energy, update_histogram) over the host's cores.  Insertion i     */
/* goes to solute type i mod ntypes and runs on the context of its OpenMP   */
/* thread for that type: ctxs[thread * ntypes + type], each holding the     */
/* frame, its own scratch and its own accumulators (the reference's threads */
/* share one insertion and split its molecules; splitting the insertions    */
/* instead needs no synchronisation and is the faster CPU schedule).        */
/* ------------------------------------------------------------------------ */
int32_t orc_bench_insertions(orc_ctx *const *ctxs, int32_t nthreads, int32_t ntypes,
                             const int32_t *solute_mol, const int32_t *nsite,
                             const double *const *bfcoord, const double *const *mass,
                             const double *cell, int64_t n, uint64_t seed, uint64_t first_index,
                             double seconds, int64_t *done_per_type) {
  if (nthreads < 1 || ntypes < 1 || n < 0) return ORC_E_ARG;
  int nsmax = 0, smax = 0;
  for (int t = 0; t < ntypes; ++t) {
    if (nsite[t] > nsmax) nsmax = nsite[t];
    if (ctxs[t]->slvmax > smax) smax = ctxs[t]->slvmax;
  }
  for (int t = 0; t < ntypes; ++t) done_per_type[t] = 0;
  int32_t status = ORC_OK;
  const double t_end = omp_get_wtime() + seconds;
  int stop = 0;
#pragma omp parallel num_threads(nthreads)
  {
    const int th = omp_get_thread_num();
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)nsmax);
    double *uv = (double *)malloc(sizeof(double) * ((size_t)smax + 1));
    int64_t *mine = (int64_t *)calloc((size_t)ntypes, sizeof(int64_t));
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < n; ++i) {
      int halt;
#pragma omp atomic read
      halt = stop;
      if (halt) continue;
      const int ty = (int)(i % ntypes);
      orc_ctx *c = ctxs[(size_t)th * ntypes + ty];
      double r[3], q[4];
      orc_random_pose(seed, first_index + (uint64_t)i, r, q);
      orc_place_solute(nsite[ty], bfcoord[ty], mass[ty], cell, r, q, pos);
      int32_t st = orc_uv_energy(c, solute_mol[ty], pos, uv);
      if (!st) st = orc_update_histogram(c, solute_mol[ty], 1.0, uv, NULL);
      if (st) {
#pragma omp critical
        status = st;
#pragma omp atomic write
        stop = 1;
      }
      ++mine[ty];
      if (omp_get_wtime() > t_end) {
#pragma omp atomic write
        stop = 1;
      }
    }
#pragma omp critical
    for (int t = 0; t < ntypes; ++t) done_per_type[t] += mine[t];
    free(pos);
    free(uv);
    free(mine);
  }
  return status;
}
