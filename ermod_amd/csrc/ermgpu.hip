// ermgpu.hip -- host side of libermgpu.so: contexts, device buffers, the
// per-frame and per-batch launch sequences and the C-ABI of ermgpu.h.
#include "ermgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ermgpu_ctx.h"
#include "ermgpu_grad.h"
#include "ermgpu_traj.h"

using namespace erm;
using namespace ermh;

// Scratch of erm_update_histogram_static: a file-scope device array, one copy
// per GPU (each at its own address), held until the process ends.  Host code
// never passes `scratch` itself to a kernel or copy: its device address is
// resolved once per device (scratch_init, from erm_create) and used instead.
// Calls share their GPU's copy and it keeps the last call's contents, so
// every call clears what must start at zero, and calls that overlap on one
// GPU race.
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[SCRATCH_DEVICES] = {};

// device address of `scratch` on `device`, which must be current
int scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= SCRATCH_DEVICES) return ERM_E_UNSUPPORTED;
  if (g_scratch[device]) return ERM_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return ERM_E_DEVICE;
  }
  g_scratch[device] = static_cast<unsigned char *>(p);
  return ERM_OK;
}
unsigned char *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < SCRATCH_DEVICES) ? g_scratch[device] : nullptr;
}
}  // namespace

namespace {
size_t zcol_lds(int ncol, int K3, size_t vb);
const double PI_REF = 3.1415926535897932;


// ---------------------------------------------------------------- host math
// spline.F90:26-59 (coefficients) -- used for the B-spline moduli
void spline_coeff(int order, double *c) {
  for (int i = 0; i <= order; ++i) {
    double f = 1.0;
    for (int k = 1; k <= i; ++k) f = f * (double)k;
    for (int k = 1; k <= order - i; ++k) f = f * (double)k;
    f = (double)order / f;
    if (i % 2 == 1) f = -f;
    c[i] = f;
  }
}
double ipow_h(double x, int n) {
  double r = 1.0, b = x;
  bool first = true;
  while (n > 0) {
    if (n & 1) {
      r = first ? b : r * b;
      first = false;
    }
    n >>= 1;
    if (n) b = b * b;
  }
  return first ? 1.0 : r;
}
double spline_value_h(const double *c, int order, double x) {
  double f = 0.0;
  if (x > 0.0 && x < (double)order) {
    int k = (int)x;
    for (int i = 0; i <= k; ++i) f = f + c[i] * ipow_h(x - (double)i, order - 1);
  }
  return f;
}

// init_spline_axis (recpcal.F90:87-109)
void spline_axis(int K, int p, const double *coeff, double *splfc) {
  for (int rci = 0; rci < K; ++rci) {
    double re = 0.0, im = 0.0;
    for (int spi = 0; spi <= p - 2; ++spi) {
      double chr = spline_value_h(coeff, p, (double)(spi + 1));
      double rtp2 = 2.0 * PI_REF * (double)(spi * rci) / (double)K;
      re = re + chr * cos(rtp2);
      im = im + chr * sin(rtp2);
    }
    splfc[rci] = re * re - im * (-im);
  }
}

// recpcal_spline_greenfunc (recpcal.F90:111-152); engfac[m3][m2][m1]
void greenfunc(const int *K, int p, const double *invcl, double volume, double screen,
               const std::vector<double> *splfc, std::vector<double> &engfac) {
  const int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
  engfac.assign((size_t)H * K2 * K3, 0.0);
  for (int rc3 = 0; rc3 < K3; ++rc3)
    for (int rc2 = 0; rc2 < K2; ++rc2)
      for (int rc1 = 0; rc1 < H; ++rc1) {
        if (rc1 == 0 && rc2 == 0 && rc3 == 0) continue;
        const int rci[3] = {rc1, rc2, rc3};
        double inm[3];
        bool skip = false;
        for (int m = 0; m < 3; ++m) {
          if ((p % 2 == 1) && (2 * std::abs(rci[m]) == K[m])) {
            skip = true;
            break;
          }
          inm[m] = (rci[m] <= K[m] / 2) ? (double)rci[m] : (double)(rci[m] - K[m]);
        }
        double factor = 0.0;
        if (!skip) {
          double xst[3];
          for (int m = 0; m < 3; ++m)
            xst[m] = invcl[3 * m + 0] * inm[0] + invcl[3 * m + 1] * inm[1] + invcl[3 * m + 2] * inm[2];
          double rtp2 = xst[0] * xst[0] + xst[1] * xst[1] + xst[2] * xst[2];
          double chr = (PI_REF * PI_REF) * rtp2 / (screen * screen);
          factor = exp(-chr) / rtp2 / PI_REF / volume;
          rtp2 = splfc[0][rc1] * splfc[1][rc2] * splfc[2][rc3];
          factor = factor / rtp2;
        }
        engfac[((size_t)rc3 * K2 + rc2) * H + rc1] = factor;
      }
  engfac[0] = 0.0;
}

// calc_gfb_pppm (recpcal.F90:168-186): coefficients b_l of the alias sum
// sum_m W^2(k + 2 pi m / h) = sum_l b_l sin^(2l)(k h / 2) for order p;
// after the inner loop its index is 0, as in the reference
void pppm_gfb(int p, double *b) {
  for (int i = 0; i < p; ++i) b[i] = 0.0;
  b[0] = 1.0;
  for (int m = 1; m < p; ++m) {
    int l = m;
    for (; l >= 1; --l)
      b[l] = 4.0 * (b[l] * (double)(l - m) * ((double)(l - m) - 0.5) -
                    b[l - 1] * (double)(l - m - 1) * (double)(l - m - 1));
    b[0] = 4.0 * (b[0] * (double)(l - m) * ((double)(l - m) - 0.5));
  }
  double f = 1.0;
  for (int i = 2; i <= 2 * p - 1; ++i) f *= (double)i;
  const double gi = 1.0 / f;
  for (int i = 0; i < p; ++i) b[i] = b[i] * gi;
}

// recpcal_pppm_greenfunc (recpcal.F90:205-325): the optimal influence
// function on the (K1/2+1, K2, K3) half spectrum, engfac[m3][m2][m1].  The
// alias range per axis is int(cell column . m_max) with m_max(d) = alpha /
// (pi K_d) (-ln 1e-7)^(1/4); the z alias offset is mz * K2 as in the
// reference (recpcal.F90:293), identical to mz * K3 when K2 = K3.
// Returns false for an empty (negative) alias range.
bool pppm_greenfunc(const int *K, int p, const double *cell, const double *invcl, double volume, double screen,
                    std::vector<double> &engfac) {
  const int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
  engfac.assign((size_t)H * K2 * K3, 0.0);
  double b[MAXP];
  pppm_gfb(p, b);
  const double hoc = std::pow(-std::log(0.0000001), 0.25);
  double mm[3];
  for (int d = 0; d < 3; ++d) mm[d] = screen / (PI_REF * (double)K[d]) * hoc;
  int lim[3];
  for (int j = 0; j < 3; ++j) lim[j] = (int)(cell[3 * j] * mm[0] + cell[3 * j + 1] * mm[1] + cell[3 * j + 2] * mm[2]);
  if (lim[0] < 0 || lim[1] < 0 || lim[2] < 0) return false;
  auto fold = [](int r, int Kd) { return (r <= Kd / 2) ? (double)r : (double)(r - Kd); };
  auto sinc2p = [p](double t) { return (t == 0.0) ? 1.0 : ipow_h(std::sin(t) / t, 2 * p); };
  auto poly = [&b, p](double x) {
    double a = 0.0;
    for (int l = p - 1; l >= 0; --l) a = b[l] + a * x;
    return a;
  };
  for (int rc3 = 0; rc3 < K3; ++rc3) {
    const double n3 = fold(rc3, K3), s3 = ipow_h(std::sin(PI_REF * n3 / (double)K3), 2);
    for (int rc2 = 0; rc2 < K2; ++rc2) {
      const double n2 = fold(rc2, K2), s2 = ipow_h(std::sin(PI_REF * n2 / (double)K2), 2);
      for (int rc1 = 0; rc1 < H; ++rc1) {
        if (rc1 == 0 && rc2 == 0 && rc3 == 0) continue;
        const double n1 = fold(rc1, K1), s1 = ipow_h(std::sin(PI_REF * n1 / (double)K1), 2);
        const double nv[3] = {n1, n2, n3};
        double kv[3];
        for (int i = 0; i < 3; ++i)
          kv[i] = 2.0 * PI_REF * (invcl[3 * i] * nv[0] + invcl[3 * i + 1] * nv[1] + invcl[3 * i + 2] * nv[2]);
        const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        if (!(k2 > 0.0)) continue;
        double acc = 0.0;
        for (int mx = -lim[0]; mx <= lim[0]; ++mx) {
          const double u1 = sinc2p(PI_REF * (n1 / (double)K1 + (double)mx));
          for (int my = -lim[1]; my <= lim[1]; ++my) {
            const double u2 = sinc2p(PI_REF * (n2 / (double)K2 + (double)my));
            for (int mz = -lim[2]; mz <= lim[2]; ++mz) {
              const double u3 = sinc2p(PI_REF * (n3 / (double)K3 + (double)mz));
              const double shift[3] = {(double)(mx * K1), (double)(my * K2), (double)(mz * K2)};
              double km[3], g[3];
              for (int i = 0; i < 3; ++i) {
                km[i] = kv[i] + 2.0 * PI_REF * (invcl[3 * i] * shift[0] + invcl[3 * i + 1] * shift[1] +
                                                invcl[3 * i + 2] * shift[2]);
                g[i] = std::exp(-0.25 * ipow_h(km[i] / screen, 2));
              }
              const double dot = kv[0] * km[0] + kv[1] * km[1] + kv[2] * km[2];
              const double kmm = km[0] * km[0] + km[1] * km[1] + km[2] * km[2];
              acc = acc + dot * (g[0] * g[1] * g[2]) * (u1 * u2 * u3) / kmm;
            }
          }
        }
        const double den = ipow_h(poly(s1) * poly(s2) * poly(s3), 2);
        engfac[((size_t)rc3 * K2 + rc2) * H + rc1] = acc * 4.0 * PI_REF / (k2 * den) / volume;
      }
    }
  }
  engfac[0] = 0.0;
  return true;
}

// ---- best fit (bestfit.F90): Horn's quaternion superposition ----------
// eigenvector of the largest eigenvalue of a symmetric 4x4 matrix (cyclic
// Jacobi rotations; the reference calls LAPACK dsyev)
void sym4_top_eigvec(const double A_in[4][4], double v[4]) {
  double A[4][4], V[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      A[i][j] = A_in[i][j];
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
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < 4; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = cs * akp - sn * akq;
          A[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 4; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = cs * apk - sn * aqk;
          A[q][k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < 4; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
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

// find_rotation_quaternion (bestfit.F90:92-146) for n points (3 x n,
// COM-centred) and weights
void fit_quaternion(int n, const double *ref, const double *mov, const double *w, double q[4]) {
  double ip[3][3];  // ip[j][i] = sum ref(i,:) mov(j,:) w
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

void com_weighted(int n, const double *x, const double *w, double c[3]) {
  double s = 0.0;
  for (int k = 0; k < n; ++k) s = s + w[k];
  for (int i = 0; i < 3; ++i) {
    double d = 0.0;
    for (int k = 0; k < n; ++k) d = d + x[3 * k + i] * w[k];
    c[i] = d / s;
  }
}

void qrot(const double *v, const double *q, double *out) {
  const double a[4] = {q[0], q[1], q[2], q[3]};
  const double b[4] = {0.0, v[0], v[1], v[2]};
  auto prod = [](const double *x, const double *y, double *p) {
    double dot = x[1] * y[1] + x[2] * y[2] + x[3] * y[3];
    p[0] = x[0] * y[0] - dot;
    double cr[3] = {x[2] * y[3] - x[3] * y[2], x[3] * y[1] - x[1] * y[3], x[1] * y[2] - x[2] * y[1]};
    for (int i = 0; i < 3; ++i) p[1 + i] = cr[i] + x[0] * y[1 + i] + y[0] * x[1 + i];
  };
  double t[4], r[4];
  const double cj[4] = {q[0], -q[1], -q[2], -q[3]};
  prod(a, b, t);
  prod(t, cj, r);
  for (int i = 0; i < 3; ++i) out[i] = r[1 + i];
}

// fit_a_rotate_b (bestfit.F90:214-241): the rotation that fits a onto refa
// (weights w, both COM-centred first) applied to b around a's COM, placed
// at refa's COM
void fit_a_rotate_b(int na, const double *refa, const double *a, const double *w, int nb, const double *b,
                    double *bout) {
  std::vector<double> wr(refa, refa + 3 * na), wa(a, a + 3 * na);
  double cr[3], ca[3];
  com_weighted(na, wr.data(), w, cr);
  for (int k = 0; k < na; ++k)
    for (int i = 0; i < 3; ++i) wr[3 * k + i] = wr[3 * k + i] - cr[i];
  com_weighted(na, wa.data(), w, ca);
  for (int k = 0; k < na; ++k)
    for (int i = 0; i < 3; ++i) wa[3 * k + i] = wa[3 * k + i] - ca[i];
  double q[4];
  fit_quaternion(na, wr.data(), wa.data(), w, q);
  for (int k = 0; k < nb; ++k) {
    double d[3] = {b[3 * k] - ca[0], b[3 * k + 1] - ca[1], b[3 * k + 2] - ca[2]}, r[3];
    qrot(d, q, r);
    for (int i = 0; i < 3; ++i) bout[3 * k + i] = r[i] + cr[i];
  }
}

// get_inverted_cell (engproc.F90:898-919), column-major
void inverted_cell(const double *c, double *inv, double *vol) {
#define M(i, j) c[3 * ((j)-1) + (i)-1]
#define I(i, j) inv[3 * ((j)-1) + (i)-1]
  double v = M(1, 1) * M(2, 2) * M(3, 3) + M(1, 2) * M(2, 3) * M(3, 1) + M(1, 3) * M(2, 1) * M(3, 2) -
             M(1, 3) * M(2, 2) * M(3, 1) - M(1, 2) * M(2, 1) * M(3, 3) - M(1, 1) * M(2, 3) * M(3, 2);
  I(1, 1) = M(2, 2) * M(3, 3) - M(2, 3) * M(3, 2);
  I(1, 2) = M(1, 3) * M(3, 2) - M(1, 2) * M(3, 3);
  I(1, 3) = M(1, 2) * M(2, 3) - M(1, 3) * M(2, 2);
  I(2, 1) = M(2, 3) * M(3, 1) - M(2, 1) * M(3, 3);
  I(2, 2) = M(1, 1) * M(3, 3) - M(1, 3) * M(3, 1);
  I(2, 3) = M(1, 3) * M(2, 1) - M(1, 1) * M(2, 3);
  I(3, 1) = M(2, 1) * M(3, 2) - M(2, 2) * M(3, 1);
  I(3, 2) = M(1, 2) * M(3, 1) - M(1, 1) * M(3, 2);
  I(3, 3) = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
#undef M
#undef I
  for (int i = 0; i < 9; ++i) inv[i] = inv[i] / v;
  *vol = v;
}

// calc_gmx_switching_force_params (realcal.F90:888-904)
void gmx_force(int pw, double lw, double up, double *A, double *B, double *C) {
  double d = up - lw;
  *A = -(double)pw * ((double)(pw + 4) * up - (double)(pw + 1) * lw) / (ipow_h(up, pw + 2) * ipow_h(d, 2)) / 3.0;
  *B = (double)pw * ((double)(pw + 3) * up - (double)(pw + 1) * lw) / (ipow_h(up, pw + 2) * ipow_h(d, 3)) / 4.0;
  *C = 1.0 / ipow_h(up, pw) - *A * ipow_h(d, 3) - *B * ipow_h(d, 4);
}

bool fft_axis(int n, FftAxis &ax) {
  ax.n = n;
  ax.nfac = 0;
  int m = n;
  const int rad[4] = {4, 2, 3, 5};
  for (int r : rad)
    while (m % r == 0 && ax.nfac < MAXFAC) {
      ax.fac[ax.nfac++] = r;
      m /= r;
    }
  return m == 1;
}

// columns per z-column workgroup: two LDS buffers of ncol x K3 complex plus
// the site chunk (ZSCH x ncol complex, ZSCH x (p + 1) words) within 56 KB
// (the kernel's static LDS on top stays under 64 KB)
int zcol_width(int K3, size_t vb) {
  int n = 64;
  while (n > 1 && zcol_lds(n, K3, vb) > (size_t)56 * 1024) --n;
  return n;
}

long fortran_nint(double x) { return (long)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }

}  // namespace

namespace {
size_t zcol_lds(int ncol, int K3, size_t vb) {
  return 2 * (size_t)ncol * K3 * vb + (size_t)ZSCH * ncol * vb + (size_t)ZSCH * (MAXP + 1) * 8 + 64;
}
}  // namespace


namespace {

template <typename T> struct TP {
  using V = typename C2<T>::type;
};

// ----------------------------------------------------------- setup helpers
int build_tw(erm_ctx *c, DBuf &b, int n) {
  if (c->fp32) {
    std::vector<float2> t(n);
    for (int j = 0; j < n; ++j) {
      double a = -2.0 * M_PI * (double)j / (double)n;
      t[j] = make_float2((float)cos(a), (float)sin(a));
    }
    return upload(c, b, t.data(), t.size());
  }
  std::vector<double2> t(n);
  for (int j = 0; j < n; ++j) {
    double a = -2.0 * M_PI * (double)j / (double)n;
    t[j] = make_double2(cos(a), sin(a));
  }
  return upload(c, b, t.data(), t.size());
}

int upload_real(erm_ctx *c, DBuf &b, const std::vector<double> &v) {
  if (c->fp32) {
    std::vector<float> f(v.begin(), v.end());
    return upload(c, b, f.data(), f.size());
  }
  return upload(c, b, v.data(), v.size());
}

// per-solute-type tables: charges and LJ rows against the solvent types
int solute_tables(erm_ctx *c) {
  int m = c->slt_template;
  int a0 = c->mol_begin[m];
  std::vector<double> q(c->ns), lj((size_t)2 * c->ns * c->nvt);
  for (int s = 0; s < c->ns; ++s) {
    q[s] = c->charge[a0 + s];
    int tu = c->ljtype[a0 + s];
    for (int v = 0; v < c->nvt; ++v) {
      int tv = c->vtypes[v];
      lj[((size_t)s * c->nvt + v) * 2] = c->lj_eps[(size_t)tv * c->ntype + tu];
      lj[((size_t)s * c->nvt + v) * 2 + 1] = c->lj_sig2[(size_t)tv * c->ntype + tu];
    }
  }
  TRY(upload(c, c->d_qslt, q.data(), q.size()));
  TRY(upload_real(c, c->d_qslt_t, q));
  TRY(upload_real(c, c->d_sltlj, lj));
  TRY(upload(c, c->d_sltlj64, lj.data(), lj.size()));
  return ERM_OK;
}

// Params fields that depend on the frame
void frame_params(erm_ctx *c) {
  Params &k = c->kp;
  for (int d = 0; d < 3; ++d) {
    k.L[d] = c->L[d];
    k.halfL[d] = 0.5 * c->L[d];
    int nb = (int)std::floor(c->L[d] / c->P.block_threshold);
    nb = std::max(1, std::min(nb, 1000));
    k.nb[d] = nb;
    k.hcell[d] = c->L[d] / (double)nb;
  }
  // keep the pillar count within what one workgroup lists
  while ((long)k.nb[1] * k.nb[2] * 2 > MAXSEG) {
    for (int d = 1; d < 3; ++d) {
      k.nb[d] = std::max(1, k.nb[d] / 2);
      k.hcell[d] = c->L[d] / (double)k.nb[d];
    }
  }
  for (int i = 0; i < 9; ++i) k.invcl[i] = c->invcl[i];
  k.volume = c->volume;
  k.stat_weight_system = c->stat_weight_system;
}

// grid lengths with compile-time register FFT kernels
template <typename T> bool reg_len(int K) {
  switch (K) {
    case 16: case 24: case 32: case 48: case 64: return true;
    case 96: return sizeof(T) == 4;
    default: return false;
  }
}

template <typename T> bool reg_path_ok(const erm_ctx *c) {
  const Params &k = c->kp;
  return reg_len<T>(k.K1) && reg_len<T>(k.K2) && reg_len<T>(k.K3) && (k.K1 % 2 == 0);
}

// the register path's z columns go through the z FFT for spline orders 4
// and 6 (two lanes per column), else they sum the solute's z spectra
bool zcol_fft(const erm_ctx *c, bool regp) { return regp && (c->kp.p == 4 || c->kp.p == 6); }
// z-column workgroups per configuration
int zcol_groups(const erm_ctx *c, bool regp, int vb) {
  const int ncol = c->kp.H * c->kp.K2;
  if (!regp) {
    const int w = zcol_width(c->kp.K3, vb);
    return (ncol + w - 1) / w;
  }
  const int zt = (c->zcol_T == 512) ? 512 : NTHR;
  // lanes per column: the spectral form one, the spread-and-FFT form
  // zfft_lanes (vb = 2 sizeof(T))
  const int L = zcol_fft(c, regp) ? zfft_lanes(vb / 2, c->kp.K3) : 1;
  return (L * ncol + zt - 1) / zt;
}

#define ERM_KSWITCH(K, CALL)                  \
  switch (K) {                                \
    case 16: { constexpr int KK = 16; CALL; } break; \
    case 24: { constexpr int KK = 24; CALL; } break; \
    case 32: { constexpr int KK = 32; CALL; } break; \
    case 48: { constexpr int KK = 48; CALL; } break; \
    case 64: { constexpr int KK = 64; CALL; } break; \
    case 96: if constexpr (sizeof(T) == 4) { constexpr int KK = 96; CALL; } break; \
    default: break;                           \
  }

// per-molecule reciprocal sums of n configurations (k_mol_sum: a thread per
// molecule and group of MOLSUM_CFG configurations)
int launch_mol_sum(erm_ctx *c, const Params &kg, int n, const double *esite, const int *sltj, double *uv) {
  if (n <= 0) return ERM_OK;
  const int S = c->slvmax;
  hipLaunchKernelGGL(k_mol_sum, dim3((S + NTHR - 1) / NTHR, (n + MOLSUM_CFG - 1) / MOLSUM_CFG), dim3(NTHR), 0,
                     c->st, kg, c->nsites, n, esite, dp<int>(c->d_zpos), dp<int>(c->d_mol_first),
                     dp<int>(c->d_mol_ns), dp<double>(c->d_mol_q), sltj, uv);
  return launch_check(c, "k_mol_sum");
}

// Slab shape of the fused PME kernels from the frame's z-sorted site counts:
// the deepest slab (at most tune_S planes) whose busiest slab fits the
// per-thread site registers (R = 16 or 32), with at most 3x halo planes;
// for the register path, tune_Q planes per group within the LDS budget and
// one row pair per thread.
template <typename T> int choose_slab(erm_ctx *c) {
  const Params &k = c->kp;
  const int K3 = k.K3;
  const std::vector<int> &zst = c->zst;
  c->slab_S = 0;
  c->slab_R = 0;
  // R = 8 only for the register path with shallow slabs (fewer live registers)
  const bool reg8 = c->tune_S <= 4 && (c->P.pme_path == 0 || c->P.pme_path == 3);
  for (int R : {8, 16, 32}) {
    if (c->P.pme_path == 2) break;
    if (R == 8 && !reg8) continue;
    for (int Sl = std::min(c->tune_S, K3); Sl >= 1 && !c->slab_S; --Sl) {
      int mx = 0;
      for (int zb = 0; zb < K3; zb += Sl) mx = std::max(mx, zst[std::min(zb + Sl, K3)] - zst[zb]);
      if (mx <= R * NTHR && (Sl + k.p - 1) <= 3 * Sl) {
        c->slab_S = Sl;
        c->slab_R = R;
      }
    }
    if (c->slab_S) break;
  }
  if ((c->P.pme_path == 1 || c->P.pme_path == 3) && !c->slab_S)
    return fail(c, ERM_E_UNSUPPORTED, "fused PME slab does not fit this grid / system");
  c->xslab_Q = 0;
  if (c->slab_S && reg_path_ok<T>(c)) {
    // one complex plane per group slot (the real rows are written in place)
    size_t plane_b = 2 * sizeof(T) * (size_t)k.H * k.K2;
    int Q = (int)std::min<size_t>((size_t)c->tune_Q, c->lds_cap / plane_b);
    if (Q >= 1) c->xslab_Q = Q;
  }
  if (c->P.pme_path == 3 && !c->xslab_Q)
    return fail(c, ERM_E_UNSUPPORTED, "register-column PME path does not support this grid");
  return ERM_OK;
}

// sorted index of every tagpt molecule (for soln-mode solutes); waits for
// the frame's preparation
int mol_index(erm_ctx *c) {
  const int S = c->slvmax;
  std::vector<int> molk(S);
  HIPCHK(c, hipMemcpyAsync(molk.data(), c->d_mol_k.p, sizeof(int) * S, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  c->jof_k.assign(S, -1);
  for (int j = 0; j < S; ++j) c->jof_k[molk[j]] = j;
  c->jof_stale = false;
  return ERM_OK;
}
// jof_k for the current frame, read back on first use (soln-mode solutes,
// host energies, the adjoint), so that frame preparation needs no readback
// for inserted test particles
int ensure_jof(erm_ctx *c) { return c->jof_stale ? mol_index(c) : ERM_OK; }

template <typename T> int frame_prep(erm_ctx *c) {
  Params &k = c->kp;
  const int ncell = k.nb[0] * k.nb[1] * k.nb[2];
  const int S = c->slvmax;
  int nsites = 0;
  for (int kk = 0; kk < S; ++kk) {
    int m = c->tagpt[kk];
    nsites += c->mol_begin[m + 1] - c->mol_begin[m];
  }
  TRY(ensure(c, c->d_cellid, sizeof(int) * S));
  TRY(ensure(c, c->d_count, sizeof(int) * (ncell + 1)));
  TRY(ensure(c, c->d_cell_start, sizeof(int) * (ncell + 1)));
  TRY(ensure(c, c->d_fill, sizeof(int) * (ncell + 1)));
  TRY(ensure(c, c->d_rm, sizeof(unsigned long long)));
  TRY(ensure(c, c->d_mol_k, sizeof(int) * S));
  TRY(ensure(c, c->d_mol_ns, sizeof(int) * S));
  TRY(ensure(c, c->d_mol_first, sizeof(int) * (S + 1)));
  TRY(ensure(c, c->d_mol_q, sizeof(double) * S));
  TRY(ensure(c, c->d_mol_spec, sizeof(int) * S));
  TRY(ensure(c, c->d_mol_m, sizeof(int) * S));
  TRY(ensure(c, c->d_mol_d1, sizeof(double) * S));
  TRY(ensure(c, c->d_jof, sizeof(int) * S));
  TRY(ensure(c, c->d_sx, sizeof(double) * nsites));
  TRY(ensure(c, c->d_sy, sizeof(double) * nsites));
  TRY(ensure(c, c->d_sz, sizeof(double) * nsites));
  TRY(ensure(c, c->d_sq, sizeof(T) * nsites));
  TRY(ensure(c, c->d_sct, sizeof(int) * nsites));
  TRY(ensure(c, c->d_sg, sizeof(int) * nsites));
  TRY(ensure(c, c->d_sw, sizeof(T) * (size_t)nsites * 3 * k.p));
  // z buckets of the site sort below: (g3, LDS bank class of the gather)
  // when they fit the workgroup-local counts, else g3
  const int K3 = k.K3;
  const bool zsort = (c->P.cltype != ERM_EL_COULOMB);
  const int KB = (K3 * ZCLASS <= ZBUCKET_MAX) ? ZCLASS : 1, nbk = K3 * KB;
  const int rowp = k.K1 + 2;
  if (zsort) TRY(ensure(c, c->d_zcount, sizeof(int) * (nbk + 1)));
  const int nclr = std::max(ncell + 1, zsort ? nbk + 1 : 0);
  hipLaunchKernelGGL(k_frame_clear, dim3((nclr + 255) / 256), dim3(256), 0, c->st, dp<int>(c->d_count), ncell + 1,
                     zsort ? dp<int>(c->d_zcount) : (int *)nullptr, zsort ? nbk + 1 : 0,
                     dp<unsigned long long>(c->d_rm));
  TRY(launch_check(c, "k_frame_clear"));
  const int tb = 256, gb = (S + tb - 1) / tb;
  hipLaunchKernelGGL(k_mol_cell, dim3(gb), dim3(tb), 0, c->st, k, dp<double>(c->d_pos), dp<int>(c->d_tagpt),
                     dp<int>(c->d_mol_begin), dp<int>(c->d_cellid), dp<int>(c->d_count),
                     dp<unsigned long long>(c->d_rm));
  TRY(launch_check(c, "k_mol_cell"));
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->st, dp<int>(c->d_count), dp<int>(c->d_cell_start), ncell,
                     dp<int>(c->d_fill), (int *)nullptr, 0, 0);
  TRY(launch_check(c, "k_scan"));
  hipLaunchKernelGGL(k_mol_scatter, dim3(gb), dim3(tb), 0, c->st, k, dp<int>(c->d_cellid), dp<int>(c->d_fill),
                     dp<int>(c->d_tagpt), dp<int>(c->d_mol_begin), dp<int>(c->d_mol_k), dp<int>(c->d_mol_ns),
                     dp<int>(c->d_jof));
  TRY(launch_check(c, "k_mol_scatter"));
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->st, dp<int>(c->d_mol_ns), dp<int>(c->d_mol_first), S,
                     (int *)nullptr, (int *)nullptr, 0, 0);
  TRY(launch_check(c, "k_scan"));
  hipLaunchKernelGGL(k_site_fill<T>, dim3(gb), dim3(tb), 0, c->st, k, dp<double>(c->d_pos), dp<int>(c->d_tagpt),
                     dp<int>(c->d_mol_begin), dp<double>(c->d_charge), dp<int>(c->d_cvt),
                     dp<double>(c->d_mol_charge), dp<int>(c->d_uvspec), dp<int>(c->d_mol_k),
                     dp<int>(c->d_mol_first), dp<double>(c->d_sx), dp<double>(c->d_sy), dp<double>(c->d_sz),
                     dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sg), dp<T>(c->d_sw), dp<double>(c->d_mol_q),
                     dp<int>(c->d_mol_spec), dp<int>(c->d_mol_m), dp<double>(c->d_mol_d1));
  TRY(launch_check(c, "k_site_fill"));
  c->nsites = nsites;
  c->jof_stale = true;
  if (!zsort) return ERM_OK;
  // sites sorted by PME plane g3, dealt over the gather's LDS bank classes
  // within a plane, for the slab kernels
  TRY(ensure(c, c->d_zfine, sizeof(int) * (nbk + 1)));
  TRY(ensure(c, c->d_zfill, sizeof(int) * (nbk + 1)));
  TRY(ensure(c, c->d_zstart, sizeof(int) * (K3 + 1)));
  TRY(ensure(c, c->d_zsite, sizeof(int) * nsites));
  TRY(ensure(c, c->d_zpos, sizeof(int) * nsites));
  TRY(ensure(c, c->d_zg, sizeof(int) * nsites));
  TRY(ensure(c, c->d_zq, sizeof(T) * nsites));
  TRY(ensure(c, c->d_zw, sizeof(T) * (size_t)nsites * 3 * k.p));
  const int gs = (nsites + tb - 1) / tb;
  hipLaunchKernelGGL(k_zcount, dim3(gs), dim3(tb), 0, c->st, nsites, nbk, KB, rowp, dp<int>(c->d_sg),
                     dp<int>(c->d_zcount));
  TRY(launch_check(c, "k_zcount"));
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->st, dp<int>(c->d_zcount), dp<int>(c->d_zfine), nbk,
                     dp<int>(c->d_zfill), dp<int>(c->d_zstart), K3, KB);
  TRY(launch_check(c, "k_scan"));
  hipLaunchKernelGGL(k_zscatter<T>, dim3(gs), dim3(tb), 0, c->st, nsites, nbk, KB, rowp, k.p, dp<int>(c->d_sg),
                     dp<T>(c->d_sq), dp<T>(c->d_sw), dp<int>(c->d_zcount), dp<int>(c->d_zfine),
                     dp<int>(c->d_zfill), dp<int>(c->d_zsite), dp<int>(c->d_zpos), dp<int>(c->d_zg),
                     dp<T>(c->d_zq), dp<T>(c->d_zw));
  TRY(launch_check(c, "k_zscatter"));
  c->zst.assign(K3 + 1, 0);
  HIPCHK(c, hipMemcpyAsync(c->zst.data(), c->d_zstart.p, sizeof(int) * (K3 + 1), hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  return choose_slab<T>(c);
}

// --------------------------------------------------------- batch pipeline
enum { DO_RECIP = 1, DO_REAL = 2, DO_RESIDUAL = 4, DO_SELF = 8, DO_HIST = 16 };

struct Batch {
  int n;
  int first;                       // global index of the first configuration
  const double *host_pos;          // n * ns * 3 (refs) or NULL
  const int *solutes;              // soln: n molecule indices, else NULL
  const double *weights;           // n or NULL
  bool random;                     // device poses
  uint64_t seed, first_index;
  const double *flex = nullptr;    // random placement of n given body configurations (n * ns * 3)
};


// solute configurations of a batch into d_spos (device poses, host
// coordinates or the frame's solute molecules) and the sorted index of each
// solute (d_sltj, -1 for a test particle)
int stage_configs(erm_ctx *c, const Batch &bt, const Params &k) {
  const int n = bt.n, ns = c->ns;
  if (bt.random) {
    if (!c->args_ready || !c->d_args.p) return fail(c, ERM_E_ENG_BUG, "random poses without generator arguments");
    if (k.ins_org == ERM_INSORG_REFSTR) {
      if (!c->d_refsite.p) return fail(c, ERM_E_ENG_BUG, "reference-structure fit missing");
      if (k.ins_pos == ERM_INSPOS_GAUSS) TRY(ensure(c, c->d_wsol, sizeof(double) * std::max(n, c->Bmax)));
      hipLaunchKernelGGL(k_pose_refstr, dim3((n + 63) / 64), dim3(64), 0, c->st, k, n, dp<uint64_t>(c->d_args),
                         (uint64_t)bt.first, dp<double>(c->d_sitemass) + c->mol_begin[c->slt_template],
                         dp<double>(c->d_cellm), dp<double>(c->d_refsite), dp<double>(c->d_refw),
                         dp<double>(c->d_spos), dp<double>(c->d_wsol), dp<int>(c->d_status));
      TRY(launch_check(c, "k_pose_refstr"));
      if (k.ins_pos == ERM_INSPOS_GAUSS) c->wsol_ones = 0;
    } else {
      if (k.ins_org == ERM_INSORG_AGGCEN && !c->d_aggcen.p)
        return fail(c, ERM_E_ENG_BUG, "aggregate centre missing");
      if (bt.flex) TRY(upload(c, c->d_flexbf, bt.flex, (size_t)n * ns * 3));
      const size_t lpose = (ns <= POSE_LDS_SITES) ? sizeof(double) * 4 * ns : 0;
      hipLaunchKernelGGL(lpose ? k_pose<true> : k_pose<false>, dim3(n), dim3(64), lpose, c->st, k, n, dp<uint64_t>(c->d_args),
                         (uint64_t)bt.first, bt.flex ? dp<double>(c->d_flexbf) : dp<double>(c->d_bf),
                         bt.flex ? 3 * ns : 0, dp<double>(c->d_sitemass) + c->mol_begin[c->slt_template],
                         dp<double>(c->d_cellm), dp<double>(c->d_aggcen), dp<double>(c->d_spos));
      TRY(launch_check(c, "k_pose"));
    }
  } else if (bt.host_pos) {
    HIPCHK(c, hipMemcpyAsync(c->d_spos.p, bt.host_pos, sizeof(double) * (size_t)n * ns * 3, hipMemcpyHostToDevice,
                             c->st));
  } else {
    for (int b = 0; b < n; ++b)
      HIPCHK(c, hipMemcpyAsync(dp<double>(c->d_spos) + (size_t)b * ns * 3,
                               dp<double>(c->d_pos) + (size_t)3 * c->mol_begin[bt.solutes[b]],
                               sizeof(double) * ns * 3, hipMemcpyDeviceToDevice, c->st));
  }
  if (bt.solutes) {
    TRY(ensure_jof(c));
    std::vector<int> sltj(n, -1);
    for (int b = 0; b < n; ++b) {
      int kk = c->kof[bt.solutes[b]];
      sltj[b] = (kk >= 0) ? c->jof_k[kk] : -1;
    }
    TRY(upload(c, c->d_sltj, sltj.data(), n));
  } else {
    // no solute among the physical molecules (test particle)
    TRY(ensure(c, c->d_sltj, sizeof(int) * n));
    hipLaunchKernelGGL(k_fill<int>, dim3((n + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, dp<int>(c->d_sltj),
                       (size_t)n, -1);
    TRY(launch_check(c, "k_fill"));
  }
  return ERM_OK;
}

// workgroups per configuration of the tiled pair kernel: nchunk chunks of 64
// candidate molecules over nwave waves, PAIR_CHUNKS_PER_WAVE chunks each
constexpr int PAIR_CHUNKS_PER_WAVE = 4;
inline int pair_splits(int nchunk, int nwave) {
  const int per = nwave * PAIR_CHUNKS_PER_WAVE;
  return std::max(1, std::min(64, (nchunk + per - 1) / per));
}

// real-space pairs of a batch: the tiled kernel with compacted pair lists,
// or the kernel with one thread per candidate molecule (pair_path 1, or
// when the tiles' LDS does not fit)
template <typename T> int launch_pair(erm_ctx *c, const Params &k, int n, bool bare) {
  const int S = c->slvmax, ns = c->ns;
  const double Rg = k.rcmax + c->slt_radius + c->mol_radius + 0.5 * (k.hcell[0] + k.hcell[1] + k.hcell[2]) / 3.0;
  const double frac = std::min(1.0, 4.18879 * Rg * Rg * Rg / c->volume);
  const int segcap = 2 * k.nb[1] * k.nb[2];
  const int nwave = NTHR / 64;
  const size_t lt = pair_tile_lds<T>(ns, c->nvt, segcap, nwave);
  if (c->pair_path == 0 && lt <= c->lds_cap) {
    // chunks of 64 candidate molecules, PAIR_CHUNKS_PER_WAVE per wave (the
    // per-configuration set-up of a workgroup is shared by its chunks)
    const int nchunk = (int)std::ceil(frac * S / 64.0);
    const int nsplit = pair_splits(nchunk, nwave);
    if (bare) {
      TRY(set_lds(c, (const void *)k_pair_tile<T, true>, lt, "k_pair_tile"));
      hipLaunchKernelGGL((k_pair_tile<T, true>), dim3(n, nsplit), dim3(NTHR), lt, c->st, k, dp<double>(c->d_spos),
                         dp<double>(c->d_sref), dp<T>(c->d_qslt_t), dp<T>(c->d_sltlj), dp<int>(c->d_cell_start),
                         dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy),
                         dp<double>(c->d_sz), dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj),
                         dp<double>(c->d_rm), dp<double>(c->d_sltlj64), dp<double>(c->d_qslt),
                         dp<double>(c->d_mol_d1), dp<double>(c->d_uv));
    } else {
      TRY(set_lds(c, (const void *)k_pair_tile<T, false>, lt, "k_pair_tile"));
      hipLaunchKernelGGL((k_pair_tile<T, false>), dim3(n, nsplit), dim3(NTHR), lt, c->st, k, dp<double>(c->d_spos),
                         dp<double>(c->d_sref), dp<T>(c->d_qslt_t), dp<T>(c->d_sltlj), dp<int>(c->d_cell_start),
                         dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy),
                         dp<double>(c->d_sz), dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj),
                         dp<double>(c->d_rm), dp<double>(c->d_sltlj64), dp<double>(c->d_qslt),
                         dp<double>(c->d_mol_d1), dp<double>(c->d_uv));
    }
    return launch_check(c, "k_pair_tile");
  }
  const int nsplit = std::max(1, std::min(16, (int)std::ceil(frac * S / (2.0 * NTHR))));
  const size_t lpair = sizeof(T) * ((size_t)7 * ns + (size_t)2 * ns * c->nvt) + sizeof(int) * (2 * MAXSEG + 2) + 64;
  TRY(set_lds(c, bare ? (const void *)k_pair<T, true> : (const void *)k_pair<T, false>, lpair, "k_pair"));
  if (bare)
    hipLaunchKernelGGL((k_pair<T, true>), dim3(n, nsplit), dim3(NTHR), lpair, c->st, k, dp<double>(c->d_spos),
                       dp<double>(c->d_sref), dp<T>(c->d_qslt_t), dp<T>(c->d_sltlj), dp<int>(c->d_cell_start),
                       dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy),
                       dp<double>(c->d_sz), dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj),
                       dp<double>(c->d_rm), dp<double>(c->d_sltlj64), dp<double>(c->d_qslt), dp<double>(c->d_uv));
  else
    hipLaunchKernelGGL((k_pair<T, false>), dim3(n, nsplit), dim3(NTHR), lpair, c->st, k, dp<double>(c->d_spos),
                       dp<double>(c->d_sref), dp<T>(c->d_qslt_t), dp<T>(c->d_sltlj), dp<int>(c->d_cell_start),
                       dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy),
                       dp<double>(c->d_sz), dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj),
                       dp<double>(c->d_rm), dp<double>(c->d_sltlj64), dp<double>(c->d_qslt), dp<double>(c->d_uv));
  return launch_check(c, "k_pair");
}

// Energies of a batch of configurations into d_uv (sorted molecule order);
// flags select the terms.  Stage events ev[0..6].
template <typename T> int run_energies(erm_ctx *c, const Batch &bt, int flags) {
  using V = typename C2<T>::type;
  // bare Coulomb: no reciprocal space, and realcal_self and residual_ene
  // return 0 (realcal.F90:294, engproc.F90:865), so uvengy(0) = 0
  const bool bare = (c->P.cltype == ERM_EL_COULOMB);
  if (bare) flags &= ~(DO_RECIP | DO_SELF | DO_RESIDUAL);
  Params k = c->kp;
  const int n = bt.n, ns = c->ns, S = c->slvmax;
  const int K1 = k.K1, K2 = k.K2, K3 = k.K3, H = k.H, p = k.p;
  const int smol = bt.solutes ? bt.solutes[0] : c->slt_template;
  k.solute_charge = c->mol_charge[smol];
  k.ns = ns;
  TRY(ensure(c, c->d_spos, sizeof(double) * (size_t)n * ns * 3));
  TRY(ensure(c, c->d_sref, sizeof(double) * 4 * n));
  TRY(ensure(c, c->d_sg3, sizeof(int) * (size_t)n * ns));
  TRY(ensure(c, c->d_sw3, sizeof(T) * (size_t)n * ns * p));
  TRY(ensure(c, c->d_sqv, sizeof(T) * (size_t)n * ns));
  TRY(ensure(c, c->d_w1h, sizeof(V) * (size_t)n * ns * H));
  TRY(ensure(c, c->d_w2h, sizeof(V) * (size_t)n * ns * K2));
  const bool regp = (flags & DO_RECIP) && c->xslab_Q > 0 && (c->P.pme_path == 0 || c->P.pme_path == 3);
  const bool zfft = zcol_fft(c, regp);  // else the preparation writes the z spectra
  if (regp && !zfft) TRY(ensure(c, c->d_w3h, sizeof(V) * (size_t)n * ns * K3));
  TRY(ensure(c, c->d_uv, sizeof(double) * (size_t)n * (S + 1)));
  TRY(ensure(c, c->d_sltj, sizeof(int) * n));
  TRY(ensure(c, c->d_selfreal, sizeof(double) * n));
  TRY(ensure(c, c->d_usreal, sizeof(double)));
  const size_t vb = sizeof(V);
  const int ncolwg = zcol_width(K3, vb);
  const int nzg = zcol_groups(c, regp, (int)vb);
  TRY(ensure(c, c->d_selfp, sizeof(double) * (size_t)n * nzg));
  if (flags & DO_RECIP) TRY(ensure(c, c->d_Hk, vb * (size_t)n * K3 * K2 * H));
  TRY(mark(c, 0));
  TRY(stage_configs(c, bt, k));
  hipLaunchKernelGGL(k_solute_prep<T>, dim3(n), dim3(128), 0, c->st, k, dp<double>(c->d_spos),
                     dp<double>(c->d_qslt), dp<V>(c->d_tw1), dp<V>(c->d_tw2), dp<V>(c->d_tw3),
                     dp<double>(c->d_sref), dp<int>(c->d_sg3), dp<T>(c->d_sw3), dp<T>(c->d_sqv),
                     dp<V>(c->d_w1h), dp<V>(c->d_w2h), (regp && !zfft) ? dp<V>(c->d_w3h) : (V *)nullptr);
  TRY(launch_check(c, "k_solute_prep"));
  TRY(mark(c, 1));
  if (regp) {
    int rz = ERM_E_UNSUPPORTED;
    if (zfft) {
      ERM_KSWITCH(K3, (rz = launch_zcol_fft<T, KK>(c, k, n, nzg)));
    } else {
      ERM_KSWITCH(K3, (rz = launch_zcol_reg<T, KK>(c, k, n, nzg)));
    }
    TRY(rz);
    TRY(launch_check(c, zfft ? "k_pme_zcol_fft" : "k_pme_zcol_reg"));
    TRY(mark(c, 2));
    if (!xslab_folds_y(c)) {
      ERM_KSWITCH(K2, (launch_ycol_reg<T, KK>(c, k, n)));
      TRY(launch_check(c, "y columns"));
    }
    TRY(mark(c, 3));
    TRY(ensure(c, c->d_esite, sizeof(double) * (size_t)n * c->nsites));
    const size_t lx = 2 * sizeof(T) * (size_t)c->xslab_Q * H * K2;
    int rc = ERM_E_UNSUPPORTED;
    ERM_KSWITCH(K1, (rc = launch_xslab<T, KK>(c, k, n, lx)));
    TRY(rc);
    TRY(launch_check(c, "k_pme_xslab"));
    TRY(mark(c, 4));
    Params kg = k;
    if (!(flags & DO_RESIDUAL)) kg.solute_charge = 0.0;
    TRY(launch_mol_sum(c, kg, n, dp<double>(c->d_esite), dp<int>(c->d_sltj), dp<double>(c->d_uv)));
  } else if (flags & DO_RECIP) {
    size_t lz = zcol_lds(ncolwg, K3, vb);
    hipLaunchKernelGGL(k_pme_zcol<T>, dim3(nzg, n), dim3(NTHR), lz, c->st, k, ncolwg, dp<T>(c->d_engfac),
                       dp<int>(c->d_sg3), dp<T>(c->d_sw3), dp<T>(c->d_sqv), dp<V>(c->d_w1h), dp<V>(c->d_w2h),
                       dp<V>(c->d_tw3), dp<V>(c->d_Hk), dp<double>(c->d_selfp));
    TRY(launch_check(c, "k_pme_zcol"));
    TRY(mark(c, 2));
    size_t lp = 2 * (size_t)K2 * H * vb;
    Params kg = k;
    if (!(flags & DO_RESIDUAL)) kg.solute_charge = 0.0;
    if (c->slab_S > 0 && c->P.pme_path != 2) {
      TRY(mark(c, 3));
      const int nslab = (K3 + c->slab_S - 1) / c->slab_S;
      TRY(ensure(c, c->d_esite, sizeof(double) * (size_t)n * c->nsites));
      if (c->slab_R == 16)
        hipLaunchKernelGGL((k_pme_slab<T, 16>), dim3(nslab, n), dim3(NTHR), lp, c->st, k, c->slab_S, c->nsites,
                           dp<V>(c->d_Hk), dp<V>(c->d_tw1), dp<V>(c->d_tw2), dp<int>(c->d_zstart),
                           dp<int>(c->d_zg), dp<T>(c->d_zq), dp<T>(c->d_zw), dp<int>(c->d_zsite),
                           dp<double>(c->d_esite));
      else
        hipLaunchKernelGGL((k_pme_slab<T, 32>), dim3(nslab, n), dim3(NTHR), lp, c->st, k, c->slab_S, c->nsites,
                           dp<V>(c->d_Hk), dp<V>(c->d_tw1), dp<V>(c->d_tw2), dp<int>(c->d_zstart),
                           dp<int>(c->d_zg), dp<T>(c->d_zq), dp<T>(c->d_zw), dp<int>(c->d_zsite),
                           dp<double>(c->d_esite));
      TRY(launch_check(c, "k_pme_slab"));
      TRY(mark(c, 4));
      TRY(launch_mol_sum(c, kg, n, dp<double>(c->d_esite), dp<int>(c->d_sltj), dp<double>(c->d_uv)));
    } else {
      TRY(ensure(c, c->d_phi, sizeof(T) * (size_t)n * K3 * K2 * K1));
      hipLaunchKernelGGL(k_pme_plane<T>, dim3(K3, n), dim3(NTHR), lp, c->st, k, dp<V>(c->d_Hk), dp<V>(c->d_tw1),
                         dp<V>(c->d_tw2), dp<T>(c->d_phi));
      TRY(launch_check(c, "k_pme_plane"));
      TRY(mark(c, 3));
      hipLaunchKernelGGL(k_gather<T>, dim3((S + NTHR - 1) / NTHR, n), dim3(NTHR), 0, c->st, kg, dp<T>(c->d_phi),
                         dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_mol_q), dp<int>(c->d_sg),
                         dp<T>(c->d_sw), dp<T>(c->d_sq), dp<int>(c->d_sltj), dp<double>(c->d_uv));
      TRY(launch_check(c, "k_gather"));
      TRY(mark(c, 4));
    }
  } else {
    const size_t nuv = (size_t)n * (S + 1);
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((nuv + NTHR - 1) / NTHR)), dim3(NTHR), 0, c->st,
                       dp<double>(c->d_uv), nuv, 0.0);
    TRY(launch_check(c, "k_fill"));
    TRY(mark(c, 2));
    TRY(mark(c, 3));
    TRY(mark(c, 4));
  }
  TRY(mark(c, 5));
  if ((flags & DO_REAL) && bare && c->P.boxshp == ERM_SYS_NONPERIODIC) {
    size_t lo = sizeof(T) * ((size_t)7 * ns + (size_t)2 * ns * c->nvt) + 64;
    TRY(set_lds(c, (const void *)k_pair_open<T>, lo, "k_pair_open"));
    hipLaunchKernelGGL(k_pair_open<T>, dim3(n, (S + NTHR - 1) / NTHR), dim3(NTHR), lo, c->st, k,
                       dp<double>(c->d_spos), dp<double>(c->d_sref), dp<T>(c->d_qslt_t), dp<T>(c->d_sltlj),
                       dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy),
                       dp<double>(c->d_sz), dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj),
                       dp<double>(c->d_sltlj64), dp<double>(c->d_qslt), dp<double>(c->d_uv));
    TRY(launch_check(c, "k_pair_open"));
  } else if (flags & DO_REAL) {
    TRY(launch_pair<T>(c, k, n, bare));
  }
  TRY(mark(c, 6));
  if (flags & DO_SELF) {
    // realcal_self: reused after the first configuration for a rigid test
    // particle (engproc.F90:712-730), recomputed per configuration otherwise
    int per_item = (c->P.slttype != 2);
    if (per_item) {
      for (int b = 0; b < n; ++b)
        hipLaunchKernelGGL(k_selfreal, dim3(1), dim3(NTHR), 0, c->st, k, dp<double>(c->d_spos), b,
                           dp<double>(c->d_qslt), dp<double>(c->d_selfreal) + b);
    } else if (!c->have_usreal) {
      hipLaunchKernelGGL(k_selfreal, dim3(1), dim3(NTHR), 0, c->st, k, dp<double>(c->d_spos), 0,
                         dp<double>(c->d_qslt), dp<double>(c->d_usreal));
      c->have_usreal = true;
    }
    TRY(launch_check(c, "k_selfreal"));
    Params ku = k;
    if (!(flags & DO_RECIP)) {
      HIPCHK(c, hipMemsetAsync(c->d_selfp.p, 0, sizeof(double) * (size_t)n * nzg, c->st));
    }
    hipLaunchKernelGGL(k_u0, dim3((n + 63) / 64), dim3(64), 0, c->st, ku, n, nzg, dp<double>(c->d_selfp),
                       per_item ? dp<double>(c->d_selfreal) : dp<double>(c->d_usreal), per_item,
                       dp<double>(c->d_uv));
    TRY(launch_check(c, "k_u0"));
  }
  TRY(mark(c, 7));
  return ERM_OK;
}

// device buffers of update_histogram for a batch of n configurations: the
// energies (sorted molecule order), the solutes' sorted indices, the
// configurations' weights in and out, their histograms, per-species sums,
// intermediate-state weights and the error status
struct HistBufs {
  double *uv, *wsol, *hist, *w, *flceng, *wi;
  int *sltj, *status;
  const int *mol_spec;  // species of the sorted molecules (the frame's owner)
};

// the context's own buffers (grown as needed)
int ctx_hist_bufs(erm_ctx *c, int n, HistBufs &hb) {
  TRY(ensure(c, c->d_hist, sizeof(double) * (size_t)n * c->ermax));
  TRY(ensure(c, c->d_w, sizeof(double) * n));
  TRY(ensure(c, c->d_flceng, sizeof(double) * (size_t)n * MAXSPEC));
  if (c->kp.intm) TRY(ensure(c, c->d_wi, sizeof(double) * n));
  hb.uv = dp<double>(c->d_uv);
  hb.wsol = dp<double>(c->d_wsol);
  hb.hist = dp<double>(c->d_hist);
  hb.w = dp<double>(c->d_w);
  hb.flceng = dp<double>(c->d_flceng);
  hb.wi = dp<double>(c->d_wi);
  hb.sltj = dp<int>(c->d_sltj);
  hb.status = dp<int>(c->d_status);
  hb.mol_spec = dp<int>(c->d_mol_spec);
  return ERM_OK;
}

// the correlation matrix part of update_histogram (engproc.F90:835-853)
int run_corr_at(erm_ctx *c, int n, const HistBufs &hb) {
  if (!c->P.corrcal) return ERM_OK;
  const int E = c->ermax;
  const int nt = (E + 15) / 16;
  const int tiles = nt * (nt + 1) / 2;  // tiles on and above the diagonal, one workgroup each
  hipLaunchKernelGGL(k_corr, dim3(tiles), dim3(256), 0, c->st, n, E, hb.hist, hb.w, dp<double>(c->d_ecorr));
  return launch_check(c, "k_corr");
}

// update_histogram for the n configurations of hb (engproc.F90:758-856)
// (without_corr: the correlation matrix is left to run_corr_at)
int run_histogram_at(erm_ctx *c, int n, int first, const HistBufs &hb, bool without_corr = false) {
  Params k = c->kp;
  const int E = c->ermax;
  const size_t lb = (size_t)E * (sizeof(double) + sizeof(int)) + sizeof(double) * (size_t)c->numslv * NTHR;
  TRY(set_lds(c, (const void *)k_bin, lb, "k_bin"));
  if (!hb.mol_spec) return fail(c, ERM_E_ENG_BUG, "histogram without the frame's molecule species");
  hipLaunchKernelGGL(k_bin, dim3(n), dim3(NTHR), lb, c->st, k, hb.uv, dp<double>(c->d_edges),
                     dp<int>(c->d_specoff), dp<int>(c->d_uvmax), dp<int>(c->d_uvsoft), hb.mol_spec, hb.sltj, hb.hist,
                     hb.flceng, dp<unsigned long long>(c->d_mnmx), hb.status, first);
  TRY(launch_check(c, "k_bin"));
  hipLaunchKernelGGL(k_scalars, dim3(1), dim3(NTHR), 0, c->st, k, n, hb.uv, hb.wsol, dp<double>(c->d_escrd),
                     dp<double>(c->d_scal), dp<double>(c->d_eself), dp<unsigned long long>(c->d_mnmx), hb.w,
                     hb.status, first);
  TRY(launch_check(c, "k_scalars"));
  hipLaunchKernelGGL(k_edens, dim3((E + 31) / 32), dim3(256), 0, c->st, n, E, hb.hist, hb.w, dp<double>(c->d_edens));
  TRY(launch_check(c, "k_edens"));
  if (c->P.slttype == 1) {
    hipLaunchKernelGGL(k_slnuv, dim3(1), dim3(MAXSPEC), 0, c->st, n, c->numslv, hb.flceng, hb.w,
                       dp<double>(c->d_slnuv));
    TRY(launch_check(c, "k_slnuv"));
  }
  if (k.intm) {
    hipLaunchKernelGGL(k_intm_weight, dim3(n), dim3(256), 0, c->st, E, hb.hist, dp<int>(c->d_zerodns), hb.w, hb.wi);
    TRY(launch_check(c, "k_intm_weight"));
    hipLaunchKernelGGL(k_edens, dim3((E + 31) / 32), dim3(256), 0, c->st, n, E, hb.hist, hb.wi,
                       dp<double>(c->d_edintm));
    TRY(launch_check(c, "k_edens"));
    hipLaunchKernelGGL(k_slnuv, dim3(1), dim3(MAXSPEC), 0, c->st, n, c->numslv, hb.flceng, hb.wi,
                       dp<double>(c->d_uvintm));
    TRY(launch_check(c, "k_slnuv"));
    hipLaunchKernelGGL(k_intm_scal, dim3(1), dim3(1), 0, c->st, n, hb.w, hb.wi, dp<double>(c->d_intm_sc));
    TRY(launch_check(c, "k_intm_scal"));
  }
  TRY(mark(c, 8));
  if (!without_corr) TRY(run_corr_at(c, n, hb));
  TRY(mark(c, 9));
  return ERM_OK;
}

// update_histogram for the n configurations in d_uv (weights in d_wsol)
int run_histogram(erm_ctx *c, int n, int first) {
  HistBufs hb;
  TRY(ctx_hist_bufs(c, n, hb));
  return run_histogram_at(c, n, first, hb);
}

int copy_out_uv(erm_ctx *c, int n, double *out) {
  const int S = c->slvmax;
  TRY(ensure(c, c->d_uvout, sizeof(double) * (size_t)n * (S + 1)));
  hipLaunchKernelGGL(k_unsort, dim3((S + 255) / 256, n), dim3(256), 0, c->st, n, S, dp<int>(c->d_mol_k),
                     dp<double>(c->d_uv), dp<double>(c->d_uvout));
  TRY(launch_check(c, "k_unsort"));
  HIPCHK(c, hipMemcpyAsync(out, c->d_uvout.p, sizeof(double) * (size_t)n * (S + 1), hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

// stage times of the last call from the per-batch events (after a sync):
// NSTAGE_EV events per batch; stage i runs from event i to event i + 1, the
// histogram stages (bin, correlation) only when the call accumulates
int stage_times(erm_ctx *c, int nb, bool hist) {
  for (int i = 0; i < NSTAGE_OUT; ++i) c->stage_ms[i] = 0;
  float ms;
  const int last = hist ? NSTAGE_EV - 1 : 7;
  for (int bi = 0; bi < nb; ++bi) {
    hipEvent_t *ev = &c->bev[NSTAGE_EV * bi];
    for (int i = 0; i < last; ++i)
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) c->stage_ms[i] += ms;
    if (hipEventElapsedTime(&ms, ev[0], ev[last]) == hipSuccess) c->stage_ms[NSTAGE_OUT - 1] += ms;
  }
  (void)hipGetLastError();
  return ERM_OK;
}

// NSTAGE_EV stage events for each of nb batches
int batch_events(erm_ctx *c, int nb) {
  while ((int)c->bev.size() < NSTAGE_EV * nb) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->bev.push_back(e);
  }
  return ERM_OK;
}

int need(erm_ctx *c, bool frame, bool bins) {
  if (!c) return ERM_E_ARG;
  if (!c->have_topo) return fail(c, ERM_E_ARG, "erm_set_topology has not been called");
  if (bins && !c->have_bins) return fail(c, ERM_E_ARG, "erm_set_bins has not been called");
  if (frame && !c->have_frame) return fail(c, ERM_E_ARG, "erm_set_frame has not been called");
  return ERM_OK;
}

// generator arguments {seed, first index of the call} for k_pose, through a
// pinned staging buffer (the previous call has synchronised before it is
// rewritten)
int upload_args(erm_ctx *c, uint64_t seed, uint64_t first_index) {
  TRY(ensure(c, c->d_args, 2 * sizeof(uint64_t)));
  if (!c->h_args) HIPCHK(c, hipHostMalloc((void **)&c->h_args, 2 * sizeof(uint64_t), hipHostMallocDefault));
  c->h_args[0] = seed;
  c->h_args[1] = first_index;
  HIPCHK(c, hipMemcpyAsync(c->d_args.p, c->h_args, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, c->st));
  c->args_ready = true;
  return ERM_OK;
}

// work of one call that precedes its batches: clear the error status, copy
// the generator arguments, unit weights when none are given
int begin_call(erm_ctx *c, const Batch &all) {
  TRY(ensure(c, c->d_status, 64));
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(32), 0, c->st, dp<int>(c->d_status), (size_t)8, 0);
  TRY(launch_check(c, "k_fill"));
  if (all.random) TRY(upload_args(c, all.seed, all.first_index));
  if (!all.weights) {
    const int need_n = std::min(all.n, c->Bmax);
    if (c->wsol_ones < need_n) {
      TRY(ensure(c, c->d_wsol, sizeof(double) * c->Bmax));
      hipLaunchKernelGGL(k_fill<double>, dim3((c->Bmax + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st,
                         dp<double>(c->d_wsol), (size_t)c->Bmax, 1.0);
      TRY(launch_check(c, "k_fill"));
      c->wsol_ones = c->Bmax;
    }
  }
  return ERM_OK;
}

// enqueues every batch of a call on the context's stream; with uv_out the
// energies are copied out per batch (synchronising)
int enqueue_batches(erm_ctx *c, const Batch &all, int flags, double *uv_out) {
  const int ns = c->ns;
  int bi = 0;
  for (int s0 = 0; s0 < all.n; s0 += c->Bmax, ++bi) {
    Batch bt = all;
    bt.n = std::min(c->Bmax, all.n - s0);
    bt.first = all.first + s0;
    if (all.host_pos) bt.host_pos = all.host_pos + (size_t)s0 * ns * 3;
    if (all.solutes) bt.solutes = all.solutes + s0;
    if (all.weights) bt.weights = all.weights + s0;
    if (all.flex) bt.flex = all.flex + (size_t)s0 * ns * 3;
    c->cur_ev = (NSTAGE_EV * (bi + 1) <= (int)c->bev.size()) ? &c->bev[NSTAGE_EV * bi] : nullptr;
    if (bt.weights) {
      TRY(upload(c, c->d_wsol, bt.weights, bt.n));
      c->wsol_ones = 0;
    }
    int st = c->fp32 ? run_energies<float>(c, bt, flags) : run_energies<double>(c, bt, flags);
    if (st) return st;
    if (flags & DO_HIST) TRY(run_histogram(c, bt.n, bt.first));
    if (uv_out) TRY(copy_out_uv(c, bt.n, uv_out + (size_t)s0 * (c->slvmax + 1)));
  }
  c->cur_ev = nullptr;
  return ERM_OK;
}

// energies (and optionally histograms) over n configurations in batches;
// one synchronisation at the end (the error status and the stage times)
int process(erm_ctx *c, const Batch &all, int flags, double *uv_out) {
  for (int i = 0; i < NSTAGE_OUT; ++i) c->stage_ms[i] = 0;
  const int nb = (all.n + c->Bmax - 1) / c->Bmax;
  TRY(batch_events(c, nb));
  TRY(begin_call(c, all));
  int st = enqueue_batches(c, all, flags, uv_out);
  c->cur_ev = nullptr;
  if (st) return st;
  TRY(check_status(c));
  return stage_times(c, nb, (flags & DO_HIST) != 0);
}

// identity of a captured call: everything its kernels take by value or by
// device address
std::string graph_key(erm_ctx *c, const Batch &all, int flags) {
  std::string k(reinterpret_cast<const char *>(&c->kp), sizeof(Params));
  char buf[256];
  snprintf(buf, sizeof buf, "|%d|%d|%d|%d|%d|%d|%d|%d|%d|%d|%d|%llu|%d", all.n, c->Bmax, flags, c->P.pme_path,
           c->slab_S, c->slab_R, c->xslab_Q, c->xslab_T, c->zcol_T, (int)c->have_usreal, c->wsol_ones,
           (unsigned long long)c->alloc_gen, (int)all.random);
  return k + buf;
}

// random insertions replayed from a captured graph: the first call with a
// new key runs eagerly and then captures its batches; later calls copy the
// generator arguments and launch the graph
int process_graph(erm_ctx *c, const Batch &all, int flags) {
  std::string key = graph_key(c, all, flags);
  if (c->gexec && key == c->gkey) {
    for (int i = 0; i < NSTAGE_OUT; ++i) c->stage_ms[i] = 0;
    TRY(begin_call(c, all));
    HIPCHK(c, hipEventRecord(c->ev[0], c->st));
    HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
    HIPCHK(c, hipEventRecord(c->ev[7], c->st));
    TRY(check_status(c));
    float ms;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[7]) == hipSuccess) c->stage_ms[NSTAGE_OUT - 1] = ms;
    return ERM_OK;
  }
  TRY(process(c, all, flags, nullptr));
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  key = graph_key(c, all, flags);
  const uint64_t gen = c->alloc_gen;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    c->use_graph = false;
    return ERM_OK;
  }
  c->capturing = true;
  int st = enqueue_batches(c, all, flags, nullptr);
  c->capturing = false;
  hipError_t e = hipStreamEndCapture(c->st, &g);
  bool ok = (st == ERM_OK && e == hipSuccess && g && c->alloc_gen == gen);
  if (ok) ok = (hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0) == hipSuccess);
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  if (!ok) {
    // the eager results of this call stand; later calls run eagerly
    c->gexec = nullptr;
    c->use_graph = false;
    c->err = "graph capture failed; eager execution";
    return ERM_OK;
  }
  c->gkey = key;
  return ERM_OK;
}

int check_solutes(erm_ctx *c, int n, const int *solutes, const double *pos) {
  if (n < 0) return fail(c, ERM_E_ARG, "negative count");
  if (c->P.slttype == 1) {
    if (!solutes) return fail(c, ERM_E_ARG, "soln mode needs solute molecule indices");
    for (int i = 0; i < n; ++i) {
      int m = solutes[i];
      if (m < 0 || m >= c->nmol || c->sluvid[m] != 1)
        return fail(c, ERM_E_ARG, "molecule %d is not a solute", m);
    }
  } else {
    if (!pos) return fail(c, ERM_E_ARG, "refs mode needs test-particle coordinates");
  }
  return ERM_OK;
}

template <typename T> int frame_prep_any(erm_ctx *c) { return frame_prep<T>(c); }

int set_frame_common(erm_ctx *c, const double *cell, double w) {
  TRY(need(c, false, false));
  if (c->P.boxshp == ERM_SYS_NONPERIODIC) {
    // no box: the cell is not read; coordinates stay raw and every
    // molecule falls in one cell of unit length
    memset(c->cell, 0, sizeof c->cell);
    memset(c->invcl, 0, sizeof c->invcl);
    c->volume = 0.0;
    for (int d = 0; d < 3; ++d) c->L[d] = 1.0;
    c->stat_weight_system = w;
    ++c->frame_serial;
    frame_params(c);
    TRY(upload(c, c->d_cellm, c->cell, 9));
    return ERM_OK;
  }
  if (!cell) return fail(c, ERM_E_ARG, "cell is NULL");
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (i != j && std::fabs(cell[3 * j + i]) > 1e-8)
        return fail(c, ERM_E_NONCUBOID, "non-cuboid cell: a non-diagonal cell is not supported");
  memcpy(c->cell, cell, sizeof c->cell);
  inverted_cell(c->cell, c->invcl, &c->volume);
  for (int d = 0; d < 3; ++d) c->L[d] = std::fabs(cell[4 * d]);
  c->stat_weight_system = w;
  ++c->frame_serial;
  Params &k = c->kp;
  // the cutoff must stay below half the box (single minimum image)
  bool changed = !c->have_prevcl;
  for (int i = 0; i < 9 && !changed; ++i) changed = (c->prevcl[i] != cell[i]);
  if (changed && c->P.cltype == ERM_EL_PPPM) {
    std::vector<double> engfac;
    if (!pppm_greenfunc(c->P.ms, k.p, c->cell, c->invcl, c->volume, c->P.screen, engfac))
      return fail(c, ERM_E_ARG, "PPPM: negative alias range for this cell");
    TRY(upload_real(c, c->d_engfac, engfac));
  } else if (changed && c->P.cltype == ERM_EL_PME) {
    std::vector<double> splfc[3];
    double coeff[MAXP + 1];
    spline_coeff(k.p, coeff);
    for (int d = 0; d < 3; ++d) {
      splfc[d].resize(c->P.ms[d]);
      spline_axis(c->P.ms[d], k.p, coeff, splfc[d].data());
    }
    std::vector<double> engfac;
    greenfunc(c->P.ms, k.p, c->invcl, c->volume, c->P.screen, splfc, engfac);
    TRY(upload_real(c, c->d_engfac, engfac));
  }
  memcpy(c->prevcl, cell, sizeof c->prevcl);
  c->have_prevcl = true;
  frame_params(c);
  if (changed) TRY(upload(c, c->d_cellm, c->cell, 9));
  return ERM_OK;
}

// ---- solute gradient ---------------------------------------------------
// PME potential of every physical molecule except sorted molecule `excl`
// on the real grid d_ggrid: spread, r2c, Green's function, c2r (the solvent
// counterpart of recpcal_prepare_solute, recpcal.F90:342-406); once per frame
template <typename T> int solvent_grid(erm_ctx *c, int excl) {
  using V = typename C2<T>::type;
  if (c->ggrid_serial == c->frame_serial && c->ggrid_excl == excl) return ERM_OK;
  const Params &k = c->kp;
  const size_t nr = (size_t)k.K1 * k.K2 * k.K3, nh = (size_t)k.H * k.K2 * k.K3;
  TRY(ensure(c, c->d_ggrid, sizeof(T) * nr));
  TRY(ensure(c, c->d_gHk, sizeof(V) * nh));
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((nr + NTHR - 1) / NTHR)), dim3(NTHR), 0, c->st,
                     dp<T>(c->d_ggrid), nr, (T)0);
  TRY(launch_check(c, "k_fill"));
  hipLaunchKernelGGL(k_spread_slv<T>, dim3((c->slvmax + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k, excl,
                     dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<int>(c->d_sg), dp<T>(c->d_sw),
                     dp<T>(c->d_sq), dp<T>(c->d_ggrid));
  TRY(launch_check(c, "k_spread_slv"));
  ERM_KSWITCH(k.K1, (launch_xr2c<T, KK>(c, k)));
  TRY(launch_check(c, "k_xrow_r2c"));
  ERM_KSWITCH(k.K2, (launch_ydense<T, KK, -1>(c, k)));
  TRY(launch_check(c, "k_pme_ycol_reg"));
  ERM_KSWITCH(k.K3, (launch_zgreen<T, KK>(c, k)));
  TRY(launch_check(c, "k_zcol_green"));
  ERM_KSWITCH(k.K2, (launch_ydense<T, KK, +1>(c, k)));
  TRY(launch_check(c, "k_pme_ycol_reg"));
  ERM_KSWITCH(k.K1, (launch_xc2r<T, KK>(c, k)));
  TRY(launch_check(c, "k_xrow_c2r"));
  c->ggrid_serial = c->frame_serial;
  c->ggrid_excl = excl;
  return ERM_OK;
}

struct GradSetup {
  std::vector<double> q, eps, sig, tab;  // per site; tab [ns][nvt][3]
  double qslt = 0.0, rf = 0.0;
  int excl = -1;
};

// one batch of the gradient: adds [sum w, sum w u, sum w grad] into acc
template <typename T> int gradient_batch(erm_ctx *c, const Batch &bt, const GradSetup &g, double *acc) {
  using V = typename C2<T>::type;
  Params k = c->kp;
  const int n = bt.n, ns = c->ns, p = k.p;
  k.ns = ns;
  TRY(ensure(c, c->d_spos, sizeof(double) * (size_t)n * ns * 3));
  TRY(ensure(c, c->d_sref, sizeof(double) * 4 * n));
  TRY(ensure(c, c->d_sg3, sizeof(int) * (size_t)n * ns));
  TRY(ensure(c, c->d_sw3, sizeof(T) * (size_t)n * ns * p));
  TRY(ensure(c, c->d_sqv, sizeof(T) * (size_t)n * ns));
  TRY(ensure(c, c->d_w1h, sizeof(V) * (size_t)n * ns * k.H));
  TRY(ensure(c, c->d_w2h, sizeof(V) * (size_t)n * ns * k.K2));
  TRY(ensure(c, c->d_sltj, sizeof(int) * n));
  TRY(stage_configs(c, bt, k));
  hipLaunchKernelGGL(k_solute_prep<T>, dim3(n), dim3(128), 0, c->st, k, dp<double>(c->d_spos),
                     dp<double>(c->d_qslt), dp<V>(c->d_tw1), dp<V>(c->d_tw2), dp<V>(c->d_tw3),
                     dp<double>(c->d_sref), dp<int>(c->d_sg3), dp<T>(c->d_sw3), dp<T>(c->d_sqv),
                     dp<V>(c->d_w1h), dp<V>(c->d_w2h), (V *)nullptr);
  TRY(launch_check(c, "k_solute_prep"));
  // real space
  double Rg = k.rcmax + 8.0;
  double frac = std::min(1.0, 4.18879 * Rg * Rg * Rg / c->volume);
  int nsplit = std::max(1, std::min(16, (int)std::ceil(frac * c->slvmax / (2.0 * NTHR))));
  const int nsl = std::max(1, NTHR / ns);
  size_t lds = sizeof(double) * 4 * (size_t)nsl * ns + sizeof(T) * ((size_t)7 * ns + (size_t)3 * ns * c->nvt) +
               sizeof(int) * (2 * MAXSEG + 2) + 64;
  if (lds > c->lds_cap) return fail(c, ERM_E_UNSUPPORTED, "solute too large for the gradient kernel's LDS");
  TRY(set_lds(c, (const void *)k_pair_grad<T>, lds, "k_pair_grad"));
  TRY(ensure(c, c->d_gpart, sizeof(double) * (size_t)n * nsplit * 4 * ns));
  hipLaunchKernelGGL(k_pair_grad<T>, dim3(n, nsplit), dim3(NTHR), lds, c->st, k, dp<double>(c->d_spos),
                     dp<double>(c->d_sref), dp<T>(c->d_gq_t), dp<T>(c->d_gtab), dp<int>(c->d_cell_start),
                     dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy),
                     dp<double>(c->d_sz), dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj),
                     dp<double>(c->d_rm), dp<double>(c->d_gpart));
  TRY(launch_check(c, "k_pair_grad"));
  // reciprocal space
  TRY(solvent_grid<T>(c, g.excl));
  TRY(ensure(c, c->d_gphi, sizeof(double) * (size_t)n * ns));
  hipLaunchKernelGGL(k_grad_recip<T>, dim3((n * ns + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k, n,
                     dp<double>(c->d_spos), dp<T>(c->d_ggrid), dp<double>(c->d_gphi));
  TRY(launch_check(c, "k_grad_recip"));
  // weights and sums
  std::vector<double> w(n, c->stat_weight_system);
  if (bt.weights)
    for (int b = 0; b < n; ++b) w[b] = c->stat_weight_system * bt.weights[b];
  TRY(ensure(c, c->d_gw, sizeof(double) * n));
  TRY(upload(c, c->d_gw, w.data(), n));
  TRY(ensure(c, c->d_gub, sizeof(double) * n));
  hipLaunchKernelGGL(k_grad_ub, dim3((n + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k, n, nsplit,
                     dp<double>(c->d_gpart), dp<double>(c->d_gphi), dp<double>(c->d_gqd), g.rf, g.qslt,
                     dp<double>(c->d_gub));
  TRY(launch_check(c, "k_grad_ub"));
  const int nout = 2 + 3 * ns;
  TRY(ensure(c, c->d_gout, sizeof(double) * nout));
  hipLaunchKernelGGL(k_grad_reduce, dim3((nout + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k, n, nsplit,
                     dp<double>(c->d_gpart), dp<double>(c->d_gphi), dp<double>(c->d_gub), dp<double>(c->d_gw),
                     dp<double>(c->d_gepsd), g.rf, dp<double>(c->d_gout));
  TRY(launch_check(c, "k_grad_reduce"));
  std::vector<double> out(nout);
  HIPCHK(c, hipMemcpyAsync(out.data(), c->d_gout.p, sizeof(double) * nout, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  for (int i = 0; i < nout; ++i) acc[i] += out[i];
  return ERM_OK;
}

// ---- several test-particle species on one solvent ------------------------
// (no counterpart in the reference, which runs one solute species per run,
// engproc.F90:89-91).  Members keep their own topology tables, bins and
// accumulators; the first member holds the frame's solvent data and the
// batch buffers.  A batch mixes configurations of every member in runs
// (largest solutes first); the solute-dependent heavy stages (z columns,
// real-space pairs) run as one persistent launch over all runs or as one
// launch per run, the solvent-side PME stages once over the whole batch,
// the light per-run stages per run on the same stream.
constexpr int GROUP_MAXRUN = 16;
}  // namespace

struct erm_group {
  std::vector<erm_ctx *> ctx;
  int Bmax = 1024;
  // persistent grids over all runs: bit 0 the z columns, bit 1 the pairs
  // (a clear bit: one launch per run)
  int persistent = 2;
  uint64_t frame_serial = 0;
  ermh::DBuf d_runs, d_counter;
  GRun *h_runs = nullptr;  // pinned run tables of the batches of a call
  int h_runs_cap = 0;
  double stage_ms[ermh::NSTAGE_OUT] = {};
  std::string err;
  // graph replay of the batch sequence: the first call of a shape runs
  // eagerly and captures; calls with the same key launch the graph
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::string gkey;
};

namespace {

// members' streams replaced by the first member's for the length of a group
// call, so every launch of the call is ordered on one stream
struct StreamSwap {
  erm_group *g;
  std::vector<hipStream_t> saved;
  explicit StreamSwap(erm_group *gg) : g(gg) {
    for (erm_ctx *c : g->ctx) saved.push_back(c->st);
    for (erm_ctx *c : g->ctx) c->st = g->ctx[0]->st;
  }
  ~StreamSwap() {
    for (size_t i = 0; i < g->ctx.size(); ++i) g->ctx[i]->st = saved[i];
  }
};

int gfail(erm_group *g, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g->err = buf;
  return code;
}

// the members describe the same solvent, grid and pair parameters
int group_compatible(erm_group *g) {
  erm_ctx *a = g->ctx[0];
  for (size_t i = 0; i < g->ctx.size(); ++i) {
    erm_ctx *c = g->ctx[i];
    if (!c->have_topo || !c->have_bins) return gfail(g, ERM_E_ARG, "member %zu lacks topology or bins", i);
    if (c->P.slttype != 2 || c->bfcoord.empty())
      return gfail(g, ERM_E_ARG, "member %zu is not a rigid test particle with bfcoord", i);
    if (c->kp.ins_org != ERM_INSORG_ORIGIN || c->kp.ins_pos != ERM_INSPOS_RANDOM || c->kp.ins_rot != ERM_INSROT_RANDOM)
      return gfail(g, ERM_E_UNSUPPORTED, "member %zu: groups insert at random positions and orientations only", i);
    if (c->P.cltype == ERM_EL_COULOMB) return gfail(g, ERM_E_UNSUPPORTED, "groups need PME or PPPM");
    if (c->P.device != a->P.device || c->fp32 != a->fp32 || c->P.splodr != a->P.splodr ||
        memcmp(c->P.ms, a->P.ms, sizeof a->P.ms) != 0 || c->P.cltype != a->P.cltype ||
        c->P.ljswitch != a->P.ljswitch || c->P.elecut != a->P.elecut || c->P.lwljcut != a->P.lwljcut ||
        c->P.upljcut != a->P.upljcut || c->P.screen != a->P.screen ||
        c->P.block_threshold != a->P.block_threshold || c->P.estype != a->P.estype || c->P.wgtslf != a->P.wgtslf)
      return gfail(g, ERM_E_ARG, "member %zu: device, precision or parameters differ from member 0", i);
    if (c->slvmax != a->slvmax || c->nvt != a->nvt || c->vtypes != a->vtypes)
      return gfail(g, ERM_E_ARG, "member %zu: solvent differs from member 0", i);
    for (int k = 0; k < a->slvmax; ++k) {
      const int m = c->tagpt[k], ma = a->tagpt[k];
      const int n = c->mol_begin[m + 1] - c->mol_begin[m];
      if (n != a->mol_begin[ma + 1] - a->mol_begin[ma] || c->uvspec[m] != a->uvspec[ma] ||
          c->mol_charge[m] != a->mol_charge[ma])
        return gfail(g, ERM_E_ARG, "member %zu: solvent molecule %d differs from member 0", i, k);
      for (int s = 0; s < n; ++s)
        if (c->charge[c->mol_begin[m] + s] != a->charge[a->mol_begin[ma] + s] ||
            c->ljtype[c->mol_begin[m] + s] != a->ljtype[a->mol_begin[ma] + s])
          return gfail(g, ERM_E_ARG, "member %zu: solvent site of molecule %d differs from member 0", i, k);
    }
  }
  return ERM_OK;
}

// one batch of the runs R[0..nrun) (configurations [0, n) of the batch)
template <typename T> int group_batch(erm_group *g, const GRun *R, int nrun, const uint64_t *offset, int n,
                                      const GRun *d_runs, int nitem_z, int nitem_p) {
  using V = typename C2<T>::type;
  erm_ctx *c = g->ctx[0];
  Params k = c->kp;
  const int S = c->slvmax, H = k.H, K2 = k.K2, K3 = k.K3, p = k.p;
  size_t sites = 0;
  int Emax = 0;
  for (int r = 0; r < nrun; ++r) sites += (size_t)R[r].ncfg * R[r].ns;
  for (erm_ctx *m : g->ctx) Emax = std::max(Emax, m->ermax);
  const int nzg = zcol_groups(c, true, (int)sizeof(V));
  TRY(ensure(c, c->d_spos, sizeof(double) * sites * 3));
  TRY(ensure(c, c->d_sref, sizeof(double) * 4 * n));
  TRY(ensure(c, c->d_sg3, sizeof(int) * sites));
  TRY(ensure(c, c->d_sw3, sizeof(T) * sites * p));
  TRY(ensure(c, c->d_sqv, sizeof(T) * sites));
  TRY(ensure(c, c->d_w1h, sizeof(V) * sites * H));
  TRY(ensure(c, c->d_w2h, sizeof(V) * sites * K2));
  TRY(ensure(c, c->d_uv, sizeof(double) * (size_t)n * (S + 1)));
  TRY(ensure(c, c->d_sltj, sizeof(int) * n));
  TRY(ensure(c, c->d_selfp, sizeof(double) * (size_t)n * nzg));
  TRY(ensure(c, c->d_Hk, sizeof(V) * (size_t)n * K3 * K2 * H));
  TRY(ensure(c, c->d_esite, sizeof(double) * (size_t)n * c->nsites));
  TRY(ensure(c, c->d_hist, sizeof(double) * (size_t)n * Emax));
  TRY(ensure(c, c->d_w, sizeof(double) * n));
  TRY(ensure(c, c->d_flceng, sizeof(double) * (size_t)n * MAXSPEC));
  TRY(ensure(c, g->d_counter, sizeof(int) * 4));
  TRY(mark(c, 0));
  hipLaunchKernelGGL(k_fill<int>, dim3((n + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, dp<int>(c->d_sltj), (size_t)n,
                     -1);
  TRY(launch_check(c, "k_fill"));
  // poses and solute splines, per run with the member's tables
  for (int r = 0; r < nrun; ++r) {
    erm_ctx *m = g->ctx[R[r].member];
    Params km = m->kp;
    km.ns = m->ns;
    const size_t so = R[r].site0;
    const size_t lpose = (m->ns <= POSE_LDS_SITES) ? sizeof(double) * 4 * m->ns : 0;
    hipLaunchKernelGGL(lpose ? k_pose<true> : k_pose<false>, dim3(R[r].ncfg), dim3(64), lpose, c->st, km, R[r].ncfg,
                       dp<uint64_t>(m->d_args),
                       offset[r], dp<double>(m->d_bf), 0, dp<double>(m->d_sitemass) + m->mol_begin[m->slt_template],
                       dp<double>(m->d_cellm), dp<double>(m->d_aggcen), dp<double>(c->d_spos) + so * 3);
    TRY(launch_check(c, "k_pose"));
    hipLaunchKernelGGL(k_solute_prep<T>, dim3(R[r].ncfg), dim3(128), 0, c->st, km, dp<double>(c->d_spos) + so * 3,
                       dp<double>(m->d_qslt), dp<V>(c->d_tw1), dp<V>(c->d_tw2), dp<V>(c->d_tw3),
                       dp<double>(c->d_sref) + 4 * (size_t)R[r].cfg0, dp<int>(c->d_sg3) + so, dp<T>(c->d_sw3) + so * p,
                       dp<T>(c->d_sqv) + so, dp<V>(c->d_w1h) + so * H, dp<V>(c->d_w2h) + so * K2, (V *)nullptr);
    TRY(launch_check(c, "k_solute_prep"));
  }
  TRY(mark(c, 1));
  // the work counters of the persistent launches ([0] z columns, [1] pairs)
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(32), 0, c->st, dp<int>(g->d_counter), (size_t)4, 0);
  TRY(launch_check(c, "k_fill"));
  // z columns over every run: one persistent launch, or one launch per run
  if (g->persistent & 1) {
    int rz = ERM_E_UNSUPPORTED;
    ERM_KSWITCH(K3, (rz = launch_zcol_fft_runs<T, KK>(c, k, nzg, d_runs, nrun, nitem_z, dp<int>(g->d_counter))));
    TRY(rz);
    TRY(launch_check(c, "k_pme_zcol_fft_runs"));
  } else {
    for (int r = 0; r < nrun; ++r) {
      Params kr = k;
      kr.ns = R[r].ns;
      int rz = ERM_E_UNSUPPORTED;
      ERM_KSWITCH(K3, (rz = launch_zcol_fft<T, KK>(c, kr, R[r].ncfg, nzg, (size_t)R[r].site0, (size_t)R[r].cfg0)));
      TRY(rz);
      TRY(launch_check(c, "k_pme_zcol_fft"));
    }
  }
  TRY(mark(c, 2));
  if (!xslab_folds_y(c)) {
    ERM_KSWITCH(K2, (launch_ycol_reg<T, KK>(c, k, n)));
    TRY(launch_check(c, "y columns"));
  }
  TRY(mark(c, 3));
  {
    const size_t lx = 2 * sizeof(T) * (size_t)c->xslab_Q * H * K2;
    int rc = ERM_E_UNSUPPORTED;
    ERM_KSWITCH(k.K1, (rc = launch_xslab<T, KK>(c, k, n, lx)));
    TRY(rc);
    TRY(launch_check(c, "k_pme_xslab"));
  }
  TRY(mark(c, 4));
  for (int r = 0; r < nrun; ++r) {
    erm_ctx *m = g->ctx[R[r].member];
    Params kg = k;
    kg.solute_charge = m->mol_charge[m->slt_template];
    const size_t c0 = R[r].cfg0;
    TRY(launch_mol_sum(c, kg, R[r].ncfg, dp<double>(c->d_esite) + c0 * c->nsites, dp<int>(c->d_sltj) + c0,
                       dp<double>(c->d_uv) + c0 * (S + 1)));
  }
  TRY(mark(c, 5));
  {
    int nsmax = 0;
    for (int r = 0; r < nrun; ++r) nsmax = std::max(nsmax, R[r].ns);
    const int segcap = 2 * k.nb[1] * k.nb[2];
    const size_t lt = pair_tile_lds<T>(nsmax, c->nvt, segcap, NTHR / 64);
    if (lt > c->lds_cap) return fail(c, ERM_E_UNSUPPORTED, "solute too large for the pair kernel's LDS");
    if (g->persistent & 2) {
      const void *f = (const void *)k_pair_tile_runs<T, false>;
      TRY(set_lds(c, f, lt, "k_pair_tile_runs"));
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, NTHR, lt) != hipSuccess || occ < 1) {
        (void)hipGetLastError();
        occ = 1;
      }
      const int grid = std::min(nitem_p, std::max(1, occ * c->ncu));
      hipLaunchKernelGGL((k_pair_tile_runs<T, false>), dim3(grid), dim3(NTHR), lt, c->st, k, dp<double>(c->d_spos),
                         dp<double>(c->d_sref), dp<int>(c->d_cell_start), dp<int>(c->d_mol_first),
                         dp<int>(c->d_mol_ns), dp<double>(c->d_sx), dp<double>(c->d_sy), dp<double>(c->d_sz),
                         dp<T>(c->d_sq), dp<int>(c->d_sct), dp<int>(c->d_sltj), dp<double>(c->d_rm), d_runs, nrun,
                         nitem_p, dp<int>(g->d_counter) + 1, dp<double>(c->d_mol_d1), dp<double>(c->d_uv));
      TRY(launch_check(c, "k_pair_tile_runs"));
    } else {
      TRY(set_lds(c, (const void *)k_pair_tile<T, false>, lt, "k_pair_tile"));
      for (int r = 0; r < nrun; ++r) {
        Params kr = k;
        kr.ns = R[r].ns;
        const size_t c0 = R[r].cfg0;
        hipLaunchKernelGGL((k_pair_tile<T, false>), dim3(R[r].ncfg, R[r].nsplit), dim3(NTHR), lt, c->st, kr,
                           dp<double>(c->d_spos) + (size_t)R[r].site0 * 3, dp<double>(c->d_sref) + 4 * c0,
                           static_cast<const T *>(R[r].q_t), static_cast<const T *>(R[r].lj_t),
                           dp<int>(c->d_cell_start), dp<int>(c->d_mol_first), dp<int>(c->d_mol_ns),
                           dp<double>(c->d_sx), dp<double>(c->d_sy), dp<double>(c->d_sz), dp<T>(c->d_sq),
                           dp<int>(c->d_sct), dp<int>(c->d_sltj) + c0, dp<double>(c->d_rm), R[r].lj64, R[r].q64,
                           dp<double>(c->d_mol_d1), dp<double>(c->d_uv) + c0 * (S + 1));
        TRY(launch_check(c, "k_pair_tile"));
      }
    }
  }
  TRY(mark(c, 6));
  // self terms: the member's realcal_self from its first configuration
  // (reused for the rigid solute, engproc.F90:712-730), then uvengy(0)
  for (int r = 0; r < nrun; ++r) {
    erm_ctx *m = g->ctx[R[r].member];
    Params km = m->kp;
    km.ns = m->ns;
    km.solute_charge = m->mol_charge[m->slt_template];
    TRY(ensure(m, m->d_usreal, sizeof(double)));
    if (!m->have_usreal) {
      hipLaunchKernelGGL(k_selfreal, dim3(1), dim3(NTHR), 0, c->st, km, dp<double>(c->d_spos) + (size_t)R[r].site0 * 3,
                         0, dp<double>(m->d_qslt), dp<double>(m->d_usreal));
      TRY(launch_check(c, "k_selfreal"));
      m->have_usreal = true;
    }
    const size_t c0 = R[r].cfg0;
    hipLaunchKernelGGL(k_u0, dim3((R[r].ncfg + 63) / 64), dim3(64), 0, c->st, km, R[r].ncfg, nzg,
                       dp<double>(c->d_selfp) + c0 * nzg, dp<double>(m->d_usreal), 0,
                       dp<double>(c->d_uv) + c0 * (S + 1));
    TRY(launch_check(c, "k_u0"));
  }
  TRY(mark(c, 7));
  // update_histogram per run into the member's accumulators, then the
  // correlation matrices
  HistBufs hbs[GROUP_MAXRUN];
  for (int r = 0; r < nrun; ++r) {
    erm_ctx *m = g->ctx[R[r].member];
    const size_t c0 = R[r].cfg0;
    HistBufs hb;
    hb.uv = dp<double>(c->d_uv) + c0 * (S + 1);
    hb.wsol = dp<double>(c->d_wsol);  // ones
    hb.hist = dp<double>(c->d_hist) + c0 * Emax;
    hb.w = dp<double>(c->d_w) + c0;
    hb.flceng = dp<double>(c->d_flceng) + c0 * MAXSPEC;
    hb.wi = nullptr;
    hb.sltj = dp<int>(c->d_sltj) + c0;
    hb.status = dp<int>(c->d_status);
    hb.mol_spec = dp<int>(c->d_mol_spec);  // the energies are in member 0's sorted order
    hipEvent_t *keep = m->cur_ev;
    m->cur_ev = nullptr;
    const int st = run_histogram_at(m, R[r].ncfg, (int)offset[r], hb, true);
    m->cur_ev = keep;
    if (st) return fail(c, st, "member %d: %s", R[r].member, m->err.c_str());
    hbs[r] = hb;
  }
  TRY(mark(c, 8));
  for (int r = 0; r < nrun; ++r) {
    erm_ctx *m = g->ctx[R[r].member];
    const int st = run_corr_at(m, R[r].ncfg, hbs[r]);
    if (st) return fail(c, st, "member %d: %s", R[r].member, m->err.c_str());
  }
  TRY(mark(c, 9));
  return ERM_OK;
}

}  // namespace

// ===================================================================== C-ABI
extern "C" {

int32_t erm_build_edges(const erm_hist *h, int32_t self_energy, double *ercrd_out, int32_t cap, int32_t *pemax_out,
                        int32_t *pesoft_out) {
  // enginit edge construction (engproc.F90:163-218)
  if (!h || !ercrd_out || !pemax_out || !pesoft_out) return ERM_E_ARG;
  const int RGLMAX = 5, LARGE = 10000;
  double ecdmin = h->ecdmin, ecfmns = h->ecfmns, ecdcen = h->ecdcen, eccore = h->eccore, ecdmax = h->ecdmax;
  double ectmvl = h->finfac * h->ecfbin;
  ecdmin = ecdmin - ectmvl;
  ecfmns = ecfmns - ectmvl;
  double ecmns0 = ecdcen - ectmvl;
  double ecpls0 = 2.0 * ecdcen - ecmns0;
  double ecfpls = 2.0 * ecdcen - ecfmns;
  eccore = eccore + ecfpls - ecdcen;
  const double cdrgvl[7] = {ecdmin, ecfmns, ecmns0, ecpls0, ecfpls, eccore, ecdmax};
  int uprgcd[7];
  uprgcd[0] = 1;
  for (int regn = 1; regn <= RGLMAX; ++regn) {
    double factor = (regn == 1 || regn == 5) ? h->eclbin : (regn == 2 || regn == 4) ? h->ecfbin : h->ec0bin;
    double incre = cdrgvl[regn] - cdrgvl[regn - 1];
    if (incre < factor) return ERM_E_ENG_ECD;
    uprgcd[regn] = uprgcd[regn - 1] + (int)fortran_nint(incre / factor);
  }
  int pesoft = uprgcd[RGLMAX] - uprgcd[0];
  int pemax = pesoft + h->pecore;
  uprgcd[RGLMAX + 1] = pemax;
  if (pemax > LARGE) return ERM_E_ENG_SIZ;
  std::vector<double> e((size_t)std::max(pemax, uprgcd[RGLMAX]) + 2, 0.0);
  for (int r = 0; r <= RGLMAX + 1; ++r) e[uprgcd[r]] = cdrgvl[r];
  int nreg = (h->pecore == 0) ? RGLMAX : RGLMAX + 1;
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
  if (n > cap) return ERM_E_ARG;
  for (int i = 0; i < n; ++i) ercrd_out[i] = e[i + 1];
  *pemax_out = n;
  *pesoft_out = pesoft;
  return ERM_OK;
}

uint64_t erm_hash_double(const double *v, int32_t n) {
  uint64_t x = 0, b;
  for (int32_t i = 0; i < n; ++i) {
    memcpy(&b, &v[i], sizeof b);
    x = ((x << 7) | (x >> 57)) ^ b;
  }
  return x;
}

int32_t erm_create(const erm_params *prm, erm_ctx **out) {
  if (!prm || !out) return ERM_E_ARG;
  *out = nullptr;
  const erm_params &P = *prm;
  if (P.slttype < 1 || P.slttype > 3) return ERM_E_ARG;
  if (P.splodr < 2 || P.splodr > MAXP) return ERM_E_UNSUPPORTED;
  if (P.ljswitch < 0 || P.ljswitch > 3) return ERM_E_ARG;
  if (P.precision != ERM_FP64 && P.precision != ERM_FP32) return ERM_E_ARG;
  if (P.pme_path < 0 || P.pme_path > 3) return ERM_E_ARG;
  // cltype / boxshp (setconf.F90:326-339, 368-372; insertion.F90:155)
  if (P.cltype == 1) return ERM_E_UNSUPPORTED;
  if (P.cltype != ERM_EL_COULOMB && P.cltype != ERM_EL_PME && P.cltype != ERM_EL_PPPM) return ERM_E_ARG;
  if (P.boxshp != ERM_SYS_NONPERIODIC && P.boxshp != ERM_SYS_PERIODIC) return ERM_E_ARG;
  if (P.boxshp == ERM_SYS_NONPERIODIC && (P.cltype != ERM_EL_COULOMB || P.estype == 2 || P.slttype != 1))
    return ERM_E_ARG;
  if (P.cltype == ERM_EL_PPPM && P.splodr % 2 == 1) return ERM_E_ARG;
  const bool bare = (P.cltype == ERM_EL_COULOMB);
  // bare Coulomb has no grid: a small placeholder keeps the buffers sized
  int msv[3];
  for (int d = 0; d < 3; ++d) msv[d] = bare ? 16 : P.ms[d];
  FftAxis ax[3];
  for (int d = 0; d < 3; ++d)
    if (msv[d] < 2 || msv[d] > 512 || msv[d] < P.splodr || !fft_axis(msv[d], ax[d])) return ERM_E_UNSUPPORTED;
  if ((!bare && !(P.screen > 0)) || !(P.elecut > 0) || !(P.upljcut >= P.lwljcut) || !(P.block_threshold > 0))
    return ERM_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ERM_E_DEVICE;
  if (P.device < 0 || P.device >= ndev) return ERM_E_ARG;
  if (hipSetDevice(P.device) != hipSuccess) return ERM_E_DEVICE;
  if (scratch_init(P.device) != ERM_OK) return ERM_E_DEVICE;
  erm_ctx *c = new erm_ctx();
  c->P = P;
  for (int d = 0; d < 3; ++d) c->P.ms[d] = msv[d];
  c->fp32 = (P.precision == ERM_FP32);
  c->Bmax = P.max_batch > 0 ? P.max_batch : 256;
  {
    // LDS a workgroup may opt into (160 KB per CU on gfx950), with room for
    // the kernels' static arrays
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, P.device) != hipSuccess)
      optin = 0;
    (void)hipGetLastError();
    c->lds_cap = std::max<size_t>(64 * 1024, std::min<size_t>(150 * 1024, optin > 8192 ? (size_t)optin - 8192 : 0));
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, P.device) == hipSuccess && ncu > 0)
      c->ncu = ncu;
    (void)hipGetLastError();
  }
  if (hipStreamCreate(&c->st) != hipSuccess) {
    delete c;
    return ERM_E_DEVICE;
  }
  for (int i = 0; i < 8; ++i)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) {
      erm_destroy(c);
      return ERM_E_DEVICE;
    }
  Params &k = c->kp;
  memset(&k, 0, sizeof k);
  k.K1 = msv[0];
  k.K2 = msv[1];
  k.K3 = msv[2];
  k.cltype = P.cltype;
  k.boxshp = P.boxshp;
  k.elecut = P.elecut;
  k.H = k.K1 / 2 + 1;
  k.p = P.splodr;
  k.ax1 = ax[0];
  k.ax2 = ax[1];
  k.ax3 = ax[2];
  spline_coeff(k.p, k.coeff);
  k.screen = P.screen;
  k.ljswitch = P.ljswitch;
  k.lwljcut = P.lwljcut;
  k.upljcut = P.upljcut;
  k.lw2 = ipow_h(P.lwljcut, 2);
  k.up2 = ipow_h(P.upljcut, 2);
  k.el2 = ipow_h(P.elecut, 2);
  k.rcmax = std::max(P.upljcut, P.elecut);
  k.cut2 = ipow_h(k.rcmax, 2);
  k.lw3 = ipow_h(P.lwljcut, 3);
  k.up3 = ipow_h(P.upljcut, 3);
  k.lw6 = k.lw3 * k.lw3;
  k.up6 = k.up3 * k.up3;
  gmx_force(12, P.lwljcut, P.upljcut, &k.repA, &k.repB, &k.repC);
  gmx_force(6, P.lwljcut, P.upljcut, &k.attA, &k.attB, &k.attC);
  k.slttype = P.slttype;
  k.wgtslf = P.wgtslf;
  k.estype = P.estype;
  k.selfcal = P.selfcal;
  k.force = P.force_calculation;
  k.temp = P.temp;
  // dynamic LDS above 64 KB for the PME plane kernel on large grids
  size_t vb = c->fp32 ? sizeof(float2) : sizeof(double2);
  size_t lp = 2 * (size_t)k.K2 * k.H * vb;
  if (lp > 160 * 1024 || (size_t)((k.K2 + 1) / 2) * k.K1 > (size_t)k.K2 * k.H) {
    erm_destroy(c);
    return ERM_E_UNSUPPORTED;
  }
  if (lp > 64 * 1024) {
    const void *fns_f[3] = {(const void *)k_pme_plane<float>, (const void *)k_pme_slab<float, 16>,
                            (const void *)k_pme_slab<float, 32>};
    const void *fns_d[3] = {(const void *)k_pme_plane<double>, (const void *)k_pme_slab<double, 16>,
                            (const void *)k_pme_slab<double, 32>};
    for (int i = 0; i < 3; ++i) {
      hipError_t e = hipFuncSetAttribute(c->fp32 ? fns_f[i] : fns_d[i], hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lp);
      if (e != hipSuccess) {
        erm_destroy(c);
        return ERM_E_UNSUPPORTED;
      }
    }
  }
  int s = build_tw(c, c->d_tw1, k.K1);
  if (!s) s = build_tw(c, c->d_tw2, k.K2);
  if (!s) s = build_tw(c, c->d_tw3, k.K3);
  if (!s) s = ensure(c, c->d_status, 64);
  if (s) {
    erm_destroy(c);
    return s;
  }
  *out = c;
  return ERM_OK;
}

void erm_destroy(erm_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->P.device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (DBuf *b : c->all()) b->release();
  for (int i = 0; i < 8; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  for (hipEvent_t e : c->bev) (void)hipEventDestroy(e);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->h_args) (void)hipHostFree(c->h_args);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char *erm_last_error(const erm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int32_t erm_set_topology(erm_ctx *c, int32_t natom, int32_t nmol, const int32_t *mol_begin, const double *charge,
                         const int32_t *ljtype, int32_t ntype, const double *lj_eps, const double *lj_sig2,
                         const int32_t *sluvid, const int32_t *uvspec, const double *mol_charge,
                         const double *sitemass, const double *bfcoord) {
  if (!c) return ERM_E_ARG;
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (natom <= 0 || nmol <= 0 || ntype <= 0 || !mol_begin || !charge || !ljtype || !lj_eps || !lj_sig2 || !sluvid ||
      !uvspec || !mol_charge || !sitemass)
    return fail(c, ERM_E_ARG, "missing topology array");
  if (mol_begin[0] != 0 || mol_begin[nmol] != natom) return fail(c, ERM_E_ARG, "mol_begin must span 0..natom");
  for (int i = 0; i < nmol; ++i)
    if (mol_begin[i + 1] <= mol_begin[i]) return fail(c, ERM_E_ARG, "molecule %d has no sites", i);
  for (int a = 0; a < natom; ++a)
    if (ljtype[a] < 0 || ljtype[a] >= ntype) return fail(c, ERM_E_ARG, "ljtype out of range at site %d", a);
  c->natom = natom;
  c->nmol = nmol;
  c->ntype = ntype;
  c->mol_begin.assign(mol_begin, mol_begin + nmol + 1);
  c->charge.assign(charge, charge + natom);
  c->ljtype.assign(ljtype, ljtype + natom);
  c->lj_eps.assign(lj_eps, lj_eps + (size_t)ntype * ntype);
  c->lj_sig2.assign(lj_sig2, lj_sig2 + (size_t)ntype * ntype);
  c->sluvid.assign(sluvid, sluvid + nmol);
  c->uvspec.assign(uvspec, uvspec + nmol);
  c->mol_charge.assign(mol_charge, mol_charge + nmol);
  c->sitemass.assign(sitemass, sitemass + natom);
  // tagpt (engproc.F90:336-347) and the solute list
  c->tagpt.clear();
  c->sltlist.clear();
  c->kof.assign(nmol, -1);
  for (int i = 0; i < nmol; ++i) {
    if (sluvid[i] < 0 || sluvid[i] > 3) return fail(c, ERM_E_ENG_BUG, "sluvid out of range");
    if (sluvid[i] == 0 || sluvid[i] == 1) {
      c->kof[i] = (int)c->tagpt.size();
      c->tagpt.push_back(i);
    }
    if (sluvid[i] > 0) c->sltlist.push_back(i);
  }
  c->slvmax = (int)c->tagpt.size();
  if (c->sltlist.empty()) return fail(c, ERM_E_ARG, "no solute or test particle");
  if (c->slvmax == 0) return fail(c, ERM_E_ARG, "no physical molecule");
  if (c->P.slttype != 1) {
    if (c->sltlist.size() != 1 || c->sltlist[0] != nmol - 1)
      return fail(c, ERM_E_ARG, "the test particle must be the single last molecule (eng_ins)");
    for (int i = 0; i < nmol; ++i)
      if (sluvid[i] == 1) return fail(c, ERM_E_ARG, "solute molecule in a refs system (eng_par)");
  } else {
    for (int i = 0; i < nmol; ++i)
      if (sluvid[i] >= 2) return fail(c, ERM_E_ARG, "test particle in a soln system (eng_par)");
  }
  c->slt_template = c->sltlist[0];
  c->ns = c->mol_begin[c->slt_template + 1] - c->mol_begin[c->slt_template];
  for (int m : c->sltlist)
    if (c->mol_begin[m + 1] - c->mol_begin[m] != c->ns) return fail(c, ERM_E_ARG, "solutes differ in size");
  if (c->ns > 4096) return fail(c, ERM_E_UNSUPPORTED, "solute with more than 4096 sites");
  // compact LJ types of the sites that act as solvent: every physical
  // molecule, except the solute itself when it is the only one
  std::vector<int> cv(ntype, -1), cvt(natom, 0);
  c->vtypes.clear();
  for (int k = 0; k < c->slvmax; ++k) {
    int m = c->tagpt[k];
    if (c->sltlist.size() == 1 && m == c->sltlist[0]) continue;
    for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a) {
      int t = ljtype[a];
      if (cv[t] < 0) {
        cv[t] = (int)c->vtypes.size();
        c->vtypes.push_back(t);
      }
      cvt[a] = cv[t];
    }
  }
  c->nvt = (int)c->vtypes.size();
  size_t tbytes = c->fp32 ? 4 : 8;
  if ((size_t)2 * c->ns * c->nvt * tbytes + 4 * (size_t)c->ns * tbytes > 100 * 1024)
    return fail(c, ERM_E_UNSUPPORTED, "solute x solvent LJ table too large for LDS");
  c->total_charge = 0.0;
  for (int i = 0; i < nmol; ++i)
    if (sluvid[i] == 0 || sluvid[i] == 1) c->total_charge = c->total_charge + mol_charge[i];
  c->kp.total_charge = c->total_charge;
  c->kp.slvmax = c->slvmax;
  c->kp.ns = c->ns;
  c->kp.nvt = c->nvt;
  c->bfcoord.clear();
  if (bfcoord) c->bfcoord.assign(bfcoord, bfcoord + 3 * c->ns);
  // test-particle radius about its centroid (sizes the pair kernel's grid);
  // solution-mode solutes keep the default
  if (bfcoord && c->ns > 0) {
    double cm[3] = {0, 0, 0}, r2 = 0.0;
    for (int s = 0; s < c->ns; ++s)
      for (int d = 0; d < 3; ++d) cm[d] += bfcoord[3 * s + d] / c->ns;
    for (int s = 0; s < c->ns; ++s) {
      double q = 0.0;
      for (int d = 0; d < 3; ++d) q += (bfcoord[3 * s + d] - cm[d]) * (bfcoord[3 * s + d] - cm[d]);
      r2 = std::max(r2, q);
    }
    c->slt_radius = std::sqrt(r2);
  }
  TRY(upload(c, c->d_mol_begin, mol_begin, nmol + 1));
  TRY(upload(c, c->d_charge, charge, natom));
  TRY(upload(c, c->d_tagpt, c->tagpt.data(), c->slvmax));
  TRY(upload(c, c->d_kof, c->kof.data(), c->nmol));
  TRY(upload(c, c->d_cvt, cvt.data(), natom));
  TRY(upload(c, c->d_mol_charge, mol_charge, nmol));
  TRY(upload(c, c->d_uvspec, uvspec, nmol));
  TRY(upload(c, c->d_sitemass, sitemass, natom));
  if (bfcoord) TRY(upload(c, c->d_bf, bfcoord, 3 * c->ns));
  TRY(solute_tables(c));
  TRY(ensure(c, c->d_pos, sizeof(double) * 3 * natom));
  TRY(sync(c));
  c->have_topo = true;
  c->have_frame = false;
  c->have_usreal = false;
  return ERM_OK;
}

int32_t erm_set_bins(erm_ctx *c, int32_t numslv, const int32_t *uvmax, const int32_t *uvsoft, const double *uvcrd,
                     int32_t esmax, const double *escrd) {
  if (!c) return ERM_E_ARG;
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (numslv <= 0 || numslv > MAXSPEC || !uvmax || !uvsoft || !uvcrd)
    return fail(c, ERM_E_ARG, "numslv must be 1..%d with edge arrays", MAXSPEC);
  if (c->P.selfcal && (esmax <= 0 || !escrd)) return fail(c, ERM_E_ARG, "selfcal needs escrd");
  c->numslv = numslv;
  c->uvmax.assign(uvmax, uvmax + numslv);
  c->uvsoft.assign(uvsoft, uvsoft + numslv);
  c->specoff.assign(numslv, 0);
  int E = 0;
  for (int i = 0; i < numslv; ++i) {
    if (uvmax[i] <= 0) return fail(c, ERM_E_ARG, "empty species grid");
    c->specoff[i] = E;
    E += uvmax[i];
  }
  // k_bin holds the edges, the counts and per-thread species sums in LDS
  if ((size_t)E * 12 + (size_t)numslv * NTHR * 8 > c->lds_cap)
    return fail(c, ERM_E_UNSUPPORTED, "ermax %d with %d species too large", E, numslv);
  c->ermax = E;
  c->uvcrd.assign(uvcrd, uvcrd + E);
  c->esmax = c->P.selfcal ? esmax : 0;
  c->escrd.clear();
  if (c->P.selfcal) c->escrd.assign(escrd, escrd + esmax);
  c->kp.ermax = E;
  c->kp.numslv = numslv;
  c->kp.esmax = c->esmax;
  TRY(upload(c, c->d_edges, uvcrd, E));
  TRY(upload(c, c->d_specoff, c->specoff.data(), numslv));
  TRY(upload(c, c->d_uvmax, uvmax, numslv));
  TRY(upload(c, c->d_uvsoft, uvsoft, numslv));
  if (c->esmax) TRY(upload(c, c->d_escrd, escrd, c->esmax));
  TRY(ensure(c, c->d_edens, sizeof(double) * E));
  if (c->P.corrcal) TRY(ensure(c, c->d_ecorr, sizeof(double) * (size_t)E * E));
  TRY(ensure(c, c->d_eself, sizeof(double) * std::max(1, c->esmax)));
  TRY(ensure(c, c->d_slnuv, sizeof(double) * (numslv + 1)));
  TRY(ensure(c, c->d_mnmx, sizeof(unsigned long long) * 2 * (numslv + 1)));
  TRY(ensure(c, c->d_scal, sizeof(double) * 8));
  c->have_bins = true;
  // first-time state as enginit leaves it (engproc.F90:255-260)
  std::vector<unsigned long long> mm(2 * (numslv + 1));
  for (int i = 0; i <= numslv; ++i) {
    mm[2 * i] = ~0ULL;
    mm[2 * i + 1] = 0ULL;
  }
  TRY(upload(c, c->d_mnmx, mm.data(), mm.size()));
  double sc[8] = {0, 0, 0, -1.0e50, 0, 0, 0, 0};
  TRY(upload(c, c->d_scal, sc, 8));
  TRY(sync(c));
  return erm_engclear(c);
}

int32_t erm_engclear(erm_ctx *c) {
  TRY(need(c, false, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  HIPCHK(c, hipMemsetAsync(c->d_edens.p, 0, sizeof(double) * c->ermax, c->st));
  if (c->P.corrcal) HIPCHK(c, hipMemsetAsync(c->d_ecorr.p, 0, sizeof(double) * (size_t)c->ermax * c->ermax, c->st));
  HIPCHK(c, hipMemsetAsync(c->d_eself.p, 0, sizeof(double) * std::max(1, c->esmax), c->st));
  HIPCHK(c, hipMemsetAsync(c->d_slnuv.p, 0, sizeof(double) * (c->numslv + 1), c->st));
  // avslf, engnorm, engsmpl (voffset persists across divisions)
  HIPCHK(c, hipMemsetAsync(c->d_scal.p, 0, sizeof(double) * 3, c->st));
  if (c->kp.intm) {
    // edintm, uvintm and norm_intm; soln_intm and soln_total run over all
    // divisions (extensions/Intermediate/engproc.F90:348-352)
    HIPCHK(c, hipMemsetAsync(c->d_edintm.p, 0, sizeof(double) * c->ermax, c->st));
    HIPCHK(c, hipMemsetAsync(c->d_uvintm.p, 0, sizeof(double) * (MAXSPEC + 1), c->st));
    HIPCHK(c, hipMemsetAsync(c->d_intm_sc.p, 0, sizeof(double), c->st));
  }
  return sync(c);
}

// aggregate centre of the frame's host molecules into d_aggcen
static int aggregate_center(erm_ctx *c) {
  TRY(ensure(c, c->d_aggcen, 3 * sizeof(double)));
  hipLaunchKernelGGL(k_aggcen, dim3(1), dim3(1), 0, c->st, c->nmol, dp<int>(c->d_mol_begin), dp<int>(c->d_hostmol),
                     dp<double>(c->d_sitemass), dp<double>(c->d_pos), dp<double>(c->d_aggcen));
  return launch_check(c, "k_aggcen");
}

// reffit (insertion.F90:598-622) for the frame in d_pos: the reference
// host structure fitted onto the frame's host atoms carries the reference
// solute along (the best-fit reference solute), and the test particle's
// body is fitted onto that; both go to d_refsite.  The GAUSS constants are
// set at the first frame (insertion.F90:240-246).
static int refstr_frame(erm_ctx *c) {
  const int nh = (int)c->ref_atoms.size(), ns = c->ns;
  TRY(ensure(c, c->d_refgather, sizeof(double) * 3 * nh));
  hipLaunchKernelGGL(k_gather_atoms, dim3((nh + 255) / 256), dim3(256), 0, c->st, nh, dp<int>(c->d_refatoms),
                     dp<double>(c->d_pos), dp<double>(c->d_refgather));
  TRY(launch_check(c, "k_gather_atoms"));
  std::vector<double> hc(3 * nh), site(6 * ns);
  HIPCHK(c, hipMemcpyAsync(hc.data(), c->d_refgather.p, sizeof(double) * 3 * nh, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  double *fitted = site.data(), *best = site.data() + 3 * ns;
  fit_a_rotate_b(nh, hc.data(), c->ref_host.data(), c->ref_hw.data(), ns, c->ref_slt.data(), best);
  fit_a_rotate_b(ns, best, c->bfcoord.data(), c->ref_sw.data(), ns, c->bfcoord.data(), fitted);
  TRY(upload(c, c->d_refsite, site.data(), site.size()));
  Params &k = c->kp;
  if (k.ins_pos == ERM_INSPOS_GAUSS && !c->gauss_init) {
    for (int d = 0; d < 3; ++d) {
      const double len = std::sqrt(c->cell[3 * d] * c->cell[3 * d] + c->cell[3 * d + 1] * c->cell[3 * d + 1] +
                                   c->cell[3 * d + 2] * c->cell[3 * d + 2]);
      k.gl[d] = (len / 2.0) / k.upreg;
    }
    k.gz0 = 8.0 * k.gl[0] * k.gl[1] * k.gl[2];
    const double sp = std::sqrt(PI_REF);
    k.gz1 = (sp * sp * sp) * std::erf(k.gl[0]) * std::erf(k.gl[1]) * std::erf(k.gl[2]);
    c->gauss_init = true;
  }
  return ERM_OK;
}

// per-frame work after the coordinates are in d_pos
static int frame_finish(erm_ctx *c) {
  if (c->kp.ins_org == ERM_INSORG_AGGCEN) TRY(aggregate_center(c));
  if (c->kp.ins_org == ERM_INSORG_REFSTR) TRY(refstr_frame(c));
  int s = c->fp32 ? frame_prep<float>(c) : frame_prep<double>(c);
  if (s) return s;
  c->have_frame = true;
  return ERM_OK;
}

static int set_frame_impl(erm_ctx *c, const double *pos, bool device_ptr, const double *cell, double w) {
  if (!c || !pos) return ERM_E_ARG;
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  TRY(set_frame_common(c, cell, w));
  HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos, sizeof(double) * 3 * c->natom,
                           device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->st));
  return frame_finish(c);
}

int32_t erm_set_insertion(erm_ctx *c, int32_t insorigin, int32_t insposition, int32_t insorient, double lwreg,
                          double upreg, int32_t nhost, const int32_t *hostspec) {
  TRY(need(c, false, false));
  if (c->P.slttype == 1) return fail(c, ERM_E_ARG, "insertion modes apply to a test particle (slttype 2 or 3)");
  if (insposition == ERM_INSPOS_NOCHANGE)
    return fail(c, ERM_E_UNSUPPORTED, "fixed positions: pass the coordinates to erm_engconst");
  // pairing of origin and position (setconf.F90:399-428)
  const bool agg = (insposition == ERM_INSPOS_SPHERE || insposition == ERM_INSPOS_SLAB_GENERIC ||
                    insposition == ERM_INSPOS_SLAB_SYMMETRIC);
  const bool refm = (insposition == ERM_INSPOS_RMSD || insposition == ERM_INSPOS_GAUSS);
  if (insposition < 0 || insposition > ERM_INSPOS_GAUSS || insorigin < 0 || insorigin > ERM_INSORG_REFSTR)
    return fail(c, ERM_E_ARG, "insorigin %d / insposition %d", insorigin, insposition);
  if (insposition == ERM_INSPOS_RANDOM ? insorigin != ERM_INSORG_ORIGIN
      : agg ? insorigin != ERM_INSORG_AGGCEN : insorigin != ERM_INSORG_REFSTR)
    return fail(c, ERM_E_ARG, "insorigin %d does not go with insposition %d (set_ins)", insorigin, insposition);
  if (refm && (lwreg < 0.0 || upreg < 0.0 || lwreg > upreg))
    return fail(c, ERM_E_ARG, "the reference-structure modes need 0 <= lwreg <= upreg (set_reg)");
  if (refm && !c->have_ref)
    return fail(c, ERM_E_ARG, "the reference-structure modes need erm_set_refstructure first");
  if (refm && c->P.slttype != 2)
    return fail(c, ERM_E_UNSUPPORTED, "reference-structure insertion of a flexible solute: place it on the host");
  if (insorient != ERM_INSROT_RANDOM && insorient != ERM_INSROT_NOCHANGE)
    return fail(c, ERM_E_ARG, "insorient %d", insorient);
  // region bounds (set_reg)
  if (insposition == ERM_INSPOS_SPHERE && (lwreg < 0.0 || upreg < 0.0 || lwreg > upreg))
    return fail(c, ERM_E_ARG, "sphere needs 0 <= lwreg <= upreg (set_reg)");
  if ((insposition == ERM_INSPOS_SLAB_GENERIC || insposition == ERM_INSPOS_SLAB_SYMMETRIC) && lwreg > upreg)
    return fail(c, ERM_E_ARG, "slab needs lwreg <= upreg (set_reg)");
  if (insposition == ERM_INSPOS_SLAB_SYMMETRIC && (lwreg < 0.0 || upreg < 0.0))
    return fail(c, ERM_E_ARG, "symmetric slab needs lwreg, upreg >= 0 (set_reg)");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (agg) {
    // host molecules: every molecule of the listed species (com_aggregate)
    if (nhost <= 0 || !hostspec) return fail(c, ERM_E_ARG, "the aggregate centre needs host species");
    c->hostmol.assign(c->nmol, 0);
    double mtot = 0.0;
    for (int m = 0; m < c->nmol; ++m) {
      if (c->sluvid[m] > 1) continue;
      for (int h = 0; h < nhost; ++h)
        if (c->uvspec[m] == hostspec[h] && hostspec[h] > 0) c->hostmol[m] = 1;
      if (c->hostmol[m])
        for (int a = c->mol_begin[m]; a < c->mol_begin[m + 1]; ++a) mtot += c->sitemass[a];
    }
    if (!(mtot > 0.0)) return fail(c, ERM_E_ARG, "host species without mass (bst_zrw)");
    TRY(upload(c, c->d_hostmol, c->hostmol.data(), c->nmol));
  }
  Params &k = c->kp;
  k.ins_org = insorigin;
  k.ins_pos = insposition;
  k.ins_rot = insorient;
  k.lwreg = lwreg;
  k.upreg = upreg;
  c->gauss_init = false;
  if (agg && c->have_frame) TRY(aggregate_center(c));
  if (refm && c->have_frame) TRY(refstr_frame(c));
  return sync(c);
}

int32_t erm_set_refstructure(erm_ctx *c, int32_t nhost, const int32_t *host_atoms, const double *host_crd,
                             const double *host_weight, const double *slt_crd, const double *slt_weight) {
  TRY(need(c, false, false));
  if (c->P.slttype == 1) return fail(c, ERM_E_ARG, "a reference structure applies to a test particle");
  if (nhost < 1 || !host_atoms || !host_crd || !host_weight || !slt_crd || !slt_weight)
    return fail(c, ERM_E_ARG, "reference structure: host atoms, coordinates and weights are required");
  const int ns = c->ns;
  double hw = 0.0, sw = 0.0;
  for (int i = 0; i < nhost; ++i) {
    if (host_atoms[i] < 0 || host_atoms[i] >= c->natom) return fail(c, ERM_E_ARG, "host atom %d out of range", i);
    hw += host_weight[i];
  }
  for (int i = 0; i < ns; ++i) sw += slt_weight[i];
  if (!(hw > 0.0) || !(sw > 0.0)) return fail(c, ERM_E_ARG, "reference weights sum to zero (bst_zrw)");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  c->ref_atoms.assign(host_atoms, host_atoms + nhost);
  c->ref_host.assign(host_crd, host_crd + 3 * nhost);
  c->ref_hw.assign(host_weight, host_weight + nhost);
  c->ref_slt.assign(slt_crd, slt_crd + 3 * ns);
  c->ref_sw.assign(slt_weight, slt_weight + ns);
  TRY(upload(c, c->d_refatoms, c->ref_atoms.data(), nhost));
  TRY(upload(c, c->d_refw, c->ref_sw.data(), ns));
  c->have_ref = true;
  if (c->have_frame && c->kp.ins_org == ERM_INSORG_REFSTR) TRY(refstr_frame(c));
  return sync(c);
}

int32_t erm_set_frame(erm_ctx *c, const double *sitepos, const double *cell, double w) {
  return set_frame_impl(c, sitepos, false, cell, w);
}

int32_t erm_set_frame_device(erm_ctx *c, const double *d_sitepos, const double *cell, double w) {
  return set_frame_impl(c, d_sitepos, true, cell, w);
}

int32_t erm_get_uv_energy(erm_ctx *c, int32_t solute, int32_t n, const double *solute_pos, double *uvengy) {
  TRY(need(c, true, false));
  if (!uvengy) return fail(c, ERM_E_ARG, "uvengy is NULL");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  std::vector<int> sl;
  if (c->P.slttype == 1) {
    if (solute_pos) return fail(c, ERM_E_UNSUPPORTED, "soln mode takes solute coordinates from the frame");
    sl.assign(std::max(n, 0), solute);
  } else if (solute != c->slt_template) {
    return fail(c, ERM_E_ARG, "molecule %d is not the test particle", solute);
  }
  TRY(check_solutes(c, n, sl.empty() ? nullptr : sl.data(), solute_pos));
  Batch bt{n, 0, solute_pos, sl.empty() ? nullptr : sl.data(), nullptr, false, 0, 0};
  return process(c, bt, DO_RECIP | DO_REAL | DO_RESIDUAL | DO_SELF, uvengy);
}

static int piece(erm_ctx *c, int32_t solute, const double *pos, int flags, double *out, double *self_out,
                 bool self_from_u0) {
  TRY(need(c, true, false));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  int sl = solute;
  if (c->P.slttype == 1) {
    if (pos) return fail(c, ERM_E_UNSUPPORTED, "soln mode takes solute coordinates from the frame");
    TRY(check_solutes(c, 1, &sl, nullptr));
  } else {
    if (solute != c->slt_template) return fail(c, ERM_E_ARG, "molecule %d is not the test particle", solute);
    TRY(check_solutes(c, 1, nullptr, pos));
  }
  Batch bt{1, 0, c->P.slttype == 1 ? nullptr : pos, c->P.slttype == 1 ? &sl : nullptr, nullptr, false, 0, 0};
  std::vector<double> uv(c->slvmax + 1);
  TRY(process(c, bt, flags, uv.data()));
  if (out) memcpy(out, uv.data() + 1, sizeof(double) * c->slvmax);
  if (self_out && self_from_u0) *self_out = uv[0];
  return ERM_OK;
}

int32_t erm_realcal_proc(erm_ctx *c, int32_t solute, const double *pos, double *ereal) {
  return piece(c, solute, pos, DO_REAL, ereal, nullptr, false);
}

int32_t erm_recpcal(erm_ctx *c, int32_t solute, const double *pos, double *erecp, double *self_recp) {
  TRY(piece(c, solute, pos, DO_RECIP, erecp, nullptr, false));
  if (self_recp) {
    // sum of the column partials of the single configuration
    const int vb = c->fp32 ? (int)sizeof(float2) : (int)sizeof(double2);
    const bool regp = c->xslab_Q > 0 && (c->P.pme_path == 0 || c->P.pme_path == 3);
    const int nzg = zcol_groups(c, regp, vb);
    std::vector<double> part(nzg);
    HIPCHK(c, hipMemcpy(part.data(), c->d_selfp.p, sizeof(double) * nzg, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double v : part) s += v;
    *self_recp = s;
  }
  return ERM_OK;
}

int32_t erm_realcal_self(erm_ctx *c, int32_t solute, const double *pos, double *self_real) {
  TRY(need(c, true, false));
  if (!self_real) return fail(c, ERM_E_ARG, "self_real is NULL");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  const int ns = c->ns;
  TRY(ensure(c, c->d_spos, sizeof(double) * ns * 3));
  TRY(ensure(c, c->d_selfreal, sizeof(double)));
  if (c->P.slttype == 1) {
    int sl = solute;
    TRY(check_solutes(c, 1, &sl, nullptr));
    HIPCHK(c, hipMemcpyAsync(c->d_spos.p, dp<double>(c->d_pos) + (size_t)3 * c->mol_begin[sl],
                             sizeof(double) * ns * 3, hipMemcpyDeviceToDevice, c->st));
  } else {
    if (!pos) return fail(c, ERM_E_ARG, "refs mode needs test-particle coordinates");
    HIPCHK(c, hipMemcpyAsync(c->d_spos.p, pos, sizeof(double) * ns * 3, hipMemcpyHostToDevice, c->st));
  }
  hipLaunchKernelGGL(k_selfreal, dim3(1), dim3(NTHR), 0, c->st, c->kp, dp<double>(c->d_spos), 0,
                     dp<double>(c->d_qslt), dp<double>(c->d_selfreal));
  TRY(launch_check(c, "k_selfreal"));
  HIPCHK(c, hipMemcpyAsync(self_real, c->d_selfreal.p, sizeof(double), hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

int32_t erm_engconst(erm_ctx *c, int32_t n, const int32_t *solutes, const double *solute_pos, const double *weights,
                     double *uvengy) {
  TRY(need(c, true, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  TRY(check_solutes(c, n, solutes, solute_pos));
  if (c->P.slttype == 1 && solute_pos) return fail(c, ERM_E_UNSUPPORTED, "soln mode takes coordinates from the frame");
  Batch bt{n, 0, c->P.slttype == 1 ? nullptr : solute_pos, c->P.slttype == 1 ? solutes : nullptr, weights, false, 0, 0};
  return process(c, bt, DO_RECIP | DO_REAL | DO_RESIDUAL | DO_SELF | DO_HIST, uvengy);
}

int32_t erm_insert_random(erm_ctx *c, int32_t n, uint64_t seed, uint64_t first_index, double *uvengy) {
  TRY(need(c, true, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (c->P.slttype != 2) return fail(c, ERM_E_ARG, "random insertion needs a rigid test particle (slttype 2)");
  if (c->bfcoord.empty()) return fail(c, ERM_E_ARG, "random insertion needs bfcoord");
  if (n < 0) return fail(c, ERM_E_ARG, "negative count");
  Batch bt{n, 0, nullptr, nullptr, nullptr, true, seed, first_index};
  const int flags = DO_RECIP | DO_REAL | DO_RESIDUAL | DO_SELF | DO_HIST;
  if (c->use_graph && !uvengy && n > 0) return process_graph(c, bt, flags);
  return process(c, bt, flags, uvengy);
}

int32_t erm_insert_flexible(erm_ctx *c, int32_t n, const double *confs, const double *weights, uint64_t seed,
                            uint64_t first_index, double *uvengy) {
  TRY(need(c, true, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (c->P.slttype != 3) return fail(c, ERM_E_ARG, "flexible insertion needs slttype 3");
  if (n < 0) return fail(c, ERM_E_ARG, "negative count");
  if (n > 0 && !confs) return fail(c, ERM_E_ARG, "no solute configurations");
  if (c->kp.ins_org == ERM_INSORG_REFSTR)
    return fail(c, ERM_E_UNSUPPORTED, "reference-structure insertion of a flexible solute: place it on the host");
  Batch bt{n, 0, nullptr, nullptr, weights, true, seed, first_index};
  bt.flex = confs;
  const int flags = DO_RECIP | DO_REAL | DO_RESIDUAL | DO_SELF | DO_HIST;
  return process(c, bt, flags, uvengy);
}

int32_t erm_insert_random_multi(erm_ctx *const *ctxs, int32_t nctx, const int32_t *n, const uint64_t *seed,
                                const uint64_t *first_index) {
  if (!ctxs || !n || !seed || !first_index || nctx < 0) return ERM_E_ARG;
  const int flags = DO_RECIP | DO_REAL | DO_RESIDUAL | DO_SELF | DO_HIST;
  for (int i = 0; i < nctx; ++i) {
    erm_ctx *c = ctxs[i];
    TRY(need(c, true, true));
    if (c->P.slttype != 2 || c->bfcoord.empty())
      return fail(c, ERM_E_ARG, "random insertion needs a rigid test particle with bfcoord");
    if (n[i] < 0) return fail(c, ERM_E_ARG, "negative count");
    if (c->P.device != ctxs[0]->P.device) return fail(c, ERM_E_ARG, "contexts on different devices");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return fail(c, ERM_E_ARG, "a context appears twice");
  }
  if (nctx == 0) return ERM_OK;
  if (hipSetDevice(ctxs[0]->P.device) != hipSuccess) return fail(ctxs[0], ERM_E_DEVICE, "hipSetDevice");
  // enqueue every context's batches on its own stream, then wait for each
  int first_err = ERM_OK;
  std::vector<int> nbs(nctx, 0), ok(nctx, 0);
  for (int i = 0; i < nctx; ++i) {
    erm_ctx *c = ctxs[i];
    Batch bt{n[i], 0, nullptr, nullptr, nullptr, true, seed[i], first_index[i]};
    for (int k = 0; k < NSTAGE_OUT; ++k) c->stage_ms[k] = 0;
    nbs[i] = (bt.n + c->Bmax - 1) / c->Bmax;
    int st = batch_events(c, nbs[i]);
    if (!st) st = begin_call(c, bt);
    if (!st) st = enqueue_batches(c, bt, flags, nullptr);
    c->cur_ev = nullptr;
    if (st && !first_err) first_err = st;
    ok[i] = (st == ERM_OK);
  }
  for (int i = 0; i < nctx; ++i) {
    erm_ctx *c = ctxs[i];
    int st = ok[i] ? check_status(c) : sync(c);
    if (ok[i] && !st) st = stage_times(c, nbs[i], true);
    if (st && !first_err) first_err = st;
  }
  return first_err;
}

// ---- several test-particle species on one solvent (erm_group_*) ---------
int32_t erm_group_create(erm_ctx *const *ctxs, int32_t nctx, erm_group **out) {
  if (!ctxs || nctx < 1 || nctx > GROUP_MAXRUN || !out) return ERM_E_ARG;
  *out = nullptr;
  for (int i = 0; i < nctx; ++i) {
    if (!ctxs[i]) return ERM_E_ARG;
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(ctxs[i], ERM_E_ARG, "a context appears twice in the group");
  }
  erm_group *g = new erm_group();
  g->ctx.assign(ctxs, ctxs + nctx);
  const int st = group_compatible(g);
  if (st) {
    ctxs[0]->err = g->err;
    delete g;
    return st;
  }
  *out = g;
  return ERM_OK;
}

void erm_group_destroy(erm_group *g) {
  if (!g) return;
  if (hipSetDevice(g->ctx[0]->P.device) == hipSuccess) {
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    g->d_runs.release();
    g->d_counter.release();
    if (g->h_runs) (void)hipHostFree(g->h_runs);
  }
  delete g;
}

const char *erm_group_last_error(const erm_group *g) { return g ? g->err.c_str() : "null group"; }

static int group_set_frame(erm_group *g, const double *pos, bool device_ptr, const double *cell, double w) {
  if (!g || !pos) return ERM_E_ARG;
  erm_ctx *c = g->ctx[0];
  int st = set_frame_impl(c, pos, device_ptr, cell, w);
  if (st) return gfail(g, st, "member 0: %s", c->err.c_str());
  for (size_t i = 1; i < g->ctx.size(); ++i) {
    erm_ctx *m = g->ctx[i];
    st = set_frame_common(m, cell, w);
    if (st) return gfail(g, st, "member %zu: %s", i, m->err.c_str());
    // the solvent's device data of this frame live in member 0 only
    m->have_frame = false;
    m->jof_stale = true;
    // the member's cell upload (its own stream) precedes the group's launches
    if (hipStreamSynchronize(m->st) != hipSuccess) {
      (void)hipGetLastError();
      return gfail(g, ERM_E_DEVICE, "member %zu: stream synchronisation", i);
    }
  }
  g->frame_serial = c->frame_serial;
  return ERM_OK;
}

int32_t erm_group_set_frame(erm_group *g, const double *sitepos, const double *cell, double w) {
  return group_set_frame(g, sitepos, false, cell, w);
}

int32_t erm_group_set_frame_device(erm_group *g, const double *d_sitepos, const double *cell, double w) {
  return group_set_frame(g, d_sitepos, true, cell, w);
}

int32_t erm_group_set_tuning(erm_group *g, int32_t max_batch, int32_t persistent) {
  if (!g) return ERM_E_ARG;
  if (max_batch < 0 || max_batch > 65536 || persistent < -1 || persistent > 3)
    return gfail(g, ERM_E_ARG, "group tuning out of range");
  if (max_batch > 0) g->Bmax = max_batch;
  if (persistent >= 0) g->persistent = persistent;
  return ERM_OK;
}

int32_t erm_group_stage_times(const erm_group *g, double *ms10) {
  if (!g || !ms10) return ERM_E_ARG;
  for (int i = 0; i < NSTAGE_OUT; ++i) ms10[i] = g->stage_ms[i];
  return ERM_OK;
}

int32_t erm_group_insert_random(erm_group *g, const int32_t *n, const uint64_t *seed, const uint64_t *first_index) {
  if (!g || !n || !seed || !first_index) return ERM_E_ARG;
  erm_ctx *c = g->ctx[0];
  const int nm = (int)g->ctx.size();
  if (hipSetDevice(c->P.device) != hipSuccess) return gfail(g, ERM_E_DEVICE, "hipSetDevice");
  if (!c->have_frame || g->frame_serial != c->frame_serial)
    return gfail(g, ERM_E_ARG, "erm_group_set_frame has not been called for the current frame");
  for (int i = 0; i < nm; ++i)
    if (n[i] < 0) return gfail(g, ERM_E_ARG, "negative count for member %d", i);
  const bool ok_path = (c->fp32 ? reg_path_ok<float>(c) : reg_path_ok<double>(c)) && zcol_fft(c, true) &&
                       c->xslab_Q > 0 && (c->P.pme_path == 0 || c->P.pme_path == 3);
  if (!ok_path) return gfail(g, ERM_E_UNSUPPORTED, "groups need the register PME path with spline order 4 or 6");
  for (int i = 0; i < NSTAGE_OUT; ++i) g->stage_ms[i] = 0;
  int64_t N = 0;
  for (int i = 0; i < nm; ++i) N += n[i];
  if (N == 0) return ERM_OK;
  StreamSwap swap(g);
  const int B = std::max(1, g->Bmax);
  const int nb = (int)((N + B - 1) / B);
  // members by decreasing size: the heavy work items are handed out first
  std::vector<int> order(nm);
  for (int i = 0; i < nm; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return g->ctx[a]->ns > g->ctx[b]->ns; });
  int st = ensure(c, c->d_status, 64);
  if (!st) {
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(32), 0, c->st, dp<int>(c->d_status), (size_t)8, 0);
    st = launch_check(c, "k_fill");
  }
  for (int i = 0; i < nm && !st; ++i) st = upload_args(g->ctx[i], seed[i], first_index[i]);
  if (!st) st = ensure(c, c->d_wsol, sizeof(double) * B);
  if (!st) {
    hipLaunchKernelGGL(k_fill<double>, dim3((B + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, dp<double>(c->d_wsol),
                       (size_t)B, 1.0);
    st = launch_check(c, "k_fill");
    c->wsol_ones = 0;  // the ones may be overwritten by a weighted call of member 0
  }
  if (!st) st = batch_events(c, nb);
  const int need_runs = nb * GROUP_MAXRUN;
  if (!st && g->h_runs_cap < need_runs) {
    if (g->h_runs) (void)hipHostFree(g->h_runs);
    g->h_runs = nullptr;
    g->h_runs_cap = 0;
    if (hipHostMalloc((void **)&g->h_runs, sizeof(GRun) * need_runs, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      st = fail(c, ERM_E_DEVICE, "hipHostMalloc of the run tables");
    } else {
      g->h_runs_cap = need_runs;
    }
  }
  if (!st) st = ensure(c, g->d_runs, sizeof(GRun) * need_runs);
  const Params &k = c->kp;
  const int vb = c->fp32 ? (int)sizeof(float2) : (int)sizeof(double2);
  const int nzg = zcol_groups(c, true, vb);
  // the run tables of every batch (host), then the batches' launches
  std::vector<int> nrun_b(nb), ncfg_b(nb), nz_b(nb), np_b(nb);
  std::vector<uint64_t> off_b((size_t)nb * GROUP_MAXRUN, 0);
  for (int bi = 0; bi < nb && !st; ++bi) {
    GRun *R = g->h_runs + (size_t)bi * GROUP_MAXRUN;
    uint64_t *offset = off_b.data() + (size_t)bi * GROUP_MAXRUN;
    int nrun = 0, cfg = 0, site = 0, zi = 0, pi = 0;
    for (int oi = 0; oi < nm; ++oi) {
      const int t = order[oi];
      erm_ctx *m = g->ctx[t];
      const int64_t a = (int64_t)n[t] * bi / nb, e = (int64_t)n[t] * (bi + 1) / nb;
      const int cnt = (int)(e - a);
      if (cnt <= 0) continue;
      const double Rg = k.rcmax + m->slt_radius + c->mol_radius + 0.5 * (k.hcell[0] + k.hcell[1] + k.hcell[2]) / 3.0;
      const double frac = std::min(1.0, 4.18879 * Rg * Rg * Rg / c->volume);
      const int nchunk = (int)std::ceil(frac * c->slvmax / 64.0);
      const int nsplit = pair_splits(nchunk, NTHR / 64);
      GRun &r = R[nrun];
      r.cfg0 = cfg;
      r.ncfg = cnt;
      r.ns = m->ns;
      r.site0 = site;
      r.zitem0 = zi;
      r.pitem0 = pi;
      r.nsplit = nsplit;
      r.member = t;
      r.q_t = m->d_qslt_t.p;
      r.lj_t = m->d_sltlj.p;
      r.q64 = dp<double>(m->d_qslt);
      r.lj64 = dp<double>(m->d_sltlj64);
      offset[nrun] = (uint64_t)a;
      cfg += cnt;
      site += cnt * m->ns;
      zi += cnt * nzg;
      pi += cnt * nsplit;
      ++nrun;
    }
    nrun_b[bi] = nrun;
    ncfg_b[bi] = cfg;
    nz_b[bi] = zi;
    np_b[bi] = pi;
  }
  auto enqueue = [&]() -> int {
    for (int bi = 0; bi < nb; ++bi) {
      GRun *R = g->h_runs + (size_t)bi * GROUP_MAXRUN;
      GRun *d_runs = dp<GRun>(g->d_runs) + (size_t)bi * GROUP_MAXRUN;
      if (hipMemcpyAsync(d_runs, R, sizeof(GRun) * nrun_b[bi], hipMemcpyHostToDevice, c->st) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, ERM_E_DEVICE, "run table upload");
      }
      c->cur_ev = c->capturing ? nullptr : &c->bev[NSTAGE_EV * bi];
      const uint64_t *offset = off_b.data() + (size_t)bi * GROUP_MAXRUN;
      TRY(c->fp32 ? group_batch<float>(g, R, nrun_b[bi], offset, ncfg_b[bi], d_runs, nz_b[bi], np_b[bi])
                  : group_batch<double>(g, R, nrun_b[bi], offset, ncfg_b[bi], d_runs, nz_b[bi], np_b[bi]));
    }
    return ERM_OK;
  };
  // everything the captured launches take by value or by address
  auto key = [&]() {
    std::string kk(reinterpret_cast<const char *>(g->h_runs), sizeof(GRun) * (size_t)nb * GROUP_MAXRUN);
    kk.append(reinterpret_cast<const char *>(off_b.data()), sizeof(uint64_t) * off_b.size());
    char buf[160];
    snprintf(buf, sizeof buf, "|%d|%d|%d|%d|%d|%d|%d|%d|%d", nb, g->persistent, c->slab_S, c->slab_R, c->xslab_Q,
             c->xslab_T, c->zcol_T, (int)c->xslab_yf, c->wsol_ones);
    kk += buf;
    for (erm_ctx *m : g->ctx) {
      kk.append(reinterpret_cast<const char *>(&m->kp), sizeof(Params));
      snprintf(buf, sizeof buf, "|%llu|%d|%d", (unsigned long long)m->alloc_gen, (int)m->have_usreal, m->ns);
      kk += buf;
    }
    return kk;
  };
  bool replayed = false;
  if (!st && g->use_graph && g->gexec && key() == g->gkey) {
    HIPCHK(c, hipEventRecord(c->ev[0], c->st));
    if (hipGraphLaunch(g->gexec, c->st) != hipSuccess) {
      (void)hipGetLastError();
      st = fail(c, ERM_E_DEVICE, "group graph launch");
    }
    HIPCHK(c, hipEventRecord(c->ev[7], c->st));
    replayed = true;
  } else if (!st) {
    st = enqueue();
    if (!st && g->use_graph) {
      // capture the same launches (the eager results above stand)
      if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
      g->gexec = nullptr;
      g->gkey = key();
      std::vector<uint64_t> gen;
      for (erm_ctx *m : g->ctx) gen.push_back(m->alloc_gen);
      hipGraph_t gr = nullptr;
      if (hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed) == hipSuccess) {
        c->capturing = true;
        const int cs = enqueue();
        c->capturing = false;
        const hipError_t e = hipStreamEndCapture(c->st, &gr);
        bool ok = (cs == ERM_OK && e == hipSuccess && gr);
        for (size_t i = 0; i < g->ctx.size() && ok; ++i) ok = (g->ctx[i]->alloc_gen == gen[i]);
        if (ok) ok = (hipGraphInstantiate(&g->gexec, gr, nullptr, nullptr, 0) == hipSuccess);
        if (gr) (void)hipGraphDestroy(gr);
        if (!ok) {
          g->gexec = nullptr;
          g->use_graph = false;  // later calls run eagerly
        }
      } else {
        g->use_graph = false;
      }
      (void)hipGetLastError();
    }
  }
  c->cur_ev = nullptr;
  if (!st) st = check_status(c);
  if (replayed) {
    for (int i = 0; i < NSTAGE_OUT; ++i) g->stage_ms[i] = 0;
    float ms;
    if (!st && hipEventElapsedTime(&ms, c->ev[0], c->ev[7]) == hipSuccess) g->stage_ms[NSTAGE_OUT - 1] = ms;
    if (st) return gfail(g, st, "%s", c->err.c_str());
    return ERM_OK;
  }
  if (!st) st = stage_times(c, nb, true);
  if (st) return gfail(g, st, "%s", c->err.c_str());
  for (int i = 0; i < NSTAGE_OUT; ++i) g->stage_ms[i] = c->stage_ms[i];
  return ERM_OK;
}

int32_t erm_group_set_graph(erm_group *g, int32_t on) {
  if (!g) return ERM_E_ARG;
  if (on < 0 || on > 1) return gfail(g, ERM_E_ARG, "graph flag must be 0 or 1");
  g->use_graph = (on == 1);
  if (!g->use_graph && g->gexec) {
    (void)hipGraphExecDestroy(g->gexec);
    g->gexec = nullptr;
  }
  return ERM_OK;
}

int32_t erm_update_histogram(erm_ctx *c, int32_t n, const int32_t *solutes, const double *weights,
                             const double *uvengy) {
  TRY(need(c, true, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (!uvengy || n < 0) return fail(c, ERM_E_ARG, "uvengy is NULL");
  if (c->P.slttype == 1 && !solutes) return fail(c, ERM_E_ARG, "soln mode needs solute indices");
  const int S = c->slvmax;
  TRY(ensure(c, c->d_status, 64));
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(32), 0, c->st, dp<int>(c->d_status), (size_t)8, 0);
  TRY(launch_check(c, "k_fill"));
  c->cur_ev = nullptr;
  for (int s0 = 0; s0 < n; s0 += c->Bmax) {
    int nb = std::min(c->Bmax, n - s0);
    // energies into the sorted order
    TRY(ensure_jof(c));
    std::vector<double> uvs((size_t)nb * (S + 1));
    for (int b = 0; b < nb; ++b) {
      const double *src = uvengy + (size_t)(s0 + b) * (S + 1);
      double *dst = uvs.data() + (size_t)b * (S + 1);
      dst[0] = src[0];
      for (int kk = 0; kk < S; ++kk) dst[1 + c->jof_k[kk]] = src[1 + kk];
    }
    TRY(ensure(c, c->d_uv, sizeof(double) * uvs.size()));
    TRY(upload(c, c->d_uv, uvs.data(), uvs.size()));
    std::vector<int> sltj(nb, -1);
    if (solutes)
      for (int b = 0; b < nb; ++b) {
        int kk = c->kof[solutes[s0 + b]];
        sltj[b] = kk >= 0 ? c->jof_k[kk] : -1;
      }
    TRY(ensure(c, c->d_sltj, sizeof(int) * nb));
    TRY(upload(c, c->d_sltj, sltj.data(), nb));
    std::vector<double> ws(nb, 1.0);
    if (weights)
      for (int b = 0; b < nb; ++b) ws[b] = weights[s0 + b];
    TRY(ensure(c, c->d_wsol, sizeof(double) * nb));
    TRY(upload(c, c->d_wsol, ws.data(), nb));
    c->wsol_ones = 0;
    TRY(run_histogram(c, nb, s0));
  }
  return check_status(c);
}

int32_t erm_update_histogram_static(erm_ctx *c, int32_t n, const int32_t *solutes, const double *weights,
                                    const double *uvengy) {
  TRY(need(c, true, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (n < 0 || (n > 0 && !uvengy)) return fail(c, ERM_E_ARG, "uvengy is NULL or n negative");
  if (c->P.slttype == 1 && n > 0 && !solutes) return fail(c, ERM_E_ARG, "soln mode needs solute indices");
  if (solutes)
    for (int b = 0; b < n; ++b)
      if (solutes[b] < 0 || solutes[b] >= c->nmol) return fail(c, ERM_E_ARG, "solute index out of range");
  unsigned char *base = scratch_of(c->P.device);
  if (!base) return fail(c, ERM_E_DEVICE, "static scratch not resolved on device %d", c->P.device);
  // layout: status words, then per configuration the energies as given and
  // sorted, its histogram, per-species sums, weights (in, out, intermediate)
  // and the solute's indices, each region 256-byte aligned
  const int S = c->slvmax, E = c->ermax;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t per = 2 * sizeof(double) * ((size_t)S + 1) + sizeof(double) * ((size_t)E + MAXSPEC + 3) +
                     2 * sizeof(int);
  const size_t fixed = 256 + 9 * 256;  // status + alignment of the nine regions
  const size_t cap = (SCRATCH_BYTES > fixed) ? (SCRATCH_BYTES - fixed) / per : 0;
  if ((size_t)n > cap)
    return fail(c, ERM_E_ARG,
                "erm_update_histogram_static: %d configurations exceed the %zu that its %zu-byte scratch holds "
                "for %d molecules and %d bins",
                n, cap, SCRATCH_BYTES, S, E);
  if (n == 0) return ERM_OK;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char *p = base + off;
    off += al(bytes);
    return p;
  };
  int *status = reinterpret_cast<int *>(carve(64));
  double *uvin = reinterpret_cast<double *>(carve(sizeof(double) * (size_t)n * (S + 1)));
  HistBufs hb;
  hb.status = status;
  hb.uv = reinterpret_cast<double *>(carve(sizeof(double) * (size_t)n * (S + 1)));
  hb.hist = reinterpret_cast<double *>(carve(sizeof(double) * (size_t)n * E));
  hb.flceng = reinterpret_cast<double *>(carve(sizeof(double) * (size_t)n * MAXSPEC));
  hb.wsol = reinterpret_cast<double *>(carve(sizeof(double) * n));
  hb.w = reinterpret_cast<double *>(carve(sizeof(double) * n));
  hb.wi = reinterpret_cast<double *>(carve(sizeof(double) * n));
  hb.sltj = reinterpret_cast<int *>(carve(sizeof(int) * n));
  int *sol = solutes ? reinterpret_cast<int *>(carve(sizeof(int) * n)) : nullptr;
  hb.mol_spec = dp<int>(c->d_mol_spec);
  if (off > SCRATCH_BYTES) return fail(c, ERM_E_ENG_BUG, "static scratch layout overflows");
  // the array keeps the previous call's contents: the status starts at zero
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(32), 0, c->st, status, (size_t)16, 0);
  TRY(launch_check(c, "k_fill"));
  HIPCHK(c, hipMemcpyAsync(uvin, uvengy, sizeof(double) * (size_t)n * (S + 1), hipMemcpyHostToDevice, c->st));
  if (sol) HIPCHK(c, hipMemcpyAsync(sol, solutes, sizeof(int) * n, hipMemcpyHostToDevice, c->st));
  if (weights) {
    HIPCHK(c, hipMemcpyAsync(hb.wsol, weights, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3((n + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, hb.wsol, (size_t)n, 1.0);
    TRY(launch_check(c, "k_fill"));
  }
  hipLaunchKernelGGL(k_sort_uv, dim3((S + 255) / 256, n), dim3(256), 0, c->st, n, S, dp<int>(c->d_mol_k),
                     dp<int>(c->d_jof), dp<int>(c->d_kof), sol, uvin, hb.uv, hb.sltj);
  TRY(launch_check(c, "k_sort_uv"));
  c->cur_ev = nullptr;
  TRY(run_histogram_at(c, n, 0, hb));
  return check_status(c, status);
}

int32_t erm_engstore(erm_ctx *c) {
  TRY(need(c, false, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  // single process: edens/engnorm, ecorr/engnorm, eself/engnorm, avslf/engnorm
  const int E = c->ermax;
  std::vector<double> sc(8), e(E);
  HIPCHK(c, hipMemcpyAsync(sc.data(), c->d_scal.p, sizeof(double) * 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(e.data(), c->d_edens.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  const double en = sc[0];
  for (double &v : e) v = v / en;
  TRY(upload(c, c->d_edens, e.data(), E));
  if (c->P.corrcal) {
    std::vector<double> m((size_t)E * E);
    HIPCHK(c, hipMemcpyAsync(m.data(), c->d_ecorr.p, sizeof(double) * m.size(), hipMemcpyDeviceToHost, c->st));
    TRY(sync(c));
    for (double &v : m) v = v / en;
    TRY(upload(c, c->d_ecorr, m.data(), m.size()));
  }
  if (c->esmax) {
    std::vector<double> s(c->esmax);
    HIPCHK(c, hipMemcpyAsync(s.data(), c->d_eself.p, sizeof(double) * s.size(), hipMemcpyDeviceToHost, c->st));
    TRY(sync(c));
    for (double &v : s) v = v / en;
    TRY(upload(c, c->d_eself, s.data(), s.size()));
  }
  sc[2] = sc[2] / en;
  TRY(upload(c, c->d_scal, sc.data(), 8));
  if (c->kp.intm) {
    // edintm / norm_intm (extensions/Intermediate/engproc.F90:656)
    std::vector<double> ed(E), isc(3);
    HIPCHK(c, hipMemcpyAsync(ed.data(), c->d_edintm.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(isc.data(), c->d_intm_sc.p, sizeof(double) * 3, hipMemcpyDeviceToHost, c->st));
    TRY(sync(c));
    for (double &v : ed) v = v / isc[0];
    TRY(upload(c, c->d_edintm, ed.data(), E));
  }
  return sync(c);
}

int32_t erm_set_intermediate(erm_ctx *c, const int32_t *zerodns) {
  TRY(need(c, false, true));
  if (c->P.slttype != 1) return fail(c, ERM_E_ARG, "the intermediate state applies to solution systems");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (!zerodns) {
    c->kp.intm = 0;
    return ERM_OK;
  }
  const int E = c->ermax;
  std::vector<int> z(zerodns, zerodns + E);
  for (int &v : z) v = v ? 1 : 0;
  TRY(upload(c, c->d_zerodns, z.data(), E));
  std::vector<double> zero(std::max(E, MAXSPEC + 1), 0.0);
  TRY(upload(c, c->d_edintm, zero.data(), E));
  TRY(upload(c, c->d_uvintm, zero.data(), MAXSPEC + 1));
  TRY(upload(c, c->d_intm_sc, zero.data(), 3));
  c->kp.intm = 1;
  return sync(c);
}

int32_t erm_fetch_intermediate(erm_ctx *c, double *edintm, double *uvintm, double *scal) {
  TRY(need(c, false, true));
  if (!c->kp.intm) return fail(c, ERM_E_ARG, "the intermediate state is off");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (edintm)
    HIPCHK(c, hipMemcpyAsync(edintm, c->d_edintm.p, sizeof(double) * c->ermax, hipMemcpyDeviceToHost, c->st));
  if (uvintm)
    HIPCHK(c, hipMemcpyAsync(uvintm, c->d_uvintm.p, sizeof(double) * (c->numslv + 1), hipMemcpyDeviceToHost, c->st));
  if (scal) HIPCHK(c, hipMemcpyAsync(scal, c->d_intm_sc.p, sizeof(double) * 3, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

int32_t erm_fetch(erm_ctx *c, double *edens, double *ecorr, double *eself, double *minuv, double *maxuv,
                  double *slnuv, double *scalars) {
  TRY(need(c, false, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  const int E = c->ermax;
  if (edens) HIPCHK(c, hipMemcpyAsync(edens, c->d_edens.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->st));
  if (ecorr) {
    if (!c->P.corrcal) return fail(c, ERM_E_ARG, "corrcal is off");
    HIPCHK(c, hipMemcpyAsync(ecorr, c->d_ecorr.p, sizeof(double) * (size_t)E * E, hipMemcpyDeviceToHost, c->st));
  }
  if (eself && c->esmax)
    HIPCHK(c, hipMemcpyAsync(eself, c->d_eself.p, sizeof(double) * c->esmax, hipMemcpyDeviceToHost, c->st));
  if (slnuv)
    HIPCHK(c, hipMemcpyAsync(slnuv, c->d_slnuv.p, sizeof(double) * (c->numslv + 1), hipMemcpyDeviceToHost, c->st));
  std::vector<unsigned long long> mm(2 * (c->numslv + 1));
  HIPCHK(c, hipMemcpyAsync(mm.data(), c->d_mnmx.p, sizeof(unsigned long long) * mm.size(), hipMemcpyDeviceToHost, c->st));
  double sc[8];
  HIPCHK(c, hipMemcpyAsync(sc, c->d_scal.p, sizeof sc, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  auto dec = [](unsigned long long u, double dflt, bool empty) {
    if (empty) return dflt;
    u = (u & 0x8000000000000000ULL) ? (u & 0x7FFFFFFFFFFFFFFFULL) : ~u;
    double d;
    memcpy(&d, &u, sizeof d);
    return d;
  };
  for (int i = 0; i <= c->numslv; ++i) {
    if (minuv) minuv[i] = dec(mm[2 * i], 1.0e50, mm[2 * i] == ~0ULL);
    if (maxuv) maxuv[i] = dec(mm[2 * i + 1], -1.0e50, mm[2 * i + 1] == 0ULL);
  }
  if (scalars) {
    scalars[0] = sc[0];
    scalars[1] = sc[1];
    scalars[2] = sc[2];
    scalars[3] = sc[3];
  }
  return ERM_OK;
}

int32_t erm_sizes(const erm_ctx *c, int32_t *slvmax, int32_t *ermax) {
  if (!c) return ERM_E_ARG;
  if (slvmax) *slvmax = c->slvmax;
  if (ermax) *ermax = c->ermax;
  return ERM_OK;
}

int32_t erm_stage_times(const erm_ctx *c, double *ms10) {
  if (!c || !ms10) return ERM_E_ARG;
  for (int i = 0; i < NSTAGE_OUT; ++i) ms10[i] = c->stage_ms[i];
  return ERM_OK;
}

int32_t erm_set_tuning(erm_ctx *c, int32_t max_batch, int32_t pme_path, int32_t use_graph, int32_t slab_planes,
                       int32_t group_planes, int32_t slab_threads, int32_t column_threads, int32_t fold_y) {
  if (!c) return ERM_E_ARG;
  if (fold_y < -1 || fold_y > 1) return fail(c, ERM_E_ARG, "fold_y must be -1, 0 or 1");
  if (fold_y >= 0) c->xslab_yf = (fold_y == 1);
  if (max_batch < 0 || max_batch > 4096 || pme_path < 0 || pme_path > 3 || use_graph < 0 || use_graph > 1 ||
      slab_planes < 0 || slab_planes > 16 || group_planes < 0 || group_planes > 16 ||
      (slab_threads != 0 && slab_threads != 256 && slab_threads != 512) ||
      (column_threads != 0 && column_threads != 256 && column_threads != 512))
    return fail(c, ERM_E_ARG, "tuning out of range");
  if (column_threads > 0) c->zcol_T = column_threads;
  if (max_batch > 0) c->Bmax = max_batch;
  if (slab_planes > 0) c->tune_S = slab_planes;
  if (group_planes > 0) c->tune_Q = group_planes;
  if (slab_threads > 0) c->xslab_T = slab_threads;
  c->P.pme_path = pme_path;
  c->use_graph = use_graph != 0;
  if (c->have_frame) TRY(c->fp32 ? choose_slab<float>(c) : choose_slab<double>(c));
  return ERM_OK;
}

int32_t erm_set_pair_path(erm_ctx *c, int32_t path) {
  if (!c) return ERM_E_ARG;
  if (path < 0 || path > 1) return fail(c, ERM_E_ARG, "pair path must be 0 (tiles) or 1 (per molecule)");
  c->pair_path = path;
  return ERM_OK;
}

int32_t erm_get_pair_path(const erm_ctx *c) { return c ? c->pair_path : -1; }

int32_t erm_get_tuning(const erm_ctx *c, int32_t *out10) {
  if (!c || !out10) return ERM_E_ARG;
  int32_t *out8 = out10, *out9 = out10;
  out10[9] = xslab_folds_y(c) ? 1 : 0;
  out9[8] = c->zcol_T;
  out8[0] = c->Bmax;
  out8[1] = c->P.pme_path;
  out8[2] = c->use_graph ? 1 : 0;
  out8[3] = c->xslab_T;
  out8[4] = c->slab_S;
  out8[5] = c->xslab_Q;
  out8[6] = c->tune_S;
  out8[7] = c->tune_Q;
  return ERM_OK;
}

int32_t erm_autotune(erm_ctx *c, int32_t n_trial, uint64_t seed, double *best_ms) {
  TRY(need(c, true, true));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (c->P.slttype != 2 || c->bfcoord.empty())
    return fail(c, ERM_E_UNSUPPORTED, "autotuning times random insertions of a rigid test particle");
  if (n_trial <= 0) return fail(c, ERM_E_ARG, "n_trial must be positive");
  const bool had_usreal = c->have_usreal;
  const int flags = DO_RECIP | DO_REAL | DO_RESIDUAL | DO_SELF;  // accumulators untouched
  std::vector<int> paths;
  if (c->xslab_Q) paths.push_back(3);
  if (c->slab_S) paths.push_back(1);
  paths.push_back(2);
  double best = 1e300;
  int bestB = c->Bmax, bestP = c->P.pme_path, bestS = c->tune_S, bestQ = c->tune_Q, bestT = c->xslab_T, bestZ = 0;
  auto timed = [&](double *ms) -> int {
    Batch bt{n_trial, 0, nullptr, nullptr, nullptr, true, seed, 0};
    TRY(process(c, bt, flags, nullptr));  // warm-up (allocations, first launches)
    HIPCHK(c, hipEventRecord(c->ev[0], c->st));
    TRY(process(c, bt, flags, nullptr));
    HIPCHK(c, hipEventRecord(c->ev[7], c->st));
    TRY(sync(c));
    float f = 0;
    HIPCHK(c, hipEventElapsedTime(&f, c->ev[0], c->ev[7]));
    *ms = f;
    return ERM_OK;
  };
  c->use_graph = false;
  auto apply = [&](int B, int pp, int Sc, int Qc, int Tc) -> int {
    c->Bmax = B;
    c->P.pme_path = pp;
    c->tune_S = Sc;
    c->tune_Q = Qc;
    c->xslab_T = Tc;
    return c->fp32 ? choose_slab<float>(c) : choose_slab<double>(c);
  };
  for (int B : {128, 256, 512, 1024}) {
    if (B > 4 * n_trial && B > 128) break;
    for (int pp : paths) {
      for (int Sc : {16, 8, 6, 5, 4, 3}) {
        for (int Qc : {16, 9, 8, 5, 4, 3}) {
          for (int Tc : {256, 512}) {
            if (pp != 3 && (Qc != 8 || Tc != 256)) continue;
            if (pp == 2 && Sc != 16) continue;
            if (apply(B, pp, Sc, Qc, Tc) != ERM_OK) {
              c->err.clear();
              continue;
            }
            double ms;
            TRY(timed(&ms));
            if (ms < best) {
              best = ms;
              bestB = B;
              bestP = pp;
              bestS = Sc;
              bestQ = Qc;
              bestT = Tc;
            }
          }
        }
      }
    }
  }
  TRY(apply(bestB, bestP, bestS, bestQ, bestT));
  if (bestP == 3) {  // z-column workgroup size for the chosen shape
    const int z0 = c->zcol_T;
    double ms;
    for (int Zc : {256, 512}) {
      c->zcol_T = Zc;
      TRY(timed(&ms));
      if (ms < best) {
        best = ms;
        bestZ = Zc;
      }
    }
    c->zcol_T = bestZ ? bestZ : z0;
    // inverse y transform inside the x slab, or its own kernel
    if (c->xslab_T == 512 && c->kp.K2 == c->kp.K1) {
      const bool y0 = c->xslab_yf;
      int besty = -1;
      for (int yf : {0, 1}) {
        c->xslab_yf = (yf == 1);
        TRY(timed(&ms));
        if (ms < best) {
          best = ms;
          besty = yf;
        }
      }
      c->xslab_yf = (besty >= 0) ? (besty == 1) : y0;
    }
  }
  // real-space kernel: compacted tiles or one thread per candidate molecule
  {
    const int p0 = c->pair_path;
    int bestp = -1;
    double ms;
    for (int pp : {0, 1}) {
      c->pair_path = pp;
      TRY(timed(&ms));
      if (ms < best) {
        best = ms;
        bestp = pp;
      }
    }
    c->pair_path = (bestp >= 0) ? bestp : p0;
  }
  // graph replay against eager launches for the chosen shape
  c->use_graph = true;
  {
    Batch bt{n_trial, 0, nullptr, nullptr, nullptr, true, seed, 0};
    TRY(process_graph(c, bt, flags));  // eager run + capture
    if (c->use_graph && c->gexec) {
      HIPCHK(c, hipEventRecord(c->ev[1], c->st));
      TRY(process_graph(c, bt, flags));
      HIPCHK(c, hipEventRecord(c->ev[2], c->st));
      TRY(sync(c));
      float f = 0;
      HIPCHK(c, hipEventElapsedTime(&f, c->ev[1], c->ev[2]));
      if (f >= best)
        c->use_graph = false;
      else
        best = f;
    }
  }
  c->have_usreal = had_usreal;  // the trial's realcal_self must not be reused
  if (best_ms) *best_ms = best;
  return ERM_OK;
}

int32_t erm_solute_gradient(erm_ctx *c, int32_t n, const int32_t *solutes, const double *solute_pos, uint64_t seed,
                            uint64_t first_index, const double *weights, int32_t cmbrule, const double *q,
                            const double *eps, const double *sigma, double *acc) {
  TRY(need(c, true, false));
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  if (!acc) return fail(c, ERM_E_ARG, "acc is NULL");
  if (n < 0) return fail(c, ERM_E_ARG, "negative count");
  if (cmbrule != 0 && cmbrule != 1) return fail(c, ERM_E_ARG, "cmbrule must be 0 or 1");
  if (c->P.wgtslf != 0 || c->P.estype != 1)
    return fail(c, ERM_E_UNSUPPORTED, "the gradient needs parameter-independent weights (wgtslf 0, estype 1)");
  if (c->P.cltype == ERM_EL_COULOMB) return fail(c, ERM_E_UNSUPPORTED, "the gradient covers PME and PPPM");
  if (c->kp.ins_pos == ERM_INSPOS_GAUSS)
    return fail(c, ERM_E_UNSUPPORTED, "the gradient with GAUSS insertion weights");
  const int ns = c->ns;
  if (ns > 1024) return fail(c, ERM_E_UNSUPPORTED, "gradient for solutes of more than 1024 sites");
  const Params &k = c->kp;
  bool regok = c->fp32 ? reg_path_ok<float>(c) : reg_path_ok<double>(c);
  if (!regok) return fail(c, ERM_E_UNSUPPORTED, "gradient needs PME grid lengths 16, 24, 32, 48, 64 (96 in FP32)");
  bool random = false;
  if (c->P.slttype == 1) {
    TRY(check_solutes(c, n, solutes, nullptr));
    if (solute_pos) return fail(c, ERM_E_UNSUPPORTED, "soln mode takes coordinates from the frame");
    if (c->sltlist.size() != 1) return fail(c, ERM_E_UNSUPPORTED, "gradient in soln mode needs a single solute");
  } else {
    if (!solute_pos) {
      if (c->P.slttype != 2 || c->bfcoord.empty())
        return fail(c, ERM_E_ARG, "random poses need a rigid test particle with bfcoord");
      random = true;
    }
  }
  // per-site parameters and the solute x solvent LJ table
  GradSetup g;
  const int a0 = c->mol_begin[c->slt_template], nt = c->ntype;
  g.q.resize(ns);
  g.eps.resize(ns);
  g.sig.resize(ns);
  for (int s = 0; s < ns; ++s) {
    int t = c->ljtype[a0 + s];
    g.q[s] = q ? q[s] : c->charge[a0 + s];
    g.eps[s] = eps ? eps[s] : c->lj_eps[(size_t)t * nt + t];
    g.sig[s] = sigma ? sigma[s] : std::sqrt(c->lj_sig2[(size_t)t * nt + t]);
    if (!(g.eps[s] >= 0.0) || !(g.sig[s] > 0.0)) return fail(c, ERM_E_ARG, "site %d: eps < 0 or sigma <= 0", s);
  }
  g.tab.resize((size_t)3 * ns * c->nvt);
  for (int s = 0; s < ns; ++s) {
    int tu = c->ljtype[a0 + s];
    for (int v = 0; v < c->nvt; ++v) {
      int tv = c->vtypes[v];
      double ev = c->lj_eps[(size_t)tv * nt + tv], sv = std::sqrt(c->lj_sig2[(size_t)tv * nt + tv]);
      double eiv = std::sqrt(g.eps[s] * ev);
      double s2 = cmbrule ? g.sig[s] * sv : ipow_h(0.5 * (g.sig[s] + sv), 2);
      if (!eps && !sigma) {
        // the topology's table must be this rule applied to its diagonals
        double te = c->lj_eps[(size_t)tv * nt + tu], ts = c->lj_sig2[(size_t)tv * nt + tu];
        if (std::fabs(te - eiv) > 1e-10 * std::fabs(eiv) + 1e-300 || std::fabs(ts - s2) > 1e-10 * s2)
          return fail(c, ERM_E_UNSUPPORTED, "LJ tables do not follow combination rule %d", cmbrule);
      }
      double* e = &g.tab[((size_t)s * c->nvt + v) * 3];
      e[0] = eiv;
      e[1] = s2;
      e[2] = cmbrule ? 1.0 / g.sig[s] : 2.0 / (g.sig[s] + sv);
    }
  }
  for (int s = 0; s < ns; ++s) g.qslt += g.q[s];
  double qsolv = c->total_charge;
  if (c->P.slttype == 1) {
    int m = c->sltlist[0];
    qsolv = qsolv - c->mol_charge[m];
    TRY(ensure_jof(c));
    g.excl = c->jof_k[c->kof[m]];
  }
  g.rf = -3.1415926535897932 * qsolv / c->P.screen / c->P.screen / c->volume;
  TRY(upload_real(c, c->d_gtab, g.tab));
  TRY(upload_real(c, c->d_gq_t, g.q));
  TRY(upload(c, c->d_gqd, g.q.data(), ns));
  TRY(upload(c, c->d_gepsd, g.eps.data(), ns));
  if (random) TRY(upload_args(c, seed, first_index));
  (void)k;
  for (int s0 = 0; s0 < n; s0 += c->Bmax) {
    // k_pose numbers insertion b of this batch first_index + s0 + b
    Batch bt{std::min(c->Bmax, n - s0), s0, nullptr, nullptr, nullptr, random, seed, first_index};
    if (c->P.slttype == 1) bt.solutes = solutes + s0;
    else if (!random) bt.host_pos = solute_pos + (size_t)s0 * ns * 3;
    if (weights) bt.weights = weights + s0;
    int st = c->fp32 ? gradient_batch<float>(c, bt, g, acc) : gradient_batch<double>(c, bt, g, acc);
    if (st) return st;
  }
  return ERM_OK;
}

// ---- frame ingestion ----------------------------------------------------
int32_t erm_dcd_open(const char *path, erm_dcd **out, int32_t *natom_file, int32_t *has_cell) {
  if (!path || !out) return ERM_E_ARG;
  *out = nullptr;
  Dcd *d = new Dcd();
  if (d->open(path)) {
    d->close();
    delete d;
    return ERM_E_ARG;
  }
  if (natom_file) *natom_file = d->natom_file;
  if (has_cell) *has_cell = d->have_cell ? 1 : 0;
  *out = reinterpret_cast<erm_dcd *>(d);
  return ERM_OK;
}

int32_t erm_dcd_read(erm_dcd *h, int32_t natom, int32_t periodic, float *xyz, double *cell) {
  if (!h || !xyz || !cell || natom <= 0) return ERM_E_ARG;
  Dcd *d = reinterpret_cast<Dcd *>(h);
  if (periodic && !d->have_cell) return ERM_E_ARG;
  int r = d->frame(natom, periodic != 0, xyz, cell);
  return r == 0 ? ERM_OK : (r == 1 ? ERM_E_EOF : ERM_E_ARG);
}

void erm_dcd_close(erm_dcd *h) {
  if (!h) return;
  Dcd *d = reinterpret_cast<Dcd *>(h);
  d->close();
  delete d;
}

int32_t erm_read_permutation(const char *path, int32_t natom_traj, int32_t *perm, int32_t *found) {
  if (!path || !perm || natom_traj <= 0) return ERM_E_ARG;
  std::string msg;
  int r = read_permutation(path, natom_traj, perm, msg);
  if (found) *found = (r == 1) ? 0 : 1;
  if (r == 1)
    for (int i = 0; i < natom_traj; ++i) perm[i] = i;
  return r < 0 ? ERM_E_ARG : ERM_OK;
}

}  // extern "C"

// one frame read ahead: pinned coordinates, cell, weight and read status
struct TrajSlot {
  float *xyz = nullptr;
  double cell[9] = {};
  double weight = 1.0;
  int status = 0;  // 0 ok, 1 end of file, -1 malformed frame, -2 bad SysWght
};

struct erm_traj {
  erm_ctx *ctx = nullptr;
  Dcd dcd;
  WeightFile wf;
  bool weights = false, periodic = true, have_perm = false;
  int natom = 0, skip = 1, cur = 0;
  TrajSlot slot[2];
  std::thread reader;
  bool reading = false;
};

// getconf_parallel's read loop for one frame (setconf.F90:917-927): skpcnf
// frames, each with its SysWght line; the last one is kept
static void traj_read(erm_traj *t, int k) {
  TrajSlot &s = t->slot[k];
  for (int i = 0; i < t->skip; ++i) {
    s.status = t->dcd.frame(t->natom, t->periodic, s.xyz, s.cell);
    if (s.status) return;
    s.weight = 1.0;
    if (t->weights && !t->wf.next(s.weight)) {
      s.status = -2;
      return;
    }
  }
}

static void traj_join(erm_traj *t) {
  if (t->reading) t->reader.join();
  t->reading = false;
}

extern "C" {

int32_t erm_traj_open(erm_ctx *c, const char *dcd_path, int32_t natom_traj, const int32_t *perm,
                      const char *weight_path, int32_t skpcnf, erm_traj **out) {
  if (!out) return ERM_E_ARG;
  *out = nullptr;
  TRY(need(c, false, false));
  if (!dcd_path) return fail(c, ERM_E_ARG, "no trajectory path");
  if (skpcnf < 1) return fail(c, ERM_E_ARG, "skpcnf must be >= 1");
  // the trajectory holds the physical molecules, which come first
  int nphys = 0;
  for (int m = 0; m < c->nmol; ++m) {
    const bool phys = (c->sluvid[m] <= 1);
    if (phys) nphys += c->mol_begin[m + 1] - c->mol_begin[m];
    if (phys ? c->mol_begin[m + 1] > natom_traj : c->mol_begin[m] < natom_traj)
      return fail(c, ERM_E_ARG, "the %d trajectory atoms must be those of the solvent and solute molecules, first",
                  natom_traj);
  }
  if (nphys != natom_traj)
    return fail(c, ERM_E_ARG, "the trajectory must hold the %d atoms of the solvent and solute molecules", nphys);
  if (perm) {
    std::vector<int> cnt(natom_traj, 0);
    for (int i = 0; i < natom_traj; ++i) {
      if (perm[i] < 0 || perm[i] >= natom_traj) return fail(c, ERM_E_ARG, "permutation entry out of range (set_pmt)");
      ++cnt[perm[i]];
    }
    for (int i = 0; i < natom_traj; ++i)
      if (cnt[i] != 1) return fail(c, ERM_E_ARG, "atom %d appears %d times in the permutation (set_pmt)", i + 1, cnt[i]);
  }
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  erm_traj *t = new erm_traj();
  t->ctx = c;
  t->natom = natom_traj;
  t->skip = skpcnf;
  t->periodic = (c->P.boxshp == ERM_SYS_PERIODIC);
  auto bail = [&](int code, const char *msg) {
    for (auto &s : t->slot)
      if (s.xyz) (void)hipHostFree(s.xyz);
    t->dcd.close();
    t->wf.close();
    delete t;
    return fail(c, code, "%s", msg);
  };
  if (const char *e = t->dcd.open(dcd_path)) return bail(ERM_E_ARG, e);
  if (t->dcd.natom_file < natom_traj) return bail(ERM_E_ARG, "the DCD file has fewer atoms than the system");
  if (t->periodic && !t->dcd.have_cell) return bail(ERM_E_ARG, "Cell info is requested, but does not exist!");
  if (weight_path) {
    if (!t->wf.open(weight_path)) return bail(ERM_E_ARG, "cannot open the weight file");
    t->weights = true;
  }
  for (auto &s : t->slot)
    if (hipHostMalloc((void **)&s.xyz, sizeof(float) * 3 * natom_traj, hipHostMallocDefault) != hipSuccess)
      return bail(ERM_E_DEVICE, "hipHostMalloc");
  if (ensure(c, c->d_tstage, sizeof(float) * 3 * natom_traj)) return bail(ERM_E_DEVICE, "staging buffer");
  if (perm) {
    if (upload(c, c->d_tperm, perm, natom_traj)) return bail(ERM_E_DEVICE, "permutation upload");
    t->have_perm = true;
  }
  t->reader = std::thread(traj_read, t, 0);
  t->reading = true;
  *out = t;
  return ERM_OK;
}

int32_t erm_traj_next(erm_traj *t, double *cell, double *weight) {
  if (!t) return ERM_E_ARG;
  erm_ctx *c = t->ctx;
  traj_join(t);
  TrajSlot &s = t->slot[t->cur];
  if (s.status == 1) return ERM_E_EOF;
  if (s.status == -1) return fail(c, ERM_E_ARG, "malformed DCD frame");
  if (s.status == -2) return fail(c, ERM_E_ARG, "the weight file is ill-formed");
  if (hipSetDevice(c->P.device) != hipSuccess) return fail(c, ERM_E_DEVICE, "hipSetDevice");
  TRY(set_frame_common(c, s.cell, s.weight));
  HIPCHK(c, hipMemcpyAsync(c->d_tstage.p, s.xyz, sizeof(float) * 3 * t->natom, hipMemcpyHostToDevice, c->st));
  hipLaunchKernelGGL(k_frame_scatter, dim3((t->natom + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, t->natom,
                     dp<float>(c->d_tstage), t->have_perm ? dp<int>(c->d_tperm) : (const int *)nullptr,
                     dp<double>(c->d_pos));
  TRY(launch_check(c, "k_frame_scatter"));
  if (cell) memcpy(cell, s.cell, sizeof s.cell);
  if (weight) *weight = s.weight;
  // the next frame is read into the other pinned buffer while the device
  // prepares this one
  const int nxt = 1 - t->cur;
  t->reader = std::thread(traj_read, t, nxt);
  t->reading = true;
  t->cur = nxt;
  return frame_finish(c);
}

void erm_traj_close(erm_traj *t) {
  if (!t) return;
  traj_join(t);
  for (auto &s : t->slot)
    if (s.xyz) (void)hipHostFree(s.xyz);
  t->dcd.close();
  t->wf.close();
  delete t;
}

}  // extern "C"
