// ermgpu_dev.h -- gfx950 device code of the solute-solvent energy pipeline.
//
// Precision is a template parameter (T = float or double) of one kernel
// family; per-molecule energies, the self energy and every accumulator are
// FP64 in both precisions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace erm {

constexpr int NTHR = 256;          // threads per workgroup of the batched kernels
constexpr int MAXFAC = 16;         // radix factors per FFT axis
constexpr int MAXSPEC = 32;        // solvent species
constexpr int MAXSEG = 2048;       // pillar segments per insertion (pair kernel)
constexpr int MAXP = 12;           // spline order
constexpr int ZSCH = 32;           // solute sites per LDS chunk of the z-column build

template <typename T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };

template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) {
  V r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename V> __device__ __forceinline__ V csub(V a, V b) {
  V r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}

// FFT plan of one axis: radices (4, 2, 3, 5 in that order) and the twiddle
// table e^{-2 pi i j / n}, j < n, in device memory.
struct FftAxis {
  int n, nfac;
  int fac[MAXFAC];
};

// Batched in-LDS Stockham FFT over `ncol` columns of length ax.n.  Element i
// of column c sits at base[c * cs + i * es].  A stage of radix R with span p
// reads a[i + q m] (m = n / R), multiplies by w^(q k) (k = i mod p, w the
// (p R)-th root), and writes the R-point DFT to b[(i - k) R + k + q p]; the
// stages ping-pong between buf0 and buf1.  Returns the buffer holding the
// result.  sign = -1 forward, +1 backward (unnormalised).  All threads of the
// block must call it.
template <typename V>
__device__ V *lds_fft(V *buf0, V *buf1, int ncol, int cs, int es, const FftAxis &ax, int sign,
                      const V *__restrict__ tw) {
  const int n = ax.n;
  V *a = buf0, *b = buf1;
  int p = 1;
  for (int f = 0; f < ax.nfac; ++f) {
    const int R = ax.fac[f];
    const int m = n / R;
    const int tstep = n / (p * R);
    const int nbf = ncol * m;
    for (int t = threadIdx.x; t < nbf; t += blockDim.x) {
      const int col = t / m;
      const int i = t - col * m;
      const int k = i % p;
      V x[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        if (q < R) {
          V v = a[col * cs + (i + q * m) * es];
          if (q && k) {
            V w = tw[(q * k * tstep) % n];
            if (sign > 0) w.y = -w.y;
            v = cmul(v, w);
          }
          x[q] = v;
        }
      }
      V *o = b + col * cs;
      const int j = (i - k) * R + k;
      if (R == 4) {
        V s0 = cadd(x[0], x[2]), d0 = csub(x[0], x[2]);
        V s1 = cadd(x[1], x[3]), d1 = csub(x[1], x[3]);
        // multiply d1 by sign * i
        V jd1;
        jd1.x = -sign * d1.y;
        jd1.y = sign * d1.x;
        o[(j)*es] = cadd(s0, s1);
        o[(j + p) * es] = cadd(d0, jd1);
        o[(j + 2 * p) * es] = csub(s0, s1);
        o[(j + 3 * p) * es] = csub(d0, jd1);
      } else if (R == 2) {
        o[(j)*es] = cadd(x[0], x[1]);
        o[(j + p) * es] = csub(x[0], x[1]);
      } else {
        const int rs = n / R;
        for (int q = 0; q < R; ++q) {
          V acc = x[0];
          for (int s = 1; s < R; ++s) {
            V w = tw[((s * q) % R) * rs];
            if (sign > 0) w.y = -w.y;
            acc = cadd(acc, cmul(x[s], w));
          }
          o[(j + q * p) * es] = acc;
        }
      }
    }
    __syncthreads();
    V *tmp = a;
    a = b;
    b = tmp;
    p *= R;
  }
  return a;
}

// B-spline M_p(x) from its explicit binomial form, evaluated exactly as the
// reference's spline_value (spline.F90:47-59) with integer powers by
// repeated squaring.
__device__ __forceinline__ double ipow_dev(double x, int n) {
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

__device__ __forceinline__ double spline_value_dev(const double *coeff, int order, double x) {
  double f = 0.0;
  if (x > 0.0 && x < (double)order) {
    int k = (int)x;
    for (int i = 0; i <= k; ++i) f = f + coeff[i] * ipow_dev(x - (double)i, order - 1);
  }
  return f;
}

// ordered-integer encoding of doubles for atomic min / max
__device__ __forceinline__ unsigned long long dbl_to_ord(double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u & 0x8000000000000000ULL) ? ~u : (u | 0x8000000000000000ULL);
}
__device__ __forceinline__ double ord_to_dbl(unsigned long long u) {
  u = (u & 0x8000000000000000ULL) ? (u & 0x7FFFFFFFFFFFFFFFULL) : ~u;
  return __longlong_as_double((long long)u);
}

}  // namespace erm

namespace erm {

// ---------------------------------------------------------------------------
// Register-resident FFT of compile-time length N (factors 4, 2, 3, 5): one
// thread transforms one column held in registers; all indices are
// compile-time, so the arrays stay in VGPRs.  Stockham stages as in lds_fft;
// twiddles come from the axis table tw (e^{-2 pi i j / N}), conjugated for
// the backward transform.
// ---------------------------------------------------------------------------
template <int N> struct Fact {
  static constexpr int nf() {
    int n = N, c = 0;
    for (int r : {4, 2, 3, 5})
      while (n % r == 0) {
        n /= r;
        ++c;
      }
    return c;
  }
  static constexpr int f(int s) {
    int n = N, c = 0;
    for (int r : {4, 2, 3, 5})
      while (n % r == 0) {
        n /= r;
        if (c == s) return r;
        ++c;
      }
    return 1;
  }
  static constexpr int span(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= f(i);
    return p;
  }
  static constexpr bool valid() {
    int n = N;
    for (int r : {4, 2, 3, 5})
      while (n % r == 0) n /= r;
    return n == 1;
  }
};

template <int R, int SIGN, typename V>
__device__ __forceinline__ void reg_butterfly(V (&x)[R], V *o, int p, const V *__restrict__ tw, int N) {
  if constexpr (R == 4) {
    V s0 = cadd(x[0], x[2]), d0 = csub(x[0], x[2]);
    V s1 = cadd(x[1], x[3]), d1 = csub(x[1], x[3]);
    V jd1;
    jd1.x = -SIGN * d1.y;
    jd1.y = SIGN * d1.x;
    o[0] = cadd(s0, s1);
    o[p] = cadd(d0, jd1);
    o[2 * p] = csub(s0, s1);
    o[3 * p] = csub(d0, jd1);
  } else if constexpr (R == 2) {
    o[0] = cadd(x[0], x[1]);
    o[p] = csub(x[0], x[1]);
  } else {
    const int rs = N / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      V acc = x[0];
#pragma unroll
      for (int s = 1; s < R; ++s) {
        V w = tw[((s * q) % R) * rs];
        if (SIGN > 0) w.y = -w.y;
        acc = cadd(acc, cmul(x[s], w));
      }
      o[q * p] = acc;
    }
  }
}

template <int N, int S, int SIGN, typename V>
__device__ __forceinline__ void reg_stages(V (&a)[N], V (&b)[N], const V *__restrict__ tw) {
  if constexpr (S < Fact<N>::nf()) {
    constexpr int R = Fact<N>::f(S);
    constexpr int p = Fact<N>::span(S);
    constexpr int m = N / R;
    constexpr int ts = N / (p * R);
#pragma unroll
    for (int i = 0; i < m; ++i) {
      const int k = i % p;
      V x[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        V v = a[i + q * m];
        if (q && k) {
          V w = tw[(q * k * ts) % N];
          if (SIGN > 0) w.y = -w.y;
          v = cmul(v, w);
        }
        x[q] = v;
      }
      reg_butterfly<R, SIGN>(x, b + (i - k) * R + k, p, tw, N);
    }
    reg_stages<N, S + 1, SIGN>(b, a, tw);
  }
}

// in-place (from the caller's view) transform of x[N]
template <int N, int SIGN, typename V> __device__ __forceinline__ void reg_fft(V (&x)[N], const V *__restrict__ tw) {
  V y[N];
  reg_stages<N, 0, SIGN>(x, y, tw);
  if constexpr (Fact<N>::nf() % 2 == 1) {
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = y[i];
  }
}

}  // namespace erm
