// ermgpu_dev.h -- gfx950 device code of the solute-solvent energy pipeline.
//
// Precision is a template parameter (T = float or double) of one kernel
// family; per-molecule energies, the self energy and every accumulator are
// FP64 in both precisions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

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
// thread transforms one column held in registers.  Mixed-radix decimation in
// frequency, in place: a stage of radix R on a block of N = R M reads
// x[n1 + M n2] (n2 < R), takes the R-point DFT, multiplies output k2 by
// w_N^(n1 k2) and writes it back to x[n1 + M k2]; each of the R blocks of
// length M is then transformed the same way.  Position p of the result holds
// X[rev_N(p)] with rev_N(p) = R rev_M(p mod M) + p / M, undone by renaming
// registers.  Every index and twiddle is a compile-time constant: twiddles
// come from a constexpr sine/cosine (octant reduction of the rational
// angle, Taylor series to order 25 on [0, pi/4]), so no table is read and
// only N complex values are live.
// ---------------------------------------------------------------------------
struct CRoot {
  double re, im;
};

constexpr double kTwoPiOver4 = 1.57079632679489661923132169163975144;  // pi / 2

constexpr double taylor_sin(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k <= 12; ++k) {
    term *= -x2 / (double)((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k <= 12; ++k) {
    term *= -x2 / (double)((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}

// e^{-2 pi i j / N}
constexpr CRoot unit_root(int j, int N) {
  j %= N;
  if (j < 0) j += N;
  const int q = (4 * j) / N;          // quadrant
  const int r = 4 * j - q * N;        // angle within the quadrant = (pi/2) r / N
  double c, s;
  if (2 * r <= N) {
    const double t = kTwoPiOver4 * (double)r / (double)N;
    c = taylor_cos(t);
    s = taylor_sin(t);
  } else {
    const double t = kTwoPiOver4 * (double)(N - r) / (double)N;
    c = taylor_sin(t);
    s = taylor_cos(t);
  }
  double cq = c, sq = s;  // angle q pi/2 + theta
  if (q == 1) {
    cq = -s;
    sq = c;
  } else if (q == 2) {
    cq = -c;
    sq = -s;
  } else if (q == 3) {
    cq = s;
    sq = -c;
  }
  return CRoot{cq, -sq};
}

constexpr int first_radix(int n) { return n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : n; }

constexpr int dif_rev(int n, int p) {
  if (n == 1) return 0;
  const int R = first_radix(n), M = n / R;
  return R * dif_rev(M, p % M) + p / M;
}

template <int N> struct Fact {
  static constexpr bool valid() {
    int n = N;
    for (int r : {4, 2, 3, 5})
      while (n % r == 0) n /= r;
    return n == 1;
  }
};

// w_N^m for the transform direction SIGN (-1 forward, +1 backward)
template <typename V, int SIGN, int M_, int N> __host__ __device__ __forceinline__ V root_v() {
  constexpr CRoot w = unit_root(M_, N);
  V r;
  r.x = (decltype(r.x))w.re;
  r.y = (decltype(r.x))(SIGN < 0 ? w.im : -w.im);
  return r;
}

template <typename V> __host__ __device__ __forceinline__ V cmul_h(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// v times w_N^m (m = 0 leaves v unchanged)
template <int SIGN, int M_, int N, typename V> __host__ __device__ __forceinline__ V rot(V v) {
  if constexpr (M_ % N == 0) {
    return v;
  } else if constexpr (4 * (M_ % N) == N) {  // w^(N/4) = SIGN i
    V r;
    r.x = -SIGN * v.y;
    r.y = SIGN * v.x;
    return r;
  } else if constexpr (2 * (M_ % N) == N) {  // w^(N/2) = -1
    V r;
    r.x = -v.x;
    r.y = -v.y;
    return r;
  } else if constexpr (4 * (M_ % N) == 3 * N) {  // w^(3N/4) = -SIGN i
    V r;
    r.x = SIGN * v.y;
    r.y = -SIGN * v.x;
    return r;
  } else {
    return cmul_h(v, root_v<V, SIGN, M_ % N, N>());
  }
}

template <typename V> __host__ __device__ __forceinline__ V cadd_h(V a, V b) {
  V r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}

// output q of an R-point DFT: sum_s v[s] w_R^(s q)
template <int R, int SIGN, int Q, typename V, int... S>
__host__ __device__ __forceinline__ V dft_out(const V (&v)[R], std::integer_sequence<int, S...>) {
  V acc = v[0];
  ((acc = (S == 0) ? acc : cadd_h(acc, rot<SIGN, (S * Q) % R, R>(v[S]))), ...);
  return acc;
}

template <int R, int SIGN, typename V, int... Q>
__host__ __device__ __forceinline__ void small_dft(const V (&v)[R], V (&o)[R], std::integer_sequence<int, Q...>) {
  if constexpr (R == 4) {
    const V s0 = cadd_h(v[0], v[2]), s1 = cadd_h(v[1], v[3]);
    V d0, d1;
    d0.x = v[0].x - v[2].x;
    d0.y = v[0].y - v[2].y;
    d1.x = v[1].x - v[3].x;
    d1.y = v[1].y - v[3].y;
    const V jd = rot<SIGN, 1, 4>(d1);
    o[0] = cadd_h(s0, s1);
    o[1] = cadd_h(d0, jd);
    o[2].x = s0.x - s1.x;
    o[2].y = s0.y - s1.y;
    o[3].x = d0.x - jd.x;
    o[3].y = d0.y - jd.y;
  } else if constexpr (R == 2) {
    o[0] = cadd_h(v[0], v[1]);
    o[1].x = v[0].x - v[1].x;
    o[1].y = v[0].y - v[1].y;
  } else {
    ((o[Q] = dft_out<R, SIGN, Q>(v, std::make_integer_sequence<int, R>{})), ...);
  }
}

template <int TOT, int N, int BASE, int SIGN, typename V> __host__ __device__ __forceinline__ void dif(V (&x)[TOT]);

// one column n1 of a DIF stage: R-point DFT of x[n1 + M k], twiddled by
// w_N^(n1 k), back into the same slots
template <int TOT, int N, int BASE, int SIGN, int N1, typename V, int... K>
__host__ __device__ __forceinline__ void dif_col(V (&x)[TOT], std::integer_sequence<int, K...>) {
  constexpr int R = sizeof...(K), M = N / R;
  const V v[R] = {x[BASE + N1 + M * K]...};
  V o[R];
  small_dft<R, SIGN>(v, o, std::make_integer_sequence<int, R>{});
  ((x[BASE + N1 + M * K] = rot<SIGN, N1 * K, N>(o[K])), ...);
}

template <int TOT, int N, int BASE, int SIGN, typename V, int... N1>
__host__ __device__ __forceinline__ void dif_cols(V (&x)[TOT], std::integer_sequence<int, N1...>) {
  (dif_col<TOT, N, BASE, SIGN, N1>(x, std::make_integer_sequence<int, first_radix(N)>{}), ...);
}

template <int TOT, int N, int BASE, int SIGN, typename V, int... K>
__host__ __device__ __forceinline__ void dif_blocks(V (&x)[TOT], std::integer_sequence<int, K...>) {
  constexpr int M = N / first_radix(N);
  (dif<TOT, M, BASE + M * K, SIGN>(x), ...);
}

template <int TOT, int N, int BASE, int SIGN, typename V> __host__ __device__ __forceinline__ void dif(V (&x)[TOT]) {
  if constexpr (N > 1) {
    constexpr int R = first_radix(N), M = N / R;
    static_assert(R * M == N && R > 1, "FFT length must factor into 2, 3, 4, 5");
    dif_cols<TOT, N, BASE, SIGN>(x, std::make_integer_sequence<int, M>{});
    dif_blocks<TOT, N, BASE, SIGN>(x, std::make_integer_sequence<int, R>{});
  }
}

template <int N, int P> struct DifRev {
  static constexpr int value = dif_rev(N, P);
};

template <int N, typename V, int... P>
__host__ __device__ __forceinline__ void dif_unscramble(V (&x)[N], std::integer_sequence<int, P...>) {
  const V y[N] = {x[P]...};
  ((x[DifRev<N, P>::value] = y[P]), ...);
}

// in-place transform of x[N] (natural order in and out), unnormalised;
// SIGN = -1 forward (e^{-i}), +1 backward
template <int N, int SIGN, typename V> __host__ __device__ __forceinline__ void reg_fft(V (&x)[N]) {
  static_assert(Fact<N>::valid(), "FFT length must factor into 2, 3, 4, 5");
  dif<N, N, 0, SIGN>(x);
  dif_unscramble<N>(x, std::make_integer_sequence<int, N>{});
}

// Inverse real transform of one Hermitian row through a half-length complex
// FFT: for k < M = N/2 with X[0] and X[M] taken as real,
//   E[k] = X[k] + conj X[M-k],  O[k] = (X[k] - conj X[M-k]) e^{+2 pi i k/N},
//   z = IFFT_M(E + i O)  gives  x[2m] = Re z[m], x[2m+1] = Im z[m]
// (unnormalised, e^{+i}), i.e. x[n] = sum_{k<N} X[k] e^{2 pi i k n / N}.
template <typename V> __host__ __device__ __forceinline__ V c2r_one(V a, V b, V w) {
  V e, d, z;
  e.x = a.x + b.x;
  e.y = a.y - b.y;
  d.x = a.x - b.x;
  d.y = a.y + b.y;
  const V o = cmul_h(d, w);
  z.x = e.x - o.y;
  z.y = e.y + o.x;
  return z;
}

template <int N, typename V, int... K>
__host__ __device__ __forceinline__ void c2r_pack(const V *__restrict__ X, V (&Z)[N / 2],
                                                  std::integer_sequence<int, K...>) {
  constexpr int M = N / 2;
  V x0 = X[0], xm = X[M];
  x0.y = 0;
  xm.y = 0;
  ((Z[K] = c2r_one(K == 0 ? x0 : X[K], K == 0 ? xm : X[M - K], root_v<V, +1, K, N>())), ...);
}

template <int N, typename V> __host__ __device__ __forceinline__ void c2r_half(const V *__restrict__ X, V (&Z)[N / 2]) {
  c2r_pack<N>(X, Z, std::make_integer_sequence<int, N / 2>{});
  reg_fft<N / 2, +1>(Z);
}

// the same with the (unused) twiddle table of the axis, for the kernels that
// hold one
template <int N, int SIGN, typename V>
__host__ __device__ __forceinline__ void reg_fft(V (&x)[N], const V *__restrict__) {
  reg_fft<N, SIGN>(x);
}

}  // namespace erm
