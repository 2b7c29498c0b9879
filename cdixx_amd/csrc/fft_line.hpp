// fft_line.hpp -- device building blocks of the 2D FFT used by the CDI loop.
//
// A line (one image row or one image column) of length N = P*Q is
// transformed by TL lanes with one LDS transpose:
//   stage 1: every lane runs P-point DFTs in registers on the samples it
//            loaded straight from global memory, multiplies by W_N^(q*k1);
//   transpose through LDS (row pitch Q+1 / P+1 keeps it bank-conflict free);
//   stage 2: every lane runs Q-point DFTs in registers.
// Register layouts (t = lane within the line):
//   Din : v[s*P + p]  = x[Q*p + t + TL*s]          s < Q/TL, p < P
//   Dout: v[s*Q + k2] = X[t + TL*s + P*k2]         s < P/TL, k2 < Q
// forward() maps Din -> Dout, inverse() maps Dout -> Din, so a
// forward/pointwise/inverse chain never leaves registers except for the
// transposes, and global loads/stores are coalesced over t in both layouts.
// The transforms are unnormalised: forward exp(-2 pi i nk/N), inverse
// exp(+2 pi i nk/N), the convention of the cuFFT plan behind myFFT/myIFFT in
// CDIxx (src/gpu/cudaConfig.cu:137-187).
#pragma once
#include <hip/hip_runtime.h>
#include <utility>

namespace cdx {

template <class T>
struct cplx {
  T x, y;
};
template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <class T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

// ---------------------------------------------------------------------------
// Compile-time roots of unity.  cos/sin(2 pi e / R) are evaluated by a
// constexpr Taylor series after exact octant reduction on the integer e, so
// every twiddle inside an in-register DFT is a literal.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2*pi*e/R, reduced to |angle| <= pi/4 with exact symmetries
// (8*e is compared against R, so no rounding enters the reduction).
struct CS {
  double c, s;
};
constexpr CS root_cs(long e, long R) {
  e %= R;
  if (e < 0) e += R;
  // angle = 2 pi e / R.  Work in units of R/8 (an octant).
  long e8 = 8 * e;  // angle in octants = e8 / R
  int oct = (int)(e8 / R);
  // angle = 2 pi e / R; choose the nearest quadrant multiple m*pi/2
  int quad = (oct + 1) / 2;  // 0..4
  // residual angle r = 2 pi e / R - quad * pi/2 = 2 pi (4 e - quad R) / (4R)
  double r = 2.0 * kPi * (double)(4 * e - quad * R) / (4.0 * (double)R);
  double c = taylor_cos(r), s = taylor_sin(r);
  switch (quad & 3) {
    case 0: return {c, s};
    case 1: return {-s, c};
    case 2: return {-c, -s};
    default: return {s, -c};
  }
}

template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  [&]<int... I>(std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
  }(std::make_integer_sequence<int, N>{});
}

// v <- v * W_R^(DIR*E) with W_R = exp(2 pi i / R); DIR=-1 forward.
template <int E, int R, int DIR, class T>
__device__ __forceinline__ cplx<T> twiddle_const(cplx<T> v) {
  constexpr long e = ((long)E % R + R) % R;
  if constexpr (e == 0) {
    return v;
  } else if constexpr (4 * e == R) {  // angle pi/2: multiply by DIR*i
    return DIR < 0 ? cplx<T>{v.y, -v.x} : cplx<T>{-v.y, v.x};
  } else if constexpr (2 * e == R) {
    return {-v.x, -v.y};
  } else if constexpr (4 * e == 3 * R) {  // angle 3pi/2
    return DIR < 0 ? cplx<T>{-v.y, v.x} : cplx<T>{v.y, -v.x};
  } else {
    constexpr CS w = root_cs(e, R);
    constexpr T c = (T)w.c;
    constexpr T s = (T)(DIR * w.s);
    return {v.x * c - v.y * s, v.x * s + v.y * c};
  }
}

// ---------------------------------------------------------------------------
// In-register DFT of size R (R = 2^a or 3*2^a, R <= 64), natural order in
// and out, strided access v[i*STRIDE].
// ---------------------------------------------------------------------------
template <int R, int DIR, int STRIDE = 1, class T>
__device__ __forceinline__ void dft(cplx<T>* v) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    cplx<T> a = v[0], b = v[STRIDE];
    v[0] = cadd(a, b);
    v[STRIDE] = csub(a, b);
  } else if constexpr (R == 3) {
    constexpr T h = (T)0.86602540378443864676372317075294;  // sin(2pi/3)
    cplx<T> a = v[0], b = v[STRIDE], c = v[2 * STRIDE];
    cplx<T> s = cadd(b, c), d = csub(b, c);
    cplx<T> m = {a.x - (T)0.5 * s.x, a.y - (T)0.5 * s.y};
    // DIR*i*h*d
    cplx<T> r = DIR < 0 ? cplx<T>{h * d.y, -h * d.x} : cplx<T>{-h * d.y, h * d.x};
    v[0] = cadd(a, s);
    v[STRIDE] = cadd(m, r);
    v[2 * STRIDE] = csub(m, r);
  } else if constexpr (R == 4) {
    cplx<T> a = v[0], b = v[STRIDE], c = v[2 * STRIDE], d = v[3 * STRIDE];
    cplx<T> s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = csub(b, d);
    // DIR*i*d1
    cplx<T> r = DIR < 0 ? cplx<T>{d1.y, -d1.x} : cplx<T>{-d1.y, d1.x};
    v[0] = cadd(s0, s1);
    v[2 * STRIDE] = csub(s0, s1);
    v[STRIDE] = cadd(d0, r);
    v[3 * STRIDE] = csub(d0, r);
  } else {
    // decimation in time: R = F*M, n = F*m + f, k = k1 + M*k2
    constexpr int F = (R % 4 == 0 && R != 12 && R != 24 && R != 48) ? 4 : (R % 3 == 0 ? 3 : 2);
    constexpr int M = R / F;
    // sub-DFTs over m for each residue f: elements v[(F*m+f)*STRIDE]
    static_for<F>([&](auto f) { dft<M, DIR, STRIDE * F>(v + f.value * STRIDE); });
    // now v[(F*k1 + f)*STRIDE] holds Y_f[k1]; combine
    cplx<T> out[R];
    static_for<M>([&](auto k1) {
      cplx<T> z[F];
      static_for<F>([&](auto f) {
        z[f.value] = twiddle_const<f.value * k1.value, R, DIR>(v[(F * k1.value + f.value) * STRIDE]);
      });
      dft<F, DIR, 1>(z);
      static_for<F>([&](auto k2) { out[k1.value + M * k2.value] = z[k2.value]; });
    });
    static_for<R>([&](auto i) { v[i.value * STRIDE] = out[i.value]; });
  }
}

// ---------------------------------------------------------------------------
// Line transform.  lds points at this line's LDS region, which must hold
// max(P*(Q+1), Q*(P+1)) elements.  twf[k1*Q+q] = W_N^(q*k1) and
// twi[n1*P+k1] = W_N^(n1*k1) (forward sign), both read through L1.
// The caller synchronises the workgroup between a transform and any reuse
// of the LDS region.
// ---------------------------------------------------------------------------
template <class T, int P, int Q, int TL>
struct Line {
  static constexpr int N = P * Q;
  static constexpr int E = N / TL;
  static constexpr int SQ = Q / TL;  // stage-1 sub-DFTs per lane (forward)
  static constexpr int SP = P / TL;  // stage-2 sub-DFTs per lane (forward)
  static constexpr int LDQ = Q + 1;
  static constexpr int LDP = P + 1;
  static constexpr int LDS_ELEMS = (P * LDQ > Q * LDP) ? P * LDQ : Q * LDP;
  static_assert(Q % TL == 0 && P % TL == 0, "TL must divide P and Q");

  // Din -> Dout
  __device__ __forceinline__ static void forward(cplx<T>* v, cplx<T>* lds, const cplx<T>* __restrict__ twf,
                                                 int t) {
    static_for<SQ>([&](auto s) {
      dft<P, -1>(v + s.value * P);
      const int q = t + TL * s.value;
      static_for<P - 1>([&](auto k) {
        constexpr int k1 = k.value + 1;
        v[s.value * P + k1] = cmul(v[s.value * P + k1], twf[k1 * Q + q]);
      });
      static_for<P>([&](auto k1) { lds[k1.value * LDQ + q] = v[s.value * P + k1.value]; });
    });
    __syncthreads();
    static_for<SP>([&](auto s) {
      const int k1 = t + TL * s.value;
      static_for<Q>([&](auto q) { v[s.value * Q + q.value] = lds[k1 * LDQ + q.value]; });
      dft<Q, -1>(v + s.value * Q);
    });
  }

  // Dout -> Din
  __device__ __forceinline__ static void inverse(cplx<T>* v, cplx<T>* lds, const cplx<T>* __restrict__ twi,
                                                 int t) {
    static_for<SP>([&](auto s) {
      dft<Q, +1>(v + s.value * Q);
      const int k1 = t + TL * s.value;
      static_for<Q - 1>([&](auto n) {
        constexpr int n1 = n.value + 1;
        v[s.value * Q + n1] = cmulc(v[s.value * Q + n1], twi[n1 * P + k1]);
      });
      static_for<Q>([&](auto n1) { lds[n1.value * LDP + k1] = v[s.value * Q + n1.value]; });
    });
    __syncthreads();
    static_for<SQ>([&](auto s) {
      const int n1 = t + TL * s.value;
      static_for<P>([&](auto k1) { v[s.value * P + k1.value] = lds[n1 * LDP + k1.value]; });
      dft<P, +1>(v + s.value * P);
    });
  }

  // global offsets (in elements along the line) of register i in each layout
  __device__ __forceinline__ static int din_pos(int i, int t) {
    int s = i / P, p = i % P;
    return Q * p + t + TL * s;
  }
  __device__ __forceinline__ static int dout_pos(int i, int t) {
    int s = i / Q, k2 = i % Q;
    return t + TL * s + P * k2;
  }
};

}  // namespace cdx
