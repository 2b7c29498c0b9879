// fft_line.hpp -- device building blocks of the 2D FFT used by the CDI loop.
//
// A line (one image row or one image column) of N points is transformed by
// TL lanes, E = N/TL points per lane, with a mixed-radix Stockham schedule:
// every pass runs radix-R DFTs in registers (E/R per lane) and the passes
// exchange data through LDS.  For a pass with radix R after passes whose
// radices multiply to Ns, butterfly j in [0, N/R) reads x[j + r*N/R],
// multiplies by W_{Ns*R}^{(j mod Ns)*r}, and writes its outputs to
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns]; the last pass leaves the result in
// natural order.  Lane t owns butterflies j = t + TL*s, so
//   Din  (input of the forward transform):   v[s*R0 + r] = x[t + TL*s + r*N/R0]
//   Dout (output of the forward transform):  v[s*RL + r] = X[t + TL*s + r*N/RL]
// with R0 / RL the first / last radix.  The inverse runs the radices in the
// opposite order, so it maps Dout back to Din: a forward -> pointwise ->
// inverse chain never reorders data outside the passes, and global loads and
// stores in both layouts are contiguous over t.
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
// constexpr Taylor series after exact quadrant reduction on the integer e,
// so every twiddle inside an in-register DFT is a literal.
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
struct CS {
  double c, s;
};
constexpr CS root_cs(long e, long R) {
  e %= R;
  if (e < 0) e += R;
  const long e8 = 8 * e;  // angle in octants = e8 / R
  const int oct = (int)(e8 / R);
  const int quad = (oct + 1) / 2;  // nearest multiple of pi/2
  const double r = 2.0 * kPi * (double)(4 * e - quad * R) / (4.0 * (double)R);  // |r| <= pi/4
  const double c = taylor_cos(r), s = taylor_sin(r);
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

// v * W_R^(DIR*E) with W_R = exp(2 pi i / R); DIR = -1 forward
template <int E, int R, int DIR, class T>
__device__ __forceinline__ cplx<T> twiddle_const(cplx<T> v) {
  constexpr long e = ((long)E % R + R) % R;
  if constexpr (e == 0) {
    return v;
  } else if constexpr (4 * e == R) {
    return DIR < 0 ? cplx<T>{v.y, -v.x} : cplx<T>{-v.y, v.x};
  } else if constexpr (2 * e == R) {
    return {-v.x, -v.y};
  } else if constexpr (4 * e == 3 * R) {
    return DIR < 0 ? cplx<T>{-v.y, v.x} : cplx<T>{v.y, -v.x};
  } else {
    constexpr CS w = root_cs(e, R);
    constexpr T c = (T)w.c;
    constexpr T s = (T)(DIR * w.s);
    return {v.x * c - v.y * s, v.x * s + v.y * c};
  }
}

// ---------------------------------------------------------------------------
// In-register DFT of size R (R = 2^a * 3^b), natural order in and out.
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
    cplx<T> r = DIR < 0 ? cplx<T>{h * d.y, -h * d.x} : cplx<T>{-h * d.y, h * d.x};
    v[0] = cadd(a, s);
    v[STRIDE] = cadd(m, r);
    v[2 * STRIDE] = csub(m, r);
  } else if constexpr (R == 4) {
    cplx<T> a = v[0], b = v[STRIDE], c = v[2 * STRIDE], d = v[3 * STRIDE];
    cplx<T> s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = csub(b, d);
    cplx<T> r = DIR < 0 ? cplx<T>{d1.y, -d1.x} : cplx<T>{-d1.y, d1.x};
    v[0] = cadd(s0, s1);
    v[2 * STRIDE] = csub(s0, s1);
    v[STRIDE] = cadd(d0, r);
    v[3 * STRIDE] = csub(d0, r);
  } else {
    // decimation in time: R = F*M, n = F*m + f, k = k1 + M*k2
    constexpr int F = (R % 4 == 0 && R % 3 != 0) ? 4 : (R % 3 == 0 ? 3 : 2);
    constexpr int M = R / F;
    static_for<F>([&](auto f) { dft<M, DIR, STRIDE * F>(v + f.value * STRIDE); });
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
// Stockham line transform.
// ---------------------------------------------------------------------------
template <int... Rs>
struct Radices {
  static constexpr int M = sizeof...(Rs);
  static constexpr int R[M] = {Rs...};
};

// PAD: one pad element per 2^PAD elements of the exchange buffer (0: none),
// which keeps strided pass outputs off a single bank
template <class T, int N, int TL, class RL, int PAD = 5>
struct Line {
  static constexpr int M = RL::M;
  static constexpr int E = N / TL;
  static constexpr int R0 = RL::R[0];
  static constexpr int RLAST = RL::R[M - 1];
  static constexpr int LDS_ELEMS = PAD ? N + (N >> PAD) + 1 : N + 1;
  static_assert(TL * E == N, "TL must divide N");

  __device__ __forceinline__ static int swz(int p) {
    if constexpr (PAD > 0) return p + (p >> PAD);
    else return p;
  }

  // radix of pass p in direction DIR (forward runs the list, inverse reverses it)
  template <int DIR>
  static constexpr int radix(int p) {
    return DIR < 0 ? RL::R[p] : RL::R[M - 1 - p];
  }
  template <int DIR>
  static constexpr int ns(int p) {
    int s = 1;
    for (int q = 0; q < p; q++) s *= radix<DIR>(q);
    return s;
  }

  // Table twiddles of a transform.  Pass p (radix R, Ns > 1) multiplies
  // butterfly j = t + TL*s by W_N^(r * b), b = (j mod Ns) * N/(Ns*R): for a
  // power-of-two R the lane holds W^(2^k b) and forms the other powers by
  // products, otherwise it holds W^(r b), r = 1..R-1.  All of a transform's
  // loads are issued together (tw_load) so their latency overlaps the data
  // loads instead of stalling every exchange.
  static constexpr bool pow2(int r) { return (r & (r - 1)) == 0; }
  static constexpr int log2c(int r) {
    int k = 0;
    while ((1 << k) < r) k++;
    return k;
  }
  template <int DIR>
  static constexpr int tw_per(int p) {
    return ns<DIR>(p) == 1 ? 0 : (pow2(radix<DIR>(p)) ? log2c(radix<DIR>(p)) : radix<DIR>(p) - 1);
  }
  template <int DIR>
  static constexpr int tw_off(int p) {  // first slot of pass p
    int o = 0;
    for (int q = 0; q < p; q++) o += (E / radix<DIR>(q)) * tw_per<DIR>(q);
    return o;
  }
  template <int DIR>
  struct Tw {
    static constexpr int K = tw_off<DIR>(M);
    cplx<T> w[K > 0 ? K : 1];
  };
  // The lane's table twiddles.  FP64 lines load each value.  FP32 lines
  // read the table's double copy (the n double pairs after the n float
  // pairs) and form in double, rounding each value once: for a power-of-two
  // R, W^(2^k b) by squaring W^b; for butterflies t + TL*s whose index mod Ns
  // equals that of s = 0, s = 0's values; where no lane's index wraps mod Ns
  // (TL*(s+1) <= Ns), s = 0's values times the compile-time root
  // W^(rr*TL*s*stride).  A pass then issues one load per distinct angle.
  template <int DIR>
  __device__ __forceinline__ static Tw<DIR> tw_load(const cplx<T>* __restrict__ tw, int t) {
    Tw<DIR> out;
    static_for<M>([&](auto pi) {
      constexpr int p = pi.value;
      constexpr int R = radix<DIR>(p);
      constexpr int Ns = ns<DIR>(p);
      if constexpr (Ns > 1) {
        constexpr int stride = N / (Ns * R);
        constexpr int per = tw_per<DIR>(p);
        constexpr int o0 = tw_off<DIR>(p);
        if constexpr (sizeof(T) == 8) {
          static_for<E / R>([&](auto s) {
            const int b = ((t + TL * s.value) % Ns) * stride;
            static_for<per>([&](auto k) {
              constexpr int rr = pow2(R) ? (1 << k.value) : k.value + 1;
              out.w[o0 + s.value * per + k.value] = tw[rr * b];
            });
          });
        } else {
          const cplx<double>* __restrict__ twd = reinterpret_cast<const cplx<double>*>(tw + N);
          cplx<double> w0[per];  // s = 0's values in double
          static_for<E / R>([&](auto s) {
            constexpr int sv = s.value;
            constexpr int o = o0 + sv * per;
            constexpr bool same = sv > 0 && (TL * sv) % Ns == 0;
            constexpr bool shifted = sv > 0 && !same && TL * (sv + 1) <= Ns;
            const int b = ((t + TL * sv) % Ns) * stride;
            cplx<double> w = {0.0, 0.0};
            static_for<per>([&](auto k) {
              constexpr int rr = pow2(R) ? (1 << k.value) : k.value + 1;
              if constexpr (same) {
                out.w[o + k.value] = out.w[o0 + k.value];
              } else {
                if constexpr (shifted) {
                  w = twiddle_const<rr * TL * sv * stride, N, -1>(w0[k.value]);
                } else if constexpr (pow2(R) && k.value > 0) {
                  w = {w.x * w.x - w.y * w.y, 2.0 * w.x * w.y};
                } else {
                  w = twd[rr * b];
                }
                if constexpr (sv == 0) w0[k.value] = w;
                out.w[o + k.value] = {(T)w.x, (T)w.y};
              }
            });
          });
        }
      }
    });
    return out;
  }
  // v[r] *= W^(DIR r b) for r = 1..R-1 from the lane's preloaded values w
  template <int R, int DIR>
  __device__ __forceinline__ static void twiddle(cplx<T>* v, const cplx<T>* wl) {
    cplx<T> w[R];
    if constexpr (pow2(R)) {
      static_for<log2c(R)>([&](auto k) { w[1 << k.value] = wl[k.value]; });
      static_for<R>([&](auto r) {
        constexpr int rr = r.value;
        if constexpr (rr > 0 && !pow2(rr)) {
          constexpr int hi = 1 << (31 - __builtin_clz(rr));
          w[rr] = cmul(w[hi], w[rr - hi]);
        }
      });
    } else {
      static_for<R - 1>([&](auto r) { w[r.value + 1] = wl[r.value]; });
    }
    static_for<R - 1>([&](auto r) {
      constexpr int rr = r.value + 1;
      v[rr] = DIR < 0 ? cmul(v[rr], w[rr]) : cmulc(v[rr], w[rr]);
    });
  }

  // Exchange indices (before the padding swizzle).  After pass p, butterfly
  // j = t + TL*s writes its output r to (j/Ns)*Ns*R + (j mod Ns) + r*Ns; pass
  // p+1 reads its input r of butterfly j from j + r*N/R2.
  static constexpr int swz_c(int p) { return PAD > 0 ? p + (p >> PAD) : p; }
  template <int DIR>
  static constexpr int w_index(int p, int t, int s, int r) {
    const int R = radix<DIR>(p), Ns = ns<DIR>(p), j = t + TL * s;
    return (j / Ns) * Ns * R + (j % Ns) + r * Ns;
  }
  template <int DIR>
  static constexpr int r_index(int p, int t, int s, int r) {
    return t + TL * s + r * (N / radix<DIR>(p + 1));
  }
  // Offsets from the lane's first element, swizzled, and whether they are
  // the same for every lane (checked over all t at compile time): then one
  // base address and immediate offsets serve the whole exchange.
  template <int DIR>
  static constexpr int w_offset(int p, int s, int r) {
    return swz_c(w_index<DIR>(p, 0, s, r)) - swz_c(w_index<DIR>(p, 0, 0, 0));
  }
  template <int DIR>
  static constexpr int r_offset(int p, int s, int r) {
    return swz_c(r_index<DIR>(p, 0, s, r)) - swz_c(r_index<DIR>(p, 0, 0, 0));
  }
  template <int DIR>
  static constexpr bool w_linear(int p) {
    const int R = radix<DIR>(p);
    for (int t = 0; t < TL; t++)
      for (int s = 0; s < E / R; s++)
        for (int r = 0; r < R; r++)
          if (swz_c(w_index<DIR>(p, t, s, r)) - swz_c(w_index<DIR>(p, t, 0, 0)) != w_offset<DIR>(p, s, r))
            return false;
    return true;
  }
  template <int DIR>
  static constexpr bool r_linear(int p) {
    const int R2 = radix<DIR>(p + 1);
    for (int t = 0; t < TL; t++)
      for (int s = 0; s < E / R2; s++)
        for (int r = 0; r < R2; r++)
          if (swz_c(r_index<DIR>(p, t, s, r)) - swz_c(r_index<DIR>(p, t, 0, 0)) != r_offset<DIR>(p, s, r))
            return false;
    return true;
  }

  // DIR = -1: Din -> Dout (forward); DIR = +1: Dout -> Din (inverse).
  // lds: this line's first exchange buffer (LDS_ELEMS), lds + alt its second;
  // ph selects the buffer of the next exchange and flips after each one.  One
  // barrier per exchange suffices: a buffer is written again two exchanges
  // later, after every thread has passed the barrier of the exchange in
  // between, which follows its reads.  tw: the lane's table twiddles
  // (tw_load of the table tw[m] = W_N^m, forward sign).
  template <int DIR>
  __device__ __forceinline__ static void run(cplx<T>* v, cplx<T>* lds0, int alt, int& ph, const Tw<DIR>& tw,
                                             int t) {
    static_for<M>([&](auto pi) {
      constexpr int p = pi.value;
      constexpr int R = radix<DIR>(p);
      constexpr int Ns = ns<DIR>(p);
      constexpr int S = E / R;
      static_assert(E % R == 0, "every radix must divide the points per lane");
      static_for<S>([&](auto s) {
        if constexpr (Ns > 1) twiddle<R, DIR>(v + s.value * R, tw.w + tw_off<DIR>(p) + s.value * tw_per<DIR>(p));
        dft<R, DIR>(v + s.value * R);
      });
      if constexpr (p < M - 1) {
        cplx<T>* lds = lds0 + ph * alt;
        ph ^= 1;
        if constexpr (w_linear<DIR>(p)) {
          // every write at the lane's base plus a compile-time offset
          cplx<T>* wb = lds + swz(w_index<DIR>(p, t, 0, 0));
          static_for<S>([&](auto s) {
            static_for<R>([&](auto r) {
              wb[w_offset<DIR>(p, s.value, r.value)] = v[s.value * R + r.value];
            });
          });
        } else {
          static_for<S>([&](auto s) {
            static_for<R>([&](auto r) { lds[swz(w_index<DIR>(p, t, s.value, r.value))] = v[s.value * R + r.value]; });
          });
        }
        __syncthreads();
        constexpr int R2 = radix<DIR>(p + 1);
        if constexpr (r_linear<DIR>(p)) {
          const cplx<T>* rb = lds + swz(r_index<DIR>(p, t, 0, 0));
          static_for<E / R2>([&](auto s) {
            static_for<R2>([&](auto r) { v[s.value * R2 + r.value] = rb[r_offset<DIR>(p, s.value, r.value)]; });
          });
        } else {
          static_for<E / R2>([&](auto s) {
            static_for<R2>([&](auto r) { v[s.value * R2 + r.value] = lds[swz(r_index<DIR>(p, t, s.value, r.value))]; });
          });
        }
      }
    });
  }
  using TwF = Tw<-1>;
  using TwI = Tw<+1>;
  __device__ __forceinline__ static void forward(cplx<T>* v, cplx<T>* lds, int alt, int& ph, const TwF& tw, int t) {
    run<-1>(v, lds, alt, ph, tw, t);
  }
  __device__ __forceinline__ static void inverse(cplx<T>* v, cplx<T>* lds, int alt, int& ph, const TwI& tw, int t) {
    run<+1>(v, lds, alt, ph, tw, t);
  }

  // element index along the line of register i in each layout
  __device__ __forceinline__ static int din_pos(int i, int t) { return t + TL * (i / R0) + (i % R0) * (N / R0); }
  __device__ __forceinline__ static int dout_pos(int i, int t) {
    return t + TL * (i / RLAST) + (i % RLAST) * (N / RLAST);
  }
};

}  // namespace cdx
