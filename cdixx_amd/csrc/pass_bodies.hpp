// pass_bodies.hpp -- device bodies of the fused FFT passes (shared by the
// single-frame kernels in passes.hip and the batched kernels in batch.hip).
//
// One iteration of CDI::phaseRetrieve (CDIxx src/post_gpu/cdi.cc:349-354,381)
// is projection -> IFFT -> 1/N -> support update -> FFT.  Here it is two
// passes that each stream the image once:
//   column pass  (COL_FWD_PROJ_INV): forward column FFT of T1 = rowFFT(g),
//                 modulus projection (+R-factor sums), inverse column FFT -> T2
//   row pass     (ROW_INV_UPDATE_FWD): inverse row FFT of T2, 1/N,
//                 HIO/RAAR/ER/... update against g and the support, store g,
//                 forward row FFT -> T1
// so F = colFFT(T1) is never stored between iterations.  Other OP values
// serve the standalone 2D FFT and shrink-wrap.
//
// The projection mode, the R-factor switch and the update rule are template
// parameters (no per-pixel branching on them), and every operand a pass
// reads besides the line itself (I; g and S) is loaded at the start so its
// latency overlaps the first transform.
//
// In the single-frame loop (TILED) T1 and the packed measurement are kept in
// column tiles (t1_index, i_index), so the column pass reads whole or half
// 64-byte pieces; T1 and T2 are moved with streaming (non-temporal) loads
// and stores, while g and S are loaded plainly (and g stored plainly up to
// 1024^2), the combination measured fastest.  (FETCH_SIZE shows g and S
// fetched from beyond L2 every iteration all the same: the gain is the load
// policy, not reuse.)
#pragma once
#include "device_ops.hpp"

namespace cdx {

// ---------------------------------------------------------------------------
// pointwise operations with compile-time selection
// ---------------------------------------------------------------------------
// applyModAbs (cudaConfig.cu:827-841) on the packed amplitude A = sqrt(max(I,0)):
// F *= A/|F|, F = (A, 0) where |F| = 0.  The FP32 build scales by A*rsqrt(|F|^2)
// (within 2 ulp of the reference's sqrt(I)/hypot(F)).
template <class T>
__device__ __forceinline__ cplx<T> mod_amp(cplx<T> f, T A) {
  if constexpr (sizeof(T) == 4) {
    const float m2 = f.x * f.x + f.y * f.y;
    if (m2 > 0.f) {
      const float s = A * rsqrtf(m2);
      return {f.x * s, f.y * s};
    }
    return {A, 0.f};
  } else {
    const double mod = dhypot(f.x, f.y);
    if (mod == 0.0) return {A, 0.0};
    const double s = A / mod;
    return {f.x * s, f.y * s};
  }
}

// The packed operand holds the amplitude for the modulus modes and the
// intensity for the others (NaN marks unmeasured pixels).
template <int MODE>
__device__ __forceinline__ bool packed_amplitude(int runtime_mode) {
  if constexpr (MODE < 0) return runtime_mode == PROJ_MOD_ABS || runtime_mode == PROJ_MOD_ABS_BS;
  else return MODE == PROJ_MOD_ABS || MODE == PROJ_MOD_ABS_BS;
}

template <class T, int MODE>
__device__ __forceinline__ cplx<T> project_m(cplx<T> f, T m, const ProjArgs<T>& pa) {
  if constexpr (MODE == PROJ_MOD_NOISE) return mod_noise(f, m, pa.noiseLevel, pa.scale, pa.rcolor);
  else if constexpr (MODE == PROJ_MOD_ABS_BS) return isnan(m) ? f : mod_amp(f, m);
  else if constexpr (MODE == PROJ_MOD_ABS) return mod_amp(f, m);
  else if constexpr (MODE == PROJ_PT_ACCURATE) return mod_accurate(cplx<T>{f.x * pa.bsnorm, f.y * pa.bsnorm}, m, pa.scale);
  else if constexpr (MODE == PROJ_PT_NOISE)
    return mod_noise(cplx<T>{f.x * pa.bsnorm, f.y * pa.bsnorm}, m, pa.noiseLevel, pa.scale, pa.rcolor);
  else {  // runtime mode, same arithmetic as the compile-time variants
    if (pa.mode == PROJ_MOD_NOISE) return mod_noise(f, m, pa.noiseLevel, pa.scale, pa.rcolor);
    if (pa.mode == PROJ_MOD_ABS_BS && isnan(m)) return f;
    return mod_amp(f, m);
  }
}

// Streaming store of one complex value that bypasses the L2 (the row pass
// writes whole rows, so the lines are complete and nothing is left dirty
// for the end-of-kernel write-back)
template <class T>
__device__ __forceinline__ void store_stream(cplx<T>* p, cplx<T> v) {
  typedef T vec2 __attribute__((ext_vector_type(2)));
  vec2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<vec2*>(p));
}

// Streaming (non-temporal) loads
template <class T>
__device__ __forceinline__ cplx<T> load_stream(const cplx<T>* p) {
  typedef T vec2 __attribute__((ext_vector_type(2)));
  const vec2 w = __builtin_nontemporal_load(reinterpret_cast<const vec2*>(p));
  return {w.x, w.y};
}
template <class T>
__device__ __forceinline__ T load_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Neighbouring column workgroups read pieces of the same cache lines
// (halves of one T1 piece; the two 32-byte measurement pieces of a line), so
// they must run on the same XCD: give each XCD a contiguous run of tiles.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  if (nb & 7) return bid;
  return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// ---------------------------------------------------------------------------
// column pass body: columns [tile << lsh, (tile+1) << lsh).  MODE/RF = -1:
// projection mode and R-factor chosen at run time.  TILED: the transformed
// input (ops that start with the forward transform) is T1 in its column-tile
// layout.  TILED_I: the packed measurement is in its column-tile layout.
// ---------------------------------------------------------------------------
template <class T, int N, int OP, int MODE, int RF, bool TILED = false, bool TILED_I = false, int V = 0,
          bool T2L = false>
__device__ __forceinline__ void col_body(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int cols,
                                         const cplx<T>* __restrict__ tw, const T* __restrict__ I,
                                         const ProjArgs<T>& pa, T gsig, int tile, int lsh, unsigned char* smem) {
  using L = LineOf<T, N, V>;
  const int pitch = lds_pitch<T, N, V>(lsh);
  const int alt = pitch << lsh;
  int ph = 0;
  constexpr int E = L::E;
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
  const int c = threadIdx.x & ((1 << lsh) - 1);
  const int t = threadIdx.x >> lsh;
  const int col = (tile << lsh) + c;
  const bool ok = col < cols;
  cplx<T>* my = lds + c * pitch;
  constexpr bool from_f = (OP == COL_INV || OP == COL_PROJ_INV || OP == COL_GAUSS_INV);
  constexpr bool fwd_only = (OP == COL_FWD || OP == COL_FWD_T1);
  constexpr bool proj = (OP == COL_FWD_PROJ_INV || OP == COL_PROJ_INV);
  cplx<T> v[E];
  T iv[proj ? E : 1];

  // Lanes past the last column load the last column (their results are
  // dropped), so the loads carry no branches: behind a branch the compiler
  // waits for every load issued so far (I included) before the first
  // transform.
  const int lc = ok ? col : cols - 1;
  if constexpr (from_f) {
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = in[(size_t)L::dout_pos(i, t) * cols + lc];
  } else if constexpr (TILED) {
    // T1 is read once per iteration: streaming loads (measured faster than
    // plain ones here)
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = load_stream(&in[t1_index(L::din_pos(i, t), lc, N)]);
  } else {
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = in[(size_t)L::din_pos(i, t) * cols + lc];
  }
  if constexpr (proj) {
#pragma unroll
    for (int i = 0; i < E; i++)
      iv[i] = I[TILED_I ? i_index(L::dout_pos(i, t), lc, N) : (size_t)L::dout_pos(i, t) * cols + lc];
  }
  // table twiddles of both transforms, loaded behind the data
  typename L::TwF twf;
  typename L::TwI twi;
  if constexpr (!from_f) twf = L::template tw_load<-1>(tw, t);
  if constexpr (!fwd_only) twi = L::template tw_load<+1>(tw, t);
  if constexpr (!from_f) L::forward(v, my, alt, ph, twf, t);

  if constexpr (fwd_only) {
    if (ok) {
#pragma unroll
      for (int i = 0; i < E; i++) out[(size_t)L::dout_pos(i, t) * cols + col] = v[i];
    }
    return;
  }

  // R-factor sums (|F| against the measured amplitude), reduced after the
  // stores so the workgroup barrier does not stall the transforms
  const bool rf = proj && (RF < 0 ? pa.rf != nullptr : RF > 0);
  T rnum = 0, rden = 0;
  if constexpr (proj) {
    const bool amp = packed_amplitude<MODE>(pa.mode);
#pragma unroll
    for (int i = 0; i < E; i++) {
      if (rf && ok && !isnan(iv[i])) {
        const T s = amp ? iv[i] : (iv[i] > (T)0 ? dsqrt(iv[i]) : (T)0);
        rnum += fabs(dsqrt(v[i].x * v[i].x + v[i].y * v[i].y) - s);
        rden += s;
      }
      if (ok) v[i] = project_m<T, MODE>(v[i], iv[i], pa);
      if constexpr (MODE < 0) {
        // patternWave after the projection, for the PRTF (cdi.cc:362,377)
        if (pa.fsave && ok) ((cplx<T>*)pa.fsave)[(size_t)L::dout_pos(i, t) * cols + col] = v[i];
      }
    }
  }
  if constexpr (OP == COL_FWD_GAUSS_INV || OP == COL_GAUSS_INV) {
    // applyGaussMult(isFreq=1), cudaConfig.cu:380-395, wrapped coordinates
    const T yrel = (T)(col >= cols / 2 ? col - cols : col);
#pragma unroll
    for (int i = 0; i < E; i++) {
      const int r = L::dout_pos(i, t);
      const T xrel = (T)(r >= N / 2 ? r - N : r);
      const T f = (T)exp(-(xrel * xrel + yrel * yrel) / ((T)2 * gsig * gsig));
      v[i].x *= f;
      v[i].y *= f;
    }
  }

  L::inverse(v, my, alt, ph, twi, t);
  if (ok) {
    // Streaming stores of the output beat leaving its partial lines in L2
    // for the end-of-kernel write-back for columns up to 1024 points (single
    // frame with tiled inputs, per iteration: 768^2 21.5 -> 20.6 us, 1024^2
    // 21.0 -> 20.1 us; batched 16-column tiles: 85.5 -> 84.2 ms) but not for
    // 2048 (57.3 -> 67.6 us).  Before the single-frame inputs were tiled they
    // made that column pass 20% slower.
    if constexpr (N <= 1024) {
#pragma unroll
      for (int i = 0; i < E; i++)
        store_stream(&out[T2L ? t2_index(L::din_pos(i, t), col, cols) : (size_t)L::din_pos(i, t) * cols + col], v[i]);
    } else {
#pragma unroll
      for (int i = 0; i < E; i++)
        out[T2L ? t2_index(L::din_pos(i, t), col, cols) : (size_t)L::din_pos(i, t) * cols + col] = v[i];
    }
  }
  if (rf) {
    __shared__ double sh[32];
    block_atomic_sum2(rnum, rden, pa.rf + (size_t)(blockIdx.x & (kRfSlots - 1)) * pa.rfs, sh);
  }
}

// ---------------------------------------------------------------------------
// row pass body: rows [tile << lsh, (tile+1) << lsh).  ALGO: support-update
// rule of ROW_INV_UPDATE_FWD (enum Algorithm).  TILED: the update ops store
// T1 in its column-tile layout.
// ---------------------------------------------------------------------------
template <class T, int N, int OP, int ALGO, bool TILED = false, int V = 0, bool T2L = false, bool GPLAIN = false,
          bool SMASK = false>
__device__ __forceinline__ void row_body(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int rows,
                                         const cplx<T>* __restrict__ tw, cplx<T>* __restrict__ g,
                                         const T* __restrict__ spt, const UpdArgs<T>& ua, T scale,
                                         T* __restrict__ sout, const ReduceSlots& rs, int tile, int lsh,
                                         unsigned char* smem) {
  using L = LineOf<T, N, V>;
  constexpr int TL = LineCfgT<T, N, V>::TL;
  const int pitch = lds_pitch<T, N, V>(lsh);
  const int alt = pitch << lsh;
  int ph = 0;
  constexpr int E = L::E;
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
  const int t = threadIdx.x % TL;
  const int rr = threadIdx.x / TL;
  const int row = (tile << lsh) + rr;
  const bool ok = row < rows;
  cplx<T>* my = lds + rr * pitch;
  const size_t base = (size_t)row * N;
  // loads of lanes past the last row read the last row (results dropped):
  // branch-free loads, as in the column pass
  const int lrow = ok ? row : rows - 1;
  const size_t lbase = (size_t)lrow * N;
  constexpr bool prtf = OP == ROW_INV_UPDATE_FWD_PRTF;
  constexpr bool upd = OP == ROW_INV_UPDATE_FWD || prtf;
  cplx<T> v[E];

  if constexpr (OP == ROW_FWD) {
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = in[lbase + L::din_pos(i, t)];
    const typename L::TwF twf = L::template tw_load<-1>(tw, t);
    L::forward(v, my, alt, ph, twf, t);
    if (ok) {
#pragma unroll
      for (int i = 0; i < E; i++) out[base + L::dout_pos(i, t)] = v[i];
    }
    return;
  } else {
    cplx<T> gv[upd ? E : 1];
    // SMASK: the support test comes as one word of bits per lane instead of
    // E floats (ua.smask, built by launch_support_mask)
    static_assert(!SMASK || E <= 32, "one mask word per lane");
    T sv[upd && !SMASK ? E : 1];
    unsigned mw = 0;
    cplx<T>* gamma0 = (cplx<T>*)ua.gamma0;
    // Plain loads of g and S with streaming loads of T2, the fastest
    // combination for 1024-point rows (per iteration at 1024^2: 20.1 us with
    // g, S and T2 all plain, 19.0 with all streamed, 17.9 with only T2
    // streamed and T1 streamed in the column pass; at 2048^2 streaming T2
    // cost 3.6 us, at smaller sizes it changed nothing).
    constexpr bool ld_nt = TILED && upd && N == 1024;
    if constexpr (ld_nt) {
#pragma unroll
      for (int i = 0; i < E; i++)
        v[i] = load_stream(&in[T2L ? t2_index(lrow, L::dout_pos(i, t), N) : lbase + L::dout_pos(i, t)]);
#pragma unroll
      for (int i = 0; i < E; i++) gv[i] = g[lbase + L::din_pos(i, t)];
      if constexpr (SMASK) {
        mw = ua.smask[(size_t)lrow * TL + t];
      } else {
#pragma unroll
        for (int i = 0; i < E; i++) sv[i] = spt[lbase + L::din_pos(i, t)];
      }
    } else {
#pragma unroll
      for (int i = 0; i < E; i++)
        v[i] = in[T2L ? t2_index(lrow, L::dout_pos(i, t), N) : lbase + L::dout_pos(i, t)];
      if constexpr (upd) {
#pragma unroll
        for (int i = 0; i < E; i++) gv[i] = g[lbase + L::din_pos(i, t)];
        if constexpr (SMASK) {
          mw = ua.smask[(size_t)lrow * TL + t];
        } else {
#pragma unroll
          for (int i = 0; i < E; i++) sv[i] = spt[lbase + L::din_pos(i, t)];
        }
      }
    }
    const typename L::TwI twi = L::template tw_load<+1>(tw, t);
    L::inverse(v, my, alt, ph, twi, t);

    if constexpr (OP == ROW_INV) {
      if (ok) {
#pragma unroll
        for (int i = 0; i < E; i++) out[base + L::din_pos(i, t)] = {v[i].x * scale, v[i].y * scale};
      }
      return;
    }
    if constexpr (OP == ROW_INV_MOD2) {
      T m = 0;
      double s = 0;
#pragma unroll
      for (int i = 0; i < E; i++) {
        const T a = v[i].x * v[i].x + v[i].y * v[i].y;
        if (ok) {
          sout[base + L::din_pos(i, t)] = a;
          m = a > m ? a : m;
          s += (double)a;
        }
      }
      __shared__ double sh[16];
      block_atomic_max(m, rs.maxbits, sh);
      block_partial_sum(s, rs, sh);
      return;
    }
    if constexpr (upd) {
      const typename L::TwF twf = L::template tw_load<-1>(tw, t);
      const T thr = (T)ua.scal[0];
      double are = 0, aim = 0;  // PRTF alignment sums (cdi.cc:358-360,373-375)
#pragma unroll
      for (int i = 0; i < E; i++) {
        const cplx<T> p = {v[i].x * ua.invN, v[i].y * ua.invN};
        bool in;
        if constexpr (SMASK) in = (mw >> i) & 1u;
        else in = sv[i] > thr;
        if constexpr (prtf) {
          if (ua.prtf == 1) {  // gamma0 = P_S(g'), sum gamma0^2
            const cplx<T> q = in && ok ? p : cplx<T>{(T)0, (T)0};
            if (ok) gamma0[base + L::din_pos(i, t)] = q;
            are += (double)q.x * q.x - (double)q.y * q.y;
            aim += 2.0 * q.x * q.y;
          } else if (in && ok) {  // sum P_S(g') conj(gamma0); loaded here, not preloaded,
                                  // to keep the pass's registers within a wave's budget
            const cplx<T> g0 = gamma0[base + L::din_pos(i, t)];
            are += (double)p.x * g0.x + (double)p.y * g0.y;
            aim += (double)p.y * g0.x - (double)p.x * g0.y;
          }
        }
        v[i] = support_update(gv[i], p, in, ALGO >= 0 ? ALGO : ua.algo, ua.beta, ua.relax);
        if (ok) {
          // GPLAIN: plain g stores measured faster in the single-frame loop
          // up to 1024^2; larger frames and batches stream it
          if constexpr (GPLAIN) g[base + L::din_pos(i, t)] = v[i];
          else store_stream(&g[base + L::din_pos(i, t)], v[i]);
        }
      }
      L::forward(v, my, alt, ph, twf, t);
      if (ok) {
#pragma unroll
        for (int i = 0; i < E; i++)
          store_stream(&out[TILED ? t1_index(row, L::dout_pos(i, t), rows) : base + L::dout_pos(i, t)], v[i]);
      }
      if constexpr (prtf) {
        __shared__ double sh[16];
        block_store_sum(are, rs.partials + 2 * blockIdx.x, sh);
        block_store_sum(aim, rs.partials + 2 * blockIdx.x + 1, sh);
        if (blockIdx.x == 0 && threadIdx.x == 0) *rs.nparts = (int)gridDim.x;
      }
    }
  }
}

}  // namespace cdx
