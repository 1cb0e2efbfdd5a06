// fft_device.hpp -- Stockham autosort FFT on LDS for gfx950 (FP32 complex).
//
// A workgroup transforms LINES lines of length N that sit in LDS.  Element i
// of line l lives at buf[l*LS + i*IS].  Every stage reads R elements per
// butterfly into registers, applies twiddles from a table of N-th roots,
// runs a radix-R DFT and writes the results in Stockham (natural-order)
// positions.  All indices and loop bounds are compile-time, so the per-thread
// register arrays stay in VGPRs.
//
// Radices: 8, 4, 2 and 3 (meshes 3*2^k need radix 3).  Forward sign
// e^{-2 pi i jk/N}, unnormalised, the convention of the reference's rustfft
// transforms (crates/gadget-ng-pm/src/fft_poisson.rs:733-813).
#pragma once
#include <hip/hip_runtime.h>

namespace gng {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct Dft<3, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    const float s3 = 0.866025403784438646763723170752936183f;  // sqrt(3)/2
    float2 a = v[0], b = v[1], c = v[2];
    float2 t1 = cadd(b, c);
    float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    float2 d = csub(b, c);
    // t3 = (b - c) * (-+ i sqrt3/2)
    float2 t3 = INV ? make_float2(-s3 * d.y, s3 * d.x) : make_float2(s3 * d.y, -s3 * d.x);
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

template <bool INV>
struct Dft<4, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[1] = cadd(b, d);
    v[2] = csub(a, c);
    v[3] = csub(b, d);
  }
};

template <bool INV>
struct Dft<8, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    const float h = 0.707106781186547524400844362104849039f;
    // split into even/odd halves: a_r = v_r + v_{r+4}, b_r = (v_r - v_{r+4}) w8^r
    float2 a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = cadd(v[r], v[r + 4]);
      b[r] = csub(v[r], v[r + 4]);
    }
    // w8 = e^{-+ i pi/4}
    b[1] = INV ? make_float2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y))
               : make_float2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
    b[2] = mul_mi<INV>(b[2]);
    b[3] = INV ? make_float2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y))
               : make_float2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
    Dft<4, INV>::run(a);
    Dft<4, INV>::run(b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  }
};

__host__ __device__ constexpr int fft_pick_radix(int n, int ns) {
  return ((n / ns) % 8 == 0) ? 8 : ((n / ns) % 4 == 0) ? 4 : ((n / ns) % 2 == 0) ? 2 : ((n / ns) % 3 == 0) ? 3 : 1;
}

__host__ __device__ constexpr bool fft_size_ok(int n) {
  int r = n;
  while (r % 2 == 0) r /= 2;
  while (r % 3 == 0) r /= 3;
  return r == 1 && n >= 2;
}

// One Stockham stage.  tw holds TWN-th roots (tw[m] = e^{-2 pi i m / TWN})
// and TWN is a multiple of N; the stage uses stride TWN/N into it.
template <int N, int R, int NS, bool INV, int LINES, int T, int LS, int IS, bool LINE_FAST, int TWN>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* __restrict__ tw, int tid) {
  constexpr int NB = N / R;
  constexpr int TOTAL = LINES * NB;
  constexpr int PER = (TOTAL + T - 1) / T;
  constexpr int TSTEP = (N / (NS * R)) * (TWN / N);
  float2 v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * T;
    if ((TOTAL % T == 0) || b < TOTAL) {
      const int l = LINE_FAST ? (b % LINES) : (b / NB);
      const int j = LINE_FAST ? (b / LINES) : (b % NB);
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = buf[l * LS + (j + r * NB) * IS];
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * TSTEP];
          if (INV) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Dft<R, INV>::run(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * T;
    if ((TOTAL % T == 0) || b < TOTAL) {
      const int l = LINE_FAST ? (b % LINES) : (b / NB);
      const int j = LINE_FAST ? (b / LINES) : (b % NB);
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[l * LS + (base + r * NS) * IS] = v[p][r];
    }
  }
  __syncthreads();
}

template <int N, int NS, bool INV, int LINES, int T, int LS, int IS, bool LINE_FAST, int TWN>
struct FftStages {
  __device__ __forceinline__ static void run(float2* buf, const float2* __restrict__ tw, int tid) {
    if constexpr (NS < N) {
      constexpr int R = fft_pick_radix(N, NS);
      static_assert(R > 1, "unsupported FFT length");
      fft_stage<N, R, NS, INV, LINES, T, LS, IS, LINE_FAST, TWN>(buf, tw, tid);
      FftStages<N, NS * R, INV, LINES, T, LS, IS, LINE_FAST, TWN>::run(buf, tw, tid);
    }
  }
};

// Transform LINES lines in LDS.  Caller must __syncthreads() before (data
// written) -- the function ends with a barrier.
template <int N, bool INV, int LINES, int T, int LS, int IS, bool LINE_FAST, int TWN>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw, int tid) {
  static_assert(fft_size_ok(N), "FFT length must be 2^a 3^b");
  FftStages<N, 1, INV, LINES, T, LS, IS, LINE_FAST, TWN>::run(buf, tw, tid);
}

// ---------------------------------------------------------------------------
// Register-resident transform
//
// The radix plan of length N is the greedy sequence of fft_pick_radix (8s
// first); REV runs it backwards.  Stage 0 takes its inputs from registers and
// stage S-1 leaves its outputs in registers, both in the "edge layout" of
// their radix R: butterfly slot b of the thread is (line l, j) and holds the
// R elements at indices j + r*N/R, r < R.  Middle stages go through LDS.
// Because the last stage's output layout equals the input layout of a first
// stage with the same radix, a transform on the reversed plan can consume
// another transform's outputs straight from registers.
// ---------------------------------------------------------------------------

__host__ __device__ constexpr int plan_len(int n) {
  int s = 0, ns = 1;
  while (ns < n) {
    ns *= fft_pick_radix(n, ns);
    ++s;
  }
  return s;
}
__host__ __device__ constexpr int plan_radix(int n, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= fft_pick_radix(n, ns);
  return fft_pick_radix(n, ns);
}
template <int N, bool REV>
__host__ __device__ constexpr int radix_at(int s) {
  return REV ? plan_radix(N, plan_len(N) - 1 - s) : plan_radix(N, s);
}
template <int N, bool REV>
__host__ __device__ constexpr int ns_at(int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= radix_at<N, REV>(i);
  return ns;
}

// Butterfly slots of a radix-R stage over LINES lines shared by T threads.
template <int N, int R, int LINES, int T>
struct Slots {
  static constexpr int NB = N / R;
  static constexpr int TOTAL = LINES * NB;
  static constexpr int PER = (TOTAL + T - 1) / T;
};

template <int N, int R, int LINES, bool LF>
__device__ __forceinline__ void slot_lj(int b, int& l, int& j) {
  constexpr int NB = N / R;
  l = LF ? b % LINES : b / NB;
  j = LF ? b / LINES : b % NB;
}

// Twiddle then radix-R DFT of one butterfly of stage NS.
template <int N, int R, int NS, bool INV, int TWN>
__device__ __forceinline__ void butterfly(float2* v, int j, const float2* __restrict__ tw) {
  if (NS > 1) {
    constexpr int TSTEP = (N / (NS * R)) * (TWN / N);
    const int k = j % NS;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[r * k * TSTEP];
      if (INV) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  Dft<R, INV>::run(v);
}

// Edge-layout element loop: f(p, r, line, index, valid) for every slot.
template <int N, int R, int LINES, int T, bool LF, typename F>
__device__ __forceinline__ void for_edge(int tid, F f) {
  using S = Slots<N, R, LINES, T>;
#pragma unroll
  for (int p = 0; p < S::PER; ++p) {
    const int b = tid + p * T;
    const bool ok = (S::TOTAL % T == 0) || b < S::TOTAL;
    int l = 0, j = 0;
    if (ok) slot_lj<N, R, LINES, LF>(b, l, j);
#pragma unroll
    for (int r = 0; r < R; ++r) f(p, r, l, j + r * S::NB, ok);
  }
}

// v is float2[Slots<N, R, LINES, T>::PER][R].
template <int N, int R, int NS, bool INV, int LINES, int T, int LS, int IS, bool LF, int TWN, typename V>
__device__ __forceinline__ void stage_regs_to_lds(V& v, float2* buf, const float2* __restrict__ tw, int tid) {
  using S = Slots<N, R, LINES, T>;
#pragma unroll
  for (int p = 0; p < S::PER; ++p) {
    const int b = tid + p * T;
    if ((S::TOTAL % T == 0) || b < S::TOTAL) {
      int l, j;
      slot_lj<N, R, LINES, LF>(b, l, j);
      butterfly<N, R, NS, INV, TWN>(v[p], j, tw);
      const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[l * LS + (base + r * NS) * IS] = v[p][r];
    }
  }
}

template <int N, int R, int NS, bool INV, int LINES, int T, int LS, int IS, bool LF, int TWN, typename V>
__device__ __forceinline__ void stage_lds_to_regs(V& v, const float2* buf, const float2* __restrict__ tw, int tid) {
  static_assert(NS * R == N, "only the last stage leaves its outputs in registers");
  using S = Slots<N, R, LINES, T>;
#pragma unroll
  for (int p = 0; p < S::PER; ++p) {
    const int b = tid + p * T;
    if ((S::TOTAL % T == 0) || b < S::TOTAL) {
      int l, j;
      slot_lj<N, R, LINES, LF>(b, l, j);
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = buf[l * LS + (j + r * S::NB) * IS];
      butterfly<N, R, NS, INV, TWN>(v[p], j, tw);
    }
  }
}

template <int N, bool REV, int S, bool INV, int LINES, int T, int LS, int IS, bool LF, int TWN>
struct MiddleStages {
  __device__ __forceinline__ static void run(float2* buf, const float2* __restrict__ tw, int tid) {
    if constexpr (S < plan_len(N) - 1) {
      fft_stage<N, radix_at<N, REV>(S), ns_at<N, REV>(S), INV, LINES, T, LS, IS, LF, TWN>(buf, tw, tid);
      MiddleStages<N, REV, S + 1, INV, LINES, T, LS, IS, LF, TWN>::run(buf, tw, tid);
    }
  }
};

// Register-resident transform of LINES lines: vin in the edge layout of the
// plan's first radix (consumed), vout in the edge layout of its last radix.
// Uses buf as exchange space; begins with a barrier, so the caller need not
// synchronise before it, and leaves LDS readable only after a barrier.
template <int N, bool INV, bool REV, int LINES, int T, int LS, int IS, bool LF, int TWN>
struct RegFft {
  static constexpr int S = plan_len(N);
  static constexpr int R0 = radix_at<N, REV>(0);
  static constexpr int RL = radix_at<N, REV>(S - 1);
  static constexpr int PER0 = Slots<N, R0, LINES, T>::PER;
  static constexpr int PERL = Slots<N, RL, LINES, T>::PER;
  static_assert(fft_size_ok(N), "FFT length must be 2^a 3^b");

  __device__ __forceinline__ static void run(float2 (&vin)[PER0][R0], float2 (&vout)[PERL][RL], float2* buf,
                                             const float2* __restrict__ tw, int tid) {
    if constexpr (S == 1) {
      using Sl = Slots<N, R0, LINES, T>;
#pragma unroll
      for (int p = 0; p < PER0; ++p) {
        const int b = tid + p * T;
        if ((Sl::TOTAL % T == 0) || b < Sl::TOTAL) {
          Dft<R0, INV>::run(vin[p]);
#pragma unroll
          for (int r = 0; r < R0; ++r) vout[p][r] = vin[p][r];
        }
      }
    } else {
      __syncthreads();
      stage_regs_to_lds<N, R0, 1, INV, LINES, T, LS, IS, LF, TWN>(vin, buf, tw, tid);
      __syncthreads();
      MiddleStages<N, REV, 1, INV, LINES, T, LS, IS, LF, TWN>::run(buf, tw, tid);
      stage_lds_to_regs<N, RL, ns_at<N, REV>(S - 1), INV, LINES, T, LS, IS, LF, TWN>(vout, buf, tw, tid);
    }
  }
};

}  // namespace gng
