// LDS-staged Stockham FFT building blocks shared by the plain per-axis transforms
// (fft_lds.hip) and the fused pad / crop / Toeplitz transforms (fft_fused.hip).
//
// Stockham autosort passes (Govindaraju et al., "High Performance Discrete Fourier
// Transforms on Graphics Processors", SC 2008): for a pass of radix r after
// sub-transform size Ns, butterfly j reads the r values x[j + q L/r], multiplies
// value q by exp(-2 pi i (j mod Ns) q / (Ns r)), takes their r-point DFT and writes
// value q to y[(j / Ns) Ns r + (j mod Ns) + q Ns]. NT transforms of length L sit in
// LDS as [L][NTP] complex values; each pass reads its inputs into registers,
// synchronises, and overwrites the same buffer.
#pragma once
#include <type_traits>

#include "common.h"

namespace mra {
namespace fft {

template <int... R> struct Radices {};

template <int L> struct Plan;  // radix list, product == L
template <> struct Plan<64> { using type = Radices<8, 8>; };
template <> struct Plan<80> { using type = Radices<5, 4, 4>; };
template <> struct Plan<96> { using type = Radices<8, 4, 3>; };
template <> struct Plan<120> { using type = Radices<8, 5, 3>; };
template <> struct Plan<128> { using type = Radices<8, 4, 4>; };
template <> struct Plan<160> { using type = Radices<8, 5, 4>; };
template <> struct Plan<192> { using type = Radices<8, 8, 3>; };
template <> struct Plan<240> { using type = Radices<8, 6, 5>; };
template <> struct Plan<256> { using type = Radices<8, 8, 4>; };
template <> struct Plan<320> { using type = Radices<8, 8, 5>; };
template <> struct Plan<384> { using type = Radices<8, 8, 6>; };
template <> struct Plan<480> { using type = Radices<8, 6, 5, 2>; };
template <> struct Plan<512> { using type = Radices<8, 8, 8>; };
template <> struct Plan<640> { using type = Radices<8, 8, 5, 2>; };

constexpr int kThreadsFFT = 256;

// transforms per block: a power of two, 8..64 (column loads are whole 64/128-byte
// segments); about 2048 points per block along the contiguous axis and 4096 along the
// others, the measured optimum at L = 160
constexpr int pow2_floor(int x) { return x >= 64 ? 64 : x >= 32 ? 32 : x >= 16 ? 16 : 8; }
template <int L, bool ROW> struct Shape {
  static constexpr int NT = pow2_floor((ROW ? 2048 : 4096) / L);
  static constexpr int NTP = NT + 1;  // LDS row pitch (odd: spreads rows over banks)
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-register r-point DFT (forward sign), v[q] -> sum_p v[p] e^{-2 pi i p q / r}.
template <int R>
__device__ __forceinline__ void dft(float2 (&v)[R]) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const float2 a0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 a1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 b0 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 b1 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    // b1 * (-i)
    const float2 b1m = make_float2(b1.y, -b1.x);
    v[0] = make_float2(a0.x + b0.x, a0.y + b0.y);
    v[2] = make_float2(a0.x - b0.x, a0.y - b0.y);
    v[1] = make_float2(a1.x + b1m.x, a1.y + b1m.y);
    v[3] = make_float2(a1.x - b1m.x, a1.y - b1m.y);
  } else if constexpr (R == 8) {
    // radix-2 split: even / odd halves as 4-point DFTs
    float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft<4>(e);
    dft<4>(o);
    constexpr float h = 0.70710678118654752f;
    const float2 w[4] = {make_float2(1.f, 0.f), make_float2(h, -h), make_float2(0.f, -1.f), make_float2(-h, -h)};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 t = cmul(o[q], w[q]);
      v[q] = make_float2(e[q].x + t.x, e[q].y + t.y);
      v[q + 4] = make_float2(e[q].x - t.x, e[q].y - t.y);
    }
  } else if constexpr (R == 6) {
    // 6 = 2 x 3: two 3-point DFTs of the even / odd samples
    float2 e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
    dft<3>(e);
    dft<3>(o);
    constexpr float c = 0.5f, s = 0.86602540378443865f;
    const float2 w[3] = {make_float2(1.f, 0.f), make_float2(c, -s), make_float2(-c, -s)};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float2 t = cmul(o[q], w[q]);
      v[q] = make_float2(e[q].x + t.x, e[q].y + t.y);
      v[q + 3] = make_float2(e[q].x - t.x, e[q].y - t.y);
    }
  } else if constexpr (R == 3) {
    constexpr float s = 0.86602540378443865f;
    const float2 a = v[0], b = v[1], cc = v[2];
    const float2 t = make_float2(b.x + cc.x, b.y + cc.y);
    const float2 d = make_float2(b.x - cc.x, b.y - cc.y);
    v[0] = make_float2(a.x + t.x, a.y + t.y);
    const float2 m = make_float2(a.x - 0.5f * t.x, a.y - 0.5f * t.y);
    // -i s d
    v[1] = make_float2(m.x + s * d.y, m.y - s * d.x);
    v[2] = make_float2(m.x - s * d.y, m.y + s * d.x);
  } else if constexpr (R == 5) {
    constexpr float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    constexpr float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    const float2 a = v[0];
    const float2 t1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y);
    const float2 t2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y);
    const float2 d1 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y);
    const float2 d2 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
    v[0] = make_float2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
    const float2 m1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const float2 m2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    // n1 = s1 d1 + s2 d2 ; n2 = s2 d1 - s1 d2 ; outputs m -/+ i n
    const float2 n1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
    const float2 n2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
    v[1] = make_float2(m1.x + n1.y, m1.y - n1.x);
    v[4] = make_float2(m1.x - n1.y, m1.y + n1.x);
    v[2] = make_float2(m2.x + n2.y, m2.y - n2.x);
    v[3] = make_float2(m2.x - n2.y, m2.y + n2.x);
  }
}

// One Stockham pass of radix R after sub-transform size NS over the NT transforms in LDS.
template <int L, class S, int R, int NS>
__device__ __forceinline__ void pass(float2* buf, const float2* __restrict__ tw) {
  constexpr int NB = (L / R) * S::NT;                     // butterflies in the block
  constexpr int MB = (NB + kThreadsFFT - 1) / kThreadsFFT;  // per thread
  constexpr int TSTEP = L / (NS * R);                     // twiddle table step
  float2 v[MB][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = tid + m * kThreadsFFT;
    if (b < NB) {
      const int c = b % S::NT, j = b / S::NT;
      const int k = j % NS;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const float2 x = buf[(j + q * (L / R)) * S::NTP + c];
        v[m][q] = q == 0 ? x : cmul(x, tw[(k * q * TSTEP) % L]);
      }
      dft<R>(v[m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = tid + m * kThreadsFFT;
    if (b < NB) {
      const int c = b % S::NT, j = b / S::NT;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[(base + q * NS) * S::NTP + c] = v[m][q];
    }
  }
  __syncthreads();
}

template <int L, class S, int NS>
__device__ __forceinline__ void run_passes(float2*, const float2*, Radices<>) {}

template <int L, class S, int NS, int R, int... Rest>
__device__ __forceinline__ void run_passes(float2* buf, const float2* tw, Radices<R, Rest...>) {
  pass<L, S, R, NS>(buf, tw);
  run_passes<L, S, NS * R>(buf, tw, Radices<Rest...>{});
}

// Forward transform of the NT columns of buf (S::NT transforms of length L).
template <int L, class S>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* tw) {
  run_passes<L, S, 1>(buf, tw, typename Plan<L>::type{});
}

// Whether a length has a compiled radix plan.
__host__ __device__ inline bool plan_supported(int L) {
  switch (L) {
    case 64: case 80: case 96: case 120: case 128: case 160: case 192: case 240: case 256: case 320:
    case 384: case 480: case 512: case 640: return true;
    default: return false;
  }
}

// Calls f(integral_constant<L>) for a supported runtime length; returns f's status or
// fails for other lengths.
template <class F>
int with_length(int L, F&& f) {
  switch (L) {
#define MRA_FFT_LEN(LL) \
  case LL: return f(std::integral_constant<int, LL>{});
    MRA_FFT_LEN(64) MRA_FFT_LEN(80) MRA_FFT_LEN(96) MRA_FFT_LEN(120) MRA_FFT_LEN(128) MRA_FFT_LEN(160)
    MRA_FFT_LEN(192) MRA_FFT_LEN(240) MRA_FFT_LEN(256) MRA_FFT_LEN(320) MRA_FFT_LEN(384) MRA_FFT_LEN(480)
    MRA_FFT_LEN(512) MRA_FFT_LEN(640)
#undef MRA_FFT_LEN
    default: return -1;
  }
}

}  // namespace fft
}  // namespace mra


