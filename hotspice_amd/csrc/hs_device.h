// Device helpers shared by the gfx950 kernels of libhotspice_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hotspice_hip.h"

#define HS_INLINE __device__ __forceinline__

// ---- Philox4x32-10 (Salmon et al., SC'11) ----------------------------------------------------
HS_INLINE uint4 hs_philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

// 53-bit integer of stream (seed, step, tag) at `index`: two 32-bit words of Philox call index>>1.
HS_INLINE uint64_t hs_bits53(uint64_t seed, uint64_t step, uint32_t tag, uint64_t index) {
  const uint4 w = hs_philox(make_uint4((uint32_t)(index >> 1), (uint32_t)step, (uint32_t)(step >> 32), tag),
                            (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint32_t a = (index & 1) ? w.z : w.x;
  const uint32_t b = (index & 1) ? w.w : w.y;
  return ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
}

// [0,1) (open_zero=0) or (0,1] (open_zero=1) uniform.
HS_INLINE double hs_u53(uint64_t bits, int open_zero) {
  return (double)(bits + (uint64_t)open_zero) * 0x1p-53;
}

// -ln(u) for u = (bits+1)*2^-53 in (0,1], FP32 with full relative precision at both ends.
HS_INLINE float hs_exp1_f32(uint64_t bits) {
  if (bits < (1ull << 52)) return -logf((float)(bits + 1) * 0x1p-53f);
  const float w = (float)((1ull << 53) - 1 - bits) * 0x1p-53f;  // 1-u, exact integer then rounded
  return -log1pf(-w);
}

// Same in FP64.
HS_INLINE double hs_exp1_f64(uint64_t bits) {
  if (bits < (1ull << 52)) return -log((double)(bits + 1) * 0x1p-53);
  return -log1p(-(double)((1ull << 53) - 1 - bits) * 0x1p-53);
}

// Neel waiting time of a cell in FP64 from its compensated fields (value - compensation):
// tau*exp(bmin*beta)/f0 == -ln(u)*exp(b*beta)/f0 (core.py:698-700), or with two channels
// -ln(u)/(exp(-b1 beta)+exp(-b2 beta))/(f0/2) (core.py:686-693).
HS_INLINE double hs_neel_dt(double e, double ep, double eb, double beta, double m_perp, int parabolic, bool perp,
                            uint64_t bits, double f0) {
  const double dE = -2.0 * e, El = e, x = hs_exp1_f64(bits);
  if (perp) {
    const double epf = m_perp * ep;
    const double b1 = fmax(dE, (eb + epf) - El), b2 = fmax(dE, (eb - epf) - El);
    const double lo = fmin(b1, b2), hi = fmax(b1, b2);
    return x * exp(lo * beta) / (1.0 + exp(-(hi - lo) * beta)) / (f0 / 2.0);
  }
  double b;
  if (!parabolic) {
    b = fmax(dE, eb - El);
  } else {
    const double q = dE * (0.25 / eb) + 1.0;
    b = dE > 4.0 * eb ? dE : fmin(eb * (q * q), dE + 4.0 * eb);
  }
  return x * exp(b * beta) / f0;
}

// ---- Kahan-compensated FP32 accumulation (value = s - c) ------------------------------------
HS_INLINE void hs_kahan_add(float &s, float &c, float x) {
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// utils.lower_than (hotspice/utils.py:90-92): x < y and not isclose(x, y, atol=0, rtol=1e-10)
// with NumPy's isclose semantics: |x-y| <= rtol*|y| for finite values, x == y for infinities.
HS_INLINE bool hs_lower_than(double x, double y) {
  if (!(x < y)) return false;
  bool close;
  if (isinf(x) || isinf(y)) close = (x == y);
  else close = fabs(x - y) <= 1e-10 * fabs(y);
  return !close;
}

// Python's modulo (result has the sign of the divisor).
HS_INLINE int hs_pmod(int a, int b) {
  const int r = a % b;
  return r < 0 ? r + b : r;
}

HS_INLINE int hs_site(const hs_instance_t &I, int y, int x) { return (y % I.uy) * I.ux + (x % I.ux); }

// ---- wavefront / block reductions (wave64) ---------------------------------------------------
// (key, index) argmin, lowest index on ties; NaN keys were mapped to +inf by the caller.
HS_INLINE void hs_argmin_pair(float &k, int &i, float k2, int i2) {
  if (k2 < k || (k2 == k && i2 < i)) {
    k = k2;
    i = i2;
  }
}

HS_INLINE void hs_wave_argmin(float &k, int &i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float k2 = __shfl_xor(k, off, 64);
    const int i2 = __shfl_xor(i, off, 64);
    hs_argmin_pair(k, i, k2, i2);
  }
}

// Same, carrying a payload (the slot the winning pair came from).
HS_INLINE void hs_wave_argmin_slot(float &k, int &i, int &slot) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float k2 = __shfl_xor(k, off, 64);
    const int i2 = __shfl_xor(i, off, 64);
    const int s2 = __shfl_xor(slot, off, 64);
    if (k2 < k || (k2 == k && i2 < i)) {
      k = k2;
      i = i2;
      slot = s2;
    }
  }
}

template <typename T>
HS_INLINE T hs_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
