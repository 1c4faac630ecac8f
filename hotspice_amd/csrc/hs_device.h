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
    // one 32x32->64 multiply (v_mad_u64_u32) per product gives both halves
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
  }
  return c;
}

// The two 32-bit words (a, b) of stream (seed, step, tag) at `index`: words 0,1 (even index) or
// 2,3 (odd index) of Philox call index>>1.
HS_INLINE void hs_words(uint64_t seed, uint64_t step, uint32_t tag, uint64_t index, uint32_t &a, uint32_t &b) {
  const uint4 w = hs_philox(make_uint4((uint32_t)(index >> 1), (uint32_t)step, (uint32_t)(step >> 32), tag),
                            (uint32_t)seed, (uint32_t)(seed >> 32));
  a = (index & 1) ? w.z : w.x;
  b = (index & 1) ? w.w : w.y;
}

// 53-bit integer (a>>5)*2^26 + (b>>6) of stream (seed, step, tag) at `index`.
HS_INLINE uint64_t hs_bits53(uint64_t seed, uint64_t step, uint32_t tag, uint64_t index) {
  uint32_t a, b;
  hs_words(seed, step, tag, index, a, b);
  return ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
}

// [0,1) (open_zero=0) or (0,1] (open_zero=1) uniform.
HS_INLINE double hs_u53(uint64_t bits, int open_zero) {
  return (double)(bits + (uint64_t)open_zero) * 0x1p-53;
}

// -ln(u) in FP32 for u = (bits53 + 1)*2^-53 in (0,1], given the two words behind bits53, without
// branches: u and w = 1-u are both formed from exact 27- and 26-bit integer parts (one rounding
// each); for w < 1/4 the series sum_{n<=11} w^n/n (truncation < 3e-8 relative), otherwise
// -log(u) from the hardware log2. Relative error < ~2e-7 over the whole range.
HS_INLINE float hs_exp1_fast(uint32_t a, uint32_t b) {
  const float u = __builtin_fmaf((float)(a >> 5), 0x1p-27f, (float)((b >> 6) + 1u) * 0x1p-53f);
  const float w = __builtin_fmaf((float)(~a >> 5), 0x1p-27f, (float)(~b >> 6) * 0x1p-53f);
  const float L = -__builtin_amdgcn_logf(u) * 0.693147180559945f;
  float p = 1.f / 11.f;
  p = __builtin_fmaf(p, w, 1.f / 10.f);
  p = __builtin_fmaf(p, w, 1.f / 9.f);
  p = __builtin_fmaf(p, w, 1.f / 8.f);
  p = __builtin_fmaf(p, w, 1.f / 7.f);
  p = __builtin_fmaf(p, w, 1.f / 6.f);
  p = __builtin_fmaf(p, w, 1.f / 5.f);
  p = __builtin_fmaf(p, w, 1.f / 4.f);
  p = __builtin_fmaf(p, w, 1.f / 3.f);
  p = __builtin_fmaf(p, w, 1.f / 2.f);
  p = __builtin_fmaf(p, w, 1.f);
  return w < 0.25f ? w * p : L;
}

// Natural log with the hardware log2 (keys only: their rounding matters for near-ties alone).
HS_INLINE float hs_log_fast(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }

// FP64 value with its low 6 mantissa bits cleared (toward zero, relative change < 2^-46 for a
// normal value; for a subnormal one, below 2.2e-308, up to 63 subnormal units), so a waiting
// time and a 6-bit round tag fit one 64-bit word; every NaN becomes the canonical NaN.
HS_INLINE double hs_dt_q(double v) {
  if (v != v) return __longlong_as_double(0x7FF8000000000000ll);
  return __longlong_as_double(__double_as_longlong(v) & ~63ll);
}

// x * exp(y) * g * s: the exponent is split as y*log2(e) = n + f in FP64, the mantissa
// x * 2^f * g is formed in FP32 (v_exp_f32, ~1 ulp) and scaled by 2^n and s in FP64. Relative
// error ~3e-7; +inf on overflow (NaN for x = 0, like x*exp(y) in FP64), NaN stays NaN.
HS_INLINE double hs_scaled_exp(float x, double y, float g, double s) {
  if (!(y < 709.78)) return y != y ? y : (double)x * (double)INFINITY * s;
  if (y < -745.2) return 0.0;
  const double z = y * 1.4426950408889634074;
  const double n = floor(z);
  const float m = x * __builtin_amdgcn_exp2f((float)(z - n)) * g;
  return ldexp((double)m, (int)n) * s;
}

// Neel waiting time of a cell in FP64 from its compensated fields (value - compensation) and its
// x = -ln(u): tau*exp(bmin*beta)/f0 == x*exp(b*beta)/f0 (core.py:698-700), or with two channels
// x/(exp(-b1 beta)+exp(-b2 beta))/(f0/2) (core.py:686-693). The barriers are formed in FP64
// (b*beta ~ 10-100 needs ~1e-9 relative), the rest as in hs_scaled_exp: relative error ~3e-7.
// Returned through hs_dt_q so every stepping path accumulates the same value.
HS_INLINE double hs_neel_dt(double e, double ep, double eb, double beta, double m_perp, int parabolic, bool perp,
                            float x, double f0) {
  const double dE = -2.0 * e, El = e;
  if (perp) {
    const double epf = m_perp * ep;
    const double b1 = fmax(dE, (eb + epf) - El), b2 = fmax(dE, (eb - epf) - El);
    const double lo = fmin(b1, b2), hi = fmax(b1, b2);
    const float g = 1.f / (1.f + __builtin_amdgcn_exp2f((float)(-(hi - lo) * beta * 1.4426950408889634074)));
    return hs_dt_q(hs_scaled_exp(x, lo * beta, g, 2.0 / f0));
  }
  double b;
  if (!parabolic) {
    b = fmax(dE, eb - El);
  } else {
    const double q = dE * (0.25 / eb) + 1.0;
    b = dE > 4.0 * eb ? dE : fmin(eb * (q * q), dE + 4.0 * eb);
  }
  return hs_dt_q(hs_scaled_exp(x, b * beta, 1.f, 1.0 / f0));
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

// hs_run_t.first_free: 0 = the steps of one progress() call (core.py:753-760), each behind its
// stop predicate; 1 = one update() call (core.py:604-622), whose step runs unconditionally;
// 2 = a sequence of update() calls: every step runs unconditionally and Neel compares each
// step's time with t_max itself. True when step `sd` of the call is such a free step.
HS_INLINE bool hs_step_free(int32_t first_free, int64_t sd) { return first_free == 2 || (first_free == 1 && sd == 0); }

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

// ---- DPP wave reductions (wave64, all lanes active) ------------------------------------------
// The minimum moves through quad, half-row, row and row-broadcast steps into lane 63 and is read
// back wave-uniform: six VALU-DPP steps instead of six LDS-crossbar shuffles.
template <int CTRL, int ROW_MASK = 0xf>
HS_INLINE uint32_t hs_dpp_u32(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xf, false);
}

HS_INLINE uint32_t hs_wave_min_u32(uint32_t v) {
  v = min(v, hs_dpp_u32<0xB1>(~0u, v));        // quad_perm [1,0,3,2]
  v = min(v, hs_dpp_u32<0x4E>(~0u, v));        // quad_perm [2,3,0,1]
  v = min(v, hs_dpp_u32<0x141>(~0u, v));       // row_half_mirror
  v = min(v, hs_dpp_u32<0x140>(~0u, v));       // row_mirror
  v = min(v, hs_dpp_u32<0x142, 0xa>(~0u, v));  // row_bcast:15 into rows 1 and 3
  v = min(v, hs_dpp_u32<0x143, 0xc>(~0u, v));  // row_bcast:31 into rows 2 and 3
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// 64-bit minimum as (min of the high words, then min of the low words among those lanes).
HS_INLINE unsigned long long hs_wave_min_u64(unsigned long long v) {
  const uint32_t hi = hs_wave_min_u32((uint32_t)(v >> 32));
  const uint32_t lo = hs_wave_min_u32((uint32_t)(v >> 32) == hi ? (uint32_t)v : ~0u);
  return ((unsigned long long)hi << 32) | lo;
}

// Wave minimum of v; `pay` becomes the payload of the lowest lane holding it.
HS_INLINE unsigned long long hs_wave_min_pay(unsigned long long v, unsigned long long &pay) {
  const unsigned long long m = hs_wave_min_u64(v);
  const int src = __ffsll((long long)__ballot(v == m)) - 1;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pay, src);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pay >> 32), src);
  pay = ((unsigned long long)hi << 32) | lo;
  return m;
}

// hs_wave_min_pay with two payloads taken from the lane that holds the minimum.
HS_INLINE unsigned long long hs_wave_min_pay2(unsigned long long v, unsigned long long &pay, unsigned long long &pay2) {
  const unsigned long long m = hs_wave_min_u64(v);
  const int src = __ffsll((long long)__ballot(v == m)) - 1;
  pay = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pay >> 32), src) << 32) |
        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pay, src);
  pay2 = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pay2 >> 32), src) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pay2, src);
  return m;
}

// Minimum over the workgroup (every thread calls it; sh holds blockDim.x / 64 floats). fminf
// drops NaN operands, as nanmin does.
HS_INLINE float hs_block_min_f32(float v, float *sh) {
  for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  float m = sh[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fminf(m, sh[q]);
  __syncthreads();
  return m;
}

// A 64-bit value of the wave's first active lane, in every lane.
HS_INLINE unsigned long long hs_uniform_u64(unsigned long long v) {
  return ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32)) << 32) |
         (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
}
