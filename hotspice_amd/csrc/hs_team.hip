// Persistent multi-instance ("team") stepping: the Neel-Arrhenius step of many spin ices in
// one cooperative launch.
//
// Every instance gets a team of G co-resident workgroups (1024 threads). The team's magnets are
// the instance's occupied cells (occ_list, row-major): owner thread t < 512 of workgroup `rank`
// holds list entries e = k*G*512 + rank*512 + t (k < MAXC) and keeps their spin, combined fields
// (+Kahan compensations), E_B and moment in registers for the whole launch. Thread 512 + t is
// owner t's drawer: it draws -ln u of the same magnets for the coming steps into LDS (Philox
// counter = cell, the streams of the single-instance path), up to two steps ahead. Per step:
//   1. owners form keys log(tau) = log(-ln u) - log(rate), packed as (key bits | position | spin)
//      so one 64-bit DPP minimum per wave gives the lowest key with nanargmin's first-index tie
//      break; the lane holding its wave's best leaves that cell's fields and draw in LDS;
//   2. wave 0 reduces the 8 owner-wave candidates and evaluates the FP64 waiting time of the
//      workgroup's best from its compensated fields. G == 1: that is the winner. G > 1: wave 0
//      writes the workgroup's record {candidate, time}, both words tagged with the round, into
//      the round's slot and polls (bounded spin, error on timeout) until all G records of the
//      round carry the tag, then reduces them. Slots rotate over three rounds; a workgroup reads
//      round k's records before it writes round k+1's, so no slot is overwritten while read.
//      Wave 0 then advances time and counters (core.py:700-707). Meanwhile (G > 1) the drawers
//      draw the next two steps;
//   3. every owner applies the rank-1 field update of the winner (energies.py:310-329 +
//      exchange stencil) to its registers, reading the winner's kernel row from L2.
// State is written back once at the end. Teams never wait for each other.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "hs_device.h"
#include "hs_internal.h"

namespace {

constexpr int TT = 1024;   // threads per workgroup
constexpr int OW = TT / 2;  // owner threads (waves 0-7); thread OW + t draws for owner t
constexpr int NBUF = 4;     // -ln(u) buffers: step s uses s % 4, drawers fill s+1 and s+2
constexpr uint64_t SPIN_LIMIT = 1ull << 22;
// a packed candidate is key bits (32) | round tag (6) | cell (24) | m+1 (2)
constexpr unsigned long long TAG_MASK = 63ull << 26;
constexpr unsigned long long NONE = ~0ull & ~TAG_MASK;  // "no candidate in this workgroup"

// One workgroup's contribution to a round: its best packed candidate and that cell's waiting
// time (low 6 mantissa bits cleared), each word carrying the round tag, so a reader that sees the
// tag in both words has the whole record even if the two words landed separately. Each record has
// a 128-byte line of its own: records that share a line serialise their device-scope stores and
// polls (64 workgroups exchange in about 1.0 us with one line each, 1.6 us at 16 bytes apart).
// With deferred times (below) the candidate is published before its time, and the time goes to
// a line of its own (dd), so storing it does not disturb the line that every workgroup polls.
// Records sit REC_BYTES apart (measured on MI355X with 64 workgroups: 2.72 us per step at 256
// bytes, 2.78 us with the deferred times in a separate array of 128-byte records).
constexpr int REC_BYTES = 256;
struct alignas(128) TeamRec {
  unsigned long long c, d, pad0[14];
  unsigned long long dd, pad1[REC_BYTES / 8 - 17];
};

struct TeamDev {
  const hs_instance_t *inst;  // device copy
  hs_run_t run;
  hs_state_t *state;
  int32_t *log_idx;
  double *log_dt;
  TeamRec *rec;  // [3][G]: round k uses slot k % 3
  int32_t G, first;
};

HS_INLINE uint32_t key_bits(float k) {
  uint32_t u = __float_as_uint(k);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// The key whose bits key_bits gave.
HS_INLINE float key_value(uint32_t b) { return __uint_as_float((b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b); }

// FP32 log-rate of a cell: -b*beta (one channel) or log(exp(-b1 beta) + exp(-b2 beta)); the
// same expression as the per-step launches use, so both paths rank cells identically.
// mpf = (float)m_perp_factor, barrier = 0 'simple' / 1 'parabolic'.
template <bool PERP>
HS_INLINE float team_log_rate(float mpf, int barrier, float e, float ep, float eb, float beta) {
  const float dE = -2.f * e, El = e;
  if (PERP) {
    const float epf = mpf * ep;
    const float b1 = fmaxf(dE, (eb + epf) - El), b2 = fmaxf(dE, (eb - epf) - El);
    const float lo = fminf(b1, b2), hi = fmaxf(b1, b2);
    return -lo * beta + hs_log_fast(1.f + __builtin_amdgcn_exp2f(-(hi - lo) * beta * 1.44269504088896f));
  }
  float b;
  if (barrier == 0) {
    b = fmaxf(dE, eb - El);
  } else {
    const float q = dE * (0.25f / eb) + 1.f;
    b = dE > 4.f * eb ? dE : fminf(eb * (q * q), dE + 4.f * eb);
  }
  return -b * beta;
}

// x % d for x < 2^24 and 1 <= d <= 64, with M = ceil(2^32 / d): q = floor(x * M / 2^32) is exact
// because x * (M d - 2^32) < 2^24 * 64 < 2^32.
HS_INLINE int mod_small(uint32_t x, uint32_t d, uint64_t M) {
  const uint32_t q = (uint32_t)(((uint64_t)x * M) >> 32);
  return (int)(x - q * d);
}

template <int MAXC, bool PERP, bool DEFER>
__global__ void __launch_bounds__(TT) k_neel_team(const TeamDev *teams, const int32_t *block_team, int32_t *err_out) {
  __shared__ hs_instance_t sI;
  __shared__ TeamDev sT;
  __shared__ unsigned long long s_pc[OW / 64];
  __shared__ float s_in[OW / 64][8];  // fields E, Ec, Ep, Epc, E_B and -ln(u) of each owner wave's best
  __shared__ unsigned long long s_win;
  __shared__ uint32_t s_ctl;
  __shared__ float s_x[NBUF][MAXC][OW];  // -ln(u) of the owned magnets, step s in buffer s % 4
  const int tid = threadIdx.x;
  const int team = block_team[blockIdx.x];
  if (tid == 0) sT = teams[team];
  __syncthreads();
  {
    const uint32_t *src = (const uint32_t *)sT.inst;
    uint32_t *dst = (uint32_t *)&sI;
    for (int i = tid; i < (int)(sizeof(hs_instance_t) / 4); i += TT) dst[i] = src[i];
  }
  __syncthreads();
  const hs_instance_t &I = sI;
  const hs_run_t R = sT.run;
  const int G = sT.G, rank = blockIdx.x - sT.first;
  const int lane = tid & 63, wv = tid >> 6;
  const bool owner = tid < OW;
  const int ot = owner ? tid : tid - OW;  // the owner this thread is or draws for
  const int nx = I.nx, nocc_i = I.n_occ;
  const int XB = nx > 1 ? 32 - __clz(nx - 1) : 0;  // positions y << XB | x order like y*nx + x
  const uint32_t XM = (1u << XB) - 1u;
  const int KW = 2 * I.kwx + 1;
  const size_t KS = (size_t)KW * (2 * I.kwy + 1);
  const int hy = I.nn_h / 2, hx = I.nn_w / 2;
  const double nocc = (double)I.n_occ;
  const float beta = I.beta[0];  // uniform over the grid (the host checks)
  // Loop invariants in registers: the descriptor sits in LDS, and the loop's own LDS stores would
  // otherwise turn every use into a fresh LDS read on the step's critical path.
  const float mpf = (float)I.m_perp_factor;
  const double mpf64 = I.m_perp_factor, dip_pf = I.dip_pf, Jx = I.J;
  const int barrier = I.barrier, kwx = I.kwx, kwy = I.kwy, ny = I.ny, pbc = I.pbc, in_plane = I.in_plane;
  const int nn_h = I.nn_h, nn_w = I.nn_w;
  const float *const K32 = I.K32, *const Kpo32 = I.Kpo32, *const mug = I.mu_src;
  const uint32_t seed_lo = (uint32_t)I.seed, seed_hi = (uint32_t)(I.seed >> 32);
  const uint32_t ux = (uint32_t)I.ux, uy = (uint32_t)I.uy;
  const uint64_t Mux = ((1ull << 32) + ux - 1) / ux, Muy = ((1ull << 32) + uy - 1) / uy;
  auto site_of = [&](int y, int x) { return mod_small((uint32_t)y, uy, Muy) * (int)ux + mod_small((uint32_t)x, ux, Mux); };

  // register-resident state of the magnets of list entries e_k = base + k*stride (k < nval).
  // Owners keep pos = y << XB | x and the fields; drawers keep the flat cell index in pos.
  // Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD b % 8); when the team
  // starts at a multiple of 8, XCD x holds ranks x, x+8, ..., and those take consecutive bands
  // of the list, so each XCD's L2 serves the kernel rows of one contiguous band of the grid.
  const int band = (G % 8 == 0 && sT.first % 8 == 0) ? (rank % 8) * (G / 8) + rank / 8 : rank;
  const int base = band * OW + ot, stride = G * OW;
  int nval = 0;
  uint32_t pos[MAXC];
  float e[MAXC], ec[MAXC], ep[MAXC], epc[MAXC], eb[MAXC], mu[MAXC];
  uint32_t occ = 0, neg = 0, cand = 0;  // bit k: occupied, m = -1, Neel candidate
#pragma unroll
  for (int k = 0; k < MAXC; ++k) {
    const int li = base + k * stride;
    pos[k] = 0;
    e[k] = ec[k] = ep[k] = epc[k] = eb[k] = mu[k] = 0.f;
    if (li < nocc_i) {
      nval = k + 1;
      const int c = I.occ_list[li];
      if (!owner) {
        pos[k] = (uint32_t)c;
        continue;
      }
      pos[k] = ((uint32_t)(c / nx) << XB) | (uint32_t)(c % nx);
      const int m = I.m[c];
      occ |= (m != 0 ? 1u : 0u) << k;
      neg |= (m < 0 ? 1u : 0u) << k;
      cand |= ((I.flags[c] & HS_FLAG_NEEL) != 0 ? 1u : 0u) << k;
      e[k] = I.E[c];
      ec[k] = I.Ec[c];
      if (PERP) {
        ep[k] = I.Ep[c];
        epc[k] = I.Epc[c];
      }
      eb[k] = I.EB[c];
      mu[k] = I.mu[c];
    }
  }
  auto spin = [&](int k) { return ((occ >> k) & 1u) ? (((neg >> k) & 1u) ? -1 : 1) : 0; };

  // time and counters: wave 0 advances them, rank 0 writes them back
  const hs_state_t S0 = sT.state[0];
  double t = S0.t, t0 = R.begin ? S0.t : S0.t0;
  int64_t att = S0.attempts, att0 = R.begin ? S0.attempts : S0.attempts0;
  int64_t sw = S0.switches, step = S0.step, sd = 0;
  int err = 0;
  uint32_t rounds = 0;
  // lower_than(att/n - att0/n, MCsteps_max) is monotone in att: find the first att where it
  // fails once, so the per-step test is an integer compare (same FP64 expression as core.py:756)
  int64_t att_stop = INT64_MAX;
  if (isfinite(R.mc_max) && R.mc_max * nocc < 9.0e18) {
    auto more = [&](int64_t a) { return hs_lower_than((double)a / nocc - (double)att0 / nocc, R.mc_max); };
    int64_t a = att0 + (int64_t)ceil(R.mc_max * nocc);
    if (a < att) a = att;
    while (more(a)) ++a;
    while (a > att && !more(a - 1)) --a;
    att_stop = a;
  }

  // -ln(u) draws by the drawers into buffers s % 4 (and (s+1) % 4 when `two`). Philox call j
  // serves cells 2j and 2j+1 (words 0,1 and 2,3). When every drawn magnet of a wave has the other
  // cell of its call in the neighbouring lane (fully occupied rows), the even lane evaluates call
  // j at step s and the odd lane at step s+1 and each hands the other the half it needs (DPP swap
  // of adjacent lanes): one call per magnet covers two steps. Otherwise every lane evaluates its
  // own call for each step.
  auto draw = [&](int64_t s, bool two) {
    const uint32_t odd = (uint32_t)lane & 1u;
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      const bool valid = k < nval;
      const uint32_t c = valid ? pos[k] : 0u, j = c >> 1;
      const uint32_t pj = hs_dpp_u32<0xB1>(~0u, valid ? j : ~0u);  // quad_perm [1,0,3,2]: lane ^ 1
      const bool paired = valid && (c & 1u) == odd && pj == j;
      if (two && __ballot(valid && !paired) == 0ull) {
        const uint64_t sk = (uint64_t)s + odd;
        const uint4 W = hs_philox(make_uint4(j, (uint32_t)sk, (uint32_t)(sk >> 32), 1u), seed_lo, seed_hi);
        const uint32_t ma = odd ? W.z : W.x, mb = odd ? W.w : W.y;  // this lane's cell in this call
        const uint32_t ra = hs_dpp_u32<0xB1>(0u, odd ? W.x : W.z);
        const uint32_t rb = hs_dpp_u32<0xB1>(0u, odd ? W.y : W.w);
        if (valid) {
          s_x[s & 3][k][ot] = hs_exp1_fast(odd ? ra : ma, odd ? rb : mb);
          s_x[(s + 1) & 3][k][ot] = hs_exp1_fast(odd ? ma : ra, odd ? mb : rb);
        }
      } else if (valid) {
        for (int q = 0; q < (two ? 2 : 1); ++q) {
          const uint64_t sk = (uint64_t)s + q;
          const uint4 W = hs_philox(make_uint4(j, (uint32_t)sk, (uint32_t)(sk >> 32), 1u), seed_lo, seed_hi);
          s_x[sk & 3][k][ot] = (c & 1u) ? hs_exp1_fast(W.z, W.w) : hs_exp1_fast(W.x, W.y);
        }
      }
    }
  };

  // hs_step_free(first_free, sd) without the step counter in the loop: the first step's flag,
  // then every later step's (only first_free == 2 frees them)
  const bool all_free = R.first_free == 2;
  bool free_now = hs_step_free(R.first_free, 0);
  bool go = sd < R.max_steps && (free_now || (att < att_stop && hs_lower_than(t - t0, R.t_max)));
  // Deferred waiting times (t_max = +inf, a team of G > 1, no step log): a step whose winning key
  // is below key_safe flips whatever its exact time, so the workgroups go on without waiting for
  // the winner's time; each adds it to t at the next round (read with that round's polls) or at
  // the end of the launch, in step order, so t is bit-identical to waiting for every time. With
  // |t| < 1e307 at the start and at most 2^20 steps of times below 1e300, t stays finite.
  const bool defer = DEFER && G > 1 && !R.log && R.t_max == INFINITY && fabs(t) < 1e307 && isfinite(t0) &&
                     R.max_steps <= (int64_t)1 << 20;
  const float key_safe = (float)(690.0 + log(R.attempt_freq / 2.0));  // exp(key) * 2 / f0 < 1e300
  int pend = -1;                   // rank whose record holds the pending time (-1: none)
  uint32_t pend_slot = 0;          // its slot
  unsigned long long pend_tag = 0; // its round tag
  if (go && !owner) draw(step, (step & 1) == 0);
  __syncthreads();
  while (go) {
    const int buf = (int)(step & 3);
    // G == 1 has no exchange to hide the draws behind: the drawers fill steps s+1, s+2 while the
    // owners work on step s (their buffers were last read at steps s-3, s-2)
    if (!owner && G == 1 && (step & 1)) draw(step + 1, true);
    // 1. keys of the owned candidates, packed (key bits | position | m+1) so one 64-bit minimum
    //    picks the lowest key and, on ties, the lowest flat index (nanargmin's first index)
    unsigned long long best = ~0ull;
    int bk = -1;
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      if ((cand >> k) & 1u) {
        float key = hs_log_fast(s_x[buf][k][ot]) - team_log_rate<PERP>(mpf, barrier, e[k], ep[k], eb[k], beta);
        if (isnan(key)) key = INFINITY;
        const unsigned long long p =
            ((unsigned long long)key_bits(key) << 32) | (pos[k] << 2) | (uint32_t)(spin(k) + 1);
        if (p < best) {
          best = p;
          bk = k;
        }
      }
    }
    // 2. the lane holding its owner wave's best leaves that cell's fields and draw for wave 0
    if (owner) {  // waves 0-7: wave-uniform, so the DPP minimum sees all lanes
      const unsigned long long wbest = hs_wave_min_u64(best);
      if (bk >= 0 && best == wbest) {
        float se = 0.f, sec = 0.f, sep = 0.f, sepc = 0.f, seb = 0.f;
#pragma unroll
        for (int k = 0; k < MAXC; ++k) {
          if (k == bk) {
            se = e[k];
            sec = ec[k];
            sep = ep[k];
            sepc = epc[k];
            seb = eb[k];
          }
        }
        s_pc[wv] = wbest;
        s_in[wv][0] = se;
        s_in[wv][1] = sec;
        s_in[wv][2] = sep;
        s_in[wv][3] = sepc;
        s_in[wv][4] = seb;
        s_in[wv][5] = s_x[buf][bk][ot];
      } else if (lane == 0 && wbest == ~0ull) {
        s_pc[wv] = ~0ull;
      }
    }
    __syncthreads();
    // 3. wave 0: the workgroup's best; with G > 1 it publishes the workgroup's record for this
    //    round (the candidate first, then its waiting time) and waits (bounded) until all G
    //    candidates of the round carry its tag; only the winner's time is needed, and a deferred
    //    step does not wait for it (below). Wave 0 then advances time and counters
    //    (core.py:700-707) and tells the others whether the winner flips and whether another
    //    step follows.
    if (wv == 0) {
      unsigned long long src = (unsigned long long)lane, pd = 0ull;
      unsigned long long pc = hs_wave_min_pay(lane < OW / 64 ? s_pc[lane] : ~0ull, src);
      if (pc == ~0ull) pc = NONE;
      // FP64 waiting time of the workgroup's best from its compensated fields (core.py:686-700)
      auto own_dt = [&]() -> unsigned long long {
        if (pc == NONE) return 0ull;
        const float *in = s_in[src];
        const double dt = hs_neel_dt((double)in[0] - (double)in[1], PERP ? (double)in[2] - (double)in[3] : 0.0,
                                     (double)in[4], I.beta64, mpf64, barrier, PERP, in[5], R.attempt_freq);
        return (unsigned long long)__double_as_longlong(dt);
      };
      bool bad = false, deferred = false;
      if (G > 1) {
        const unsigned long long tg = rounds & 63u;
        const uint32_t sl = rounds % 3u;
        TeamRec *slot = sT.rec + (size_t)sl * G;
        if (defer) {  // the candidate at once, its time (needed only if it wins) after it
          if (lane == 0) __hip_atomic_store(&slot[rank].c, pc | (tg << 26), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          pd = own_dt();
          if (lane == 0) __hip_atomic_store(&slot[rank].dd, pd | tg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          pd = own_dt();
          if (lane == 0) {
            __hip_atomic_store(&slot[rank].c, pc | (tg << 26), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&slot[rank].d, pd | tg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        // the previous round's deferred time, loaded together with the first polls
        unsigned long long dprev = 0ull;
        if (pend >= 0 && lane == 0)
          dprev = __hip_atomic_load(&sT.rec[(size_t)pend_slot * G + pend].dd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long c2 = ~0ull, d2 = 0ull, r2 = 0ull;
        for (int r = lane; r < G; r += 64) {
          unsigned long long cr, dr;
          uint64_t spins = 0;
          for (;;) {
            cr = __hip_atomic_load(&slot[r].c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            dr = defer ? 0ull : __hip_atomic_load(&slot[r].d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (((cr & TAG_MASK) >> 26) == tg && (defer || (dr & 63ull) == tg)) break;
            if (++spins > SPIN_LIMIT) {
              bad = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          if (cr < c2) {
            c2 = cr;
            d2 = dr;
            r2 = (unsigned long long)r;
          }
        }
        pc = hs_wave_min_pay2(c2, d2, r2) & ~TAG_MASK;
        bad = __ballot(bad) != 0ull;
        if (pend >= 0) {  // settle the previous round's time first: t accumulates in step order
          unsigned long long dp = hs_uniform_u64(dprev);
          uint64_t spins = 0;
          while ((dp & 63ull) != pend_tag && !bad) {
            if (++spins > SPIN_LIMIT) bad = true;
            dp = __hip_atomic_load(&sT.rec[(size_t)pend_slot * G + pend].dd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          t += __longlong_as_double((long long)(dp & ~63ull));
          pend = -1;
        }
        // a deferred step: t_max = +inf and a finite key below key_safe, whose waiting time is
        // finite and below 1e300, so the step flips and the stop predicate does not depend on t
        // (t stays below 1e307 over the launch); every other step waits for the winner's time
        if (defer && pc != NONE && !bad) {
          const float key = key_value((uint32_t)(pc >> 32));
          deferred = key < key_safe && key > -INFINITY;
        }
        if (defer && !deferred && pc != NONE && !bad) {  // a step that needs the winner's time now
          uint64_t spins = 0;
          d2 = __hip_atomic_load(&slot[r2].dd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          while ((d2 & 63ull) != tg) {
            if (++spins > SPIN_LIMIT) {
              bad = true;
              break;
            }
            d2 = __hip_atomic_load(&slot[r2].dd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          bad = __ballot(bad) != 0ull;
        }
        if (deferred) {
          pend = (int)r2;
          pend_slot = sl;
          pend_tag = tg;
        }
        pd = d2 & ~63ull;
      } else {
        pd = own_dt();
      }
      ++rounds;
      uint32_t ctl = 4u;  // stop without a step
      if (bad) {
        err = 3;  // a record of the round never arrived
      } else if (pc == NONE) {
        err = 1;  // nanargmin over an all-NaN grid
      } else if (deferred) {
        att += 1;
        sd += 1;
        sw += 1;
        free_now = all_free;
        const bool more = sd < R.max_steps && (all_free || att < att_stop);
        ctl = 1u | (more ? 2u : 0u);
      } else {
        const double dt = __longlong_as_double((long long)pd);
        // progress passes t_max - dt (core.py:757); update() passes t_max itself, also when t
        // has overflowed to +inf (where t - t0 would be NaN)
        const double tmax_step = free_now ? R.t_max : R.t_max - (t - t0);
        free_now = all_free;
        att += 1;
        sd += 1;
        const bool flip = !(dt > tmax_step || isnan(dt));
        t += flip ? dt : tmax_step;
        sw += flip ? 1 : 0;
        if (rank == 0 && lane == 0 && R.log) {
          const uint32_t wp = (uint32_t)(pc >> 2) & 0xFFFFFFu;
          sT.log_idx[sd - 1] = flip ? (int)(wp >> XB) * nx + (int)(wp & XM) : -1;
          sT.log_dt[sd - 1] = flip ? dt : tmax_step;
        }
        const bool more = sd < R.max_steps && (all_free || (att < att_stop && hs_lower_than(t - t0, R.t_max)));
        ctl = (flip ? 1u : 0u) | (more ? 2u : 0u);
      }
      if (lane == 0) {
        s_win = pc;
        s_ctl = ctl;
      }
    } else if (!owner && G > 1 && (step & 1)) {
      // meanwhile the drawers fill the next two steps' buffers
      draw(step + 1, true);
    }
    __syncthreads();
    const uint32_t ctl = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_ctl);
    if (ctl & 4u) break;
    step += 1;
    go = (ctl & 2u) != 0;
    if (!(ctl & 1u) || !owner) continue;
    // 4. rank-1 field update of the winner (energies.py:310-329 + exchange stencil)
    const unsigned long long win = s_win;
    const uint32_t wpos = (uint32_t)__builtin_amdgcn_readfirstlane((int)((uint32_t)(win >> 2) & 0xFFFFFFu));
    const int mw = __builtin_amdgcn_readfirstlane((int)(win & 3ull)) - 1;
    const int yw = (int)(wpos >> XB), xw = (int)(wpos & XM);
    const int w = yw * nx + xw, sitew = site_of(yw, xw);
    const int nw = dip_pf != 0.0 ? I.sub_of[sitew] : -1;
    const double mnew = -(double)mw;
    const float A = (float)(2.0 * dip_pf * (mnew * (double)mug[w]));
    const float Ax = (float)(-2.0 * Jx * mnew);
    const float oxw = (float)I.ox_uc[sitew], oyw = (float)I.oy_uc[sitew];
    const float *Kw = nw >= 0 ? K32 + (size_t)nw * KS : nullptr;
    const float *Kpw = (PERP && nw >= 0) ? Kpo32 + (size_t)nw * KS : nullptr;
    constexpr int CH = MAXC < 4 ? MAXC : 4;  // cells whose kernel entries are loaded together
#pragma unroll
    for (int k0 = 0; k0 < MAXC; k0 += CH) {
      float kv[CH], kpv[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int k = k0 + j;
        kv[j] = 0.f;
        kpv[j] = 0.f;
        if (k < nval && Kw) {
          const int dy = (int)(pos[k] >> XB) - yw, dx = (int)(pos[k] & XM) - xw;
          if (dy >= -kwy && dy <= kwy && dx >= -kwx && dx <= kwx) {
            const int o = (kwy + dy) * KW + (kwx + dx);
            kv[j] = Kw[o];
            if (PERP) kpv[j] = Kpw[o];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int k = k0 + j;
        if (k >= nval) continue;
        if (pos[k] == wpos) {
          neg ^= 1u << k;
          e[k] = -e[k];
          ec[k] = -ec[k];
          if (PERP) {
            ep[k] = -ep[k];
            epc[k] = -epc[k];
          }
          continue;
        }
        const int mk = spin(k);
        if (mk == 0) continue;
        const int y = (int)(pos[k] >> XB), x = (int)(pos[k] & XM);
        const int dy = y - yw, dx = x - xw;
        const float mmu = (float)mk * mu[k];
        float de = A * (mmu * kv[j]);
        float dep = PERP ? A * (mmu * kpv[j]) : 0.f;
        if (Ax != 0.f) {
          int qy = dy, qx = dx;
          if (pbc) {
            qy = hs_pmod(dy, ny);
            qy = qy > ny / 2 ? qy - ny : qy;
            qx = hs_pmod(dx, nx);
            qx = qx > nx / 2 ? qx - nx : qx;
          }
          qy += hy;
          qx += hx;
          if (qy >= 0 && qy < nn_h && qx >= 0 && qx < nn_w) {
            const float wq = I.nn_mask[qy * nn_w + qx];
            if (wq != 0.f) {
              if (in_plane) {
                const int s = site_of(y, x);
                const float oxi = (float)I.ox_uc[s], oyi = (float)I.oy_uc[s];
                de += Ax * (float)mk * wq * (oxi * oxw + oyi * oyw);
                if (PERP) dep += Ax * (float)mk * wq * (oyw * oxi - oxw * oyi);
              } else {
                de += Ax * (float)mk * wq;
              }
            }
          }
        }
        if (de != 0.f) hs_kahan_add(e[k], ec[k], de);
        if (PERP && dep != 0.f) hs_kahan_add(ep[k], epc[k], dep);
      }
    }
  }
  // write back
#pragma unroll
  for (int k = 0; k < MAXC; ++k) {
    if (!owner || k >= nval) continue;
    const int c = (int)(pos[k] >> XB) * nx + (int)(pos[k] & XM);
    I.m[c] = (int8_t)spin(k);
    I.E[c] = e[k];
    I.Ec[c] = ec[k];
    if (PERP) {
      I.Ep[c] = ep[k];
      I.Epc[c] = epc[k];
    }
  }
  if (rank == 0 && wv == 0 && pend >= 0) {  // the last deferred time
    uint64_t spins = 0;
    unsigned long long dp;
    for (;;) {
      dp = __hip_atomic_load(&sT.rec[(size_t)pend_slot * G + pend].dd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((dp & 63ull) == pend_tag) break;
      if (++spins > SPIN_LIMIT) {
        err = 3;
        break;
      }
    }
    if (err != 3) t += __longlong_as_double((long long)(dp & ~63ull));
  }
  if (rank == 0 && tid == 0) {
    hs_state_t o = S0;
    o.t = t;
    o.t0 = t0;
    o.attempts = att;
    o.attempts0 = att0;
    o.switches = sw;
    o.step = step;
    o.steps_done = sd;
    o.done = 1;
    o.err = err;
    sT.state[0] = o;
  }
  if (err && tid == 0) atomicMax(err_out, err);
}

// Clears the error word and marks every record slot stale (tag 63 in both words: rounds 0-2 use
// tags 0-2 and every slot is rewritten before tag 63 comes round).
__global__ void k_team_reset(const TeamDev *teams, int32_t *err) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *err = 0;
  const TeamDev d = teams[blockIdx.x];
  for (int i = threadIdx.x; i < 3 * d.G; i += blockDim.x) {
    d.rec[i].c = ~0ull;
    d.rec[i].d = ~0ull;
    d.rec[i].dd = ~0ull;
  }
}

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
  size_t teams_off, inst_off, map_off, sync_off, err_off, total;
  std::vector<size_t> team_sync;  // offset of each team's sync block
  std::vector<int> G, first;
  int nblocks, maxc;
};

// Default team size: one magnet per owner thread, at most one workgroup per CU of the device
// (more magnets per thread than 4 spill the register-resident state).
int team_blocks(const hs_team_t &t, int cap) {
  if (t.nblocks > 0) return t.nblocks;
  int g = (t.inst->n_occ + OW - 1) / OW;
  return g < 1 ? 1 : (g > cap ? cap : g);
}

// CUs of the current device (the cap of a default team), 128 if the query fails.
int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    return 128;
  return cus;
}

int make_layout(const hs_team_t *teams, int n, Layout &L) {
  const int cap = device_cus();
  L.teams_off = 0;
  L.inst_off = al256(sizeof(TeamDev) * n);
  L.map_off = L.inst_off + al256(sizeof(hs_instance_t) * n);
  L.nblocks = 0;
  L.maxc = 1;
  L.G.resize(n);
  L.first.resize(n);
  for (int i = 0; i < n; ++i) {
    const int g = team_blocks(teams[i], cap);
    L.G[i] = g;
    L.first[i] = L.nblocks;
    L.nblocks += g;
    const int64_t per = (int64_t)g * OW;
    const int c = (int)((teams[i].inst->n_occ + per - 1) / per);
    while (L.maxc < c) L.maxc *= 2;
  }
  L.sync_off = L.map_off + al256(sizeof(int32_t) * (size_t)L.nblocks);
  size_t off = L.sync_off;
  L.team_sync.resize(n);
  for (int i = 0; i < n; ++i) {
    L.team_sync[i] = off;
    off += al256(sizeof(TeamRec) * 3 * (size_t)L.G[i]);
  }
  L.err_off = off;
  L.total = off + 256;
  return 0;
}

template <int MAXC>
const void *team_kernel(bool perp, bool defer) {
  if (defer) return perp ? (const void *)&k_neel_team<MAXC, true, true> : (const void *)&k_neel_team<MAXC, false, true>;
  return perp ? (const void *)&k_neel_team<MAXC, true, false> : (const void *)&k_neel_team<MAXC, false, false>;
}

// The k_neel_team instance for the largest owned-magnet count, the E_perp setting and whether
// deferred waiting times are compiled in (only worth it when every team runs with t_max = +inf
// and no step log; the kernel without them keeps the shorter step of the other calls).
const void *team_kernel_of(int maxc, bool perp, bool defer) {
  switch (maxc) {
    case 1: return team_kernel<1>(perp, defer);
    case 2: return team_kernel<2>(perp, defer);
    case 4: return team_kernel<4>(perp, defer);
    case 8: return team_kernel<8>(perp, defer);
    default: return team_kernel<16>(perp, defer);
  }
}

}  // namespace

extern "C" int64_t hs_neel_teams_ws_bytes(const hs_team_t *teams, int32_t nteams) {
  if (!teams || nteams <= 0) return -1;
  for (int i = 0; i < nteams; ++i)
    if (!teams[i].inst) return -1;
  Layout L;
  make_layout(teams, nteams, L);
  return (int64_t)L.total;
}

extern "C" int hs_neel_run_teams(const hs_team_t *teams, int32_t nteams, void *ws, int64_t ws_bytes, void *stream) {
  hs_clear_error();
  if (!teams || nteams <= 0 || !ws) return hs_set_error(HS_ERR_ARG, "missing teams / workspace");
  bool perp = false;
  for (int i = 0; i < nteams; ++i) {
    const hs_team_t &t = teams[i];
    int rc = hs_check_instance(t.inst);
    if (rc) return rc;
    rc = hs_check_f32_fields(t.inst, "hs_neel_run_teams");
    if (rc) return rc;
    if (!t.state) return hs_set_error(HS_ERR_ARG, "team %d has no state", i);
    if (t.run.max_steps < 0) return hs_set_error(HS_ERR_ARG, "negative step budget");
    if (t.run.log && (!t.log_idx || !t.log_dt)) return hs_set_error(HS_ERR_ARG, "team %d: log without arrays", i);
    if (!t.inst->uniform_beta) return hs_set_error(HS_ERR_ARG, "Neel stepping needs a uniform temperature");
    if (!(t.inst->beta64 > 0.0)) return hs_set_error(HS_ERR_ARG, "team %d: beta64 must be > 0 (FP64 1/(kB T))", i);
    if (i > 0 && (bool)t.inst->use_perp != perp) return hs_set_error(HS_ERR_ARG, "teams of one launch must agree on use_perp");
    perp = t.inst->use_perp;
    if (t.nblocks < 0 || t.nblocks > 4096) return hs_set_error(HS_ERR_ARG, "team %d: bad nblocks", i);
  }
  Layout L;
  make_layout(teams, nteams, L);
  if (ws_bytes < (int64_t)L.total) return hs_set_error(HS_ERR_ARG, "workspace %lld < %lld bytes", (long long)ws_bytes, (long long)L.total);
  if (L.maxc > 16) return hs_set_error(HS_ERR_ARG, "more than 16 magnets per owner thread: give the team more workgroups");
  for (int i = 0; i < nteams; ++i) {
    // packed candidates hold the position y << ceil(log2 nx) | x in 24 bits
    const hs_instance_t &in = *teams[i].inst;
    int xb = 0;
    while ((1 << xb) < in.nx) ++xb;
    if (xb > 23 || ((int64_t)in.ny << xb) > (1 << 24))
      return hs_set_error(HS_ERR_ARG, "team %d: grid %dx%d too large for the persistent path (ny * 2^ceil(log2 nx) > 2^24)",
                          i, in.nx, in.ny);
  }
  hipStream_t s = (hipStream_t)stream;
  int dev = 0;
  HS_HIP_TRY(hipGetDevice(&dev));
  int cus = 0, per = 0;
  HS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  bool defer = true;
  for (int i = 0; i < nteams; ++i) defer = defer && teams[i].run.t_max == INFINITY && !teams[i].run.log;
  const void *fn = team_kernel_of(L.maxc, perp, defer);
  HS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, TT, 0));
  if ((int64_t)per * cus < L.nblocks)
    return hs_set_error(HS_ERR_ARG, "%d workgroups requested but only %d can be co-resident (%d per CU x %d CUs)",
                        L.nblocks, per * cus, per, cus);
  // descriptors
  char *base = (char *)ws;
  std::vector<char> host(L.sync_off, 0);
  TeamDev *td = (TeamDev *)(host.data() + L.teams_off);
  hs_instance_t *ti = (hs_instance_t *)(host.data() + L.inst_off);
  int32_t *map = (int32_t *)(host.data() + L.map_off);
  for (int i = 0; i < nteams; ++i) {
    ti[i] = *teams[i].inst;
    TeamDev d{};
    d.inst = (const hs_instance_t *)(base + L.inst_off) + i;
    d.run = teams[i].run;
    d.state = teams[i].state;
    d.log_idx = teams[i].log_idx;
    d.log_dt = teams[i].log_dt;
    d.rec = (TeamRec *)(base + L.team_sync[i]);
    d.G = L.G[i];
    d.first = L.first[i];
    td[i] = d;
    for (int b = 0; b < L.G[i]; ++b) map[L.first[i] + b] = i;
  }
  if (int rc = hs_upload_async(base, host.data(), host.size(), s)) return rc;
  int32_t *err = (int32_t *)(base + L.err_off);
  k_team_reset<<<nteams, 256, 0, s>>>((const TeamDev *)base, err);
  HS_LAUNCH_CHECK("k_team_reset");
  const TeamDev *dtd = (const TeamDev *)(base + L.teams_off);
  const int32_t *dmap = (const int32_t *)(base + L.map_off);
  void *args[] = {(void *)&dtd, (void *)&dmap, (void *)&err};
  const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(L.nblocks), dim3(TT), args, 0, s);
  if (e != hipSuccess) return hs_set_error(HS_ERR_HIP, "cooperative launch of k_neel_team failed: %s", hipGetErrorString(e));
  return HS_OK;
}

// ---- the exchange round alone ----------------------------------------------------------------
namespace {

// G workgroups of one wave exchange a tagged candidate per round through the same records,
// stores and bounded polls as k_neel_team's step (device-scope relaxed atomics, one REC_BYTES
// record per workgroup, three slots), and reduce them with the same 64-bit minimum; nothing else.
// A round whose records do not all arrive within SPIN_LIMIT polls counts in *fails (the kernel
// always ends).
__global__ void __launch_bounds__(64) k_exchange_probe(TeamRec *rec, int G, int rounds, int *fails,
                                                       unsigned long long *sink) {
  const int rank = blockIdx.x, lane = threadIdx.x;
  unsigned long long acc = 0ull;
  int bad_rounds = 0;
  for (int round = 0; round < rounds; ++round) {
    const unsigned long long tg = (unsigned long long)(round & 63);
    TeamRec *slot = rec + (size_t)(round % 3) * G;
    const unsigned long long pc = ((unsigned long long)(rank * 2654435761u) << 32) | ((unsigned long long)rank << 2);
    if (lane == 0) __hip_atomic_store(&slot[rank].c, (pc & ~TAG_MASK) | (tg << 26), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long c2 = ~0ull;
    bool bad = false;
    for (int r = lane; r < G; r += 64) {
      uint64_t spins = 0;
      unsigned long long cr;
      for (;;) {
        cr = __hip_atomic_load(&slot[r].c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (((cr & TAG_MASK) >> 26) == tg) break;
        if (++spins > SPIN_LIMIT) {
          bad = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      c2 = cr < c2 ? cr : c2;
    }
    acc ^= hs_wave_min_u64(c2);
    if (__ballot(bad) != 0ull) ++bad_rounds;
  }
  if (lane == 0) {
    sink[rank] = acc;
    if (bad_rounds) atomicAdd(fails, bad_rounds);
  }
}

}  // namespace

extern "C" int hs_exchange_round_us(int32_t G, int32_t rounds, double *us_per_round, int32_t *failed_rounds) {
  hs_clear_error();
  if (!us_per_round || G < 1 || rounds < 1) return hs_set_error(HS_ERR_ARG, "need G >= 1, rounds >= 1 and an output");
  int dev = 0, cus = 0;
  HS_HIP_TRY(hipGetDevice(&dev));
  HS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (G > cus) return hs_set_error(HS_ERR_ARG, "G = %d exceeds the %d co-resident workgroups of one per CU", G, cus);
  const size_t rec_bytes = sizeof(TeamRec) * 3 * (size_t)G;
  char *buf = nullptr;
  HS_HIP_TRY(hipMalloc(&buf, rec_bytes + 256 + 8 * (size_t)G));
  TeamRec *rec = (TeamRec *)buf;
  int *fails = (int *)(buf + rec_bytes);
  unsigned long long *sink = (unsigned long long *)(buf + rec_bytes + 256);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = HS_OK;
  float ms = 0.f;
  int hf = 0;
  auto fail = [&](hipError_t e, const char *what) {
    if (rc == HS_OK && e != hipSuccess) rc = hs_set_error(HS_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
  };
  fail(hipMemset(buf, 0xff, rec_bytes), "hipMemset");
  fail(hipMemset(fails, 0, 4), "hipMemset");
  fail(hipEventCreate(&e0), "hipEventCreate");
  fail(hipEventCreate(&e1), "hipEventCreate");
  if (rc == HS_OK) {
    void *args[] = {(void *)&rec, (void *)&G, (void *)&rounds, (void *)&fails, (void *)&sink};
    fail(hipEventRecord(e0, nullptr), "hipEventRecord");
    fail(hipLaunchCooperativeKernel((const void *)&k_exchange_probe, dim3(G), dim3(64), args, 0, nullptr),
         "cooperative launch of k_exchange_probe");
    fail(hipEventRecord(e1, nullptr), "hipEventRecord");
    fail(hipEventSynchronize(e1), "hipEventSynchronize");
    fail(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    fail(hipMemcpy(&hf, fails, 4, hipMemcpyDeviceToHost), "hipMemcpy");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  if (rc != HS_OK) return rc;
  *us_per_round = (double)ms * 1000.0 / rounds;
  if (failed_rounds) *failed_rounds = hf;
  return HS_OK;
}
