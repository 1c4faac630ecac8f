// Persistent multi-instance ("team") stepping: the Neel-Arrhenius step of many spin ices in
// one cooperative launch.
//
// Every instance gets a team of G co-resident workgroups (256 threads). Each thread owns up to
// MAXC grid cells (c = rank*256 + tid + k*G*256) and keeps their spin, combined fields (+Kahan
// compensations), E_B and beta in registers for the whole launch. Per step:
//   1. keys log(tau) = log(-ln u) - log(rate) for the owned cells (Philox counter = cell, same
//      streams as the single-instance path), best (key, cell) per thread, then per workgroup;
//   2. G == 1: the workgroup's best is the winner. G > 1: thread 0 publishes the workgroup's
//      winner time/spin in a slot and folds its (key, cell) into one 64-bit atomic minimum,
//      then arrives at the team's counter barrier (bounded spin, error on timeout) and reads
//      the team winner back; slots rotate over three steps so no reset races a reader;
//   3. every thread applies the identical time/counter update (core.py:700-707) and the
//      rank-1 field update of the winner (energies.py:310-329 + exchange stencil) to its
//      registers, reading the winner's kernel row from L2.
// State is written back once at the end. Teams never wait for each other.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "hs_device.h"
#include "hs_internal.h"

namespace {

constexpr int TT = 1024;           // threads per workgroup (16 waves: latency hiding)
constexpr uint64_t SPIN_LIMIT = 1ull << 26;

struct TeamDev {
  const hs_instance_t *inst;  // device copy
  hs_run_t run;
  hs_state_t *state;
  int32_t *log_idx;
  double *log_dt;
  uint32_t *bar;
  unsigned long long *minslot;  // [3]
  double *dtslot;               // [3][G] waiting time of each workgroup's best cell
  uint32_t *dtflag;             // [3][G] round tag: the slot holds this round's value
  int32_t G, first;
};

HS_INLINE uint32_t key_bits(float k) {
  uint32_t u = __float_as_uint(k);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}



template <bool PERP>
HS_INLINE float team_log_rate(const hs_instance_t &I, float e, float ep, float eb, float beta) {
  const float dE = -2.f * e, El = e;
  if (PERP) {
    const float epf = (float)I.m_perp_factor * ep;
    const float b1 = fmaxf(dE, (eb + epf) - El), b2 = fmaxf(dE, (eb - epf) - El);
    const float lo = fminf(b1, b2), hi = fmaxf(b1, b2);
    return -lo * beta + log1pf(expf(-(hi - lo) * beta));
  }
  float b;
  if (I.barrier == 0) {
    b = fmaxf(dE, eb - El);
  } else {
    const float q = dE * (0.25f / eb) + 1.f;
    b = dE > 4.f * eb ? dE : fminf(eb * (q * q), dE + 4.f * eb);
  }
  return -b * beta;
}

HS_INLINE unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)v, off, 64), hi = __shfl_xor((uint32_t)(v >> 32), off, 64);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    v = o < v ? o : v;
  }
  return v;
}

template <int MAXC, bool PERP>
__global__ void __launch_bounds__(TT) k_neel_team(const TeamDev *teams, const int32_t *block_team, int32_t *err_out) {
  __shared__ hs_instance_t sI;
  __shared__ TeamDev sT;
  __shared__ unsigned long long s_part[TT / 64];
  __shared__ unsigned long long s_block, s_win;
  __shared__ double s_dtw;
  __shared__ int s_bad;
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
  const int nx = I.nx, ncell = I.nx * I.ny;
  const int KW = 2 * I.kwx + 1;
  const size_t KS = (size_t)KW * (2 * I.kwy + 1);
  const int hy = I.nn_h / 2, hx = I.nn_w / 2;
  const double nocc = (double)I.n_occ;

  // register-resident state of the owned cells
  int cell[MAXC], mm[MAXC];
  float e[MAXC], ec[MAXC], ep[MAXC], epc[MAXC], eb[MAXC], bt[MAXC], mu[MAXC];
  bool cand[MAXC];
#pragma unroll
  for (int k = 0; k < MAXC; ++k) {
    const int c = rank * TT + tid + k * G * TT;
    cell[k] = c < ncell ? c : -1;
    mm[k] = 0;
    e[k] = ec[k] = ep[k] = epc[k] = eb[k] = bt[k] = mu[k] = 0.f;
    cand[k] = false;
    if (c < ncell) {
      mm[k] = I.m[c];
      e[k] = I.E[c];
      ec[k] = I.Ec[c];
      if (PERP) {
        ep[k] = I.Ep[c];
        epc[k] = I.Epc[c];
      }
      eb[k] = I.EB[c];
      bt[k] = I.beta[c];
      mu[k] = I.mu[c];
      cand[k] = (I.flags[c] & HS_FLAG_NEEL) != 0;
    }
  }
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
  const int lane = tid & 63, wv = tid >> 6;

  for (;;) {
    if (!(sd < R.max_steps && att < att_stop && hs_lower_than(t - t0, R.t_max))) break;
    // 1. keys of the owned cells, packed (key bits | cell | m+1) so one 64-bit minimum picks the
    //    lowest key and, on ties, the lowest flat index (nanargmin's first index)
    unsigned long long best = ~0ull;
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      if (cell[k] >= 0 && cand[k]) {
        const uint64_t bits = hs_bits53(I.seed, (uint64_t)step, 1u, (uint64_t)cell[k]);
        float key = logf(hs_exp1_f32(bits)) - team_log_rate<PERP>(I, e[k], ep[k], eb[k], bt[k]);
        if (isnan(key)) key = INFINITY;
        const unsigned long long p = ((unsigned long long)key_bits(key) << 32) | ((uint32_t)cell[k] << 2) | (uint32_t)(mm[k] + 1);
        best = p < best ? p : best;
      }
    }
    best = wave_min_u64(best);
    if (lane == 0) s_part[wv] = best;
    __syncthreads();
    if (wv == 0) {
      unsigned long long v = lane < TT / 64 ? s_part[lane] : ~0ull;
      v = wave_min_u64(v);
      if (lane == 0) s_block = v;
    }
    __syncthreads();
    const unsigned long long bb = s_block;
    const int bcell = bb == ~0ull ? -1 : (int)((bb & 0xffffffffull) >> 2);
    const int sl = (int)(sd % 3);
    // 2a. the owner of the workgroup's best cell evaluates its FP64 waiting time from the
    //     compensated fields and publishes it (program order before the exchange: a wave never
    //     waits on its own lanes)
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      if (cell[k] == bcell) {
        const uint64_t bits = hs_bits53(I.seed, (uint64_t)step, 1u, (uint64_t)cell[k]);
        const double dtk = hs_neel_dt((double)e[k] - (double)ec[k], PERP ? (double)ep[k] - (double)epc[k] : 0.0,
                                      (double)eb[k], (double)bt[k], I.m_perp_factor, I.barrier, PERP, bits,
                                      R.attempt_freq);
        if (G > 1) {
          __hip_atomic_store(&sT.dtslot[sl * G + rank], dtk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __builtin_amdgcn_s_waitcnt(0);
          __hip_atomic_store(&sT.dtflag[sl * G + rank], rounds + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          s_dtw = dtk;
        }
      }
    }
    // 2. team winner: one 64-bit atomic minimum + counter barrier (bounded spins)
    if (tid == 0) {
      unsigned long long win = bb;
      int bad = 0;
      if (G > 1) {
        __hip_atomic_fetch_min(&sT.minslot[sl], bb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (rank == 0) __hip_atomic_store(&sT.minslot[(sl + 1) % 3], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // all team data travels through device-scope atomics performed at the coherence point, so
        // the arrival only waits for them to complete: no agent-scope release/acquire, which would
        // write back and invalidate the L2 holding the kernels every step
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(sT.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t target = (rounds + 1) * (uint32_t)G;
        uint64_t spins = 0;
        while (__hip_atomic_load(sT.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
          if (++spins > SPIN_LIMIT) {
            bad = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        win = __hip_atomic_load(&sT.minslot[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!bad && win != ~0ull) {  // the winner's time, once its owner's slot shows this round
          const int owner = ((int)((win & 0xffffffffull) >> 2) / TT) % G;
          while (__hip_atomic_load(&sT.dtflag[sl * G + owner], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != rounds + 1u) {
            if (++spins > SPIN_LIMIT) {
              bad = 1;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          s_dtw = __hip_atomic_load(&sT.dtslot[sl * G + owner], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      s_win = win;
      s_bad = bad;
    }
    __syncthreads();
    const unsigned long long win = s_win;
    ++rounds;
    if (s_bad) {
      err = 3;  // team barrier timed out
      break;
    }
    if (win == ~0ull) {
      err = 1;  // nanargmin over an all-NaN grid
      break;
    }
    const int w = (int)((win & 0xffffffffull) >> 2);
    const int mw = (int)(win & 3ull) - 1;
    const double dt = s_dtw;
    // 3. time, counters (core.py:700-707) and the field update of the winner
    const double tmax_step = R.t_max - (t - t0);
    att += 1;
    step += 1;
    sd += 1;
    const bool flip = !(dt > tmax_step || isnan(dt));
    t += flip ? dt : tmax_step;
    if (rank == 0 && tid == 0 && R.log) {
      sT.log_idx[sd - 1] = flip ? w : -1;
      sT.log_dt[sd - 1] = flip ? dt : tmax_step;
    }
    if (flip) {
      sw += 1;
      const int yw = w / nx, xw = w % nx, sitew = hs_site(I, yw, xw);
      const int nw = I.dip_pf != 0.0 ? I.sub_of[sitew] : -1;
      const double mnew = -(double)mw;
      const float A = (float)(2.0 * I.dip_pf * (mnew * (double)I.mu[w]));
      const float Ax = (float)(-2.0 * I.J * mnew);
      const float oxw = (float)I.ox_uc[sitew], oyw = (float)I.oy_uc[sitew];
      const float *Kw = nw >= 0 ? I.K32 + (size_t)nw * KS : nullptr;
      const float *Kpw = (PERP && nw >= 0) ? I.Kpo32 + (size_t)nw * KS : nullptr;
      float kv[MAXC], kpv[MAXC];  // the winner's kernel entries for every owned cell, loaded together
#pragma unroll
      for (int k = 0; k < MAXC; ++k) {
        kv[k] = 0.f;
        kpv[k] = 0.f;
        const int c = cell[k];
        if (c >= 0 && Kw) {
          const int dy = c / nx - yw, dx = c % nx - xw;
          if (dy >= -I.kwy && dy <= I.kwy && dx >= -I.kwx && dx <= I.kwx) {
            const size_t o = (size_t)(I.kwy + dy) * KW + (I.kwx + dx);
            kv[k] = Kw[o];
            if (PERP) kpv[k] = Kpw[o];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < MAXC; ++k) {
        const int c = cell[k];
        if (c < 0) continue;
        if (c == w) {
          mm[k] = -mm[k];
          e[k] = -e[k];
          ec[k] = -ec[k];
          if (PERP) {
            ep[k] = -ep[k];
            epc[k] = -epc[k];
          }
          continue;
        }
        if (mm[k] == 0) continue;
        const int y = c / nx, x = c % nx;
        const int dy = y - yw, dx = x - xw;
        float de = 0.f, dep = 0.f;
        const float mmu = (float)mm[k] * mu[k];
        de = A * (mmu * kv[k]);
        if (PERP) dep = A * (mmu * kpv[k]);
        if (Ax != 0.f) {
          int qy = dy, qx = dx;
          if (I.pbc) {
            qy = hs_pmod(dy, I.ny);
            qy = qy > I.ny / 2 ? qy - I.ny : qy;
            qx = hs_pmod(dx, I.nx);
            qx = qx > I.nx / 2 ? qx - I.nx : qx;
          }
          qy += hy;
          qx += hx;
          if (qy >= 0 && qy < I.nn_h && qx >= 0 && qx < I.nn_w) {
            const float wq = I.nn_mask[qy * I.nn_w + qx];
            if (wq != 0.f) {
              if (I.in_plane) {
                const int s = hs_site(I, y, x);
                const float oxi = (float)I.ox_uc[s], oyi = (float)I.oy_uc[s];
                de += Ax * (float)mm[k] * wq * (oxi * oxw + oyi * oyw);
                if (PERP) dep += Ax * (float)mm[k] * wq * (oyw * oxi - oxw * oyi);
              } else {
                de += Ax * (float)mm[k] * wq;
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
    const int c = cell[k];
    if (c < 0) continue;
    I.m[c] = (int8_t)mm[k];
    I.E[c] = e[k];
    I.Ec[c] = ec[k];
    if (PERP) {
      I.Ep[c] = ep[k];
      I.Epc[c] = epc[k];
    }
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

__global__ void k_team_reset(TeamDev *teams, int nteams, int32_t *err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *err = 0;
  if (t >= nteams) return;
  TeamDev d = teams[t];
  *d.bar = 0u;
  d.minslot[0] = d.minslot[1] = d.minslot[2] = ~0ull;
  for (int i = 0; i < 3 * d.G; ++i) d.dtflag[i] = 0u;
}

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
  size_t teams_off, inst_off, map_off, sync_off, err_off, total;
  std::vector<size_t> team_sync;  // offset of each team's sync block
  std::vector<int> G, first;
  int nblocks, maxc;
};

int team_blocks(const hs_team_t &t) {
  const int ncell = t.inst->nx * t.inst->ny;
  if (t.nblocks > 0) return t.nblocks;
  // default: one cell per thread, at most 128 workgroups
  int g = (ncell + TT - 1) / TT;
  return g < 1 ? 1 : (g > 128 ? 128 : g);
}

int make_layout(const hs_team_t *teams, int n, Layout &L) {
  L.teams_off = 0;
  L.inst_off = al256(sizeof(TeamDev) * n);
  L.map_off = L.inst_off + al256(sizeof(hs_instance_t) * n);
  L.nblocks = 0;
  L.maxc = 1;
  L.G.resize(n);
  L.first.resize(n);
  for (int i = 0; i < n; ++i) {
    const int g = team_blocks(teams[i]);
    L.G[i] = g;
    L.first[i] = L.nblocks;
    L.nblocks += g;
    const int ncell = teams[i].inst->nx * teams[i].inst->ny;
    const int c = (ncell + g * TT - 1) / (g * TT);
    while (L.maxc < c) L.maxc *= 2;
  }
  L.sync_off = L.map_off + al256(sizeof(int32_t) * (size_t)L.nblocks);
  size_t off = L.sync_off;
  L.team_sync.resize(n);
  for (int i = 0; i < n; ++i) {
    L.team_sync[i] = off;
    off += 256 + al256(3 * 8) + al256(3 * 8 * (size_t)L.G[i]) + al256(3 * 4 * (size_t)L.G[i]);
  }
  L.err_off = off;
  L.total = off + 256;
  return 0;
}

template <int MAXC>
hipError_t launch_neel(bool perp, int nblocks, hipStream_t s, const TeamDev *td, const int32_t *map, int32_t *err) {
  void *args[] = {(void *)&td, (void *)&map, (void *)&err};
  const void *fn = perp ? (const void *)&k_neel_team<MAXC, true> : (const void *)&k_neel_team<MAXC, false>;
  return hipLaunchCooperativeKernel(fn, dim3(nblocks), dim3(TT), args, 0, s);
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
    if (!t.state) return hs_set_error(HS_ERR_ARG, "team %d has no state", i);
    if (t.run.max_steps < 0) return hs_set_error(HS_ERR_ARG, "negative step budget");
    if (t.run.log && (!t.log_idx || !t.log_dt)) return hs_set_error(HS_ERR_ARG, "team %d: log without arrays", i);
    if (!t.inst->uniform_beta) return hs_set_error(HS_ERR_ARG, "Neel stepping needs a uniform temperature");
    if (i > 0 && (bool)t.inst->use_perp != perp) return hs_set_error(HS_ERR_ARG, "teams of one launch must agree on use_perp");
    perp = t.inst->use_perp;
    if (t.nblocks < 0 || t.nblocks > 4096) return hs_set_error(HS_ERR_ARG, "team %d: bad nblocks", i);
  }
  Layout L;
  make_layout(teams, nteams, L);
  if (ws_bytes < (int64_t)L.total) return hs_set_error(HS_ERR_ARG, "workspace %lld < %lld bytes", (long long)ws_bytes, (long long)L.total);
  if (L.maxc > 16) return hs_set_error(HS_ERR_ARG, "more than 16 cells per thread: give the team more workgroups");
  for (int i = 0; i < nteams; ++i) {
    // the barrier counter counts rounds*G in 32 bits
    if ((double)teams[i].run.max_steps * (double)L.G[i] >= 4.0e9)
      return hs_set_error(HS_ERR_ARG, "team %d: max_steps*nblocks exceeds the 32-bit barrier counter", i);
  }
  hipStream_t s = (hipStream_t)stream;
  int dev = 0;
  HS_HIP_TRY(hipGetDevice(&dev));
  int cus = 0, per = 0;
  HS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const void *fn = nullptr;
  switch (L.maxc) {
    case 1: fn = perp ? (const void *)&k_neel_team<1, true> : (const void *)&k_neel_team<1, false>; break;
    case 2: fn = perp ? (const void *)&k_neel_team<2, true> : (const void *)&k_neel_team<2, false>; break;
    case 4: fn = perp ? (const void *)&k_neel_team<4, true> : (const void *)&k_neel_team<4, false>; break;
    case 8: fn = perp ? (const void *)&k_neel_team<8, true> : (const void *)&k_neel_team<8, false>; break;
    default: fn = perp ? (const void *)&k_neel_team<16, true> : (const void *)&k_neel_team<16, false>; break;
  }
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
    char *sy = base + L.team_sync[i];
    d.bar = (uint32_t *)sy;
    d.minslot = (unsigned long long *)(sy + 256);
    d.dtslot = (double *)(sy + 256 + al256(24));
    d.dtflag = (uint32_t *)(sy + 256 + al256(24) + al256(24 * (size_t)L.G[i]));
    d.G = L.G[i];
    d.first = L.first[i];
    td[i] = d;
    for (int b = 0; b < L.G[i]; ++b) map[L.first[i] + b] = i;
  }
  HS_HIP_TRY(hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, s));
  HS_HIP_TRY(hipStreamSynchronize(s));  // the host descriptor buffer must outlive the copy
  int32_t *err = (int32_t *)(base + L.err_off);
  k_team_reset<<<(nteams + 255) / 256, 256, 0, s>>>((TeamDev *)base, nteams, err);
  HS_LAUNCH_CHECK("k_team_reset");
  hipError_t e;
  const TeamDev *dtd = (const TeamDev *)(base + L.teams_off);
  const int32_t *dmap = (const int32_t *)(base + L.map_off);
  switch (L.maxc) {
    case 1: e = launch_neel<1>(perp, L.nblocks, s, dtd, dmap, err); break;
    case 2: e = launch_neel<2>(perp, L.nblocks, s, dtd, dmap, err); break;
    case 4: e = launch_neel<4>(perp, L.nblocks, s, dtd, dmap, err); break;
    case 8: e = launch_neel<8>(perp, L.nblocks, s, dtd, dmap, err); break;
    default: e = launch_neel<16>(perp, L.nblocks, s, dtd, dmap, err); break;
  }
  if (e != hipSuccess) return hs_set_error(HS_ERR_HIP, "cooperative launch of k_neel_team failed: %s", hipGetErrorString(e));
  return HS_OK;
}
