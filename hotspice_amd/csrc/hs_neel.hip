// Neel-Arrhenius stepping (Magnets._update_Neel, hotspice/core.py:679-709).
//
// One launch per step, fused across two consecutive steps:
//   1. every block reduces the per-block argmin partials of the previous launch to the global
//      winner (lowest key, lowest flat index on ties == nanargmin's first index);
//   2. it applies the switch to the cells it owns: flip + negate the winner's fields, and the
//      rank-1 field update E_i += 2*pf*(m mu)_w*(m mu)_i*K_{n_w}[i-w] (update_single,
//      energies.py:310-329) plus the exchange stencil (energies.py:771-787) and, for E_perp,
//      the K_perpother term;
//   3. it draws -ln(u) per owned cell (Philox, counter = cell) and computes the waiting-time key
//      log(tau) = log(-ln u) + b*beta (or the two-channel form), and writes the block's best.
// The grid-wide minimum of b (core.py:688, 696) only rescales all taus when beta is uniform, so
// it is not needed for the argmin; the winner's dt = -ln(u)*exp(b*beta)/f0 is evaluated in FP64.
// With a temperature that varies over the grid the minimum does reorder the taus
// (tau_i = -ln(u_i) exp((b_i - b_min) beta_i)): each step then takes two launches, the first
// (update, as above) leaving per-block minima of the barriers, the second (k_neel_keys_nu)
// reducing them to b_min and forming the keys. b_min cancels in the winner's time.
// Blocks never wait for each other inside a launch, so there is no grid barrier to hang.
#include <hip/hip_runtime.h>

#include <cmath>

#include "hs_device.h"
#include "hs_internal.h"

namespace {

struct __attribute__((aligned(32))) NeelPartial {
  float key;
  int32_t idx;
  double dt;
  int32_t mw;   // m of the candidate before its switch
  int32_t pad[3];
};

constexpr int NB_THREADS = 256;

// FP32 log-rate of a cell: -b*beta (one channel) or log(exp(-b1 beta) + exp(-b2 beta)).
template <bool PERP>
HS_INLINE float log_rate_f32(const hs_instance_t &I, float e, float ep, float eb, float beta) {
  const float dE = -2.f * e;  // switch_energy = -2E (energies.py:29-35)
  const float El = e;         // E = -delta_E/2 (core.py:635)
  if (PERP) {
    const float epf = (float)I.m_perp_factor * ep;
    const float b1 = fmaxf(dE, (eb + epf) - El), b2 = fmaxf(dE, (eb - epf) - El);  // core.py:651-653
    const float lo = fminf(b1, b2), hi = fmaxf(b1, b2);
    return -lo * beta + hs_log_fast(1.f + __builtin_amdgcn_exp2f(-(hi - lo) * beta * 1.44269504088896f));
  }
  float b;
  if (I.barrier == 0) {
    b = fmaxf(dE, eb - El);  // 'simple' core.py:643
  } else {                   // 'parabolic' core.py:645
    const float q = dE * (0.25f / eb) + 1.f;
    b = dE > 4.f * eb ? dE : fminf(eb * (q * q), dE + 4.f * eb);
  }
  return -b * beta;
}

// FP32 barriers of a cell (core.py:641-655): lo = min(b1, b2) and hi = max(b1, b2) (E_perp),
// else lo = hi = b.
template <bool PERP>
HS_INLINE void barriers_f32(const hs_instance_t &I, float e, float ep, float eb, float &lo, float &hi) {
  const float dE = -2.f * e, El = e;
  if (PERP) {
    const float epf = (float)I.m_perp_factor * ep;
    const float b1 = fmaxf(dE, (eb + epf) - El), b2 = fmaxf(dE, (eb - epf) - El);
    lo = fminf(b1, b2);
    hi = fmaxf(b1, b2);
    return;
  }
  if (I.barrier == 0) {
    lo = fmaxf(dE, eb - El);
  } else {
    const float q = dE * (0.25f / eb) + 1.f;
    lo = dE > 4.f * eb ? dE : fminf(eb * (q * q), dE + 4.f * eb);
  }
  hi = lo;
}

template <bool PERP>
__global__ void __launch_bounds__(NB_THREADS) k_neel_step(hs_instance_t I, hs_run_t R, hs_state_t *st,
                                                        NeelPartial *parts, int32_t *log_idx, double *log_dt,
                                                        int j, int last, float *bparts) {
  __shared__ float s_key[NB_THREADS / 64];
  __shared__ int s_idx[NB_THREADS / 64];
  __shared__ int s_slot[NB_THREADS / 64];
  __shared__ NeelPartial s_win;
  const int P = gridDim.x;
  const int tid = threadIdx.x;
  const hs_state_t S = st[j & 1];
  double t = S.t;
  int64_t att = S.attempts, sw = S.switches, step = S.step, sd = S.steps_done;
  int done = S.done, err = S.err;
  bool flip = false;
  int w = -1, mw = 0;

  // 1. the previous launch's winner: (key, index) argmin over the partials, carrying the slot
  if (!done && j > 0) {
    const NeelPartial *pp = parts + (size_t)((j - 1) & 1) * P;
    float k = INFINITY;
    int idx = 0x7fffffff, slot = -1;
    for (int b = tid; b < P; b += NB_THREADS) {
      const float kb = pp[b].key;
      const int ib = pp[b].idx < 0 ? 0x7fffffff : pp[b].idx;
      if (kb < k || (kb == k && ib < idx)) {
        k = kb;
        idx = ib;
        slot = b;
      }
    }
    hs_wave_argmin_slot(k, idx, slot);
    if ((tid & 63) == 0) {
      s_key[tid >> 6] = k;
      s_idx[tid >> 6] = idx;
      s_slot[tid >> 6] = slot;
    }
    __syncthreads();
    if (tid == 0) {
      float bk = s_key[0];
      int bi = s_idx[0], bs = s_slot[0];
      for (int q = 1; q < NB_THREADS / 64; ++q)
        if (s_key[q] < bk || (s_key[q] == bk && s_idx[q] < bi)) {
          bk = s_key[q];
          bi = s_idx[q];
          bs = s_slot[q];
        }
      NeelPartial win{};
      win.idx = -1;
      if (bi != 0x7fffffff && bs >= 0) win = pp[bs];
      s_win = win;
    }
    __syncthreads();
    const NeelPartial win = s_win;
    if (win.idx < 0) {
      err = 1;  // nanargmin over an all-NaN grid
      done = 1;
    } else {
      // progress passes t_max - dt (core.py:757); update() passes t_max itself, also when t has
      // overflowed to +inf (where t - t0 would be NaN)
      const double tmax_step = hs_step_free(R.first_free, sd) ? R.t_max : R.t_max - (t - S.t0);
      att += 1;
      step += 1;
      sd += 1;
      double add;
      if (win.dt > tmax_step || isnan(win.dt)) {  // core.py:702-704
        add = tmax_step;
      } else {
        add = win.dt;
        sw += 1;
        flip = true;
        w = win.idx;
        mw = win.mw;
      }
      t += add;
      if (R.log && blockIdx.x == 0 && tid == 0) {
        log_idx[sd - 1] = flip ? w : -1;
        log_dt[sd - 1] = add;
      }
    }
  }
  // stop predicate of Magnets._progress_stepwise (core.py:756)
  bool compute = false;
  if (!done) {
    const double n = (double)I.n_occ;
    compute = !last && sd < R.max_steps &&
              (hs_step_free(R.first_free, sd) || (hs_lower_than(t - S.t0, R.t_max) &&
                                              hs_lower_than((double)att / n - (double)S.attempts0 / n, R.mc_max)));
    if (!compute) done = 1;
  }
  if (blockIdx.x == 0 && tid == 0) {
    hs_state_t o = S;
    o.t = t;
    o.attempts = att;
    o.switches = sw;
    o.step = step;
    o.steps_done = sd;
    o.done = done;
    o.err = err;
    st[(j + 1) & 1] = o;
  }
  if (!flip && !compute) return;

  // 2+3. fused field update and waiting-time keys over the cells this block owns
  const int nx = I.nx, ncell = I.nx * I.ny;
  const int KW = 2 * I.kwx + 1;
  const size_t KS = (size_t)KW * (2 * I.kwy + 1);
  int yw = 0, xw = 0, sitew = 0, nw = -1;
  float A = 0.f, Ax = 0.f, oxw = 0.f, oyw = 0.f;
  const float *Kw = nullptr, *Kpw = nullptr;
  if (flip) {
    yw = w / nx;
    xw = w % nx;
    sitew = hs_site(I, yw, xw);
    nw = I.dip_pf != 0.0 ? I.sub_of[sitew] : -1;
    const double mnew = -(double)mw;
    A = (float)(2.0 * I.dip_pf * (mnew * (double)I.mu_src[w]));  // 2*prefactor*(m w)_w
    Ax = (float)(-2.0 * I.J * mnew);                          // exchange: -J m_i * (2 m_w)
    oxw = I.ox_uc[sitew];
    oyw = I.oy_uc[sitew];
    if (nw >= 0) {
      Kw = I.K32 + (size_t)nw * KS;
      if (PERP) Kpw = I.Kpo32 + (size_t)nw * KS;
    }
  }
  const int hy = I.nn_h / 2, hx = I.nn_w / 2;
  float bk = INFINITY;
  int bi = 0x7fffffff;
  for (int c = blockIdx.x * NB_THREADS + tid; c < ncell; c += P * NB_THREADS) {
    int mc = I.m[c];
    float e = I.E[c], ec = I.Ec[c];
    float ep = 0.f, epc = 0.f;
    if (PERP) {
      ep = I.Ep[c];
      epc = I.Epc[c];
    }
    if (flip) {
      if (c == w) {  // m *= -1; E[w] *= -1 for every energy (energies.py:121-122, 324, 329)
        mc = -mc;
        I.m[c] = (int8_t)mc;
        e = -e;
        ec = -ec;
        I.E[c] = e;
        I.Ec[c] = ec;
        if (PERP) {
          ep = -ep;
          epc = -epc;
          I.Ep[c] = ep;
          I.Epc[c] = epc;
        }
      } else if (mc != 0) {
        const int y = c / nx, x = c % nx;
        const int dy = y - yw, dx = x - xw;
        float de = 0.f, dep = 0.f;
        const float mmu = (float)mc * I.mu[c];
        if (Kw && dy >= -I.kwy && dy <= I.kwy && dx >= -I.kwx && dx <= I.kwx) {
          const size_t o = (size_t)(I.kwy + dy) * KW + (I.kwx + dx);
          de = A * (mmu * Kw[o]);
          if (PERP) dep = A * (mmu * Kpw[o]);
        }
        if (Ax != 0.f) {
          int qy = dy + hy, qx = dx + hx;  // NN[i + h - w]: q = (i - w) + h
          if (I.pbc) {
            qy = hs_pmod(dy, I.ny);
            qy = qy > I.ny / 2 ? qy - I.ny : qy;
            qx = hs_pmod(dx, I.nx);
            qx = qx > I.nx / 2 ? qx - I.nx : qx;
            qy += hy;
            qx += hx;
          }
          if (qy >= 0 && qy < I.nn_h && qx >= 0 && qx < I.nn_w) {
            const float wq = I.nn_mask[qy * I.nn_w + qx];
            if (wq != 0.f) {
              if (I.in_plane) {
                const int s = hs_site(I, y, x);
                const float oxi = I.ox_uc[s], oyi = I.oy_uc[s];
                de += Ax * (float)mc * wq * (oxi * oxw + oyi * oyw);
                if (PERP) dep += Ax * (float)mc * wq * (oyw * oxi - oxw * oyi);
              } else {
                de += Ax * (float)mc * wq;
              }
            }
          }
        }
        if (de != 0.f) {
          hs_kahan_add(e, ec, de);
          I.E[c] = e;
          I.Ec[c] = ec;
        }
        if (PERP && dep != 0.f) {
          hs_kahan_add(ep, epc, dep);
          I.Ep[c] = ep;
          I.Epc[c] = epc;
        }
      }
    }
    if (compute && (I.flags[c] & HS_FLAG_NEEL)) {
      if (bparts) {  // varying temperature: this step's keys wait for the grid's minimum barrier
        float lo, hi;
        barriers_f32<PERP>(I, e, ep, I.EB[c], lo, hi);
        bk = fminf(bk, lo);  // nanmin: NaN barriers drop out
        continue;
      }
      uint32_t wa, wb;
      hs_words(I.seed, (uint64_t)step, 1u, (uint64_t)c, wa, wb);
      const float x = hs_exp1_fast(wa, wb);
      float key = hs_log_fast(x) - log_rate_f32<PERP>(I, e, ep, I.EB[c], I.beta[c]);
      if (isnan(key)) key = INFINITY;
      hs_argmin_pair(bk, bi, key, c);
    }
  }
  if (!compute) return;
  if (bparts) {
    bk = hs_block_min_f32(bk, s_key);
    if (tid == 0) bparts[blockIdx.x] = bk;
    return;
  }
  hs_wave_argmin(bk, bi);
  __syncthreads();
  if ((tid & 63) == 0) {
    s_key[tid >> 6] = bk;
    s_idx[tid >> 6] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    float k = s_key[0];
    int i = s_idx[0];
    for (int q = 1; q < NB_THREADS / 64; ++q) hs_argmin_pair(k, i, s_key[q], s_idx[q]);
    NeelPartial p{};
    p.key = k;
    p.idx = (i == 0x7fffffff) ? -1 : i;
    p.dt = NAN;
    p.mw = 0;
    if (p.idx >= 0) {
      const int c = p.idx;
      uint32_t wa, wb;
      hs_words(I.seed, (uint64_t)step, 1u, (uint64_t)c, wa, wb);
      const double ep64 = PERP ? (double)I.Ep[c] - (double)I.Epc[c] : 0.0;
      p.dt = hs_neel_dt((double)I.E[c] - (double)I.Ec[c], ep64, (double)I.EB[c], I.beta64, I.m_perp_factor,
                        I.barrier, PERP, hs_exp1_fast(wa, wb), R.attempt_freq);
      p.mw = I.m[c];
    }
    parts[(size_t)(j & 1) * P + blockIdx.x] = p;
  }
}

// Second launch of a step with a varying temperature: b_min from the first launch's per-block
// minima, then the keys log(tau_i) = log(-ln u_i) + (b_i - b_min) beta_i (one channel) or
// log(-ln u_i) + (lo_i - b_min) beta_i - log(1 + exp(-(hi_i - lo_i) beta_i)) (two channels,
// core.py:686-699) and the block's argmin, with the candidate's FP64 time from its own beta.
template <bool PERP>
__global__ void __launch_bounds__(NB_THREADS) k_neel_keys_nu(hs_instance_t I, hs_run_t R, const hs_state_t *st,
                                                           NeelPartial *parts, const float *bparts, int j) {
  __shared__ float s_key[NB_THREADS / 64];
  __shared__ int s_idx[NB_THREADS / 64];
  const int P = gridDim.x, tid = threadIdx.x;
  const hs_state_t S = st[(j + 1) & 1];
  if (S.done) return;
  float bmin = INFINITY;
  for (int b = tid; b < P; b += NB_THREADS) bmin = fminf(bmin, bparts[b]);
  bmin = hs_block_min_f32(bmin, s_key);
  const uint64_t step = (uint64_t)S.step;
  const int ncell = I.nx * I.ny;
  float bk = INFINITY;
  int bi = 0x7fffffff;
  for (int c = blockIdx.x * NB_THREADS + tid; c < ncell; c += P * NB_THREADS) {
    if (!(I.flags[c] & HS_FLAG_NEEL)) continue;
    const float e = I.E[c], ep = PERP ? I.Ep[c] : 0.f, beta = I.beta[c];
    float lo, hi;
    barriers_f32<PERP>(I, e, ep, I.EB[c], lo, hi);
    uint32_t wa, wb;
    hs_words(I.seed, step, 1u, (uint64_t)c, wa, wb);
    float key = hs_log_fast(hs_exp1_fast(wa, wb)) + (lo - bmin) * beta;
    if (PERP) key -= hs_log_fast(1.f + __builtin_amdgcn_exp2f(-(hi - lo) * beta * 1.44269504088896f));
    if (isnan(key)) key = INFINITY;
    hs_argmin_pair(bk, bi, key, c);
  }
  hs_wave_argmin(bk, bi);
  __syncthreads();
  if ((tid & 63) == 0) {
    s_key[tid >> 6] = bk;
    s_idx[tid >> 6] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    float k = s_key[0];
    int i = s_idx[0];
    for (int q = 1; q < NB_THREADS / 64; ++q) hs_argmin_pair(k, i, s_key[q], s_idx[q]);
    NeelPartial p{};
    p.key = k;
    p.idx = (i == 0x7fffffff) ? -1 : i;
    p.dt = NAN;
    p.mw = 0;
    if (p.idx >= 0) {
      const int c = p.idx;
      uint32_t wa, wb;
      hs_words(I.seed, step, 1u, (uint64_t)c, wa, wb);
      const double ep64 = PERP ? (double)I.Ep[c] - (double)I.Epc[c] : 0.0;
      p.dt = hs_neel_dt((double)I.E[c] - (double)I.Ec[c], ep64, (double)I.EB[c], I.beta64c[c], I.m_perp_factor,
                        I.barrier, PERP, hs_exp1_fast(wa, wb), R.attempt_freq);
      p.mw = I.m[c];
    }
    parts[(size_t)(j & 1) * P + blockIdx.x] = p;
  }
}

}  // namespace

__global__ void k_state_begin(hs_state_t *st, int begin) {
  hs_state_t s = st[0];
  if (begin) {
    s.t0 = s.t;
    s.attempts0 = s.attempts;
  }
  s.done = 0;
  s.steps_done = 0;
  s.err = 0;
  st[0] = s;
}

__global__ void k_state_copy(hs_state_t *st, int from) { st[0] = st[from]; }

int hs_neel_blocks(const hs_instance_t *I, int requested) {
  const int ncell = I->nx * I->ny;
  int nb = requested > 0 ? requested : (ncell + NB_THREADS - 1) / NB_THREADS;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  return nb;
}

extern "C" int hs_neel_run(const hs_instance_t *inst, const hs_run_t *run, hs_state_t *state, void *partials,
                           int32_t *log_idx, double *log_dt, int32_t *final_copy, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  rc = hs_check_f32_fields(inst, "hs_neel_run");
  if (rc) return rc;
  if (!run || !state || !partials) return hs_set_error(HS_ERR_ARG, "missing run/state/partials");
  if (run->max_steps < 0) return hs_set_error(HS_ERR_ARG, "negative step budget");
  if (run->log && (!log_idx || !log_dt)) return hs_set_error(HS_ERR_ARG, "log requested without arrays");
  if (run->nblocks < 0 || run->nblocks > 1024) return hs_set_error(HS_ERR_ARG, "nblocks must be in [0, 1024]");
  const bool nu = !inst->uniform_beta;
  if (nu && !inst->beta64c)
    return hs_set_error(HS_ERR_ARG, "Neel stepping with a varying temperature needs beta64c (FP64 1/(kB T) per cell)");
  if (!nu && !(inst->beta64 > 0.0)) return hs_set_error(HS_ERR_ARG, "Neel stepping needs beta64 > 0 (FP64 1/(kB T))");
  hipStream_t s = (hipStream_t)stream;
  const int P = hs_neel_blocks(inst, run->nblocks);
  // the barrier minima of a varying-temperature step follow the two partial buffers
  float *bparts = nu ? (float *)((NeelPartial *)partials + 2 * (size_t)P) : nullptr;
  k_state_begin<<<1, 1, 0, s>>>(state, run->begin ? 1 : 0);
  HS_LAUNCH_CHECK("k_state_begin");
  const hs_run_t R = *run;
  const int64_t N = run->max_steps;
  for (int64_t j = 0; j <= N; ++j) {
    const int last = (j == N);
    if (inst->use_perp)
      k_neel_step<true><<<P, NB_THREADS, 0, s>>>(*inst, R, state, (NeelPartial *)partials, log_idx, log_dt, (int)j, last,
                                                 bparts);
    else
      k_neel_step<false><<<P, NB_THREADS, 0, s>>>(*inst, R, state, (NeelPartial *)partials, log_idx, log_dt, (int)j, last,
                                                  bparts);
    HS_LAUNCH_CHECK("k_neel_step");
    if (nu && !last) {
      if (inst->use_perp)
        k_neel_keys_nu<true><<<P, NB_THREADS, 0, s>>>(*inst, R, state, (NeelPartial *)partials, bparts, (int)j);
      else
        k_neel_keys_nu<false><<<P, NB_THREADS, 0, s>>>(*inst, R, state, (NeelPartial *)partials, bparts, (int)j);
      HS_LAUNCH_CHECK("k_neel_keys_nu");
    }
  }
  const int from = (int)((N + 1) & 1);
  if (from != 0) {
    k_state_copy<<<1, 1, 0, s>>>(state, from);
    HS_LAUNCH_CHECK("k_state_copy");
  }
  if (final_copy) *final_copy = 0;
  return HS_OK;
}
