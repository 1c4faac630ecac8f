// Metropolis stepping with grid multisampling (Magnets._update_Metropolis, core.py:659-677).
//
// Two launches per step, for one instance or for a batch of independent instances ("jobs"):
//   k_metro_select (one workgroup per job): the stop predicate of progress(), the supergrid
//     selection of _select_grid (core.py:496-558) with its retries (core.py:484-487) and the
//     _select_single fallback (core.py:489-494, 555-558), acceptance min(1, e^{-beta dE})
//     clipped to [1e-10, 1e10] (core.py:666, 671), optional METROPOLIS_TIME (core.py:667-669),
//     the flips with E[w] *= -1 (energies.py:124-126, 332-334), and the order-preserving list of
//     accepted switches with per-sublattice counts (energies.py:335-343);
//   k_metro_update (32x8 output tiles): the sparse-source correlation
//     of the accepted switches with the dipolar kernel -- full window, or the (2k+1)^2
//     REDUCED_KERNEL_SIZE window with fill/wrap boundary when a sublattice has more than
//     SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF switches (energies.py:343-361) -- plus the
//     exchange stencil, added once per cell with Kahan compensation (energies.py:362-366).
//     A tile finds its sources directly from the supercell grid (the samples of the supercells
//     whose cells lie within the window of the tile), or, where that is not shorter (full
//     kernels, PBC, single-sample steps), by filtering the whole list through LDS. Both visit
//     the sources in selection order, so every cell sums the same terms in the same order.
//     The selection marks the tiles that the windows of its accepted switches reach (and the
//     tile of each switch, which applies its flip) and appends them to a worklist; the update
//     launch is a fixed grid that works through that list, so a step costs in proportion to the
//     tiles it changes, not to the grid.
// A batch runs every job's step in the same two launches; each job keeps its own stop
// predicate, so a job alone and a job inside a batch execute identical arithmetic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "hs_device.h"
#include "hs_internal.h"

namespace {

constexpr int MS_THREADS = 1024;
constexpr int MU_TX = 32, MU_TY = 8, MU_THREADS = MU_TX * MU_TY;
constexpr int MAX_ATTEMPTS = 4096;
// instances up to this many grid cells step in the persistent one-workgroup kernel
constexpr int PERSIST_CELLS = 80 * 80;
// workgroups of an update launch over the worklist when the caller leaves it automatic
// (hs_run_t.nblocks = 0; 6 per CU of a 256-CU device), and the waves per SIMD the update kernels
// are compiled for, so that all of them are resident at once (6: 80 VGPRs, no spills; measured
// faster than 8 waves with spills, 5 or 4 waves)
constexpr int UPDATE_WGS = 1536;
constexpr int UPDATE_WAVES = 6;

// Workgroups of an update launch over `ntiles` tiles: `requested` (hs_run_t.nblocks) when
// positive, else UPDATE_WGS; never more than one per tile.
inline unsigned update_grid(int64_t ntiles, int64_t requested) {
  const int64_t g = requested > 0 ? requested : UPDATE_WGS;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, std::max<int64_t>(1, ntiles)));
}
// sources whose kernel entries a cell of the persistent kernel's update loads together
constexpr int PERSIST_SRC_BATCH = 8;
// tile bitmaps of at most this many words (32768 tiles of 32x8 cells) are listed by the last
// presampling workgroup (a few words per thread); larger grids list their tiles with k_metro_tiles
constexpr int BITMAP_MAX_WORDS = 1024;
// a switch whose window spans more tiles than this lists every tile of its job instead
constexpr int MARK_MAX_TILES = 256;
// the selection marks tiles in an LDS bitmap of this many 32-bit words (jobs with more tiles list
// every tile)
constexpr int MARK_WORDS = 4096;
// samples per thread whose loads the selection kernels issue together (and keep in registers)
constexpr int SELECT_BATCH = 2;
// workgroup size of the presampling kernel (one sample per thread)
constexpr int PRESAMPLE_THREADS = 256;
// presampled candidates per selection thread kept in registers (up to 4096 samples)
constexpr int PRESELECT_BATCH = 4;
// steps with at most this many samples (one per selection thread) select in one workgroup: below
// it the extra launches of the presampled step cost more than they save
constexpr int PRESAMPLE_MIN = MS_THREADS;

struct MetroWS {
  int32_t *cand;     // [nsc] candidate cell, -1 if not a valid sample
  int8_t *accf;      // [nsc] accepted flag
  int32_t *acc_cell; // [nsc] accepted cells, selection order
  float *acc_mmu;    // [nsc] (m mu) after the flip
  int32_t *acc_n;    // [1]  number accepted this step
  int32_t *trunc;    // [HS_MAX_UC] per unit-cell site: 1 = truncated window (convolution branch)
  int32_t *info;     // [5] grid offset ox, oy of the selection that succeeded, 1 = single sample;
                     // [3] the step a presampling launch finished (FUSED_MARK), [4] 1 = the update
                     // takes its sources from the sample array only (no ordered list was written)
  int32_t *fz;       // [8 + HS_MAX_UC] presampling launch: ticket, in-grid / valid / accepted
                     // samples, then accepted samples per unit-cell site (zero between steps)
  float *pax, *pay;  // [nsc] Poisson: accepted point of each background cell (NaN = none)
  float *pcx, *pcy;  // [nsc] Poisson: this round's candidate point
  uint32_t *pcp;     // [nsc] Poisson: its priority (0xffffffff = no candidate)
  int32_t *pwin;     // [nsc] Poisson: 1 = this round's candidate is accepted
  float *cmmu;       // [nsc] (m mu) of each candidate after its flip (read where accf is set)
  long long *part;   // [nsc/PRESAMPLE_THREADS] presampling: in-grid samples << 32 | valid samples, per workgroup
};

// The step's (job, tile) worklist of an update launch: count[par] entries in wl. The selection
// of step parity `par` appends; the update of that step reads them and zeroes count[par ^ 1]
// for the next step.
struct WorkList {
  int2 *wl;
  int32_t *count;
  // Tile bitmaps of a presampled single-instance step (two, by step parity; nwords words each), or
  // NULL: the presampling launch marks the tiles its accepted switches reach and its last
  // workgroup lists them, so the step needs no tile-listing launch.
  uint32_t *bits = nullptr;
  int32_t nwords = 0;
};

// Everything one job's step needs (passed by value for one instance, from an array for a batch).
struct MetroJob {
  hs_instance_t I;
  hs_select_t SEL;
  hs_run_t R;
  hs_state_t *st;
  MetroWS ws;
  int32_t *log_cells, *log_count;
  int32_t tiles_x, tiles_y;
};

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

MetroWS carve(void *base, int nsc) {
  char *p = (char *)base;
  MetroWS w;
  const size_t n = (size_t)(nsc > 0 ? nsc : 1);
  w.cand = (int32_t *)p;
  p += align256(n * 4);
  w.accf = (int8_t *)p;
  p += align256(n);
  w.acc_cell = (int32_t *)p;
  p += align256(n * 4);
  w.acc_mmu = (float *)p;
  p += align256(n * 4);
  w.acc_n = (int32_t *)p;
  p += 256;
  w.trunc = (int32_t *)p;
  p += align256(HS_MAX_UC * 4);
  w.info = (int32_t *)p;
  p += 256;
  w.fz = (int32_t *)p;
  p += 512;
  w.pax = (float *)p;
  p += align256(n * 4);
  w.pay = (float *)p;
  p += align256(n * 4);
  w.pcx = (float *)p;
  p += align256(n * 4);
  w.pcy = (float *)p;
  p += align256(n * 4);
  w.pcp = (uint32_t *)p;
  p += align256(n * 4);
  w.pwin = (int32_t *)p;
  p += align256(n * 4);
  w.cmmu = (float *)p;
  p += align256(n * 4);
  w.part = (long long *)p;
  p += align256(8 * ((n + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS));
  return w;
}

// Per-job scratch (the MetroWS arrays).
size_t ws_bytes(int nsc) {
  const size_t n = (size_t)(nsc > 0 ? nsc : 1);
  return align256(n * 4) + align256(n) + align256(n * 4) + align256(n * 4) + 256 + align256(HS_MAX_UC * 4) + 256 + 512 +
         7 * align256(n * 4) + align256(8 * ((n + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS));
}

// A worklist for `ntiles` tiles: the entries, the two step counters and the two tile bitmaps.
inline int64_t tile_words(int64_t ntiles) { return (ntiles + 31) / 32; }
size_t wl_bytes(int64_t ntiles) {
  return align256(8 * (size_t)(ntiles > 0 ? ntiles : 1)) + 256 + 2 * align256(4 * (size_t)std::max<int64_t>(1, tile_words(ntiles)));
}

int tiles_of(const hs_instance_t &in) { return ((in.nx + MU_TX - 1) / MU_TX) * ((in.ny + MU_TY - 1) / MU_TY); }

// Block-wide sum of an int (1024 threads).
HS_INLINE int block_sum_int(int v, int *s_tmp) {
  v = hs_wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_tmp[w];
  return t;
}

HS_INLINE long long block_sum_i64(long long v, long long *s_tmp) {
  v = hs_wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = v;
  __syncthreads();
  long long t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_tmp[w];
  return t;
}

HS_INLINE double block_sum_double(double v, double *s_tmp) {
  v = hs_wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_tmp[w];
  return t;
}

// Exclusive prefix sum over the block (in thread order).
HS_INLINE int block_exclusive_scan(int v, int *s_tmp, int &total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  __syncthreads();
  if (lane == 63) s_tmp[wid] = incl;
  __syncthreads();
  int base = 0;
  total = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    if (w < wid) base += s_tmp[w];
    total += s_tmp[w];
  }
  return base + incl - v;
}

// min-barrier of a sample for METROPOLIS_TIME (E_barrier(idx, min_only=True), core.py:625-656), FP64.
HS_INLINE double barrier_min_f64(const hs_instance_t &I, int c) {
  const double e = I.E64 ? I.E64[c] : (double)I.E[c] - (double)I.Ec[c];
  const double dE = -2.0 * e, El = e, eb = (double)I.EB[c];
  if (I.use_perp) {
    const double ep = I.m_perp_factor * ((double)I.Ep[c] - (double)I.Epc[c]);
    return fmin(fmax(dE, (eb + ep) - El), fmax(dE, (eb - ep) - El));
  }
  if (I.barrier == 0) return fmax(dE, eb - El);
  const double q = dE * (0.25 / eb) + 1.0;
  return dE > 4.0 * eb ? dE : fmin(eb * (q * q), dE + 4.0 * eb);
}

// Poisson-disc selection (the scheme of Magnets._select_Poisson, core.py:560-570, with the
// minimal distance r of its docstring, core.py:464-467): parallel dart throwing on a background
// grid of cells (r_x, r_y)/sqrt(2), so a cell holds at most one sample and every conflict lies
// within +-2 cells. Each round, every empty background cell proposes a uniform point with a
// Philox priority; a proposal survives if no accepted sample and no proposal of lower priority
// (lower cell index on ties) lies within the ellipse (dx/r_x)^2 + (dy/r_y)^2 < 1 (wrapped with
// PBC). Four rounds; the samples are then snapped to the nearest grid cell and kept where
// occupied (Magnets.select, core.py:484-487). Writes cand[b] for every background cell b and
// returns the number of valid samples (block-wide).
__device__ int poisson_select(const hs_instance_t &I, const hs_select_t &SEL, const MetroWS &ws, uint64_t step,
                              int attempt, int *s_tmp) {
  const int tid = threadIdx.x;
  const int gx = SEL.snx, gy = SEL.sny, nbg = SEL.nsc;
  const float rx = (float)SEL.r_x, ry = (float)SEL.r_y;
  const float sx = rx * 0.70710678f, sy = ry * 0.70710678f;
  const float nxf = (float)I.nx, nyf = (float)I.ny;
  auto close = [&](float x1, float y1, float x2, float y2) {
    float dx = x1 - x2, dy = y1 - y2;
    if (I.pbc) {
      dx -= nxf * rintf(dx / nxf);
      dy -= nyf * rintf(dy / nyf);
    }
    const float qx = dx / rx, qy = dy / ry;
    return qx * qx + qy * qy < 1.f;
  };
  for (int b = tid; b < nbg; b += blockDim.x) ws.pax[b] = NAN;
  __syncthreads();
  for (int round = 0; round < 4; ++round) {
    for (int b = tid; b < nbg; b += blockDim.x) {
      uint32_t pr = 0xffffffffu;
      float cx = 0.f, cy = 0.f;
      if (isnan(ws.pax[b])) {
        const uint4 w = hs_philox(make_uint4((uint32_t)b, (uint32_t)step, (uint32_t)(step >> 32),
                                             6u | ((uint32_t)attempt << 8) | ((uint32_t)round << 24)),
                                  (uint32_t)I.seed, (uint32_t)(I.seed >> 32));
        cx = ((float)(b % gx) + (float)w.x * 0x1p-32f) * sx;
        cy = ((float)(b / gx) + (float)w.y * 0x1p-32f) * sy;
        if (cx < nxf && cy < nyf) pr = w.z == 0xffffffffu ? 0xfffffffeu : w.z;
      }
      ws.pcx[b] = cx;
      ws.pcy[b] = cy;
      ws.pcp[b] = pr;
    }
    __syncthreads();
    for (int b = tid; b < nbg; b += blockDim.x) {
      const uint32_t pr = ws.pcp[b];
      bool win = pr != 0xffffffffu;
      const int bx = b % gx, by = b / gx;
      const float cx = ws.pcx[b], cy = ws.pcy[b];
      for (int oy = -2; oy <= 2 && win; ++oy)
        for (int ox = -2; ox <= 2 && win; ++ox) {
          int qx = bx + ox, qy = by + oy;
          if (I.pbc) {
            qx = hs_pmod(qx, gx);
            qy = hs_pmod(qy, gy);
          } else if (qx < 0 || qx >= gx || qy < 0 || qy >= gy) {
            continue;
          }
          const int q = qy * gx + qx;
          if (q == b) continue;
          const float ax = ws.pax[q];
          if (!isnan(ax) && close(cx, cy, ax, ws.pay[q])) win = false;
          const uint32_t pq = ws.pcp[q];
          if (pq != 0xffffffffu && (pq < pr || (pq == pr && q < b)) && close(cx, cy, ws.pcx[q], ws.pcy[q])) win = false;
        }
      ws.pwin[b] = win ? 1 : 0;  // applied after every cell has read this round's points
    }
    __syncthreads();
    for (int b = tid; b < nbg; b += blockDim.x)
      if (ws.pwin[b]) {
        ws.pax[b] = ws.pcx[b];
        ws.pay[b] = ws.pcy[b];
      }
    __syncthreads();
  }
  int val = 0;
  for (int b = tid; b < nbg; b += blockDim.x) {
    int c = -1;
    const float ax = ws.pax[b];
    if (!isnan(ax)) {
      const int x = min(I.nx - 1, (int)floorf(ax + 0.5f)), y = min(I.ny - 1, (int)floorf(ws.pay[b] + 0.5f));
      if (I.flags[y * I.nx + x] & HS_FLAG_OCC) c = y * I.nx + x;
    }
    ws.cand[b] = c;
    val += c >= 0;
  }
  return block_sum_int(val, s_tmp);
}

// Tile ranges [t0, t1] (up to three, the images k*n with wrap) of the cells within w of s along
// one axis of n cells and tiles of T cells: exactly the tiles for which metro_update's
// reaches(s, w, n, wrap, ...) holds with an actual cell in reach.
HS_INLINE int tile_ranges(int s, int w, int n, bool wrap, int T, int t0[3], int t1[3]) {
  int cnt = 0;
  if (w < 0) return 0;
  for (int k = wrap ? -1 : 0; k <= (wrap ? 1 : 0); ++k) {
    const int lo = max(0, s + k * n - w), hi = min(n - 1, s + k * n + w);
    if (lo > hi) continue;
    t0[cnt] = lo / T;
    t1[cnt] = hi / T;
    ++cnt;
  }
  return cnt;
}

// Marks the tiles of the rectangle (sx +- wx) x (sy +- wy) in the bitmap; returns false (marking
// nothing) when it spans more than MARK_MAX_TILES tiles.
HS_INLINE bool mark_rect(uint32_t *bits, int tiles_x, int nx, int ny, int sx, int wx, int sy, int wy, bool wrap) {
  int x0[3], x1[3], y0[3], y1[3];
  const int nxr = tile_ranges(sx, wx, nx, wrap, MU_TX, x0, x1);
  const int nyr = tile_ranges(sy, wy, ny, wrap, MU_TY, y0, y1);
  int span = 0;
  for (int a = 0; a < nyr; ++a)
    for (int b = 0; b < nxr; ++b) span += (y1[a] - y0[a] + 1) * (x1[b] - x0[b] + 1);
  if (span > MARK_MAX_TILES) return false;
  for (int a = 0; a < nyr; ++a)
    for (int ty = y0[a]; ty <= y1[a]; ++ty)
      for (int b = 0; b < nxr; ++b)
        for (int tx = x0[b]; tx <= x1[b]; ++tx) {
          const int t = ty * tiles_x + tx;
          atomicOr(&bits[t >> 5], 1u << (t & 31));
        }
  return true;
}

// One Metropolis step's selection for one job (one workgroup). Each thread takes samples
// [j0, j1) and works on them in batches of QB: within a batch the independent loads of all its
// samples are issued together, and when a thread has at most QB samples their candidate cells,
// acceptance flags and (m mu) stay in registers from phase to phase instead of being re-read.
// The grid selection's supergrid offset of one attempt (core.py:527-531).
HS_INLINE void grid_offsets(const hs_instance_t &I, const hs_select_t &SEL, uint64_t b0, uint64_t b1, int &ox,
                            int &oy) {
  if (I.pbc) {  // np.random.randint(nx), randint(ny) (core.py:527)
    ox = (int)floor(hs_u53(b0, 0) * I.nx);
    oy = (int)floor(hs_u53(b1, 0) * I.ny);
  } else {  // randint(-Rx, Rx), randint(-Ry, Ry) (core.py:529)
    ox = -SEL.Rx + (int)floor(hs_u53(b0, 0) * (2.0 * SEL.Rx));
    oy = -SEL.Ry + (int)floor(hs_u53(b1, 0) * (2.0 * SEL.Ry));
  }
  if (SEL.Rx == I.nx) ox = 0;  // core.py:530-531
  if (SEL.Ry == I.ny) oy = 0;
}

// The cell of sample j (chosen supercell j of the every-other grid) for pattern cell pc, and
// whether it lies in the grid (core.py:540-552).
HS_INLINE int grid_cell(const hs_instance_t &I, const hs_select_t &SEL, int ox, int oy, int j, int pc, bool &ok) {
  const int nx = I.nx, ny = I.ny;
  const int sy = 2 * (j / SEL.ncols), sx = 2 * (j % SEL.ncols);
  int x = sx * SEL.Rx + (pc & 0xffff), y = sy * SEL.Ry + (pc >> 16);
  if (I.pbc) {
    ok = (x < max(SEL.Rx, nx - SEL.Rx)) && (y < max(SEL.Ry, ny - SEL.Ry));  // core.py:545
    x = (x + ox) % nx;
    y = (y + oy) % ny;
  } else {
    x += ox;
    y += oy;
    ok = x >= 0 && x < nx && y >= 0 && y < ny;  // core.py:552
  }
  return y * nx + x;
}

// Metropolis acceptance of a sample at cell c (core.py:662-671): exp(-switch_energy*beta) with
// switch_energy = -2E from the compensated field, relative error ~3e-7 (FP64 exponent split, FP32
// mantissa), clipped to [1e-10, 1e10], against the Philox uniform (cell, step, tag 4).
HS_INLINE bool metro_accepts_e(const hs_instance_t &I, uint64_t step, int c, double e, float beta);
HS_INLINE bool metro_accepts(const hs_instance_t &I, uint64_t step, int c, float E, float Ec, float beta) {
  return metro_accepts_e(I, step, c, (double)E - (double)Ec, beta);
}
// The same from the field's value e (FP64 fields: E64).
// With FP64 fields the probability is the FP64 exp of the FP64 product (beta64 / beta64c when
// present), as the reference evaluates it.
HS_INLINE bool metro_accepts_e(const hs_instance_t &I, uint64_t step, int c, double e, float beta) {
  double p;
  if (I.E64)
    p = exp(2.0 * e * (I.beta64c ? I.beta64c[c] : (I.uniform_beta ? I.beta64 : (double)beta)));
  else
    p = hs_scaled_exp(1.f, 2.0 * e * (double)beta, 1.f, 1.0);
  p = fmin(fmax(p, 1e-10), 1e10);
  const double ua = hs_u53(hs_bits53(I.seed, step, 4u, (uint64_t)c), 0);
  return ua < p;
}

// info[3] == FUSED_DONE: the presampling launch of this step completed it (that launch clears the
// mark first thing, so a mark never outlives its step).
constexpr int FUSED_DONE = 1;

// LIST_TILES (with PRE): the selection lists the reached tiles itself, for the fallback run by
// k_metro_tiles when the presampled attempt did not complete the step. NT: the workgroup size
// (the selection kernels use MS_THREADS; the tile kernel's fallback runs with its own 256).
template <int QB, bool PRE = false, bool LIST_TILES = false, int NT = MS_THREADS>
__device__ __forceinline__ int metro_select(const MetroJob &J, WorkList WL = WorkList{nullptr, nullptr}, int par = 0,
                                             int job = 0) {
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const hs_run_t &R = J.R;
  hs_state_t *st = J.st;
  const MetroWS &ws = J.ws;
  int32_t *log_cells = J.log_cells, *log_count = J.log_count;
  __shared__ int s_i[NT / 64];
  __shared__ long long s_l[NT / 64];
  __shared__ double s_d[NT / 64];
  __shared__ int s_site[HS_MAX_UC];
  __shared__ int s_single;
  __shared__ float s_single_mmu;
  __shared__ int s_all, s_base;
  __shared__ uint32_t s_bits[MARK_WORDS];
  const int tid = threadIdx.x;
  hs_state_t S = st[0];
  const double n = (double)I.n_occ;
  if (PRE && ws.info[3] == FUSED_DONE) return *ws.acc_n;  // the presampling launch completed this step
  if (tid < HS_MAX_UC) s_site[tid] = 0;
  bool go = !S.done && S.steps_done < R.max_steps &&
            (hs_step_free(R.first_free, S.steps_done) ||
             (hs_lower_than(S.t - S.t0, R.t_max) && hs_lower_than((double)S.attempts / n - (double)S.attempts0 / n, R.mc_max)));
  if (!go) {
    __syncthreads();
    if (tid == 0) {
      *ws.acc_n = 0;
      if (!S.done) {
        S.done = 1;
        st[0] = S;
      }
    }
    return -1;
  }
  const uint64_t step = (uint64_t)S.step;
  const int nsc = SEL.nsc;
  const int q = (nsc + NT - 1) / NT;
  const int j0 = min(nsc, tid * q), j1 = min(nsc, j0 + q);
  const int nx = I.nx, ny = I.ny;
  // all of this thread's grid samples fit one batch: they stay in registers (the Poisson scheme
  // writes its candidates to the workspace only)
  const bool regs = q <= QB && SEL.scheme == 0;
  int candr[QB], accr[QB];
  float mmur[QB];
#pragma unroll
  for (int u = 0; u < QB; ++u) {
    candr[u] = -1;
    accr[u] = 0;
    mmur[u] = 0.f;
  }
  int nvalid = 0;
  bool single = false;
  int attempt = 0;
  int ox = 0, oy = 0;
  // PRE: k_metro_presample drew attempt 0's samples and their acceptance already (cand, accf,
  // cmmu in the workspace, per-workgroup counts in part); used when that attempt has valid samples
  bool pre = false;
  if (PRE) {
    long long mine = 0;
    for (int b = tid; b < (nsc + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS; b += NT) mine += ws.part[b];
    if (regs) {  // this thread's presampled candidates, loaded together with the counts
#pragma unroll
      for (int u = 0; u < QB; ++u) {
        const int j = j0 + u;
        if (j < j1) {
          candr[u] = ws.cand[j];
          accr[u] = ws.accf[j];
          mmur[u] = ws.cmmu[j];
        }
      }
    }
    const long long both = block_sum_i64(mine, s_l);
    const int tot_inb = (int)(both >> 32), tot_val = (int)(both & 0xffffffffll);
    pre = tot_inb > 0 && tot_val > 0;
    if (pre) {
      nvalid = tot_val;
      grid_offsets(I, SEL, hs_bits53(I.seed, step, 2u, 0ull), hs_bits53(I.seed, step, 2u, 1ull), ox, oy);
    }
  }
  bool bad_list = false;
  for (; !pre && attempt < MAX_ATTEMPTS; ++attempt) {
    const uint64_t b0 = hs_bits53(I.seed, step, 2u, (uint64_t)(4 * attempt + 0));
    const uint64_t b1 = hs_bits53(I.seed, step, 2u, (uint64_t)(4 * attempt + 1));
    const uint64_t b2 = hs_bits53(I.seed, step, 2u, (uint64_t)(4 * attempt + 2));
    if (SEL.scheme == 3) {  // the caller's samples (update(idx=...), core.py:660-661), in its order
      int bad = 0;
      for (int j = j0; j < j1; ++j) {
        const int c = SEL.pat_cells[j];
        const bool ok = c >= 0 && c < nx * ny && (I.flags[c] & HS_FLAG_OCC);
        ws.cand[j] = ok ? c : -1;
        bad += ok ? 0 : 1;
      }
      bad_list = block_sum_i64(bad, s_l) > 0;
      nvalid = nsc;
      break;
    }
    if (SEL.scheme == 2) {
      const int tot_val = poisson_select(I, SEL, ws, step, attempt, s_i);
      if (tot_val > 0) {
        nvalid = tot_val;
        break;
      }
      continue;  // no occupied sample: select() again (core.py:487)
    }
    if (SEL.scheme == 0) {
      grid_offsets(I, SEL, b0, b1, ox, oy);
      const int pidx = hs_pmod(oy, I.uy) * I.ux + hs_pmod(ox, I.ux);
      const int p0 = SEL.pat_off[pidx], cnt = SEL.pat_off[pidx + 1] - p0;
      int inb = 0, val = 0;
      for (int jb = j0; jb < j1; jb += QB) {
        int pcs[QB], cl[QB], fl[QB];
        bool oks[QB];
#pragma unroll
        for (int u = 0; u < QB; ++u) {  // the picks, and their pattern loads issued together
          const int j = jb + u;
          pcs[u] = 0;
          if (j < j1) {
            const int pick =
                (int)floor(hs_u53(hs_bits53(I.seed, step, 3u | ((uint32_t)attempt << 8), (uint64_t)j), 0) * cnt);
            pcs[u] = SEL.pat_cells[p0 + pick];
          }
        }
#pragma unroll
        for (int u = 0; u < QB; ++u) {  // the sample cells, and their flag loads issued together
          const int j = jb + u;
          oks[u] = false;
          cl[u] = -1;
          fl[u] = 0;
          if (j < j1) {
            bool ok;
            cl[u] = grid_cell(I, SEL, ox, oy, j, pcs[u], ok);
            oks[u] = ok;
            if (ok) fl[u] = I.flags[cl[u]];
          }
        }
#pragma unroll
        for (int u = 0; u < QB; ++u) {
          const int j = jb + u;
          if (j >= j1) continue;
          const bool v = oks[u] && (fl[u] & HS_FLAG_OCC);
          const int c = v ? cl[u] : -1;
          ws.cand[j] = c;
          candr[u] = c;  // (one batch when regs)
          inb += oks[u];
          val += v;
        }
      }
      // one block reduction of both counts (in-range samples in the high word, valid ones in the low)
      const long long both = block_sum_i64(((long long)inb << 32) + val, s_l);
      const int tot_inb = (int)(both >> 32), tot_val = (int)(both & 0xffffffffll);
      if (tot_inb > 0) {
        if (tot_val > 0) {
          nvalid = tot_val;
          break;
        }
        continue;  // nothing occupied: select() again (core.py:487)
      }
    }
    // _select_single: np.random.randint(self.n) over the initial nonzero list (core.py:493)
    const int i = (int)floor(hs_u53(b2, 0) * n);
    const int cell = I.occ_list[min(i, I.n_occ - 1)];
    if (I.flags[cell] & HS_FLAG_OCC) {
      if (tid == 0) s_single = cell;
      single = true;
      nvalid = 1;
      break;
    }
  }
  __syncthreads();
  if (attempt >= MAX_ATTEMPTS || bad_list) {
    if (tid == 0) {
      S.err = bad_list ? 4 : 2;  // 4: a listed sample is outside the grid or not a magnet
      S.done = 1;
      st[0] = S;
      *ws.acc_n = 0;
    }
    return -1;
  }
  // acceptance (core.py:662-671); with a single sample thread 0 plays candidate 0
  int my_acc = 0;
  double tsum = 0.0;
  const int a0 = single ? (tid == 0 ? 0 : 1) : j0, a1 = single ? 1 : j1;
  if (pre) {
    if (regs) {
#pragma unroll
      for (int u = 0; u < QB; ++u) my_acc += j0 + u < j1 ? accr[u] : 0;
    } else {
      for (int j = a0; j < a1; ++j) my_acc += ws.accf[j];
    }
  }
  for (int jb = a0; !pre && jb < a1; jb += QB) {
    int cs[QB], mv[QB];
    float ev[QB], ecv[QB], btv[QB], muv[QB];
#pragma unroll
    for (int u = 0; u < QB; ++u) {  // the samples' state, loaded together
      const int j = jb + u;
      cs[u] = j < a1 ? (single ? s_single : (regs ? candr[u] : ws.cand[j])) : -1;
      mv[u] = 0;
      ev[u] = ecv[u] = btv[u] = muv[u] = 0.f;
      if (cs[u] >= 0) {
        // FP64 fields: the value in E64 (ev - ecv below then stands for it exactly)
        ev[u] = I.E64 ? 0.f : I.E[cs[u]];
        ecv[u] = I.E64 ? 0.f : I.Ec[cs[u]];
        btv[u] = I.beta[cs[u]];
        mv[u] = (int)I.m[cs[u]];
        muv[u] = I.mu_src[cs[u]];  // source weight (1 for DiMonopolarEnergy)
      }
    }
#pragma unroll
    for (int u = 0; u < QB; ++u) {
      const int j = jb + u;
      if (j >= a1) continue;
      const int c = cs[u];
      bool acc = false;
      float mmu = 0.f;
      if (c >= 0) {
        mmu = (float)(-mv[u]) * muv[u];  // (m mu) after a flip
        if (SEL.metropolis_time) {
          const double ut = hs_u53(hs_bits53(I.seed, step, 5u, (uint64_t)c), 1);
          // FP64 beta when T is uniform: barrier*beta reaches a few hundred here as in Neel times
          const double b64 = I.uniform_beta ? I.beta64 : (double)btv[u];
          tsum += -log(ut) / R.attempt_freq / n * exp(barrier_min_f64(I, c) * b64);
        }
        acc = I.E64 ? metro_accepts_e(I, step, c, I.E64[c], btv[u]) : metro_accepts(I, step, c, ev[u], ecv[u], btv[u]);
      }
      if (single) {
        if (tid == 0) {
          s_single = acc ? c : -1;
          s_single_mmu = mmu;
        }
      } else {
        ws.accf[j] = acc ? 1 : 0;
        ws.cmmu[j] = mmu;
        accr[u] = acc ? 1 : 0;
        mmur[u] = mmu;
      }
      my_acc += acc;
    }
  }
  // this thread's accepted cell of sample slot u of batch jb (-1 if none), and its (m mu)
  auto accepted = [&](int jb, int u, float &mmu) {
    const int j = jb + u;
    if (single) {
      mmu = s_single_mmu;
      return s_single;
    }
    if (regs) {
      mmu = mmur[u];
      return accr[u] ? candr[u] : -1;
    }
    mmu = ws.cmmu[j];
    return ws.accf[j] ? ws.cand[j] : -1;
  };
  if (SEL.metropolis_time) tsum = block_sum_double(tsum, s_d);
  int total = 0;
  int pos = block_exclusive_scan(my_acc, s_i, total);
  __syncthreads();
  // the accepted switches in selection order; the flips themselves (m *= -1, E *= -1,
  // energies.py:124-126, 332-334) are applied by the update pass, by the tile holding the cell
  for (int jb = a0; jb < a1; jb += QB) {
#pragma unroll
    for (int u = 0; u < QB; ++u) {
      if (jb + u >= a1) continue;
      float mmu;
      const int c = accepted(jb, u, mmu);
      if (c < 0) continue;
      ws.acc_cell[pos] = c;
      ws.acc_mmu[pos] = mmu;
      atomicAdd(&s_site[hs_site(I, c / nx, c % nx)], 1);
      if (R.log) log_cells[(size_t)S.steps_done * (size_t)(nsc > 0 ? nsc : 1) + pos] = c;
      ++pos;
    }
  }
  __syncthreads();
  if (tid < I.ux * I.uy) {
    // update_single for one switch (core.py:250-251) uses the full kernel; update_multiple
    // truncates per sublattice above the cutoff when REDUCED_KERNEL_SIZE != 0 (energies.py:343-349)
    ws.trunc[tid] = (total > 1 && SEL.reduced_k != 0 && s_site[tid] > SEL.cutoff) ? 1 : 0;
  }
  // the presampled grid scheme without PBC leaves the tiles to k_metro_tiles (many workgroups)
  const bool defer_tiles = PRE && !LIST_TILES && !single && !I.pbc;
  if (WL.wl && total > 0 && !defer_tiles) {
    // the tiles this step changes: the windows metro_update adds (dipolar: stored or truncated
    // window, wrapped when truncated with PBC; exchange: the NN stencil, wrapped with PBC) and the
    // tile of every switch (its flip), appended once each to the worklist
    const int tiles_x = J.tiles_x, ntile = J.tiles_x * J.tiles_y;
    const bool has_dip = I.dip_pf != 0.0, has_ex = I.J != 0.0 && I.nn_w > 0;
    const int tkx = min(min(SEL.reduced_k, nx - 1), I.kwx), tky = min(min(SEL.reduced_k, ny - 1), I.kwy);
    const int nword = (ntile + 31) / 32;
    if (tid == 0) s_all = nword > MARK_WORDS ? 1 : 0;
    for (int i = tid; i < min(nword, MARK_WORDS); i += NT) s_bits[i] = 0u;
    __syncthreads();
    if (s_all == 0) {
      for (int jb = a0; jb < a1; jb += QB) {
#pragma unroll
        for (int u = 0; u < QB; ++u) {
          if (jb + u >= a1) continue;
          float mmu;
          const int c = accepted(jb, u, mmu);
          if (c < 0) continue;
          const int sy = c / nx, sx = c % nx, ss = hs_site(I, sy, sx);
          const int t0 = (sy / MU_TY) * tiles_x + sx / MU_TX;
          atomicOr(&s_bits[t0 >> 5], 1u << (t0 & 31));
          bool ok = true;
          if (has_dip && I.sub_of[ss] >= 0) {
            const bool tr = total > 1 && SEL.reduced_k != 0 && s_site[ss] > SEL.cutoff;
            ok = mark_rect(s_bits, tiles_x, nx, ny, sx, tr ? tkx : I.kwx, sy, tr ? tky : I.kwy, tr && I.pbc);
          }
          if (ok && has_ex) ok = mark_rect(s_bits, tiles_x, nx, ny, sx, I.nn_w / 2, sy, I.nn_h / 2, I.pbc != 0);
          if (!ok) s_all = 1;
        }
      }
      __syncthreads();
    }
    // list the marked tiles (every tile when a window was too wide or the bitmap too small)
    const bool all = s_all != 0;
    auto word_bits = [&](int w) {  // bit b: tile 32*w + b is listed
      const uint32_t v = all ? 0xffffffffu : s_bits[w];
      const int valid = ntile - 32 * w;
      return valid >= 32 ? v : v & ((1u << valid) - 1u);
    };
    int mine = 0;
    for (int w = tid; w < nword; w += NT) mine += __popc(word_bits(w));
    int nlist = 0;
    int off = block_exclusive_scan(mine, s_i, nlist);
    if (tid == 0) s_base = atomicAdd(&WL.count[par], nlist);
    __syncthreads();
    off += s_base;
    for (int w = tid; w < nword; w += NT) {
      uint32_t bits = word_bits(w);
      while (bits) {
        const int b = __ffs((int)bits) - 1;
        bits &= bits - 1u;
        WL.wl[off++] = make_int2(job, 32 * w + b);
      }
    }
  }
  if (tid == 0) {
    ws.info[0] = ox;
    ws.info[1] = oy;
    ws.info[2] = single ? 1 : 0;
    ws.info[4] = 0;
    *ws.acc_n = total;
    if (R.log) log_count[S.steps_done] = total;
    S.attempts += nvalid;
    S.switches += total;
    S.step += 1;
    S.steps_done += 1;
    if (SEL.metropolis_time) S.t += tsum;
    st[0] = S;
  }
  return total;
}

// Displacement candidates d = i - s (+k*n) with |d| <= w: wrap enumerates images, fill only k = 0.
HS_INLINE int disp_candidates(int i, int s, int n, int w, bool wrap, int out[3]) {
  int cnt = 0;
  const int d = i - s;
  if (!wrap) {
    if (d >= -w && d <= w) out[cnt++] = d;
    return cnt;
  }
  for (int k = -1; k <= 1; ++k) {
    const int dd = d + k * n;
    if (dd >= -w && dd <= w) out[cnt++] = dd;
  }
  return cnt;
}

// Can a source at s with half-width w (optionally wrapping over n) reach [lo, hi]?
HS_INLINE bool reaches(int s, int w, int n, bool wrap, int lo, int hi) {
  if (w < 0) return false;
  if (!wrap) return s + w >= lo && s - w <= hi;
  for (int k = -1; k <= 1; ++k)
    if (s + k * n + w >= lo && s + k * n - w <= hi) return true;
  return false;
}

HS_INLINE int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Update of one TX x TY output tile by a workgroup of TX*TY threads (one cell per thread).
// PRELOAD: load the cell's own state (m, mu, fields) at the start, with the step's counts, for
// tiles known to be reached (the worklist); otherwise only once a source reaches the tile (the
// persistent kernel visits every tile and keeps its registers for the selection).
// F64 (an instance with FP64 fields, hs_instance_t.E64): the increments are FP64 sums of
// K64 * (m mu64) and exchange terms, added to E64 without compensation.
template <int TX, int TY, bool PRELOAD, bool F64 = false>
__device__ __forceinline__ void metro_update(const MetroJob &J, int tile_x, int tile_y) {
  constexpr int MU_TX = TX, MU_TY = TY, MU_THREADS = TX * TY;
  using FT = typename std::conditional<F64, double, float>::type;
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const MetroWS &ws = J.ws;
  __shared__ int s_cell[MU_THREADS];
  __shared__ float s_mmu[MU_THREADS];
  __shared__ int s_site_of[MU_THREADS];
  __shared__ int s_cnt[MU_THREADS / 64 + 1];
  __shared__ int s_sub[HS_MAX_UC], s_trunc[HS_MAX_UC];
  __shared__ float s_oxu[HS_MAX_UC], s_oyu[HS_MAX_UC], s_nn[HS_MAX_NN];
  __shared__ double s_mmu64[F64 ? MU_THREADS : 1];  // F64: the sources' (m mu64) after their flips
  // decoded per staged source once, for the whole tile (grids without PBC): (sy << 16) | sx, the
  // kernel offset n*KS + (kwy - sy)*KW + (kwx - sx) (target (y, x) reads entry offset + y*KW + x),
  // and the window half-widths wy | wx << 16 (-1: no dipolar term)
  __shared__ int s_yx[MU_THREADS], s_win[MU_THREADS];
  __shared__ long long s_koff[MU_THREADS];
  const int tid = threadIdx.x;
  const int x0 = tile_x * MU_TX, y0 = tile_y * MU_TY;
  const int x = x0 + (tid % MU_TX), y = y0 + (tid / MU_TX);
  const int nx = I.nx, ny = I.ny;
  const bool inside = x < nx && y < ny;
  const int c = inside ? y * nx + x : 0;
  // everything that does not depend on the step's sources is loaded up front, together: the
  // step's counts, the per-site tables and the cell's own state
  const int a = *ws.acc_n;
  int info_ox = 0, info_oy = 0, info_single = 0, info_samples = 0;
  if (PRELOAD) {
    info_ox = ws.info[0];
    info_oy = ws.info[1];
    info_single = ws.info[2];
    info_samples = ws.info[4];
  }
  if (PRELOAD && tid < I.ux * I.uy) {
    s_sub[tid] = I.sub_of[tid];
    s_trunc[tid] = ws.trunc[tid];
    s_oxu[tid] = (float)I.ox_uc[tid];
    s_oyu[tid] = (float)I.oy_uc[tid];
  }
  if (PRELOAD && tid < I.nn_w * I.nn_h) s_nn[tid] = I.nn_mask[tid];
  // per-site tables: from LDS when preloaded (short dependency chains), else straight from the
  // descriptor (fewer live registers in the persistent kernel)
  auto sub_of = [&](int ss) { return PRELOAD ? s_sub[ss] : I.sub_of[ss]; };
  auto trunc_of = [&](int ss) { return PRELOAD ? s_trunc[ss] : ws.trunc[ss]; };
  auto ox_of = [&](int ss) { return F64 ? (FT)I.ox_uc[ss] : (FT)(PRELOAD ? s_oxu[ss] : (float)I.ox_uc[ss]); };
  auto oy_of = [&](int ss) { return F64 ? (FT)I.oy_uc[ss] : (FT)(PRELOAD ? s_oyu[ss] : (float)I.oy_uc[ss]); };
  int m0 = 0;
  float mu0 = 0.f, E0 = 0.f, Ec0 = 0.f, Ep0 = 0.f, Epc0 = 0.f;
  if (PRELOAD && inside) {
    m0 = (int)I.m[c];
    mu0 = I.mu[c];
    E0 = I.E[c];
    Ec0 = I.Ec[c];
    if (I.use_perp) {
      Ep0 = I.Ep[c];
      Epc0 = I.Epc[c];
    }
  }
  if (a == 0) return;
  if (PRELOAD) __syncthreads();
  int mi = 2;  // the cell's spin, taken once some source reaches the tile (2 = not yet)
  const bool has_dip = I.dip_pf != 0.0;
  const int hy = I.nn_h / 2, hx = I.nn_w / 2;
  const bool has_ex = I.J != 0.0 && I.nn_w > 0;
  const int KW = 2 * I.kwx + 1;
  const size_t KS = (size_t)KW * (2 * I.kwy + 1);
  const int tkx = min(min(SEL.reduced_k, nx - 1), I.kwx), tky = min(min(SEL.reduced_k, ny - 1), I.kwy);
  // the sources' coordinates, kernel offsets and windows are decoded once per staged source (no
  // wrapped images without PBC; coordinates packed in 16 bits)
  const bool decoded = !I.pbc && nx <= 32767 && ny <= 32767;
  const long long cellk = (long long)y * KW + x;
  const int yx_self = (y << 16) | x;
  FT oxi = 0, oyi = 0;
  FT acc = 0, ex = 0;
  float accp = 0.f, exp_ = 0.f;
  bool touched = false, flipped = false;

  // the contribution of one switched source (cell scell, (m mu) after its flip, unit-cell site ss,
  // staged at LDS slot r)
  auto add_source = [&](int scell, float mm, int ss, int r) {
    const int sy = scell / nx, sx = scell % nx;
    if (has_dip) {
      const int n = sub_of(ss);
      if (n >= 0) {
        const bool tr = trunc_of(ss) != 0;
        const bool wrap = tr && I.pbc;
        const int wx = tr ? tkx : I.kwx, wy = tr ? tky : I.kwy;
        int dys[3], dxs[3];
        const int ndy = disp_candidates(y, sy, ny, wy, wrap, dys);
        const int ndx = disp_candidates(x, sx, nx, wx, wrap, dxs);
        const float *Kn = I.K32 + (size_t)n * KS;
        const float *Kp = I.use_perp ? I.Kpo32 + (size_t)n * KS : nullptr;
        for (int a1 = 0; a1 < ndy; ++a1)
          for (int a2 = 0; a2 < ndx; ++a2) {
            const size_t o = (size_t)(I.kwy + dys[a1]) * KW + (I.kwx + dxs[a2]);
            if (F64) {
              acc += s_mmu64[r] * I.K64[(size_t)n * KS + o];
            } else {
              acc += mm * Kn[o];
              if (Kp) accp += mm * Kp[o];
            }
            touched = true;
          }
      }
    }
    if (has_ex) {
      int dy = y - sy, dx = x - sx;
      if (I.pbc) {
        dy = hs_pmod(dy, ny);
        dy = dy > ny / 2 ? dy - ny : dy;
        dx = hs_pmod(dx, nx);
        dx = dx > nx / 2 ? dx - nx : dx;
      }
      if (dy >= -hy && dy <= hy && dx >= -hx && dx <= hx) {
        const int q = (dy + hy) * I.nn_w + (dx + hx);
        const float wq = PRELOAD ? s_nn[q] : I.nn_mask[q];
        if (wq != 0.f) {
          const FT msn = mm > 0.f ? 1 : -1;  // m of the source after its flip
          if (I.in_plane) {
            const FT oxs = ox_of(ss), oys = oy_of(ss);
            ex += msn * (FT)wq * (oxi * oxs + oyi * oys);
            if (!F64) exp_ += (float)(msn * (FT)wq * (oys * oxi - oxs * oyi));
          } else {
            ex += msn * (FT)wq;
          }
          touched = true;
        }
      }
    }
  };

  // Direct lookup: the sample of chosen supercell j = rs*ncols + cs lies in rows
  // [2 rs Ry + oy, +Ry) and columns [2 cs Rx + ox, +Rx) (core.py:540-552), so only supercells
  // within the largest window of the tile can reach it.
  bool lookup = false;
  int rs0 = 0, rs1 = -1, cs0 = 0, cs1 = -1;
  if (!PRELOAD) {
    info_single = ws.info[2];
    info_samples = ws.info[4];
  }
  if (SEL.scheme == 0 && !I.pbc && info_single == 0) {
    const int ox = PRELOAD ? info_ox : ws.info[0], oy = PRELOAD ? info_oy : ws.info[1];
    const int Wx = max(has_dip ? I.kwx : 0, has_ex ? hx : 0), Wy = max(has_dip ? I.kwy : 0, has_ex ? hy : 0);
    const int nrows = SEL.nsc / SEL.ncols;
    cs0 = max(0, -floor_div(-(x0 - Wx - ox - SEL.Rx + 1), 2 * SEL.Rx));
    cs1 = min(SEL.ncols - 1, floor_div(x0 + MU_TX - 1 + Wx - ox, 2 * SEL.Rx));
    rs0 = max(0, -floor_div(-(y0 - Wy - oy - SEL.Ry + 1), 2 * SEL.Ry));
    rs1 = min(nrows - 1, floor_div(y0 + MU_TY - 1 + Wy - oy, 2 * SEL.Ry));
    // the sample array only, when the step wrote no ordered list (same sources, same order)
    lookup = info_samples != 0 || (int64_t)max(0, cs1 - cs0 + 1) * max(0, rs1 - rs0 + 1) <= (int64_t)a;
  }
  // stage the sources that reach the tile in LDS (in selection order), chunk by chunk, and let
  // every cell of the tile add their contributions
  const int ncs = max(0, cs1 - cs0 + 1);
  const int nsrc = lookup ? ncs * max(0, rs1 - rs0 + 1) : a;
  for (int base = 0; base < nsrc; base += MU_THREADS) {
    const int sidx = base + tid;
    bool rel = false;
    int sc = 0, ssite = 0, sy = 0, sx = 0;
    float smmu = 0.f;
    if (sidx < nsrc) {
      bool have;
      if (lookup) {
        const int j = (rs0 + sidx / ncs) * SEL.ncols + cs0 + sidx % ncs;
        const int af = ws.accf[j], cj = ws.cand[j];  // loaded together (cand/cmmu are set for every j)
        const float mj = ws.cmmu[j];
        have = af != 0;
        if (have) {
          sc = cj;
          smmu = mj;  // (m mu) after the flip, as in the selection order list
        }
      } else {
        have = true;
        sc = ws.acc_cell[sidx];
        smmu = ws.acc_mmu[sidx];
      }
      if (have) {
        sy = sc / nx;
        sx = sc % nx;
        ssite = hs_site(I, sy, sx);
        if (has_dip && sub_of(ssite) >= 0) {
          const bool tr = trunc_of(ssite) != 0;
          const int wx = tr ? tkx : I.kwx, wy = tr ? tky : I.kwy;
          const bool wrap = tr && I.pbc;
          rel = reaches(sx, wx, nx, wrap, x0, x0 + MU_TX - 1) && reaches(sy, wy, ny, wrap, y0, y0 + MU_TY - 1);
        }
        if (!rel && has_ex)
          rel = reaches(sx, hx, nx, I.pbc, x0, x0 + MU_TX - 1) && reaches(sy, hy, ny, I.pbc, y0, y0 + MU_TY - 1);
        // a switch always reaches its own tile, which applies the flip
        rel = rel || (sx >= x0 && sx < x0 + MU_TX && sy >= y0 && sy < y0 + MU_TY);
      }
    }
    const unsigned long long ball = __ballot(rel);
    const int lane = tid & 63, wid = tid >> 6;
    __syncthreads();
    if (lane == 0) s_cnt[wid] = __popcll(ball);
    __syncthreads();
    int off = 0, nrel = 0;
    for (int w = 0; w < MU_THREADS / 64; ++w) {
      if (w < wid) off += s_cnt[w];
      nrel += s_cnt[w];
    }
    if (rel) {
      const int p = off + __popcll(ball & ((1ull << lane) - 1ull));
      s_cell[p] = sc;
      s_mmu[p] = smmu;
      s_site_of[p] = ssite;
      if (F64) s_mmu64[p] = (smmu > 0.f ? 1.0 : -1.0) * I.mu64[sc];
      if (decoded) {
        int win = -1;
        long long ko = 0;
        const int n = has_dip ? sub_of(ssite) : -1;
        if (n >= 0) {
          const bool tr = trunc_of(ssite) != 0;
          win = (tr ? tky : I.kwy) | ((tr ? tkx : I.kwx) << 16);
          ko = (long long)n * (long long)KS + (long long)(I.kwy - sy) * KW + (I.kwx - sx);
        }
        s_yx[p] = (sy << 16) | sx;
        s_win[p] = win;
        s_koff[p] = ko;
      }
    }
    __syncthreads();
    if (nrel == 0) continue;
    if (mi == 2) {  // first chunk with sources: the cell's own spin and site
      mi = PRELOAD ? m0 : (inside ? (int)I.m[c] : 0);
      if (mi != 0) {
        const int site_i = hs_site(I, y, x);
        oxi = ox_of(site_i);
        oyi = oy_of(site_i);
      }
    }
    if (!inside || mi == 0) continue;
    if (!decoded) {
      for (int r = 0; r < nrel; ++r) {
        flipped = flipped || s_cell[r] == c;
        add_source(s_cell[r], s_mmu[r], s_site_of[r], r);
      }
      continue;
    }
    if constexpr (!PRELOAD) {
      // the persistent kernel (one workgroup, every tile, mostly windows that cover the tile):
      // the kernel entries of PERSIST_SRC_BATCH sources are loaded together, then added in list
      // order (a source outside the window, or past the list's end, loads entry 0 and adds nothing)
      constexpr int SB = PERSIST_SRC_BATCH;
      for (int r0 = 0; r0 < nrel; r0 += SB) {
        int syxv[SB];
        bool inv[SB];
        float mmv[SB], kpv[SB];
        FT kvv[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int r = r0 + u;
          const bool live = r < nrel;
          const int rr = live ? r : 0;
          const int syx = s_yx[rr], win = s_win[rr];
          const int sy = syx >> 16, sx = syx & 0xffff;
          const int wy = win & 0xffff, wx = win >> 16;
          const bool in = live && has_dip && win >= 0 && (unsigned)(y - sy + wy) <= (unsigned)(2 * wy) &&
                          (unsigned)(x - sx + wx) <= (unsigned)(2 * wx);
          const size_t o = in ? (size_t)(s_koff[rr] + cellk) : 0;
          syxv[u] = syx;
          inv[u] = in;
          mmv[u] = s_mmu[rr];
          kvv[u] = 0;
          kpv[u] = 0.f;
          if (has_dip) {  // the kernel arrays exist only with a pair energy
            if (F64) {
              kvv[u] = (FT)I.K64[o];
            } else {
              kvv[u] = (FT)I.K32[o];
              if (I.use_perp) kpv[u] = I.Kpo32[o];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int r = r0 + u;
          if (r >= nrel) break;
          const int syx = syxv[u];
          const float mm = mmv[u];
          flipped = flipped || syx == yx_self;
          if (has_dip) {
            if (inv[u]) {
              if (F64) {
                acc += s_mmu64[r] * kvv[u];
              } else {
                acc += mm * kvv[u];
                if (I.use_perp) accp += mm * kpv[u];
              }
            }
            touched = touched || inv[u];
          }
          if (has_ex) {
            const int dy = y - (syx >> 16), dx = x - (syx & 0xffff);
            if (dy >= -hy && dy <= hy && dx >= -hx && dx <= hx) {
              const int q = (dy + hy) * I.nn_w + (dx + hx);
              const float wq = I.nn_mask[q];
              if (wq != 0.f) {
                const FT msn = mm > 0.f ? 1 : -1;  // m of the source after its flip
                if (I.in_plane) {
                  const int ss = s_site_of[r];
                  const FT oxs = ox_of(ss), oys = oy_of(ss);
                  ex += msn * (FT)wq * (oxi * oxs + oyi * oys);
                  if (!F64) exp_ += (float)(msn * (FT)wq * (oys * oxi - oxs * oyi));
                } else {
                  ex += msn * (FT)wq;
                }
                touched = true;
              }
            }
          }
        }
      }
      continue;
    }
    // the same terms as add_source, in the same order and arithmetic, from the decoded sources:
    // a source outside the target's window loads entry 0 and adds nothing (no branch per source)
    for (int r = 0; r < nrel; ++r) {
      const int syx = s_yx[r], win = s_win[r];
      const float mm = s_mmu[r];
      flipped = flipped || syx == yx_self;
      const int sy = syx >> 16, sx = syx & 0xffff;
      if (has_dip) {
        const int wy = win & 0xffff, wx = win >> 16;
        const bool in = win >= 0 && (unsigned)(y - sy + wy) <= (unsigned)(2 * wy) &&
                        (unsigned)(x - sx + wx) <= (unsigned)(2 * wx);
        const size_t o = in ? (size_t)(s_koff[r] + cellk) : 0;
        if (F64) {
          const double kv = I.K64[o];
          if (in) acc += s_mmu64[r] * kv;
        } else {
          const float kv = I.K32[o];
          if (in) acc += mm * kv;
          if (I.use_perp) {
            const float kp = I.Kpo32[o];
            if (in) accp += mm * kp;
          }
        }
        touched = touched || in;
      }
      if (has_ex) {
        const int dy = y - sy, dx = x - sx;
        if (dy >= -hy && dy <= hy && dx >= -hx && dx <= hx) {
          const int q = (dy + hy) * I.nn_w + (dx + hx);
          const float wq = PRELOAD ? s_nn[q] : I.nn_mask[q];
          if (wq != 0.f) {
            const FT msn = mm > 0.f ? 1 : -1;  // m of the source after its flip
            if (I.in_plane) {
              const int ss = s_site_of[r];
              const FT oxs = ox_of(ss), oys = oy_of(ss);
              ex += msn * (FT)wq * (oxi * oxs + oyi * oys);
              if (!F64) exp_ += (float)(msn * (FT)wq * (oys * oxi - oxs * oyi));
            } else {
              ex += msn * (FT)wq;
            }
            touched = true;
          }
        }
      }
    }
  }
  if (!inside || mi == 0 || mi == 2 || !(touched || flipped)) return;
  if (flipped) {  // the flip first (E[w] *= -1 before the sum, energies.py:332-334)
    mi = -mi;
    I.m[c] = (int8_t)mi;
  }
  if (F64) {  // E[w] *= -1 for the flip, then the FP64 increment (no compensation needed)
    double e = (flipped ? -1.0 : 1.0) * I.E64[c];
    if (touched) e += 2.0 * I.dip_pf * ((double)mi * I.mu64[c]) * (double)acc + (-2.0 * I.J) * ((double)mi * (double)ex);
    I.E64[c] = e;
    return;
  }
  const float sgn = flipped ? -1.f : 1.f;
  if (!PRELOAD) {
    mu0 = I.mu[c];
    E0 = I.E[c];
    Ec0 = I.Ec[c];
  }
  const float mmui = (float)mi * mu0;
  // E += 2*prefactor*(m mu)_i*conv (energies.py:362-363); exchange: -J m_i * 2 m_s * NN
  float e = sgn * E0, ec = sgn * Ec0;
  if (touched) {
    const float de = (float)(2.0 * I.dip_pf) * (mmui * (float)acc) + (float)(-2.0 * I.J) * ((float)mi * (float)ex);
    hs_kahan_add(e, ec, de);
  }
  I.E[c] = e;
  I.Ec[c] = ec;
  if (I.use_perp) {
    if (!PRELOAD) {
      Ep0 = I.Ep[c];
      Epc0 = I.Epc[c];
    }
    float p = sgn * Ep0, pc = sgn * Epc0;
    if (touched) {
      const float dep = (float)(2.0 * I.dip_pf) * (mmui * accp) + (float)(-2.0 * I.J) * ((float)mi * exp_);
      hs_kahan_add(p, pc, dep);
    }
    I.Ep[c] = p;
    I.Epc[c] = pc;
  }
}

__global__ void __launch_bounds__(MS_THREADS) k_metro_select_one(MetroJob J, WorkList WL, int par) {
  metro_select<SELECT_BATCH>(J, WL, par, 0);
}

// Grid selection of large instances, first half: attempt 0's samples and their acceptance, one
// sample per thread over many workgroups (the single-workgroup selection would walk them 3+ per
// thread through chains of dependent loads). Writes cand / accf / cmmu of every sample and the
// workgroup's counts; k_metro_select_pre lists them (or redoes the step when attempt 0 has no
// valid sample). Same draws and arithmetic as metro_select.
__device__ __forceinline__ void metro_presample(const MetroJob &J, int blk, WorkList WL = WorkList{nullptr, nullptr},
                                                int par = 0) {
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const hs_run_t &R = J.R;
  const MetroWS &ws = J.ws;
  __shared__ long long s_l[PRESAMPLE_THREADS / 64];
  __shared__ int s_site[HS_MAX_UC];
  __shared__ bool s_last;
  const hs_state_t S = J.st[0];
  const double n = (double)I.n_occ;
  // the step is completed here rather than by k_metro_select_pre: without PBC (that selection
  // lists the tiles itself) and without a step log (which wants the ordered list)
  const bool fuse = !I.pbc && !R.log;
  // tile bitmap mode (single instance, fused): the accepted switches mark the tiles they reach.
  // The launch then also owns the selection's retries: an attempt without a valid sample leaves
  // the step to the next launch, which draws the next attempt (ws.info[5])
  const bool mark = fuse && WL.bits != nullptr;
  if (fuse && threadIdx.x < HS_MAX_UC) s_site[threadIdx.x] = 0;
  // the previous step's mark goes first (before this workgroup's ticket, so before the last
  // workgroup sets it again)
  if (fuse && blk == 0 && threadIdx.x == 0) __hip_atomic_store(&ws.info[3], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const bool go = !S.done && S.steps_done < R.max_steps &&
                  (hs_step_free(R.first_free, S.steps_done) ||
                   (hs_lower_than(S.t - S.t0, R.t_max) &&
                    hs_lower_than((double)S.attempts / n - (double)S.attempts0 / n, R.mc_max)));
  if (!go) {  // the run is over: as the selection records it (no switch, done)
    if (mark && blk == 0 && threadIdx.x == 0) {
      *ws.acc_n = 0;
      if (!S.done) {
        hs_state_t o = S;
        o.done = 1;
        J.st[0] = o;
      }
    }
    return;
  }
  const uint64_t step = (uint64_t)S.step;
  const int attempt = mark ? ws.info[5] : 0;  // the selection's attempt index (core.py:487)
  int ox, oy;
  grid_offsets(I, SEL, hs_bits53(I.seed, step, 2u, (uint64_t)(4 * attempt + 0)),
               hs_bits53(I.seed, step, 2u, (uint64_t)(4 * attempt + 1)), ox, oy);
  const int pidx = hs_pmod(oy, I.uy) * I.ux + hs_pmod(ox, I.ux);
  const int p0 = SEL.pat_off[pidx], cnt = SEL.pat_off[pidx + 1] - p0;
  const int j = blk * PRESAMPLE_THREADS + threadIdx.x;
  int inb = 0, val = 0, nacc = 0;
  int c = -1;
  bool acc = false;
  float mmu = 0.f;
  if (fuse) __syncthreads();  // s_site cleared
  if (j < SEL.nsc) {
    const int pick =
        (int)floor(hs_u53(hs_bits53(I.seed, step, 3u | ((uint32_t)attempt << 8), (uint64_t)j), 0) * cnt);
    bool ok;
    const int cell = grid_cell(I, SEL, ox, oy, j, SEL.pat_cells[p0 + pick], ok);
    const bool v = ok && (I.flags[cell] & HS_FLAG_OCC);
    c = v ? cell : -1;
    inb = ok;
    val = v;
    if (c >= 0) {
      const float E = I.E[c], Ec = I.Ec[c], beta = I.beta[c];
      const int mc = (int)I.m[c];
      const float muv = I.mu_src[c];
      mmu = (float)(-mc) * muv;
      acc = metro_accepts(I, step, c, E, Ec, beta);
    }
    nacc = acc ? 1 : 0;
    if (acc && fuse) atomicAdd(&s_site[hs_site(I, c / I.nx, c % I.nx)], 1);
    if (acc && mark) {
      // the tiles within the largest window of the switch (dipolar or exchange; the switch's own
      // tile included): a superset of those the update's reach rules select, whose extra tiles
      // find no source in reach and write nothing. Read by the update after this launch.
      const bool has_dip = I.dip_pf != 0.0, has_ex = I.J != 0.0 && I.nn_w > 0;
      const int Wx = max(has_dip ? I.kwx : 0, has_ex ? I.nn_w / 2 : 0), Wy = max(has_dip ? I.kwy : 0, has_ex ? I.nn_h / 2 : 0);
      const int sy = c / I.nx, sx = c % I.nx;
      const int tx0 = max(0, sx - Wx) / MU_TX, tx1 = min(I.nx - 1, sx + Wx) / MU_TX;
      const int ty0 = max(0, sy - Wy) / MU_TY, ty1 = min(I.ny - 1, sy + Wy) / MU_TY;
      uint32_t *bits = WL.bits + (size_t)par * WL.nwords;
      for (int ty = ty0; ty <= ty1; ++ty) {
        const int t0 = ty * J.tiles_x + tx0, t1 = ty * J.tiles_x + tx1;
        for (int w = t0 >> 5; w <= (t1 >> 5); ++w) {
          const int b0 = max(t0, w << 5) & 31, b1 = min(t1, (w << 5) + 31) & 31;
          const uint32_t hi = b1 == 31 ? 0xffffffffu : ((1u << (b1 + 1)) - 1u);
          atomicOr(&bits[w], hi & ~((1u << b0) - 1u));
        }
      }
    }
  }
  // in-grid, valid and accepted samples of the workgroup, 21 bits each (at most 256)
  const long long sums = block_sum_i64(((long long)inb << 42) + ((long long)val << 21) + nacc, s_l);
  const int b_inb = (int)(sums >> 42), b_val = (int)((sums >> 21) & 0x1fffff), b_acc = (int)(sums & 0x1fffff);
  // the samples and the workgroup's counts, read by later launches only (with the listing fused,
  // they are stored after the ticket, so that the wait before it covers the counts and marks alone)
  auto store_samples = [&]() {
    if (j < SEL.nsc) {
      ws.cand[j] = c;
      ws.accf[j] = acc ? 1 : 0;
      ws.cmmu[j] = mmu;
    }
    if (threadIdx.x == 0) ws.part[blk] = ((long long)b_inb << 32) + b_val;
  };
  if (!fuse) {
    store_samples();
    return;
  }
  // Fused listing: the workgroups add their counts to the job's counters; the last one to finish
  // (ticket) completes the step as the selection workgroup would (truncation flags, offsets,
  // counters and state), and k_metro_select_pre then finds the step done. The update takes its
  // sources from the sample array (supercell order is selection order), so no ordered list is
  // needed. A step whose presampled attempt has no valid sample is left to k_metro_select_pre.
  const int tid = threadIdx.x, nsites = I.ux * I.uy;
  if (tid < nsites && s_site[tid]) atomicAdd(&ws.fz[8 + tid], s_site[tid]);
  if (tid == 0) {
    atomicAdd(&ws.fz[1], b_inb);
    atomicAdd(&ws.fz[2], b_val);
    atomicAdd(&ws.fz[3], b_acc);
  }
  // The counts are device-scope atomics and the mark's clear a device-scope store: each thread
  // waits until they are performed (s_waitcnt, behind a workgroup fence for the compiler) before
  // the ticket, and the last workgroup reads them with device-scope loads. No device-scope fence:
  // on this GPU that writes back and invalidates the whole L2 of the XCD, per workgroup, and the
  // following launches would find their data evicted.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) s_last = atomicAdd(&ws.fz[0], 1) == (SEL.nsc + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS - 1;
  store_samples();
  __syncthreads();
  if (!s_last) return;
  const int tot_inb = __hip_atomic_load(&ws.fz[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int tot_val = __hip_atomic_load(&ws.fz[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int tot_acc = __hip_atomic_load(&ws.fz[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int site_acc = tid < nsites ? __hip_atomic_load(&ws.fz[8 + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  __syncthreads();
  if (tid < 8 + HS_MAX_UC) ws.fz[tid] = 0;  // zero for the next step
  if (tot_inb <= 0 || tot_val <= 0) {
    // This attempt has no valid sample. Without the bitmap, the selection launch after this one
    // redraws the step. With it, the step (its counters, state and worklist) is left as it is and
    // the next launch draws the next attempt, as the selection's loop would (the same draws);
    // a call then completes one step fewer. No sample in range at all (the selection's
    // _select_single, core.py:493) cannot happen on a grid of more than 1024 supercells.
    if (mark && tid == 0) {
      *ws.acc_n = 0;
      if (tot_inb <= 0 || attempt + 1 >= MAX_ATTEMPTS) {
        hs_state_t o = S;
        o.err = 2;
        o.done = 1;
        J.st[0] = o;
      } else {
        ws.info[5] = attempt + 1;
      }
    }
    return;
  }
  if (tid < nsites) ws.trunc[tid] = (tot_acc > 1 && SEL.reduced_k != 0 && site_acc > SEL.cutoff) ? 1 : 0;
  if (mark) {
    // the marked tiles (every workgroup's marks were performed before its ticket) become the
    // step's worklist, in tile order: each thread takes a contiguous run of words
    const uint32_t *bits = WL.bits + (size_t)par * WL.nwords;
    const int nw = WL.nwords, per = (nw + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS;
    const int w0 = min(nw, tid * per), w1 = min(nw, w0 + per);
    uint32_t wv[(BITMAP_MAX_WORDS + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < (BITMAP_MAX_WORDS + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS; ++k) {
      wv[k] = w0 + k < w1 ? __hip_atomic_load(&bits[w0 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
      mine += __popc(wv[k]);
    }
    __shared__ int s_scan[PRESAMPLE_THREADS / 64];
    int ntl = 0;
    int off = block_exclusive_scan(mine, s_scan, ntl);
#pragma unroll
    for (int k = 0; k < (BITMAP_MAX_WORDS + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS; ++k) {
      uint32_t v = wv[k];
      while (v) {
        const int b = __ffs((int)v) - 1;
        v &= v - 1u;
        WL.wl[off++] = make_int2(0, 32 * (w0 + k) + b);
      }
    }
    if (tid == 0) WL.count[par] = ntl;
  }
  if (tid == 0) {
    ws.info[0] = ox;
    ws.info[1] = oy;
    ws.info[2] = 0;
    ws.info[4] = 1;
    ws.info[5] = 0;  // the next step starts at attempt 0
    *ws.acc_n = tot_acc;
    hs_state_t o = S;
    o.attempts += tot_val;
    o.switches += tot_acc;
    o.step += 1;
    o.steps_done += 1;
    J.st[0] = o;
    __hip_atomic_store(&ws.info[3], FUSED_DONE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void __launch_bounds__(PRESAMPLE_THREADS) k_metro_presample(MetroJob J, WorkList WL, int par) {
  metro_presample(J, blockIdx.x, WL, par);
}

// A batch: block b presamples block map[b].y of job map[b].x.
__global__ void __launch_bounds__(PRESAMPLE_THREADS) k_metro_presample_many(const MetroJob *__restrict__ jobs,
                                                                          const int2 *__restrict__ map) {
  const int2 e = map[blockIdx.x];
  metro_presample(jobs[e.x], e.y);
}

// Second half: metro_select on the presampled attempt (listing, truncation flags, state), one
// workgroup.
__global__ void __launch_bounds__(MS_THREADS) k_metro_select_pre(MetroJob J, WorkList WL, int par) {
  metro_select<PRESELECT_BATCH, true>(J, WL, par, 0);
}

__global__ void __launch_bounds__(MS_THREADS) k_metro_select_pre_many(const MetroJob *__restrict__ jobs,
                                                                      const int32_t *__restrict__ which, WorkList WL,
                                                                      int par) {
  const int j = which[blockIdx.x];
  metro_select<PRESELECT_BATCH, true>(jobs[j], WL, par, j);
}

// The worklist of a presampled grid step without PBC: one thread per update tile decides whether
// a switch reaches it, from the chosen supercells whose samples can (the update's direct lookup),
// with the update's reach rules (dipolar window, truncated per site; exchange stencil; the tile of
// the switch itself). Reached tiles are appended to the worklist, one atomic per wave (the order
// of the entries does not matter: every tile's update is independent of the others).
// Returns false (doing nothing) when `fallback` is set and the presampling launch did not
// complete the step.
__device__ __forceinline__ bool metro_tiles(const MetroJob &J, WorkList WL, int par, int job, int t,
                                            bool fallback = false) {
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const MetroWS &ws = J.ws;
  // the per-site truncation flags go to LDS, their loads issued with the step's counts, offsets
  // and completion mark
  __shared__ int s_tr[HS_MAX_UC];
  if ((int)threadIdx.x < I.ux * I.uy) s_tr[threadIdx.x] = ws.trunc[threadIdx.x];
  const int a = *ws.acc_n, single = ws.info[2], mark = ws.info[3];
  const int ox = ws.info[0], oy = ws.info[1];
  __syncthreads();
  if (fallback && mark != FUSED_DONE) return false;
  if (a == 0 || single != 0 || I.pbc) return true;  // nothing switched, the single-sample step or PBC (listed already)
  const int nx = I.nx, ny = I.ny, tiles_x = J.tiles_x, ntile = J.tiles_x * J.tiles_y;
  bool touched = false;
  if (t < ntile) {
    const int x0 = (t % tiles_x) * MU_TX, y0 = (t / tiles_x) * MU_TY;
    const int x1 = min(nx - 1, x0 + MU_TX - 1), y1 = min(ny - 1, y0 + MU_TY - 1);
    const bool has_dip = I.dip_pf != 0.0, has_ex = I.J != 0.0 && I.nn_w > 0;
    const int hy = I.nn_h / 2, hx = I.nn_w / 2;
    const int tkx = min(min(SEL.reduced_k, nx - 1), I.kwx), tky = min(min(SEL.reduced_k, ny - 1), I.kwy);
    const int Wx = max(has_dip ? I.kwx : 0, has_ex ? hx : 0), Wy = max(has_dip ? I.kwy : 0, has_ex ? hy : 0);
    const int nrows = SEL.nsc / SEL.ncols;
    const int cs0 = max(0, -floor_div(-(x0 - Wx - ox - SEL.Rx + 1), 2 * SEL.Rx));
    const int cs1 = min(SEL.ncols - 1, floor_div(x0 + MU_TX - 1 + Wx - ox, 2 * SEL.Rx));
    const int rs0 = max(0, -floor_div(-(y0 - Wy - oy - SEL.Ry + 1), 2 * SEL.Ry));
    const int rs1 = min(nrows - 1, floor_div(y0 + MU_TY - 1 + Wy - oy, 2 * SEL.Ry));
    const int ncs = max(0, cs1 - cs0 + 1), nq = ncs * max(0, rs1 - rs0 + 1);
    // the accepted flags and cells of the first TILE_BATCH candidate supercells are loaded
    // together (independent loads, not a chain) and then tested; any further ones in turn
    constexpr int TILE_BATCH = 12;
    int8_t af[TILE_BATCH];
    int cc[TILE_BATCH];
#pragma unroll
    for (int q = 0; q < TILE_BATCH; ++q) {
      af[q] = 0;
      cc[q] = 0;
      if (q < nq) {
        const int j = (rs0 + q / ncs) * SEL.ncols + cs0 + q % ncs;
        af[q] = ws.accf[j];
        cc[q] = ws.cand[j];
      }
    }
    for (int q = 0; q < nq && !touched; ++q) {
      int8_t acc_q = 0;
      int sc = 0;
      if (q < TILE_BATCH) {
#pragma unroll
        for (int u = 0; u < TILE_BATCH; ++u)
          if (u == q) {
            acc_q = af[u];
            sc = cc[u];
          }
      } else {
        const int j = (rs0 + q / ncs) * SEL.ncols + cs0 + q % ncs;
        acc_q = ws.accf[j];
        sc = ws.cand[j];
      }
      if (!acc_q) continue;
      const int sy = sc / nx, sx = sc % nx, ss = hs_site(I, sy, sx);
      bool rel = sx >= x0 && sx <= x1 && sy >= y0 && sy <= y1;  // the switch's own tile (its flip)
      if (!rel && has_dip && I.sub_of[ss] >= 0) {
        const bool tr = s_tr[ss] != 0;
        rel = reaches(sx, tr ? tkx : I.kwx, nx, false, x0, x0 + MU_TX - 1) &&
              reaches(sy, tr ? tky : I.kwy, ny, false, y0, y0 + MU_TY - 1);
      }
      if (!rel && has_ex)
        rel = reaches(sx, hx, nx, false, x0, x0 + MU_TX - 1) && reaches(sy, hy, ny, false, y0, y0 + MU_TY - 1);
      touched = rel;
    }
  }
  const unsigned long long ball = __ballot(touched);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && ball) base = atomicAdd(&WL.count[par], __popcll(ball));
  base = __shfl(base, 0, 64);
  if (touched) WL.wl[base + __popcll(ball & ((1ull << lane) - 1ull))] = make_int2(job, t);
  return true;
}

// One thread per tile. With `fallback` (no k_metro_select_pre launch in the step): when the
// presampling launch did not complete the step (its attempt had no valid sample, or the stop
// predicate ended the run), workgroup 0 runs the selection instead, listing the tiles itself.
__global__ void __launch_bounds__(256) k_metro_tiles(MetroJob J, WorkList WL, int par, int fallback) {
  if (!metro_tiles(J, WL, par, 0, blockIdx.x * blockDim.x + threadIdx.x, fallback != 0) && blockIdx.x == 0)
    metro_select<PRESELECT_BATCH, true, true, 256>(J, WL, par, 0);
}

// A batch: block b covers tiles 256*map[b].y .. +255 of job map[b].x.
__global__ void __launch_bounds__(256) k_metro_tiles_many(const MetroJob *__restrict__ jobs, const int2 *__restrict__ map,
                                                          WorkList WL, int par) {
  const int2 e = map[blockIdx.x];
  metro_tiles(jobs[e.x], WL, par, e.x, e.y * 256 + threadIdx.x);
}

// One selection workgroup per job of the list `which`.
__global__ void __launch_bounds__(MS_THREADS) k_metro_select_many(const MetroJob *__restrict__ jobs,
                                                                  const int32_t *__restrict__ which, WorkList WL,
                                                                  int par) {
  const int j = which[blockIdx.x];
  metro_select<SELECT_BATCH>(jobs[j], WL, par, j);
}

// The update of the step's listed tiles (one tile at a time per workgroup); then the counter and
// the tile bitmap of the next step are cleared.
__global__ void __launch_bounds__(MU_THREADS, UPDATE_WAVES) k_metro_update_one(MetroJob J, WorkList WL, int par) {
  const int n = WL.count[par];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int t = WL.wl[i].y;
    metro_update<MU_TX, MU_TY, true>(J, t % J.tiles_x, t / J.tiles_x);
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) WL.count[par ^ 1] = 0;
    if (WL.bits)
      for (int w = threadIdx.x; w < WL.nwords; w += MU_THREADS) WL.bits[(size_t)(par ^ 1) * WL.nwords + w] = 0u;
  }
}

__global__ void __launch_bounds__(MU_THREADS, UPDATE_WAVES) k_metro_update_many(const MetroJob *__restrict__ jobs, WorkList WL,
                                                                  int par) {
  const int n = WL.count[par];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int2 e = WL.wl[i];
    const MetroJob &J = jobs[e.x];
    metro_update<MU_TX, MU_TY, true>(J, e.y % J.tiles_x, e.y / J.tiles_x);
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) WL.count[par ^ 1] = 0;
}

// Clears both worklist counters (a count left by an earlier call is discarded) and the fused
// presampling counters and marks of the call's job(s): the workspace may have held other data.
__global__ void k_metro_clear(WorkList WL, MetroWS ws) {
  if (threadIdx.x < 2) WL.count[threadIdx.x] = 0;
  if (WL.bits)
    for (int w = threadIdx.x; w < 2 * WL.nwords; w += blockDim.x) WL.bits[w] = 0u;
  for (int i = threadIdx.x; i < 8 + HS_MAX_UC; i += blockDim.x) ws.fz[i] = 0;
  if (threadIdx.x == 0) {
    ws.info[3] = 0;
    ws.info[4] = 0;
    ws.info[5] = 0;
  }
}

__global__ void k_metro_clear_many(WorkList WL, const MetroJob *__restrict__ jobs, const int32_t *__restrict__ which,
                                   int n) {
  if (blockIdx.x == 0 && threadIdx.x < 2) WL.count[threadIdx.x] = 0;
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const MetroWS &ws = jobs[which[k]].ws;
    for (int i = threadIdx.x; i < 8 + HS_MAX_UC; i += blockDim.x) ws.fz[i] = 0;
    if (threadIdx.x == 0) {
      ws.info[3] = 0;
      ws.info[4] = 0;
    }
  }
}

// Persistent Metropolis stepping of small instances: one 1024-thread workgroup per job runs the
// selection, then the update of every 32x32 tile of its grid, step after step, until the job's
// stop predicate or step budget ends it. The arithmetic is that of the two-launch step (the same
// selection code; every cell adds the same sources in the same order whatever the tile shape).
template <bool F64 = false>
__device__ __forceinline__ void metro_persist(const MetroJob &J) {
  const int tx = (J.I.nx + 31) / 32, ty = (J.I.ny + 31) / 32;
  for (int64_t s = 0; s < J.R.max_steps; ++s) {
    const int accepted = metro_select<1>(J);
    __syncthreads();
    if (accepted < 0) break;
    if (accepted == 0) continue;  // nothing switched: no field changes this step
    for (int t = 0; t < tx * ty; ++t) metro_update<32, 32, false, F64>(J, t % tx, t / tx);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(MS_THREADS) k_metro_persist_one(MetroJob J) { metro_persist(J); }

// An instance with FP64 fields (E64).
__global__ void __launch_bounds__(MS_THREADS) k_metro_persist_one_f64(MetroJob J) { metro_persist<true>(J); }

__global__ void __launch_bounds__(MS_THREADS) k_metro_persist_many(const MetroJob *__restrict__ jobs,
                                                                   const int32_t *__restrict__ which) {
  metro_persist(jobs[which[blockIdx.x]]);
}

__global__ void k_state_begin_many(const MetroJob *__restrict__ jobs, int njobs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  hs_state_t s = jobs[j].st[0];
  if (jobs[j].R.begin) {
    s.t0 = s.t;
    s.attempts0 = s.attempts;
  }
  s.done = 0;
  s.steps_done = 0;
  s.err = 0;
  jobs[j].st[0] = s;
}

// ---- zero-temperature sweeps (Magnets._minimize_all, core.py:785-806) --------------------------
// One subset = the cells (y0 + i*sy, x0 + j*sx). Every magnet of it that gains energy by switching
// (switch_energy < 0, or switch_energy/2 + E_B < 0 with the barrier) switches, all together, and
// the fields follow as update_multiple. The selection lists them in the reference's index order
// (numpy.meshgrid(ys, xs) flattened: x outer, y inner); the update is metro_update over every tile.
__global__ void __launch_bounds__(MS_THREADS) k_relax_select(MetroJob J, int y0, int x0, int sy, int sx,
                                                             int with_barrier) {
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const MetroWS &ws = J.ws;
  __shared__ int s_i[MS_THREADS / 64 + 1];
  __shared__ int s_site[HS_MAX_UC];
  const int tid = threadIdx.x, nx = I.nx, ny = I.ny;
  if (tid < HS_MAX_UC) s_site[tid] = 0;
  __syncthreads();
  // core.py:797: a subset whose first cell is empty is skipped
  const bool live = (I.flags[y0 * nx + x0] & HS_FLAG_OCC) != 0;
  const int ky = live ? (ny - y0 + sy - 1) / sy : 0, kx = live ? (nx - x0 + sx - 1) / sx : 0;
  const int N = kx * ky;
  int base = 0;
  for (int b = 0; b < N; b += MS_THREADS) {
    const int j = b + tid;
    int c = -1, acc = 0;
    float mmu = 0.f;
    if (j < N) {
      c = (y0 + (j % ky) * sy) * nx + (x0 + (j / ky) * sx);
      const int mc = (int)I.m[c];
      if (mc != 0) {
        const double e = (double)I.E[c] - (double)I.Ec[c];  // E of the magnet; switch_energy = -2E
        acc = with_barrier ? ((double)I.EB[c] - e < 0.0) : (-2.0 * e < 0.0);
        mmu = (float)(-mc) * I.mu_src[c];  // (m w) after the flip
      }
    }
    int total = 0;
    const int pos = base + block_exclusive_scan(acc, s_i, total);
    if (acc) {
      ws.acc_cell[pos] = c;
      ws.acc_mmu[pos] = mmu;
      atomicAdd(&s_site[hs_site(I, c / nx, c % nx)], 1);
    }
    base += total;
  }
  __syncthreads();
  if (tid < I.ux * I.uy) ws.trunc[tid] = (base > 1 && SEL.reduced_k != 0 && s_site[tid] > SEL.cutoff) ? 1 : 0;
  if (tid == 0) {
    ws.info[0] = 0;
    ws.info[1] = 0;
    ws.info[2] = 1;  // not a supercell grid: the update filters the whole list
    *ws.acc_n = base;
    hs_state_t S = J.st[0];
    S.switches += base;  // core.py:803-804
    S.attempts += base;
    J.st[0] = S;
  }
}

// The update of a sweep's switches over every tile of the grid.
__global__ void __launch_bounds__(MU_THREADS) k_relax_update(MetroJob J) {
  const int nt = J.tiles_x * J.tiles_y;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    metro_update<MU_TX, MU_TY, true>(J, t % J.tiles_x, t / J.tiles_x);
    __syncthreads();
  }
}

// ---- sequential minimisation (Magnets.minimize, core.py:763-783) -------------------------------
// The scan of one iteration: the barrier of every cell (switch_energy = -2E, 0 where empty; or
// the smaller E_barrier of an occupied magnet), the cells with barrier < -kT in row-major order
// (numpy.where order), and the first cell of the smallest barrier (numpy.argmin).
__global__ void __launch_bounds__(MS_THREADS) k_min_scan(hs_instance_t I, int with_barrier, int32_t *cells,
                                                         int64_t *result) {
  __shared__ int s_i[MS_THREADS / 64 + 1];
  __shared__ double s_v[MS_THREADS];
  __shared__ int s_c[MS_THREADS];
  const int tid = threadIdx.x, ncell = I.nx * I.ny;
  double best = INFINITY;
  int best_c = 0x7fffffff, base = 0;
  for (int b = 0; b < ncell; b += MS_THREADS) {
    const int c = b + tid;
    int below = 0;
    if (c < ncell) {
      const int mc = (int)I.m[c];
      double bar;
      if (with_barrier)
        bar = mc != 0 ? barrier_min_f64(I, c) : INFINITY;
      else
        bar = mc != 0 ? -2.0 * ((double)I.E[c] - (double)I.Ec[c]) : 0.0;
      below = bar < -1.0 / (double)I.beta[c];  // barrier < -kT
      if (bar < best) {  // cells come in ascending order per thread: the first minimum stays
        best = bar;
        best_c = c;
      }
    }
    int total = 0;
    const int pos = base + block_exclusive_scan(below, s_i, total);
    if (below) cells[pos] = c;
    base += total;
  }
  s_v[tid] = best;
  s_c[tid] = best_c;
  __syncthreads();
  for (int w = MS_THREADS / 2; w > 0; w >>= 1) {
    if (tid < w) {
      const double v2 = s_v[tid + w];
      const int c2 = s_c[tid + w];
      if (v2 < s_v[tid] || (v2 == s_v[tid] && c2 < s_c[tid])) {
        s_v[tid] = v2;
        s_c[tid] = c2;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    result[0] = base;
    result[1] = s_c[0] == 0x7fffffff ? 0 : s_c[0];
    result[2] = (int64_t)__double_as_longlong(s_v[0]);
  }
}

// k given cells (distinct, device array) switch together: the selection list of an update whose
// fields follow as update_single (one cell) or update_multiple (several), counters += k.
__global__ void __launch_bounds__(MS_THREADS) k_flip_select(MetroJob J, const int32_t *__restrict__ cells, int k) {
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const MetroWS &ws = J.ws;
  __shared__ int s_site[HS_MAX_UC];
  const int tid = threadIdx.x, nx = I.nx, ncell = I.nx * I.ny;
  if (tid < HS_MAX_UC) s_site[tid] = 0;
  int bad = 0;
  for (int i = tid; i < k; i += MS_THREADS) bad |= (cells[i] < 0 || cells[i] >= ncell);
  if (__syncthreads_or(bad)) {  // an index outside the grid: nothing switches, the state reports it
    if (tid == 0) {
      *ws.acc_n = 0;
      J.st[0].err = 1;
    }
    return;
  }
  for (int i = tid; i < k; i += MS_THREADS) {
    const int c = cells[i];
    ws.acc_cell[i] = c;
    ws.acc_mmu[i] = (float)(-(int)I.m[c]) * I.mu_src[c];
    atomicAdd(&s_site[hs_site(I, c / nx, c % nx)], 1);
  }
  __syncthreads();
  if (tid < I.ux * I.uy) ws.trunc[tid] = (k > 1 && SEL.reduced_k != 0 && s_site[tid] > SEL.cutoff) ? 1 : 0;
  if (tid == 0) {
    ws.info[0] = 0;
    ws.info[1] = 0;
    ws.info[2] = 1;
    *ws.acc_n = k;
    hs_state_t S = J.st[0];
    S.switches += k;
    S.attempts += k;
    J.st[0] = S;
  }
}

// ---- Wolff cluster steps (Magnets._update_Wolff / _select_cluster_exchange, core.py:588-602,
// 711-718) -----------------------------------------------------------------------------------
struct WolffWS {
  int8_t *cl;  // [ncell] in the cluster
  int8_t *ck;  // [ncell] checked (had its chance, or is anti-aligned with the seed)
  float *fr;   // [ncell] this iteration's frontier value: sum of NN weights of cluster neighbours
};

// One workgroup grows the cluster of one step and lists it (row-major, numpy.where order) as the
// step's switches. The seed is _select_single's draw; every iteration, each unchecked cell with
// cluster neighbours joins with probability 1 - exp(-2 J/kT * frontier) (Swendsen-Wang bonds) on
// the Philox uniform (cell, step, tag 7 | iteration << 8).
__global__ void __launch_bounds__(MS_THREADS) k_wolff_select(MetroJob J, WolffWS W) {
  const hs_instance_t &I = J.I;
  const hs_select_t &SEL = J.SEL;
  const hs_run_t &R = J.R;
  const MetroWS &ws = J.ws;
  __shared__ int s_i[MS_THREADS / 64 + 1];
  __shared__ int s_site[HS_MAX_UC];
  __shared__ float s_nn[HS_MAX_NN];
  const int tid = threadIdx.x, nx = I.nx, ny = I.ny, ncell = nx * ny;
  hs_state_t S = J.st[0];
  const double n = (double)I.n_occ;
  if (tid < HS_MAX_UC) s_site[tid] = 0;
  if (tid < I.nn_w * I.nn_h) s_nn[tid] = I.nn_mask[tid];
  const bool go = !S.done && S.steps_done < R.max_steps &&
                  (hs_step_free(R.first_free, S.steps_done) ||
                   (hs_lower_than(S.t - S.t0, R.t_max) &&
                    hs_lower_than((double)S.attempts / n - (double)S.attempts0 / n, R.mc_max)));
  if (!go) {
    if (tid == 0) {
      *ws.acc_n = 0;
      if (!S.done) {
        S.done = 1;
        J.st[0] = S;
      }
    }
    return;
  }
  const uint64_t step = (uint64_t)S.step;
  // _select_single: np.random.randint(self.n) over the initial nonzero list (core.py:489-494)
  const int pick = (int)floor(hs_u53(hs_bits53(I.seed, step, 2u, 2ull), 0) * n);
  const int seed = I.occ_list[min(pick, I.n_occ - 1)];
  const int mseed = (int)I.m[seed];
  for (int c = tid; c < ncell; c += MS_THREADS) {
    W.cl[c] = c == seed;
    W.ck[c] = c == seed || (int)I.m[c] != mseed;
  }
  __syncthreads();
  const int hy = I.nn_h / 2, hx = I.nn_w / 2;
  const double J2 = 2.0 * I.J;
  for (uint32_t it = 0;; ++it) {
    // frontier = (1 - checked) * convolve2d(cluster, NN, 'same', wrap if PBC else fill)
    int any = 0;
    for (int c = tid; c < ncell; c += MS_THREADS) {
      float f = 0.f;
      if (!W.ck[c]) {
        const int y = c / nx, x = c % nx;
        for (int qy = 0; qy < I.nn_h; ++qy)
          for (int qx = 0; qx < I.nn_w; ++qx) {
            const float w = s_nn[qy * I.nn_w + qx];
            if (w == 0.f) continue;
            int yy = y + hy - qy, xx = x + hx - qx;
            if (I.pbc) {
              yy = hs_pmod(yy, ny);
              xx = hs_pmod(xx, nx);
            } else if (yy < 0 || yy >= ny || xx < 0 || xx >= nx) {
              continue;
            }
            if (W.cl[yy * nx + xx]) f += w;
          }
      }
      W.fr[c] = f;
      any |= f != 0.f;
    }
    if (!__syncthreads_or(any)) break;
    for (int c = tid; c < ncell; c += MS_THREADS) {
      const float f = W.fr[c];
      if (f == 0.f) continue;
      W.ck[c] = 1;
      const double p = 1.0 - exp(-J2 * (double)I.beta[c] * (double)f);  // bond probability
      const double u = hs_u53(hs_bits53(I.seed, step, 7u | (it << 8), (uint64_t)c), 0);
      if (!(u > p)) W.cl[c] = 1;
    }
    __syncthreads();
  }
  // the cluster in row-major order: the step's switches
  int base = 0;
  for (int b = 0; b < ncell; b += MS_THREADS) {
    const int c = b + tid;
    const int in = c < ncell ? (int)W.cl[c] : 0;
    int total = 0;
    const int pos = base + block_exclusive_scan(in, s_i, total);
    if (in) {
      ws.acc_cell[pos] = c;
      ws.acc_mmu[pos] = (float)(-(int)I.m[c]) * I.mu_src[c];
      atomicAdd(&s_site[hs_site(I, c / nx, c % nx)], 1);
      if (R.log) J.log_cells[pos] = c;
    }
    base += total;
  }
  __syncthreads();
  if (tid < I.ux * I.uy) ws.trunc[tid] = (base > 1 && SEL.reduced_k != 0 && s_site[tid] > SEL.cutoff) ? 1 : 0;
  if (tid == 0) {
    ws.info[0] = 0;
    ws.info[1] = 0;
    ws.info[2] = 1;
    *ws.acc_n = base;
    if (R.log) J.log_count[0] = base;
    S.switches += base;  // core.py:713-715
    S.attempts += base;
    S.step += 1;
    S.steps_done += 1;
    J.st[0] = S;
  }
}

int check_select(const hs_select_t *sel) {
  if (sel->scheme == 0) {
    if (sel->Rx < 1 || sel->Ry < 1 || sel->ncols < 1 || sel->nsc < 1 || !sel->pat_off || !sel->pat_cells)
      return hs_set_error(HS_ERR_ARG, "bad selection geometry");
    if (sel->Rx > 0xffff || sel->Ry > 0x7fff) return hs_set_error(HS_ERR_ARG, "supercell too large");
    if (sel->nsc % sel->ncols) return hs_set_error(HS_ERR_ARG, "nsc is not a multiple of ncols");
  } else if (sel->scheme == 2) {
    if (sel->snx < 1 || sel->sny < 1 || (int64_t)sel->snx * sel->sny != sel->nsc || !(sel->r_x >= 1.5) ||
        !(sel->r_y >= 1.5))
      return hs_set_error(HS_ERR_ARG, "bad Poisson geometry (background %dx%d, r %g x %g cells)", sel->snx, sel->sny,
                          sel->r_x, sel->r_y);
  } else if (sel->scheme == 3) {
    if (sel->nsc < 1 || !sel->pat_cells) return hs_set_error(HS_ERR_ARG, "an explicit sample list needs nsc >= 1 cells");
  } else if (sel->scheme != 1) {
    return hs_set_error(HS_ERR_ARG, "unsupported MULTISAMPLING_SCHEME");
  }
  return HS_OK;
}

MetroJob make_job(const hs_instance_t *inst, const hs_select_t *sel, const hs_run_t *run, hs_state_t *state, void *ws,
                  int32_t *log_cells, int32_t *log_count) {
  MetroJob J;
  J.I = *inst;
  J.SEL = *sel;
  J.SEL.nsc = sel->scheme != 1 ? sel->nsc : 1;
  J.R = *run;
  J.st = state;
  J.ws = carve(ws, J.SEL.nsc);
  J.log_cells = log_cells;
  J.log_count = log_count;
  J.tiles_x = (inst->nx + MU_TX - 1) / MU_TX;
  J.tiles_y = (inst->ny + MU_TY - 1) / MU_TY;
  return J;
}

}  // namespace

extern "C" int64_t hs_relax_ws_bytes(const hs_instance_t *inst, int32_t step_x, int32_t step_y) {
  if (!inst || inst->nx < 1 || inst->ny < 1 || step_x < 1 || step_y < 1) return -1;
  const int64_t n = (int64_t)((inst->nx + step_x - 1) / step_x) * ((inst->ny + step_y - 1) / step_y);
  return (int64_t)ws_bytes((int)n);
}

extern "C" int hs_relax_sweep(const hs_instance_t *inst, const int32_t *order, int32_t n_order, int32_t step_x,
                              int32_t step_y, int32_t with_barrier, int32_t cutoff, int32_t reduced_k,
                              hs_state_t *state, void *ws, int64_t ws_bytes_, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  rc = hs_check_f32_fields(inst, "hs_relax_sweep");
  if (rc) return rc;
  if (!state || !ws || (n_order > 0 && !order) || n_order < 0) return hs_set_error(HS_ERR_ARG, "missing order/state/ws");
  if (step_x < 1 || step_y < 1) return hs_set_error(HS_ERR_ARG, "bad subset steps %d,%d", step_x, step_y);
  if (cutoff < 0 || reduced_k < 0) return hs_set_error(HS_ERR_ARG, "negative cutoff / reduced kernel size");
  const int64_t need = hs_relax_ws_bytes(inst, step_x, step_y);
  if (ws_bytes_ < need) return hs_set_error(HS_ERR_ARG, "relax workspace of %lld bytes < %lld", (long long)ws_bytes_,
                                            (long long)need);
  for (int i = 0; i < n_order; ++i)
    if (order[i] < 0 || order[i] >= step_x * step_y) return hs_set_error(HS_ERR_ARG, "order[%d] out of range", i);
  const int ncells = (int)((inst->nx + step_x - 1) / step_x) * ((inst->ny + step_y - 1) / step_y);
  MetroJob J;
  J.I = *inst;
  J.SEL = hs_select_t{};
  J.SEL.scheme = 1;
  J.SEL.nsc = ncells;
  J.SEL.cutoff = cutoff;
  J.SEL.reduced_k = reduced_k;
  J.R = hs_run_t{};
  J.st = state;
  J.ws = carve(ws, ncells);
  J.log_cells = nullptr;
  J.log_count = nullptr;
  J.tiles_x = (inst->nx + MU_TX - 1) / MU_TX;
  J.tiles_y = (inst->ny + MU_TY - 1) / MU_TY;
  const int nt = J.tiles_x * J.tiles_y;
  const unsigned wgs = (unsigned)(nt < UPDATE_WGS ? nt : UPDATE_WGS);
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n_order; ++i) {
    const int y0 = order[i] / step_x, x0 = order[i] % step_x;
    if (y0 >= inst->ny || x0 >= inst->nx) continue;  // no cell of the grid in this subset
    k_relax_select<<<1, MS_THREADS, 0, s>>>(J, y0, x0, step_y, step_x, with_barrier ? 1 : 0);
    HS_LAUNCH_CHECK("k_relax_select");
    k_relax_update<<<wgs, MU_THREADS, 0, s>>>(J);
    HS_LAUNCH_CHECK("k_relax_update");
  }
  return HS_OK;
}

extern "C" int hs_minimize_scan(const hs_instance_t *inst, int32_t with_barrier, int32_t *cells, int64_t *result,
                                void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  rc = hs_check_f32_fields(inst, "hs_minimize_scan");
  if (rc) return rc;
  if (!cells || !result) return hs_set_error(HS_ERR_ARG, "missing cells/result");
  k_min_scan<<<1, MS_THREADS, 0, (hipStream_t)stream>>>(*inst, with_barrier ? 1 : 0, cells, result);
  HS_LAUNCH_CHECK("k_min_scan");
  return HS_OK;
}

extern "C" int64_t hs_flip_ws_bytes(const hs_instance_t *inst) {
  if (!inst || inst->nx < 1 || inst->ny < 1) return -1;
  return (int64_t)ws_bytes(inst->nx * inst->ny);
}

extern "C" int hs_flip_cells(const hs_instance_t *inst, const int32_t *cells, int32_t k, int32_t cutoff,
                             int32_t reduced_k, hs_state_t *state, void *ws, int64_t ws_bytes_, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  rc = hs_check_f32_fields(inst, "hs_flip_cells");
  if (rc) return rc;
  const int ncell = inst->nx * inst->ny;
  if (!state || !ws || k < 0 || k > ncell || (k > 0 && !cells)) return hs_set_error(HS_ERR_ARG, "bad flip arguments");
  if (cutoff < 0 || reduced_k < 0) return hs_set_error(HS_ERR_ARG, "negative cutoff / reduced kernel size");
  if (ws_bytes_ < hs_flip_ws_bytes(inst)) return hs_set_error(HS_ERR_ARG, "flip workspace too small");
  if (k == 0) return HS_OK;
  MetroJob J;
  J.I = *inst;
  J.SEL = hs_select_t{};
  J.SEL.scheme = 1;
  J.SEL.nsc = ncell;
  J.SEL.cutoff = cutoff;
  J.SEL.reduced_k = reduced_k;
  J.R = hs_run_t{};
  J.st = state;
  J.ws = carve(ws, ncell);
  J.log_cells = nullptr;
  J.log_count = nullptr;
  J.tiles_x = (inst->nx + MU_TX - 1) / MU_TX;
  J.tiles_y = (inst->ny + MU_TY - 1) / MU_TY;
  const int nt = J.tiles_x * J.tiles_y;
  hipStream_t s = (hipStream_t)stream;
  k_flip_select<<<1, MS_THREADS, 0, s>>>(J, cells, k);
  HS_LAUNCH_CHECK("k_flip_select");
  k_relax_update<<<update_grid(nt, 0), MU_THREADS, 0, s>>>(J);
  HS_LAUNCH_CHECK("k_relax_update");
  return HS_OK;
}

extern "C" int64_t hs_wolff_ws_bytes(const hs_instance_t *inst) {
  if (!inst || inst->nx < 1 || inst->ny < 1) return -1;
  const size_t ncell = (size_t)inst->nx * inst->ny;
  return (int64_t)(ws_bytes((int)ncell) + 2 * align256(ncell) + align256(4 * ncell));
}

extern "C" int hs_wolff_run(const hs_instance_t *inst, const hs_run_t *run, int32_t cutoff, int32_t reduced_k,
                            hs_state_t *state, void *ws, int64_t ws_bytes_, int32_t *log_cells, int32_t *log_count,
                            void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  rc = hs_check_f32_fields(inst, "hs_wolff_run");
  if (rc) return rc;
  if (!run || !state || !ws) return hs_set_error(HS_ERR_ARG, "missing run/state/ws");
  if (run->max_steps < 0) return hs_set_error(HS_ERR_ARG, "negative step budget");
  if (run->log && (!log_cells || !log_count || run->max_steps > 1))
    return hs_set_error(HS_ERR_ARG, "the cluster log needs its arrays and a single step");
  if (inst->J == 0.0 || inst->nn_w <= 0) return hs_set_error(HS_ERR_ARG, "Wolff clusters need an exchange energy");
  if (cutoff < 0 || reduced_k < 0) return hs_set_error(HS_ERR_ARG, "negative cutoff / reduced kernel size");
  if (ws_bytes_ < hs_wolff_ws_bytes(inst)) return hs_set_error(HS_ERR_ARG, "Wolff workspace too small");
  const int ncell = inst->nx * inst->ny;
  MetroJob J;
  J.I = *inst;
  J.SEL = hs_select_t{};
  J.SEL.scheme = 1;
  J.SEL.nsc = ncell;
  J.SEL.cutoff = cutoff;
  J.SEL.reduced_k = reduced_k;
  J.R = *run;
  J.st = state;
  J.ws = carve(ws, ncell);
  J.log_cells = log_cells;
  J.log_count = log_count;
  J.tiles_x = (inst->nx + MU_TX - 1) / MU_TX;
  J.tiles_y = (inst->ny + MU_TY - 1) / MU_TY;
  char *p = (char *)ws + ws_bytes(ncell);
  WolffWS W;
  W.cl = (int8_t *)p;
  W.ck = (int8_t *)(p + align256(ncell));
  W.fr = (float *)(p + 2 * align256(ncell));
  const int nt = J.tiles_x * J.tiles_y;
  hipStream_t s = (hipStream_t)stream;
  k_state_begin<<<1, 1, 0, s>>>(state, run->begin ? 1 : 0);
  HS_LAUNCH_CHECK("k_state_begin");
  for (int64_t j = 0; j < run->max_steps; ++j) {
    k_wolff_select<<<1, MS_THREADS, 0, s>>>(J, W);
    HS_LAUNCH_CHECK("k_wolff_select");
    k_relax_update<<<update_grid(nt, 0), MU_THREADS, 0, s>>>(J);
    HS_LAUNCH_CHECK("k_relax_update");
  }
  return HS_OK;
}

extern "C" int64_t hs_metropolis_ws_bytes(const hs_instance_t *inst, const hs_select_t *sel) {
  if (!sel || !inst || inst->nx < 1 || inst->ny < 1) return -1;
  const int nt = tiles_of(*inst);
  return (int64_t)(ws_bytes(sel->scheme != 1 ? sel->nsc : 1) + wl_bytes(nt));
}

extern "C" int hs_metropolis_run(const hs_instance_t *inst, const hs_select_t *sel, const hs_run_t *run,
                                 hs_state_t *state, void *ws, int32_t *log_cells, int32_t *log_count, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  if (!sel || !run || !state || !ws) return hs_set_error(HS_ERR_ARG, "missing sel/run/state/ws");
  if (run->max_steps < 0) return hs_set_error(HS_ERR_ARG, "negative step budget");
  if (run->log && (!log_cells || !log_count)) return hs_set_error(HS_ERR_ARG, "log requested without arrays");
  rc = check_select(sel);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (inst->E64 && inst->nx * inst->ny > PERSIST_CELLS)
    return hs_set_error(HS_ERR_ARG, "FP64 fields (E64) step on grids of at most %d cells (this one has %d)", PERSIST_CELLS,
                        inst->nx * inst->ny);
  const MetroJob J = make_job(inst, sel, run, state, ws, log_cells, log_count);
  k_state_begin<<<1, 1, 0, s>>>(state, run->begin ? 1 : 0);
  HS_LAUNCH_CHECK("k_state_begin");
  if (inst->E64) {
    k_metro_persist_one_f64<<<1, MS_THREADS, 0, s>>>(J);
    HS_LAUNCH_CHECK("k_metro_persist_one_f64");
    return HS_OK;
  }
  if (inst->nx * inst->ny <= PERSIST_CELLS) {
    k_metro_persist_one<<<1, MS_THREADS, 0, s>>>(J);
    HS_LAUNCH_CHECK("k_metro_persist_one");
    return HS_OK;
  }
  const int nt = J.tiles_x * J.tiles_y;
  char *wlp = (char *)ws + ws_bytes(J.SEL.nsc);
  // grid selection without METROPOLIS_TIME and with more samples than one batch per selection
  // thread: the samples are drawn and accepted by a wide launch
  const bool presample = J.SEL.scheme == 0 && !J.SEL.metropolis_time && J.SEL.nsc > PRESAMPLE_MIN;
  const bool fused = !inst->pbc && !run->log;
  // the presampling launch completes the step itself without PBC or log; it then also marks the
  // step's tiles in a bitmap (grids of at most BITMAP_MAX_WORDS * 32 tiles) and its last workgroup
  // lists them for the update
  const int nwords = (int)tile_words(nt);
  const bool bitmap = presample && fused && nwords <= BITMAP_MAX_WORDS;
  WorkList WL{(int2 *)wlp, (int32_t *)(wlp + align256(8 * (size_t)nt))};
  if (bitmap) {
    WL.bits = (uint32_t *)(wlp + align256(8 * (size_t)nt) + 256);
    WL.nwords = nwords;
  }
  k_metro_clear<<<1, 64, 0, s>>>(WL, J.ws);
  HS_LAUNCH_CHECK("k_metro_clear");
  const unsigned wgs = update_grid(nt, run->nblocks);
  const unsigned pblocks = (unsigned)((J.SEL.nsc + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS);
  for (int64_t j = 0; j < run->max_steps; ++j) {
    const int par = (int)(j & 1);
    if (presample) {
      k_metro_presample<<<pblocks, PRESAMPLE_THREADS, 0, s>>>(J, WL, par);
      HS_LAUNCH_CHECK("k_metro_presample");
      // without the bitmap, the tile launch runs the selection only when the presampling launch
      // did not complete the step (the rare attempt without a valid sample)
      if (!fused) {
        k_metro_select_pre<<<1, MS_THREADS, 0, s>>>(J, WL, par);
        HS_LAUNCH_CHECK("k_metro_select_pre");
      }
      if (!inst->pbc && !bitmap) {
        k_metro_tiles<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(J, WL, par, fused ? 1 : 0);
        HS_LAUNCH_CHECK("k_metro_tiles");
      }
    } else {
      k_metro_select_one<<<1, MS_THREADS, 0, s>>>(J, WL, par);
      HS_LAUNCH_CHECK("k_metro_select");
    }
    k_metro_update_one<<<wgs, MU_THREADS, 0, s>>>(J, WL, par);
    HS_LAUNCH_CHECK("k_metro_update");
  }
  return HS_OK;
}

// ---- the allocation-free variant -----------------------------------------------------------
// hs_metropolis_run_static takes its scratch (the MetroWS arrays and the worklist) from this
// file-scope array instead of a caller's workspace. Every HIP device the library runs on holds
// its own copy at its own address; host code reaches it only through the address that
// hipGetSymbolAddress reports on that device (the array's name in host code is not a device
// address).
static __device__ __align__(256) unsigned char scratch[HS_METRO_STATIC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up on the first call there.
int static_scratch(void **out) {
  int dev = 0;
  HS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return hs_set_error(HS_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hs_set_error(HS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return HS_OK;
}

}  // namespace

// Zeroes n16 16-byte words (a kernel rather than hipMemsetAsync, so the call can be captured).
__global__ void k_zero16(uint4 *p, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

extern "C" int hs_metropolis_run_static(const hs_instance_t *inst, const hs_select_t *sel, const hs_run_t *run,
                                        hs_state_t *state, int32_t *log_cells, int32_t *log_count, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  if (!sel) return hs_set_error(HS_ERR_ARG, "missing sel");
  const int64_t need = hs_metropolis_ws_bytes(inst, sel);
  if (need < 0) return hs_set_error(HS_ERR_ARG, "bad selection geometry");
  if (need > (int64_t)HS_METRO_STATIC_SCRATCH_BYTES)
    return hs_set_error(HS_ERR_ARG,
                        "this instance and selection need %lld scratch bytes; the static scratch holds %lld "
                        "(use hs_metropolis_run with a workspace)",
                        (long long)need, (long long)HS_METRO_STATIC_SCRATCH_BYTES);
  void *ws = nullptr;
  rc = static_scratch(&ws);
  if (rc) return rc;
  // the scratch keeps the last call's contents: start from zeros, as a fresh workspace does
  const int64_t n16 = (need + 15) / 16;
  const unsigned blocks = (unsigned)std::min<int64_t>((n16 + 255) / 256, 1024);
  k_zero16<<<blocks, 256, 0, (hipStream_t)stream>>>((uint4 *)ws, n16);
  HS_LAUNCH_CHECK("k_zero16");
  return hs_metropolis_run(inst, sel, run, state, ws, log_cells, log_count, stream);
}

namespace {

struct MetroLayout {
  size_t jobs_off, small_off, large_off, pmap_off, tmap_off, wl_off, total;
  std::vector<size_t> ws_off;
  std::vector<int32_t> small, large;
  int64_t ntiles;  // tiles of the two-launch (large) jobs
  bool pre;        // every large job is a grid selection without METROPOLIS_TIME: presampled steps
  std::vector<int2> pmap, tmap;  // presample blocks (job, block) and tile blocks (job, block) of the large jobs
};

bool persistent(const hs_instance_t &in) { return in.nx * in.ny <= PERSIST_CELLS; }

// A batch of small instances only runs in the persistent kernel (one launch for all steps).
// Once the batch has a large instance, its two launches per step pace the batch anyway, and the
// small instances ride along in them at no extra cost.
void metro_layout(const hs_team_t *teams, int n, MetroLayout &L) {
  L.jobs_off = 0;
  L.ntiles = 0;
  L.small.clear();
  L.large.clear();
  bool all_small = true;
  for (int i = 0; i < n; ++i) all_small = all_small && persistent(*teams[i].inst);
  for (int i = 0; i < n; ++i) {
    if (all_small) {
      L.small.push_back(i);
      continue;
    }
    L.large.push_back(i);
    L.ntiles += tiles_of(*teams[i].inst);
  }
  L.pre = !L.large.empty();
  int max_nsc = 0;
  for (int i : L.large) {
    L.pre = L.pre && teams[i].sel->scheme == 0 && !teams[i].sel->metropolis_time;
    max_nsc = std::max(max_nsc, (int)teams[i].sel->nsc);
  }
  L.pre = L.pre && max_nsc > PRESAMPLE_MIN;
  L.pmap.clear();
  L.tmap.clear();
  if (L.pre)
    for (int i : L.large) {
      for (int b = 0; b < (teams[i].sel->nsc + PRESAMPLE_THREADS - 1) / PRESAMPLE_THREADS; ++b)
        L.pmap.push_back(make_int2(i, b));
      if (!teams[i].inst->pbc)
        for (int b = 0; b < (tiles_of(*teams[i].inst) + 255) / 256; ++b) L.tmap.push_back(make_int2(i, b));
    }
  L.small_off = align256(sizeof(MetroJob) * (size_t)n);
  L.large_off = L.small_off + align256(4 * (size_t)n);
  L.pmap_off = L.large_off + align256(4 * (size_t)n);
  L.tmap_off = L.pmap_off + align256(8 * (L.pmap.size() + 1));
  L.wl_off = L.tmap_off + align256(8 * (L.tmap.size() + 1));
  size_t off = L.wl_off + wl_bytes(L.ntiles);
  L.ws_off.resize(n);
  for (int i = 0; i < n; ++i) {
    L.ws_off[i] = off;
    off += ws_bytes(teams[i].sel->scheme != 1 ? teams[i].sel->nsc : 1);
  }
  L.total = off;
}

}  // namespace

extern "C" int64_t hs_metro_teams_ws_bytes(const hs_team_t *teams, int32_t nteams) {
  if (!teams || nteams <= 0) return -1;
  for (int i = 0; i < nteams; ++i)
    if (!teams[i].inst || !teams[i].sel) return -1;
  MetroLayout L;
  metro_layout(teams, nteams, L);
  return (int64_t)L.total;
}

extern "C" int hs_metro_run_teams(const hs_team_t *teams, int32_t nteams, void *ws, int64_t ws_bytes_, void *stream) {
  hs_clear_error();
  if (!teams || nteams <= 0 || !ws) return hs_set_error(HS_ERR_ARG, "missing teams / workspace");
  int64_t steps = 0;
  for (int i = 0; i < nteams; ++i) {
    const hs_team_t &t = teams[i];
    int rc = hs_check_instance(t.inst);
    if (rc) return rc;
    rc = hs_check_f32_fields(t.inst, "hs_metro_run_teams");
    if (rc) return rc;
    if (!t.sel || !t.state) return hs_set_error(HS_ERR_ARG, "team %d has no selection / state", i);
    rc = check_select(t.sel);
    if (rc) return rc;
    if (t.run.max_steps < 0) return hs_set_error(HS_ERR_ARG, "negative step budget");
    if (t.run.log && (!t.log_cells || !t.log_idx)) return hs_set_error(HS_ERR_ARG, "team %d: log without arrays", i);
    steps = t.run.max_steps > steps ? t.run.max_steps : steps;
  }
  MetroLayout L;
  metro_layout(teams, nteams, L);
  if (ws_bytes_ < (int64_t)L.total)
    return hs_set_error(HS_ERR_ARG, "workspace %lld < %lld bytes", (long long)ws_bytes_, (long long)L.total);
  char *base = (char *)ws;
  std::vector<char> host(L.wl_off, 0);
  MetroJob *jobs = (MetroJob *)(host.data() + L.jobs_off);
  int32_t *small = (int32_t *)(host.data() + L.small_off);
  int64_t large_steps = 0;
  for (int i = 0; i < nteams; ++i) {
    const hs_team_t &t = teams[i];
    jobs[i] = make_job(t.inst, t.sel, &t.run, t.state, base + L.ws_off[i], t.log_cells, t.log_idx);
    if (L.large.empty() || std::find(L.large.begin(), L.large.end(), i) == L.large.end()) continue;
    large_steps = t.run.max_steps > large_steps ? t.run.max_steps : large_steps;
  }
  for (size_t k = 0; k < L.small.size(); ++k) small[k] = L.small[k];
  int32_t *large = (int32_t *)(host.data() + L.large_off);
  for (size_t k = 0; k < L.large.size(); ++k) large[k] = L.large[k];
  int2 *pmap = (int2 *)(host.data() + L.pmap_off);
  for (size_t k = 0; k < L.pmap.size(); ++k) pmap[k] = L.pmap[k];
  int2 *tmap = (int2 *)(host.data() + L.tmap_off);
  for (size_t k = 0; k < L.tmap.size(); ++k) tmap[k] = L.tmap[k];
  hipStream_t s = (hipStream_t)stream;
  if (int rc = hs_upload_async(base, host.data(), host.size(), s)) return rc;
  const MetroJob *dj = (const MetroJob *)(base + L.jobs_off);
  k_state_begin_many<<<(nteams + 255) / 256, 256, 0, s>>>(dj, nteams);
  HS_LAUNCH_CHECK("k_state_begin_many");
  if (!L.small.empty()) {  // a batch of small instances: one persistent workgroup each
    k_metro_persist_many<<<(unsigned)L.small.size(), MS_THREADS, 0, s>>>(dj, (const int32_t *)(base + L.small_off));
    HS_LAUNCH_CHECK("k_metro_persist_many");
  }
  steps = L.ntiles > 0 ? large_steps : 0;
  if (steps > 0) {
    char *wlp = base + L.wl_off;
    const WorkList WL{(int2 *)wlp, (int32_t *)(wlp + align256(8 * (size_t)L.ntiles))};
    const int32_t *lj = (const int32_t *)(base + L.large_off);
    k_metro_clear_many<<<(unsigned)std::min<size_t>(L.large.size(), 1024), 64, 0, s>>>(WL, dj, lj, (int)L.large.size());
    HS_LAUNCH_CHECK("k_metro_clear_many");
    int64_t req = 0;  // the largest update grid a team asks for (0: automatic)
    for (int i : L.large) req = std::max<int64_t>(req, teams[i].run.nblocks);
    const unsigned wgs = update_grid(L.ntiles, req);
    // the launches of steps j0 .. j0+count-1 on stream `st`
    auto record = [&](hipStream_t st, int64_t j0, int64_t count) -> int {
      for (int64_t j = j0; j < j0 + count; ++j) {
        const int par = (int)(j & 1);
        if (L.pre) {  // as hs_metropolis_run's large steps, for every job of the batch
          k_metro_presample_many<<<(unsigned)L.pmap.size(), PRESAMPLE_THREADS, 0, st>>>(
              dj, (const int2 *)(base + L.pmap_off));
          HS_LAUNCH_CHECK("k_metro_presample_many");
          k_metro_select_pre_many<<<(unsigned)L.large.size(), MS_THREADS, 0, st>>>(dj, lj, WL, par);
          HS_LAUNCH_CHECK("k_metro_select_pre_many");
          if (!L.tmap.empty()) {
            k_metro_tiles_many<<<(unsigned)L.tmap.size(), 256, 0, st>>>(dj, (const int2 *)(base + L.tmap_off), WL,
                                                                         par);
            HS_LAUNCH_CHECK("k_metro_tiles_many");
          }
        } else {
          k_metro_select_many<<<(unsigned)L.large.size(), MS_THREADS, 0, st>>>(dj, lj, WL, par);
          HS_LAUNCH_CHECK("k_metro_select_many");
        }
        k_metro_update_many<<<wgs, MU_THREADS, 0, st>>>(dj, WL, par);
        HS_LAUNCH_CHECK("k_metro_update_many");
      }
      return HS_OK;
    };
    // whole groups of U steps (U even, so that every group starts at parity 0) replay one captured
    // graph; the launches and arguments depend only on the workspace layout, which the key holds
    const int64_t U = hs_graph_steps() & ~1;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HS_HIP_TRY(hipStreamIsCapturing(s, &cap));
    int64_t j = 0;
    if (U >= 2 && steps >= U && cap == hipStreamCaptureStatusNone) {
      const std::vector<int64_t> key{1, (int64_t)(uintptr_t)base, nteams, (int64_t)L.large.size(),
                                     (int64_t)L.pmap.size(), (int64_t)L.tmap.size(), L.ntiles, (int64_t)L.pre,
                                     (int64_t)wgs, U};
      for (; j + U <= steps; j += U)
        if (int rc = hs_graph_launch(key, s, [&](hipStream_t cs) { return record(cs, 0, U); })) return rc;
    }
    if (int rc = record(s, j, steps - j)) return rc;
  }
  return HS_OK;
}
