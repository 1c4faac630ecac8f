// libhotspice_hip.so: error state, Philox streams, dipolar kernel construction, FP64 full
// recompute of the combined local fields, per-energy fields, totals and the Zeeman field step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "hs_device.h"
#include "hs_internal.h"

static thread_local std::string g_err;

int hs_set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

void hs_clear_error() { g_err.clear(); }

// ---- descriptor uploads ----------------------------------------------------------------------
// Launch descriptors are built on the host and copied into the caller's workspace ahead of the
// kernels that read them. A copy from pageable memory would have to be followed by a stream
// synchronisation (the buffer must outlive the copy), which makes the host wait for every kernel
// queued before it. Instead the bytes go through a ring of pinned buffers: a slot is refilled
// only after the event recorded behind its last copy has completed, which for a ring of 8 is
// long past in any ordinary sequence of launches.
namespace {
struct StageSlot {
  void *host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
  int device = -1;
  bool pending = false;
};
constexpr int STAGE_SLOTS = 8;
std::mutex g_stage_mu;
StageSlot g_stage[STAGE_SLOTS];
int g_stage_next = 0;
}  // namespace

int hs_upload_async(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return HS_OK;
  std::lock_guard<std::mutex> lock(g_stage_mu);
  StageSlot &slot = g_stage[g_stage_next];
  g_stage_next = (g_stage_next + 1) % STAGE_SLOTS;
  if (slot.pending) {
    HS_HIP_TRY(hipEventSynchronize(slot.done));
    slot.pending = false;
  }
  int dev = 0;
  HS_HIP_TRY(hipGetDevice(&dev));
  if (slot.done && slot.device != dev) {  // an event is recorded on streams of its own device
    HS_HIP_TRY(hipEventDestroy(slot.done));
    slot.done = nullptr;
  }
  if (!slot.done) {
    HS_HIP_TRY(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming));
    slot.device = dev;
  }
  if (slot.cap < bytes) {
    if (slot.host) HS_HIP_TRY(hipHostFree(slot.host));
    slot.host = nullptr;
    slot.cap = 0;
    const size_t cap = bytes < 65536 ? 65536 : bytes + bytes / 2;
    HS_HIP_TRY(hipHostMalloc(&slot.host, cap, hipHostMallocDefault));
    slot.cap = cap;
  }
  std::memcpy(slot.host, src, bytes);
  HS_HIP_TRY(hipMemcpyAsync(dst, slot.host, bytes, hipMemcpyHostToDevice, s));
  HS_HIP_TRY(hipEventRecord(slot.done, s));
  slot.pending = true;
  return HS_OK;
}

// ---- captured step graphs -------------------------------------------------------------------
// A stepping loop enqueues the same few kernels for every step, with arguments that depend only
// on the workspace layout and the step's parity. hs_graph_launch captures such a run of launches
// once, on a private stream of the device (capture records the launches without running them),
// instantiates it, and replays the executable graph on the caller's stream for every later call
// with the same key. The cache holds the GRAPH_CACHE most recently used graphs over all devices.
namespace {
struct GraphEntry {
  std::vector<int64_t> key;
  int device;
  hipGraphExec_t exec;
  uint64_t used;
};
constexpr size_t GRAPH_CACHE = 16;
constexpr int GRAPH_DEVICES = 64;
std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs;
uint64_t g_graph_clock = 0;
hipStream_t g_capture_stream[GRAPH_DEVICES] = {};
std::atomic<int> g_graph_steps{HS_GRAPH_STEPS_DEFAULT};
}  // namespace

int hs_graph_steps() { return g_graph_steps.load(std::memory_order_relaxed); }

extern "C" int hs_set_graph_steps(int32_t steps) { return g_graph_steps.exchange(steps > 0 ? steps : 0); }

int hs_graph_launch(const std::vector<int64_t> &key, hipStream_t s, const std::function<int(hipStream_t)> &record) {
  int dev = 0;
  HS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= GRAPH_DEVICES) return hs_set_error(HS_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_graph_mu);
  GraphEntry *hit = nullptr;
  for (GraphEntry &e : g_graphs)
    if (e.device == dev && e.key == key) hit = &e;
  if (!hit) {
    if (!g_capture_stream[dev]) HS_HIP_TRY(hipStreamCreateWithFlags(&g_capture_stream[dev], hipStreamNonBlocking));
    hipStream_t cs = g_capture_stream[dev];
    HS_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int rc = record(cs);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc ? rc : hs_set_error(HS_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return hs_set_error(HS_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    if (g_graphs.size() >= GRAPH_CACHE) {  // drop the least recently used graph, once it has run
      size_t lru = 0;
      for (size_t i = 1; i < g_graphs.size(); ++i)
        if (g_graphs[i].used < g_graphs[lru].used) lru = i;
      HS_HIP_TRY(hipSetDevice(g_graphs[lru].device));
      const hipError_t es = hipDeviceSynchronize();
      HS_HIP_TRY(hipSetDevice(dev));
      if (es != hipSuccess) return hs_set_error(HS_ERR_HIP, "hipDeviceSynchronize failed: %s", hipGetErrorString(es));
      HS_HIP_TRY(hipGraphExecDestroy(g_graphs[lru].exec));
      g_graphs.erase(g_graphs.begin() + (std::ptrdiff_t)lru);
    }
    g_graphs.push_back(GraphEntry{key, dev, exec, 0});
    hit = &g_graphs.back();
  }
  hit->used = ++g_graph_clock;
  HS_HIP_TRY(hipGraphLaunch(hit->exec, s));
  return HS_OK;
}

extern "C" int hs_graph_cache_clear(void) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  int cur = 0;
  HS_HIP_TRY(hipGetDevice(&cur));
  for (GraphEntry &e : g_graphs) {  // once its device has drained, nothing of the graph runs any more
    HS_HIP_TRY(hipSetDevice(e.device));
    HS_HIP_TRY(hipDeviceSynchronize());
    HS_HIP_TRY(hipGraphExecDestroy(e.exec));
  }
  g_graphs.clear();
  for (int d = 0; d < GRAPH_DEVICES; ++d)
    if (g_capture_stream[d]) {
      HS_HIP_TRY(hipSetDevice(d));
      HS_HIP_TRY(hipStreamDestroy(g_capture_stream[d]));
      g_capture_stream[d] = nullptr;
    }
  HS_HIP_TRY(hipSetDevice(cur));
  return HS_OK;
}

extern "C" int hs_version(void) { return 100; }

extern "C" const char *hs_last_error(void) { return g_err.c_str(); }

extern "C" int hs_device_info(int32_t *count, char *name, int32_t name_len) {
  int n = 0;
  HS_HIP_TRY(hipGetDeviceCount(&n));
  if (count) *count = n;
  if (name && name_len > 0) {
    name[0] = 0;
    if (n > 0) {
      hipDeviceProp_t p;
      HS_HIP_TRY(hipGetDeviceProperties(&p, 0));
      snprintf(name, name_len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    }
  }
  return HS_OK;
}

#include <cstddef>

// sizeof / offsetof of the ABI structs, so bindings can verify their mirrors without a device.
extern "C" int hs_abi_layout(int64_t *out, int32_t n) {
  const int64_t v[] = {
      (int64_t)sizeof(hs_instance_t), (int64_t)offsetof(hs_instance_t, seed), (int64_t)offsetof(hs_instance_t, nx),
      (int64_t)offsetof(hs_instance_t, sub_of), (int64_t)offsetof(hs_instance_t, nn_mask),
      (int64_t)sizeof(hs_kernel_geom_t), (int64_t)offsetof(hs_kernel_geom_t, nx), (int64_t)offsetof(hs_kernel_geom_t, uc_o),
      (int64_t)sizeof(hs_run_t), (int64_t)offsetof(hs_run_t, max_steps), (int64_t)offsetof(hs_run_t, begin),
      (int64_t)sizeof(hs_state_t), (int64_t)offsetof(hs_state_t, done),
      (int64_t)sizeof(hs_select_t), (int64_t)offsetof(hs_select_t, metropolis_time),
      (int64_t)sizeof(hs_team_t), (int64_t)offsetof(hs_team_t, run), (int64_t)offsetof(hs_team_t, nblocks),
      (int64_t)offsetof(hs_instance_t, pair_src_unit), (int64_t)sizeof(hs_monopole_geom_t)};
  const int cnt = (int)(sizeof(v) / sizeof(v[0]));
  if (!out || n < cnt) return cnt;
  for (int i = 0; i < cnt; ++i) out[i] = v[i];
  return cnt;
}

int hs_check_instance(const hs_instance_t *I) {
  if (!I) return hs_set_error(HS_ERR_ARG, "instance is NULL");
  if (I->nx <= 0 || I->ny <= 0) return hs_set_error(HS_ERR_ARG, "bad grid %dx%d", I->nx, I->ny);
  if ((int64_t)I->nx * I->ny > (1ll << 30)) return hs_set_error(HS_ERR_ARG, "grid too large");
  if (I->ux <= 0 || I->uy <= 0 || I->ux * I->uy > HS_MAX_UC)
    return hs_set_error(HS_ERR_ARG, "unit cell %dx%d exceeds %d sites", I->ux, I->uy, HS_MAX_UC);
  if (I->nn_w < 0 || I->nn_h < 0 || I->nn_w * I->nn_h > HS_MAX_NN || (I->nn_w && !(I->nn_w & 1)) ||
      (I->nn_h && !(I->nn_h & 1)))
    return hs_set_error(HS_ERR_ARG, "exchange mask %dx%d must be odd and <= %d entries", I->nn_w, I->nn_h, HS_MAX_NN);
  if (!I->m || !I->flags || !I->E || !I->Ec || !I->EB || !I->beta || !I->mu || !I->mu64 || !I->occ_list)
    return hs_set_error(HS_ERR_ARG, "missing state array");
  if (I->use_perp && (!I->Ep || !I->Epc)) return hs_set_error(HS_ERR_ARG, "use_perp without E_perp arrays");
  if (I->dip_pf != 0.0) {
    if (!I->K32 || !I->K64) return hs_set_error(HS_ERR_ARG, "dipolar energy without kernels");
    if (!I->pw64 || !I->mu_src) return hs_set_error(HS_ERR_ARG, "pair energy without its weights");
    if (I->use_perp && (!I->Kpo32 || !I->Kps64)) return hs_set_error(HS_ERR_ARG, "perp fields without perp kernels");
    if (I->kwx < 0 || I->kwx > I->nx - 1 || I->kwy < 0 || I->kwy > I->ny - 1)
      return hs_set_error(HS_ERR_ARG, "kernel window %d,%d outside grid", I->kwx, I->kwy);
    if (I->n_sub <= 0) return hs_set_error(HS_ERR_ARG, "no dipolar kernels");
    for (int s = 0; s < I->ux * I->uy; ++s)
      if (I->sub_of[s] >= I->n_sub) return hs_set_error(HS_ERR_ARG, "sub_of[%d] out of range", s);
  }
  if (I->n_occ <= 0 || I->n_occ > I->nx * I->ny) return hs_set_error(HS_ERR_ARG, "bad n_occ %d", I->n_occ);
  if (I->E64 && (I->use_perp || I->pair_src_unit))
    return hs_set_error(HS_ERR_ARG, "FP64 fields (E64) hold neither E_perp nor DiMonopolarEnergy pair terms");
  return HS_OK;
}

int hs_check_f32_fields(const hs_instance_t *I, const char *who) {
  if (I && I->E64) return hs_set_error(HS_ERR_ARG, "%s steps FP32 fields only; this instance holds FP64 fields (E64)", who);
  return HS_OK;
}

// ---- Philox streams --------------------------------------------------------------------------
__global__ void k_philox_u53(uint64_t seed, uint64_t step, uint32_t tag, int64_t index0, int64_t n,
                             int open_zero, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = hs_u53(hs_bits53(seed, step, tag, (uint64_t)(index0 + i)), open_zero);
}

extern "C" int hs_philox_uniform53(uint64_t seed, uint64_t step, uint32_t tag, int64_t index0, int64_t n,
                                   int32_t open_zero, double *out, void *stream) {
  if (n < 0 || index0 < 0 || (n > 0 && !out)) return hs_set_error(HS_ERR_ARG, "bad philox arguments");
  if (n == 0) return HS_OK;
  const int bs = 256;
  k_philox_u53<<<(unsigned)((n + bs - 1) / bs), bs, 0, (hipStream_t)stream>>>(seed, step, tag, index0, n,
                                                                                open_zero, out);
  HS_LAUNCH_CHECK("k_philox_u53");
  return HS_OK;
}

// ---- dipolar kernels (energies.py:176-288) -----------------------------------------------------
struct KernVals {
  double k, ps, po;
};

// Unfolded kernel entry at full-matrix position (R, C) of kernel n, before the 1e-7 factor.
HS_INLINE KernVals kern_entry(const hs_kernel_geom_t &g, int n, int R, int C) {
  KernVals v{0.0, 0.0, 0.0};
  if (R == g.ny - 1 && C == g.nx - 1) return v;  // centre: rr_sq = inf (energies.py:218)
  const double rx = g.rrx[(size_t)n * (2 * g.nx - 1) + C];
  const double ry = g.rry[(size_t)n * (2 * g.ny - 1) + R];
  const double rr_sq = rx * rx + ry * ry;
  if (rr_sq > g.cutoff * g.cutoff) return v;  // energies.py:219
  const double rr_inv = pow(rr_sq, -0.5);
  const double rinv3 = pow(rr_inv, -g.decay);
  const double rinv5 = pow(rr_inv, -g.decay + 2.0);
  const double ux = rx * rr_inv, uy = ry * rr_inv;
  const int site1 = g.site_y[n] * g.ux + g.site_x[n];
  const int site2 = ((R + g.shift_y[n]) % g.uy) * g.ux + ((C + g.shift_x[n]) % g.ux);
  if (g.in_plane) {
    const double ox1 = g.uc_ox[site1], oy1 = g.uc_oy[site1];
    const double ox2 = g.uc_ox[site2], oy2 = g.uc_oy[site2];
    const double a3x = 3 * (ux * ux) - 1, a3y = 3 * (uy * uy) - 1, a5x = 5 * (ux * ux) - 1, a5y = 5 * (uy * uy) - 1;
    const double uxy = ux * uy, c32 = 3 * g.I / 2;
    auto pair = [&](double p1, double q1, double p2, double q2) {
      const double k = p1 * p2 * a3x + q1 * q2 * a3y + 3 * uxy * (p1 * q2 + q1 * p2);
      const double kc = p1 * p2 * a5x + q1 * q2 * a5y + 5 * uxy * (p1 * q2 + q1 * p2);
      return -rinv3 * k - c32 * rinv5 * kc;
    };
    v.k = pair(ox1, oy1, ox2, oy2);
    v.ps = pair(-oy1, ox1, ox2, oy2);
    v.po = pair(ox1, oy1, -oy2, ox2);
  } else {
    const double o2 = g.uc_o[site2];
    v.k = o2 * (rinv3 + 9 * g.I / 2 * rinv5);
  }
  return v;
}

HS_INLINE void kv_add(KernVals &a, const KernVals &b) {
  a.k += b.k;
  a.ps += b.ps;
  a.po += b.po;
}

__global__ void k_build_kernels(hs_kernel_geom_t g, double *K64, double *Kps64, double *Kpo64, float *K32,
                                float *Kpo32) {
  const int W = 2 * g.kwx + 1, H = 2 * g.kwy + 1;
  const int64_t per = (int64_t)W * H;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * g.n_sub) return;
  const int n = (int)(i / per);
  const int rem = (int)(i % per);
  const int r = rem / W, c = rem % W;
  const int R = g.ny - 1 - g.kwy + r, C = g.nx - 1 - g.kwx + c;
  const int nx = g.nx, ny = g.ny;
  KernVals v = kern_entry(g, n, R, C);
  if (g.pbc) {  // energies.py:264-273, statement order
    if (C >= nx) kv_add(v, kern_entry(g, n, R, C - nx));
    if (R >= ny && C >= nx) kv_add(v, kern_entry(g, n, R - ny, C - nx));
    if (R >= ny) kv_add(v, kern_entry(g, n, R - ny, C));
    if (R >= ny && C <= nx - 2) kv_add(v, kern_entry(g, n, R - ny, C + nx));
    if (C <= nx - 2) kv_add(v, kern_entry(g, n, R, C + nx));
    if (R <= ny - 2 && C <= nx - 2) kv_add(v, kern_entry(g, n, R + ny, C + nx));
    if (R <= ny - 2) kv_add(v, kern_entry(g, n, R + ny, C));
    if (R <= ny - 2 && C >= nx) kv_add(v, kern_entry(g, n, R + ny, C - nx));
  }
  v.k *= 1e-7;
  v.ps *= 1e-7;
  v.po *= 1e-7;
  K64[i] = v.k;
  K32[i] = (float)v.k;
  if (Kps64) Kps64[i] = v.ps;
  if (Kpo64) Kpo64[i] = v.po;
  if (Kpo32) Kpo32[i] = (float)v.po;
}

extern "C" int hs_build_dipolar_kernels(const hs_kernel_geom_t *g, double *K64, double *Kps64, double *Kpo64,
                                        float *K32, float *Kpo32, void *stream) {
  hs_clear_error();
  if (!g || !g->rrx || !g->rry || !K64 || !K32) return hs_set_error(HS_ERR_ARG, "missing kernel arrays");
  if (g->n_sub <= 0 || g->n_sub > HS_MAX_UC || g->ux <= 0 || g->uy <= 0 || g->ux * g->uy > HS_MAX_UC)
    return hs_set_error(HS_ERR_ARG, "bad unit cell / n_sub");
  if (g->kwx < 0 || g->kwx > g->nx - 1 || g->kwy < 0 || g->kwy > g->ny - 1)
    return hs_set_error(HS_ERR_ARG, "kernel window outside grid");
  if (g->pbc && (g->kwx != g->nx - 1 || g->kwy != g->ny - 1))
    return hs_set_error(HS_ERR_ARG, "PBC needs the full kernel window");
  for (int n = 0; n < g->n_sub; ++n)
    if (g->site_x[n] < 0 || g->site_x[n] >= g->ux || g->site_y[n] < 0 || g->site_y[n] >= g->uy ||
        g->shift_x[n] < 0 || g->shift_y[n] < 0)
      return hs_set_error(HS_ERR_ARG, "bad site %d", n);
  const int64_t total = (int64_t)(2 * g->kwx + 1) * (2 * g->kwy + 1) * g->n_sub;
  const int bs = 256;
  k_build_kernels<<<(unsigned)((total + bs - 1) / bs), bs, 0, (hipStream_t)stream>>>(*g, K64, Kps64, Kpo64, K32, Kpo32);
  HS_LAUNCH_CHECK("k_build_kernels");
  return HS_OK;
}

// ---- DiMonopolarEnergy kernels (energies.py:406-641) -------------------------------------------
// One magnet = charge +1 at (x + h cos a, y + h sin a) and -1 at (x - h cos a, y - h sin a). The
// positions, differences and squares are evaluated in the reference's order without contraction,
// so the FP64 squared distances (and their FP32 roundings for K) are the reference's.
struct Dumbbell {
  double nx_, ny_, sx_, sy_;
};

HS_INLINE Dumbbell dumbbell(double x, double y, double h, double c, double s) {
#pragma clang fp contract(off)
  const double hc = h * c, hs = h * s;
  return Dumbbell{x + hc, y + hs, x - hc, y - hs};
}

// 1/|a-b|, 0 where the charges coincide (energies.py:516-525): FP32 square root of the FP32-rounded
// square for K, FP64 for the perpendicular kernels.
template <bool F32>
HS_INLINE double inv_dist(double ax, double ay, double bx, double by) {
#pragma clang fp contract(off)
  const double dx = ax - bx, dy = ay - by;
  const double sq = dx * dx + dy * dy;
  if (F32) {
    const float f = (float)sq;
    if (f == 0.f) return 0.0;
    return (double)(float)(1.0 / sqrt((double)f));
  }
  if (sq == 0.0) return 0.0;
  return pow(sq, -0.5);
}

// NN + SS - NS - SN between the target magnet t and the centre magnet c; the cross terms of the
// centre with itself are dropped (energies.py:494-498).
template <bool F32>
HS_INLINE double pair_sum(const Dumbbell &t, const Dumbbell &c, bool centre) {
#pragma clang fp contract(off)
  const double nn = inv_dist<F32>(t.nx_, t.ny_, c.nx_, c.ny_);
  const double ss = inv_dist<F32>(t.sx_, t.sy_, c.sx_, c.sy_);
  const double ns = centre ? 0.0 : inv_dist<F32>(t.nx_, t.ny_, c.sx_, c.sy_);
  const double sn = centre ? 0.0 : inv_dist<F32>(t.sx_, t.sy_, c.nx_, c.ny_);
  return nn + ss + -ns + -sn;
}

HS_INLINE KernVals mono_entry(const hs_kernel_geom_t &g, const hs_monopole_geom_t &mg, int n, int R, int C) {
  KernVals v{0.0, 0.0, 0.0};
  const int s1 = g.site_y[n] * g.ux + g.site_x[n];
  const int s2 = ((R + g.shift_y[n]) % g.uy) * g.ux + ((C + g.shift_x[n]) % g.ux);
  const double occ = g.uc_o[s2];
  if (occ == 0.0) return v;
  const bool centre = R == g.ny - 1 && C == g.nx - 1;
  const double *X = g.rrx + (size_t)n * (2 * g.nx - 1), *Y = g.rry + (size_t)n * (2 * g.ny - 1);
  const double xc = X[g.nx - 1], yc = Y[g.ny - 1], xt = X[C], yt = Y[R];
  const Dumbbell cN = dumbbell(xc, yc, mg.h[s1], mg.ca[s1], mg.sa[s1]);     // centre, as it is
  const Dumbbell cP = dumbbell(xc, yc, mg.h_small, mg.cp[s1], mg.sp[s1]);  // centre, perpendicular
  const Dumbbell tN = dumbbell(xt, yt, mg.h[s2], mg.ca[s2], mg.sa[s2]);
  v.k = pair_sum<true>(tN, cN, centre) * occ;
  if (centre) {  // every term of the centre entry vanishes
    v.k = v.ps = v.po = 0.0;
    return v;
  }
  v.ps = pair_sum<false>(tN, cP, false) * occ;
  const Dumbbell tP = dumbbell(xt, yt, mg.h_small, mg.cp[s2], mg.sp[s2]);
  v.po = pair_sum<false>(tP, cN, false) * occ;
  return v;
}

__global__ void k_build_mono_kernels(hs_kernel_geom_t g, const hs_monopole_geom_t *__restrict__ mgp, double *K64,
                                     double *Kps64, double *Kpo64, float *K32, float *Kpo32) {
  const hs_monopole_geom_t &mg = *mgp;
  const int W = 2 * g.nx - 1, H = 2 * g.ny - 1;
  const int64_t per = (int64_t)W * H;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * g.n_sub) return;
  const int n = (int)(i / per);
  const int rem = (int)(i % per);
  const int R = rem / W, C = rem % W;
  const int nx = g.nx, ny = g.ny;
  KernVals v = mono_entry(g, mg, n, R, C);
  if (g.pbc) {  // energies.py:610-623, statement order
    if (C >= nx) kv_add(v, mono_entry(g, mg, n, R, C - nx));
    if (R >= ny && C >= nx) kv_add(v, mono_entry(g, mg, n, R - ny, C - nx));
    if (R >= ny) kv_add(v, mono_entry(g, mg, n, R - ny, C));
    if (R >= ny && C <= nx - 2) kv_add(v, mono_entry(g, mg, n, R - ny, C + nx));
    if (C <= nx - 2) kv_add(v, mono_entry(g, mg, n, R, C + nx));
    if (R <= ny - 2 && C <= nx - 2) kv_add(v, mono_entry(g, mg, n, R + ny, C + nx));
    if (R <= ny - 2) kv_add(v, mono_entry(g, mg, n, R + ny, C));
    if (R <= ny - 2 && C >= nx) kv_add(v, mono_entry(g, mg, n, R + ny, C - nx));
  }
  v.k *= 1e-7;
  v.ps *= 1e-7;
  v.po *= 1e-7;
  K64[i] = v.k;
  K32[i] = (float)v.k;
  if (Kps64) Kps64[i] = v.ps;
  if (Kpo64) Kpo64[i] = v.po;
  if (Kpo32) Kpo32[i] = (float)v.po;
}

extern "C" int hs_build_dimonopolar_kernels(const hs_kernel_geom_t *g, const hs_monopole_geom_t *mg, double *K64,
                                            double *Kps64, double *Kpo64, float *K32, float *Kpo32, void *stream) {
  hs_clear_error();
  if (!g || !mg || !g->rrx || !g->rry || !K64 || !K32) return hs_set_error(HS_ERR_ARG, "missing kernel arrays");
  if (g->nx <= 0 || g->ny <= 0) return hs_set_error(HS_ERR_ARG, "bad grid %dx%d", g->nx, g->ny);
  if (g->n_sub <= 0 || g->n_sub > HS_MAX_UC || g->ux <= 0 || g->uy <= 0 || g->ux * g->uy > HS_MAX_UC)
    return hs_set_error(HS_ERR_ARG, "bad unit cell / n_sub");
  if (g->kwx != g->nx - 1 || g->kwy != g->ny - 1)
    return hs_set_error(HS_ERR_ARG, "DiMonopolar kernels use the full window (kwx = nx-1, kwy = ny-1)");
  for (int n = 0; n < g->n_sub; ++n)
    if (g->site_x[n] < 0 || g->site_x[n] >= g->ux || g->site_y[n] < 0 || g->site_y[n] >= g->uy ||
        g->shift_x[n] < 0 || g->shift_y[n] < 0)
      return hs_set_error(HS_ERR_ARG, "bad site %d", n);
  const int64_t total = (int64_t)(2 * g->nx - 1) * (2 * g->ny - 1) * g->n_sub;
  const int bs = 256;
  k_build_mono_kernels<<<(unsigned)((total + bs - 1) / bs), bs, 0, (hipStream_t)stream>>>(*g, mg, K64, Kps64, Kpo64,
                                                                                         K32, Kpo32);
  HS_LAUNCH_CHECK("k_build_mono_kernels");
  return HS_OK;
}

// ---- per-cell closed forms: Zeeman factors and exchange sums ------------------------------------
HS_INLINE void zeeman_factors(const hs_instance_t &I, int c, int site, double Bx, double By, double &f,
                              double &fp) {
  const double mu = I.mu64[c];
  if (I.in_plane) {
    const double ox = I.ox_uc[site], oy = I.oy_uc[site];
    f = -mu * (Bx * ox + By * oy);   // energies.py:108
    fp = -mu * (By * ox - Bx * oy);  // energies.py:109
  } else {
    f = -mu * Bx;  // energies.py:112
    fp = 0.0;
  }
}

HS_INLINE void field_at(const hs_instance_t &I, int c, double &Bx, double &By) {
  Bx = I.zee_bx ? (double)I.zee_bx[c] : I.zee_Bx;
  By = I.zee_by ? (double)I.zee_by[c] : I.zee_By;
}

// convolve2d(., NN, 'same', fill|wrap) sums at cell (y, x): out = sum_q NN[q] * v[i + h - q].
HS_INLINE void exchange_sums(const hs_instance_t &I, int y, int x, double &sx, double &sy) {
  sx = 0.0;
  sy = 0.0;
  const int hy = I.nn_h / 2, hx = I.nn_w / 2;
  for (int qy = 0; qy < I.nn_h; ++qy)
    for (int qx = 0; qx < I.nn_w; ++qx) {
      const float wq = I.nn_mask[qy * I.nn_w + qx];
      if (wq == 0.f) continue;
      int yy = y + hy - qy, xx = x + hx - qx;
      if (I.pbc) {
        yy = hs_pmod(yy, I.ny);
        xx = hs_pmod(xx, I.nx);
      } else if (yy < 0 || yy >= I.ny || xx < 0 || xx >= I.nx) {
        continue;
      }
      const int cc = yy * I.nx + xx;
      const double mj = (double)I.m[cc];
      if (I.in_plane) {
        const int s = hs_site(I, yy, xx);
        sx += wq * (I.ox_uc[s] * mj);
        sy += wq * (I.oy_uc[s] * mj);
      } else {
        sx += wq * mj;
      }
    }
}

// Exchange energy and E_perp of cell c (energies.py:771-781).
HS_INLINE void exchange_fields(const hs_instance_t &I, int y, int x, double mi, int site, double &e, double &ep) {
  e = 0.0;
  ep = 0.0;
  if (I.J == 0.0 || mi == 0.0) return;
  double sx, sy;
  exchange_sums(I, y, x, sx, sy);
  if (I.in_plane) {
    const double mx = I.ox_uc[site] * mi, my = I.oy_uc[site] * mi;
    e = -I.J * (sx * mx + sy * my);
    ep = -I.J * (sy * mx - sx * my);
  } else {
    e = -I.J * (sx * mi);
  }
}

// ---- FP64 full recompute (energies.py:290-308 + Zeeman/Exchange update) ------------------------
//
// DipolarEnergy.update computes, per target sublattice p (the unit-cell site of kernel n_p),
// E_i = prefactor * mu_i * m_i * sum_j K_{n_p}[centre + (j - i)] * (m mu)_j. Restricted to one
// source sublattice q, targets and sources sit on coarse grids a, a' (y = site_y + uy*a,
// x = site_x + ux*b), so each (p, q) pair is a dense 2D correlation with the decimated kernel
// D_pq[a'-a][b'-b]. One workgroup owns a 32x64 coarse target tile of one p and a group of
// source tiles (split for parallelism); it stages the source tile (m mu) and the 63x127 kernel
// block in LDS and each thread accumulates 8 consecutive outputs with a register window that
// slides along the kernel row (15 LDS reads per 64 FMAs). Per-group partial sums are reduced
// in a fixed order by k_recompute_finish, which also adds the Zeeman and exchange closed forms.
constexpr int RA = 32, RB = 64, RR = 8;  // coarse tile rows, cols, outputs per thread (256 threads)
constexpr int KBR = 2 * RA - 1, KBC = 2 * RB - 1;

struct RecPlan {
  int Ay, Bx;      // coarse grid extents ceil(ny/uy), ceil(nx/ux)
  int tA, tB;      // target tiles along a, b
  int F;           // source-tile groups per target tile
  int sites_y[HS_MAX_UC], sites_x[HS_MAX_UC];  // unit-cell site of kernel n
};

HS_INLINE int div_floor(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
HS_INLINE int div_ceil(int a, int b) { return -div_floor(-a, b); }

// Coarse offset range da with |dy0 + u*da| <= kw.
HS_INLINE void coarse_range(int dy0, int u, int kw, int &lo, int &hi) {
  lo = div_ceil(-kw - dy0, u);
  hi = div_floor(kw - dy0, u);
}

__device__ __forceinline__ void rec_dip(const hs_instance_t &I, const RecPlan &plan, const double *Kbase,
                                        double *part, int tile, int p, int g) {
  __shared__ double sK[KBR * KBC];
  __shared__ double sS[RA * RB];
  const int tid = threadIdx.x;
  const int a0 = (tile / plan.tB) * RA, b0 = (tile % plan.tB) * RB;
  const int py = plan.sites_y[p], px = plan.sites_x[p];
  const int KW = 2 * I.kwx + 1;
  const size_t KS = (size_t)KW * (2 * I.kwy + 1);
  const double *Kn = Kbase + (size_t)p * KS;
  // consecutive lanes take consecutive target rows: their kernel reads are one row (127 doubles)
  // apart, which spreads a wave over all LDS banks (column groups would hit the same banks)
  const int ta = tid % RA, tbg = tid / RA;
  double acc[RR];
#pragma unroll
  for (int j = 0; j < RR; ++j) acc[j] = 0.0;
  // count the (q, source tile) chunks of this target tile; group g takes an even slice
  int total = 0;
  for (int q = 0; q < I.n_sub; ++q) {
    int dlo, dhi, elo, ehi;
    coarse_range(plan.sites_y[q] - py, I.uy, I.kwy, dlo, dhi);
    coarse_range(plan.sites_x[q] - px, I.ux, I.kwx, elo, ehi);
    const int r0 = max(0, a0 + dlo), r1 = min(plan.Ay - 1, a0 + RA - 1 + dhi);
    const int c0 = max(0, b0 + elo), c1 = min(plan.Bx - 1, b0 + RB - 1 + ehi);
    if (r0 <= r1 && c0 <= c1) total += ((r1 - r0) / RA + 1) * ((c1 - c0) / RB + 1);
  }
  const int lo = (int)((int64_t)total * g / plan.F), hi = (int)((int64_t)total * (g + 1) / plan.F);
  int idx = 0;
  for (int q = 0; q < I.n_sub && idx < hi; ++q) {
    const int qy = plan.sites_y[q], qx = plan.sites_x[q];
    const int dy0 = qy - py, dx0 = qx - px;
    int dlo, dhi, elo, ehi;
    coarse_range(dy0, I.uy, I.kwy, dlo, dhi);
    coarse_range(dx0, I.ux, I.kwx, elo, ehi);
    const int r0 = max(0, a0 + dlo), r1 = min(plan.Ay - 1, a0 + RA - 1 + dhi);
    const int c0 = max(0, b0 + elo), c1 = min(plan.Bx - 1, b0 + RB - 1 + ehi);
    if (r0 > r1 || c0 > c1) continue;
    for (int as0 = r0; as0 <= r1 && idx < hi; as0 += RA)
      for (int bs0 = c0; bs0 <= c1 && idx < hi; bs0 += RB, ++idx) {
        if (idx < lo) continue;
        __syncthreads();
        // source tile (m mu) on sublattice q
        for (int e = tid; e < RA * RB; e += 256) {
          const int r = e / RB, c = e % RB;
          const int a = as0 + r, b = bs0 + c;
          const int y = qy + I.uy * a, x = qx + I.ux * b;
          double v = 0.0;
          if (a <= r1 && b <= c1 && y < I.ny && x < I.nx) {
            const int cc = y * I.nx + x;
            v = I.pair_src_unit ? (double)I.m[cc] : (double)I.m[cc] * I.pw64[cc];
          }
          sS[e] = v;
        }
        // kernel block: row ra <-> da = as0 - a0 - (RA-1) + ra, col cb <-> db = bs0 - b0 - (RB-1) + cb
        for (int e = tid; e < KBR * KBC; e += 256) {
          const int ra = e / KBC, cb = e % KBC;
          const int dy = dy0 + I.uy * (as0 - a0 - (RA - 1) + ra);
          const int dx = dx0 + I.ux * (bs0 - b0 - (RB - 1) + cb);
          double v = 0.0;
          if (dy >= -I.kwy && dy <= I.kwy && dx >= -I.kwx && dx <= I.kwx)
            v = Kn[(size_t)(I.kwy + dy) * KW + (I.kwx + dx)];
          sK[e] = v;
        }
        __syncthreads();
        // acc[j] += sum_{r,c} K[r - ta + RA-1][c - (8 tbg + j) + RB-1] * S[r][c]
        for (int r = 0; r < RA; ++r) {
          const double *krow = sK + (r - ta + RA - 1) * KBC + (RB - 1) - RR * tbg;
          const double *srow = sS + r * RB;
          for (int c0b = 0; c0b < RB; c0b += RR) {
            double w[2 * RR - 1];  // w[t] = krow[c0b - (RR-1) + t]
#pragma unroll
            for (int t = 0; t < 2 * RR - 1; ++t) w[t] = krow[c0b - (RR - 1) + t];
#pragma unroll
            for (int cc = 0; cc < RR; ++cc) {
              const double s = srow[c0b + cc];
#pragma unroll
              for (int j = 0; j < RR; ++j) acc[j] = fma(w[cc - j + RR - 1], s, acc[j]);
            }
          }
        }
      }
  }
  // partial sums of this group for the tile's targets
  const int a = a0 + ta;
  const int y = py + I.uy * a;
#pragma unroll
  for (int j = 0; j < RR; ++j) {
    const int b = b0 + RR * tbg + j;
    const int x = px + I.ux * b;
    if (a < plan.Ay && b < plan.Bx && y < I.ny && x < I.nx) part[(size_t)g * I.nx * I.ny + (size_t)y * I.nx + x] = acc[j];
  }
}

__global__ void __launch_bounds__(256) k_recompute_dip(hs_instance_t I, RecPlan plan, const double *Kbase,
                                                       double *part) {
  rec_dip(I, plan, Kbase, part, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Several recomputes in two launches, one for the small grids and one for the tiled ones: block b
// works for job block_job[b] (a tiled job's blocks in the order tile, then sublattice p, then
// source group g).
__device__ __forceinline__ void rec_small(const hs_instance_t &I, const double *Kbase, double *part, int c);

struct RecJob {
  hs_instance_t I;
  RecPlan plan;
  const double *K;
  double *part;
  int32_t first, small;
};

// Tiled jobs: blocks offset + blockIdx.x of the block map (the small jobs' blocks come first and
// run in k_recompute_small_many, whose LDS footprint differs).
__global__ void __launch_bounds__(256) k_recompute_dip_many(const RecJob *__restrict__ jobs,
                                                            const int32_t *__restrict__ block_job, int offset) {
  const int b = offset + (int)blockIdx.x;
  const RecJob &J = jobs[block_job[b]];
  const int lb = b - J.first;
  const int nt = J.plan.tA * J.plan.tB;
  rec_dip(J.I, J.plan, J.K, J.part, lb % nt, (lb / nt) % J.I.n_sub, lb / (nt * J.I.n_sub));
}

__global__ void __launch_bounds__(1024) k_recompute_small_lds_many(const RecJob *__restrict__ jobs,
                                                                   const int32_t *__restrict__ job_of_block);

__global__ void __launch_bounds__(256) k_recompute_small_many(const RecJob *__restrict__ jobs,
                                                              const int32_t *__restrict__ block_job) {
  const RecJob &J = jobs[block_job[blockIdx.x]];
  rec_small(J.I, J.K, J.part, ((int)blockIdx.x - J.first) * 256 + threadIdx.x);
}

__device__ __forceinline__ void rec_finish(const hs_instance_t &I, int F, const double *partE, const double *partP,
                                           int which, int c) {
  const int ncell = I.nx * I.ny;
  if (c >= ncell) return;
  const int y = c / I.nx, x = c % I.nx, site = hs_site(I, y, x);
  const int n = I.dip_pf != 0.0 ? I.sub_of[site] : -1;
  const double mi = (double)I.m[c];
  const bool do_e = which & 1, do_p = (which & 2) && I.use_perp;
  double e = 0.0, ep = 0.0;
  if (mi != 0.0) {
    // prefactor*moment*(partial_m*conv) (energies.py:303, 307); DiMonopolarEnergy:
    // prefactor*(partial_m*conv(m))*charge_sq (energies.py:654-659)
    if (n >= 0) {
      const double mu = I.pw64[c];
      double acc = 0.0, accp = 0.0;
      for (int g = 0; g < F; ++g) {
        if (do_e) acc += partE[(size_t)g * ncell + c];
        if (do_p) accp += partP[(size_t)g * ncell + c];
      }
      e += (I.dip_pf * mu) * (mi * acc);
      ep += (I.dip_pf * mu) * (mi * accp);
      // the shadow of the raw correlations, when this pass has both components
      if (I.C64 && F > 0 && do_e && (do_p || !I.use_perp)) {
        I.C64[c] = acc;
        if (do_p) I.C64[(size_t)ncell + c] = accp;
      }
    }
    if (I.has_zeeman) {
      double Bx, By, f, fp;
      field_at(I, c, Bx, By);
      zeeman_factors(I, c, site, Bx, By, f, fp);
      e += mi * f;
      ep += mi * fp;
    }
    double xe, xp;
    exchange_fields(I, y, x, mi, site, xe, xp);
    e += xe;
    ep += xp;
  }
  if (I.m_ref && F > 0 && do_e && (do_p || !I.use_perp)) I.m_ref[c] = I.m[c];
  if (do_e && I.E64) {
    I.E64[c] = e;
  } else if (do_e) {
    const float h = (float)e;
    I.E[c] = h;
    I.Ec[c] = (float)((double)h - e);
  }
  if (do_p) {
    const float h = (float)ep;
    I.Ep[c] = h;
    I.Epc[c] = (float)((double)h - ep);
  }
}

__global__ void k_recompute_finish(hs_instance_t I, int F, const double *partE, const double *partP, int which) {
  rec_finish(I, F, partE, partP, which, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- refresh from the FP64 shadow (hs_refresh_fields) -------------------------------------------
// The raw correlations C_i = sum_j K_n(i)[j - i] (m w)_j are linear in the sources, so the shadow of
// the last full recompute plus the change of the sources since then, (m_j - m_ref_j) w_j over their
// windows, is the recompute of the current spins (up to FP64 rounding). The changed cells are
// listed in cell order (row counts, then an ordered compaction), so the additions to each
// target run in a fixed order and the result does not depend on scheduling.
constexpr int DIFF_THREADS = 256;

// One workgroup per grid row: the number of cells of the row whose spin differs from m_ref.
__global__ void __launch_bounds__(DIFF_THREADS) k_diff_count(hs_instance_t I, int32_t *row_count) {
  __shared__ int s[DIFF_THREADS / 64];
  const int y = blockIdx.x;
  int v = 0;
  for (int x = threadIdx.x; x < I.nx; x += DIFF_THREADS) v += I.m[y * I.nx + x] != I.m_ref[y * I.nx + x];
  v = hs_wave_sum(v);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < DIFF_THREADS / 64; ++w) t += s[w];
    row_count[y] = t;
  }
}

// Row y writes its changed cells after those of the rows before it, in cell order, with their
// source change (m - m_ref) w, takes m_ref := m for them and records where it starts (row_off[y]);
// the last row also writes the end (row_off[ny]).
__global__ void __launch_bounds__(DIFF_THREADS) k_diff_list(hs_instance_t I, const int32_t *row_count,
                                                            int32_t *row_off, int32_t *list_cell, double *list_w) {
  __shared__ int s[DIFF_THREADS / 64];
  __shared__ int s_off;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, y = blockIdx.x;
  int before = 0;
  for (int r = tid; r < y; r += DIFF_THREADS) before += row_count[r];
  before = hs_wave_sum(before);
  if (lane == 0) s[wv] = before;
  __syncthreads();
  if (tid == 0) {
    int t = 0;
    for (int w = 0; w < DIFF_THREADS / 64; ++w) t += s[w];
    s_off = t;
    row_off[y] = t;
    if (y == I.ny - 1) row_off[I.ny] = t + row_count[y];
  }
  for (int x0 = 0; x0 < I.nx; x0 += DIFF_THREADS) {
    __syncthreads();  // s and s_off of the previous chunk are read by every thread
    const int x = x0 + tid, c = y * I.nx + x;
    const bool ch = x < I.nx && I.m[c] != I.m_ref[c];
    const unsigned long long ball = __ballot(ch);
    if (lane == 0) s[wv] = __popcll(ball);
    __syncthreads();
    int off = s_off, total = 0;
    for (int w = 0; w < DIFF_THREADS / 64; ++w) {
      off += w < wv ? s[w] : 0;
      total += s[w];
    }
    if (ch) {
      const int pos = off + __popcll(ball & ((1ull << lane) - 1ull));
      const int m = I.m[c];
      const double w = I.pair_src_unit ? 1.0 : I.pw64[c];
      list_cell[pos] = (y << 16) | x;  // packed (y, x): no division in the additions
      list_w[pos] = (double)(m - I.m_ref[c]) * w;
      I.m_ref[c] = (int8_t)m;
    }
    __syncthreads();
    if (tid == 0) s_off += total;
  }
}

// One thread per target cell (32 x 8 tiles). The listed sources in the rows the tile's window
// reaches (a contiguous part of the list, from row_off) are staged in LDS chunk by chunk and add
// K_n(target)[source - target] * change to the shadow (and K_perpself for E_perp), in list order.
// A source outside a target's window adds an exact zero (its load is masked), so the loop has no
// branch and its kernel loads can be issued several sources ahead.
__global__ void __launch_bounds__(256) k_shadow_add(hs_instance_t I, const int32_t *row_off, const int32_t *list_cell,
                                                    const double *list_w, int do_p) {
  __shared__ int s_c[256];
  __shared__ double s_w[256];
  const int ntx = (I.nx + 31) / 32;
  const int x0 = (blockIdx.x % ntx) * 32, y0 = (blockIdx.x / ntx) * 8;
  const int x = x0 + (threadIdx.x & 31), y = y0 + (threadIdx.x >> 5);
  const bool inside = x < I.nx && y < I.ny;
  const int c = inside ? y * I.nx + x : 0;
  const int n = inside ? I.sub_of[hs_site(I, y, x)] : -1;
  const int lo = row_off[max(0, y0 - I.kwy)], hi = row_off[min(I.ny, y0 + 8 + I.kwy)];
  const int KW = 2 * I.kwx + 1;
  const size_t KS = (size_t)KW * (2 * I.kwy + 1);
  const double *Kn = I.K64 + (size_t)(n >= 0 ? n : 0) * KS;
  const double *Kp = I.Kps64 ? I.Kps64 + (size_t)(n >= 0 ? n : 0) * KS : nullptr;
  double acc = 0.0, accp = 0.0;
  for (int base = lo; base < hi; base += 256) {
    __syncthreads();
    if (base + (int)threadIdx.x < hi) {
      s_c[threadIdx.x] = list_cell[base + threadIdx.x];
      s_w[threadIdx.x] = list_w[base + threadIdx.x];
    }
    __syncthreads();
    const int cnt = min(256, hi - base);
    if (n < 0) continue;
#pragma unroll 4
    for (int r = 0; r < cnt; ++r) {
      const int v = s_c[r];
      const int dy = (v >> 16) - y, dx = (v & 0xffff) - x;
      const bool in = dy >= -I.kwy && dy <= I.kwy && dx >= -I.kwx && dx <= I.kwx;
      const int o = in ? (I.kwy + dy) * KW + (I.kwx + dx) : 0;
      const double w = in ? s_w[r] : 0.0;
      acc = fma(Kn[o], w, acc);
      if (do_p) accp = fma(Kp[o], w, accp);
    }
  }
  if (n < 0 || hi == lo) return;
  I.C64[c] += acc;
  if (do_p) I.C64[(size_t)I.nx * I.ny + c] += accp;
}

extern "C" int64_t hs_refresh_scratch_bytes(const hs_instance_t *inst) {
  if (!inst || inst->nx <= 0 || inst->ny <= 0) return -1;
  const int64_t ncell = (int64_t)inst->nx * inst->ny;
  return 8 * ncell + 4 * ncell + 4 * (int64_t)inst->ny + 4 * ((int64_t)inst->ny + 1);
}

extern "C" int hs_refresh_fields(const hs_instance_t *inst, int32_t which, void *scratch, int64_t scratch_bytes,
                                 void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  if (!inst->C64 || !inst->m_ref) return hs_set_error(HS_ERR_ARG, "no FP64 shadow (C64, m_ref) to refresh from");
  if (inst->nx > 32767 || inst->ny > 32767) return hs_set_error(HS_ERR_ARG, "refresh of a grid side above 32767");
  const bool do_p = (which & 2) && inst->use_perp;
  if (!(which & 1) || (inst->use_perp && !do_p))
    return hs_set_error(HS_ERR_ARG, "a refresh covers every field component the shadow holds (which = %s)",
                        inst->use_perp ? "3" : "1 or 3");
  hipStream_t s = (hipStream_t)stream;
  const int ncell = inst->nx * inst->ny;
  if (inst->dip_pf != 0.0) {
    const int64_t need = hs_refresh_scratch_bytes(inst);
    if (!scratch || scratch_bytes < need || ((uintptr_t)scratch & 7))
      return hs_set_error(HS_ERR_ARG, "refresh scratch of %lld bytes (8-byte aligned) is below the %lld needed",
                          (long long)scratch_bytes, (long long)need);
    if (do_p && !inst->Kps64) return hs_set_error(HS_ERR_ARG, "E_perp refresh without K_perpself");
    char *p = (char *)scratch;
    double *list_w = (double *)p;
    p += 8 * (size_t)ncell;
    int32_t *list_cell = (int32_t *)p;
    p += 4 * (size_t)ncell;
    int32_t *row_count = (int32_t *)p;
    p += 4 * (size_t)inst->ny;
    int32_t *row_off = (int32_t *)p;
    k_diff_count<<<inst->ny, DIFF_THREADS, 0, s>>>(*inst, row_count);
    HS_LAUNCH_CHECK("k_diff_count");
    k_diff_list<<<inst->ny, DIFF_THREADS, 0, s>>>(*inst, row_count, row_off, list_cell, list_w);
    HS_LAUNCH_CHECK("k_diff_list");
    const unsigned tiles = (unsigned)(((inst->nx + 31) / 32) * ((inst->ny + 7) / 8));
    k_shadow_add<<<tiles, 256, 0, s>>>(*inst, row_off, list_cell, list_w, do_p ? 1 : 0);
    HS_LAUNCH_CHECK("k_shadow_add");
  }
  const double *pe = inst->C64, *pp = do_p ? inst->C64 + (size_t)ncell : nullptr;
  k_recompute_finish<<<(ncell + 255) / 256, 256, 0, s>>>(*inst, inst->dip_pf != 0.0 ? 1 : 0, pe, pp, which);
  HS_LAUNCH_CHECK("k_recompute_finish");
  return HS_OK;
}

// Small grids (at most SMALL_REC_CELLS cells): one thread per target cell sums
// K[source - target] * (m mu)[source] over the stored window directly (the tiled kernel's 32x64
// tiles would be mostly padding there). The workgroup (256 threads) first stages (m mu) of the
// whole grid in LDS, so the inner loop reads only the kernel entry (coalesced over a row of
// targets). Only occupied targets on a kernel's sublattice are computed, as the finish pass reads
// no others.
__device__ __forceinline__ void rec_small(const hs_instance_t &I, const double *Kbase, double *part, int c) {
  __shared__ double s_v[4096];
  const int nx = I.nx, ny = I.ny, ncell = nx * ny;
  for (int s = threadIdx.x; s < ncell; s += blockDim.x)
    s_v[s] = I.pair_src_unit ? (double)I.m[s] : (double)I.m[s] * I.pw64[s];
  __syncthreads();
  if (c >= ncell) return;
  const int y = c / nx, x = c % nx, n = I.sub_of[hs_site(I, y, x)];
  double acc = 0.0;
  if (n >= 0 && I.m[c] != 0) {
    const int KW = 2 * I.kwx + 1;
    const double *Kn = Kbase + (size_t)n * KW * (2 * I.kwy + 1);
    const int y0 = max(0, y - I.kwy), y1 = min(ny - 1, y + I.kwy);
    const int x0 = max(0, x - I.kwx), x1 = min(nx - 1, x + I.kwx);
    for (int ys = y0; ys <= y1; ++ys) {
      const double *krow = Kn + (size_t)(ys - y + I.kwy) * KW + (I.kwx - x);
      const double *vrow = s_v + (size_t)ys * nx;
      for (int xs = x0; xs <= x1; ++xs) {
        const double v = vrow[xs];
        if (v != 0.0) acc = fma(krow[xs], v, acc);
      }
    }
  }
  part[c] = acc;
}

__global__ void k_recompute_small(hs_instance_t I, const double *Kbase, double *part) {
  rec_small(I, Kbase, part, blockIdx.x * blockDim.x + threadIdx.x);
}

// Small grids whose kernels (every sublattice's window) fit in SMALL_LDS_K doubles: one
// 1024-thread workgroup per grid stages (m w) and the kernels in LDS once, so the inner loop reads
// LDS only (rec_small re-reads every kernel entry from L2 for every target). Same arithmetic and
// order as rec_small.
constexpr int SMALL_LDS_K = 10240;
constexpr int SMALL_LDS_THREADS = 1024;

__device__ __forceinline__ void rec_small_lds(const hs_instance_t &I, const double *Kbase, double *part) {
  __shared__ double s_v[4096];
  __shared__ double s_k[SMALL_LDS_K];
  const int nx = I.nx, ny = I.ny, ncell = nx * ny;
  const int KW = 2 * I.kwx + 1, KS = KW * (2 * I.kwy + 1);
  for (int s = threadIdx.x; s < ncell; s += blockDim.x)
    s_v[s] = I.pair_src_unit ? (double)I.m[s] : (double)I.m[s] * I.pw64[s];
  for (int e = threadIdx.x; e < KS * I.n_sub; e += blockDim.x) s_k[e] = Kbase[e];
  __syncthreads();
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int y = c / nx, x = c % nx, n = I.sub_of[hs_site(I, y, x)];
    double acc = 0.0;
    if (n >= 0 && I.m[c] != 0) {
      const double *Kn = s_k + n * KS;
      const int y0 = max(0, y - I.kwy), y1 = min(ny - 1, y + I.kwy);
      const int x0 = max(0, x - I.kwx), x1 = min(nx - 1, x + I.kwx);
      for (int ys = y0; ys <= y1; ++ys) {
        const double *krow = Kn + (ys - y + I.kwy) * KW + (I.kwx - x);
        const double *vrow = s_v + ys * nx;
        for (int xs = x0; xs <= x1; ++xs) {
          const double v = vrow[xs];
          if (v != 0.0) acc = fma(krow[xs], v, acc);
        }
      }
    }
    part[c] = acc;
  }
}

__global__ void __launch_bounds__(1024) k_recompute_small_lds_many(const RecJob *__restrict__ jobs,
                                                                   const int32_t *__restrict__ job_of_block) {
  const RecJob &J = jobs[job_of_block[blockIdx.x]];
  rec_small_lds(J.I, J.K, J.part);
}

static bool small_lds_rec(const hs_instance_t *I) {
  return I->nx * I->ny <= 4096 && (int64_t)(2 * I->kwx + 1) * (2 * I->kwy + 1) * I->n_sub <= SMALL_LDS_K;
}

__global__ void __launch_bounds__(SMALL_LDS_THREADS) k_recompute_small_lds(hs_instance_t I, const double *Kbase,
                                                                          double *part) {
  rec_small_lds(I, Kbase, part);
}

struct FinJob {
  hs_instance_t I;
  const double *partE, *partP;
  int32_t F, which, first;
};

__global__ void k_recompute_finish_many(const FinJob *__restrict__ jobs, const int32_t *__restrict__ block_job) {
  const FinJob &J = jobs[block_job[blockIdx.x]];
  rec_finish(J.I, J.F, J.partE, J.partP, J.which, (blockIdx.x - J.first) * blockDim.x + threadIdx.x);
}

constexpr int SMALL_REC_CELLS = 4096;  // grids up to this many cells use rec_small

static bool small_rec(const hs_instance_t *I) { return I->nx * I->ny <= SMALL_REC_CELLS; }

static RecPlan make_plan(const hs_instance_t *I) {
  RecPlan pl{};
  pl.Ay = (I->ny + I->uy - 1) / I->uy;
  pl.Bx = (I->nx + I->ux - 1) / I->ux;
  pl.tA = (pl.Ay + RA - 1) / RA;
  pl.tB = (pl.Bx + RB - 1) / RB;
  const int tiles = pl.tA * pl.tB * (I->n_sub > 0 ? I->n_sub : 1);
  int F = (768 + tiles - 1) / tiles;
  pl.F = F < 1 ? 1 : (F > 64 ? 64 : F);
  if (small_rec(I)) pl.F = 1;
  for (int s = 0; s < I->ux * I->uy; ++s) {
    const int n = I->sub_of[s];
    if (n >= 0 && n < HS_MAX_UC) {
      pl.sites_y[n] = s / I->ux;
      pl.sites_x[n] = s % I->ux;
    }
  }
  return pl;
}

extern "C" int64_t hs_recompute_scratch_bytes(const hs_instance_t *inst) {
  if (!inst || inst->nx <= 0 || inst->ny <= 0 || inst->ux <= 0 || inst->uy <= 0) return -1;
  const RecPlan pl = make_plan(inst);
  return (int64_t)2 * pl.F * inst->nx * inst->ny * (int64_t)sizeof(double);
}

extern "C" int hs_recompute_fields(const hs_instance_t *inst, int32_t which, void *scratch, int64_t scratch_bytes,
                                   void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const RecPlan pl = make_plan(inst);
  const int ncell = inst->nx * inst->ny;
  const bool dip = inst->dip_pf != 0.0;
  const bool do_p = (which & 2) && inst->use_perp;
  double *partE = nullptr, *partP = nullptr;
  if (dip) {
    const int64_t need = hs_recompute_scratch_bytes(inst);
    if (!scratch || scratch_bytes < need)
      return hs_set_error(HS_ERR_ARG, "recompute scratch of %lld bytes is below the %lld needed", (long long)scratch_bytes,
                          (long long)need);
    for (int n = 0; n < inst->n_sub; ++n) {
      bool found = false;
      for (int q = 0; q < inst->ux * inst->uy; ++q) found |= inst->sub_of[q] == n;
      if (!found) return hs_set_error(HS_ERR_ARG, "kernel %d has no unit-cell site", n);
    }
    partE = (double *)scratch;
    partP = partE + (size_t)pl.F * ncell;
    const dim3 grid(pl.tA * pl.tB, inst->n_sub, pl.F);
    const unsigned sb = (unsigned)((ncell + 255) / 256);
    if (which & 1) {
      if (small_lds_rec(inst)) k_recompute_small_lds<<<1, SMALL_LDS_THREADS, 0, s>>>(*inst, inst->K64, partE);
      else if (small_rec(inst)) k_recompute_small<<<sb, 256, 0, s>>>(*inst, inst->K64, partE);
      else k_recompute_dip<<<grid, 256, 0, s>>>(*inst, pl, inst->K64, partE);
      HS_LAUNCH_CHECK("k_recompute_dip");
    }
    if (do_p) {
      if (small_lds_rec(inst)) k_recompute_small_lds<<<1, SMALL_LDS_THREADS, 0, s>>>(*inst, inst->Kps64, partP);
      else if (small_rec(inst)) k_recompute_small<<<sb, 256, 0, s>>>(*inst, inst->Kps64, partP);
      else k_recompute_dip<<<grid, 256, 0, s>>>(*inst, pl, inst->Kps64, partP);
      HS_LAUNCH_CHECK("k_recompute_dip");
    }
  }
  k_recompute_finish<<<(ncell + 255) / 256, 256, 0, s>>>(*inst, dip ? pl.F : 0, partE, partP, which);
  HS_LAUNCH_CHECK("k_recompute_finish");
  return HS_OK;
}

namespace {

struct ManyLayout {
  size_t dip_off, fin_off, dmap_off, fmap_off, lmap_off, total;
  int ndip, dip_blocks, fin_blocks, nlds;
};

size_t al256c(size_t v) { return (v + 255) & ~(size_t)255; }

ManyLayout many_layout(const hs_instance_t *const *insts, const int32_t *which, int n) {
  ManyLayout L{};
  for (int i = 0; i < n; ++i) {
    const hs_instance_t &I = *insts[i];
    const RecPlan pl = make_plan(&I);
    if (I.dip_pf != 0.0) {
      const bool lds = small_lds_rec(&I);
      const int per = lds ? 0 : small_rec(&I) ? (I.nx * I.ny + 255) / 256 : pl.tA * pl.tB * I.n_sub * pl.F;
      for (int w = 1; w <= 2; w <<= 1) {
        if (!(which[i] & w) || (w == 2 && !I.use_perp)) continue;
        ++L.ndip;
        L.dip_blocks += per;
        L.nlds += lds ? 1 : 0;
      }
    }
    L.fin_blocks += (I.nx * I.ny + 255) / 256;
  }
  L.dip_off = 0;
  L.fin_off = al256c(sizeof(RecJob) * (size_t)(L.ndip > 0 ? L.ndip : 1));
  L.dmap_off = L.fin_off + al256c(sizeof(FinJob) * (size_t)n);
  L.fmap_off = L.dmap_off + al256c(4 * (size_t)(L.dip_blocks > 0 ? L.dip_blocks : 1));
  L.lmap_off = L.fmap_off + al256c(4 * (size_t)L.fin_blocks);
  L.total = L.lmap_off + al256c(4 * (size_t)(L.nlds > 0 ? L.nlds : 1));
  return L;
}

}  // namespace

extern "C" int64_t hs_recompute_many_ws_bytes(const hs_instance_t *const *insts, const int32_t *which, int32_t n) {
  if (!insts || !which || n <= 0) return -1;
  for (int i = 0; i < n; ++i)
    if (!insts[i] || insts[i]->nx <= 0 || insts[i]->ny <= 0 || insts[i]->ux <= 0 || insts[i]->uy <= 0) return -1;
  return (int64_t)many_layout(insts, which, n).total;
}

extern "C" int hs_recompute_fields_many(const hs_instance_t *const *insts, const int32_t *which, void *const *scratch,
                                        const int64_t *scratch_bytes, int32_t n, void *ws, int64_t ws_bytes,
                                        void *stream) {
  hs_clear_error();
  if (!insts || !which || !scratch || !scratch_bytes || n <= 0 || !ws) return hs_set_error(HS_ERR_ARG, "missing arrays");
  for (int i = 0; i < n; ++i) {
    int rc = hs_check_instance(insts[i]);
    if (rc) return rc;
    const hs_instance_t &I = *insts[i];
    if (I.dip_pf != 0.0) {
      const int64_t need = hs_recompute_scratch_bytes(&I);
      if (!scratch[i] || scratch_bytes[i] < need)
        return hs_set_error(HS_ERR_ARG, "instance %d: recompute scratch of %lld bytes is below the %lld needed", i,
                            (long long)scratch_bytes[i], (long long)need);
      for (int k = 0; k < I.n_sub; ++k) {
        bool found = false;
        for (int q = 0; q < I.ux * I.uy; ++q) found |= I.sub_of[q] == k;
        if (!found) return hs_set_error(HS_ERR_ARG, "instance %d: kernel %d has no unit-cell site", i, k);
      }
    }
  }
  const ManyLayout L = many_layout(insts, which, n);
  if (ws_bytes < (int64_t)L.total)
    return hs_set_error(HS_ERR_ARG, "workspace %lld < %lld bytes", (long long)ws_bytes, (long long)L.total);
  std::string host(L.total, '\0');
  RecJob *dj = (RecJob *)&host[L.dip_off];
  FinJob *fj = (FinJob *)&host[L.fin_off];
  int32_t *dmap = (int32_t *)&host[L.dmap_off];
  int32_t *fmap = (int32_t *)&host[L.fmap_off];
  int32_t *lmap = (int32_t *)&host[L.lmap_off];
  // block numbering: every small job's blocks first, then the tiled jobs' blocks (jobs whose
  // kernels fit in LDS take one workgroup each, listed in lmap)
  int dbs = 0, nl = 0;
  for (int i = 0; i < n; ++i) {
    const hs_instance_t &I = *insts[i];
    if (I.dip_pf == 0.0 || !small_rec(&I) || small_lds_rec(&I)) continue;
    const int per = (I.nx * I.ny + 255) / 256;
    dbs += ((which[i] & 1) ? per : 0) + (((which[i] & 2) && I.use_perp) ? per : 0);
  }
  int nd = 0, db_small = 0, db = dbs, fb = 0;
  for (int i = 0; i < n; ++i) {
    const hs_instance_t &I = *insts[i];
    const RecPlan pl = make_plan(&I);
    const int ncell = I.nx * I.ny;
    const bool dip = I.dip_pf != 0.0;
    double *partE = dip ? (double *)scratch[i] : nullptr;
    double *partP = dip ? partE + (size_t)pl.F * ncell : nullptr;
    if (dip) {
      const bool lds = small_lds_rec(&I);
      const int per = small_rec(&I) ? (ncell + 255) / 256 : pl.tA * pl.tB * I.n_sub * pl.F;
      for (int w = 1; w <= 2; w <<= 1) {
        if (!(which[i] & w) || (w == 2 && !I.use_perp)) continue;
        RecJob &J = dj[nd];
        J.I = I;
        J.plan = pl;
        J.K = w == 1 ? I.K64 : I.Kps64;
        J.part = w == 1 ? partE : partP;
        J.small = small_rec(&I) ? 1 : 0;
        if (lds) {
          J.first = 0;
          lmap[nl++] = nd;
        } else {
          int &next = J.small ? db_small : db;
          J.first = next;
          for (int b = 0; b < per; ++b) dmap[next + b] = nd;
          next += per;
        }
        ++nd;
      }
    }
    FinJob &F = fj[i];
    F.I = I;
    F.partE = partE;
    F.partP = partP;
    F.F = dip ? pl.F : 0;
    F.which = which[i];
    F.first = fb;
    const int nb = (ncell + 255) / 256;
    for (int b = 0; b < nb; ++b) fmap[fb + b] = i;
    fb += nb;
  }
  hipStream_t s = (hipStream_t)stream;
  char *base = (char *)ws;
  if (int rc = hs_upload_async(base, host.data(), L.total, s)) return rc;
  if (nl > 0) {
    k_recompute_small_lds_many<<<nl, SMALL_LDS_THREADS, 0, s>>>((const RecJob *)(base + L.dip_off),
                                                                (const int32_t *)(base + L.lmap_off));
    HS_LAUNCH_CHECK("k_recompute_small_lds_many");
  }
  if (dbs > 0) {
    k_recompute_small_many<<<dbs, 256, 0, s>>>((const RecJob *)(base + L.dip_off), (const int32_t *)(base + L.dmap_off));
    HS_LAUNCH_CHECK("k_recompute_small_many");
  }
  if (db > dbs) {
    k_recompute_dip_many<<<db - dbs, 256, 0, s>>>((const RecJob *)(base + L.dip_off),
                                                  (const int32_t *)(base + L.dmap_off), dbs);
    HS_LAUNCH_CHECK("k_recompute_dip_many");
  }
  k_recompute_finish_many<<<fb, 256, 0, s>>>((const FinJob *)(base + L.fin_off), (const int32_t *)(base + L.fmap_off));
  HS_LAUNCH_CHECK("k_recompute_finish_many");
  return HS_OK;
}

// ---- per-energy fields -------------------------------------------------------------------------
__global__ void k_field_components(hs_instance_t I, double *out) {
  const int ncell = I.nx * I.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int y = c / I.nx, x = c % I.nx, site = hs_site(I, y, x);
  const double mi = (double)I.m[c];
  double zf = 0.0, zp = 0.0, xe = 0.0, xp = 0.0;
  if (I.has_zeeman && mi != 0.0) {
    double Bx, By, f, fp;
    field_at(I, c, Bx, By);
    zeeman_factors(I, c, site, Bx, By, f, fp);
    zf = mi * f;
    zp = mi * fp;
  }
  exchange_fields(I, y, x, mi, site, xe, xp);
  out[0 * (size_t)ncell + c] = zf;
  out[1 * (size_t)ncell + c] = I.use_perp ? zp : 0.0;
  out[2 * (size_t)ncell + c] = xe;
  out[3 * (size_t)ncell + c] = I.use_perp ? xp : 0.0;
  out[4 * (size_t)ncell + c] = I.E64 ? I.E64[c] : (double)I.E[c] - (double)I.Ec[c];
  out[5 * (size_t)ncell + c] = I.use_perp ? (double)I.Ep[c] - (double)I.Epc[c] : 0.0;
}

extern "C" int hs_field_components(const hs_instance_t *inst, double *out, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  if (!out) return hs_set_error(HS_ERR_ARG, "out is NULL");
  const int ncell = inst->nx * inst->ny;
  k_field_components<<<(ncell + 255) / 256, 256, 0, (hipStream_t)stream>>>(*inst, out);
  HS_LAUNCH_CHECK("k_field_components");
  return HS_OK;
}

// ---- total energy: 0.5*sum(E) over pair energies + sum(E_zeeman) = 0.5*sum(E) + 0.5*sum(E_z) ----
__global__ void __launch_bounds__(1024) k_energy_total(hs_instance_t I, double *out) {
  __shared__ double s[16];
  const int ncell = I.nx * I.ny;
  double acc = 0.0;
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const double mi = (double)I.m[c];
    double v = I.E64 ? I.E64[c] : (double)I.E[c] - (double)I.Ec[c];
    if (I.has_zeeman && mi != 0.0) {
      double Bx, By, f, fp;
      field_at(I, c, Bx, By);
      zeeman_factors(I, c, hs_site(I, c / I.nx, c % I.nx), Bx, By, f, fp);
      v += mi * f;
    }
    acc += 0.5 * v;
  }
  acc = hs_wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s[w];
    out[0] = t;
  }
}

extern "C" int hs_energy_total(const hs_instance_t *inst, double *out_host, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  if (!out_host) return hs_set_error(HS_ERR_ARG, "out is NULL");
  double *d = nullptr;
  HS_HIP_TRY(hipMallocAsync((void **)&d, sizeof(double), (hipStream_t)stream));
  k_energy_total<<<1, 1024, 0, (hipStream_t)stream>>>(*inst, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out_host, d, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream);
  hipError_t e2 = hipFreeAsync(d, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hs_set_error(HS_ERR_HIP, "hs_energy_total: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return hs_set_error(HS_ERR_HIP, "hs_energy_total free: %s", hipGetErrorString(e2));
  return HS_OK;
}

// ---- Zeeman field step (energies.py:101-118): E += m*(f_new - f_old) ----------------------------
__global__ void k_zeeman_swap(hs_instance_t I, double Bx0, double By0, const float *bx0, const float *by0) {
  const int ncell = I.nx * I.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const double mi = (double)I.m[c];
  if (mi == 0.0) return;
  const int site = hs_site(I, c / I.nx, c % I.nx);
  double Bx, By, f1, fp1, f0, fp0;
  field_at(I, c, Bx, By);
  zeeman_factors(I, c, site, Bx, By, f1, fp1);
  zeeman_factors(I, c, site, bx0 ? (double)bx0[c] : Bx0, by0 ? (double)by0[c] : By0, f0, fp0);
  if (I.E64) {
    I.E64[c] += mi * (f1 - f0);
  } else {
    float s = I.E[c], cc = I.Ec[c];
    hs_kahan_add(s, cc, (float)(mi * (f1 - f0)));
    I.E[c] = s;
    I.Ec[c] = cc;
  }
  if (I.use_perp) {
    float sp = I.Ep[c], cp = I.Epc[c];
    hs_kahan_add(sp, cp, (float)(mi * (fp1 - fp0)));
    I.Ep[c] = sp;
    I.Epc[c] = cp;
  }
}

extern "C" int hs_zeeman_swap(const hs_instance_t *inst, double Bx0, double By0, const float *bx_old,
                              const float *by_old, void *stream) {
  hs_clear_error();
  int rc = hs_check_instance(inst);
  if (rc) return rc;
  const int ncell = inst->nx * inst->ny;
  k_zeeman_swap<<<(ncell + 255) / 256, 256, 0, (hipStream_t)stream>>>(*inst, Bx0, By0, bx_old, by_old);
  HS_LAUNCH_CHECK("k_zeeman_swap");
  return HS_OK;
}
