// Readouts on the device: sums of m * w over rectangular bins of the grid, the operation behind
// the reference's magnetisation averages (m_avg_x / m_avg_y / m_avg, core.py:413-424) and its
// regional output readers (RegionalOutputReader.read_state io.py:335-380,
// OOPSquareChessOutputReader.read_state io.py:404-445), so a readout downloads a few numbers
// instead of the spin array.
//
// Bins are contiguous column ranges [x_start[b], x_start[b+1]) and row ranges [y_start[b],
// y_start[b+1]) (the reference bins with a monotone linspace, so every bin is such a range). Two
// passes, each summing in a fixed order, so a readout is the same on every call:
//   1. one thread per (weight k, row y, column bin bx) sums m[y][x] * w[k][y][x] over the bin's
//      columns in ascending x;
//   2. one thread per (weight k, bin) sums the row partials of its bin's rows in ascending y.
#include <hip/hip_runtime.h>

#include "hs_device.h"
#include "hs_internal.h"

namespace {

__global__ void k_bin_rows(hs_instance_t I, int nbx, const int32_t *__restrict__ x_start, const double *__restrict__ w,
                           int nw, double *__restrict__ rowpart) {
  const int ny = I.ny, nx = I.nx;
  const int64_t total = (int64_t)nw * ny * nbx;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int bx = (int)(i % nbx), y = (int)((i / nbx) % ny), k = (int)(i / ((int64_t)nbx * ny));
    const int8_t *mrow = I.m + (size_t)y * nx;
    const double *wrow = w ? w + ((size_t)k * ny + y) * nx : nullptr;
    double acc = 0.0;
    const int x0 = max(0, x_start[bx]), x1 = min(nx, x_start[bx + 1]);  // in bounds whatever the bins hold
    for (int x = x0; x < x1; ++x) {
      const double m = (double)mrow[x];
      acc += wrow ? m * wrow[x] : m;
    }
    rowpart[i] = acc;
  }
}

__global__ void k_bin_cols(int ny, int nbx, int nby, const int32_t *__restrict__ y_start, int nw,
                           const double *__restrict__ rowpart, double *__restrict__ out) {
  const int total = nw * nby * nbx;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int bx = i % nbx, by = (i / nbx) % nby, k = i / (nbx * nby);
    double acc = 0.0;
    const int y0 = max(0, y_start[by]), y1 = min(ny, y_start[by + 1]);
    for (int y = y0; y < y1; ++y) acc += rowpart[((size_t)k * ny + y) * nbx + bx];
    out[i] = acc;
  }
}

}  // namespace

extern "C" int64_t hs_binned_sums_scratch_bytes(const hs_instance_t *inst, int32_t nbx, int32_t nw) {
  if (!inst || nbx < 1 || nw < 1) return -1;
  return (int64_t)sizeof(double) * nw * (int64_t)inst->ny * nbx;
}

extern "C" int hs_binned_sums(const hs_instance_t *inst, int32_t nbx, int32_t nby, const int32_t *x_start,
                              const int32_t *y_start, const double *w, int32_t nw, double *scratch, int64_t scratch_bytes,
                              double *out, void *stream) {
  hs_clear_error();
  if (!inst || !inst->m || inst->nx < 1 || inst->ny < 1) return hs_set_error(HS_ERR_ARG, "bad instance");
  if (nbx < 1 || nby < 1 || nbx > inst->nx || nby > inst->ny || nw < 1 || (!w && nw != 1))
    return hs_set_error(HS_ERR_ARG, "bad bins (%d x %d) or weights (%d)", nbx, nby, nw);
  if (!x_start || !y_start || !scratch || !out) return hs_set_error(HS_ERR_ARG, "missing arrays");
  if (scratch_bytes < hs_binned_sums_scratch_bytes(inst, nbx, nw))
    return hs_set_error(HS_ERR_ARG, "scratch of %lld bytes is below the %lld needed", (long long)scratch_bytes,
                        (long long)hs_binned_sums_scratch_bytes(inst, nbx, nw));
  hipStream_t s = (hipStream_t)stream;
  const int64_t rows = (int64_t)nw * inst->ny * nbx;
  const unsigned g1 = (unsigned)((rows + 255) / 256 < 4096 ? (rows + 255) / 256 : 4096);
  k_bin_rows<<<g1, 256, 0, s>>>(*inst, nbx, x_start, w, nw, scratch);
  HS_LAUNCH_CHECK("k_bin_rows");
  const int bins = nw * nbx * nby;
  const unsigned g2 = (unsigned)((bins + 255) / 256 < 4096 ? (bins + 255) / 256 : 4096);
  k_bin_cols<<<g2, 256, 0, s>>>(inst->ny, nbx, nby, y_start, nw, scratch, out);
  HS_LAUNCH_CHECK("k_bin_cols");
  return HS_OK;
}
