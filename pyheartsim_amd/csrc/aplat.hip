// Activation markers and local activation times of sampled action potentials on the device
// (legacy/calculate_local_activation_time.py:18-66 execute_on_action_potential and :67-136
// find_action_potential_marker).
//
// Per signal s (one column of the T x n sample array), the reference takes the sorted, unique
// boundary points of the runs of samples above the threshold (each run's first and last
// index; a one-sample run gives one point), pairs them in order (an odd last point is
// dropped), and marks in each pair [a, b] the first index of the largest forward difference
// s[j+1] - s[j], j in [a, b - 1], deleting markers whose difference is <= 0.  A one-sample run
// shifts the pairing of every later point, so a pair may span samples below the threshold;
// that is reproduced here.  NumPy's max/argmax propagate NaN (the first NaN difference wins
// and is kept), and so does this scan.
//
// One thread per signal streams its column once per pass (the loads of a wavefront are
// consecutive signals of one sample row, so they coalesce):
//   pass 1: marker count, first and last marker, whether any sample is above the threshold;
//   reduce: mean cycle length CL = sum(last - first) / sum(count - 1) over the signals (the
//           cycle lengths of a signal telescope), 180 if there is none; the window of interest
//           [t_start, t_start + rint(CL * fraction)]; exclusive offsets of the markers and of
//           the cycle lengths;
//   pass 2: markers and cycle lengths at their offsets (optional), and the first marker in
//           the window of interest;
//   shift:  lat - min(lat) + 1 over the signals that have one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "phs_internal.h"

namespace {

constexpr int kApThreads = 256;
constexpr int kApReduceThreads = 1024;

// Calls on_marker(index) for each kept marker of the column s (stride `stride` between
// samples), in time order; returns whether any sample exceeds the threshold.
template <typename F>
__device__ __forceinline__ bool scan_markers(const float* __restrict__ s, int64_t T,
                                             int64_t stride, float thr, F&& on_marker) {
  bool open = false, any = false, prev_above = false;
  int64_t start = 0, best_i = -1, nan_i = -1;
  float best = -INFINITY, prev = 0.0f;
  float cur = T > 0 ? s[0] : 0.0f;
  for (int64_t t = 0; t < T; ++t) {
    const bool has_next = t + 1 < T;
    const float next = has_next ? s[(t + 1) * stride] : 0.0f;
    const bool above = cur > thr;
    const bool next_above = has_next && next > thr;
    any |= above;
    if (open && t > start) {  // forward difference d[t - 1] = s[t] - s[t - 1]
      const float d = cur - prev;
      if (d != d) {
        if (nan_i < 0) nan_i = t - 1;
      } else if (d > best) {
        best = d;
        best_i = t - 1;
      }
    }
    // at most one boundary point per index: a run's start, its end, or both for a one-sample run
    if (above && (!prev_above || !next_above)) {
      if (!open) {
        open = true;
        start = t;
        best = -INFINITY;
        best_i = -1;
        nan_i = -1;
      } else {
        open = false;
        if (nan_i >= 0) on_marker(nan_i);
        else if (best > 0.0f) on_marker(best_i);
      }
    }
    prev = cur;
    prev_above = above;
    cur = next;
  }
  return any;
}

struct ApWork {
  int32_t* count;   // markers per signal
  int32_t* first;
  int32_t* last;
  int64_t* moff;    // exclusive offsets of the markers
  int64_t* coff;    // exclusive offsets of the cycle lengths
  double* stats;    // [0] CL, [1] woi start, [2] woi end, [3] markers, [4] cycle lengths
  double* scratch;  // per-signal latency column of the marker pass
  uint8_t* any;     // some sample above the threshold
};

__global__ void __launch_bounds__(kApThreads) ap_count_k(const float* __restrict__ ap, int64_t T,
                                                          int64_t n, int64_t stride, float thr,
                                                          ApWork w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t k = 0, f = -1, l = -1;
  const bool any = scan_markers(ap + i, T, stride, thr, [&](int64_t m) {
    if (k == 0) f = (int32_t)m;
    l = (int32_t)m;
    ++k;
  });
  w.count[i] = k;
  w.first[i] = f;
  w.last[i] = l;
  w.any[i] = any ? 1 : 0;
}

// One workgroup: totals, exclusive offsets (each thread owns a contiguous range of signals),
// the mean cycle length and the window of interest.
__global__ void __launch_bounds__(kApReduceThreads) ap_reduce_k(int64_t n, double t_start,
                                                                double fraction, ApWork w) {
  __shared__ int64_t sm[kApReduceThreads], sc[kApReduceThreads], ss[kApReduceThreads];
  const int tid = threadIdx.x;
  const int64_t per = (n + kApReduceThreads - 1) / kApReduceThreads;
  const int64_t lo = per * tid < n ? per * tid : n;
  const int64_t hi = lo + per < n ? lo + per : n;
  int64_t m = 0, c = 0, span = 0;
  for (int64_t i = lo; i < hi; ++i) {
    const int32_t k = w.count[i];
    m += k;
    if (k >= 1) c += k - 1;
    if (k >= 2) span += (int64_t)w.last[i] - w.first[i];
  }
  sm[tid] = m;
  sc[tid] = c;
  ss[tid] = span;
  __syncthreads();
  if (tid == 0) {  // exclusive scan of the per-thread totals, in order
    int64_t am = 0, ac = 0, as = 0;
    for (int j = 0; j < kApReduceThreads; ++j) {
      const int64_t xm = sm[j], xc = sc[j], xs = ss[j];
      sm[j] = am;
      sc[j] = ac;
      am += xm;
      ac += xc;
      as += xs;
    }
    const double cl = ac > 0 ? (double)as / (double)ac : 180.0;
    w.stats[0] = cl;
    w.stats[1] = t_start;
    w.stats[2] = t_start + rint(cl * fraction);
    w.stats[3] = (double)am;
    w.stats[4] = (double)ac;
  }
  __syncthreads();
  int64_t om = sm[tid], oc = sc[tid];
  for (int64_t i = lo; i < hi; ++i) {
    const int32_t k = w.count[i];
    w.moff[i] = om;
    w.coff[i] = oc;
    om += k;
    if (k >= 1) oc += k - 1;
  }
}

__global__ void __launch_bounds__(kApThreads) ap_lat_k(const float* __restrict__ ap, int64_t T,
                                                        int64_t n, int64_t stride, float thr,
                                                        ApWork w, int32_t* __restrict__ markers,
                                                        double* __restrict__ cl,
                                                        double* __restrict__ lat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double w0 = w.stats[1], w1 = w.stats[2];
  int64_t j = 0, prev = -1;
  double first_in = NAN;
  int64_t mo = markers ? w.moff[i] : 0, co = cl ? w.coff[i] : 0;
  scan_markers(ap + i, T, stride, thr, [&](int64_t m) {
    if (markers) markers[mo + j] = (int32_t)m;
    if (cl && j > 0) cl[co + j - 1] = (double)(m - prev);
    if (first_in != first_in && (double)m >= w0 && (double)m <= w1) first_in = (double)m;
    prev = m;
    ++j;
  });
  lat[i] = first_in;
}

// lat - min(lat) + 1 over the signals with a marker in the window (NaN stays NaN).
__global__ void __launch_bounds__(kApReduceThreads) ap_shift_k(int64_t n, double* __restrict__ lat) {
  __shared__ double sm[kApReduceThreads];
  const int tid = threadIdx.x;
  double m = INFINITY;
  for (int64_t i = tid; i < n; i += kApReduceThreads) {
    const double v = lat[i];
    if (v == v) m = fmin(m, v);
  }
  sm[tid] = m;
  __syncthreads();
  for (int o = kApReduceThreads / 2; o > 0; o >>= 1) {
    if (tid < o) sm[tid] = fmin(sm[tid], sm[tid + o]);
    __syncthreads();
  }
  const double lo = sm[0];
  if (lo == INFINITY) return;
  for (int64_t i = tid; i < n; i += kApReduceThreads) lat[i] = lat[i] - lo + 1.0;
}

ApWork carve(void* workspace, int64_t n) {
  ApWork w;
  char* p = (char*)workspace;
  w.stats = (double*)p;
  p += 8 * 8;
  w.moff = (int64_t*)p;
  p += 8 * n;
  w.coff = (int64_t*)p;
  p += 8 * n;
  w.count = (int32_t*)p;
  p += 4 * n;
  w.first = (int32_t*)p;
  p += 4 * n;
  w.last = (int32_t*)p;
  p += 4 * n;
  p += (8 - ((uintptr_t)p & 7)) & 7;
  w.scratch = (double*)p;
  p += 8 * n;
  w.any = (uint8_t*)p;
  return w;
}

}  // namespace

extern "C" {

int phs_ap_lat_workspace(int64_t n_signal, int64_t* bytes) {
  if (!bytes || n_signal <= 0) return phs::fail(PHS_ERR_ARG, "invalid AP-LAT workspace query");
  *bytes = 64 + 8 + 37 * n_signal;  // stats, offsets, counts, scratch, flags (+ alignment)
  return PHS_OK;
}

int phs_ap_lat(const float* ap, int64_t n_time, int64_t n_signal, int64_t stride, double threshold,
               double t_start, double cycle_length_fraction, void* workspace, double* lat,
               double* stats, uint8_t* any_above, void* stream) {
  if (!ap || !workspace || !lat || n_time <= 0 || n_signal <= 0 || stride < 1 ||
      n_time > INT32_MAX)
    return phs::fail(PHS_ERR_ARG, "invalid AP-LAT arguments");
  hipStream_t s = (hipStream_t)stream;
  const ApWork w = carve(workspace, n_signal);
  const unsigned grid = (unsigned)((n_signal + kApThreads - 1) / kApThreads);
  const float thr = (float)threshold;
  ap_count_k<<<grid, kApThreads, 0, s>>>(ap, n_time, n_signal, stride, thr, w);
  ap_reduce_k<<<1, kApReduceThreads, 0, s>>>(n_signal, t_start, cycle_length_fraction, w);
  ap_lat_k<<<grid, kApThreads, 0, s>>>(ap, n_time, n_signal, stride, thr, w, nullptr, nullptr, lat);
  ap_shift_k<<<1, kApReduceThreads, 0, s>>>(n_signal, lat);
  PHS_HIP_CHECK(hipGetLastError());
  if (stats) PHS_HIP_CHECK(hipMemcpyAsync(stats, w.stats, 5 * sizeof(double),
                                          hipMemcpyDeviceToDevice, s));
  if (any_above) PHS_HIP_CHECK(hipMemcpyAsync(any_above, w.any, (size_t)n_signal,
                                              hipMemcpyDeviceToDevice, s));
  return PHS_OK;
}

int phs_ap_markers(const float* ap, int64_t n_time, int64_t n_signal, int64_t stride,
                   double threshold, void* workspace, int32_t* markers, double* cycle_lengths,
                   void* stream) {
  if (!ap || !workspace || n_time <= 0 || n_signal <= 0 || stride < 1 || n_time > INT32_MAX)
    return phs::fail(PHS_ERR_ARG, "invalid AP-marker arguments");
  hipStream_t s = (hipStream_t)stream;
  const ApWork w = carve(workspace, n_signal);
  const unsigned grid = (unsigned)((n_signal + kApThreads - 1) / kApThreads);
  // the offsets and the window come from phs_ap_lat on the same workspace; the latency
  // column of this pass goes to the workspace's scratch
  ap_lat_k<<<grid, kApThreads, 0, s>>>(ap, n_time, n_signal, stride, (float)threshold, w, markers,
                                       cycle_lengths, w.scratch);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

}  // extern "C"
