// mcw_eigen.hip -- the end of an eigenvalue batch on the device.
//
// After transport (which banks fission sites in arrival order, each with the
// key history << 24 | progeny, and stores every history's three keff scores):
//   1. keff scores: each of the three per-history arrays is summed in a fixed
//      order (k_tree_sum: blocks of 1024 values folded by 256 lanes, then
//      halved pairwise in LDS, level by level), so the generation keff does
//      not depend on scheduling and equals a CPU sum taken in the same order;
//   2. the bank is put in OpenMC's order (parent history, then progeny) by a
//      radix sort of (key, slot) pairs;
//   3. OpenMCStudy::synchronizeBanks (src/userobjects/OpenMCStudy.C:454-607)
//      as data-parallel steps: site i (in sorted order) draws the i-th number
//      of the tracking stream of seed id `batch` -- reached by a skip-ahead,
//      not by walking the stream -- and is copied floor(n / total) times when
//      fewer than n were banked plus once when the draw is below
//      p = (n mod total, or n) / total (k_sync_count); an exclusive scan of the
//      copy counts gives each site's first output position; every output
//      position then finds its site by binary search over those offsets, and
//      positions past the sampled total take the last sites of the bank, as
//      the reference fills a short source (k_sync_scatter).  The next batch's
//      tracking seeds are written with the sites.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "../../include/macaw_amd.h"
#include "mcw_device.h"
#include "mcw_types.h"

namespace mcw {

// One level of the fixed-order sum: out[b] = sum of in[1024 b .. 1024 b + 1023].
extern "C" __global__ void __launch_bounds__(256) k_tree_sum(const double *in, int64_t n, double *out)
{
  __shared__ double a[256];
  const int t = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * 1024 + t;
  const double v0 = i < n ? in[i] : 0.0;
  const double v1 = i + 256 < n ? in[i + 256] : 0.0;
  const double v2 = i + 512 < n ? in[i + 512] : 0.0;
  const double v3 = i + 768 < n ? in[i + 768] : 0.0;
  a[t] = ((v0 + v1) + v2) + v3;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) a[t] = a[t] + a[t + s];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = a[0];
}

// counts[i] = base + (i-th draw of the stream from seed0 < p_sample)
extern "C" __global__ void k_sync_count(int64_t total, int64_t base, double p_sample, uint64_t seed0, int64_t *counts)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    uint64_t s = future_seed((uint64_t)i, seed0);
    counts[i] = base + (prn(s) < p_sample ? 1 : 0);
  }
}

// Output position j of the next source: the sorted site whose copies cover
// j, or, past the sampled total `fin`, sorted site max(0, total - (n - fin) +
// (j - fin)).
extern "C" __global__ void k_sync_scatter(FissionView F, const uint32_t *perm, const int64_t *counts,
                                          const int64_t *offsets, int64_t total, int64_t n, SourceBank out,
                                          uint64_t *out_key, int64_t next_batch, uint64_t master, int roundtrip)
{
  const int64_t fin = offsets[total - 1] + counts[total - 1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    int64_t src;
    if (j < fin) {
      int64_t lo = 0, hi = total; // last i with offsets[i] <= j
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= j) lo = mid;
        else hi = mid;
      }
      src = lo;
    } else {
      src = total - (n - fin) + (j - fin);
      if (src < 0) src = 0;
    }
    const int64_t q = perm[src];
    out.x[j] = F.x[q]; out.y[j] = F.y[q]; out.z[j] = F.z[q];
    out.u[j] = F.u[q]; out.v[j] = F.v[q]; out.w[j] = F.w[q];
    out.E[j] = F.E[q];
    out.elem[j] = F.elem[q];
    out_key[j] = F.key[q];
    uint64_t seed = init_seed((next_batch - 1) * n + j + 1, 0, master);
    if (roundtrip) seed = seed_roundtrip(seed);
    out.seed[j] = seed;
  }
}

extern "C" __global__ void k_iota(uint32_t *v, int64_t n)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (uint32_t)i;
}

// Sites (through perm when given) as mcw_site records in caller element ids.
extern "C" __global__ void k_sites_out(const double *x, const double *y, const double *z, const double *u,
                                       const double *v, const double *w, const double *E, const int32_t *elem,
                                       const uint64_t *key, const uint32_t *perm, int64_t n,
                                       const int32_t *orig_of_int, mcw_site *out)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const int64_t q = perm ? (int64_t)perm[j] : j;
    mcw_site s;
    s.r[0] = x[q]; s.r[1] = y[q]; s.r[2] = z[q];
    s.u[0] = u[q]; s.u[1] = v[q]; s.u[2] = w[q];
    s.E = E[q];
    s.parent = (int64_t)(key[q] >> PROGENY_BITS);
    s.progeny = (int32_t)(key[q] & ((1ull << PROGENY_BITS) - 1));
    s.element = elem[q] < 0 ? -1 : orig_of_int[elem[q]];
    out[j] = s;
  }
}

static unsigned grid_of(int64_t n, int64_t cap = 8192)
{
  int64_t b = (n + 255) / 256;
  if (b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

// Fixed-order sum of in[0..n) into *out; scratch holds 2 ceil(n / 1024) values.
hipError_t launch_tree_sum(hipStream_t s, const double *in, int64_t n, double *scratch, double *out)
{
  const int64_t nb0 = (n + 1023) / 1024;
  double *buf[2] = {scratch, scratch + nb0};
  int k = 0;
  for (;;) {
    const int64_t nb = (n + 1023) / 1024;
    double *dst = nb == 1 ? out : buf[k];
    hipLaunchKernelGGL(k_tree_sum, dim3((unsigned)nb), dim3(256), 0, s, in, n, dst);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 1) return e;
    in = dst;
    n = nb;
    k ^= 1;
  }
}

hipError_t fission_sort_temp_bytes(int64_t n, size_t *bytes)
{
  *bytes = 0;
  return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                            (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 64);
}

// perm <- slots of the n banked sites in key order.
hipError_t launch_fission_sort(hipStream_t s, const FissionView &F, int64_t n, int end_bit, uint64_t *keys_tmp,
                               uint32_t *perm, uint32_t *idx, void *tmp, size_t tmp_bytes)
{
  hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(256), 0, s, idx, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, (const uint64_t *)F.key, keys_tmp, (const uint32_t *)idx,
                                            perm, (int)n, 0, end_bit, s);
}

hipError_t launch_sync_count(hipStream_t s, int64_t total, int64_t base, double p_sample, uint64_t seed0,
                             int64_t *counts)
{
  hipLaunchKernelGGL(k_sync_count, dim3(grid_of(total)), dim3(256), 0, s, total, base, p_sample, seed0, counts);
  return hipGetLastError();
}

hipError_t scan_temp_bytes(int64_t n, size_t *bytes)
{
  *bytes = 0;
  return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, (const int64_t *)nullptr, (int64_t *)nullptr, (int)n);
}

hipError_t launch_scan(hipStream_t s, const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes)
{
  return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int)n, s);
}

hipError_t launch_sync_scatter(hipStream_t s, const FissionView &F, const uint32_t *perm, const int64_t *counts,
                               const int64_t *offsets, int64_t total, int64_t n, const SourceBank &out,
                               uint64_t *out_key, int64_t next_batch, uint64_t master, int roundtrip)
{
  hipLaunchKernelGGL(k_sync_scatter, dim3(grid_of(n)), dim3(256), 0, s, F, perm, counts, offsets, total, n, out,
                     out_key, next_batch, master, roundtrip);
  return hipGetLastError();
}

hipError_t launch_sites_out(hipStream_t s, const double *x, const double *y, const double *z, const double *u,
                            const double *v, const double *w, const double *E, const int32_t *elem,
                            const uint64_t *key, const uint32_t *perm, int64_t n, const int32_t *orig_of_int,
                            mcw_site *out)
{
  hipLaunchKernelGGL(k_sites_out, dim3(grid_of(n)), dim3(256), 0, s, x, y, z, u, v, w, E, elem, key, perm, n,
                     orig_of_int, out);
  return hipGetLastError();
}

} // namespace mcw
