// mcw_sort.hip -- ordering particles by (material, element).
//
// Sorting a bank by the key (material << shift | element) puts particles
// that will read the same element records, face planes and material data
// next to each other, so a wavefront's gathers share cache lines.  Internal
// element ids follow a Morton (Z-order) curve, so key order is also spatial
// order within a material.  The sort is a radix sort of (key, slot) pairs
// (hipCUB) followed by a gather of the particle fields into the other bank.
// Per-history results do not depend on bank order: every history draws from
// its own random-number stream.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "mcw_types.h"

namespace mcw {

__device__ __forceinline__ uint32_t sort_key(const MeshView &M, int32_t e, int shift, uint32_t mat_mask)
{
  if (e < 0) return 0xffffffffu; // lost at birth: last
  const uint32_t mat = (uint32_t)M.erec[2 * (int64_t)e + 1].z & mat_mask;
  return (mat << shift) | (uint32_t)e;
}

extern "C" __global__ void k_key_sites(MeshView M, const int32_t *elem, int64_t n, int shift, uint32_t mat_mask,
                                       uint32_t *key, uint32_t *slot)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = sort_key(M, elem[i], shift, mat_mask);
  slot[i] = (uint32_t)i;
}

// out[j] = in[perm[j]]; out.slot[j] = the original slot (in.slot composed
// when the input was itself sorted).
extern "C" __global__ void k_gather_sites(SourceBank in, SourceBank out, const uint32_t *perm, int64_t n)
{
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t i = perm[j];
  out.x[j] = in.x[i]; out.y[j] = in.y[i]; out.z[j] = in.z[i];
  out.u[j] = in.u[i]; out.v[j] = in.v[i]; out.w[j] = in.w[i];
  out.E[j] = in.E[i];
  out.seed[j] = in.seed[i];
  out.elem[j] = in.elem[i];
  out.slot[j] = in.slot ? in.slot[i] : (int32_t)i;
}

extern "C" __global__ void k_key_bank(MeshView M, BankView B, int64_t n, int shift, uint32_t mat_mask,
                                      uint32_t *key, uint32_t *slot)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = sort_key(M, B.elem[i], shift, mat_mask);
  slot[i] = (uint32_t)i;
}

// out[j] = in[perm[j]] over every field of the event bank; the queue count
// of `out` becomes n.
extern "C" __global__ void k_gather_bank(BankView in, BankView out, const uint32_t *perm, int64_t n, int rec,
                                         Counters *C, int qout)
{
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) C->q_count[qout] = (unsigned long long)n;
  if (j >= n) return;
  const int64_t i = perm[j];
  out.x[j] = in.x[i]; out.y[j] = in.y[i]; out.z[j] = in.z[i];
  out.u[j] = in.u[i]; out.v[j] = in.v[i]; out.w[j] = in.w[i];
  out.E[j] = in.E[i]; out.sigt[j] = in.sigt[i];
  out.seed[j] = in.seed[i]; out.hid[j] = in.hid[i];
  out.elem[j] = in.elem[i]; out.inface[j] = in.inface[i]; out.mat[j] = in.mat[i];
  out.k[j] = in.k[i]; out.nseg[j] = in.nseg[i];
  if (rec) {
    out.digest[j] = in.digest[i];
    out.ncoll[j] = in.ncoll[i];
  }
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// Temporary storage the radix sort of n pairs needs.
hipError_t sort_temp_bytes(int64_t n, size_t *bytes)
{
  *bytes = 0;
  return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                            (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 32);
}

static hipError_t sort_pairs(hipStream_t s, int64_t n, int bits, uint32_t *k0, uint32_t *k1, uint32_t *v0,
                             uint32_t *v1, void *tmp, size_t tmp_bytes)
{
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, v0, v1, (int)n, 0, bits, s);
}

// Sort the n sites of `in` into `out` (which must have a slot array).
hipError_t launch_sort_sites(hipStream_t s, const MeshView &M, const SourceBank &in, const SourceBank &out,
                             int64_t n, int shift, int bits, uint32_t mat_mask, uint32_t *k0, uint32_t *k1,
                             uint32_t *v0, uint32_t *v1, void *tmp, size_t tmp_bytes)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_key_sites, dim3(blocks_for(n)), dim3(256), 0, s, M, in.elem, n, shift, mat_mask, k0, v0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = sort_pairs(s, n, bits, k0, k1, v0, v1, tmp, tmp_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather_sites, dim3(blocks_for(n)), dim3(256), 0, s, in, out, v1, n);
  return hipGetLastError();
}

// Sort the n live particles of event bank `in` into `out`; q_count[qout] = n.
hipError_t launch_sort_bank(hipStream_t s, const MeshView &M, const BankView &in, const BankView &out, int64_t n,
                            bool rec, Counters *C, int qout, int shift, int bits, uint32_t mat_mask, uint32_t *k0,
                            uint32_t *k1, uint32_t *v0, uint32_t *v1, void *tmp, size_t tmp_bytes)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_key_bank, dim3(blocks_for(n)), dim3(256), 0, s, M, in, n, shift, mat_mask, k0, v0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = sort_pairs(s, n, bits, k0, k1, v0, v1, tmp, tmp_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather_bank, dim3(blocks_for(n)), dim3(256), 0, s, in, out, v1, n, rec ? 1 : 0, C, qout);
  return hipGetLastError();
}

} // namespace mcw
