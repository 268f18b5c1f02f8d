// sort.hip -- the energy order of a chunk and the active list for the
// host photon list (marx_create_photons and the drop-in stages).
//
// marx_create_photons must leave sorted_index / sorted_energies in energy
// order with the index as tie-break (source.c:361-371, dsort.c:28-61).  On
// the device that is one stable radix sort of (energy, index) pairs of the
// resident store: the index order of equal energies is the input order, as
// the reference's comparator has it.  After a drop-in stage, the pruned
// sorted_index is the stage's output list itself (the compaction keeps the
// input order), so it and its energies are copied back instead of being
// rebuilt on the host.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "dev_store.h"

namespace mxg {

__global__ __launch_bounds__ (kBlock)
void k_energy_pairs (const double *__restrict__ energy, uint32_t n, double *__restrict__ keys,
                     uint32_t *__restrict__ vals)
{
   uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   keys[i] = energy[i];
   vals[i] = i;
}

__global__ __launch_bounds__ (kBlock)
void k_gather_energies (const double *__restrict__ energy, const uint32_t *__restrict__ idx,
                        const uint32_t *__restrict__ n, double *__restrict__ out)
{
   uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= *n) return;
   out[i] = energy[idx[i]];
}

// Sorts the store's first n energies with their indices: keys_out /
// vals_out receive them in ascending energy order (stable).  tmp / *tmp_bytes:
// rocPRIM's temporary storage; with tmp == nullptr only *tmp_bytes is set.
hipError_t sort_energies (const double *energy, uint32_t n, double *keys_in, uint32_t *vals_in,
                          double *keys_out, uint32_t *vals_out, void *tmp, size_t *tmp_bytes,
                          hipStream_t st)
{
   if (tmp == nullptr)
     return rocprim::radix_sort_pairs (nullptr, *tmp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0,
                                       8 * (unsigned int) sizeof (double), st);
   if (n == 0) return hipSuccess;
   hipLaunchKernelGGL (k_energy_pairs, dim3 ((n + kBlock - 1) / kBlock), dim3 (kBlock), 0, st, energy, n,
                       keys_in, vals_in);
   hipError_t e = hipGetLastError ();
   if (e != hipSuccess) return e;
   return rocprim::radix_sort_pairs (tmp, *tmp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0,
                                     8 * (unsigned int) sizeof (double), st);
}

hipError_t gather_energies (const double *energy, const uint32_t *idx, const uint32_t *n_dev, uint32_t n_upper,
                            double *out, hipStream_t st)
{
   if (n_upper == 0) return hipSuccess;
   hipLaunchKernelGGL (k_gather_energies, dim3 ((n_upper + kBlock - 1) / kBlock), dim3 (kBlock), 0, st, energy,
                       idx, n_dev, out);
   return hipGetLastError ();
}

} // namespace mxg
