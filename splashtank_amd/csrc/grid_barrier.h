// grid_barrier.h - a barrier over the workgroups of one cooperative launch
// (every workgroup resident at once), shared by the solves that run all their
// sweeps in a single launch.
//
// The n workgroups taking part are numbered w = 0 .. n-1 (blockIdx.x and
// gridDim.x for a whole launch, or a range of it).  Words: flags[w] is
// workgroup w's arrival epoch (n <= blockDim.x, so workgroup 0 can wait for
// every flag with one thread each), *release the epoch
// workgroup 0 has let go, *fails the count of time-outs.  Every wave waits
// until its stores are done; workgroup w raises flags[w] to `epoch` after an
// agent-scope release fence; workgroup 0 waits for every flag and raises
// *release; every other workgroup waits for *release with a single thread; an
// agent-scope acquire fence follows, so plain stores before the barrier are
// seen by plain loads after it on every XCD.  Each wait gives up after
// GB_SPIN_LIMIT polls, so every wave always leaves: a time-out sets *sFail (a
// shared word, 0 on entry), is counted in *fails, and the barrier returns false.
// The caller clears flags and *release before the launch; epochs start at 1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GB_SPIN_LIMIT (1u << 21)

__device__ __forceinline__ bool gb_wait(const uint32_t* word, uint32_t epoch, uint32_t* sFail) {
  uint32_t spins = 0;
  while (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
    if (++spins > GB_SPIN_LIMIT) {
      *sFail = 1u;
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

__device__ __forceinline__ bool gb_barrier(uint32_t* flags, uint32_t* release, uint32_t* fails, uint32_t epoch, uint32_t* sFail,
                                           uint32_t w, uint32_t n) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(&flags[w], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (w == 0) {
    if (threadIdx.x < n) gb_wait(&flags[threadIdx.x], epoch, sFail);
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      __hip_atomic_store(release, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (threadIdx.x == 0) {
    gb_wait(release, epoch, sFail);
  }
  __syncthreads();
  if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  __syncthreads();
  if (*sFail) {
    if (threadIdx.x == 0) atomicAdd(fails, 1u);
    return false;
  }
  return true;
}
