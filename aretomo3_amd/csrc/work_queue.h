// work_queue.h -- device helpers shared by the persistent (batched)
// back-projection and forward-projection kernels (gfx950).
#pragma once

#include "atr_internal.h"

namespace atr {

// A value every lane holds alike, moved to an SGPR so that the compiler
// treats what is derived from it as wave-uniform.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Block-wide work queue of the persistent launches.  Thread 0 claims item
// indices with an atomic on the queue counter; every thread sees the value
// through LDS.  The item after the current one is claimed when the current
// one starts, so the atomic's round trip overlaps the item's own loads
// instead of stalling the block between items.  Indices are handed out in
// increasing order, each to exactly one block, and a block leaves when it
// holds an index past the end (every block therefore exits).
__device__ __forceinline__ int publish_item(int v)
{
    __shared__ int s_item;
    if (threadIdx.x == 0) s_item = v;
    __syncthreads();
    const int it = uniform(s_item);
    __syncthreads();
    return it;
}

__device__ __forceinline__ int claim_item(int* __restrict__ counter)
{
    return threadIdx.x == 0 ? atomicAdd(counter, 1) : 0;
}

// Entry of queue index `it`.  Entries are ordered by their first item and a
// block's indices only grow, so the search resumes where the last one ended.
template <typename Item>
__device__ __forceinline__ int entry_of(cptr<Item> E, int n_e, int it, int s)
{
    while (s + 1 < n_e && E[s + 1].item0 <= it) s++;
    return s;
}

}  // namespace atr
