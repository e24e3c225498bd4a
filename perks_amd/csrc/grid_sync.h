// grid_sync.h -- the grid-wide barrier of the persistent (cooperative) kernels.
//
// State lives in one device block per subsystem and device, zeroed once
// when it is allocated.  The barrier resets itself: the last workgroup to
// arrive returns the arrival count to zero and then advances the generation,
// so a launch needs no clearing kernel or memset beforehand (and a launch
// captured in a hipGraph carries none).
//
// Every wait is bounded.  A workgroup that waits longer than `timeout` ticks
// of the device wall clock sets the abort word and leaves; every other
// workgroup sees the word in its own wait (or at its next barrier) and leaves
// too, so the grid always drains.  The host checks the word after the launch
// (grid_sync_check) and reports PERKS_ERR_BARRIER_TIMEOUT.
#pragma once

#include <hip/hip_runtime.h>

namespace perks {

// Layout: the root counter and the abort word on lines of their own, then
// kSyncGroups sub-groups, each with an arrival counter and a release word,
// 4352 bytes apart (17 x 256 B: the groups fall on different HBM channels).
// Workgroup w belongs to group w % kSyncGroups; the last arrival of a group
// arrives at the root; the last root arrival releases every group.  At most
// about G/32 atomics meet on one address, and each group polls its own word.
// The last arrival of a group resets the group counter, and the releasing
// arrival resets the root, each with an atomic exchange that completes before
// the next step is issued, so every barrier leaves all counters at zero and
// the next launch may use any grid size.
constexpr int kSyncGroups = 32;
constexpr int kSyncStride = 4352 / 8;  // in 64-bit words

struct GridSync {
    unsigned long long w[kSyncStride * (kSyncGroups + 1)];
};
static_assert(sizeof(GridSync) == 4352 * 33, "GridSync layout");
constexpr int kSyncAbortWord = 256;  // index of the abort word (2 KiB into the root block)

__device__ __forceinline__ unsigned int* gs_root(GridSync* gs) { return (unsigned int*)gs->w; }
__device__ __forceinline__ unsigned int* gs_abort(GridSync* gs) { return (unsigned int*)(gs->w + kSyncAbortWord); }
__device__ __forceinline__ unsigned int* gs_count(GridSync* gs, int g)
{
    return (unsigned int*)(gs->w + kSyncStride * (g + 1));
}

// Group layout of a grid of nblocks workgroups: ng = min(nblocks, kSyncGroups)
// groups, workgroup w in group w % ng, so group g has
// ceil((nblocks - g) / ng) members.  Host-callable so the arithmetic can be
// checked off the device.
__host__ __device__ constexpr unsigned int gs_groups(unsigned int nblocks)
{
    return nblocks < (unsigned int)kSyncGroups ? nblocks : (unsigned int)kSyncGroups;
}
__host__ __device__ constexpr unsigned int gs_members(unsigned int nblocks, unsigned int g)
{
    return (nblocks - g + gs_groups(nblocks) - 1) / gs_groups(nblocks);
}

// Waits until this thread's outstanding memory operations (here: an atomic
// exchange) have completed, and keeps the compiler from moving memory
// operations across the wait.
__device__ __forceinline__ void gs_complete()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ unsigned int* gs_gen(GridSync* gs, int g)
{
    return (unsigned int*)(gs->w + kSyncStride * (g + 1) + 16);
}

// All threads of the workgroup call this.  Every wave waits for its own loads
// and stores to complete, the workgroup meets, and thread 0 then writes back
// the L2 at agent scope (so the other XCDs see what the workgroup stored),
// arrives, waits for the release, and invalidates L1/L2 at agent scope before
// the workgroup meets again and reads.  Returns false when the grid is
// aborting: the caller must leave the kernel.
__device__ __forceinline__ bool grid_sync(GridSync* gs, unsigned int nblocks, unsigned long long timeout)
{
    __shared__ int s_ok;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned int ng = gs_groups(nblocks);
        const int g = (int)(blockIdx.x % ng);
        const unsigned int members = gs_members(nblocks, (unsigned int)g);
        unsigned int* gen = gs_gen(gs, g);
        // The release word is read before arriving (it cannot change without this
        // arrival); the release fence below also waits for this load.
        const unsigned int g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned int before =
            __hip_atomic_fetch_add(gs_count(gs, g), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool released = false;
        if (before == members - 1) {
            __hip_atomic_exchange(gs_count(gs, g), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            gs_complete();
            const unsigned int rb = __hip_atomic_fetch_add(gs_root(gs), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (rb == ng - 1) {
                __hip_atomic_exchange(gs_root(gs), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                gs_complete();
                for (unsigned int j = 0; j < ng; ++j)
                    __hip_atomic_fetch_add(gs_gen(gs, (int)j), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
        if (!released) {
            const unsigned long long t0 = wall_clock64();
            unsigned int spins = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                // The abort word and the clock are looked at every 64th poll only.
                if ((++spins & 63u) != 0u) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                if (__hip_atomic_load(gs_abort(gs), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > timeout) {
                    __hip_atomic_store(gs_abort(gs), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        if (ok && __hip_atomic_load(gs_abort(gs), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Cooperative grids are sized to at most this many waves per CU.  On MI355X
// a grid of 8 four-wave workgroups per CU (every wave slot the occupancy
// query reports) did not reach its first barrier within the wait limit,
// while grids of up to 16 waves per CU meet every barrier; the persistent
// kernels' measured optimum is at or below that anyway.
constexpr int kCoopMaxWavesPerCu = 16;

// ---- host side ----------------------------------------------------------------
// Barrier blocks come from a per-device pool.  A launch leases a block of its
// own for the time from the launch until grid_sync_check has seen the launch
// complete, so launches from different host threads or streams never share
// counters or abort words.  A batch plan owns a block for its whole lifetime
// (grid_sync_alloc / grid_sync_free), so perks_release() cannot free a block
// that a captured graph still names.

// Takes a zeroed block from the current device's pool (allocating one when
// the pool is empty).
int grid_sync_acquire(GridSync** out);
// Returns a block taken with grid_sync_acquire to the pool of the device it
// came from.  The block must be idle (its launch completed and checked).
void grid_sync_return(GridSync* gs);
// A block owned by the caller (allocated and zeroed on the current device).
int grid_sync_alloc(GridSync** out);
int grid_sync_free(GridSync* gs);
// Wait limit in wall-clock ticks for one barrier (10 s of the current device).
unsigned long long grid_sync_timeout();
// After a launch has completed: PERKS_OK, or PERKS_ERR_BARRIER_TIMEOUT (the
// block is then cleared for the next launch).  Synchronises `stream`.
int grid_sync_check(GridSync* gs, hipStream_t stream);
// Frees the current device's idle pooled blocks (leased and plan-owned blocks
// stay valid).
int grid_sync_release();

// Launches a persistent grid (at most CUs x measured occupancy workgroups)
// cooperatively.  The HIP runtime refuses a cooperative launch with
// hipErrorIllegalState while another process on the device holds its
// cooperative queue; the same grid then goes out as an ordinary launch, which
// an otherwise idle device places whole, and whose barrier's bounded wait
// turns a grid that is not co-resident into PERKS_ERR_BARRIER_TIMEOUT rather
// than a hang.
hipError_t launch_persistent(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t stream);

// Scope guard for a leased block.
struct SyncLease {
    GridSync* gs = nullptr;
    SyncLease() = default;
    SyncLease(const SyncLease&) = delete;
    SyncLease& operator=(const SyncLease&) = delete;
    int acquire() { return gs ? 0 : grid_sync_acquire(&gs); }
    ~SyncLease()
    {
        if (gs) grid_sync_return(gs);
    }
};

}  // namespace perks
