// grid_sync.hip -- host side of the grid barrier (grid_sync.h): a per-device
// pool of barrier blocks leased one per launch.
#include "grid_sync.h"

#include <map>
#include <mutex>
#include <vector>

#include "api_common.h"

namespace perks {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::vector<GridSync*> g_free[kMaxDevices];  // idle pooled blocks per device
std::map<GridSync*, int> g_owner;           // every live block -> its device
unsigned long long g_timeout[kMaxDevices] = {};

int current_device(int* dev)
{
    PERKS_HIP_TRY(hipGetDevice(dev));
    if (*dev < 0 || *dev >= kMaxDevices) return PERKS_ERR_INVALID_ARG;
    return PERKS_OK;
}

// Allocates and zeroes a block on device `dev` (current); g_mu held.
int new_block(int dev, GridSync** out)
{
    GridSync* p = nullptr;
    PERKS_HIP_TRY(hipMalloc(&p, sizeof(GridSync)));
    if (hipMemset(p, 0, sizeof(GridSync)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(p);
        return PERKS_ERR_HIP;
    }
    if (!g_timeout[dev]) {
        int khz = 0;
        PERKS_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) khz = 100000;
        g_timeout[dev] = (unsigned long long)khz * 1000ull * 10ull;
    }
    g_owner[p] = dev;
    *out = p;
    return PERKS_OK;
}
}  // namespace

int grid_sync_acquire(GridSync** out)
{
    int dev = 0;
    int rc = current_device(&dev);
    if (rc != PERKS_OK) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_free[dev].empty()) {
        *out = g_free[dev].back();
        g_free[dev].pop_back();
        return PERKS_OK;
    }
    return new_block(dev, out);
}

void grid_sync_return(GridSync* gs)
{
    if (!gs) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_owner.find(gs);
    if (it != g_owner.end()) g_free[it->second].push_back(gs);
}

int grid_sync_alloc(GridSync** out)
{
    int dev = 0;
    int rc = current_device(&dev);
    if (rc != PERKS_OK) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    return new_block(dev, out);
}

int grid_sync_free(GridSync* gs)
{
    if (!gs) return PERKS_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    g_owner.erase(gs);
    PERKS_HIP_TRY(hipFree(gs));
    return PERKS_OK;
}

unsigned long long grid_sync_timeout()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices || !g_timeout[dev]) return 1000000000ull;
    return g_timeout[dev];
}

int grid_sync_check(GridSync* gs, hipStream_t stream)
{
    unsigned int abort_word = 0;
    PERKS_HIP_TRY(hipMemcpyAsync(&abort_word, (unsigned int*)(gs->w + kSyncAbortWord), sizeof(abort_word), hipMemcpyDeviceToHost, stream));
    PERKS_HIP_TRY(hipStreamSynchronize(stream));
    if (!abort_word) return PERKS_OK;
    PERKS_HIP_TRY(hipMemsetAsync(gs, 0, sizeof(GridSync), stream));
    PERKS_HIP_TRY(hipStreamSynchronize(stream));
    std::fprintf(stderr, "perks_amd: a grid barrier wait timed out; the launch was abandoned\n");
    return PERKS_ERR_BARRIER_TIMEOUT;
}

hipError_t launch_persistent(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t stream)
{
    hipError_t e = hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, stream);
    if (e != hipErrorIllegalState) return e;
    (void)hipGetLastError();  // clear the refusal
    static std::once_flag note;
    std::call_once(note, [] {
        std::fprintf(stderr, "perks_amd: the cooperative queue is held by another process; persistent grids run as "
                             "ordinary launches of the same size\n");
    });
    return hipLaunchKernel(fn, grid, block, args, lds, stream);
}

int grid_sync_release()
{
    int dev = 0;
    int rc = current_device(&dev);
    if (rc != PERKS_OK) return PERKS_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    while (!g_free[dev].empty()) {
        GridSync* p = g_free[dev].back();
        g_free[dev].pop_back();
        g_owner.erase(p);
        PERKS_HIP_TRY(hipFree(p));
    }
    return PERKS_OK;
}

}  // namespace perks
