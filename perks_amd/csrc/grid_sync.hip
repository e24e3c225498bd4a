// grid_sync.hip -- host side of the grid barrier (grid_sync.h).
#include "grid_sync.h"

#include <mutex>

#include "api_common.h"

namespace perks {

namespace {
std::mutex g_mu;
GridSync* g_blocks[64][kSyncUsers] = {};
unsigned long long g_timeout[64] = {};
}  // namespace

int grid_sync_get(int user, GridSync** out)
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64 || user < 0 || user >= kSyncUsers) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_blocks[dev][user]) {
        GridSync* p = nullptr;
        PERKS_HIP_TRY(hipMalloc(&p, sizeof(GridSync)));
        PERKS_HIP_TRY(hipMemset(p, 0, sizeof(GridSync)));
        PERKS_HIP_TRY(hipDeviceSynchronize());
        g_blocks[dev][user] = p;
    }
    if (!g_timeout[dev]) {
        int khz = 0;
        PERKS_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) khz = 100000;
        g_timeout[dev] = (unsigned long long)khz * 1000ull * 10ull;
    }
    *out = g_blocks[dev][user];
    return PERKS_OK;
}

unsigned long long grid_sync_timeout()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || !g_timeout[dev]) return 1000000000ull;
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

int grid_sync_release()
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    for (int u = 0; u < kSyncUsers; ++u)
        if (g_blocks[dev][u]) {
            PERKS_HIP_TRY(hipFree(g_blocks[dev][u]));
            g_blocks[dev][u] = nullptr;
        }
    return PERKS_OK;
}

}  // namespace perks
