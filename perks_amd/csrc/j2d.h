// j2d.h -- 2D Jacobi (single grid and batches of independent grids).
#pragma once

#include <hip/hip_runtime.h>

#include "grid_sync.h"
#include "j3d_common.h"

namespace perks {

// One grid of a run: buffers a (input) and b, both overwritten; the result is
// in b when result_in_b is set on return, else in a.
struct J2dSpec {
    void* a;
    void* b;
    int ny, nx;
    int steps;
    int result_in_b;  // out
};

// Device-side descriptor (sorted by pass count, item_end cumulative).
struct J2dGrid {
    void* a;
    void* b;
    int ny, nx;
    int steps;
    int tiles;
    int item_end;
    int seg;  // rows per segment (streaming kernel, tile variant 2)
};

struct J2dArgs {
    const J2dGrid* grids;
    const int* active;      // grids still running in each pass (a prefix of grids)
    int npass;              // passes this launch runs
    int step_offset;        // index of this launch's first pass (per-step launches)
    int levels;             // time steps per full pass
    GridSync* sync;               // grid barrier state (grid_sync.h)
    unsigned long long timeout;   // barrier wait limit, wall-clock ticks
    int sr;                       // stochastic rounding of 16-bit storage
    uint32_t seed;
};

int j2d_run(int dtype, J2dSpec* specs, int n, const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep);
// j2d_run with the launch's scratch in a fixed device array: no device allocation.
int j2d_run_noalloc(int dtype, J2dSpec* specs, int n, const perks_stencil_opts* o, hipStream_t stream,
                    perks_run_report* rep);
int j2d_pass_levels(int dtype, int shape, const perks_stencil_opts* o, int tile);
int release_j2d_scratch();

// Kernel entry points per storage type (j2d_<type>.hip); nullptr if not built.
// tile: 0 or 1 (LDS tiles, TileOf in j2d_kernel.h).
void* j2d_kernel_f64(int shape, int bc, bool has_div, int tile);
void* j2d_kernel_f32(int shape, int bc, bool has_div, int tile);
void* j2d_kernel_f16(int shape, int bc, bool has_div, int tile);
void* j2d_kernel_bf16(int shape, int bc, bool has_div, int tile);

// Deepest pass the row-streaming kernel is built for, by stencil radius.
__host__ __device__ constexpr int stream_max_levels(int radius)
{
    return radius == 1 ? 4 : (radius == 2 ? 2 : 1);
}

// Ghost columns per side of a two-column streaming strip: R*L rounded up to even.
__host__ __device__ constexpr int pair_ghost(int radius, int levels) { return 2 * ((radius * levels + 1) / 2); }

// The row-streaming kernel (j2d_stream_kernel.h) per storage type: tile
// variant 2 (one column per lane) or, with pair, 4 (two columns per lane;
// FP32/FP16/BF16 storage and even grid widths; nullptr for FP64).
void* j2d_stream_kernel_f64(int shape, int bc, bool has_div, bool pair);
void* j2d_stream_kernel_f32(int shape, int bc, bool has_div, bool pair);
void* j2d_stream_kernel_f16(int shape, int bc, bool has_div, bool pair);
void* j2d_stream_kernel_bf16(int shape, int bc, bool has_div, bool pair);

}  // namespace perks
