// Host-side bookkeeping shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/fdtdpic.h"

struct SimHost {
    fdtd_sim_desc_t d;          // scalars; host pointers cleared after upload
    void* block = nullptr;      // one device allocation: coefficients then state
    size_t bytes = 0;
    size_t lay_off[64] = {0};   // byte offsets of the sub-arrays inside block
    int ns = 0, nstrip = 0;     // slab cells, strip cells
    long long launch_begin = 0; // context launch index at admission (ping-pong parity base)
    long long t_begin = 0;      // context global step at admission
    long long steps_done = 0;
    long long launches = 0;     // launches this simulation took part in
};

struct TileGroup {
    int kind, offset, count;    // tiles [offset, offset+count) run the kernel for this slab kind
};

struct fdtd_ctx {
    int prec = FDTD_F64;
    hipStream_t st = nullptr;
    int K = 2, S = 16;
    std::vector<SimHost*> sims;
    void* d_sims = nullptr;
    size_t d_sims_cap = 0;
    int2* d_tiles = nullptr;
    size_t d_tiles_cap = 0;
    int ntiles = 0;
    std::vector<int> tile_sims;  // simulations covered by the current tile list
    std::vector<TileGroup> groups;
    long long launch_idx = 0, gstep = 0;
    bool dirty = true;
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t ev_used = 0;
    long long launches = 0;
    long long dev_bytes = 0;
};

int fdtd_fail(int code, const std::string& msg);

// Default steps per launch for K cells per lane: the tile keeps 75% of its
// cells as output (U = 64K - 2S = 48K).
inline int default_S(int K) { return 8 * K; }
