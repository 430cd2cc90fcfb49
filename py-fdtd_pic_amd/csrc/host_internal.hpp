// Host-side bookkeeping shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <cstdint>
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
    // energy accumulator (fdtd_set_energy): H' before a launch's last H update
    // (N+1 FP64) then the recorded W values (rows FP64), one allocation
    void* eblock = nullptr;
    size_t ebytes = 0;
    int e_every = 0, e_rows = 0, e_count = 0;
    long long e_launches = 0;   // launches in which this simulation stepped, since energy was set
    std::vector<int64_t> e_steps;   // step index after which each row was taken
};

// One particle beam attached to a simulation (at most one per simulation).
struct BeamHost {
    int sim = -1;
    int N = 0;                   // grid size of the simulation
    int64_t np = 0;
    int G = 1;                   // cells per push/count group
    int nwaves = 1, nnode = 1;   // push/count waves, scatter node blocks
    int nmove = 1;               // scatter blocks for staged movers and parked particles
    int xr_want = 1;             // one-cell beams: LDS rows per cell that hold its fullest cell at attach
    int32_t range0[2] = {0, -1}; // live cell range at attach
    double q = 0, m = 0, w = 0;  // charge, mass, weight per unit transverse area
    double a = 0, jscale = 0, Sc = 0;
    void* block = nullptr;
    size_t bytes = 0;
    size_t off[32] = {0};
    std::vector<int32_t> cell0, id0, start0;   // initial sorted state (kept for reset)
    std::vector<double> frac0, u0;             // frac, then ux uy uz interleaved
};

struct EvPair {
    hipEvent_t first = nullptr, second = nullptr;
    int kind = 0;               // FDTD_KERNEL_*
    int launches = 1;           // launches between first and second
};

struct TileGroup {
    int kind, offset, count;    // tiles [offset, offset+count) run the kernel for this slab kind
};

struct fdtd_ctx {
    int prec = FDTD_F64;
    int device = 0;                          // HIP device current at fdtd_ctx_create
    hipStream_t st = nullptr;
    hipStream_t up = nullptr;                // uploads of new simulations and beams (never waits on st);
                                             // one per device, shared by its contexts
    // pinned staging of the descriptor / tile / work-list uploads; reused
    // once the event recorded after the last upload has completed
    void* stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t stage_ev = nullptr;
    bool stage_pending = false;
    int K = 2, S = 16;
    int K_slab = 2;             // cells per lane of tiles meeting a D-form slab (K or K/2)
    int halo = 0;               // requested workgroup halo exchange period (0: off)
    bool grouped = false;       // the last rebuild's tile list is in workgroup windows
    int S_run = 16;             // steps per launch of the current tile list (1 with a beam)
    std::vector<SimHost*> sims;
    void* d_sims = nullptr;
    size_t d_sims_cap = 0;
    int2* d_tiles = nullptr;
    size_t d_tiles_cap = 0;     // bytes (all *_cap members)
    int ntiles = 0;
    std::vector<int> tile_sims;  // simulations covered by the current tile list
    std::vector<TileGroup> groups;
    long long launch_idx = 0, gstep = 0;
    bool dirty = true;
    int profile = 0;            // stamp every profile-th launch with events (0: off)
    std::vector<EvPair> events;   // profiling: start/stop per launch, tagged by kernel kind
    size_t ev_used = 0;
    long long bracket = -1;       // field-only profiling: the open run of launches (index into events)
    long long launches = 0;
    long long dev_bytes = 0;
    // particle beams
    std::vector<BeamHost*> beams;            // indexed like sims (nullptr: no beam)
    void* d_beams = nullptr;
    size_t d_beams_cap = 0;
    int2* d_pblocks = nullptr;               // scatter blocks (beam, first particle | -(node block + 1))
    size_t d_pblocks_cap = 0;
    int2* d_pitems = nullptr;                // push/count items: npitems multi-cell waves (beam, wave
    size_t d_pitems_cap = 0;                 // index), then npcell one-cell workgroups (beam, workgroup)
    int npblocks = 0, npitems = 0, npcell = 0;
    int npnode = 0;                          // scatter blocks: npnode node blocks, then particle blocks
    bool has_pic = false;                    // some active simulation carries a beam
    bool fuse_pic = true;                    // scatter in the same launch as the last tile group's field step
    // bound on queued work: an event every kThrottleEvery launches; the host
    // waits on the one recorded kThrottleDepth events back before queueing more
    std::vector<hipEvent_t> throttle;
    size_t throttle_head = 0;
    long long throttle_launches = 0;
    int32_t* d_err = nullptr;                // index-check flag of the sort kernel
    // persistent field launches (fdtd_set_persistent)
    bool persist = false;                    // requested
    bool persist_ok = false;                 // the current tile list qualifies
    int cu_count = 0;
    int lds_max = 0;            // shared memory a workgroup may use (bytes; 0: unknown)
    int pc_rows = 0;            // forced LDS rows per cell of the one-cell PIC kernel (0: chosen per launch)
    int pc_xr = 1;              // the rows of the current tile list's launches
    int pc_single = -1;         // the one-cell items' beam when there is one, else -1
    unsigned long long* d_flags = nullptr;   // per tile: last block finished (global launch index + 1)
    size_t d_flags_cap = 0;
    int2* d_nbr = nullptr;                   // per tile: left, right neighbour tile
    size_t d_nbr_cap = 0;
    int32_t* d_perr = nullptr;               // set by a tile whose neighbour wait timed out
    void* d_scratch = nullptr;               // spectra: pointer table, frequencies, results
    size_t d_scratch_cap = 0;
};

int fdtd_fail(int code, const std::string& msg);

// Largest field launch (tiles) that takes the particle scatter into its grid.
constexpr int kFuseMaxTiles = 2048;

// Blocks of S steps in one persistent field launch (bounded so that a long
// run still returns control to the host now and then), and the bound on one
// neighbour wait inside it (2 s of the 100 MHz real-time counter).
constexpr int kPersistMaxBlocks = 4096;
constexpr long long kPersistSpinTicks = 200000000LL;

constexpr int kThrottleEvery = 128;
constexpr size_t kThrottleDepth = 4;

// Default steps per launch for K cells per lane: the tile keeps 75% of its
// cells as output (U = 64K - 2S = 48K).
inline int default_S(int K) { return 8 * K; }
