// Library handle: tuning knobs, workspaces, the lookahead side stream and GEMM timing records.
//
// Every host thread works through the handle bound to it (mnp_handle_bind); a thread that binds
// none gets its own default handle on first use. Two threads therefore never share workspaces,
// barrier counters or knobs unless they bind the same handle. Workspaces grow on demand and are
// kept until mnp_handle_destroy or mnp_release.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace mnp {

// One timed matrix-core GEMM launch (HIP events on its stream).
struct ProfRec {
    hipEvent_t a, b;
    double flops, bytes;
    int plane;                              // 1: the split-BF16 plane kernel
};

// Plane workspace of one stream (the LU runs GEMMs on two streams at once).
struct PlaneWs {
    hipStream_t s;
    void* p;
    size_t bytes;
};

struct Ctx {
    // LU / GEMM knobs (mnp_set_tuning keys 0-6)
    int nbo;               // 0: outer block width of the blocked LU and right-side solves
    bool lookahead;        // 1: trailing update split over two streams
    bool panel_reg;        // 2: register-resident complex-float panel
    int coop_min_rows;     // 3: panel height from which the multi-workgroup panel is used
    int gemm_split;        // 4: complex-float GEMM mode
    int plane_tile;        // 5: plane GEMM tile (11: 64 x 64, 22: 128 x 128 with direct LDS copies)
    int plane_group;       // 6: row tiles per group of the plane GEMM's XCD-aware order
    int asm_rows;          // 7: rows per workgroup of the Green-function assembly
    int asm_lchunk;        // 8: wavelengths per assembly thread (0: all)
    int tall_panel;        // 9: tall panels by tournament pivoting (1) or cooperative partial pivoting (0)
    int lu_recursive;      // 10: outer panels factored by recursive halving (1) or panel by panel (0)
    int f16_fused;         // 11: fp16-pair operands scaled and split in one pass (1) or two (0)
    // cooperative-panel scratch (candidates, diagonal rows, barrier counters)
    void* coop_ws = nullptr;
    size_t coop_bytes = 0;
    // split-BF16 plane workspaces, one per stream
    std::vector<PlaneWs> plane_ws;
    // lookahead side stream (default priority) and its two events
    hipStream_t side = nullptr;
    hipEvent_t side_ready = nullptr, side_done = nullptr;
    // GEMM timing
    bool prof_on = false;
    std::vector<ProfRec> prof;
    ProfRec prof_pending{};
    std::vector<hipEvent_t> pool;

    Ctx();
    // Waits for the handle's streams and frees its device memory, stream and events; the
    // handle stays usable (they are created again on demand). False when it held none (no HIP
    // call was made).
    bool release();
    hipEvent_t event();
};

// The calling thread's handle.
Ctx& ctx();
Ctx* ctx_create();
// Destroys h (after release()); a thread bound to it falls back to its default handle.
void ctx_destroy(Ctx* h);
void ctx_bind(Ctx* h);
bool ctx_valid(Ctx* h);
// release() of every handle the library made; false when none held anything.
bool ctx_release_all();

}  // namespace mnp
