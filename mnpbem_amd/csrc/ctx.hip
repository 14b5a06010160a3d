// Library handles (host code only): knobs, workspaces and streams per handle; a registry of
// every handle for mnp_release.
#include "ctx.hpp"

#include <stdlib.h>

#include <algorithm>
#include <mutex>

namespace mnp {

static int env_int(const char* name, int dflt) {
    const char* e = getenv(name);
    return e ? atoi(e) : dflt;
}

Ctx::Ctx()
    : nbo(env_int("MNP_LU_NBO", 128)),
      lookahead(env_int("MNP_LU_LOOKAHEAD", 1) != 0),
      panel_reg(env_int("MNP_PANEL_REG", 1) != 0),
      coop_min_rows(env_int("MNP_PANEL_COOP_MIN_ROWS", 4096)),
      gemm_split(env_int("MNP_GEMM_SPLIT", 0)),
      plane_tile(env_int("MNP_GEMM_TILE", 11)),
      plane_group(env_int("MNP_GEMM_GROUP", 8)),
      asm_rows(env_int("MNP_ASM_ROWS", 64)),
      asm_lchunk(env_int("MNP_ASM_LCHUNK", 0)),
      tall_panel(env_int("MNP_TALL_PANEL", 1)),
      lu_recursive(env_int("MNP_LU_RECURSIVE", 1)),
      f16_fused(env_int("MNP_F16_FUSED", 1)) {}

hipEvent_t Ctx::event() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

bool Ctx::release() {
    const bool any = side || coop_ws || !plane_ws.empty() || !prof.empty() || !pool.empty();
    if (!any) return false;
    // the side stream and every plane-workspace stream may still run work that uses the buffers
    if (side) hipStreamSynchronize(side);
    for (auto& w : plane_ws) hipStreamSynchronize(w.s);
    if (coop_ws || !plane_ws.empty()) hipDeviceSynchronize();
    for (auto& w : plane_ws)
        if (w.p) hipFree(w.p);
    plane_ws.clear();
    if (coop_ws) hipFree(coop_ws);
    coop_ws = nullptr;
    coop_bytes = 0;
    for (auto& r : prof) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    prof.clear();
    for (auto e : pool) hipEventDestroy(e);
    pool.clear();
    if (side_ready) hipEventDestroy(side_ready);
    if (side_done) hipEventDestroy(side_done);
    if (side) hipStreamDestroy(side);
    side = nullptr;
    side_ready = side_done = nullptr;
    return true;
}

static std::mutex g_reg_mu;
static std::vector<Ctx*>& registry() {
    static std::vector<Ctx*>* r = new std::vector<Ctx*>();   // outlives static destructors
    return *r;
}

static thread_local Ctx* t_bound = nullptr;
static thread_local Ctx* t_default = nullptr;

Ctx* ctx_create() {
    Ctx* h = new Ctx();
    std::lock_guard<std::mutex> lk(g_reg_mu);
    registry().push_back(h);
    return h;
}

bool ctx_valid(Ctx* h) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto& r = registry();
    return std::find(r.begin(), r.end(), h) != r.end();
}

Ctx& ctx() {
    if (t_bound) return *t_bound;
    if (!t_default) t_default = ctx_create();
    return *t_default;
}

void ctx_bind(Ctx* h) { t_bound = h; }

void ctx_destroy(Ctx* h) {
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        auto& r = registry();
        r.erase(std::remove(r.begin(), r.end(), h), r.end());
    }
    h->release();
    if (t_bound == h) t_bound = nullptr;
    if (t_default == h) t_default = nullptr;
    delete h;
}

bool ctx_release_all() {
    std::vector<Ctx*> all;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        all = registry();
    }
    bool any = false;
    for (Ctx* h : all) any |= h->release();
    return any;
}

}  // namespace mnp
