// frep_api.hip -- host driver behind the C ABI of include/frep.h.
//
// Owns the HIP stream, device buffers and handle tables.  A voxelization is
// one init kernel plus one persistent kernel per tree level, enqueued back to
// back; the host reads the counter block once at the end.  If a level ran
// out of node or tape-pool capacity the kernels flag it (and later levels
// return at once), the host grows the buffers to the sizes the counters
// requested and runs again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/frep.h"
#include "frep_asm.hpp"
#include "frep_host.hpp"
#include "frep_internal.hpp"

using namespace frep;

// Output staging of frep_render_noalloc: a fixed array in the library's code
// object, one copy per GPU, held until the process ends.  In host code
// `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress (scratch_address), which is what kernels, copies and
// fills receive.  Every allocation-free render on a GPU shares that GPU's
// copy, so such calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[FREP_NOALLOC_STAGING_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up once per
// device.
hipError_t scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

namespace frep {

uint32_t maxs_for(uint32_t slot_count) {
    uint32_t m = 8;
    while (m < slot_count) m *= 2;
    return m;
}

// Values: 0..3 = x, y, z, t; 4 + i = result of instruction i.  Instructions
// the result does not depend on are left out.  Each instruction is placed
// as late as its readers allow (level = longest chain to the result counted
// back from the deepest), which keeps values short-lived; a level of more
// than lpc instructions takes several passes.  Value slots are reused by
// values first written after the previous value's last read; constants
// come from a table instead of slots.
bool build_root_schedule(const uint32_t* w, uint32_t lpc, size_t lds_per_block,
                         std::vector<uint32_t>* out, RootSched* rs) {
    const uint32_t n = w[0];
    if (n == 0 || n > kRootMaxInstr) return false;
    int32_t cur[256];
    for (uint32_t s = 0; s < 256; ++s) cur[s] = s < 4 ? (int32_t)s : -1;
    std::vector<uint32_t> va(n, 0), vb(n, 0);
    std::vector<uint8_t> is_const(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t x = w[1 + i];
        const uint32_t op = x & 0xFFu, o = (x >> 8) & 0xFFu, a = (x >> 16) & 0xFFu, b = x >> 24;
        if (op == OP_CONST) {
            is_const[i] = 1;
        } else {
            const bool binary = op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
                                op == OP_MIN || op == OP_MAX;
            if (cur[a] < 0 || (binary && cur[b] < 0)) return false;
            va[i] = (uint32_t)cur[a];
            vb[i] = binary ? (uint32_t)cur[b] : va[i];
        }
        cur[o] = (int32_t)(4 + i);
    }
    const uint32_t res = (uint32_t)cur[0];
    auto is_op = [&](uint32_t v) { return v >= 4 && !is_const[v - 4]; };
    // instructions the result depends on (operands precede their readers)
    std::vector<uint8_t> need(n, 0);
    if (res >= 4) need[res - 4] = 1;
    for (uint32_t i = n; i-- > 0;)
        if (need[i] && !is_const[i]) {
            if (va[i] >= 4) need[va[i] - 4] = 1;
            if (vb[i] >= 4) need[vb[i] - 4] = 1;
        }
    std::vector<uint32_t> depth(n, 0), height(n, 0);
    uint32_t D = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (need[i] && !is_const[i]) {
            const uint32_t da = is_op(va[i]) ? depth[va[i] - 4] : 0u;
            const uint32_t db = is_op(vb[i]) ? depth[vb[i] - 4] : 0u;
            depth[i] = 1 + std::max(da, db);
            D = std::max(D, depth[i]);
        }
    for (uint32_t i = n; i-- > 0;)
        if (need[i] && !is_const[i])
            for (const uint32_t v : {va[i], vb[i]})
                if (is_op(v)) height[v - 4] = std::max(height[v - 4], height[i] + 1);
    std::vector<std::vector<uint32_t>> by_level(D + 1);
    for (uint32_t i = 0; i < n; ++i)
        if (need[i] && !is_const[i]) by_level[D - height[i]].push_back(i);
    std::vector<std::vector<uint32_t>> passes;
    for (uint32_t L = 1; L <= D; ++L)
        for (size_t k = 0; k < by_level[L].size(); k += lpc)
            passes.emplace_back(by_level[L].begin() + k,
                                by_level[L].begin() + std::min(by_level[L].size(), k + lpc));
    if (2 * passes.size() > n) return false;
    while (passes.size() % 4) passes.emplace_back();   // empty passes: the kernel runs 4 at a time
    const uint32_t n_pass = (uint32_t)passes.size();
    // last pass reading each value; the result is read after the last pass
    std::vector<int32_t> last(n, -1);
    for (uint32_t p = 0; p < n_pass; ++p)
        for (const uint32_t i : passes[p])
            for (const uint32_t v : {va[i], vb[i]})
                if (is_op(v)) last[v - 4] = std::max(last[v - 4], (int32_t)p);
    if (is_op(res)) last[res - 4] = (int32_t)n_pass;
    std::vector<uint32_t> slot(n, 0), free_slots;
    std::vector<std::pair<int32_t, uint32_t>> active, still;   // (last read pass, slot)
    uint32_t n_slot = 4;
    for (uint32_t p = 0; p < n_pass; ++p) {
        still.clear();
        for (const auto& ls : active) {
            if (ls.first < (int32_t)p) free_slots.push_back(ls.second);
            else still.push_back(ls);
        }
        active.swap(still);
        for (const uint32_t i : passes[p]) {
            uint32_t sl;
            if (!free_slots.empty()) {
                sl = free_slots.back();
                free_slots.pop_back();
            } else {
                sl = n_slot++;
            }
            slot[i] = sl;
            active.emplace_back(std::max(last[i], (int32_t)p), sl);
        }
    }
    if (n_slot > kRootMaxSlots + 1) return false;
    std::vector<uint32_t> kidx(n, 0), consts;
    for (uint32_t i = 0; i < n; ++i)
        if (need[i] && is_const[i]) {
            kidx[i] = (uint32_t)consts.size();
            consts.push_back(i | (((w[1 + i] >> 16) & 0xFFFFu) << 16));
        }
    if (consts.size() >= kRootConst) return false;
    auto enc = [&](uint32_t v) -> uint32_t {
        if (v < 4) return v;
        return is_const[v - 4] ? (kRootConst | kidx[v - 4]) : slot[v - 4];
    };
    const uint32_t nk = (uint32_t)consts.size();
    const uint32_t table_bytes = (10 * nk + 15u) & ~15u;
    const uint32_t cell_bytes = (8 * n_slot + ((n + 3u) & ~3u) + n_slot + 15u) & ~15u;
    const uint32_t cpb = root_cells_per_block(table_bytes, cell_bytes, lds_per_block);
    if (cpb == 0) return false;
    out->clear();
    rs->c0 = 0;
    out->insert(out->end(), consts.begin(), consts.end());
    rs->m0 = (uint32_t)out->size();
    for (const auto& ps : passes) {
        uint32_t m = 0;
        for (const uint32_t i : ps) m |= 1u << (w[1 + i] & 0xFu);
        out->push_back(m);
    }
    if (out->size() & 1u) out->push_back(0);
    rs->e0 = (uint32_t)out->size();
    for (const auto& ps : passes) {
        for (const uint32_t i : ps) {
            out->push_back((w[1 + i] & 0xFu) | (slot[i] << 4) | (i << 16));
            out->push_back(enc(va[i]) | (enc(vb[i]) << 16));
        }
        for (size_t k = ps.size(); k < lpc; ++k) {
            out->push_back(kRootNop);
            out->push_back(0);
        }
    }
    rs->ok = true;
    rs->lpc = lpc;
    rs->pass = n_pass;
    rs->nconst = nk;
    rs->res = enc(res);
    rs->slots = n_slot;
    rs->table_bytes = table_bytes;
    rs->cell_bytes = cell_bytes;
    rs->cpb = cpb;
    return true;
}

RootArgs root_args(const TapeObj* tp, const TreeObj* t, const RootSched& rs, uint32_t mm_stride,
                   uint32_t cells_per_block) {
    RootArgs r{};
    r.a = level_args(t, 0, mm_stride);
    r.sched = tp->d_sched + rs.p0;
    r.tape = tp->d_words;
    r.n = tp->words[0];
    r.n_pass = rs.pass;
    r.n_const = rs.nconst;
    r.c0 = rs.c0;
    r.m0 = rs.m0;
    r.e0 = rs.e0;
    r.n_slot = rs.slots;
    r.res = rs.res;
    r.table_bytes = rs.table_bytes;
    r.cell_bytes = rs.cell_bytes;
    r.cells_per_block = cells_per_block;
    r.lanes_per_cell = rs.lpc;
    return r;
}

// A pass is bound by one wavefront's latency, and about 4 wavefronts per
// SIMD overlap it for free: cost = passes x rounds of 4 wavefronts per SIMD.
double root_cost(const RootSched& rs, uint32_t cells, uint32_t cu_count) {
    const double slots = 16.0 * std::max<uint32_t>(cu_count, 1);
    const double waves = (double)cells * rs.lpc / 64.0;
    return (double)rs.pass * std::max(1.0, waves / slots);
}

int pick_root_schedule(const TapeObj* tp, uint32_t cells, uint32_t cu_count) {
    int best = -1;
    double best_cost = 0.0;
    for (uint32_t v = 0; v < kRootVariants; ++v) {
        const RootSched& rs = tp->root[v];
        if (!rs.ok) continue;
        const double cost = root_cost(rs, cells, cu_count);
        if (best < 0 || cost < best_cost) {
            best = (int)v;
            best_cost = cost;
        }
    }
    return best;
}

frep_status fail(frep_ctx* c, frep_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

TapeObj* get_tape(frep_ctx* ctx, frep_tape_t h) {
    if (h == 0 || h > ctx->tapes.size()) return nullptr;
    return ctx->tapes[h - 1].get();
}

TreeObj* get_tree(frep_ctx* ctx, frep_tree_t h) {
    if (h == 0 || h > ctx->trees.size()) return nullptr;
    return ctx->trees[h - 1].get();
}

// Returns 0 when the tape is well formed (same checks as the oracle's
// orc_tape_validate), otherwise a code naming the first problem.
int validate_tape(const uint32_t* w, uint32_t n_words, uint32_t n_consts, uint32_t slots,
                  uint32_t format) {
    if (n_words < 1) return 1;
    const uint32_t n = w[0];
    if ((uint64_t)n + 1 > n_words) return 2;
    if (slots < 4 || slots > 256) return 3;
    if (format == FREP_TAPE_FORMAT_REF && n_consts > 256) return 9;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t x = w[1 + k];
        const uint32_t op = x & 0xFFu, o = (x >> 8) & 0xFFu, a = (x >> 16) & 0xFFu, b = x >> 24;
        if (op > OP_COPY) return 4;
        if (o >= slots) return 5;
        if (op == OP_CONST) {
            if ((a | (b << 8)) >= n_consts) return 6;
            continue;
        }
        if (a >= slots) return 7;
        const bool binary = op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
                            op == OP_MIN || op == OP_MAX;
        if (binary && b >= slots) return 8;
    }
    return 0;
}

frep_status check_grid(frep_ctx* ctx, const frep_grid* g) {
    if (!g) return fail(ctx, FREP_INVALID_ARG, "grid is null");
    if (g->n_levels < 1 || g->n_levels > kMaxLevels)
        return fail(ctx, FREP_INVALID_ARG, "n_levels must be in 1..8");
    uint64_t fine = 1;
    for (uint32_t L = 0; L < g->n_levels; ++L) {
        const uint32_t d = g->dims[L];
        if (d < 4 || d > 32 || (d * d * d) % 64 != 0)
            return fail(ctx, FREP_UNSUPPORTED, "dims must satisfy 4 <= d <= 32, d^3 % 64 == 0");
        fine *= d;
    }
    if (fine > (1u << 20))
        return fail(ctx, FREP_UNSUPPORTED, "finest resolution above 2^20 per axis");
    if (!(g->world_size > 0.0f) || !std::isfinite(g->world_size))
        return fail(ctx, FREP_INVALID_ARG, "world_size must be positive and finite");
    return FREP_OK;
}

// Tree buffers come from the context's device pool; the callers have
// synchronized with every launch that used them.
void free_tree_buffers(frep_ctx* ctx, TreeObj* t) {
    for (uint32_t L = 0; L < kMaxLevels; ++L) {
        ctx->dpool.free(t->d_nodes[L]);
        t->d_nodes[L] = nullptr;
        t->node_cap[L] = 0;
    }
    ctx->dpool.free(t->d_pool);
    t->d_pool = nullptr;
    t->pool_cap = 0;
    ctx->dpool.free(t->d_consts);
    t->d_consts = nullptr;
    t->consts_cap = 0;
    t->consts_uid = 0;
    if (t->vexec) (void)hipGraphExecDestroy(t->vexec);
    t->vexec = nullptr;
    t->vkey = 0;
}

frep_status ensure_level(frep_ctx* ctx, TreeObj* t, uint32_t L, uint32_t cap) {
    if (t->d_nodes[L] && t->node_cap[L] >= cap) return FREP_OK;
    ++t->dev_gen;
    ctx->dpool.free(t->d_nodes[L]);
    t->d_nodes[L] = nullptr;
    t->node_cap[L] = 0;
    void* p = nullptr;
    HIP_TRY(ctx->dpool.alloc((size_t)cap * t->ns[L] * sizeof(uint32_t), &p));
    t->d_nodes[L] = static_cast<uint32_t*>(p);
    t->node_cap[L] = cap;
    return FREP_OK;
}

frep_status ensure_pool(frep_ctx* ctx, TreeObj* t, uint64_t words) {
    if (t->d_pool && t->pool_cap >= words) return FREP_OK;
    ++t->dev_gen;
    ctx->dpool.free(t->d_pool);
    t->d_pool = nullptr;
    t->pool_cap = 0;
    void* p = nullptr;
    HIP_TRY(ctx->dpool.alloc((size_t)(words + kPoolPad) * sizeof(uint32_t), &p));
    t->d_pool = static_cast<uint32_t*>(p);
    t->pool_cap = words;
    return FREP_OK;
}

uint32_t clamp_cap(uint64_t v, uint64_t bound) {
    v = std::min(v, bound);
    v = std::min<uint64_t>(v, kMaxLevelNodes);
    return (uint32_t)std::max<uint64_t>(v, 1);
}

uint64_t host_alloc(const TreeObj* t, uint32_t L) {
    uint64_t v;
    std::memcpy(&v, t->h_counters + CNT_ALLOC + 2 * L, 8);
    return v;
}

// Node count of level L from the counter mirror (level 0 is the root).
uint32_t host_nodes(const TreeObj* t, uint32_t L) {
    return L == 0 ? 1u : alloc_nodes(host_alloc(t, L));
}

// First pool word past the tapes of every level.
uint64_t host_pool_top(const TreeObj* t) {
    uint64_t top = t->h_counters[CNT_BASE1];
    for (uint32_t L = 1; L < t->grid.n_levels; ++L) top += alloc_words(host_alloc(t, L));
    return top;
}

uint64_t tree_device_bytes(const TreeObj* t) {
    uint64_t bytes = 0;
    for (uint32_t L = 0; L < kMaxLevels; ++L)
        if (t->d_nodes[L]) bytes += (uint64_t)t->node_cap[L] * t->ns[L] * 4;
    if (t->d_pool) bytes += (t->pool_cap + kPoolPad) * 4;
    return bytes + (uint64_t)t->consts_cap * 4;
}

frep_status new_tree(frep_ctx* ctx, frep_tree_t* out) {
    std::unique_ptr<TreeObj> t(new TreeObj());
    void* d = nullptr;
    void* h = nullptr;
    HIP_TRY(ctx->dpool.alloc(CNT_WORDS * 4, &d));
    HIP_TRY(ctx->hpool.alloc(CNT_WORDS * 4, &h));
    t->d_counters = static_cast<uint32_t*>(d);
    t->h_counters = static_cast<uint32_t*>(h);
    std::memset(t->h_counters, 0, CNT_WORDS * 4);
    ctx->trees.push_back(std::move(t));
    *out = (frep_tree_t)ctx->trees.size();
    return FREP_OK;
}

}  // namespace frep

extern "C" {

frep_status frep_init(int device, frep_ctx** out) {
    if (!out) return FREP_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FREP_DEVICE;
    if (device < 0 || device >= n) return FREP_INVALID_ARG;
    std::unique_ptr<frep_ctx> c(new frep_ctx());
    frep_ctx* ctx = c.get();
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->cu_count = (uint32_t)prop.multiProcessorCount;
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
                hipSuccess && lds > 0)
            c->lds_per_block = std::min<size_t>((size_t)lds, 160 * 1024);
    }
    // persistent level kernels: 8 blocks of 4 wavefronts per CU
    c->level_grid = c->cu_count * 8;
    const char* na = std::getenv("FREP_NO_ASM");
    c->no_asm = na && na[0] == '1';
    const char* nt = std::getenv("FREP_NO_AUTOTUNE");
    c->autotune = !(nt && nt[0] == '1');
    // plain builds launch directly unless FREP_GRAPH=1: with the init and
    // counter-copy launches gone, three direct launches submit faster than a
    // graph replay of them
    const char* gr = std::getenv("FREP_GRAPH");
    c->graphs = gr && gr[0] == '1';
    const char* ro = std::getenv("FREP_ROOT_OPS");
    c->root_ops = !(ro && ro[0] == '0');
    c->root_ops_required = ro && ro[0] == '2';
    // the plain last level's wavefronts lower their issue priority as they
    // progress, so a SIMD's wavefronts finish together
    const char* pb = std::getenv("FREP_POINT_BALANCE");
    if (pb) {
        const int v = std::atoi(pb);
        if (v >= 0 && v <= 64) c->point_balance = (uint32_t)v;
    }
    if (const char* bl = std::getenv("FREP_BATCH_LANES")) {
        const int v = std::atoi(bl);
        if (v >= 1 && v <= (int)frep::kBatchLanes) c->batch_lanes = (uint32_t)v;
    }
    const char* rb = std::getenv("FREP_ROOT_BALANCE");
    c->root_balance = !(rb && rb[0] == '0');
    HIP_TRY(scratch_address(device, &c->noalloc_staging));
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_TRY(hipEventCreate(&e));
    for (auto& st : c->lanes.st) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->lanes.fork, hipEventDisableTiming));
    for (auto& e : c->lanes.join) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *out = c.release();
    return FREP_OK;
}

void frep_shutdown(frep_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& t : ctx->tapes) {
        if (!t) continue;
        if (t->d_words) (void)hipFree(t->d_words);
        if (t->d_consts) (void)hipFree(t->d_consts);
        if (t->d_sched) (void)hipFree(t->d_sched);
    }
    for (auto& t : ctx->trees) {
        if (!t) continue;
        free_tree_buffers(ctx, t.get());
        ctx->dpool.free(t->d_counters);
        ctx->hpool.free(t->h_counters);
    }
    if (ctx->d_out) (void)hipFree(ctx->d_out);
    for (auto& tg : ctx->targets)
        if (tg) ctx->dpool.free(tg->d);
    free_stage(&ctx->stage);
    for (auto& kv : ctx->tune)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto& e : ctx->bev) (void)hipEventDestroy(e);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& st : ctx->lanes.st)
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    if (ctx->lanes.fork) (void)hipEventDestroy(ctx->lanes.fork);
    for (auto& e : ctx->lanes.join)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* frep_last_error(const frep_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

frep_status frep_set_timing(frep_ctx* ctx, int enable) {
    if (!ctx) return FREP_INVALID_ARG;
    ctx->timing = enable != 0;
    return FREP_OK;
}

frep_status frep_get_timing(frep_ctx* ctx, frep_timing* out) {
    if (!ctx || !out) return FREP_INVALID_ARG;
    *out = ctx->last;
    return FREP_OK;
}

frep_status frep_device_info(frep_ctx* ctx, uint32_t* cu_count, uint32_t* level_grid) {
    if (!ctx) return FREP_INVALID_ARG;
    if (cu_count) *cu_count = ctx->cu_count;
    if (level_grid) *level_grid = ctx->level_grid;
    return FREP_OK;
}

frep_status frep_tape_upload(frep_ctx* ctx, const uint32_t* words, uint32_t n_words,
                             const float* consts, uint32_t n_consts, uint32_t slot_count,
                             uint32_t format, frep_tape_t* out) {
    if (!ctx || !words || !out) return fail(ctx, FREP_INVALID_ARG, "null argument");
    if (n_consts && !consts) return fail(ctx, FREP_INVALID_ARG, "consts is null");
    if (format != FREP_TAPE_FORMAT_REF && format != FREP_TAPE_FORMAT_WIDE)
        return fail(ctx, FREP_INVALID_ARG, "unknown tape format");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int v = validate_tape(words, n_words, n_consts, slot_count, format);
    if (v) return fail(ctx, FREP_BAD_TAPE, "malformed tape (code " + std::to_string(v) + ")");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<TapeObj> t(new TapeObj());
    t->uid = ctx->next_uid++;
    const uint32_t n = words[0];
    t->words.assign(words, words + n + 1);
    t->consts.assign(consts, consts + n_consts);
    t->slot_count = slot_count;
    t->maxs = maxs_for(slot_count);
    t->asm_ok = true;
    // does any instruction read slot 3 while it still holds the input t?
    // Shortening keeps the instruction that last wrote each operand a kept
    // instruction reads, so a shortened tape reads t only if its root does.
    bool t_written = false;
    t->reads_t = false;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t w = words[1 + k], op = w & 0xFFu;
        t->n_mm += (op == OP_MIN || op == OP_MAX);
        t->asm_ok = t->asm_ok && asm_op_ok(op);
        if (op != OP_CONST && !t_written) {
            const bool binary = op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
                                op == OP_MIN || op == OP_MAX;
            if (((w >> 16) & 0xFFu) == 3u || (binary && (w >> 24) == 3u)) t->reads_t = true;
        }
        if (((w >> 8) & 0xFFu) == 3u) t_written = true;
    }
    // padded device copies: scalar loads may read a few words past the end
    HIP_TRY(hipMalloc(&t->d_words, (size_t)(n + 1 + kPoolPad) * 4));
    HIP_TRY(hipMemset(t->d_words, 0, (size_t)(n + 1 + kPoolPad) * 4));
    HIP_TRY(hipMemcpy(t->d_words, words, (size_t)(n + 1) * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&t->d_consts, (size_t)(n_consts + kPoolPad) * 4));
    if (n_consts)
        HIP_TRY(hipMemcpy(t->d_consts, consts, (size_t)n_consts * 4, hipMemcpyHostToDevice));
    {
        std::vector<uint32_t> all, one;
        bool any = false;
        for (uint32_t v = 0; v < kRootVariants; ++v) {
            if (!build_root_schedule(words, kRootLanes[v], ctx->lds_per_block, &one, &t->root[v]))
                continue;
            any = true;
            if (all.size() & 1u) all.push_back(0);     // entries stay 8-byte aligned
            t->root[v].p0 = (uint32_t)all.size();
            all.insert(all.end(), one.begin(), one.end());
        }
        if (any) {
            if (all.empty()) all.push_back(0);         // no passes: the result is an input

            HIP_TRY(hipMalloc(&t->d_sched, all.size() * 4));
            HIP_TRY(hipMemcpy(t->d_sched, all.data(), all.size() * 4, hipMemcpyHostToDevice));
        }
    }
    ctx->tapes.push_back(std::move(t));
    *out = (frep_tape_t)ctx->tapes.size();
    return FREP_OK;
}

frep_status frep_tape_free(frep_ctx* ctx, frep_tape_t h) {
    if (!ctx) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TapeObj* t = get_tape(ctx, h);
    if (!t) return fail(ctx, FREP_INVALID_ARG, "unknown tape handle");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t->d_words) HIP_TRY(hipFree(t->d_words));
    if (t->d_consts) HIP_TRY(hipFree(t->d_consts));
    if (t->d_sched) HIP_TRY(hipFree(t->d_sched));
    ctx->tapes[h - 1].reset();
    return FREP_OK;
}

frep_status frep_tree_create(frep_ctx* ctx, frep_tree_t* out) {
    if (!ctx || !out) return fail(ctx, FREP_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return new_tree(ctx, out);
}

frep_status frep_tree_free(frep_ctx* ctx, frep_tree_t h) {
    if (!ctx) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = get_tree(ctx, h);
    if (!t) return fail(ctx, FREP_INVALID_ARG, "unknown tree handle");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    free_tree_buffers(ctx, t);
    ctx->dpool.free(t->d_counters);
    ctx->hpool.free(t->h_counters);
    ctx->trees[h - 1].reset();
    return FREP_OK;
}

}  // extern "C"

namespace frep {

// Sets up tree t for tape tp over grid: record sizes, cell sizes, first
// capacity guesses, the tree's copy of the constant pool (enqueued).
// First capacities of a voxelization: what the tree already holds, the
// hint, or a default guess, each level capped by the most nodes it can have
// (n_{L+1} <= n_L * dim_L^3).
void initial_caps(const TapeObj* tp, const frep_grid* grid, const TreeObj* t, const CapHint* hint,
                  uint32_t* want, uint64_t* pool_want) {
    uint64_t bound = 1;
    for (uint32_t L = 0; L < grid->n_levels; ++L) {
        uint64_t guess = 1u << 16;
        if (hint && hint->nodes[L]) guess = (uint64_t)hint->nodes[L] + hint->nodes[L] / 4 + 64;
        if (t) guess = std::max<uint64_t>(guess, t->node_cap[L]);
        want[L] = clamp_cap(L == 0 ? 1 : guess, bound);
        const uint64_t cells = (uint64_t)grid->dims[L] * grid->dims[L] * grid->dims[L];
        bound = std::min<uint64_t>(bound * cells, 0x7FFFFFFFull);
    }
    uint64_t pw = (uint64_t)kRootOff + (tp->words[0] + 1) * 64ull + (1u << 20);
    if (hint && hint->pool_words) pw = std::max<uint64_t>(pw, hint->pool_words + hint->pool_words / 4);
    *pool_want = std::max<uint64_t>(t ? t->pool_cap : 0, pw);
}

uint64_t estimate_job_bytes(const TapeObj* tp, const frep_grid* grid, const CapHint* hint) {
    uint32_t want[kMaxLevels];
    uint64_t pool_want = 0;
    initial_caps(tp, grid, nullptr, hint, want, &pool_want);
    uint64_t bytes = (pool_want + kPoolPad + tp->consts.size() + kPoolPad) * 4;
    for (uint32_t L = 0; L < grid->n_levels; ++L)
        bytes += (uint64_t)want[L] * node_words(grid->dims[L], L + 1 == grid->n_levels) * 4;
    return bytes;
}

frep_status prepare_job(frep_ctx* ctx, TapeObj* tp, TreeObj* t, const frep_grid* grid, Job* j,
                        hipStream_t st, const CapHint* hint) {
    const uint32_t nl = grid->n_levels;
    const bool same_shape = t->built && t->grid.n_levels == nl &&
                            std::equal(grid->dims, grid->dims + nl, t->grid.dims);
    if (!same_shape) {
        // record sizes change with the dims: drop the level buffers
        ++t->dev_gen;
        for (uint32_t L = 0; L < kMaxLevels; ++L) {
            ctx->dpool.free(t->d_nodes[L]);
            t->d_nodes[L] = nullptr;
            t->node_cap[L] = 0;
        }
    }
    t->built = false;
    t->grid = *grid;
    t->slot_count = tp->slot_count;
    t->maxs = tp->maxs;
    t->asm_ok = tp->asm_ok && !ctx->no_asm;
    t->reads_t = tp->reads_t;
    // cell sizes by sequential float division (voxelizer.cpp:210-216)
    float c = grid->world_size;
    for (uint32_t L = 0; L < nl; ++L) {
        c /= (float)grid->dims[L];
        t->cs[L] = c;
        t->ns[L] = node_words(grid->dims[L], L + 1 == nl);
    }
    initial_caps(tp, grid, t, hint, j->want, &j->pool_want);
    const uint32_t nc = (uint32_t)tp->consts.size();
    if (!t->d_consts || t->consts_cap < nc + kPoolPad) {
        ctx->dpool.free(t->d_consts);
        t->d_consts = nullptr;
        t->consts_uid = 0;
        void* p = nullptr;
        HIP_TRY(ctx->dpool.alloc((size_t)(nc + kPoolPad) * 4, &p));
        t->d_consts = static_cast<float*>(p);
        t->consts_cap = nc + kPoolPad;
    }
    // the copy is enqueued once per (upload, buffer): it is complete before
    // any later work on the tree, which is ordered after this call's work
    if (nc && t->consts_uid != tp->uid)
        HIP_TRY(hipMemcpyAsync(t->d_consts, tp->d_consts, (size_t)nc * 4, hipMemcpyDeviceToDevice,
                               st));
    t->consts_uid = tp->uid;
    t->n_consts = nc;
    // min/max choice rows live in the interval kernel's LDS
    j->mm_stride = std::max<uint32_t>(tp->n_mm, 1);
    if (nl > 1 && j->mm_stride > max_minmax(ctx->lds_per_block))
        return fail(ctx, FREP_UNSUPPORTED,
                    "tape has more than " + std::to_string(max_minmax(ctx->lds_per_block)) +
                        " MIN/MAX instructions");
    t->retries = 0;
    j->tp = tp;
    j->t = t;
    return FREP_OK;
}

frep_status ensure_job(frep_ctx* ctx, Job* j) {
    for (uint32_t L = 0; L < j->t->grid.n_levels; ++L) {
        const frep_status s = ensure_level(ctx, j->t, L, j->want[L]);
        if (s != FREP_OK) return s;
    }
    return ensure_pool(ctx, j->t, j->pool_want);
}

// Arguments of level L of tree t (pk / use_asm are set by the caller).
LevelArgs level_args(const TreeObj* t, uint32_t L, uint32_t mm_stride) {
    const frep_grid& g = t->grid;
    const bool last = (L + 1 == g.n_levels);
    LevelArgs a{};
    a.nodes = t->d_nodes[L];
    a.child_nodes = last ? nullptr : t->d_nodes[L + 1];
    a.pool = t->d_pool;
    a.consts = t->d_consts;
    a.counters = t->d_counters;
    a.pool_cap = t->pool_cap;
    a.level = L;
    a.dim = g.dims[L];
    a.dim_shift = 0;
    if ((a.dim & (a.dim - 1)) == 0)
        while ((1u << a.dim_shift) < a.dim) ++a.dim_shift;
    a.node_cap = t->node_cap[L];
    a.child_cap = last ? 0 : t->node_cap[L + 1];
    a.ns = t->ns[L];
    a.child_ns = last ? 0 : t->ns[L + 1];
    a.mm_stride = mm_stride;
    a.cs = t->cs[L];
    a.low[0] = g.low[0];
    a.low[1] = g.low[1];
    a.low[2] = g.low[2];
    a.tape_time = g.tape_time;
    a.pk = 1;
    a.use_asm = 0;
    a.reads_t = t->reads_t ? 1u : 0u;
    return a;
}

// After a flagged overflow (counters in t->h_counters): grow what the
// counters asked for.  Fails when nothing can grow.
frep_status regrow_job(frep_ctx* ctx, Job* j, uint32_t attempt) {
    TreeObj* t = j->t;
    const uint32_t nl = t->grid.n_levels;
    if (attempt > nl + 4) return fail(ctx, FREP_OUT_OF_MEMORY, "tree did not fit after regrowth");
    ++t->retries;
    bool grew = false;
    for (uint32_t L = 1; L < nl; ++L) {
        const uint32_t req = host_nodes(t, L);
        if (req > t->node_cap[L]) {
            if (req >= kMaxLevelNodes)
                return fail(ctx, FREP_OUT_OF_MEMORY, "more than 2^26 nodes in one level");
            j->want[L] = clamp_cap((uint64_t)req + req / 4 + 64, ~0ull);
            grew = true;
        }
    }
    const uint64_t top = host_pool_top(t);
    if (top > t->pool_cap) {
        j->pool_want = top + top / 4 + 4096;
        grew = true;
    }
    // a flagged overflow with every request within capacity means a packed
    // counter wrapped: the tree cannot be represented
    if (!grew) return fail(ctx, FREP_OUT_OF_MEMORY, "node or tape counter overflow");
    if (j->pool_want > kMaxPoolWords)
        return fail(ctx, FREP_OUT_OF_MEMORY, "tape pool exceeds 2^32 words");
    return FREP_OK;
}

size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// ---------------------------------------------------------------------------
// Batched voxelization driver.
//
// A batch of m trees is cut into groups of kBatchMax.  Per group: one init
// launch, then per level one interval launch over the trees for which the
// level is not the last and one point launch over those for which it is,
// then a gather of the group's counter blocks.  All argument tables go to the
// device with one copy; all counters come back with one copy.
//
// Autotuner: the work of a batch is fixed by the bytes of its argument
// tables, so their hash keys a tuning record.  The first kTuneTrials calls
// with a key time every launch under each candidate persistent-grid size
// (blocks per CU) and keep, per launch, the fastest; later calls run the
// chosen sizes from a captured hipGraph (replayed with one launch).
// ---------------------------------------------------------------------------
// Candidates: blocks per CU (4 wavefronts per block) x cells per lane of
// the interval launches (the tile edge: 64 or 128 cells per wavefront).
constexpr uint32_t kTuneBlocksPerCu[] = {8, 4, 2};
constexpr uint32_t kTuneGrids = sizeof(kTuneBlocksPerCu) / sizeof(kTuneBlocksPerCu[0]);
constexpr uint32_t kTuneTrials = 2 * kTuneGrids;
constexpr size_t kTuneMaxEntries = 64;

// A batched root launch uses one schedule variant for all its trees (one
// block shape): trees with both variants, the variant of least total cost.
bool root_job(const BatchPlan& P, const Job& j) {
    if (!P.root_ops || j.t->grid.n_levels < 2) return false;
    for (uint32_t v = 0; v < kRootVariants; ++v)
        if (!j.tp->root[v].ok) return false;
    return true;
}

// A shape whose last level (level L) runs 8 voxels per lane: a slot file of
// at most 16 and nodes of a multiple of 8 wavefront chunks (its point
// launch then takes 8 voxels per lane for every shape in it).
bool point8_shape(const Job& j, uint32_t L) {
    const uint32_t d = j.t->grid.dims[L];
    return j.t->maxs <= 16 && (d * d * d / 64) % 8 == 0;
}

BatchPlan plan_batch(const std::vector<Job>& jobs, const std::vector<uint32_t>& idx,
                     uint32_t lanes, bool root_ops, uint32_t cu_count) {
    BatchPlan P;
    P.m = (uint32_t)idx.size();
    P.root_ops = root_ops;
    // group boundaries: lane cuts (long tapes first), then chunks of kBatchMax
    std::vector<std::pair<uint32_t, uint32_t>> spans;    // (end, lane)
    if (lanes > 1 && P.m >= 8) {
        const uint32_t b1 = P.m / 8, b2 = P.m * 3 / 8;
        spans = {{b1, 0u}, {b2, 1u}, {P.m, 2u}};
    } else {
        spans = {{P.m, 0u}};
    }
    std::vector<BatchGroup> bounds;
    uint32_t begin = 0;
    for (const auto& sp : spans) {
        for (uint32_t f = begin; f < sp.first; f += kBatchMax) {
            BatchGroup G;
            G.first = f;
            G.count = std::min<uint32_t>(kBatchMax, sp.first - f);
            G.lane = sp.second;
            bounds.push_back(G);
        }
        begin = sp.first;
    }
    for (uint32_t g = 0; g < bounds.size(); ++g) {
        BatchGroup G = bounds[g];
        G.init_off = P.bytes;
        P.bytes = align16(P.bytes + G.count * sizeof(InitArgs));
        uint32_t max_nl = 0;
        for (uint32_t k = 0; k < G.count; ++k)
            max_nl = std::max(max_nl, jobs[idx[G.first + k]].t->grid.n_levels);
        for (uint32_t L = 0; L < max_nl; ++L) {
            if (L == 0) {
                BatchLaunch la{g, 0, false, 0, 8, 1, 0};
                la.root = true;
                la.cpb = 16;
                double cost[kRootVariants] = {};
                for (uint32_t k = 0; k < G.count; ++k) {
                    const Job& j = jobs[idx[G.first + k]];
                    if (!root_job(P, j)) continue;
                    ++la.count;
                    la.mm = std::max(la.mm, j.mm_stride);
                    const uint32_t d = j.t->grid.dims[0];
                    for (uint32_t v = 0; v < kRootVariants; ++v)
                        cost[v] += root_cost(j.tp->root[v], d * d * d, cu_count);
                }
                for (uint32_t v = 1; v < kRootVariants; ++v)
                    if (cost[v] < cost[la.variant]) la.variant = v;
                for (uint32_t k = 0; k < G.count; ++k) {
                    const Job& j = jobs[idx[G.first + k]];
                    if (root_job(P, j)) la.cpb = std::min(la.cpb, j.tp->root[la.variant].cpb);
                }
                if (la.count) {
                    for (uint32_t k = 0; k < G.count; ++k) {
                        const Job& j = jobs[idx[G.first + k]];
                        if (!root_job(P, j)) continue;
                        const uint32_t d = j.t->grid.dims[0];
                        la.blocks += d * d * d / la.cpb;
                        const RootSched& rs = j.tp->root[la.variant];
                        la.lds = std::max(la.lds, (size_t)rs.table_bytes +
                                                      (size_t)la.cpb * rs.cell_bytes);
                    }
                    la.off = P.bytes;
                    P.bytes = align16(P.bytes + la.count * sizeof(RootArgs));
                    la.starts_off = P.bytes;
                    P.bytes = align16(P.bytes + (la.count + 1) * sizeof(uint32_t));
                    P.launches.push_back(la);
                }
            }
            // interval launch, then the point launches: the shapes whose last
            // level runs 8 voxels per lane, then the others (one or four per
            // lane) on a kernel without the 8-voxel paths, whose few registers
            // let many more wavefronts hide the interpreter's latency
            for (int pt = 0; pt < 3; ++pt) {
                BatchLaunch la{g, L, pt >= 1, 0, 8, 1, P.bytes};
                la.pkset = (uint32_t)pt;
                for (uint32_t k = 0; k < G.count; ++k) {
                    const Job& j = jobs[idx[G.first + k]];
                    const uint32_t nl = j.t->grid.n_levels;
                    if (L >= nl || (L + 1 == nl) != (pt >= 1)) continue;
                    if (L == 0 && root_job(P, j)) continue;
                    if (pt >= 1 && point8_shape(j, L) != (pt == 1)) continue;
                    ++la.count;
                    la.maxs = std::max(la.maxs, j.t->maxs);
                    la.slots = std::max(la.slots, j.t->slot_count);
                    la.mm = std::max(la.mm, j.mm_stride);
                }
                if (la.count && !la.point && la.maxs <= 16) {
                    // two cells per lane for the shapes whose nodes allow it
                    bool any = false;
                    for (uint32_t k = 0; k < G.count && !any; ++k) {
                        const Job& j = jobs[idx[G.first + k]];
                        const uint32_t nl = j.t->grid.n_levels;
                        if (L >= nl || L + 1 == nl || (L == 0 && root_job(P, j))) continue;
                        const uint32_t d = j.t->grid.dims[L];
                        any = (d * d * d) % 128u == 0;
                    }
                    la.kc2_ok = any;
                }
                if (la.count && la.point) {
                    // the assembly-only kernel when every shape takes the
                    // assembly interpreter (fill_batch's use_asm rule)
                    bool all_asm = la.maxs <= 16;
                    uint32_t slots = 0;
                    for (uint32_t k = 0; k < G.count && all_asm; ++k) {
                        const Job& j = jobs[idx[G.first + k]];
                        const uint32_t nl = j.t->grid.n_levels;
                        if (L >= nl || L + 1 != nl) continue;
                        if (point8_shape(j, L) != (la.pkset == 1)) continue;
                        all_asm = j.t->asm_ok && point_k_used(la.maxs, j.t->grid.dims[L]) == 8;
                        slots = std::max(slots, j.t->slot_count);
                    }
                    la.asm_slots = all_asm ? slots : 0u;
                }
                if (la.count) {
                    P.bytes = align16(P.bytes + la.count * sizeof(LevelArgs));
                    P.launches.push_back(la);
                }
            }
        }
        P.groups.push_back(G);
    }
    return P;
}

void fill_batch(const BatchPlan& P, const std::vector<Job>& jobs, const std::vector<uint32_t>& idx,
                uint8_t* hs) {
    std::memset(hs, 0, P.bytes);
    for (const BatchGroup& G : P.groups) {
        InitArgs* ia = reinterpret_cast<InitArgs*>(hs + G.init_off);
        for (uint32_t k = 0; k < G.count; ++k) {
            const Job& j = jobs[idx[G.first + k]];
            ++j.t->dev_gen;   // the batched build rewrites the tree's buffers
            ia[k] = InitArgs{j.t->d_counters, j.t->d_nodes[0], j.t->d_pool, j.tp->d_words, kRootOff,
                             0};
        }
    }
    for (const BatchLaunch& la : P.launches) {
        const BatchGroup& G = P.groups[la.group];
        if (la.root) {
            RootArgs* ra = reinterpret_cast<RootArgs*>(hs + la.off);
            uint32_t* starts = reinterpret_cast<uint32_t*>(hs + la.starts_off);
            uint32_t c = 0, b = 0;
            starts[0] = 0;
            for (uint32_t k = 0; k < G.count; ++k) {
                const Job& j = jobs[idx[G.first + k]];
                if (!root_job(P, j)) continue;
                ra[c] = root_args(j.tp, j.t, j.tp->root[la.variant], la.mm, la.cpb);
                const uint32_t d = j.t->grid.dims[0];
                b += d * d * d / la.cpb;
                starts[++c] = b;
            }
            continue;
        }
        LevelArgs* out = reinterpret_cast<LevelArgs*>(hs + la.off);
        uint32_t c = 0;
        for (uint32_t k = 0; k < G.count; ++k) {
            const Job& j = jobs[idx[G.first + k]];
            const uint32_t nl = j.t->grid.n_levels;
            if (la.level >= nl || (la.level + 1 == nl) != la.point) continue;
            if (la.level == 0 && root_job(P, j)) continue;
            if (la.point && point8_shape(j, la.level) != (la.pkset == 1)) continue;
            LevelArgs a = level_args(j.t, la.level, la.mm);
            if (la.point) {
                a.pk = point_k_used(la.maxs, a.dim);
                // 8 voxels per lane in the kernel with the 8-voxel paths, one
                // voxel per lane in the other (la.pkset 2)
                a.use_asm = (j.t->asm_ok && la.maxs <= 16 && a.pk == (la.pkset == 2 ? 1u : 8u))
                                ? 1u
                                : 0u;
            }
            out[c++] = a;
        }
    }
}

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}

uint64_t batch_key(const BatchPlan& P, const uint8_t* hs) {
    uint64_t h = fnv1a(hs, P.bytes);
    for (const BatchLaunch& la : P.launches) {
        const uint32_t v[9] = {la.group, la.level, la.point ? 1u : 0u, la.count, la.maxs, la.mm,
                               la.root ? la.cpb : 0u, la.variant, la.asm_slots};
        h = fnv1a(v, sizeof(v), h);
    }
    return fnv1a(&P.m, sizeof(P.m), h);
}

void drop_graphs(frep_ctx* ctx) {
    for (auto& kv : ctx->tune)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    ctx->tune.clear();
}

frep_status ensure_stage(frep_ctx* ctx, BatchStage* S, const BatchPlan& P) {
    if (S->stage_bytes < P.bytes) {
        if (S->h_stage) HIP_TRY(hipHostFree(S->h_stage));
        if (S->d_stage) HIP_TRY(hipFree(S->d_stage));
        S->h_stage = nullptr;
        S->d_stage = nullptr;
        S->stage_bytes = 0;
        const size_t cap = P.bytes + P.bytes / 2 + 4096;
        HIP_TRY(hipHostMalloc(&S->h_stage, cap, hipHostMallocDefault));
        HIP_TRY(hipMalloc(&S->d_stage, cap));
        S->stage_bytes = cap;
    }
    if (S->cnt_words < (size_t)P.m * CNT_WORDS) {
        if (S->h_cnt) HIP_TRY(hipHostFree(S->h_cnt));
        if (S->d_cnt) HIP_TRY(hipFree(S->d_cnt));
        S->h_cnt = nullptr;
        S->d_cnt = nullptr;
        S->cnt_words = 0;
        const size_t cap = (size_t)std::max<uint32_t>(P.m, kBatchMax) * CNT_WORDS;
        HIP_TRY(hipHostMalloc(&S->h_cnt, cap * 4, hipHostMallocDefault));
        HIP_TRY(hipMalloc(&S->d_cnt, cap * 4));
        S->cnt_words = cap;
    }
    return FREP_OK;
}

void free_stage(BatchStage* S) {
    if (S->h_stage) (void)hipHostFree(S->h_stage);
    if (S->d_stage) (void)hipFree(S->d_stage);
    if (S->h_cnt) (void)hipHostFree(S->h_cnt);
    if (S->d_cnt) (void)hipFree(S->d_cnt);
    *S = BatchStage{};
}

// With ev != nullptr an event is recorded before the first level launch and
// after each launch.
frep_status enqueue_batch(frep_ctx* ctx, BatchStage* S, hipStream_t st, const BatchPlan& P,
                          const std::vector<uint32_t>& grid, hipEvent_t* ev,
                          const LaneSet* lanes, const std::vector<uint32_t>* kcs) {
    uint8_t* ds = static_cast<uint8_t*>(S->d_stage);
    HIP_TRY(hipMemcpyAsync(ds, S->h_stage, P.bytes, hipMemcpyHostToDevice, st));
    bool used[kBatchLanes] = {};
    for (const BatchGroup& G : P.groups) used[G.lane] = true;
    if (!lanes) std::fill(used + 1, used + kBatchLanes, false);
    // fork: the side lanes start after the argument copy
    const bool fork = used[1] || used[2];
    if (fork) {
        HIP_TRY(hipEventRecord(lanes->fork, st));
        for (uint32_t k = 1; k < kBatchLanes; ++k)
            if (used[k]) HIP_TRY(hipStreamWaitEvent(lanes->st[k - 1], lanes->fork, 0));
    }
    size_t li = 0;
    for (uint32_t g = 0; g < P.groups.size(); ++g) {
        const BatchGroup& G = P.groups[g];
        const hipStream_t gs = (lanes && G.lane > 0) ? lanes->st[G.lane - 1] : st;
        const InitArgs* d_init = reinterpret_cast<const InitArgs*>(ds + G.init_off);
        HIP_TRY(launch_init_batch(gs, d_init, G.count));
        if (ev && g == 0) HIP_TRY(hipEventRecord(ev[0], gs));
        for (; li < P.launches.size() && P.launches[li].group == g; ++li) {
            const BatchLaunch& la = P.launches[li];
            BatchArgs b{reinterpret_cast<const LevelArgs*>(ds + la.off), la.count, la.mm};
            b.kc = (la.kc2_ok && kcs && (*kcs)[li] == 2) ? 2u : 1u;
            if (la.root)
                HIP_TRY(launch_root_interval_batch(
                    gs, reinterpret_cast<const RootArgs*>(ds + la.off),
                    reinterpret_cast<const uint32_t*>(ds + la.starts_off), la.count, la.blocks,
                    la.cpb, kRootLanes[la.variant], la.lds));
            else if (la.point)
                HIP_TRY(launch_level_point_batch(gs, b, la.maxs, ctx->cu_count, grid[li],
                                                 la.asm_slots, la.pkset == 2));
            else
                HIP_TRY(launch_level_interval_batch(gs, b, la.maxs, la.slots, ctx->cu_count,
                                                    grid[li]));
            if (ev) HIP_TRY(hipEventRecord(ev[1 + li], gs));
        }
        HIP_TRY(launch_gather_counters(gs, d_init, G.count, S->d_cnt + (size_t)G.first * CNT_WORDS));
    }
    // join: the readback waits for every lane
    if (fork)
        for (uint32_t k = 1; k < kBatchLanes; ++k)
            if (used[k]) {
                HIP_TRY(hipEventRecord(lanes->join[k - 1], lanes->st[k - 1]));
                HIP_TRY(hipStreamWaitEvent(st, lanes->join[k - 1], 0));
            }
    HIP_TRY(hipMemcpyAsync(S->h_cnt, S->d_cnt, (size_t)P.m * CNT_WORDS * 4, hipMemcpyDeviceToHost,
                           st));
    return FREP_OK;
}

frep_status ensure_events(frep_ctx* ctx, size_t n) {
    while (ctx->bev.size() < n) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        ctx->bev.push_back(e);
    }
    return FREP_OK;
}

// Runs the batched voxelization of jobs[idx[0..m)] and waits; the counter
// block of every tree is left in its host mirror.  tune: this is a full
// batch whose work may be tuned and replayed from a graph.
frep_status run_batch(frep_ctx* ctx, std::vector<Job>& jobs, const std::vector<uint32_t>& idx_in,
                      bool tune) {
    // longest tapes first: the plan's lanes then separate the few shapes
    // whose long level-0/1 chains set the latency from the many short ones,
    // which run beside them instead of after them
    std::vector<uint32_t> idx(idx_in);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
        return jobs[x].tp->words[0] > jobs[y].tp->words[0];
    });
    if (ctx->root_ops_required)
        for (const uint32_t k : idx)
            if (jobs[k].t->grid.n_levels > 1 && !root_job(BatchPlan{0, true}, jobs[k]))
                return fail(ctx, FREP_UNSUPPORTED, "tape has no level-0 schedule (FREP_ROOT_OPS=2)");
    const BatchPlan P = plan_batch(jobs, idx, ctx->batch_lanes, ctx->root_ops, ctx->cu_count);
    if (P.m == 0) return FREP_OK;
    BatchStage* S = &ctx->stage;
    if (S->stage_bytes < P.bytes || S->cnt_words < (size_t)P.m * CNT_WORDS) {
        drop_graphs(ctx);   // graphs hold the staging addresses
        const frep_status s = ensure_stage(ctx, S, P);
        if (s != FREP_OK) return s;
    }
    fill_batch(P, jobs, idx, static_cast<uint8_t*>(S->h_stage));
    const size_t nl = P.launches.size();
    std::vector<uint32_t> grid(nl, ctx->level_grid);
    std::vector<uint32_t> kcs(nl, 1u);
    TuneRecord* rec = nullptr;
    if (tune && ctx->autotune) {
        const uint64_t key = batch_key(P, static_cast<const uint8_t*>(S->h_stage));
        auto it = ctx->tune.find(key);
        if (it == ctx->tune.end()) {
            if (ctx->tune.size() >= kTuneMaxEntries) drop_graphs(ctx);
            it = ctx->tune.emplace(key, TuneRecord{}).first;
            it->second.best_ms.assign(nl, 1e30f);
            it->second.grid.assign(nl, ctx->level_grid);
            it->second.kc.assign(nl, 1u);
        }
        rec = &it->second;
    }
    hipEvent_t* ev = nullptr;
    bool timed_trial = false;
    // trials that would repeat an earlier one (no launch can take two cells
    // per lane) are skipped
    bool any_kc2 = false;
    for (const BatchLaunch& la : P.launches) any_kc2 = any_kc2 || la.kc2_ok;
    if (rec && rec->trial >= kTuneGrids && !any_kc2) rec->trial = kTuneTrials;
    if (rec && rec->trial < kTuneTrials) {
        const uint32_t cap = ctx->cu_count * kTuneBlocksPerCu[rec->trial % kTuneGrids];
        for (auto& g : grid) g = std::min(cap, ctx->level_grid);
        for (auto& k : kcs) k = rec->trial < kTuneGrids ? 1u : 2u;
        const frep_status s = ensure_events(ctx, nl + 1);
        if (s != FREP_OK) return s;
        ev = ctx->bev.data();
        timed_trial = true;
    } else if (rec) {
        grid = rec->grid;
        kcs = rec->kc;
    }
    // a plan with side lanes is launched directly: a graph replays forked
    // branches one after another, losing the overlap the lanes are for
    bool laned = false;
    for (const BatchGroup& G : P.groups) laned = laned || G.lane > 0;
    if (ctx->timing) HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    if (rec && !timed_trial && !laned) {
        if (!rec->exec) {
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const frep_status s = enqueue_batch(ctx, S, ctx->stream, P, grid, nullptr, nullptr, &kcs);
            const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            if (s != FREP_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return s;
            }
            HIP_TRY(e);
            const hipError_t ie = hipGraphInstantiate(&rec->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_TRY(ie);
        }
        HIP_TRY(hipGraphLaunch(rec->exec, ctx->stream));
    } else {
        const frep_status s =
            enqueue_batch(ctx, S, ctx->stream, P, grid, ev, ev ? nullptr : &ctx->lanes, &kcs);
        if (s != FREP_OK) return s;
    }
    if (ctx->timing) HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (uint32_t k = 0; k < P.m; ++k)
        std::memcpy(jobs[idx[k]].t->h_counters, S->h_cnt + (size_t)k * CNT_WORDS, CNT_WORDS * 4);
    if (timed_trial) {
        for (size_t li = 0; li < nl; ++li) {
            float ms = 0.0f;
            // the first launch of a later group is timed from the previous
            // launch, which includes that group's init and gather (small)
            HIP_TRY(hipEventElapsedTime(&ms, ev[li], ev[1 + li]));
            if (ms < rec->best_ms[li]) {
                rec->best_ms[li] = ms;
                rec->grid[li] = grid[li];
                rec->kc[li] = kcs[li];
            }
        }
        ++rec->trial;
    }
    if (ctx->timing) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->last.batch_ms += ms;
    }
    return FREP_OK;
}

frep_status check_camera(frep_ctx* ctx, const frep_camera* cam) {
    if (cam->width == 0 || cam->height == 0 || cam->width > 16384 || cam->height > 16384)
        return fail(ctx, FREP_INVALID_ARG, "image size out of range");
    if (cam->light_count > 4) return fail(ctx, FREP_INVALID_ARG, "at most 4 lights");
    return FREP_OK;
}

RenderArgs render_args(const TreeObj* t, const frep_camera* cam, uint8_t* rgba,
                       unsigned long long* hit_ids, float* normals) {
    RenderArgs a{};
    const uint32_t nl = t->grid.n_levels;
    for (uint32_t L = 0; L < nl; ++L) {
        a.lv[L].nodes = t->d_nodes[L];
        a.lv[L].ns = t->ns[L];
        a.lv[L].dim = t->grid.dims[L];
        a.lv[L].cs = t->cs[L];
        a.lv[L].cap = t->node_cap[L];
    }
    a.n_levels = nl;
    a.pool = t->d_pool;
    a.pool_words = t->pool_cap + kPoolPad;
    a.consts = t->d_consts;
    a.counters = t->d_counters;
    for (int c = 0; c < 3; ++c) {
        a.low[c] = t->grid.low[c];
        a.pos[c] = cam->pos[c];
        a.fwd[c] = cam->forward[c];
        a.right[c] = cam->right[c];
        a.up[c] = cam->up[c];
    }
    a.world_size = t->grid.world_size;
    a.tape_time = t->grid.tape_time;
    a.sw[0] = cam->screen_world[0];
    a.sw[1] = cam->screen_world[1];
    a.width = cam->width;
    a.height = cam->height;
    a.spp = cam->spp ? cam->spp : 1;
    a.ortho = cam->ortho;
    a.jitter = cam->jitter;
    a.seed = cam->seed;
    a.light_count = cam->light_count;
    for (int c = 0; c < 4; ++c) {
        a.background[c] = cam->background[c];
        a.ambient[c] = cam->ambient[c];
        for (int l = 0; l < 4; ++l) {
            a.light_dir[l][c] = cam->light_dir[l][c];
            a.light_color[l][c] = cam->light_color[l][c];
        }
    }
    a.rgba = rgba;
    a.hit_ids = hit_ids;
    a.normals = normals;
    return a;
}

}  // namespace frep

extern "C" {

frep_status frep_root_schedule(const uint32_t* words, uint32_t n_words, uint32_t lanes_per_cell,
                               uint32_t lds_per_block, uint32_t* info, uint32_t* out,
                               uint32_t out_cap) {
    if (!words || !info || n_words < 1 || (lanes_per_cell != 16 && lanes_per_cell != 64))
        return FREP_INVALID_ARG;
    if ((uint64_t)words[0] + 1 > n_words) return FREP_INVALID_ARG;
    for (uint32_t k = 0; k < words[0]; ++k)
        if ((words[1 + k] & 0xFFu) > OP_COPY) return FREP_BAD_TAPE;
    std::vector<uint32_t> sched;
    RootSched rs;
    const bool ok = build_root_schedule(words, lanes_per_cell, lds_per_block, &sched, &rs);
    const uint32_t v[12] = {ok ? 1u : 0u, rs.pass,        rs.nconst,     rs.c0,
                            rs.m0,        rs.e0,          rs.res,        rs.slots,
                            rs.table_bytes, rs.cell_bytes, rs.cpb,       (uint32_t)sched.size()};
    std::memcpy(info, v, sizeof(v));
    if (ok && out && out_cap >= sched.size()) std::memcpy(out, sched.data(), sched.size() * 4);
    return FREP_OK;
}

frep_status frep_voxelize(frep_ctx* ctx, frep_tape_t th, const frep_grid* grid, frep_tree_t h) {
    if (!ctx) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TapeObj* tp = get_tape(ctx, th);
    TreeObj* t = get_tree(ctx, h);
    if (!tp) return fail(ctx, FREP_INVALID_ARG, "unknown tape handle");
    if (!t) return fail(ctx, FREP_INVALID_ARG, "unknown tree handle");
    const frep_status gs = check_grid(ctx, grid);
    if (gs != FREP_OK) return gs;
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t nl = grid->n_levels;
    Job j;
    {
        const frep_status s = prepare_job(ctx, tp, t, grid, &j, ctx->stream, nullptr);
        if (s != FREP_OK) return s;
    }
    for (uint32_t attempt = 0;; ++attempt) {
        {
            const frep_status s = ensure_job(ctx, &j);
            if (s != FREP_OK) return s;
        }
        LevelArgs la[kMaxLevels];
        for (uint32_t L = 0; L < nl; ++L) la[L] = level_args(t, L, j.mm_stride);
        // node counts of the tree's previous build (zero for a new tree):
        // a hint for the interval levels' work shape only
        uint32_t nodes_hint[kMaxLevels] = {};
        for (uint32_t L = 0; L < nl; ++L) nodes_hint[L] = host_nodes(t, L);
        la[nl - 1].finish = 1;
        la[nl - 1].host_counters = t->h_counters;
        la[nl - 1].balance = ctx->point_balance;
        // the buffers are as k_init_tree leaves them for this tape when the
        // last plain build of it into them completed (finish_tree)
        const bool skip_init = t->ready_gen == t->dev_gen && t->ready_tape == tp->uid;
        t->ready_gen = 0;   // set again only when this build completes
        t->h_counters[CNT_DONE] = 0;   // finish_tree marks the mirror written
        // level 0 in passes over the tape's instructions when the tape has
        // a schedule (k_root_interval)
        const uint32_t d0 = grid->dims[0];
        const int rv = (ctx->root_ops && tp->d_sched && nl > 1)
                           ? pick_root_schedule(tp, d0 * d0 * d0, ctx->cu_count)
                           : -1;
        const bool use_root = rv >= 0;
        if (ctx->root_ops_required && nl > 1 && !use_root)
            return fail(ctx, FREP_UNSUPPORTED, "tape has no level-0 schedule (FREP_ROOT_OPS=2)");
        RootArgs ra{};
        if (use_root) {
            ra = root_args(tp, t, tp->root[rv], j.mm_stride, tp->root[rv].cpb);
            ra.a.balance = ctx->root_balance ? 1u : 0u;
        }
        // init, the level launches and the counter readback; with timing on,
        // an event before the init and after each launch
        auto enqueue = [&](bool events) -> frep_status {
            if (events) HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
            if (!skip_init)
                HIP_TRY(launch_init_tree(ctx->stream, t->d_counters, t->d_nodes[0], t->d_pool,
                                         tp->d_words, kRootOff, nl));
            if (events) HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
            for (uint32_t L = 0; L < nl; ++L) {
                if (L + 1 == nl)
                    HIP_TRY(launch_level_point(ctx->stream, la[L], t->maxs, t->slot_count,
                                               ctx->cu_count, ctx->level_grid, t->asm_ok));
                else if (L == 0 && use_root)
                    HIP_TRY(launch_root_interval(ctx->stream, ra));
                else
                    HIP_TRY(launch_level_interval(ctx->stream, la[L], t->maxs, t->slot_count,
                                                  ctx->cu_count, ctx->level_grid, nodes_hint[L]));
                if (events) HIP_TRY(hipEventRecord(ctx->ev[2 + L], ctx->stream));
            }
            return FREP_OK;
        };
        // a repeated voxelization (same buffers, tape, grid and time) is
        // replayed as one graph: its launches then follow each other without
        // the per-launch gaps of separate submissions.  A call whose
        // arguments differ from the previous one's is launched directly, so
        // animated calls never pay for a capture.
        uint64_t key = 0;
        if (!ctx->timing && ctx->graphs) {
            key = fnv1a(la, sizeof(LevelArgs) * nl);
            const uintptr_t v[10] = {(uintptr_t)tp->d_words, (uintptr_t)t->h_counters, nl,
                                     t->maxs, t->slot_count, t->asm_ok ? 1u : 0u,
                                     ctx->cu_count, ctx->level_grid, (uintptr_t)(rv + 1),
                                     skip_init ? 1u : 0u};
            key = fnv1a(v, sizeof(v), key);
        }
        const bool replay = key != 0 && ((t->vexec && t->vkey == key) || t->vkey_seen == key);
        if (!replay) {
            const frep_status s = enqueue(ctx->timing);
            if (s != FREP_OK) return s;
            t->vkey_seen = key;
        } else {
            if (!t->vexec || t->vkey != key) {
                if (t->vexec) (void)hipGraphExecDestroy(t->vexec);
                t->vexec = nullptr;
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                const frep_status s = enqueue(false);
                const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
                if (s != FREP_OK) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return s;
                }
                HIP_TRY(e);
                const hipError_t ie = hipGraphInstantiate(&t->vexec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_TRY(ie);
                t->vkey = key;
            }
            HIP_TRY(hipGraphLaunch(t->vexec, ctx->stream));
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (t->h_counters[CNT_DONE] != kCountersWritten)
            return fail(ctx, FREP_DEVICE, "the build's counters did not reach the host");
        if (t->h_counters[CNT_OVERFLOW] == 0) {
            t->ready_gen = t->dev_gen;
            t->ready_tape = tp->uid;
            break;
        }
        t->ready_gen = 0;
        const frep_status s = regrow_job(ctx, &j, attempt);
        if (s != FREP_OK) return s;
    }
    if (ctx->timing) {
        ctx->last = frep_timing{};
        ctx->last.n_levels = nl;
        HIP_TRY(hipEventElapsedTime(&ctx->last.init_ms, ctx->ev[0], ctx->ev[1]));
        for (uint32_t L = 0; L < nl; ++L)
            HIP_TRY(hipEventElapsedTime(&ctx->last.level_ms[L], ctx->ev[1 + L], ctx->ev[2 + L]));
    }
    t->built = true;
    return FREP_OK;
}

frep_status frep_voxelize_batch(frep_ctx* ctx, uint32_t n, const frep_tape_t* tapes,
                                const frep_grid* grids, const frep_tree_t* trees) {
    if (!ctx) return FREP_INVALID_ARG;
    if (n && (!tapes || !grids || !trees)) return fail(ctx, FREP_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<Job> jobs(n);
    std::vector<frep_tree_t> seen(trees, trees + n);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
        return fail(ctx, FREP_INVALID_ARG, "a tree appears twice in one batch");
    for (uint32_t k = 0; k < n; ++k) {
        if (!get_tape(ctx, tapes[k])) return fail(ctx, FREP_INVALID_ARG, "unknown tape handle");
        if (!get_tree(ctx, trees[k])) return fail(ctx, FREP_INVALID_ARG, "unknown tree handle");
        const frep_status gs = check_grid(ctx, &grids[k]);
        if (gs != FREP_OK) return gs;
        if (grids[k].n_levels > 1 && std::max<uint32_t>(get_tape(ctx, tapes[k])->n_mm, 1) >
                                         max_minmax(ctx->lds_per_block))
            return fail(ctx, FREP_UNSUPPORTED,
                        "tape has more than " + std::to_string(max_minmax(ctx->lds_per_block)) +
                            " MIN/MAX instructions");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    for (uint32_t k = 0; k < n; ++k) {
        const frep_status s =
            prepare_job(ctx, get_tape(ctx, tapes[k]), get_tree(ctx, trees[k]), &grids[k], &jobs[k],
                        ctx->stream, nullptr);
        if (s != FREP_OK) return s;
    }
    if (ctx->timing) ctx->last = frep_timing{};
    std::vector<uint32_t> pending(n);
    for (uint32_t k = 0; k < n; ++k) pending[k] = k;
    for (uint32_t attempt = 0; !pending.empty(); ++attempt) {
        for (uint32_t k : pending) {
            const frep_status s = ensure_job(ctx, &jobs[k]);
            if (s != FREP_OK) return s;
        }
        {
            const frep_status s = run_batch(ctx, jobs, pending, attempt == 0);
            if (s != FREP_OK) return s;
        }
        std::vector<uint32_t> again;
        for (uint32_t k : pending) {
            if (jobs[k].t->h_counters[CNT_OVERFLOW] == 0) {
                jobs[k].t->built = true;
                continue;
            }
            const frep_status s = regrow_job(ctx, &jobs[k], attempt);
            if (s != FREP_OK) return s;
            again.push_back(k);
        }
        pending.swap(again);
    }
    return FREP_OK;
}

frep_status frep_tree_info_get(frep_ctx* ctx, frep_tree_t h, frep_tree_info* out) {
    if (!ctx || !out) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = get_tree(ctx, h);
    if (!t || !t->built) return fail(ctx, FREP_INVALID_ARG, "unknown or unbuilt tree");
    *out = frep_tree_info{};
    out->n_levels = t->grid.n_levels;
    HIP_TRY(hipSetDevice(ctx->device));
    // tape-length stats (what the reference's stats buffer records,
    // vulkan/shaders/voxelizer/main.comp:146-147) and the work of each
    // level, reduced on the device
    const uint64_t top = host_pool_top(t);
    const uint32_t nl = t->grid.n_levels;
    void* acc = nullptr;
    HIP_TRY(ctx->dpool.alloc(kMaxLevels * 3 * 8, &acc));
    unsigned long long* d_acc = static_cast<unsigned long long*>(acc);
    HIP_TRY(hipMemsetAsync(d_acc, 0, kMaxLevels * 3 * 8, ctx->stream));
    for (uint32_t L = 0; L < nl; ++L)
        HIP_TRY(launch_tree_stats(ctx->stream, t->d_nodes[L], t->ns[L], host_nodes(t, L), t->d_pool,
                                  t->pool_cap + kPoolPad, L + 1 < nl, d_acc + 3 * L));
    unsigned long long h_acc[kMaxLevels * 3] = {};
    HIP_TRY(hipMemcpyAsync(h_acc, d_acc, sizeof(h_acc), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->dpool.free(acc);
    for (uint32_t L = 0; L < nl; ++L) {
        const uint64_t cells = (uint64_t)t->grid.dims[L] * t->grid.dims[L] * t->grid.dims[L];
        out->dims[L] = t->grid.dims[L];
        out->node_count[L] = host_nodes(t, L);
        out->tape_len_sum[L] = h_acc[3 * L];
        out->tape_len_max[L] = (uint32_t)h_acc[3 * L + 1];
        out->op_evals[L] = h_acc[3 * L] * cells;
        out->work_flops[L] = h_acc[3 * L + 2] * cells;
    }
    out->tape_pool_words = top;
    out->device_bytes = tree_device_bytes(t);
    out->slot_count = t->slot_count;
    out->retries = t->retries;
    return FREP_OK;
}

}  // extern "C"

namespace {

// The canonical blob of a built tree (layout in the oracle's header and the
// integration notes): levels in (x, y, z) order, child links as canonical
// indices, tapes by content.
frep_status canonical_blob(frep_ctx* ctx, TreeObj* t, std::vector<uint32_t>& v) {
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t nl = t->grid.n_levels;
    const uint64_t top = host_pool_top(t);
    std::vector<uint32_t> pool(top);
    HIP_TRY(hipMemcpy(pool.data(), t->d_pool, top * 4, hipMemcpyDeviceToHost));
    std::vector<std::vector<uint32_t>> recs(nl);
    std::vector<uint32_t> cnt(nl);
    for (uint32_t L = 0; L < nl; ++L) {
        cnt[L] = host_nodes(t, L);
        recs[L].resize((size_t)cnt[L] * t->ns[L]);
        if (cnt[L])
            HIP_TRY(hipMemcpy(recs[L].data(), t->d_nodes[L], recs[L].size() * 4,
                              hipMemcpyDeviceToHost));
    }
    // canonical order: sort each level by node coords (x, y, z)
    std::vector<std::vector<uint32_t>> order(nl), rank(nl);
    for (uint32_t L = 0; L < nl; ++L) {
        const uint32_t ns = t->ns[L];
        const uint32_t* r = recs[L].data();
        order[L].resize(cnt[L]);
        for (uint32_t j = 0; j < cnt[L]; ++j) order[L][j] = j;
        std::sort(order[L].begin(), order[L].end(), [&](uint32_t a, uint32_t b) {
            const uint32_t* ca = r + (size_t)a * ns + 1;
            const uint32_t* cb = r + (size_t)b * ns + 1;
            if (ca[0] != cb[0]) return ca[0] < cb[0];
            if (ca[1] != cb[1]) return ca[1] < cb[1];
            return ca[2] < cb[2];
        });
        rank[L].resize(cnt[L]);
        for (uint32_t k = 0; k < cnt[L]; ++k) rank[L][order[L][k]] = k;
    }
    v.push_back(0x54505246u);
    v.push_back(1u);
    v.push_back(nl);
    for (uint32_t L = 0; L < nl; ++L) v.push_back(t->grid.dims[L]);
    for (uint32_t L = 0; L < nl; ++L) v.push_back(cnt[L]);
    for (uint32_t L = 0; L < nl; ++L) {
        const uint32_t dim = t->grid.dims[L], cells = dim * dim * dim, W = cells / 32;
        const bool last = (L + 1 == nl);
        for (uint32_t k = 0; k < cnt[L]; ++k) {
            const uint32_t* r = recs[L].data() + (size_t)order[L][k] * t->ns[L];
            v.insert(v.end(), r + 1, r + 4);
            const uint32_t toff = r[0];
            if (toff >= top || (uint64_t)toff + pool[toff] + 1 > top)
                return fail(ctx, FREP_DEVICE, "corrupt tape offset in tree");
            v.insert(v.end(), pool.begin() + toff, pool.begin() + toff + pool[toff] + 1);
            v.insert(v.end(), r + 4, r + 4 + W);
            if (!last) {
                v.insert(v.end(), r + 4 + W, r + 4 + 2 * W);
                for (uint32_t i = 0; i < cells; ++i) {
                    const uint32_t ch = r[4 + 2 * W + i];
                    if (ch == kNoChild) { v.push_back(ch); continue; }
                    if (ch >= cnt[L + 1]) return fail(ctx, FREP_DEVICE, "corrupt child id in tree");
                    v.push_back(rank[L + 1][ch]);
                }
            }
        }
    }
    return FREP_OK;
}

// Tree file: a header with the grid, slot count and constant pool, then the
// canonical blob.
constexpr uint32_t kTreeFileMagic = 0x46505246u;   // 'FRPF'
constexpr uint32_t kTreeBlobMagic = 0x54505246u;   // 'FRPT'
constexpr uint32_t kTreeFileHeader = 18;

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

}  // namespace

extern "C" {

frep_status frep_tree_export(frep_ctx* ctx, frep_tree_t h, uint32_t* buf, uint64_t cap,
                             uint64_t* n_words) {
    if (!ctx || !n_words) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = get_tree(ctx, h);
    if (!t || !t->built) return fail(ctx, FREP_INVALID_ARG, "unknown or unbuilt tree");
    std::vector<uint32_t> v;
    {
        const frep_status s = canonical_blob(ctx, t, v);
        if (s != FREP_OK) return s;
    }
    *n_words = v.size();
    if (buf && cap >= v.size()) std::memcpy(buf, v.data(), v.size() * 4);
    return FREP_OK;
}

frep_status frep_tree_save(frep_ctx* ctx, frep_tree_t h, uint32_t* buf, uint64_t cap,
                           uint64_t* n_words) {
    if (!ctx || !n_words) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = get_tree(ctx, h);
    if (!t || !t->built) return fail(ctx, FREP_INVALID_ARG, "unknown or unbuilt tree");
    std::vector<uint32_t> v(kTreeFileHeader, 0u);
    v[0] = kTreeFileMagic;
    v[1] = 1u;
    v[2] = t->grid.n_levels;
    for (uint32_t L = 0; L < t->grid.n_levels; ++L) v[3 + L] = t->grid.dims[L];
    for (int c = 0; c < 3; ++c) v[11 + c] = fbits(t->grid.low[c]);
    v[14] = fbits(t->grid.world_size);
    v[15] = fbits(t->grid.tape_time);
    v[16] = t->slot_count;
    std::vector<float> consts(t->consts_cap);
    uint32_t nc = t->n_consts;
    v[17] = nc;
    if (nc) {
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipMemcpy(consts.data(), t->d_consts, (size_t)nc * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nc; ++i) v.push_back(fbits(consts[i]));
    }
    {
        const frep_status s = canonical_blob(ctx, t, v);
        if (s != FREP_OK) return s;
    }
    *n_words = v.size();
    if (buf && cap >= v.size()) std::memcpy(buf, v.data(), v.size() * 4);
    return FREP_OK;
}

frep_status frep_tree_load(frep_ctx* ctx, const uint32_t* buf, uint64_t n_words,
                           frep_tree_t* out) {
    if (!ctx || !buf || !out) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto bad = [&](const std::string& why) {
        return fail(ctx, FREP_INVALID_ARG, "tree file: " + why);
    };
    if (n_words < kTreeFileHeader || buf[0] != kTreeFileMagic || buf[1] != 1u)
        return bad("not a version-1 tree file");
    frep_grid g{};
    g.n_levels = buf[2];
    if (g.n_levels < 1 || g.n_levels > kMaxLevels) return bad("level count");
    for (uint32_t L = 0; L < g.n_levels; ++L) g.dims[L] = buf[3 + L];
    for (int c = 0; c < 3; ++c) g.low[c] = bitsf(buf[11 + c]);
    g.world_size = bitsf(buf[14]);
    g.tape_time = bitsf(buf[15]);
    {
        const frep_status s = check_grid(ctx, &g);
        if (s != FREP_OK) return s;
    }
    const uint32_t nl = g.n_levels;
    const uint32_t slots = buf[16], nc = buf[17];
    if (slots < 4 || slots > 256) return bad("slot count");
    if (nc > 65536) return bad("constant count");
    uint64_t pos = kTreeFileHeader + (uint64_t)nc;
    if (pos + 3 + 2 * (uint64_t)nl > n_words) return bad("truncated");
    std::vector<float> consts(nc);
    for (uint32_t i = 0; i < nc; ++i) consts[i] = bitsf(buf[kTreeFileHeader + i]);
    const uint32_t* b = buf;
    if (b[pos] != kTreeBlobMagic || b[pos + 1] != 1u || b[pos + 2] != nl)
        return bad("canonical blob header");
    pos += 3;
    for (uint32_t L = 0; L < nl; ++L)
        if (b[pos + L] != g.dims[L]) return bad("blob dims differ from the grid");
    std::vector<uint32_t> cnt(nl);
    for (uint32_t L = 0; L < nl; ++L) cnt[L] = b[pos + nl + L];
    pos += 2 * (uint64_t)nl;
    if (cnt[0] != 1) return bad("level 0 must hold one node");
    // records in the device layout, tapes appended level by level
    std::vector<std::vector<uint32_t>> recs(nl);
    std::vector<uint32_t> pool(kRootOff, 0u);
    std::vector<uint64_t> words(nl, 0);
    std::vector<std::vector<uint32_t>> coords(nl);
    std::vector<std::vector<uint8_t>> seen(nl);
    uint64_t base1 = 0;
    bool asm_ok = !ctx->no_asm, reads_t = false;
    uint32_t n_mm = 0;
    for (uint32_t L = 0; L < nl; ++L) {
        if (cnt[L] > kMaxLevelNodes) return bad("too many nodes in a level");
        const uint32_t dim = g.dims[L], cells = dim * dim * dim, W = cells / 32;
        const bool last = (L + 1 == nl);
        const uint32_t ns = node_words(dim, last);
        recs[L].assign((size_t)cnt[L] * ns, 0u);
        coords[L].resize((size_t)cnt[L] * 3);
        seen[L].assign(cnt[L], 0);
        for (uint32_t k = 0; k < cnt[L]; ++k) {
            if (pos + 4 > n_words) return bad("truncated");
            uint32_t* r = recs[L].data() + (size_t)k * ns;
            for (int c = 0; c < 3; ++c) coords[L][3 * k + c] = r[1 + c] = b[pos + c];
            if (k > 0) {
                const uint32_t* p = &coords[L][3 * (k - 1)];
                const uint32_t* q = &coords[L][3 * k];
                if (!(p[0] < q[0] || (p[0] == q[0] && (p[1] < q[1] || (p[1] == q[1] && p[2] < q[2])))))
                    return bad("nodes out of canonical order");
            }
            pos += 3;
            const uint32_t tk = b[pos];
            if (pos + 1 + tk > n_words) return bad("truncated tape");
            if (validate_tape(b + pos, tk + 1, nc, slots, FREP_TAPE_FORMAT_WIDE) != 0)
                return bad("malformed tape");
            r[0] = (uint32_t)pool.size();
            pool.insert(pool.end(), b + pos, b + pos + tk + 1);
            if (L == 0) {
                while (pool.size() % 8) pool.push_back(0u);
                base1 = pool.size();
                bool t_written = false;
                for (uint32_t i = 0; i < tk; ++i) {
                    const uint32_t w = b[pos + 1 + i], op = w & 0xFFu;
                    n_mm += (op == OP_MIN || op == OP_MAX);
                    asm_ok = asm_ok && asm_op_ok(op);
                    if (op != OP_CONST && !t_written) {
                        const bool binary = op == OP_ADD || op == OP_SUB || op == OP_MUL ||
                                            op == OP_DIV || op == OP_MIN || op == OP_MAX;
                        if (((w >> 16) & 0xFFu) == 3u || (binary && (w >> 24) == 3u)) reads_t = true;
                    }
                    if (((w >> 8) & 0xFFu) == 3u) t_written = true;
                }
            } else {
                words[L] += tk + 1;
            }
            pos += tk + 1;
            const uint64_t need = (uint64_t)W * (last ? 1 : 2) + (last ? 0 : cells);
            if (pos + need > n_words) return bad("truncated node");
            std::memcpy(r + 4, b + pos, (size_t)W * 4);
            pos += W;
            if (!last) {
                std::memcpy(r + 4 + W, b + pos, (size_t)W * 4);
                pos += W;
                for (uint32_t i = 0; i < cells; ++i) {
                    const uint32_t ch = b[pos + i];
                    const bool leaf = (r[4 + (i >> 5)] >> (i & 31)) & 1u;
                    const bool inner = (r[4 + W + (i >> 5)] >> (i & 31)) & 1u;
                    if (leaf && inner) return bad("cell both leaf and interior");
                    if (inner != (ch != kNoChild)) return bad("interior bit and child link disagree");
                    r[4 + 2 * W + i] = ch;
                }
                pos += cells;
            }
        }
    }
    if (pos != n_words) return bad("trailing words");
    // tree structure: every node below the root has exactly one parent, at
    // the cell its coordinates name
    for (uint32_t L = 0; L + 1 < nl; ++L) {
        const uint32_t dim = g.dims[L], cells = dim * dim * dim, W = cells / 32;
        const uint32_t ns = node_words(dim, false);
        for (uint32_t k = 0; k < cnt[L]; ++k) {
            const uint32_t* r = recs[L].data() + (size_t)k * ns;
            for (uint32_t i = 0; i < cells; ++i) {
                const uint32_t ch = r[4 + 2 * W + i];
                if (ch == kNoChild) continue;
                if (ch >= cnt[L + 1] || seen[L + 1][ch]) return bad("child link");
                seen[L + 1][ch] = 1;
                const uint32_t cx = i / (dim * dim), cy = (i / dim) % dim, cz = i % dim;
                const uint32_t* q = &coords[L + 1][3 * (size_t)ch];
                if (q[0] != r[1] * dim + cx || q[1] != r[2] * dim + cy || q[2] != r[3] * dim + cz)
                    return bad("child coordinates");
            }
        }
        for (uint32_t k = 0; k < cnt[L + 1]; ++k)
            if (!seen[L + 1][k]) return bad("node without a parent");
    }
    if (coords[0][0] || coords[0][1] || coords[0][2]) return bad("root coordinates");
    if (nl > 1 && std::max<uint32_t>(n_mm, 1) > max_minmax(ctx->lds_per_block))
        return fail(ctx, FREP_UNSUPPORTED, "tape has too many MIN/MAX instructions");
    // device copy
    HIP_TRY(hipSetDevice(ctx->device));
    frep_tree_t hnd = 0;
    {
        const frep_status s = new_tree(ctx, &hnd);
        if (s != FREP_OK) return s;
    }
    TreeObj* t = get_tree(ctx, hnd);
    auto drop = [&](frep_status s) {
        free_tree_buffers(ctx, t);
        ctx->dpool.free(t->d_counters);
        ctx->hpool.free(t->h_counters);
        ctx->trees[hnd - 1].reset();
        return s;
    };
    t->grid = g;
    t->slot_count = slots;
    t->maxs = maxs_for(slots);
    t->asm_ok = asm_ok;
    t->reads_t = reads_t;
    float c = g.world_size;
    for (uint32_t L = 0; L < nl; ++L) {
        c /= (float)g.dims[L];
        t->cs[L] = c;
        t->ns[L] = node_words(g.dims[L], L + 1 == nl);
        const frep_status s = ensure_level(ctx, t, L, std::max<uint32_t>(cnt[L], 1));
        if (s != FREP_OK) return drop(s);
        if (cnt[L])
            if (hipMemcpy(t->d_nodes[L], recs[L].data(), recs[L].size() * 4,
                          hipMemcpyHostToDevice) != hipSuccess)
                return drop(fail(ctx, FREP_DEVICE, "tree upload"));
    }
    {
        const frep_status s = ensure_pool(ctx, t, pool.size());
        if (s != FREP_OK) return drop(s);
    }
    if (hipMemset(t->d_pool, 0, (pool.size() + kPoolPad) * 4) != hipSuccess ||
        hipMemcpy(t->d_pool, pool.data(), pool.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return drop(fail(ctx, FREP_DEVICE, "tape pool upload"));
    void* p = nullptr;
    if (ctx->dpool.alloc((size_t)(nc + kPoolPad) * 4, &p) != hipSuccess)
        return drop(fail(ctx, FREP_OUT_OF_MEMORY, "constant pool"));
    t->d_consts = static_cast<float*>(p);
    t->consts_cap = nc + kPoolPad;
    t->n_consts = nc;
    t->consts_uid = 0;
    if (nc && hipMemcpy(t->d_consts, consts.data(), (size_t)nc * 4, hipMemcpyHostToDevice) !=
                  hipSuccess)
        return drop(fail(ctx, FREP_DEVICE, "constant upload"));
    // counters: node counts and tape words per level on the host mirror; the
    // device block as a finished build leaves it (zero but the level-1 base)
    std::memset(t->h_counters, 0, CNT_WORDS * 4);
    t->h_counters[CNT_BASE1] = (uint32_t)base1;
    for (uint32_t L = 1; L < nl; ++L) {
        const uint64_t v = ((uint64_t)cnt[L] << kAllocShift) | words[L];
        std::memcpy(t->h_counters + CNT_ALLOC + 2 * L, &v, 8);
    }
    std::vector<uint32_t> dc(CNT_WORDS, 0u);
    dc[CNT_BASE1] = (uint32_t)base1;
    if (hipMemcpy(t->d_counters, dc.data(), CNT_WORDS * 4, hipMemcpyHostToDevice) != hipSuccess)
        return drop(fail(ctx, FREP_DEVICE, "counter upload"));
    ++t->dev_gen;
    t->ready_gen = 0;
    t->built = true;
    *out = hnd;
    return FREP_OK;
}

frep_status frep_camera_look(frep_camera* cam, const float initial_up[3], float fov_deg) {
    if (!cam || !initial_up) return FREP_INVALID_ARG;
    const float* f = cam->forward;
    const float cr[3] = {f[1] * initial_up[2] - f[2] * initial_up[1],
                         f[2] * initial_up[0] - f[0] * initial_up[2],
                         f[0] * initial_up[1] - f[1] * initial_up[0]};
    const float l = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const float r[3] = {cr[0] / l, cr[1] / l, cr[2] / l};
    cam->right[0] = r[0]; cam->right[1] = r[1]; cam->right[2] = r[2];
    cam->up[0] = r[1] * f[2] - r[2] * f[1];
    cam->up[1] = r[2] * f[0] - r[0] * f[2];
    cam->up[2] = r[0] * f[1] - r[1] * f[0];
    if (!cam->ortho) {
        const float rad = fov_deg * (3.14159265358979323846f / 180.0f);
        cam->screen_world[0] = 2.0f * std::tan(rad / 2.0f);
        cam->screen_world[1] = cam->screen_world[0] * ((float)cam->height / (float)cam->width);
    }
    return FREP_OK;
}

}  // extern "C"

namespace {

// Staging layout of a render's outputs: rgba, then hit ids, then normals,
// each 256-byte aligned.
struct RenderStaging {
    size_t rgba_b, hit_b, nrm_b, off_hit, off_nrm, total;
};

RenderStaging render_staging(const frep_camera* cam, bool rgba, bool hits, bool normals) {
    RenderStaging r;
    const uint32_t spp = cam->spp ? cam->spp : 1;
    const size_t px = (size_t)cam->width * cam->height;
    r.rgba_b = rgba ? px * 4 : 0;
    r.hit_b = hits ? px * spp * 8 : 0;
    r.nrm_b = normals ? px * spp * 12 : 0;
    r.off_hit = (r.rgba_b + 255) & ~(size_t)255;
    r.off_nrm = (r.off_hit + r.hit_b + 255) & ~(size_t)255;
    r.total = r.off_nrm + r.nrm_b + 256;
    return r;
}

// Renders into the device staging at `base` (laid out by render_staging)
// and copies the requested outputs to the caller's host arrays.
frep_status render_into(frep_ctx* ctx, TreeObj* t, const frep_camera* cam, char* base,
                        const RenderStaging& S, uint8_t* rgba, uint64_t* hit_ids,
                        float* normals) {
    const RenderArgs a = render_args(t, cam, rgba ? (uint8_t*)base : nullptr,
                                     hit_ids ? (unsigned long long*)(base + S.off_hit) : nullptr,
                                     normals ? (float*)(base + S.off_nrm) : nullptr);
    if (ctx->timing) HIP_TRY(hipEventRecord(ctx->ev[kMaxLevels + 2], ctx->stream));
    HIP_TRY(launch_render(ctx->stream, a, t->maxs, t->slot_count));
    if (ctx->timing) HIP_TRY(hipEventRecord(ctx->ev[kMaxLevels + 3], ctx->stream));
    if (rgba) HIP_TRY(hipMemcpyAsync(rgba, base, S.rgba_b, hipMemcpyDeviceToHost, ctx->stream));
    if (hit_ids)
        HIP_TRY(hipMemcpyAsync(hit_ids, base + S.off_hit, S.hit_b, hipMemcpyDeviceToHost,
                               ctx->stream));
    if (normals)
        HIP_TRY(hipMemcpyAsync(normals, base + S.off_nrm, S.nrm_b, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->timing)
        HIP_TRY(hipEventElapsedTime(&ctx->last.render_ms, ctx->ev[kMaxLevels + 2],
                                    ctx->ev[kMaxLevels + 3]));
    return FREP_OK;
}

frep_status render_checks(frep_ctx* ctx, frep_tree_t h, const frep_camera* cam, TreeObj** out) {
    TreeObj* t = get_tree(ctx, h);
    if (!t || !t->built) return fail(ctx, FREP_INVALID_ARG, "unknown or unbuilt tree");
    const frep_status s = check_camera(ctx, cam);
    if (s != FREP_OK) return s;
    *out = t;
    return FREP_OK;
}

}  // namespace

extern "C" {

frep_status frep_render(frep_ctx* ctx, frep_tree_t h, const frep_camera* cam, uint8_t* rgba,
                        uint64_t* hit_ids, float* normals) {
    if (!ctx || !cam) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = nullptr;
    {
        const frep_status s = render_checks(ctx, h, cam, &t);
        if (s != FREP_OK) return s;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const RenderStaging S = render_staging(cam, rgba, hit_ids, normals);
    if (S.total > ctx->out_bytes) {
        if (ctx->d_out) HIP_TRY(hipFree(ctx->d_out));
        ctx->d_out = nullptr;
        ctx->out_bytes = 0;
        HIP_TRY(hipMalloc(&ctx->d_out, S.total));
        ctx->out_bytes = S.total;
    }
    return render_into(ctx, t, cam, (char*)ctx->d_out, S, rgba, hit_ids, normals);
}

frep_status frep_target_create(frep_ctx* ctx, uint32_t width, uint32_t height, uint32_t layers,
                               frep_target_t* out) {
    if (!ctx || !out) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (width == 0 || height == 0 || width > 16384 || height > 16384 || layers == 0 ||
        layers > 64)
        return fail(ctx, FREP_INVALID_ARG, "target size: 1..16384 pixels a side, 1..64 layers");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<TargetObj> tg(new TargetObj());
    tg->width = width;
    tg->height = height;
    tg->layers = layers;
    const size_t bytes = (size_t)width * height * layers * 4;
    HIP_TRY(ctx->dpool.alloc(bytes, &tg->d));
    HIP_TRY(hipMemsetAsync(tg->d, 0, bytes, ctx->stream));   // RenderTarget::ZeroOutImage
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->targets.push_back(std::move(tg));
    *out = (frep_target_t)ctx->targets.size();
    return FREP_OK;
}

frep_status frep_target_free(frep_ctx* ctx, frep_target_t h) {
    if (!ctx) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (h == 0 || h > ctx->targets.size() || !ctx->targets[h - 1])
        return fail(ctx, FREP_INVALID_ARG, "unknown target");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->dpool.free(ctx->targets[h - 1]->d);
    ctx->targets[h - 1].reset();
    return FREP_OK;
}

frep_status frep_render_layer(frep_ctx* ctx, frep_tree_t h, const frep_camera* cam,
                              frep_target_t target, uint32_t layer) {
    if (!ctx || !cam) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = nullptr;
    {
        const frep_status s = render_checks(ctx, h, cam, &t);
        if (s != FREP_OK) return s;
    }
    if (target == 0 || target > ctx->targets.size() || !ctx->targets[target - 1])
        return fail(ctx, FREP_INVALID_ARG, "unknown target");
    const TargetObj* tg = ctx->targets[target - 1].get();
    if (cam->width != tg->width || cam->height != tg->height)
        return fail(ctx, FREP_INVALID_ARG, "camera size differs from the target's");
    if ((cam->spp ? cam->spp : 1) != 1)
        return fail(ctx, FREP_INVALID_ARG, "a target layer holds one sample per pixel (spp 1)");
    if (layer >= tg->layers) return fail(ctx, FREP_INVALID_ARG, "layer out of range");
    HIP_TRY(hipSetDevice(ctx->device));
    uint8_t* dst = (uint8_t*)tg->d + (size_t)layer * tg->width * tg->height * 4;
    const RenderArgs a = render_args(t, cam, dst, nullptr, nullptr);
    if (ctx->timing) HIP_TRY(hipEventRecord(ctx->ev[kMaxLevels + 2], ctx->stream));
    HIP_TRY(launch_render(ctx->stream, a, t->maxs, t->slot_count));
    if (ctx->timing) HIP_TRY(hipEventRecord(ctx->ev[kMaxLevels + 3], ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->timing)
        HIP_TRY(hipEventElapsedTime(&ctx->last.render_ms, ctx->ev[kMaxLevels + 2],
                                    ctx->ev[kMaxLevels + 3]));
    return FREP_OK;
}

frep_status frep_target_resolve(frep_ctx* ctx, frep_target_t target, uint8_t* rgba) {
    if (!ctx || !rgba) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (target == 0 || target > ctx->targets.size() || !ctx->targets[target - 1])
        return fail(ctx, FREP_INVALID_ARG, "unknown target");
    const TargetObj* tg = ctx->targets[target - 1].get();
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t px = (size_t)tg->width * tg->height;
    if (px * 4 + 256 > ctx->out_bytes) {
        if (ctx->d_out) HIP_TRY(hipFree(ctx->d_out));
        ctx->d_out = nullptr;
        ctx->out_bytes = 0;
        HIP_TRY(hipMalloc(&ctx->d_out, px * 4 + 256));
        ctx->out_bytes = px * 4 + 256;
    }
    HIP_TRY(launch_resolve(ctx->stream, tg->d, (uint32_t)px, tg->layers, ctx->d_out));
    HIP_TRY(hipMemcpyAsync(rgba, ctx->d_out, px * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FREP_OK;
}

frep_status frep_render_noalloc(frep_ctx* ctx, frep_tree_t h, const frep_camera* cam,
                                uint8_t* rgba, uint64_t* hit_ids, float* normals) {
    if (!ctx || !cam) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TreeObj* t = nullptr;
    {
        const frep_status s = render_checks(ctx, h, cam, &t);
        if (s != FREP_OK) return s;
    }
    const RenderStaging S = render_staging(cam, rgba, hit_ids, normals);
    if (S.total > (size_t)FREP_NOALLOC_STAGING_BYTES)
        return fail(ctx, FREP_INVALID_ARG,
                    "frep_render_noalloc: the requested outputs need " + std::to_string(S.total) +
                        " bytes of staging, more than its fixed " +
                        std::to_string((size_t)FREP_NOALLOC_STAGING_BYTES) + " bytes");
    if (!ctx->noalloc_staging) return fail(ctx, FREP_DEVICE, "no staging address for this device");
    HIP_TRY(hipSetDevice(ctx->device));
    return render_into(ctx, t, cam, (char*)ctx->noalloc_staging, S, rgba, hit_ids, normals);
}

}  // extern "C"

namespace {

// Runs one batch evaluation: copies `in` arrays to the device, launches, and
// copies `out` arrays back.  Temporary buffers live for the call only.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" {

frep_status frep_eval_points(frep_ctx* ctx, frep_tape_t th, const float* xyz, uint32_t n, float t,
                             float* out) {
    if (!ctx || (n && (!xyz || !out))) return fail(ctx, FREP_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    TapeObj* tp = get_tape(ctx, th);
    if (!tp) return fail(ctx, FREP_INVALID_ARG, "unknown tape handle");
    if (n == 0) return FREP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf din, dout;
    HIP_TRY(hipMalloc(&din.p, (size_t)n * 12));
    HIP_TRY(hipMalloc(&dout.p, (size_t)n * 4));
    HIP_TRY(hipMemcpyAsync(din.p, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_eval_points(ctx->stream, tp->d_words, tp->d_consts, (const float*)din.p, n, t,
                               (float*)dout.p, tp->maxs, tp->asm_ok && !ctx->no_asm));
    HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FREP_OK;
}

frep_status frep_eval_intervals(frep_ctx* ctx, frep_tape_t th, const float* lo, const float* hi,
                                uint32_t n, float t, float* out, uint8_t* choices) {
    if (!ctx || (n && (!lo || !hi || !out))) return fail(ctx, FREP_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    TapeObj* tp = get_tape(ctx, th);
    if (!tp) return fail(ctx, FREP_INVALID_ARG, "unknown tape handle");
    if (n == 0) return FREP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf dlo, dhi, dout, dch;
    HIP_TRY(hipMalloc(&dlo.p, (size_t)n * 12));
    HIP_TRY(hipMalloc(&dhi.p, (size_t)n * 12));
    HIP_TRY(hipMalloc(&dout.p, (size_t)n * 8));
    const size_t chb = (size_t)n * tp->n_mm;
    if (choices && chb) HIP_TRY(hipMalloc(&dch.p, chb));
    HIP_TRY(hipMemcpyAsync(dlo.p, lo, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dhi.p, hi, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_eval_intervals(ctx->stream, tp->d_words, tp->d_consts, (const float*)dlo.p,
                                  (const float*)dhi.p, n, t, (float*)dout.p, (uint8_t*)dch.p,
                                  tp->n_mm, tp->maxs));
    HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (choices && chb)
        HIP_TRY(hipMemcpyAsync(choices, dch.p, chb, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FREP_OK;
}

frep_status frep_eval_gradients(frep_ctx* ctx, frep_tape_t th, const float* xyz, uint32_t n,
                                float t, float* out4) {
    if (!ctx || (n && (!xyz || !out4))) return fail(ctx, FREP_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    TapeObj* tp = get_tape(ctx, th);
    if (!tp) return fail(ctx, FREP_INVALID_ARG, "unknown tape handle");
    if (n == 0) return FREP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf din, dout;
    HIP_TRY(hipMalloc(&din.p, (size_t)n * 12));
    HIP_TRY(hipMalloc(&dout.p, (size_t)n * 16));
    HIP_TRY(hipMemcpyAsync(din.p, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_eval_gradients(ctx->stream, tp->d_words, tp->d_consts, (const float*)din.p, n,
                                  t, (float*)dout.p, tp->maxs));
    HIP_TRY(hipMemcpyAsync(out4, dout.p, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FREP_OK;
}

frep_status frep_math(frep_ctx* ctx, uint32_t fn, const float* x, uint32_t n, float* out) {
    if (!ctx || fn > 2 || (n && (!x || !out))) return fail(ctx, FREP_INVALID_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (n == 0) return FREP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf din, dout;
    HIP_TRY(hipMalloc(&din.p, (size_t)n * 4));
    HIP_TRY(hipMalloc(&dout.p, (size_t)n * 4));
    HIP_TRY(hipMemcpyAsync(din.p, x, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_math(ctx->stream, fn, (const float*)din.p, n, (float*)dout.p));
    HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FREP_OK;
}

}  // extern "C"
