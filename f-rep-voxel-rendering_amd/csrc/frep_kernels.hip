// frep_kernels.hip -- HIP kernels for gfx950: tree levels, render, batch evals.
//
// Level kernels are persistent: a fixed grid of 256-thread blocks loops over
// "chunks" (64 consecutive mask bits of one node = one wavefront of work) up
// to the node count that the previous level left in device memory, so the
// whole voxelization is enqueued without a host round trip per level (the
// reference waits on a fence and reads the child count back after every
// level: vulkan/src/engine/voxelizer.cpp:323-328).
#include "frep_asm.hpp"
#include "frep_device.hpp"
#include "frep_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <utility>

namespace frep {

namespace {

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive prefix sum of v over the lanes of the wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, (unsigned)d);
        if (lane >= (uint32_t)d) v += u;
    }
    return v;
}

// Counter words read by the kernels (level node counts, tape bases, the
// overflow bits of earlier levels) are final before the reading launch
// starts: earlier launches on the stream wrote them, and the launch
// boundary makes those writes visible.  They are read with plain loads,
// served by the L2: agent-scope atomic loads would go to the memory side,
// to the same line as the allocation atomics of the running level, and
// queue behind them (half the level-1 throughput on the 1024^3 random CSG).
__device__ __forceinline__ uint32_t load_counter(const uint32_t* p) { return *p; }

__device__ __forceinline__ uint64_t load_alloc(const uint32_t* counters, uint32_t level) {
    return *(const unsigned long long*)(counters + CNT_ALLOC + 2 * level);
}

// Node count of a level (level 0 is the root alone), written by the
// previous level's kernel.
__device__ __forceinline__ uint32_t level_nodes(const uint32_t* counters, uint32_t level) {
    return level == 0 ? 1u : uni(alloc_nodes(load_alloc(counters, level)));
}

// Child reservation of one wavefront: `req` = (nodes << kAllocShift) | words
// (0 when the chunk has no ambiguous cell) added to the packed counter of
// the next level; every lane gets the counter's previous value.
struct WaveReserve {
    __device__ __forceinline__ uint64_t operator()(unsigned long long* ctr, uint64_t req,
                                                   uint32_t lane) const {
        unsigned long long old = 0;
        if (req != 0 && lane == 0) old = atomicAdd(ctr, (unsigned long long)req);
        return ((uint64_t)uni((uint32_t)__shfl((int)(uint32_t)(old >> 32), 0)) << 32) |
               uni((uint32_t)__shfl((int)(uint32_t)old, 0));
    }
};

// Child reservation of a whole block: the wavefronts post their requests in
// LDS, one atomic reserves their sum and each wavefront's share starts after
// those of the wavefronts before it, so ids stay dense with a quarter of the
// same-address atomics.  Every wavefront of the block calls it once per
// round (two barriers); `req`/`base` alternate between two LDS copies by
// round parity, so a round's posts never overwrite values the previous
// round still reads.
struct BlockReserve {
    unsigned long long* req;    // LDS, one entry per wavefront of the block
    unsigned long long* base;   // LDS, the block's reserved start
    uint32_t wib;               // wavefront index in the block
    __device__ __forceinline__ uint64_t operator()(unsigned long long* ctr, uint64_t req_w,
                                                   uint32_t lane) const {
        if (lane == 0) req[wib] = req_w;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
            for (uint32_t k = 0; k < kLevelBlock / 64; ++k) t += req[k];
            *base = t != 0 ? atomicAdd(ctr, t) : 0ull;
        }
        __syncthreads();
        unsigned long long old = *base;
        for (uint32_t k = 0; k < wib; ++k) old += req[k];
        return old;
    }
};

// Write the 64-bit ballot as two mask words (lanes 0 and 1, one word each).
__device__ __forceinline__ void store_mask64(uint32_t* dst, uint64_t m, uint32_t lane) {
    if (lane < 2) dst[lane] = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
}

// Eight consecutive 64-bit masks as one store: word i of m[0..7] is placed in
// lane i, and lanes 0..15 write it.
__device__ __forceinline__ void store_mask64x8(uint32_t* dst, const uint64_t (&m)[8],
                                               uint32_t lane) {
    uint32_t v;
    // the masks were just written by VALU compares: wait states before VALU
    // reads them
    asm volatile(
        "s_nop 4\n"
        "v_writelane_b32 %0, %1, 0\n v_writelane_b32 %0, %2, 1\n"
        "v_writelane_b32 %0, %3, 2\n v_writelane_b32 %0, %4, 3\n"
        "v_writelane_b32 %0, %5, 4\n v_writelane_b32 %0, %6, 5\n"
        "v_writelane_b32 %0, %7, 6\n v_writelane_b32 %0, %8, 7\n"
        "v_writelane_b32 %0, %9, 8\n v_writelane_b32 %0, %10, 9\n"
        "v_writelane_b32 %0, %11, 10\n v_writelane_b32 %0, %12, 11\n"
        "v_writelane_b32 %0, %13, 12\n v_writelane_b32 %0, %14, 13\n"
        "v_writelane_b32 %0, %15, 14\n v_writelane_b32 %0, %16, 15\n"
        : "=v"(v)
        : "s"((uint32_t)m[0]), "s"((uint32_t)(m[0] >> 32)), "s"((uint32_t)m[1]),
          "s"((uint32_t)(m[1] >> 32)), "s"((uint32_t)m[2]), "s"((uint32_t)(m[2] >> 32)),
          "s"((uint32_t)m[3]), "s"((uint32_t)(m[3] >> 32)), "s"((uint32_t)m[4]),
          "s"((uint32_t)(m[4] >> 32)), "s"((uint32_t)m[5]), "s"((uint32_t)(m[5] >> 32)),
          "s"((uint32_t)m[6]), "s"((uint32_t)(m[6] >> 32)), "s"((uint32_t)m[7]),
          "s"((uint32_t)(m[7] >> 32)));
    if (lane < 16) dst[lane] = v;
}

// Cell (x, y, z) of mask bit idx: INDEX3D = z + dim*y + dim*dim*x
// (vulkan/shaders/utils.comp:6).  Shifts when dim is a power of two.
__device__ __forceinline__ void cell_of(uint32_t idx, uint32_t dim, uint32_t sh, uint32_t* x,
                                        uint32_t* y, uint32_t* z) {
    if (sh) {
        *x = idx >> (2 * sh);
        *y = (idx >> sh) & (dim - 1);
        *z = idx & (dim - 1);
    } else {
        *x = idx / (dim * dim);
        *y = (idx / dim) % dim;
        *z = idx % dim;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Tree initialisation: root tape copied into the pool, root record, counters.
// ---------------------------------------------------------------------------
__global__ void k_init_tree(uint32_t* counters, uint32_t* nodes0, uint32_t* pool,
                            const uint32_t* root_tape, uint32_t root_off, uint32_t n_levels) {
    const uint32_t tid = threadIdx.x;
    const uint32_t n = root_tape[0];
    for (uint32_t i = tid; i <= n; i += blockDim.x) pool[root_off + i] = root_tape[i];
    if (tid < CNT_WORDS) counters[tid] = tid == CNT_BASE1 ? root_off + ((n + 1 + 7u) & ~7u) : 0u;
    if (tid == 0) {
        nodes0[0] = root_off;
        nodes0[1] = nodes0[2] = nodes0[3] = 0;
    }
    (void)n_levels;
}

// ---------------------------------------------------------------------------
// Tape shortening of a wavefront's ambiguous cells (shorten.comp:164-223),
// for G groups of 64 cells (lane l owns cell l of each group).
//
// One walk over the parent tape from its end serves every cell: a cell
// keeps word i iff the word's output slot is live there (the output slot
// is cleared, then the operands its min/max choice reads are set).  The
// walk counts each cell's kept words, sorts the cells of a group into
// classes that decide the same at every word (the same kept words and the
// same min/max rewrites, so the same shortened tape: cells that decided the
// same so far have the same live slots, so a class splits only at a live
// MIN/MAX whose choices differ), and logs in LDS, for every word, the
// ballots of the cells keeping it.  The first cell of each class owns the
// class's tape.  One reservation (per wavefront, or per block) takes the
// child ids and exactly the words of those tapes ([k, k words] each; none
// where nothing was removed: the child then reuses the parent's tape,
// shorten.comp:188-191).  Each owner's tape is written in tape order from
// the log by ballot and prefix-sum compaction, 64 words at a time.  A tape
// longer than the log is walked a second time instead, the owners writing
// their kept words from the end of their region.
// ---------------------------------------------------------------------------

// A kept word as the shortened tape holds it: a MIN/MAX with one chosen
// operand becomes a COPY of that operand (shorten.comp:199-214).
__device__ __forceinline__ uint32_t shortened_word(uint32_t w, bool is_mm, uint32_t ch,
                                                   uint32_t o, uint32_t ib) {
    const uint32_t copy_a = (w & ~0xFFu) | OP_COPY;
    const uint32_t copy_b = OP_COPY | (o << 8) | (ib << 16) | (ib << 24);
    return !is_mm ? w : (ch == MM_FIRST ? copy_a : (ch == MM_SECOND ? copy_b : w));
}

// Backward walk over the parent tape tw[0..n) of a wavefront whose min/max
// choice rows are scr[m * G + g] (m-th MIN/MAX, group g; mm of them):
// on_word(i, m, is_mm, live, L, ch, ow) for every word some cell keeps,
// with, per group, this lane's cell's liveness, the ballot of it, the
// cell's choice (MM_BOTH for other ops) and the word as the cell keeps it.
// With LOG, the ballots of word i are also stored as log entry n - 1 - i,
// logL[(n - 1 - i) * G + g] (zero for a word no cell keeps): each window's
// ballots are gathered in registers, one lane per word, and stored by every
// lane at the window's end.
template <int MAXS, int G, bool LOG, class F>
__device__ __forceinline__ void backward_walk(const uint32_t* tw, uint32_t n, uint32_t mm,
                                              const uint4* scr, const bool (&amb)[G],
                                              uint32_t lane, uint2* logL, F&& on_word) {
    constexpr int AW = (MAXS + 31) / 32;
    uint32_t act[G][AW];
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int k = 0; k < AW; ++k) act[g][k] = 0;
        act[g][0] = amb[g] ? 1u : 0u;                  // slot 0 is the result
    }
    int32_t jm = (int32_t)mm - 1;
    uint4 row[G];                                      // next choice rows, read ahead
#pragma unroll
    for (int g = 0; g < G; ++g) row[g] = scr[(jm > 0 ? jm : 0) * G + g];
    int32_t e = (int32_t)n;                            // window = words [e-64, e)
    uint32_t bw0 = window_load(tw, e - 64, n, lane);
    uint32_t bw1 = window_load(tw, e - 128, n, lane);
    uint32_t bw2 = window_load(tw, e - 192, n, lane);
    // nothing in flight when the loop starts, so that the compiler's waits
    // for the loads issued in the loop fall at the window's end
    __builtin_amdgcn_s_waitcnt(0);
    uint32_t acc_lo[G], acc_hi[G];                     // the window's ballots (LOG)
    while (e > 0) {
        // the window after next, loaded now and taken at this window's end
        // (a load taken right after it is issued would wait for it there)
        const uint32_t bw3 = window_load(tw, e - 256, n, lane);
        const int32_t s = e - 64;
        const int32_t lo = s > 0 ? s : 0;
        if constexpr (LOG) {
#pragma unroll
            for (int g = 0; g < G; ++g) acc_lo[g] = acc_hi[g] = 0u;
        }
        for (int32_t i = e - 1; i >= lo; --i) {
            const uint32_t w = rdlane(bw0, (uint32_t)(i - s));
            const uint32_t op = w & 0xFFu, o = slot_field<MAXS>((w >> 8) & 0xFFu);
            const bool is_mm = (op == OP_MIN || op == OP_MAX);
            const uint32_t mi = jm > 0 ? (uint32_t)jm : 0u;
            uint32_t ch[G];
#pragma unroll
            for (int g = 0; g < G; ++g) ch[g] = MM_BOTH;
            if (is_mm) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const uint4 v = row[g];
                    const uint64_t Fm = ((uint64_t)v.y << 32) | v.x;
                    const uint64_t Sm = ((uint64_t)v.w << 32) | v.z;
                    ch[g] = (uint32_t)((Fm >> lane) & 1u) | ((uint32_t)((Sm >> lane) & 1u) << 1);
                }
                --jm;
#pragma unroll
                for (int g = 0; g < G; ++g) row[g] = scr[(jm > 0 ? jm : 0) * G + g];
            }
            const uint32_t obit = 1u << (o & 31u);
            bool live[G];
            uint64_t L[G], any = 0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                live[g] = (act[g][o >> 5] & obit) != 0;
                L[g] = __ballot(live[g]);
                any |= L[g];
            }
            // a word no cell keeps changes nothing (common once the cells of
            // a chunk agree on their choices)
            if (any == 0) continue;
            if constexpr (LOG) {
                // lane e - 1 - i takes the word's ballots; the lane select
                // goes through M0 (one scalar operand per instruction besides
                // it; the compiler sets M0 before each of its own uses, so it
                // is not listed).  A VALU-written SGPR needs wait states only
                // as a lane select, not as the value written.
                const uint32_t slot = (uint32_t)(e - 1 - i);
                if constexpr (G == 1) {
                    asm volatile("s_mov_b32 m0, %4\n"
                                 "v_writelane_b32 %0, %2, m0\n"
                                 "v_writelane_b32 %1, %3, m0\n"
                                 : "+v"(acc_lo[0]), "+v"(acc_hi[0])
                                 : "s"((uint32_t)L[0]), "s"((uint32_t)(L[0] >> 32)), "s"(slot));
                } else {
                    static_assert(G == 2, "one or two cell groups");
                    asm volatile("s_mov_b32 m0, %8\n"
                                 "v_writelane_b32 %0, %4, m0\n"
                                 "v_writelane_b32 %1, %5, m0\n"
                                 "v_writelane_b32 %2, %6, m0\n"
                                 "v_writelane_b32 %3, %7, m0\n"
                                 : "+v"(acc_lo[0]), "+v"(acc_hi[0]), "+v"(acc_lo[1]),
                                   "+v"(acc_hi[1])
                                 : "s"((uint32_t)L[0]), "s"((uint32_t)(L[0] >> 32)),
                                   "s"((uint32_t)L[1]), "s"((uint32_t)(L[1] >> 32)), "s"(slot));
                }
            }
            const uint32_t ia = slot_field<MAXS>((w >> 16) & 0xFFu);
            const uint32_t ib = slot_field<MAXS>(w >> 24);
            uint32_t ow[G];
            if (!is_mm) {
                // operands a kept word reads, the same for every cell: none
                // for CONST, both for ADD/SUB/MUL/DIV, else A
                const uint32_t rd = op == OP_CONST ? MM_NONE
                                    : (op - OP_ADD <= OP_DIV - OP_ADD) ? MM_BOTH : MM_FIRST;
                const uint32_t abit = (rd & MM_FIRST) ? 1u << (ia & 31u) : 0u;
                const uint32_t bbit = (rd & MM_SECOND) ? 1u << (ib & 31u) : 0u;
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    // a kept word clears its output slot, then sets its
                    // operand slots (shorten.comp:179)
                    if constexpr (AW == 1) {
                        act[g][0] = live[g] ? ((act[g][0] & ~obit) | abit | bbit) : act[g][0];
                    } else {
                        const uint32_t keep = live[g] ? ~0u : 0u;
                        act[g][o >> 5] &= ~(obit & keep);
                        act[g][ia >> 5] |= abit & keep;
                        act[g][ib >> 5] |= bbit & keep;
                    }
                    ow[g] = w;
                }
            } else {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    // MIN/MAX: the operands this cell's choice reads; lane
                    // masks instead of branches, output slot cleared first
                    const uint32_t sel = ch[g];
                    const uint32_t keep = live[g] ? ~0u : 0u;
                    act[g][o >> 5] &= ~(obit & keep);
                    act[g][ia >> 5] |= ((sel & MM_FIRST) ? (1u << (ia & 31u)) : 0u) & keep;
                    act[g][ib >> 5] |= ((sel & MM_SECOND) ? (1u << (ib & 31u)) : 0u) & keep;
                    ow[g] = shortened_word(w, is_mm, ch[g], o, ib);
                }
            }
            on_word((uint32_t)i, mi, is_mm, live, L, ch, ow);
        }
        if constexpr (LOG) {
            if (lane < (uint32_t)(e - lo)) {
                const uint32_t entry = n - (uint32_t)e + lane;
#pragma unroll
                for (int g = 0; g < G; ++g) logL[entry * G + g] = make_uint2(acc_lo[g], acc_hi[g]);
            }
        }
        bw0 = bw1;
        bw1 = bw2;
        bw2 = bw3;
        e -= 64;
    }
}

// Shortened tapes and child records of a wavefront's ambiguous cells
// (Bamb[g]: the ambiguous cells of group g; this lane's cell of group g is
// mask bit cell0 + 64 g + lane of the node at ncx, ncy, ncz).  The parent
// tape is at toff
// (n words, mm MIN/MAX whose choice rows are in scr); logL is the
// wavefront's log (cap entries of G ballots, one entry per word).
// child[g] gets the child id of this lane's cell of group g (kNoChild for
// cells that are not ambiguous or when the level ran out of capacity).
template <int MAXS, int G, class Reserve>
__device__ __forceinline__ void shorten_children(const LevelArgs& a, uint32_t lane,
                                                 uint32_t toff, uint32_t n, uint32_t mm,
                                                 const uint4* scr, uint2* logL, uint32_t cap,
                                                 const uint64_t (&Bamb)[G], uint32_t cell0,
                                                 uint32_t ncx, uint32_t ncy, uint32_t ncz,
                                                 uint64_t child_base, const Reserve& reserve,
                                                 uint32_t (&child)[G]) {
    const uint32_t* tw = a.pool + toff + 1;
    bool amb[G];
    uint32_t namb = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        amb[g] = ((Bamb[g] >> lane) & 1u) != 0;
        child[g] = kNoChild;
        namb += (uint32_t)__popcll(Bamb[g]);
    }
    // ---- the walk: kept counts, classes, log ----
    uint32_t kept[G];
    uint64_t eq[G];                                    // cells deciding as this lane's so far
#pragma unroll
    for (int g = 0; g < G; ++g) {
        kept[g] = 0;
        eq[g] = Bamb[g];
    }
    const bool logged = n <= cap;                      // the log holds every word
    auto count = [&](uint32_t, uint32_t, bool is_mm, const bool (&live)[G], const uint64_t (&)[G],
                     const uint32_t (&ch)[G], const uint32_t (&)[G]) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            kept[g] += live[g] ? 1u : 0u;
            // cells of one class have the same live slots, so they can part
            // only where a live MIN/MAX chose differently
            if (is_mm) {
                const bool f = live[g] && (ch[g] & MM_FIRST) != 0;
                const bool sc = live[g] && (ch[g] & MM_SECOND) != 0;
                const uint64_t K1 = __ballot(f), K2 = __ballot(sc);
                if (live[g]) eq[g] &= (f ? K1 : ~K1) & (sc ? K2 : ~K2);
            }
        }
    };
    if (namb != 0) {
        if (logged) backward_walk<MAXS, G, true>(tw, n, mm, scr, amb, lane, logL, count);
        else backward_walk<MAXS, G, false>(tw, n, mm, scr, amb, lane, logL, count);
    }
    // ---- owners and one exact reservation ----
    bool own[G];
    uint32_t lead[G], off[G], words = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        lead[g] = amb[g] ? (uint32_t)__builtin_ctzll(eq[g] & Bamb[g]) : lane;
        own[g] = amb[g] && lead[g] == lane && kept[g] < n;
        const uint32_t need = own[g] ? kept[g] + 1u : 0u;
        const uint32_t incl = wave_incl_scan(need, lane);
        off[g] = words + incl - need;
        words += rdlane(incl, 63);
    }
    const uint64_t old =
        reserve((unsigned long long*)(a.counters + CNT_ALLOC + 2 * (a.level + 1)),
                ((uint64_t)namb << kAllocShift) | (uint64_t)words, lane);
    if (namb == 0) return;
    const uint32_t cbase = alloc_nodes(old);
    const uint64_t pbase = child_base + alloc_words(old);
    const bool nodes_ok = (uint64_t)cbase + namb <= a.child_cap;
    const bool pool_ok = pbase + words <= a.pool_cap;
    if (!(nodes_ok && pool_ok)) {
        if (lane == 0)
            atomicOr(a.counters + CNT_OVERFLOW,
                     ((nodes_ok ? 0u : OVF_NODES) | (pool_ok ? 0u : OVF_POOL)) << (2 * a.level));
        return;
    }
    uint32_t reg[G], lreg[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        reg[g] = (uint32_t)(pbase + off[g]);
        lreg[g] = (uint32_t)__shfl((int)reg[g], (int)lead[g]);
    }
    // ---- the owners' tapes ----
    if (logged) {
        // tape order, 64 words at a time: lane j holds word w0 + j and its
        // MIN/MAX row; each owner takes the words its log bit marks, at the
        // positions its running count (lane ld of pos[g]) and the ballot's
        // prefix give
        uint64_t M[G];
        uint32_t pos[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            M[g] = __ballot(own[g]);
            pos[g] = 1u;
            if (own[g]) a.pool[reg[g]] = kept[g];
        }
        uint32_t mm0 = 0;
        for (uint32_t w0 = 0; w0 < n; w0 += 64) {
            const uint32_t i = w0 + lane;
            const uint32_t ic = i < n ? i : n - 1u;
            const uint32_t wd = tw[ic];
            const uint32_t op = wd & 0xFFu;
            const bool is_mm = i < n && (op == OP_MIN || op == OP_MAX);
            const uint64_t mmb = __ballot(is_mm);
            const uint32_t mi = mm0 + lane_rank(mmb);
            mm0 += (uint32_t)__popcll(mmb);
            const uint32_t o = slot_field<MAXS>((wd >> 8) & 0xFFu), ib = slot_field<MAXS>(wd >> 24);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint2 el = logL[(n - 1u - ic) * G + g];
                const uint64_t Lm = ((uint64_t)el.y << 32) | el.x;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (is_mm) v = scr[mi * G + g];
                const uint64_t Fm = ((uint64_t)v.y << 32) | v.x;
                const uint64_t Sm = ((uint64_t)v.w << 32) | v.z;
                for (uint64_t Mo = M[g]; Mo != 0; Mo &= Mo - 1u) {
                    const uint32_t ld = (uint32_t)__builtin_ctzll(Mo);
                    const bool kp = i < n && ((Lm >> ld) & 1u) != 0;
                    const uint64_t b = __ballot(kp);
                    const uint32_t r = rdlane(pos[g], ld);
                    if (kp) {
                        const uint32_t c = (uint32_t)((Fm >> ld) & 1u) |
                                           ((uint32_t)((Sm >> ld) & 1u) << 1);
                        a.pool[rdlane(reg[g], ld) + r + lane_rank(b)] =
                            shortened_word(wd, is_mm, c, o, ib);
                    }
                    if (lane == ld) pos[g] = r + (uint32_t)__popcll(b);
                }
            }
        }
    } else {
        uint32_t r[G];
#pragma unroll
        for (int g = 0; g < G; ++g) r[g] = 0;
        backward_walk<MAXS, G, false>(
            tw, n, mm, scr, amb, lane, logL,
            [&](uint32_t, uint32_t, bool, const bool (&live)[G], const uint64_t (&)[G],
                const uint32_t (&)[G], const uint32_t (&ow)[G]) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (own[g] && live[g]) a.pool[reg[g] + kept[g] - r[g]] = ow[g];
                    r[g] += live[g] ? 1u : 0u;
                }
            });
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (own[g]) a.pool[reg[g]] = kept[g];
    }
    // ---- child records ----
    uint32_t rank0 = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (amb[g]) {
            child[g] = cbase + rank0 + lane_rank(Bamb[g]);
            uint32_t cx, cy, cz;
            cell_of(cell0 + 64u * g + lane, a.dim, a.dim_shift, &cx, &cy, &cz);
            uint32_t* crec = a.child_nodes + (size_t)child[g] * a.child_ns;
            crec[0] = kept[g] < n ? lreg[g] : toff;
            crec[1] = ncx * a.dim + cx;
            crec[2] = ncy * a.dim + cy;
            crec[3] = ncz * a.dim + cz;
        }
        rank0 += (uint32_t)__popcll(Bamb[g]);
    }
}

// ---------------------------------------------------------------------------
// Non-last level: interval classification of every cell, min/max choice
// recording, and tape shortening of ambiguous cells
// (vulkan/shaders/voxelizer/main.comp:88-149, vulkan/shaders/tape/
// interval.comp:121-160, vulkan/shaders/tape/shorten.comp:148-223).
//
// One wavefront owns 64 consecutive mask bits of one node; lane l evaluates
// the cell whose bit index is 64*sub + l.  Mask words are built by ballot and
// written once (no atomics).  Each MIN/MAX records two 64-bit ballots (lanes
// whose choice includes FIRST / SECOND) in the wave's scratch row; the
// backward liveness pass of the shortening reads them back
// (shorten_children: one tape per class of cells that keep the same words,
// written by ballot and prefix-sum compaction into exactly reserved
// regions).  If nothing was removed the child reuses the parent's tape
// (shorten.comp:188-191).
// ---------------------------------------------------------------------------
template <int MAXS, class Reserve>
__device__ __forceinline__ void interval_chunk(const LevelArgs& a, uint32_t node, uint32_t sub,
                                               uint64_t child_base, uint32_t lane, uint4* scr,
                                               uint2* logL, uint32_t cap,
                                               const Reserve& reserve) {
    const uint32_t dim = a.dim, cells = dim * dim * dim, W = cells / 32;
    uint32_t* rec = a.nodes + (size_t)node * a.ns;
    ctape_t hdr = as_const(rec);
    const uint32_t toff = hdr[0], ncx = hdr[1], ncy = hdr[2], ncz = hdr[3];
    const uint32_t idx = sub * 64 + lane;
    uint32_t cx, cy, cz;
    cell_of(idx, dim, a.dim_shift, &cx, &cy, &cz);
    const uint32_t gx = ncx * dim + cx, gy = ncy * dim + cy, gz = ncz * dim + cz;
    // main.comp:113-116, 124-126: box [pos, pos + cell_size]
    const float lx = a.low[0] + (float)gx * a.cs;
    const float ly = a.low[1] + (float)gy * a.cs;
    const float lz = a.low[2] + (float)gz * a.cs;
    const uint32_t n = as_const(a.pool + toff)[0];
    const uint32_t* tw = a.pool + toff + 1;             // instruction words
    // window 0 goes out with the length instead of after it: a tape starts
    // below the pool's end and 64 words of padding follow the pool
    const uint32_t raw0 = tw[lane];
    // The level ends with its longest tapes: their wavefronts take issue
    // priority over the shorter ones on the same SIMD (restored at the end).
    if (n >= 96) __builtin_amdgcn_s_setprio(3);
    else if (n >= 64) __builtin_amdgcn_s_setprio(2);
    else if (n >= 32) __builtin_amdgcn_s_setprio(1);

    // ---- forward pass: interval evaluation (interval.comp:121-160) ----
    float ilo, ihi;
    uint32_t mm = 0;
    float slo[MAXS], shi[MAXS];
    slo[0] = lx; shi[0] = lx + a.cs;
    slo[1] = ly; shi[1] = ly + a.cs;
    slo[2] = lz; shi[2] = lz + a.cs;
    slo[3] = a.tape_time; shi[3] = a.tape_time;
    // three windows in flight: the current one (with its constants), the
    // next (constants requested), the one after (words requested)
    uint32_t fw0 = lane < n ? raw0 : 0xFFFFFFFFu;
    float fc0 = window_consts(fw0, a.consts);               // windows 1 and 2 beside it
    uint32_t fw1 = window_load(tw, 64, n, lane), fw2 = window_load(tw, 128, n, lane);
    // nothing in flight when the loop starts but window 1's constants, so
    // that the compiler's waits for the loads issued in the loop fall at the
    // window's end
    __builtin_amdgcn_s_waitcnt(0);
    float fc1 = window_consts(fw1, a.consts);
    for (uint32_t s = 0; s < n; s += 64) {
        // the constants of the window after next and the words of the one
        // after that go out now and are taken at this window's end (a load
        // taken right after it is issued would wait for it there)
        const uint32_t fw3 = window_load(tw, (int32_t)s + 192, n, lane);
        const float fc2 = window_consts(fw2, a.consts);
        const uint32_t m = uni(n - s < 64 ? n - s : 64);   // a scalar loop bound
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t w = rdlane(fw0, j);
            // operand fields are slot indices except for CONST (pool
            // index); masking keeps every register-file access in range
            const uint32_t op = w & 0xFFu, o = slot_field<MAXS>((w >> 8) & 0xFFu);
            const uint32_t ia = slot_field<MAXS>((w >> 16) & 0xFFu), ib = slot_field<MAXS>(w >> 24);
            // One uniform compare per test, the ops inner levels run most
            // first (CONST, COPY, ADD/SUB/MUL make up ~90% of their words);
            // op codes: 0 CONST 1 SIN 2 COS 3 ADD 4 SUB 5 MUL 6 DIV 7 MIN
            //           8 MAX 9 EXP 10 NEG 11 SQRT 12 COPY
            Itv r;
            if (op == OP_CONST) {
                const float c = rdlanef(fc0, j);
                r = Itv{c, c};
            } else if (op == OP_COPY) {
                if (o == ia) continue;
                r = Itv{slo[ia], shi[ia]};
                FREP_FENCE();
            } else if (op - OP_ADD <= OP_MUL - OP_ADD) {
                const Itv A{slo[ia], shi[ia]};
                FREP_FENCE();
                const Itv B{slo[ib], shi[ib]};
                FREP_FENCE();
                if (op == OP_MUL) r = imul(A, B);
                else if (op == OP_ADD) r = Itv{A.lo + B.lo, A.hi + B.hi};
                else r = Itv{A.lo - B.hi, A.hi - B.lo};
            } else if (op == OP_MIN || op == OP_MAX) {
                const Itv A{slo[ia], shi[ia]};
                FREP_FENCE();
                const Itv B{slo[ib], shi[ib]};
                FREP_FENCE();
                r = op == OP_MIN ? Itv{fmin_glm(A.lo, B.lo), fmin_glm(A.hi, B.hi)}
                                 : Itv{fmax_glm(A.lo, B.lo), fmax_glm(A.hi, B.hi)};
                const uint32_t ch = mm_choice(op, A, B);
                const uint64_t F = __ballot((ch & MM_FIRST) != 0);
                const uint64_t S = __ballot((ch & MM_SECOND) != 0);
                scr[mm] = make_uint4((uint32_t)F, (uint32_t)(F >> 32), (uint32_t)S,
                                     (uint32_t)(S >> 32));
                ++mm;
            } else if (op == OP_NEG) {
                const Itv A{slo[ia], shi[ia]};
                FREP_FENCE();
                r = Itv{-A.hi, -A.lo};
            } else if (op == OP_SQRT) {
                const Itv A{slo[ia], shi[ia]};
                FREP_FENCE();
                r = Itv{__builtin_sqrtf(fmax_glm(0.0f, A.lo)),
                        __builtin_sqrtf(fmax_glm(0.0f, A.hi))};
            } else {                                   // DIV, SIN, COS, EXP
                const Itv A{slo[ia], shi[ia]};
                FREP_FENCE();
                const Itv B{slo[ib], shi[ib]};
                FREP_FENCE();
                r = interval_op(op, A, B, 0.0f);
            }
            FREP_FENCE();
            slo[o] = r.lo;
            shi[o] = r.hi;
        }
        fw0 = fw1; fc0 = fc1;
        fw1 = fw2; fc1 = fc2;
        fw2 = fw3;
    }
    ilo = slo[0];
    ihi = shi[0];
    // main.comp:88-108: lo > 0 empty, hi < 0 full, else ambiguous
    const bool full = ihi < 0.0f;
    const bool empty = ilo > 0.0f;
    const bool amb = !full && !empty;
    const uint64_t Bfull = __ballot(full);
    const uint64_t Bamb = __ballot(amb);
    store_mask64(rec + 4 + sub * 2, Bfull, lane);
    store_mask64(rec + 4 + W + sub * 2, Bamb, lane);

    const uint64_t Bambs[1] = {Bamb};
    uint32_t child[1];
    shorten_children<MAXS, 1>(a, lane, toff, n, mm, scr, logL, cap, Bambs, sub * 64, ncx,
                              ncy, ncz, child_base, reserve, child);
    rec[4 + 2 * W + idx] = child[0];
    __builtin_amdgcn_s_setprio(0);
}

// The same rules with two cells per lane: a wavefront owns 128 consecutive
// mask bits (two 64-bit groups, lane l evaluating bits 128*sub + l and
// 128*sub + 64 + l).  Tape words are decoded, dispatched and their windows
// and constants loaded once for both cells, and a wavefront's header and
// window loads are waited on once per 128 cells; each MIN/MAX writes one
// choice row per group.  Everything else -- classification, reservation,
// shortening, child records -- is interval_chunk's, per group.
template <int MAXS, class Reserve>
__device__ __forceinline__ void interval_chunk2(const LevelArgs& a, uint32_t node, uint32_t sub,
                                                uint64_t child_base, uint32_t lane, uint4* scr,
                                                uint2* logL, uint32_t cap,
                                                const Reserve& reserve) {
    const uint32_t dim = a.dim, cells = dim * dim * dim, W = cells / 32;
    uint32_t* rec = a.nodes + (size_t)node * a.ns;
    ctape_t hdr = as_const(rec);
    const uint32_t toff = hdr[0], ncx = hdr[1], ncy = hdr[2], ncz = hdr[3];
    const uint32_t idx0 = sub * 128 + lane, idx1 = idx0 + 64;
    uint32_t cx, cy, cz;
    cell_of(idx0, dim, a.dim_shift, &cx, &cy, &cz);
    const uint32_t gx0 = ncx * dim + cx, gy0 = ncy * dim + cy, gz0 = ncz * dim + cz;
    cell_of(idx1, dim, a.dim_shift, &cx, &cy, &cz);
    const uint32_t gx1 = ncx * dim + cx, gy1 = ncy * dim + cy, gz1 = ncz * dim + cz;
    const uint32_t n = as_const(a.pool + toff)[0];
    const uint32_t* tw = a.pool + toff + 1;
    // window 0 goes out with the length instead of after it: a tape starts
    // below the pool's end and 64 words of padding follow the pool
    const uint32_t raw0 = tw[lane];
    if (n >= 96) __builtin_amdgcn_s_setprio(3);
    else if (n >= 64) __builtin_amdgcn_s_setprio(2);
    else if (n >= 32) __builtin_amdgcn_s_setprio(1);

    // ---- forward pass: interval evaluation (interval.comp:121-160) ----
    uint32_t mm = 0;
    float slo0[MAXS], shi0[MAXS], slo1[MAXS], shi1[MAXS];
    {
        // main.comp:113-116, 124-126: box [pos, pos + cell_size]
        const float lx0 = a.low[0] + (float)gx0 * a.cs, lx1 = a.low[0] + (float)gx1 * a.cs;
        const float ly0 = a.low[1] + (float)gy0 * a.cs, ly1 = a.low[1] + (float)gy1 * a.cs;
        const float lz0 = a.low[2] + (float)gz0 * a.cs, lz1 = a.low[2] + (float)gz1 * a.cs;
        slo0[0] = lx0; shi0[0] = lx0 + a.cs; slo1[0] = lx1; shi1[0] = lx1 + a.cs;
        slo0[1] = ly0; shi0[1] = ly0 + a.cs; slo1[1] = ly1; shi1[1] = ly1 + a.cs;
        slo0[2] = lz0; shi0[2] = lz0 + a.cs; slo1[2] = lz1; shi1[2] = lz1 + a.cs;
        slo0[3] = shi0[3] = slo1[3] = shi1[3] = a.tape_time;
    }
    uint32_t fw0 = lane < n ? raw0 : 0xFFFFFFFFu;
    float fc0 = window_consts(fw0, a.consts);               // windows 1 and 2 beside it
    uint32_t fw1 = window_load(tw, 64, n, lane), fw2 = window_load(tw, 128, n, lane);
    // nothing in flight when the loop starts but window 1's constants, so
    // that the compiler's waits for the loads issued in the loop fall at the
    // window's end
    __builtin_amdgcn_s_waitcnt(0);
    float fc1 = window_consts(fw1, a.consts);
    for (uint32_t s = 0; s < n; s += 64) {
        // the constants of the window after next and the words of the one
        // after that go out now and are taken at this window's end (a load
        // taken right after it is issued would wait for it there)
        const uint32_t fw3 = window_load(tw, (int32_t)s + 192, n, lane);
        const float fc2 = window_consts(fw2, a.consts);
        const uint32_t m = uni(n - s < 64 ? n - s : 64);   // a scalar loop bound
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t w = rdlane(fw0, j);
            const uint32_t op = w & 0xFFu, o = slot_field<MAXS>((w >> 8) & 0xFFu);
            const uint32_t ia = slot_field<MAXS>((w >> 16) & 0xFFu), ib = slot_field<MAXS>(w >> 24);
            Itv r0, r1;
            if (op == OP_CONST) {
                const float c = rdlanef(fc0, j);
                r0 = r1 = Itv{c, c};
            } else if (op == OP_COPY) {
                if (o == ia) continue;
                r0 = Itv{slo0[ia], shi0[ia]};
                r1 = Itv{slo1[ia], shi1[ia]};
                FREP_FENCE();
            } else if (op - OP_ADD <= OP_MUL - OP_ADD) {
                const Itv A0{slo0[ia], shi0[ia]}, A1{slo1[ia], shi1[ia]};
                FREP_FENCE();
                const Itv B0{slo0[ib], shi0[ib]}, B1{slo1[ib], shi1[ib]};
                FREP_FENCE();
                if (op == OP_MUL) {
                    r0 = imul(A0, B0);
                    r1 = imul(A1, B1);
                } else if (op == OP_ADD) {
                    r0 = Itv{A0.lo + B0.lo, A0.hi + B0.hi};
                    r1 = Itv{A1.lo + B1.lo, A1.hi + B1.hi};
                } else {
                    r0 = Itv{A0.lo - B0.hi, A0.hi - B0.lo};
                    r1 = Itv{A1.lo - B1.hi, A1.hi - B1.lo};
                }
            } else if (op == OP_MIN || op == OP_MAX) {
                const Itv A0{slo0[ia], shi0[ia]}, A1{slo1[ia], shi1[ia]};
                FREP_FENCE();
                const Itv B0{slo0[ib], shi0[ib]}, B1{slo1[ib], shi1[ib]};
                FREP_FENCE();
                if (op == OP_MIN) {
                    r0 = Itv{fmin_glm(A0.lo, B0.lo), fmin_glm(A0.hi, B0.hi)};
                    r1 = Itv{fmin_glm(A1.lo, B1.lo), fmin_glm(A1.hi, B1.hi)};
                } else {
                    r0 = Itv{fmax_glm(A0.lo, B0.lo), fmax_glm(A0.hi, B0.hi)};
                    r1 = Itv{fmax_glm(A1.lo, B1.lo), fmax_glm(A1.hi, B1.hi)};
                }
                const uint32_t ch0 = mm_choice(op, A0, B0), ch1 = mm_choice(op, A1, B1);
                const uint64_t F0 = __ballot((ch0 & MM_FIRST) != 0);
                const uint64_t S0 = __ballot((ch0 & MM_SECOND) != 0);
                const uint64_t F1 = __ballot((ch1 & MM_FIRST) != 0);
                const uint64_t S1 = __ballot((ch1 & MM_SECOND) != 0);
                scr[2 * mm] = make_uint4((uint32_t)F0, (uint32_t)(F0 >> 32), (uint32_t)S0,
                                         (uint32_t)(S0 >> 32));
                scr[2 * mm + 1] = make_uint4((uint32_t)F1, (uint32_t)(F1 >> 32), (uint32_t)S1,
                                             (uint32_t)(S1 >> 32));
                ++mm;
            } else if (op == OP_NEG) {
                const Itv A0{slo0[ia], shi0[ia]}, A1{slo1[ia], shi1[ia]};
                FREP_FENCE();
                r0 = Itv{-A0.hi, -A0.lo};
                r1 = Itv{-A1.hi, -A1.lo};
            } else if (op == OP_SQRT) {
                const Itv A0{slo0[ia], shi0[ia]}, A1{slo1[ia], shi1[ia]};
                FREP_FENCE();
                r0 = Itv{__builtin_sqrtf(fmax_glm(0.0f, A0.lo)),
                         __builtin_sqrtf(fmax_glm(0.0f, A0.hi))};
                r1 = Itv{__builtin_sqrtf(fmax_glm(0.0f, A1.lo)),
                         __builtin_sqrtf(fmax_glm(0.0f, A1.hi))};
            } else {                                   // DIV, SIN, COS, EXP
                const Itv A0{slo0[ia], shi0[ia]}, A1{slo1[ia], shi1[ia]};
                FREP_FENCE();
                const Itv B0{slo0[ib], shi0[ib]}, B1{slo1[ib], shi1[ib]};
                FREP_FENCE();
                r0 = interval_op(op, A0, B0, 0.0f);
                r1 = interval_op(op, A1, B1, 0.0f);
            }
            FREP_FENCE();
            slo0[o] = r0.lo;
            shi0[o] = r0.hi;
            slo1[o] = r1.lo;
            shi1[o] = r1.hi;
        }
        fw0 = fw1; fc0 = fc1;
        fw1 = fw2; fc1 = fc2;
        fw2 = fw3;
    }
    // main.comp:88-108: lo > 0 empty, hi < 0 full, else ambiguous
    const bool full0 = shi0[0] < 0.0f, empty0 = slo0[0] > 0.0f, amb0 = !full0 && !empty0;
    const bool full1 = shi1[0] < 0.0f, empty1 = slo1[0] > 0.0f, amb1 = !full1 && !empty1;
    const uint64_t Bfull0 = __ballot(full0), Bamb0 = __ballot(amb0);
    const uint64_t Bfull1 = __ballot(full1), Bamb1 = __ballot(amb1);
    store_mask64(rec + 4 + sub * 4, Bfull0, lane);
    store_mask64(rec + 4 + sub * 4 + 2, Bfull1, lane);
    store_mask64(rec + 4 + W + sub * 4, Bamb0, lane);
    store_mask64(rec + 4 + W + sub * 4 + 2, Bamb1, lane);

    // group 0's children first
    const uint64_t Bambs[2] = {Bamb0, Bamb1};
    uint32_t child[2];
    shorten_children<MAXS, 2>(a, lane, toff, n, mm, scr, logL, cap, Bambs, sub * 128, ncx,
                              ncy, ncz, child_base, reserve, child);
    rec[4 + 2 * W + idx0] = child[0];
    rec[4 + 2 * W + idx1] = child[1];
    __builtin_amdgcn_s_setprio(0);
}

// Node count of `level` and first pool word of level+1's tapes (tapes of
// level+1 start after those of levels 1..level).
__device__ __forceinline__ uint64_t child_tape_base(const uint32_t* counters, uint32_t level) {
    uint64_t b = load_counter(counters + CNT_BASE1);
    for (uint32_t j = 1; j <= level; ++j) b += alloc_words(load_alloc(counters, j));
    return b;
}

// One-cell chunks with slot files of up to 12 slots are built for 7
// wavefronts per SIMD (at most 72 VGPRs and a scalar budget to match), two-
// cell chunks for 4; larger files keep the compiler's choice.
template <int MAXS, int KC = 1>
__global__ __launch_bounds__(kLevelBlock)
__attribute__((amdgpu_waves_per_eu(KC == 2 ? 4 : (MAXS <= 12 ? 7 : 1)))) void k_level_interval(LevelArgs a) {
    // min/max choice ballots of this wavefront's current chunk: mm_stride rows
    // of {FIRST lanes, SECOND lanes} (two 64-bit masks) per wavefront
    extern __shared__ uint4 lds_choice[];
    if ((load_counter(a.counters + CNT_OVERFLOW) & ovf_before(a.level)) != 0) return;
    const uint32_t lane = lane_id();
    const uint32_t wib = uni(threadIdx.x >> 6);
    uint32_t n_nodes = level_nodes(a.counters, a.level);
    if (n_nodes > a.node_cap) n_nodes = a.node_cap;
    uint64_t child_base = child_tape_base(a.counters, a.level);
    child_base = ((uint64_t)uni((uint32_t)(child_base >> 32)) << 32) | uni((uint32_t)child_base);
    const uint32_t cpn = a.dim * a.dim * a.dim / (64 * KC);
    const uint32_t total = n_nodes * cpn;
    uint4* scr = lds_choice + (size_t)wib * a.mm_stride * KC;
    // the shortening log after every wavefront's choice rows (log_entry_bytes)
    uint2* logL = (uint2*)(lds_choice + (size_t)(kLevelBlock / 64) * a.mm_stride * KC) +
                  (size_t)wib * a.log_cap * KC;
    // rounds of kLevelBlock/64 consecutive chunks per block, one reservation
    // atomic per round; a wavefront past the end still takes part
    constexpr uint32_t WPB = kLevelBlock / 64;
    __shared__ unsigned long long blk_req[2][WPB], blk_base[2];
    unsigned long long* ctr = (unsigned long long*)(a.counters + CNT_ALLOC + 2 * (a.level + 1));
    uint32_t par = 0;
    for (uint32_t g = blockIdx.x; g * WPB < total; g += gridDim.x, par ^= 1u) {
        const BlockReserve br{blk_req[par], &blk_base[par], wib};
        const uint32_t chunk = g * WPB + wib;
        if (chunk < total) {
            const uint32_t node = chunk / cpn, sub = chunk - node * cpn;
            if constexpr (KC == 2)
                interval_chunk2<MAXS>(a, node, sub, child_base, lane, scr, logL, a.log_cap,
                                      br);
            else
                interval_chunk<MAXS>(a, node, sub, child_base, lane, scr, logL, a.log_cap,
                                     br);
        } else {
            br(ctr, 0, lane);
        }
    }
}

// ---------------------------------------------------------------------------
// Root level, parallel over the tape's instructions.  The root is one node
// of dim^3 cells evaluated with the full tape; the per-cell kernel above
// gives it dim^3/64 wavefronts, each walking the whole tape one instruction
// at a time.  Here a group of LPC lanes owns one cell (64/LPC cells per
// wavefront) and evaluates up to LPC independent instructions together (the
// same interval rules, interval.comp:36-160, on the same operand values).
//
// The host renames the tape's slots into values and groups the
// instructions the result depends on into passes of at most LPC, so that
// every operand of a pass was written by an earlier pass (or is an input or
// a constant).  Values live in LDS slots, reused once a value's last reader
// has run; constants come from a table the block loads once.
//
// Shortening (shorten.comp:164-223) keeps instruction i iff slot out(i) is
// active when the backward pass reaches i: iff a kept later instruction
// reads, through its min/max choice, the value i wrote.  Here that is the
// same walk over the passes in reverse: a lane reads and clears the live
// flag of its output slot (the slot's later values were handled in later
// passes), and if it was set, marks the instruction kept and sets the flags
// of its chosen operands.  Kept instructions are then written in tape
// order, LPC at a time, by ballot and prefix-sum compaction.
//
// A block of C cells owns C consecutive mask bits (a 16- or 8-bit slice of
// a mask word, stored as such) and reserves the child ids and the exact
// tape words of its ambiguous cells ([k, k words] each, none where nothing
// was removed) with one atomic once their backward passes have counted k.
//
// LDS: constant table {c, c}[n_const] then instruction ids u16[n_const];
// per cell, values float2[n_slot], flags u8[n] (choice | kept << 2, rounded
// to 4 bytes), live u8[n_slot].
// ---------------------------------------------------------------------------
template <int C>
__device__ __forceinline__ void store_mask_slice(__attribute__((address_space(1))) uint32_t* words,
                                                 uint32_t slice, uint32_t bits) {
    if constexpr (C == 16)
        ((__attribute__((address_space(1))) uint16_t*)words)[slice] = (uint16_t)bits;
    else
        ((__attribute__((address_space(1))) uint8_t*)words)[slice] = (uint8_t)bits;
}

// Block `blk` of one tree's root: cells blk*C .. blk*C + C - 1.
template <int C, int LPC>
__device__ __forceinline__ void root_block(const RootArgs& r, uint32_t blk) {
    static_assert(C == 8 || C == 16, "mask slices are bytes or half-words");
    static_assert(LPC == 16 || LPC == 64, "lanes per cell");
    constexpr uint32_t G = 64 / LPC;                     // cells per wavefront
    extern __shared__ __align__(16) unsigned char root_lds[];
    __shared__ uint32_t blk_flags[C];
    __shared__ uint32_t blk_need[C];                     // words of each cell's tape, then its offset
    __shared__ uint32_t blk_amb, blk_words;
    __shared__ unsigned long long blk_old;
    const LevelArgs& a = r.a;
    // global-typed views of the argument pointers: plain (not flat) memory
    // instructions also when the arguments were read from memory
    typedef __attribute__((address_space(1))) uint32_t gword;
    const gword* gsched = (const gword*)r.sched;
    const gword* gtape = (const gword*)r.tape;
    const __attribute__((address_space(1))) float* gconsts =
        (const __attribute__((address_space(1))) float*)a.consts;
    const uint32_t lane = lane_id();
    const uint32_t sl = lane % LPC;                      // lane within the cell's group
    const uint32_t gshift = lane - sl;                   // first lane of the group
    const uint32_t cib = uni(threadIdx.x >> 6) * G + lane / LPC;   // cell within the block
    const uint32_t idx = blk * C + cib;                  // mask bit of this cell
    const uint32_t n = r.n;
    float2* Kv = (float2*)root_lds;
    uint16_t* Ki = (uint16_t*)(Kv + r.n_const);
    unsigned char* cell = root_lds + r.table_bytes + (size_t)cib * r.cell_bytes;
    float2* V = (float2*)cell;
    uint8_t* flags = cell + 8u * r.n_slot;
    uint8_t* live = flags + ((n + 3u) & ~3u);
    auto opnd = [&](uint32_t v) -> float2 {
        const float2* q = (v & kRootConst) ? Kv + (v & (kRootConst - 1u)) : V + v;
        return *q;
    };
    auto mark = [&](uint32_t v) {
        if (v & kRootConst) flags[Ki[v & (kRootConst - 1u)]] = 4u;
        else live[v] = 1u;
    };
    // the group's bits of a ballot, at bit 0
    auto group_bits = [&](uint64_t b) -> uint64_t {
        return LPC == 64 ? b : (b >> gshift) & ((1ull << (LPC % 64)) - 1ull);
    };

    // constant table: kDeep entries per thread in flight (entry, then the
    // constant it names)
    constexpr uint32_t kDeep = 8;
    for (uint32_t k0 = threadIdx.x; k0 < r.n_const; k0 += kDeep * C * LPC) {
        uint32_t e[kDeep];
        float c[kDeep];
#pragma unroll
        for (uint32_t u = 0; u < kDeep; ++u) {
            const uint32_t k = k0 + u * C * LPC;
            e[u] = gsched[r.c0 + (k < r.n_const ? k : r.n_const - 1u)];   // i | pool index << 16
        }
#pragma unroll
        for (uint32_t u = 0; u < kDeep; ++u)
            c[u] = gconsts[e[u] >> 16];
#pragma unroll
        for (uint32_t u = 0; u < kDeep; ++u) {
            const uint32_t k = k0 + u * C * LPC;
            if (k < r.n_const) {
                Kv[k] = make_float2(c[u], c[u]);
                Ki[k] = (uint16_t)(e[u] & 0xFFFFu);
            }
        }
    }
    const uint32_t dim = a.dim, W = dim * dim * dim / 32;
    gword* rec = (gword*)a.nodes;
    ctape_t hdr = as_const(a.nodes);
    const uint32_t toff = hdr[0], ncx = hdr[1], ncy = hdr[2], ncz = hdr[3];
    uint32_t cx, cy, cz;
    cell_of(idx, dim, a.dim_shift, &cx, &cy, &cz);
    const uint32_t gx = ncx * dim + cx, gy = ncy * dim + cy, gz = ncz * dim + cz;
    // main.comp:113-116, 124-126: box [pos, pos + cell_size]
    const float lx = a.low[0] + (float)gx * a.cs;
    const float ly = a.low[1] + (float)gy * a.cs;
    const float lz = a.low[2] + (float)gz * a.cs;
    for (uint32_t q = sl; q < (n + 3u) / 4u; q += LPC) ((uint32_t*)flags)[q] = 0u;
    for (uint32_t v = sl; v < r.n_slot; v += LPC) live[v] = 0;
    if (sl < 4) {
        const float lo = sl == 0 ? lx : (sl == 1 ? ly : (sl == 2 ? lz : a.tape_time));
        const float hi = sl == 3 ? a.tape_time : lo + a.cs;
        V[sl] = make_float2(lo, hi);
    }
    __syncthreads();                                     // constant table complete

    // ---- forward pass ----
    // Pass p's entry of this lane is ent[p * LPC] (kRootNop where the pass
    // has fewer instructions).  The pass count is a multiple of 4 and each
    // entry is loaded 3 passes ahead, from an index clamped to the last pass
    // (never a conditional address).
    typedef const __attribute__((address_space(1))) unsigned long long gpair;
    gpair* ent = (gpair*)(gsched + r.e0) + sl;  // {op | out slot << 4 | i << 16, A | B << 16}
    const uint32_t P = r.n_pass;
    auto ld = [&](uint32_t p) -> uint2 {
        const unsigned long long v = ent[(p < P ? p : P - 1u) * LPC];
        return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    };
    auto forward = [&](const uint2 e, uint32_t p) {
        // op types of the pass (bit op): the common ones are evaluated for
        // every lane under a uniform test and selected; others take the
        // general rule on their own lanes
        const uint32_t m = gsched[r.m0 + p];
        const uint32_t op = e.x & 0xFu, so = (e.x >> 4) & 0xFFFu, i = e.x >> 16;
        const float2 av = opnd(e.y & 0xFFFFu), bv = opnd(e.y >> 16);
        const Itv A{av.x, av.y}, B{bv.x, bv.y};
        Itv res = A;                                     // COPY
        if (m & (1u << OP_ADD)) {
            const Itv t{A.lo + B.lo, A.hi + B.hi};
            if (op == OP_ADD) res = t;
        }
        if (m & (1u << OP_SUB)) {
            const Itv t{A.lo - B.hi, A.hi - B.lo};
            if (op == OP_SUB) res = t;
        }
        if (m & (1u << OP_MUL)) {
            const Itv t = imul(A, B);
            if (op == OP_MUL) res = t;
        }
        if (m & ((1u << OP_MIN) | (1u << OP_MAX))) {
            const bool mn = op == OP_MIN;
            const Itv t = mn ? Itv{fmin_glm(A.lo, B.lo), fmin_glm(A.hi, B.hi)}
                             : Itv{fmax_glm(A.lo, B.lo), fmax_glm(A.hi, B.hi)};
            if (mn || op == OP_MAX) {
                res = t;
                flags[i] = (uint8_t)mm_choice(op, A, B);
            }
        }
        if (m & (1u << OP_NEG)) {
            const Itv t{-A.hi, -A.lo};
            if (op == OP_NEG) res = t;
        }
        if (m & (1u << OP_SQRT)) {
            const Itv t{__builtin_sqrtf(fmax_glm(0.0f, A.lo)), __builtin_sqrtf(fmax_glm(0.0f, A.hi))};
            if (op == OP_SQRT) res = t;
        }
        if (m & ((1u << OP_SIN) | (1u << OP_COS) | (1u << OP_DIV) | (1u << OP_EXP))) {
            if (op == OP_SIN || op == OP_COS || op == OP_DIV || op == OP_EXP)
                res = interval_op(op, A, B, 0.0f);
        }
        if (op != kRootNop) V[so] = make_float2(res.lo, res.hi);
    };
    if (P) {
        uint2 q0 = ld(0), q1 = ld(1), q2 = ld(2), q3 = ld(3);
        // with balance on, a wavefront steps its issue priority down as it
        // progresses (3 at the start, 2 past half the forward passes, 1 after
        // them, 0 past half the backward ones), so the SIMD's wavefronts
        // move together instead of oldest first
        const uint32_t half = (P / 2u) & ~3u;
        if (a.balance) __builtin_amdgcn_s_setprio(3);
        // each entry is reloaded once its pass has used it, into the same
        // registers (no copies that would wait for the loads in flight)
        for (uint32_t p = 0; p < P; p += 4) {
            if (a.balance && p == half) __builtin_amdgcn_s_setprio(2);
            forward(q0, p);
            q0 = ld(p + 4);
            forward(q1, p + 1);
            q1 = ld(p + 5);
            forward(q2, p + 2);
            q2 = ld(p + 6);
            forward(q3, p + 3);
            q3 = ld(p + 7);
        }
    }
    if (a.balance) __builtin_amdgcn_s_setprio(1);
    const float2 R = opnd(r.res);
    // main.comp:88-108: lo > 0 empty, hi < 0 full, else ambiguous
    const bool full = R.y < 0.0f;
    const bool empty = R.x > 0.0f;
    const bool amb = !full && !empty;

    // ---- backward pass: instructions the result reaches through its choices ----
    uint32_t k = 0;                                      // kept instructions of this cell
    if (amb) {
        if (sl == 0) mark(r.res);
        // passes P-1 .. 0, entries loaded 3 passes ahead
        auto backward = [&](const uint2 e) {
            if ((e.x & 0xFu) != kRootNop) {
                const uint32_t op = e.x & 0xFu, so = (e.x >> 4) & 0xFFFu, i = e.x >> 16;
                const uint32_t lv = live[so];
                live[so] = 0;
                if (lv) {
                    const uint32_t f = flags[i];
                    flags[i] = (uint8_t)(f | 4u);
                    uint32_t sel = MM_FIRST;
                    if (op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV)
                        sel = MM_BOTH;
                    else if (op == OP_MIN || op == OP_MAX)
                        sel = f & 3u;
                    if (sel & MM_FIRST) mark(e.y & 0xFFFFu);
                    if (sel & MM_SECOND) mark(e.y >> 16);
                }
            }
        };
        if (P) {
            auto ldr = [&](uint32_t t) { return ld(t < P ? P - 1u - t : 0u); };
            uint2 q0 = ldr(0), q1 = ldr(1), q2 = ldr(2), q3 = ldr(3);
            const uint32_t half = (P / 2u) & ~3u;
            for (uint32_t t = 0; t < P; t += 4) {
                if (a.balance && t == half) __builtin_amdgcn_s_setprio(0);
                backward(q0);
                q0 = ldr(t + 4);
                backward(q1);
                q1 = ldr(t + 5);
                backward(q2);
                q2 = ldr(t + 6);
                backward(q3);
                q3 = ldr(t + 7);
            }
        }
        for (uint32_t s = 0; s < n; s += LPC)
            k += (uint32_t)__popcll(group_bits(
                __ballot(s + sl < n && (flags[s + sl < n ? s + sl : n - 1u] & 4u) != 0)));
    }

    // ---- block: mask slices and one exact reservation for the block's
    // children: a child id per ambiguous cell and [k, k words] per shortened
    // tape (none when nothing was removed: the child reuses the root tape,
    // shorten.comp:188-191) ----
    if (sl == 0) {
        blk_flags[cib] = (full ? 1u : 0u) | (amb ? 2u : 0u);
        blk_need[cib] = (amb && k < n) ? k + 1u : 0u;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t f = lane < C ? blk_flags[lane] : 0u;
        const uint64_t bf = __ballot((f & 1u) != 0), ba = __ballot((f & 2u) != 0);
        const uint64_t namb = (uint64_t)__popcll(ba);
        const uint32_t nd = lane < C ? blk_need[lane] : 0u;
        const uint32_t incl = wave_incl_scan(nd, lane);
        const uint64_t words = rdlane(incl, 63);
        if (lane < C) blk_need[lane] = incl - nd;        // the cell's offset in the block's words
        if (lane == 0) {
            store_mask_slice<C>(rec + 4, blk, (uint32_t)bf);
            store_mask_slice<C>(rec + 4 + W, blk, (uint32_t)ba);
            blk_amb = (uint32_t)ba;
            blk_words = (uint32_t)words;
            unsigned long long* ctr = (unsigned long long*)(a.counters + CNT_ALLOC + 2);
            blk_old = namb ? atomicAdd(ctr, (namb << kAllocShift) | words) : 0ull;
        }
    }
    __syncthreads();

    uint32_t child_id = kNoChild;
    if (amb) {
        const uint32_t bam = blk_amb;
        const uint32_t namb = (uint32_t)__popc(bam);
        const uint32_t rank = (uint32_t)__popc(bam & ((1u << cib) - 1u));
        const uint64_t old = blk_old;
        const uint32_t cbase = alloc_nodes(old);
        const uint64_t pbase = child_tape_base(a.counters, 0) + alloc_words(old);
        const bool nodes_ok = (uint64_t)cbase + namb <= a.child_cap;
        const bool pool_ok = pbase + blk_words <= a.pool_cap;
        if (!(nodes_ok && pool_ok)) {
            if (sl == 0)
                atomicOr(a.counters + CNT_OVERFLOW,
                         (nodes_ok ? 0u : OVF_NODES) | (pool_ok ? 0u : OVF_POOL));
        } else {
            // ---- kept instructions in tape order (shorten.comp:61-88, 188-191) ----
            const uint64_t mybase = pbase + blk_need[cib];
            uint32_t final_off = toff;
            if (k < n) {
                gword* out = (gword*)a.pool + mybase;
                if (sl == 0) out[0] = k;
                // the tape words of kDeep windows are loaded together
                uint32_t j = 1;
                for (uint32_t s0 = 0; s0 < n; s0 += kDeep * LPC) {
                    uint32_t words[kDeep];
#pragma unroll
                    for (uint32_t u = 0; u < kDeep; ++u) {
                        const uint32_t i = s0 + u * LPC + sl;
                        words[u] = gtape[1 + (i < n ? i : n - 1u)];
                    }
#pragma unroll
                    for (uint32_t u = 0; u < kDeep; ++u) {
                    const uint32_t i = s0 + u * LPC + sl;
                    const uint32_t f = flags[i < n ? i : n - 1u];
                    const bool kp = i < n && (f & 4u) != 0;
                    const uint64_t b = group_bits(__ballot(kp));
                    if (kp) {
                        uint32_t wd = words[u];
                        const uint32_t op = wd & 0xFFu;
                        if (op == OP_MIN || op == OP_MAX) {
                            const uint32_t c = f & 3u;
                            if (c == MM_FIRST) {
                                wd = (wd & ~0xFFu) | OP_COPY;
                            } else if (c == MM_SECOND) {
                                const uint32_t o = (wd >> 8) & 0xFFu, ib = wd >> 24;
                                wd = OP_COPY | (o << 8) | (ib << 16) | (ib << 24);
                            }
                        }
                        out[j + (uint32_t)__popcll(b & ((1ull << sl) - 1ull))] = wd;
                    }
                    j += (uint32_t)__popcll(b);
                    }
                }
                final_off = (uint32_t)mybase;
            }
            child_id = cbase + rank;
            if (sl == 0) {
                gword* crec = (gword*)a.child_nodes + (size_t)child_id * a.child_ns;
                crec[0] = final_off;
                crec[1] = gx;
                crec[2] = gy;
                crec[3] = gz;
            }
        }
    }
    if (sl == 0) rec[4 + 2 * W + idx] = child_id;
    __builtin_amdgcn_s_setprio(0);
}

template <int C, int LPC>
__global__ __launch_bounds__(C * LPC) void k_root_interval(RootArgs r) {
    root_block<C, LPC>(r, blockIdx.x);
}

__device__ __forceinline__ RootArgs load_root_args(const RootArgs* p) {
    static_assert(sizeof(RootArgs) % 4 == 0, "RootArgs is copied as words");
    typedef const __attribute__((address_space(4))) uint32_t* cw_t;
    const cw_t w = (cw_t)p;
    RootArgs r;
    uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(RootArgs) / 4; ++i) d[i] = w[i];
    return r;
}

// Roots of several trees in one launch: tree j owns blocks
// starts[j] .. starts[j+1]-1 (n <= 64 trees, one ballot finds j).
template <int C, int LPC>
__global__ __launch_bounds__(C * LPC) void k_root_interval_batch(const RootArgs* args,
                                                                 const uint32_t* starts,
                                                                 uint32_t n) {
    const uint32_t lane = lane_id();
    const uint32_t end = lane < n ? starts[lane + 1] : 0xFFFFFFFFu;
    const uint32_t j = uni((uint32_t)__popcll(__ballot(blockIdx.x >= end)));
    const RootArgs r = load_root_args(args + j);
    root_block<C, LPC>(r, blockIdx.x - starts[j]);
}

// ---------------------------------------------------------------------------
// Last level: point evaluation at each voxel's low corner, leaf iff < 0
// (vulkan/shaders/voxelizer/main.comp:74-83).  Each lane evaluates K voxels
// of the same node (same tape) so instruction decode is shared K ways.
// ---------------------------------------------------------------------------
// Chunk `sub` of a last-level node through the assembly interpreter (8
// voxels per lane), given the node's header words and the tape's window 0
// with its constants (w0, c0 as run_point_tape_asm8 takes them).  With PF
// the interpreter also loads pf_win[lane], pf_len[0] and pf_hdr[0..3] into
// *w1, *n1 and h2 while it runs, and gathers the constants of the window
// nx_w (nx_n words) into *nx_c.
template <int MAXS, bool PF = false>
__device__ __forceinline__ void point_chunk_asm(const LevelArgs& a, uint32_t node, uint32_t sub,
                                                uint32_t lane, uint32_t toff, uint32_t ncx,
                                                uint32_t ncy, uint32_t ncz, uint32_t w0,
                                                float c0, const uint32_t* pf_win = nullptr,
                                                const uint32_t* pf_len = nullptr,
                                                const uint32_t* pf_hdr = nullptr,
                                                uint32_t* w1 = nullptr, uint32_t* n1 = nullptr,
                                                uint32_t* h2 = nullptr, uint32_t n = 0,
                                                uint32_t nx_w = 0, uint32_t nx_n = 0,
                                                float* nx_c = nullptr) {
    const uint32_t dim = a.dim;
    uint32_t* rec = a.nodes + (size_t)node * a.ns;
    v8f xs, ys, zs;
    // main.comp:77-79: grid low + float(node.coords * dim + cell) * cell size
    if (dim == 8) {
        // one chunk per node: bit index k*64 + lane is cell (k, lane / 8,
        // lane % 8), so y and z are the same for the lane's 8 voxels
        const float y = a.low[1] + (float)(ncy * 8u + (lane >> 3)) * a.cs;
        const float z = a.low[2] + (float)(ncz * 8u + (lane & 7u)) * a.cs;
        // x of voxel k is the same in every lane: lanes 0..7 of each row of
        // 16 compute the eight values, and a DPP row_newbcast:k move (lane k
        // of its row to the whole row) fills register k
        const int xl = __builtin_bit_cast(int, a.low[0] + (float)(ncx * 8u + (lane & 7u)) * a.cs);
#define FREP_XBCAST(k) \
    xs[k] = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(xl, 0x150 + k, 0xF, 0xF, false));
        FREP_XBCAST(0) FREP_XBCAST(1) FREP_XBCAST(2) FREP_XBCAST(3)
        FREP_XBCAST(4) FREP_XBCAST(5) FREP_XBCAST(6) FREP_XBCAST(7)
#undef FREP_XBCAST
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ys[k] = y;
            zs[k] = z;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t idx = (sub * 8 + (uint32_t)k) * 64 + lane;
            uint32_t cx, cy, cz;
            cell_of(idx, dim, a.dim_shift, &cx, &cy, &cz);
            xs[k] = a.low[0] + (float)(ncx * dim + cx) * a.cs;
            ys[k] = a.low[1] + (float)(ncy * dim + cy) * a.cs;
            zs[k] = a.low[2] + (float)(ncz * dim + cz) * a.cs;
        }
    }
    v8f d;
    if constexpr (PF)
        d = run_point_tape_asm8_pf2<MAXS>(a.pool + toff, a.consts, a.tape_time, xs, ys, zs, w0, c0,
                                          n, pf_win, pf_len, pf_hdr, w1, n1, h2, nx_w, nx_n,
                                          nx_c, a.reads_t);
    else
        d = run_point_tape_asm8<MAXS>(a.pool + toff, a.consts, a.tape_time, xs, ys, zs, w0, c0,
                                      a.reads_t);
    uint64_t B[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) B[k] = __ballot(d[k] < 0.0f);
    store_mask64x8(rec + 4 + sub * 16, B, lane);
}

template <int MAXS, int K, bool ASM>
__device__ __forceinline__ void point_chunk(const LevelArgs& a, uint32_t node, uint32_t sub,
                                            uint32_t lane) {
    static_assert(!ASM || (K == 8 && (MAXS == 8 || MAXS == 10 || MAXS == 16)) ||
                      (K == 1 && (MAXS == 8 || MAXS == 16)),
                  "assembly path shape");
    const uint32_t dim = a.dim;
    uint32_t* rec = a.nodes + (size_t)node * a.ns;
    ctape_t hdr = as_const(rec);
    const uint32_t toff = hdr[0], ncx = hdr[1], ncy = hdr[2], ncz = hdr[3];
    auto init = [&](int k) {
        const uint32_t idx = (sub * K + (uint32_t)k) * 64 + lane;
        uint32_t cx, cy, cz;
        cell_of(idx, dim, a.dim_shift, &cx, &cy, &cz);
        // main.comp:77-79: grid low + float(node.coords * dim + cell) * cell size
        return make_float3(a.low[0] + (float)(ncx * dim + cx) * a.cs,
                           a.low[1] + (float)(ncy * dim + cy) * a.cs,
                           a.low[2] + (float)(ncz * dim + cz) * a.cs);
    };
    auto out = [&](int k, float d) {
        const uint64_t B = __ballot(d < 0.0f);
        store_mask64(rec + 4 + (sub * K + (uint32_t)k) * 2, B, lane);
    };
    if constexpr (ASM && K == 8) {
        uint32_t w0;
        float c0;
        point_window0(a.pool + toff, a.consts, lane, &w0, &c0);
        point_chunk_asm<MAXS>(a, node, sub, lane, toff, ncx, ncy, ncz, w0, c0);
    } else if constexpr (ASM) {
        uint32_t w0;
        float c0;
        point_window0(a.pool + toff, a.consts, lane, &w0, &c0);
        const float3 p = init(0);
        out(0, run_point_tape_asm1<MAXS>(a.pool + toff, a.consts, a.tape_time, p.x, p.y, p.z, w0,
                                         c0, a.reads_t));
    } else {
        run_point_tape<MAXS, K>(as_const(a.pool + toff), as_const(a.consts), a.tape_time, init,
                                out);
    }
}

template <int MAXS, int K, bool ASM>
__device__ __forceinline__ void level_point_body(const LevelArgs& a) {
    if ((load_counter(a.counters + CNT_OVERFLOW) & ovf_before(a.level)) != 0) return;
    const uint32_t lane = lane_id();
    const uint32_t gwave = blockIdx.x * (kLevelBlock / 64) + uni(threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (kLevelBlock / 64);
    uint32_t n_nodes = level_nodes(a.counters, a.level);
    if (n_nodes > a.node_cap) n_nodes = a.node_cap;
    const uint32_t cpw = a.dim * a.dim * a.dim / (64 * K);
    const uint32_t total = n_nodes * cpw;
    if constexpr (ASM && K == 8) {
        // Software pipeline over this wavefront's chunks c, c1, c2, c3: while
        // the interpreter runs chunk c it gathers c1's constants (from c1's
        // window, loaded during the previous chunk), loads c2's tape window
        // (64 words from the tape's first instruction; the pool has 64 words
        // of padding past its end) and length, and c3's header.
        uint32_t c = gwave;
        if (c >= total) return;
        // node of a chunk: one chunk per node for 8-wide last levels (the
        // common case) skips the integer division
        auto node_of = [&](uint32_t ch) { return cpw == 1 ? ch : ch / cpw; };
        auto header = [&](uint32_t ch, uint32_t h[4]) {
            ctape_t r = as_const(a.nodes + (size_t)node_of(ch) * a.ns);
            h[0] = r[0]; h[1] = r[1]; h[2] = r[2]; h[3] = r[3];
        };
        uint32_t h0[4], h1[4], h2[4];
        header(c, h0);
        header(c + nwaves < total ? c + nwaves : c, h1);
        header(c + 2 * nwaves < total ? c + 2 * nwaves : c, h2);
        uint32_t w0, n0;
        float c0;
        point_window0(a.pool + h0[0], a.consts, lane, &w0, &c0, &n0);
        uint32_t w1 = a.pool[h1[0] + 1 + lane];
        uint32_t n1 = c + nwaves < total ? as_const(a.pool + h1[0])[0] : 0u;
        // A SIMD issues from its oldest wavefront first, so with equal shares
        // its wavefronts would finish one after another and the last ones run
        // alone.  With balance on, a wavefront steps its priority down from 3
        // as it completes each quarter of its share and the others catch up.
        const uint32_t q =
            a.balance ? uni(max(1u, (total + a.balance * nwaves - 1u) / (a.balance * nwaves))) : 0u;
        if (q) __builtin_amdgcn_s_setprio(3);
        uint32_t done = 0;
        for (;;) {
            const uint32_t c1 = c + nwaves, c2 = c1 + nwaves, c3 = c2 + nwaves;
            const uint32_t* pf_len = a.pool + (c2 < total ? h2[0] : h0[0]);
            const uint32_t* pf_hdr = a.nodes + (size_t)node_of(c3 < total ? c3 : c) * a.ns;
            uint32_t w2, n2, h3[4];
            float c0n;
            const uint32_t node = node_of(c);
            point_chunk_asm<MAXS, true>(a, node, c - node * cpw, lane, h0[0], h0[1], h0[2], h0[3],
                                        w0, c0, pf_len + 1, pf_len, pf_hdr, &w2, &n2, h3, n0, w1,
                                        n1, &c0n);
            if (c1 >= total) break;
            ++done;
            if (q) {
                if (done == q) __builtin_amdgcn_s_setprio(2);
                else if (done == 2u * q) __builtin_amdgcn_s_setprio(1);
                else if (done == 3u * q) __builtin_amdgcn_s_setprio(0);
            }
            // the assembly's scalar outputs are wave-uniform: say so, so that
            // they stay in SGPRs for the next block's scalar inputs
            w0 = w1;
            n0 = n1;
            c0 = c0n;
            w1 = w2;
            n1 = c2 < total ? uni(n2) : 0u;
            c = c1;
            h0[0] = h1[0]; h0[1] = h1[1]; h0[2] = h1[2]; h0[3] = h1[3];
            h1[0] = h2[0]; h1[1] = h2[1]; h1[2] = h2[2]; h1[3] = h2[3];
            h2[0] = uni(h3[0]); h2[1] = uni(h3[1]); h2[2] = uni(h3[2]); h2[3] = uni(h3[3]);
        }
        __builtin_amdgcn_s_setprio(0);
    } else {
        for (uint32_t chunk = gwave; chunk < total; chunk += nwaves) {
            const uint32_t node = chunk / cpw, sub = chunk - node * cpw;
            point_chunk<MAXS, K, ASM>(a, node, sub, lane);
        }
    }
}

// End of a tree's last level (a.finish): every block arrives once; the last
// to arrive writes the counters to the host mirror a.host_counters (system
// scope, word CNT_DONE replaced by kCountersWritten so the host can tell the
// copy arrived) and, when the build did not overflow, resets them to what
// k_init_tree writes for the same tape (all zero but CNT_BASE1), so the next
// build of that tape into the same buffers needs no init launch.  The
// arrival count CNT_DONE is zero again afterwards either way.  No fence on
// arrival: the counters are read-only during this launch (its waves read
// them before arriving), and the reset is visible to later launches at the
// launch boundary.
__device__ __forceinline__ void finish_tree(const LevelArgs& a) {
    __shared__ uint32_t last;
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(a.counters + CNT_DONE, 1u) == gridDim.x - 1 ? 1u : 0u;
    __syncthreads();
    if (last == 0 || threadIdx.x >= 64) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t v = lane < CNT_WORDS ? a.counters[lane] : 0u;
    const uint32_t ovf = (uint32_t)__shfl((int)v, CNT_OVERFLOW);
    if (lane < CNT_WORDS) {
        __hip_atomic_store(a.host_counters + lane, lane == CNT_DONE ? kCountersWritten : v,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if ((ovf == 0 && lane != CNT_BASE1) || lane == CNT_DONE) a.counters[lane] = 0u;
    }
    __threadfence_system();
}

template <int MAXS, int K, bool ASM = false>
__global__ __launch_bounds__(kLevelBlock) void k_level_point(LevelArgs a) {
    level_point_body<MAXS, K, ASM>(a);
    if (a.finish) finish_tree(a);
}

// ---------------------------------------------------------------------------
// Batched levels: one persistent grid walks the chunks of many shapes (each
// with its own tree, tape pool, constants, dims and resolution).  Every
// wavefront first builds the plan in registers -- lane s holds the inclusive
// prefix sum of the chunk counts of shapes 0..s, read from each tree's device
// counters -- and maps a global chunk index to its shape with one ballot.
// The shape's arguments are reloaded (scalar loads) only when a wavefront
// moves on to another shape.
// ---------------------------------------------------------------------------
namespace {

typedef const __attribute__((address_space(4))) LevelArgs* cargs_t;

// Chunks of shape `lane` at this launch's level, and the first pool word of
// its child tapes.  Only counts that no wavefront of this launch modifies
// enter the plan (the node count of this level was final when the previous
// launch ended), so every wavefront maps chunk indices to shapes the same
// way; overflow flags, which this launch may set, are checked per shape
// while walking the chunks.
// Cells per lane of a shape whose nodes hold `cells` cells in an interval
// launch of kc cells per lane: two only for nodes of a multiple of 128 cells.
__device__ __forceinline__ uint32_t interval_kc(uint32_t kc, uint32_t cells) {
    return kc == 2 && cells % 128u == 0 ? 2u : 1u;
}

template <bool POINT>
__device__ __forceinline__ uint32_t plan_lane(const BatchArgs& b, uint32_t lane, uint64_t* cbase) {
    *cbase = 0;
    if (lane >= b.n_shapes) return 0u;
    const LevelArgs* A = b.shapes + lane;
    const uint32_t* cnt = A->counters;
    uint32_t nn = A->level == 0 ? 1u : alloc_nodes(load_alloc(cnt, A->level));
    nn = nn < A->node_cap ? nn : A->node_cap;
    const uint32_t cells = A->dim * A->dim * A->dim;
    if (!POINT) *cbase = child_tape_base(cnt, A->level);
    return nn * (POINT ? cells / (64u * A->pk) : cells / (64u * interval_kc(b.kc, cells)));
}

// Scalar (constant address space) copy of one shape's arguments.
__device__ __forceinline__ LevelArgs load_args(cargs_t p) {
    static_assert(sizeof(LevelArgs) % 4 == 0, "LevelArgs is copied as words");
    typedef const __attribute__((address_space(4))) uint32_t* cw_t;
    const cw_t w = (cw_t)p;
    LevelArgs a;
    uint32_t* d = reinterpret_cast<uint32_t*>(&a);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(LevelArgs) / 4; ++i) d[i] = w[i];
    return a;
}

__device__ __forceinline__ uint64_t rdlane64(uint64_t v, uint32_t l) {
    return ((uint64_t)rdlane((uint32_t)(v >> 32), l) << 32) | rdlane((uint32_t)v, l);
}

}  // namespace

template <int MAXS, int KC = 1>
__global__ __launch_bounds__(kLevelBlock)
__attribute__((amdgpu_waves_per_eu(KC == 2 ? 4 : (MAXS <= 12 ? 7 : 1)))) void k_level_interval_batch(BatchArgs b) {
    extern __shared__ uint4 lds_choice[];
    const uint32_t lane = lane_id();
    const uint32_t wib = uni(threadIdx.x >> 6);
    const uint32_t gwave = blockIdx.x * (kLevelBlock / 64) + wib;
    const uint32_t nwaves = gridDim.x * (kLevelBlock / 64);
    uint64_t cbase;
    const uint32_t incl = wave_incl_scan(plan_lane<false>(b, lane, &cbase), lane);
    const uint32_t total = rdlane(incl, 63);
    uint4* scr = lds_choice + (size_t)wib * b.mm_stride * KC;
    uint2* logL = (uint2*)(lds_choice + (size_t)(kLevelBlock / 64) * b.mm_stride * KC) +
                  (size_t)wib * b.log_cap * KC;
    cargs_t shapes = (cargs_t)b.shapes;
    uint32_t cur = 0xFFFFFFFFu, start = 0, cpn = 1, kcs = 1;
    uint64_t cb = 0;
    bool skip = false;
    LevelArgs a{};
    for (uint32_t chunk = gwave; chunk < total; chunk += nwaves) {
        const uint32_t s = (uint32_t)__popcll(__ballot(incl <= chunk));
        if (s != cur) {
            cur = s;
            a = load_args(shapes + s);
            start = s ? rdlane(incl, s - 1) : 0u;
            cb = rdlane64(cbase, s);
            kcs = interval_kc(KC, a.dim * a.dim * a.dim);
            cpn = a.dim * a.dim * a.dim / (64 * kcs);
            // a shape that ran out of capacity at an earlier level is rebuilt
            // by the host; one that runs out during this launch goes on, so
            // its counters report what the whole level needs
            skip = (uni(load_counter(a.counters + CNT_OVERFLOW)) & ovf_before(a.level)) != 0;
        }
        if (skip) continue;
        const uint32_t local = chunk - start;
        const uint32_t node = local / cpn, sub = local - node * cpn;
        // a shape of 64-cell nodes in a two-cell launch walks one cell per
        // lane, in the first half of its wavefront's LDS rows and log
        if constexpr (KC == 2) {
            if (kcs == 2) {
                interval_chunk2<MAXS>(a, node, sub, cb, lane, scr, logL, b.log_cap,
                                      WaveReserve{});
                continue;
            }
        }
        interval_chunk<MAXS>(a, node, sub, cb, lane, scr, logL, b.log_cap, WaveReserve{});
    }
}

// The same walk when every shape of the launch takes the assembly
// interpreter (8 voxels per lane, slot file MS): without the C++ paths the
// kernel needs only the assembly path's registers (more wavefronts per SIMD).
template <int MS>
__global__ __launch_bounds__(kLevelBlock) void k_level_point_batch_asm(BatchArgs b) {
    const uint32_t lane = lane_id();
    const uint32_t gwave = blockIdx.x * (kLevelBlock / 64) + uni(threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (kLevelBlock / 64);
    uint64_t unused;
    const uint32_t incl = wave_incl_scan(plan_lane<true>(b, lane, &unused), lane);
    const uint32_t total = rdlane(incl, 63);
    cargs_t shapes = (cargs_t)b.shapes;
    uint32_t cur = 0xFFFFFFFFu, start = 0, cpw = 1;
    bool skip = false;
    LevelArgs a{};
    for (uint32_t chunk = gwave; chunk < total; chunk += nwaves) {
        const uint32_t s = (uint32_t)__popcll(__ballot(incl <= chunk));
        if (s != cur) {
            cur = s;
            a = load_args(shapes + s);
            start = s ? rdlane(incl, s - 1) : 0u;
            cpw = a.dim * a.dim * a.dim / (64u * 8u);
            skip = (uni(load_counter(a.counters + CNT_OVERFLOW)) & ovf_before(a.level)) != 0;
        }
        if (skip) continue;
        const uint32_t local = chunk - start;
        const uint32_t node = local / cpw, sub = local - node * cpw;
        point_chunk<MS, 8, true>(a, node, sub, lane);
    }
}

// SMALL: the launch holds only shapes evaluated with fewer than 8 voxels per
// lane; the 8-voxel paths are left out, and with them most registers.
template <int MAXS, bool SMALL = false>
__global__ __launch_bounds__(kLevelBlock) void k_level_point_batch(BatchArgs b) {
    constexpr int KM = SMALL ? (MAXS <= 32 ? 4 : 1) : (MAXS <= 16 ? 8 : (MAXS <= 32 ? 4 : 1));
    const uint32_t lane = lane_id();
    const uint32_t gwave = blockIdx.x * (kLevelBlock / 64) + uni(threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (kLevelBlock / 64);
    uint64_t unused;
    const uint32_t incl = wave_incl_scan(plan_lane<true>(b, lane, &unused), lane);
    const uint32_t total = rdlane(incl, 63);
    cargs_t shapes = (cargs_t)b.shapes;
    uint32_t cur = 0xFFFFFFFFu, start = 0, cpw = 1;
    bool skip = false;
    LevelArgs a{};
    for (uint32_t chunk = gwave; chunk < total; chunk += nwaves) {
        const uint32_t s = (uint32_t)__popcll(__ballot(incl <= chunk));
        if (s != cur) {
            cur = s;
            a = load_args(shapes + s);
            start = s ? rdlane(incl, s - 1) : 0u;
            cpw = a.dim * a.dim * a.dim / (64u * a.pk);
            skip = (uni(load_counter(a.counters + CNT_OVERFLOW)) & ovf_before(a.level)) != 0;
        }
        if (skip) continue;
        const uint32_t local = chunk - start;
        const uint32_t node = local / cpw, sub = local - node * cpw;
        if constexpr (MAXS <= 16 && !SMALL) {
            if (a.use_asm) {
                point_chunk<MAXS, 8, true>(a, node, sub, lane);
                continue;
            }
        }
        if constexpr (MAXS <= 16 && SMALL) {
            if (a.use_asm) {
                point_chunk<MAXS, 1, true>(a, node, sub, lane);
                continue;
            }
        }
        if constexpr (KM >= 8) {
            if (a.pk == 8) {
                point_chunk<MAXS, 8, false>(a, node, sub, lane);
                continue;
            }
        }
        if constexpr (KM >= 4) {
            if (a.pk == 4) {
                point_chunk<MAXS, 4, false>(a, node, sub, lane);
                continue;
            }
        }
        point_chunk<MAXS, 1, false>(a, node, sub, lane);
    }
}

__global__ void k_init_batch(const InitArgs* args) {
    const InitArgs a = args[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const uint32_t n = a.root_tape[0];
    for (uint32_t i = tid; i <= n; i += blockDim.x) a.pool[a.root_off + i] = a.root_tape[i];
    if (tid < CNT_WORDS)
        a.counters[tid] = tid == CNT_BASE1 ? a.root_off + ((n + 1 + 7u) & ~7u) : 0u;
    if (tid == 0) {
        a.nodes0[0] = a.root_off;
        a.nodes0[1] = a.nodes0[2] = a.nodes0[3] = 0;
    }
}

__global__ void k_gather_counters(const InitArgs* args, uint32_t n, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * CNT_WORDS) out[i] = args[i / CNT_WORDS].counters[i % CNT_WORDS];
}

// ---------------------------------------------------------------------------
// Render: ray generation (raytracer/main.comp:245-267), raygrid_intersect +
// kd-restart dda (main.comp:24-111), gradient shading at the hit
// (main.comp:115-139).  One wavefront per 8x8 pixel tile.
// ---------------------------------------------------------------------------
namespace {

constexpr float kEps = 0.0001f;            // utils.comp:19
constexpr uint32_t kMaxSteps = 1u << 14;

__device__ __forceinline__ uint32_t hash32(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

__device__ __forceinline__ float jitter01(uint32_t seed, uint32_t x, uint32_t y, uint32_t sample,
                                          uint32_t axis) {
    const uint32_t h = hash32(seed ^ hash32(x ^ hash32(y ^ hash32(sample * 2u + axis))));
    return (float)(h >> 8) * 5.9604644775390625e-08f;
}

struct Hit {
    float t;
    uint32_t node, level;
    uint32_t gc[3];
};

__device__ bool dda(const RenderArgs& a, const float o[3], const float d[3], Hit* hit) {
    float inv[3], t0[3], t1[3], sgn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        inv[c] = 1.0f / d[c];
        t0[c] = ((a.low[c] + kEps) - o[c]) * inv[c];
        t1[c] = (((a.low[c] + a.world_size) - kEps) - o[c]) * inv[c];
        if (inv[c] < 0.0f) { const float tmp = t0[c]; t0[c] = t1[c]; t1[c] = tmp; }
        sgn[c] = inv[c] >= 0.0f ? 1.0f : 0.0f;
    }
    const float t_enter = fmax_glm(t0[0], fmax_glm(t0[1], t0[2]));
    const float t_leave = fmin_glm(t1[0], fmin_glm(t1[1], t1[2]));
    if (!(t_enter <= t_leave && t_leave >= 0.0f)) return false;
    float t_curr = fmax_glm(t_enter, 0.0f) + kEps;
    float t_max = t_curr;
    uint32_t steps = 0;
    // the current level's parameters stay in registers: a step's only
    // memory round trip is the cell's words (with the next level's
    // parameters beside them), not the per-lane level table first
    const RenderLevel L0 = a.lv[0];
    while (t_max < t_leave - kEps) {
        uint32_t lvl = 0, node = 0;
        RenderLevel L = L0;
        float np[3] = {a.low[0], a.low[1], a.low[2]};
        t_curr = t_max + kEps;
        t_max = t_leave;
        while (t_curr < t_max - kEps) {
            if (++steps > kMaxSteps) return false;
            const float cs = L.cs;
            const uint32_t dim = L.dim;
            const float s = t_curr + kEps;
            int cell[3];
            float tn[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float p = (o[c] + d[c] * s) - np[c];
                const float norm = p / cs;
                int ci = (int)__builtin_floorf(norm);
                ci = ci < 0 ? 0 : (ci > (int)dim - 1 ? (int)dim - 1 : ci);
                cell[c] = ci;
                tn[c] = s + ((((float)ci + sgn[c]) - norm) * cs) * inv[c];
            }
            const uint32_t idx = (uint32_t)cell[2] + dim * (uint32_t)cell[1] +
                                 dim * dim * (uint32_t)cell[0];
            const uint32_t* rec = L.nodes + (size_t)node * L.ns;
            const uint32_t W = dim * dim * dim / 32;
            const uint32_t bit = 1u << (idx & 31u);
            const bool last = (lvl + 1 == a.n_levels);
            // the cell's leaf word, interior word and child entry are loaded
            // together (one memory round trip per step instead of up to three
            // dependent ones); a last-level record has only the leaf mask
            const uint32_t leaf_w = rec[4 + (idx >> 5)];
            uint32_t int_w = 0, child = 0;
            RenderLevel Ln = L;
            if (!last) {
                int_w = rec[4 + W + (idx >> 5)];
                child = rec[4 + 2 * W + idx];
                Ln = a.lv[lvl + 1];
            }
            if (leaf_w & bit) {
                hit->t = t_curr;
                hit->node = node;
                hit->level = lvl;
#pragma unroll
                for (int c = 0; c < 3; ++c) hit->gc[c] = rec[1 + c] * dim + (uint32_t)cell[c];
                return true;
            } else if (!last && (int_w & bit)) {
                node = child;
                // a child id outside the next level's buffer (a tree whose
                // build ran out of capacity) ends the ray instead of reading
                // past the buffer
                if (node >= Ln.cap) return false;
#pragma unroll
                for (int c = 0; c < 3; ++c) np[c] = np[c] + (float)cell[c] * cs;
                t_max = fmin_glm(tn[0], fmin_glm(tn[1], tn[2]));
                ++lvl;
                L = Ln;
            } else {
                t_curr = fmin_glm(tn[0], fmin_glm(tn[1], tn[2]));
            }
        }
    }
    return false;
}

// Forward-mode gradient (vulkan/shaders/tape/gradient.comp:4-88) of the tape
// at tape_words (wave-uniform) at each lane's point.  The slot file is four
// per-component arrays that the compiler keeps in VGPRs (indexed by the
// uniform slot number), and the tape words arrive through VGPR windows.
// Called with every lane active; n_cap bounds the words read.
template <int MAXS>
__device__ __forceinline__ V4 eval_gradient(const uint32_t* tape_words, uint64_t n_cap,
                                            const float* consts, float x, float y, float z,
                                            float t, uint32_t lane) {
    // window 0 goes out with the length instead of after it (clamped to
    // the readable words)
    const uint32_t raw0 = tape_words[1u + lane < n_cap ? 1u + lane : (n_cap ? n_cap - 1u : 0u)];
    uint32_t n = as_const(tape_words)[0];
    if ((uint64_t)n + 1 > n_cap) n = n_cap ? (uint32_t)(n_cap - 1) : 0u;
    const uint32_t* tw = tape_words + 1;
    float gx[MAXS], gy[MAXS], gz[MAXS], gw[MAXS];
    gx[0] = 1.0f; gy[0] = 0.0f; gz[0] = 0.0f; gw[0] = x;
    gx[1] = 0.0f; gy[1] = 1.0f; gz[1] = 0.0f; gw[1] = y;
    gx[2] = 0.0f; gy[2] = 0.0f; gz[2] = 1.0f; gw[2] = z;
    gx[3] = 0.0f; gy[3] = 0.0f; gz[3] = 0.0f; gw[3] = t;
    // the current window with its constants and the next window's words
    // (the short tapes of a last level take one window)
    uint32_t fw0 = lane < n ? raw0 : 0xFFFFFFFFu;
    float fc0 = window_consts(fw0, consts);               // window 1 beside it
    uint32_t fw1 = window_load(tw, 64, n, lane);
    // nothing in flight when the loop starts, so that the compiler's waits
    // for the loads issued in the loop fall at the window's end
    __builtin_amdgcn_s_waitcnt(0);
    for (uint32_t s = 0; s < n; s += 64) {
        // the next window's constants and the words of the one after go out
        // now and are taken at this window's end (a load taken right after
        // it is issued would wait for it there)
        const float fc1 = window_consts(fw1, consts);
        const uint32_t fw2 = window_load(tw, (int32_t)s + 128, n, lane);
        const uint32_t m = uni(n - s < 64 ? n - s : 64);   // a scalar loop bound
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t w = rdlane(fw0, j);
            const uint32_t op = w & 0xFFu, o = slot_in<MAXS>((w >> 8) & 0xFFu);
            const uint32_t ia = slot_in<MAXS>((w >> 16) & 0xFFu), ib = slot_in<MAXS>(w >> 24);
            const V4 A{gx[ia], gy[ia], gz[ia], gw[ia]};
            FREP_FENCE();
            V4 B = A;
            float c = 0.0f;
            if (op >= OP_ADD && op <= OP_MAX) {
                B = V4{gx[ib], gy[ib], gz[ib], gw[ib]};
                FREP_FENCE();
            } else if (op == OP_CONST) {
                c = rdlanef(fc0, j);
            }
            const V4 r = gradient_op(op, A, B, c);
            FREP_FENCE();
            gx[o] = r.x;
            gy[o] = r.y;
            gz[o] = r.z;
            gw[o] = r.w;
        }
        fw0 = fw1; fc0 = fc1;
        fw1 = fw2;
    }
    return V4{gx[0], gy[0], gz[0], gw[0]};
}

__device__ __forceinline__ void normalize3(float v[3]) {
    const float l = __builtin_sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
}

}  // namespace

template <int MAXS>
__global__ __launch_bounds__(64) void k_render(RenderArgs a) {
    // a tree whose build ran out of capacity is incomplete: render nothing
    // (the caller rebuilds it and renders again)
    if (a.counters && load_counter(a.counters + CNT_OVERFLOW) != 0) return;
    const uint32_t lane = lane_id();
    const uint32_t tiles_x = (a.width + 7) / 8;
    const uint32_t tile = blockIdx.x;
    const uint32_t px = (tile % tiles_x) * 8 + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8 + (lane >> 3);
    const bool valid = px < a.width && py < a.height;
    const uint32_t spp = a.spp ? a.spp : 1;
    float acc[4] = {0, 0, 0, 0};
    for (uint32_t smp = 0; smp < spp; ++smp) {
        float ox = 0.0f, oy = 0.0f;
        if (a.jitter) {
            ox = 2.0f * (jitter01(a.seed, px, py, smp, 0) - 0.5f);
            oy = 2.0f * (jitter01(a.seed, px, py, smp, 1) - 0.5f);
        }
        const float dx = 2.0f * (((float)px + ox) / (float)a.width) - 1.0f;
        const float dy = 2.0f * (((float)py + oy) / (float)a.height) - 1.0f;
        float orig[3], dir[3];
        if (a.ortho) {
            const float hx = (dx * a.sw[0]) * 0.5f, hy = (dy * a.sw[1]) * 0.5f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                orig[c] = (a.pos[c] + a.right[c] * hx) + a.up[c] * hy;
                dir[c] = a.fwd[c];
            }
        } else {
            const float sx = dx * a.sw[0], sy = dy * a.sw[1];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                orig[c] = a.pos[c];
                dir[c] = (a.fwd[c] + a.right[c] * sx) + a.up[c] * sy;
            }
            normalize3(dir);
        }
        Hit h{};
        const bool hit = valid && dda(a, orig, dir, &h);
        uint32_t toff = 0;
        if (hit) toff = a.lv[h.level].nodes[(size_t)h.node * a.lv[h.level].ns];
        // Waterfall over the distinct tapes hit in this wavefront: each round
        // evaluates one tape, wave-uniform, for the lanes that hit it.
        V4 g{0, 0, 0, 0};
        bool need = hit && toff < a.pool_words;
        float p[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = orig[c] + dir[c] * h.t;
        for (;;) {
            const uint64_t m = __ballot(need);
            if (m == 0) break;
            const uint32_t leader = (uint32_t)__builtin_ctzll(m);
            const uint32_t tu = uni((uint32_t)__shfl((int)toff, (int)leader));
            // every lane runs the tape (windows are loaded by the whole
            // wavefront); the lanes that hit this tape keep the result
            const V4 gv = eval_gradient<MAXS>(a.pool + tu, a.pool_words - tu, a.consts, p[0], p[1],
                                              p[2], a.tape_time, lane);
            if (need && toff == tu) {
                g = gv;
                need = false;
            }
        }
        const size_t sidx = ((size_t)py * a.width + px) * spp + smp;
        float col[4];
        if (hit) {
            float nrm[3] = {g.x, g.y, g.z};
            normalize3(nrm);
            if (a.hit_ids)
                a.hit_ids[sidx] = ((unsigned long long)h.level << 60) |
                                  ((unsigned long long)h.gc[0] << 40) |
                                  ((unsigned long long)h.gc[1] << 20) | (unsigned long long)h.gc[2];
            if (a.normals) {
                a.normals[sidx * 3 + 0] = nrm[0];
                a.normals[sidx * 3 + 1] = nrm[1];
                a.normals[sidx * 3 + 2] = nrm[2];
            }
            if (h.t < 0.0f) {
                col[0] = col[1] = col[2] = col[3] = 1.0f;
            } else {
                float diffuse[4] = {0, 0, 0, 0};
                for (uint32_t l = 0; l < a.light_count; ++l) {
                    const float it = fmax_glm(-a.light_dir[l][0] * nrm[0] +
                                                  -a.light_dir[l][1] * nrm[1] +
                                                  -a.light_dir[l][2] * nrm[2], 0.0f);
#pragma unroll
                    for (int c = 0; c < 4; ++c) diffuse[c] = diffuse[c] + it * a.light_color[l][c];
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) col[c] = a.ambient[c] + diffuse[c];
            }
        } else {
            if (valid && a.hit_ids) a.hit_ids[sidx] = ~0ull;
            if (valid && a.normals) {
                const float nan = __builtin_nanf("");
                a.normals[sidx * 3 + 0] = nan;
                a.normals[sidx * 3 + 1] = nan;
                a.normals[sidx * 3 + 2] = nan;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) col[c] = a.background[c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = acc[c] + col[c];
    }
    if (valid && a.rgba) {
        uchar4 v;
        float f[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x = acc[c] / (float)spp;
            x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
            f[c] = x * 255.0f + 0.5f;
        }
        v.x = (unsigned char)f[0]; v.y = (unsigned char)f[1];
        v.z = (unsigned char)f[2]; v.w = (unsigned char)f[3];
        reinterpret_cast<uchar4*>(a.rgba)[(size_t)py * a.width + px] = v;
    }
}

// ---------------------------------------------------------------------------
// Batch primitives: one lane per point / box, one tape for the whole launch.
// ---------------------------------------------------------------------------
template <int MAXS>
__global__ __launch_bounds__(256) void k_eval_points(const uint32_t* tape, const float* consts,
                                                     const float* xyz, uint32_t n, float t,
                                                     float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = i < n ? i : n - 1;
    auto init = [&](int) { return make_float3(xyz[3 * j + 0], xyz[3 * j + 1], xyz[3 * j + 2]); };
    auto put = [&](int, float d) { if (i < n) out[i] = d; };
    run_point_tape<MAXS, 1>(as_const(tape), as_const(consts), t, init, put);
}

// Same through the assembly interpreter: 8 points per lane, point
// blockIdx.x*2048 + k*256 + threadIdx.x for k = 0..7.
template <int MAXS>
__global__ __launch_bounds__(256) void k_eval_points_asm(const uint32_t* tape, const float* consts,
                                                         const float* xyz, uint32_t n, float t,
                                                         float* out) {
    v8f xs, ys, zs;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t p = blockIdx.x * 2048u + (uint32_t)k * 256u + threadIdx.x;
        const uint32_t q = p < n ? p : n - 1;
        xs[k] = xyz[3 * q]; ys[k] = xyz[3 * q + 1]; zs[k] = xyz[3 * q + 2];
    }
    uint32_t w0;
    float c0;
    point_window0(tape, consts, threadIdx.x & 63u, &w0, &c0);
    const v8f d = run_point_tape_asm8<MAXS>(tape, consts, t, xs, ys, zs, w0, c0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t p = blockIdx.x * 2048u + (uint32_t)k * 256u + threadIdx.x;
        if (p < n) out[p] = d[k];
    }
}

template <int MAXS>
__global__ __launch_bounds__(256) void k_eval_intervals(const uint32_t* tape_, const float* consts_,
                                                        const float* lo, const float* hi,
                                                        uint32_t n, float t, float* out,
                                                        uint8_t* choices, uint32_t n_mm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = i < n ? i : n - 1;
    ctape_t tape = as_const(tape_);
    cconst_t consts = as_const(consts_);
    float slo[MAXS], shi[MAXS];
#pragma unroll
    for (int c = 0; c < 3; ++c) { slo[c] = lo[3 * j + c]; shi[c] = hi[3 * j + c]; }
    slo[3] = t; shi[3] = t;
    const uint32_t nt = tape[0];
    uint32_t mm = 0;
    for (uint32_t k = 0; k < nt; ++k) {
        const uint32_t w = tape[1 + k];
        const uint32_t op = w & 0xFFu, o = ((w >> 8) & 0xFFu) & (MAXS - 1);
        const uint32_t ia = ((w >> 16) & 0xFFu) & (MAXS - 1), ib = (w >> 24) & (MAXS - 1);
        const Itv A{slo[ia], shi[ia]}, B{slo[ib], shi[ib]};
        const float c = (op == OP_CONST) ? consts[w >> 16] : 0.0f;
        const Itv r = interval_op(op, A, B, c);
        if (op == OP_MIN || op == OP_MAX) {
            if (choices && i < n) choices[(size_t)i * n_mm + mm] = (uint8_t)mm_choice(op, A, B);
            ++mm;
        }
        slo[o] = r.lo;
        shi[o] = r.hi;
    }
    if (i < n) { out[2 * i] = slo[0]; out[2 * i + 1] = shi[0]; }
}

template <int MAXS>
__global__ __launch_bounds__(256) void k_eval_gradients(const uint32_t* tape, const float* consts,
                                                        const float* xyz, uint32_t n, float t,
                                                        float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = i < n ? i : n - 1;
    const V4 g = eval_gradient<MAXS>(tape, ~0ull, consts, xyz[3 * j], xyz[3 * j + 1],
                                     xyz[3 * j + 2], t, lane_id());
    if (i < n) {
        out[4 * i + 0] = g.x; out[4 * i + 1] = g.y; out[4 * i + 2] = g.z; out[4 * i + 3] = g.w;
    }
}

__global__ __launch_bounds__(256) void k_math(uint32_t fn, const float* x, uint32_t n, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = x[i];
    out[i] = fn == 0 ? dev_sinf(v) : (fn == 1 ? dev_cosf(v) : dev_expf(v));
}

// ---------------------------------------------------------------------------
// Launchers: template dispatch on the slot-file size.
// ---------------------------------------------------------------------------
#define FREP_DISPATCH_MAXS(maxs, CALL)                 \
    switch (maxs) {                                     \
        case 8: { constexpr int M = 8; CALL; } break;   \
        case 16: { constexpr int M = 16; CALL; } break; \
        case 32: { constexpr int M = 32; CALL; } break; \
        case 64: { constexpr int M = 64; CALL; } break; \
        case 128: { constexpr int M = 128; CALL; } break; \
        case 256: { constexpr int M = 256; CALL; } break; \
        default: return hipErrorInvalidValue;           \
    }

hipError_t launch_init_tree(hipStream_t st, uint32_t* counters, uint32_t* nodes0, uint32_t* pool,
                            const uint32_t* root_tape, uint32_t root_off, uint32_t n_levels) {
    k_init_tree<<<1, 256, 0, st>>>(counters, nodes0, pool, root_tape, root_off, n_levels);
    return hipGetLastError();
}

// Persistent grid of a level kernel: as many 256-thread blocks as can be
// resident at once (occupancy of that instantiation x CU count), so every
// block starts immediately and no block waits for another to retire.
template <class Kern>
uint32_t resident_grid(Kern kern, uint32_t cu_count, uint32_t max_grid, size_t lds) {
    int bpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kern, kLevelBlock, lds) != hipSuccess ||
        bpc < 1)
        bpc = 1;
    const uint32_t g = cu_count * (uint32_t)bpc;
    return g < max_grid ? g : max_grid;
}

// The occupancy of an interval kernel depends on its dynamic LDS (choice
// rows); the last answer is kept per instantiation and thread.
template <int M, int KC = 1>
uint32_t interval_grid(uint32_t cu_count, uint32_t max_grid, size_t lds) {
    thread_local size_t last_lds = ~(size_t)0;
    thread_local uint32_t last_cu = 0, last_max = 0, g = 0;
    if (lds != last_lds || cu_count != last_cu || max_grid != last_max) {
        g = resident_grid(k_level_interval<M, KC>, cu_count, max_grid, lds);
        last_lds = lds;
        last_cu = cu_count;
        last_max = max_grid;
    }
    return g;
}

template <int M, int K, bool ASM>
uint32_t point_grid(uint32_t cu_count, uint32_t max_grid) {
    thread_local uint32_t last_cu = 0, last_max = 0, g = 0;
    if (cu_count != last_cu || max_grid != last_max) {
        g = resident_grid(k_level_point<M, K, ASM>, cu_count, max_grid, 0);
        last_cu = cu_count;
        last_max = max_grid;
    }
    return g;
}

// FREP_INTERVAL_K1=1 keeps one cell per lane on every interval level
// (read at each launch, so a test can compare both forms in one process).
bool two_cells_per_lane() {
    const char* e = std::getenv("FREP_INTERVAL_K1");
    return !(e && e[0] == '1');
}

// Shortening log of a wavefront: KC ballots per entry (one entry per word of
// the parent tape).
size_t log_entry_bytes(uint32_t kc) { return 8u * kc; }

size_t interval_lds_bytes(uint32_t mm_stride, uint32_t kc, uint32_t cap) {
    return (size_t)(kLevelBlock / 64) *
           ((size_t)mm_stride * kc * sizeof(uint4) + (size_t)cap * log_entry_bytes(kc));
}

int blocks_per_cu(const void* kern, size_t lds);

// LDS a block of kern may take without lowering the blocks per CU that its
// registers allow: the CU's 160 KiB shared among them, less the kernel's
// static LDS and a margin.
size_t lds_share(const void* kern) {
    thread_local std::map<const void*, size_t> cache;
    const auto it = cache.find(kern);
    if (it != cache.end()) return it->second;
    constexpr size_t kLdsPerCu = 160 * 1024;
    hipFuncAttributes at{};
    const size_t st = hipFuncGetAttributes(&at, kern) == hipSuccess ? at.sharedSizeBytes : 1024;
    const size_t per = kLdsPerCu / (size_t)blocks_per_cu(kern, 0);
    const size_t share = per > st + 256 ? per - st - 256 : 0;
    cache[kern] = share;
    return share;
}

// Log entries per wavefront: what fits in block_budget bytes of LDS per
// block after the choice rows, between 64 and 4096, and within the device's
// LDS per block.  FREP_LOG_CAP=n sets it (read at each launch; a tiny log
// makes every chunk take the second walk).
uint32_t log_cap_for(uint32_t mm_stride, uint32_t kc, size_t block_budget) {
    const size_t per_wave = block_budget / (kLevelBlock / 64);
    const size_t rows = (size_t)mm_stride * kc * sizeof(uint4);
    size_t cap = per_wave > rows ? (per_wave - rows) / log_entry_bytes(kc) : 0;
    cap = std::min<size_t>(std::max<size_t>(cap, 64), 4096);
    if (const char* e = std::getenv("FREP_LOG_CAP")) {
        const long v = std::atol(e);
        if (v >= 1 && v <= 4096) cap = (size_t)v;
    }
    while (cap > 1 && interval_lds_bytes(mm_stride, kc, (uint32_t)cap) > kMaxChoiceLds) cap /= 2;
    return (uint32_t)cap;
}

// One interval launch: the log sized to the kernel's LDS share.
template <int M, int KC>
hipError_t launch_interval_kc(hipStream_t st, LevelArgs la, uint32_t cu_count, uint32_t max_grid) {
    la.log_cap = log_cap_for(la.mm_stride, KC, lds_share((const void*)k_level_interval<M, KC>));
    const size_t lds = interval_lds_bytes(la.mm_stride, KC, la.log_cap);
    if (lds > kMaxChoiceLds) return hipErrorInvalidValue;
    k_level_interval<M, KC><<<interval_grid<M, KC>(cu_count, max_grid, lds), kLevelBlock, lds,
                              st>>>(la);
    return hipGetLastError();
}

hipError_t launch_level_interval(hipStream_t st, const LevelArgs& a, uint32_t maxs,
                                 uint32_t slot_count, uint32_t cu_count, uint32_t max_grid,
                                 uint32_t nodes_hint) {
    if (interval_lds_bytes(a.mm_stride, 1, 1) > kMaxChoiceLds) return hipErrorInvalidValue;
    // two cells per lane (one choice row per group, so twice the LDS) for
    // files of up to 16 slots on nodes of a multiple of 128 cells, unless
    // the level is known to be small: below 16 one-cell chunks per CU, the
    // wavefronts would not fill the SIMDs (a level of 122 nodes of 512
    // cells ran 50 us with one cell per lane, 55.5 us with two)
    const uint32_t cells = a.dim * a.dim * a.dim;
    const bool big = nodes_hint == 0 || (uint64_t)nodes_hint * (cells / 64) >= 16ull * cu_count;
    if (maxs <= 16 && cells % 128 == 0 && big && two_cells_per_lane() &&
        interval_lds_bytes(a.mm_stride, 2, 1) <= kMaxChoiceLds) {
        if (maxs == 8) return launch_interval_kc<8, 2>(st, a, cu_count, max_grid);
        if (slot_count <= 12) return launch_interval_kc<12, 2>(st, a, cu_count, max_grid);
        return launch_interval_kc<16, 2>(st, a, cu_count, max_grid);
    }
    // a 12-slot file frees 8 registers per lane over the 16-slot one, which
    // is what 8 wavefronts per SIMD need
    if (maxs == 16 && slot_count <= 12) return launch_interval_kc<12, 1>(st, a, cu_count, max_grid);
    FREP_DISPATCH_MAXS(maxs, return (launch_interval_kc<M, 1>(st, a, cu_count, max_grid)));
    return hipGetLastError();
}
uint32_t root_cells_per_block(uint32_t table_bytes, uint32_t cell_bytes, size_t lds_per_block) {
    if (lds_per_block <= kRootStaticLds) return 0;
    const size_t cap = lds_per_block - kRootStaticLds;
    if (table_bytes + (size_t)cell_bytes * 16 <= std::min<size_t>(cap, 64 * 1024 - kRootStaticLds))
        return 16;
    if (table_bytes + (size_t)cell_bytes * 8 <= cap) return 8;
    return 0;
}

hipError_t launch_root_interval(hipStream_t st, const RootArgs& r) {
    const uint32_t c = r.cells_per_block, lpc = r.lanes_per_cell;
    const uint32_t cells = r.a.dim * r.a.dim * r.a.dim;
    const bool res_ok = (r.res & kRootConst) ? (r.res & (kRootConst - 1u)) < r.n_const
                                              : r.res < r.n_slot;
    if ((c != 16 && c != 8) || (lpc != 16 && lpc != 64) || (c * lpc) % 64 != 0 || r.n_pass % 4 ||
        cells % c != 0 || (r.e0 & 1u) || r.n_slot < 4 || r.n_slot > kRootMaxSlots + 1 ||
        r.n == 0 || r.n > kRootMaxInstr || r.n_const >= kRootConst || !res_ok ||
        r.table_bytes < 10 * r.n_const ||
        r.cell_bytes < 8 * r.n_slot + ((r.n + 3u) & ~3u) + r.n_slot)
        return hipErrorInvalidValue;
    const size_t lds = r.table_bytes + (size_t)c * r.cell_bytes;
    const uint32_t blocks = cells / c;
    if (c == 16 && lpc == 64) k_root_interval<16, 64><<<blocks, 1024, lds, st>>>(r);
    else if (c == 8 && lpc == 64) k_root_interval<8, 64><<<blocks, 512, lds, st>>>(r);
    else if (c == 16) k_root_interval<16, 16><<<blocks, 256, lds, st>>>(r);
    else k_root_interval<8, 16><<<blocks, 128, lds, st>>>(r);
    return hipGetLastError();
}

hipError_t launch_root_interval_batch(hipStream_t st, const RootArgs* d_args,
                                      const uint32_t* d_starts, uint32_t n, uint32_t blocks,
                                      uint32_t cells_per_block, uint32_t lanes_per_cell,
                                      size_t lds) {
    if (n == 0 || n > kBatchMax || blocks == 0) return hipErrorInvalidValue;
    const uint32_t c = cells_per_block, lpc = lanes_per_cell;
    if (c == 16 && lpc == 64)
        k_root_interval_batch<16, 64><<<blocks, 1024, lds, st>>>(d_args, d_starts, n);
    else if (c == 8 && lpc == 64)
        k_root_interval_batch<8, 64><<<blocks, 512, lds, st>>>(d_args, d_starts, n);
    else if (c == 16 && lpc == 16)
        k_root_interval_batch<16, 16><<<blocks, 256, lds, st>>>(d_args, d_starts, n);
    else if (c == 8 && lpc == 16)
        k_root_interval_batch<8, 16><<<blocks, 128, lds, st>>>(d_args, d_starts, n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Largest voxels-per-lane count for a slot-file size: K banks of MAXS
// registers each, at most 128 slot registers per lane (MAXS >= 64 banks do
// not fit in registers, so one bank is used).
__host__ __device__ constexpr int point_k_for(int maxs) {
    return maxs <= 16 ? 8 : (maxs <= 32 ? 4 : 1);
}

uint32_t point_k_used(uint32_t maxs, uint32_t dim) {
    const uint32_t cpn = dim * dim * dim / 64;   // wavefront chunks per node
    uint32_t k = (uint32_t)point_k_for((int)maxs);
    while (k > 1 && cpn % k != 0) k /= 2;
    return k;
}

hipError_t launch_level_point(hipStream_t st, const LevelArgs& a, uint32_t maxs,
                              uint32_t slots, uint32_t cu_count, uint32_t max_grid, bool asm_ok) {
    const uint32_t k = point_k_used(maxs, a.dim);
    // hand-scheduled interpreter: 8 points per lane, slot file of 8, 10 or
    // 16 (the 10-slot file keeps 4 wavefronts per SIMD where 16 allows 3)
    if (asm_ok && k == 8 && maxs == 8) {
        k_level_point<8, 8, true><<<point_grid<8, 8, true>(cu_count, max_grid), kLevelBlock, 0,
                                    st>>>(a);
        return hipGetLastError();
    }
    if (asm_ok && k == 8 && maxs == 16 && slots <= 10) {
        k_level_point<10, 8, true><<<point_grid<10, 8, true>(cu_count, max_grid), kLevelBlock, 0,
                                     st>>>(a);
        return hipGetLastError();
    }
    if (asm_ok && k == 8 && maxs == 16) {
        k_level_point<16, 8, true><<<point_grid<16, 8, true>(cu_count, max_grid), kLevelBlock, 0,
                                     st>>>(a);
        return hipGetLastError();
    }
    // one voxel per lane (nodes of one wavefront chunk): the same
    // interpreter with one-register slots
    if (asm_ok && k == 1 && maxs == 8) {
        k_level_point<8, 1, true><<<point_grid<8, 1, true>(cu_count, max_grid), kLevelBlock, 0,
                                    st>>>(a);
        return hipGetLastError();
    }
    if (asm_ok && k == 1 && maxs == 16) {
        k_level_point<16, 1, true><<<point_grid<16, 1, true>(cu_count, max_grid), kLevelBlock, 0,
                                     st>>>(a);
        return hipGetLastError();
    }
    // C++ interpreter, 8 points per lane: a 10-slot file for tapes of at
    // most 10 slots (4 wavefronts per SIMD where the 16-slot file allows 3)
    if (k == 8 && maxs == 16 && slots <= 10) {
        k_level_point<10, 8, false><<<point_grid<10, 8, false>(cu_count, max_grid), kLevelBlock,
                                      0, st>>>(a);
        return hipGetLastError();
    }
#define FREP_POINT_LAUNCH(KK)                                                                 \
    k_level_point<M, KK><<<point_grid<M, KK, false>(cu_count, max_grid), kLevelBlock, 0, st>>>(a)
#define FREP_POINT_CALL                                                                       \
    {                                                                                         \
        constexpr int KM = point_k_for(M);                                                    \
        if (k == 8) FREP_POINT_LAUNCH((KM >= 8 ? 8 : 1));                                     \
        else if (k == 4) FREP_POINT_LAUNCH((KM >= 4 ? 4 : 1));                                \
        else if (k == 2) FREP_POINT_LAUNCH((KM >= 2 ? 2 : 1));                                \
        else FREP_POINT_LAUNCH(1);                                                            \
    }
    FREP_DISPATCH_MAXS(maxs, FREP_POINT_CALL);
#undef FREP_POINT_CALL
#undef FREP_POINT_LAUNCH
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Work statistics of one level of a built tree: one wavefront per node reads
// the node's tape (lanes stride over its words) and adds its length and the
// weighted cost of its instructions to out[0..2] = {sum of lengths, max
// length, sum of weights}.  Weights are FP32 operations per evaluation:
// interval rules for levels above the last, point rules for the last.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tree_stats(const uint32_t* nodes, uint32_t ns,
                                                    uint32_t n_nodes, const uint32_t* pool,
                                                    uint64_t pool_words, uint32_t interval,
                                                    unsigned long long* out) {
    const uint32_t lane = lane_id();
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    for (uint32_t node = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); node < n_nodes;
         node += nwaves) {
        const uint32_t toff = nodes[(size_t)node * ns];
        if ((uint64_t)toff + 1 > pool_words) continue;
        uint32_t n = pool[toff];
        if ((uint64_t)toff + 1 + n > pool_words) n = (uint32_t)(pool_words - toff - 1);
        unsigned long long w = 0;
        for (uint32_t i = lane; i < n; i += 64) {
            const uint32_t op = pool[(size_t)toff + 1 + i] & 0xFFu;
            uint32_t c = 0;
            if (interval) {
                c = (op == OP_SIN || op == OP_COS || op == OP_DIV) ? 12u
                    : op == OP_MUL                                   ? 10u
                    : (op == OP_MIN || op == OP_MAX || op == OP_SQRT) ? 4u
                    : (op == OP_ADD || op == OP_SUB || op == OP_NEG || op == OP_EXP) ? 2u
                                                                                     : 0u;
            } else {
                c = (op == OP_CONST || op == OP_COPY || op > OP_COPY) ? 0u : 1u;
            }
            w += c;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) w += (unsigned long long)__shfl_xor((long long)w, d);
        if (lane == 0) {
            atomicAdd(out + 0, (unsigned long long)n);
            atomicMax(out + 1, (unsigned long long)n);
            atomicAdd(out + 2, w);
        }
    }
}

hipError_t launch_tree_stats(hipStream_t st, const uint32_t* nodes, uint32_t ns, uint32_t n_nodes,
                             const uint32_t* pool, uint64_t pool_words, bool interval,
                             unsigned long long* out) {
    if (n_nodes == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n_nodes + 3) / 4, 65536);
    k_tree_stats<<<blocks, 256, 0, st>>>(nodes, ns, n_nodes, pool, pool_words, interval ? 1u : 0u,
                                         out);
    return hipGetLastError();
}

// Resident blocks per CU of a kernel at a given dynamic LDS size, cached.
int blocks_per_cu(const void* kern, size_t lds) {
    thread_local std::map<std::pair<const void*, size_t>, int> cache;
    const auto key = std::make_pair(kern, lds);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int bpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kern, kLevelBlock, lds) != hipSuccess ||
        bpc < 1)
        bpc = 1;
    cache[key] = bpc;
    return bpc;
}

uint32_t batch_grid(const void* kern, size_t lds, uint32_t cu_count, uint32_t max_grid) {
    const uint32_t g = cu_count * (uint32_t)blocks_per_cu(kern, lds);
    return g < max_grid ? g : max_grid;
}

hipError_t launch_init_batch(hipStream_t st, const InitArgs* d_args, uint32_t n) {
    if (n == 0) return hipSuccess;
    k_init_batch<<<n, 256, 0, st>>>(d_args);
    return hipGetLastError();
}

hipError_t launch_gather_counters(hipStream_t st, const InitArgs* d_args, uint32_t n,
                                  uint32_t* out) {
    if (n == 0) return hipSuccess;
    const uint32_t threads = n * CNT_WORDS;
    k_gather_counters<<<(threads + 255) / 256, 256, 0, st>>>(d_args, n, out);
    return hipGetLastError();
}

// One batched interval launch: the log sized to the kernel's LDS share.
template <int M, int KC>
hipError_t launch_interval_batch_kc(hipStream_t st, BatchArgs bb, uint32_t cu_count,
                                    uint32_t max_grid) {
    const void* kern = (const void*)k_level_interval_batch<M, KC>;
    bb.kc = KC;
    bb.log_cap = log_cap_for(bb.mm_stride, KC, lds_share(kern));
    const size_t lds = interval_lds_bytes(bb.mm_stride, KC, bb.log_cap);
    if (lds > kMaxChoiceLds) return hipErrorInvalidValue;
    k_level_interval_batch<M, KC><<<batch_grid(kern, lds, cu_count, max_grid), kLevelBlock, lds,
                                    st>>>(bb);
    return hipGetLastError();
}

hipError_t launch_level_interval_batch(hipStream_t st, const BatchArgs& b, uint32_t maxs,
                                       uint32_t slot_count, uint32_t cu_count, uint32_t max_grid) {
    if (b.n_shapes == 0 || b.n_shapes > kBatchMax) return hipErrorInvalidValue;
    if (interval_lds_bytes(b.mm_stride, 1, 1) > kMaxChoiceLds) return hipErrorInvalidValue;
    // two cells per lane when asked for and the slot file allows it (shapes
    // whose nodes hold a multiple of 128 cells; the others take one)
    if (b.kc == 2 && maxs <= 16 && interval_lds_bytes(b.mm_stride, 2, 1) <= kMaxChoiceLds) {
        if (maxs == 8) return launch_interval_batch_kc<8, 2>(st, b, cu_count, max_grid);
        if (slot_count <= 12) return launch_interval_batch_kc<12, 2>(st, b, cu_count, max_grid);
        return launch_interval_batch_kc<16, 2>(st, b, cu_count, max_grid);
    }
    if (maxs == 16 && slot_count <= 12)
        return launch_interval_batch_kc<12, 1>(st, b, cu_count, max_grid);
    FREP_DISPATCH_MAXS(maxs, return (launch_interval_batch_kc<M, 1>(st, b, cu_count, max_grid)));
    return hipGetLastError();
}

hipError_t launch_level_point_batch(hipStream_t st, const BatchArgs& b, uint32_t maxs,
                                    uint32_t cu_count, uint32_t max_grid, uint32_t asm_slots,
                                    bool small) {
    if (b.n_shapes == 0 || b.n_shapes > kBatchMax) return hipErrorInvalidValue;
    if (small) {
        FREP_DISPATCH_MAXS(maxs, (k_level_point_batch<M, true><<<
                                      batch_grid((const void*)k_level_point_batch<M, true>, 0,
                                                 cu_count, max_grid),
                                      kLevelBlock, 0, st>>>(b)));
        return hipGetLastError();
    }
    if (asm_slots && asm_slots <= 16) {
        if (asm_slots <= 8)
            k_level_point_batch_asm<8><<<batch_grid((const void*)k_level_point_batch_asm<8>, 0,
                                                    cu_count, max_grid),
                                         kLevelBlock, 0, st>>>(b);
        else if (asm_slots <= 10)
            k_level_point_batch_asm<10><<<batch_grid((const void*)k_level_point_batch_asm<10>, 0,
                                                     cu_count, max_grid),
                                          kLevelBlock, 0, st>>>(b);
        else
            k_level_point_batch_asm<16><<<batch_grid((const void*)k_level_point_batch_asm<16>, 0,
                                                     cu_count, max_grid),
                                          kLevelBlock, 0, st>>>(b);
        return hipGetLastError();
    }
    FREP_DISPATCH_MAXS(maxs, (k_level_point_batch<M><<<
                                  batch_grid((const void*)k_level_point_batch<M>, 0, cu_count,
                                             max_grid),
                                  kLevelBlock, 0, st>>>(b)));
    return hipGetLastError();
}

hipError_t launch_render(hipStream_t st, const RenderArgs& a, uint32_t maxs, uint32_t slots) {
    const uint32_t tiles = ((a.width + 7) / 8) * ((a.height + 7) / 8);
    // tapes of at most 10 slots on a 10-slot gradient file: 4 wavefronts per
    // SIMD where the 16-slot file allows 3, more rays in flight to hide the
    // walk's dependent loads
    if (maxs == 16 && slots <= 10) {
        k_render<10><<<tiles, 64, 0, st>>>(a);
        return hipGetLastError();
    }
    FREP_DISPATCH_MAXS(maxs, (k_render<M><<<tiles, 64, 0, st>>>(a)));
    return hipGetLastError();
}

// Presentation of a layered render target (vulkan/shaders/renderer/
// triangle.frag:14-23): the mean over the layers of each pixel's rgb as
// unorm values, alpha 1, written back as rgba8 rounded to nearest.  Layers
// are summed in order in float, as the fragment shader's loop does.  One
// thread per pixel; 4 bytes read per layer and 4 written per pixel.
__global__ __launch_bounds__(256) void k_resolve(const uchar4* layers, uint32_t pixels,
                                                 uint32_t n_layers, uchar4* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    for (uint32_t l = 0; l < n_layers; ++l) {
        const uchar4 q = layers[(size_t)l * pixels + i];
        r = r + (float)q.x / 255.0f;
        g = g + (float)q.y / 255.0f;
        b = b + (float)q.z / 255.0f;
    }
    const float n = (float)n_layers;
    auto u8 = [](float v) {
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        return (unsigned char)(v * 255.0f + 0.5f);
    };
    out[i] = make_uchar4(u8(r / n), u8(g / n), u8(b / n), u8(n / n));
}

hipError_t launch_resolve(hipStream_t st, const void* layers, uint32_t pixels, uint32_t n_layers,
                          void* out) {
    if (pixels == 0) return hipSuccess;
    k_resolve<<<(pixels + 255) / 256, 256, 0, st>>>((const uchar4*)layers, pixels, n_layers,
                                                    (uchar4*)out);
    return hipGetLastError();
}

hipError_t launch_eval_points(hipStream_t st, const uint32_t* tape, const float* consts,
                              const float* xyz, uint32_t n, float t, float* out, uint32_t maxs,
                              bool asm_ok) {
    if (asm_ok && (maxs == 8 || maxs == 16)) {
        const uint32_t blocks = (n + 2047) / 2048;
        if (maxs == 8) k_eval_points_asm<8><<<blocks, 256, 0, st>>>(tape, consts, xyz, n, t, out);
        else k_eval_points_asm<16><<<blocks, 256, 0, st>>>(tape, consts, xyz, n, t, out);
        return hipGetLastError();
    }
    const uint32_t blocks = (n + 255) / 256;
    FREP_DISPATCH_MAXS(maxs, (k_eval_points<M><<<blocks, 256, 0, st>>>(tape, consts, xyz, n, t, out)));
    return hipGetLastError();
}

hipError_t launch_eval_intervals(hipStream_t st, const uint32_t* tape, const float* consts,
                                 const float* lo, const float* hi, uint32_t n, float t,
                                 float* out, uint8_t* choices, uint32_t n_mm, uint32_t maxs) {
    const uint32_t blocks = (n + 255) / 256;
    FREP_DISPATCH_MAXS(maxs, (k_eval_intervals<M><<<blocks, 256, 0, st>>>(tape, consts, lo, hi, n, t,
                                                                         out, choices, n_mm)));
    return hipGetLastError();
}

hipError_t launch_eval_gradients(hipStream_t st, const uint32_t* tape, const float* consts,
                                 const float* xyz, uint32_t n, float t, float* out, uint32_t maxs) {
    const uint32_t blocks = (n + 255) / 256;
    FREP_DISPATCH_MAXS(maxs, (k_eval_gradients<M><<<blocks, 256, 0, st>>>(tape, consts, xyz, n, t, out)));
    return hipGetLastError();
}

hipError_t launch_math(hipStream_t st, uint32_t fn, const float* x, uint32_t n, float* out) {
    k_math<<<(n + 255) / 256, 256, 0, st>>>(fn, x, n, out);
    return hipGetLastError();
}

}  // namespace frep
