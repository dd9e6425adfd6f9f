// The C ABI of include/nwgrad_hip.h: contexts, device-resident batches, the batch
// planner, kernel launches and the gradient reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/nwgrad_hip.h"
#include "hard_kernels.hpp"
#include "hb_host.hpp"
#include "long_kernels.hpp"
#include "reduce_kernels.hpp"
#include "soft_kernels.hpp"

using nwg::TileDesc;

struct nwg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr;   // a stream the context created (nwg_ctx_own_stream)
    int n_cu = 256;
    std::string err;
    ~nwg_ctx() {
        if (own) (void)hipStreamDestroy(own);
    }
};

namespace {

constexpr int kS = 16;                 // columns per lane: hard and score-only fills
constexpr int kSN = 8;                 // ... narrow: batches with too few tiles at kS to fill the SIMDs
// Columns per lane of long pairs (column panels of 64 lanes): 8 in float32 (twice the waves of
// 16, at 103 VGPRs), 16 in float64 (16 pairs of 36,000^2: 270 against 225 GCUPS at 8, half
// the panel hand-offs at the same one to two waves a SIMD).
constexpr int long_cols(bool f64) { return f64 ? 16 : 8; }
constexpr int kSS = 8;                 // columns per lane: soft pass (float32 and double)
constexpr int kMaxG = 64;              // lanes per pair at most: one wavefront
constexpr int kMaxN = kS * kMaxG;      // longest B a lane group holds (hard / score-only); wider: column panels
constexpr int kHistBytes = 16384;      // LDS histogram budget per wave (hard walk, uint16 counts)
constexpr int kSoftHistBytes = 13440;  // LDS histogram budget per wave (soft, double expected counts)
constexpr size_t kSoftLdsMax = 65536;  // LDS of one soft wavefront, at most (tracked: K sets)
constexpr int kBlocksPerCU = 8;        // persistent grid of the hard / score fills: waves per CU
constexpr int kSoftBlocksPerCU = 8;    // fallback persistent grid of the soft pass
constexpr int64_t kSumBlock = 4096;    // the reference's WEIGHTED_GRAD_BLOCK
constexpr size_t kScratchBudget = size_t(96) << 30;   // soft row-segment scratch of one launch, at most
constexpr size_t kTableBudget = size_t(64) << 30;     // long-pair pointer codes of one launch group, at most
constexpr int64_t kHardTableBudget = int64_t(16) << 30;   // short-pair pointer codes of one fill (default)
constexpr int64_t kHardMinLaunch = 1024;   // tiles of a per-bucket launch the budget cannot split further
constexpr int kRowBlock = 64;          // long pairs: rows a panel publishes at a time
constexpr uint64_t kWaitTicks = 200000000;   // long pairs: a row-block wait, at most (2 s of the 100 MHz s_memrealtime)

int fail(nwg_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define NWG_HIP(ctx, call)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), NWG_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_));       \
    } while (0)

template <typename P>
struct DevPtr {
    P* p = nullptr;
    size_t n = 0;
    DevPtr() = default;
    DevPtr(const DevPtr&) = delete;
    DevPtr& operator=(const DevPtr&) = delete;
    ~DevPtr() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        reset();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(P));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
    // grow to at least `count` elements (contents are not kept)
    hipError_t reserve(size_t count) { return n >= count && p ? hipSuccess : alloc(count); }
    // the same with headroom, for buffers a reloaded batch regrows
    hipError_t grow(size_t count) { return n >= count && p ? hipSuccess : alloc(count + count / 2); }
};

// Persistent grid: as many one-wave blocks as the kernel keeps resident per CU (registers
// and LDS permitting), times the CUs -- never more blocks than tiles.
template <typename K>
int64_t resident_blocks(nwg_ctx* ctx, K kern, size_t lds, int fallback_per_cu) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds) != hipSuccess || per_cu <= 0)
        per_cu = fallback_per_cu;
    return (int64_t)per_cu * ctx->n_cu;
}

int next_pow2(int x) {
    int g = 1;
    while (g < x) g <<= 1;
    return g;
}

struct Bucket {
    int G;
    int tile_begin, tile_end;  // tiles [begin, end) of the plain order
};

// The planner's output for a set of pairs: tiles of 64/G pairs with one lane-group width
// G each (G from len(B), at least gmin), pairs sorted by len(A) within a width so a
// tile's rows are even; the plain order (bucket by bucket) and the persistent order
// (longest tile first) of the same tiles.
// One plain launch of the hard / score-only fill: tiles [begin, end) of the plain order, each
// block's pointer-code slot sized for the longest of them.
struct HardLaunch {
    int32_t begin, end;
    int64_t slot_dwords;
};

struct Plan {
    int S = kS;              // columns per lane the tiles were cut for
    std::vector<TileDesc> tiles;
    std::vector<Bucket> buckets;
    std::vector<int32_t> order;
    int64_t max_steps = 1;   // max over tiles of max_m + G - 1
    int panels = 1;          // max over tiles of the column panels of its pair
    bool built = false;
    DevPtr<int32_t> d_order;
    DevPtr<TileDesc> d_tiles, d_ptiles;
    // the per-bucket fill's launches: each bucket in spans whose pointer codes fit the batch's
    // table budget, and the table the largest needs
    std::vector<HardLaunch> bucket_launches;
    int64_t bucket_table_dwords = 0;
};

// The soft planner's output: tiles of one or more pairs side by side (nwg::SoftTile).
struct SoftPlan {
    std::vector<nwg::SoftTile> tiles;   // plan order (plain fill: bucket by bucket)
    std::vector<Bucket> buckets;
    std::vector<int32_t> order;
    int64_t rows = 1;        // max over tiles of fsteps + 1
    int panels = 1;          // max over tiles of the column panels
    int64_t work = 0;        // wave-steps over all tiles: 2 x panels x fsteps (forward and backward)
    std::vector<nwg::SoftTile> ptiles;   // the persistent order (longest first)
    std::vector<int64_t> pwork;          // per tile of it: wave-steps
    int64_t sync_words = 0;  // pipelined wide tiles: error word, done flags, each tile's words
    bool built = false;
    DevPtr<int32_t> d_order;
    DevPtr<nwg::SoftTile> d_tiles, d_ptiles;
    // persistent wide tiles: the queue of items and the pool regions' wait lists, built for
    // one count of pipelined tiles, pool size and lag
    std::vector<int2> items;
    std::vector<int32_t> waits;
    DevPtr<int2> d_items;
    DevPtr<int32_t> d_waits;
    int64_t items_tiles = -1, items_pooled = -1, items_base = -1, items_pool = 0, items_lag = -1;
    int64_t pipelined_panels = 0;
};

// Long pairs (len(B) > kMaxN): one item per (pair, column panel), in launch groups whose
// pointer tables fit kTableBudget.  Offsets in the items are relative to their group.
struct LongGroup {
    int pair_begin, pair_end, item_begin, item_end;
    int64_t table_dwords, bnd_elems;
};
struct LongPlan {
    std::vector<nwg::PanelItem> items;
    std::vector<nwg::LongPair> lpairs;
    std::vector<LongGroup> groups;
    int64_t max_items = 0, max_table = 0, max_bnd = 0;   // over groups
    DevPtr<nwg::PanelItem> d_items;
    DevPtr<nwg::LongPair> d_lpairs;
};

}  // namespace

struct nwg_batch {
    nwg_ctx* ctx = nullptr;
    nwg_problem prob{};
    int na = 0;
    int64_t N = 0;
    std::vector<int32_t> mlen, nlen;
    // device-resident pairs and results
    DevPtr<uint8_t> ca, cb;
    DevPtr<int64_t> oa, ob;
    DevPtr<double> scores, grads;
    // plans: the pairs one lane group holds, the wider ones (column panels), and the pairs
    // a float32 soft pass refused
    int gmin = 1, hist_slots = 64;
    Plan plan;               // hard / score-only: len(B) within one wavefront's columns
    LongPlan lpl;            // hard / score-only: wider pairs (column panels)
    // soft: packed tiles (persistent fill) and per-bucket tiles (plain fill), each split into
    // one-panel and wide (column-panel) tiles; the pairs a float32 pass refused
    SoftPlan sp_main, sp_wide, sq_main, sq_wide, rr_main, rr_wide;
    int64_t cells = 0;
    DevPtr<int32_t> queue;
    // hard: pointer codes per wave slot
    DevPtr<uint32_t> table;
    // long pairs: pointer codes, boundary columns, per-item counters and candidates, error word
    DevPtr<uint32_t> ltable;
    DevPtr<float> lbnd32;
    DevPtr<double> lbnd64;
    DevPtr<int32_t> lprog, lerr;
    DevPtr<unsigned char> lcand;
    // soft: row segments of the main and the long-pair pools, panel columns, weights
    DevPtr<float> seg32, lseg32, pb32, pf32, gw32;
    DevPtr<double> seg64, lseg64, pb64, pf64, gw64;
    DevPtr<double> theta;    // params on the device: matrix, go_a, ge_a, go_b, ge_b
    DevPtr<int32_t> pexp;
    DevPtr<int32_t> lqueue;
    int64_t soft_grid = 0, soft_lslots = 0, soft_occ = 0;   // the last soft launch's shape
    int64_t soft_pool = 0, soft_lag = 0, soft_ptiles = 0, soft_qtiles = 0;   // and its wide pool's
    DevPtr<double> lzpart, lpart;                            // pipelined wide tiles' parts
    DevPtr<int32_t> lsync;                                   // and words
    bool soft_piped = false;                                 // a launch since the last check
    size_t soft_scratch = kScratchBudget;                    // row-segment scratch of one launch
    int64_t table_budget = kHardTableBudget;                 // short-pair pointer codes of one fill
    int64_t hard_grid = 0;                                   // the last hard fill's persistent grid (0: plain)
    int soft_occ_cache[8] = {};   // resident blocks per CU: float, double; untracked, tracked; full, banded
    // hard: the guides of the last pass (pair p's m + 1 entries at off_a[p] + p), and its stored
    // paths when kept (pair p's ops end at off_a[p+1] + off_b[p+1]; op counts; end cells)
    DevPtr<int32_t> guide;
    DevPtr<uint32_t> ops;
    DevPtr<int32_t> ops_n, ends;
    bool keep_paths = false;     // the next passes store their paths
    bool guides_valid = false, paths_valid = false;
    bool banded = false;         // while a banded pass is launched: its half-width
    int band = 0;
    int lanes_S = 0;             // columns per lane of the hard / score-only tiles: 0 = chosen per batch
    DevPtr<uint8_t> soft_ok;
    std::vector<uint8_t> host_ok;
    int64_t soft_reruns = 0;
    DevPtr<unsigned long long> srec;   // soft: refusal record of the main passes (reduce_kernels.hpp)
    // params
    std::vector<double> matrix;           // K matrices, n_alpha^2 each
    std::vector<double> gaps = std::vector<double>(4, 0.0);   // K x (go_a, ge_a, go_b, ge_b)
    int K = 1;                            // parameter sets
    bool have_params = false;
    // tracks: set per residue of A (index off_a[p] + i - 1) and per DP row (off_a[p] + p + i)
    bool tracked = false, bx_dirty = true;
    std::vector<int64_t> hoa;             // off_a, host copy
    std::vector<uint16_t> htrk, hgtrk;
    DevPtr<uint16_t> dtrk, dgtrk;
    DevPtr<float> bx32, gt32;             // column-0 borders per DP row; the K x 4 gap table
    DevPtr<double> bx64, gt64;
    DevPtr<float> sub_f32;
    DevPtr<double> sub_f64;
    // reduction scratch
    DevPtr<double> partial, wdev, outdev;
    DevPtr<nwg::ScoreStats> stats;   // the score total's parts, then the result
    double* host_scores = nullptr;   // pinned
    int64_t host_cap = 0;
    nwg::ScoreStats* host_stats = nullptr;   // pinned
    int64_t total_fast = 0, total_ordered = 0;   // totals taken from the device sum / in pair order
    // settings
    int fill = 0;
    double temperature = 1.0;
    int soft_precision = 0;
    bool results_valid = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // traceback (0 auto, 1 pointers, 2 scores, 3 hirschberg, 4 hirschberg_pmax) and the
    // Hirschberg base case; the pairs a hard pass runs through the recursion, and whether the
    // current plans leave them out
    int traceback = 0;
    int hb_cutoff = 512;
    std::vector<int32_t> hb_pairs;
    bool plan_excl_hb = false;
    std::vector<int64_t> hob;             // off_b, host copy
    std::unique_ptr<nwg_hb::Work, void (*)(nwg_hb::Work*)> hbw{nullptr, nwg_hb::work_free};

    ~nwg_batch() {
        if (host_scores) (void)hipHostFree(host_scores);
        if (host_stats) (void)hipHostFree(host_stats);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

namespace {

// Lane-group width of a pair: enough lanes of S columns for len(B) (at least gmin, at most
// one wavefront); the hard walk's cross-lane reductions need a power of two.
int group_width(int n, int S, int gmin, bool pow2) {
    int g = (n + S - 1) / S;
    if (pow2) g = next_pow2(g);
    return std::min(std::max(g, gmin), kMaxG);
}

// The per-bucket fill's launches: each bucket in spans whose pointer codes (one slot per tile,
// sized for the span's longest tile) fit the batch's table budget, never fewer than
// kHardMinLaunch tiles; score-only fills keep no codes and take each bucket whole.
void split_launches(nwg_batch* b, Plan& pl) {
    const bool codes = b->prob.grad_mode == 1;
    const int S = pl.S;
    pl.bucket_launches.clear();
    pl.bucket_table_dwords = 0;
    for (const Bucket& bk : pl.buckets) {
        int c0 = bk.tile_begin;
        while (c0 < bk.tile_end) {
            int64_t steps = 1;
            int c1 = c0;
            while (c1 < bk.tile_end) {
                const TileDesc& t = pl.tiles[(size_t)c1];
                const int64_t s = std::max<int64_t>(steps, (int64_t)t.max_m + t.G - 1);
                const int64_t bytes = (int64_t)(c1 - c0 + 1) * s * (S / 4) * 64 * 4;
                if (codes && c1 - c0 >= kHardMinLaunch && bytes > b->table_budget) break;
                steps = s;
                ++c1;
            }
            const int64_t sd = steps * (S / 4) * 64;
            pl.bucket_launches.push_back(HardLaunch{c0, c1, sd});
            pl.bucket_table_dwords = std::max<int64_t>(pl.bucket_table_dwords, (int64_t)(c1 - c0) * sd);
            c0 = c1;
        }
    }
}

// Whether the default fill of a plan runs persistent: its pointer codes, a slot per tile of
// the plan's longest steps, exceed the batch's table budget.
bool hard_persistent(const nwg_batch* b, const Plan& pl) {
    const int64_t slot_bytes = pl.max_steps * (pl.S / 4) * 64 * 4;
    return b->prob.grad_mode == 1 && (int64_t)pl.tiles.size() * slot_bytes > b->table_budget;
}

// The tiles of `pairs` (lengths mlen / nlen by pair index): lane-group width G from len(B)
// (S columns a lane, at least gmin lanes, a power of two with pow2), pairs sorted by G and then
// longest len(A) first, cut into tiles of 64 / G pairs of one width.  Fills order, tiles and
// buckets (the plain order), ptiles (the same tiles, most row steps first), max_steps and
// panels.  Host work only: the batch planner and the no-allocation entry share it.
void cut_tiles(const std::vector<int32_t>& mlen, const std::vector<int32_t>& nlen, const std::vector<int32_t>& pairs,
               int S, int gmin, bool pow2, std::vector<int32_t>& order, std::vector<TileDesc>& tiles,
               std::vector<Bucket>& buckets, std::vector<TileDesc>& ptiles, int64_t& max_steps, int& panels) {
    tiles.clear();
    buckets.clear();
    max_steps = 1;
    panels = 1;
    const int64_t N = (int64_t)pairs.size();
    std::vector<int32_t> gsel(std::max<size_t>(mlen.size(), 1));
    for (int32_t i : pairs) {
        gsel[(size_t)i] = group_width(nlen[(size_t)i], S, gmin, pow2);
        panels = std::max(panels, (nlen[(size_t)i] + S * kMaxG - 1) / (S * kMaxG));
    }
    order = pairs;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (gsel[(size_t)x] != gsel[(size_t)y]) return gsel[(size_t)x] < gsel[(size_t)y];
        return mlen[(size_t)x] > mlen[(size_t)y];
    });
    int64_t pos = 0;
    while (pos < N) {
        const int G = gsel[(size_t)order[(size_t)pos]];
        Bucket bk{G, (int)tiles.size(), 0};
        while (pos < N && gsel[(size_t)order[(size_t)pos]] == G) {
            const int per = 64 / G;
            TileDesc td{};
            td.first = (int32_t)pos;
            td.G = (int16_t)G;
            int np = 0, mm = 0, mn = 0;
            while (np < per && pos < N && gsel[(size_t)order[(size_t)pos]] == G) {
                mm = std::max(mm, mlen[(size_t)order[(size_t)pos]]);
                mn = std::max(mn, nlen[(size_t)order[(size_t)pos]]);
                ++np;
                ++pos;
            }
            td.npairs = (int16_t)np;
            td.max_m = mm;
            td.max_n = mn;
            max_steps = std::max<int64_t>(max_steps, (int64_t)mm + G - 1);
            tiles.push_back(td);
        }
        bk.tile_end = (int)tiles.size();
        buckets.push_back(bk);
    }
    // persistent order: most row steps first (a wave's time is its steps times its panels)
    ptiles = tiles;
    auto work = [S](const TileDesc& t) {
        const int pw = t.G * S;
        const int64_t np = t.max_n > pw ? (t.max_n + pw - 1) / pw : 1;
        return ((int64_t)t.max_m + t.G - 1) * np;
    };
    std::stable_sort(ptiles.begin(), ptiles.end(),
                     [&](const TileDesc& a, const TileDesc& c) { return work(a) > work(c); });
}

int build_plan(nwg_batch* b, const std::vector<int32_t>& pairs, Plan& pl, int S, bool pow2) {
    nwg_ctx* ctx = b->ctx;
    pl.S = S;
    const int64_t N = (int64_t)pairs.size();
    std::vector<TileDesc> ptiles;
    cut_tiles(b->mlen, b->nlen, pairs, S, b->gmin, pow2, pl.order, pl.tiles, pl.buckets, ptiles, pl.max_steps,
              pl.panels);
    split_launches(b, pl);
    NWG_HIP(ctx, pl.d_order.reserve(std::max<size_t>(pl.order.size(), 1)));
    NWG_HIP(ctx, pl.d_tiles.reserve(std::max<size_t>(pl.tiles.size(), 1)));
    NWG_HIP(ctx, pl.d_ptiles.reserve(std::max<size_t>(pl.tiles.size(), 1)));
    if (N)
        NWG_HIP(ctx, hipMemcpyAsync(pl.d_order.p, pl.order.data(), sizeof(int32_t) * (size_t)N, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (!pl.tiles.empty()) {
        NWG_HIP(ctx, hipMemcpyAsync(pl.d_tiles.p, pl.tiles.data(), sizeof(TileDesc) * pl.tiles.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        NWG_HIP(ctx, hipMemcpyAsync(pl.d_ptiles.p, ptiles.data(), sizeof(TileDesc) * ptiles.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    pl.built = true;
    return NWG_OK;
}

// Long pairs: items in order of cells (largest first), a pair's panels consecutive and in
// panel order -- the order the fill's waits rely on.  Groups close when the next pair's
// pointer table would pass kTableBudget.
int build_long_plan(nwg_batch* b, std::vector<int32_t> pairs, LongPlan& L) {
    nwg_ctx* ctx = b->ctx;
    const bool hard = b->prob.grad_mode == 1;
    const int LS = long_cols(b->prob.precision == 1);
    const int PW = LS * kMaxG;
    const int64_t SD = (LS / 4) * kMaxG;   // table dwords per step
    L.items.clear();
    L.lpairs.clear();
    L.groups.clear();
    L.max_items = L.max_table = L.max_bnd = 0;
    std::stable_sort(pairs.begin(), pairs.end(), [&](int32_t x, int32_t y) {
        return (int64_t)b->mlen[(size_t)x] * b->nlen[(size_t)x] > (int64_t)b->mlen[(size_t)y] * b->nlen[(size_t)y];
    });
    LongGroup gr{0, 0, 0, 0, 0, 0};
    auto close = [&]() {
        gr.pair_end = (int)L.lpairs.size();
        gr.item_end = (int)L.items.size();
        if (gr.pair_end > gr.pair_begin) {
            L.groups.push_back(gr);
            L.max_items = std::max<int64_t>(L.max_items, gr.item_end - gr.item_begin);
            L.max_table = std::max(L.max_table, gr.table_dwords);
            L.max_bnd = std::max(L.max_bnd, gr.bnd_elems);
        }
        gr = LongGroup{gr.pair_end, 0, gr.item_end, 0, 0, 0};
    };
    for (int32_t p : pairs) {
        const int m = b->mlen[(size_t)p], n = b->nlen[(size_t)p];
        const int P = (n + PW - 1) / PW;
        const int64_t per_panel = hard ? ((int64_t)m + kMaxG - 1) * SD : 0;
        const int64_t tbl = per_panel * P;
        if ((size_t)tbl * 4 > kTableBudget)
            return fail(ctx, NWG_EINVAL,
                        "nwgrad_hip: pair %d (%d x %d) needs a %.1f GiB pointer table; this build keeps at most %.0f GiB "
                        "of codes per launch",
                        (int)p, m, n, (double)tbl * 4 / (1 << 30), (double)(kTableBudget >> 30));
        if (gr.table_dwords + tbl > (int64_t)(kTableBudget / 4)) close();
        nwg::LongPair lp{p, (int32_t)(L.items.size() - gr.item_begin), P, 0};
        L.lpairs.push_back(lp);
        for (int k = 0; k < P; ++k) {
            nwg::PanelItem it{};
            it.pair = p;
            it.panel = k;
            it.table_off = gr.table_dwords + per_panel * k;
            it.bnd_in = k == 0 ? -1 : gr.bnd_elems - 3 * ((int64_t)m + 1);
            it.bnd_out = -1;
            if (k + 1 < P) {
                it.bnd_out = gr.bnd_elems;
                gr.bnd_elems += 3 * ((int64_t)m + 1);
            }
            L.items.push_back(it);
        }
        gr.table_dwords += tbl;
    }
    close();
    NWG_HIP(ctx, L.d_items.reserve(std::max<size_t>(L.items.size(), 1)));
    NWG_HIP(ctx, L.d_lpairs.reserve(std::max<size_t>(L.lpairs.size(), 1)));
    if (!L.items.empty()) {
        NWG_HIP(ctx, hipMemcpyAsync(L.d_items.p, L.items.data(), sizeof(nwg::PanelItem) * L.items.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        NWG_HIP(ctx, hipMemcpyAsync(L.d_lpairs.p, L.lpairs.data(), sizeof(nwg::LongPair) * L.lpairs.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return NWG_OK;
}

// The lane-width table of the hard / score-only fills (nwg_set_lane_table): rules (n_max,
// tiles16_max, S), first match wins.  A batch whose median len(B) is at most n_max and whose
// tile count at 16 columns a lane is at most tiles16_max / 16 per CU cuts its tiles at S
// columns a lane.  Without a matching rule the built-in choice holds (choose_lanes).
struct LaneRule {
    int32_t n_max, tiles16_max, S;
};
std::mutex g_lane_mu;
std::vector<LaneRule> g_lane_rules;

// Columns a lane for the tiles of `sp`, given their tile count at kS.  Fewer tiles than two
// per SIMD leave the fill latency-bound (one wave a SIMD has nothing to hide a step's latency
// behind): kSN columns a lane gives about twice as many waves, when every pair fits one wave
// at that width (protein_global_hard, 1,182 tiles at 16 columns: 266 -> 328 GCUPS;
// dna_local_short, 65,536 tiles, keeps 16, where 8 measured 18 % slower).  A rule of the
// lane-width table, or the batch's own setting, overrides that choice.
int choose_lanes(const std::vector<int32_t>& nlen, const std::vector<int32_t>& sp, int64_t tiles_at_kS, int n_cu,
                 int forced) {
    int64_t widest = 0;
    for (int32_t i : sp) widest = std::max<int64_t>(widest, nlen[(size_t)i]);
    const bool narrow_fits = widest <= kSN * kMaxG;
    if (forced == kS || (forced == kSN && !narrow_fits)) return kS;
    if (forced == kSN) return kSN;
    {
        std::lock_guard<std::mutex> lk(g_lane_mu);
        if (!g_lane_rules.empty() && !sp.empty()) {
            std::vector<int32_t> ns;
            ns.reserve(sp.size());
            for (int32_t i : sp) ns.push_back(nlen[(size_t)i]);
            std::nth_element(ns.begin(), ns.begin() + ns.size() / 2, ns.end());
            const int32_t med = ns[ns.size() / 2];
            const double per_cu16 = 16.0 * (double)tiles_at_kS / (double)std::max(1, n_cu);
            for (const LaneRule& r : g_lane_rules)
                if (med <= r.n_max && per_cu16 <= (double)r.tiles16_max)
                    return (r.S == kSN && narrow_fits) ? kSN : kS;
        }
    }
    return (tiles_at_kS > 0 && tiles_at_kS < 2 * 4 * (int64_t)n_cu && narrow_fits) ? kSN : kS;
}

// The pairs of `pairs`: those within one wavefront's columns into the tile plan, the wider
// ones into the long-pair plan.
int build_plans(nwg_batch* b, const std::vector<int32_t>& pairs, Plan& main, LongPlan& wide) {
    std::vector<int32_t> sp, lp;
    for (int32_t i : pairs) (b->nlen[(size_t)i] > kMaxN ? lp : sp).push_back(i);
    int rc = build_plan(b, sp, main, kS, true);
    if (rc) return rc;
    if (choose_lanes(b->nlen, sp, (int64_t)main.tiles.size(), b->ctx->n_cu, b->lanes_S) == kSN) {
        rc = build_plan(b, sp, main, kSN, true);
        if (rc) return rc;
    }
    return build_long_plan(b, lp, wide);
}

// The pairs a hard pass takes through the Hirschberg recursion: under traceback "hirschberg"
// or "hirschberg_pmax", those of more than hb_cutoff rows (hb_splits, aligner.hpp:136-141:
// an explicit Hirschberg mode has no pointer budget); the shorter ones run as Pointers.
void find_split_pairs(nwg_batch* b) {
    b->hb_pairs.clear();
    if (b->prob.grad_mode != 1 || (b->traceback != 3 && b->traceback != 4)) return;
    for (int64_t p = 0; p < b->N; ++p)
        if (b->mlen[(size_t)p] > b->hb_cutoff) b->hb_pairs.push_back((int32_t)p);
}

// (Re)builds the hard / score-only plans: every pair for a banded pass (the band bounds memory
// already, and the reference's banded fill takes no Hirschberg mode), the others without the
// pairs the recursion takes.
int replan(nwg_batch* b) {
    find_split_pairs(b);
    const bool excl = !b->hb_pairs.empty() && !b->banded;
    std::vector<int32_t> pairs;
    pairs.reserve((size_t)b->N);
    size_t k = 0;
    for (int64_t p = 0; p < b->N; ++p) {
        if (excl && k < b->hb_pairs.size() && b->hb_pairs[k] == (int32_t)p) {
            ++k;
            continue;
        }
        pairs.push_back((int32_t)p);
    }
    b->plan_excl_hb = excl;
    return build_plans(b, pairs, b->plan, b->lpl);
}

// Soft tiles.  A pair needs ceil(n / (64 S)) column panels at least; in a tile of NP panels
// it takes ceil(n / (S NP)) lanes.  Packed (persistent fill): pairs in order of len(A),
// longest first, go into the first open tile with enough free lanes, a free histogram slot
// and a len(A) within 1/16 of the tile's first (so the tile's rows stay even); a pair that
// fits one panel may also join an open wide tile.  Plain fill: per-bucket tiles of one
// power-of-two lane-group width (64 / G pairs each), and one wide pair per tile.
struct SoftBin {
    nwg::SoftTile t;
    std::vector<int32_t> pairs;
    int used = 0;
};

int upload_soft_plan(nwg_batch* b, SoftPlan& pl, std::vector<SoftBin>& bins, std::vector<Bucket> buckets) {
    nwg_ctx* ctx = b->ctx;
    pl.tiles.clear();
    pl.order.clear();
    pl.items.clear();
    pl.waits.clear();
    pl.items_tiles = -1;
    pl.items_pooled = -1;
    pl.items_base = -1;
    pl.items_pool = 0;
    pl.items_lag = -1;
    pl.pipelined_panels = 0;
    pl.rows = 1;
    pl.panels = 1;
    pl.work = 0;
    pl.pwork.clear();
    for (SoftBin& sb : bins) {
        nwg::SoftTile t = sb.t;
        t.first = (int32_t)pl.order.size();
        t.npairs = (int16_t)sb.pairs.size();
        for (int32_t i : sb.pairs) pl.order.push_back(i);
        pl.rows = std::max<int64_t>(pl.rows, (int64_t)t.fsteps + 1);
        pl.panels = std::max<int>(pl.panels, t.panels);
        pl.work += 2 * (int64_t)t.panels * t.fsteps;
        t.sync = 0;
        pl.tiles.push_back(t);
    }
    pl.buckets = std::move(buckets);
    std::vector<nwg::SoftTile> ptiles(pl.tiles);
    std::stable_sort(ptiles.begin(), ptiles.end(), [](const nwg::SoftTile& a, const nwg::SoftTile& c) {
        return (int64_t)a.fsteps * a.panels > (int64_t)c.fsteps * c.panels;
    });
    // the persistent order's words for pipelined wide tiles: after the error word and one
    // done flag per tile, per tile: forward and backward steps done per panel, forward and
    // backward items done
    int64_t words = 1 + (int64_t)ptiles.size();
    for (nwg::SoftTile& t : ptiles) {
        t.sync = (int32_t)words;
        words += 2 * (int64_t)t.panels + 2;
        pl.pwork.push_back(2 * (int64_t)t.panels * t.fsteps);
        t.rbase = 0;
        t.wait0 = t.nwait = t.poff = t.pad = 0;
    }
    pl.sync_words = words;
    pl.ptiles = ptiles;
    const size_t N = pl.order.size();
    NWG_HIP(ctx, pl.d_order.reserve(std::max<size_t>(N, 1)));
    NWG_HIP(ctx, pl.d_tiles.reserve(std::max<size_t>(pl.tiles.size(), 1)));
    NWG_HIP(ctx, pl.d_ptiles.reserve(std::max<size_t>(pl.tiles.size(), 1)));
    if (N)
        NWG_HIP(ctx, hipMemcpyAsync(pl.d_order.p, pl.order.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (!pl.tiles.empty()) {
        NWG_HIP(ctx, hipMemcpyAsync(pl.d_tiles.p, pl.tiles.data(), sizeof(nwg::SoftTile) * pl.tiles.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        NWG_HIP(ctx, hipMemcpyAsync(pl.d_ptiles.p, ptiles.data(), sizeof(nwg::SoftTile) * ptiles.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    pl.built = true;
    return NWG_OK;
}

void bin_add(nwg_batch* b, SoftBin& sb, int32_t i, int G) {
    nwg::SoftTile& t = sb.t;
    const int k = (int)sb.pairs.size();
    t.G[k] = (uint8_t)G;
    t.lo[k] = (uint8_t)sb.used;
    sb.used += G;
    sb.pairs.push_back(i);
    t.npairs = (int16_t)sb.pairs.size();
    t.max_m = std::max(t.max_m, b->mlen[(size_t)i]);
    t.max_n = std::max(t.max_n, b->nlen[(size_t)i]);
    t.fsteps = std::max(t.fsteps, b->mlen[(size_t)i] + G - 1);
}

SoftBin new_bin(int panels) {
    SoftBin sb;
    std::memset(&sb.t, 0, sizeof sb.t);
    sb.t.panels = (int16_t)panels;
    sb.t.fsteps = 1;
    return sb;
}

int build_soft_plans(nwg_batch* b, const std::vector<int32_t>& pairs, SoftPlan& main, SoftPlan& wide, bool packed) {
    const int S = kSS;
    const int slots = std::min(b->hist_slots, nwg::kTilePairs);
    auto min_panels = [&](int32_t i) { return std::max(1, (b->nlen[(size_t)i] + S * kMaxG - 1) / (S * kMaxG)); };
    auto lanes = [&](int32_t i, int np) { return std::max(1, (b->nlen[(size_t)i] + S * np - 1) / (S * np)); };
    std::vector<SoftBin> mbins, wbins;
    std::vector<Bucket> mb, wb;
    if (packed) {
        std::vector<int32_t> ord(pairs);
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int32_t x, int32_t y) { return b->mlen[(size_t)x] > b->mlen[(size_t)y]; });
        constexpr size_t kWindow = 32;
        std::vector<size_t> mopen, wopen;   // indices of open bins
        auto close_full = [&](std::vector<size_t>& open, std::vector<SoftBin>& bins) {
            open.erase(std::remove_if(open.begin(), open.end(),
                                      [&](size_t k) {
                                          return bins[k].used >= kMaxG || (int)bins[k].pairs.size() >= slots;
                                      }),
                       open.end());
            while (open.size() > kWindow) open.erase(open.begin());
        };
        auto near = [&](const SoftBin& sb, int32_t i) {
            const int m0 = sb.t.max_m, m = b->mlen[(size_t)i];
            return m0 - m <= std::max(4, m0 >> 4);
        };
        auto try_open = [&](std::vector<size_t>& open, std::vector<SoftBin>& bins, int32_t i, int need_np) {
            for (size_t k : open) {
                SoftBin& sb = bins[k];
                if (sb.t.panels < need_np || !near(sb, i)) continue;
                const int G = lanes(i, sb.t.panels);
                if (sb.used + G <= kMaxG && (int)sb.pairs.size() < slots) {
                    bin_add(b, sb, i, G);
                    return true;
                }
            }
            return false;
        };
        for (int32_t i : ord) {
            const int np = min_panels(i);
            if (np > 1) {
                if (!try_open(wopen, wbins, i, np)) {
                    wbins.push_back(new_bin(np));
                    bin_add(b, wbins.back(), i, lanes(i, np));
                    wopen.push_back(wbins.size() - 1);
                }
                close_full(wopen, wbins);
            } else {
                if (!try_open(wopen, wbins, i, 1) && !try_open(mopen, mbins, i, 1)) {
                    mbins.push_back(new_bin(1));
                    bin_add(b, mbins.back(), i, lanes(i, 1));
                    mopen.push_back(mbins.size() - 1);
                }
                close_full(mopen, mbins);
                close_full(wopen, wbins);
            }
        }
        mb.push_back(Bucket{0, 0, (int)mbins.size()});
        wb.push_back(Bucket{0, 0, (int)wbins.size()});
    } else {
        // plain: power-of-two lane groups, 64 / G pairs per tile, no more than the slots
        int gmin = 1;
        while (64 / gmin > slots) gmin <<= 1;
        std::vector<int32_t> ms, ws;
        for (int32_t i : pairs) (min_panels(i) > 1 ? ws : ms).push_back(i);
        auto gsel = [&](int32_t i) { return std::min(std::max(next_pow2(lanes(i, 1)), gmin), kMaxG); };
        std::stable_sort(ms.begin(), ms.end(), [&](int32_t x, int32_t y) {
            if (gsel(x) != gsel(y)) return gsel(x) < gsel(y);
            return b->mlen[(size_t)x] > b->mlen[(size_t)y];
        });
        size_t pos = 0;
        while (pos < ms.size()) {
            const int G = gsel(ms[pos]);
            Bucket bk{G, (int)mbins.size(), 0};
            while (pos < ms.size() && gsel(ms[pos]) == G) {
                mbins.push_back(new_bin(1));
                for (int k = 0; k < 64 / G && pos < ms.size() && gsel(ms[pos]) == G; ++k) bin_add(b, mbins.back(), ms[pos++], G);
            }
            bk.tile_end = (int)mbins.size();
            mb.push_back(bk);
        }
        std::stable_sort(ws.begin(), ws.end(), [&](int32_t x, int32_t y) {
            if (min_panels(x) != min_panels(y)) return min_panels(x) < min_panels(y);
            return b->mlen[(size_t)x] > b->mlen[(size_t)y];
        });
        pos = 0;
        while (pos < ws.size()) {
            const int np = min_panels(ws[pos]);
            Bucket bk{np, (int)wbins.size(), 0};
            while (pos < ws.size() && min_panels(ws[pos]) == np) {
                wbins.push_back(new_bin(np));
                bin_add(b, wbins.back(), ws[pos], lanes(ws[pos], np));
                ++pos;
            }
            bk.tile_end = (int)wbins.size();
            wb.push_back(bk);
        }
    }
    int rc = upload_soft_plan(b, main, mbins, mb);
    if (rc) return rc;
    return upload_soft_plan(b, wide, wbins, wb);
}

// LDS slots of the hard walk's histograms (K sets of uint16 counts per pair slot, plus K x 4
// gap counters when tracked) and of the soft pass (n_alpha x (n_alpha + 1) doubles): the slot
// count fixes the narrowest lane group.  Returns whether it changed.
bool set_slots(nwg_batch* b) {
    const int na = b->na, nn = na * na;
    int gmin = 1, slots = 64;
    if (b->prob.grad_mode == 2) {
        const int K = b->tracked ? b->K : 1;
        slots = std::max(1, std::min(nwg::kTilePairs, kSoftHistBytes / (K * na * (na + 1) * 8)));
        while (64 / gmin > slots) ++gmin;
    } else {
        const int per_slot = b->prob.grad_mode == 1 ? b->K * nn * 2 + (b->tracked ? 16 * b->K : 0) : 0;
        while (gmin < kMaxG && (64 / gmin) * per_slot > kHistBytes) gmin <<= 1;
        slots = 64 / gmin;
    }
    const bool changed = gmin != b->gmin || slots != b->hist_slots;
    b->gmin = gmin;
    b->hist_slots = slots;
    if (b->prob.grad_mode == 2) {
        // the soft tiles hold at most `slots` pairs and the LDS follows K: plan and query again
        if (changed)
            for (SoftPlan* sp : {&b->sp_main, &b->sp_wide, &b->sq_main, &b->sq_wide}) sp->built = false;
        for (int& o : b->soft_occ_cache) o = 0;
    }
    return changed;
}

// Tracks on the host: `trk` one set per residue of A (index off_a[p] + i - 1), `gtrk` one
// per DP row (index off_a[p] + p + i); NULL: the reference's defaults (setup_tracks_,
// aligner.hpp:1956-1967: set 0; row 0 the first residue's set, row i the set of residue i).
int set_tracks_host(nwg_batch* b, const uint16_t* trk, const uint16_t* gtrk) {
    nwg_ctx* ctx = b->ctx;
    const int64_t N = b->N, la = N ? b->hoa[(size_t)N] : 0;
    std::vector<uint16_t> t((size_t)std::max<int64_t>(la, 1), 0), g((size_t)std::max<int64_t>(la + N, 1), 0);
    for (int64_t p = 0; p < N; ++p) {
        const int64_t a0 = b->hoa[(size_t)p], m = b->mlen[(size_t)p];
        for (int64_t i = 0; i < m; ++i) {
            const uint16_t v = trk ? trk[a0 + i] : 0;
            if (v >= b->K)
                return fail(ctx, NWG_EINVAL, "nwgrad: track[%lld] = %d of pair %lld is not a parameter set (params has %d)",
                            (long long)i, (int)v, (long long)p, b->K);
            t[(size_t)(a0 + i)] = v;
        }
        for (int64_t i = 0; i <= m; ++i) {
            uint16_t v;
            if (gtrk) v = gtrk[a0 + p + i];
            else v = i == 0 ? (m ? t[(size_t)a0] : 0) : t[(size_t)(a0 + i - 1)];
            if (v >= b->K)
                return fail(ctx, NWG_EINVAL,
                            "nwgrad: gap_track[%lld] = %d of pair %lld is not a parameter set (params has %d)",
                            (long long)i, (int)v, (long long)p, b->K);
            g[(size_t)(a0 + p + i)] = v;
        }
    }
    b->htrk.swap(t);
    b->hgtrk.swap(g);
    NWG_HIP(ctx, b->dtrk.grow(b->htrk.size()));
    NWG_HIP(ctx, b->dgtrk.grow(b->hgtrk.size()));
    NWG_HIP(ctx, hipMemcpyAsync(b->dtrk.p, b->htrk.data(), 2 * b->htrk.size(), hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipMemcpyAsync(b->dgtrk.p, b->hgtrk.data(), 2 * b->hgtrk.size(), hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->tracked = true;
    b->bx_dirty = true;
    return NWG_OK;
}

// The per-row column-0 borders of every pair and the K x 4 gap table, in precision T, as
// the reference forms them (setup_tracks_, aligner.hpp:1968-2007): gaps cast once per set;
// the border of row i is -(go_b(1) + C_i), C_i the sum of ge_b over rows 1..i taken by runs of
// bitwise-equal extends (linear: -C_i).
template <typename T>
int upload_track_rows(nwg_batch* b, DevPtr<T>& bx, DevPtr<T>& gt) {
    nwg_ctx* ctx = b->ctx;
    const int K = b->K;
    std::vector<T> g4((size_t)4 * K);
    for (int k = 0; k < 4 * K; ++k) g4[(size_t)k] = static_cast<T>(b->gaps[(size_t)k]);
    const int64_t N = b->N;
    std::vector<T> out(b->hgtrk.size(), T(0));
    const bool linear = b->prob.gap_model == 0;
    for (int64_t p = 0; p < N; ++p) {
        const int64_t a0 = b->hoa[(size_t)p], m = b->mlen[(size_t)p];
        if (m == 0) continue;
        T* o = out.data() + a0 + p;
        const uint16_t* tk = b->htrk.data() + a0;
        const T go = g4[(size_t)4 * tk[0] + 2];
        T csum = T(0);
        int64_t s0 = 1;
        for (int64_t i = 1; i <= m; ++i) {
            const T ge = g4[(size_t)4 * tk[i - 1] + 3];
            if (i > 1) {
                const T gp = g4[(size_t)4 * tk[i - 2] + 3];
                if (std::memcmp(&ge, &gp, sizeof(T)) != 0) {
                    csum = (s0 == 1) ? static_cast<T>(i - 1) * gp : csum + static_cast<T>(i - s0) * gp;
                    s0 = i;
                }
            }
            if (linear) o[i] = (s0 == 1) ? -static_cast<T>(i) * ge : -(csum + static_cast<T>(i - s0 + 1) * ge);
            else o[i] = (s0 == 1) ? -(go + static_cast<T>(i) * ge) : -(go + (csum + static_cast<T>(i - s0 + 1) * ge));
        }
    }
    NWG_HIP(ctx, bx.grow(std::max<size_t>(out.size(), 1)));
    NWG_HIP(ctx, gt.grow((size_t)4 * K));
    if (!out.empty())
        NWG_HIP(ctx, hipMemcpyAsync(bx.p, out.data(), sizeof(T) * out.size(), hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipMemcpyAsync(gt.p, g4.data(), sizeof(T) * g4.size(), hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NWG_OK;
}

int prepare_tracks(nwg_batch* b) {
    if (!b->tracked || !b->bx_dirty) return NWG_OK;
    const int rc = b->prob.precision == 0 ? upload_track_rows<float>(b, b->bx32, b->gt32)
                                          : upload_track_rows<double>(b, b->bx64, b->gt64);
    if (rc) return rc;
    b->bx_dirty = false;
    return NWG_OK;
}

// Validate and upload N pairs into the batch, (re)build the hard / score-only plan and mark
// the soft plans stale.  Nothing of the batch changes when validation fails.  `keep`: grow
// the device buffers with headroom and never shrink them (a reloaded batch).
int load_pairs(nwg_batch* b, int64_t N, const uint8_t* codes_a, const int64_t* off_a, const uint8_t* codes_b,
               const int64_t* off_b, bool keep) {
    nwg_ctx* ctx = b->ctx;
    const int n_alpha = b->na, nn = n_alpha * n_alpha;
    if (N < 0 || N > INT32_MAX) return fail(ctx, NWG_EINVAL, "nwgrad_hip: invalid pair count");
    if (N > 0 && (!off_a || !off_b)) return fail(ctx, NWG_EINVAL, "nwgrad_hip: missing offsets");
    std::vector<int32_t> mlen((size_t)N), nlen((size_t)N);
    int64_t cells = 0;
    const int64_t la = N ? off_a[N] : 0, lb = N ? off_b[N] : 0;
    if (N && (off_a[0] < 0 || off_b[0] < 0)) return fail(ctx, NWG_EINVAL, "nwgrad_hip: negative offset");
    for (int64_t i = 0; i < N; ++i) {
        const int64_t m = off_a[i + 1] - off_a[i], n = off_b[i + 1] - off_b[i];
        if (m < 0 || n < 0) return fail(ctx, NWG_EINVAL, "nwgrad_hip: offsets must be non-decreasing");
        if (m > (1 << 24) || n > (1 << 24)) return fail(ctx, NWG_EINVAL, "nwgrad_hip: pair %lld is too long", (long long)i);
        mlen[(size_t)i] = (int32_t)m;
        nlen[(size_t)i] = (int32_t)n;
        cells += m * n;
    }
    for (int64_t k = 0; k < la; ++k)
        if (codes_a[k] >= n_alpha) return fail(ctx, NWG_EINVAL, "nwgrad_hip: code out of range in A");
    for (int64_t k = 0; k < lb; ++k)
        if (codes_b[k] >= n_alpha) return fail(ctx, NWG_EINVAL, "nwgrad_hip: code out of range in B");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    auto size = [&](auto& d, size_t count) { return keep ? d.grow(count) : d.alloc(count); };
    NWG_HIP(ctx, size(b->ca, (size_t)std::max<int64_t>(la, 1)));
    NWG_HIP(ctx, size(b->cb, (size_t)std::max<int64_t>(lb, 1)));
    NWG_HIP(ctx, size(b->oa, (size_t)N + 1));
    NWG_HIP(ctx, size(b->ob, (size_t)N + 1));
    NWG_HIP(ctx, size(b->scores, (size_t)std::max<int64_t>(N, 1)));
    if (b->prob.grad_mode != 0)
        NWG_HIP(ctx, size(b->grads, (size_t)std::max<int64_t>(N, 1) * (size_t)b->K * (nn + 4)));
    if (b->prob.grad_mode == 1) NWG_HIP(ctx, size(b->guide, (size_t)std::max<int64_t>(la + N, 1)));
    if (b->prob.grad_mode == 1 && b->keep_paths) {
        NWG_HIP(ctx, size(b->ops, (size_t)std::max<int64_t>(la + lb, 1)));
        NWG_HIP(ctx, size(b->ops_n, (size_t)std::max<int64_t>(N, 1)));
        NWG_HIP(ctx, size(b->ends, (size_t)std::max<int64_t>(2 * N, 1)));
    }
    if (la) NWG_HIP(ctx, hipMemcpyAsync(b->ca.p, codes_a, (size_t)la, hipMemcpyHostToDevice, ctx->stream));
    if (lb) NWG_HIP(ctx, hipMemcpyAsync(b->cb.p, codes_b, (size_t)lb, hipMemcpyHostToDevice, ctx->stream));
    const int64_t zero[1] = {0};
    NWG_HIP(ctx, hipMemcpyAsync(b->oa.p, N ? off_a : zero, sizeof(int64_t) * (size_t)(N + 1), hipMemcpyHostToDevice,
                                ctx->stream));
    NWG_HIP(ctx, hipMemcpyAsync(b->ob.p, N ? off_b : zero, sizeof(int64_t) * (size_t)(N + 1), hipMemcpyHostToDevice,
                                ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->host_cap < std::max<int64_t>(N, 1)) {
        if (b->host_scores) NWG_HIP(ctx, hipHostFree(b->host_scores));
        b->host_scores = nullptr;
        b->host_cap = 0;
        const int64_t cap = keep ? std::max<int64_t>(N, 1) + N / 2 : std::max<int64_t>(N, 1);
        NWG_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&b->host_scores), sizeof(double) * (size_t)cap,
                                   hipHostMallocDefault));
        b->host_cap = cap;
    }
    b->N = N;
    b->mlen.swap(mlen);
    b->nlen.swap(nlen);
    b->cells = cells;
    b->hoa.assign(off_a ? off_a : zero, (off_a ? off_a : zero) + (N ? N + 1 : 1));
    b->hob.assign(off_b ? off_b : zero, (off_b ? off_b : zero) + (N ? N + 1 : 1));
    if (b->tracked) {   // new pairs: the default tracks (set 0) until the caller gives others
        int rc = set_tracks_host(b, nullptr, nullptr);
        if (rc) return rc;
    }
    b->results_valid = false;
    b->guides_valid = b->paths_valid = false;
    b->soft_reruns = 0;
    for (SoftPlan* sp : {&b->sp_main, &b->sp_wide, &b->sq_main, &b->sq_wide}) sp->built = false;
    if (b->prob.grad_mode != 2) {
        const int rc = replan(b);
        if (rc) return rc;
    }
    return NWG_OK;
}

}  // namespace

extern "C" {

int nwg_init(int device, nwg_ctx** out) {
    if (!out) return NWG_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return NWG_EDEVICE;
    if (device < 0 || device >= count) return NWG_EINVAL;
    auto* c = new (std::nothrow) nwg_ctx();
    if (!c) return NWG_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return NWG_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->n_cu = prop.multiProcessorCount;
    *out = c;
    return NWG_OK;
}

void nwg_destroy(nwg_ctx* ctx) { delete ctx; }

const char* nwg_last_error(const nwg_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int nwg_set_stream(nwg_ctx* ctx, void* s) {
    if (!ctx) return NWG_EINVAL;
    ctx->stream = static_cast<hipStream_t>(s);
    return NWG_OK;
}

int nwg_ctx_own_stream(nwg_ctx* ctx) {
    if (!ctx) return NWG_EINVAL;
    if (ctx->own) {
        ctx->stream = ctx->own;
        return NWG_OK;
    }
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    NWG_HIP(ctx, hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking));
    ctx->stream = ctx->own;
    return NWG_OK;
}

int nwg_batch_create(nwg_ctx* ctx, const nwg_problem* prob, int n_alpha, int64_t N, const uint8_t* codes_a,
                     const int64_t* off_a, const uint8_t* codes_b, const int64_t* off_b, nwg_batch** out) {
    if (!ctx || !prob || !out) return NWG_EINVAL;
    *out = nullptr;
    if (prob->precision < 0 || prob->precision > 1 || prob->gap_model < 0 || prob->gap_model > 1 || prob->mode < 0 ||
        prob->mode > 2 || prob->grad_mode < 0 || prob->grad_mode > 2)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: invalid problem type");
    if (n_alpha < 1 || n_alpha > 64)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: alphabet size %d outside [1, 64]", n_alpha);
    if (N < 0 || N > INT32_MAX) return fail(ctx, NWG_EINVAL, "nwgrad_hip: invalid pair count");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<nwg_batch> b(new (std::nothrow) nwg_batch());
    if (!b) return fail(ctx, NWG_ENOMEM, "nwgrad_hip: out of host memory");
    b->ctx = ctx;
    b->prob = *prob;
    b->na = n_alpha;
    b->N = N;
    set_slots(b.get());
    NWG_HIP(ctx, b->queue.alloc(4));
    if (prob->grad_mode == 2) {
        NWG_HIP(ctx, b->srec.alloc(3));
        NWG_HIP(ctx, hipMemset(b->srec.p, 0, 3 * sizeof(unsigned long long)));
    }
    const int rc = load_pairs(b.get(), N, codes_a, off_a, codes_b, off_b, false);
    if (rc) return rc;
    NWG_HIP(ctx, hipEventCreate(&b->ev0));
    NWG_HIP(ctx, hipEventCreate(&b->ev1));
    *out = b.release();
    return NWG_OK;
}

void nwg_batch_free(nwg_batch* b) { delete b; }

int nwg_batch_reload(nwg_batch* b, int64_t N, const uint8_t* codes_a, const int64_t* off_a, const uint8_t* codes_b,
                     const int64_t* off_b) {
    if (!b) return NWG_EINVAL;
    return load_pairs(b, N, codes_a, off_a, codes_b, off_b, true);
}

int nwg_batch_query(nwg_batch* b, int* done) {
    if (!b || !done) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    const hipError_t e = hipStreamQuery(ctx->stream);
    if (e == hipSuccess) {
        *done = 1;
        return NWG_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return NWG_OK;
    }
    return fail(ctx, NWG_EDEVICE, "hipStreamQuery failed: %s", hipGetErrorString(e));
}

int nwg_batch_set_params(nwg_batch* b, const double* matrix, const double* gaps4) {
    return nwg_batch_set_params_k(b, 1, matrix, gaps4);
}

int nwg_batch_set_params_k(nwg_batch* b, int K, const double* matrices, const double* gaps) {
    if (!b || !matrices || !gaps) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (K < 1 || K > 4096) return fail(ctx, NWG_EINVAL, "nwgrad: the number of parameter sets must be in [1, 4096]");
    const int nn = b->na * b->na;
    if (b->tracked) {   // the tracks must stay within the new sets
        for (uint16_t v : b->htrk)
            if (v >= K) return fail(ctx, NWG_EINVAL, "nwgrad: a track names set %d; params has %d sets", (int)v, K);
        for (uint16_t v : b->hgtrk)
            if (v >= K) return fail(ctx, NWG_EINVAL, "nwgrad: a gap track names set %d; params has %d sets", (int)v, K);
    }
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    // Work still queued on the context's stream (an asynchronous pass, a set_params_dev copy)
    // reads these buffers: let it finish, then upload on the same stream and wait for the copies
    // (their host sources are this call's own vectors).
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int oldK = b->K;
    b->matrix.assign(matrices, matrices + (size_t)K * nn);
    b->gaps.assign(gaps, gaps + (size_t)4 * K);
    b->K = K;
    const size_t kn = (size_t)K * nn;
    std::vector<float> sv;
    if (b->prob.precision == 0) {
        sv.resize(kn);
        for (size_t k = 0; k < kn; ++k) sv[k] = static_cast<float>(matrices[k]);
        NWG_HIP(ctx, b->sub_f32.reserve(kn));
        NWG_HIP(ctx, hipMemcpyAsync(b->sub_f32.p, sv.data(), sizeof(float) * kn, hipMemcpyHostToDevice, ctx->stream));
    } else {
        NWG_HIP(ctx, b->sub_f64.reserve(kn));
        NWG_HIP(ctx, hipMemcpyAsync(b->sub_f64.p, matrices, sizeof(double) * kn, hipMemcpyHostToDevice, ctx->stream));
    }
    // the soft pass's parameter vector: K matrices, then K rows of gaps
    NWG_HIP(ctx, b->theta.reserve((size_t)K * (nn + 4)));
    std::vector<double> th(b->matrix);
    th.insert(th.end(), gaps, gaps + 4 * (size_t)K);
    NWG_HIP(ctx, hipMemcpyAsync(b->theta.p, th.data(), sizeof(double) * th.size(), hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (K != oldK && b->prob.grad_mode != 0)
        NWG_HIP(ctx, b->grads.alloc((size_t)std::max<int64_t>(b->N, 1) * kn + (size_t)std::max<int64_t>(b->N, 1) * 4 * K));
    if (K > 1 && !b->tracked) {   // several sets make the problem tracked (default tracks)
        const int rc = set_tracks_host(b, nullptr, nullptr);
        if (rc) return rc;
    }
    b->bx_dirty = true;
    if (set_slots(b) && b->prob.grad_mode != 2) {
        const int rc = replan(b);
        if (rc) return rc;
    }
    b->have_params = true;
    b->results_valid = false;
    b->paths_valid = false;   // guides stay: a banded pass re-aligns around them
    return NWG_OK;
}

int nwg_batch_set_tracks(nwg_batch* b, const uint16_t* tracks, const uint16_t* gap_tracks) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    const int rc = set_tracks_host(b, tracks, gap_tracks);
    if (rc) return rc;
    if (set_slots(b) && b->prob.grad_mode != 2) {
        const int rc2 = replan(b);
        if (rc2) return rc2;
    }
    b->results_valid = false;
    return NWG_OK;
}

int nwg_batch_num_sets(nwg_batch* b, int* K, int* tracked) {
    if (!b) return NWG_EINVAL;
    if (K) *K = b->K;
    if (tracked) *tracked = b->tracked ? 1 : 0;
    return NWG_OK;
}

int nwg_batch_set_params_dev(nwg_batch* b, const double* theta_dev) {
    if (!b || !theta_dev) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->prob.grad_mode != 2)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: set_params_dev serves soft batches (grad_mode=\"soft\") only");
    const size_t th = (size_t)b->K * (b->na * b->na + 4);   // K matrices, then K rows of gaps
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    if (b->theta.n < th) NWG_HIP(ctx, b->theta.reserve(th));
    NWG_HIP(ctx, hipMemcpyAsync(b->theta.p, theta_dev, sizeof(double) * th, hipMemcpyDeviceToDevice, ctx->stream));
    b->have_params = true;
    b->results_valid = false;
    return NWG_OK;
}

int nwg_batch_set_fill(nwg_batch* b, int fill) {
    if (!b || fill < 0 || fill > 1) return NWG_EINVAL;
    b->fill = fill;
    return NWG_OK;
}

int nwg_batch_set_lanes(nwg_batch* b, int columns) {
    if (!b) return NWG_EINVAL;
    if (columns != 0 && columns != kS && columns != kSN)
        return fail(b->ctx, NWG_EINVAL, "nwgrad_hip: columns per lane must be 0 (chosen per batch), %d or %d", kSN, kS);
    b->lanes_S = columns;
    if (b->prob.grad_mode != 2) {
        NWG_HIP(b->ctx, hipSetDevice(b->ctx->device));
        const int rc = replan(b);
        if (rc) return rc;
    }
    return NWG_OK;
}

int nwg_batch_lanes(nwg_batch* b, int* columns) {
    if (!b || !columns) return NWG_EINVAL;
    *columns = b->prob.grad_mode == 2 ? kSS : b->plan.S;
    return NWG_OK;
}

int nwg_set_lane_table(int n_rules, const int32_t* rules) {
    if (n_rules < 0 || (n_rules > 0 && !rules)) return NWG_EINVAL;
    std::vector<LaneRule> v;
    for (int k = 0; k < n_rules; ++k) {
        const LaneRule r{rules[3 * k], rules[3 * k + 1], rules[3 * k + 2]};
        if (r.S != kS && r.S != kSN) return NWG_EINVAL;
        v.push_back(r);
    }
    std::lock_guard<std::mutex> lk(g_lane_mu);
    g_lane_rules.swap(v);
    return NWG_OK;
}

int nwg_batch_set_soft(nwg_batch* b, double temperature, int precision) {
    if (!b) return NWG_EINVAL;
    if (!(temperature > 0.0)) return fail(b->ctx, NWG_EINVAL, "nwgrad_hip: soft temperature must be > 0");
    if (precision < 0 || precision > 1) return fail(b->ctx, NWG_EINVAL, "nwgrad_hip: soft precision is 0 or 1");
    if (precision == 0 && b->prob.precision == 1)
        return fail(b->ctx, NWG_EINVAL,
                    "nwgrad: soft_precision=\"float32\" needs a float32 SeqPair / SeqPairBatch (the *Double "
                    "classes run double)");
    b->temperature = temperature;
    b->soft_precision = precision;
    b->results_valid = false;
    return NWG_OK;
}

}  // extern "C"

namespace {

// ── hard and score-only fills ───────────────────────────────────────────────
template <typename T> struct TrackBufs;
template <> struct TrackBufs<float> {
    static const float* bx(nwg_batch* b) { return b->bx32.p; }
    static const float* gt(nwg_batch* b) { return b->gt32.p; }
};
template <> struct TrackBufs<double> {
    static const double* bx(nwg_batch* b) { return b->bx64.p; }
    static const double* gt(nwg_batch* b) { return b->gt64.p; }
};
template <typename T>
nwg::TrackArgs<T> track_args(nwg_batch* b) {
    nwg::TrackArgs<T> t{};
    if (b->tracked) {
        t.trk = b->dtrk.p;
        t.gtrk = b->dgtrk.p;
        t.bx0 = TrackBufs<T>::bx(b);
        t.gaps = TrackBufs<T>::gt(b);
        t.K = b->K;
    }
    return t;
}

// The hard / score-only fill kernel of a problem type and lane width (kS or kSN columns).
template <typename T>
void (*hard_kernel_for(int gm, int am, bool hard, bool tr, int S, bool band = false))(nwg::HardArgs<T>) {
    auto pick = [&](auto gmc, auto amc, auto hardc, auto trc) -> void (*)(nwg::HardArgs<T>) {
        constexpr int G = decltype(gmc)::value, AM = decltype(amc)::value;
        constexpr bool H = decltype(hardc)::value, TRC = decltype(trc)::value;
        if constexpr (H) {   // banded passes are hard passes
            if (band)
                return S == kSN ? nwg::hard_fill_kernel<T, G, AM, kSN, H, TRC, true>
                                : nwg::hard_fill_kernel<T, G, AM, kS, H, TRC, true>;
        }
        return S == kSN ? nwg::hard_fill_kernel<T, G, AM, kSN, H, TRC> : nwg::hard_fill_kernel<T, G, AM, kS, H, TRC>;
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using BT = std::true_type;
    using BF = std::false_type;
    void (*kern)(nwg::HardArgs<T>) = nullptr;
#define NWG_PICK(GMV, AMV)                                                                      \
    if (gm == GMV::value && am == AMV::value)                                                   \
        kern = hard ? (tr ? pick(GMV{}, AMV{}, BT{}, BT{}) : pick(GMV{}, AMV{}, BT{}, BF{}))     \
                    : (tr ? pick(GMV{}, AMV{}, BF{}, BT{}) : pick(GMV{}, AMV{}, BF{}, BF{}));
    NWG_PICK(I0, I0) NWG_PICK(I0, I1) NWG_PICK(I0, I2) NWG_PICK(I1, I0) NWG_PICK(I1, I1) NWG_PICK(I1, I2)
#undef NWG_PICK
    return kern;
}

template <typename T>
int launch_hard(nwg_batch* b, const Plan& pl) {
    nwg_ctx* ctx = b->ctx;
    const bool hard = b->prob.grad_mode == 1, tr = b->tracked;
    const size_t lds = (size_t)nwg::hard_lds_sub_bytes<T>(b->na, b->K, tr) +
                       (hard ? (size_t)nwg::hard_lds_hist_bytes(b->na, b->K, tr, b->hist_slots) : 0);
    if (lds > 65536)
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: %d parameter sets of %d x %d do not fit the fill's local memory",
                    b->K, b->na, b->na);
    void (*kern)(nwg::HardArgs<T>) = hard_kernel_for<T>(b->prob.gap_model, b->prob.mode, hard, tr, pl.S, b->banded);
    if (!kern) return fail(ctx, NWG_EINVAL, "nwgrad_hip: unsupported problem type");
    // fill 0: one plain launch over the persistent order when a slot per tile fits the table
    // budget, else one persistent launch with as many slots (waves) as fit it, at most the
    // resident blocks; fill 1: the per-bucket launches
    const int64_t slot_dwords = pl.max_steps * (pl.S / 4) * 64;
    const bool persistent = b->fill == 0 && hard_persistent(b, pl);
    int64_t grid = 0;
    if (persistent)
        grid = std::max<int64_t>(1, std::min<int64_t>({(int64_t)pl.tiles.size(),
                                                       resident_blocks(ctx, kern, lds, kBlocksPerCU),
                                                       b->table_budget / (slot_dwords * 4)}));
    const int64_t table_dwords = b->fill != 0 ? pl.bucket_table_dwords
                                              : (persistent ? grid : (int64_t)pl.tiles.size()) * slot_dwords;
    if (hard && b->table.n < (size_t)table_dwords) {
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NWG_HIP(ctx, b->table.alloc((size_t)table_dwords));
    }
    nwg::HardArgs<T> A{};
    A.codes_a = b->ca.p;
    A.codes_b = b->cb.p;
    A.off_a = b->oa.p;
    A.off_b = b->ob.p;
    A.order = pl.d_order.p;
    A.sub = std::is_same<T, float>::value ? reinterpret_cast<const T*>(b->sub_f32.p)
                                           : reinterpret_cast<const T*>(b->sub_f64.p);
    A.na = b->na;
    A.hist_slots = b->hist_slots;
    A.go_a = static_cast<T>(b->gaps[0]);
    A.ge_a = static_cast<T>(b->gaps[1]);
    A.go_b = static_cast<T>(b->gaps[2]);
    A.ge_b = static_cast<T>(b->gaps[3]);
    A.tr = track_args<T>(b);
    A.scores = b->scores.p;
    A.grads = b->grads.p;
    A.table = b->table.p;
    if (hard) {
        A.guide = b->guide.p;
        A.band = b->band;
        if (b->keep_paths) {
            A.ops = b->ops.p;
            A.ops_n = b->ops_n.p;
            A.ends = b->ends.p;
        }
    }
    b->hard_grid = grid;
    if (b->fill == 0) {
        A.tiles = pl.d_ptiles.p;   // most row steps first
        A.tile_base = 0;
        A.n_tiles = (int32_t)pl.tiles.size();
        A.slot_dwords = slot_dwords;
        if (persistent) {
            hipLaunchKernelGGL(nwg::clear_words_kernel, dim3(1), dim3(256), 0, ctx->stream, b->queue.p, (int64_t)1);
            NWG_HIP(ctx, hipGetLastError());
            A.queue = b->queue.p;
        }
        hipLaunchKernelGGL(kern, dim3((unsigned)(persistent ? grid : (int64_t)pl.tiles.size())), dim3(64), lds,
                           ctx->stream, A);
        NWG_HIP(ctx, hipGetLastError());
        return NWG_OK;
    }
    A.tiles = pl.d_tiles.p;
    for (const HardLaunch& L : pl.bucket_launches) {
        A.tile_base = L.begin;
        A.n_tiles = L.end;
        A.slot_dwords = L.slot_dwords;
        hipLaunchKernelGGL(kern, dim3((unsigned)(L.end - L.begin)), dim3(64), lds, ctx->stream, A);
        NWG_HIP(ctx, hipGetLastError());
    }
    return NWG_OK;
}

// ── long pairs: column panels ───────────────────────────────────────────────
template <typename T> struct LongBnd;
template <> struct LongBnd<float> { static DevPtr<float>& get(nwg_batch* b) { return b->lbnd32; } };
template <> struct LongBnd<double> { static DevPtr<double>& get(nwg_batch* b) { return b->lbnd64; } };

template <typename T>
int launch_long(nwg_batch* b) {
    nwg_ctx* ctx = b->ctx;
    LongPlan& L = b->lpl;
    const bool hard = b->prob.grad_mode == 1, tr = b->tracked;
    const int gm = b->prob.gap_model, am = b->prob.mode;
    using FillK = void (*)(nwg::LongArgs<T>);
    FillK fill = nullptr, walk = nullptr;
    const bool band = b->banded;
    constexpr int kLS = long_cols(sizeof(T) == 8);
    auto pick = [&](auto gmc, auto amc, auto hardc) {
        constexpr int G = decltype(gmc)::value, AM = decltype(amc)::value;
        constexpr bool H = decltype(hardc)::value;
        if (tr) {
            fill = nwg::long_fill_kernel<T, G, AM, kLS, H, kRowBlock, true>;
            walk = nwg::long_walk_kernel<T, G, AM, kLS, H, true>;
        } else {
            fill = nwg::long_fill_kernel<T, G, AM, kLS, H, kRowBlock, false>;
            walk = nwg::long_walk_kernel<T, G, AM, kLS, H, false>;
        }
        if constexpr (H) {   // banded passes are hard passes
            if (band)
                fill = tr ? nwg::long_fill_kernel<T, G, AM, kLS, H, kRowBlock, true, true>
                          : nwg::long_fill_kernel<T, G, AM, kLS, H, kRowBlock, false, true>;
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using BT = std::true_type;
    using BF = std::false_type;
#define NWG_PICK(GMV, AMV) \
    if (gm == GMV::value && am == AMV::value) { if (hard) pick(GMV{}, AMV{}, BT{}); else pick(GMV{}, AMV{}, BF{}); }
    NWG_PICK(I0, I0) NWG_PICK(I0, I1) NWG_PICK(I0, I2) NWG_PICK(I1, I0) NWG_PICK(I1, I1) NWG_PICK(I1, I2)
#undef NWG_PICK
    if (!fill) return fail(ctx, NWG_EINVAL, "nwgrad_hip: unsupported problem type");
    const size_t lds_fill = (size_t)nwg::hard_lds_sub_bytes<T>(b->na, b->K, tr) + 256;   // + the A-code ring
    const size_t lds_walk = hard ? (size_t)nwg::long_walk_lds_bytes(b->na, b->K, tr, kLS)
                                 : 16;
    if (lds_fill > 65536 || lds_walk > 65536)
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: %d parameter sets of %d x %d do not fit the long-pair kernels' "
                    "local memory", b->K, b->na, b->na);
    DevPtr<T>& bnd = LongBnd<T>::get(b);
    const bool grow = b->lprog.n < (size_t)L.max_items || b->lcand.n < (size_t)L.max_items * sizeof(nwg::CandRec<T>) ||
                      bnd.n < (size_t)std::max<int64_t>(L.max_bnd, 1) ||
                      (hard && b->ltable.n < (size_t)std::max<int64_t>(L.max_table, 1)) || b->lerr.n < 1;
    if (grow) {
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NWG_HIP(ctx, b->lprog.reserve((size_t)L.max_items));
        NWG_HIP(ctx, b->lcand.reserve((size_t)L.max_items * sizeof(nwg::CandRec<T>)));
        NWG_HIP(ctx, bnd.reserve((size_t)std::max<int64_t>(L.max_bnd, 1)));
        if (hard) NWG_HIP(ctx, b->ltable.reserve((size_t)std::max<int64_t>(L.max_table, 1)));
        if (b->lerr.n < 1) NWG_HIP(ctx, b->lerr.alloc(1));
    }
    // A launch starts with a clear error word: a word left set would make every wait of this
    // launch give up at once.  Cleared by a kernel, so a captured pass holds no memset node.
    hipLaunchKernelGGL(nwg::clear_words_kernel, dim3(1), dim3(256), 0, ctx->stream, b->lerr.p, (int64_t)1);
    NWG_HIP(ctx, hipGetLastError());
    nwg::LongArgs<T> A{};
    A.codes_a = b->ca.p;
    A.codes_b = b->cb.p;
    A.off_a = b->oa.p;
    A.off_b = b->ob.p;
    A.sub = std::is_same<T, float>::value ? reinterpret_cast<const T*>(b->sub_f32.p)
                                           : reinterpret_cast<const T*>(b->sub_f64.p);
    A.na = b->na;
    A.go_a = static_cast<T>(b->gaps[0]);
    A.ge_a = static_cast<T>(b->gaps[1]);
    A.go_b = static_cast<T>(b->gaps[2]);
    A.ge_b = static_cast<T>(b->gaps[3]);
    A.tr = track_args<T>(b);
    A.progress = b->lprog.p;
    A.bnd = bnd.p;
    A.cand = reinterpret_cast<nwg::CandRec<T>*>(b->lcand.p);
    A.table = hard ? b->ltable.p : nullptr;
    A.scores = b->scores.p;
    A.grads = b->grads.p;
    A.err = b->lerr.p;
    A.wait_ticks = kWaitTicks;
    A.queue = b->queue.p + 1;
    if (hard) {
        A.guide = b->guide.p;
        A.band = b->band;
        if (b->keep_paths) {
            A.ops = b->ops.p;
            A.ops_n = b->ops_n.p;
            A.ends = b->ends.p;
        }
    }
    const int64_t resident = resident_blocks(ctx, fill, lds_fill, kBlocksPerCU);
    for (const LongGroup& gr : L.groups) {
        const int items = gr.item_end - gr.item_begin, pairs = gr.pair_end - gr.pair_begin;
        A.items = L.d_items.p + gr.item_begin;
        A.n_items = items;
        A.lpairs = L.d_lpairs.p + gr.pair_begin;
        hipLaunchKernelGGL(nwg::clear_words_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, ctx->stream,
                           b->lprog.p, (int64_t)items);
        NWG_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(nwg::clear_words_kernel, dim3(1), dim3(256), 0, ctx->stream, A.queue, (int64_t)1);
        NWG_HIP(ctx, hipGetLastError());
        const int64_t grid = std::max<int64_t>(std::min<int64_t>(items, resident), 1);
        hipLaunchKernelGGL(fill, dim3((unsigned)grid), dim3(64), lds_fill, ctx->stream, A);
        NWG_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(walk, dim3((unsigned)pairs), dim3(nwg::kWalkThreads), lds_walk, ctx->stream, A);
        NWG_HIP(ctx, hipGetLastError());
    }
    return NWG_OK;
}

// A failed long-pair fill leaves bits in the error word (1: a wait passed its bound, 2: a walk
// found no valid end cell): report and clear them.
int check_long_err(nwg_batch* b) {
    if (b->lpl.items.empty() || b->lerr.n < 1) return NWG_OK;
    nwg_ctx* ctx = b->ctx;
    int32_t e = 0;
    NWG_HIP(ctx, hipMemcpyAsync(&e, b->lerr.p, sizeof e, hipMemcpyDeviceToHost, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (e == 0) return NWG_OK;
    NWG_HIP(ctx, hipMemsetAsync(b->lerr.p, 0, sizeof(int32_t), ctx->stream));
    b->results_valid = false;
    b->guides_valid = b->paths_valid = false;
    return fail(ctx, NWG_EDEVICE,
                "nwgrad_hip: the long-pair fill failed (error word %d: 1 = a column panel waited past its bound for "
                "its left neighbour, 2 = a pair had no valid end cell); the results of this call are not valid",
                (int)e);
}

// ── soft pass ──────────────────────────────────────────────────────────────
template <typename SF> struct SoftScratch;
template <> struct SoftScratch<float> {
    static DevPtr<float>& gw(nwg_batch* b) { return b->gw32; }
    static DevPtr<float>& seg(nwg_batch* b) { return b->seg32; }
    static DevPtr<float>& lseg(nwg_batch* b) { return b->lseg32; }
    static DevPtr<float>& pf(nwg_batch* b) { return b->pf32; }
    static DevPtr<float>& pb(nwg_batch* b) { return b->pb32; }
};
template <> struct SoftScratch<double> {
    static DevPtr<double>& gw(nwg_batch* b) { return b->gw64; }
    static DevPtr<double>& seg(nwg_batch* b) { return b->seg64; }
    static DevPtr<double>& lseg(nwg_batch* b) { return b->lseg64; }
    static DevPtr<double>& pf(nwg_batch* b) { return b->pf64; }
    static DevPtr<double>& pb(nwg_batch* b) { return b->pb64; }
};

using nwg::SoftTile;

template <typename SF>
using SoftKernel = void (*)(nwg::SoftArgs<SF>);

template <typename SF>
SoftKernel<SF> soft_kernel_for(int mode, bool tracked, bool band) {
    if (band)   // banded passes: untracked batches (the caller refuses tracked ones)
        return mode == 1 ? nwg::soft_kernel<SF, nwg::kLocal, kSS, false, true>
                         : mode == 2 ? nwg::soft_kernel<SF, nwg::kGlocal, kSS, false, true>
                                     : nwg::soft_kernel<SF, nwg::kGlobal, kSS, false, true>;
    if (tracked)
        return mode == 1 ? nwg::soft_kernel<SF, nwg::kLocal, kSS, true>
                         : mode == 2 ? nwg::soft_kernel<SF, nwg::kGlocal, kSS, true>
                                     : nwg::soft_kernel<SF, nwg::kGlobal, kSS, true>;
    return mode == 1 ? nwg::soft_kernel<SF, nwg::kLocal, kSS, false>
                     : mode == 2 ? nwg::soft_kernel<SF, nwg::kGlocal, kSS, false>
                                 : nwg::soft_kernel<SF, nwg::kGlobal, kSS, false>;
}

// The wide pool's queue and regions: tiles 0 .. Tq-1 of the persistent order (longest first),
// the first Tp of them pipelined.  Tile t takes the next panels x (fsteps + 1) rows of the
// circular pool [base, base + pool), from its start again when they do not fit, and waits
// for the tiles that hold rows of its region to be done.  A pipelined tile's forward items
// (panels 0 .. NP-1) come at the tile's turn, its backward items (panels NP-1 .. 0) `lag`
// tiles later; another tile is one whole item.  Every item comes after the items it waits
// for: the tile's earlier forward panels, its forward items and later backward panels, and
// the items that finish the tiles whose regions it reuses (a pipelined tile's backward items
// are queued early for that when the lag would place them later).
int build_soft_items(nwg_batch* b, SoftPlan& wl, int64_t Tp, int64_t Tq, int64_t base, int64_t pool, int64_t lag) {
    nwg_ctx* ctx = b->ctx;
    std::vector<nwg::SoftTile>& pt = wl.ptiles;
    wl.items.clear();
    wl.waits.clear();
    std::map<int64_t, std::pair<int64_t, int32_t>> held;   // region start -> (end, tile)
    std::deque<int32_t> pending;                          // pipelined tiles, backward not queued
    auto queue_backward = [&](int32_t j) {
        for (int pp = pt[(size_t)j].panels - 1; pp >= 0; --pp) wl.items.push_back(int2{j, 2 * pp + 1});
    };
    int64_t pos = 0;
    int32_t poff = 0;
    std::vector<int32_t> w;
    for (int64_t t = 0; t < Tq; ++t) {
        nwg::SoftTile& tt = pt[(size_t)t];
        const int64_t size = (int64_t)tt.panels * (tt.fsteps + 1);
        if (pos + size > pool) pos = 0;
        const int64_t a = pos, e = pos + size;
        pos = e;
        w.clear();
        auto it = held.lower_bound(a);
        if (it != held.begin() && std::prev(it)->second.first > a) it = std::prev(it);
        while (it != held.end() && it->first < e) {
            const int64_t s0 = it->first, s1 = it->second.first;
            const int32_t j = it->second.second;
            w.push_back(j);
            it = held.erase(it);
            if (s0 < a) held[s0] = {a, j};
            if (s1 > e) {
                held[e] = {s1, j};
                break;
            }
        }
        held[a] = {e, (int32_t)t};
        std::sort(w.begin(), w.end());
        w.erase(std::unique(w.begin(), w.end()), w.end());
        tt.rbase = base + a;
        tt.wait0 = (int32_t)wl.waits.size();
        tt.nwait = (int32_t)w.size();
        wl.waits.insert(wl.waits.end(), w.begin(), w.end());
        const int32_t last_wait = w.empty() ? -1 : w.back();
        while (!pending.empty() && pending.front() <= last_wait) {
            queue_backward(pending.front());
            pending.pop_front();
        }
        if (t < Tp) {
            tt.poff = poff;
            poff += tt.panels;
            for (int pp = 0; pp < tt.panels; ++pp) wl.items.push_back(int2{(int)t, 2 * pp});
            pending.push_back((int32_t)t);
        } else {
            tt.poff = 0;
            wl.items.push_back(int2{(int)t, -1});
        }
        while (!pending.empty() && pending.front() <= t - lag) {
            queue_backward(pending.front());
            pending.pop_front();
        }
    }
    while (!pending.empty()) {
        queue_backward(pending.front());
        pending.pop_front();
    }
    NWG_HIP(ctx, wl.d_items.reserve(std::max<size_t>(wl.items.size(), 1)));
    NWG_HIP(ctx, wl.d_waits.reserve(std::max<size_t>(wl.waits.size(), 1)));
    NWG_HIP(ctx, hipMemcpyAsync(wl.d_items.p, wl.items.data(), sizeof(int2) * wl.items.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    if (!wl.waits.empty())
        NWG_HIP(ctx, hipMemcpyAsync(wl.d_waits.p, wl.waits.data(), sizeof(int32_t) * wl.waits.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipMemcpyAsync(wl.d_ptiles.p, pt.data(), sizeof(nwg::SoftTile) * pt.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wl.items_tiles = Tp;
    wl.items_pooled = Tq;
    wl.items_base = base;
    wl.items_pool = pool;
    wl.items_lag = lag;
    wl.pipelined_panels = poff;
    return NWG_OK;
}

// One soft pass over the tiles of `pl` (one panel) and `wl` (column panels).  Persistent:
// one launch.  A wide tile whose panels in sequence would take one wave longer than the
// launch's work per resident wave is pipelined: its panel items spread it over as many waves
// as it has panels (the longest tiles, so a prefix of the longest-first order).  The other
// wide tiles run whole, one wave each, in the block's own slot, with no waits.  Blocks below
// n_pwaves first drain the pipelined items, blocks below n_lslots the whole wide tiles, then
// blocks below n_mslots the main queue.  Plain: one launch per bucket, main plan then wide
// plan (whole tiles).
template <typename SF>
int launch_soft(nwg_batch* b, const SoftPlan& pl, SoftPlan& wl) {
    using SS = SoftScratch<SF>;
    using LY = nwg::Seg<SF, kSS>;
    nwg_ctx* ctx = b->ctx;
    const int nn = b->na * b->na;
    const bool lin = b->prob.gap_model == 0;
    if (pl.tiles.empty() && wl.tiles.empty()) return NWG_OK;
    DevPtr<SF>& dps = SS::pf(b);
    const bool tr = b->tracked;
    const int K = tr ? b->K : 1;
    NWG_HIP(ctx, dps.reserve((size_t)K * nn));
    NWG_HIP(ctx, SS::gw(b).reserve((size_t)4 * K));
    const bool band = b->banded;
    SoftKernel<SF> kern = soft_kernel_for<SF>(b->prob.mode, tr, band);
    const size_t lds = nwg::soft_lds_bytes<SF>(b->na, K, b->hist_slots);
    if (lds > kSoftLdsMax)
        return fail(ctx, NWG_EINVAL,
                    "nwgrad_amd: the soft pass with %d parameter sets over a %d-letter alphabet needs %zu bytes of LDS "
                    "per wavefront (at most %zu)", K, b->na, lds, kSoftLdsMax);
    // resident blocks per CU: queried once per kernel and LDS size (no query inside a captured step)
    int& occ = b->soft_occ_cache[(std::is_same<SF, float>::value ? 0 : 1) + (tr ? 2 : 0) + (band ? 4 : 0)];
    if (occ <= 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64, lds) != hipSuccess || occ <= 0))
        occ = kSoftBlocksPerCU;
    b->soft_occ = occ;
    const int64_t resident = (int64_t)occ * ctx->n_cu;
    const bool persistent = b->fill == 0;
    // Slots: the main pool holds one panel region of pl.rows rows per slot (one per block
    // that runs main tiles); the wide pool one slot of panels x rows per block that runs whole
    // wide tiles (the largest whole tile's) and, persistent, after them the circular pool of
    // the pipelined tiles (or of every wide tile), each taking its own panels x (fsteps + 1)
    // rows.
    const size_t row_bytes = (size_t)LY::row * sizeof(SF);
    const size_t mslot = (size_t)pl.rows * row_bytes;
    const size_t budget = b->soft_scratch;
    const int64_t wrows_cap = (int64_t)(budget / 2 / row_bytes);   // the wide pool: half the budget
    const int64_t T = (int64_t)wl.tiles.size();
    int64_t Tp = 0, Tq = 0, pool = 0, lslots = 0, pwaves = 0, lag = 0;
    int64_t wpanels = wl.panels, wrows = wl.rows;   // a whole tile's block slot
    if (T > 0 && persistent) {
        // Pipelined: the tiles longer than the launch's work per resident wave.
        const double share = (double)(pl.work + wl.work) / (double)resident;
        while (Tp < T && (double)wl.pwork[(size_t)Tp] > share) ++Tp;
        auto size_of = [&](int64_t t) {
            const nwg::SoftTile& tt = wl.ptiles[(size_t)t];
            return (int64_t)tt.panels * (tt.fsteps + 1);
        };
        int64_t need_p = 0, big_p = 0, need = 0, big = 0, work_w = 0;
        wpanels = 1;
        wrows = 1;
        for (int64_t t = 0; t < T; ++t) {
            const int64_t size = size_of(t);
            need += size;
            big = std::max(big, size);
            if (t < Tp) {
                need_p += size;
                big_p = std::max(big_p, size);
            } else {
                wpanels = std::max<int64_t>(wpanels, wl.ptiles[(size_t)t].panels);
                wrows = std::max<int64_t>(wrows, (int64_t)wl.ptiles[(size_t)t].fsteps + 1);
                work_w += wl.pwork[(size_t)t];
            }
        }
        // Whole tiles run in block slots (no waits, no cache writeback), beside the pipelined
        // tiles' pool (all of it, or at least half the wide budget), when the slots that fit
        // get through the whole tiles' work within the launch's work per resident wave (the
        // other blocks run main tiles meanwhile).  Otherwise -- slots sized for the largest
        // whole tile hold too few of them -- every wide tile goes to the pool, whose regions
        // fit each tile's own size.
        const int64_t wslot = wpanels * wrows;
        const int64_t keep_p = Tp > 0 ? std::max(big_p, std::min(need_p, wrows_cap / 2)) : 0;
        const int64_t n_b = std::min<int64_t>(std::min<int64_t>(T - Tp, resident),
                                              std::max<int64_t>(0, wrows_cap - keep_p) / wslot);
        if (Tp == T || (n_b >= 1 && (double)work_w / (double)n_b <= share)) {
            lslots = Tp < T ? n_b : 0;
            Tq = Tp;
            pool = Tp > 0 ? std::max(big_p, std::min(need_p, wrows_cap - lslots * wslot)) : 0;
        } else {
            lslots = 0;
            Tq = T;
            need_p = need;
            big_p = big;
            pool = std::max(big, std::min(need, wrows_cap));
        }
        if (Tq > 0) {
            // Blocks that take pool items: the tiles that fit in the pool at once, times their
            // waves (a pipelined tile's forward and backward items both); a pipelined tile's
            // backward items come half of that many tiles after its forward ones.
            double waves = 0.0;
            for (int64_t t = 0; t < Tq; ++t) waves += t < Tp ? 2.0 * wl.ptiles[(size_t)t].panels : 1.0;
            const double fit = std::min((double)Tq, (double)pool * (double)Tq / (double)need_p);
            pwaves = std::max<int64_t>(1, std::min<int64_t>(resident, (int64_t)std::ceil(fit * waves / (double)Tq)));
            lag = std::max<int64_t>(0, (int64_t)std::ceil(fit / 2.0));
            const int64_t base = lslots * wslot;
            if (wl.items_tiles != Tp || wl.items_pooled != Tq || wl.items_base != base || wl.items_pool != pool ||
                wl.items_lag != lag) {
                const int rc = build_soft_items(b, wl, Tp, Tq, base, pool, lag);
                if (rc) return rc;
            }
        }
    } else if (T > 0) {
        const int64_t wslot = wpanels * wrows;
        lslots = std::max<int64_t>(1, std::min<int64_t>({T, resident, wrows_cap / wslot}));
    }
    const size_t wide_rows = (size_t)(lslots * wpanels * wrows + pool);
    int64_t mslots = pl.tiles.empty() ? 0 : std::min<int64_t>((int64_t)pl.tiles.size(), resident);
    if (mslots > 0) {
        const size_t used = wide_rows * row_bytes;
        const size_t left = budget > used ? budget - used : 0;
        mslots = std::max<int64_t>(1, std::min<int64_t>(mslots, (int64_t)(left / std::max<size_t>(mslot, 1))));
    }
    const bool piped = Tq > 0;
    const size_t mseg = (size_t)mslots * pl.rows * LY::row;
    const size_t lseg = wide_rows * LY::row;
    const size_t bnd = wide_rows * nwg::kTilePairs;
    const size_t zpart = piped ? (size_t)wl.pipelined_panels * nwg::kTilePairs * 4 : 0;
    const size_t part = piped ? (size_t)wl.pipelined_panels * b->hist_slots * ((size_t)K * (nn + 4) + 2) : 0;
    const size_t nsync = piped ? (size_t)wl.sync_words : 0;
    if (SS::seg(b).n < mseg || SS::lseg(b).n < lseg || SS::pb(b).n < bnd * 6 || b->pexp.n < bnd * 2 ||
        b->lzpart.n < zpart || b->lpart.n < part || b->lsync.n < nsync) {
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NWG_HIP(ctx, SS::seg(b).reserve(std::max<size_t>(mseg, 1)));
        NWG_HIP(ctx, SS::lseg(b).reserve(std::max<size_t>(lseg, 1)));
        NWG_HIP(ctx, SS::pb(b).reserve(std::max<size_t>(bnd * 6, 1)));
        NWG_HIP(ctx, b->pexp.reserve(std::max<size_t>(bnd * 2, 1)));
        NWG_HIP(ctx, b->lzpart.reserve(std::max<size_t>(zpart, 1)));
        NWG_HIP(ctx, b->lpart.reserve(std::max<size_t>(part, 1)));
        NWG_HIP(ctx, b->lsync.reserve(std::max<size_t>(nsync, 1)));
    }
    NWG_HIP(ctx, b->soft_ok.reserve((size_t)std::max<int64_t>(b->N, 1)));
    NWG_HIP(ctx, b->lqueue.reserve(2));
    nwg::SoftArgs<SF> A{};
    A.codes_a = b->ca.p;
    A.codes_b = b->cb.p;
    A.off_a = b->oa.p;
    A.off_b = b->ob.p;
    A.view = nwg::SoftView<SF>{SS::seg(b).p, SS::pb(b).p, b->pexp.p, pl.rows, 1};
    A.lview = nwg::SoftView<SF>{SS::lseg(b).p, SS::pb(b).p, b->pexp.p, wrows, (int32_t)wpanels};
    A.psub = dps.p;
    A.na = b->na;
    A.hist_slots = b->hist_slots;
    A.gapw = SS::gw(b).p;
    if (tr) {
        A.trk = b->dtrk.p;
        A.gtrk = b->dgtrk.p;
        A.K = K;
    }
    A.linear = lin ? 1 : 0;
    A.temperature = b->temperature;
    A.scores = b->scores.p;
    A.grads = b->grads.p;
    A.ok = b->soft_ok.p;
    A.tol = std::is_same<SF, float>::value ? 0x1p-20 : 0x1p-45;   // the reference's SoftRange<F>::tol
    A.vmin = (double)std::numeric_limits<SF>::min();
    A.vtiny = (double)std::numeric_limits<SF>::denorm_min();
    A.guide = band ? b->guide.p : nullptr;
    A.band = band ? b->band : 0;
    // the weights (and, persistent, the queue counters) from the device params
    hipLaunchKernelGGL(nwg::soft_weights_kernel<SF>, dim3(1), dim3(256), 0, ctx->stream, b->theta.p, nn, K,
                       1.0 / b->temperature, lin ? 1 : 0, dps.p, SS::gw(b).p, persistent ? b->queue.p : nullptr,
                       persistent ? b->lqueue.p : nullptr, piped ? b->lsync.p : nullptr, (int64_t)nsync);
    NWG_HIP(ctx, hipGetLastError());
    if (persistent) {
        const int64_t grid = std::max<int64_t>({mslots, lslots, pwaves});
        A.order = pl.d_order.p;
        A.tiles = pl.d_ptiles.p;
        A.n_tiles = (int32_t)pl.tiles.size();
        A.queue = b->queue.p;
        A.n_mslots = (int32_t)mslots;
        A.lorder = wl.d_order.p;
        A.ltiles = wl.d_ptiles.p;
        A.n_ltiles = (int32_t)T;
        A.n_ptiles = (int32_t)Tq;
        A.n_lslots = (int32_t)lslots;
        A.n_pwaves = (int32_t)pwaves;
        A.lqueue = b->lqueue.p;
        if (piped) {
            A.litems = wl.d_items.p;
            A.n_litems = (int32_t)wl.items.size();
            A.lwait = wl.d_waits.p;
            A.lsync = b->lsync.p;
            A.lzpart = b->lzpart.p;
            A.lpart = b->lpart.p;
            A.wait_ticks = kWaitTicks;
            b->soft_piped = true;
        }
        b->soft_grid = grid;
        b->soft_lslots = std::max(lslots, pwaves);
        b->soft_pool = pool;
        b->soft_lag = lag;
        b->soft_ptiles = Tp;
        b->soft_qtiles = Tq;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), lds, ctx->stream, A);
        NWG_HIP(ctx, hipGetLastError());
        return NWG_OK;
    }
    // plain: one launch per bucket, at most `slots` tiles per launch
    A.queue = nullptr;
    A.lqueue = nullptr;
    b->soft_grid = 0;
    b->soft_lslots = 0;
    b->soft_pool = 0;
    b->soft_lag = 0;
    b->soft_ptiles = 0;
    b->soft_qtiles = 0;
    for (int w = 0; w < 2; ++w) {
        const SoftPlan& P = w == 0 ? pl : wl;
        const int64_t slots = w == 0 ? mslots : lslots;
        if (P.tiles.empty()) continue;
        nwg::SoftArgs<SF> B = A;
        B.order = P.d_order.p;
        B.tiles = P.d_tiles.p;
        if (w == 1) B.view = A.lview;
        for (const Bucket& bk : P.buckets) {
            for (int t0 = bk.tile_begin; t0 < bk.tile_end; t0 += (int)slots) {
                const int t1 = std::min<int>(bk.tile_end, t0 + (int)slots);
                B.tile_base = t0;
                B.n_tiles = t1;
                hipLaunchKernelGGL(kern, dim3(t1 - t0), dim3(64), lds, ctx->stream, B);
                NWG_HIP(ctx, hipGetLastError());
            }
        }
    }
    return NWG_OK;
}

// A batch's own soft pass: float32 or double, packed tiles (persistent) or per-bucket tiles
// (plain), planned on first use.
int launch_soft_batch(nwg_batch* b, bool record) {
    const bool packed = b->fill == 0;
    SoftPlan& mp = packed ? b->sp_main : b->sq_main;
    SoftPlan& wp = packed ? b->sp_wide : b->sq_wide;
    if (!mp.built) {
        std::vector<int32_t> all((size_t)b->N);
        std::iota(all.begin(), all.end(), 0);
        const int rc = build_soft_plans(b, all, mp, wp, packed);
        if (rc) return rc;
    }
    const int rc = (b->prob.precision == 0 && b->soft_precision == 0) ? launch_soft<float>(b, mp, wp)
                                                                       : launch_soft<double>(b, mp, wp);
    if (rc || b->N == 0 || !record) return rc;
    // an asynchronous pass's refusals into the batch's record (a captured step records them too)
    nwg_ctx* ctx = b->ctx;
    hipLaunchKernelGGL(nwg::soft_refusal_step_kernel, dim3(1), dim3(64), 0, ctx->stream, b->srec.p);
    NWG_HIP(ctx, hipGetLastError());
    const int64_t blocks = std::min<int64_t>((b->N + 255) / 256, 1024);
    hipLaunchKernelGGL(nwg::soft_refusal_scan_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       b->soft_ok.p, b->N, b->srec.p);
    NWG_HIP(ctx, hipGetLastError());
    return NWG_OK;
}

// After a soft pass: a pair whose float32 pass lost range or mass beyond 2^-20 reruns in
// double, and one that fails in double too fails the call (the reference's policy,
// aligner.hpp:6554-6586: float32, then double, then std::domain_error under "scaled").
// The pipelined wide tiles' error word, after a synchronisation: a wait that passed its bound
// (the results of that launch are not to be used) fails the call.
int check_soft_err(nwg_batch* b, const int32_t& word) {
    if (!b->soft_piped) return NWG_OK;
    b->soft_piped = false;
    if (word == 0) return NWG_OK;
    b->results_valid = false;
    return fail(b->ctx, NWG_EDEVICE,
                "nwgrad_hip: soft pass: a wide pair's panel waited past its bound for another panel (error word %d)",
                (int)word);
}

int soft_finish(nwg_batch* b) {
    nwg_ctx* ctx = b->ctx;
    b->soft_reruns = 0;
    if (b->N == 0) return NWG_OK;
    b->host_ok.resize((size_t)b->N);
    int32_t serr = 0;
    NWG_HIP(ctx, hipMemcpyAsync(b->host_ok.data(), b->soft_ok.p, (size_t)b->N, hipMemcpyDeviceToHost, ctx->stream));
    if (b->soft_piped)
        NWG_HIP(ctx, hipMemcpyAsync(&serr, b->lsync.p, sizeof serr, hipMemcpyDeviceToHost, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc = check_soft_err(b, serr);
    if (rc) return rc;
    std::vector<int32_t> refused;
    for (int64_t i = 0; i < b->N; ++i)
        if (!b->host_ok[(size_t)i]) refused.push_back((int32_t)i);
    if (refused.empty()) return NWG_OK;
    const bool ran_f32 = b->prob.precision == 0 && b->soft_precision == 0;
    if (ran_f32) {
        b->soft_reruns = (int64_t)refused.size();
        rc = build_soft_plans(b, refused, b->rr_main, b->rr_wide, b->fill == 0);
        if (rc) return rc;
        rc = launch_soft<double>(b, b->rr_main, b->rr_wide);
        if (rc) return rc;
        NWG_HIP(ctx, hipMemcpyAsync(b->host_ok.data(), b->soft_ok.p, (size_t)b->N, hipMemcpyDeviceToHost, ctx->stream));
        if (b->soft_piped)
            NWG_HIP(ctx, hipMemcpyAsync(&serr, b->lsync.p, sizeof serr, hipMemcpyDeviceToHost, ctx->stream));
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        rc = check_soft_err(b, serr);
        if (rc) return rc;
        refused.clear();
        for (int64_t i = 0; i < b->N; ++i)
            if (!b->host_ok[(size_t)i]) refused.push_back((int32_t)i);
        if (refused.empty()) return NWG_OK;
    }
    b->results_valid = false;
    return fail(ctx, NWG_ERANGE,
                "nwgrad: soft path: %lld pair(s) (first: %lld) do not fit the scaled probability-space "
                "forward-backward (mass lost to underflow could move the result by more than 2^-45, or a weight "
                "overflows -- typically a very low temperature or extreme scores); the reference's "
                "soft_impl=\"scaled_or_log\" would rerun them in log space, which this GPU path does not have",
                (long long)refused.size(), (long long)refused[0]);
}

// What the Hirschberg path needs of a pass (aligner.hpp:3295-3330): Global affine untracked
// pairs, non-negative gap penalties (the reference raises for an explicit Hirschberg mode
// otherwise), and a pass that may synchronise (the host cuts each recursion level).
int check_hb(nwg_batch* b, bool capturing) {
    nwg_ctx* ctx = b->ctx;
    if (b->prob.gap_model != 1 || b->prob.mode != 0)
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: traceback=\"hirschberg\" / \"hirschberg_pmax\" is on the GPU path for "
                    "Global affine pairs only");
    if (b->tracked)
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: traceback=\"hirschberg\" / \"hirschberg_pmax\" of tracked pairs is not "
                    "on the GPU path");
    for (int k = 0; k < 4; ++k)
        if (b->gaps[(size_t)k] < 0.0)
            return fail(ctx, NWG_EINVAL, "nwgrad: traceback=\"hirschberg\" / \"hirschberg_pmax\" needs non-negative gap "
                        "penalties (got a negative one); use traceback=\"pointers\" or \"auto\"");
    if (capturing)
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: a pass with Hirschberg pairs synchronises once per recursion level and "
                    "cannot be captured into a graph");
    return NWG_OK;
}

// The pairs of b->hb_pairs through the recursion (hirschberg.hip), results into the batch's
// scores, gradient rows, guides and stored paths.
int launch_hb(nwg_batch* b) {
    nwg_ctx* ctx = b->ctx;
    if (!b->hbw) b->hbw.reset(nwg_hb::work_new());
    if (!b->hbw) return fail(ctx, NWG_ENOMEM, "nwgrad_hip: out of host memory");
    nwg_hb::Job J;
    J.stream = ctx->stream;
    J.precision = b->prob.precision;
    J.pmax = b->traceback == 4 ? 1 : 0;
    J.cutoff = b->hb_cutoff;
    J.na = b->na;
    J.codes_a = b->ca.p;
    J.codes_b = b->cb.p;
    J.sub = b->prob.precision ? static_cast<const void*>(b->sub_f64.p) : static_cast<const void*>(b->sub_f32.p);
    J.go_a = b->gaps[0];
    J.ge_a = b->gaps[1];
    J.go_b = b->gaps[2];
    J.ge_b = b->gaps[3];
    J.scores = b->scores.p;
    J.grads = b->grads.p;
    J.guide = b->guide.p;
    if (b->keep_paths) {
        J.ops = b->ops.p;
        J.ops_n = b->ops_n.p;
        J.ends = b->ends.p;
    }
    J.pairs = &b->hb_pairs;
    J.off_a = &b->hoa;
    J.off_b = &b->hob;
    J.table_budget = b->table_budget;
    std::string err;
    const int rc = nwg_hb::run(b->hbw.get(), J, err);
    if (rc) return fail(ctx, rc, "%s", err.c_str());
    return NWG_OK;
}

// record: the pass is asynchronous (its caller may consume results before any rerun), so its
// float32 refusals go into the batch's refusal record.
int run_async(nwg_batch* b, bool record) {
    nwg_ctx* ctx = b->ctx;
    if (!b->have_params) return fail(ctx, NWG_ESTATE, "nwgrad_hip: set_params() has not been called");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    NWG_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cap));
    const bool timed = cap == hipStreamCaptureStatusNone;   // no event nodes in a captured graph
    if (timed) NWG_HIP(ctx, hipEventRecord(b->ev0, ctx->stream));
    int rc = NWG_OK;
    if (b->prob.grad_mode == 2) {
        rc = launch_soft_batch(b, record);
        if (rc) return rc;
    } else {
        rc = prepare_tracks(b);
        if (rc) return rc;
        const bool hb = !b->hb_pairs.empty() && !b->banded;
        if (hb != b->plan_excl_hb) {
            rc = replan(b);
            if (rc) return rc;
        }
        if (hb) {
            rc = check_hb(b, !timed);
            if (rc) return rc;
        }
        if (!b->plan.tiles.empty()) {
            rc = b->prob.precision == 0 ? launch_hard<float>(b, b->plan) : launch_hard<double>(b, b->plan);
            if (rc) return rc;
        }
        if (!b->lpl.items.empty()) {
            rc = b->prob.precision == 0 ? launch_long<float>(b) : launch_long<double>(b);
            if (rc) return rc;
        }
        if (hb) {
            rc = launch_hb(b);
            if (rc) return rc;
        }
    }
    if (timed) NWG_HIP(ctx, hipEventRecord(b->ev1, ctx->stream));
    b->results_valid = true;
    if (b->prob.grad_mode == 1) {
        b->guides_valid = true;
        b->paths_valid = b->keep_paths;
    }
    return NWG_OK;
}

int weighted_dev(nwg_batch* b, const double* w_dev, double* out_dev) {
    nwg_ctx* ctx = b->ctx;
    const int gs = b->K * (b->na * b->na + 4);
    const int64_t nblocks = (b->N + kSumBlock - 1) / kSumBlock;
    NWG_HIP(ctx, b->partial.reserve((size_t)(nblocks * gs)));
    const int threads = 64;
    const int64_t total = nblocks * gs;
    const int grid = (int)((total + threads - 1) / threads);
    const uint8_t* ok = b->prob.grad_mode == 2 ? b->soft_ok.p : nullptr;
    hipLaunchKernelGGL(nwg::block_sum_kernel, dim3(std::max(grid, 1)), dim3(threads), 0, ctx->stream, b->grads.p,
                       w_dev, ok, b->N, gs, kSumBlock, b->partial.p);
    NWG_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(nwg::block_total_kernel, dim3((gs + 63) / 64), dim3(64), 0, ctx->stream, b->partial.p, nblocks,
                       gs, out_dev);
    NWG_HIP(ctx, hipGetLastError());
    return NWG_OK;
}

// The sum of the scores in pair-index order, as the reference forms the total
// (batch_engine.hpp:1754-1757), after the results are complete on the stream.  Where no
// partial sum can round (ScoreStats), every order gives that double and the device's sum is
// taken (32 bytes back); otherwise the scores come back and are summed in pair order here.
int score_total(nwg_batch* b, double* total_out) {
    nwg_ctx* ctx = b->ctx;
    double t = 0.0;
    if (b->N > 0) {
        const int parts = (int)std::min<int64_t>((b->N + nwg::kStatsThreads - 1) / nwg::kStatsThreads, 1024);
        NWG_HIP(ctx, b->stats.reserve((size_t)parts + 1));
        if (!b->host_stats)
            NWG_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&b->host_stats), sizeof(nwg::ScoreStats),
                                       hipHostMallocDefault));
        hipLaunchKernelGGL(nwg::score_stats_kernel, dim3(parts), dim3(nwg::kStatsThreads), 0, ctx->stream,
                           b->scores.p, b->N, b->stats.p);
        NWG_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(nwg::score_stats_final_kernel, dim3(1), dim3(nwg::kStatsThreads), 0, ctx->stream,
                           b->stats.p, parts, b->stats.p + parts);
        NWG_HIP(ctx, hipGetLastError());
        NWG_HIP(ctx, hipMemcpyAsync(b->host_stats, b->stats.p + parts, sizeof(nwg::ScoreStats),
                                    hipMemcpyDeviceToHost, ctx->stream));
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const nwg::ScoreStats st = *b->host_stats;
        // st.mag is a sum of N non-negative doubles in some order: the true sum of magnitudes
        // is below st.mag * (1 + 2N * 2^-53); the extra 4 * 2^-52 covers this product's rounding.
        const double bound = st.mag * (1.0 + (double)(b->N + 4) * 0x1p-52);
        const bool exact = st.bad == 0 && (st.qmin == nwg::kNoBit || bound < std::ldexp(1.0, 53 + st.qmin));
        if (exact) {
            t = st.sum + 0.0;
            ++b->total_fast;
        } else {
            NWG_HIP(ctx, hipMemcpyAsync(b->host_scores, b->scores.p, sizeof(double) * (size_t)b->N,
                                        hipMemcpyDeviceToHost, ctx->stream));
            NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (int64_t i = 0; i < b->N; ++i) t += b->host_scores[i];
            ++b->total_ordered;
        }
    }
    *total_out = t;
    return NWG_OK;
}

}  // namespace

extern "C" {

int nwg_batch_run_async(nwg_batch* b) {
    if (!b) return NWG_EINVAL;
    return run_async(b, true);
}

int nwg_batch_soft_finish(nwg_batch* b) {
    if (!b) return NWG_EINVAL;
    if (b->prob.grad_mode != 2) return NWG_OK;
    if (!b->results_valid) return fail(b->ctx, NWG_ESTATE, "nwgrad: no soft pass to finish");
    NWG_HIP(b->ctx, hipSetDevice(b->ctx->device));
    return soft_finish(b);
}

int nwg_batch_run(nwg_batch* b, double* total_out) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    int rc = run_async(b, false);
    if (rc) return rc;
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
    rc = check_long_err(b);
    if (rc) return rc;
    if (b->prob.grad_mode == 2) {
        rc = soft_finish(b);
        if (rc) return rc;
    }
    if (total_out) return score_total(b, total_out);
    return NWG_OK;
}

int nwg_batch_scores(nwg_batch* b, double* out) {
    if (!b || !out) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (!b->results_valid) return fail(ctx, NWG_ESTATE, "nwgrad: score not valid; compute an alignment first");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    if (b->N)
        NWG_HIP(ctx, hipMemcpyAsync(out, b->scores.p, sizeof(double) * (size_t)b->N, hipMemcpyDeviceToHost,
                                    ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return check_long_err(b);   // an asynchronous pass's long-pair failure surfaces here
}

int nwg_batch_pair_result(nwg_batch* b, int64_t i, double* score, double* grad) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (i < 0 || i >= b->N)
        return fail(ctx, NWG_EINDEX, "nwgrad: pair index %lld out of range for %lld pairs", (long long)i,
                    (long long)b->N);
    if (!b->results_valid) return fail(ctx, NWG_ESTATE, "nwgrad: score not valid; compute an alignment first");
    if (grad && b->prob.grad_mode == 0) return fail(ctx, NWG_ESTATE, "nwgrad: grad_mode \"none\" computes no gradient");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    const int gs = b->K * (b->na * b->na + 4);
    if (score) NWG_HIP(ctx, hipMemcpyAsync(score, b->scores.p + i, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (grad)
        NWG_HIP(ctx, hipMemcpyAsync(grad, b->grads.p + i * gs, sizeof(double) * (size_t)gs, hipMemcpyDeviceToHost,
                                    ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return check_long_err(b);
}

int nwg_batch_scores_dev(nwg_batch* b, const double** out_dev) {
    if (!b || !out_dev) return NWG_EINVAL;
    *out_dev = b->scores.p;
    return NWG_OK;
}

int nwg_batch_scores_copy_dev(nwg_batch* b, double* dst_dev) {
    if (!b || !dst_dev) return NWG_EINVAL;
    if (!b->results_valid) return fail(b->ctx, NWG_ESTATE, "nwgrad: score not valid; compute an alignment first");
    if (b->N == 0) return NWG_OK;
    NWG_HIP(b->ctx, hipSetDevice(b->ctx->device));
    NWG_HIP(b->ctx, hipMemcpyAsync(dst_dev, b->scores.p, sizeof(double) * (size_t)b->N, hipMemcpyDeviceToDevice,
                                   b->ctx->stream));
    return NWG_OK;
}

int nwg_batch_grads(nwg_batch* b, double* mats, double* gaps) {
    if (!b || !mats || !gaps) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->prob.grad_mode == 0) return fail(ctx, NWG_ESTATE, "nwgrad: grad_mode \"none\" computes no gradient");
    if (!b->results_valid) return fail(ctx, NWG_ESTATE, "nwgrad: gradient not valid; compute an alignment first");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    // per pair: K matrices (mats[(i*K + k)*nn ..]) and K gap rows (gaps[(i*K + k)*4 ..]),
    // the reference's grads_into layout (batch_engine.hpp:719-746)
    const int K = b->K, nn = b->na * b->na, knn = K * nn, gs = K * (nn + 4);
    if (b->N == 0) return NWG_OK;
    std::vector<double> tmp((size_t)b->N * gs);
    NWG_HIP(ctx, hipMemcpyAsync(tmp.data(), b->grads.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int rc = check_long_err(b);
    if (rc) return rc;
    for (int64_t i = 0; i < b->N; ++i) {
        std::memcpy(mats + i * knn, tmp.data() + i * gs, sizeof(double) * (size_t)knn);
        std::memcpy(gaps + i * 4 * K, tmp.data() + i * gs + knn, sizeof(double) * 4 * (size_t)K);
    }
    return NWG_OK;
}

int nwg_batch_weighted_grad(nwg_batch* b, const double* w, int64_t n, double* out) {
    if (!b || !out) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->N == 0) return fail(ctx, NWG_EINVAL, "nwgrad: SeqPairBatch is empty");
    if (w && n != b->N)
        return fail(ctx, NWG_EINVAL,
                    "nwgrad: weighted_grad() needs one weight per pair (got %lld weights for %lld pairs)",
                    (long long)n, (long long)b->N);
    if (b->prob.grad_mode == 0) return fail(ctx, NWG_ESTATE, "nwgrad: grad_mode \"none\" computes no gradient");
    if (!b->results_valid) return fail(ctx, NWG_ESTATE, "nwgrad: gradient not valid; compute an alignment first");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    const int gs = b->K * (b->na * b->na + 4);
    const double* wd = nullptr;
    if (w) {
        NWG_HIP(ctx, b->wdev.reserve((size_t)b->N));
        NWG_HIP(ctx, hipMemcpyAsync(b->wdev.p, w, sizeof(double) * (size_t)b->N, hipMemcpyHostToDevice, ctx->stream));
        wd = b->wdev.p;
    }
    NWG_HIP(ctx, b->outdev.reserve((size_t)gs));
    int rc = weighted_dev(b, wd, b->outdev.p);
    if (rc) return rc;
    NWG_HIP(ctx, hipMemcpyAsync(out, b->outdev.p, sizeof(double) * (size_t)gs, hipMemcpyDeviceToHost, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return check_long_err(b);
}

int nwg_batch_weighted_grad_dev(nwg_batch* b, const double* w_dev, double* out_dev) {
    if (!b || !out_dev) return NWG_EINVAL;
    if (b->N == 0) return fail(b->ctx, NWG_EINVAL, "nwgrad: SeqPairBatch is empty");
    if (b->prob.grad_mode == 0) return fail(b->ctx, NWG_ESTATE, "nwgrad: grad_mode \"none\" computes no gradient");
    if (!b->results_valid) return fail(b->ctx, NWG_ESTATE, "nwgrad: gradient not valid; compute an alignment first");
    return weighted_dev(b, w_dev, out_dev);
}

int nwg_batch_plan_info(nwg_batch* b, int64_t* n_tiles, int64_t* n_cells) {
    if (!b) return NWG_EINVAL;
    if (n_tiles) {
        if (b->prob.grad_mode == 2) {
            SoftPlan& mp = b->fill == 0 ? b->sp_main : b->sq_main;
            SoftPlan& wp = b->fill == 0 ? b->sp_wide : b->sq_wide;
            if (!mp.built) {
                std::vector<int32_t> all((size_t)b->N);
                std::iota(all.begin(), all.end(), 0);
                const int rc = build_soft_plans(b, all, mp, wp, b->fill == 0);
                if (rc) return rc;
            }
            *n_tiles = (int64_t)(mp.tiles.size() + wp.tiles.size());
        } else {
            *n_tiles = (int64_t)(b->plan.tiles.size() + b->lpl.items.size());
        }
    }
    if (n_cells) *n_cells = b->cells;
    return NWG_OK;
}

int nwg_batch_last_kernel_ms(nwg_batch* b, double* ms) {
    if (!b || !ms) return NWG_EINVAL;
    *ms = b->last_ms;
    return NWG_OK;
}

int nwg_batch_launch_info(nwg_batch* b, int64_t* grid, int64_t* wide_slots, int64_t* per_cu) {
    if (!b) return NWG_EINVAL;
    if (grid) *grid = b->soft_grid;
    if (wide_slots) *wide_slots = b->soft_lslots;
    if (per_cu) *per_cu = b->soft_occ;
    return NWG_OK;
}

int nwg_batch_soft_reruns(nwg_batch* b, int64_t* n) {
    if (!b || !n) return NWG_EINVAL;
    *n = b->soft_reruns;
    return NWG_OK;
}

int nwg_batch_set_keep_paths(nwg_batch* b, int keep) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (keep && b->prob.grad_mode != 1)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: stored paths come from hard passes (grad_mode=\"hard\")");
    if (keep && b->ops.n == 0) {
        int64_t cap = 0;
        for (int64_t i = 0; i < b->N; ++i) cap += (int64_t)b->mlen[(size_t)i] + b->nlen[(size_t)i];
        NWG_HIP(ctx, hipSetDevice(ctx->device));
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NWG_HIP(ctx, b->ops.alloc((size_t)std::max<int64_t>(cap, 1)));
        NWG_HIP(ctx, b->ops_n.alloc((size_t)std::max<int64_t>(b->N, 1)));
        NWG_HIP(ctx, b->ends.alloc((size_t)std::max<int64_t>(2 * b->N, 1)));
    }
    b->keep_paths = keep != 0;
    return NWG_OK;
}

int nwg_batch_drop_paths(nwg_batch* b) {
    if (!b) return NWG_EINVAL;
    NWG_HIP(b->ctx, hipSetDevice(b->ctx->device));
    NWG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    b->ops.reset();
    b->ops_n.reset();
    b->ends.reset();
    b->keep_paths = false;
    b->paths_valid = false;
    return NWG_OK;
}

int nwg_batch_guides(nwg_batch* b, int32_t* out) {
    if (!b || !out) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->prob.grad_mode != 1)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: guides come from hard passes (grad_mode=\"hard\")");
    if (!b->guides_valid) return fail(ctx, NWG_ESTATE, "nwgrad: alignment not computed; call align() first");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t n = b->N ? b->hoa[(size_t)b->N] + b->N : 0;
    if (n) NWG_HIP(ctx, hipMemcpyAsync(out, b->guide.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return check_long_err(b);
}

int nwg_batch_pair_guide(nwg_batch* b, int64_t i, int32_t* out) {
    if (!b || !out) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (i < 0 || i >= b->N)
        return fail(ctx, NWG_EINDEX, "nwgrad: pair index %lld out of range for %lld pairs", (long long)i, (long long)b->N);
    if (b->prob.grad_mode != 1)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: guides come from hard passes (grad_mode=\"hard\")");
    if (!b->guides_valid) return fail(ctx, NWG_ESTATE, "nwgrad: alignment not computed; call align() first");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    NWG_HIP(ctx, hipMemcpyAsync(out, b->guide.p + b->hoa[(size_t)i] + i, sizeof(int32_t) * (size_t)(b->mlen[(size_t)i] + 1),
                                hipMemcpyDeviceToHost, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return check_long_err(b);
}

int nwg_batch_paths(nwg_batch* b, uint32_t* ops, int32_t* n_ops, int32_t* ends) {
    if (!b || !ops || !n_ops || !ends) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (!b->paths_valid)
        return fail(ctx, NWG_ESTATE,
                    "nwgrad: stored path unavailable; use align(keep_paths=True) or banded_align(keep_paths=True)");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    int64_t cap = 0;
    for (int64_t i = 0; i < b->N; ++i) cap += (int64_t)b->mlen[(size_t)i] + b->nlen[(size_t)i];
    if (cap) NWG_HIP(ctx, hipMemcpyAsync(ops, b->ops.p, sizeof(uint32_t) * (size_t)cap, hipMemcpyDeviceToHost, ctx->stream));
    if (b->N) {
        NWG_HIP(ctx, hipMemcpyAsync(n_ops, b->ops_n.p, sizeof(int32_t) * (size_t)b->N, hipMemcpyDeviceToHost, ctx->stream));
        NWG_HIP(ctx, hipMemcpyAsync(ends, b->ends.p, sizeof(int32_t) * 2 * (size_t)b->N, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return check_long_err(b);
}

int nwg_batch_run_banded(nwg_batch* b, int bandwidth, double* total_out) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->prob.grad_mode == 0)
        return fail(ctx, NWG_EINVAL,
                    "nwgrad_amd: a score-only banded pass is its hard batch's (grad_mode=\"hard\") banded scores");
    if (b->prob.grad_mode == 2 && b->tracked)
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: the banded soft pass of tracked pairs is not on the GPU path");
    if (bandwidth < 0) return fail(ctx, NWG_EINVAL, "nwgrad: band must be >= 0, got %d", bandwidth);
    if (!b->guides_valid)
        return fail(ctx, NWG_ESTATE, "nwgrad: banded_grad() needs a guide path; run score_and_grad() first");
    b->banded = true;
    b->band = bandwidth;
    const int rc = nwg_batch_run(b, total_out);
    b->banded = false;
    b->band = 0;
    return rc;
}

int nwg_batch_set_guides(nwg_batch* b, const int32_t* guides) {
    if (!b || !guides) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->prob.grad_mode == 0)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: guides belong to hard and soft batches");
    for (int64_t p = 0; p < b->N; ++p) {   // the reference's validate_guide (aligner.hpp:2622-2642)
        const int32_t* g = guides + b->hoa[(size_t)p] + p;
        const int m = b->mlen[(size_t)p], n = b->nlen[(size_t)p];
        for (int i = 0; i <= m; ++i) {
            if (g[i] < 0 || g[i] > n)
                return fail(ctx, NWG_EINVAL, "nwgrad: guide_j[%d] = %d is outside [0, %d] for sequence B of length %d", i,
                            (int)g[i], n, n);
            if (i > 0 && g[i] < g[i - 1])
                return fail(ctx, NWG_EINVAL, "nwgrad: guide_j decreases at row %d (%d -> %d)", i, (int)g[i - 1], (int)g[i]);
        }
    }
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t n = b->N ? b->hoa[(size_t)b->N] + b->N : 0;
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NWG_HIP(ctx, b->guide.reserve((size_t)std::max<int64_t>(n, 1)));   // soft batches: the band's guides
    if (n) NWG_HIP(ctx, hipMemcpyAsync(b->guide.p, guides, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->guides_valid = true;
    return NWG_OK;
}

int nwg_batch_soft_refusals(nwg_batch* b, int64_t* pair_passes, int64_t* first_pass, int64_t* passes, int reset) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (b->prob.grad_mode != 2) return fail(ctx, NWG_EINVAL, "nwgrad_hip: refusal records belong to soft batches");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    unsigned long long r[3] = {0, 0, 0};
    NWG_HIP(ctx, hipMemcpyAsync(r, b->srec.p, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (pair_passes) *pair_passes = (int64_t)r[0];
    if (first_pass) *first_pass = (int64_t)r[1];
    if (passes) *passes = (int64_t)r[2];
    if (reset) {   // refusal counts start again; the pass counter keeps counting
        r[0] = r[1] = 0;
        NWG_HIP(ctx, hipMemcpyAsync(b->srec.p, r, 2 * sizeof(unsigned long long), hipMemcpyHostToDevice, ctx->stream));
        NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return NWG_OK;
}

int nwg_batch_set_traceback(nwg_batch* b, int traceback, int hb_cutoff) {
    if (!b) return NWG_EINVAL;
    nwg_ctx* ctx = b->ctx;
    if (traceback < 0 || traceback > 4) return fail(ctx, NWG_EINVAL, "nwgrad: unknown traceback %d", traceback);
    if (hb_cutoff < 1) return fail(ctx, NWG_EINVAL, "nwgrad: hb_cutoff must be >= 1");
    if ((traceback == 3 || traceback == 4) && b->prob.grad_mode == 1 && (b->prob.gap_model != 1 || b->prob.mode != 0))
        return fail(ctx, NWG_EINVAL, "nwgrad_amd: traceback=\"hirschberg\" / \"hirschberg_pmax\" is on the GPU path for "
                    "Global affine pairs only");
    const bool changed = traceback != b->traceback || hb_cutoff != b->hb_cutoff;
    b->traceback = traceback;
    b->hb_cutoff = hb_cutoff;
    if (changed && b->prob.grad_mode != 2) {
        NWG_HIP(ctx, hipSetDevice(ctx->device));
        const int rc = replan(b);
        if (rc) return rc;
    }
    return NWG_OK;
}

int nwg_batch_hb_info(nwg_batch* b, int64_t* pairs, int64_t* levels, int64_t* base_blocks, int64_t* scratch_bytes,
                      int64_t* filled_cells) {
    if (!b || !pairs || !levels || !base_blocks || !scratch_bytes || !filled_cells) return NWG_EINVAL;
    *pairs = (int64_t)b->hb_pairs.size();
    const nwg_hb::Stats z{};
    const nwg_hb::Stats& st = b->hbw ? nwg_hb::work_stats(b->hbw.get()) : z;
    *levels = st.levels;
    *base_blocks = st.base_blocks;
    *scratch_bytes = st.peak_bytes;
    *filled_cells = st.sweep_cells + st.base_cells;
    return NWG_OK;
}

int nwg_batch_set_soft_scratch(nwg_batch* b, int64_t bytes) {
    if (!b) return NWG_EINVAL;
    if (bytes < 0) return fail(b->ctx, NWG_EINVAL, "nwgrad_hip: soft scratch bytes must be >= 0 (0: the default)");
    b->soft_scratch = bytes == 0 ? kScratchBudget : (size_t)bytes;
    return NWG_OK;
}

int nwg_batch_set_table_budget(nwg_batch* b, int64_t bytes) {
    if (!b) return NWG_EINVAL;
    if (bytes < 0) return fail(b->ctx, NWG_EINVAL, "nwgrad_hip: table budget bytes must be >= 0 (0: the default)");
    b->table_budget = bytes == 0 ? kHardTableBudget : bytes;
    split_launches(b, b->plan);
    return NWG_OK;
}

int nwg_batch_hard_launches(nwg_batch* b, int64_t* launches, int64_t* persistent, int64_t* last_grid) {
    if (!b) return NWG_EINVAL;
    if (launches) *launches = b->plan.tiles.empty() ? 0 : b->fill == 0 ? 1 : (int64_t)b->plan.bucket_launches.size();
    if (persistent) *persistent = b->fill == 0 && hard_persistent(b, b->plan) ? 1 : 0;
    if (last_grid) *last_grid = b->hard_grid;
    return NWG_OK;
}

int nwg_batch_wide_info(nwg_batch* b, int64_t* tiles, int64_t* pooled, int64_t* pool_rows, int64_t* lag) {
    if (!b) return NWG_EINVAL;
    if (tiles) *tiles = b->soft_ptiles;
    if (pooled) *pooled = b->soft_qtiles;
    if (pool_rows) *pool_rows = b->soft_pool;
    if (lag) *lag = b->soft_lag;
    return NWG_OK;
}

int nwg_batch_total_info(nwg_batch* b, int64_t* from_device, int64_t* in_pair_order) {
    if (!b) return NWG_EINVAL;
    if (from_device) *from_device = b->total_fast;
    if (in_pair_order) *in_pair_order = b->total_ordered;
    return NWG_OK;
}

}  // extern "C"

#include "noalloc.hpp"
