// Host driver of the Hirschberg traceback (hb_kernels.hpp): the recursion level by level.
//
// Every pair that splits starts as one node (rows (0, m], columns (0, n]).  A level sweeps
// both halves of each of its nodes in one persistent launch (panel items pulled from a
// device queue), joins each node in a second launch, and returns each node's crossing
// column and span flag to the host, which cuts the children.  The row structure of the
// recursion is fixed by m and hb_cutoff; only the columns depend on the data, so a level's
// panel count is known only after the previous join: one stream synchronisation per level
// (about log2(m / hb_cutoff) of them; 7 for 36,000 rows at 512).  Children of at most
// hb_cutoff rows are base cases, filled and walked one wave each after the last level; the
// replay then walks each pair's base blocks in row order (the reference's recursion emits
// the halves in that order, aligner.hpp:4514-4517).
//
// Memory is linear in the pair lengths: per level, the halves' last rows (3 (W + 1) values a
// node and half) and the panel boundary columns (3 or 4 values a row and panel); the base
// cases' codes, at most hb_cutoff + 64 rows by the block width, a byte a cell.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nwgrad_hip.h"
#include "hb_host.hpp"
#include "hb_kernels.hpp"

namespace nwg_hb {

namespace {

constexpr uint64_t kWaitTicks = 200000000;   // a row-block wait, at most (2 s of the 100 MHz s_memrealtime)

template <typename P>
struct Buf {
    P* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    // at least `count` elements, contents not kept; grows with headroom
    hipError_t reserve(size_t count) {
        if (count == 0) count = 1;
        if (p && n >= count) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        const size_t want = count + count / 4;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(P));
        if (e == hipSuccess) n = want;
        else p = nullptr;
        return e;
    }
    size_t bytes() const { return n * sizeof(P); }
};

__global__ __launch_bounds__(256) void hb_clear_kernel(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

struct Node {
    int32_t pr;               // index into the job's pairs
    int32_t i0, i1, j0, j1;
    uint8_t in_x, out_x;
};

}  // namespace

struct Work {
    Buf<nwg::HbItem> items;
    Buf<nwg::HbJoin> joins;
    Buf<nwg::HbBase> blocks;
    Buf<nwg::HbPair> hpairs;
    Buf<int32_t> prog, res, words, hcount;
    Buf<float> bnd32, out32, bbnd32;
    Buf<double> bnd64, out64, bbnd64;
    Buf<uint8_t> table, hops;
    // host staging, alive until the next pass (uploads are asynchronous)
    std::vector<nwg::HbItem> h_items;
    std::vector<nwg::HbJoin> h_joins;
    std::vector<nwg::HbBase> h_blocks;
    std::vector<nwg::HbPair> h_pairs;
    std::vector<int32_t> h_res;
    Stats st;
};

Work* work_new() { return new Work(); }
void work_free(Work* w) { delete w; }
const Stats& work_stats(const Work* w) { return w->st; }

namespace {

template <typename T> struct Bufs;
template <> struct Bufs<float> {
    static Buf<float>& bnd(Work* w) { return w->bnd32; }
    static Buf<float>& out(Work* w) { return w->out32; }
    static Buf<float>& bbnd(Work* w) { return w->bbnd32; }
};
template <> struct Bufs<double> {
    static Buf<double>& bnd(Work* w) { return w->bnd64; }
    static Buf<double>& out(Work* w) { return w->out64; }
    static Buf<double>& bbnd(Work* w) { return w->bbnd64; }
};

#define HB_HIP(call)                                                                           \
    do {                                                                                       \
        const hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                                \
            err = std::string(#call " failed: ") + hipGetErrorString(e_);                       \
            return NWG_EDEVICE;                                                                \
        }                                                                                      \
    } while (0)

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename T, bool PMAX>
int run_t(Work* w, const Job& J, std::string& err) {
    const std::vector<int32_t>& pairs = *J.pairs;
    const std::vector<int64_t>& oa = *J.off_a;
    const std::vector<int64_t>& ob = *J.off_b;
    hipStream_t s = J.stream;
    const T go_a = static_cast<T>(J.go_a), ge_a = static_cast<T>(J.ge_a);
    const T go_b = static_cast<T>(J.go_b), ge_b = static_cast<T>(J.ge_b);
    const T* sub = static_cast<const T*>(J.sub);
    constexpr int NB = PMAX ? 4 : 3;
    Stats st;
    int dev = 0, n_cu = 256;
    HB_HIP(hipGetDevice(&dev));
    HB_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds_sub = 256 + (size_t)J.na * J.na * sizeof(T);   // the A-code ring, then the matrix
    int per_cu = 0;
    HB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, nwg::hb_sweep_kernel<T, PMAX>, 64, lds_sub));
    const int64_t resident = (int64_t)std::max(per_cu, 1) * n_cu;
    HB_HIP(w->words.reserve(2));   // [0] the sweep queue, [1] the error word

    std::vector<Node> cur, next, bases;
    for (size_t k = 0; k < pairs.size(); ++k) {
        const int32_t p = pairs[k];
        cur.push_back(Node{(int32_t)k, 0, (int32_t)(oa[p + 1] - oa[p]), 0, (int32_t)(ob[p + 1] - ob[p]), 0, 0});
    }
    while (!cur.empty()) {
        ++st.levels;
        auto& items = w->h_items;
        auto& joins = w->h_joins;
        items.clear();
        joins.clear();
        int64_t out_elems = 0, bnd_elems = 0;
        for (size_t k = 0; k < cur.size(); ++k) {
            const Node& nd = cur[k];
            const int32_t p = pairs[(size_t)nd.pr];
            const int H = nd.i1 - nd.i0, mid = nd.i0 + H / 2, W = nd.j1 - nd.j0;
            nwg::HbJoin jn{};
            jn.W = W;
            for (int half = 0; half < 2; ++half) {
                const int Hh = half ? nd.i1 - mid : mid - nd.i0;
                const int64_t outoff = out_elems;
                out_elems += 3 * ((int64_t)W + 1);
                (half ? jn.rev : jn.fwd) = outoff;
                const int P = (int)std::max<int64_t>(1, ceil_div(W, nwg::kHbSweepPanel));
                for (int pp = 0; pp < P; ++pp) {
                    nwg::HbItem it{};
                    it.node = (int32_t)k;
                    it.panel = pp;
                    it.H = Hh;
                    it.W = W;
                    it.flag = half ? nd.out_x : nd.in_x;
                    it.a_step = half ? -1 : 1;
                    it.b_step = half ? -1 : 1;
                    it.a_base = half ? oa[p] + nd.i1 - 1 : oa[p] + nd.i0;
                    it.b_base = half ? ob[p] + nd.j1 - 1 : ob[p] + nd.j0;
                    it.bnd_in = pp == 0 ? -1 : bnd_elems - (int64_t)NB * (Hh + 1);
                    it.bnd_out = -1;
                    if (pp + 1 < P) {
                        it.bnd_out = bnd_elems;
                        bnd_elems += (int64_t)NB * (Hh + 1);
                    }
                    it.out = outoff;
                    items.push_back(it);
                }
            }
            joins.push_back(jn);
            st.sweep_cells += (int64_t)H * W;
        }
        st.sweep_items += (int64_t)items.size();
        st.nodes += (int64_t)cur.size();
        Buf<T>& bnd = Bufs<T>::bnd(w);
        Buf<T>& out = Bufs<T>::out(w);
        HB_HIP(w->items.reserve(items.size()));
        HB_HIP(w->joins.reserve(joins.size()));
        HB_HIP(w->prog.reserve(items.size()));
        HB_HIP(w->res.reserve(2 * joins.size()));
        HB_HIP(bnd.reserve((size_t)std::max<int64_t>(bnd_elems, 1)));
        HB_HIP(out.reserve((size_t)out_elems));
        HB_HIP(hipMemcpyAsync(w->items.p, items.data(), sizeof(nwg::HbItem) * items.size(), hipMemcpyHostToDevice, s));
        HB_HIP(hipMemcpyAsync(w->joins.p, joins.data(), sizeof(nwg::HbJoin) * joins.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(hb_clear_kernel, dim3((unsigned)ceil_div((int64_t)items.size(), 256)), dim3(256), 0, s,
                           w->prog.p, (int64_t)items.size());
        HB_HIP(hipGetLastError());
        hipLaunchKernelGGL(hb_clear_kernel, dim3(1), dim3(256), 0, s, w->words.p, (int64_t)2);
        HB_HIP(hipGetLastError());
        nwg::HbSweepArgs<T> A{};
        A.codes_a = J.codes_a;
        A.codes_b = J.codes_b;
        A.sub = sub;
        A.na = J.na;
        A.go_a = go_a; A.ge_a = ge_a; A.go_b = go_b; A.ge_b = ge_b;
        A.items = w->items.p;
        A.n_items = (int32_t)items.size();
        A.queue = w->words.p;
        A.progress = w->prog.p;
        A.bnd = bnd.p;
        A.out = out.p;
        A.err = w->words.p + 1;
        A.wait_ticks = kWaitTicks;
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)items.size(), resident));
        hipLaunchKernelGGL((nwg::hb_sweep_kernel<T, PMAX>), dim3((unsigned)grid), dim3(64), lds_sub, s, A);
        HB_HIP(hipGetLastError());
        nwg::HbJoinArgs<T> JA{};
        JA.nodes = w->joins.p;
        JA.out = out.p;
        JA.go_b = go_b;
        JA.res = w->res.p;
        hipLaunchKernelGGL(nwg::hb_join_kernel<T>, dim3((unsigned)joins.size()), dim3(256), 0, s, JA);
        HB_HIP(hipGetLastError());
        w->h_res.resize(2 * joins.size() + 2);
        HB_HIP(hipMemcpyAsync(w->h_res.data(), w->res.p, sizeof(int32_t) * 2 * joins.size(), hipMemcpyDeviceToHost, s));
        HB_HIP(hipMemcpyAsync(w->h_res.data() + 2 * joins.size(), w->words.p, sizeof(int32_t) * 2,
                              hipMemcpyDeviceToHost, s));
        HB_HIP(hipStreamSynchronize(s));
        st.peak_bytes = std::max<int64_t>(st.peak_bytes, (int64_t)(bnd.bytes() + out.bytes() + w->items.bytes() +
                                                                   w->prog.bytes()));
        if (w->h_res[2 * joins.size() + 1] != 0) {
            err = "nwgrad_hip: a Hirschberg sweep's wait for a neighbouring panel passed its time bound";
            return NWG_EDEVICE;
        }
        next.clear();
        for (size_t k = 0; k < cur.size(); ++k) {
            const Node& nd = cur[k];
            const int H = nd.i1 - nd.i0, mid = nd.i0 + H / 2;
            const int cstar = w->h_res[2 * k], span = w->h_res[2 * k + 1];
            if (cstar < 0 || cstar > nd.j1 - nd.j0) {
                err = "nwgrad_hip: a Hirschberg join returned a column outside its block";
                return NWG_EDEVICE;
            }
            const int jstar = nd.j0 + cstar;
            const Node lo{nd.pr, nd.i0, mid, nd.j0, jstar, nd.in_x, (uint8_t)span};
            const Node hi{nd.pr, mid, nd.i1, jstar, nd.j1, (uint8_t)span, nd.out_x};
            for (const Node& c : {lo, hi}) (c.i1 - c.i0 <= J.cutoff ? bases : next).push_back(c);
        }
        cur.swap(next);
    }

    // Base cases, in pair then row order: the replay reads each pair's blocks in sequence.
    std::stable_sort(bases.begin(), bases.end(),
                     [](const Node& x, const Node& y) { return x.pr != y.pr ? x.pr < y.pr : x.i0 < y.i0; });
    auto& blocks = w->h_blocks;
    blocks.clear();
    struct Group { size_t first, count; int64_t table; };
    std::vector<Group> groups;
    int64_t gtable = 0, max_table = 0, bbnd = 0, hops = 0;
    Group g{0, 0, 0};
    for (const Node& nd : bases) {
        const int32_t p = pairs[(size_t)nd.pr];
        nwg::HbBase B{};
        B.H = nd.i1 - nd.i0;
        B.W = nd.j1 - nd.j0;
        B.in_x = nd.in_x;
        B.out_x = nd.out_x;
        B.a_base = oa[p] + nd.i0;
        B.b_base = ob[p] + nd.j0;
        const int64_t P = B.W > 0 ? ceil_div(B.W, nwg::kHbPanel) : 0;
        const int64_t tbytes = P * ((int64_t)B.H + nwg::kWave) * nwg::kHbPanel;
        if (g.count > 0 && gtable + tbytes > J.table_budget) {
            g.table = gtable;
            groups.push_back(g);
            max_table = std::max(max_table, gtable);
            g = Group{blocks.size(), 0, 0};
            gtable = 0;
        }
        B.table = gtable;
        gtable += tbytes;
        B.bnd = bbnd;
        bbnd += (P > 1 ? P - 1 : 0) * 3 * ((int64_t)B.H + 1);
        B.hops = hops;
        hops += (int64_t)B.H + B.W;
        blocks.push_back(B);
        ++g.count;
        st.base_cells += (int64_t)B.H * B.W;
    }
    if (g.count > 0) {
        g.table = gtable;
        groups.push_back(g);
        max_table = std::max(max_table, gtable);
    }
    st.base_blocks = (int64_t)blocks.size();
    auto& hp = w->h_pairs;
    hp.clear();
    for (size_t b = 0; b < bases.size(); ++b) {
        const int32_t k = bases[b].pr;
        if (hp.empty() || hp.back().pair != pairs[(size_t)k]) {
            const int32_t p = pairs[(size_t)k];
            hp.push_back(nwg::HbPair{p, (int32_t)b, 0, (int32_t)(oa[p + 1] - oa[p]), (int32_t)(ob[p + 1] - ob[p]), oa[p],
                                     ob[p]});
        }
        ++hp.back().count;
    }
    Buf<T>& bb = Bufs<T>::bbnd(w);
    HB_HIP(w->blocks.reserve(blocks.size()));
    HB_HIP(w->hcount.reserve(blocks.size()));
    HB_HIP(w->hpairs.reserve(hp.size()));
    HB_HIP(w->table.reserve((size_t)std::max<int64_t>(max_table, 1)));
    HB_HIP(bb.reserve((size_t)std::max<int64_t>(bbnd, 1)));
    HB_HIP(w->hops.reserve((size_t)std::max<int64_t>(hops, 1)));
    HB_HIP(hipMemcpyAsync(w->blocks.p, blocks.data(), sizeof(nwg::HbBase) * blocks.size(), hipMemcpyHostToDevice, s));
    HB_HIP(hipMemcpyAsync(w->hpairs.p, hp.data(), sizeof(nwg::HbPair) * hp.size(), hipMemcpyHostToDevice, s));
    nwg::HbBaseArgs<T> BA{};
    BA.codes_a = J.codes_a;
    BA.codes_b = J.codes_b;
    BA.sub = sub;
    BA.na = J.na;
    BA.go_a = go_a; BA.ge_a = ge_a; BA.go_b = go_b; BA.ge_b = ge_b;
    BA.table = w->table.p;
    BA.bnd = bb.p;
    BA.hops = w->hops.p;
    for (const Group& gr : groups) {   // one launch per table budget's worth of blocks, in stream order
        BA.blocks = w->blocks.p + gr.first;
        BA.n_blocks = (int32_t)gr.count;
        BA.hcount = w->hcount.p + gr.first;
        hipLaunchKernelGGL(nwg::hb_base_kernel<T>, dim3((unsigned)gr.count), dim3(64), lds_sub + 4 * sizeof(T), s, BA);
        HB_HIP(hipGetLastError());
    }
    nwg::HbReplayArgs<T> RA{};
    RA.codes_a = J.codes_a;
    RA.codes_b = J.codes_b;
    RA.sub = sub;
    RA.na = J.na;
    RA.go_a = go_a; RA.ge_a = ge_a; RA.go_b = go_b; RA.ge_b = ge_b;
    RA.pairs = w->hpairs.p;
    RA.blocks = w->blocks.p;
    RA.hops = w->hops.p;
    RA.hcount = w->hcount.p;
    RA.scores = J.scores;
    RA.grads = J.grads;
    RA.guide = J.guide;
    RA.ops = J.ops;
    RA.ops_n = J.ops_n;
    RA.ends = J.ends;
    hipLaunchKernelGGL(nwg::hb_replay_kernel<T>, dim3((unsigned)hp.size()), dim3(64),
                       (size_t)nwg::hb_replay_lds_bytes<T>(J.na), s, RA);
    HB_HIP(hipGetLastError());
    st.peak_bytes = std::max<int64_t>(st.peak_bytes, (int64_t)(w->table.bytes() + bb.bytes() + w->hops.bytes()));
    w->st = st;
    return NWG_OK;
}

}  // namespace

int run(Work* w, const Job& J, std::string& err) {
    if (!J.pairs || J.pairs->empty()) return NWG_OK;
    if ((size_t)J.na * J.na * (J.precision ? 8 : 4) > 32768) {
        err = "nwgrad_amd: the Hirschberg kernels hold the substitution matrix in 32 KiB of local memory";
        return NWG_EINVAL;
    }
    if (J.precision == 0) return J.pmax ? run_t<float, true>(w, J, err) : run_t<float, false>(w, J, err);
    return J.pmax ? run_t<double, true>(w, J, err) : run_t<double, false>(w, J, err);
}

}  // namespace nwg_hb
