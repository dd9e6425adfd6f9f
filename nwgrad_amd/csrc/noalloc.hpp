// The no-allocation entry, nwg_hard_noalloc (include/nwgrad_hip.h): a hard or score-only pass
// that calls no hipMalloc / hipFree.  Included at the end of nwg_api.hip, whose planner
// (cut_tiles), kernel picker (hard_kernel_for) and helpers it uses.
//
// Its device memory is `scratch`, one array per device held for the life of the process.  On
// the host the array's name is only a stand-in, never a device address: host code passes the
// address hipGetSymbolAddress returns for it on each device -- looked up once per device, on
// the entry's first call there, and kept in g_scratch -- to kernels and copies.  Every call on
// a device shares that device's array, so calls that overlap on one device race.

#include <mutex>

static __device__ __align__(256) unsigned char scratch[size_t(64) << 20];

namespace {

constexpr size_t kNoAllocBytes = size_t(64) << 20;    // the size of `scratch`
constexpr size_t kNoAllocQueue = 256;                  // [0, 256): the persistent fill's tile counter
constexpr size_t kNoAllocSub = 64 * 64 * 8;            // then the matrix in the working precision
constexpr int64_t kNoAllocMaxPairs = int64_t(1) << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// This device's address of `scratch` (the context's device is current).
int scratch_address(nwg_ctx* ctx, void** out) {
    const int d = ctx->device;
    if (d < 0 || d >= kMaxDevices) return fail(ctx, NWG_EDEVICE, "nwgrad_hip: device %d out of range", d);
    std::lock_guard<std::mutex> hold(g_scratch_mu);
    if (!g_scratch_done[d]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[d] = e == hipSuccess ? p : nullptr;
        g_scratch_err[d] = e;
        g_scratch_done[d] = true;
    }
    if (!g_scratch[d])
        return fail(ctx, NWG_EDEVICE, "nwgrad_hip: hipGetSymbolAddress(scratch) failed on device %d: %s", d,
                    hipGetErrorString(g_scratch_err[d]));
    *out = g_scratch[d];
    return NWG_OK;
}

size_t round256(size_t x) { return (x + 255) & ~size_t(255); }

// One hard or score-only pass over short pairs already on the device, every buffer beyond the
// caller's inputs and outputs in this device's `scratch`: the tile counter, the matrix in the
// working precision, the plan (pair order and tiles) and, hard, the pointer-code slots of a
// persistent grid sized to what is left.  The launch is the batch path's fill kernel.
template <typename T>
int noalloc_run(nwg_ctx* ctx, void* scratch_ptr, const nwg_problem& prob, int na, int64_t N, const uint8_t* ca, const int64_t* oa,
                const uint8_t* cb, const int64_t* ob, const int64_t* hoa, const int64_t* hob, const double* matrix,
                const double* gaps4, double* scores, double* grads) {
    const bool hard = prob.grad_mode == 1;
    std::vector<int32_t> mlen((size_t)N), nlen((size_t)N), pairs((size_t)N);
    int widest = 0;
    for (int64_t i = 0; i < N; ++i) {
        const int64_t m = hoa[i + 1] - hoa[i], n = hob[i + 1] - hob[i];
        if (m < 0 || n < 0) return fail(ctx, NWG_EINVAL, "nwgrad_hip: offsets must be non-decreasing");
        if (n > kMaxN)
            return fail(ctx, NWG_EINVAL,
                        "nwgrad_hip: nwg_hard_noalloc takes pairs of at most %d columns (pair %lld has %lld); wider "
                        "pairs need the batch path's column panels", kMaxN, (long long)i, (long long)n);
        if (m > (1 << 24)) return fail(ctx, NWG_EINVAL, "nwgrad_hip: pair %lld is too long", (long long)i);
        mlen[(size_t)i] = (int32_t)m;
        nlen[(size_t)i] = (int32_t)n;
        pairs[(size_t)i] = (int32_t)i;
        widest = std::max<int>(widest, (int)n);
    }
    // histogram slots and the narrowest lane group, as an untracked batch of this problem has them
    const int per_slot = hard ? na * na * 2 : 0;
    int gmin = 1;
    while (gmin < kMaxG && (64 / gmin) * per_slot > kHistBytes) gmin <<= 1;
    const int slots = 64 / gmin;
    std::vector<int32_t> order;
    std::vector<TileDesc> tiles, ptiles;
    std::vector<Bucket> buckets;
    int64_t max_steps = 1;
    int panels = 1;
    cut_tiles(mlen, nlen, pairs, kS, gmin, true, order, tiles, buckets, ptiles, max_steps, panels);
    int S = kS;
    if (choose_lanes(nlen, pairs, (int64_t)tiles.size(), ctx->n_cu, 0) == kSN) {   // as a batch of these pairs
        S = kSN;
        cut_tiles(mlen, nlen, pairs, kSN, gmin, true, order, tiles, buckets, ptiles, max_steps, panels);
    }
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr);
    const size_t o_sub = kNoAllocQueue, o_order = o_sub + kNoAllocSub;
    const size_t o_tiles = o_order + round256(sizeof(int32_t) * std::max<size_t>(order.size(), 1));
    const size_t o_table = o_tiles + round256(sizeof(TileDesc) * std::max<size_t>(ptiles.size(), 1));
    if (o_table > kNoAllocBytes)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: the plan of %lld pairs does not fit the no-allocation scratch",
                    (long long)N);
    const int64_t slot_dwords = max_steps * (S / 4) * 64;
    const size_t slot_bytes = (size_t)slot_dwords * 4;
    const int64_t fit = (int64_t)((kNoAllocBytes - o_table) / slot_bytes);
    if (hard && fit < 1)
        return fail(ctx, NWG_EINVAL,
                    "nwgrad_hip: a tile of %lld row steps needs %.1f MiB of pointer codes; the no-allocation scratch "
                    "has %.1f MiB left after the plan", (long long)max_steps, slot_bytes / 1048576.0,
                    (kNoAllocBytes - o_table) / 1048576.0);
    void (*kern)(nwg::HardArgs<T>) = hard_kernel_for<T>(prob.gap_model, prob.mode, hard, false, S);
    if (!kern) return fail(ctx, NWG_EINVAL, "nwgrad_hip: unsupported problem type");
    const size_t lds = (size_t)nwg::hard_lds_sub_bytes<T>(na, 1, false) +
                       (hard ? (size_t)nwg::hard_lds_hist_bytes(na, 1, false, slots) : 0);
    std::vector<T> sub((size_t)na * na);
    for (size_t k = 0; k < sub.size(); ++k) sub[k] = static_cast<T>(matrix[k]);
    // uploads into the scratch (the host vectors live until the synchronisation below)
    NWG_HIP(ctx, hipMemcpyAsync(base + o_sub, sub.data(), sizeof(T) * sub.size(), hipMemcpyHostToDevice, ctx->stream));
    if (!order.empty())
        NWG_HIP(ctx, hipMemcpyAsync(base + o_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
    if (!ptiles.empty())
        NWG_HIP(ctx, hipMemcpyAsync(base + o_tiles, ptiles.data(), sizeof(TileDesc) * ptiles.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
    nwg::HardArgs<T> A{};
    A.codes_a = ca;
    A.codes_b = cb;
    A.off_a = oa;
    A.off_b = ob;
    A.order = reinterpret_cast<const int32_t*>(base + o_order);
    A.tiles = reinterpret_cast<const TileDesc*>(base + o_tiles);
    A.n_tiles = (int32_t)ptiles.size();
    A.tile_base = 0;
    A.sub = reinterpret_cast<const T*>(base + o_sub);
    A.na = na;
    A.hist_slots = slots;
    A.go_a = static_cast<T>(gaps4[0]);
    A.ge_a = static_cast<T>(gaps4[1]);
    A.go_b = static_cast<T>(gaps4[2]);
    A.ge_b = static_cast<T>(gaps4[3]);
    A.scores = scores;
    A.grads = grads;
    A.slot_dwords = slot_dwords;
    if (!ptiles.empty()) {
        int64_t grid = (int64_t)ptiles.size();
        if (hard) {   // a persistent grid: one pointer-code slot per wave, as many as fit
            A.table = reinterpret_cast<uint32_t*>(base + o_table);
            A.queue = reinterpret_cast<int32_t*>(base);
            grid = std::max<int64_t>(1, std::min<int64_t>({grid, resident_blocks(ctx, kern, lds, kBlocksPerCU), fit}));
            hipLaunchKernelGGL(nwg::clear_words_kernel, dim3(1), dim3(256), 0, ctx->stream, A.queue, (int64_t)1);
            NWG_HIP(ctx, hipGetLastError());
        }
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), lds, ctx->stream, A);
        NWG_HIP(ctx, hipGetLastError());
    }
    NWG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NWG_OK;
}

}  // namespace

extern "C" {

int nwg_hard_noalloc(nwg_ctx* ctx, const nwg_problem* prob, int n_alpha, int64_t n_pairs, const uint8_t* codes_a_dev,
                     const int64_t* off_a_dev, const uint8_t* codes_b_dev, const int64_t* off_b_dev,
                     const int64_t* off_a, const int64_t* off_b, const double* matrix, const double* gaps4,
                     double* scores_dev, double* grads_dev) {
    if (!ctx || !prob || !matrix || !gaps4) return NWG_EINVAL;
    if (prob->precision < 0 || prob->precision > 1 || prob->gap_model < 0 || prob->gap_model > 1 || prob->mode < 0 ||
        prob->mode > 2 || prob->grad_mode < 0 || prob->grad_mode > 1)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: nwg_hard_noalloc runs hard (grad_mode 1) and score-only (0) passes");
    if (n_alpha < 1 || n_alpha > 64)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: alphabet size %d outside [1, 64]", n_alpha);
    if (n_pairs < 0 || n_pairs > kNoAllocMaxPairs)
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: nwg_hard_noalloc takes at most %lld pairs (got %lld)",
                    (long long)kNoAllocMaxPairs, (long long)n_pairs);
    if (n_pairs > 0 && (!codes_a_dev || !codes_b_dev || !off_a_dev || !off_b_dev || !off_a || !off_b || !scores_dev ||
                        (prob->grad_mode == 1 && !grads_dev)))
        return fail(ctx, NWG_EINVAL, "nwgrad_hip: nwg_hard_noalloc: missing buffer");
    NWG_HIP(ctx, hipSetDevice(ctx->device));
    void* ptr = nullptr;
    const int rc = scratch_address(ctx, &ptr);
    if (rc) return rc;
    if (n_pairs == 0) return NWG_OK;
    return prob->precision == 0
               ? noalloc_run<float>(ctx, ptr, *prob, n_alpha, n_pairs, codes_a_dev, off_a_dev, codes_b_dev, off_b_dev, off_a,
                                    off_b, matrix, gaps4, scores_dev, grads_dev)
               : noalloc_run<double>(ctx, ptr, *prob, n_alpha, n_pairs, codes_a_dev, off_a_dev, codes_b_dev, off_b_dev,
                                     off_a, off_b, matrix, gaps4, scores_dev, grads_dev);
}

}  // extern "C"
