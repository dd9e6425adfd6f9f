// nwgrad_amd engine: the header-only C++20 core above the C ABI (include/nwgrad_hip.h).
//
// It mirrors the reference's drop-in surface for the batched hot path (relative to the
// reference's src/nwgrad/cpp/nwgrad/):
//   Alphabet      alphabet.hpp:33-198    (interned symbol sets, encode with validation)
//   SubstMatrix   subst_matrix.hpp:35-183 (one N x N block per set, row-major)
//   AlignParams   align_params.hpp:45-203 (matrix + go_a, ge_a, go_b, ge_b; + - * as a gradient)
//   SeqPairBatchT seq_pair_batch.hpp:56-192 and BatchEngine (batch_engine.hpp:244-3737):
//                 ctor, set_params, score_and_grad, scores, grads_into, weighted_grad,
//                 compute_grad, fill / soft settings
// with the same argument meanings and the same exception types.  All alignment work runs
// on the GPU through the C ABI; this layer only validates, encodes and converts.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "../../../include/nwgrad_hip.h"

namespace nwgrad_amd {

inline constexpr std::string_view kAAOrder = "ACDEFGHIKLMNPQRSTVWY";

// An ordered set of ASCII symbols; symbol k encodes as the byte k.  Alphabets are interned --
// get() returns one object per distinct symbol string for the life of the process -- so two
// alphabets are equal exactly when they are the same object, and params, matrices and batches
// compare alphabets by address.
class Alphabet {
public:
    static const Alphabet& get(std::string_view symbols) {
        static Interned table;
        return table.find_or_add(symbols);
    }
    int size() const noexcept { return static_cast<int>(symbols_.size()); }
    const std::string& symbols() const noexcept { return symbols_; }
    // the symbol's code, or -1 when c is not a symbol
    int index_of(char c) const noexcept {
        const uint8_t v = code_[static_cast<unsigned char>(c)];
        return v == kNotASymbol ? -1 : static_cast<int>(v);
    }
    // Codes of s into out (s.size() bytes).  The whole string is translated first; a string
    // with a character outside the alphabet then raises, naming the first such character.
    void encode_into(std::string_view s, uint8_t* out) const {
        uint8_t miss = 0;
        for (size_t k = 0; k < s.size(); ++k) {
            out[k] = code_[static_cast<unsigned char>(s[k])];
            miss |= static_cast<uint8_t>(out[k] == kNotASymbol);
        }
        if (!miss) return;
        size_t at = 0;
        while (out[at] != kNotASymbol) ++at;
        char msg[96];
        const auto u = static_cast<unsigned char>(s[at]);
        std::snprintf(msg, sizeof msg, "nwgrad: character '%c' (0x%02x) at position %zu", s[at], u, at);
        throw std::invalid_argument(std::string(msg) + " is not in alphabet \"" + symbols_ + "\"");
    }
    Alphabet(const Alphabet&) = delete;
    Alphabet& operator=(const Alphabet&) = delete;

private:
    static constexpr uint8_t kNotASymbol = 0xff;

    // The process's alphabets, each built once; lookups under a mutex (alphabets are few and
    // looked up once per params or batch, never per residue).
    class Interned {
    public:
        const Alphabet& find_or_add(std::string_view symbols) {
            std::lock_guard<std::mutex> hold(mu_);
            for (const auto& a : made_)
                if (a->symbols_ == symbols) return *a;
            made_.emplace_back(new Alphabet(symbols));
            return *made_.back();
        }

    private:
        std::mutex mu_;
        std::vector<std::unique_ptr<const Alphabet>> made_;
    };

    // The reason `symbols` cannot be an alphabet, or an empty string when it can.  One pass in
    // symbol order; the first offending symbol decides the message.
    static std::string rejection(std::string_view symbols) {
        if (symbols.empty()) return "nwgrad: alphabet must be non-empty";
        if (symbols.size() > 127) return "nwgrad: alphabet may have at most 127 symbols";
        std::array<bool, 128> seen{};
        for (char ch : symbols) {
            const auto u = static_cast<unsigned char>(ch);
            if (u > 127)
                return "nwgrad: alphabet symbols must be ASCII; got a byte >= 0x80 (a non-ASCII character encoded "
                       "as UTF-8?)";
            if (ch == '-')
                return "nwgrad: '-' cannot be an alphabet symbol; it is reserved as the gap marker in aligned "
                       "sequences";
            if (seen[u]) return "nwgrad: duplicate symbol '" + std::string(1, ch) + "' in alphabet \"" +
                                std::string(symbols) + "\"";
            seen[u] = true;
        }
        return {};
    }

    explicit Alphabet(std::string_view symbols) : symbols_(symbols) {
        const std::string why = rejection(symbols);
        if (!why.empty()) throw std::invalid_argument(why);
        code_.fill(kNotASymbol);
        for (size_t k = 0; k < symbols_.size(); ++k) code_[static_cast<unsigned char>(symbols_[k])] = static_cast<uint8_t>(k);
    }

    std::string symbols_;
    std::array<uint8_t, 256> code_{};
};

// An N x N substitution matrix over an interned alphabet, row-major (subst_matrix.hpp:29-33).
class SubstMatrix {
public:
    // nsets blocks of N x N, back to back (subst_matrix.hpp:29-33: set k at k*N*N).
    SubstMatrix(const double* src, const Alphabet& a, int nsets = 1)
        : alpha_(&a), sets_(nsets), blk_(src, src + static_cast<size_t>(nsets) * a.size() * a.size()) {}
    explicit SubstMatrix(const Alphabet& a, int nsets = 1)
        : alpha_(&a), sets_(nsets), blk_(static_cast<size_t>(nsets) * a.size() * a.size(), 0.0) {}
    const Alphabet& alphabet() const noexcept { return *alpha_; }
    int size() const noexcept { return alpha_->size(); }
    int sets() const noexcept { return sets_; }
    // set k alone, as a one-set matrix
    SubstMatrix slice(int k) const {
        return SubstMatrix(blk_.data() + static_cast<size_t>(k) * size() * size(), *alpha_);
    }
    const double* data() const noexcept { return blk_.data(); }
    double* data() noexcept { return blk_.data(); }
    double at(int i, int j) const { return blk_[static_cast<size_t>(i * size() + j)]; }
    double score(char a, char b) const {
        const int ia = alpha_->index_of(a), ib = alpha_->index_of(b);
        if (ia < 0 || ib < 0) {
            const char bad = ia < 0 ? a : b;
            throw std::invalid_argument(std::string("nwgrad: character '") + bad + "' is not in alphabet \"" +
                                        alpha_->symbols() + "\"");
        }
        return at(ia, ib);
    }
    void check_same(const SubstMatrix& o) const {
        if (alpha_ != o.alpha_)
            throw std::invalid_argument("nwgrad: cannot combine matrices over different alphabets (\"" +
                                        alpha_->symbols() + "\" vs \"" + o.alpha_->symbols() + "\")");
        if (sets_ != o.sets_)
            throw std::invalid_argument("nwgrad: cannot combine params with " + std::to_string(sets_) + " and " +
                                        std::to_string(o.sets_) + " parameter sets");
    }
    SubstMatrix& operator+=(const SubstMatrix& o) {
        check_same(o);
        for (size_t k = 0; k < blk_.size(); ++k) blk_[k] += o.blk_[k];
        return *this;
    }
    SubstMatrix& operator*=(double s) noexcept {
        for (auto& v : blk_) v *= s;
        return *this;
    }

private:
    const Alphabet* alpha_;
    int sets_ = 1;
    std::vector<double> blk_;
};

// Parameter point and gradient (align_params.hpp:10-43: gap fields of a gradient are <= 0),
// with K parameter sets (align_params.hpp:55-152): the matrix holds K blocks, the plain gap
// fields are set 0's and `more_gaps` holds sets 1..K-1 (go_a, ge_a, go_b, ge_b each).
struct AlignParams {
    SubstMatrix matrix;
    double gap_open_a = 0.0, gap_extend_a = 0.0, gap_open_b = 0.0, gap_extend_b = 0.0;
    std::vector<double> more_gaps;

    enum : int { kOpenA = 0, kExtendA = 1, kOpenB = 2, kExtendB = 3 };

    explicit AlignParams(const SubstMatrix& m, double goa = 0, double gea = 0, double gob = 0, double geb = 0)
        : matrix(m), gap_open_a(goa), gap_extend_a(gea), gap_open_b(gob), gap_extend_b(geb),
          more_gaps(static_cast<size_t>(4 * (m.sets() - 1)), 0.0) {}
    // K sets: `stack` the K matrices, `gaps` K rows of go_a, ge_a, go_b, ge_b
    AlignParams(const SubstMatrix& stack, const double* gaps)
        : matrix(stack), gap_open_a(gaps[0]), gap_extend_a(gaps[1]), gap_open_b(gaps[2]), gap_extend_b(gaps[3]),
          more_gaps(gaps + 4, gaps + 4 * stack.sets()) {}
    explicit AlignParams(const Alphabet& a, int nsets = 1)
        : matrix(a, nsets), more_gaps(static_cast<size_t>(4 * (nsets - 1)), 0.0) {}

    const Alphabet& alphabet() const noexcept { return matrix.alphabet(); }
    int num_sets() const noexcept { return matrix.sets(); }
    void gaps4(double* g) const { g[0] = gap_open_a; g[1] = gap_extend_a; g[2] = gap_open_b; g[3] = gap_extend_b; }
    double gap(int k, int f) const { return const_cast<AlignParams*>(this)->gap_ref(k, f); }
    double& gap_ref(int k, int f) {
        if (k == 0) return f == kOpenA ? gap_open_a : f == kExtendA ? gap_extend_a : f == kOpenB ? gap_open_b
                                                                                                : gap_extend_b;
        return more_gaps[static_cast<size_t>(4 * (k - 1) + f)];
    }
    // every set's gap fields, K rows of 4
    std::vector<double> all_gaps() const {
        std::vector<double> g(4);
        gaps4(g.data());
        g.insert(g.end(), more_gaps.begin(), more_gaps.end());
        return g;
    }
    // set k as a one-set AlignParams
    AlignParams set(int k) const {
        return AlignParams(matrix.slice(k), gap(k, 0), gap(k, 1), gap(k, 2), gap(k, 3));
    }
    void assign_set(int k, const AlignParams& p) {
        if (&p.alphabet() != &alphabet())
            throw std::invalid_argument("nwgrad: the set's alphabet does not match the params'");
        if (p.num_sets() != 1) throw std::invalid_argument("nwgrad: assign one parameter set at a time");
        const size_t nn = static_cast<size_t>(matrix.size()) * static_cast<size_t>(matrix.size());
        std::memcpy(matrix.data() + static_cast<size_t>(k) * nn, p.matrix.data(), nn * sizeof(double));
        for (int f = 0; f < 4; ++f) gap_ref(k, f) = p.gap(0, f);
    }
    // the sets of `parts`, in order, as one params (align_params.hpp:83-96)
    static AlignParams stack(const std::vector<AlignParams>& parts) {
        if (parts.empty()) throw std::invalid_argument("nwgrad: stack() needs at least one AlignParams");
        const Alphabet& a = parts[0].alphabet();
        int K = 0;
        for (const AlignParams& p : parts) {
            if (&p.alphabet() != &a) throw std::invalid_argument("nwgrad: stack() needs params over one alphabet");
            K += p.num_sets();
        }
        AlignParams r(a, K);
        int k = 0;
        for (const AlignParams& p : parts)
            for (int s = 0; s < p.num_sets(); ++s) r.assign_set(k++, p.set(s));
        return r;
    }

    AlignParams& operator+=(const AlignParams& o) {
        matrix += o.matrix;
        gap_open_a += o.gap_open_a; gap_extend_a += o.gap_extend_a;
        gap_open_b += o.gap_open_b; gap_extend_b += o.gap_extend_b;
        for (size_t k = 0; k < more_gaps.size(); ++k) more_gaps[k] += o.more_gaps[k];
        return *this;
    }
    AlignParams& operator*=(double s) noexcept {
        matrix *= s;
        gap_open_a *= s; gap_extend_a *= s; gap_open_b *= s; gap_extend_b *= s;
        for (auto& v : more_gaps) v *= s;
        return *this;
    }
    AlignParams operator+(const AlignParams& o) const { AlignParams r(*this); r += o; return r; }
    AlignParams operator*(double s) const { AlignParams r(*this); r *= s; return r; }
    AlignParams operator-() const { return *this * -1.0; }
    AlignParams operator-(const AlignParams& o) const { return *this + (-1.0 * o); }
    AlignParams& operator-=(const AlignParams& o) { return *this += (-1.0 * o); }
    friend AlignParams operator*(double s, const AlignParams& p) { return p * s; }
};

// The guide of an alignment given as two gapped strings (the reference's guide_j_from_aligned,
// aligner.hpp:159-180): entry i is the column reached once i residues of A are consumed; a
// column of two residues or a gap in B consumes one of A, a gap in A moves the column only.
inline std::vector<int> guide_from_aligned(std::string_view a, std::string_view b) {
    if (a.size() != b.size()) throw std::invalid_argument("aligned strings must have equal length");
    std::vector<int> g{0};
    int col = 0;
    for (size_t c = 0; c < a.size(); ++c) {
        const bool ga = a[c] == '-', gb = b[c] == '-';
        if (ga && gb) throw std::invalid_argument("aligned strings have a column with two gaps");
        if (!gb) ++col;
        if (!ga) g.push_back(col);
    }
    return g;
}

enum class GapModel { Linear = 0, Affine = 1 };
enum class AlignMode { Global = 0, Local = 1, Glocal = 2 };
enum class GradMode { None = 0, Hard = 1, Soft = 2 };

// C ABI status -> the reference's exception types (py_exports / batch_engine conventions).
[[noreturn]] inline void throw_status(int rc, const char* msg) {
    const std::string m = msg ? msg : "nwgrad_hip: error";
    switch (rc) {
        case NWG_EINVAL: throw std::invalid_argument(m);
        case NWG_ESTATE: throw std::logic_error(m);
        case NWG_ERANGE: throw std::domain_error(m);
        case NWG_EINDEX: throw std::out_of_range(m);
        default: throw std::runtime_error(m);
    }
}

// One context per process and device, created on first use.
class Device {
public:
    static nwg_ctx* ctx(int device = 0) {
        static std::mutex mu;
        static std::map<int, nwg_ctx*> ctxs;
        std::lock_guard<std::mutex> lock(mu);
        auto it = ctxs.find(device);
        if (it != ctxs.end()) return it->second;
        nwg_ctx* c = nullptr;
        const int rc = nwg_init(device, &c);
        if (rc != NWG_OK)
            throw std::runtime_error("nwgrad_amd: no usable HIP device " + std::to_string(device) +
                                     " (nwg_init status " + std::to_string(rc) +
                                     "); this build runs its alignment kernels on an MI355X (gfx950) only");
        ctxs[device] = c;
        return c;
    }
};

// What a streaming align() returns (batch_result.hpp:69-77): one score per pair and the gradient
// summed over them.
struct BatchResult {
    std::vector<double> scores;
    AlignParams grad;
};

// SeqPairBatchT<T>: T = float (SeqPairBatch) or double (SeqPairBatchDouble).
template <class T>
class SeqPairBatchT {
public:
    SeqPairBatchT(const std::vector<std::string_view>& seqs_a, const std::vector<std::string_view>& seqs_b,
                  const AlignParams& params, GapModel gm, AlignMode am, GradMode gd, int n_threads = 0,
                  int device = 0)
        : alpha_(&params.alphabet()), params_(params), gm_(gm), am_(am), gd_(gd), n_threads_(n_threads) {
        if (n_threads < 0) throw std::invalid_argument("nwgrad: n_threads must be >= 0");
        if (seqs_a.size() != seqs_b.size())
            throw std::invalid_argument("nwgrad: seqs_a and seqs_b must have the same length");
        const size_t N = seqs_a.size();
        std::vector<int64_t> oa(N + 1, 0), ob(N + 1, 0);
        for (size_t i = 0; i < N; ++i) {
            oa[i + 1] = oa[i] + static_cast<int64_t>(seqs_a[i].size());
            ob[i + 1] = ob[i] + static_cast<int64_t>(seqs_b[i].size());
        }
        std::vector<uint8_t> ca(static_cast<size_t>(oa[N])), cb(static_cast<size_t>(ob[N]));
        for (size_t i = 0; i < N; ++i) {
            alpha_->encode_into(seqs_a[i], ca.data() + oa[i]);
            alpha_->encode_into(seqs_b[i], cb.data() + ob[i]);
        }
        create_(ca.data(), oa.data(), cb.data(), ob.data(), static_cast<int64_t>(N), device);
    }
    // Pre-encoded pairs (codes in [0, alphabet size)).
    SeqPairBatchT(const uint8_t* ca, const int64_t* oa, const uint8_t* cb, const int64_t* ob, int64_t N,
                  const AlignParams& params, GapModel gm, AlignMode am, GradMode gd, int device = 0)
        : alpha_(&params.alphabet()), params_(params), gm_(gm), am_(am), gd_(gd) {
        create_(ca, oa, cb, ob, N, device);
    }
    SeqPairBatchT(const SeqPairBatchT&) = delete;
    SeqPairBatchT& operator=(const SeqPairBatchT&) = delete;
    ~SeqPairBatchT() {
        if (b_) nwg_batch_free(b_);
    }

    size_t size() const noexcept { return static_cast<size_t>(N_); }
    const Alphabet& alphabet() const {
        if (N_ == 0) throw std::logic_error("nwgrad: SeqPairBatch is empty; it has no alphabet");
        return *alpha_;
    }
    GapModel gap_model() const noexcept { return gm_; }
    AlignMode align_mode() const noexcept { return am_; }
    GradMode grad_mode() const noexcept { return gd_; }
    int n_threads() const noexcept { return n_threads_; }
    void set_n_threads(int v) {
        if (v < 0) throw std::invalid_argument("nwgrad: n_threads must be >= 0");
        n_threads_ = v;
    }

    // batch_engine.hpp:597-606: a copy of params, over the batch's alphabet.  Clears scores,
    // gradients and stored paths; guides stay (banded_grad() re-aligns around them).
    void set_params(const AlignParams& p) {
        if (&p.alphabet() != alpha_ && N_ > 0)
            throw std::invalid_argument("nwgrad: params alphabet \"" + p.alphabet().symbols() +
                                        "\" does not match the batch's \"" + alpha_->symbols() + "\"");
        AlignParams old = params_;
        params_ = p;
        try {
            upload_params_();
        } catch (...) {
            params_ = old;
            throw;
        }
        stored_ = false;
        paths_cached_ = false;
    }
    const AlignParams& params() const noexcept { return params_; }

    // fill: "interpair" = the batch's tiles in one launch, most work first (short hard / score
    // tiles one per block in dispatch order; soft and long-pair work from device queues on a
    // persistent grid); "striped" / "rowwise" = plain per-bucket launches.  Results are the
    // same whichever runs.
    void set_fill(const std::string& f) {
        int v;
        if (f == "interpair") v = 0;
        else if (f == "striped" || f == "rowwise") v = 1;
        else throw std::invalid_argument("nwgrad: unknown fill \"" + f +
                                         "\" (expected \"interpair\", \"striped\" or \"rowwise\")");
        check_(nwg_batch_set_fill(b_, v));
        fill_ = f;
    }
    const std::string& fill() const noexcept { return fill_; }
    // Columns a lane of the hard / score-only tiles: 0 chosen per batch (the lane-width table,
    // else the built-in rule), 8 or 16 forced; results do not depend on it.
    void set_columns_per_lane(int c) { check_(nwg_batch_set_lanes(b_, c)); }
    int columns_per_lane() const {
        int c = 0;
        check_(nwg_batch_lanes(b_, &c));
        return c;
    }

    void set_soft_temperature(double t) {
        if (!(t > 0.0)) throw std::invalid_argument("nwgrad: soft_temperature must be > 0");
        soft_temp_ = t;
        check_(nwg_batch_set_soft(b_, soft_temp_, soft_f32_ ? 0 : 1));
    }
    double soft_temperature() const noexcept { return soft_temp_; }
    void set_soft_precision(const std::string& v) {
        if (v == "float32") {
            if (!std::is_same<T, float>::value)
                throw std::invalid_argument("nwgrad: soft_precision=\"float32\" needs a float32 SeqPair / "
                                            "SeqPairBatch (the *Double classes run double)");
            soft_f32_ = true;
        } else if (v == "double") {
            soft_f32_ = false;
        } else {
            throw std::invalid_argument("nwgrad: soft_precision must be \"float32\" or \"double\", got \"" + v + "\"");
        }
        check_(nwg_batch_set_soft(b_, soft_temp_, soft_f32_ ? 0 : 1));
    }
    std::string soft_precision() const { return soft_f32_ ? "float32" : "double"; }

    // batch_engine.hpp:615-629: every pair's score (and gradient) and guide; the sum in pair
    // order.  keep_paths: also each pair's path (aligned(), coordinates(), cigar()).  Hard
    // batches write guides and paths in their own pass; score-only and soft batches take them
    // from a hard pass over the same pairs under the params scored -- the reference's eager
    // guide, the same hard fill -- run here when paths are kept, else on the first guide read.
    double score_and_grad(bool keep_paths = false) {
        double total = 0.0;
        if (gd_ == GradMode::Hard) check_(nwg_batch_set_keep_paths(b_, keep_paths ? 1 : 0));
        gbscores_ = false;
        check_(nwg_batch_run(b_, &total));
        valid_ = true;
        gparams_ = params_;
        gfresh_ = false;
        paths_cached_ = false;
        stored_ = keep_paths;
        guides_ = true;
        if (keep_paths && gd_ != GradMode::Hard) resolve_guides_(true);
        return total;
    }
    // batch_engine.hpp:633-653: a banded re-alignment and gradient around every pair's guide,
    // the training-loop step after set_params; new guides.  Hard batches run it on their own
    // device batch.  A score-only batch's banded pass is the banded Viterbi score with its guide
    // and path (batch_engine.hpp:2188-2205, 2125-2134), which is the companion hard batch's
    // banded pass around the same guides: its scores are this batch's until the next pass.  A
    // soft batch's is the banded forward-backward around the guides (score and gradient, on its
    // own device batch), its new guides and paths the banded Viterbi's, the companion's.
    double banded_grad(int bandwidth, bool keep_paths = false) {
        if (N_ == 0) return 0.0;
        if (bandwidth <= 0)
            throw std::invalid_argument("nwgrad: banded_grad() needs bandwidth > 0 (got " + std::to_string(bandwidth) +
                                        "); use score_and_grad() for full DP");
        if (!guides_) throw std::logic_error("nwgrad: banded_grad() needs a guide path; run score_and_grad() first");
        double total = 0.0;
        if (gd_ == GradMode::Soft) {
            if (!gfresh_) resolve_guides_(false);   // the guides of the last pass, under its params
            const std::vector<int32_t> g = gb_->all_guides_();
            check_(nwg_batch_set_guides(b_, g.data()));
            gbscores_ = false;
            check_(nwg_batch_run_banded(b_, bandwidth, &total));
            gb_->set_params(params_);
            gb_->banded_grad(bandwidth, keep_paths);   // the new guides (and paths): the banded Viterbi's
        } else if (gd_ == GradMode::None) {
            if (!gfresh_) resolve_guides_(false);   // the guides of the last pass, under its params
            gb_->set_params(params_);
            total = gb_->banded_grad(bandwidth, keep_paths);
            gbscores_ = true;
        } else {
            check_(nwg_batch_set_keep_paths(b_, keep_paths ? 1 : 0));
            check_(nwg_batch_run_banded(b_, bandwidth, &total));
        }
        valid_ = true;
        gparams_ = params_;
        paths_cached_ = false;
        stored_ = keep_paths;
        return total;
    }
    // Free the stored paths; scores, gradients and guides stay.
    void drop_paths() {
        if (gd_ == GradMode::Hard) check_(nwg_batch_drop_paths(b_));
        if (gb_) gb_->drop_paths();
        stored_ = false;
        paths_cached_ = false;
        std::vector<uint32_t>().swap(pops_);
    }
    bool path_valid() const noexcept { return guides_; }
    bool stored_path_valid() const noexcept { return stored_; }
    // Pair i's guide (the column after each row of A, m + 1 entries).
    std::vector<int> guide_j(size_t i) {
        if (i >= size()) throw std::out_of_range("SeqPairBatch index out of range");
        if (!guides_) throw std::logic_error("nwgrad: alignment not computed; call align() first");
        std::vector<int> g(static_cast<size_t>(len_a(i) + 1));
        if (gd_ == GradMode::Hard) {
            check_(nwg_batch_pair_guide(b_, static_cast<int64_t>(i), g.data()));
            return g;
        }
        if (!gfresh_) resolve_guides_(false);
        return gb_->guide_j(i);
    }
    // Stored paths (keep_paths), rendered as the reference renders them (batch_engine.hpp:
    // 513-590): gapped strings, Biopython-style coordinates and SAM CIGARs (seq_a the query).
    std::pair<std::string, std::string> aligned(size_t i) const {
        const PathView v = path_(i);
        const std::string& sym = alpha_->symbols();
        std::string a, b;
        a.reserve(static_cast<size_t>(v.cols));
        b.reserve(static_cast<size_t>(v.cols));
        const uint8_t* ca = codes_a(i);
        const uint8_t* cb = codes_b(i);
        int64_t x = v.i0, y = v.j0;
        for (int64_t q = 0; q < v.n; ++q) {
            const uint32_t t = v.ops[q] >> 30, len = v.ops[q] & 0x3fffffffu;
            for (uint32_t k = 0; k < len; ++k) {
                a.push_back(t == 2 ? '-' : sym[ca[x++]]);
                b.push_back(t == 1 ? '-' : sym[cb[y++]]);
            }
        }
        return {std::move(a), std::move(b)};
    }
    std::pair<std::vector<int64_t>, std::vector<int64_t>> coordinates(size_t i) const {
        const PathView v = path_(i);
        std::vector<int64_t> ci{v.i0}, cj{v.j0};
        int64_t x = v.i0, y = v.j0;
        for (int64_t q = 0; q < v.n; ++q) {
            const uint32_t t = v.ops[q] >> 30, len = v.ops[q] & 0x3fffffffu;
            x += t != 2 ? len : 0;
            y += t != 1 ? len : 0;
            ci.push_back(x);
            cj.push_back(y);
        }
        if (v.n == 0) {   // an empty path still has its end column
            ci.push_back(x);
            cj.push_back(y);
        }
        return {std::move(ci), std::move(cj)};
    }
    std::string cigar(size_t i, bool extended = false, bool soft_clip = false) const {
        const PathView v = path_(i);
        const uint8_t* ca = codes_a(i);
        const uint8_t* cb = codes_b(i);
        std::string out;
        char cur = 0;
        uint64_t run = 0;
        auto put = [&](char op, uint64_t cnt) {
            if (cnt == 0) return;
            if (op != cur && run) out += std::to_string(run) + cur;
            if (op != cur) { cur = op; run = 0; }
            run += cnt;
        };
        if (soft_clip) put('S', static_cast<uint64_t>(v.i0));
        int64_t x = v.i0, y = v.j0;
        for (int64_t q = 0; q < v.n; ++q) {
            const uint32_t t = v.ops[q] >> 30, len = v.ops[q] & 0x3fffffffu;
            if (t == 1) { put('I', len); x += len; }
            else if (t == 2) { put('D', len); y += len; }
            else if (!extended) { put('M', len); x += len; y += len; }
            else for (uint32_t k = 0; k < len; ++k, ++x, ++y) put(ca[x] == cb[y] ? '=' : 'X', 1);
        }
        if (soft_clip) put('S', static_cast<uint64_t>(len_a(i) - x));
        if (run) out += std::to_string(run) + cur;
        return out;
    }
    std::vector<std::string> cigars(bool extended = false, bool soft_clip = false) const {
        std::vector<std::string> out(size());
        for (size_t i = 0; i < size(); ++i) out[i] = cigar(i, extended, soft_clip);
        return out;
    }
    // Enqueue only (no total, no synchronisation): for device-side pipelines.
    void score_and_grad_async() {
        gbscores_ = false;
        check_(nwg_batch_run_async(b_));
        valid_ = true;
    }

    std::vector<double> scores() const {
        std::vector<double> out(static_cast<size_t>(N_));
        if (!valid_) throw std::logic_error("nwgrad: score not valid; compute an alignment first");
        if (gbscores_) return gb_->scores();
        check_(nwg_batch_scores(b_, out.data()));
        return out;
    }
    void grads_into(double* mats, double* gaps) const {
        alphabet();
        if (gd_ == GradMode::None) throw std::logic_error("nwgrad: grad_mode \"none\" computes no gradient");
        if (!valid_) throw std::logic_error("nwgrad: gradient not valid; compute an alignment first");
        check_(nwg_batch_grads(b_, mats, gaps));
    }
    AlignParams weighted_grad(const double* w, size_t n) const {
        const Alphabet& a = alphabet();
        if (n != size())
            throw std::invalid_argument("nwgrad: weighted_grad() needs one weight per pair (got " + std::to_string(n) +
                                        " weights for " + std::to_string(size()) + " pairs)");
        return reduce_(a, w);
    }
    AlignParams compute_grad() const { return reduce_(alphabet(), nullptr); }

    // Tracks (batch_engine.hpp:284-350): `tracks` one parameter set per residue of A, pair p at
    // off_a[p] .. off_a[p+1]; `gap_tracks` one per DP row 0..m, pair p at off_a[p] + p; either may
    // be null (the reference's defaults).  The batch is then tracked: grads() has num_sets() sets.
    void set_tracks(const uint16_t* tracks, const uint16_t* gap_tracks) {
        check_(nwg_batch_set_tracks(b_, tracks, gap_tracks));
        const int64_t la = oa_.empty() ? 0 : oa_.back();
        if (tracks) trk_.assign(tracks, tracks + la);
        else trk_.assign(static_cast<size_t>(la), 0);
        gtrk_.resize(static_cast<size_t>(la + N_));
        for (int64_t p = 0; p < N_; ++p) {   // the reference's default gap track where none is given
            const int64_t a0 = oa_[static_cast<size_t>(p)], m = oa_[static_cast<size_t>(p) + 1] - a0;
            uint16_t* g = gtrk_.data() + a0 + p;
            if (gap_tracks) {
                std::copy(gap_tracks + a0 + p, gap_tracks + a0 + p + m + 1, g);
            } else {
                g[0] = m ? trk_[static_cast<size_t>(a0)] : 0;
                for (int64_t i = 1; i <= m; ++i) g[i] = trk_[static_cast<size_t>(a0 + i - 1)];
            }
        }
        valid_ = false;
        gbscores_ = false;
        gb_.reset();   // the companion hard batch is rebuilt with these tracks when next needed
        gfresh_ = false;
    }
    int num_sets() const {
        int K = 1, tr = 0;
        check_(nwg_batch_num_sets(b_, &K, &tr));
        return K;
    }
    bool tracked() const {
        int K = 1, tr = 0;
        check_(nwg_batch_num_sets(b_, &K, &tr));
        return tr != 0;
    }

    // Pair i (the reference's SeqPairT view, seq_pair.hpp): its sequences, its tracks, and its
    // cached results.
    std::string seq_a(size_t i) const { return decode_(ca_, oa_, i); }
    std::string seq_b(size_t i) const { return decode_(cb_, ob_, i); }
    int64_t len_a(size_t i) const { return oa_[i + 1] - oa_[i]; }
    int64_t len_b(size_t i) const { return ob_[i + 1] - ob_[i]; }
    const uint8_t* codes_a(size_t i) const { return ca_.data() + oa_[i]; }
    const uint8_t* codes_b(size_t i) const { return cb_.data() + ob_[i]; }
    std::vector<uint16_t> track(size_t i) const {
        if (trk_.empty() && len_a(i) > 0) return std::vector<uint16_t>(static_cast<size_t>(len_a(i)), 0);
        return std::vector<uint16_t>(trk_.begin() + oa_[i], trk_.begin() + oa_[i + 1]);
    }
    std::vector<uint16_t> gap_track(size_t i) const {
        if (gtrk_.empty()) return std::vector<uint16_t>(static_cast<size_t>(len_a(i) + 1), 0);
        const auto at = gtrk_.begin() + oa_[i] + static_cast<int64_t>(i);
        return std::vector<uint16_t>(at, at + len_a(i) + 1);
    }
    bool results_valid() const noexcept { return valid_; }
    double pair_score(size_t i) const {
        double v = 0.0;
        if (gbscores_) return gb_->pair_score(i);
        check_(nwg_batch_pair_result(b_, static_cast<int64_t>(i), &v, nullptr));
        return v;
    }
    AlignParams pair_grad(size_t i) const {
        const int nn = alpha_->size() * alpha_->size(), K = num_sets();
        std::vector<double> row(static_cast<size_t>(K) * (nn + 4));
        check_(nwg_batch_pair_result(b_, static_cast<int64_t>(i), nullptr, row.data()));
        return AlignParams(SubstMatrix(row.data(), *alpha_, K), row.data() + static_cast<size_t>(K) * nn);
    }
    // Soft: the float32 refusals the batch's passes left (pair-passes, first pass, passes).
    std::array<int64_t, 3> soft_refusals(bool reset) {
        std::array<int64_t, 3> r{0, 0, 0};
        check_(nwg_batch_soft_refusals(b_, &r[0], &r[1], &r[2], reset ? 1 : 0));
        return r;
    }

    // Device-side plumbing for autograd and serving.
    nwg_batch* handle() const noexcept { return b_; }
    const double* scores_dev() const {
        const double* p = nullptr;
        check_(nwg_batch_scores_dev(b_, &p));
        return p;
    }
    void weighted_grad_dev(const double* w_dev, double* out_dev) const {
        if (gd_ == GradMode::None) throw std::logic_error("nwgrad: grad_mode \"none\" computes no gradient");
        check_(nwg_batch_weighted_grad_dev(b_, w_dev, out_dev));
    }
    // Params from a device buffer (matrix, go_a, ge_a, go_b, ge_b as doubles), enqueued on the
    // device's stream: soft batches only.
    void set_params_dev(const double* theta_dev) { check_(nwg_batch_set_params_dev(b_, theta_dev)); }
    // Device-to-device copy of the scores into a caller buffer of size() doubles.
    void scores_copy_dev(double* dst_dev) const { check_(nwg_batch_scores_copy_dev(b_, dst_dev)); }
    // After score_and_grad_async: rerun refused float32 pairs in double; throws as
    // score_and_grad would when a pair fits neither.
    void soft_finish() { check_(nwg_batch_soft_finish(b_)); }
    // Work of this device's context goes to `stream` (a hipStream_t, e.g. torch's current one).
    void set_stream(void* stream) { check_(nwg_set_stream(ctx_, stream)); }
    double last_kernel_ms() const {
        double ms = 0.0;
        check_(nwg_batch_last_kernel_ms(b_, &ms));
        return ms;
    }
    int64_t cells() const {
        int64_t t = 0, c = 0;
        check_(nwg_batch_plan_info(b_, &t, &c));
        return c;
    }
    // grid, wide-queue blocks, resident blocks per CU of the last soft launch
    // Soft scratch: device memory the soft pass may take for stored forward rows (0: default).
    void set_soft_scratch(int64_t bytes) { check_(nwg_batch_set_soft_scratch(b_, bytes)); }
    // Pointer-code memory of the hard fill (0: default); its launches, whether fill 0 runs
    // persistent, the last fill's persistent grid (0: plain).
    void set_table_budget(int64_t bytes) { check_(nwg_batch_set_table_budget(b_, bytes)); }
    // traceback: 0 auto, 1 pointers, 2 scores, 3 hirschberg, 4 hirschberg_pmax (the
    // reference's TracebackMode; auto, pointers, scores and grid all give Pointers' result)
    void set_traceback(int traceback, int hb_cutoff) { check_(nwg_batch_set_traceback(b_, traceback, hb_cutoff)); }
    // the last hard pass's recursion: pairs that split, levels, base blocks, scratch bytes,
    // cells filled by its sweeps and base cases
    std::array<int64_t, 5> hb_info() const {
        std::array<int64_t, 5> r{};
        check_(nwg_batch_hb_info(b_, &r[0], &r[1], &r[2], &r[3], &r[4]));
        return r;
    }
    std::array<int64_t, 3> hard_launches() const {
        std::array<int64_t, 3> r{0, 0, 0};
        check_(nwg_batch_hard_launches(b_, &r[0], &r[1], &r[2]));
        return r;
    }
    std::array<int64_t, 4> wide_info() const {
        std::array<int64_t, 4> r{0, 0, 0, 0};
        check_(nwg_batch_wide_info(b_, &r[0], &r[1], &r[2], &r[3]));
        return r;
    }
    // Totals formed so far from the device's exact sum / in pair order on the host.
    std::array<int64_t, 2> total_info() const {
        std::array<int64_t, 2> r{0, 0};
        check_(nwg_batch_total_info(b_, &r[0], &r[1]));
        return r;
    }
    std::array<int64_t, 3> launch_info() const {
        std::array<int64_t, 3> r{0, 0, 0};
        check_(nwg_batch_launch_info(b_, &r[0], &r[1], &r[2]));
        return r;
    }
    int64_t soft_reruns() const {
        int64_t n = 0;
        check_(nwg_batch_soft_reruns(b_, &n));
        return n;
    }
    int64_t tiles() const {
        int64_t t = 0, c = 0;
        check_(nwg_batch_plan_info(b_, &t, &c));
        return t;
    }

    // Streaming align() (batch_engine.hpp:748-1273): the pairs (seqs_a[i], seqs_b[i]) under
    // `params`, this batch's problem type and settings, nothing kept: every score and the
    // gradient summed over the pairs (zero for grad_mode "none").  band > 0 (hard batches)
    // bands every pair around its guide -- `guides` (one per pair, from aligned strings) or
    // the diagonal round(i * n / m); guides with band 0 band at width 0.
    BatchResult align(const std::vector<std::string_view>& seqs_a, const std::vector<std::string_view>& seqs_b,
                       const AlignParams& params, int band, const std::vector<std::vector<int>>& guides,
                       const uint16_t* tracks, const uint16_t* gap_tracks) const {
        if (seqs_a.size() != seqs_b.size()) throw std::invalid_argument("nwgrad: align() needs seqs_a and seqs_b of equal length");
        if (N_ > 0 && &params.alphabet() != alpha_)
            throw std::invalid_argument("nwgrad: align() params over alphabet \"" + params.alphabet().symbols() +
                                        "\" in a batch over alphabet \"" + alpha_->symbols() + "\"");
        if (band < 0) throw std::invalid_argument("nwgrad: band must be >= 0, got " + std::to_string(band));
        const bool banded = band > 0 || !guides.empty();
        if (!guides.empty() && guides.size() != seqs_a.size())
            throw std::invalid_argument("nwgrad: aligned_a / aligned_b need one entry per pair");
        BatchResult r{std::vector<double>(seqs_a.size(), 0.0), AlignParams(params.alphabet(), params.num_sets())};
        if (seqs_a.empty()) return r;
        // a banded score-only pass is the banded Viterbi's scores: a hard batch's, gradient unread
        SeqPairBatchT<T> tmp(seqs_a, seqs_b, params, gm_, am_, banded && gd_ == GradMode::None ? GradMode::Hard : gd_,
                             n_threads_, device_);
        if (tracks || gap_tracks) tmp.set_tracks(tracks, gap_tracks);
        if (gd_ == GradMode::Soft) {
            tmp.set_soft_temperature(soft_temp_);
            tmp.set_soft_precision(soft_precision());
        }
        if (banded) {
            std::vector<int32_t> g;
            for (size_t i = 0; i < seqs_a.size(); ++i) {
                const int m = static_cast<int>(seqs_a[i].size()), n = static_cast<int>(seqs_b[i].size());
                if (!guides.empty()) {
                    if (guides[i].size() != static_cast<size_t>(m) + 1)
                        throw std::invalid_argument("nwgrad: guide_j has " + std::to_string(guides[i].size()) +
                                                    " entries, but sequence A of length " + std::to_string(m) +
                                                    " needs " + std::to_string(m + 1) +
                                                    " (do the aligned strings spell the input sequences?)");
                    g.insert(g.end(), guides[i].begin(), guides[i].end());
                } else {   // the diagonal (aligner.hpp:1781-1790)
                    for (int k = 0; k <= m; ++k)
                        g.push_back(m > 0 ? static_cast<int32_t>(std::llround(static_cast<double>(k) * n / m)) : 0);
                }
            }
            tmp.check_(nwg_batch_set_guides(tmp.b_, g.data()));
            double total = 0.0;
            tmp.check_(nwg_batch_run_banded(tmp.b_, band, &total));
            tmp.valid_ = true;
        } else {
            tmp.score_and_grad();
        }
        r.scores = tmp.scores();
        if (gd_ != GradMode::None) r.grad = tmp.compute_grad();
        return r;
    }

private:
    // The stored path of pair i: its ops (in path order), their columns, and the start cell (the
    // end cell less the path's moves, batch_engine.hpp:1441-1449).
    struct PathView {
        const uint32_t* ops;
        int64_t n, cols, i0, j0;
    };
    PathView path_(size_t i) const {
        if (i >= size()) throw std::out_of_range("SeqPairBatch index out of range");
        if (!guides_) throw std::logic_error("nwgrad: alignment not computed; call align() first");
        if (!stored_)
            throw std::logic_error(
                "nwgrad: stored path unavailable; use align(keep_paths=True) or banded_align(keep_paths=True)");
        if (gd_ != GradMode::Hard) return gb_->path_(i);
        if (!paths_cached_) {
            int64_t cap = oa_.back() + ob_.back();
            pops_.assign(static_cast<size_t>(std::max<int64_t>(cap, 1)), 0u);
            pn_.assign(static_cast<size_t>(std::max<int64_t>(N_, 1)), 0);
            pend_.assign(static_cast<size_t>(std::max<int64_t>(2 * N_, 1)), 0);
            check_(nwg_batch_paths(b_, pops_.data(), pn_.data(), pend_.data()));
            paths_cached_ = true;
        }
        PathView v{};
        const int64_t end = oa_[i + 1] + ob_[i + 1];
        v.n = pn_[i];
        v.ops = pops_.data() + end - v.n;
        int64_t x = pend_[2 * i], y = pend_[2 * i + 1], cols = 0;
        for (int64_t q = 0; q < v.n; ++q) {
            const uint32_t t = v.ops[q] >> 30, len = v.ops[q] & 0x3fffffffu;
            x -= t != 2 ? len : 0;
            y -= t != 1 ? len : 0;
            cols += len;
        }
        v.i0 = x;
        v.j0 = y;
        v.cols = cols;
        return v;
    }
    // The guides (and, keep, the paths) of a score-only or soft batch: a hard pass over the same
    // pairs under the params of the last full pass.
    void resolve_guides_(bool keep) {
        if (!gb_) {
            gb_ = std::make_unique<SeqPairBatchT<T>>(ca_.data(), oa_.data(), cb_.data(), ob_.data(), N_, gparams_, gm_,
                                                     am_, GradMode::Hard, device_);
            if (!trk_.empty()) gb_->set_tracks(trk_.data(), gtrk_.data());
        } else {
            gb_->set_params(gparams_);
        }
        gb_->score_and_grad(keep);
        gfresh_ = true;
    }
    // Every pair's guide, pair p's m + 1 entries at off_a[p] + p (a hard batch's).
    std::vector<int32_t> all_guides_() const {
        std::vector<int32_t> g(static_cast<size_t>(std::max<int64_t>(oa_.back() + N_, 1)));
        check_(nwg_batch_guides(b_, g.data()));
        return g;
    }
    std::string decode_(const std::vector<uint8_t>& codes, const std::vector<int64_t>& off, size_t i) const {
        if (i >= static_cast<size_t>(N_)) throw std::out_of_range("SeqPairBatch index out of range");
        const std::string& sym = alpha_->symbols();
        std::string out(static_cast<size_t>(off[i + 1] - off[i]), '\0');
        for (size_t k = 0; k < out.size(); ++k) out[k] = sym[codes[static_cast<size_t>(off[i]) + k]];
        return out;
    }
    void create_(const uint8_t* ca, const int64_t* oa, const uint8_t* cb, const int64_t* ob, int64_t N, int device) {
        N_ = N;
        device_ = device;
        // host copies: the views' sequences and tracks (the reference keeps its codes on the host too)
        oa_.assign(oa, oa + N + 1);
        ob_.assign(ob, ob + N + 1);
        ca_.assign(ca, ca + oa_.back());
        cb_.assign(cb, cb + ob_.back());
        ctx_ = Device::ctx(device);
        nwg_problem pr{};
        pr.precision = std::is_same<T, float>::value ? 0 : 1;
        pr.gap_model = static_cast<int32_t>(gm_);
        pr.mode = static_cast<int32_t>(am_);
        pr.grad_mode = static_cast<int32_t>(gd_);
        soft_f32_ = std::is_same<T, float>::value;
        check_(nwg_batch_create(ctx_, &pr, alpha_->size(), N, ca, oa, cb, ob, &b_));
        check_(nwg_batch_set_soft(b_, soft_temp_, soft_f32_ ? 0 : 1));
        upload_params_();
    }
    void upload_params_() {
        const std::vector<double> g = params_.all_gaps();
        check_(nwg_batch_set_params_k(b_, params_.num_sets(), params_.matrix.data(), g.data()));
        valid_ = false;
        gbscores_ = false;
    }
    AlignParams reduce_(const Alphabet& a, const double* w) const {
        if (gd_ == GradMode::None) throw std::logic_error("nwgrad: grad_mode \"none\" computes no gradient");
        if (!valid_) throw std::logic_error("nwgrad: gradient not valid; compute an alignment first");
        const int nn = a.size() * a.size(), K = num_sets();
        std::vector<double> out(static_cast<size_t>(K) * (nn + 4));
        check_(nwg_batch_weighted_grad(b_, w, w ? static_cast<int64_t>(size()) : 0, out.data()));
        return AlignParams(SubstMatrix(out.data(), a, K), out.data() + static_cast<size_t>(K) * nn);
    }
    void check_(int rc) const {
        if (rc != NWG_OK) throw_status(rc, nwg_last_error(ctx_));
    }

    const Alphabet* alpha_;
    std::vector<uint8_t> ca_, cb_;
    std::vector<int64_t> oa_, ob_;
    std::vector<uint16_t> trk_, gtrk_;   // tracked: per residue of A / per DP row (pair p at oa_[p] + p)
    AlignParams params_;
    GapModel gm_;
    AlignMode am_;
    GradMode gd_;
    int n_threads_ = 0;
    int64_t N_ = 0;
    nwg_ctx* ctx_ = nullptr;
    nwg_batch* b_ = nullptr;
    bool valid_ = false;
    int device_ = 0;
    // guides and stored paths: whether the last full pass left guides / kept paths; the host
    // copy of the paths; for score-only and soft batches the hard batch that holds them and the
    // params it must run under
    bool guides_ = false, stored_ = false, gfresh_ = false;
    bool gbscores_ = false;   // a score-only batch's banded scores: held by the companion
    mutable bool paths_cached_ = false;
    mutable std::vector<uint32_t> pops_;
    mutable std::vector<int32_t> pn_, pend_;
    std::unique_ptr<SeqPairBatchT<T>> gb_;
    AlignParams gparams_{Alphabet::get("A")};
    std::string fill_ = "interpair";
    double soft_temp_ = 1.0;
    bool soft_f32_ = true;
};

}  // namespace nwgrad_amd
