// TEST INFRASTRUCTURE ONLY.  A C-ABI driver over the reference nwgrad engine
// (SeqPairBatchT<T>, seq_pair_batch.hpp), compiled from the reference's own sources by
// oracle/Makefile.ref into oracle/_ref/libnwgref.so.  It is the second oracle (next to the
// restatement in oracle/nwg_oracle.c) and the CPU baseline of bench.py.  Nothing in the
// product (nwgrad_amd/) links, loads or calls it.
//
// The calls mirror the reference's Python surface (py_exports.cpp:787-1263):
//   SeqPairBatch(seqs_a, seqs_b, params, n_threads=, traceback=, gap_model=, mode=,
//                grad_mode=)  -> nwgref_create
//   .fill / .soft_temperature / .soft_precision / .soft_impl        -> nwgref_set_*
//   .score_and_grad() (soft: nwgrad_soft_score_and_grad)             -> nwgref_score_and_grad
//   .scores() / .grads() / .weighted_grad(w) / .compute_grad()       -> nwgref_scores / ...
//   .set_params(p)                                                   -> nwgref_set_params
// Exceptions are mapped to status codes: 1 invalid_argument (ValueError), 2 logic_error
// (RuntimeError), 3 domain_error (ValueError), 4 out_of_range (IndexError), 5 other.
#include "seq_pair_batch.hpp"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace {

thread_local std::string g_err;

struct Handle {
    int dbl = 0;
    std::string chars_a, chars_b;
    std::unique_ptr<SeqPairBatchT<float>> f;
    std::unique_ptr<SeqPairBatchT<double>> d;
    int grad_mode = 1;
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::domain_error& e) {
        g_err = e.what();
        return 3;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::out_of_range& e) {
        g_err = e.what();
        return 4;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return 2;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 5;
    }
}

AlignParams make_params(const char* alphabet, const double* matrix, const double* gaps4) {
    const Alphabet& al = Alphabet::get(alphabet);
    return AlignParams(SubstMatrix(matrix, al), gaps4[0], gaps4[1], gaps4[2], gaps4[3]);
}

template <class B, class F>
auto with_batch(Handle* h, F&& f) {
    if (h->dbl) return f(*h->d);
    return f(*h->f);
}

}  // namespace

extern "C" {

const char* nwgref_last_error(void) { return g_err.c_str(); }

// gap_model 0 linear / 1 affine; mode 0 global / 1 local / 2 glocal; grad_mode 0 none /
// 1 hard / 2 soft; traceback 0 auto / 1 pointers / 2 scores / 3 hirschberg /
// 4 hirschberg_pmax.  Sequences are given as characters, concatenated, with N+1 offsets.
int nwgref_create(int dbl, const char* alphabet, const double* matrix, const double* gaps4,
                  long n_pairs, const char* seq_a, const long* off_a, const char* seq_b,
                  const long* off_b, int gap_model, int mode, int grad_mode, int n_threads,
                  int traceback, void** out) {
    return guarded([&] {
        auto h = std::make_unique<Handle>();
        h->dbl = dbl;
        h->grad_mode = grad_mode;
        h->chars_a.assign(seq_a, static_cast<size_t>(off_a[n_pairs]));
        h->chars_b.assign(seq_b, static_cast<size_t>(off_b[n_pairs]));
        std::vector<std::string_view> va, vb;
        va.reserve(static_cast<size_t>(n_pairs));
        vb.reserve(static_cast<size_t>(n_pairs));
        for (long i = 0; i < n_pairs; ++i) {
            va.emplace_back(h->chars_a.data() + off_a[i], static_cast<size_t>(off_a[i + 1] - off_a[i]));
            vb.emplace_back(h->chars_b.data() + off_b[i], static_cast<size_t>(off_b[i + 1] - off_b[i]));
        }
        const AlignParams p = make_params(alphabet, matrix, gaps4);
        const GapModel gm = gap_model ? GapModel::Affine : GapModel::Linear;
        const AlignMode am = mode == 1 ? AlignMode::Local : mode == 2 ? AlignMode::Glocal : AlignMode::Global;
        const GradMode gd = grad_mode == 0 ? GradMode::None : grad_mode == 2 ? GradMode::Soft : GradMode::Hard;
        const TracebackMode tbs[] = {TracebackMode::Default, TracebackMode::Pointers, TracebackMode::Scores,
                                     TracebackMode::Hirschberg, TracebackMode::HirschbergPmax};
        const TracebackMode tb = tbs[(traceback >= 0 && traceback <= 4) ? traceback : 0];
        if (dbl)
            h->d = std::make_unique<SeqPairBatchT<double>>(va, vb, p, gm, am, gd, n_threads, tb);
        else
            h->f = std::make_unique<SeqPairBatchT<float>>(va, vb, p, gm, am, gd, n_threads, tb);
        *out = h.release();
    });
}

// The same with K parameter sets (matrices K*n*n, gaps K*4) and tracks: `tracks` one set per
// residue of A, pair i at off_a[i] .. off_a[i+1]; `gap_tracks` one per DP row, pair i at
// off_a[i] + i, m + 1 entries (NULL: the reference's default).  The reference's flat
// PairTracks form (batch_result.hpp:23-34).
int nwgref_create_tracked(int dbl, const char* alphabet, int K, const double* matrices, const double* gaps,
                          long n_pairs, const char* seq_a, const long* off_a, const char* seq_b, const long* off_b,
                          const uint16_t* tracks, const uint16_t* gap_tracks, int gap_model, int mode,
                          int grad_mode, int n_threads, void** out) {
    return guarded([&] {
        auto h = std::make_unique<Handle>();
        h->dbl = dbl;
        h->grad_mode = grad_mode;
        h->chars_a.assign(seq_a, static_cast<size_t>(off_a[n_pairs]));
        h->chars_b.assign(seq_b, static_cast<size_t>(off_b[n_pairs]));
        std::vector<std::string_view> va, vb;
        for (long i = 0; i < n_pairs; ++i) {
            va.emplace_back(h->chars_a.data() + off_a[i], static_cast<size_t>(off_a[i + 1] - off_a[i]));
            vb.emplace_back(h->chars_b.data() + off_b[i], static_cast<size_t>(off_b[i + 1] - off_b[i]));
        }
        std::vector<size_t> ot(static_cast<size_t>(n_pairs) + 1), og(static_cast<size_t>(n_pairs) + 1);
        for (long i = 0; i <= n_pairs; ++i) {
            ot[static_cast<size_t>(i)] = static_cast<size_t>(off_a[i]);
            og[static_cast<size_t>(i)] = static_cast<size_t>(off_a[i] + i);
        }
        const AlignParams p(SubstMatrix(matrices, Alphabet::get(alphabet), K), gaps);
        const GapModel gm = gap_model ? GapModel::Affine : GapModel::Linear;
        const AlignMode am = mode == 1 ? AlignMode::Local : mode == 2 ? AlignMode::Glocal : AlignMode::Global;
        const GradMode gd = grad_mode == 0 ? GradMode::None : grad_mode == 2 ? GradMode::Soft : GradMode::Hard;
        const PairTracks tk = tracks ? PairTracks(tracks, ot.data(), static_cast<size_t>(n_pairs)) : PairTracks();
        const PairTracks gk = gap_tracks ? PairTracks(gap_tracks, og.data(), static_cast<size_t>(n_pairs)) : PairTracks();
        if (dbl)
            h->d = std::make_unique<SeqPairBatchT<double>>(va, vb, p, gm, am, gd, n_threads, TracebackMode::Default,
                                                           kBackendAuto, tk, gk);
        else
            h->f = std::make_unique<SeqPairBatchT<float>>(va, vb, p, gm, am, gd, n_threads, TracebackMode::Default,
                                                          kBackendAuto, tk, gk);
        *out = h.release();
    });
}

void nwgref_destroy(void* hp) { delete static_cast<Handle*>(hp); }

// fill 0 interpair / 1 striped / 2 rowwise
int nwgref_set_fill(void* hp, int fill) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] { with_batch<void>(h, [&](auto& b) { b.set_fill(fill == 2, fill == 0); }); });
}

// Rows of the Hirschberg base case (SeqPairBatch.hb_cutoff, seq_pair_batch.hpp:153).
int nwgref_set_hb_cutoff(void* hp, int rows) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] { with_batch<void>(h, [&](auto& b) { b.set_hb_cutoff(rows); }); });
}

// precision 0 float32 / 1 double; impl 0 scaled / 1 scaled_or_log / 2 log
int nwgref_set_soft(void* hp, double temperature, int precision, int impl) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        with_batch<void>(h, [&](auto& b) {
            b.set_soft_temperature(temperature);
            b.set_soft_impl(impl == 2 ? SoftImpl::Log : impl == 1 ? SoftImpl::ScaledOrLog : SoftImpl::Scaled);
        });
        if (!h->dbl) h->f->set_soft_f32(precision == 0);
    });
}

int nwgref_set_params(void* hp, const char* alphabet, const double* matrix, const double* gaps4) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        const AlignParams p = make_params(alphabet, matrix, gaps4);
        with_batch<void>(h, [&](auto& b) { b.set_params(p); });
    });
}

int nwgref_score_and_grad(void* hp, double* total) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        *total = with_batch<double>(h, [&](auto& b) -> double {
            if (h->grad_mode == 2) return nwgrad_soft_score_and_grad(b, false);
            return b.score_and_grad(false);
        });
    });
}

int nwgref_scores(void* hp, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        const std::vector<double> s = with_batch<std::vector<double>>(h, [](auto& b) { return b.scores(); });
        std::memcpy(out, s.data(), s.size() * sizeof(double));
    });
}

// mats: N*K*n*n, gaps: N*K*4
int nwgref_grads(void* hp, double* mats, double* gaps) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] { with_batch<void>(h, [&](auto& b) { b.grads_into(mats, gaps); }); });
}

// K sets: the K matrices row-major, then K rows of go_a, ge_a, go_b, ge_b (K = 1: n*n + 4).
static void params_out(const AlignParams& g, double* out) {
    const size_t nn = static_cast<size_t>(g.matrix.size()) * static_cast<size_t>(g.matrix.size());
    const size_t K = static_cast<size_t>(g.num_sets());
    std::memcpy(out, g.matrix.data(), K * nn * sizeof(double));
    const std::vector<double> gaps = g.all_gaps();
    std::memcpy(out + K * nn, gaps.data(), gaps.size() * sizeof(double));
}

// out: K*(n*n + 4) doubles (see params_out)
int nwgref_weighted_grad(void* hp, const double* w, long n, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        with_batch<void>(h, [&](auto& b) { params_out(b.weighted_grad(w, static_cast<size_t>(n)), out); });
    });
}

int nwgref_compute_grad(void* hp, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] { with_batch<void>(h, [&](auto& b) { params_out(b.compute_grad(), out); }); });
}

// score_and_grad(keep_paths) / banded_grad(bandwidth, keep_paths) (batch_engine.hpp:615-653)
int nwgref_score_and_grad_keep(void* hp, int keep, double* total) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        *total = with_batch<double>(h, [&](auto& b) -> double {
            if (h->grad_mode == 2) return nwgrad_soft_score_and_grad(b, keep != 0);
            return b.score_and_grad(keep != 0);
        });
    });
}

int nwgref_banded_grad(void* hp, int bandwidth, int keep, double* total) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        *total = with_batch<double>(h, [&](auto& b) -> double { return b.banded_grad(bandwidth, keep != 0); });
    });
}

// Every pair's guide_j (batch[i].guide_j), concatenated: pair i's m + 1 entries at off_a[i] + i.
int nwgref_guides(void* hp, int32_t* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        with_batch<void>(h, [&](auto& b) {
            size_t at = 0;
            for (size_t i = 0; i < b.size(); ++i) {
                const std::vector<int> g = b[i].guide_j();
                for (int v : g) out[at++] = v;
            }
        });
    });
}

// Pair i's stored path as the view renders it: aligned() (two strings of `cols` characters into
// a and b, cap bytes each), coordinates() (2 x k int64 into coords, cap_k columns), cigar() and
// cigar(extended=True, soft_clip=True) (NUL-terminated, 8 * cap bytes each).
int nwgref_path(void* hp, long i, char* a, char* b, long cap, long* cols, int64_t* coords, long cap_k, long* k,
                char* cigar, char* cigar_ext) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        with_batch<void>(h, [&](auto& bt) {
            auto v = bt[static_cast<size_t>(i)];
            const auto al = v.aligned();
            if ((long)al.first.size() > cap) throw std::out_of_range("nwgref: aligned() longer than the buffer");
            std::memcpy(a, al.first.data(), al.first.size());
            std::memcpy(b, al.second.data(), al.second.size());
            *cols = static_cast<long>(al.first.size());
            const auto cc = v.coordinates();
            if ((long)cc.first.size() > cap_k) throw std::out_of_range("nwgref: coordinates() wider than the buffer");
            *k = static_cast<long>(cc.first.size());
            for (size_t q = 0; q < cc.first.size(); ++q) {
                coords[q] = cc.first[q];
                coords[cap_k + static_cast<long>(q)] = cc.second[q];
            }
            const std::string c0 = v.cigar(false, false), c1 = v.cigar(true, true);
            if ((long)c0.size() >= 8 * cap || (long)c1.size() >= 8 * cap) throw std::out_of_range("nwgref: cigar too long");
            std::memcpy(cigar, c0.c_str(), c0.size() + 1);
            std::memcpy(cigar_ext, c1.c_str(), c1.size() + 1);
        });
    });
}

// Streaming align() (batch_engine.hpp:748-1273) on the handle's problem type: pairs given as
// concatenated characters with offsets, params over the handle's alphabet; band 0 (the full DP).
// scores: n_pairs doubles; grad: n*n + 4 (the summed gradient, params_out's layout).
int nwgref_align(void* hp, const char* alphabet, const double* matrix, const double* gaps4, long n_pairs,
                 const char* seq_a, const long* off_a, const char* seq_b, const long* off_b, int band, double* scores,
                 double* grad) {
    Handle* h = static_cast<Handle*>(hp);
    return guarded([&] {
        std::vector<ProblemInstance> probs(static_cast<size_t>(n_pairs));
        for (long q = 0; q < n_pairs; ++q) {
            probs[static_cast<size_t>(q)].seq_a = std::string_view(seq_a + off_a[q], static_cast<size_t>(off_a[q + 1] - off_a[q]));
            probs[static_cast<size_t>(q)].seq_b = std::string_view(seq_b + off_b[q], static_cast<size_t>(off_b[q + 1] - off_b[q]));
        }
        const AlignParams p = make_params(alphabet, matrix, gaps4);
        with_batch<void>(h, [&](auto& b) {
            const BatchResult r = b.align(probs, p, band);
            std::memcpy(scores, r.scores.data(), r.scores.size() * sizeof(double));
            if (grad) params_out(r.grad, grad);
        });
    });
}

}  // extern "C"
