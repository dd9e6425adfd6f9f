// nwgrad_amd serving loop: continuous batching of alignment jobs over the C ABI.
//
// No counterpart in the reference, which is a synchronous library (one BatchEngine call at a
// time, batch_engine.hpp:2005-2014).  A job is a set of pairs of one kind -- score only, hard
// gradient or soft gradient -- under the server's params and problem type.  The server keeps
// `slots_per_kind` slots per kind; a slot is a context with its own HIP stream and one batch
// that is reloaded for every launch (nwg_batch_reload: device buffers grow, never shrink, so a
// warm slot allocates and frees nothing).  poll() is the whole scheduler:
//   1. every busy slot whose stream is idle is finished -- soft refusals rerun in double
//      (nwg_batch_soft_finish), scores and gradient rows are copied out and handed back to
//      the slot's jobs -- and its memory returns to the budget;
//   2. while a slot is free and jobs wait, the kind whose oldest job has waited longest is
//      served: its waiting jobs, in arrival order, are coalesced into one launch up to
//      max_cells_per_launch / max_pairs_per_launch, as long as the bytes in flight stay within
//      hbm_budget_bytes, and the launch is enqueued (nwg_batch_run_async) on the slot's stream.
// Jobs that arrive while others run are thus admitted as soon as a slot and memory are free,
// and launches of different slots overlap on the device.  A pair's result does not depend on
// the launch it lands in (hard and score-only: bit for bit; soft: the per-pair sums are
// formed within the pair's tile, so up to float rounding of a different lane split).
//
// Params may hold K parameter sets (AlignParams::stack): the server is then tracked, every job
// carries a set per residue of A (tracks) and per DP row (gap tracks; the reference's default
// when not given, aligner.hpp:1956-1967), and a pair's gradient row is K matrices, then K gap
// rows (the reference's grads_into layout, batch_engine.hpp:719-746).
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"

namespace nwgrad_amd {

enum class JobKind : int { Score = 0, Hard = 1, Soft = 2 };

struct ServeOptions {
    double hbm_budget_bytes = 200e9;            // device memory the jobs in flight may hold
    int slots_per_kind = 2;                     // concurrent launches per kind
    int64_t max_cells_per_launch = 4000000000;  // coalescing limit (a single larger job runs alone)
    int64_t max_pairs_per_launch = 1 << 20;
    double soft_temperature = 1.0;
};

class AlignServer {
public:
    struct Job {
        JobKind kind = JobKind::Score;
        int64_t n = 0, cells = 0;
        double bytes = 0.0;
        std::vector<uint8_t> ca, cb;
        std::vector<int64_t> oa, ob;
        std::vector<uint16_t> trk, gtrk;   // tracked server: per residue of A, per DP row (pair p at oa[p] + p)
        int state = 0;   // 0 waiting, 1 running, 2 done, 3 failed
        double t_submit = 0.0, t_launch = 0.0, t_done = 0.0;
        std::vector<double> scores, grads;   // grads: n rows of K * (n_alpha^2 + 4) (hard and soft)
        std::string error;
    };

    AlignServer(const AlignParams& params, GapModel gm, AlignMode am, bool double_precision, ServeOptions opt = {},
                int device = 0)
        : params_(params), gm_(gm), am_(am), f64_(double_precision), opt_(opt), device_(device) {
        if (opt_.slots_per_kind < 1) throw std::invalid_argument("nwgrad: slots_per_kind must be >= 1");
        if (!(opt_.hbm_budget_bytes > 0)) throw std::invalid_argument("nwgrad: hbm_budget_bytes must be > 0");
        if (!(opt_.soft_temperature > 0)) throw std::invalid_argument("nwgrad: soft_temperature must be > 0");
        na_ = static_cast<int>(params_.alphabet().size());
        K_ = params_.num_sets();
        t0_ = std::chrono::steady_clock::now();
        for (int k = 0; k < 3; ++k)
            for (int s = 0; s < opt_.slots_per_kind; ++s) {
                auto sl = std::make_unique<Slot>();
                sl->kind = static_cast<JobKind>(k);
                slots_.push_back(std::move(sl));
            }
    }
    AlignServer(const AlignServer&) = delete;
    AlignServer& operator=(const AlignServer&) = delete;
    ~AlignServer() {
        for (auto& sl : slots_) {
            if (sl->b) nwg_batch_free(sl->b);
            if (sl->ctx) nwg_destroy(sl->ctx);
        }
    }

    // Seconds since the server started (the clock of every job timestamp).
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    // Queue a job of pre-encoded pairs (codes in [0, n_alpha), offsets of n + 1 entries from 0).
    // A tracked server (K > 1 sets) takes `tracks`, one set per residue of A (oa[n] entries), and
    // `gap_tracks`, one per DP row (oa[n] + n entries, pair p's at oa[p] + p); either may be null
    // for the reference's defaults (every residue set 0; a gap track following the track).
    int64_t submit(JobKind kind, const uint8_t* ca, const int64_t* oa, const uint8_t* cb, const int64_t* ob,
                   int64_t n, const uint16_t* tracks = nullptr, const uint16_t* gap_tracks = nullptr) {
        if ((tracks || gap_tracks) && K_ == 1)
            throw std::invalid_argument("nwgrad: tracks need params with more than one parameter set");
        if (n < 0) throw std::invalid_argument("nwgrad: a job needs n >= 0 pairs");
        auto j = std::make_unique<Job>();
        j->kind = kind;
        j->n = n;
        if (n > 0 && (oa[0] != 0 || ob[0] != 0)) throw std::invalid_argument("nwgrad: job offsets must start at 0");
        const int64_t la = n ? oa[n] : 0, lb = n ? ob[n] : 0;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t m = oa[i + 1] - oa[i], w = ob[i + 1] - ob[i];
            if (m < 0 || w < 0) throw std::invalid_argument("nwgrad: job offsets must be non-decreasing");
            j->cells += m * w;
            if (kind == JobKind::Hard && w > 1024) {
                // the long-pair pointer table: 1 B per cell, panels of 1024 columns, 63 skew rows
                j->bytes += static_cast<double>((w + 1023) / 1024) * static_cast<double>(m + 63) * 1024.0;
            }
        }
        for (int64_t k = 0; k < la; ++k)
            if (ca[k] >= na_) throw std::invalid_argument("nwgrad: job code out of range in A");
        for (int64_t k = 0; k < lb; ++k)
            if (cb[k] >= na_) throw std::invalid_argument("nwgrad: job code out of range in B");
        const int gs = K_ * (na_ * na_ + 4);
        j->bytes += static_cast<double>(la + lb) + 24.0 * static_cast<double>(n) +
                    (kind == JobKind::Score ? 0.0 : 8.0 * gs * static_cast<double>(n));
        if (K_ > 1) {   // the job's tracks, defaults filled in
            j->trk.assign(static_cast<size_t>(la), 0);
            if (tracks) std::copy(tracks, tracks + la, j->trk.begin());
            j->gtrk.assign(static_cast<size_t>(la + n), 0);
            for (int64_t p = 0; p < n; ++p) {
                const int64_t a0 = oa[p], m = oa[p + 1] - a0;
                uint16_t* g = j->gtrk.data() + a0 + p;
                if (gap_tracks) {
                    std::copy(gap_tracks + a0 + p, gap_tracks + a0 + p + m + 1, g);
                } else {
                    g[0] = m ? j->trk[static_cast<size_t>(a0)] : 0;
                    for (int64_t i = 1; i <= m; ++i) g[i] = j->trk[static_cast<size_t>(a0 + i - 1)];
                }
            }
            for (uint16_t v : j->trk)
                if (v >= K_) throw std::invalid_argument("nwgrad: a track names a set outside [0, K)");
            for (uint16_t v : j->gtrk)
                if (v >= K_) throw std::invalid_argument("nwgrad: a gap track names a set outside [0, K)");
            j->bytes += 2.0 * static_cast<double>(2 * la + n) + 8.0 * static_cast<double>(la + n);
        }
        j->ca.assign(ca, ca + la);
        j->cb.assign(cb, cb + lb);
        j->oa.assign(oa, oa + n + 1);
        j->ob.assign(ob, ob + n + 1);
        if (n == 0) {
            j->oa.assign(1, 0);
            j->ob.assign(1, 0);
        }
        j->t_submit = now();
        const int64_t id = next_id_++;
        waiting_[static_cast<int>(kind)].push_back(id);
        jobs_[id] = std::move(j);
        return id;
    }

    // One scheduler pass (see the file comment); the number of jobs that completed in it.
    int poll() {
        int completed = 0;
        for (auto& sl : slots_)
            if (sl->busy) completed += try_finish_(*sl);
        for (;;) {
            int best = -1;
            double oldest = 0.0;
            for (int k = 0; k < 3; ++k) {
                if (waiting_[k].empty() || !free_slot_(static_cast<JobKind>(k))) continue;
                const double t = jobs_.at(waiting_[k].front())->t_submit;
                if (best < 0 || t < oldest) {
                    best = k;
                    oldest = t;
                }
            }
            if (best < 0 || !launch_(static_cast<JobKind>(best))) break;
        }
        return completed;
    }

    const Job& job(int64_t id) const {
        auto it = jobs_.find(id);
        if (it == jobs_.end()) throw std::out_of_range("nwgrad: no job " + std::to_string(id));
        return *it->second;
    }
    void release(int64_t id) {
        auto it = jobs_.find(id);
        if (it == jobs_.end()) throw std::out_of_range("nwgrad: no job " + std::to_string(id));
        if (it->second->state < 2) throw std::logic_error("nwgrad: job " + std::to_string(id) + " has not finished");
        jobs_.erase(it);
    }
    int64_t waiting() const { return (int64_t)(waiting_[0].size() + waiting_[1].size() + waiting_[2].size()); }
    int64_t running_launches() const {
        int64_t r = 0;
        for (auto& sl : slots_) r += sl->busy ? 1 : 0;
        return r;
    }
    double bytes_in_flight() const { return in_flight_bytes_; }
    int64_t launches() const { return n_launches_; }
    int n_alpha() const { return na_; }
    int num_sets() const { return K_; }

private:
    struct Slot {
        JobKind kind = JobKind::Score;
        nwg_ctx* ctx = nullptr;
        nwg_batch* b = nullptr;
        bool busy = false;
        std::vector<int64_t> jobs;
        double bytes = 0.0;
    };

    Slot* free_slot_(JobKind k) {
        for (auto& sl : slots_)
            if (sl->kind == k && !sl->busy) return sl.get();
        return nullptr;
    }

    void check_(Slot& sl, int rc) {
        if (rc != NWG_OK) throw_status(rc, sl.ctx ? nwg_last_error(sl.ctx) : "nwgrad_hip: no context");
    }

    bool launch_(JobKind k) {
        std::deque<int64_t>& q = waiting_[static_cast<int>(k)];
        Slot* sl = free_slot_(k);
        if (!sl || q.empty()) return false;
        std::vector<int64_t> take;
        int64_t cells = 0, pairs = 0;
        double bytes = 0.0;
        for (int64_t id : q) {
            const Job& j = *jobs_.at(id);
            const bool first = take.empty();
            if (!first && (cells + j.cells > opt_.max_cells_per_launch || pairs + j.n > opt_.max_pairs_per_launch))
                break;
            if (in_flight_bytes_ + bytes + j.bytes > opt_.hbm_budget_bytes && !(first && in_flight_bytes_ == 0.0))
                break;   // wait for memory (a job larger than the budget runs when nothing else does)
            take.push_back(id);
            cells += j.cells;
            pairs += j.n;
            bytes += j.bytes;
        }
        if (take.empty()) return false;
        q.erase(q.begin(), q.begin() + static_cast<std::ptrdiff_t>(take.size()));
        // concatenate the jobs' pairs (and tracks: concatenation keeps the per-row layout, pair P's
        // gap track at oa[P] + P)
        std::vector<uint8_t> ca, cb;
        std::vector<int64_t> oa(1, 0), ob(1, 0);
        std::vector<uint16_t> trk, gtrk;
        oa.reserve((size_t)pairs + 1);
        ob.reserve((size_t)pairs + 1);
        for (int64_t id : take) {
            const Job& j = *jobs_.at(id);
            const int64_t ba = oa.back(), bb = ob.back();
            ca.insert(ca.end(), j.ca.begin(), j.ca.end());
            cb.insert(cb.end(), j.cb.begin(), j.cb.end());
            trk.insert(trk.end(), j.trk.begin(), j.trk.end());
            gtrk.insert(gtrk.end(), j.gtrk.begin(), j.gtrk.end());
            for (int64_t i = 1; i <= j.n; ++i) {
                oa.push_back(ba + j.oa[(size_t)i]);
                ob.push_back(bb + j.ob[(size_t)i]);
            }
        }
        const double t = now();
        try {
            prepare_(*sl, ca, oa, cb, ob, pairs, trk, gtrk);
            check_(*sl, nwg_batch_run_async(sl->b));
        } catch (const std::exception& e) {
            for (int64_t id : take) {
                Job& j = *jobs_.at(id);
                j.state = 3;
                j.error = e.what();
                j.t_launch = j.t_done = t;
            }
            return true;
        }
        for (int64_t id : take) {
            Job& j = *jobs_.at(id);
            j.state = 1;
            j.t_launch = t;
        }
        sl->busy = true;
        sl->jobs = std::move(take);
        sl->bytes = bytes;
        in_flight_bytes_ += bytes;
        ++n_launches_;
        return true;
    }

    void prepare_(Slot& sl, const std::vector<uint8_t>& ca, const std::vector<int64_t>& oa,
                  const std::vector<uint8_t>& cb, const std::vector<int64_t>& ob, int64_t n,
                  const std::vector<uint16_t>& trk, const std::vector<uint16_t>& gtrk) {
        const uint8_t* pa = ca.empty() ? &kZero_ : ca.data();
        const uint8_t* pb = cb.empty() ? &kZero_ : cb.data();
        if (!sl.ctx) {
            nwg_ctx* c = nullptr;
            const int rc = nwg_init(device_, &c);
            if (rc != NWG_OK) throw std::runtime_error("nwgrad_amd: no usable HIP device for the server");
            sl.ctx = c;
            check_(sl, nwg_ctx_own_stream(sl.ctx));
        }
        if (!sl.b) {
            nwg_problem pr{f64_ ? 1 : 0, static_cast<int32_t>(gm_), static_cast<int32_t>(am_),
                           static_cast<int32_t>(sl.kind)};
            check_(sl, nwg_batch_create(sl.ctx, &pr, na_, n, pa, oa.data(), pb, ob.data(), &sl.b));
            const std::vector<double> gaps = params_.all_gaps();
            check_(sl, nwg_batch_set_params_k(sl.b, K_, params_.matrix.data(), gaps.data()));
            if (sl.kind == JobKind::Soft) check_(sl, nwg_batch_set_soft(sl.b, opt_.soft_temperature, f64_ ? 1 : 0));
        } else {
            check_(sl, nwg_batch_reload(sl.b, n, pa, oa.data(), pb, ob.data()));
        }
        if (K_ > 1) {
            static const uint16_t kNone = 0;
            check_(sl, nwg_batch_set_tracks(sl.b, trk.empty() ? &kNone : trk.data(), gtrk.empty() ? &kNone : gtrk.data()));
        }
    }

    int try_finish_(Slot& sl) {
        int done = 0;
        check_(sl, nwg_batch_query(sl.b, &done));
        if (!done) return 0;
        const double t = now();
        int64_t total = 0;
        for (int64_t id : sl.jobs) total += jobs_.at(id)->n;
        std::vector<double> scores((size_t)total), grads;
        const int nn = na_ * na_, knn = K_ * nn, gs = K_ * (nn + 4);
        std::string err;
        try {
            if (sl.kind == JobKind::Soft) check_(sl, nwg_batch_soft_finish(sl.b));
            if (total > 0) {
                check_(sl, nwg_batch_scores(sl.b, scores.data()));
                if (sl.kind != JobKind::Score) {
                    std::vector<double> mats((size_t)total * knn), gaps((size_t)total * 4 * K_);
                    check_(sl, nwg_batch_grads(sl.b, mats.data(), gaps.data()));
                    grads.resize((size_t)total * gs);
                    for (int64_t i = 0; i < total; ++i) {
                        std::copy(mats.begin() + i * knn, mats.begin() + (i + 1) * knn, grads.begin() + i * gs);
                        std::copy(gaps.begin() + i * 4 * K_, gaps.begin() + (i + 1) * 4 * K_,
                                  grads.begin() + i * gs + knn);
                    }
                }
            }
        } catch (const std::exception& e) {
            err = e.what();
        }
        int64_t at = 0;
        for (int64_t id : sl.jobs) {
            Job& j = *jobs_.at(id);
            j.t_done = t;
            if (!err.empty()) {
                j.state = 3;
                j.error = err;
            } else {
                j.state = 2;
                j.scores.assign(scores.begin() + at, scores.begin() + at + j.n);
                if (!grads.empty()) j.grads.assign(grads.begin() + at * gs, grads.begin() + (at + j.n) * gs);
            }
            std::vector<uint8_t>().swap(j.ca);
            std::vector<uint8_t>().swap(j.cb);
            std::vector<uint16_t>().swap(j.trk);
            std::vector<uint16_t>().swap(j.gtrk);
            at += j.n;
        }
        in_flight_bytes_ -= sl.bytes;
        if (in_flight_bytes_ < 0.5) in_flight_bytes_ = 0.0;
        const int n = static_cast<int>(sl.jobs.size());
        sl.jobs.clear();
        sl.bytes = 0.0;
        sl.busy = false;
        return n;
    }

    static constexpr uint8_t kZero_ = 0;
    AlignParams params_;
    GapModel gm_;
    AlignMode am_;
    bool f64_;
    ServeOptions opt_;
    int device_;
    int na_ = 0;
    int K_ = 1;
    std::chrono::steady_clock::time_point t0_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::deque<int64_t> waiting_[3];
    std::map<int64_t, std::unique_ptr<Job>> jobs_;
    int64_t next_id_ = 0;
    int64_t n_launches_ = 0;
    double in_flight_bytes_ = 0.0;
};

}  // namespace nwgrad_amd
