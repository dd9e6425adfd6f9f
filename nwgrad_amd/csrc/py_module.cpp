// Python bindings of the nwgrad_amd engine (pybind11; the reference binds the same
// surface with nanobind, py_exports.cpp:787-1263 and :1324-1535, which is not available
// in this image).  Names, argument meanings and exceptions follow the reference:
//   SeqPairBatch (float32) / SeqPairBatchDouble (float64), AlignParams, SubstMatrix.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <string>
#include <vector>

#include "../cpp/nwgrad_amd/engine.hpp"
#include "../cpp/nwgrad_amd/serving.hpp"

namespace py = pybind11;
using namespace nwgrad_amd;

using F64Arr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using U8Arr = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;
using I64Arr = py::array_t<int64_t, py::array::c_style | py::array::forcecast>;

namespace {

GapModel parse_gap_model(const std::string& s) {
    if (s == "linear") return GapModel::Linear;
    if (s == "affine") return GapModel::Affine;
    throw py::value_error("nwgrad: unknown gap_model \"" + s + "\" (expected \"linear\" or \"affine\")");
}
AlignMode parse_mode(const std::string& s) {
    if (s == "global") return AlignMode::Global;
    if (s == "local") return AlignMode::Local;
    if (s == "glocal") return AlignMode::Glocal;
    if (s == "overlap" || s == "semiglobal")
        throw py::value_error("nwgrad_amd: mode \"" + s + "\" (semi-global free ends) is not part of this build's "
                              "GPU path; use \"global\", \"local\" or \"glocal\"");
    throw py::value_error("nwgrad: unknown mode \"" + s +
                          "\" (expected \"global\", \"local\", \"glocal\", \"overlap\" or \"semiglobal\")");
}
GradMode parse_grad_mode(const std::string& s) {
    if (s == "hard") return GradMode::Hard;
    if (s == "soft") return GradMode::Soft;
    if (s == "none") return GradMode::None;
    throw py::value_error("nwgrad: unknown grad_mode \"" + s + "\" (expected \"hard\", \"soft\" or \"none\")");
}
// The reference's TracebackMode names (py_exports.cpp:355-366) as the ABI's codes: auto,
// pointers, scores and grid give Pointers' result (0-2); hirschberg 3, hirschberg_pmax 4.
int parse_traceback(const std::string& s) {
    if (s == "auto" || s == "grid") return 0;
    if (s == "pointers") return 1;
    if (s == "scores") return 2;
    if (s == "hirschberg") return 3;
    if (s == "hirschberg_pmax") return 4;
    throw py::value_error("nwgrad: unknown traceback \"" + s +
                          "\" (expected \"auto\", \"pointers\", \"scores\", \"hirschberg\", \"hirschberg_pmax\" or "
                          "\"grid\")");
}
JobKind parse_kind(const std::string& s) {
    if (s == "score" || s == "none") return JobKind::Score;
    if (s == "hard") return JobKind::Hard;
    if (s == "soft") return JobKind::Soft;
    throw py::value_error("nwgrad_amd: unknown job kind \"" + s + "\" (expected \"score\", \"hard\" or \"soft\")");
}

void bind_server(py::module_& m) {
    py::class_<AlignServer>(m, "AlignServer")
        .def(py::init([](const AlignParams& p, const std::string& gap_model, const std::string& mode, bool dbl,
                         double hbm_budget_bytes, int slots_per_kind, int64_t max_cells, int64_t max_pairs,
                         double soft_temperature, int device) {
                 ServeOptions o;
                 o.hbm_budget_bytes = hbm_budget_bytes;
                 o.slots_per_kind = slots_per_kind;
                 o.max_cells_per_launch = max_cells;
                 o.max_pairs_per_launch = max_pairs;
                 o.soft_temperature = soft_temperature;
                 return std::make_unique<AlignServer>(p, parse_gap_model(gap_model), parse_mode(mode), dbl, o, device);
             }),
             py::arg("params"), py::kw_only(), py::arg("gap_model") = "affine", py::arg("mode") = "global",
             py::arg("double") = false, py::arg("hbm_budget_bytes") = 200e9, py::arg("slots_per_kind") = 2,
             py::arg("max_cells_per_launch") = int64_t(4000000000), py::arg("max_pairs_per_launch") = int64_t(1) << 20,
             py::arg("soft_temperature") = 1.0, py::arg("device") = 0)
        .def("submit",
             [](AlignServer& s, const std::string& kind, U8Arr ca, I64Arr oa, U8Arr cb, I64Arr ob, py::object tracks,
                py::object gap_tracks) {
                 if (oa.ndim() != 1 || ob.ndim() != 1 || oa.size() < 1 || oa.size() != ob.size())
                     throw py::value_error("nwgrad_amd: off_a and off_b need n + 1 entries each");
                 const int64_t n = static_cast<int64_t>(oa.size()) - 1;
                 if (oa.data()[n] > static_cast<int64_t>(ca.size()) || ob.data()[n] > static_cast<int64_t>(cb.size()))
                     throw py::value_error("nwgrad_amd: offsets run past the codes");
                 // flat tracks: oa[n] entries (one per residue of A), gap tracks oa[n] + n (one per DP row)
                 auto flat = [&](const py::object& o, int64_t want, const char* what) {
                     std::vector<uint16_t> v;
                     if (o.is_none()) return v;
                     auto a = py::array_t<int64_t, py::array::c_style | py::array::forcecast>::ensure(o);
                     if (!a || a.ndim() != 1 || a.shape(0) != want)
                         throw py::value_error(std::string("nwgrad_amd: ") + what + " must be a flat integer array of " +
                                               std::to_string(want) + " entries");
                     v.resize(static_cast<size_t>(want));
                     for (int64_t k = 0; k < want; ++k) {
                         const int64_t x = a.data()[k];
                         if (x < 0 || x > 65535)
                             throw py::value_error(std::string("nwgrad_amd: ") + what + " entry " + std::to_string(x) +
                                                   " is not a parameter set index");
                         v[static_cast<size_t>(k)] = static_cast<uint16_t>(x);
                     }
                     return v;
                 };
                 const int64_t la = oa.data()[n];
                 const std::vector<uint16_t> t = flat(tracks, la, "tracks");
                 const std::vector<uint16_t> g = flat(gap_tracks, la + n, "gap_tracks");
                 return s.submit(parse_kind(kind), ca.data(), oa.data(), cb.data(), ob.data(), n,
                                 tracks.is_none() ? nullptr : t.data(), gap_tracks.is_none() ? nullptr : g.data());
             },
             py::arg("kind"), py::arg("codes_a"), py::arg("off_a"), py::arg("codes_b"), py::arg("off_b"),
             py::arg("tracks") = py::none(), py::arg("gap_tracks") = py::none())
        .def("poll", &AlignServer::poll)
        .def("now", &AlignServer::now)
        .def("state", [](const AlignServer& s, int64_t id) {
            static const char* names[] = {"waiting", "running", "done", "failed"};
            return std::string(names[s.job(id).state]);
        })
        .def("result", [](const AlignServer& s, int64_t id) {
            const AlignServer::Job& j = s.job(id);
            py::dict d;
            d["state"] = j.state;
            d["n"] = j.n;
            d["cells"] = j.cells;
            d["t_submit"] = j.t_submit;
            d["t_launch"] = j.t_launch;
            d["t_done"] = j.t_done;
            d["error"] = j.error;
            py::array_t<double> sc(static_cast<py::ssize_t>(j.scores.size()));
            if (!j.scores.empty()) std::memcpy(sc.mutable_data(), j.scores.data(), sizeof(double) * j.scores.size());
            d["scores"] = sc;
            if (!j.grads.empty()) {
                const py::ssize_t gs = s.num_sets() * (s.n_alpha() * s.n_alpha() + 4);
                py::array_t<double> g({static_cast<py::ssize_t>(j.n), gs});
                std::memcpy(g.mutable_data(), j.grads.data(), sizeof(double) * j.grads.size());
                d["grads"] = g;
            } else {
                d["grads"] = py::none();
            }
            return d;
        })
        .def("release", &AlignServer::release)
        .def_property_readonly("num_sets", &AlignServer::num_sets)
        .def_property_readonly("waiting", &AlignServer::waiting)
        .def_property_readonly("running_launches", &AlignServer::running_launches)
        .def_property_readonly("launches", &AlignServer::launches)
        .def_property_readonly("bytes_in_flight", &AlignServer::bytes_in_flight);
}

const char* gm_name(GapModel g) { return g == GapModel::Linear ? "linear" : "affine"; }
const char* am_name(AlignMode a) {
    return a == AlignMode::Local ? "local" : a == AlignMode::Glocal ? "glocal" : "global";
}
const char* gd_name(GradMode g) { return g == GradMode::Hard ? "hard" : g == GradMode::Soft ? "soft" : "none"; }

AlignParams params_from_array(F64Arr arr, const std::string& alphabet, double goa, double gea, double gob,
                              double geb) {
    if (arr.ndim() != 2 || arr.shape(0) != arr.shape(1)) throw std::invalid_argument("matrix must be square");
    if (static_cast<py::ssize_t>(alphabet.size()) != arr.shape(0))
        throw std::invalid_argument("alphabet length must match matrix size");
    return AlignParams(SubstMatrix(arr.data(), Alphabet::get(alphabet)), goa, gea, gob, geb);
}

py::array_t<double> mat_to_numpy(const SubstMatrix& m) {
    const py::ssize_t n = m.size();
    py::array_t<double> out({n, n});
    std::memcpy(out.mutable_data(), m.data(), sizeof(double) * static_cast<size_t>(n * n));
    return out;
}

// A tracks argument (py_exports.cpp:187-250): None; one sequence per pair, an entry None for
// that pair's default; or one flat 1-D integer array, every pair's entries concatenated.
// Pair p takes lens[p] entries.  Returns the concatenation (the C ABI's layout) and, per pair,
// whether it was given; `given` is false for None.
std::vector<uint16_t> flat_tracks(const py::object& obj, const std::vector<int64_t>& lens, const char* what,
                                  bool& given, std::vector<char>& pair_given) {
    const size_t n = lens.size();
    int64_t total = 0;
    for (int64_t l : lens) total += l;
    std::vector<uint16_t> out(static_cast<size_t>(total), 0);
    pair_given.assign(n, 0);
    given = !obj.is_none();
    if (!given) return out;
    auto put = [&](const py::handle& h, int64_t at, int64_t len, size_t p) {
        auto arr = py::array_t<int64_t, py::array::c_style | py::array::forcecast>::ensure(h);
        if (!arr || arr.ndim() != 1)
            throw py::type_error(std::string("nwgrad: ") + what + " entries must be 1-D integer sequences");
        if (arr.shape(0) != len)
            throw py::value_error(std::string("nwgrad: ") + what + "[" + std::to_string(p) + "] has " +
                                  std::to_string(arr.shape(0)) + " entries, expected " + std::to_string(len));
        for (int64_t k = 0; k < len; ++k) {
            const int64_t v = arr.data()[k];
            if (v < 0 || v > 65535)
                throw py::value_error(std::string("nwgrad: ") + what + "[" + std::to_string(k) + "] = " +
                                      std::to_string(v) + " is not a parameter set index");
            out[static_cast<size_t>(at + k)] = static_cast<uint16_t>(v);
        }
    };
    if (py::isinstance<py::array>(obj)) {
        py::array a = obj.cast<py::array>();
        if (a.ndim() == 1 && a.dtype().kind() != 'O') {
            if (a.shape(0) != total)
                throw py::value_error(std::string("nwgrad: flat ") + what + " has " + std::to_string(a.shape(0)) +
                                      " entries, expected " + std::to_string(total) +
                                      " (every pair's entries, concatenated)");
            put(obj, 0, total, 0);
            pair_given.assign(n, 1);
            return out;
        }
    }
    const py::sequence seq = obj.cast<py::sequence>();
    if (static_cast<size_t>(py::len(seq)) != n)
        throw py::value_error(std::string("nwgrad: ") + what + " needs one entry per pair (got " +
                              std::to_string(py::len(seq)) + " for " + std::to_string(n) + " pairs)");
    int64_t at = 0;
    for (size_t p = 0; p < n; ++p) {
        const py::object e = seq[p];
        if (!e.is_none()) {
            put(e, at, lens[p], p);
            pair_given[p] = 1;
        }
        at += lens[p];
    }
    return out;
}

// Apply tracks / gap_tracks to a batch of pairs of lengths la (len(A) per pair).
template <class T>
void apply_tracks(SeqPairBatchT<T>& b, const std::vector<int64_t>& la, const py::object& tracks,
                  const py::object& gap_tracks) {
    if (tracks.is_none() && gap_tracks.is_none()) return;
    std::vector<int64_t> lg(la);
    for (auto& v : lg) v += 1;
    bool tg = false, gg = false;
    std::vector<char> tpg, gpg;
    std::vector<uint16_t> t = flat_tracks(tracks, la, "tracks", tg, tpg);
    std::vector<uint16_t> g = flat_tracks(gap_tracks, lg, "gap_tracks", gg, gpg);
    if (gg) {   // a pair whose gap track is None takes the default: row 0 the first residue's set
        int64_t at = 0, ag = 0;
        for (size_t p = 0; p < la.size(); ++p) {
            if (!gpg[p]) {
                g[static_cast<size_t>(ag)] = la[p] ? t[static_cast<size_t>(at)] : 0;
                for (int64_t i = 1; i <= la[p]; ++i) g[static_cast<size_t>(ag + i)] = t[static_cast<size_t>(at + i - 1)];
            }
            at += la[p];
            ag += la[p] + 1;
        }
    }
    b.set_tracks(t.data(), gg ? g.data() : nullptr);
}

std::vector<std::string_view> views(const std::vector<std::string>& v) {
    return std::vector<std::string_view>(v.begin(), v.end());
}

struct BatchOpts {
    std::string traceback = "auto";
    std::string soft_impl = "scaled";
    std::string soft_guide = "eager";
    int hb_cutoff = 512;
};

template <class T>
struct PyBatch {
    std::unique_ptr<SeqPairBatchT<T>> b;
    BatchOpts o;
};

// batch[i]: a view of pair i that keeps its batch alive (the reference's SeqPairT view,
// py_exports.cpp:895-905 and :566-765).  Results are read from the batch's device arrays when
// asked for, so a view always shows the batch's current results.
template <class T>
struct PairView {
    py::object owner;
    PyBatch<T>* pb = nullptr;
    size_t i = 0;
    SeqPairBatchT<T>& b() const { return *pb->b; }
};

template <class T>
void bind_view(py::module_& m, const char* name) {
    using PV = PairView<T>;
    py::class_<PV>(m, name)
        .def_property_readonly("score", [](const PV& v) -> py::object {
            if (!v.b().results_valid()) return py::none();
            return py::float_(v.b().pair_score(v.i));
        }, "Alignment score, or None if not yet computed.")
        .def_property_readonly("grad", [](const PV& v) -> py::object {
            if (!v.b().results_valid() || v.b().grad_mode() == GradMode::None) return py::none();
            return py::cast(v.b().pair_grad(v.i));
        }, "Gradient as an AlignParams object, or None if not computed.")
        .def_property_readonly("score_valid", [](const PV& v) { return v.b().results_valid(); })
        .def_property_readonly("grad_valid", [](const PV& v) {
            return v.b().results_valid() && v.b().grad_mode() != GradMode::None;
        })
        .def_property_readonly("seq_a", [](const PV& v) { return v.b().seq_a(v.i); })
        .def_property_readonly("seq_b", [](const PV& v) { return v.b().seq_b(v.i); })
        .def_property_readonly("track", [](const PV& v) -> py::object {
            if (!v.b().tracked()) return py::none();
            return py::cast(v.b().track(v.i));
        }, "The pair's parameter set per residue of seq_a (None when not tracked).")
        .def_property_readonly("gap_track", [](const PV& v) -> py::object {
            if (!v.b().tracked()) return py::none();
            return py::cast(v.b().gap_track(v.i));
        }, "The pair's parameter set per DP row 0..len(seq_a) for gaps in seq_a (None when not tracked).")
        .def_property_readonly("gap_model", [](const PV& v) { return gm_name(v.b().gap_model()); })
        .def_property_readonly("mode", [](const PV& v) { return am_name(v.b().align_mode()); })
        .def_property_readonly("grad_mode", [](const PV& v) { return gd_name(v.b().grad_mode()); })
        .def_property_readonly("index", [](const PV& v) { return v.i; })
        .def_property_readonly("guide_j", [](const PV& v) -> py::object {
            if (!v.b().path_valid()) return py::none();
            return py::cast(v.b().guide_j(v.i));
        }, "Current alignment as a guide_j vector (length m+1), or None.")
        .def_property_readonly("path_valid", [](const PV& v) { return v.b().path_valid(); })
        .def_property_readonly("stored_path_valid", [](const PV& v) { return v.b().stored_path_valid(); },
                               "Whether a stored path is available (aligned(), coordinates(), cigar()).")
        .def("aligned", [](const PV& v) { return v.b().aligned(v.i); },
             "The alignment as a pair of gapped strings (seq_a, seq_b); needs score_and_grad(keep_paths=True).")
        .def("coordinates", [](const PV& v) {
            auto cc = v.b().coordinates(v.i);
            const py::ssize_t k = static_cast<py::ssize_t>(cc.first.size());
            py::array_t<int64_t> out({static_cast<py::ssize_t>(2), k});
            std::memcpy(out.mutable_data(), cc.first.data(), sizeof(int64_t) * cc.first.size());
            std::memcpy(out.mutable_data() + k, cc.second.data(), sizeof(int64_t) * cc.second.size());
            return out;
        }, "The alignment as Biopython-style coordinates: int64 (2, k), row 0 positions in seq_a, row 1 in seq_b.")
        .def("cigar", [](const PV& v, bool extended, bool soft_clip) { return v.b().cigar(v.i, extended, soft_clip); },
             py::kw_only(), py::arg("extended") = false, py::arg("soft_clip") = false,
             "The alignment as a SAM CIGAR string (seq_a the query, seq_b the reference).")
        .def("formatted", [](const PV& v, int width) {
            auto al = v.b().aligned(v.i);
            const std::string& a = al.first;
            const std::string& b = al.second;
            std::string mid(a.size(), ' ');
            for (size_t k = 0; k < a.size(); ++k)
                mid[k] = (a[k] == '-' || b[k] == '-') ? ' ' : (a[k] == b[k] ? '|' : '.');
            const size_t w = width > 0 ? static_cast<size_t>(width) : std::max<size_t>(a.size(), 1);
            std::string out;
            for (size_t at = 0; at < a.size(); at += w) {
                const size_t len = std::min(w, a.size() - at);
                out += a.substr(at, len) + "\n" + mid.substr(at, len) + "\n" + b.substr(at, len);
                if (at + len < a.size()) out += "\n\n";
            }
            return out;
        }, py::arg("width") = 60, "Pretty-printed alignment block, wrapped at `width` columns (0 = no wrap).")
        .def("__repr__", [name](const PV& v) {
            return std::string(name) + "(batch pair " + std::to_string(v.i) + ", len(a)=" +
                   std::to_string(v.b().len_a(v.i)) + ", len(b)=" + std::to_string(v.b().len_b(v.i)) + ")";
        });
}

template <class T>
void bind_batch(py::module_& m, const char* name) {
    using PB = PyBatch<T>;
    py::class_<PB>(m, name, py::dynamic_attr())
        .def(py::init([](const std::vector<std::string>& seqs_a, const std::vector<std::string>& seqs_b,
                         const AlignParams& params, int n_threads, const std::string& traceback,
                         const std::string& gap_model, const std::string& mode, const std::string& grad_mode,
                         const std::string& kernel, py::object tracks, py::object gap_tracks, py::object free_ends,
                         int device) {
                 const GapModel gm = parse_gap_model(gap_model);
                 const AlignMode am = parse_mode(mode);
                 const GradMode gd = parse_grad_mode(grad_mode);
                 const int tb = parse_traceback(traceback);
                 if (n_threads < 0) throw py::value_error("nwgrad: n_threads must be >= 0");
                 if (kernel != "auto" && kernel != "gpu" && kernel != "gfx950")
                     throw py::value_error("nwgrad_amd: kernel \"" + kernel + "\": this build's only backend is the "
                                           "gfx950 GPU (\"auto\")");
                 if (!free_ends.is_none())
                     throw py::value_error("nwgrad_amd: free_ends (semi-global) are not part of this build's GPU path");
                 auto pb = std::make_unique<PB>();
                 {
                     py::gil_scoped_release nogil;
                     pb->b = std::make_unique<SeqPairBatchT<T>>(views(seqs_a), views(seqs_b), params, gm, am, gd,
                                                                n_threads, device);
                 }
                 std::vector<int64_t> la(seqs_a.size());
                 for (size_t i = 0; i < seqs_a.size(); ++i) la[i] = static_cast<int64_t>(seqs_a[i].size());
                 apply_tracks(*pb->b, la, tracks, gap_tracks);
                 pb->o.traceback = traceback;
                 pb->b->set_traceback(tb, pb->o.hb_cutoff);
                 return pb;
             }),
             py::arg("seqs_a"), py::arg("seqs_b"), py::arg("params"), py::kw_only(), py::arg("n_threads") = 0,
             py::arg("traceback") = "auto", py::arg("gap_model") = "affine", py::arg("mode") = "global",
             py::arg("grad_mode") = "hard", py::arg("kernel") = "auto", py::arg("tracks") = py::none(),
             py::arg("gap_tracks") = py::none(), py::arg("free_ends") = py::none(), py::arg("device") = 0,
             "A batch of the pairs (seqs_a[i], seqs_b[i]) under params, of ONE problem type, aligned on the\n"
             "GPU.  Sequences are validated and encoded once and stay resident in device memory; the batch\n"
             "keeps an owned copy of params (set_params() replaces it).")
        .def_static(
            "from_codes",
            [](py::array_t<uint8_t, py::array::c_style | py::array::forcecast> codes_a,
               py::array_t<int64_t, py::array::c_style | py::array::forcecast> off_a,
               py::array_t<uint8_t, py::array::c_style | py::array::forcecast> codes_b,
               py::array_t<int64_t, py::array::c_style | py::array::forcecast> off_b, const AlignParams& params,
               const std::string& gap_model, const std::string& mode, const std::string& grad_mode, int device,
               py::object tracks, py::object gap_tracks, const std::string& traceback) {
                const GapModel gm = parse_gap_model(gap_model);
                const AlignMode am = parse_mode(mode);
                const GradMode gd = parse_grad_mode(grad_mode);
                const int tb = parse_traceback(traceback);
                if (off_a.ndim() != 1 || off_b.ndim() != 1 || off_a.shape(0) != off_b.shape(0) || off_a.shape(0) < 1)
                    throw py::value_error("nwgrad_amd: offsets must be 1-D arrays of N+1 entries");
                const int64_t N = off_a.shape(0) - 1;
                const int64_t* oa = off_a.data();
                const int64_t* ob = off_b.data();
                if (oa[0] != 0 || ob[0] != 0 || oa[N] > codes_a.size() || ob[N] > codes_b.size())
                    throw py::value_error("nwgrad_amd: offsets do not match the code arrays");
                auto pb = std::make_unique<PB>();
                {
                    py::gil_scoped_release nogil;
                    pb->b = std::make_unique<SeqPairBatchT<T>>(codes_a.data(), oa, codes_b.data(), ob, N, params, gm,
                                                               am, gd, device);
                }
                std::vector<int64_t> la(static_cast<size_t>(N));
                for (int64_t i = 0; i < N; ++i) la[static_cast<size_t>(i)] = oa[i + 1] - oa[i];
                apply_tracks(*pb->b, la, tracks, gap_tracks);
                pb->o.traceback = traceback;
                pb->b->set_traceback(tb, pb->o.hb_cutoff);
                return pb;
            },
            py::arg("codes_a"), py::arg("off_a"), py::arg("codes_b"), py::arg("off_b"), py::arg("params"),
            py::kw_only(), py::arg("gap_model") = "affine", py::arg("mode") = "global", py::arg("grad_mode") = "hard",
            py::arg("device") = 0, py::arg("tracks") = py::none(), py::arg("gap_tracks") = py::none(),
            py::arg("traceback") = "auto",
            "A batch from pre-encoded pairs: uint8 codes in alphabet order and int64 offsets (N+1), pair i\n"
            "being codes[off[i]:off[i+1]].")
        .def("__len__", [](const PB& s) { return s.b->size(); })
        .def(
            "__getitem__",
            [](py::object self, int64_t i) {
                PB& s = self.cast<PB&>();
                const int64_t n = static_cast<int64_t>(s.b->size());
                if (i < 0) i += n;
                if (i < 0 || i >= n) throw py::index_error("SeqPairBatch index out of range");
                PairView<T> v;
                v.owner = self;
                v.pb = &s;
                v.i = static_cast<size_t>(i);
                return v;
            },
            py::arg("i"), "Pair i as a view (score, grad, seq_a, seq_b, track, ...) that keeps the batch alive.")
        .def("set_params", [](PB& s, const AlignParams& p) { s.b->set_params(p); }, py::arg("params"))
        .def(
            "score_and_grad",
            [](PB& s, bool keep_paths) {
                py::gil_scoped_release nogil;
                return s.b->score_and_grad(keep_paths);
            },
            py::arg("keep_paths") = false,
            "Full DP on every pair: score, guide and (unless grad_mode \"none\") gradient, cached per pair on\n"
            "the device.  keep_paths=True also stores each alignment path (batch[i].aligned() /\n"
            ".coordinates() / .cigar(), cigars()).  Returns the sum of scores, in pair order.")
        .def(
            "banded_grad",
            [](PB& s, int bandwidth, bool keep_paths) {
                py::gil_scoped_release nogil;
                return s.b->banded_grad(bandwidth, keep_paths);
            },
            py::arg("bandwidth"), py::arg("keep_paths") = false,
            "Banded re-align + gradient around each pair's cached guide path: run score_and_grad() once,\n"
            "then set_params() + banded_grad(bw) after each update.  Returns the sum of scores.  Hard and\n"
            "score-only batches (a soft batch's banded forward-backward raises ValueError).")
        .def("drop_paths", [](PB& s) { s.b->drop_paths(); },
             "Free the stored paths; scores, gradients and guides survive.")
        .def("cigars", [](const PB& s, bool extended, bool soft_clip) { return s.b->cigars(extended, soft_clip); },
             py::kw_only(), py::arg("extended") = false, py::arg("soft_clip") = false,
             "Every pair's cigar() as a list of str, in pair order; needs stored paths.")
        .def(
            "align",
            [](const PB& s, const std::vector<std::string>& seqs_a, const std::vector<std::string>& seqs_b,
               const AlignParams& params, int band, const std::vector<std::string>& aligned_a,
               const std::vector<std::string>& aligned_b, const std::string& kernel, py::object tracks,
               py::object gap_tracks) {
                if (kernel != "auto" && kernel != "gpu" && kernel != "gfx950")
                    throw py::value_error("nwgrad_amd: kernel \"" + kernel + "\": this build's only backend is the "
                                          "gfx950 GPU (\"auto\")");
                if (seqs_a.size() != seqs_b.size())
                    throw py::value_error("nwgrad: align() needs seqs_a and seqs_b of equal length");
                const bool guided = !aligned_a.empty() || !aligned_b.empty();
                if (guided && (aligned_a.size() != seqs_a.size() || aligned_b.size() != seqs_a.size()))
                    throw py::value_error("nwgrad: aligned_a / aligned_b need one entry per pair");
                std::vector<std::vector<int>> guides;
                if (guided)
                    for (size_t i = 0; i < aligned_a.size(); ++i) guides.push_back(guide_from_aligned(aligned_a[i], aligned_b[i]));
                std::vector<int64_t> la(seqs_a.size());
                for (size_t i = 0; i < seqs_a.size(); ++i) la[i] = static_cast<int64_t>(seqs_a[i].size());
                bool tg = false, gg = false;
                std::vector<char> tpg, gpg;
                std::vector<int64_t> lg(la);
                for (auto& v : lg) v += 1;
                const std::vector<uint16_t> t = flat_tracks(tracks, la, "tracks", tg, tpg);
                std::vector<uint16_t> g = flat_tracks(gap_tracks, lg, "gap_tracks", gg, gpg);
                if (gg) {   // a pair whose gap track is None takes the default
                    int64_t at = 0, ag = 0;
                    for (size_t p = 0; p < la.size(); ++p) {
                        if (!gpg[p]) {
                            g[static_cast<size_t>(ag)] = la[p] ? t[static_cast<size_t>(at)] : 0;
                            for (int64_t i = 1; i <= la[p]; ++i) g[static_cast<size_t>(ag + i)] = t[static_cast<size_t>(at + i - 1)];
                        }
                        at += la[p];
                        ag += la[p] + 1;
                    }
                }
                auto va = views(seqs_a), vb = views(seqs_b);
                py::gil_scoped_release nogil;
                return s.b->align(va, vb, params, band, guides, tg || gg ? t.data() : nullptr, gg ? g.data() : nullptr);
            },
            py::arg("seqs_a"), py::arg("seqs_b"), py::arg("params"), py::arg("band") = 0,
            py::arg("aligned_a") = std::vector<std::string>{}, py::arg("aligned_b") = std::vector<std::string>{},
            py::arg("kernel") = "auto", py::kw_only(), py::arg("tracks") = py::none(), py::arg("gap_tracks") = py::none(),
            "Align (seqs_a[i], seqs_b[i]) under `params` WITHOUT adding them to the batch: returns a\n"
            "BatchResult (scores, gradient summed over the pairs), storing nothing per pair.  Uses the\n"
            "batch's problem type, grad mode and settings.  band > 0 bands every pair (hard batches;\n"
            "aligned_a / aligned_b, one gapped string per pair, give the guides; the diagonal otherwise).")
        .def("score_and_grad_async", [](PB& s) { s.b->score_and_grad_async(); },
             "Enqueue score_and_grad() on the batch's stream without waiting (no total).")
        .def("compute_grad", [](const PB& s) {
            py::gil_scoped_release nogil;
            return s.b->compute_grad();
        })
        .def("scores", [](const PB& s) {
            std::vector<double> v;
            {
                py::gil_scoped_release nogil;
                v = s.b->scores();
            }
            py::array_t<double> out(static_cast<py::ssize_t>(v.size()));
            if (!v.empty()) std::memcpy(out.mutable_data(), v.data(), sizeof(double) * v.size());
            return out;
        })
        .def(
            "weighted_grad",
            [](const PB& s, py::array_t<double, py::array::c_style> w) {
                if (w.ndim() != 1) throw py::value_error("nwgrad: weights must be a 1-D float64 array");
                const double* p = w.data();
                const size_t n = static_cast<size_t>(w.shape(0));
                py::gil_scoped_release nogil;
                return s.b->weighted_grad(p, n);
            },
            py::arg("weights"))
        .def("grads", [](const PB& s) {
            // (N, n, n) and (N, 4); tracked: (N, K, n, n) and (N, K, 4) (py_exports.cpp:1003-1030)
            const py::ssize_t N = static_cast<py::ssize_t>(s.b->size());
            const py::ssize_t n = s.b->alphabet().size();
            const bool tr = s.b->tracked();
            const py::ssize_t K = s.b->num_sets();
            py::array_t<double> mats = tr ? py::array_t<double>({N, K, n, n}) : py::array_t<double>({N, n, n});
            py::array_t<double> gaps = tr ? py::array_t<double>({N, K, static_cast<py::ssize_t>(4)})
                                          : py::array_t<double>({N, static_cast<py::ssize_t>(4)});
            {
                double* mp = mats.mutable_data();
                double* gp = gaps.mutable_data();
                py::gil_scoped_release nogil;
                s.b->grads_into(mp, gp);
            }
            return py::make_tuple(mats, gaps);
        })
        .def_property_readonly("tracked", [](const PB& s) { return s.b->tracked(); },
                               "Whether the pairs carry tracks (tracks=... or K > 1 parameter sets).")
        .def_property_readonly("num_sets", [](const PB& s) { return s.b->num_sets(); })
        .def_property(
            "fill", [](const PB& s) { return s.b->fill(); }, [](PB& s, const std::string& v) { s.b->set_fill(v); },
            "\"interpair\" (default): the batch's tiles in one launch, most work first -- short hard /\n"
            "score tiles one per block in dispatch order, soft and long-pair work pulled from device queues\n"
            "by a persistent grid; \"striped\" / \"rowwise\": plain per-bucket launches.  Results are identical.")
        .def_property(
            "columns_per_lane", [](const PB& s) { return s.b->columns_per_lane(); },
            [](PB& s, int v) { s.b->set_columns_per_lane(v); },
            "Columns a lane of the hard / score-only tiles (8 or 16; the soft pass's 8 for soft batches).\n"
            "Set 0 to choose per batch (the lane-width table of nwgrad_amd.tuning, else the built-in rule),\n"
            "8 or 16 to force.  Results are identical.  No counterpart in the reference.")
        .def_property(
            "soft_temperature", [](const PB& s) { return s.b->soft_temperature(); },
            [](PB& s, double v) { s.b->set_soft_temperature(v); })
        .def_property(
            "soft_precision", [](const PB& s) { return s.b->soft_precision(); },
            [](PB& s, const std::string& v) { s.b->set_soft_precision(v); })
        .def_property(
            "soft_impl", [](const PB& s) { return s.o.soft_impl; },
            [](PB& s, const std::string& v) {
                if (v == "log")
                    throw py::value_error("nwgrad_amd: soft_impl \"log\" (the log-space forward-backward) is not part "
                                          "of the GPU path; \"scaled\" (float32, then double, then an error) and "
                                          "\"scaled_or_log\" are");
                if (v != "scaled" && v != "scaled_or_log")
                    throw py::value_error("nwgrad: unknown soft_impl \"" + v + "\"");
                s.o.soft_impl = v;
            },
            "\"scaled\" (default): the scaled probability-space pass, a float32 refusal rerun in double, a\n"
            "double refusal a ValueError.  \"scaled_or_log\": the same; where the reference would fall back\n"
            "to its log-space pass this build raises the same ValueError (its GPU pass keeps per-segment\n"
            "exponents, so it refuses far fewer pairs).  \"log\" is refused.")
        .def_property(
            "soft_guide", [](const PB& s) { return s.o.soft_guide; },
            [](PB& s, const std::string& v) {
                if (v != "eager" && v != "lazy" && v != "posterior")
                    throw py::value_error("nwgrad: unknown soft_guide \"" + v +
                                          "\" (expected \"eager\", \"lazy\" or \"posterior\")");
                s.o.soft_guide = v;
            })
        .def_property(
            "hb_cutoff", [](const PB& s) { return s.o.hb_cutoff; },
            [](PB& s, int v) {
                if (v < 1) throw py::value_error("nwgrad: hb_cutoff must be >= 1");
                s.b->set_traceback(parse_traceback(s.o.traceback), v);
                s.o.hb_cutoff = v;
            },
            "Hirschberg base case in rows: under traceback \"hirschberg\" / \"hirschberg_pmax\" a hard pass\n"
            "takes every pair of more than hb_cutoff rows through the linear-memory recursion on the GPU.")
        .def_property_readonly(
            "hb_info", [](const PB& s) { auto r = s.b->hb_info(); return py::make_tuple(r[0], r[1], r[2], r[3], r[4]); },
            "(pairs split, recursion levels, base-case blocks, device scratch bytes, cells filled) of the last\n"
            "hard pass's Hirschberg recursion.")
        .def_property(
            "n_threads", [](const PB& s) { return s.b->n_threads(); }, [](PB& s, int v) { s.b->set_n_threads(v); })
        .def_property_readonly("alphabet", [](const PB& s) { return s.b->alphabet().symbols(); })
        .def_property_readonly("gap_model", [](const PB& s) { return gm_name(s.b->gap_model()); })
        .def_property_readonly("mode", [](const PB& s) { return am_name(s.b->align_mode()); })
        .def_property_readonly("grad_mode", [](const PB& s) { return gd_name(s.b->grad_mode()); })
        .def_property_readonly("traceback", [](const PB& s) { return s.o.traceback; })
        .def_property_readonly("cells", [](const PB& s) { return s.b->cells(); },
                               "Sum of len(A)*len(B) over the pairs (the GCUPS numerator).")
        .def_property_readonly("tiles", [](const PB& s) { return s.b->tiles(); })
        .def_property_readonly("wide_info", [](const PB& s) { auto r = s.b->wide_info(); return py::make_tuple(r[0], r[1], r[2], r[3]); },
                               "(pipelined, pooled, pool_rows, lag) of the last persistent soft launch's wide tiles: "
                               "how many ran pipelined, as panel items across waves, how many took a region of the "
                               "circular wide pool, the pool's rows and the lag, in tiles, of a pipelined tile's "
                               "backward items behind its forward ones")
        .def_property_readonly("total_info", [](const PB& s) { auto r = s.b->total_info(); return py::make_tuple(r[0], r[1]); },
                               "(from_device, in_pair_order): how score_and_grad() has formed the totals so far -- "
                               "the device's sum, taken where no partial sum can round (every order gives the "
                               "pair-order total), or the scores summed in pair order on the host")
        .def("set_table_budget", [](PB& s, int64_t bytes) { s.b->set_table_budget(bytes); }, py::arg("bytes"),
             "Device memory the hard fill (pairs up to 1,024 columns) may take for its pointer codes (0: 16 GiB). "
             "Over it, fill \"interpair\" runs one persistent launch with as many waves as fit it and the "
             "per-bucket fill splits its launches; results never change")
        .def_property_readonly("hard_launches", [](const PB& s) { auto r = s.b->hard_launches(); return py::make_tuple(r[0], r[1], r[2]); },
                               "(launches, persistent, last_grid) of the hard / score-only fill of pairs up to "
                               "1,024 columns under the current fill mode and table budget: its launches, whether "
                               "the default fill runs persistent, and the last fill's persistent grid (0: plain)")
        .def("set_soft_scratch", [](PB& s, int64_t bytes) { s.b->set_soft_scratch(bytes); }, py::arg("bytes"),
             "Device memory the soft pass may take for stored forward rows per launch (0: 96 GiB); "
             "fewer bytes run fewer wide pairs at once, never a different result")
        .def_property_readonly("launch_info", [](const PB& s) { auto r = s.b->launch_info(); return py::make_tuple(r[0], r[1], r[2]); },
                               "(grid, wide-queue blocks, resident blocks per CU) of the last soft launch")
        .def_property_readonly("soft_reruns", [](const PB& s) { return s.b->soft_reruns(); },
                               "Pairs of the last soft pass rerun in double after the float32 pass refused them.")
        .def_property_readonly("last_kernel_ms", [](const PB& s) { return s.b->last_kernel_ms(); },
                               "Device time of the last pass's kernels (HIP events), milliseconds.")
        .def("_handle", [](const PB& s) { return reinterpret_cast<uintptr_t>(s.b->handle()); })
        .def("_scores_dev_ptr", [](const PB& s) { return reinterpret_cast<uintptr_t>(s.b->scores_dev()); })
        .def("_set_params_dev", [](PB& s, uintptr_t theta) { s.b->set_params_dev(reinterpret_cast<const double*>(theta)); },
             py::arg("theta_dev"))
        .def("_run_async", [](PB& s) { s.b->score_and_grad_async(); })
        .def("_scores_copy_dev", [](const PB& s, uintptr_t dst) { s.b->scores_copy_dev(reinterpret_cast<double*>(dst)); },
             py::arg("dst_dev"))
        .def("_soft_finish", [](PB& s) { s.b->soft_finish(); return s.b->soft_reruns(); })
        .def("_soft_refusals", [](PB& s, bool reset) { auto r = s.b->soft_refusals(reset); return py::make_tuple(r[0], r[1], r[2]); },
             py::arg("reset") = false,
             "(pair_passes, first_pass, passes): float32 refusals the batch's passes left before any rerun -- "
             "pair-passes refused, the first pass (1-based) with one (0: none), passes run.  reset clears the counts.")
        .def("_set_stream", [](PB& s, uintptr_t stream) { s.b->set_stream(reinterpret_cast<void*>(stream)); },
             py::arg("stream"))
        .def(
            "_weighted_grad_dev",
            [](const PB& s, uintptr_t w_dev, uintptr_t out_dev) {
                s.b->weighted_grad_dev(reinterpret_cast<const double*>(w_dev), reinterpret_cast<double*>(out_dev));
            },
            py::arg("w_dev"), py::arg("out_dev"),
            "sum_i w[i] * grad_i on device buffers (float64), asynchronous on the batch's stream.");
}

}  // namespace

PYBIND11_MODULE(_nwgrad_amd, m) {
    m.doc() = "nwgrad_amd: the MI355X (gfx950) alignment hot path";

    py::class_<SubstMatrix>(m, "SubstMatrix")
        .def(py::init([](F64Arr arr, const std::string& alphabet) {
                 if (arr.ndim() != 2 || arr.shape(0) != arr.shape(1))
                     throw std::invalid_argument("matrix must be square");
                 if (static_cast<py::ssize_t>(alphabet.size()) != arr.shape(0))
                     throw std::invalid_argument("alphabet length must match matrix size");
                 return SubstMatrix(arr.data(), Alphabet::get(alphabet));
             }),
             py::arg("matrix"), py::arg("alphabet") = std::string(kAAOrder))
        .def("score", [](const SubstMatrix& s, const std::string& a, const std::string& b) {
            if (a.size() != 1 || b.size() != 1) throw std::invalid_argument("score() expects single-character strings");
            return s.score(a[0], b[0]);
        })
        .def("to_matrix", &mat_to_numpy)
        .def_property_readonly("size", &SubstMatrix::size)
        .def_property_readonly("alphabet", [](const SubstMatrix& s) { return s.alphabet().symbols(); });

    py::class_<AlignParams>(m, "AlignParams")
        .def(py::init(&params_from_array), py::arg("matrix"), py::arg("alphabet") = std::string(kAAOrder),
             py::arg("gap_open_a") = 0.0, py::arg("gap_extend_a") = 0.0, py::arg("gap_open_b") = 0.0,
             py::arg("gap_extend_b") = 0.0)
        .def(py::init([](const SubstMatrix& mat, double goa, double gea, double gob, double geb) {
                 return AlignParams(mat, goa, gea, gob, geb);
             }),
             py::arg("matrix"), py::arg("gap_open_a") = 0.0, py::arg("gap_extend_a") = 0.0,
             py::arg("gap_open_b") = 0.0, py::arg("gap_extend_b") = 0.0)
        .def_static(
            "from_arrays",
            [](py::array_t<double, py::array::c_style | py::array::forcecast> mats,
               py::array_t<double, py::array::c_style | py::array::forcecast> gaps, const std::string& alphabet) {
                if (mats.ndim() != 3 || mats.shape(1) != mats.shape(2))
                    throw std::invalid_argument("nwgrad: matrices must have shape (K, N, N)");
                const py::ssize_t K = mats.shape(0), n = mats.shape(1);
                if (static_cast<py::ssize_t>(alphabet.size()) != n)
                    throw std::invalid_argument("nwgrad: alphabet length must match matrix size");
                if (gaps.ndim() != 2 || gaps.shape(0) != K || gaps.shape(1) != 4)
                    throw std::invalid_argument("nwgrad: gaps must have shape (K, 4) -- gap_open_a, gap_extend_a, "
                                                "gap_open_b, gap_extend_b per set -- for K = " + std::to_string(K) +
                                                " matrices");
                if (K < 1 || K > 4096)
                    throw std::invalid_argument("nwgrad: the number of parameter sets must be in [1, 4096]");
                return AlignParams(SubstMatrix(mats.data(), Alphabet::get(alphabet), static_cast<int>(K)), gaps.data());
            },
            py::arg("matrices"), py::arg("gaps"), py::arg("alphabet") = std::string(kAAOrder),
            "K parameter sets from arrays: matrices (K, N, N) and gaps (K, 4).  Set k is what a track\n"
            "entry k selects (SeqPairBatch(..., tracks=...)).")
        .def_static("stack", &AlignParams::stack, py::arg("sets"),
                    "Stack params over one alphabet into one params with their sets in order.")
        .def_property(
            "matrices",
            [](const AlignParams& p) {
                const py::ssize_t K = p.num_sets(), n = p.matrix.size();
                py::array_t<double> out({K, n, n});
                std::memcpy(out.mutable_data(), p.matrix.data(), sizeof(double) * static_cast<size_t>(K * n * n));
                return out;
            },
            [](AlignParams& p, py::array_t<double, py::array::c_style | py::array::forcecast> mats) {
                const py::ssize_t K = p.num_sets(), n = p.matrix.size();
                if (mats.ndim() != 3 || mats.shape(0) != K || mats.shape(1) != n || mats.shape(2) != n)
                    throw std::invalid_argument("nwgrad: matrices must have shape (K, N, N) = (" + std::to_string(K) +
                                                ", " + std::to_string(n) + ", " + std::to_string(n) + ")");
                std::memcpy(p.matrix.data(), mats.data(), sizeof(double) * static_cast<size_t>(K * n * n));
            })
        .def_property(
            "gaps",
            [](const AlignParams& p) {
                const std::vector<double> g = p.all_gaps();
                py::array_t<double> out({static_cast<py::ssize_t>(p.num_sets()), static_cast<py::ssize_t>(4)});
                std::memcpy(out.mutable_data(), g.data(), sizeof(double) * g.size());
                return out;
            },
            [](AlignParams& p, py::array_t<double, py::array::c_style | py::array::forcecast> g) {
                const py::ssize_t K = p.num_sets();
                if (g.ndim() != 2 || g.shape(0) != K || g.shape(1) != 4)
                    throw std::invalid_argument("nwgrad: gaps must have shape (K, 4) = (" + std::to_string(K) + ", 4)");
                for (int k = 0; k < K; ++k)
                    for (int f = 0; f < 4; ++f) p.gap_ref(k, f) = g.data()[k * 4 + f];
            })
        .def("__getitem__", [](const AlignParams& p, int64_t k) {
            const int64_t K = p.num_sets();
            if (k < 0) k += K;
            if (k < 0 || k >= K) throw py::index_error("nwgrad: parameter set index out of range");
            return p.set(static_cast<int>(k));
        })
        .def("__setitem__", [](AlignParams& p, int64_t k, const AlignParams& v) {
            const int64_t K = p.num_sets();
            if (k < 0) k += K;
            if (k < 0 || k >= K) throw py::index_error("nwgrad: parameter set index out of range");
            p.assign_set(static_cast<int>(k), v);
        })
        .def_property(
            "matrix", [](const AlignParams& p) { return p.matrix.slice(0); },
            [](AlignParams& p, const SubstMatrix& mat) {
                if (&mat.alphabet() != &p.matrix.alphabet())
                    throw std::invalid_argument("nwgrad: cannot assign a matrix over alphabet \"" +
                                                mat.alphabet().symbols() + "\" to params over \"" +
                                                p.matrix.alphabet().symbols() + "\"");
                const size_t nn = static_cast<size_t>(mat.size()) * static_cast<size_t>(mat.size());
                std::memcpy(p.matrix.data(), mat.data(), sizeof(double) * nn);   // set 0
            })
        .def_readwrite("gap_open_a", &AlignParams::gap_open_a)
        .def_readwrite("gap_extend_a", &AlignParams::gap_extend_a)
        .def_readwrite("gap_open_b", &AlignParams::gap_open_b)
        .def_readwrite("gap_extend_b", &AlignParams::gap_extend_b)
        .def_property_readonly("num_sets", &AlignParams::num_sets)
        .def_property_readonly("alphabet", [](const AlignParams& p) { return p.alphabet().symbols(); })
        .def("to_dict", [](const AlignParams& p) {
            py::dict d;
            d["matrix"] = mat_to_numpy(p.matrix);
            d["alphabet"] = p.matrix.alphabet().symbols();
            d["gap_open_a"] = p.gap_open_a;
            d["gap_extend_a"] = p.gap_extend_a;
            d["gap_open_b"] = p.gap_open_b;
            d["gap_extend_b"] = p.gap_extend_b;
            const int K = p.num_sets();
            if (K > 1) {   // every set too (py_exports.cpp:1488-1510)
                const py::ssize_t n = p.matrix.size();
                d["num_sets"] = K;
                py::array_t<double> mats({static_cast<py::ssize_t>(K), n, n});
                std::memcpy(mats.mutable_data(), p.matrix.data(), sizeof(double) * static_cast<size_t>(K * n * n));
                d["matrices"] = mats;
                const std::vector<double> g = p.all_gaps();
                py::array_t<double> gg({static_cast<py::ssize_t>(K), static_cast<py::ssize_t>(4)});
                std::memcpy(gg.mutable_data(), g.data(), sizeof(double) * g.size());
                d["gaps"] = gg;
            }
            return d;
        })
        .def("__add__", [](const AlignParams& a, const AlignParams& b) { return a + b; })
        .def("__iadd__", [](AlignParams& a, const AlignParams& b) -> AlignParams& { a += b; return a; },
             py::return_value_policy::reference)
        .def("__mul__", [](const AlignParams& a, double s) { return a * s; })
        .def("__rmul__", [](const AlignParams& a, double s) { return a * s; })
        .def("__imul__", [](AlignParams& a, double s) -> AlignParams& { a *= s; return a; },
             py::return_value_policy::reference)
        .def("__neg__", [](const AlignParams& a) { return -a; })
        .def("__sub__", [](const AlignParams& a, const AlignParams& b) { return a - b; })
        .def("__isub__", [](AlignParams& a, const AlignParams& b) -> AlignParams& { a -= b; return a; },
             py::return_value_policy::reference);

    using SR = BatchResult;
    py::class_<SR>(m, "BatchResult")
        .def_property_readonly("scores", [](const SR& r) {
            py::array_t<double> out(static_cast<py::ssize_t>(r.scores.size()));
            if (!r.scores.empty()) std::memcpy(out.mutable_data(), r.scores.data(), sizeof(double) * r.scores.size());
            return out;
        }, "Alignment scores, shape (N,).")
        .def_property_readonly("grad", [](const SR& r) { return r.grad; },
                               "Summed gradient over the batch as an AlignParams object.");
    bind_view<float>(m, "SeqPair");
    bind_view<double>(m, "SeqPairDouble");
    bind_batch<float>(m, "SeqPairBatch");
    bind_batch<double>(m, "SeqPairBatchDouble");
    bind_server(m);

    m.def(
        "_set_lane_table",
        [](const std::vector<std::array<int32_t, 3>>& rules) {
            std::vector<int32_t> flat;
            for (const auto& r : rules) flat.insert(flat.end(), r.begin(), r.end());
            const int rc = nwg_set_lane_table(static_cast<int>(rules.size()), flat.data());
            if (rc != NWG_OK) throw py::value_error("nwgrad_amd: lane-width rules are (n_max, tiles_per_cu_max, 8 or 16)");
        },
        py::arg("rules"), "Install the process-wide lane-width table (see nwgrad_amd.tuning).");
    m.def("device_available", []() {
        try {
            Device::ctx(0);
            return true;
        } catch (const std::exception&) {
            return false;
        }
    });
}
