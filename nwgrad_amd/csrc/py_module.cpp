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
void check_traceback(const std::string& s) {
    if (s == "auto" || s == "pointers" || s == "scores") return;
    if (s == "hirschberg" || s == "hirschberg_pmax")
        throw py::value_error("nwgrad_amd: traceback \"" + s + "\" is not available on the GPU path; \"auto\", "
                              "\"pointers\" and \"scores\" (identical results) are");
    throw py::value_error("nwgrad: unknown traceback \"" + s + "\"");
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
             [](AlignServer& s, const std::string& kind, U8Arr ca, I64Arr oa, U8Arr cb, I64Arr ob) {
                 if (oa.ndim() != 1 || ob.ndim() != 1 || oa.size() < 1 || oa.size() != ob.size())
                     throw py::value_error("nwgrad_amd: off_a and off_b need n + 1 entries each");
                 const int64_t n = static_cast<int64_t>(oa.size()) - 1;
                 if (oa.data()[n] > static_cast<int64_t>(ca.size()) || ob.data()[n] > static_cast<int64_t>(cb.size()))
                     throw py::value_error("nwgrad_amd: offsets run past the codes");
                 return s.submit(parse_kind(kind), ca.data(), oa.data(), cb.data(), ob.data(), n);
             },
             py::arg("kind"), py::arg("codes_a"), py::arg("off_a"), py::arg("codes_b"), py::arg("off_b"))
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
                const py::ssize_t gs = s.n_alpha() * s.n_alpha() + 4;
                py::array_t<double> g({static_cast<py::ssize_t>(j.n), gs});
                std::memcpy(g.mutable_data(), j.grads.data(), sizeof(double) * j.grads.size());
                d["grads"] = g;
            } else {
                d["grads"] = py::none();
            }
            return d;
        })
        .def("release", &AlignServer::release)
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
                 check_traceback(traceback);
                 if (n_threads < 0) throw py::value_error("nwgrad: n_threads must be >= 0");
                 if (kernel != "auto" && kernel != "gpu" && kernel != "gfx950")
                     throw py::value_error("nwgrad_amd: kernel \"" + kernel + "\": this build's only backend is the "
                                           "gfx950 GPU (\"auto\")");
                 if (!tracks.is_none() || !gap_tracks.is_none())
                     throw py::value_error("nwgrad_amd: tracks / gap_tracks are not part of this build's GPU path");
                 if (!free_ends.is_none())
                     throw py::value_error("nwgrad_amd: free_ends (semi-global) are not part of this build's GPU path");
                 auto pb = std::make_unique<PB>();
                 {
                     py::gil_scoped_release nogil;
                     pb->b = std::make_unique<SeqPairBatchT<T>>(views(seqs_a), views(seqs_b), params, gm, am, gd,
                                                                n_threads, device);
                 }
                 pb->o.traceback = traceback;
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
               const std::string& gap_model, const std::string& mode, const std::string& grad_mode, int device) {
                const GapModel gm = parse_gap_model(gap_model);
                const AlignMode am = parse_mode(mode);
                const GradMode gd = parse_grad_mode(grad_mode);
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
                return pb;
            },
            py::arg("codes_a"), py::arg("off_a"), py::arg("codes_b"), py::arg("off_b"), py::arg("params"),
            py::kw_only(), py::arg("gap_model") = "affine", py::arg("mode") = "global", py::arg("grad_mode") = "hard",
            py::arg("device") = 0,
            "A batch from pre-encoded pairs: uint8 codes in alphabet order and int64 offsets (N+1), pair i\n"
            "being codes[off[i]:off[i+1]].")
        .def("__len__", [](const PB& s) { return s.b->size(); })
        .def("set_params", [](PB& s, const AlignParams& p) { s.b->set_params(p); }, py::arg("params"))
        .def(
            "score_and_grad",
            [](PB& s, bool keep_paths) {
                if (keep_paths)
                    throw py::value_error("nwgrad_amd: keep_paths (stored alignment paths) is not part of this "
                                          "build's GPU path");
                py::gil_scoped_release nogil;
                return s.b->score_and_grad();
            },
            py::arg("keep_paths") = false,
            "Full DP on every pair: score and (unless grad_mode \"none\") gradient, cached per pair on the\n"
            "device.  Returns the sum of scores, in pair order.")
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
            const size_t N = s.b->size();
            const py::ssize_t n = s.b->alphabet().size();
            py::array_t<double> mats({static_cast<py::ssize_t>(N), n, n});
            py::array_t<double> gaps({static_cast<py::ssize_t>(N), static_cast<py::ssize_t>(4)});
            {
                double* mp = mats.mutable_data();
                double* gp = gaps.mutable_data();
                py::gil_scoped_release nogil;
                s.b->grads_into(mp, gp);
            }
            return py::make_tuple(mats, gaps);
        })
        .def_property(
            "fill", [](const PB& s) { return s.b->fill(); }, [](PB& s, const std::string& v) { s.b->set_fill(v); },
            "\"interpair\" (default): one persistent launch pulling length-sorted tiles from a device work\n"
            "queue; \"striped\" / \"rowwise\": plain per-bucket launches.  Results are identical.")
        .def_property(
            "soft_temperature", [](const PB& s) { return s.b->soft_temperature(); },
            [](PB& s, double v) { s.b->set_soft_temperature(v); })
        .def_property(
            "soft_precision", [](const PB& s) { return s.b->soft_precision(); },
            [](PB& s, const std::string& v) { s.b->set_soft_precision(v); })
        .def_property(
            "soft_impl", [](const PB& s) { return s.o.soft_impl; },
            [](PB& s, const std::string& v) {
                if (v != "scaled" && v != "scaled_or_log" && v != "log")
                    throw py::value_error("nwgrad: unknown soft_impl \"" + v + "\"");
                s.o.soft_impl = v;
            })
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
                s.o.hb_cutoff = v;
            })
        .def_property(
            "n_threads", [](const PB& s) { return s.b->n_threads(); }, [](PB& s, int v) { s.b->set_n_threads(v); })
        .def_property_readonly("alphabet", [](const PB& s) { return s.b->alphabet().symbols(); })
        .def_property_readonly("gap_model", [](const PB& s) { return gm_name(s.b->gap_model()); })
        .def_property_readonly("mode", [](const PB& s) { return am_name(s.b->align_mode()); })
        .def_property_readonly("grad_mode", [](const PB& s) { return gd_name(s.b->grad_mode()); })
        .def_property_readonly("traceback", [](const PB& s) { return s.o.traceback; })
        .def_property_readonly("tracked", [](const PB&) { return false; })
        .def_property_readonly("num_sets", [](const PB& s) { return s.b->size() ? 1 : 0; })
        .def_property_readonly("cells", [](const PB& s) { return s.b->cells(); },
                               "Sum of len(A)*len(B) over the pairs (the GCUPS numerator).")
        .def_property_readonly("tiles", [](const PB& s) { return s.b->tiles(); })
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
        .def_property(
            "matrix", [](const AlignParams& p) { return p.matrix; },
            [](AlignParams& p, const SubstMatrix& mat) {
                if (&mat.alphabet() != &p.matrix.alphabet())
                    throw std::invalid_argument("nwgrad: cannot assign a matrix over alphabet \"" +
                                                mat.alphabet().symbols() + "\" to params over \"" +
                                                p.matrix.alphabet().symbols() + "\"");
                p.matrix = mat;
            })
        .def_readwrite("gap_open_a", &AlignParams::gap_open_a)
        .def_readwrite("gap_extend_a", &AlignParams::gap_extend_a)
        .def_readwrite("gap_open_b", &AlignParams::gap_open_b)
        .def_readwrite("gap_extend_b", &AlignParams::gap_extend_b)
        .def_property_readonly("num_sets", [](const AlignParams&) { return 1; })
        .def_property_readonly("alphabet", [](const AlignParams& p) { return p.alphabet().symbols(); })
        .def("to_dict", [](const AlignParams& p) {
            py::dict d;
            d["matrix"] = mat_to_numpy(p.matrix);
            d["alphabet"] = p.matrix.alphabet().symbols();
            d["gap_open_a"] = p.gap_open_a;
            d["gap_extend_a"] = p.gap_extend_a;
            d["gap_open_b"] = p.gap_open_b;
            d["gap_extend_b"] = p.gap_extend_b;
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

    bind_batch<float>(m, "SeqPairBatch");
    bind_batch<double>(m, "SeqPairBatchDouble");
    bind_server(m);

    m.def("device_available", []() {
        try {
            Device::ctx(0);
            return true;
        } catch (const std::exception&) {
            return false;
        }
    });
}
