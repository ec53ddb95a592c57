// pymodule.cpp -- the drop-in Python module `Tagore` (pybind11) over the C ABI.
//
// Mirrors the reference module (pybind.cu:393-399): GNN_descent, Pruning, Pruning_with_CFS
// and largeIndex (one-GPU devicelist) with the same names, argument order and defaults.
// Differences, all deliberate:
//   * failures raise Python exceptions instead of exit() (Tagore_src.cuh:51-60);
//   * unknown index types / modes / metrics and layout-mixing CFS combinations raise
//     ValueError instead of silently doing nothing (pybind.cu:177-222, 264-284);
//   * the 8 capsules share ownership of one device handle (any of them keeps it alive),
//     and the kNN graph is not overwritten by pruning, so one build can be pruned many times;
//   * capsule 2 holds FP32 vectors (the reference stores scaled FP16).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tagore.h"

namespace py = pybind11;

namespace {

struct Holder {
  tg_state* s = nullptr;
  ~Holder() { tg_free(s); }
};

const char* kCapsuleName = "tagore.device_buffer";

void check(int rc) {
  if (rc == TG_OK) return;
  const std::string msg = tg_last_error();
  if (rc == TG_ERR_ARG) throw py::value_error(msg);
  if (rc == TG_ERR_IO) throw std::ios_base::failure(msg);
  throw std::runtime_error(msg);
}

void capsule_destructor(PyObject* cap) {
  auto* sp = static_cast<std::shared_ptr<Holder>*>(PyCapsule_GetContext(cap));
  delete sp;
}

py::object make_capsule(const std::shared_ptr<Holder>& h, void* dev_ptr) {
  // a capsule may not hold a null pointer: unused buffers point at the data buffer's base
  PyObject* cap = PyCapsule_New(dev_ptr, kCapsuleName, capsule_destructor);
  if (!cap) throw py::error_already_set();
  if (PyCapsule_SetContext(cap, new std::shared_ptr<Holder>(h)) != 0) {
    Py_DECREF(cap);
    throw py::error_already_set();
  }
  return py::reinterpret_steal<py::object>(cap);
}

std::shared_ptr<Holder> holder_of(const py::tuple& ptrs) {
  if (ptrs.size() != 8) throw py::value_error("ptrs must be the 8-tuple returned by GNN_descent");
  PyObject* cap = ptrs[0].ptr();
  if (!PyCapsule_IsValid(cap, kCapsuleName))
    throw py::value_error("ptrs[0] is not a capsule produced by Tagore.GNN_descent");
  auto* sp = static_cast<std::shared_ptr<Holder>*>(PyCapsule_GetContext(cap));
  if (!sp || !*sp || !(*sp)->s) throw py::value_error("ptrs refer to a released handle");
  return *sp;
}

void check_shape(tg_state* s, unsigned K, unsigned POINTS, unsigned DIM) {
  void* p;
  size_t bytes_g = 0, bytes_x = 0;
  check(tg_buffer(s, 0, &p, &bytes_g));
  check(tg_buffer(s, 2, &p, &bytes_x));
  if (bytes_x != size_t(POINTS) * DIM * 4)
    throw py::value_error("POINTS/DIM do not match the kNN graph passed in ptrs");
  if (bytes_g != size_t(POINTS) * K * 4)
    throw py::value_error("K does not match the kNN graph passed in ptrs");
}

// GNN_descent (pybind.cu:19-142)
py::tuple GNN_descent(unsigned K, unsigned POINTS, unsigned DIM, unsigned iter, std::string fname,
                      int precision, uint64_t seed, int device, std::string init) {
  int init_kind = TG_INIT_RANDOM;
  if (init == "learned") init_kind = TG_INIT_LEARNED;
  else if (init != "random") throw py::value_error("init must be 'random' or 'learned'");
  float* rows = nullptr;
  uint32_t n = 0, d = 0;
  check(tg_read_fvecs(fname.c_str(), &rows, &n, &d));
  std::unique_ptr<float, void (*)(void*)> guard(rows, tg_host_free);
  if (n != POINTS || d != DIM)
    throw py::value_error("file holds " + std::to_string(n) + " x " + std::to_string(d) +
                          " vectors, expected POINTS x DIM = " + std::to_string(POINTS) +
                          " x " + std::to_string(DIM));
  std::cout << "Points: " << n << ", Dim: " << d << std::endl;
  auto h = std::make_shared<Holder>();
  tg_knn_params p{};
  p.K = K;
  p.iters = iter;
  p.seed = seed;
  p.precision = precision;
  p.init = init_kind;
  double ms = 0.0;
  {
    py::gil_scoped_release nogil;
    check(tg_create(n, d, device, &h->s));
    check(tg_set_data(h->s, rows));
    check(tg_knn_build(h->s, &p));
    tg_stats st{};
    check(tg_get_stats(h->s, &st));
    ms = st.knn_ms;
  }
  std::cout << "time of GNN-Descent: " << ms / 1000.0 << "s" << std::endl;
  std::vector<py::object> caps;
  for (int i = 0; i < 8; ++i) {
    void* ptr = nullptr;
    size_t bytes = 0;
    check(tg_buffer(h->s, i, &ptr, &bytes));
    if (!ptr) check(tg_buffer(h->s, 2, &ptr, &bytes));
    caps.push_back(make_capsule(h, ptr));
  }
  return py::make_tuple(caps[0], caps[1], caps[2], caps[3], caps[4], caps[5], caps[6], caps[7]);
}

void run_prune(tg_state* s, const tg_prune_params& p, const char* store_type,
               const std::string& path) {
  double ms = 0.0;
  {
    py::gil_scoped_release nogil;
    const auto t0 = std::chrono::high_resolution_clock::now();
    check(tg_prune(s, &p));
    if (store_type) check(tg_store_index(s, path.c_str(), store_type, 0x5EEDull));
    const auto t1 = std::chrono::high_resolution_clock::now();
    ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  std::cout << "time of pruning: " << ms / 1000.0 << "s" << std::endl;
}

// Pruning (pybind.cu:145-230)
void Pruning(unsigned K, unsigned POINTS, unsigned DIM, unsigned FINAL_DEGREE, unsigned TOPM,
             py::tuple ptrs, std::string index_type, float thre, std::string index_path) {
  auto h = holder_of(ptrs);
  check_shape(h->s, K, POINTS, DIM);
  tg_prune_params p{};
  p.precision = TG_PREC_BF16;  // BF16 rows only when exact for this data set
  p.FD = FINAL_DEGREE;
  p.TOPM = TOPM;
  p.thre = thre;
  if (index_type == "NSG") {
    std::cout << "begin to construct NSG" << std::endl;
    p.kind = TG_KIND_PATH;
    p.fwd_rule = p.rev_rule = TG_RULE_DIST;
    p.thre = 1.0f;  // pybind.cu:181,183 hard-wire 1.0
  } else if (index_type == "Vamana") {
    std::cout << "begin to construct Vamana" << std::endl;
    p.kind = TG_KIND_PATH;
    p.fwd_rule = p.rev_rule = TG_RULE_DIST;
  } else if (index_type == "NSSG") {
    std::cout << "begin to construct NSSG" << std::endl;
    p.kind = TG_KIND_TWOHOP;
    p.fwd_rule = p.rev_rule = TG_RULE_ANGLE;
  } else if (index_type == "CAGRA") {
    std::cout << "begin to construct CAGRA" << std::endl;
    p.kind = TG_KIND_CAGRA;
    p.rev_rule = TG_RULE_RANK;
  } else if (index_type == "DPG") {
    std::cout << "begin to construct DPG" << std::endl;
    p.kind = TG_KIND_DPG;
    p.rev_rule = TG_RULE_RANK;
  } else {
    throw py::value_error("index_type must be one of NSG, Vamana, NSSG, CAGRA, DPG");
  }
  run_prune(h->s, p, index_type.c_str(), index_path);
}

// Pruning_with_CFS (pybind.cu:232-295); no index file, like the reference
void Pruning_with_CFS(unsigned K, unsigned POINTS, unsigned DIM, unsigned FINAL_DEGREE,
                      unsigned TOPM, py::tuple ptrs, std::string mode, std::string metric,
                      float thre) {
  auto h = holder_of(ptrs);
  check_shape(h->s, K, POINTS, DIM);
  tg_prune_params p{};
  p.precision = TG_PREC_BF16;  // BF16 rows only when exact for this data set
  p.FD = FINAL_DEGREE;
  p.TOPM = TOPM;
  p.thre = thre;
  if (mode == "path") {
    p.kind = TG_KIND_PATH;
    p.fwd_rule = TG_RULE_DIST;
  } else if (mode == "2-hop") {
    p.kind = TG_KIND_TWOHOP;
    p.fwd_rule = TG_RULE_ANGLE;
  } else if (mode == "1-hop") {
    p.kind = TG_KIND_CAGRA;
  } else {
    throw py::value_error("mode must be one of path, 2-hop, 1-hop");
  }
  if (metric == "dist") p.rev_rule = TG_RULE_DIST;
  else if (metric == "angle") p.rev_rule = TG_RULE_ANGLE;
  else if (metric == "rank") p.rev_rule = TG_RULE_RANK;
  else throw py::value_error("metric must be one of dist, angle, rank");
  run_prune(h->s, p, nullptr, "");
}

// Test accessor with no counterpart in the reference (whose Pruning_with_CFS leaves the
// pruned rows in device memory only): the last pruned graph of `ptrs` as (deg [POINTS],
// ids [POINTS, FINAL_DEGREE], entry point).
py::tuple pruned_graph(unsigned K, unsigned POINTS, unsigned DIM, py::tuple ptrs) {
  auto h = holder_of(ptrs);
  check_shape(h->s, K, POINTS, DIM);
  uint32_t FD = 0, ep = 0;
  check(tg_pruned_width(h->s, &FD));
  py::array_t<uint32_t> deg(POINTS);
  py::array_t<uint32_t> ids({size_t(POINTS), size_t(FD)});
  check(tg_get_pruned(h->s, deg.mutable_data(), ids.mutable_data(), &ep));
  return py::make_tuple(deg, ids, ep);
}

// largeIndex (pybind.cu:297-391) on a one-GPU devicelist
void largeIndex(std::vector<int> devicelist, unsigned DIM, unsigned K, unsigned cluster_num,
                unsigned max_points, std::string file_prefix, std::string local_index_path_prefix,
                std::string index_store_path, unsigned max_degree, unsigned buffer_size,
                unsigned all_node_num, std::string centers_path, float thre,
                unsigned FINAL_DEGREE, unsigned TOPM, uint64_t seed, int precision,
                uint64_t batch_rows, unsigned iters, bool tune) {
  std::vector<int32_t> dev(devicelist.begin(), devicelist.end());
  tg_large_params p{};
  p.devicelist = dev.data();
  p.ndevices = uint32_t(dev.size());
  p.DIM = DIM;
  p.K = K;
  p.cluster_num = cluster_num;
  p.max_points = max_points;
  p.file_prefix = file_prefix.c_str();
  p.local_index_path_prefix = local_index_path_prefix.c_str();
  p.index_store_path = index_store_path.c_str();
  p.max_degree = max_degree;
  p.buffer_size = buffer_size;
  p.all_node_num = all_node_num;
  p.centers_path = centers_path.c_str();
  p.thre = thre;
  p.FINAL_DEGREE = FINAL_DEGREE;
  p.TOPM = TOPM;
  p.iters = iters;
  p.seed = seed;
  p.precision = precision;
  p.batch_rows = batch_rows;
  p.verbose = 1;
  p.tune = tune ? 1 : 0;
  tg_large_stats st{};
  {
    py::gil_scoped_release nogil;
    check(tg_large_index(&p, &st));
  }
  std::cout << "Clusters: " << st.clusters << " in " << st.batches << " batched builds; build "
            << st.build_ms / 1000.0 << "s (device " << st.device_ms / 1000.0 << "s), merge "
            << st.merge_ms / 1000.0 << "s" << std::endl;
}

}  // namespace

PYBIND11_MODULE(Tagore, m) {
  m.doc() = "Tagore GPU Indexing";
  m.def("GNN_descent", &GNN_descent, "GNN_descent", py::arg("K"), py::arg("POINTS"),
        py::arg("DIM"), py::arg("iter"), py::arg("fname"), py::kw_only(),
        py::arg("precision") = int(TG_PREC_BF16), py::arg("seed") = uint64_t(0x7A600000ull),
        py::arg("device") = 0, py::arg("init") = "random");
  m.def("Pruning", &Pruning, "Pruning", py::arg("K"), py::arg("POINTS"), py::arg("DIM"),
        py::arg("FINAL_DEGREE"), py::arg("TOPM") = 64, py::arg("ptrs"), py::arg("index_type"),
        py::arg("thre") = 1.0f, py::arg("index_path") = "index.data");
  m.def("Pruning_with_CFS", &Pruning_with_CFS, "Pruning_with_CFS");
  m.def("pruned_graph", &pruned_graph, "the last pruned graph of ptrs: (deg, ids, ep)",
        py::arg("K"), py::arg("POINTS"), py::arg("DIM"), py::arg("ptrs"));
  m.def("largeIndex", &largeIndex, "largeIndex", py::arg("devicelist"), py::arg("DIM"),
        py::arg("K"), py::arg("cluster_num"), py::arg("max_points"), py::arg("file_prefix"),
        py::arg("local_index_path_prefix"), py::arg("index_store_path"), py::arg("max_degree"),
        py::arg("buffer_size"), py::arg("all_node_num"), py::arg("centers_path"),
        py::arg("thre"), py::arg("FINAL_DEGREE"), py::arg("TOPM"), py::kw_only(),
        py::arg("seed") = uint64_t(0x7A600000ull), py::arg("precision") = int(TG_PREC_BF16),
        py::arg("batch_rows") = uint64_t(0), py::arg("iters") = 16u, py::arg("tune") = true);
}
