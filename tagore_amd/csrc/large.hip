// large.hip -- largeIndex on one GPU (pybind.cu:297-391, large_index.cu:9-522).
//
// Host side restated from the reference:
//   * clusters from the two-centre file, each cluster larger than max_points split
//     round-robin into parts (pybind.cu:311-355);
//   * the cluster-aware dispatch order of Algorithm 4 (reorder, large_index.cu:9-81, 84-106);
//   * the streaming union merge through a buffer of buffer_size cluster graphs with spill
//     files (order_merge_new, large_index.cu:287-421);
//   * the final merged index in the DiskANN layout (large_index.cu:423-521).
// The GPU side differs from the reference's one-cluster-at-a-time loop (gpu_construct,
// large_index.cu:121-280): clusters are packed back to back into segmented handles, so each
// GNN-Descent and pruning kernel covers a whole batch of uneven clusters in one launch.
// Each cluster's graph is independent of the batching (its seed is keyed by its id).
// As in the reference's builder / merger threads, the host work overlaps the GPU: the next
// batch's cluster files are read while a batch builds, and a merger thread runs the
// streaming merge over every cluster whose batch is done.
//
// Decisions where the reference is undefined (all documented in tagore.h):
//   * reorder() reads an uninitialised max_id / max_evict when no candidate overlaps the
//     buffer; here the smallest unvisited cluster is taken and slot 0 is evicted;
//   * buffer_size is clamped to the cluster count (the reference indexes past it);
//   * a point whose two centres are equal is rejected (the reference's final merge reads its
//     shard records out of step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <future>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tagore.h"
#include "errors.hpp"

namespace {

constexpr uint32_t INV = 0xFFFFFFFFu;
constexpr uint32_t TUNE_ROWS = 1u << 20;  // rows the autotuner's probes build

int lfail(int code, const std::string& msg) { return tg::set_error(code, msg); }

// Algorithm 4 (large_index.cu:9-81): greedy order of clusters by overlap with the buffer.
void reorder(const std::vector<std::map<uint32_t, uint32_t>>& inter, uint32_t C, uint32_t B,
             std::vector<uint32_t>& result, std::vector<uint32_t>& position) {
  std::vector<uint32_t> buffer;
  std::vector<bool> visited(C, false);
  result.assign(1, 0u);
  visited[0] = true;
  buffer.push_back(0);
  position[0] = 0;
  auto overlap = [&](uint32_t j, uint32_t other) -> uint32_t {
    auto it = inter[j].find(other);
    return it == inter[j].end() ? 0u : it->second;
  };
  auto first_unvisited = [&]() {
    for (uint32_t j = 0; j < C; ++j)
      if (!visited[j]) return j;
    return INV;
  };
  for (uint32_t i = 1; i < B; ++i) {
    uint32_t max_sum = 0, max_id = INV;
    for (uint32_t j = 0; j < C; ++j) {
      if (visited[j]) continue;
      uint32_t t = 0;
      for (uint32_t r : result) t += overlap(j, r);
      if (t > max_sum) {
        max_sum = t;
        max_id = j;
      }
    }
    if (max_id == INV) max_id = first_unvisited();
    result.push_back(max_id);
    buffer.push_back(max_id);
    position[max_id] = i;
    visited[max_id] = true;
  }
  for (uint32_t i = 0; i + B < C; ++i) {
    uint32_t max_sum = 0, max_id = INV, max_evict = 0;
    for (uint32_t j = 0; j < C; ++j) {
      if (visited[j]) continue;
      uint32_t local_sum = 0, local_evict = 0;
      for (uint32_t k = 0; k < B; ++k) {
        uint32_t t = 0;
        for (uint32_t l = k + 1; (l % B) != k; ++l) t += overlap(j, buffer[l % B]);
        if (t > local_sum) {
          local_sum = t;
          local_evict = k;
        }
      }
      if (local_sum > max_sum) {
        max_sum = local_sum;
        max_id = j;
        max_evict = local_evict;
      }
    }
    if (max_id == INV) {
      max_id = first_unvisited();
      max_evict = 0;
    }
    buffer[max_evict] = max_id;
    result.push_back(max_id);
    visited[max_id] = true;
    position[max_id] = max_evict;
  }
}

struct Spill {
  // rows of a spilled cluster as written to its file: degree then ids
  std::vector<uint32_t> deg;
  std::vector<uint32_t> off;
  std::vector<uint32_t> ids;
  size_t cur = 0;  // read cursor (record index)
};

}  // namespace

static int large_index(const tg_large_params* p, tg_large_stats* out_stats) {
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  if (!p) return lfail(TG_ERR_ARG, "null params");
  if (!p->devicelist || p->ndevices != 1)
    return lfail(TG_ERR_ARG, "this build runs largeIndex on one GPU: devicelist must hold one id");
  const uint32_t K = p->K, D = p->DIM, FD = p->FINAL_DEGREE;
  if (K < 2 || K > 128) return lfail(TG_ERR_ARG, "K must be in [2, 128]");
  if (FD < 1 || FD >= K) return lfail(TG_ERR_ARG, "FINAL_DEGREE must be in [1, K)");
  if (p->max_degree < 1 || p->max_degree >= K) return lfail(TG_ERR_ARG, "max_degree must be in [1, K)");
  if (p->cluster_num < 1 || p->all_node_num < 1)
    return lfail(TG_ERR_ARG, "cluster_num and all_node_num must be positive");
  // with one slot the reference never releases the second cluster's slot (its builder waits
  // on has_merged forever, large_index.cu:255 / 398-412)
  if (p->buffer_size < 2) return lfail(TG_ERR_ARG, "buffer_size must be at least 2");
  if (p->tile > 1) return lfail(TG_ERR_ARG, "tile must be 0 or 1");
  if (!p->file_prefix || !p->local_index_path_prefix || !p->index_store_path || !p->centers_path)
    return lfail(TG_ERR_ARG, "null path");
  const uint32_t Npts = p->all_node_num;
  // ---- centres (pybind.cu:311-325)
  std::vector<uint32_t> cen(Npts), cen2(Npts);
  {
    FILE* f = fopen(p->centers_path, "rb");
    if (!f) return lfail(TG_ERR_IO, std::string("cannot open ") + p->centers_path);
    std::vector<uint32_t> buf(size_t(Npts) * 2);
    const size_t got = fread(buf.data(), 4, buf.size(), f);
    fclose(f);
    if (got != buf.size()) return lfail(TG_ERR_IO, "centres file holds fewer than all_node_num pairs");
    for (uint32_t i = 0; i < Npts; ++i) {
      cen[i] = buf[2 * size_t(i)];
      cen2[i] = buf[2 * size_t(i) + 1];
      if (cen[i] >= p->cluster_num || cen2[i] >= p->cluster_num)
        return lfail(TG_ERR_ARG, "centre id out of range in the centres file");
      if (cen[i] == cen2[i])
        return lfail(TG_ERR_ARG, "point " + std::to_string(i) + " has equal first and second centres");
    }
  }
  // ---- clusters and the split of oversized ones (pybind.cu:326-355)
  uint32_t C = p->cluster_num;
  std::vector<std::vector<uint32_t>> clusters(C);
  for (uint32_t i = 0; i < Npts; ++i) {
    clusters[cen[i]].push_back(i);
    clusters[cen2[i]].push_back(i);
  }
  uint32_t extra = 0;
  for (uint32_t i = 0; i < C; ++i) {
    if (clusters[i].size() <= p->max_points) continue;
    uint32_t div = 2;
    while (clusters[i].size() / div > p->max_points) ++div;
    std::vector<uint32_t> cand = {i};
    for (uint32_t j = 0; j + 1 < div; ++j) cand.push_back(C + extra + j);
    for (size_t j = 0; j < clusters[i].size(); ++j) {
      const uint32_t node = clusters[i][j];
      if (cen[node] == i) cen[node] = cand[j % div];
      else cen2[node] = cand[j % div];
    }
    extra += div - 1;
  }
  C += extra;
  clusters.assign(C, {});
  for (uint32_t i = 0; i < Npts; ++i) {
    clusters[cen[i]].push_back(i);
    clusters[cen2[i]].push_back(i);
  }
  for (uint32_t c = 0; c < C; ++c)
    if (clusters[c].size() <= K)
      return lfail(TG_ERR_ARG, "cluster " + std::to_string(c) + " holds " +
                                   std::to_string(clusters[c].size()) + " points, not more than K");
  if (p->verbose) printf("Cluster num after being adjusted:%u\n", C);
  // ---- dispatch order (large_index.cu:84-106)
  const uint32_t B = std::min(p->buffer_size, C);
  std::vector<std::map<uint32_t, uint32_t>> inter(C);
  for (uint32_t i = 0; i < Npts; ++i) {
    ++inter[cen[i]][cen2[i]];
    ++inter[cen2[i]][cen[i]];
  }
  std::vector<uint32_t> order, position(C, 0);
  reorder(inter, C, B, order, position);
  // ---- batched construction on the GPU
  const uint32_t stride = std::max(FD, p->max_degree) + 1;  // graph row: [deg, ids...]
  std::vector<std::vector<uint32_t>> graph(C);
  std::vector<uint32_t> metroid(C, 0);  // indexed by build position
  size_t batch_rows = p->batch_rows;
  if (batch_rows == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(p->devicelist[0]) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
      return lfail(TG_ERR_HIP, "cannot query device memory");
    // about 24 KB of device memory per row (data, lists, push buffers) at K <= 128
    batch_rows = std::max<size_t>(1, size_t(double(free_b) * 0.6 / (24.0 * 1024.0 + 8.0 * D)));
  }
  const auto t_build = clk::now();
  double device_ms = 0.0;
  int32_t precision = p->precision;
  uint32_t tile = p->tile;
  bool tuned = !p->tune;
  // batches: consecutive positions [b0, b1) of the dispatch order
  std::vector<std::pair<uint32_t, uint32_t>> batches;
  for (uint32_t b0 = 0; b0 < C;) {
    uint32_t b1 = b0;
    size_t rows = 0;
    while (b1 < C && (b1 == b0 || rows + clusters[order[b1]].size() <= batch_rows))
      rows += clusters[order[b1++]].size();
    batches.push_back({b0, b1});
    b0 = b1;
  }
  const uint32_t nbatches = uint32_t(batches.size());
  // one batch's rows, read from its cluster files (status, message)
  struct Rows {
    std::vector<float> X;
    std::vector<uint32_t> starts, keys;
    int rc = TG_OK;
    std::string msg;
  };
  auto read_batch = [&](uint32_t bi) {
    Rows r;
    const uint32_t b0 = batches[bi].first, b1 = batches[bi].second;
    size_t rows = 0;
    for (uint32_t k = b0; k < b1; ++k) rows += clusters[order[k]].size();
    r.X.resize(rows * D);
    r.starts.assign(1, 0);
    size_t r0 = 0;
    for (uint32_t k = b0; k < b1; ++k) {
      const uint32_t c = order[k];
      const size_t n = clusters[c].size();
      const std::string path = std::string(p->file_prefix) + std::to_string(c);
      std::ifstream in(path, std::ios::binary);
      if (!in) {
        r.rc = TG_ERR_IO;
        r.msg = "cannot open cluster file " + path;
        return r;
      }
      uint32_t header = 0;
      in.read(reinterpret_cast<char*>(&header), 4);
      in.read(reinterpret_cast<char*>(r.X.data() + r0 * D), std::streamsize(n * D * 4));
      if (!in) {
        r.rc = TG_ERR_IO;
        r.msg = "cluster file " + path + " holds fewer than its points";
        return r;
      }
      r0 += n;
      r.starts.push_back(uint32_t(r0));
      r.keys.push_back(c);
    }
    return r;
  };
  // ---- streaming merge (order_merge_new, large_index.cu:287-421), in its own thread: the
  // cluster at dispatch position i is merged once its batch is built (built > i)
  std::vector<uint32_t> pointers(Npts, 0), buffer(B, INV);
  std::vector<uint8_t> visited(Npts, 0), node_set(Npts, 0), cached(Npts, 0);
  std::vector<Spill> spill(C);
  std::mutex mu;
  std::condition_variable cv;
  uint32_t built = 0;
  bool stop = false;
  int merge_rc = TG_OK;
  std::string merge_msg;
  auto wait_built = [&](uint32_t pos) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return built > pos || stop; });
    return built > pos;
  };
  auto write_spill = [&](uint32_t clu) -> bool {
    const std::string path = std::string(p->local_index_path_prefix) + std::to_string(clu);
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
      merge_rc = TG_ERR_IO;
      merge_msg = "cannot write " + path;
      return false;
    }
    Spill& sp = spill[clu];
    const std::vector<uint32_t>& g = graph[clu];
    std::vector<uint32_t> rec;
    rec.reserve(g.size());
    for (size_t j = 0; j < clusters[clu].size(); ++j) {
      const uint32_t node = clusters[clu][j];
      const uint32_t d = g[j * stride];
      rec.push_back(d);
      sp.deg.push_back(d);
      sp.off.push_back(uint32_t(sp.ids.size()));
      for (uint32_t t = 0; t < d; ++t) {
        rec.push_back(g[j * stride + 1 + t]);
        sp.ids.push_back(g[j * stride + 1 + t]);
      }
      visited[node] = 1;
      pointers[node] = INV;
    }
    const bool ok = fwrite(rec.data(), 4, rec.size(), f) == rec.size();
    fclose(f);
    if (!ok) {
      merge_rc = TG_ERR_IO;
      merge_msg = "short write to " + path;
    }
    return ok;
  };
  auto stream_merge = [&]() {
    if (!wait_built(0)) return;
    for (size_t i = 0; i < clusters[order[0]].size(); ++i) {
      visited[clusters[order[0]][i]] = 1;
      pointers[clusters[order[0]][i]] = uint32_t(i);
    }
    buffer[0] = order[0];
    std::vector<uint32_t> nhood;
    for (uint32_t i = 1; i < C; ++i) {
      if (!wait_built(i)) return;
      const uint32_t c = order[i];
      std::vector<uint32_t>& gc = graph[c];
      for (size_t j = 0; j < clusters[c].size(); ++j) {
        const uint32_t node = clusters[c][j];
        if (!visited[node]) {
          visited[node] = 1;
          pointers[node] = uint32_t(j);
          continue;
        }
        if (pointers[node] == INV) continue;
        const uint32_t ano = cen[node] == c ? cen2[node] : cen[node];
        uint32_t* ra = graph[ano].data() + size_t(pointers[node]) * stride;
        uint32_t* rc = gc.data() + j * stride;
        const uint32_t* first = cen[node] != c ? ra : rc;
        const uint32_t* second = cen[node] != c ? rc : ra;
        for (uint32_t t = 1; t <= first[0]; ++t) {
          node_set[first[t]] = 1;
          nhood.push_back(first[t]);
        }
        for (uint32_t t = 1; t <= second[0]; ++t)
          if (!node_set[second[t]]) {
            node_set[second[t]] = 1;
            nhood.push_back(second[t]);
          }
        const uint32_t nd = std::min<uint32_t>(uint32_t(nhood.size()), p->max_degree);
        for (uint32_t t = 0; t < nd; ++t) ra[1 + t] = nhood[t];
        ra[0] = nd;
        rc[0] = 0;
        for (uint32_t v : nhood) node_set[v] = 0;
        nhood.clear();
        cached[node] = 1;
      }
      buffer[position[c]] = c;
      if (i + 1 >= B && i + 1 < C) {
        const uint32_t clu = buffer[position[order[i + 1]]];
        if (clu != INV && !write_spill(clu)) return;
      }
    }
    for (uint32_t k = 0; k < B; ++k)
      if (buffer[k] != INV && !write_spill(buffer[k])) return;
  };
  std::thread merger([&] {
    try {
      stream_merge();
    } catch (const std::bad_alloc&) {
      merge_rc = TG_ERR_CAPACITY;
      merge_msg = "largeIndex merge: host memory exhausted";
    } catch (const std::exception& e) {
      merge_rc = TG_ERR_STATE;
      merge_msg = std::string("largeIndex merge: ") + e.what();
    }
  });
  // every way out of this function (errors and exceptions included) stops and joins it
  struct MergerStop {
    std::thread& t;
    std::mutex& mu;
    std::condition_variable& cv;
    bool& stop;
    ~MergerStop() {
      if (!t.joinable()) return;
      {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
      }
      cv.notify_all();
      t.join();
    }
  } merger_stop{merger, mu, cv, stop};
  // a batch's handle, freed on every way out of its iteration
  struct Handle {
    tg_state* s = nullptr;
    ~Handle() {
      if (s) tg_free(s);
    }
  };
  // ---- batched construction on the GPU; the next batch's files are read meanwhile
  std::future<Rows> pending = std::async(std::launch::async, read_batch, 0u);
  for (uint32_t bi = 0; bi < nbatches; ++bi) {
    Rows cur = pending.get();
    if (bi + 1 < nbatches) pending = std::async(std::launch::async, read_batch, bi + 1);
    Handle h;
    auto bail = [&](int code, const std::string& m, tg_state*) {
      if (pending.valid()) pending.wait();
      return lfail(code, m);
    };
    if (cur.rc != TG_OK) return bail(cur.rc, cur.msg, nullptr);
    const uint32_t b0 = batches[bi].first, b1 = batches[bi].second;
    const size_t rows = cur.X.size() / D;
    int rc = TG_OK;
    if (!tuned) {
      // Every variant builds the same graphs, so the probes time them on the batch's first
      // clusters, about a million rows (at least one cluster), in a handle of their own that
      // is freed before the batch's.
      size_t nt = 1;
      while (nt < cur.keys.size() && cur.starts[nt + 1] <= TUNE_ROWS) ++nt;
      Handle t;
      if ((rc = tg_create(cur.starts[nt], D, p->devicelist[0], &t.s)) != TG_OK)
        return bail(rc, tg_last_error(), nullptr);
      if ((rc = tg_set_data(t.s, cur.X.data())) != TG_OK ||
          (rc = tg_set_segments(t.s, uint32_t(nt), cur.starts.data(), cur.keys.data())) != TG_OK)
        return bail(rc, tg_last_error(), nullptr);
      tg_tune_result tr{};
      if ((rc = tg_autotune_knn(t.s, K, p->seed, &tr)) != TG_OK)
        return bail(rc, tg_last_error(), nullptr);
      precision = tr.precision;
      tile = tr.tile;
      tuned = true;
      if (p->verbose)
        for (uint32_t c = 0; c < tr.ncand; ++c)
          printf("autotune: precision %d tile %u: %.2f ms\n", tr.cand_precision[c],
                 tr.cand_tile[c], tr.cand_ms[c]);
    }
    if ((rc = tg_create(uint32_t(rows), D, p->devicelist[0], &h.s)) != TG_OK)
      return bail(rc, tg_last_error(), nullptr);
    tg_state* s = h.s;
    if ((rc = tg_set_data(s, cur.X.data())) != TG_OK) return bail(rc, tg_last_error(), s);
    if ((rc = tg_set_segments(s, b1 - b0, cur.starts.data(), cur.keys.data())) != TG_OK)
      return bail(rc, tg_last_error(), s);
    tg_knn_params kp{};
    kp.K = K;
    kp.iters = p->iters ? p->iters : 16;
    kp.seed = p->seed;
    kp.precision = precision;
    kp.tile = tile;
    kp.init = TG_INIT_RANDOM;
    if ((rc = tg_knn_build(s, &kp)) != TG_OK) return bail(rc, tg_last_error(), s);
    tg_prune_params pp{};
    pp.kind = TG_KIND_PATH;
    pp.fwd_rule = pp.rev_rule = TG_RULE_DIST;
    pp.FD = FD;
    pp.TOPM = p->TOPM;
    pp.thre = p->thre;
    pp.precision = TG_PREC_FP32;
    if ((rc = tg_prune(s, &pp)) != TG_OK) return bail(rc, tg_last_error(), s);
    {
      tg_stats ts{};
      if ((rc = tg_get_stats(s, &ts)) != TG_OK) return bail(rc, tg_last_error(), s);
      device_ms += ts.knn_ms + ts.prune_ms;
    }
    std::vector<uint32_t> deg(rows), ids(rows * FD), eps(b1 - b0);
    uint32_t ep0 = 0;
    if ((rc = tg_get_pruned(s, deg.data(), ids.data(), &ep0)) != TG_OK)
      return bail(rc, tg_last_error(), s);
    if ((rc = tg_get_entry_points(s, eps.data())) != TG_OK) return bail(rc, tg_last_error(), s);
    tg_free(h.s);  // device memory back before the next batch's handle
    h.s = nullptr;
    // local -> global ids (large_index.cu:265-271)
    for (uint32_t k = b0; k < b1; ++k) {
      const uint32_t c = order[k];
      const uint32_t lo = cur.starts[k - b0];
      const std::vector<uint32_t>& mem = clusters[c];
      std::vector<uint32_t>& g = graph[c];
      g.assign(mem.size() * stride, 0u);
      for (size_t j = 0; j < mem.size(); ++j) {
        const uint32_t d = deg[lo + j];
        g[j * stride] = d;
        for (uint32_t t = 0; t < d; ++t) g[j * stride + 1 + t] = mem[ids[(lo + j) * FD + t] - lo];
      }
      metroid[k] = mem[eps[k - b0] - lo];
    }
    if (p->verbose) printf("built clusters %u..%u of %u (%zu rows)\n", b0, b1 - 1, C, rows);
    {
      std::lock_guard<std::mutex> lk(mu);
      built = b1;
    }
    cv.notify_all();
  }
  const auto t_merge = clk::now();
  merger.join();
  if (merge_rc != TG_OK) return lfail(merge_rc, merge_msg);
  // ---- final merge into the DiskANN layout (large_index.cu:423-521)
  FILE* out = fopen(p->index_store_path, "wb");
  if (!out) return lfail(TG_ERR_IO, std::string("cannot write ") + p->index_store_path);
  std::vector<uint32_t> obuf;
  obuf.reserve(1 << 20);
  uint64_t index_size = 0;
  const uint64_t zero64 = 0;
  const uint32_t md = p->max_degree;
  std::vector<uint32_t> nhood;
  bool ok = fwrite(&zero64, 8, 1, out) == 1 && fwrite(&md, 4, 1, out) == 1 &&
            fwrite(&md, 4, 1, out) == 1 && fwrite(&zero64, 8, 1, out) == 1;
  auto next = [&](uint32_t clu, const uint32_t** ids) -> uint32_t {
    Spill& sp = spill[clu];
    const size_t r = sp.cur++;
    if (r >= sp.deg.size()) return INV;
    *ids = sp.ids.data() + sp.off[r];
    return sp.deg[r];
  };
  auto flush = [&]() {
    if (!obuf.empty()) ok = ok && fwrite(obuf.data(), 4, obuf.size(), out) == obuf.size();
    obuf.clear();
  };
  for (uint32_t i = 0; i < Npts && ok; ++i) {
    const uint32_t* a = nullptr;
    const uint32_t* b = nullptr;
    if (cached[i]) {
      uint32_t d = next(cen[i], &a);
      if (d == 0) {
        d = next(cen2[i], &a);
      } else {
        const uint32_t d2 = next(cen2[i], &b);
        if (d2 != 0) {
          fclose(out);
          return lfail(TG_ERR_STATE, "merged point with two non-empty shard records");
        }
      }
      if (d == INV) break;
      obuf.push_back(d);
      obuf.insert(obuf.end(), a, a + d);
      index_size += uint64_t(1 + d) * 4;
    } else {
      const uint32_t d1 = next(cen[i], &a), d2 = next(cen2[i], &b);
      if (d1 == INV || d2 == INV) break;
      for (uint32_t t = 0; t < d1; ++t) {
        node_set[a[t]] = 1;
        nhood.push_back(a[t]);
      }
      for (uint32_t t = 0; t < d2; ++t)
        if (!node_set[b[t]]) {
          node_set[b[t]] = 1;
          nhood.push_back(b[t]);
        }
      for (uint32_t v : nhood) node_set[v] = 0;
      const uint32_t d = std::min<uint32_t>(uint32_t(nhood.size()), md);
      obuf.push_back(d);
      obuf.insert(obuf.end(), nhood.begin(), nhood.begin() + d);
      index_size += uint64_t(1 + d) * 4;
      nhood.clear();
    }
    if (obuf.size() > (1u << 20)) flush();
  }
  flush();
  index_size += 24;
  ok = ok && fseek(out, 0, SEEK_SET) == 0 && fwrite(&index_size, 8, 1, out) == 1 &&
       fwrite(&md, 4, 1, out) == 1 && fwrite(&metroid[C - 1], 4, 1, out) == 1;
  fclose(out);
  if (!ok) return lfail(TG_ERR_IO, std::string("short write to ") + p->index_store_path);
  if (out_stats) {
    const auto t_end = clk::now();
    out_stats->clusters = C;
    out_stats->batches = nbatches;
    out_stats->build_ms = std::chrono::duration<double, std::milli>(t_merge - t_build).count();
    out_stats->merge_ms = std::chrono::duration<double, std::milli>(t_end - t_merge).count();
    out_stats->total_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    out_stats->precision = precision;
    out_stats->tile = tile;
    out_stats->device_ms = device_ms;
  }
  return TG_OK;
}

extern "C" int tg_large_index(const tg_large_params* p, tg_large_stats* out_stats) {
  // host buffers of GB size (a batch's rows, the cluster graphs) can fail to allocate: that
  // and any other exception of the host side becomes a status, never an unwinding C call
  try {
    return large_index(p, out_stats);
  } catch (const std::bad_alloc&) {
    return lfail(TG_ERR_CAPACITY, "largeIndex: host memory exhausted");
  } catch (const std::exception& e) {
    return lfail(TG_ERR_STATE, std::string("largeIndex: ") + e.what());
  }
}
