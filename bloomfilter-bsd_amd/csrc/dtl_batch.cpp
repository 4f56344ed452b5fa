// dtl_batch.cpp -- batch plans: many filters, each with its own probe batch,
// probed by one persistent kernel (dtl_batched.hpp) plus the shared
// compaction phases.  The plan owns every device buffer the run needs, so a
// run allocates nothing and can be captured into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "dtl_hip.h"
#include "dtl_internal.hpp"

using namespace dtlh;

// One (configuration, size class) share of a plan: its tasks, tile table and scratch.
struct PlanGroup {
  const BbfEntry* entry = nullptr;
  uint32_t size_class = 0;
  uint32_t max_filter_bytes = 0;
  uint32_t n_tasks = 0;
  uint32_t total_tiles = 0;
  BatchTaskDev* d_tasks = nullptr;
  uint32_t* d_tile_task = nullptr;
  uint64_t* d_bits = nullptr;
  uint32_t* d_tile_cnt = nullptr;  // total_tiles counts + 1 work counter
  BbfBatchLaunch launch{};
  uint64_t keys = 0;               // probe keys in the group (largest group first)
};

struct dtl_batch_plan {
  int device = 0;
  std::vector<PlanGroup> groups;   // one per distinct ((word_bits, w, s, z, k), size class)
  // groups 1.. run on side streams forked from the caller's stream (default
  // priority), joined back before the run returns; created on the first run
  std::vector<hipStream_t> side;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
};

namespace {

void group_free(PlanGroup& g) {
  if (g.d_tasks) (void)hipFree(g.d_tasks);
  if (g.d_tile_task) (void)hipFree(g.d_tile_task);
  if (g.d_bits) (void)hipFree(g.d_bits);
  if (g.d_tile_cnt) (void)hipFree(g.d_tile_cnt);
  g = PlanGroup{};
}

void plan_free(dtl_batch_plan* p) {
  if (!p) return;
  for (PlanGroup& g : p->groups) group_free(g);
  for (hipStream_t s : p->side)
    if (s) (void)hipStreamDestroy(s);
  for (hipEvent_t e : p->join)
    if (e) (void)hipEventDestroy(e);
  if (p->fork) (void)hipEventDestroy(p->fork);
  delete p;
}

// Side streams and fork / join events for the plan's groups 1.. (once).
// Groups after the first run on at most kSideStreams side streams (round
// robin), so that a plan of hundreds of groups does not create hundreds of
// streams: HIP spreads streams over 4 hardware queues per process by default.
constexpr size_t kSideStreams = 3;

hipError_t plan_streams(dtl_batch_plan* p) {
  const size_t need = std::min(p->groups.size() > 1 ? p->groups.size() - 1 : 0, kSideStreams);
  if (p->side.size() == need && (need == 0 || p->fork)) return hipSuccess;
  hipError_t e = hipSuccess;
  if (!p->fork) e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
  while (e == hipSuccess && p->side.size() < need) {
    hipStream_t s = nullptr;
    hipEvent_t j = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if (s) p->side.push_back(s);
    if (j) p->join.push_back(j);
  }
  return e;
}

bool same_config(const dtl_bbf_desc& a, const dtl_bbf_desc& b) {
  return a.word_bits == b.word_bits && a.w == b.w && a.s == b.s && a.z == b.z && a.k == b.k;
}

// Builds one group from tasks that share a configuration and size class.
int group_create(const dtl_bbf_task* tasks, const std::vector<uint32_t>& ids, const dtl_probe_params* params,
                 uint32_t cus, uint32_t max_lds, uint32_t cls, PlanGroup& g) {
  const dtl_bbf_desc& d0 = tasks[ids[0]].desc;
  const BbfEntry* e = bbf_find(d0.word_bits, d0.w, d0.s, d0.z, d0.k);
  if (!e || !e->probe_batched) return DTL_EINVAL;
  const uint32_t n_tasks = static_cast<uint32_t>(ids.size());
  std::vector<BatchTaskDev> host(n_tasks);
  uint64_t tiles = 0, keys = 0;
  uint32_t max_fb = 0;
  for (uint32_t t = 0; t < n_tasks; ++t) {
    const dtl_bbf_task& T = tasks[ids[t]];
    const dtl_bbf_desc& d = T.desc;
    if (d.block_cnt < 2 || !T.d_words || !T.d_count || T.n > 0xFFFFFFFFull) return DTL_EINVAL;
    if (T.n && (!T.d_keys || !T.d_pos)) return DTL_EINVAL;
    const uint32_t bb = d.w * d.word_bits / 8;
    if (reinterpret_cast<uintptr_t>(T.d_words) % std::min(16u, bb)) return DTL_EINVAL;
    BatchTaskDev& h = host[t];
    h.words = T.d_words;
    h.keys = T.d_keys;
    h.pos = T.d_pos;
    h.count = T.d_count;
    h.n = static_cast<uint32_t>(T.n);
    h.tile_begin = static_cast<uint32_t>(tiles);
    h.tile_count = static_cast<uint32_t>((T.n + kTileKeys - 1) / kTileKeys);
    h.match_offset = T.match_offset;
    if (d.addr_mode == DTL_ADDR_POW2) {
      h.addr_magic = 0;
      h.addr_mshift = 0;
      h.addr_divisor = 0;
      h.addr_post = d.block_shift;
    } else {
      h.addr_magic = d.magic;
      h.addr_mshift = d.magic_shift;
      h.addr_divisor = d.block_cnt;
      h.addr_post = 0;
    }
    h.filter_bytes = static_cast<uint32_t>(std::min<uint64_t>(d.size_bytes, 0xFFFFFFFFull));
    max_fb = std::max(max_fb, h.filter_bytes);
    tiles += h.tile_count;
    keys += T.n;
    if (tiles > 0xFFFFFFFFull) return DTL_EINVAL;
  }
  std::vector<uint32_t> tile_task(std::max<uint64_t>(tiles, 1), 0);
  for (uint32_t t = 0; t < n_tasks; ++t)
    std::fill(tile_task.begin() + host[t].tile_begin,
              tile_task.begin() + host[t].tile_begin + host[t].tile_count, t);

  g.entry = e;
  g.size_class = cls;
  g.max_filter_bytes = max_fb;
  g.n_tasks = n_tasks;
  g.total_tiles = static_cast<uint32_t>(tiles);
  g.keys = keys;
  const size_t nt = std::max<size_t>(tiles, 1);
  if (hipMalloc(&g.d_tasks, sizeof(BatchTaskDev) * n_tasks) != hipSuccess ||
      hipMalloc(&g.d_tile_task, sizeof(uint32_t) * nt) != hipSuccess ||
      hipMalloc(&g.d_bits, nt * kTileWords * 8) != hipSuccess ||
      hipMalloc(&g.d_tile_cnt, (nt + 1) * sizeof(uint32_t)) != hipSuccess)
    return DTL_ENOMEM;
  if (hipMemcpy(g.d_tasks, host.data(), sizeof(BatchTaskDev) * n_tasks, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(g.d_tile_task, tile_task.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice) != hipSuccess)
    return DTL_EHIP;
  BbfBatchLaunch& L = g.launch;
  L.tasks = g.d_tasks;
  L.tile_task = g.d_tile_task;
  L.n_tasks = n_tasks;
  L.total_tiles = g.total_tiles;
  L.bits = g.d_bits;
  L.tile_cnt = g.d_tile_cnt;
  L.work_counter = g.d_tile_cnt + nt;
  const uint32_t lds_cap = max_lds > 2048 ? max_lds - 2048 : 0;
  // the class's own shape (the size of its largest filter staged, or nothing
  // staged), unless the caller sets one
  uint32_t lds_max = cls == 2 ? 0u : (max_fb + 15u) & ~15u;
  if (params && params->lds_max_bytes) lds_max = params->lds_max_bytes;
  if (params && (params->flags & DTL_PROBE_FORCE_GLOBAL)) lds_max = 0;
  L.lds_max_bytes = std::min(lds_max, lds_cap);
  uint32_t threads = (cls == 1 && d0.word_bits == 32) ? 1024u : 256u;
  if (params && params->block_threads) threads = params->block_threads >= 1024 ? 1024u : 256u;
  L.threads = threads;
  L.grid = params ? params->grid_blocks : 0;
  L.chunk = (params && params->chunk_tiles) ? params->chunk_tiles : 1;
  L.half_in_flight = (params && (params->flags & DTL_PROBE_HALF_IN_FLIGHT)) ? 1u : 0u;
  L.cu_count = cus;
  return DTL_OK;
}

hipError_t group_run(const PlanGroup& g, hipStream_t st) {
  BbfBatchLaunch L = g.launch;
  L.stream = st;
  hipError_t e = zero_u32(L.work_counter, st);
  if (e == hipSuccess && g.total_tiles > 0) e = g.entry->probe_batched(L);
  if (e == hipSuccess) e = compact_tasks(L.bits, L.tile_cnt, L.tasks, L.tile_task, g.n_tasks, g.total_tiles, st);
  return e;
}

void copy_shape(const BbfBatchLaunch& L, dtl_probe_params* out) {
  std::memset(out, 0, sizeof(*out));
  out->lds_max_bytes = L.lds_max_bytes;
  out->block_threads = L.threads;
  out->grid_blocks = L.grid;
  out->chunk_tiles = L.chunk;
  out->flags = (L.lds_max_bytes == 0 ? DTL_PROBE_FORCE_GLOBAL : 0u) |
               (L.half_in_flight ? DTL_PROBE_HALF_IN_FLIGHT : 0u);
}

}  // namespace

extern "C" {

// Tasks may mix configurations and filter sizes: they are grouped by
// ((word_bits, w, s, z, k), size class), in order of first appearance, and
// each group gets its own persistent launch; every task's list is the same as
// bbf_probe's for it either way.
int bbf_batch_plan_create(const dtl_bbf_task* tasks, uint32_t n_tasks,
                          const dtl_probe_params* params, dtl_batch_plan** out) {
  if (!out || !tasks || n_tasks == 0) return DTL_EINVAL;
  *out = nullptr;
  int dev = 0, cus = 0;
  uint32_t max_lds = 0;
  int rc = current_device(&dev, &cus, &max_lds);
  if (rc != DTL_OK) return rc;
  const uint32_t lds_cap = max_lds > 2048 ? max_lds - 2048 : 0;
  std::vector<std::vector<uint32_t>> ids;
  std::vector<uint32_t> rep, cls;  // first task and size class of each group
  for (uint32_t t = 0; t < n_tasks; ++t) {
    const uint32_t c = size_class(tasks[t].desc.size_bytes, lds_cap);
    size_t g = 0;
    while (g < rep.size() && !(cls[g] == c && same_config(tasks[rep[g]].desc, tasks[t].desc))) ++g;
    if (g == rep.size()) {
      rep.push_back(t);
      cls.push_back(c);
      ids.emplace_back();
    }
    ids[g].push_back(t);
  }
  dtl_batch_plan* p = new dtl_batch_plan();
  p->device = dev;
  p->groups.resize(ids.size());
  for (size_t g = 0; g < ids.size(); ++g) {
    rc = group_create(tasks, ids[g], params, static_cast<uint32_t>(cus), max_lds, cls[g], p->groups[g]);
    if (rc != DTL_OK) {
      plan_free(p);
      return rc;
    }
  }
  *out = p;
  return DTL_OK;
}

// The groups are independent (each has its own tasks, bitmap, counts and
// work counter), so groups 1.. run on side streams forked from `stream` by an
// event and joined back into it: a group's persistent kernel takes the CUs
// that another group's tail leaves idle, and a group's scan and expand
// overlap the others' probing.  Captured into a hipGraph, the fork and join
// become graph dependencies.
int bbf_batch_plan_run(dtl_batch_plan* p, dtl_stream_t stream) {
  if (!p) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  (void)hipGetLastError();  // report only this call's failures
  hipError_t e = plan_streams(p);
  const size_t ng = p->groups.size();
  const size_t ns = p->side.size();
  if (e == hipSuccess && ns > 0) e = hipEventRecord(p->fork, st);
  for (size_t i = 0; i < ns && e == hipSuccess; ++i) e = hipStreamWaitEvent(p->side[i], p->fork, 0);
  for (size_t g = 1; g < ng && e == hipSuccess; ++g) e = group_run(p->groups[g], p->side[(g - 1) % ns]);
  if (e == hipSuccess && ng > 0) e = group_run(p->groups[0], st);
  for (size_t i = 0; i < ns && e == hipSuccess; ++i) {
    e = hipEventRecord(p->join[i], p->side[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, p->join[i], 0);
  }
  if (e != hipSuccess) {
    set_last_error(e);
    return DTL_EHIP;
  }
  return DTL_OK;
}

uint32_t bbf_batch_plan_group_count(const dtl_batch_plan* p) {
  return p ? static_cast<uint32_t>(p->groups.size()) : 0u;
}

int bbf_batch_plan_group_shape(const dtl_batch_plan* p, uint32_t group, dtl_probe_params* shape,
                               uint32_t* n_tasks, uint64_t* keys, uint32_t* max_filter_bytes) {
  if (!p || group >= p->groups.size()) return DTL_EINVAL;
  const PlanGroup& g = p->groups[group];
  if (shape) copy_shape(g.launch, shape);
  if (n_tasks) *n_tasks = g.n_tasks;
  if (keys) *keys = g.keys;
  if (max_filter_bytes) *max_filter_bytes = g.max_filter_bytes;
  return DTL_OK;
}

void bbf_batch_plan_destroy(dtl_batch_plan* p) { plan_free(p); }

// Autotuner: for each group, times every candidate launch shape of the
// persistent probe kernel on the group's own tasks (median of `reps` runs,
// HIP events on `stream`) and keeps the fastest.  Candidates: LDS staging
// threshold {0, the group's largest filter} x workgroup {256, 1024} threads x
// keys per lane in flight {the kernel's full count, half of it} x tiles per
// grab {1, 2, 4, 8, 16} x grid {CUs x occupancy, half of it (at least one
// workgroup per CU)}.  *best_out receives the shape of the group with the
// most keys.
int bbf_batch_plan_tune(dtl_batch_plan* p, dtl_stream_t stream, uint32_t reps,
                        dtl_probe_params* best_out) {
  if (!p || reps == 0) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0, cus = 0;
  uint32_t max_lds = 0;
  int rc = current_device(&dev, &cus, &max_lds);
  if (rc != DTL_OK) return rc;
  (void)hipGetLastError();
  const uint32_t lds_cap = max_lds > 2048 ? max_lds - 2048 : 0;
  const uint32_t thr_opts[] = {256, 1024};
  const uint32_t chunk_opts[] = {1, 2, 4, 8, 16};
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    return DTL_EHIP;
  }
  size_t largest = 0;
  for (size_t gi = 0; gi < p->groups.size() && rc == DTL_OK; ++gi) {
    PlanGroup& g = p->groups[gi];
    if (g.keys > p->groups[largest].keys) largest = gi;
    const BbfBatchLaunch saved = g.launch;
    float best_ms = 1e30f;
    BbfBatchLaunch best = saved;
    const uint32_t staged = (g.max_filter_bytes + 15u) & ~15u;
    const uint32_t lds_opts[] = {0u, staged <= lds_cap ? staged : 0u};
    for (uint32_t li = 0; li < 2; ++li) {
      const uint32_t lds = lds_opts[li];
      if (rc != DTL_OK || (li == 1 && lds == 0)) continue;
      for (uint32_t thr : thr_opts)
      for (uint32_t half = 0; half < 2; ++half) {
        // the full grid (CUs x occupancy) and half of it
        uint32_t full = 0;
        {
          BbfBatchLaunch probe = g.launch;
          probe.lds_max_bytes = lds;
          probe.threads = thr;
          probe.half_in_flight = half;
          full = g.entry->batched_grid(probe);
        }
        const uint32_t grid_opts[] = {full, std::max(static_cast<uint32_t>(cus), full / 2)};
        for (uint32_t gi2 = 0; gi2 < 2 && rc == DTL_OK; ++gi2) {
          if (gi2 == 1 && grid_opts[1] >= grid_opts[0]) continue;
        for (uint32_t ch : chunk_opts) {
          g.launch.lds_max_bytes = lds;
          g.launch.threads = thr;
          g.launch.half_in_flight = half;
          g.launch.chunk = ch;
          g.launch.grid = grid_opts[gi2];
          std::vector<float> t;
          for (uint32_t r = 0; r < reps + 1 && rc == DTL_OK; ++r) {
            if (hipEventRecord(a, st) != hipSuccess || group_run(g, st) != hipSuccess ||
                hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess) {
              rc = DTL_EHIP;
              break;
            }
            float ms = 0;
            if (hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = DTL_EHIP;
            if (r > 0) t.push_back(ms);  // the first run warms up
          }
          if (rc != DTL_OK) break;
          std::sort(t.begin(), t.end());
          const float med = t[t.size() / 2];
          if (med < best_ms) {
            best_ms = med;
            best = g.launch;
          }
        }
        }
      }
    }
    g.launch = (rc == DTL_OK) ? best : saved;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc == DTL_EHIP) set_last_error(hipGetLastError());
  if (best_out && !p->groups.empty()) copy_shape(p->groups[largest].launch, best_out);
  return rc;
}

int bbf_probe_batched(const dtl_bbf_task* tasks, uint32_t n_tasks, const dtl_probe_params* params,
                      dtl_stream_t stream) {
  dtl_batch_plan* p = nullptr;
  int rc = bbf_batch_plan_create(tasks, n_tasks, params, &p);
  if (rc != DTL_OK) return rc;
  rc = bbf_batch_plan_run(p, stream);
  if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess && rc == DTL_OK)
    rc = DTL_EHIP;
  plan_free(p);
  return rc;
}

}  // extern "C"
