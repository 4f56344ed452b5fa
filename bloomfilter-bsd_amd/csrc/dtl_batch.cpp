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

struct dtl_batch_plan {
  const BbfEntry* entry = nullptr;
  int device = 0;
  uint32_t n_tasks = 0;
  uint32_t total_tiles = 0;
  BatchTaskDev* d_tasks = nullptr;
  uint32_t* d_tile_task = nullptr;
  uint64_t* d_bits = nullptr;
  uint32_t* d_tile_cnt = nullptr;  // total_tiles counts + 1 work counter
  BbfBatchLaunch launch{};
};

static void plan_free(dtl_batch_plan* p) {
  if (!p) return;
  if (p->d_tasks) (void)hipFree(p->d_tasks);
  if (p->d_tile_task) (void)hipFree(p->d_tile_task);
  if (p->d_bits) (void)hipFree(p->d_bits);
  if (p->d_tile_cnt) (void)hipFree(p->d_tile_cnt);
  delete p;
}

extern "C" {

int bbf_batch_plan_create(const dtl_bbf_task* tasks, uint32_t n_tasks,
                          const dtl_probe_params* params, dtl_batch_plan** out) {
  if (!out || !tasks || n_tasks == 0) return DTL_EINVAL;
  *out = nullptr;
  const dtl_bbf_desc& d0 = tasks[0].desc;
  const BbfEntry* e = bbf_find(d0.word_bits, d0.w, d0.s, d0.z, d0.k);
  if (!e || !e->probe_batched) return DTL_EINVAL;
  int dev = 0, cus = 0;
  uint32_t max_lds = 0;
  int rc = current_device(&dev, &cus, &max_lds);
  if (rc != DTL_OK) return rc;

  std::vector<BatchTaskDev> host(n_tasks);
  uint64_t tiles = 0;
  for (uint32_t t = 0; t < n_tasks; ++t) {
    const dtl_bbf_task& T = tasks[t];
    const dtl_bbf_desc& d = T.desc;
    if (d.word_bits != d0.word_bits || d.w != d0.w || d.s != d0.s || d.z != d0.z || d.k != d0.k)
      return DTL_EINVAL;
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
    tiles += h.tile_count;
    if (tiles > 0xFFFFFFFFull) return DTL_EINVAL;
  }
  std::vector<uint32_t> tile_task(std::max<uint64_t>(tiles, 1), 0);
  for (uint32_t t = 0; t < n_tasks; ++t)
    std::fill(tile_task.begin() + host[t].tile_begin,
              tile_task.begin() + host[t].tile_begin + host[t].tile_count, t);

  dtl_batch_plan* p = new dtl_batch_plan();
  p->entry = e;
  p->device = dev;
  p->n_tasks = n_tasks;
  p->total_tiles = static_cast<uint32_t>(tiles);
  const size_t nt = std::max<size_t>(tiles, 1);
  if (hipMalloc(&p->d_tasks, sizeof(BatchTaskDev) * n_tasks) != hipSuccess ||
      hipMalloc(&p->d_tile_task, sizeof(uint32_t) * nt) != hipSuccess ||
      hipMalloc(&p->d_bits, nt * kTileWords * 8) != hipSuccess ||
      hipMalloc(&p->d_tile_cnt, (nt + 1) * sizeof(uint32_t)) != hipSuccess) {
    plan_free(p);
    return DTL_ENOMEM;
  }
  if (hipMemcpy(p->d_tasks, host.data(), sizeof(BatchTaskDev) * n_tasks, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(p->d_tile_task, tile_task.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice) !=
          hipSuccess) {
    plan_free(p);
    return DTL_EHIP;
  }
  BbfBatchLaunch& L = p->launch;
  L.tasks = p->d_tasks;
  L.tile_task = p->d_tile_task;
  L.n_tasks = n_tasks;
  L.total_tiles = p->total_tiles;
  L.bits = p->d_bits;
  L.tile_cnt = p->d_tile_cnt;
  L.work_counter = p->d_tile_cnt + nt;
  const uint32_t lds_cap = max_lds > 2048 ? max_lds - 2048 : 0;
  uint32_t lds_max = 32 * 1024;
  if (params && params->lds_max_bytes) lds_max = params->lds_max_bytes;
  if (params && (params->flags & DTL_PROBE_FORCE_GLOBAL)) lds_max = 0;
  L.lds_max_bytes = std::min(lds_max, lds_cap);
  L.threads = (params && params->block_threads >= 1024) ? 1024 : 256;
  L.grid = params ? params->grid_blocks : 0;
  L.chunk = (params && params->chunk_tiles) ? params->chunk_tiles : 1;
  L.cu_count = cus;
  *out = p;
  return DTL_OK;
}

int bbf_batch_plan_run(dtl_batch_plan* p, dtl_stream_t stream) {
  if (!p) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  BbfBatchLaunch L = p->launch;
  L.stream = st;
  (void)hipGetLastError();  // report only this call's failures
  hipError_t e = zero_u32(L.work_counter, st);
  if (e == hipSuccess && p->total_tiles > 0) e = p->entry->probe_batched(L);
  if (e == hipSuccess)
    e = compact_tasks(L.bits, L.tile_cnt, L.tasks, L.tile_task, p->n_tasks, p->total_tiles, st);
  if (e != hipSuccess) {
    set_last_error(e);
    return DTL_EHIP;
  }
  return DTL_OK;
}

void bbf_batch_plan_destroy(dtl_batch_plan* p) { plan_free(p); }

// Autotuner: times every candidate launch shape of the persistent probe
// kernel on the plan's own tasks (median of `reps` runs, HIP events on
// `stream`) and keeps the fastest.  Candidates: LDS staging threshold
// {0, 8, 16, 32, 64, 96, 150 KiB} x workgroup {256, 1024} threads x tiles per
// grab {1, 2, 4, 8, 16}; the grid is CUs x occupancy.
int bbf_batch_plan_tune(dtl_batch_plan* p, dtl_stream_t stream, uint32_t reps,
                        dtl_probe_params* best_out) {
  if (!p || reps == 0) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0, cus = 0;
  uint32_t max_lds = 0;
  int rc = current_device(&dev, &cus, &max_lds);
  if (rc != DTL_OK) return rc;
  const uint32_t lds_cap = max_lds > 2048 ? max_lds - 2048 : 0;
  const uint32_t lds_opts[] = {0, 8 << 10, 16 << 10, 32 << 10, 64 << 10, 96 << 10, 150 << 10};
  const uint32_t thr_opts[] = {256, 1024};
  const uint32_t chunk_opts[] = {1, 2, 4, 8, 16};
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return DTL_EHIP;
  const BbfBatchLaunch saved = p->launch;
  float best_ms = 1e30f;
  BbfBatchLaunch best = saved;
  rc = DTL_OK;
  for (uint32_t lds : lds_opts) {
    if (lds > lds_cap) continue;
    for (uint32_t thr : thr_opts) {
      for (uint32_t ch : chunk_opts) {
        p->launch.lds_max_bytes = lds;
        p->launch.threads = thr;
        p->launch.chunk = ch;
        p->launch.grid = 0;
        std::vector<float> t;
        for (uint32_t r = 0; r < reps + 1 && rc == DTL_OK; ++r) {
          if (hipEventRecord(a, st) != hipSuccess) rc = DTL_EHIP;
          if (rc == DTL_OK) rc = bbf_batch_plan_run(p, stream);
          if (rc == DTL_OK && hipEventRecord(b, st) != hipSuccess) rc = DTL_EHIP;
          if (rc == DTL_OK && hipEventSynchronize(b) != hipSuccess) rc = DTL_EHIP;
          float ms = 0;
          if (rc == DTL_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = DTL_EHIP;
          if (r > 0) t.push_back(ms);  // the first run warms up
        }
        if (rc != DTL_OK) break;
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (med < best_ms) {
          best_ms = med;
          best = p->launch;
        }
      }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  p->launch = (rc == DTL_OK) ? best : saved;
  if (best_out) {
    std::memset(best_out, 0, sizeof(*best_out));
    best_out->lds_max_bytes = p->launch.lds_max_bytes;
    best_out->block_threads = p->launch.threads;
    best_out->grid_blocks = p->launch.grid;
    best_out->chunk_tiles = p->launch.chunk;
    best_out->flags = p->launch.lds_max_bytes == 0 ? DTL_PROBE_FORCE_GLOBAL : 0;
  }
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
