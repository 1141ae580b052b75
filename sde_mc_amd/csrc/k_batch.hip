// k_batch.hip -- the batched persistent launch: one kernel runs a segment
// list spanning any number of jobs (different models, grids, path counts,
// step counts and MLMC levels).  Each wavefront repeatedly takes the next
// tile index from a device counter and finds its segment by binary search,
// so long and short tiles balance across the 256 CUs without host
// involvement; the host orders segments costliest-first.
#include <string.h>

#include <map>
#include <mutex>

#include "k_batch_kernels.h"

namespace sdemc {

int validate_job(const sdemc_job* job);

#define SDEMC_EXTERN_GROUP(CLS, U) extern template SDEMC_LAUNCH_GROUP_SIG(CLS, U)
SDEMC_BATCH_INSTANCES(SDEMC_EXTERN_GROUP)
#undef SDEMC_EXTERN_GROUP

// Zero the launch's n tile counters and its fault word (counter[SDEMC_MAX_GROUPS]).
__global__ void k_zero_counter(int32_t* counter, int32_t n) {
  if ((int32_t)threadIdx.x < n) counter[threadIdx.x] = 0;
  if (threadIdx.x == 0) counter[SDEMC_MAX_GROUPS] = 0;
}

// Fixed-order finalisation of the slots flagged this launch.  A slot that is
// not prepared or whose chunks lie outside the partial pool gets NaN sums and
// flag SDEMC_FLAG_TABLE; so does every slot when the launch's fault word is set.
// The chunk flags it reads are reset to zero (the group kernels write only
// set flags).
__global__ __launch_bounds__(256) void k_finalize_slots(const KJob* __restrict__ jobs,
                                                        const int32_t* __restrict__ mask,
                                                        const double* partials,
                                                        int32_t* cflags,
                                                        const int32_t* __restrict__ fault,
                                                        int64_t partial_chunks, double* stats) {
  const int j = blockIdx.x;
  if (mask[j] == 0) return;
  const KJob& job = jobs[j];
  double* st = stats + 4 * (int64_t)j;
  const int64_t nc = (int64_t)((job.n_paths + WAVE - 1) / WAVE);
  if (job.magic != KJOB_MAGIC || job.chunk_base < 0 || job.chunk_base + nc > partial_chunks) {
    if (threadIdx.x == 0) {
      st[0] = __builtin_nan("");
      st[1] = __builtin_nan("");
      st[2] = 0.0;
      st[3] = (double)SDEMC_FLAG_TABLE;
    }
    return;
  }
  finalize_block(partials, cflags, job.chunk_base, nc, job.n_paths, st,
                 *fault ? SDEMC_FLAG_TABLE : 0, cflags);
}

static int check_batch(const sdemc_batch_dev* b, const sdemc_batch_tune* tune) {
  if (!b || !tune) return SDEMC_ERR_ARG;
  if (!b->jobs_dev || !b->segs_dev || !b->seg_hdr_dev || !b->finalize_dev || !b->partials_dev ||
      !b->chunk_flags_dev || !b->stats_dev || !b->counter_dev)
    return SDEMC_ERR_ARG;
  if (b->max_jobs < 1 || b->max_segments < 1 || b->partial_chunks < 0) return SDEMC_ERR_ARG;
  if (b->n_groups < 1 || b->n_groups > SDEMC_MAX_GROUPS) return SDEMC_ERR_ARG;
  for (int g = 0; g < b->n_groups; ++g)
    if (b->group_class[g] < 0 || b->group_class[g] > SDEMC_CLASS_GENERIC) return SDEMC_ERR_ARG;
  const int w = tune->waves_per_block;
  if (w != 1 && w != 2 && w != 4 && w != 8) return SDEMC_ERR_ARG;
  if (tune->blocks_per_cu < 1 || tune->blocks_per_cu > 64) return SDEMC_ERR_ARG;
  if (tune->step_unroll < 0 || tune->step_unroll > 2) return SDEMC_ERR_ARG;
  return SDEMC_OK;
}

static int cu_count_cached() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    cus = prop.multiProcessorCount;
  }
  return cus;
}


// Streams and events that fork a round's group kernels off the caller's
// stream and join them back (default priority, non-blocking).
struct Fork {
  hipStream_t side[SDEMC_MAX_GROUPS - 1];
  hipEvent_t fork;
  hipEvent_t join[SDEMC_MAX_GROUPS - 1];
  bool ok;
};

static bool fork_create(Fork& f) {
  memset(&f, 0, sizeof(f));
  f.ok = hipEventCreateWithFlags(&f.fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; f.ok && i < SDEMC_MAX_GROUPS - 1; ++i) {
    f.ok = hipStreamCreateWithFlags(&f.side[i], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&f.join[i], hipEventDisableTiming) == hipSuccess;
  }
  return f.ok;
}

static void fork_destroy(Fork& f) {
  for (int i = 0; i < SDEMC_MAX_GROUPS - 1; ++i) {
    if (f.join[i]) (void)hipEventDestroy(f.join[i]);
    if (f.side[i]) (void)hipStreamDestroy(f.side[i]);
  }
  if (f.fork) (void)hipEventDestroy(f.fork);
  memset(&f, 0, sizeof(f));
}

// The process-wide fork of each device for launches outside a graph.
static Fork* device_fork() {
  static std::mutex mu;
  static std::map<int, Fork> forks;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = forks.find(dev);
  if (it == forks.end()) {
    Fork f;
    if (!fork_create(f)) {
      fork_destroy(f);
      return nullptr;
    }
    it = forks.emplace(dev, f).first;
  }
  return &it->second;
}

static void launch_class(const sdemc_batch_dev& b, int g, unsigned grid, int wpb, int unroll,
                         hipStream_t s) {
  switch (b.group_class[g]) {
    case 0:
      if (unroll == 2) launch_group<0, 2>(b, g, grid, wpb, s);
      else launch_group<0, 1>(b, g, grid, wpb, s);
      break;
    case 1: launch_group<1, 1>(b, g, grid, wpb, s); break;
    case 2: launch_group<2, 1>(b, g, grid, wpb, s); break;
    case 3: launch_group<3, 1>(b, g, grid, wpb, s); break;
    case 4:
      if (unroll == 2) launch_group<4, 2>(b, g, grid, wpb, s);
      else launch_group<4, 1>(b, g, grid, wpb, s);
      break;
    case 5: launch_group<5, 1>(b, g, grid, wpb, s); break;
    case 6: launch_group<6, 1>(b, g, grid, wpb, s); break;
    case 7: launch_group<7, 1>(b, g, grid, wpb, s); break;
    case SDEMC_CLASS_INDEP2_CLOCK:
      if (unroll == 2) launch_group<SDEMC_CLASS_INDEP2_CLOCK, 2>(b, g, grid, wpb, s);
      else launch_group<SDEMC_CLASS_INDEP2_CLOCK, 1>(b, g, grid, wpb, s);
      break;
    default: launch_group<SDEMC_CLASS_GENERIC, 1>(b, g, grid, wpb, s); break;
  }
}

// Zero the tile counters, run the group kernels -- one after another on s,
// or (fork != null) group 0 on s and group g > 0 on fork->side[g - 1], all
// after the counter reset and joined back into s -- then finalise.
static int enqueue_batch(const sdemc_batch_dev& b, const sdemc_batch_tune& tune, hipStream_t s,
                         Fork* fork) {
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, b.counter_dev, b.n_groups);
  const int wpb = tune.waves_per_block;
  const int unroll = tune.step_unroll;
  const unsigned grid = (unsigned)(cu_count_cached() * tune.blocks_per_cu);
  if (fork && b.n_groups > 1) {
    if (hipEventRecord(fork->fork, s) != hipSuccess) return SDEMC_ERR_LAUNCH;
    for (int g = 1; g < b.n_groups; ++g)
      if (hipStreamWaitEvent(fork->side[g - 1], fork->fork, 0) != hipSuccess) return SDEMC_ERR_LAUNCH;
    launch_class(b, 0, grid, wpb, unroll, s);
    for (int g = 1; g < b.n_groups; ++g) {
      launch_class(b, g, grid, wpb, unroll, fork->side[g - 1]);
      if (hipEventRecord(fork->join[g - 1], fork->side[g - 1]) != hipSuccess) return SDEMC_ERR_LAUNCH;
    }
    for (int g = 1; g < b.n_groups; ++g)
      if (hipStreamWaitEvent(s, fork->join[g - 1], 0) != hipSuccess) return SDEMC_ERR_LAUNCH;
  } else {
    for (int g = 0; g < b.n_groups; ++g) launch_class(b, g, grid, wpb, unroll, s);
  }
  hipLaunchKernelGGL(k_finalize_slots, dim3((unsigned)b.max_jobs), dim3(256), 0, s,
                     (const KJob*)b.jobs_dev, b.finalize_dev, b.partials_dev, b.chunk_flags_dev,
                     b.counter_dev + SDEMC_MAX_GROUPS, b.partial_chunks, b.stats_dev);
  return SDEMC_OK;
}

}  // namespace sdemc

using namespace sdemc;

struct sdemc_batch_plan {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" {

size_t sdemc_sizeof_prepared_job(void) { return sizeof(KJob); }

int sdemc_job_class(const sdemc_job* job) {
  if (validate_job(job) != SDEMC_OK) return -1;
  return job_class_of(*job);
}

int sdemc_prepare_jobs(const sdemc_job* jobs, int32_t n, const int64_t* chunk_base,
                       void* prepared_out, int32_t* classes) {
  if (n < 0 || (n > 0 && (!jobs || !chunk_base || !prepared_out))) return SDEMC_ERR_ARG;
  for (int32_t i = 0; i < n; ++i) {
    const int st = validate_job(&jobs[i]);
    if (st != SDEMC_OK) return st;
    if (chunk_base[i] < 0) return SDEMC_ERR_ARG;
    KJob k;
    memset(&k, 0, sizeof(k));
    make_kjob(jobs[i], k);
    k.chunk_base = chunk_base[i];
    memcpy((char*)prepared_out + (size_t)i * sizeof(KJob), &k, sizeof(KJob));
    if (classes)
      classes[i] = job_class_of(jobs[i]);
  }
  return SDEMC_OK;
}

int sdemc_batch_check(const void* prepared_host, int32_t max_jobs, const sdemc_segment* segs,
                      int32_t max_segments, const int32_t* hdr, int32_t n_groups,
                      const int32_t* group_class, int64_t partial_chunks) {
  if (!prepared_host || !segs || !hdr || !group_class) return SDEMC_ERR_ARG;
  if (max_jobs < 1 || max_segments < 1 || partial_chunks < 0) return SDEMC_ERR_ARG;
  if (n_groups < 1 || n_groups > SDEMC_MAX_GROUPS) return SDEMC_ERR_ARG;
  const KJob* jobs = (const KJob*)prepared_host;
  for (int g = 0; g < n_groups; ++g) {
    const int32_t cls = group_class[g];
    if (cls < 0 || cls > SDEMC_CLASS_GENERIC) return SDEMC_ERR_ARG;
    const int32_t first = hdr[4 * g], n = hdr[4 * g + 1], tiles = hdr[4 * g + 2];
    if (n == 0 && tiles == 0) continue;
    if (first < 0 || n < 1 || first > max_segments - n || tiles < 1) return SDEMC_ERR_ARG;
    int64_t next_tile = 0;
    for (int32_t i = first; i < first + n; ++i) {
      const sdemc_segment& sg = segs[i];
      if (sg.tile_begin != next_tile) return SDEMC_ERR_ARG;
      if (sg.job < 0 || sg.job >= max_jobs || sg.per_tile < 1 || sg.count < 1 || sg.first < 0)
        return SDEMC_ERR_ARG;
      const KJob& k = jobs[sg.job];
      if (k.magic != KJOB_MAGIC) return SDEMC_ERR_ARG;
      if (cls != SDEMC_CLASS_GENERIC && kjob_class_of(k) != cls) return SDEMC_ERR_ARG;
      const int64_t nc = (int64_t)((k.n_paths + WAVE - 1) / WAVE);
      const int64_t end = (int64_t)sg.first + sg.count;
      if (end > nc || k.chunk_base < 0 || k.chunk_base + end > partial_chunks) return SDEMC_ERR_ARG;
      next_tile += (sg.count + sg.per_tile - 1) / sg.per_tile;
    }
    if (next_tile != tiles) return SDEMC_ERR_ARG;
  }
  return SDEMC_OK;
}

int sdemc_batch_launch(const sdemc_batch_dev* b, const sdemc_batch_tune* tune, void* stream) {
  int st = check_batch(b, tune);
  if (st != SDEMC_OK) return st;
  Fork* fork = nullptr;
  if (tune->concurrent_groups && b->n_groups > 1) {
    fork = device_fork();
    if (!fork) return SDEMC_ERR_LAUNCH;
  }
  st = enqueue_batch(*b, *tune, (hipStream_t)stream, fork);
  const hipError_t e = hipGetLastError();
  if (st != SDEMC_OK) return st;
  return e == hipSuccess ? SDEMC_OK : SDEMC_ERR_LAUNCH;
}

int sdemc_batch_plan_create(const sdemc_batch_dev* b, const sdemc_batch_tune* tune,
                            sdemc_batch_plan** plan) {
  if (!plan) return SDEMC_ERR_ARG;
  *plan = nullptr;
  int st = check_batch(b, tune);
  if (st != SDEMC_OK) return st;
  (void)cu_count_cached();
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return SDEMC_ERR_LAUNCH;
  Fork fork;
  const bool forked = tune->concurrent_groups && b->n_groups > 1;
  if (forked && !fork_create(fork)) {
    fork_destroy(fork);
    (void)hipStreamDestroy(cs);
    return SDEMC_ERR_LAUNCH;
  }
  sdemc_batch_plan* p = new sdemc_batch_plan();
  p->graph = nullptr;
  p->exec = nullptr;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const int qs = enqueue_batch(*b, *tune, cs, forked ? &fork : nullptr);
    e = hipStreamEndCapture(cs, &p->graph);
    if (e == hipSuccess && qs != SDEMC_OK) e = hipErrorUnknown;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (forked) fork_destroy(fork);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    delete p;
    return SDEMC_ERR_LAUNCH;
  }
  *plan = p;
  return SDEMC_OK;
}

int sdemc_batch_plan_launch(sdemc_batch_plan* plan, void* stream) {
  if (!plan || !plan->exec) return SDEMC_ERR_ARG;
  const hipError_t e = hipGraphLaunch(plan->exec, (hipStream_t)stream);
  return e == hipSuccess ? SDEMC_OK : SDEMC_ERR_LAUNCH;
}

int sdemc_batch_plan_destroy(sdemc_batch_plan* plan) {
  if (!plan) return SDEMC_OK;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  delete plan;
  return SDEMC_OK;
}

}  // extern "C"
