// batch.hip — batched solves with continuous admission (nz_batch_*, include/nozzle.h).
//
// Resident solves of any grid size share every launch: the stage kernels run one
// workgroup per tile over the concatenated tile ranges of all residents (each workgroup
// reads its instance from a per-tile table), the spray kernel over their concatenated
// parcel blocks, the close kernel one workgroup per resident. Per-instance stop tests
// (simulated time, residual, step count) run on the device at the start of each step's
// stage 1 and in the close kernel after the run; a stopped instance's workgroups return
// at once, so nothing waits on the host between steps. The step sequence of a run is
// captured in a HIP graph and replayed while the resident set and step count stay the same.
//
// The host stays off the device's critical path. Every resident's buffers come from one
// device slab allocated at nz_batch_create (host bookkeeping, no allocation call per
// solve). Admissions upload on a stream of their own. A run call enqueues its chunk of
// steps first and only then waits for the previous chunk's stop flags (copied with it
// into pinned memory), so the device works on one chunk while the host reads the last
// one, retires the solves that stopped and admits queued jobs. A stopped solve's
// integrals and kept state are enqueued behind the chunk in flight and collected at a
// later call, when its slot and memory (which that chunk still names) are released.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "nz_inst.h"
#include "nz_launch.h"

using nzk::Consts;
using nzk::DevInst;
using nzk::InstState;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// a submitted job: the caller's arrays copied, until it is admitted
struct Job {
  int64_t id = 0;
  int Nx = 0, Nr = 0;
  std::vector<uint8_t> mask;
  std::vector<float> geoA, geoG, geoM, jWall, A, B, tab;
  nz_params P{};
  int parcels = 0;
  nz_probes probes{};
  double tEnd = 0, resTol = 0;
  int minSteps = 0, maxSteps = 0;
  bool keep = false;
  size_t bytes = 0;
  double tSubmit = 0;
};

struct Resident {
  std::unique_ptr<Job> job;  // inputs are dropped after admission; metadata kept
  nz::Inst I;
  int slot = 0;              // index into the batch's state-info array
  double tAdmit = 0;
};

struct Done {
  nz_batch_result r{};
  std::vector<float> A, B;   // kept final state
};

// a solve that stopped: its integrals and kept state are on their way to the host
struct Retiring {
  std::unique_ptr<Resident> res;
  InstState is{};            // step state at the stop
  hipEvent_t ev = nullptr;   // after its copies (and after the chunk that last named it)
  float* keepDev = nullptr;  // slab staging of the packed state
  float* keepHost = nullptr; // pinned, from the batch's pool
  size_t keepBytes = 0;
  double tStop = 0;          // host clock when the stop was read
};

// one chunk of steps in flight: its stop flags land in stHost; t0 / t1 time it
struct Chunk {
  hipEvent_t t0 = nullptr, t1 = nullptr;
  InstState* stHost = nullptr;
  std::vector<std::pair<int64_t, int>> members;  // (job id, slot) of the residents it ran
  bool pending = false;
};

}  // namespace

struct nz_batch {
  int prec = 32;
  uint64_t budget = 0, used = 0;
  int maxRes = 1;
  hipStream_t stream = nullptr;  // the chunks, retirement copies, table uploads
  hipStream_t up = nullptr;      // admissions' uploads
  hipEvent_t upEv = nullptr;     // after the last admission's uploads
  bool upWait = false;           // the next chunk waits for upEv
  nz::DevArena arena;            // every resident's device buffers
  InstState* st = nullptr;       // device, one per slot
  Chunk chunk[2];
  int cur = 0;                   // the chunk record the next run fills
  int pend = -1;                 // the chunk in flight whose flags are not read yet
  double busy = 0;               // device seconds of the chunks read so far
  std::vector<std::unique_ptr<Retiring>> retiring;
  std::vector<std::pair<float*, size_t>> keepPool;  // free pinned buffers for kept states
  std::map<float*, size_t> keepSizes;               // every pinned kept-state buffer and its size
  double* integHost = nullptr;   // pinned, 16 doubles per slot: integrals at retirement
  DevInst* dtab = nullptr;       // device descriptor table (resident order)
  DevInst* htab[2] = {nullptr, nullptr};  // pinned host copies, used in turn
  unsigned short* dti = nullptr; // device: the resident index of every launch tile
  unsigned short* hti[2] = {nullptr, nullptr};
  int dtiCap = 0, htiCap = 0, tabPar = 0;
  std::vector<int> freeSlots;
  std::vector<std::unique_ptr<Resident>> res;
  std::deque<std::unique_ptr<Job>> queue;
  std::map<int64_t, Done> done;
  std::vector<int64_t> finished;
  int64_t nextId = 1;
  bool tableDirty = true;
  int ntiles = 0, npblk = 0;
  bool phys = false;
  bool useGraph = true;
  int shape = 0;                 // tile shape of every resident
  hipGraphExec_t graph = nullptr;
  int graphSteps = 0;
};

namespace {

void drop_graph(nz_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graphSteps = 0;
}

// Admit one queued job: 0 done, 1 the slab has no room for it now (it stays queued), else
// an error. Its uploads go on the admissions' stream (the host waits on them, not on the
// chunk in flight); its slot and memory were released only once no queued work names them.
int admit(nz_batch* b, std::unique_ptr<Job>& job) {
  auto r = std::make_unique<Resident>();
  r->slot = b->freeSlots.back();
  nz::Inst& I = r->I;
  I.NP = job->tab.empty() ? 4 : job->parcels;
  b->arena.clear_failed();
  // a fresh solve: the slot's state info starts zeroed with spray seed 1 (spray.js:221)
  int rc = nz::inst_alloc(I, job->Nx, job->Nr, b->prec, b->st + r->slot, b->shape, &b->arena);
  if (!rc)
    rc = nz::inst_set_geometry(I, job->mask.data(), job->geoA.data(), job->geoG.data(), job->geoM.data(),
                               job->jWall.empty() ? nullptr : job->jWall.data());
  if (!rc && !job->tab.empty()) rc = nz::inst_set_injector(I, job->tab.data());
  if (!rc) rc = nz::inst_spray_alloc(I, I.NP, b->up);
  if (!rc) {
    nz::fill_consts(I, job->P, I.c, job->tEnd, job->resTol, job->minSteps, job->maxSteps);
    rc = nz::inst_set_state(I, job->A.data(), job->B.data(), b->up);
  }
  if (rc) {
    nz::inst_free(I);
    return b->arena.failed() ? 1 : rc;
  }
  b->freeSlots.pop_back();
  b->used += job->bytes;
  // the inputs are on the device now
  job->mask.clear(); job->mask.shrink_to_fit();
  for (auto* v : {&job->geoA, &job->geoG, &job->geoM, &job->jWall, &job->A, &job->B, &job->tab}) {
    v->clear();
    v->shrink_to_fit();
  }
  r->tAdmit = now_s();
  r->job = std::move(job);
  b->res.push_back(std::move(r));
  b->tableDirty = true;
  return 0;
}

int admit_queued(nz_batch* b) {
  bool any = false;
  for (auto it = b->queue.begin(); it != b->queue.end();) {
    if ((int)b->res.size() >= b->maxRes || b->freeSlots.empty()) break;
    if (b->used + (*it)->bytes > b->budget) { ++it; continue; }
    const int rc = admit(b, *it);
    if (rc == 1) {
      if (b->res.empty() && b->retiring.empty())
        return nz_fail(-7, "nz_batch_run: a queued job does not fit the device slab even with nothing resident");
      ++it;
      continue;
    }
    if (rc) return rc;
    it = b->queue.erase(it);
    any = true;
  }
  if (any) {
    HIPCHK(hipEventRecord(b->upEv, b->up), "batch upload event");
    b->upWait = true;
  }
  return 0;
}

// Descriptor and tile tables of the resident set, copied on the chunks' stream (so they
// land after the chunk in flight, which still reads the old ones) from pinned buffers
// used in turn: the one written here was last copied two uploads ago, before a chunk whose
// flags the host has read since.
int upload_table(nz_batch* b) {
  const int par = b->tabPar;
  b->tabPar ^= 1;
  DevInst* ht = b->htab[par];
  int t = 0, pb = 0;
  b->phys = false;
  for (size_t n = 0; n < b->res.size(); ++n) {
    nz::Inst& I = b->res[n]->I;
    DevInst d = nz::inst_desc(I);
    d.c = I.c;
    d.c.NP = I.NP;
    d.tile0 = t;
    d.pblk0 = pb;
    d.npblk = d.c.spray ? (I.NP + 255) / 256 : 0;
    t += I.tilesX * I.tilesY;
    pb += d.npblk;
    b->phys |= (d.c.visc || d.c.chem);
    ht[n] = d;
  }
  b->ntiles = t;
  b->npblk = pb;
  if (t > b->dtiCap || t > b->htiCap) {
    // grow the tile tables (rare): nothing in flight may still read the old ones
    HIPCHK(hipStreamSynchronize(b->stream), "batch tile table sync");
    const int cap = std::max(t, 2 * std::max(b->dtiCap, b->htiCap));
    if (b->dti) (void)hipFree(b->dti);
    for (auto& h : b->hti) { if (h) (void)hipHostFree(h); h = nullptr; }
    b->dti = nullptr;
    b->dtiCap = b->htiCap = 0;
    HIPCHK(hipMalloc(&b->dti, (size_t)cap * sizeof(unsigned short)), "batch tile table");
    for (auto& h : b->hti) HIPCHK(hipHostMalloc(&h, (size_t)cap * sizeof(unsigned short), hipHostMallocDefault), "batch tile table");
    b->dtiCap = b->htiCap = cap;
  }
  unsigned short* hti = b->hti[par];
  for (size_t n = 0; n < b->res.size(); ++n) {
    const DevInst& d = ht[n];
    std::fill(hti + d.tile0, hti + d.tile0 + d.tilesX * d.tilesY, (unsigned short)n);
  }
  if (!b->res.empty()) {
    HIPCHK(hipMemcpyAsync(b->dtab, ht, b->res.size() * sizeof(DevInst), hipMemcpyHostToDevice, b->stream), "batch table");
    HIPCHK(hipMemcpyAsync(b->dti, hti, (size_t)t * sizeof(unsigned short), hipMemcpyHostToDevice, b->stream),
           "batch tile table");
  }
  b->tableDirty = false;
  drop_graph(b);
  return 0;
}

int enqueue_steps(nz_batch* b, int steps) {
  const bool fp64 = b->prec == 64;
  const int n = (int)b->res.size();
  // a resident with more tile partials than stage 1 folds itself needs the close kernel
  // after every step; else the close at the end makes the step state current
  bool closeEach = false;
  for (const auto& r : b->res) closeEach |= (unsigned)(r->I.tilesX * r->I.tilesY) > nzk::FUSED_CLOSE_MAX;
  for (int k = 0; k < steps; ++k) {
    const int tx0 = n ? b->res[0]->I.tilesX : 1;
    HIPCHK(nzk::launch_stage(1, false, fp64, b->shape, b->dtab, n, b->ntiles, tx0, b->dti, b->stream), "batch stage 1");
    if (b->npblk) HIPCHK(nzk::launch_spray(fp64, b->dtab, n, b->npblk, b->stream), "batch spray");
    HIPCHK(nzk::launch_stage(2, b->phys, fp64, b->shape, b->dtab, n, b->ntiles, tx0, b->dti, b->stream), "batch stage 2");
    if (closeEach || k == steps - 1) HIPCHK(nzk::launch_close(b->dtab, n, b->stream), "batch close");
  }
  return 0;
}

int run_steps(nz_batch* b, int steps) {
  if (!b->useGraph) return enqueue_steps(b, steps);
  if (!b->graph || b->graphSteps != steps) {
    drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "batch capture begin");
    const int rc = enqueue_steps(b, steps);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return nz_fail(-10, std::string("batch capture end: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) { b->graph = nullptr; return nz_fail(-10, std::string("batch graph: ") + hipGetErrorString(ei)); }
    b->graphSteps = steps;
  }
  HIPCHK(hipGraphLaunch(b->graph, b->stream), "batch graph launch");
  return 0;
}

float* keep_buffer(nz_batch* b, size_t bytes) {
  for (auto it = b->keepPool.begin(); it != b->keepPool.end(); ++it)
    if (it->second >= bytes) {
      float* p = it->first;
      b->keepPool.erase(it);
      return p;
    }
  float* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  b->keepSizes[p] = bytes;
  return p;
}

// A resident whose stop flag the host has read leaves the resident set (the next table
// drops it); its integrals and kept state are enqueued behind the chunk in flight, which
// still runs it (stopped, its workgroups return at once, so its state is final), and
// collected by finalize_retired.
int start_retire(nz_batch* b, size_t n, const InstState& is) {
  auto rt = std::make_unique<Retiring>();
  rt->res = std::move(b->res[n]);
  b->res.erase(b->res.begin() + (long)n);
  b->tableDirty = true;
  rt->is = is;
  rt->tStop = now_s();
  Resident& r = *rt->res;
  const bool fp64 = b->prec == 64;
  if (r.job->probes.dx > 0) {
    DevInst desc = nz::inst_desc(r.I);
    desc.c = r.I.c;
    const nzk::IntegCtx q = nz::integ_ctx(desc, r.job->probes);
    HIPCHK(nzk::launch_integrals(fp64, desc, q, r.I.dint + 16, r.I.dint, b->stream), "batch integrals");
    HIPCHK(hipMemcpyAsync(b->integHost + (size_t)r.slot * 16, r.I.dint, 13 * 8, hipMemcpyDeviceToHost, b->stream),
           "batch integrals copy");
  }
  if (r.job->keep) {
    // packed into U1, which a stopped solve no longer reads or writes
    const long long N = (long long)r.I.Nx * r.I.Nr;
    rt->keepBytes = (size_t)N * 32;
    rt->keepHost = keep_buffer(b, rt->keepBytes);
    if (!rt->keepHost) return nz_fail(-10, "batch: pinned memory for a kept state");
    float* pk = static_cast<float*>(r.I.U1);
    HIPCHK(nzk::launch_pack(fp64, r.I.U0, r.I.Ns, r.I.Nx, 0, 0, 0, r.I.Nx, r.I.Nr, pk, b->stream), "batch keep pack");
    HIPCHK(nzk::launch_pack(fp64, r.I.U0, r.I.Ns, r.I.Nx, 4, 0, 0, r.I.Nx, r.I.Nr, pk + N * 4, b->stream), "batch keep pack");
    HIPCHK(hipMemcpyAsync(rt->keepHost, pk, rt->keepBytes, hipMemcpyDeviceToHost, b->stream), "batch keep copy");
  }
  HIPCHK(hipEventCreateWithFlags(&rt->ev, hipEventDisableTiming), "batch retire event");
  HIPCHK(hipEventRecord(rt->ev, b->stream), "batch retire event");
  b->retiring.push_back(std::move(rt));
  return 0;
}

// Retirements whose copies have landed (all of them, waiting, when `block`): results
// filed, memory and slot released.
int finalize_retired(nz_batch* b, bool block) {
  for (size_t k = 0; k < b->retiring.size();) {
    Retiring& rt = *b->retiring[k];
    if (block) {
      HIPCHK(hipEventSynchronize(rt.ev), "batch retire wait");
    } else {
      const hipError_t e = hipEventQuery(rt.ev);
      if (e == hipErrorNotReady) { ++k; continue; }
      if (e != hipSuccess) return nz_fail(-10, std::string("batch retire: ") + hipGetErrorString(e));
    }
    Resident& r = *rt.res;
    const InstState& is = rt.is;
    Done d;
    nz_batch_result& o = d.r;
    o.state = NZ_JOB_DONE;
    o.stop = (is.flags & 2u) ? NZ_STOP_RESIDUAL : (is.flags & 4u) ? NZ_STOP_TIME : (is.flags & 8u) ? NZ_STOP_STEPS : NZ_STOP_NONE;
    o.steps = is.steps;
    o.t_grid = is.last[2];
    o.dt = is.last[0];
    o.residual = is.resLast;
    o.t_submit = r.job->tSubmit;
    o.t_admit = r.tAdmit;
    o.t_done = rt.tStop;
    if (r.job->probes.dx > 0) {
      nz::fill_integrals(b->integHost + (size_t)r.slot * 16, r.job->probes, &o.integ);
      o.have_integrals = 1;
    }
    if (rt.keepHost) {
      const size_t n4 = rt.keepBytes / 8;  // floats per array
      d.A.assign(rt.keepHost, rt.keepHost + n4);
      d.B.assign(rt.keepHost + n4, rt.keepHost + 2 * n4);
      b->keepPool.emplace_back(rt.keepHost, b->keepSizes[rt.keepHost]);
    }
    (void)hipEventDestroy(rt.ev);
    b->used -= r.job->bytes;
    b->freeSlots.push_back(r.slot);
    b->finished.push_back(r.job->id);
    b->done[r.job->id] = std::move(d);
    nz::inst_free(r.I);
    b->retiring.erase(b->retiring.begin() + (long)k);
  }
  return 0;
}

// Read a chunk's stop flags (waiting for it) and start the retirement of the solves that
// stopped in it.
int process_chunk(nz_batch* b, int c) {
  Chunk& ch = b->chunk[c];
  HIPCHK(hipEventSynchronize(ch.t1), "batch chunk wait");
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ch.t0, ch.t1), "batch chunk time");
  b->busy += ms * 1e-3;
  ch.pending = false;
  for (const auto& m : ch.members) {
    for (size_t n = 0; n < b->res.size(); ++n) {
      if (b->res[n]->job->id != m.first) continue;
      const InstState& is = ch.stHost[m.second];
      if (is.flags & 1u) { if (int rc = start_retire(b, n, is)) return rc; }
      break;
    }
  }
  return 0;
}

// Nothing in flight or on its way: the chunk in flight read, every retirement filed.
int drain(nz_batch* b) {
  if (b->pend >= 0) {
    const int c = b->pend;
    b->pend = -1;
    if (int rc = process_chunk(b, c)) return rc;
  }
  return finalize_retired(b, true);
}

Resident* find_res(nz_batch* b, int64_t id) {
  for (auto& r : b->res)
    if (r->job->id == id) return r.get();
  return nullptr;
}

bool is_retiring(const nz_batch* b, int64_t id) {
  for (const auto& rt : b->retiring)
    if (rt->res->job->id == id) return true;
  return false;
}

void free_host(nz_batch* b) {
  for (auto& c : b->chunk) {
    if (c.t0) (void)hipEventDestroy(c.t0);
    if (c.t1) (void)hipEventDestroy(c.t1);
    if (c.stHost) (void)hipHostFree(c.stHost);
    c.t0 = c.t1 = nullptr;
    c.stHost = nullptr;
  }
  for (auto& h : b->htab) { if (h) (void)hipHostFree(h); h = nullptr; }
  for (auto& h : b->hti) { if (h) (void)hipHostFree(h); h = nullptr; }
  for (auto& kv : b->keepSizes) (void)hipHostFree(kv.first);
  b->keepSizes.clear();
  b->keepPool.clear();
  if (b->integHost) (void)hipHostFree(b->integHost);
  b->integHost = nullptr;
  if (b->upEv) (void)hipEventDestroy(b->upEv);
  b->upEv = nullptr;
}

}  // namespace

extern "C" {

uint64_t nz_batch_instance_bytes(int Nx, int Nr, int precision, int parcels) {
  return (uint64_t)nz::inst_bytes(Nx, Nr, precision, parcels > 0 ? parcels : 4);
}

int nz_batch_create(nz_batch** out, int precision, uint64_t hbm_budget, int max_resident) {
  if (!out) return nz_fail(-1, "nz_batch_create: null output");
  *out = nullptr;
  if (precision != 32 && precision != 64) return nz_fail(-1, "nz_batch_create: precision must be 32 or 64");
  if (max_resident < 1 || max_resident > 4096) return nz_fail(-1, "nz_batch_create: max_resident must be in [1, 4096]");
  if (hbm_budget == 0) return nz_fail(-1, "nz_batch_create: zero memory budget");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nz_fail(-5, "nz_batch_create: no HIP device available");
  auto b = std::make_unique<nz_batch>();
  b->prec = precision;
  b->budget = hbm_budget;
  b->maxRes = max_resident;
  auto fail = [&](const char* what) {
    std::string msg = what;
    if (b->st) (void)hipFree(b->st);
    if (b->dtab) (void)hipFree(b->dtab);
    b->arena.destroy();
    free_host(b.get());
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->up) (void)hipStreamDestroy(b->up);
    return nz_fail(-10, msg);
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail("nz_batch_create: stream");
  if (hipStreamCreateWithFlags(&b->up, hipStreamNonBlocking) != hipSuccess) return fail("nz_batch_create: upload stream");
  if (hipEventCreateWithFlags(&b->upEv, hipEventDisableTiming) != hipSuccess) return fail("nz_batch_create: event");
  for (auto& c : b->chunk) {
    if (hipEventCreate(&c.t0) != hipSuccess || hipEventCreate(&c.t1) != hipSuccess) return fail("nz_batch_create: event");
    if (hipHostMalloc(&c.stHost, sizeof(InstState) * max_resident, hipHostMallocDefault) != hipSuccess)
      return fail("nz_batch_create: pinned state");
  }
  // the slab: the budget, plus room for what the per-solve estimate leaves out (block
  // alignment, the merged-cell face lists)
  const size_t slab = hbm_budget + (64ull << 20) + (size_t)max_resident * (64u << 10);
  if (b->arena.create(slab)) return fail("nz_batch_create: device slab");
  if (hipMalloc(&b->st, sizeof(InstState) * max_resident) != hipSuccess) return fail("nz_batch_create: state info");
  if (hipMalloc(&b->dtab, sizeof(DevInst) * max_resident) != hipSuccess) return fail("nz_batch_create: table");
  for (auto& h : b->htab)
    if (hipHostMalloc(&h, sizeof(DevInst) * max_resident, hipHostMallocDefault) != hipSuccess) return fail("nz_batch_create: pinned table");
  if (hipHostMalloc(&b->integHost, sizeof(double) * 16 * max_resident, hipHostMallocDefault) != hipSuccess)
    return fail("nz_batch_create: pinned integrals");
  for (int k = max_resident - 1; k >= 0; --k) b->freeSlots.push_back(k);
  *out = b.release();
  return 0;
}

int nz_batch_destroy(nz_batch* b) {
  if (!b) return 0;
  (void)hipStreamSynchronize(b->stream);
  (void)hipStreamSynchronize(b->up);
  drop_graph(b);
  for (auto& rt : b->retiring) {
    if (rt->ev) (void)hipEventDestroy(rt->ev);
    nz::inst_free(rt->res->I);
  }
  b->retiring.clear();
  for (auto& r : b->res) nz::inst_free(r->I);
  b->res.clear();
  b->arena.destroy();
  (void)hipFree(b->st);
  (void)hipFree(b->dtab);
  if (b->dti) (void)hipFree(b->dti);
  free_host(b);
  (void)hipStreamDestroy(b->stream);
  (void)hipStreamDestroy(b->up);
  delete b;
  return 0;
}

int nz_batch_submit(nz_batch* b, const nz_batch_job* jb, int64_t* id) {
  if (!b || !jb || !id) return nz_fail(-1, "nz_batch_submit: null argument");
  if (jb->Nx < 2 || jb->Nr < 2 || (long long)jb->Nx * jb->Nr > (1LL << 28)) return nz_fail(-1, "nz_batch_submit: grid size out of range");
  if (!jb->mask || !jb->geoA || !jb->geoG || !jb->geoM || !jb->A || !jb->B) return nz_fail(-1, "nz_batch_submit: null array");
  if (jb->injector && (jb->parcels < 4 || jb->parcels > (1 << 22) || jb->parcels % 4 != 0))
    return nz_fail(-1, "nz_batch_submit: parcel count must be a multiple of 4 in [4, 2^22]");
  auto j = std::make_unique<Job>();
  const size_t N = (size_t)jb->Nx * jb->Nr, NA = (size_t)(jb->Nx + 1) * (jb->Nr + 1) * 4;
  j->Nx = jb->Nx;
  j->Nr = jb->Nr;
  j->mask.assign(jb->mask, jb->mask + N);
  j->geoA.assign(jb->geoA, jb->geoA + NA);
  j->geoG.assign(jb->geoG, jb->geoG + N * 4);
  j->geoM.assign(jb->geoM, jb->geoM + N * 4);
  if (jb->jWall) j->jWall.assign(jb->jWall, jb->jWall + (size_t)jb->Nx * 4);
  j->A.assign(jb->A, jb->A + N * 4);
  j->B.assign(jb->B, jb->B + N * 4);
  if (jb->injector) j->tab.assign(jb->injector, jb->injector + 256 * 3 * 4);
  j->P = jb->P;
  j->parcels = jb->injector ? jb->parcels : 0;
  j->probes = jb->probes;
  j->tEnd = jb->t_end;
  j->resTol = jb->res_tol;
  j->minSteps = jb->min_steps;
  j->maxSteps = jb->max_steps;
  j->keep = jb->keep_state != 0;
  j->bytes = nz::inst_bytes(jb->Nx, jb->Nr, b->prec, j->parcels > 0 ? j->parcels : 4);
  if (j->bytes > b->budget) return nz_fail(-7, "nz_batch_submit: the job alone exceeds the memory budget");
  if (jb->probes.dx > 0) {
    const nz_probes& pr = jb->probes;
    if (pr.iExit < 0 || pr.iExit + 1 > jb->Nx || pr.jExit < 0 || pr.jExit > jb->Nr || pr.iThroat < 0 ||
        pr.iThroat >= jb->Nx || pr.iChamber < 0 || pr.iChamber >= jb->Nx)
      return nz_fail(-1, "nz_batch_submit: probe outside the grid");
  }
  j->tSubmit = now_s();
  j->id = b->nextId++;
  *id = j->id;
  b->queue.push_back(std::move(j));
  return 0;
}

int nz_batch_run(nz_batch* b, int steps, int* resident, int* queued) {
  if (!b) return nz_fail(-1, "nz_batch_run: null batch");
  if (steps < 0) return nz_fail(-1, "nz_batch_run: negative step count");
  if (steps > 0) {
    b->finished.clear();
    if (int rc = finalize_retired(b, false)) return rc;
  }
  if (int rc = admit_queued(b)) return rc;
  if (b->tableDirty) { if (int rc = upload_table(b)) return rc; }
  const int n = (int)b->res.size();
  int enq = -1;
  if (n > 0 && steps > 0) {
    // this call's chunk goes on the device before the host waits for anything
    Chunk& ch = b->chunk[b->cur];
    if (b->upWait) {
      HIPCHK(hipStreamWaitEvent(b->stream, b->upEv, 0), "batch upload wait");
      b->upWait = false;
    }
    HIPCHK(hipEventRecord(ch.t0, b->stream), "batch chunk event");
    if (int rc = run_steps(b, steps)) return rc;
    HIPCHK(hipMemcpyAsync(ch.stHost, b->st, sizeof(InstState) * b->maxRes, hipMemcpyDeviceToHost, b->stream), "batch poll");
    HIPCHK(hipEventRecord(ch.t1, b->stream), "batch chunk event");
    ch.members.clear();
    for (const auto& r : b->res) ch.members.emplace_back(r->job->id, r->slot);
    ch.pending = true;
    enq = b->cur;
    b->cur ^= 1;
  }
  if (steps > 0) {
    // the previous chunk's flags (the device works on this call's chunk meanwhile)
    if (b->pend >= 0) {
      const int c = b->pend;
      b->pend = -1;
      if (int rc = process_chunk(b, c)) return rc;
    }
    if (enq >= 0) b->pend = enq;
    else if (int rc = finalize_retired(b, true)) return rc;  // nothing to run: file what is on its way
  }
  if (resident) *resident = (int)(b->res.size() + b->retiring.size());
  if (queued) *queued = (int)b->queue.size();
  return 0;
}

int nz_batch_gpu_busy(const nz_batch* b, double* seconds) {
  if (!b || !seconds) return nz_fail(-1, "nz_batch_gpu_busy: null argument");
  *seconds = b->busy;
  return 0;
}

int nz_batch_result_get(nz_batch* b, int64_t id, nz_batch_result* out) {
  if (!b || !out) return nz_fail(-1, "nz_batch_result_get: null argument");
  std::memset(out, 0, sizeof(*out));
  if (is_retiring(b, id)) { if (int rc = finalize_retired(b, true)) return rc; }
  auto it = b->done.find(id);
  if (it != b->done.end()) { *out = it->second.r; return 0; }
  for (auto& j : b->queue)
    if (j->id == id) { out->state = NZ_JOB_QUEUED; out->t_submit = j->tSubmit; return 0; }
  if (Resident* r = find_res(b, id)) {
    InstState is{};
    HIPCHK(hipStreamSynchronize(b->stream), "batch result sync");
    HIPCHK(hipMemcpy(&is, b->st + r->slot, sizeof(is), hipMemcpyDeviceToHost), "batch result");
    out->state = NZ_JOB_RUNNING;
    out->steps = is.steps;
    out->t_grid = is.last[2];
    out->dt = is.last[0];
    out->residual = is.resLast;
    out->t_submit = r->job->tSubmit;
    out->t_admit = r->tAdmit;
    return 0;
  }
  return nz_fail(-1, "nz_batch_result_get: unknown job id");
}

int nz_batch_read_state(nz_batch* b, int64_t id, float* A, float* B) {
  if (!b || !A || !B) return nz_fail(-1, "nz_batch_read_state: null argument");
  if (is_retiring(b, id)) { if (int rc = finalize_retired(b, true)) return rc; }
  auto it = b->done.find(id);
  if (it != b->done.end()) {
    if (it->second.A.empty()) return nz_fail(-4, "nz_batch_read_state: the job was not submitted with keep_state");
    std::memcpy(A, it->second.A.data(), it->second.A.size() * 4);
    std::memcpy(B, it->second.B.data(), it->second.B.size() * 4);
    return 0;
  }
  Resident* r = find_res(b, id);
  if (!r) return nz_fail(-1, "nz_batch_read_state: job is not resident or finished");
  const long long N = (long long)r->I.Nx * r->I.Nr;
  float* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, (size_t)N * 32), "batch read staging");
  const bool fp64 = b->prec == 64;
  hipError_t e = nzk::launch_pack(fp64, r->I.U0, r->I.Ns, r->I.Nx, 0, 0, 0, r->I.Nx, r->I.Nr, tmp, b->stream);
  if (e == hipSuccess) e = nzk::launch_pack(fp64, r->I.U0, r->I.Ns, r->I.Nx, 4, 0, 0, r->I.Nx, r->I.Nr, tmp + N * 4, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(A, tmp, (size_t)N * 16, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(B, tmp + N * 4, (size_t)N * 16, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return nz_fail(-10, std::string("nz_batch_read_state: ") + hipGetErrorString(e));
  return 0;
}

int nz_batch_release(nz_batch* b, int64_t id) {
  if (!b) return nz_fail(-1, "nz_batch_release: null batch");
  if (!b->done.erase(id)) return nz_fail(-1, "nz_batch_release: job is not finished");
  return 0;
}

int nz_batch_set_graph(nz_batch* b, int on) {
  if (!b) return nz_fail(-1, "nz_batch_set_graph: null batch");
  b->useGraph = on != 0;
  drop_graph(b);
  return 0;
}

int nz_batch_set_tile_shape(nz_batch* b, int shape) {
  if (!b) return nz_fail(-1, "nz_batch_set_tile_shape: null batch");
  int tx = 0, ty = 0;
  if (!nzk::tile_shape(b->prec == 64, shape, &tx, &ty)) return nz_fail(-1, "nz_batch_set_tile_shape: unknown tile shape");
  if (int rc = drain(b)) return rc;
  HIPCHK(hipStreamSynchronize(b->stream), "nz_batch_set_tile_shape: sync");
  for (auto& r : b->res)
    if (int rc = nz::inst_set_shape(r->I, shape)) return rc;
  b->shape = shape;
  b->tableDirty = true;
  return 0;
}

int nz_batch_autotune(nz_batch* b, int steps, int* best, double* us_per_step) {
  if (!b) return nz_fail(-1, "nz_batch_autotune: null batch");
  if (steps < 1) return nz_fail(-1, "nz_batch_autotune: step count must be positive");
  if (int rc = drain(b)) return rc;
  if (int rc = admit_queued(b)) return rc;
  if (b->upWait) {
    HIPCHK(hipStreamWaitEvent(b->stream, b->upEv, 0), "batch upload wait");
    b->upWait = false;
  }
  const int n = nzk::tile_shape_count(b->prec == 64);
  if (b->res.empty()) {
    if (best) *best = b->shape;
    return 0;
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0), "event");
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return nz_fail(-10, "event"); }
  int pick = b->shape, rc = 0;
  double bestUs = 1e300;
  for (int k = 0; k < n && !rc; ++k) {
    if ((rc = nz_batch_set_tile_shape(b, k)) || (rc = upload_table(b))) break;
    // one untimed step loads the shape's code object; every step is real work
    if ((rc = enqueue_steps(b, 1))) break;
    if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = nz_fail(-10, "event"); break; }
    if ((rc = enqueue_steps(b, steps))) break;
    float ms = 0.f;
    if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = nz_fail(-10, "batch autotune timing"); break; }
    const double us = ms * 1e3 / steps;
    if (us_per_step) us_per_step[k] = us;
    if (us < bestUs) { bestUs = us; pick = k; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (best) *best = pick;
  return nz_batch_set_tile_shape(b, pick);
}

int nz_batch_memory(const nz_batch* b, uint64_t* used, uint64_t* budget) {
  if (!b) return nz_fail(-1, "nz_batch_memory: null batch");
  if (used) *used = b->used;
  if (budget) *budget = b->budget;
  return 0;
}

int nz_batch_finished(const nz_batch* b, int64_t* ids, int cap) {
  if (!b) return nz_fail(-1, "nz_batch_finished: null batch");
  const int n = (int)b->finished.size();
  for (int k = 0; k < n && k < cap && ids; ++k) ids[k] = b->finished[k];
  return n;
}

}  // extern "C"
