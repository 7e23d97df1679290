// engine.hpp -- the PDLP engines' host code: the single-LP engine, the batch and
// the serving slot table, templated on the iterate type.  engine_inst.hip
// instantiates it once per precision (two translation units built side by
// side), engine_api.hip holds the C ABI.
#pragma once
#include "common.hpp"

using sov::LpState;
using sov::StepCtl;

// ====================================================================== engines
struct sovlp_pdlp {
  int precision, W, m, n;
  void* lpdev;               // LpDev<T> (device)
  LpState* st;
  StepCtl* ctl;
  unsigned* done;
  double *xo, *yo, *dout, *hist, *rhist;
  long long* pcount;         // [kPhaseSlots]
  long long* pticks;         // [kPhaseSlots]
  int phases_per_replay;
  long long hist_cap;
  double hz;
  bool started;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t cap;
  long long ticks_set;
  LpState* hstate;           // [2] pinned host copies of the state
  long long* hkstop;         // pinned: the iteration stop of the next replays
  double *rowf, *colf;       // scaling factors (device, FP64)
  hipEvent_t ev[2];
};

struct sovlp_batch {
  int count, precision, block;
  void* lps;                 // LpDev<T>[count] (device)
  StepCtl* ctl;
  unsigned* done;
  long long* pcount;         // [count][kPhaseSlots]
  long long* pticks;         // [count][kPhaseSlots]
  std::vector<double*> xo, yo, dout, rhist;
  std::vector<LpState*> st;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t cap;
  double hz;
  long long ticks_set;
  unsigned* hdone;           // [2] pinned host copies of the done counter
  hipEvent_t ev[2];
  hipStream_t side[3];       // capture branches of the groups after the first
  hipEvent_t fork, join[3];
};


// ====================================================================== server
// A slot table for continuous batching: `slots` LPs share every launch of one
// recorded graph (one workgroup per slot, up to `phases` phases each).  An
// LP is admitted into a free slot between launches (its descriptor and its
// fresh state are copied in stream order), runs until its state reaches
// DONE, and is released by the host after the result is read.  Empty and
// finished slots point at a DONE state in the server's own workspace, so
// their workgroups return at once and never touch a released LP's memory.
struct sovlp_server {
  int slots, precision, block, phases;
  int team;                  // workgroups per slot and phase (1: run kernel)
  double* partials;          // [slots][team][kMaxRed] (team > 1)
  int32_t* parts;            // [slots][2][team + 1] tile partitions of A and A^T (team > 1)
  int32_t* hparts;           // pinned staging of the partitions
  void* lps;                 // LpDev<T>[slots] (device)
  LpState* states;           // [slots] (device)
  long long* pcount;         // [slots][kPhaseSlots] (device)
  long long* pticks;         // [slots][kPhaseSlots] (device)
  StepCtl* ctl;
  unsigned* done;
  char* hstage;              // pinned LpDev<T>[slots]: descriptors being copied in
  LpState* hinit;            // pinned [slots]: fresh states being copied in
  LpState* hsnap;            // pinned [slots]: states after the last step
  std::vector<char> busy, seen_done;
  std::vector<double*> xo, yo, dout, rhist, hist;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t cap;
  double hz;
};


namespace sovh {


constexpr int kTnnzSmall = 512;   // 64-thread workgroups: 2 x 512 x 8 B = 8 KiB LDS
constexpr int kPairsPerGraph = 64;     // W > 1: step launches (one phase each) per graph

// Workgroup sizes of the batched engine, and their tile nonzero caps (16 KiB
// of LDS for 128 threads, 32 KiB for 256, 64 KiB for 512: the largest LPs
// of a batch, whose phases one 256-thread workgroup runs tile after tile)
constexpr int kBatchBlocks = 4;
constexpr int kBatchBlock[kBatchBlocks] = {64, 128, 256, 512};
inline bool batch_block_ok(int block) {
  return block == 64 || block == 128 || block == 256 || block == 512;
}
inline int batch_tnnz(int precision, int block) {
  if (block == 64) return kTnnzSmall;
  if (block == 128) return precision == 64 ? kTnnz64 / 2 : kTnnz32 / 2;
  if (block == 512) return precision == 64 ? kTnnz64 * 2 : kTnnz32 * 2;
  return precision == 64 ? kTnnz64 : kTnnz32;
}

// Tile nonzero cap of the run kernel with `block` threads per LP
template <class T>
constexpr int run_tnnz(int block) {
  return block == 64 ? kTnnzSmall
       : block == 128 ? (sizeof(T) == 8 ? kTnnz64 : kTnnz32) / 2
       : block == 512 ? (sizeof(T) == 8 ? kTnnz64 : kTnnz32) * 2
                      : (sizeof(T) == 8 ? kTnnz64 : kTnnz32);
}

// One launch of the run kernel over `count` LP descriptors with BLOCK threads
// each.  Defined in engine_run.hip, which compiles once per (iterate type,
// workgroup size) so that the six run kernels build side by side.
template <class T, int BLOCK>
void launch_run_block(const LpDev<T>* lps, int count, const StepCtl* ctl, unsigned* done,
                      int phases, hipStream_t st);

template <class T>
void launch_run(int block, const LpDev<T>* lps, int count, const StepCtl* ctl, unsigned* done,
                int phases, hipStream_t st) {
  if (block == 64) launch_run_block<T, 64>(lps, count, ctl, done, phases, st);
  else if (block == 128) launch_run_block<T, 128>(lps, count, ctl, done, phases, st);
  else if (block == 512) launch_run_block<T, 512>(lps, count, ctl, done, phases, st);
  else launch_run_block<T, 256>(lps, count, ctl, done, phases, st);
}

// run kernel: phases per LP per launch. Every launch ends when its slowest LP
// has run its phases, so shorter launches leave CUs idle at each end (4096
// uneven LPs: 1024 phases 8470 LPs/s, 16384 11170, 65536 12012, 262144 12101);
// time and iteration limits are enforced on the device, not by this count.
constexpr int kPhasesPerLaunch = 65536;

template <class T>
void fill_lp(LpDev<T>& L, const sovlp_lp& lp) {
  L.m = lp.m; L.n = lp.n; L.ntA = lp.ntiles_a; L.ntT = lp.ntiles_t;
  L.flip = lp.maximise ? 1 : 0;
  L.offset0 = lp.obj_offset;
  L.ap = lp.rp; L.ai = lp.ri; L.av = lp.rx;
  L.tp = lp.cp; L.ti = lp.ci; L.tv = lp.cx;
  L.c0 = lp.c; L.lo0 = lp.lo; L.hi0 = lp.hi; L.rl0 = lp.rl; L.ru0 = lp.ru; L.v0 = lp.v0;
  L.tilesA = lp.tiles_a;
  L.tilesT = lp.tiles_t;
  if (L.xacc) {              // the rounds' bounds live in the workspace
    L.crl = L.wrl; L.cru = L.wru; L.clo = L.wlo; L.chi = L.whi;
  } else {                   // no refinement: the original bounds
    L.crl = lp.rl; L.cru = lp.ru; L.clo = lp.lo; L.chi = lp.hi;
  }
}

// workgroups sharing one LP's phases: one per ~4 tiles, at most what is resident
template <class T>
int pick_width(const sovlp_lp* lp, int team, int* W) {
  int dev = 0, cus = 0, nb = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, pdlp_step_kernel<T, kBlock, sizeof(T) == 8 ? kTnnz64 : kTnnz32>, kBlock, 0));
  const long long tiles = (long long)lp->ntiles_a + lp->ntiles_t;
  const long long resident = (long long)std::max(nb, 1) * cus;
  long long w = std::max<long long>(1, tiles / 4);
  if (team == 1) w = 1;
  if (team == 0 && lp->nnz < 200000) w = 1;
  *W = (int)std::min(w, resident);
  return 0;
}

template <class T>
size_t single_ws_T(const sovlp_lp* lp, long long hist_cap, int W) {
  Carver cv{nullptr, 0, 0};
  cv.take<LpDev<T>>(1);
  cv.take<StepCtl>(1);
  cv.take<unsigned>(64);
  cv.take<double>((size_t)2 * W * kMaxRed);
  cv.take<LpState>(1);
  cv.take<int32_t>(W + 1);
  cv.take<int32_t>(W + 1);
  cv.take<long long>(kPhaseSlots);
  cv.take<long long>(kPhaseSlots);
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  return cv.off + 256;
}

template <class T>
int single_create_T(const sovlp_lp* lp, const sovlp_params* prm, void* ws, int64_t ws_bytes,
                    int64_t hist_cap, int team, hipStream_t st, sovlp_pdlp* h) {
  int W = 1;
  int rc = pick_width<T>(lp, team, &W);
  if (rc) return rc;
  if ((size_t)ws_bytes < single_ws_T<T>(lp, hist_cap, W)) return fail(1, "workspace too small");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T>* d_lp = cv.take<LpDev<T>>(1);
  StepCtl* ctl = cv.take<StepCtl>(1);
  unsigned* done = cv.take<unsigned>(64);
  double* partials = cv.take<double>((size_t)2 * W * kMaxRed);
  LpState* st2 = cv.take<LpState>(1);
  int32_t* partA = cv.take<int32_t>(W + 1);
  int32_t* partT = cv.take<int32_t>(W + 1);
  long long* pcount = cv.take<long long>(kPhaseSlots);
  long long* pticks = cv.take<long long>(kPhaseSlots);
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  fill_lp<T>(L, *lp);
  L.partials = partials;
  L.st2 = st2;
  L.pcount = pcount;
  L.pticks = pticks;
  if (W > 1) {
    // equal-cost split of both tile lists (needs the pointer values at tile starts)
    std::vector<int32_t> ta(lp->ntiles_a + 1), tt(lp->ntiles_t + 1), rp(lp->m + 1), cp(lp->n + 1);
    HIPCHK(hipMemcpyAsync(ta.data(), lp->tiles_a, sizeof(int32_t) * ta.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(tt.data(), lp->tiles_t, sizeof(int32_t) * tt.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rp.data(), lp->rp, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cp.data(), lp->cp, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int64_t> pa(ta.size()), pt(tt.size());
    for (size_t t = 0; t < ta.size(); ++t) pa[t] = rp[ta[t]];
    for (size_t t = 0; t < tt.size(); ++t) pt[t] = cp[tt[t]];
    std::vector<int32_t> ha(W + 1), ht(W + 1);
    host_partition(lp->ntiles_a, ta.data(), pa.data(), W, ha.data());
    host_partition(lp->ntiles_t, tt.data(), pt.data(), W, ht.data());
    HIPCHK(hipMemcpyAsync(partA, ha.data(), sizeof(int32_t) * (W + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(partT, ht.data(), sizeof(int32_t) * (W + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    L.partA = partA;
    L.partT = partT;
  }
  LpState s0{};
  s0.phase = PH_INIT;
  s0.status = ST_RUNNING;
  StepCtl c0{};
  c0.kstop = kNoStop;
  std::memcpy(&c0.P, prm, sizeof(Params));
  c0.ticks = (long long)9.0e18;
  HIPCHK(hipMemcpyAsync(L.st, &s0, sizeof s0, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ctl, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(done, 0, sizeof(unsigned), st));
  HIPCHK(hipMemsetAsync(pcount, 0, sizeof(long long) * kPhaseSlots, st));
  HIPCHK(hipMemsetAsync(pticks, 0, sizeof(long long) * kPhaseSlots, st));
  HIPCHK(hipMemcpyAsync(d_lp, &L, sizeof L, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));

  h->precision = sizeof(T) == 8 ? 64 : 32;
  h->W = W;
  h->m = lp->m;
  h->n = lp->n;
  h->lpdev = d_lp;
  h->st = L.st;
  h->ctl = ctl;
  h->done = done;
  h->xo = L.xo; h->yo = L.yo; h->dout = L.dout; h->hist = L.hist; h->rhist = L.rhist;
  h->rowf = L.rowf;
  h->colf = L.colf;
  h->pcount = pcount;
  h->pticks = pticks;
  h->hist_cap = hist_cap;
  h->ticks_set = c0.ticks;
  // record the launches of one replay into a graph on a private capture
  // stream: W == 1, one run launch (kPhasesPerLaunch phases); W > 1,
  // kPairsPerGraph step launches (one phase each, alternating state copies),
  // then the settle launch that takes the last phase's transition
  HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  constexpr int TN = sizeof(T) == 8 ? kTnnz64 : kTnnz32;
  if (W == 1) {
    launch_run<T>(kBlock, d_lp, 1, ctl, done, kPhasesPerLaunch, h->cap);
    h->phases_per_replay = kPhasesPerLaunch;
  } else {
    for (int i = 0; i < kPairsPerGraph; ++i)
      pdlp_step_kernel<T, kBlock, TN><<<W, kBlock, 0, h->cap>>>(d_lp, i & 1, ctl, done);
    pdlp_settle_kernel<T, kBlock><<<1, kBlock, 0, h->cap>>>(d_lp, W, kPairsPerGraph & 1,
                                                            (kPairsPerGraph - 1) & 1, ctl, done);
    h->phases_per_replay = kPairsPerGraph;
  }
  HIPCHK(hipStreamEndCapture(h->cap, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  HIPCHK(hipHostMalloc((void**)&h->hstate, 2 * sizeof(LpState), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&h->hkstop, sizeof(long long), hipHostMallocDefault));
  for (hipEvent_t& e : h->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return clock_hz(&h->hz);
}

template <class T>
size_t batch_ws_T(int count, const sovlp_lp* lps, long long hist_cap) {
  Carver cv{nullptr, 0, 0};
  cv.take<LpDev<T>>(count);
  cv.take<StepCtl>(1);
  cv.take<unsigned>(64);
  cv.take<long long>((size_t)count * kPhaseSlots);
  cv.take<long long>((size_t)count * kPhaseSlots);
  for (int i = 0; i < count; ++i) {
    LpDev<T> L{};
    carve_lp<T>(cv, L, lps[i].m, lps[i].n, lps[i].nnz, hist_cap);
  }
  return cv.off + 256;
}

template <class T>
int batch_create_T(int count, const sovlp_lp* lps, const sovlp_params* prm, const int32_t* blocks,
                   void* ws, int64_t ws_bytes, int64_t hist_cap, hipStream_t st, sovlp_batch* h) {
  if ((size_t)ws_bytes < batch_ws_T<T>(count, lps, hist_cap)) return fail(1, "workspace too small");
  // descriptors grouped by workgroup size (one launch per group), largest LP
  // first within a group, so that the longest workgroups of a launch start
  // first and do not extend its tail; pos[i] is LP i's place in the device
  // array (every output stays indexed by i)
  const int* kBlocks = kBatchBlock;
  std::vector<int> order, pos(count), gcount(kBatchBlocks, 0);
  for (int g = 0; g < kBatchBlocks; ++g) {
    std::vector<int> idx;
    for (int i = 0; i < count; ++i)
      if (blocks[i] == kBlocks[g]) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return lps[a].nnz > lps[b].nnz; });
    for (int i : idx) {
      pos[i] = (int)order.size();
      order.push_back(i);
      ++gcount[g];
    }
  }
  if ((int)order.size() != count) return fail(1, "every LP's block must be 64, 128, 256 or 512");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T>* d_lps = cv.take<LpDev<T>>(count);
  StepCtl* ctl = cv.take<StepCtl>(1);
  unsigned* done = cv.take<unsigned>(64);
  long long* pcount = cv.take<long long>((size_t)count * kPhaseSlots);
  long long* pticks = cv.take<long long>((size_t)count * kPhaseSlots);
  std::vector<LpDev<T>> host(count);
  h->xo.resize(count);
  h->yo.resize(count);
  h->dout.resize(count);
  h->rhist.resize(count);
  h->st.resize(count);
  for (int i = 0; i < count; ++i) {
    if (lps[i].nnz >= (int64_t)INT32_MAX) return fail(4, "an LP has more than 2^31-1 nonzeros");
    LpDev<T>& L = host[pos[i]];
    L = LpDev<T>{};
    carve_lp<T>(cv, L, lps[i].m, lps[i].n, lps[i].nnz, hist_cap);
    fill_lp<T>(L, lps[i]);
    L.pcount = pcount + (size_t)i * kPhaseSlots;
    L.pticks = pticks + (size_t)i * kPhaseSlots;
    h->xo[i] = L.xo; h->yo[i] = L.yo; h->dout[i] = L.dout; h->st[i] = L.st;
    h->rhist[i] = L.rhist;
  }
  StepCtl c0{};
  c0.kstop = kNoStop;
  std::memcpy(&c0.P, prm, sizeof(Params));
  c0.ticks = (long long)9.0e18;
  HIPCHK(hipMemcpyAsync(d_lps, host.data(), sizeof(LpDev<T>) * count, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ctl, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  h->count = count;
  h->precision = sizeof(T) == 8 ? 64 : 32;
  h->block = blocks[0];
  h->lps = d_lps;
  h->ctl = ctl;
  h->done = done;
  h->pcount = pcount;
  h->pticks = pticks;
  h->ticks_set = c0.ticks;
  HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  // one launch per group; the groups are independent branches of the graph
  // (forked from and joined to the capture stream), so they can run side by
  // side.  The largest workgroups are launched first: groups of smaller ones
  // dispatched ahead of them would take the free slots and hold back the
  // batch's largest LPs, which set the step.
  int goff[kBatchBlocks];
  for (int g = 0, off = 0; g < kBatchBlocks; off += gcount[g], ++g) goff[g] = off;
  int first = -1, nside = 0;
  for (int g = kBatchBlocks - 1; g >= 0; --g)
    if (gcount[g] > 0 && first < 0) first = g;
  HIPCHK(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
  for (int k = 0; k < kBatchBlocks - 1; ++k) {
    HIPCHK(hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->join[k], hipEventDisableTiming));
  }
  HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  HIPCHK(hipEventRecord(h->fork, h->cap));
  for (int g = kBatchBlocks - 1; g >= 0; --g) {
    const int off = goff[g];
    if (gcount[g] == 0) continue;
    if (g == first) {
      launch_run<T>(kBlocks[g], d_lps + off, gcount[g], ctl, done, kPhasesPerLaunch, h->cap);
    } else {
      hipStream_t sd = h->side[nside];
      HIPCHK(hipStreamWaitEvent(sd, h->fork, 0));
      launch_run<T>(kBlocks[g], d_lps + off, gcount[g], ctl, done, kPhasesPerLaunch, sd);
      HIPCHK(hipEventRecord(h->join[nside], sd));
      ++nside;
    }
  }
  for (int k = 0; k < nside; ++k) HIPCHK(hipStreamWaitEvent(h->cap, h->join[k], 0));
  HIPCHK(hipStreamEndCapture(h->cap, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  HIPCHK(hipHostMalloc((void**)&h->hdone, 2 * sizeof(unsigned), hipHostMallocDefault));
  for (hipEvent_t& e : h->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return clock_hz(&h->hz);
}

// Puts every LP of a batch back at the start of its solve.
template <class T>
int batch_reset_T(sovlp_batch* h, hipStream_t st) {
  state_reset_kernel<T><<<(h->count + 255) / 256, 256, 0, st>>>(
      (const LpDev<T>*)h->lps, h->count, h->done);
  HIPCHK(hipGetLastError());
  return 0;
}

inline int set_ticks(StepCtl* ctl, long long* cached, double seconds, double hz, hipStream_t st) {
  const double secs = seconds > 0.0 ? seconds : 0.0;
  long long ticks = (long long)std::min(secs * hz, 9.0e18);
  if (ticks != *cached) {
    *cached = ticks;
    HIPCHK(hipMemcpyAsync(&ctl->ticks, cached, sizeof(long long), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));   // the source is host memory owned by the handle
  }
  return 0;
}

inline void fill_info(const LpState& t, sovlp_info* f) {
  std::memset(f, 0, sizeof *f);
  const bool opt = t.status == ST_OPTIMAL;
  const double nan = std::nan("");
  f->status = t.status;
  f->stage = t.phase == PH_DONE ? 3
           : t.phase < PH_KKT0 ? 0
           : (t.phase == PH_FINAL || t.phase == PH_FINAL_D) ? 2 : 1;
  f->iterations = t.k;
  f->restarts = t.restarts;
  f->rejected_steps = t.rejected;
  f->history_len = t.hist_len;
  f->primal_res = opt ? t.fin[0] : nan;
  f->dual_res = opt ? t.fin[1] : nan;
  f->gap = opt ? t.fin[2] : nan;
  f->primal_obj = opt ? t.fin[3] : nan;
  f->dual_obj = opt ? t.fin[4] : nan;
  f->step_size = t.eta;
  f->primal_weight = t.omega;
  f->objective = t.objective;
  f->scale_obj = t.sobj0;
  f->device_ticks = t.ticks;
  f->iterations = t.k_done + t.k;
  f->refine_rows = t.ref_rows;
  f->refine_round = t.round;
  f->refine_scale = t.ref_scale;
  f->start_tick = t.tick0;
}

inline int check_params(const sovlp_params* p, int precision) {
  if (p->check_every < 1 || p->stepsize_every < 1) return fail(1, "check/stepsize period < 1");
  if (p->refine_rounds < 0 || p->refine_rounds > kMaxRounds)
    return fail(1, "refine_rounds must be in [0, 16]");
  if (p->refine_rounds > 0 && precision != 32)
    return fail(4, "iterative refinement is provided for precision 32 only");
  if (p->refine_rounds > 0 && !(p->refine_eps > 0.0 && p->refine_tol >= 0.0))
    return fail(1, "refine_eps must be > 0 and refine_tol >= 0");
  if (p->refine_stall < 0 || p->refine_stall_open < 0)
    return fail(1, "refine_stall and refine_stall_open must be >= 0");
  if (p->flags & ~kKnownFlags) return fail(1, "unknown bits in flags");
  return 0;
}



template <class T>
size_t server_ws_T(int slots, int team) {
  Carver cv{nullptr, 0, 0};
  cv.take<LpDev<T>>(slots);
  cv.take<LpState>(slots);
  cv.take<long long>((size_t)slots * kPhaseSlots);
  cv.take<long long>((size_t)slots * kPhaseSlots);
  cv.take<StepCtl>(1);
  cv.take<unsigned>(64);
  if (team > 1) {
    cv.take<LpState>(slots);
    cv.take<double>((size_t)slots * 2 * team * kMaxRed);
    cv.take<int32_t>((size_t)slots * 2 * (team + 1));
  }
  return cv.off + 256;
}

template <class T>
size_t server_lp_ws_T(const sovlp_lp* lp, long long hist_cap) {
  Carver cv{nullptr, 0, 0};
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  return cv.off + 256;
}

template <class T>
int server_create_T(int slots, int team, const sovlp_params* prm, int block, int phases, void* ws,
                    int64_t ws_bytes, hipStream_t st, sovlp_server* h) {
  if ((size_t)ws_bytes < server_ws_T<T>(slots, team)) return fail(1, "workspace too small");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T>* d_lps = cv.take<LpDev<T>>(slots);
  LpState* states = cv.take<LpState>(slots);
  long long* pcount = cv.take<long long>((size_t)slots * kPhaseSlots);
  long long* pticks = cv.take<long long>((size_t)slots * kPhaseSlots);
  StepCtl* ctl = cv.take<StepCtl>(1);
  unsigned* done = cv.take<unsigned>(64);
  h->team = team;
  LpState* states2 = nullptr;
  if (team > 1) {
    states2 = cv.take<LpState>(slots);
    h->partials = cv.take<double>((size_t)slots * 2 * team * kMaxRed);
    h->parts = cv.take<int32_t>((size_t)slots * 2 * (team + 1));
    HIPCHK(hipHostMalloc((void**)&h->hparts, sizeof(int32_t) * (size_t)slots * 2 * (team + 1),
                         hipHostMallocDefault));
  }
  h->slots = slots;
  h->precision = sizeof(T) == 8 ? 64 : 32;
  h->block = block;
  h->phases = phases;
  h->lps = d_lps;
  h->states = states;
  h->pcount = pcount;
  h->pticks = pticks;
  h->ctl = ctl;
  h->done = done;
  h->busy.assign(slots, 0);
  h->seen_done.assign(slots, 0);
  h->xo.assign(slots, nullptr);
  h->yo.assign(slots, nullptr);
  h->dout.assign(slots, nullptr);
  h->rhist.assign(slots, nullptr);
  h->hist.assign(slots, nullptr);
  HIPCHK(hipHostMalloc((void**)&h->hstage, sizeof(LpDev<T>) * (size_t)slots, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&h->hinit, sizeof(LpState) * (size_t)slots, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&h->hsnap, sizeof(LpState) * (size_t)slots, hipHostMallocDefault));
  LpDev<T>* stage = reinterpret_cast<LpDev<T>*>(h->hstage);
  for (int i = 0; i < slots; ++i) {
    stage[i] = LpDev<T>{};
    stage[i].st = states + i;
    stage[i].st2 = states2 ? states2 + i : nullptr;
    stage[i].pcount = pcount + (size_t)i * kPhaseSlots;
    stage[i].pticks = pticks + (size_t)i * kPhaseSlots;
    h->hinit[i] = LpState{};
    h->hinit[i].phase = PH_DONE;
    h->hinit[i].status = ST_RUNNING;
  }
  StepCtl c0{};
  c0.kstop = kNoStop;
  std::memcpy(&c0.P, prm, sizeof(Params));
  const double secs = prm->time_limit > 0.0 ? prm->time_limit : 0.0;
  double hz = 0.0;
  if (int rc = clock_hz(&hz)) return rc;
  h->hz = hz;
  c0.ticks = (long long)std::min(secs * hz, 9.0e18);
  HIPCHK(hipMemcpyAsync(d_lps, stage, sizeof(LpDev<T>) * (size_t)slots, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(states, h->hinit, sizeof(LpState) * (size_t)slots, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(pcount, 0, sizeof(long long) * kPhaseSlots * (size_t)slots, st));
  HIPCHK(hipMemsetAsync(pticks, 0, sizeof(long long) * kPhaseSlots * (size_t)slots, st));
  HIPCHK(hipMemcpyAsync(ctl, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(done, 0, sizeof(unsigned), st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  if (team > 1) {
    // `phases` step launches: one phase of every slot over `team` workgroups,
    // each workgroup first taking the slot's previous transition; then one
    // workgroup per slot takes the last one
    constexpr int TN = sizeof(T) == 8 ? kTnnz64 : kTnnz32;
    for (int i = 0; i < phases; ++i)
      pdlp_slot_step_kernel<T, kBlock, TN><<<slots * team, kBlock, 0, h->cap>>>(d_lps, team, i & 1,
                                                                              ctl, done);
    pdlp_settle_kernel<T, kBlock><<<slots, kBlock, 0, h->cap>>>(d_lps, team, phases & 1,
                                                                (phases - 1) & 1, ctl, done);
  } else {
    launch_run<T>(block, d_lps, slots, ctl, done, phases, h->cap);
  }
  HIPCHK(hipStreamEndCapture(h->cap, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  return 0;
}

template <class T>
int server_admit_T(sovlp_server* h, int slot, const sovlp_lp* lp, const double* x0,
                   const double* y0, double omega0, void* ws, int64_t ws_bytes,
                   int64_t hist_cap, hipStream_t st) {
  if ((size_t)ws_bytes < server_lp_ws_T<T>(lp, hist_cap)) return fail(1, "LP workspace too small");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  fill_lp<T>(L, *lp);
  L.st = h->states + slot;   // the state lives in the server's table, not the LP's workspace
  L.pcount = h->pcount + (size_t)slot * kPhaseSlots;
  L.pticks = h->pticks + (size_t)slot * kPhaseSlots;
  if (h->team > 1) {
    // equal-cost split of both tile lists over the team (host plan from the
    // device copies of the tile starts and row pointers)
    const int W = h->team;
    std::vector<int32_t> ta(lp->ntiles_a + 1), tt(lp->ntiles_t + 1), rp(lp->m + 1), cp(lp->n + 1);
    HIPCHK(hipMemcpyAsync(ta.data(), lp->tiles_a, sizeof(int32_t) * ta.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(tt.data(), lp->tiles_t, sizeof(int32_t) * tt.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rp.data(), lp->rp, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cp.data(), lp->cp, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int64_t> pa(ta.size()), pt(tt.size());
    for (size_t t = 0; t < ta.size(); ++t) pa[t] = rp[ta[t]];
    for (size_t t = 0; t < tt.size(); ++t) pt[t] = cp[tt[t]];
    int32_t* hp = h->hparts + (size_t)slot * 2 * (W + 1);
    host_partition(lp->ntiles_a, ta.data(), pa.data(), W, hp);
    host_partition(lp->ntiles_t, tt.data(), pt.data(), W, hp + W + 1);
    int32_t* dp = h->parts + (size_t)slot * 2 * (W + 1);
    HIPCHK(hipMemcpyAsync(dp, hp, sizeof(int32_t) * 2 * (W + 1), hipMemcpyHostToDevice, st));
    L.partA = dp;
    L.partT = dp + W + 1;
    L.partials = h->partials + (size_t)slot * 2 * W * kMaxRed;
    L.st2 = reinterpret_cast<LpDev<T>*>(h->hstage)[slot].st2;
  }
  L.wx0 = x0;
  L.wy0 = y0;
  L.womega = omega0 > 0.0 ? omega0 : 0.0;
  LpDev<T>* stage = reinterpret_cast<LpDev<T>*>(h->hstage);
  stage[slot] = L;
  h->hinit[slot] = LpState{};
  h->hinit[slot].phase = PH_INIT;
  h->hinit[slot].status = ST_RUNNING;
  LpDev<T>* d_lps = reinterpret_cast<LpDev<T>*>(h->lps);
  HIPCHK(hipMemcpyAsync(d_lps + slot, &stage[slot], sizeof(LpDev<T>), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(h->states + slot, &h->hinit[slot], sizeof(LpState), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(h->pcount + (size_t)slot * kPhaseSlots, 0, sizeof(long long) * kPhaseSlots, st));
  HIPCHK(hipMemsetAsync(h->pticks + (size_t)slot * kPhaseSlots, 0, sizeof(long long) * kPhaseSlots, st));
  h->xo[slot] = L.xo;
  h->yo[slot] = L.yo;
  h->dout[slot] = L.dout;
  h->rhist[slot] = L.rhist;
  h->hist[slot] = L.hist;
  return 0;
}


}  // namespace sovh
