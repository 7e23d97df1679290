// engine_api.hip -- C ABI of the PDLP engines (include/sovlp.h): the single-LP
// engine, the batch and the serving slot table.  The templated host code and
// the kernels it launches live in engine.hpp, instantiated per precision by
// engine_inst.hip.
#include "engine.hpp"

namespace sovh {
// Instantiated per precision in engine_inst.hip (two translation units built
// side by side); declared here so that this unit instantiates no kernel.
#define SOVLP_ENGINE_EXTERN(T)                                                              \
  extern template int pick_width<T>(const sovlp_lp*, int, int*);                          \
  extern template size_t single_ws_T<T>(const sovlp_lp*, long long, int);                 \
  extern template int single_create_T<T>(const sovlp_lp*, const sovlp_params*, void*,     \
                                         int64_t, int64_t, int, hipStream_t, sovlp_pdlp*); \
  extern template size_t batch_ws_T<T>(int, const sovlp_lp*, long long);                  \
  extern template int batch_create_T<T>(int, const sovlp_lp*, const sovlp_params*,        \
                                        const int32_t*, void*, int64_t, int64_t,          \
                                        hipStream_t, sovlp_batch*);                       \
  extern template int batch_reset_T<T>(sovlp_batch*, hipStream_t);                        \
  extern template size_t server_ws_T<T>(int, int);                                        \
  extern template size_t server_lp_ws_T<T>(const sovlp_lp*, long long);                   \
  extern template int server_create_T<T>(int, int, const sovlp_params*, int, int, void*,  \
                                         int64_t, hipStream_t, sovlp_server*);           \
  extern template int server_admit_T<T>(sovlp_server*, int, const sovlp_lp*,              \
                                        const double*, const double*, double, void*,      \
                                        int64_t, int64_t, hipStream_t);
SOVLP_ENGINE_EXTERN(double)
SOVLP_ENGINE_EXTERN(float)
}  // namespace sovh

using namespace sovh;

extern "C" {

int sovlp_pdlp_workspace_bytes(const sovlp_lp* lp, int precision, int64_t history_cap,
                               int64_t* bytes) {
  if (!lp || !bytes) return fail(1, "null argument");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  int W = 1;
  int rc = precision == 64 ? pick_width<double>(lp, 2, &W) : pick_width<float>(lp, 2, &W);
  if (rc) return rc;
  *bytes = (int64_t)(precision == 64 ? single_ws_T<double>(lp, history_cap, W)
                                     : single_ws_T<float>(lp, history_cap, W));
  return 0;
}

int sovlp_pdlp_create(const sovlp_lp* lp, const sovlp_params* params, int precision,
                      void* workspace, int64_t workspace_bytes, int64_t history_cap, int team,
                      sovlp_stream_t stream, sovlp_pdlp** out) {
  static_assert(sizeof(sovlp_params) == sizeof(Params), "params layout");
  if (!lp || !params || !workspace || !out) return fail(1, "null argument");
  if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return fail(1, "negative dimension");
  if (lp->nnz >= (int64_t)INT32_MAX) return fail(4, "more than 2^31-1 nonzeros");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (int rc = check_params(params, precision)) return rc;
  if (team < 0 || team > 2) return fail(1, "team must be 0 (auto), 1 (one workgroup) or 2 (many)");
  sovlp_pdlp* h = new sovlp_pdlp{};
  int rc = precision == 64
               ? single_create_T<double>(lp, params, workspace, workspace_bytes, history_cap, team,
                                         S(stream), h)
               : single_create_T<float>(lp, params, workspace, workspace_bytes, history_cap, team,
                                        S(stream), h);
  if (rc) {
    sovlp_pdlp_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int sovlp_pdlp_advance(sovlp_pdlp* h, int64_t iteration_budget, double seconds_left,
                       sovlp_stream_t stream) {
  if (!h) return fail(1, "null handle");
  hipStream_t st = S(stream);
  if (!h->started) {
    int rc = set_ticks(h->ctl, &h->ticks_set, seconds_left, h->hz, st);
    if (rc) return rc;
    h->started = true;
  }
  HIPCHK(hipMemcpyAsync(&h->hstate[0], h->st, sizeof(LpState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // iterations of every refinement round so far
  const long long k0 = h->hstate[0].k_done + h->hstate[0].k;
  const long long budget = iteration_budget < 1 ? 1 : iteration_budget;
  if (h->W == 1) {
    // the run kernel leaves the LP once it has run the budget, so that a
    // replay of up to kPhasesPerLaunch phases does not overshoot it
    *h->hkstop = k0 + budget;
    HIPCHK(hipMemcpyAsync(&h->ctl->kstop, h->hkstop, sizeof(long long), hipMemcpyHostToDevice, st));
  }
  // every replay runs up to phases_per_replay phases; an iteration takes at
  // most 61 phases (60 rejected trials), so this bound is never the one that
  // stops a healthy solve
  const long long max_replays = (budget * 61 + 200) / h->phases_per_replay + 16;
  // one replay stays queued while the host reads the state the previous one
  // left (pinned copy + event), so the device does not idle on the host
  long long replays = 0;
  int cur = 0;
  if (h->hstate[0].phase == PH_DONE) return 0;
  HIPCHK(hipGraphLaunch(h->exec, st));
  HIPCHK(hipMemcpyAsync(&h->hstate[0], h->st, sizeof(LpState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(h->ev[0], st));
  ++replays;
  for (;;) {
    const int nxt = cur ^ 1;
    const bool more = replays < max_replays;
    if (more) {
      HIPCHK(hipGraphLaunch(h->exec, st));
      HIPCHK(hipMemcpyAsync(&h->hstate[nxt], h->st, sizeof(LpState), hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(h->ev[nxt], st));
      ++replays;
    }
    HIPCHK(hipEventSynchronize(h->ev[cur]));
    const LpState& s = h->hstate[cur];
    if (s.phase == PH_DONE || s.k_done + s.k - k0 >= budget || !more) break;
    cur = nxt;
  }
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int sovlp_pdlp_info(sovlp_pdlp* h, sovlp_info* info, sovlp_stream_t stream) {
  if (!h || !info) return fail(1, "null argument");
  LpState s;
  HIPCHK(hipMemcpyAsync(&s, h->st, sizeof s, hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  fill_info(s, info);
  return 0;
}

int sovlp_pdlp_outputs(sovlp_pdlp* h, double** x, double** y, double** d, double** history) {
  if (!h) return fail(1, "null handle");
  if (x) *x = h->xo;
  if (y) *y = h->yo;
  if (d) *d = h->dout;
  if (history) *history = h->hist;
  return 0;
}

int sovlp_pdlp_refinement(sovlp_pdlp* h, double** rows) {
  if (!h || !rows) return fail(1, "null argument");
  *rows = h->rhist;
  return 0;
}

int sovlp_pdlp_scaling(sovlp_pdlp* h, double** row, double** col) {
  if (!h) return fail(1, "null handle");
  if (row) *row = h->rowf;
  if (col) *col = h->colf;
  return 0;
}

int sovlp_phases_per_launch(int* phases) {
  if (!phases) return fail(1, "null argument");
  *phases = kPhasesPerLaunch;
  return 0;
}

int sovlp_pdlp_width(sovlp_pdlp* h, int* workgroups) {
  if (!h || !workgroups) return fail(1, "null argument");
  *workgroups = h->W;
  return 0;
}

int sovlp_phase_slots(int* slots) {
  if (!slots) return fail(1, "null argument");
  *slots = kPhaseSlots;
  return 0;
}

const char* sovlp_phase_name(int slot) {
  static const char* names[kPhaseSlots] = {
      "INIT", "RUIZ_MAX", "RUIZ_APPLY", "PC_SUM", "PC_APPLY", "SCALE", "AX0", "POW_A",
      "POW_T", "KKT0", "P_TRIAL", "A_TRIAL", "B_ADAPT", "A_FUSED", "B_FUSED", "CHECK",
      "RESTART", "COPY_AVG", "FINAL", "FINAL_D", "REF_X", "REF_KKT", "REF_SHIFT", "DONE",
      "B_ADAPT_NEXT", "B_FUSED_NEXT", "", ""};
  return slot >= 0 && slot < kPhaseSlots ? names[slot] : "";
}

int sovlp_pdlp_phase_counts(sovlp_pdlp* h, int64_t* counts, sovlp_stream_t stream) {
  if (!h || !counts) return fail(1, "null argument");
  static_assert(sizeof(long long) == sizeof(int64_t), "counter width");
  HIPCHK(hipMemcpyAsync(counts, h->pcount, sizeof(long long) * kPhaseSlots,
                        hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_pdlp_phase_ticks(sovlp_pdlp* h, int64_t* ticks, sovlp_stream_t stream) {
  if (!h || !ticks) return fail(1, "null argument");
  HIPCHK(hipMemcpyAsync(ticks, h->pticks, sizeof(long long) * kPhaseSlots,
                        hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_pdlp_destroy(sovlp_pdlp* h) {
  if (!h) return 0;
  for (hipEvent_t e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->hstate) (void)hipHostFree(h->hstate);
  if (h->hkstop) (void)hipHostFree(h->hkstop);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
  return 0;
}

int sovlp_batch_geometry(int precision, int block, int* tile_rows, int* tile_nnz) {
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128, 256 or 512");
  *tile_rows = kRowsPerLane * block;
  *tile_nnz = batch_tnnz(precision, block);
  return 0;
}

int sovlp_lds_resident(int precision, int block, int m, int n, int64_t nnz, int tiles_a,
                       int tiles_t, int* resident) {
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128, 256 or 512");
  if (!resident || m < 0 || n < 0 || nnz < 0 || tiles_a < 0 || tiles_t < 0)
    return fail(1, "bad arguments");
  const int tn = batch_tnnz(precision, block);
  *resident = precision == 64
                  ? run_kernel_lds_mode<double>(m, n, nnz, tiles_a, tiles_t, block, tn)
                  : run_kernel_lds_mode<float>(m, n, nnz, tiles_a, tiles_t, block, tn);
  return 0;
}

int sovlp_batch_workspace_bytes(int count, const sovlp_lp* lps, int precision,
                                int64_t history_cap, int64_t* bytes) {
  if (count < 1 || !lps || !bytes) return fail(1, "bad batch arguments");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *bytes = (int64_t)(precision == 64 ? batch_ws_T<double>(count, lps, history_cap)
                                     : batch_ws_T<float>(count, lps, history_cap));
  return 0;
}

int sovlp_batch_create_mixed(int count, const sovlp_lp* lps, const sovlp_params* params,
                             int precision, const int32_t* blocks, void* workspace,
                             int64_t workspace_bytes, int64_t history_cap, sovlp_stream_t stream,
                             sovlp_batch** out) {
  if (count < 1 || !lps || !params || !blocks || !workspace || !out)
    return fail(1, "bad batch arguments");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  for (int i = 0; i < count; ++i)
    if (!batch_block_ok(blocks[i])) return fail(1, "block must be 64, 128, 256 or 512");
  if (int rc = check_params(params, precision)) return rc;
  sovlp_batch* h = new sovlp_batch{};
  int rc = precision == 64
               ? batch_create_T<double>(count, lps, params, blocks, workspace, workspace_bytes,
                                        history_cap, S(stream), h)
               : batch_create_T<float>(count, lps, params, blocks, workspace, workspace_bytes,
                                       history_cap, S(stream), h);
  if (rc) {
    sovlp_batch_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int sovlp_batch_create(int count, const sovlp_lp* lps, const sovlp_params* params,
                       int precision, int block, void* workspace, int64_t workspace_bytes,
                       int64_t history_cap, sovlp_stream_t stream, sovlp_batch** out) {
  if (count < 1) return fail(1, "bad batch arguments");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128, 256 or 512");
  std::vector<int32_t> blocks(count, block);
  return sovlp_batch_create_mixed(count, lps, params, precision, blocks.data(), workspace,
                                  workspace_bytes, history_cap, stream, out);
}

int sovlp_batch_run(sovlp_batch* h, double seconds_per_lp, int64_t max_graphs,
                    sovlp_stream_t stream, int64_t* graphs_run) {
  if (!h) return fail(1, "null handle");
  hipStream_t st = S(stream);
  int rc = set_ticks(h->ctl, &h->ticks_set, seconds_per_lp, h->hz, st);
  if (rc) return rc;
  rc = h->precision == 64 ? batch_reset_T<double>(h, st) : batch_reset_T<float>(h, st);
  if (rc) return rc;
  // one replay stays queued while the host reads the done counter the
  // previous one left (pinned copy + event)
  long long g = 0;
  int cur = 0;
  HIPCHK(hipGraphLaunch(h->exec, st));
  HIPCHK(hipMemcpyAsync(&h->hdone[0], h->done, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(h->ev[0], st));
  ++g;
  for (;;) {
    const int nxt = cur ^ 1;
    const bool more = max_graphs <= 0 || g < max_graphs;
    if (more) {
      HIPCHK(hipGraphLaunch(h->exec, st));
      HIPCHK(hipMemcpyAsync(&h->hdone[nxt], h->done, sizeof(unsigned), hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(h->ev[nxt], st));
      ++g;
    }
    HIPCHK(hipEventSynchronize(h->ev[cur]));
    if (h->hdone[cur] >= (unsigned)h->count || !more) break;
    cur = nxt;
  }
  HIPCHK(hipStreamSynchronize(st));
  if (graphs_run) *graphs_run = g;
  return 0;
}

int sovlp_batch_infos(sovlp_batch* h, sovlp_info* infos, sovlp_stream_t stream) {
  if (!h || !infos) return fail(1, "null argument");
  std::vector<LpState> s(h->count);
  for (int i = 0; i < h->count; ++i)
    HIPCHK(hipMemcpyAsync(&s[i], h->st[i], sizeof(LpState), hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  for (int i = 0; i < h->count; ++i) fill_info(s[i], &infos[i]);
  return 0;
}

int sovlp_batch_outputs(sovlp_batch* h, int index, double** x, double** y, double** d) {
  if (!h || index < 0 || index >= h->count) return fail(1, "bad batch index");
  if (x) *x = h->xo[index];
  if (y) *y = h->yo[index];
  if (d) *d = h->dout[index];
  return 0;
}

int sovlp_batch_refinement(sovlp_batch* h, int index, double** rows) {
  if (!h || !rows || index < 0 || index >= h->count) return fail(1, "bad batch index");
  *rows = h->rhist[index];
  return 0;
}

int sovlp_batch_phase_counts(sovlp_batch* h, int64_t* counts, sovlp_stream_t stream) {
  if (!h || !counts) return fail(1, "null argument");
  HIPCHK(hipMemcpyAsync(counts, h->pcount, sizeof(long long) * kPhaseSlots * (size_t)h->count,
                        hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_batch_phase_ticks(sovlp_batch* h, int64_t* ticks, sovlp_stream_t stream) {
  if (!h || !ticks) return fail(1, "null argument");
  HIPCHK(hipMemcpyAsync(ticks, h->pticks, sizeof(long long) * kPhaseSlots * (size_t)h->count,
                        hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_batch_destroy(sovlp_batch* h) {
  if (!h) return 0;
  for (int k = 0; k < kBatchBlocks - 1; ++k) {
    if (h->side[k]) (void)hipStreamDestroy(h->side[k]);
    if (h->join[k]) (void)hipEventDestroy(h->join[k]);
  }
  if (h->fork) (void)hipEventDestroy(h->fork);
  for (hipEvent_t e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->hdone) (void)hipHostFree(h->hdone);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
  return 0;
}

}  // extern "C"

extern "C" {

int sovlp_server_workspace_bytes(int slots, int precision, int64_t* bytes) {
  return sovlp_server_team_workspace_bytes(slots, 1, precision, bytes);
}

int sovlp_server_team_workspace_bytes(int slots, int team, int precision, int64_t* bytes) {
  if (slots < 1 || team < 1 || !bytes) return fail(1, "slots and team must be >= 1");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *bytes = (int64_t)(precision == 64 ? server_ws_T<double>(slots, team)
                                     : server_ws_T<float>(slots, team));
  return 0;
}

int sovlp_server_lp_bytes(const sovlp_lp* lp, int precision, int64_t history_cap,
                          int64_t* bytes) {
  if (!lp || !bytes) return fail(1, "null argument");
  if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return fail(1, "negative dimension");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *bytes = (int64_t)(precision == 64 ? server_lp_ws_T<double>(lp, history_cap)
                                     : server_lp_ws_T<float>(lp, history_cap));
  return 0;
}

static int server_create(int slots, int team, const sovlp_params* params, int precision,
                         int block, int phases_per_launch, void* workspace,
                         int64_t workspace_bytes, sovlp_stream_t stream, sovlp_server** out) {
  if (slots < 1 || team < 1 || !params || !workspace || !out)
    return fail(1, "bad server arguments");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128, 256 or 512");
  if (phases_per_launch < 1 || phases_per_launch > 4096)
    return fail(1, "phases_per_launch must be in [1, 4096]");
  if ((long long)slots * team > (1LL << 24)) return fail(1, "slots x team too large");
  if (int rc = check_params(params, precision)) return rc;
  sovlp_server* h = new sovlp_server{};
  int rc = precision == 64
               ? server_create_T<double>(slots, team, params, block, phases_per_launch, workspace,
                                         workspace_bytes, S(stream), h)
               : server_create_T<float>(slots, team, params, block, phases_per_launch, workspace,
                                        workspace_bytes, S(stream), h);
  if (rc) {
    sovlp_server_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int sovlp_server_create(int slots, const sovlp_params* params, int precision, int block,
                        int phases_per_launch, void* workspace, int64_t workspace_bytes,
                        sovlp_stream_t stream, sovlp_server** out) {
  return server_create(slots, 1, params, precision, block, phases_per_launch, workspace,
                       workspace_bytes, stream, out);
}

int sovlp_server_create_team(int slots, int team, const sovlp_params* params, int precision,
                             int pairs_per_launch, void* workspace, int64_t workspace_bytes,
                             sovlp_stream_t stream, sovlp_server** out) {
  return server_create(slots, team, params, precision, kBlock, pairs_per_launch, workspace,
                       workspace_bytes, stream, out);
}

int sovlp_server_admit(sovlp_server* h, int slot, const sovlp_lp* lp, const double* x0,
                       const double* y0, double omega0, void* workspace,
                       int64_t workspace_bytes, int64_t history_cap, sovlp_stream_t stream) {
  if (!h || !lp || !workspace) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots) return fail(1, "slot out of range");
  if (h->busy[slot]) return fail(1, "slot is occupied (release it first)");
  if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return fail(1, "negative dimension");
  if (lp->nnz >= (int64_t)INT32_MAX) return fail(4, "more than 2^31-1 nonzeros");
  int rc = h->precision == 64
               ? server_admit_T<double>(h, slot, lp, x0, y0, omega0, workspace, workspace_bytes,
                                        history_cap, S(stream))
               : server_admit_T<float>(h, slot, lp, x0, y0, omega0, workspace, workspace_bytes,
                                       history_cap, S(stream));
  if (rc) return rc;
  h->busy[slot] = 1;
  h->seen_done[slot] = 0;
  return 0;
}

int sovlp_server_step(sovlp_server* h, int launches, sovlp_stream_t stream, int32_t* finished,
                      int32_t* running) {
  if (!h) return fail(1, "null handle");
  hipStream_t st = S(stream);
  for (int i = 0; i < launches; ++i) HIPCHK(hipGraphLaunch(h->exec, st));
  HIPCHK(hipMemcpyAsync(h->hsnap, h->states, sizeof(LpState) * (size_t)h->slots,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  int run = 0;
  for (int i = 0; i < h->slots; ++i) {
    const bool fin = h->busy[i] && h->hsnap[i].phase == PH_DONE;
    if (fin) h->seen_done[i] = 1;
    if (finished) finished[i] = fin ? 1 : 0;
    run += (h->busy[i] && !fin) ? 1 : 0;
  }
  if (running) *running = run;
  return 0;
}

int sovlp_server_info(sovlp_server* h, int slot, sovlp_info* info) {
  if (!h || !info) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  fill_info(h->hsnap[slot], info);
  return 0;
}

int sovlp_server_outputs(sovlp_server* h, int slot, double** x, double** y, double** d) {
  if (!h) return fail(1, "null handle");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  if (x) *x = h->xo[slot];
  if (y) *y = h->yo[slot];
  if (d) *d = h->dout[slot];
  return 0;
}

int sovlp_server_refinement(sovlp_server* h, int slot, double** rows) {
  if (!h || !rows) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  *rows = h->rhist[slot];
  return 0;
}

int sovlp_server_history(sovlp_server* h, int slot, double** rows) {
  if (!h || !rows) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  *rows = h->hist[slot];
  return 0;
}

int sovlp_server_phase_counts(sovlp_server* h, int slot, int64_t* counts,
                              sovlp_stream_t stream) {
  if (!h || !counts) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots) return fail(1, "slot out of range");
  HIPCHK(hipMemcpyAsync(counts, h->pcount + (size_t)slot * kPhaseSlots,
                        sizeof(long long) * kPhaseSlots, hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_server_release(sovlp_server* h, int slot) {
  if (!h) return fail(1, "null handle");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  if (!h->seen_done[slot]) return fail(1, "the slot's LP has not finished (step until it has)");
  h->busy[slot] = 0;
  h->seen_done[slot] = 0;
  h->xo[slot] = h->yo[slot] = h->dout[slot] = h->rhist[slot] = h->hist[slot] = nullptr;
  return 0;
}

int sovlp_server_clock_hz(sovlp_server* h, double* hz) {
  if (!h || !hz) return fail(1, "null argument");
  *hz = h->hz;
  return 0;
}

int sovlp_server_destroy(sovlp_server* h) {
  if (!h) return 0;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  if (h->hstage) (void)hipHostFree(h->hstage);
  if (h->hinit) (void)hipHostFree(h->hinit);
  if (h->hsnap) (void)hipHostFree(h->hsnap);
  if (h->hparts) (void)hipHostFree(h->hparts);
  delete h;
  return 0;
}

}  // extern "C"
