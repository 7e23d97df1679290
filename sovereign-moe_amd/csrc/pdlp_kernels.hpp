// pdlp_kernels.hpp -- the PDLP launch kernels (device code only).
//
// Each launch runs phases (pdlp_solver.hpp) of the LPs it serves:
//   pdlp_run_kernel      one workgroup per LP runs up to `max_phases` phases of
//                        its LP back to back (state in LDS, workgroup barriers
//                        between phases; no workgroup waits for another)
//   pdlp_step_kernel     W workgroups share one LP's phase: each first takes
//                        the previous phase's transition from the W partial
//                        sums (fixed order, the same in every workgroup), then
//                        runs the phase and writes its partial sums
//   pdlp_settle_kernel   one workgroup takes the last pending transition
//   state_reset_kernel   puts a batch back at the start of its solves
#pragma once
#include "pdlp_solver.hpp"

namespace sov {

// LDS of a phase: the products of one tile, two vectors of T (the CHECK
// phase) or one vector of doubles (scaling phases, FP64 check passes)
template <class T>
constexpr int lds_doubles_of(int tnnz) {
  return sizeof(T) == 8 ? 2 * tnnz : tnnz;
}
template <class T, int TNNZ>
constexpr int lds_doubles() {
  return lds_doubles_of<T>(TNNZ);
}

// The run kernel's staging LDS (doubles): a phase's tile products, and an LP's
// resident copy when it fits.  A 512-thread workgroup holds half of a CU's
// waves (4 per SIMD under the register cap), so two share a CU and nothing
// else fits beside them: each gets half of the CU's 160 KB less the kernel's
// other LDS (under 3 KB).  Smaller workgroups keep the staging size, so that
// as many of them fit a CU as its waves allow.
constexpr int kRunLds512 = 9728;   // 76 KB
template <class T>
constexpr int run_lds_doubles_of(int block, int tnnz) {
  return block == 512 && kRunLds512 > lds_doubles_of<T>(tnnz) ? kRunLds512
                                                              : lds_doubles_of<T>(tnnz);
}

// What the run kernel of BLOCK threads (tile cap tnnz) holds in LDS for an LP
// of this shape: 1 its matrix and step vectors, 2 only its step vectors, 0
// nothing -- the rule it applies to each LP, for the host.
template <class T>
inline int run_kernel_lds_mode(int m, int n, long long nnz, int ntA, int ntT, int block,
                               int tnnz) {
  const long long cap = 8LL * run_lds_doubles_of<T>(block, tnnz);
  if (resident_fits<T>(m, n, nnz, ntA, ntT, cap, tnnz)) return 1;
  return vec_fits<T>(m, n, ntA, ntT, cap, tnnz) ? 2 : 0;
}

constexpr long long kNoStop = 0x7fffffffffffffffLL;

struct StepCtl {             // device-resident control block (read by every launch)
  Params P;
  long long ticks;           // time budget in device clock ticks, taken when a solve starts
  long long kstop;           // a run launch leaves an LP once it has run this many
                             // iterations (all refinement rounds); kNoStop: never
};

__device__ __forceinline__ void init_part(int ph, double (&part)[kMaxRed]) {
#pragma unroll
  for (int k = 0; k < kMaxRed; ++k) part[k] = 0.0;
  if (ph == PH_PC_APPLY) part[1] = -__builtin_inf();
}

// Whether a phase's transition reads its sums: the fused A phase reduces
// nothing, and the fused B phase only ||dx||^2 for a step-size trial that
// follows it (the transition reads it then only).
__device__ __forceinline__ bool phase_reduces(int ph, int next_mode) {
  return !(ph == PH_A_FUSED || (ph == PH_B_FUSED && next_mode != NEXT_TRIAL));
}

// Workgroup totals of a phase's partial sums: CHECK has NCHK values, every
// other phase at most 6.  Leaves the totals in red[(BLOCK/64)*kMaxRed + k].
template <int BLOCK>
__device__ __forceinline__ void reduce_phase_to_lds(int ph, const double (&part)[kMaxRed],
                                                    double* red) {
  const uint32_t mm = phase_maxmask(ph);
  if (ph == PH_CHECK) {
    block_reduce_to_lds<BLOCK, kMaxRed>(part, mm, red);
  } else if (ph == PH_P_TRIAL || ph == PH_B_ADAPT || ph == PH_B_FUSED) {
    // the loop's transitions read one total (||dx||^2) ...
    double v[1] = {part[0]};
    block_reduce_to_lds<BLOCK, 1>(v, mm, red);
  } else if (ph == PH_A_TRIAL) {
    // ... or two (dy.adx, ||dy||^2): each total is reduced on its own, so
    // leaving the unread ones out changes none of the others
    double v[2] = {part[0], part[1]};
    block_reduce_to_lds<BLOCK, 2>(v, mm, red);
  } else {
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = part[k];
    block_reduce_to_lds<BLOCK, 6>(v, mm, red);
  }
}

// Thread 0: apply the transition of phase `ph` (totals `t` in LDS) to the
// state and store it; count the LP as done when it terminates.
template <class T>
__device__ __forceinline__ void take_transition(const LpDev<T>& L, const LpState& s, int ph,
                                                const double* t, const StepCtl* ctl,
                                                unsigned* done, LpState& ss) {
  ss = s;
  if (ph == PH_INIT) {
    ss.tick0 = (long long)wall_clock64();
    ss.ticks_left = ctl->ticks;
  }
  L.pcount[ph] += 1;
  if ((ph == PH_B_ADAPT || ph == PH_B_FUSED) && s.next_mode != NEXT_NONE)
    L.pcount[ph == PH_B_ADAPT ? kSlotBAdaptNext : kSlotBFusedNext] += 1;
  Transition tr{ctl->P, L.offset0, L.flip};
  ss.phase = tr.next(ss, ph, t, L.hist, L.hist_cap, L.rhist);
  if (ss.phase == PH_DONE) {
    ss.ticks = (long long)wall_clock64() - ss.tick0;
    atomicAdd(done, 1u);
  }
  *L.st = ss;
}

// Thread 0: apply the transition of phase `ph` to the state held in LDS;
// `pc` are the launch's phase counters (LDS) and `P` the parameters (LDS).
template <class T>
__device__ __forceinline__ void transition_in_lds(const LpDev<T>& L, int ph, const double* t,
                                                  const StepCtl* ctl, const Params& P,
                                                  unsigned* done, int* pc, LpState& ss) {
  if (ph == PH_INIT) {
    ss.tick0 = (long long)wall_clock64();
    ss.ticks_left = ctl->ticks;
  }
  pc[ph] += 1;
  if ((ph == PH_B_ADAPT || ph == PH_B_FUSED) && ss.next_mode != NEXT_NONE)
    pc[ph == PH_B_ADAPT ? kSlotBAdaptNext : kSlotBFusedNext] += 1;
  Transition tr{P, L.offset0, L.flip};
  ss.phase = tr.next(ss, ph, t, L.hist, L.hist_cap, L.rhist);
  if (ss.phase == PH_DONE) {
    ss.ticks = (long long)wall_clock64() - ss.tick0;
    atomicAdd(done, 1u);
  }
}

// Thread 0: the end of an iteration (after a B phase), the most frequent
// transition, on the state and parameters in LDS: LDS accesses only (the LP's
// offset and sense, which only the other transitions read, are not loaded).
// It never ends the solve.
__device__ __forceinline__ void end_iteration_in_lds(int ph, const double* t, const Params& P,
                                                     int* pc, LpState& ss) {
  pc[ph] += 1;
  if (ss.next_mode != NEXT_NONE) pc[ph == PH_B_ADAPT ? kSlotBAdaptNext : kSlotBFusedNext] += 1;
  const Transition tr{P, 0.0, 0};
  ss.phase = tr.end_iteration(ss, t[0]);
}

// The phases of the PDHG loop: they read the matrix only through tile passes
// over the scaled copies and write dx / dy only in their epilogues, so they
// can run on an LP's resident copy in LDS (ResidentLds).
__device__ __forceinline__ bool loop_phase(int ph) {
  return ph == PH_P_TRIAL || ph == PH_A_TRIAL || ph == PH_B_ADAPT || ph == PH_A_FUSED ||
         ph == PH_B_FUSED;
}

// One workgroup per LP (blockIdx.x): up to `max_phases` phases of the LP in a
// row.  Register use is capped for 4 waves per SIMD (a few spills, measured
// faster at the batch's p50 than 3 waves without them).  Phase k+1 reads what phase k wrote to global memory; the workgroup
// barrier between them orders those accesses (all of a workgroup's waves share
// one CU).  The loop is bounded, and the state goes back to HBM at the end.
// An LP whose resident copy fits the staging LDS runs the loop phases on it:
// the copy is filled before the first loop phase of the launch and again after
// any other phase (which may stage its tiles in the same LDS).
template <class T, int BLOCK, int TNNZ>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8)))
    pdlp_run_kernel(const LpDev<T>* __restrict__ lps, const StepCtl* __restrict__ ctl,
                    unsigned* done, int max_phases) {
  __shared__ __align__(16) double lds_raw[run_lds_doubles_of<T>(BLOCK, TNNZ)];
  __shared__ double red[(BLOCK / kWave + 2) * kMaxRed];
  __shared__ LpState ss;
  __shared__ Params sp;                  // the parameters, read once per launch
  __shared__ long long kstop;
  __shared__ int pc[kPhaseSlots];        // phase counts of this launch
  __shared__ long long pt[kPhaseSlots];  // device clock ticks of each phase in this launch
  const LpDev<T>& L = lps[blockIdx.x];
  if (threadIdx.x == 0) {
    ss = *L.st;
    sp = ctl->P;
    kstop = ctl->kstop;
  }
  if (threadIdx.x < kPhaseSlots) {
    pc[threadIdx.x] = 0;
    pt[threadIdx.x] = 0;
  }
  __syncthreads();
  if (uniform(ss.phase) == PH_DONE) return;
  const int m = uniform(L.m), n = uniform(L.n), ntA = uniform(L.ntA), ntT = uniform(L.ntT);
  const int nnz = uniform(L.ap[m]);
  const bool resident =
      uniform(!(sp.flags & kFlagNoLdsMatrix) &&
              resident_fits<T>(m, n, nnz, ntA, ntT, (long long)sizeof(lds_raw), TNNZ));
  // else only dx and dy, above the loop phases' staging area
  const bool vec = uniform(!resident && !(sp.flags & kFlagNoLdsMatrix) &&
                           vec_fits<T>(m, n, ntA, ntT, (long long)sizeof(lds_raw), TNNZ));
  ResidentLds<T> rl{};
  if (resident) rl = carve_resident<T>(lds_raw, m, n, nnz, ntA);
  if (vec) rl = carve_vec<T>(lds_raw, m, n, ntA, ntT, TNNZ);
  bool res_ok = false;       // the LDS copy holds the current matrix (or only dx and dy)
  // thread 0: when the current phase started (the end of the previous one)
  long long tp = threadIdx.x == 0 ? (long long)wall_clock64() : 0;
  for (int it = 0; it < max_phases; ++it) {
    const int ph = uniform(ss.phase);
    if (ph == PH_DONE) break;
    if (ss.k_done + ss.k >= kstop) break;   // the caller's iteration budget
    // the fused A phase reduces nothing, and the fused B phase only ||dx||^2
    // for a step-size trial that follows it.  Read here, while no thread
    // writes the state: thread 0 changes it right after the phase.
    const bool reduce = phase_reduces(ph, uniform(ss.next_mode));
    const bool on_res = resident && loop_phase(ph);
    const bool on_vec = vec && loop_phase(ph);
    if ((on_res || on_vec) && !res_ok) {
      if (on_res) load_resident<T, BLOCK>(L, nnz, rl);
      else load_vec<T, BLOCK>(L, rl);
      __syncthreads();
      res_ok = true;
    }
    if (!on_res && !on_vec) res_ok = false;   // this phase may stage tiles over the copy
    double part[kMaxRed];
    init_part(ph, part);
    PhaseWork<T, BLOCK, TNNZ> W(L, ss, reinterpret_cast<T*>(lds_raw), 0, 1, rl, on_res, on_vec);
    W.run(ph, part);
    if (reduce) reduce_phase_to_lds<BLOCK>(ph, part, red);
    if (threadIdx.x == 0) {
      if (ph == PH_A_FUSED) {
        pc[PH_A_FUSED] += 1;
        ss.phase = PH_B_FUSED;
      } else if (ph == PH_B_FUSED || ph == PH_B_ADAPT) {
        end_iteration_in_lds(ph, red + (BLOCK / kWave) * kMaxRed, sp, pc, ss);
      } else if (ph == PH_P_TRIAL) {     // the P_TRIAL case of Transition::apply
        pc[PH_P_TRIAL] += 1;
        ss.dxdx = red[(BLOCK / kWave) * kMaxRed];
        ss.phase = PH_A_TRIAL;
      } else if (ph == PH_A_TRIAL) {     // the step-size rule on the LDS state
        pc[PH_A_TRIAL] += 1;
        const Transition tr{sp, 0.0, 0};
        ss.phase = tr.trial_step(ss, red + (BLOCK / kWave) * kMaxRed);
      } else {
        transition_in_lds(L, ph, red + (BLOCK / kWave) * kMaxRed, ctl, sp, done, pc, ss);
      }
      const long long now = (long long)wall_clock64();
      pt[ph] += now - tp;
      tp = now;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *L.st = ss;
  if (threadIdx.x < kPhaseSlots && pc[threadIdx.x] != 0) {
    L.pcount[threadIdx.x] += pc[threadIdx.x];
    L.pticks[threadIdx.x] += pt[threadIdx.x];
  }
}

// Sum of the W partials of a phase in a fixed order (lane t: partials t,
// t + BLOCK, ...; then the workgroup tree), the K values a phase's transition
// reads: the totals land in red[(BLOCK/64)*kMaxRed + k].
template <int BLOCK, int K>
__device__ __forceinline__ void sum_partials(const double* __restrict__ pin, int W, uint32_t mm,
                                             double* red) {
  double part[K];
#pragma unroll
  for (int k = 0; k < K; ++k) part[k] = ((mm >> k) & 1u) ? -__builtin_inf() : 0.0;
  for (int b = threadIdx.x; b < W; b += BLOCK) {
    const double* pb = pin + (size_t)b * kMaxRed;
#pragma unroll
    for (int k = 0; k < K; ++k)
      part[k] = ((mm >> k) & 1u) ? nanmax(part[k], pb[k]) : part[k] + pb[k];
  }
  block_reduce_to_lds<BLOCK, K>(part, mm, red);
}

// W workgroups of one LP: the step launches replace the phase + decide
// launch pair by one launch.  Every workgroup first takes the pending
// transition of the phase before it itself (the same partials summed in the
// same order, so every workgroup computes the same state; only the team's
// lead, rank 0, writes history rows, counts the phase and stores the state),
// then runs the new phase and writes its partials.  State and partials
// alternate between two copies by launch parity: launch k reads state copy
// k % 2 and partial half (k + 1) % 2 and writes the other ones, so no launch
// reads what another workgroup of the same launch writes.  L.partials holds
// the two halves, [2][W][kMaxRed].
template <class T, int BLOCK, int TNNZ>
__device__ __forceinline__ void team_step(const LpDev<T>& L, int rank, int W, int parity,
                                          const StepCtl* __restrict__ ctl, unsigned* done,
                                          T* lds, double* red, LpState& ss, int& go) {
  LpState* in = parity ? L.st2 : L.st;
  LpState* out = parity ? L.st : L.st2;
  const double* pin = L.partials + (size_t)(parity ^ 1) * W * kMaxRed;
  double* pout = L.partials + (size_t)parity * W * kMaxRed;
  const bool lead = rank == 0;
  if (threadIdx.x == 0) ss = *in;
  __syncthreads();
  if (uniform(ss.pending)) {
    const int ph = uniform(ss.phase);
    const uint32_t mm = phase_maxmask(ph);
    if (ph == PH_CHECK) sum_partials<BLOCK, kMaxRed>(pin, W, mm, red);
    else if (phase_reduces(ph, uniform(ss.next_mode))) sum_partials<BLOCK, 6>(pin, W, mm, red);
    if (threadIdx.x == 0) {
      const double* t = red + (BLOCK / kWave) * kMaxRed;
      if (ph == PH_INIT) {
        ss.tick0 = (long long)wall_clock64();
        ss.ticks_left = ctl->ticks;
      }
      if (lead) {
        L.pcount[ph] += 1;
        if ((ph == PH_B_ADAPT || ph == PH_B_FUSED) && ss.next_mode != NEXT_NONE)
          L.pcount[ph == PH_B_ADAPT ? kSlotBAdaptNext : kSlotBFusedNext] += 1;
      }
      Transition tr{ctl->P, L.offset0, L.flip, lead};
      ss.phase = tr.next(ss, ph, t, L.hist, L.hist_cap, L.rhist);
      ss.pending = 0;
      if (ss.phase == PH_DONE && lead) {
        ss.ticks = (long long)wall_clock64() - ss.tick0;
        atomicAdd(done, 1u);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) go = ss.phase != PH_DONE;
  __syncthreads();
  if (!uniform(go)) {
    if (lead && threadIdx.x == 0) *out = ss;
    return;
  }
  const int ph = uniform(ss.phase);
  const bool reduce = phase_reduces(ph, uniform(ss.next_mode));
  double part[kMaxRed];
  init_part(ph, part);
  PhaseWork<T, BLOCK, TNNZ> Wk(L, ss, lds, rank, W);
  Wk.run(ph, part);
  if (reduce) {
    reduce_phase_to_lds<BLOCK>(ph, part, red);
    constexpr int NW = BLOCK / kWave;
    if (threadIdx.x < kMaxRed)
      pout[(size_t)rank * kMaxRed + threadIdx.x] = red[NW * kMaxRed + threadIdx.x];
  } else {
    __syncthreads();      // every thread is done with `ss` before the lead moves it
  }
  if (lead && threadIdx.x == 0) {
    ss.pending = 1;
    *out = ss;
  }
}

// One LP over all gridDim.x workgroups.  Registers are capped for 4 waves per
// SIMD, as in the run kernel (the FP64 instance otherwise takes 132 and runs
// 3 workgroups per CU).
template <class T, int BLOCK, int TNNZ>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8)))
    pdlp_step_kernel(const LpDev<T>* __restrict__ lp, int parity, const StepCtl* __restrict__ ctl,
                     unsigned* done) {
  __shared__ __align__(16) double lds_raw[lds_doubles<T, TNNZ>()];
  __shared__ double red[(BLOCK / kWave + 2) * kMaxRed];
  __shared__ LpState ss;
  __shared__ int go;
  team_step<T, BLOCK, TNNZ>(*lp, (int)blockIdx.x, (int)gridDim.x, parity, ctl, done,
                            reinterpret_cast<T*>(lds_raw), red, ss, go);
}

// Served LPs: slot blockIdx.x / team over `team` workgroups each (rank
// blockIdx.x % team).
template <class T, int BLOCK, int TNNZ>
__global__ void __launch_bounds__(BLOCK)
    pdlp_slot_step_kernel(const LpDev<T>* __restrict__ lps, int team, int parity,
                          const StepCtl* __restrict__ ctl, unsigned* done) {
  __shared__ __align__(16) double lds_raw[lds_doubles<T, TNNZ>()];
  __shared__ double red[(BLOCK / kWave + 2) * kMaxRed];
  __shared__ LpState ss;
  __shared__ int go;
  const int slot = (int)blockIdx.x / team, rank = (int)blockIdx.x % team;
  team_step<T, BLOCK, TNNZ>(lps[slot], rank, team, parity, ctl, done,
                            reinterpret_cast<T*>(lds_raw), red, ss, go);
}

// After a run of step launches: one workgroup per LP (blockIdx.x) takes the
// last phase's pending transition from state copy `from` (partial half
// `half`) and leaves the state in L.st, current between replays.
template <class T, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    pdlp_settle_kernel(const LpDev<T>* __restrict__ lps, int W, int from, int half,
                       const StepCtl* __restrict__ ctl, unsigned* done) {
  __shared__ double red[(BLOCK / kWave + 2) * kMaxRed];
  __shared__ LpState ss;
  const LpDev<T>& L = lps[blockIdx.x];
  if (threadIdx.x == 0) ss = from ? *L.st2 : *L.st;
  __syncthreads();
  if (!uniform(ss.pending)) {
    if (from && threadIdx.x == 0) *L.st = ss;
    return;
  }
  const int ph = uniform(ss.phase);
  const uint32_t mm = phase_maxmask(ph);
  const double* pin = L.partials + (size_t)half * W * kMaxRed;
  if (ph == PH_CHECK) sum_partials<BLOCK, kMaxRed>(pin, W, mm, red);
  else if (phase_reduces(ph, uniform(ss.next_mode))) sum_partials<BLOCK, 6>(pin, W, mm, red);
  if (threadIdx.x == 0) {
    LpState s = ss;
    s.pending = 0;
    take_transition(L, s, ph, red + (BLOCK / kWave) * kMaxRed, ctl, done, ss);
  }
}

// Puts every LP of a batch back at the start of its solve (a kernel rather
// than a memset, so it can sit in a replayed graph).
template <class T>
__global__ void state_reset_kernel(const LpDev<T>* __restrict__ lps, int count, unsigned* done) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *done = 0u;
  if (i >= count) return;
  LpState s{};
  s.phase = PH_INIT;
  s.status = ST_RUNNING;
  *lps[i].st = s;
  for (int k = 0; k < kPhaseSlots; ++k) {
    lps[i].pcount[k] = 0;
    lps[i].pticks[k] = 0;
  }
}

}  // namespace sov
