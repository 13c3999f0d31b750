// Time-step runtime: the monodomain time loop in native code, pipelined.
//
// Per step the host enqueues (all on one stream, nothing waits):
//   1. the model's fused ionic kernel (states, w, stimulus, LAT/RT of the
//      previous step's solution),
//   2. the fused RHS + PCG start pass,
//   3. K speculative PCG iterations (spmv + update passes), where K is the
//      largest iteration count among the last steps,
//   4. a one-block stopping check that stores the step status and, when the
//      solve is not finished, raises ctl->stall,
//   5. an optional snapshot copy: into a device frame buffer, or into one
//      slot of a small device ring that a copy stream drains into pinned
//      host memory while the next steps run (tc_sim_advance_ring).
// The check kernel writes the step status straight into pinned host memory
// (two 8-byte words, each tagged with the record's sequence number), so no
// per-step event, copy or fence is needed; the host polls that record.
// The host keeps up to `depth` steps in flight and validates the oldest.
// While ctl->stall is set every later kernel returns at once, so a step that
// needed more than K iterations leaves the device exactly where the plain
// loop would be; the host then launches the remaining iterations, clears the
// flag and re-issues the later steps from that state.
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include <atomic>

#include "pcg_kernels.h"

namespace tc {
namespace {

constexpr int kRing = 64;
constexpr int kTimers = 5;  // ionic, rhs+start, spmv, update, check
constexpr int kHistoryDefault = 4, kHistoryMax = 256;
constexpr int kMaxRingSlots = 16;

struct HostState {
  int64_t steps = 0;
  double t = 0.0;
  int xc = 0;  // desc.x[xc] holds the current iterate
  int pc = 0;  // desc.xp[pc] holds x_prev
  int64_t frames = 0;
  bool lat_pending = false;
  double lat_t = 0.0;
};

struct Plan {
  int64_t step;
  int slot;
  int K;
  long long seq;  // sequence number the check pass writes for this step
  bool timed;
  bool fused;     // the stopping test ran in the last update pass
  int evi;        // event set of a timed step (index into tc_sim::ev_sets)
  int64_t frame;  // frame written after this step or -1
  HostState after;
};

}  // namespace
}  // namespace tc

struct tc_sim {
  tc_sim_desc d;
  union {
    tc_ms2v_params ms2v;
    tc_ttp_params ttp;
    tc_crn_params crn;
  } params;
  tc::StatusRec* host_status;   // pinned, written by the device
  tc::StatusRec* dev_view;      // device address of host_status
  tc::StepStatus status[tc::kRing];  // records as read by wait_status
  long long seq;
  tc::HostState vs;  // validated
  std::deque<int> recent;
  int timing;
  // Event sets of timed steps: a step takes a free set at enqueue; when its
  // status is consumed the set waits in ev_done, with the iterations it
  // timed, until the end of the advance call, where the elapsed times are
  // read (the events are complete by then, so no query waits and the host
  // loop that keeps the pipeline fed does no timing work).
  std::vector<std::vector<hipEvent_t>> ev_sets;
  std::vector<int> ev_free;
  struct EvDone {
    int evi, kk, fused, K;
  };
  std::vector<EvDone> ev_done;
  // the stopping test of a step runs in its last update pass (1, default)
  // or in a one-block check pass (0)
  int fuse_check;
  int history;  // steps whose iteration counts set K
  double acc_ms[tc::kTimers];
  int64_t acc_n[tc::kTimers];
  // frame buffer of the current advance call (host memory in ring mode)
  void* frames;
  int64_t tpf, max_frames;
  // ring mode: frame k goes to device slot k % ring_slots, then by DMA on
  // the copy stream cs to frames + k*n.  written[j] orders the copy after
  // the slot's write; copied[j] orders the slot's next write after the copy.
  void* ring;
  int ring_slots;
  hipStream_t cs;
  hipEvent_t written[tc::kMaxRingSlots], copied[tc::kMaxRingSlots];
  bool slot_busy[tc::kMaxRingSlots];
  tc::GraphCache graphs;  // graph mode: each step's fixed RHS + PCG launches
};

namespace tc {
namespace {

uint32_t active_bits(const tc_sim_desc& d, double t) { return stimulus_bits(d.stim, d.n_stim, t); }

int ev_take(tc_sim* S, size_t count, int* evi) {
  if (S->ev_free.empty()) {
    S->ev_sets.emplace_back();
    S->ev_free.push_back((int)S->ev_sets.size() - 1);
  }
  const int slot = S->ev_free.back();
  S->ev_free.pop_back();
  *evi = slot;
  while (S->ev_sets[slot].size() < count) {
    hipEvent_t e;
    // timing only: no system-scope fence (no cache write-back per record)
    TC_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    S->ev_sets[slot].push_back(e);
  }
  return TC_OK;
}

int ionic_launch(tc_sim* S, const tc_ionic_io* io, hipStream_t s, hipEvent_t e0,
                 hipEvent_t e1) {
  const tc_sim_desc& d = S->d;
  switch (d.model) {
    case TC_MODEL_MS2V:
      return ms2v_step(&S->params.ms2v, d.states, io, d.dtype, s, e0, e1);
    case TC_MODEL_TTP:
      return ttp_step(&S->params.ttp, d.states, d.tab[0], d.tab[1], d.tab[2], io, d.dtype, s, e0,
                      e1);
    case TC_MODEL_CRN:
      return crn_step(&S->params.crn, d.states, d.tab[0], d.tab[1], d.tab[2], io, d.dtype, s, e0,
                      e1);
  }
  set_error("unknown model %d", d.model);
  return TC_ERR_ARG;
}

template <class T, class CI>
void launch_spmv(tc_pcg* P, const CI* sc, T* pin, T* pout, int k, hipStream_t s, hipEvent_t e0,
                 hipEvent_t e1) {
  launch(k_pcg_spmv<T, CI>, dim3(P->G), dim3(kPcgBlock), s, e0, e1, P->n, hot_slices(P->A), sc,
         P->A.pat,
         (const T*)P->Av, (const T*)P->Dinv, (const T*)P->r, pin, pout, (T*)P->Ap, k, P->ctl,
         (const double*)P->part[0], (const double*)P->part[2], P->part[1]);
}

// The plan's stage applies: built for its block count, on an operator
// without shared slices.
inline bool staged(const tc_pcg* P) {
  return P->stage.cols && P->stage.blocks == P->G && !P->A.slice_tag;
}

// The spmv pass of iteration k: staged gathers when the plan has a stage,
// else int16 offsets or int32 columns.
template <class T>
int launch_spmv_at(tc_pcg* P, int k, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  T* pp[2] = {(T*)P->p[0], (T*)P->p[1]};
  T* pin = pp[k & 1];
  T* pout = pp[(k + 1) & 1];
  if (staged(P)) {
    const tc_stage& st = P->stage;
    launch_lds(st.chunk ? k_pcg_spmv_staged<T, true> : k_pcg_spmv_staged<T, false>,
               dim3(P->G), dim3(kPcgBlock),
               (size_t)st.max_cols * sizeof(T), s, e0, e1, P->n, P->A.slice_ptr, st.lidx,
               st.col_ptr, st.own, st.cols, st.chunk, (const T*)P->Av, (const T*)P->Dinv,
               (const T*)P->r,
               (const T*)pin, pout, (T*)P->Ap, k, P->ctl, (const double*)P->part[0],
               (const double*)P->part[2], P->part[1]);
  } else if (P->A.dcol) {
    launch_spmv<T>(P, P->A.dcol, pin, pout, k, s, e0, e1);
  } else {
    launch_spmv<T>(P, P->A.col, pin, pout, k, s, e0, e1);
  }
  TC_LAUNCH();
  return TC_OK;
}

// The update pass of iteration k with the stopping test, publishing record
// seq in status slot `slot`.
template <class T, class XT>
int launch_check_update(tc_sim* S, int k, XT* xs, int slot, long long seq, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1) {
  tc_pcg* P = S->d.pcg;
  T* pout = (T*)P->p[(k + 1) & 1];
  launch(k_pcg_update<T, XT, true>, dim3(P->G), dim3(kPcgBlock), s, e0, e1, P->n,
         (const T*)P->Dinv, (const T*)pout, (const T*)P->Ap, xs, (T*)P->r, k, P->ctl,
         (const double*)P->part[1], P->part[2], S->dev_view + slot, 1, S->d.max_iter, seq);
  TC_LAUNCH();
  return TC_OK;
}

// Iterations k0 .. k1-1 (spmv and update passes).  With check_slot >= 0
// the last update also runs the stopping test and publishes record seq in
// that status slot (no separate check pass); its events are then
// (*ev)[4], (*ev)[5], the check pass's.
template <class T, class XT>
int launch_iters(tc_sim* S, int k0, int k1, XT* xs, hipStream_t s, std::vector<hipEvent_t>* ev,
                 size_t ev_base, int check_slot = -1, long long seq = 0) {
  tc_pcg* P = S->d.pcg;
  const dim3 g(P->G), blk(kPcgBlock);
  T* pp[2] = {(T*)P->p[0], (T*)P->p[1]};
  for (int k = k0; k < k1; ++k) {
    T* pout = pp[(k + 1) & 1];
    const size_t e = ev_base + 4 * (size_t)(k - k0);
    hipEvent_t ev0 = ev ? (*ev)[e] : nullptr, ev1 = ev ? (*ev)[e + 1] : nullptr;
    int rc = launch_spmv_at<T>(P, k, s, ev0, ev1);
    if (rc) return rc;
    if (check_slot >= 0 && k == k1 - 1) {
      launch(k_pcg_update<T, XT, true>, g, blk, s, ev ? (*ev)[4] : nullptr,
             ev ? (*ev)[5] : nullptr, P->n, (const T*)P->Dinv, (const T*)pout, (const T*)P->Ap,
             xs, (T*)P->r, k, P->ctl, (const double*)P->part[1], P->part[2],
             S->dev_view + check_slot, 1, S->d.max_iter, seq);
    } else {
      launch(k_pcg_update<T, XT>, g, blk, s, ev ? (*ev)[e + 2] : nullptr,
             ev ? (*ev)[e + 3] : nullptr, P->n, (const T*)P->Dinv, (const T*)pout,
             (const T*)P->Ap, xs, (T*)P->r, k, P->ctl, (const double*)P->part[1], P->part[2],
             (StatusRec*)nullptr, 0, 0, 0LL);
    }
    TC_LAUNCH();
  }
  return TC_OK;
}

int check_launch(tc_sim* S, int k, int slot, long long seq, hipStream_t s,
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  tc_pcg* P = S->d.pcg;
  launch(k_pcg_check, dim3(1), dim3(kPcgBlock), s, e0, e1, k, P->G, P->ctl,
         (const double*)P->part[2], S->dev_view + slot, 1, S->d.max_iter, seq);
  TC_LAUNCH();
  return TC_OK;
}

// Wait until the check pass has published record `seq` in slot `slot`.
void accumulate_timing(tc_sim* S, const std::vector<hipEvent_t>& ev, int kk, int fused, int K);

// Reads the elapsed times of one consumed timed step whose events are
// complete (false when there is none), so that the host does its timing
// work while it waits for the device anyway.
bool drain_one_timing(tc_sim* S) {
  if (S->ev_done.empty()) return false;
  const tc_sim::EvDone dn = S->ev_done.back();
  const std::vector<hipEvent_t>& ev = S->ev_sets[dn.evi];
  if (hipEventQuery(ev[5]) != hipSuccess) return false;
  accumulate_timing(S, ev, dn.kk, dn.fused, dn.K);
  S->ev_free.push_back(dn.evi);
  S->ev_done.pop_back();
  return true;
}

int wait_status(tc_sim* S, int slot, long long seq) {
  const auto t0 = std::chrono::steady_clock::now();
  for (long spins = 0; !status_take(S->host_status + slot, seq, S->status + slot); ++spins) {
    if (drain_one_timing(S)) continue;
    if ((spins & 1023) == 1023) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) {
        set_error("time-step runtime: no status for record %lld after 120 s", seq);
        return TC_ERR_STATE;
      }
      std::this_thread::yield();
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return TC_OK;
}

template <class T, class XT>
int copy_frame(tc_sim* S, const void* src, int64_t frame, const int* gate, hipStream_t s) {
  const int64_t n = S->d.n;
  const dim3 g((int)ceil_div(n, 256)), b(256);
  if (!S->ring) {
    T* dst = (T*)S->frames + frame * n;
    hipLaunchKernelGGL((k_copy_gated<T, XT>), g, b, 0, s, gate, (const XT*)src, dst, n);
    TC_LAUNCH();
    return TC_OK;
  }
  // A gated copy (its step stalled) leaves the slot as it was and the DMA
  // below then carries stale values; the stalled step's redo and the
  // re-issued later steps copy the same frames again, later on the same copy
  // stream, so the host frame ends up with the right values.
  const int j = (int)(frame % S->ring_slots);
  if (S->slot_busy[j]) TC_HIP(hipStreamWaitEvent(s, S->copied[j], 0));
  T* slot = (T*)S->ring + (int64_t)j * n;
  hipLaunchKernelGGL((k_copy_gated<T, XT>), g, b, 0, s, gate, (const XT*)src, slot, n);
  TC_LAUNCH();
  TC_HIP(hipEventRecord(S->written[j], s));
  TC_HIP(hipStreamWaitEvent(S->cs, S->written[j], 0));
  TC_HIP(hipMemcpyAsync((T*)S->frames + frame * n, slot, (size_t)n * sizeof(T),
                        hipMemcpyDeviceToHost, S->cs));
  TC_HIP(hipEventRecord(S->copied[j], S->cs));
  S->slot_busy[j] = true;
  return TC_OK;
}

int desired_K(tc_sim* S) {
  int K = 1;
  for (int v : S->recent) K = v > K ? v : K;
  if (K > S->d.max_iter) K = S->d.max_iter;
  return K;
}

// Staged gathers in the RHS + PCG start pass: the plan's stage applies and
// the set's x, w and x_prev fit the LDS of one block (gfx950: 160 KB a
// workgroup, the whole CU's LDS).
constexpr size_t kStageLdsMax = 163840;

template <class T, class XT, class WT, bool LG>
void rhs_init_staged_launch(tc_sim* S, int xc, int pc, double coef, double tdt, hipStream_t s,
                            hipEvent_t e0, hipEvent_t e1) {
  const tc_sim_desc& d = S->d;
  tc_pcg* P = d.pcg;
  const tc_stage& st = P->stage;
  // blocks of at most 16 slices run as 16 waves, one slice each (handing
  // their rows' factors over takes 3 doubles per lane of 8 waves, 12 KB,
  // of the set's LDS)
  const int64_t ns = (d.n + 63) / 64;
  // 16 waves when a block's range holds at most 16 slices, or its units
  // 16 slices (units of 8 slices: one slice per wave of 8)
  const bool wide = st.chunk ? st.chunk == 2 * kWaves : (ns + P->G - 1) / P->G <= 2 * kWaves;
  constexpr size_t kHandOver = 3 * kWaves * 64 * sizeof(double);
  size_t lds = rhs_stage_bytes<T, XT, WT, LG>(st.max_cols);
  if (wide && lds < kHandOver) lds = kHandOver;
  auto kern = st.chunk ? (wide ? k_rhs_init_staged<T, XT, WT, LG, 2 * kPcgBlock, true>
                               : k_rhs_init_staged<T, XT, WT, LG, kPcgBlock, true>)
                       : (wide ? k_rhs_init_staged<T, XT, WT, LG, 2 * kPcgBlock, false>
                               : k_rhs_init_staged<T, XT, WT, LG, kPcgBlock, false>);
  launch_lds(kern, dim3(P->G), dim3(wide ? 2 * kPcgBlock : kPcgBlock), lds, s, e0, e1, d.n,
             P->A.slice_ptr, st.lidx, st.col_ptr, st.own, st.cols, st.chunk, (const T*)d.Mv,
             (const T*)d.Kv, (const T*)P->Dinv, (const WT*)d.w, (const XT*)d.x[xc],
             LG ? (const T*)d.xp[pc] : (const T*)nullptr, LG ? (XT*)d.x[1 - xc] : (XT*)nullptr,
             LG ? (T*)d.xp[1 - pc] : (T*)nullptr, (T*)P->r, coef, d.Cm, tdt, P->ctl, d.a_tol,
             d.r_tol, P->part[0]);
}

template <class T, class XT, class WT, bool LG, class CI>
void rhs_init_launch(tc_sim* S, const CI* sc, int xc, int pc, double coef, double tdt, dim3 g,
                     dim3 blk, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const tc_sim_desc& d = S->d;
  tc_pcg* P = d.pcg;
  if (staged(P) && rhs_stage_bytes<T, XT, WT, LG>(P->stage.max_cols) <= kStageLdsMax) {
    rhs_init_staged_launch<T, XT, WT, LG>(S, xc, pc, coef, tdt, s, e0, e1);
    return;
  }
  launch(k_rhs_init<T, XT, WT, LG, CI>, g, blk, s, e0, e1, d.n, hot_slices(P->A), sc, P->A.pat, (const T*)d.Mv, (const T*)d.Kv, (const T*)P->Dinv, (const WT*)d.w,
         (const XT*)d.x[xc], LG ? (const T*)d.xp[pc] : (const T*)nullptr,
         LG ? (XT*)d.x[1 - xc] : (XT*)nullptr, LG ? (T*)d.xp[1 - pc] : (T*)nullptr, (T*)P->r,
         coef, d.Cm, tdt, P->ctl, d.a_tol, d.r_tol, P->part[0]);
}

// The fused RHS + PCG start pass of a step whose solution is in x[xc]
// (x_prev in xp[pc]).
template <class T, class XT, class WT>
int rhs_init_seq(tc_sim* S, int xc, int pc, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const tc_sim_desc& d = S->d;
  tc_pcg* P = d.pcg;
  const dim3 g(P->G), blk(kPcgBlock);
  const double coef = (1.0 - d.theta) * d.dt;
  const double tdt = d.theta * d.dt;
  if (d.linear_guess) {
    if (P->A.dcol)
      rhs_init_launch<T, XT, WT, true>(S, P->A.dcol, xc, pc, coef, tdt, g, blk, s, e0, e1);
    else
      rhs_init_launch<T, XT, WT, true>(S, P->A.col, xc, pc, coef, tdt, g, blk, s, e0, e1);
  } else {
    if (P->A.dcol)
      rhs_init_launch<T, XT, WT, false>(S, P->A.dcol, xc, pc, coef, tdt, g, blk, s, e0, e1);
    else
      rhs_init_launch<T, XT, WT, false>(S, P->A.col, xc, pc, coef, tdt, g, blk, s, e0, e1);
  }
  TC_LAUNCH();
  return TC_OK;
}

// Enqueue the step after planned state ps; ps becomes the state after it.
template <class T, class XT, class WT>
int enqueue(tc_sim* S, HostState& ps, Plan& plan, hipStream_t s) {
  const tc_sim_desc& d = S->d;
  tc_pcg* P = d.pcg;
  const int64_t step = ps.steps + 1;
  const double t = ps.t + d.dt;
  const int slot = (int)(step % kRing);
  const int K = desired_K(S);
  const int xc = ps.xc, pc = ps.pc;
  std::vector<hipEvent_t>* ev = nullptr;
  plan.timed = S->timing > 0 && step % S->timing == 0;
  plan.evi = -1;
  if (plan.timed) {
    int rc = ev_take(S, 6 + 4 * (size_t)K, &plan.evi);
    if (rc) return rc;
    ev = &S->ev_sets[plan.evi];
  }
  // 1. ionic
  tc_ionic_io io{};
  io.n = d.n;
  io.dt = d.dt_ionic;
  io.dt_rhs = d.dt;
  io.Cm = d.Cm;
  io.V = d.x[xc];
  io.v_f64 = (sizeof(XT) == 8 && sizeof(T) == 4) ? 1 : 0;
  io.w_f64 = (sizeof(WT) == 8 && sizeof(T) == 4) ? 1 : 0;
  io.w = d.w;
  io.dV = nullptr;
  const uint32_t bits = active_bits(d, t);
  io.stim_mask = bits ? d.stim_mask : nullptr;
  io.stim_active = bits;
  for (int k = 0; k < d.n_stim; ++k) io.stim_amp[k] = d.stim[k].amp;
  if (d.AT && ps.lat_pending) {
    io.AT = d.AT;
    io.RT = d.RT;
    io.lat_t = ps.lat_t;
  }
  io.gate = &P->ctl->stall;
  io.stream_states = 1;
  io.state_stride = d.state_stride;
  io.state_block = d.state_block;
  int rc = ionic_launch(S, &io, s, ev ? (*ev)[0] : nullptr, ev ? (*ev)[1] : nullptr);
  if (rc) return rc;
  // 2. RHS + PCG start; 3. speculative iterations; 4. the stopping test,
  // in the last update pass (or a check pass of its own when no iteration
  // is launched or fusing is off)
  XT* xs = (XT*)d.x[d.linear_guess ? 1 - xc : xc];
  plan.seq = ++S->seq;
  plan.fused = S->fuse_check && K > 0;
  if (d.graph && !ev) {
    // the launches fixed by the buffer parity and K (all but the last
    // update when it runs the test, which takes this step's record) replayed
    // as one graph
    const int kg = plan.fused ? K - 1 : K;
    // K < 2^30 (the status word's iteration field); the plan's block count
    // and version (bumped by tc_pcg_set_blocks / tc_pcg_set_stage) in key2
    const long long key = (long long)((xc << 1) | pc) | ((long long)K << 2);
    const long long key2 = (long long)P->G | ((long long)P->version << 16);
    rc = graph_replay(S->graphs, key, key2, s, [&](hipStream_t cs) {
      int r = rhs_init_seq<T, XT, WT>(S, xc, pc, cs, nullptr, nullptr);
      if (r == TC_OK) r = launch_iters<T, XT>(S, 0, kg, xs, cs, nullptr, 0);
      if (r == TC_OK && plan.fused) r = launch_spmv_at<T>(P, K - 1, cs, nullptr, nullptr);
      return r;
    });
    if (rc) return rc;
    if (plan.fused) {
      rc = launch_check_update<T, XT>(S, K - 1, xs, slot, plan.seq, s, nullptr, nullptr);
      if (rc) return rc;
    }
  } else {
    rc = rhs_init_seq<T, XT, WT>(S, xc, pc, s, ev ? (*ev)[2] : nullptr, ev ? (*ev)[3] : nullptr);
    if (rc) return rc;
    rc = launch_iters<T, XT>(S, 0, K, xs, s, ev, 6, plan.fused ? slot : -1, plan.seq);
    if (rc) return rc;
  }
  if (!plan.fused) {
    rc = check_launch(S, K - 1, slot, plan.seq, s, ev ? (*ev)[4] : nullptr,
                      ev ? (*ev)[5] : nullptr);
    if (rc) return rc;
  }
  // state after the step
  HostState nx = ps;
  nx.steps = step;
  nx.t = t;
  if (d.linear_guess) {
    nx.xc = 1 - xc;
    nx.pc = 1 - pc;
  }
  nx.lat_pending = true;
  nx.lat_t = t;
  // 5. snapshot
  plan.frame = -1;
  if (S->frames && S->tpf > 0 && step % S->tpf == 0 && ps.frames < S->max_frames) {
    rc = copy_frame<T, XT>(S, d.x[nx.xc], ps.frames, &P->ctl->stall, s);
    if (rc) return rc;
    plan.frame = ps.frames;
    nx.frames = ps.frames + 1;
  }
  plan.step = step;
  plan.slot = slot;
  plan.K = K;
  plan.after = nx;
  ps = nx;
  return TC_OK;
}

// The step of `plan` stalled: run its remaining iterations until the stop
// rule or max_iter, then redo its snapshot copy (which the stall gated).
template <class T, class XT>
int resolve_stall(tc_sim* S, const Plan& plan, hipStream_t s) {
  const tc_sim_desc& d = S->d;
  tc_pcg* P = d.pcg;
  StepStatus* st = S->status + plan.slot;
  XT* xs = (XT*)d.x[plan.after.xc];
  int it = st->iter;
  int chunk = it > 0 ? it : 1;
  while (st->stalled) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, &P->ctl->stall, 0);
    TC_LAUNCH();
    const int c = chunk < d.max_iter - it ? chunk : d.max_iter - it;
    int rc = launch_iters<T, XT>(S, it, it + c, xs, s, nullptr, 0);
    if (rc) return rc;
    it += c;
    const long long seq = ++S->seq;
    rc = check_launch(S, it - 1, plan.slot, seq, s);
    if (rc) return rc;
    rc = wait_status(S, plan.slot, seq);
    if (rc) return rc;
    chunk = chunk * 2 < 64 ? chunk * 2 : 64;
  }
  if (plan.frame >= 0) {
    int rc = copy_frame<T, XT>(S, xs, plan.frame, nullptr, s);
    if (rc) return rc;
  }
  return TC_OK;
}

// Elapsed times of one timed step: ionic, rhs_init and check, and the
// first kk of its K speculative iterations (those that ran).  With the
// check fused into the last update pass, events 4-5 bracket that pass: it
// counts as an update when all K iterations ran, else as the check (the
// solve had stopped and the pass only ran the test).
void accumulate_timing(tc_sim* S, const std::vector<hipEvent_t>& ev, int kk, int fused, int K) {
  (void)hipEventSynchronize(ev[5]);
  float ms = 0.f;
  const bool last_fused = fused && kk >= K;
  const int pairs[3][2] = {{0, 1}, {2, 3}, {4, 5}};
  const int which[3] = {0, 1, last_fused ? 3 : 4};
  for (int k = 0; k < 3; ++k) {
    if (hipEventElapsedTime(&ms, ev[pairs[k][0]], ev[pairs[k][1]]) == hipSuccess) {
      S->acc_ms[which[k]] += ms;
      S->acc_n[which[k]] += 1;
    }
  }
  for (int k = 0; k < kk; ++k) {
    const size_t e = 6 + 4 * (size_t)k;
    if (hipEventElapsedTime(&ms, ev[e], ev[e + 1]) == hipSuccess) {
      S->acc_ms[2] += ms;
      S->acc_n[2] += 1;
    }
    if (last_fused && k == K - 1) continue;  // timed as the 4-5 pair
    if (hipEventElapsedTime(&ms, ev[e + 2], ev[e + 3]) == hipSuccess) {
      S->acc_ms[3] += ms;
      S->acc_n[3] += 1;
    }
  }
}

template <class T, class XT, class WT>
int advance(tc_sim* S, int64_t nsteps, int32_t* iters, hipStream_t s) {
  const int depth = S->d.depth < 1 ? 1 : (S->d.depth > kRing / 2 ? kRing / 2 : S->d.depth);
  std::deque<Plan> inflight;
  HostState ps = S->vs;
  int64_t enq = 0, done = 0;
  while (done < nsteps) {
    while ((int)inflight.size() < depth && enq < nsteps) {
      Plan plan;
      int rc = enqueue<T, XT, WT>(S, ps, plan, s);
      if (rc) return rc;
      inflight.push_back(plan);
      ++enq;
    }
    Plan plan = inflight.front();
    int rc = wait_status(S, plan.slot, plan.seq);
    if (rc) return rc;
    StepStatus* st = S->status + plan.slot;
    if (plan.timed)
      S->ev_done.push_back(
          {plan.evi, st->iter < plan.K ? st->iter : plan.K, plan.fused ? 1 : 0, plan.K});
    if (st->stalled) {
      int rc = resolve_stall<T, XT>(S, plan, s);
      if (rc) return rc;
      enq -= (int64_t)inflight.size() - 1;
      // the later steps are re-issued: their event sets are free again
      for (size_t j = 1; j < inflight.size(); ++j)
        if (inflight[j].timed) S->ev_free.push_back(inflight[j].evi);
      inflight.clear();
      ps = plan.after;
    } else {
      inflight.pop_front();
    }
    iters[done] = st->n_iter;
    S->recent.push_back(st->iter > 0 ? st->iter : 1);
    if ((int)S->recent.size() > S->history) S->recent.pop_front();
    S->vs = plan.after;
    ++done;
  }
  TC_HIP(hipStreamSynchronize(s));
  for (const auto& dn : S->ev_done) {
    accumulate_timing(S, S->ev_sets[dn.evi], dn.kk, dn.fused, dn.K);
    S->ev_free.push_back(dn.evi);
  }
  S->ev_done.clear();
  if (S->ring) {
    TC_HIP(hipStreamSynchronize(S->cs));
    for (int j = 0; j < kMaxRingSlots; ++j) S->slot_busy[j] = false;
  }
  return TC_OK;
}

// Copy stream and events of ring mode, made on first use.
int ring_setup(tc_sim* S) {
  if (S->cs) return TC_OK;
  TC_HIP(hipStreamCreateWithFlags(&S->cs, hipStreamNonBlocking));
  for (int j = 0; j < kMaxRingSlots; ++j) {
    TC_HIP(hipEventCreateWithFlags(&S->written[j], hipEventDisableTiming));
    TC_HIP(hipEventCreateWithFlags(&S->copied[j], hipEventDisableTiming));
    S->slot_busy[j] = false;
  }
  return TC_OK;
}

int run(tc_sim* S, int64_t nsteps, int32_t* iters, hipStream_t s) {
  if (S->d.dtype == TC_F64) return advance<double, double, double>(S, nsteps, iters, s);
  if (S->d.x_f64 && S->d.w_f64) return advance<float, double, double>(S, nsteps, iters, s);
  if (S->d.x_f64) return advance<float, double, float>(S, nsteps, iters, s);
  return advance<float, float, float>(S, nsteps, iters, s);
}

}  // namespace
}  // namespace tc

using namespace tc;

extern "C" int tc_sim_create(tc_sim** out, const tc_sim_desc* d) {
  TC_CHECK(out && d && d->pcg && d->states && d->Mv && d->Kv && d->w && d->n > 0, TC_ERR_ARG,
           "tc_sim_create: bad arguments");
  TC_CHECK(d->dtype == TC_F32 || d->dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", d->dtype);
  TC_CHECK(d->x[0] && d->x[1] && (!d->linear_guess || (d->xp[0] && d->xp[1])), TC_ERR_ARG,
           "tc_sim_create: need two x buffers (and two x_prev buffers with the linear guess)");
  TC_CHECK(d->n_stim >= 0 && d->n_stim <= TC_MAX_STIMULI, TC_ERR_ARG, "bad stimulus count");
  TC_CHECK(d->n_stim == 0 || d->stim_mask, TC_ERR_ARG, "stimuli need a mask");
  TC_CHECK(d->state_block == 0 || d->state_block == 1, TC_ERR_ARG, "state_block %d not 0 or 1",
           d->state_block);
  TC_CHECK(d->state_block || d->state_stride == 0 || d->state_stride >= d->n, TC_ERR_ARG,
           "tc_sim_create: state stride below n");
  TC_CHECK(d->k_history >= 0 && d->k_history <= kHistoryMax, TC_ERR_ARG,
           "k_history %d outside 0..%d", d->k_history, kHistoryMax);
  TC_CHECK(d->max_iter >= 0 && d->max_iter <= kMaxIterLimit, TC_ERR_ARG,
           "max_iter %d outside 0..%d", d->max_iter, kMaxIterLimit);
  TC_CHECK(!d->w_f64 || d->x_f64 || d->dtype == TC_F64, TC_ERR_ARG, "w_f64 needs x_f64");
  TC_CHECK(!d->AT || d->RT, TC_ERR_ARG, "AT without RT");
  TC_CHECK(d->pcg->n == d->n && d->pcg->dtype == d->dtype, TC_ERR_ARG,
           "tc_sim_create: PCG plan does not match n/dtype");
  tc_sim* S = new tc_sim();
  S->d = *d;
  switch (d->model) {
    case TC_MODEL_MS2V: S->params.ms2v = *(const tc_ms2v_params*)d->params; break;
    case TC_MODEL_TTP: S->params.ttp = *(const tc_ttp_params*)d->params; break;
    case TC_MODEL_CRN: S->params.crn = *(const tc_crn_params*)d->params; break;
    default:
      delete S;
      set_error("unknown model %d", d->model);
      return TC_ERR_ARG;
  }
  S->d.params = nullptr;
  hipError_t e = hipHostMalloc((void**)&S->host_status, sizeof(StatusRec) * kRing,
                               hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&S->dev_view, S->host_status, 0);
  if (e != hipSuccess) {
    set_error("tc_sim_create: %s", hipGetErrorString(e));
    tc_sim_destroy(S);
    return TC_ERR_HIP;
  }
  status_clear(S->host_status, kRing);
  S->seq = 0;
  // The stall flag lives in the shared PCG control block; start clear.
  e = hipMemset(&d->pcg->ctl->stall, 0, sizeof(int));
  if (e != hipSuccess) {
    set_error("tc_sim_create: %s", hipGetErrorString(e));
    return TC_ERR_HIP;
  }
  S->timing = 0;
  S->fuse_check = d->separate_check ? 0 : 1;
  S->history = d->k_history > 0 ? d->k_history : kHistoryDefault;
  for (int k = 0; k < kTimers; ++k) {
    S->acc_ms[k] = 0.0;
    S->acc_n[k] = 0;
  }
  *out = S;
  return TC_OK;
}

extern "C" int tc_sim_destroy(tc_sim* S) {
  if (!S) return TC_OK;
  S->graphs.release();
  for (auto& set : S->ev_sets)
    for (hipEvent_t e : set) (void)hipEventDestroy(e);
  if (S->cs) {
    (void)hipStreamSynchronize(S->cs);
    for (int j = 0; j < kMaxRingSlots; ++j) {
      if (S->written[j]) (void)hipEventDestroy(S->written[j]);
      if (S->copied[j]) (void)hipEventDestroy(S->copied[j]);
    }
    (void)hipStreamDestroy(S->cs);
  }
  if (S->host_status) (void)hipHostFree(S->host_status);
  delete S;
  return TC_OK;
}

extern "C" int tc_sim_advance(tc_sim* S, int64_t nsteps, int32_t* iters, void* frames,
                              int64_t tpf, int64_t max_frames, void* stream) {
  TC_CHECK(S && iters && nsteps >= 0, TC_ERR_ARG, "tc_sim_advance: bad arguments");
  S->frames = frames;
  S->ring = nullptr;
  S->tpf = tpf;
  S->max_frames = frames ? max_frames : 0;
  return run(S, nsteps, iters, (hipStream_t)stream);
}

extern "C" int tc_sim_advance_ring(tc_sim* S, int64_t nsteps, int32_t* iters, void* frames_host,
                                   void* ring, int ring_slots, int64_t tpf, int64_t max_frames,
                                   void* stream) {
  TC_CHECK(S && iters && nsteps >= 0 && frames_host && ring, TC_ERR_ARG,
           "tc_sim_advance_ring: bad arguments");
  TC_CHECK(ring_slots >= 1 && ring_slots <= kMaxRingSlots, TC_ERR_ARG,
           "tc_sim_advance_ring: ring_slots %d outside 1..%d", ring_slots, kMaxRingSlots);
  int rc = ring_setup(S);
  if (rc) return rc;
  S->frames = frames_host;
  S->ring = ring;
  S->ring_slots = ring_slots;
  S->tpf = tpf;
  S->max_frames = max_frames;
  rc = run(S, nsteps, iters, (hipStream_t)stream);
  if (rc) (void)hipStreamSynchronize(S->cs);  // no DMA left in flight into frames_host
  S->ring = nullptr;
  return rc;
}

extern "C" int tc_sim_finish(tc_sim* S, void* stream) {
  TC_CHECK(S, TC_ERR_ARG, "tc_sim_finish: null");
  if (!(S->d.AT && S->vs.lat_pending)) return TC_OK;
  const int64_t n = S->d.n;
  const dim3 g((int)ceil_div(n, 256)), b(256);
  hipStream_t s = (hipStream_t)stream;
  const void* x = S->d.x[S->vs.xc];
  if (S->d.dtype == TC_F64)
    hipLaunchKernelGGL((k_lat_x<double, double>), g, b, 0, s, (const double*)x,
                       (double*)S->d.AT, (double*)S->d.RT, n, S->vs.lat_t);
  else if (S->d.x_f64)
    hipLaunchKernelGGL((k_lat_x<float, double>), g, b, 0, s, (const double*)x,
                       (float*)S->d.AT, (float*)S->d.RT, n, S->vs.lat_t);
  else
    hipLaunchKernelGGL((k_lat_x<float, float>), g, b, 0, s, (const float*)x,
                       (float*)S->d.AT, (float*)S->d.RT, n, S->vs.lat_t);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_sim_status(tc_sim* S, double* t, int64_t* steps, int* x_index,
                             int64_t* frames) {
  TC_CHECK(S && t && steps && x_index && frames, TC_ERR_ARG, "tc_sim_status: bad arguments");
  *t = S->vs.t;
  *steps = S->vs.steps;
  *x_index = S->vs.xc;
  *frames = S->vs.frames;
  return TC_OK;
}

extern "C" int tc_sim_set_timing(tc_sim* S, int on) {
  TC_CHECK(S, TC_ERR_ARG, "tc_sim_set_timing: null");
  S->timing = on > 0 ? on : 0;
  for (int k = 0; k < kTimers; ++k) {
    S->acc_ms[k] = 0.0;
    S->acc_n[k] = 0;
  }
  return TC_OK;
}

extern "C" int tc_sim_graph_stats(const tc_sim* S, int64_t* captures, int64_t* replays) {
  TC_CHECK(S && captures && replays, TC_ERR_ARG, "tc_sim_graph_stats: bad arguments");
  *captures = S->graphs.captures;
  *replays = S->graphs.replays;
  return TC_OK;
}

extern "C" int tc_sim_get_timing(tc_sim* S, double* ms, int64_t* cnt) {
  TC_CHECK(S && ms && cnt, TC_ERR_ARG, "tc_sim_get_timing: bad arguments");
  for (int k = 0; k < kTimers; ++k) {
    ms[k] = S->acc_ms[k];
    cnt[k] = S->acc_n[k];
  }
  return TC_OK;
}
