// hz_krylov.cpp -- flexible GMRES for lockstep batches of right-hand sides,
// driven from one host thread for any number of problems.
//
// Arithmetic follows the FGMRES of FGMRES.cpp:458-661 with a flexible basis
// Z_j = L0^{-1} V_j (right preconditioning), the Arnoldi coefficients
// h_ij = V_i^H w, complex Givens rotations on the Hessenberg columns
// (in place of the reference's zgels on the whole Hessenberg matrix,
// FGMRES.cpp:563-569) and the true residual at every restart.
//
// Every Arnoldi step runs on the device: the batched preconditioner and
// operator applies, Gram-Schmidt (inverse compact-WY, or sequential), the
// normalisation v_{j+1} = w / h_{j+1,j} and the Givens update, which writes
// each member's residual estimate and a status word.  The host reads those
// words once per chunk of steps (pinned copies behind events) and decides
// which members go on; at the end of a cycle the device solves R y = g and
// adds Z y to x.  Chunks of one cycle are queued two deep, so the host's
// decisions overlap the next chunk's device work.  With opts.graph, a chunk
// is captured as one HIP graph on its second use and replayed afterwards.
//
// Nothing is allocated inside the loop: the basis, the device scalars, the
// reduction partials and the pinned status buffers are sized before it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "hz_internal.h"

namespace hz {

namespace {

// Device scalars of one lockstep member (double2 slots), in this order:
// H (ldh x m, column-major), y (m), r = V^H w (m + 1), l = V^H V_j (m + 1),
// the ICWY factor L (ldh x ldh), Givens cosines (m) and sines (m), g (m + 1),
// the basis scales s_k (m + 1; .x, v_k = s_k U_k for the stored vectors U_k).
struct ScalarLayout {
  int ldh, offH, offY, offR, offL, offLg, offCS, offSN, offG, offS, size;
  explicit ScalarLayout(int m) {
    ldh = m + 1;
    offH = 0;
    offY = offH + ldh * m;
    offR = offY + m;
    offL = offR + m + 1;
    offLg = offL + m + 1;
    offCS = offLg + ldh * ldh;
    offSN = offCS + m;
    offG = offSN + m;
    offS = offG + m + 1;
    size = ((offS + m + 1 + 7) / 8) * 8;
  }
};

// device scalar block: [0, 64) misc | R members | norms (2R) | ||f|| (R) | est (R m)
size_t scalar_slots(int m, int R) { return 64 + (size_t)ScalarLayout(m).size * R + 3 * (size_t)R + (size_t)R * m + 64; }

// one pinned slot: est (R m) | norms (2R) | ||f|| (R)
size_t pin_slot(int m, int R) { return (size_t)R * m + 3 * (size_t)R; }
constexpr int NSLOT = 4;  // at most three waits are outstanding

constexpr int MD_GROUP = 12;   // vectors per group of the multi-dot pass (k_vec.hip)
constexpr int RED_BLOCKS = 1024;

struct Member {
  double norm_f = 0, rel = 0, res = 0;
  int jj = 0;
  bool done = false, in_cycle = false, conv = false, brk = false;
};

struct HistW {
  double* buf = nullptr;
  int cap = 0;
  int n = 0;
  void push(double v) {
    if (buf && n < cap) buf[n] = v;
    ++n;
  }
};

enum class Wait { NORMS, RESID, CHUNK, FINAL };
struct Pending {
  Wait kind;
  int slot;
  int cyc;
  int j0, K;
};

class Lockstep {
 public:
  Lockstep(hz_problem* p, int R, const double2* f, double2* x, const hz_krylov_opts& o, hz_krylov_stats* st,
           bool adjoint)
      : p_(p), R_(R), f_(f), x_(x), o_(o), st_(st), adj_(adjoint), m_(o.m), N_(p->N), L_(o.m) {
    HZ_CHECK(R >= 1 && R <= LOCKSTEP_MAX, HZ_ERR_ARG, "lockstep batch of %d (1..%d)", R, LOCKSTEP_MAX);
    ensure_krylov(p, m_, R);
    p->pin.reserve(NSLOT * pin_slot(m_, R) * sizeof(double2));
    for (auto& e : ev_) HZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    seq_ = o.ortho == HZ_ORTHO_MGS_SEQ || m_ > ICWY_MAX;
    chunk_ = o.graph ? 5 : 1;
    if (o.chunk > 0) chunk_ = std::min(o.chunk, m_);
    M_.resize(R);
    hist_ = HistW{o.hist, o.hist_cap, 0};
    for (int q = 0; q < R; ++q) memset(&st_[q], 0, sizeof(st_[q]));
  }

  ~Lockstep() {
    // nothing of this solve may still run when its graphs and events go
    (void)hipStreamSynchronize(p_->stream);
    for (auto& kv : graphs_)
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto e : ev_)
      if (e) (void)hipEventDestroy(e);
  }

  bool finished() const { return done_; }

  bool ready() const {
    if (q_.empty()) return true;
    const hipError_t e = hipEventQuery(ev_[q_.front().slot]);
    if (e == hipErrorNotReady) return false;
    HZ_HIP(e);
    return true;
  }

  void wait() const {
    if (!q_.empty()) HZ_HIP(hipEventSynchronize(ev_[q_.front().slot]));
  }

  // ||f_q|| and x = 0
  void begin() {
    t0_ = std::chrono::steady_clock::now();
    for (int q = 0; q < R_; ++q)
      launch_dot(p_, N_, f_ + q * N_, nullptr, HZ_MASK_FULL, dNf() + q, p_->stream);
    launch_fill((int64_t)R_ * N_, make_double2(0, 0), x_, p_->stream);
    const int s = take_slot();
    HZ_HIP(hipMemcpyAsync(pinNf(s), dNf(), R_ * sizeof(double2), hipMemcpyDeviceToHost, p_->stream));
    post(Wait::NORMS, s);
  }

  // Handles the oldest wait (its event has completed) and queues what follows.
  void process() {
    const Pending w = q_.front();
    q_.pop_front();
    switch (w.kind) {
      case Wait::NORMS: on_norms(w); break;
      case Wait::RESID: on_resid(w); break;
      case Wait::CHUNK: on_chunk(w); break;
      case Wait::FINAL: on_final(w); break;
    }
  }

 private:
  hz_problem* p_;
  int R_;
  const double2* f_;
  double2* x_;
  hz_krylov_opts o_;
  hz_krylov_stats* st_;
  bool adj_;
  int m_;
  int64_t N_;
  ScalarLayout L_;
  bool seq_ = false;
  int chunk_ = 1;
  std::vector<Member> M_;
  HistW hist_;
  hipEvent_t ev_[NSLOT] = {};
  int slot_next_ = 0;
  std::deque<Pending> q_;
  int cyc_ = 0;
  int next_j_ = 0;
  uint64_t mask_ = 0;
  bool done_ = false;
  std::chrono::steady_clock::time_point t0_;
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    int uses = 0;
  };
  std::map<std::pair<int, uint64_t>, GraphSlot> graphs_;

  // ---- device scalar addresses
  double2* dTop() const { return dscal(p_, 64); }
  double2* dMem(int q) const { return dTop() + (size_t)L_.size * q; }
  double2* dH(int q) const { return dMem(q) + L_.offH; }
  double2* dNrm() const { return dTop() + (size_t)L_.size * R_; }  // 2R: mask norms, full norms
  double2* dNf() const { return dNrm() + 2 * R_; }
  double2* dEst() const { return dNf() + R_; }
  double2* Vb(int j, int q) const { return p_->kV.as<double2>() + ((int64_t)j * R_ + q) * N_; }
  double2* Zb(int j, int q) const { return p_->kZ.as<double2>() + ((int64_t)j * R_ + q) * N_; }
  double2* W() const { return p_->kw.as<double2>(); }
  double2* Rv() const { return p_->kr.as<double2>(); }
  int64_t vstride() const { return (int64_t)R_ * N_; }

  // ---- pinned slots
  double2* pinBase(int s) const { return p_->pin.as<double2>() + (size_t)s * pin_slot(m_, R_); }
  double2* pinEst(int s) const { return pinBase(s); }
  double2* pinNrm(int s) const { return pinBase(s) + (size_t)R_ * m_; }
  double2* pinNf(int s) const { return pinNrm(s) + 2 * R_; }

  int take_slot() {
    HZ_CHECK(q_.size() < NSLOT - 1, HZ_ERR_STATE, "lockstep FGMRES: too many outstanding waits");
    const int s = slot_next_;
    slot_next_ = (slot_next_ + 1) % NSLOT;
    return s;
  }
  void post(Wait k, int s, int j0 = 0, int K = 0) {
    HZ_HIP(hipEventRecord(ev_[s], p_->stream));
    q_.push_back(Pending{k, s, cyc_, j0, K});
  }
  bool any_running() const {
    for (const auto& a : M_)
      if (!a.done) return true;
    return false;
  }
  bool any_in_cycle() const {
    for (const auto& a : M_)
      if (a.in_cycle) return true;
    return false;
  }
  int chunks_pending() const {
    int n = 0;
    for (const auto& w : q_)
      if (w.kind == Wait::CHUNK && w.cyc == cyc_) ++n;
    return n;
  }

  // true residual norms (mask, full) of the members (all, or those still running)
  void queue_residuals(Wait kind, bool only_running) {
    residual(p_, R_, f_, x_, Rv(), adj_);
    for (int q = 0; q < R_; ++q) {
      if (only_running && M_[q].done) continue;
      launch_dot(p_, N_, Rv() + q * N_, nullptr, o_.mask, dNrm() + q, p_->stream);
      launch_dot(p_, N_, Rv() + q * N_, nullptr, HZ_MASK_FULL, dNrm() + R_ + q, p_->stream);
    }
    const int s = take_slot();
    HZ_HIP(hipMemcpyAsync(pinNrm(s), dNrm(), 2 * R_ * sizeof(double2), hipMemcpyDeviceToHost, p_->stream));
    post(kind, s);
  }

  void cycle_begin() {
    if (cyc_ >= o_.max_restarts || !any_running()) queue_residuals(Wait::FINAL, false);
    else queue_residuals(Wait::RESID, true);
  }

  void on_norms(const Pending& w) {
    for (int q = 0; q < R_; ++q) {
      M_[q].norm_f = pinNf(w.slot)[q].x;
      if (M_[q].norm_f == 0.0) M_[q].done = M_[q].conv = true;
    }
    cycle_begin();
  }

  void on_resid(const Pending& w) {
    const double2* h = pinNrm(w.slot);
    mask_ = 0;
    for (int q = 0; q < R_; ++q) {
      Member& a = M_[q];
      if (a.done) continue;
      a.res = h[q].x;
      a.rel = a.res / a.norm_f;
      if (!std::isfinite(a.rel)) {
        set_error("non-finite residual in FGMRES (right-hand side %d)", q);
        throw Error{HZ_ERR_NONFINITE};
      }
      if (a.rel < o_.tol) {
        a.conv = a.done = true;
        continue;
      }
      st_[q].restarts++;
      launch_scale_inv_dev(N_, dNrm() + R_ + q, Rv() + q * N_, Vb(0, q), p_->stream);
      a.jj = 0;
      a.in_cycle = true;
      mask_ |= 1ull << q;
    }
    if (!mask_) {
      queue_residuals(Wait::FINAL, false);
      return;
    }
    next_j_ = 0;
    queue_chunks();
  }

  // keep two chunks of this cycle in flight
  void queue_chunks() {
    while (next_j_ < m_ && chunks_pending() < 2) queue_chunk();
  }

  // Arnoldi steps [j0, j0 + K) for the members of `mask`
  void steps(int j0, int K, uint64_t mask) {
    const int ldh = L_.ldh;
    for (int j = j0; j < j0 + K; ++j) {
      precond_apply(p_, o_.prec, R_, Vb(j, 0), Zb(j, 0), adj_, false);
      op_apply(p_, R_, Zb(j, 0), W(), adj_);
      if (seq_) {
        for (int q = 0; q < R_; ++q) {
          if (!((mask >> q) & 1ull)) continue;
          double2* wq = W() + q * N_;
          double2* hc = dH(q) + (size_t)ldh * j;
          // modified Gram-Schmidt in sequence: each update fused with the next
          // coefficient's dot, the last with ||w||
          launch_dot(p_, N_, Vb(0, q), wq, HZ_MASK_FULL, hc, p_->stream);
          for (int i = 0; i <= j; ++i) {
            const double2* vnext = i < j ? Vb(i + 1, q) : nullptr;
            launch_axpy_dot(p_, N_, hc + i, Vb(i, q), wq, vnext, hc + i + 1, p_->stream);
          }
          // v_{j+1} = w / h_{j+1,j}, zero on breakdown
          launch_scale_inv_guard(N_, hc + j + 1, o_.breakdown, wq, Vb(j + 1, q), p_->stream);
        }
        launch_givens(j, m_, ldh, mask, dTop(), L_.size, L_.offH, L_.offCS, L_.offSN, L_.offG, dNrm() + R_,
                      dNf(), o_.tol, o_.breakdown, dEst(), p_->stream);
      } else {
        // inverse compact-WY for every member of the mask at once: r = V^H w
        // and l = V^H v_j in one pass, h = (I + L)^{-1} r, then the new basis
        // vector w - V h and h_{j+1,j} = its norm in a second pass; the vector
        // is stored unnormalised with its scale 1 / h_{j+1,j} (0 on breakdown);
        // the Givens update closes the step
        launch_gs_members(p_, N_, R_, mask, j, m_, ldh, Vb(0, 0), vstride(), W(), dTop(), (int64_t)L_.size,
                          L_.offR, L_.offL, L_.offLg, L_.offH, L_.offS, L_.offCS, L_.offSN, L_.offG, dNrm() + R_,
                          dNf(), o_.tol, o_.breakdown, dEst(), Vb(j + 1, 0), p_->stream);
      }
    }
  }

  void queue_chunk() {
    const int j0 = next_j_, K = std::min(chunk_, m_ - j0);
    next_j_ += K;
    bool eager = true;
    if (o_.graph) {
      // captured on its second use: the first runs eagerly and sets every
      // kernel attribute, so the capture makes no attribute calls
      GraphSlot& g = graphs_[{j0, mask_}];
      if (g.uses++ > 0) {
        eager = false;
        if (!g.exec) {
          hipGraph_t gr = nullptr;
          HZ_HIP(hipStreamBeginCapture(p_->stream, hipStreamCaptureModeThreadLocal));
          try {
            steps(j0, K, mask_);
          } catch (...) {
            (void)hipStreamEndCapture(p_->stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
          }
          HZ_HIP(hipStreamEndCapture(p_->stream, &gr));
          const hipError_t e = hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          HZ_HIP(e);
        }
        HZ_HIP(hipGraphLaunch(g.exec, p_->stream));
      }
    }
    if (eager) steps(j0, K, mask_);
    const int s = take_slot();
    HZ_HIP(hipMemcpyAsync(pinEst(s), dEst(), (size_t)R_ * m_ * sizeof(double2), hipMemcpyDeviceToHost, p_->stream));
    post(Wait::CHUNK, s, j0, K);
  }

  void on_chunk(const Pending& w) {
    if (w.cyc != cyc_) return;  // queued before its cycle ended: its steps are unused
    const double2* est = pinEst(w.slot);
    for (int j = w.j0; j < w.j0 + w.K; ++j)
      for (int q = 0; q < R_; ++q) {
        Member& a = M_[q];
        if (!a.in_cycle) continue;
        st_[q].iters++;
        a.jj = j + 1;
        const double2 e = est[(size_t)q * m_ + j];
        if (q == 0) hist_.push(e.x);
        const int status = (int)e.y;
        if (status == 3) {
          set_error("non-finite Arnoldi coefficient in FGMRES (right-hand side %d)", q);
          throw Error{HZ_ERR_NONFINITE};
        }
        if (status == 1) a.in_cycle = false;
        else if (status == 2) a.brk = true, a.in_cycle = false;
      }
    mask_ = 0;
    for (int q = 0; q < R_; ++q)
      if (M_[q].in_cycle) mask_ |= 1ull << q;
    if (mask_ && next_j_ < m_) {
      queue_chunks();
      return;
    }
    if (mask_ && chunks_pending() > 0) return;  // the cycle's last chunk is still running
    end_cycle();
  }

  // x_q += Z_q y_q, R y = g, for every member that ran this cycle
  void end_cycle() {
    for (int q = 0; q < R_; ++q) {
      Member& a = M_[q];
      a.in_cycle = false;
      if (a.done || a.jj == 0) continue;
      launch_backsub(a.jj, L_.ldh, dMem(q) + L_.offH, dMem(q) + L_.offG, dMem(q) + L_.offY,
                     seq_ ? nullptr : dMem(q) + L_.offS, p_->stream);
      launch_lincomb(N_, a.jj, dMem(q) + L_.offY, Zb(0, q), vstride(), x_ + q * N_, true, p_->stream);
      a.jj = 0;
      if (a.brk) a.done = true;
    }
    mask_ = 0;
    ++cyc_;
    cycle_begin();
  }

  void on_final(const Pending& w) {
    const double2* h = pinNrm(w.slot);
    const double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    for (int q = 0; q < R_; ++q) {
      Member& a = M_[q];
      st_[q].t_solve_s = ts;
      st_[q].n_hist = q == 0 ? hist_.n : 0;
      if (a.norm_f == 0.0) {
        st_[q].converged = 1;
        continue;
      }
      a.res = h[q].x;
      a.rel = a.res / a.norm_f;
      a.conv = a.rel < o_.tol;
      st_[q].converged = a.conv;
      st_[q].breakdown = a.brk && !a.conv;
      st_[q].rel_res = a.rel;
      st_[q].res = a.res;
    }
    done_ = true;
  }
};

struct JobState {
  int q0 = 0;
  std::unique_ptr<Lockstep> cur;
  bool active = true;
};

}  // namespace

void ensure_krylov(hz_problem* p, int m, int R) {
  const size_t vb = (size_t)p->N * sizeof(double2) * R;
  const bool fresh = p->kV.bytes < vb * (m + 1) || p->kZ.bytes < vb * m;
  grow(p->kV, vb * (m + 1));
  grow(p->kZ, vb * m);
  grow(p->kr, vb);
  grow(p->kw, vb);
  grow(p->kx, vb);
  grow(p->kaux, vb);
  grow(p->scal, std::max<size_t>(4096, scalar_slots(m, R)) * sizeof(double2));
  // reduction partials of the largest multi-dot of the cycle for R members at
  // once (and of any dot)
  const int ng = (m + 1 + MD_GROUP - 1) / MD_GROUP;
  grow(p->red, (size_t)RED_BLOCKS * std::max(2 * MD_GROUP * ng, 4) * sizeof(double2) * R);
  ensure_work(p, R);
  if (fresh) {
    // basis slots a batch member never reaches still enter the batched
    // preconditioner: keep them finite
    launch_fill((int64_t)(p->kV.bytes / sizeof(double2)), make_double2(0, 0), p->kV.as<double2>(), p->stream);
    launch_fill((int64_t)(p->kZ.bytes / sizeof(double2)), make_double2(0, 0), p->kZ.as<double2>(), p->stream);
  }
}

void run_lockstep(std::vector<LockJob>& jobs) {
  const int nj = (int)jobs.size();
  std::vector<JobState> js(nj);
  auto fail = [&](int i, int code) {
    LockJob& J = jobs[i];
    J.rc = code;
    J.msg = hz_last_error();
    js[i].cur.reset();  // synchronises the job's stream
    js[i].active = false;
  };
  // Set-up before any solve starts: launch choices timed on an idle device,
  // every buffer of the largest batch allocated.
  for (int i = 0; i < nj; ++i) {
    LockJob& J = jobs[i];
    try {
      HZ_CHECK(J.p && J.nrhs >= 1 && J.f && J.u && J.st, HZ_ERR_ARG, "bad lockstep job");
      J.batch = std::max(1, std::min(J.batch, J.nrhs));
      HZ_CHECK(J.batch <= LOCKSTEP_MAX, HZ_ERR_ARG, "lockstep batch %d exceeds %d", J.batch, LOCKSTEP_MAX);
      if (J.o.autotune)
        for (int r : {J.batch, J.nrhs % J.batch})
          if (r > 0 && !J.p->tuned.count({J.o.prec, r})) autotune(J.p, J.o.prec, r, nullptr);
      ensure_krylov(J.p, J.o.m, J.batch);
      J.p->pin.reserve(NSLOT * pin_slot(J.o.m, J.batch) * sizeof(double2));
      HZ_HIP(hipStreamSynchronize(J.p->stream));
    } catch (const Error& e) {
      fail(i, e.code);
    }
  }
  auto start_next = [&](int i) {
    LockJob& J = jobs[i];
    JobState& s = js[i];
    s.cur.reset();
    if (s.q0 >= J.nrhs) {
      s.active = false;
      return;
    }
    const int R = std::min(J.batch, J.nrhs - s.q0);
    hz_krylov_opts o = J.o;
    if (s.q0 > 0) {
      o.hist = nullptr;
      o.hist_cap = 0;
    }
    s.cur.reset(new Lockstep(J.p, R, J.f + (size_t)s.q0 * J.p->N, J.u + (size_t)s.q0 * J.p->N, o, J.st + s.q0,
                             J.adjoint));
    s.q0 += R;
    s.cur->begin();
  };
  for (int i = 0; i < nj; ++i) {
    if (!js[i].active) continue;
    try {
      start_next(i);
    } catch (const Error& e) {
      fail(i, e.code);
    }
  }
  for (;;) {
    int active = 0, last = -1;
    bool progressed = false;
    for (int i = 0; i < nj; ++i) {
      JobState& s = js[i];
      if (!s.active) continue;
      try {
        while (s.active && s.cur->ready()) {
          s.cur->process();
          progressed = true;
          if (s.cur->finished()) start_next(i);
        }
      } catch (const Error& e) {
        fail(i, e.code);
      }
      if (s.active) {
        ++active;
        last = i;
      }
    }
    if (!active) break;
    if (progressed) continue;
    if (active == 1) {
      try {
        js[last].cur->wait();
      } catch (const Error& e) {
        fail(last, e.code);
      }
    } else {
      std::this_thread::yield();
    }
  }
}

}  // namespace hz
