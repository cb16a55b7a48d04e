// serve.cpp -- batched and served solves: a request loop over worker
// threads, each with its own HIP stream and resident device workspace.
//
// Concurrency comes from the device: every worker runs whole solves on its
// own stream, so the kernels of different problems overlap on the GPU while
// each worker's host-side steps (tree sizes read back, launches) proceed
// independently.  Memory is governed by the workers' workspaces: a worker
// keeps the workspace of its largest solve so far (no allocation per
// request), and a request is admitted only while the sum of all workers'
// workspaces, the chosen worker's grown to the request's estimate, stays
// within the budget.  Idle workers hand their workspace back when the head
// of the queue would not fit.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace fmm {
namespace {

// Workspace bytes of one solve, an upper estimate fitted to measured
// solves (2e3 to 2e6 points, eps 1e-4 to 1e-10, the three kernels): per
// source ~360 B (keys and sort buffers, sorted copies, gathered strengths,
// sorted outputs, staging of host arrays) plus 64 B with dipoles, per
// target ~190 B, and the two expansion arrays over the boxes (at most
// 4 ns / ndiv + 64 of them: refinement follows the sources).  The order
// comes from the library's own rules: l2dterms for Laplace; for Helmholtz
// the largest h2dterms order over levels 2..6 of the points' root box (the
// unit square when the points are in device memory).  A 25% margin on top;
// the server refines it at run time by the largest measured/estimated
// ratio of each kernel.
double estimate_bytes(const fmm2dhip_problem& p) {
  const int ndiv = p.ndiv > 0 ? p.ndiv : ndiv_rule(p.eps, p.ns, p.nt);
  const double nb = 4.0 * p.ns / std::max(ndiv, 1) + 64.0;
  const double ndc = std::max(p.nd, 1) * (p.kernel == FMM2DHIP_LOG ? 2.0 : 1.0);
  double ncoef;
  if (p.kernel == FMM2DHIP_HELMHOLTZ) {
    double bs0 = 1.0;
    if (!p.device_ptrs && p.sources) {
      double x0 = p.sources[0], x1 = x0, y0 = p.sources[1], y1 = y0;
      auto scan = [&](const double* a, int n) {
        for (int i = 0; i < n; ++i) {
          x0 = std::min(x0, a[2 * i]); x1 = std::max(x1, a[2 * i]);
          y0 = std::min(y0, a[2 * i + 1]); y1 = std::max(y1, a[2 * i + 1]);
        }
      };
      scan(p.sources, p.ns);
      if (p.targ && p.nt > 0) scan(p.targ, p.nt);
      bs0 = std::max(x1 - x0, y1 - y0);
    }
    int pmax = 1;
    for (int l = 2; l <= 6; ++l) {
      int e = 0;
      pmax = std::max(pmax, h2dterms_host(bs0 / (1 << l), std::complex<double>(p.zk_re, p.zk_im),
                                          p.eps, &e));
    }
    ncoef = 2.0 * std::min(pmax, 2000) + 1.0;
  } else {
    ncoef = l2dterms(p.eps) + 1.0;
  }
  const double dip = p.ifdipole ? 64.0 : 0.0;
  const double bytes = 64.0 * (1 << 20) + p.ns * (360.0 + dip + 100.0 * (ndc - 1.0)) +
                       p.nt * (190.0 + 60.0 * (ndc - 1.0)) + nb * 2.0 * 16.0 * ndc * ncoef;
  return 1.25 * bytes;
}

double now_s() {
  using C = std::chrono::steady_clock;
  return std::chrono::duration<double>(C::now().time_since_epoch()).count();
}

}  // namespace

class Server {
 public:
  Server(int nworkers, double budget) : budget_(budget) {
    if (hipGetDevice(&device_) != hipSuccess) device_ = 0;
    nworkers = std::min(std::max(nworkers, 1), 32);
    cap_.assign(nworkers, 0.0);
    hold_.assign(nworkers, 0.0);
    busy_.assign(nworkers, 0);
    for (int w = 0; w < nworkers; ++w) threads_.emplace_back([this, w] { worker(w); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
  }

  int workers() const { return (int)threads_.size(); }

  long long submit(fmm2dhip_problem* p) {
    const double est = estimate_bytes(*p) * ratio(p->kernel);
    std::lock_guard<std::mutex> g(mu_);
    if (stop_ || grown(0.0, est) > budget_) return -1;
    const long long t = next_ticket_++;
    const double ts = now_s();
    if (t_first_ < 0) t_first_ = ts;
    queue_.push_back({p, t, ts, est});
    state_[t] = 0;
    cv_work_.notify_all();
    return t;
  }

  // A ticket is forgotten once a wait has seen it done, so a long-running
  // server holds state only for requests still queued, running, or done
  // and not yet waited for.
  int wait(long long ticket, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    if (state_.find(ticket) == state_.end()) return -1;
    auto done = [&] {
      auto it = state_.find(ticket);
      return it == state_.end() || it->second == 1;
    };
    if (timeout_s < 0) {
      cv_done_.wait(lk, done);
    } else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
      return 1;
    }
    state_.erase(ticket);
    return 0;
  }

  int stats(double* out, int n) {
    std::lock_guard<std::mutex> g(mu_);
    std::vector<double> lat = lat_;
    std::sort(lat.begin(), lat.end());
    auto q = [&](double f) {
      if (lat.empty()) return 0.0;
      size_t i = (size_t)std::ceil(f * lat.size()) - 1;
      return lat[std::min(i, lat.size() - 1)];
    };
    const double span = t_last_ > t_first_ ? t_last_ - t_first_ : 0.0;
    double v[11] = {(double)lat.size(), (double)failed_, points_, span,
                    span > 0 ? points_ / span / 1e6 : 0.0, q(0.50), q(0.99),
                    lat.empty() ? 0.0 : lat.back(), peak_bytes_, budget_,
                    (double)state_.size()};
    for (int k = 0; k < n && k < 11; ++k) out[k] = v[k];
    return 11;
  }

 private:
  struct Item {
    fmm2dhip_problem* p;
    long long ticket;
    double t_submit;
    double est;
  };

  double ratio(int kernel) {
    std::lock_guard<std::mutex> g(ratio_mu_);
    return kernel >= 0 && kernel < kKernels ? ratio_[kernel] : 1.0;
  }

  // workspace a worker holding cap ends up with after a solve that takes
  // est bytes: its blocks when they suffice, else one block of 1.25 est
  // (at least the 64 MiB minimum block)
  static double grown(double cap, double est) {
    return est <= cap ? cap : std::max(1.25 * est, 64.0 * (1 << 20));
  }
  // workspace held by all workers if worker w takes an item of estimate est
  double held_if(int w, double est) const {
    double s = 0;
    for (size_t v = 0; v < hold_.size(); ++v) s += (int)v == w ? grown(cap_[v], est) : hold_[v];
    return s;
  }

  void worker(int w) {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      if (shrink_ && !busy_[w] && cap_[w] > 0) {
        // hand the workspace back so the head of the queue can fit
        lk.unlock();
        release_workspace();
        lk.lock();
        cap_[w] = hold_[w] = 0;
        cv_work_.notify_all();
        continue;
      }
      if (!queue_.empty()) {
        Item it = queue_.front();
        if (held_if(w, it.est) <= budget_) {
          queue_.pop_front();
          shrink_ = false;
          busy_[w] = 1;
          hold_[w] = grown(cap_[w], it.est);
          peak_bytes_ = std::max(peak_bytes_, held_if(-1, 0.0));
          lk.unlock();
          solve_problem(*it.p);
          const double t_end = now_s();
          const double cap = (double)workspace_capacity();
          const double used = last_stats().workspace_bytes;
          const double est0 = estimate_bytes(*it.p);
          if (est0 > 0 && it.p->kernel >= 0 && it.p->kernel < kKernels) {
            std::lock_guard<std::mutex> g(ratio_mu_);
            ratio_[it.p->kernel] = std::max(ratio_[it.p->kernel], 1.1 * used / est0);
          }
          lk.lock();
          busy_[w] = 0;
          cap_[w] = hold_[w] = cap;
          peak_bytes_ = std::max(peak_bytes_, held_if(-1, 0.0));
          it.p->latency = t_end - it.t_submit;
          lat_.push_back(it.p->latency);
          if (it.p->ier != 0) ++failed_;
          points_ += (double)it.p->ns + (double)it.p->nt;
          t_last_ = std::max(t_last_, t_end);
          state_[it.ticket] = 1;
          cv_done_.notify_all();
          cv_work_.notify_all();
          continue;
        }
        // the head does not fit while workspaces are held: idle workers
        // (this one included) give theirs back
        bool idle_held = false;
        for (size_t v = 0; v < cap_.size(); ++v) idle_held |= !busy_[v] && cap_[v] > 0;
        if (idle_held && !shrink_) {
          shrink_ = true;
          cv_work_.notify_all();
          continue;
        }
      } else if (stop_) {
        break;
      }
      cv_work_.wait(lk);
    }
    lk.unlock();
    release_thread();
  }

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Item> queue_;
  std::unordered_map<long long, int> state_;  // 0 queued or running, 1 done, not waited for
  std::vector<std::thread> threads_;
  std::vector<double> cap_, hold_;
  std::vector<char> busy_;
  bool stop_ = false, shrink_ = false;
  long long next_ticket_ = 0;
  double budget_;
  int device_ = 0;
  std::vector<double> lat_;
  double points_ = 0, peak_bytes_ = 0, t_first_ = -1, t_last_ = 0;
  long long failed_ = 0;
  std::mutex ratio_mu_;
  // per kernel (FMM2DHIP_CAUCHY .. FMM2DHIP_REAL): largest measured /
  // estimated workspace so far
  static constexpr int kKernels = FMM2DHIP_REAL + 1;
  double ratio_[kKernels] = {1.0, 1.0, 1.0, 1.0};
};

namespace {

// the pool behind fmm2dhip_batch, kept between calls so its workers keep
// their streams and workspaces; a call that asks for more workers replaces
// it (the old pool is joined when its last user drops it).  The holder is
// never destroyed: the last pool's workers stay parked until the process
// ends instead of being joined while the runtime shuts down.
std::mutex g_pool_mu;
std::shared_ptr<Server>* g_pool = new std::shared_ptr<Server>();

std::shared_ptr<Server> batch_pool(int nworkers) {
  std::lock_guard<std::mutex> g(g_pool_mu);
  std::shared_ptr<Server>& pool = *g_pool;
  if (!pool || pool->workers() < nworkers) {
    size_t free_b = 0, total_b = 0;
    double budget = 1e15;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0) budget = 0.9 * (double)total_b;
    pool.reset();
    pool = std::make_shared<Server>(nworkers, budget);
  }
  return pool;
}

}  // namespace
}  // namespace fmm

struct fmm2dhip_server {
  fmm::Server* s;
};

extern "C" {

int fmm2dhip_batch(int n, fmm2dhip_problem* probs, int nworkers) {
  if (n <= 0) return 0;
  if (!probs) return n;
  try {
    std::shared_ptr<fmm::Server> pool = fmm::batch_pool(std::min(std::max(nworkers, 1), 32));
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (double)probs[a].ns + probs[a].nt > (double)probs[b].ns + probs[b].nt;
    });
    std::vector<long long> tickets(n, -1);
    int failed = 0;
    for (int i : order) {
      tickets[i] = pool->submit(&probs[i]);
      if (tickets[i] < 0) {
        probs[i].ier = fmm::IER_ALLOC;
        ++failed;
      }
    }
    for (int i = 0; i < n; ++i) {
      if (tickets[i] < 0) continue;
      pool->wait(tickets[i], -1.0);
      if (probs[i].ier != 0) ++failed;
    }
    return failed;
  } catch (...) {
    return n;
  }
}

fmm2dhip_server* fmm2dhip_server_create(int nworkers, double hbm_budget_bytes) {
  try {
    return new fmm2dhip_server{new fmm::Server(nworkers, hbm_budget_bytes)};
  } catch (...) {
    return nullptr;
  }
}

long long fmm2dhip_server_submit(fmm2dhip_server* s, fmm2dhip_problem* p) {
  if (!s || !p) return -1;
  try {
    return s->s->submit(p);
  } catch (...) {
    return -1;
  }
}

int fmm2dhip_server_wait(fmm2dhip_server* s, long long ticket, double timeout_s) {
  if (!s) return -1;
  return s->s->wait(ticket, timeout_s);
}

int fmm2dhip_server_stats(fmm2dhip_server* s, double* out, int n) {
  if (!s || !out) return 0;
  return s->s->stats(out, n);
}

void fmm2dhip_server_destroy(fmm2dhip_server* s) {
  if (!s) return;
  delete s->s;
  delete s;
}

int fmm2dhip_batch_fused(int n, fmm2dhip_problem* probs, int nworkers) {
  if (n <= 0) return 0;
  if (!probs) return n;
  try {
    const std::vector<std::vector<int>> groups = fmm::fused_groups(n, probs);
    std::vector<char> grouped(n, 0);
    for (const auto& g : groups)
      for (int i : g) grouped[i] = 1;
    // the others go to the worker pool first and run beside the forests
    std::vector<int> rest;
    for (int i = 0; i < n; ++i)
      if (!grouped[i]) rest.push_back(i);
    std::stable_sort(rest.begin(), rest.end(), [&](int a, int b) {
      return (double)probs[a].ns + probs[a].nt > (double)probs[b].ns + probs[b].nt;
    });
    std::shared_ptr<fmm::Server> pool;
    std::vector<long long> tickets(rest.size(), -1);
    int failed = 0;
    if (!rest.empty()) {
      pool = fmm::batch_pool(std::min(std::max(nworkers, 1), 32));
      for (size_t k = 0; k < rest.size(); ++k) {
        tickets[k] = pool->submit(&probs[rest[k]]);
        if (tickets[k] < 0) {
          probs[rest[k]].ier = fmm::IER_ALLOC;
          ++failed;
        }
      }
    }
    for (const auto& g : groups) failed += fmm::solve_forest(probs, g);
    for (size_t k = 0; k < rest.size(); ++k) {
      if (tickets[k] < 0) continue;
      pool->wait(tickets[k], -1.0);
      if (probs[rest[k]].ier != 0) ++failed;
    }
    return failed;
  } catch (...) {
    return n;
  }
}

}  // extern "C"
