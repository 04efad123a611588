// Continuous-batching scheduler of the serving loop (host code).
//
// Replicas arrive over time.  The loop admits them in arrival order (FIFO,
// no overtaking, so a large replica is never starved by smaller ones behind
// it) when a batch slot and a contiguous range of the device-memory arena are
// free, lets the executor prepare them asynchronously, primes the ones whose
// state has become resident, and advances every active replica together in
// chunks of at most chunk_steps steps (a chunk ends early when a replica
// reaches its step count, so replicas retire on time and their slot and
// memory go to the next in line).  The executor is an interface: the GPU
// server implements it with device memory and batched kernels, and a
// simulated executor with a virtual clock and a fake memory budget lets the
// policy be tested without a GPU.
#ifndef MMD_SCHED_H
#define MMD_SCHED_H

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <limits>
#include <map>
#include <vector>

namespace mmd {

// First-fit allocator over [0, size) in 256-byte granules; freed ranges are
// merged with their neighbours.
class ArenaAlloc {
 public:
  explicit ArenaAlloc(size_t size) : size_(size & ~(size_t)255) {
    if (size_) free_[0] = size_;
  }
  static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
  bool alloc(size_t bytes, size_t *off) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      live_[*off] = bytes;
      used_ += bytes;
      return true;
    }
    return false;
  }
  bool release(size_t off) {
    auto lv = live_.find(off);
    if (lv == live_.end()) return false;
    size_t beg = off, len = lv->second;
    used_ -= len;
    live_.erase(lv);
    auto nx = free_.lower_bound(beg);
    if (nx != free_.end() && nx->first == beg + len) {
      len += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == beg) {
        beg = pv->first;
        len += pv->second;
        free_.erase(pv);
      }
    }
    free_[beg] = len;
    return true;
  }
  size_t size() const { return size_; }
  size_t used() const { return used_; }

 private:
  size_t size_, used_ = 0;
  std::map<size_t, size_t> free_, live_;  // offset -> bytes
};

enum RepState { kWaiting, kQueued, kLoading, kActive, kDone, kFailed };

struct Rep {
  // request
  uint64_t id = 0, seed = 0;
  int32_t cells = 0, steps = 0;
  double temperature = 0.0, arrival = 0.0;
  int64_t atoms = 0;  // lattice sites of the box
  size_t bytes = 0;   // device bytes
  // schedule
  RepState state = kWaiting;
  int status = 0;
  int slot = -1;
  size_t off = 0;
  int32_t done = 0;
  double t_admit = 0.0, t_ready = 0.0, t_done = 0.0;
  double e[4] = {0.0, 0.0, 0.0, 0.0};  // {epot, ekin} after the last step, then at admission
};

class Executor {
 public:
  virtual ~Executor() = default;
  virtual double now() = 0;
  // begins preparing replica r (asynchronous; its slot and arena range are assigned)
  virtual int start_load(int r) = 0;
  // appends replicas whose preparation ended (Rep::status tells how); when
  // none has ended, waits until one does or the clock reaches `deadline`
  virtual int wait_loaded(std::vector<int> &out, double deadline) = 0;
  // computes forces and admission energies of freshly loaded replicas
  virtual int prime(const std::vector<int> &reps) = 0;
  // advances the active replicas by `steps` steps (energies recorded at the last)
  virtual int run(const std::vector<int> &active, int steps) = 0;
  // collects the results of replicas that reached their step count
  virtual int finish(const std::vector<int> &reps) = 0;
  // the replica's slot and memory are about to be reused
  virtual void release(int r) { (void)r; }
};

struct SchedStats {
  double wall_s = 0.0, atom_steps = 0.0;
  int64_t batched_steps = 0;
  int32_t completed = 0, failed = 0, max_active = 0;
};

// Runs the whole trace; returns 0, or the first executor error (replicas not
// finished by then are marked failed with that status).
inline int serve_loop(std::vector<Rep> &R, int slots, size_t arena_bytes, int chunk_steps, Executor &X,
                      SchedStats &st, int nomem_status) {
  const int n = (int)R.size();
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return R[a].arrival < R[b].arrival; });
  ArenaAlloc A(arena_bytes);
  std::vector<int> free_slots;
  for (int s = slots - 1; s >= 0; s--) free_slots.push_back(s);
  std::deque<int> queue;
  std::vector<int> active, loaded, fresh, finished;
  int next = 0, n_loading = 0, closed = 0, err = 0;
  const double inf = std::numeric_limits<double>::infinity();

  auto retire = [&](int r, RepState s, int status) {
    X.release(r);
    R[r].state = s;
    R[r].status = status;
    R[r].t_done = X.now();
    if (R[r].slot >= 0) free_slots.push_back(R[r].slot);
    A.release(R[r].off);
    closed++;
    (s == kDone ? st.completed : st.failed)++;
  };

  while (closed < n && !err) {
    double now = X.now();
    while (next < n && R[order[next]].arrival <= now) {
      R[order[next]].state = kQueued;
      queue.push_back(order[next++]);
    }
    // FIFO admission by slot and arena space
    while (!queue.empty()) {
      const int r = queue.front();
      if (ArenaAlloc::round(R[r].bytes) > A.size()) {
        queue.pop_front();
        R[r].state = kFailed;
        R[r].status = nomem_status;
        R[r].t_done = now;
        closed++, st.failed++;
        continue;
      }
      if (free_slots.empty() || !A.alloc(R[r].bytes, &R[r].off)) break;
      queue.pop_front();
      R[r].slot = free_slots.back();
      free_slots.pop_back();
      R[r].state = kLoading;
      R[r].t_admit = now;
      n_loading++;
      if ((err = X.start_load(r))) break;
    }
    if (err) break;
    // wait for loads only when the device has nothing else to do
    const double deadline = !active.empty() ? -inf : (next < n ? R[order[next]].arrival : inf);
    if (active.empty() && n_loading == 0 && deadline == inf) {
      if (closed < n) err = -1;  // nothing can make progress (cannot happen with a consistent trace)
      break;
    }
    loaded.clear();
    if ((err = X.wait_loaded(loaded, deadline))) break;
    fresh.clear();
    finished.clear();
    for (int r : loaded) {
      n_loading--;
      if (R[r].status) {
        retire(r, kFailed, R[r].status);
        continue;
      }
      fresh.push_back(r);
    }
    if (!fresh.empty()) {
      if ((err = X.prime(fresh))) break;
      now = X.now();
      for (int r : fresh) {
        R[r].state = kActive;
        R[r].t_ready = now;
        if (R[r].steps == 0) {
          R[r].e[0] = R[r].e[2], R[r].e[1] = R[r].e[3];
          finished.push_back(r);
        } else {
          active.push_back(r);
        }
      }
    }
    if (!active.empty()) {
      st.max_active = std::max<int32_t>(st.max_active, (int32_t)active.size());
      int steps = chunk_steps;
      double atoms = 0.0;
      for (int r : active) {
        steps = std::min(steps, R[r].steps - R[r].done);
        atoms += (double)R[r].atoms;
      }
      if ((err = X.run(active, steps))) break;
      st.batched_steps += steps;
      st.atom_steps += atoms * steps;
      size_t keep = 0;
      for (size_t k = 0; k < active.size(); k++) {
        const int r = active[k];
        R[r].done += steps;
        if (R[r].done >= R[r].steps)
          finished.push_back(r);
        else
          active[keep++] = r;
      }
      active.resize(keep);
    }
    if (!finished.empty()) {
      if ((err = X.finish(finished))) break;
      for (int r : finished) retire(r, kDone, 0);
    }
  }
  if (err) {
    for (auto &r : R)
      if (r.state != kDone && r.state != kFailed) {
        r.state = kFailed;
        r.status = err;
        st.failed++;
      }
  }
  st.wall_s = X.now();
  return err;
}

// Nearest-rank percentile of a sample (q in [0, 1]).
inline double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t k = (size_t)std::ceil(q * (double)v.size());
  return v[std::min(v.size() - 1, k == 0 ? 0 : k - 1)];
}

}  // namespace mmd

#endif
