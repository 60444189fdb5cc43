// Serving loop over per-world engines (see server.h). One worker thread owns every HIP call of the served worlds; the
// caller's threads only queue requests, wait, and hand reads to the worker.
#include "server.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

namespace avbd {

namespace {

#define SRV_HIP(x)                                                                                            \
  do {                                                                                                        \
    hipError_t err_ = (x);                                                                                    \
    if (err_ != hipSuccess) throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err_) + " at " #x); \
  } while (0)

template <class T> double percentile(std::vector<T> v, double q) {
  if (v.empty()) return 0;
  size_t k = size_t(std::floor(q / 100.0 * double(v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + k, v.end());
  return double(v[k]);
}

}  // namespace

Server::Server(const ServerOptions &opt) : opt_(opt) {
  if (opt_.maxActive == 0) throw std::invalid_argument("maxActive must be positive");
  if (opt_.lookahead == 0) throw std::invalid_argument("lookahead must be positive");
  if (opt_.budgetBytes == 0) throw std::invalid_argument("the HBM budget must be positive");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw std::runtime_error("no HIP device is available");
  if (opt_.device < 0 || opt_.device >= n) throw std::invalid_argument("device ordinal out of range");
  opt_.engine.adaptive = false;  // a served world's run must not depend on when counters happen to be read
  opt_.engine.device = opt_.device;
  worker_ = std::thread([this] { loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (worker_.joinable()) worker_.join();
  (void)hipSetDevice(opt_.device);
  for (auto &kv : requests_)
    for (auto &e : kv.second->inflight) eventPool_.push_back(e);
  for (auto &e : eventPool_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (auto &w : worlds_) w->engine.reset();
}

double Server::now() const {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

uint32_t Server::addWorld(const HostWorld &w) {
  if (w.bodies.empty()) throw std::invalid_argument("a served world needs at least one body");
  std::lock_guard<std::mutex> lk(mu_);
  if (!error_.empty()) throw std::runtime_error(error_);
  auto rec = std::make_unique<World>();
  rec->scene = w;
  worlds_.push_back(std::move(rec));
  return uint32_t(worlds_.size() - 1);
}

uint32_t Server::bodyCount(uint32_t world) {
  std::lock_guard<std::mutex> lk(mu_);
  if (world >= worlds_.size()) throw std::out_of_range("no such world");
  return uint32_t(worlds_[world]->scene.bodies.size());
}

uint64_t Server::submit(uint32_t world, uint32_t steps) {
  const double t = now();
  if (steps == 0) throw std::invalid_argument("a request needs at least one step");
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (!error_.empty()) throw std::runtime_error(error_);
    if (world >= worlds_.size()) throw std::out_of_range("no such world");
    id = nextId_++;
    auto r = std::make_unique<Request>();
    r->id = id;
    r->world = world;
    r->steps = steps;
    r->submitted = t;
    requests_[id] = std::move(r);
    pending_.push_back(id);
    worlds_[world]->queued++;
    counters_.submitted++;
    if (statsStart_ < 0) statsStart_ = t;
  }
  cv_.notify_all();
  return id;
}

bool Server::done(uint64_t request) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(request);
  if (it == requests_.end()) throw std::out_of_range("no such request");
  return it->second->done;
}

double Server::wait(uint64_t request) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = requests_.find(request);
  if (it == requests_.end()) throw std::out_of_range("no such request");
  Request *r = it->second.get();
  doneCv_.wait(lk, [&] { return r->done || !error_.empty(); });
  if (!r->done) throw std::runtime_error(error_);
  return (r->finished - r->submitted) * 1e3;
}

void Server::readBodies(uint32_t world, std::vector<Body> &out) {
  std::unique_lock<std::mutex> lk(mu_);
  if (world >= worlds_.size()) throw std::out_of_range("no such world");
  World &w = *worlds_[world];
  doneCv_.wait(lk, [&] { return w.queued == 0 || !error_.empty(); });
  if (!error_.empty()) throw std::runtime_error(error_);
  // The world is idle: the worker only touches it to evict it, which happens under this lock
  if (w.engine) {
    out.resize(w.engine->bodyCount());
    SRV_HIP(hipSetDevice(opt_.device));
    w.engine->readBodies(0, w.engine->bodyCount(), out.data());
  } else if (w.hasSaved) {
    out = w.saved.bodies;
  } else {
    out = w.scene.bodies;
  }
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  ServerStats s = counters_;
  s.stepMsP50 = float(percentile(stepMs_, 50));
  s.stepMsP99 = float(percentile(stepMs_, 99));
  s.stepMsMax = stepMs_.empty() ? 0.f : *std::max_element(stepMs_.begin(), stepMs_.end());
  s.requestMsP50 = percentile(requestMs_, 50);
  s.requestMsP99 = percentile(requestMs_, 99);
  s.wallSeconds = statsStart_ < 0 ? 0 : std::max(0.0, lastFinish_ - statsStart_);
  s.residentBytes = resident_;
  s.peakResidentBytes = peak_;
  s.worlds = uint32_t(worlds_.size());
  s.residentWorlds = 0;
  for (auto &w : worlds_) s.residentWorlds += w->engine ? 1u : 0u;
  return s;
}

void Server::resetStats() {
  std::lock_guard<std::mutex> lk(mu_);
  counters_ = ServerStats{};
  stepMs_.clear();
  requestMs_.clear();
  statsStart_ = -1;
  lastFinish_ = 0;
  peak_ = resident_;
}

size_t Server::estimateBytes(const World &w) const {
  // The engine's allocations per body and per pair slot (pair capacity 8 per body, at least 4096), with the BVH
  const size_t n = w.scene.bodies.size();
  const size_t pairs = opt_.engine.pairCapacity ? opt_.engine.pairCapacity : std::max<size_t>(4096, 8 * n);
  return n * (460 + 180) + pairs * 1240 + (size_t(1) << 16);
}

void Server::makeResident(World &w) {
  std::vector<HostWorld> one{w.scene};
  if (w.hasSaved) {
    one[0].bodies = w.saved.bodies;
    one[0].joints = w.saved.joints;
  }
  // A small world's iterations in one workgroup: one launch instead of one per colour and iteration, so the device's
  // few hardware queues carry more worlds' work at once (the same results either way)
  EngineOptions eo = opt_.engine;
  if (w.scene.bodies.size() <= opt_.smallWorldBodies) {
    eo.solver = 3;
    eo.worldBlock = opt_.smallWorldBlock;
  } else if (opt_.largeWorldBodies && w.scene.bodies.size() >= opt_.largeWorldBodies) {
    // A large world's iterations in one persistent launch holding a share of the device's workgroups: as many of
    // them fit at once as the shares allow, so solves running together never wait for workgroups that cannot start
    eo.solver = 2;
    eo.dataflowShare = opt_.largeWorldShare;
  }
  w.engine = std::make_unique<Engine>(one, eo);
  if (w.hasSaved) {
    w.engine->restore(w.saved);
    w.saved = SavedState{};
    w.hasSaved = false;
    counters_.restores++;
  }
  w.bytes = w.engine->deviceBytes();
  resident_ += w.bytes;
  peak_ = std::max(peak_, resident_);
}

void Server::evict(World &w) {
  w.saved = w.engine->save();
  w.hasSaved = true;
  w.engine.reset();
  resident_ -= w.bytes;
  w.bytes = 0;
  counters_.evictions++;
}

/** Start a waiting request if its world is free and fits the budget (evicting idle worlds, least recent first). */
bool Server::admit(Request &r) {
  World &w = *worlds_[r.world];
  if (w.active) return false;
  if (!w.engine) {
    const size_t need = estimateBytes(w);
    while (resident_ + need > opt_.budgetBytes) {
      World *victim = nullptr;
      for (auto &c : worlds_)
        if (c->engine && !c->active && (!victim || c->lastUse < victim->lastUse)) victim = c.get();
      if (!victim) break;
      evict(*victim);
    }
    // Over budget with nothing left to evict: wait for running requests, unless none runs (then it must go alone)
    if (resident_ + need > opt_.budgetBytes && !active_.empty()) return false;
    makeResident(w);
  }
  w.active = true;
  w.lastUse = ++tick_;
  counters_.admissions++;
  return true;
}

void Server::loop() {
  try {
    SRV_HIP(hipSetDevice(opt_.device));
    std::vector<float> stepTimes;
    while (true) {
      std::vector<Request *> running;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !pending_.empty() || !active_.empty(); });
        if (stop_) return;
        // Admission, in arrival order; requests of a busy world wait for it without blocking the others
        for (auto it = pending_.begin(); it != pending_.end() && active_.size() < opt_.maxActive;) {
          Request &r = *requests_[*it];
          if (admit(r)) {
            active_.push_back(r.id);
            it = pending_.erase(it);
          } else if (worlds_[r.world]->active) {
            ++it;
          } else {
            break;  // out of memory until something finishes
          }
        }
        for (uint64_t id : active_) running.push_back(requests_[id].get());
      }
      // Launch: keep `lookahead` steps in flight per request, each bracketed by events on the world's stream
      for (Request *r : running) {
        Engine &e = *worlds_[r->world]->engine;
        while (r->inflight.size() < opt_.lookahead && r->launched < r->steps) {
          std::pair<hipEvent_t, hipEvent_t> ev;
          if (!eventPool_.empty()) {
            ev = eventPool_.back();
            eventPool_.pop_back();
          } else {
            SRV_HIP(hipEventCreate(&ev.first));
            SRV_HIP(hipEventCreate(&ev.second));
          }
          SRV_HIP(hipEventRecord(ev.first, e.stream()));
          e.step(1);
          SRV_HIP(hipEventRecord(ev.second, e.stream()));
          r->inflight.push_back(ev);
          r->launched++;
        }
      }
      // Retire finished steps; a request whose last step finished completes
      bool progressed = false;
      stepTimes.clear();
      std::vector<Request *> finished;
      for (Request *r : running) {
        while (!r->inflight.empty()) {
          auto ev = r->inflight.front();
          hipError_t q = hipEventQuery(ev.second);
          if (q == hipErrorNotReady) break;
          SRV_HIP(q);
          float ms = 0;
          SRV_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
          stepTimes.push_back(ms);
          eventPool_.push_back(ev);
          r->inflight.pop_front();
          r->retired++;
          progressed = true;
        }
        if (r->retired == r->steps) finished.push_back(r);
      }
      // A finished request's world reports whether a dataflow wait gave up in its steps (reading the counters also
      // switches that world to colour launches)
      uint64_t aborts = 0;
      for (Request *r : finished) {
        Engine &e = *worlds_[r->world]->engine;
        if (e.solverKind() == 2) aborts += e.readCounters().solveAborts;
      }
      if (!stepTimes.empty() || !finished.empty()) {
        const double t = now();
        std::lock_guard<std::mutex> lk(mu_);
        stepMs_.insert(stepMs_.end(), stepTimes.begin(), stepTimes.end());
        counters_.solveAborts += aborts;
        for (Request *r : finished) {
          World &w = *worlds_[r->world];
          r->done = true;
          r->finished = t;
          w.active = false;
          w.queued--;
          w.lastUse = ++tick_;
          counters_.completed++;
          counters_.steps += r->steps;
          counters_.bodySteps += uint64_t(r->steps) * w.scene.bodies.size();
          requestMs_.push_back((r->finished - r->submitted) * 1e3);
          lastFinish_ = t;
          active_.erase(std::find(active_.begin(), active_.end(), r->id));
        }
        if (!finished.empty()) doneCv_.notify_all();
      }
      if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  } catch (const std::exception &ex) {
    std::lock_guard<std::mutex> lk(mu_);
    error_ = std::string("serving loop stopped: ") + ex.what();
    doneCv_.notify_all();
  }
}

}  // namespace avbd
