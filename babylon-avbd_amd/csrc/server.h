// Serving loop: step/rollout requests for many worlds of different sizes, admitted as they arrive, with world state kept
// resident in HBM under a byte budget (least recently used idle worlds are evicted to host memory and restored exactly
// when next requested). Admitted requests run concurrently, each world on its own HIP stream: every turn of the loop
// launches the next step of every active request (continuous batching at step granularity), retires finished steps
// and admits waiting requests as slots and memory free up. Per-step device latency is timed with HIP events.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"

namespace avbd {

struct ServerOptions {
  size_t budgetBytes = size_t(200) << 30;  // HBM the resident worlds may hold
  uint32_t maxActive = 32;                 // requests running at once (one stream each)
  uint32_t lookahead = 2;                  // steps in flight per active request
  int device = 0;
  EngineOptions engine;                    // per-world options (not adaptive: a run must not depend on timing)
  uint32_t smallWorldBodies = 0;           // worlds of at most this many bodies: one workgroup solves each step
  uint32_t smallWorldBlock = 1024;         // its threads
  uint32_t largeWorldBodies = 0;           // worlds of at least this many bodies (0: none): the dataflow solve, ...
  uint32_t largeWorldShare = 4;            // ... each grid this share of the device
};

struct ServerStats {
  uint64_t submitted = 0, completed = 0, admissions = 0, evictions = 0, restores = 0;
  uint64_t steps = 0, bodySteps = 0;
  uint64_t solveAborts = 0;                // dataflow waits given up (read as each request finishes)
  double wallSeconds = 0;                  // since the first submission (or the last reset) until the last completion
  float stepMsP50 = 0, stepMsP99 = 0, stepMsMax = 0;
  double requestMsP50 = 0, requestMsP99 = 0;
  size_t residentBytes = 0, peakResidentBytes = 0;
  uint32_t residentWorlds = 0, worlds = 0;
};

class Server {
 public:
  explicit Server(const ServerOptions &opt);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  /** Register a world (kept on the host until a request needs it). */
  uint32_t addWorld(const HostWorld &w);
  uint32_t bodyCount(uint32_t world);
  /** Queue `steps` steps of a world; returns the request id. Requests of one world run in submission order. */
  uint64_t submit(uint32_t world, uint32_t steps);
  /** Block until the request has finished; returns its latency (submission to last step done) in ms. */
  double wait(uint64_t request);
  bool done(uint64_t request);
  /** The world's bodies once its queued requests have finished (from HBM, or the host copy if evicted). */
  void readBodies(uint32_t world, std::vector<Body> &out);
  ServerStats stats();
  void resetStats();

 private:
  struct World {
    HostWorld scene;
    std::unique_ptr<Engine> engine;  // resident
    SavedState saved;                // evicted (or never run: empty)
    bool hasSaved = false;
    size_t bytes = 0;
    uint64_t lastUse = 0;
    uint32_t queued = 0;             // requests submitted and not finished
    bool active = false;
  };
  struct Request {
    uint64_t id;
    uint32_t world, steps, launched = 0, retired = 0;
    double submitted = 0, finished = 0;
    bool done = false;
    std::deque<std::pair<hipEvent_t, hipEvent_t>> inflight;
  };

  void loop();
  bool admit(Request &r);
  void makeResident(World &w);
  void evict(World &w);
  size_t estimateBytes(const World &w) const;
  double now() const;

  ServerOptions opt_;
  std::mutex mu_;
  std::condition_variable cv_, doneCv_;
  std::vector<std::unique_ptr<World>> worlds_;
  std::deque<uint64_t> pending_;
  std::map<uint64_t, std::unique_ptr<Request>> requests_;
  std::vector<uint64_t> active_;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> eventPool_;
  uint64_t nextId_ = 1, tick_ = 0;
  size_t resident_ = 0, peak_ = 0;
  ServerStats counters_;
  std::vector<float> stepMs_;
  std::vector<double> requestMs_;
  double statsStart_ = -1, lastFinish_ = 0;
  bool stop_ = false;
  std::string error_;
  std::thread worker_;
};

}  // namespace avbd
