// server.h -- continuous-batching solve server over one Batch.
//
// Requests (a lattice, its NodeTypes and settings, a step budget and stop
// criteria) arrive from any thread.  A worker thread owns the Batch: at every
// chunk boundary it admits queued requests while the arena (the HBM budget)
// has room, advances every resident request together by one chunk, evaluates
// their stop criteria and retires the finished ones, freeing their memory.
// Chunks end at the next event of any resident request (a stop check or the
// end of a step budget), so every request runs exactly its own steps.
//
// Stop criteria follow the reference's <Stop> handler
// (src/Handlers/cbStop.cpp:12-104): <Global>Change, <Global>PercentChange,
// <Global>Above, <Global>Below, checked every `interval` steps; a check
// passes when every criterion holds, and the request stops after `times`
// consecutive passing checks.  There is no serving layer in the reference.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"

namespace tclbx {

enum StopType { STOP_CHANGE = 0, STOP_PERCENTCHANGE = 1, STOP_ABOVE = 2, STOP_BELOW = 3 };

struct StopCriterion {
    int global;
    int type;
    double limit;
};

// cbStop state: the previous value of each criterion and the score
struct StopRule {
    std::vector<StopCriterion> crit;
    int interval = 0;  // steps between checks (0: no stop rule)
    int times = 1;
    std::vector<double> old;
    int score = 0;
    void reset() { old.assign(crit.size(), -12341234.0); score = 0; }
    // one check on the Globals; true when the request should stop
    bool check(const std::vector<double>& g);
};

struct Request {
    int nx = 0, ny = 0, nz = 1;
    std::vector<uint16_t> flags;
    std::vector<std::pair<int, double>> settings;
    struct Zonal { int id, zone; double v; };
    std::vector<Zonal> zonal;
    long long max_steps = 0;
    StopRule stop;
};

enum RequestState { RQ_REJECTED = -1, RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE_BUDGET = 2, RQ_DONE_STOP = 3 };

struct RequestResult {
    int state = RQ_QUEUED;
    long long steps = 0;
    std::vector<double> globals;
    double t_submit = 0, t_admit = -1, t_finish = -1;  // seconds since the server started
    std::string error;  // why a rejected request was rejected
};

class Server {
public:
    Server(const ModelDef* m, int storage, size_t hbm_budget, int max_chunk);
    ~Server();
    int Submit(Request r);                 // request id
    RequestResult Poll(int id);
    bool WaitAll(double timeout_s);        // true when nothing is queued or running
    struct Stats {
        long long completed = 0, rejected = 0;
        long long node_steps = 0;
        double busy_s = 0;
        int active = 0, queued = 0;
        long long chunks = 0;
    };
    Stats stats();
    void SetTuning(const Tuning& t);
    // stochastic rounding of the requests' stores (Batch::SetStochasticRounding)
    void SetStochasticRounding(bool on, unsigned seed);
    double now() const;
    std::string Error();  // message of the error that stopped the worker, "" if none

    const ModelDef* model;

private:
    void loop();
    void run();

    std::unique_ptr<Batch> batch_;
    int max_chunk_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::pair<int, Request>> queue_;
    std::map<int, RequestResult> results_;
    struct Active {
        int rid;
        long long steps = 0, max_steps = 0;
        StopRule stop;
    };
    std::map<int, Active> active_;  // case id -> request
    int next_id_ = 0;
    long long inflight_ = 0;  // submitted and not yet finished or rejected
    bool quit_ = false;
    std::string error_;
    Tuning pending_tuning_;
    bool has_pending_tuning_ = false;
    bool has_pending_sr_ = false, pending_sr_ = false;
    unsigned pending_sr_seed_ = 0;
    Stats stats_;
    std::chrono::steady_clock::time_point t0_;
    std::thread worker_;
};

}  // namespace tclbx
