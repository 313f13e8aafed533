// server.cpp -- see server.h.
#include "server.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace tclbx {

// cbStop::DoIt, src/Handlers/cbStop.cpp:53-96
bool StopRule::check(const std::vector<double>& g) {
    int any = 0;
    for (size_t i = 0; i < crit.size(); i++) {
        const double v = g.at(crit[i].global);
        switch (crit[i].type) {
        case STOP_CHANGE:
            if (std::fabs(old[i] - v) > crit[i].limit) any++;
            break;
        case STOP_PERCENTCHANGE:
            if (std::fabs((old[i] - v) / v) > crit[i].limit) any++;
            break;
        case STOP_ABOVE:
            if (v < crit[i].limit) any++;
            break;
        case STOP_BELOW:
            if (v > crit[i].limit) any++;
            break;
        }
        old[i] = v;
    }
    if (!any) score++;
    else score = 0;
    if (score >= times) {
        reset();
        return true;
    }
    return false;
}

Server::Server(const ModelDef* m, int storage, size_t hbm_budget, int max_chunk)
    : model(m), max_chunk_(max_chunk), t0_(std::chrono::steady_clock::now()) {
    if (max_chunk < 1) throw std::invalid_argument("chunk must be at least one step");
    batch_.reset(new Batch(m, storage, hbm_budget));
    // requests have arbitrary sizes: flat tiles leave no lane idle at the
    // end of a row; four rows per tile (a block walks four runs of 256
    // nodes) measured 64.6k MLUPS on the d3q19 f16s serving mix against
    // 51.6k with one, as fast as one 256^3 lattice alone (66.4k)
    Tuning t;
    t.flat = 1;
    t.block = 256;
    t.rpt = 4;
    batch_->SetTuning(t);
    worker_ = std::thread([this]() { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        quit_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

std::string Server::Error() {
    std::lock_guard<std::mutex> lk(mu_);
    return error_;
}

int Server::Submit(Request r) {
    // the checks Batch::AddCase makes, so that a malformed request fails
    // here and never reaches the worker
    check_lattice_size(model, r.nx, r.ny, r.nz, batch_->storage);
    if ((long long)r.flags.size() != (long long)r.nx * r.ny * r.nz)
        throw std::invalid_argument("flags do not match the lattice size");
    if (r.max_steps < 1) throw std::invalid_argument("step budget must be positive");
    if (r.stop.interval < 0 || r.stop.times < 1) throw std::invalid_argument("bad stop rule");
    for (auto& c : r.stop.crit)
        if (c.global < 0 || c.global >= (int)model->globals.size() || c.type < 0 || c.type > 3)
            throw std::invalid_argument("bad stop criterion");
    r.stop.reset();
    std::lock_guard<std::mutex> lk(mu_);
    if (!error_.empty()) throw std::runtime_error("server stopped after an error: " + error_);
    const int id = next_id_++;
    RequestResult res;
    res.t_submit = now();
    results_[id] = res;
    inflight_++;
    queue_.emplace_back(id, std::move(r));
    stats_.queued = (int)queue_.size();
    cv_.notify_all();
    return id;
}

RequestResult Server::Poll(int id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = results_.find(id);
    if (it == results_.end()) throw std::out_of_range("no such request");
    return it->second;
}

bool Server::WaitAll(double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    return done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), [&]() { return inflight_ == 0; });
}

Server::Stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    Stats s = stats_;
    s.active = (int)active_.size();
    s.queued = (int)queue_.size();
    return s;
}

void Server::SetTuning(const Tuning& t) {
    std::lock_guard<std::mutex> lk(mu_);
    pending_tuning_ = t;
    has_pending_tuning_ = true;
}

void Server::SetStochasticRounding(bool on, unsigned seed) {
    if (on && batch_->storage != ST_F16S && batch_->storage != ST_BF16S)
        throw std::invalid_argument("stochastic rounding applies to the 16-bit formats f16s and bf16s");
    std::lock_guard<std::mutex> lk(mu_);
    pending_sr_ = on;
    pending_sr_seed_ = seed;
    has_pending_sr_ = true;
}

// A device or host error ends the worker: every request not finished is
// marked rejected and the error is kept for Error().
void Server::loop() {
    try {
        run();
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu_);
        error_ = e.what();
        const double t = now();
        for (auto& kv : results_)
            if (kv.second.state == RQ_QUEUED || kv.second.state == RQ_RUNNING) {
                kv.second.state = RQ_REJECTED;
                kv.second.t_finish = t;
            }
        queue_.clear();
        active_.clear();
        inflight_ = 0;
        quit_ = true;
        done_cv_.notify_all();
    }
}

void Server::run() {
    for (;;) {
        std::deque<std::pair<int, Request>> pending;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&]() { return quit_ || !queue_.empty() || !active_.empty(); });
            if (quit_) return;
            pending.swap(queue_);
            if (has_pending_tuning_) {
                batch_->SetTuning(pending_tuning_);
                has_pending_tuning_ = false;
            }
            if (has_pending_sr_) {
                batch_->SetStochasticRounding(pending_sr_, pending_sr_seed_);
                has_pending_sr_ = false;
            }
        }
        // admission: in arrival order, every request that fits the arena now
        // (a later, smaller one may start before an earlier, larger one)
        std::deque<std::pair<int, Request>> waiting;
        for (auto& pr : pending) {
            Request& r = pr.second;
            const size_t need = batch_->CaseBytes(r.nx, r.ny, r.nz);
            if (need > batch_->arena_capacity()) {
                std::lock_guard<std::mutex> lk(mu_);
                results_[pr.first].state = RQ_REJECTED;
                results_[pr.first].t_finish = now();
                results_[pr.first].error = "the request needs " + std::to_string(need) +
                                           " bytes of the arena, which holds " +
                                           std::to_string(batch_->arena_capacity());
                stats_.rejected++;
                inflight_--;
                continue;
            }
            // an error while admitting one request rejects that request
            // only; an error of the device shows again at the next step,
            // which ends the worker
            int cid = -1;
            try {
                cid = batch_->AddCase(r.nx, r.ny, r.nz);
                if (cid >= 0) {
                    batch_->FlagOverwrite(cid, r.flags.data());
                    for (auto& s : r.settings) batch_->SetSetting(cid, s.first, s.second);
                    for (auto& z : r.zonal) batch_->SetZoneSetting(cid, z.id, z.zone, z.v);
                    batch_->InitCase(cid);
                }
            } catch (const std::exception& e) {
                if (cid >= 0 && batch_->HasCase(cid)) batch_->RemoveCase(cid);
                std::lock_guard<std::mutex> lk(mu_);
                results_[pr.first].state = RQ_REJECTED;
                results_[pr.first].t_finish = now();
                results_[pr.first].error = e.what();
                stats_.rejected++;
                inflight_--;
                continue;
            }
            if (cid < 0) {
                waiting.push_back(std::move(pr));
                continue;
            }
            Active a;
            a.rid = pr.first;
            a.max_steps = r.max_steps;
            a.stop = r.stop;
            std::lock_guard<std::mutex> lk(mu_);
            active_[cid] = std::move(a);
            results_[pr.first].state = RQ_RUNNING;
            results_[pr.first].t_admit = now();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto it = waiting.rbegin(); it != waiting.rend(); ++it) queue_.push_front(std::move(*it));
            if (active_.empty()) {
                done_cv_.notify_all();
                continue;
            }
        }
        // one chunk: up to the next event of any resident request
        long long chunk = max_chunk_;
        for (auto& kv : active_) {
            const Active& a = kv.second;
            chunk = std::min(chunk, a.max_steps - a.steps);
            if (a.stop.interval > 0) chunk = std::min<long long>(chunk, a.stop.interval - a.steps % a.stop.interval);
        }
        chunk = std::max(1LL, chunk);
        const long long nodes = batch_->nodes();
        batch_->Iterate((int)chunk, ITER_NORM | ITER_LASTGLOB);
        std::vector<int> finished;
        std::vector<std::pair<int, int>> states;
        for (auto& kv : active_) {
            Active& a = kv.second;
            a.steps += chunk;
            int st = RQ_RUNNING;
            if (a.stop.interval > 0 && a.steps % a.stop.interval == 0 && a.stop.check(batch_->Globals(kv.first)))
                st = RQ_DONE_STOP;
            else if (a.steps >= a.max_steps)
                st = RQ_DONE_BUDGET;
            if (st != RQ_RUNNING) {
                finished.push_back(kv.first);
                states.emplace_back(kv.first, st);
            }
        }
        std::vector<std::pair<int, RequestResult>> done;
        for (auto& cs : states) {
            const Active& a = active_.at(cs.first);
            RequestResult r;
            r.state = cs.second;
            r.steps = a.steps;
            r.globals = batch_->Globals(cs.first);
            done.emplace_back(a.rid, r);
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            stats_.node_steps += nodes * chunk;
            stats_.busy_s += batch_->last_iterate_ms() * 1e-3;
            stats_.chunks++;
            const double t = now();
            for (auto& d : done) {
                RequestResult& res = results_[d.first];
                res.state = d.second.state;
                res.steps = d.second.steps;
                res.globals = d.second.globals;
                res.t_finish = t;
                stats_.completed++;
                inflight_--;
            }
            for (int cid : finished) active_.erase(cid);
        }
        for (int cid : finished) batch_->RemoveCase(cid);
        done_cv_.notify_all();
    }
}

}  // namespace tclbx
