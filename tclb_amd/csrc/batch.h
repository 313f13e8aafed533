// batch.h -- many independent lattices of one model and storage format in
// one ragged device arena, advanced together by one persistent kernel per
// step over a tile work-list.
//
// There is no counterpart in the reference (one Lattice per process there,
// src/Lattice.h.Rt:39-191); every per-case operation mirrors the Lattice one
// (settings, zonal settings, flags, Init, Globals, quantities, densities).
// A case is computed exactly as the same case alone: the node arithmetic is
// shared with the single-lattice kernel, and a case's Globals are reduced
// over its own tiles only, in a fixed order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "kernels.h"
#include "lattice.h"
#include "model.h"

namespace tclbx {

// First-fit allocator over one device allocation (offsets 256-byte aligned).
class Arena {
public:
    explicit Arena(size_t capacity);
    ~Arena();
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    // offset of a free block of `bytes`, or SIZE_MAX if none fits
    size_t alloc(size_t bytes);
    void release(size_t offset);
    char* base() const { return base_; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }

private:
    char* base_ = nullptr;
    size_t cap_, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    std::map<size_t, size_t> live_;  // offset -> size
};

struct Tuning {
    int block = 256;  // x-segment width of a tile (threads per block)
    int rpt = 1;      // rows per plain tile (nodes per thread)
    int grid = 0;     // persistent blocks (0: one block per tile)
    int flat = 0;     // 1: tiles are runs of the linear node index (no idle lanes at row ends)
    int order = 0;    // 1: tiles walked in XCD-contiguous order (xcd_tile)
    double us_per_step = 0;
};

class Batch {
public:
    // capacity_bytes: HBM held by the arena (snapshots and flags of all
    // cases); admission fails when a case does not fit
    Batch(const ModelDef* m, int storage, size_t capacity_bytes);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // a new case; returns its id, or -1 if the arena has no room for it
    int AddCase(int nx, int ny, int nz);
    void RemoveCase(int id);
    bool HasCase(int id) const;
    size_t CaseBytes(int nx, int ny, int nz) const;  // arena bytes one case needs
    int ncases() const { return (int)active_.size(); }
    std::vector<int> case_ids() const;

    void SetSetting(int id, int setting, double v);   // id -1: every case
    double GetSetting(int id, int setting) const;
    void SetZoneSetting(int id, int zs, int zone, double v);
    void FlagOverwrite(int id, const uint16_t* flags);  // whole lattice of the case
    void GetFlags(int id, uint16_t* out);
    void InitCase(int id);  // Init() of one case
    void Iterate(int niter, int iter_type);
    // rebuild the device tables and capture the step loop now, so that the
    // next Iterate runs steps only
    void Prepare();
    const std::vector<double>& Globals(int id) const;
    long long CaseIteration(int id) const;
    void GetQuantity(int id, int q, double* out, double scale);
    void GetDensity(int id, int q, double* out);
    void SetDensity(int id, int q, const double* in);

    // time every tile shape / grid candidate on the current cases (steps
    // that rewrite only the next snapshot, so the state is untouched) and
    // keep the fastest
    Tuning Autotune(int steps);
    void SetTuning(const Tuning& t);
    // stochastic rounding of every case's stores (16-bit formats only); a
    // case is keyed by its own node index, iteration and `seed`, as a lone
    // Lattice with the same seed keys it, so its result is the same alone and
    // in a batch.  Steps are then launched one by one (no captured loop).
    void SetStochasticRounding(bool on, unsigned seed);
    bool stochastic_rounding() const { return sr_; }
    const Tuning& tuning() const { return tune_; }
    std::vector<Tuning> last_trials() const { return trials_; }

    double last_iterate_ms() const { return last_ms_; }
    double total_ms() const { return total_ms_; }
    long long nodes() const;  // nodes of all active cases
    size_t bytes_per_node_step() const;
    size_t arena_used() const { return arena_.used(); }
    size_t arena_capacity() const { return arena_.capacity(); }
    void set_graph_steps(int k) { graph_steps_ = k & ~1; drop_graph(); }

    const ModelDef* model;
    const int storage;

private:
    struct Case {
        int nx, ny, nz;
        size_t N, pitch;
        size_t snap_off[2], flags_off;
        double* d_zone = nullptr;
        std::vector<double> settings;
        std::vector<double> zone;
        bool zone_dirty = true;
        int parity = 0;
        long long iter = 0;
        std::vector<double> globals;
    };
    StepArgs case_args(const Case& c) const;
    void rebuild();
    void upload_zones();
    // one step; k = its index in the current Iterate (stochastic rounding
    // keys a case's step k by the case's iteration + k)
    void launch(int phase, bool glob, int k = 0);
    void upload_sr_iterations();
    void drop_graph();
    void capture_graph();
    Case& get(int id);
    const Case& get(int id) const;

    Arena arena_;
    std::map<int, Case> active_;
    int next_id_ = 0;
    int phase_ = 0;
    bool dirty_ = true;
    bool sr_ = false;
    unsigned sr_seed_ = 0;
    std::vector<BatchCase> descs_;  // host copy of the case table
    Tuning tune_;
    std::vector<Tuning> trials_;
    // device tables, rebuilt when the case set or tuning changes
    BatchCase* d_cases_ = nullptr;
    int2* d_tiles_ = nullptr;
    int2* d_gtiles_ = nullptr;
    int* d_first_ = nullptr;
    int* d_count_ = nullptr;
    double* d_gpart_ = nullptr;
    double* d_glob_ = nullptr;
    int ntiles_ = 0, ngtiles_ = 0, ncases_dev_ = 0;
    std::vector<int> order_;  // case id of device slot i
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double last_ms_ = 0, total_ms_ = 0;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    int graph_steps_ = 32;
    unsigned version_ = 0, graph_version_ = ~0u;
    int ncu_ = 256;
};

}  // namespace tclbx
