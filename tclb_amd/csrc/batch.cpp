// batch.cpp -- see batch.h.
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace tclbx {

// ---------------------------------------------------------------- arena ---
Arena::Arena(size_t capacity) : cap_((capacity + 255) / 256 * 256) {
    if (cap_ == 0) throw std::invalid_argument("arena capacity must be positive");
    hip_check(hipMalloc((void**)&base_, cap_), "hipMalloc(arena)");
    free_[0] = cap_;
}

Arena::~Arena() {
    if (base_) (void)hipFree(base_);
}

size_t Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, size = it->second;
        free_.erase(it);
        if (size > bytes) free_[off + bytes] = size - bytes;
        live_[off] = bytes;
        used_ += bytes;
        return off;
    }
    return std::numeric_limits<size_t>::max();
}

void Arena::release(size_t offset) {
    auto it = live_.find(offset);
    if (it == live_.end()) throw std::invalid_argument("arena: unknown block");
    size_t off = it->first, size = it->second;
    live_.erase(it);
    used_ -= size;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + size) {
        size += next->second;
        free_.erase(next);
    }
    auto prev = free_.lower_bound(off);
    if (prev != free_.begin()) {
        --prev;
        if (prev->first + prev->second == off) {
            off = prev->first;
            size += prev->second;
            free_.erase(prev);
        }
    }
    free_[off] = size;
}

// ---------------------------------------------------------------- batch ---
namespace {
// the arena holds distributions, flags and zone tables: a model with
// densities that are not streamed (the AVG builds' averages, the bc builds'
// BC) runs on a Lattice.  Checked before the arena is allocated.
const ModelDef* batchable(const ModelDef* m, int storage) {
    if (!m) throw std::invalid_argument("unknown model");
    if (storage < 0 || storage >= ST_COUNT) throw std::invalid_argument("unknown storage format");
    if (m->nextra() > 0)
        throw std::invalid_argument("model " + m->name +
                                    " keeps densities that are not streamed, which a batch does not hold; use a lattice");
    return m;
}
}  // namespace

Batch::Batch(const ModelDef* m, int storage_, size_t capacity_bytes)
    : model(batchable(m, storage_)), storage(storage_), arena_(capacity_bytes) {
    graph_steps_ = default_graph_steps();
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&ev0_), "hipEventCreate");
    hip_check(hipEventCreate(&ev1_), "hipEventCreate");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    ncu_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
}

Batch::~Batch() {
    // nothing may still run from the captured graph when it is destroyed
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (auto& kv : active_)
        if (kv.second.d_zone) (void)hipFree(kv.second.d_zone);
    for (void* p : {(void*)d_cases_, (void*)d_tiles_, (void*)d_gtiles_, (void*)d_first_, (void*)d_count_,
                    (void*)d_gpart_, (void*)d_glob_})
        if (p) (void)hipFree(p);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

size_t Batch::CaseBytes(int nx, int ny, int nz) const {
    const size_t N = (size_t)nx * ny * nz;
    const int S = storage_bytes(storage);
    const size_t align = 256 / S;
    const size_t pitch = (N + align - 1) / align * align;
    const size_t snap = (size_t)model->Q * pitch * S;
    return 2 * snap + (N * sizeof(uint16_t) + 255) / 256 * 256;
}

int Batch::AddCase(int nx, int ny, int nz) {
    check_lattice_size(model, nx, ny, nz, storage);
    const size_t N = (size_t)nx * ny * nz;
    const size_t bytes = CaseBytes(nx, ny, nz);
    const size_t off = arena_.alloc(bytes);
    if (off == std::numeric_limits<size_t>::max()) return -1;
    Case c;
    c.nx = nx; c.ny = ny; c.nz = nz; c.N = N;
    const int S = storage_bytes(storage);
    const size_t align = 256 / S;
    c.pitch = (N + align - 1) / align * align;
    const size_t snap = (size_t)model->Q * c.pitch * S;
    c.snap_off[0] = off;
    c.snap_off[1] = off + snap;
    c.flags_off = off + 2 * snap;
    c.settings.resize(model->settings.size());
    for (size_t i = 0; i < model->settings.size(); i++)
        c.settings[i] = model->settings[i].derived < 0 ? model->settings[i].def : 0.0;
    const int zcols = (int)model->zonal.size(), nzones = model->zone_max + 1;
    c.zone.assign((size_t)nzones * zcols, 0.0);
    for (int z = 0; z < nzones; z++)
        for (int k = 0; k < zcols; k++) c.zone[(size_t)z * zcols + k] = model->zonal[k].def;
    hipError_t e = hipMalloc((void**)&c.d_zone, std::max<size_t>(1, c.zone.size()) * sizeof(double));
    if (e != hipSuccess) {
        arena_.release(off);
        hip_check(e, "hipMalloc(zone table)");
    }
    c.globals.assign(kernel_nglobals(model->id), 0.0);
    hip_check(hipMemsetAsync(arena_.base() + off, 0, bytes, stream_), "clear case");
    hip_check(hipStreamSynchronize(stream_), "AddCase");
    const int id = next_id_++;
    active_.emplace(id, std::move(c));
    dirty_ = true;
    return id;
}

void Batch::RemoveCase(int id) {
    Case& c = get(id);
    hip_check(hipStreamSynchronize(stream_), "RemoveCase");
    if (c.d_zone) (void)hipFree(c.d_zone);
    arena_.release(c.snap_off[0]);
    active_.erase(id);
    dirty_ = true;
}

bool Batch::HasCase(int id) const { return active_.count(id) != 0; }

std::vector<int> Batch::case_ids() const {
    std::vector<int> ids;
    for (auto& kv : active_) ids.push_back(kv.first);
    return ids;
}

Batch::Case& Batch::get(int id) {
    auto it = active_.find(id);
    if (it == active_.end()) throw std::out_of_range("no such case in the batch");
    return it->second;
}

const Batch::Case& Batch::get(int id) const {
    auto it = active_.find(id);
    if (it == active_.end()) throw std::out_of_range("no such case in the batch");
    return it->second;
}

void Batch::SetSetting(int id, int setting, double v) {
    if (setting < 0 || setting >= (int)model->settings.size()) throw std::out_of_range("setting id");
    std::vector<Case*> cs;
    if (id == -1) for (auto& kv : active_) cs.push_back(&kv.second);
    else cs.push_back(&get(id));
    for (Case* c : cs) {
        int s = setting;
        double val = v;
        while (s >= 0) {  // derived settings, as Lattice::SetSetting
            c->settings[s] = val;
            const SettingDef& d = model->settings[s];
            if (d.derived < 0) break;
            val = derived_value(d, val);
            s = d.derived;
        }
    }
    dirty_ = true;
}

double Batch::GetSetting(int id, int setting) const {
    if (setting < 0 || setting >= (int)model->settings.size()) throw std::out_of_range("setting id");
    return get(id).settings[setting];
}

void Batch::SetZoneSetting(int id, int zs, int zone, double v) {
    const int zcols = (int)model->zonal.size(), nzones = model->zone_max + 1;
    if (zs < 0 || zs >= zcols) throw std::out_of_range("zonal setting id");
    if (zone < -1 || zone >= nzones) throw std::out_of_range("zone");
    Case& c = get(id);
    if (zone == -1) for (int z = 0; z < nzones; z++) c.zone[(size_t)z * zcols + zs] = v;
    else c.zone[(size_t)zone * zcols + zs] = v;
    c.zone_dirty = true;
}

void Batch::FlagOverwrite(int id, const uint16_t* flags) {
    Case& c = get(id);
    hip_check(hipMemcpy(arena_.base() + c.flags_off, flags, c.N * sizeof(uint16_t), hipMemcpyHostToDevice),
              "FlagOverwrite");
}

void Batch::GetFlags(int id, uint16_t* out) {
    Case& c = get(id);
    hip_check(hipMemcpy(out, arena_.base() + c.flags_off, c.N * sizeof(uint16_t), hipMemcpyDeviceToHost),
              "GetFlags");
}

StepArgs Batch::case_args(const Case& c) const {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.fin = arena_.base() + c.snap_off[0];
    a.fout = arena_.base() + c.snap_off[1];
    a.flags = (const uint16_t*)(arena_.base() + c.flags_off);
    a.zone = c.d_zone;
    a.pitch = (long long)c.pitch;
    a.nx = c.nx; a.ny = c.ny; a.nz = c.nz;
    a.flat = tune_.flat;
    a.inv_nx = 1.0 / c.nx;
    a.inv_ny = 1.0 / c.ny;
    if (tune_.flat) {
        // a "row" of a flat tile is a run of `block` consecutive nodes
        a.bpr = 1;
        a.rpb = globals_rows_per_block(((long long)c.N + tune_.block - 1) / tune_.block, 1);
    } else {
        a.bpr = (c.nx + tune_.block - 1) / tune_.block;
        a.rpb = globals_rows_per_block(c.ny * c.nz, a.bpr);
    }
    a.rpt = tune_.rpt;
    a.zcols = (int)model->zonal.size();
    a.nzones = model->zone_max + 1;
    const NodeTypeDef* ib = model->nodetype("IB");
    a.ib_mask = ib ? ib->value : 0u;   // a batch holds no cuts: the branch never runs
    model_params(model, c.settings, a.par, a.fpar);
    return a;
}

void Batch::upload_zones() {
    bool any = false;
    for (auto& kv : active_) {
        Case& c = kv.second;
        if (!c.zone_dirty) continue;
        if (!c.zone.empty())
            hip_check(hipMemcpyAsync(c.d_zone, c.zone.data(), c.zone.size() * sizeof(double),
                                     hipMemcpyHostToDevice, stream_), "upload zone table");
        c.zone_dirty = false;
        any = true;
    }
    if (any) hip_check(hipStreamSynchronize(stream_), "upload zone tables");
}

void Batch::InitCase(int id) {
    Case& c = get(id);
    upload_zones();
    StepArgs a = case_args(c);
    a.fout = arena_.base() + c.snap_off[0];
    hip_check(launch_init(model->id, storage, a, tune_.block, stream_), "init kernel");
    hip_check(hipStreamSynchronize(stream_), "InitCase");
    c.parity = phase_;  // snapshot 0 is read at the next step
    c.iter = 0;
    std::fill(c.globals.begin(), c.globals.end(), 0.0);
    dirty_ = true;
}

// device tables: one descriptor per case, the plain and the Globals tile
// lists, and per case the slice of Globals tiles it reduces
void Batch::rebuild() {
    drop_graph();
    for (void* p : {(void*)d_cases_, (void*)d_tiles_, (void*)d_gtiles_, (void*)d_first_, (void*)d_count_,
                    (void*)d_gpart_, (void*)d_glob_})
        if (p) (void)hipFree(p);
    d_cases_ = nullptr; d_tiles_ = d_gtiles_ = nullptr; d_first_ = d_count_ = nullptr;
    d_gpart_ = d_glob_ = nullptr;
    order_.clear();
    std::vector<BatchCase> descs;
    std::vector<int2> tiles, gtiles;
    std::vector<int> first, count;
    for (auto& kv : active_) {
        const Case& c = kv.second;
        const int slot = (int)order_.size();
        order_.push_back(kv.first);
        BatchCase bc;
        std::memset(&bc, 0, sizeof(bc));
        bc.args = case_args(c);
        bc.parity = c.parity;
        bc.sr_iter = (unsigned)c.iter;
        bc.sr_seed = sr_seed_;
        descs.push_back(bc);
        const int nt = step_blocks(bc.args, false, tune_.block);
        for (int t = 0; t < nt; t++) tiles.push_back(make_int2(slot, t));
        const int ng = step_blocks(bc.args, true, tune_.block);
        first.push_back((int)gtiles.size());
        count.push_back(ng);
        for (int t = 0; t < ng; t++) gtiles.push_back(make_int2(slot, t));
    }
    ntiles_ = (int)tiles.size();
    ngtiles_ = (int)gtiles.size();
    ncases_dev_ = (int)descs.size();
    const int G = kernel_nglobals(model->id);
    if (ncases_dev_ > 0) {
        hip_check(hipMalloc((void**)&d_cases_, descs.size() * sizeof(BatchCase)), "hipMalloc(cases)");
        hip_check(hipMalloc((void**)&d_tiles_, tiles.size() * sizeof(int2)), "hipMalloc(tiles)");
        hip_check(hipMalloc((void**)&d_gtiles_, gtiles.size() * sizeof(int2)), "hipMalloc(tiles)");
        hip_check(hipMalloc((void**)&d_first_, first.size() * sizeof(int)), "hipMalloc(first)");
        hip_check(hipMalloc((void**)&d_count_, count.size() * sizeof(int)), "hipMalloc(count)");
        hip_check(hipMalloc((void**)&d_gpart_, (size_t)ngtiles_ * G * 2 * sizeof(double)), "hipMalloc(partials)");
        hip_check(hipMalloc((void**)&d_glob_, (size_t)ncases_dev_ * G * sizeof(double)), "hipMalloc(globals)");
        hip_check(hipMemcpy(d_cases_, descs.data(), descs.size() * sizeof(BatchCase), hipMemcpyHostToDevice),
                  "upload cases");
        descs_ = descs;
        hip_check(hipMemcpy(d_tiles_, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice),
                  "upload tiles");
        hip_check(hipMemcpy(d_gtiles_, gtiles.data(), gtiles.size() * sizeof(int2), hipMemcpyHostToDevice),
                  "upload tiles");
        hip_check(hipMemcpy(d_first_, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice),
                  "upload first");
        hip_check(hipMemcpy(d_count_, count.data(), count.size() * sizeof(int), hipMemcpyHostToDevice),
                  "upload count");
        hip_check(hipMemset(d_glob_, 0, (size_t)ncases_dev_ * G * sizeof(double)), "clear globals");
    }
    dirty_ = false;
    version_++;
}

void Batch::launch(int phase, bool glob, int k) {
    if (ncases_dev_ == 0) return;
    const int nt = glob ? ngtiles_ : ntiles_;
    const int grid = glob ? std::min(nt, 8 * ncu_) : (tune_.grid > 0 ? std::min(tune_.grid, nt) : nt);
    // the phase word: bit 0 the snapshot phase, bit 1 the tile order, the
    // bits above the step's index k in the Iterate (stochastic rounding);
    // with stochastic rounding on, the kernels of its instantiation
    const int st = !sr_ ? storage : (storage == ST_F16S ? ST_F16S_SR : ST_BF16S_SR);
    hip_check(launch_batch_step(model->id, st, glob, d_cases_, glob ? d_gtiles_ : d_tiles_, nt,
                                (phase & 1) | (tune_.order ? 2 : 0) | (k << 2), d_gpart_, grid, tune_.block,
                                stream_),
              "batch step kernel");
    if (glob) {
        hip_check(launch_reduce_globals_cases(d_gpart_, ngtiles_, kernel_nglobals(model->id), model->sum_globals,
                                              ncases_dev_, d_first_, d_count_, d_glob_, stream_),
                  "batch globals reduction");
    }
}

void Batch::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_version_ = ~0u;
}

// graph_steps_ plain steps from phase 0, captured once per set of cases
void Batch::capture_graph() {
    if (graph_exec_ != nullptr && graph_version_ == version_) return;
    drop_graph();
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int k = 0; k < graph_steps_; k++) {
        const int nt = ntiles_;
        const int grid = tune_.grid > 0 ? std::min(tune_.grid, nt) : nt;
        hipError_t e = launch_batch_step(model->id, storage, false, d_cases_, d_tiles_, nt,
                                         (k & 1) | (tune_.order ? 2 : 0), d_gpart_, grid, tune_.block, stream_);
        if (e != hipSuccess) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            hip_check(e, "batch step kernel (capture)");
        }
    }
    hip_check(hipStreamEndCapture(stream_, &graph_), "end capture");
    hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "graph instantiate");
    graph_version_ = version_;
}

void Batch::SetStochasticRounding(bool on, unsigned seed) {
    if (on && storage != ST_F16S && storage != ST_BF16S)
        throw std::invalid_argument(std::string("stochastic rounding applies to the 16-bit formats f16s and bf16s, not ") +
                                    storage_name(storage));
    sr_ = on;
    sr_seed_ = seed;
    dirty_ = true;  // the case table carries the seed
}

// each case's iteration at the start of this Iterate into the case table
void Batch::upload_sr_iterations() {
    if (ncases_dev_ == 0) return;
    for (int slot = 0; slot < ncases_dev_; slot++) descs_[slot].sr_iter = (unsigned)get(order_[slot]).iter;
    hip_check(hipMemcpyAsync(d_cases_, descs_.data(), descs_.size() * sizeof(BatchCase), hipMemcpyHostToDevice,
                             stream_),
              "upload case iterations");
}

void Batch::Prepare() {
    if (dirty_) rebuild();
    upload_zones();
    if (ncases_dev_ > 0 && graph_steps_ >= 2 && !sr_) capture_graph();
}

void Batch::Iterate(int niter, int iter_type) {
    if (niter < 0) throw std::invalid_argument("negative iteration count");
    if (dirty_) rebuild();
    upload_zones();
    if (ncases_dev_ == 0) return;
    int last_glob = iter_type & ITER_LASTGLOB;
    if (iter_type & ITER_INTEG) last_glob = 0;
    const bool every_glob = (iter_type & ITER_GLOBS) != 0;
    const int G = kernel_nglobals(model->id);
    hip_check(hipEventRecord(ev0_, stream_), "event record");
    int i = 0;
    const int plain_end = niter - (last_glob ? 1 : 0);
    const bool captured = graph_exec_ != nullptr && graph_version_ == version_;
    const int lead = phase_ != 0 ? 1 : 0;
    if (sr_) upload_sr_iterations();
    if (!every_glob && !sr_ && graph_steps_ >= 2 && plain_end - lead >= (captured ? 1 : 2) * graph_steps_) {
        if (phase_ != 0) {
            launch(phase_, false);
            phase_ ^= 1;
            i++;
        }
        capture_graph();
        const int reps = (plain_end - i) / graph_steps_;
        for (int r = 0; r < reps; r++) hip_check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
        i += reps * graph_steps_;
    }
    for (; i < niter; i++) {
        const bool glob = every_glob || (last_glob && i == niter - 1);
        if (last_glob && i == niter - 1)
            hip_check(hipMemsetAsync(d_glob_, 0, (size_t)ncases_dev_ * G * sizeof(double), stream_),
                      "clear globals");
        launch(phase_, glob, i);
        phase_ ^= 1;
    }
    hip_check(hipEventRecord(ev1_, stream_), "event record");
    hip_check(hipStreamSynchronize(stream_), "Iterate");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
    last_ms_ = ms;
    total_ms_ += ms;
    for (auto& kv : active_) kv.second.iter += niter;
    if (last_glob || (iter_type & ITER_INTEG)) {
        std::vector<double> tab((size_t)ncases_dev_ * G);
        hip_check(hipMemcpyAsync(tab.data(), d_glob_, tab.size() * sizeof(double), hipMemcpyDeviceToHost, stream_),
                  "read globals");
        hip_check(hipMemsetAsync(d_glob_, 0, tab.size() * sizeof(double), stream_), "clear globals");
        hip_check(hipStreamSynchronize(stream_), "read globals");
        for (int slot = 0; slot < ncases_dev_; slot++) {
            Case& c = active_.at(order_[slot]);
            if (last_glob) std::fill(c.globals.begin(), c.globals.end(), 0.0);
            for (int g = 0; g < G; g++) c.globals[g] += tab[(size_t)slot * G + g];
        }
    }
}

const std::vector<double>& Batch::Globals(int id) const { return get(id).globals; }
long long Batch::CaseIteration(int id) const { return get(id).iter; }

void Batch::GetQuantity(int id, int q, double* out, double scale) {
    if (q < 0 || q >= (int)model->quantities.size()) throw std::out_of_range("quantity id");
    Case& c = get(id);
    upload_zones();
    StepArgs a = case_args(c);
    if ((phase_ ^ c.parity) & 1) a.fin = arena_.base() + c.snap_off[1];
    const size_t bytes = c.N * (model->quantities[q].vector ? 3 : 1) * sizeof(double);
    double* d = nullptr;
    hip_check(hipMalloc((void**)&d, bytes), "hipMalloc(quantity)");
    hipError_t e = launch_quantity(model->id, storage, q, a, 0, 0, 0, c.nx, c.ny, c.nz, scale, d, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(d);
    hip_check(e, "GetQuantity");
}

void Batch::GetDensity(int id, int q, double* out) {
    if (q < 0 || q >= model->Q) throw std::out_of_range("density id");
    Case& c = get(id);
    const int cur = (phase_ ^ c.parity) & 1;
    double* d = nullptr;
    hip_check(hipMalloc((void**)&d, c.N * sizeof(double)), "hipMalloc(density)");
    hipError_t e = launch_get_density(model->id, storage, arena_.base() + c.snap_off[cur], (long long)c.pitch,
                                      (long long)c.N, q, d, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, c.N * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(d);
    hip_check(e, "GetDensity");
}

void Batch::SetDensity(int id, int q, const double* in) {
    if (q < 0 || q >= model->Q) throw std::out_of_range("density id");
    Case& c = get(id);
    const int cur = (phase_ ^ c.parity) & 1;
    double* d = nullptr;
    hip_check(hipMalloc((void**)&d, c.N * sizeof(double)), "hipMalloc(density)");
    hipError_t e = hipMemcpyAsync(d, in, c.N * sizeof(double), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess)
        e = launch_set_density(model->id, storage, arena_.base() + c.snap_off[cur], (long long)c.pitch,
                               (long long)c.N, q, d, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(d);
    hip_check(e, "SetDensity");
}

long long Batch::nodes() const {
    long long n = 0;
    for (auto& kv : active_) n += (long long)kv.second.N;
    return n;
}

size_t Batch::bytes_per_node_step() const {
    return (size_t)2 * model->Q * storage_bytes(storage) + sizeof(uint16_t);
}

void Batch::SetTuning(const Tuning& t) {
    if (t.block != 64 && t.block != 128 && t.block != 256) throw std::invalid_argument("block must be 64, 128 or 256");
    if (t.rpt < 1 || t.rpt > 64) throw std::invalid_argument("rows per tile must be 1..64");
    if (t.grid < 0) throw std::invalid_argument("grid must be >= 0");
    if (t.order != 0 && t.order != 1) throw std::invalid_argument("tile order must be 0 or 1");
    tune_ = t;
    dirty_ = true;
}

// Every candidate runs `steps` plain steps from the current snapshot into the
// other one (the same phase each time), so the state the next Iterate reads
// is never touched.
Tuning Batch::Autotune(int steps) {
    if (steps < 1) throw std::invalid_argument("autotune needs at least one step");
    if (active_.empty()) return tune_;
    const Tuning saved = tune_;
    trials_.clear();
    Tuning best = saved;
    best.us_per_step = 1e300;
    for (int flat : {0, 1})
    for (int block : {256, 128, 64})
        for (int rpt : {1, 2, 4})
            for (int gmul : {0, 8})
            for (int order : {0, 1}) {
                Tuning t;
                t.flat = flat;
                t.block = block;
                t.rpt = rpt;
                t.grid = gmul * ncu_;
                t.order = order;
                tune_ = t;
                rebuild();
                upload_zones();
                launch(phase_, false);  // warm up
                hip_check(hipEventRecord(ev0_, stream_), "event record");
                for (int s = 0; s < steps; s++) launch(phase_, false);
                hip_check(hipEventRecord(ev1_, stream_), "event record");
                hip_check(hipStreamSynchronize(stream_), "autotune");
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
                t.us_per_step = ms * 1e3 / steps;
                trials_.push_back(t);
                if (t.us_per_step < best.us_per_step) best = t;
            }
    tune_ = best;
    dirty_ = true;
    return best;
}

}  // namespace tclbx
