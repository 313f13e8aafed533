// lattice.cpp -- see lattice.h.
#include "lattice.h"

#include "pri.h"
#include "storage.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace tclbx {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }
}

namespace {
// x-segment width: the wavefront multiple (64..256) that wastes the fewest
// lanes on a row of nx nodes, preferring wider blocks
int pick_block(int nx) {
    int best = 256;
    long long best_waste = -1;
    for (int b : {256, 128, 64}) {
        long long waste = (long long)((nx + b - 1) / b) * b - nx;
        if (best_waste < 0 || waste < best_waste) { best = b; best_waste = waste; }
    }
    return best;
}
}  // namespace

int default_graph_steps() {
    // a captured loop starts and ends on snapshot 0, so its length is even:
    // an odd value is rounded down (1 -> 0, no graph)
    if (const char* e = std::getenv("TCLBX_GRAPH_STEPS")) return std::max(0, std::atoi(e)) & ~1;
    return 32;
}

void check_lattice_size(const ModelDef* m, int nx, int ny, int nz, int storage) {
    if (!m) throw std::invalid_argument("unknown model");
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::invalid_argument("lattice size must be positive");
    if (m->dim == 2 && nz != 1) throw std::invalid_argument("2D model needs nz = 1");
    if (storage < 0 || storage >= ST_COUNT) throw std::invalid_argument("unknown storage format");
    const unsigned long long N = (unsigned long long)nx * ny * nz;
    if (N >= (1ull << 32)) throw std::invalid_argument("lattice exceeds 2^32 nodes");
    // the step kernel addresses a density plane with 32-bit byte offsets
    const unsigned long long S = storage_bytes(storage);
    const unsigned long long align = 256 / S;
    const unsigned long long plane = (N + align - 1) / align * align * S;
    if (plane > (1ull << 32))
        throw std::invalid_argument("lattice " + std::to_string(nx) + "x" + std::to_string(ny) + "x" +
                                    std::to_string(nz) + " in " + storage_name(storage) + " has a density plane of " +
                                    std::to_string(plane) + " bytes; the limit is 4 GiB per plane (" +
                                    std::to_string((1ull << 32) / S) + " nodes in this format)");
}

Lattice::Lattice(const ModelDef* m, int nx_, int ny_, int nz_, int storage_)
    : model(m), nx(nx_), ny(ny_), nz(nz_), storage(storage_), N((size_t)nx_ * ny_ * nz_) {
    graph_steps_ = default_graph_steps();
    check_lattice_size(m, nx_, ny_, nz_, storage_);
    const int S = storage_bytes(storage);
    const size_t align = 256 / S;
    pitch = (N + align - 1) / align * align;
    // optional extra spacing between density planes (bytes, a multiple of
    // 256): keeps the Q streams of a step off the same HBM channels when the
    // plane size is a large power of two
    if (const char* e = std::getenv("TCLBX_PLANE_PAD_BYTES")) {
        const long long pad = std::atoll(e);
        if (pad > 0) pitch += ((size_t)pad + 255) / 256 * 256 / S;
    }

    settings_.resize(m->settings.size());
    // defaults of the non-deriving settings (src/Solver.cpp.Rt:277-280)
    for (size_t i = 0; i < m->settings.size(); i++)
        settings_[i] = m->settings[i].derived < 0 ? m->settings[i].def : 0.0;
    zcols_ = (int)m->zonal.size();
    nzones_ = m->zone_max + 1;
    zone_.assign((size_t)nzones_ * zcols_, 0.0);
    for (int z = 0; z < nzones_; z++)
        for (int c = 0; c < zcols_; c++) zone_[(size_t)z * zcols_ + c] = m->zonal[c].def;

    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&ev0_), "hipEventCreate");
    hip_check(hipEventCreate(&ev1_), "hipEventCreate");
    const size_t bytes = (size_t)m->Q * pitch * S;
    for (int i = 0; i < 2; i++) {
        hip_check(hipMalloc(&snaps_[i], bytes), "hipMalloc(snapshot)");
        hip_check(hipMemsetAsync(snaps_[i], 0, bytes, stream_), "hipMemset(snapshot)");
    }
    hip_check(hipMalloc((void**)&d_flags_, N * sizeof(uint16_t)), "hipMalloc(flags)");
    hip_check(hipMemsetAsync(d_flags_, 0, N * sizeof(uint16_t), stream_), "hipMemset(flags)");
    if (m->nextra() > 0) {
        const size_t ab = (size_t)m->nextra() * N * sizeof(double);
        hip_check(hipMalloc((void**)&d_extra_, ab), "hipMalloc(extra densities)");
        hip_check(hipMemsetAsync(d_extra_, 0, ab, stream_), "hipMemset(extra densities)");
    }
    hip_check(hipMalloc((void**)&d_zone_, std::max<size_t>(1, zone_.size()) * sizeof(double)), "hipMalloc(zone)");
    set_block(pick_block(nx));
    const int G = kernel_nglobals(m->id);
    hip_check(hipMalloc((void**)&d_glob_, MAX_GLOBALS * sizeof(double)), "hipMalloc(globals)");
    hip_check(hipMemsetAsync(d_glob_, 0, MAX_GLOBALS * sizeof(double), stream_), "hipMemset(globals)");
    globals_.assign(G, 0.0);
    // the class of a node: its NodeType bits below the zone; the bulk class
    // is the plain collision node (MRT, or BGK in d2q9_SRT)
    class_mask_ = (1u << m->zone_shift) - 1u;
    const NodeTypeDef* coll = m->nodetype("MRT");
    if (!coll) coll = m->nodetype("BGK");
    bulk_val_ = coll ? coll->value : 0u;
    if (const char* e = std::getenv("TCLBX_SORTED_DISPATCH")) sorted_ = std::atoi(e) != 0 && coll != nullptr;
    tb_auto_ = temporal_blocking_model(m->id);
    if (const char* e = std::getenv("TCLBX_TEMPORAL_BLOCKING"))
        if (tb_auto_ && std::atoi(e) >= 0) {
            tb_k_ = std::min(TB_TILE / 2 - 1, std::atoi(e));
            tb_auto_ = false;
        }
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

Lattice::~Lattice() {
    // nothing may still run from the captured graph when it is destroyed
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (void* p : snaps_) if (p) (void)hipFree(p);
    if (d_flags_) (void)hipFree(d_flags_);
    if (d_cuts_) (void)hipFree(d_cuts_);
    if (d_extra_) (void)hipFree(d_extra_);
    if (d_zone_) (void)hipFree(d_zone_);
    if (d_gpart_) (void)hipFree(d_gpart_);
    if (d_glob_) (void)hipFree(d_glob_);
    if (d_list_) (void)hipFree(d_list_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Lattice::set_block(int b) {
    if (b != 64 && b != 128 && b != 256) throw std::invalid_argument("block must be 64, 128 or 256");
    block_ = b;
    bpr_ = (nx + b - 1) / b;
    nblocks_ = bpr_ * ny * nz;
    rpb_ = globals_rows_per_block((long long)ny * nz, bpr_);
    if (d_gpart_) (void)hipFree(d_gpart_);
    d_gpart_ = nullptr;
    // compensated partial sums: two doubles per Global per block
    hip_check(hipMalloc((void**)&d_gpart_, (size_t)nblocks_ * kernel_nglobals(model->id) * 2 * sizeof(double)),
              "hipMalloc(partials)");
    drop_graph();
}

size_t Lattice::bytes_per_node_step() const {
    // the averages are read and written in place on collision nodes; the
    // parameter densities are read (only wall nodes write theirs)
    return (size_t)2 * model->Q * storage_bytes(storage) + sizeof(uint16_t) +
           (2 * model->averages.size() + model->params.size()) * sizeof(double);
}

void Lattice::SetSetting(int id, double v) {
    if (id < 0 || id >= (int)settings_.size()) throw std::out_of_range("setting id");
    settings_[id] = v;
    params_version_++;
    const SettingDef& s = model->settings[id];
    if (s.derived >= 0) SetSetting(s.derived, derived_value(s, v));
}

void Lattice::SetZoneSetting(int id, int zone, double v) {
    if (id < 0 || id >= zcols_) throw std::out_of_range("zonal setting id");
    if (zone < -1 || zone >= nzones_) throw std::out_of_range("zone");
    if (zone == -1) {
        for (int z = 0; z < nzones_; z++) zone_[(size_t)z * zcols_ + id] = v;
    } else {
        zone_[(size_t)zone * zcols_ + id] = v;
    }
    zone_dirty_ = true;
}

double Lattice::GetZoneSetting(int id, int zone) const {
    if (id < 0 || id >= zcols_ || zone < 0 || zone >= nzones_) throw std::out_of_range("zonal setting");
    return zone_[(size_t)zone * zcols_ + id];
}

void Lattice::upload_zone() {
    if (!zone_dirty_) return;
    if (!zone_.empty())
        hip_check(hipMemcpyAsync(d_zone_, zone_.data(), zone_.size() * sizeof(double), hipMemcpyHostToDevice,
                                 stream_), "upload zone table");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    zone_dirty_ = false;
}

static bool region_in(const Region& r, int nx, int ny, int nz) {
    return r.dx >= 0 && r.dy >= 0 && r.dz >= 0 && r.nx >= 0 && r.ny >= 0 && r.nz >= 0 &&
           r.dx + r.nx <= nx && r.dy + r.ny <= ny && r.dz + r.nz <= nz;
}

void Lattice::FlagOverwrite(const uint16_t* flags, const Region& over) {
    if (!region_in(over, nx, ny, nz)) throw std::out_of_range("flag region outside the lattice");
    flags_version_++;
    if (over.dx == 0 && over.dy == 0 && over.dz == 0 && over.nx == nx && over.ny == ny && over.nz == nz) {
        hip_check(hipMemcpy(d_flags_, flags, N * sizeof(uint16_t), hipMemcpyHostToDevice), "FlagOverwrite");
        return;
    }
    for (int z = 0; z < over.nz; z++)
        for (int y = 0; y < over.ny; y++) {
            const size_t dst = (size_t)over.dx + (size_t)nx * ((size_t)(over.dy + y) + (size_t)ny * (over.dz + z));
            const size_t src = (size_t)over.nx * ((size_t)y + (size_t)over.ny * z);
            hip_check(hipMemcpy(d_flags_ + dst, flags + src, over.nx * sizeof(uint16_t), hipMemcpyHostToDevice),
                      "FlagOverwrite");
        }
}

void Lattice::GetFlags(const Region& over, uint16_t* out) {
    if (!region_in(over, nx, ny, nz)) throw std::out_of_range("flag region outside the lattice");
    for (int z = 0; z < over.nz; z++)
        for (int y = 0; y < over.ny; y++) {
            const size_t src = (size_t)over.dx + (size_t)nx * ((size_t)(over.dy + y) + (size_t)ny * (over.dz + z));
            const size_t dst = (size_t)over.nx * ((size_t)y + (size_t)over.ny * z);
            hip_check(hipMemcpy(out + dst, d_flags_ + src, over.nx * sizeof(uint16_t), hipMemcpyDeviceToHost),
                      "GetFlags");
        }
}

void Lattice::CutsOverwrite(const uint16_t* cuts, const Region& over) {
    if (!region_in(over, nx, ny, nz)) throw std::out_of_range("cut region outside the lattice");
    if (d_cuts_ == nullptr) {
        hip_check(hipMalloc((void**)&d_cuts_, 26 * N * sizeof(uint16_t)), "hipMalloc(cuts)");
        hip_check(hipMemset(d_cuts_, 0, 26 * N * sizeof(uint16_t)), "hipMemset(cuts)");
        params_version_++;  // a captured step loop holds the old (null) pointer
    }
    const size_t osize = (size_t)over.nx * over.ny * over.nz;
    if (osize == N) {  // the whole lattice: one copy
        hip_check(hipMemcpy(d_cuts_, cuts, 26 * N * sizeof(uint16_t), hipMemcpyHostToDevice), "CutsOverwrite");
        return;
    }
    for (int d = 0; d < 26; d++)
        for (int z = 0; z < over.nz; z++)
            for (int y = 0; y < over.ny; y++) {
                const size_t dst = (size_t)d * N + (size_t)over.dx +
                                   (size_t)nx * ((size_t)(over.dy + y) + (size_t)ny * (over.dz + z));
                const size_t src = (size_t)d * osize + (size_t)over.nx * ((size_t)y + (size_t)over.ny * z);
                hip_check(hipMemcpy(d_cuts_ + dst, cuts + src, over.nx * sizeof(uint16_t), hipMemcpyHostToDevice),
                          "CutsOverwrite");
            }
}

void Lattice::GetCuts(const Region& over, uint16_t* out) {
    if (!region_in(over, nx, ny, nz)) throw std::out_of_range("cut region outside the lattice");
    const size_t osize = (size_t)over.nx * over.ny * over.nz;
    if (d_cuts_ == nullptr) {  // no cuts: every link reads NO_CUT (src/LatticeAccess.inc.cpp.Rt:279-283)
        std::fill(out, out + 26 * osize, (uint16_t)65535);
        return;
    }
    if (osize == N) {
        hip_check(hipMemcpy(out, d_cuts_, 26 * N * sizeof(uint16_t), hipMemcpyDeviceToHost), "GetCuts");
        return;
    }
    for (int d = 0; d < 26; d++)
        for (int z = 0; z < over.nz; z++)
            for (int y = 0; y < over.ny; y++) {
                const size_t src = (size_t)d * N + (size_t)over.dx +
                                   (size_t)nx * ((size_t)(over.dy + y) + (size_t)ny * (over.dz + z));
                const size_t dst = (size_t)d * osize + (size_t)over.nx * ((size_t)y + (size_t)over.ny * z);
                hip_check(hipMemcpy(out + dst, d_cuts_ + src, over.nx * sizeof(uint16_t), hipMemcpyDeviceToHost),
                          "GetCuts");
            }
}

StepArgs Lattice::args(int from, int to, bool glob) const {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.fin = snaps_[from];
    a.fout = snaps_[to];
    a.flags = d_flags_;
    a.cuts = d_cuts_;
    a.extra = d_extra_;
    a.avg_n = (double)(iter_ - reset_iter_);
    a.zone = d_zone_;
    a.gpart = glob ? d_gpart_ : nullptr;
    a.pitch = (long long)pitch;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.bpr = bpr_;
    a.rpb = rpb_;
    a.rpt = rpt_;
    a.flat = 0;
    a.inv_nx = 1.0 / nx;
    a.inv_ny = 1.0 / ny;
    a.zcols = zcols_;
    a.nzones = nzones_;
    const NodeTypeDef* ib = model->nodetype("IB");
    a.ib_mask = ib ? ib->value : 0u;
    a.mode = 0;
    a.class_mask = class_mask_;
    a.bulk_val = bulk_val_;
    a.list = d_list_;
    a.list_n = (unsigned)list_n_;
    a.sr = sr_ ? 1 : 0;
    a.sr_key = 0;
    model_params(model, settings_, a.par, a.fpar);
    return a;
}

void Lattice::set_stochastic_rounding(bool on, unsigned seed) {
    if (on && storage != ST_F16S && storage != ST_BF16S)
        throw std::invalid_argument(std::string("stochastic rounding applies to the 16-bit formats f16s and bf16s, not ") +
                                    storage_name(storage));
    sr_ = on;
    sr_seed_ = seed;
    drop_graph();
}

bool temporal_blocking_model(int id) {
    return id == MODEL_D2Q9_SRT || id == MODEL_D2Q9_SRT_DRAG || id == MODEL_D2Q9 || id == MODEL_D2Q9_DRAG;
}

void Lattice::set_temporal_blocking(int k) {
    if (k < -1 || k > TB_TILE / 2 - 1)
        throw std::invalid_argument("temporal blocking takes -1 (timed) or 0-" + std::to_string(TB_TILE / 2 - 1) +
                                    " steps per launch");
    if ((k > 1 || k == -1) && !temporal_blocking_model(model->id))
        throw std::invalid_argument("temporal blocking applies to d2q9_SRT, d2q9_SRT_drag, d2q9 and d2q9_drag, not " +
                                    model->name);
    tb_auto_ = k == -1;
    tb_k_ = k == -1 ? 0 : k;
    drop_graph();
}

// Chooses the temporal blocking by timing real steps: the plain steps
// [i, i + 270) of the current Iterate run in three rounds of 30 steps per
// candidate (single steps, 3 and 5 per launch), each candidate's best
// round counting, and the fastest per step is kept; returns the index of
// the next step.  Every candidate gives the same bits, so the steps count
// as the run's own.  (Trial steps that leave the state as it is, reading
// one snapshot over and over, keep their input in the L2 and time the
// single step at two thirds of its real cost on small lattices.)  Blocking
// wins where a lattice is too small for one step to fill the GPU or its
// snapshots sit in the Infinity Cache, and loses on lattices whose steps
// stream from HBM (the halo is computed again, at lower occupancy); which
// k is best depends on the storage format (the FP64 tile allows one block
// per CU) and on how the tiles divide over the CUs.
int Lattice::tune_blocking(int i) {
    tb_auto_ = false;
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    hip_check(hipEventCreate(&ev.e[0]), "event create");
    hip_check(hipEventCreate(&ev.e[1]), "event create");
    const int ks[3] = {0, 3, 5};
    float best[3] = {0, 0, 0};
    for (int round = 0; round < TB_TUNE_ROUNDS; round++) {
        for (int c = 0; c < 3; c++) {
            const int k = ks[c];
            tb_k_ = k;
            hip_check(hipEventRecord(ev.e[0], stream_), "event record");
            for (int done = 0; done < TB_TUNE_SAMPLE;) {
                if (k > 1)
                    launch_blocked(snap_, 1 - snap_, iter_ + i);
                else
                    launch(snap_, 1 - snap_, false, iter_ + i);
                snap_ = 1 - snap_;
                i += k > 1 ? k : 1;
                done += k > 1 ? k : 1;
            }
            hip_check(hipEventRecord(ev.e[1], stream_), "event record");
            hip_check(hipEventSynchronize(ev.e[1]), "tune blocking");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "event elapsed");
            const float per_step = ms / (float)TB_TUNE_SAMPLE;
            if (round == 0 || per_step < best[c]) best[c] = per_step;
        }
    }
    int bi = 0;
    for (int c = 1; c < 3; c++)
        if (best[c] < best[bi]) bi = c;
    tb_k_ = ks[bi];
    if (const char* e = std::getenv("TCLBX_TUNE_LOG"); e && std::atoi(e) != 0)
        std::fprintf(stderr, "tclbx: %s %dx%d %s, us per step: single %.3f, 3 per launch %.3f, 5 per launch %.3f -> %d\n",
                     model->name.c_str(), nx, ny, storage_name(storage), 1e3 * best[0], 1e3 * best[1], 1e3 * best[2],
                     tb_k_);
    drop_graph();
    return i;
}

// steps of one replay of the captured loop: graph_steps_, or with temporal
// blocking an even number (at least two) of blocked launches
int Lattice::graph_len() const {
    const int k = blocked_k();
    if (k == 1) return graph_steps_;
    return std::max(2, (graph_steps_ / k) & ~1) * k;
}

// blocked_k() plain steps in one launch, snapshot `from` -> `to`; the first
// computes iteration `step` (the key of stochastic rounding)
void Lattice::launch_blocked(int from, int to, int step) {
    StepArgs a = args(from, to, false);
    a.mode = 3;
    a.tb_k = tb_k_;
    a.sr_step = (unsigned)step;
    a.sr_seed = sr_seed_;
    hip_check(launch_step(model->id, storage, false, a, block_, stream_), "temporally blocked step");
}

void Lattice::set_sorted_dispatch(bool on) {
    if (on && !model->nodetype("MRT") && !model->nodetype("BGK"))
        throw std::invalid_argument("model " + model->name + " has no plain collision NodeType to sort around");
    sorted_ = on;
    drop_graph();
}

// The nodes outside the bulk class, sorted by class and, within a class, by
// index: a counting sort of the flags read back from the device.
void Lattice::ensure_list() {
    if (!sorted_ || list_version_ == flags_version_) return;
    std::vector<uint16_t> fl(N);
    hip_check(hipMemcpy(fl.data(), d_flags_, N * sizeof(uint16_t), hipMemcpyDeviceToHost), "read flags");
    std::vector<size_t> start((size_t)class_mask_ + 2, 0);
    for (size_t n = 0; n < N; n++) {
        const unsigned c = fl[n] & class_mask_;
        if (c != bulk_val_) start[c + 1]++;
    }
    for (size_t c = 1; c < start.size(); c++) start[c] += start[c - 1];
    const size_t count = start.back();
    std::vector<unsigned> list(std::max<size_t>(count, 1));
    for (size_t n = 0; n < N; n++) {
        const unsigned c = fl[n] & class_mask_;
        if (c != bulk_val_) list[start[c]++] = (unsigned)n;
    }
    if (count > list_cap_) {
        if (d_list_) (void)hipFree(d_list_);
        d_list_ = nullptr;
        list_cap_ = 0;
        hip_check(hipMalloc((void**)&d_list_, count * sizeof(unsigned)), "hipMalloc(node list)");
        list_cap_ = count;
    }
    if (count > 0)
        hip_check(hipMemcpy(d_list_, list.data(), count * sizeof(unsigned), hipMemcpyHostToDevice), "upload node list");
    list_n_ = count;
    list_version_ = flags_version_;
    params_version_++;  // a captured step loop holds the old list
}

void model_params(const ModelDef* model, const std::vector<double>& settings, double* par, float* fpar) {
    auto S = [&](const char* n) { int i = model->setting_id(n); return i >= 0 ? settings[i] : 0.0; };
    for (int i = 0; i < MAX_PAR; i++) par[i] = 0.0;
    for (int i = 0; i < MAX_PAR; i++) fpar[i] = 0.0f;
    if (model->id == MODEL_D3Q19 || model->id == MODEL_D3Q19_LID) {
        const char* names[] = {"omega", "InletVelocity", "InletDensity", "ForceX", "ForceY", "ForceZ"};
        for (int i = 0; i < 6; i++) par[i] = S(names[i]);
        // relaxation factors of CollisionMRT (models/flow/d3q19/Dynamics.c.Rt:
        // 246-252): gamma1 = 1 - omega on e, eps, nu, pi moments, gamma2 =
        // 1 - 8(2-omega)/(8-omega) on q and m moments; a positive MRT_S*
        // setting gives its group 1 - S instead
        const double omega = par[0];
        const double gamma1 = 1 - omega;
        const double gamma2 = 1 - 8.0 * (2.0 - omega) / (8.0 - omega);
        auto rate = [&](const char* n, double dflt) { double v = S(n); return v > 0 ? 1 - v : dflt; };
        par[6] = rate("MRT_Se", gamma1);
        par[7] = rate("MRT_Seps", gamma1);
        par[8] = rate("MRT_Sq", gamma2);
        par[9] = rate("MRT_Spi", gamma1);
        par[10] = rate("MRT_Sm", gamma2);
        par[11] = gamma1;
        par[12] = S("MovingWallVelocityX");
        par[13] = S("MovingWallVelocityY");
        par[14] = S("MovingWallVelocityZ");
        // the float collision's factors, each divided by its moment's
        // colSums(M^2) (2394, 252, 40, 36, 72, 12, 24, 4, 8), and the force
        const double scaled[9] = {par[6] / 2394, par[7] / 252, par[8] / 40, gamma1 / 36, par[9] / 72,
                                  gamma1 / 12, par[9] / 24, gamma1 / 4, par[10] / 8};
        for (int i = 0; i < 9; i++) fpar[i] = (float)scaled[i];
        for (int i = 0; i < 3; i++) fpar[9 + i] = (float)par[3 + i];
    } else if (model->name.compare(0, 14, "d3q27_cumulant") == 0) {
        // [nu, nubuffer, GalileanCorrection, ForceX, ForceY, ForceZ, Omega, Smag]
        const char* names[] = {"nu", "nubuffer", "GalileanCorrection", "ForceX", "ForceY", "ForceZ", "Omega", "Smag"};
        for (int i = 0; i < 8; i++) par[i] = S(names[i]);
    } else if (model->id == MODEL_D3Q27) {
        // [omega, Smag, ForceX, ForceY, ForceZ]
        const char* names[] = {"omega", "Smag", "ForceX", "ForceY", "ForceZ"};
        for (int i = 0; i < 5; i++) par[i] = S(names[i]);
    } else if (model->id == MODEL_D2Q9 || model->id == MODEL_D2Q9_DRAG || model->id == MODEL_D2Q9_AUTOSYM ||
               model->id == MODEL_D2Q9_AUTOSYM2 || model->id == MODEL_D2Q9_BC || model->id == MODEL_D2Q9_BC_AUTOSYM ||
               model->id == MODEL_D2Q9_BC_AUTOSYM2) {
        // [S2, S3, S4, GravitationX, GravitationY]: the relaxation factors of
        // the 2nd, 3rd and 4th order moments and the body force
        par[0] = S("S2");
        par[1] = S("S3");
        par[2] = S("S4");
        par[3] = S("GravitationX");
        par[4] = S("GravitationY");
    } else {
        par[0] = S("omega");
        par[1] = S("Density");
    }
}

int globals_rows_per_block(long long rows, int bpr) {
    // enough blocks to fill the GPU several times over, each reducing
    // several rows at once
    return (int)std::max(1LL, std::min(16LL, rows * bpr / 8192));
}

void Lattice::launch(int from, int to, bool glob, int step) {
    StepArgs a = args(from, to, glob);
    a.sr_key = sr_key((unsigned)step, sr_seed_);
    if (sorted_ && !glob) {
        a.mode = 1;
        hip_check(launch_step(model->id, storage, false, a, block_, stream_), "step kernel (bulk)");
        a.mode = 2;
        hip_check(launch_step(model->id, storage, false, a, block_, stream_), "step kernel (sorted list)");
        return;
    }
    hip_check(launch_step(model->id, storage, glob, a, block_, stream_), "step kernel");
    if (glob) {
        hip_check(launch_reduce_globals(d_gpart_, step_blocks(a, true, block_), kernel_nglobals(model->id),
                                        model->sum_globals, d_glob_, stream_), "globals reduction");
    }
}

void Lattice::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_version_ = ~0u;
}

// replay a captured loop of graph_steps_ plain steps (0->1->0 ...); the
// kernel arguments are baked in, so a settings change re-captures it
void Lattice::run_graph_steps(int nsteps) {
    if (graph_exec_ == nullptr || graph_version_ != params_version_) {
        drop_graph();
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
        const int k = blocked_k();
        for (int i = 0; i < graph_len() / k; i++) {
            StepArgs a = args(i % 2, (i + 1) % 2, false);
            hipError_t e = hipSuccess;
            if (k > 1) {
                a.mode = 3;
                a.tb_k = k;
                e = launch_step(model->id, storage, false, a, block_, stream_);
            } else if (sorted_) {
                a.mode = 1;
                e = launch_step(model->id, storage, false, a, block_, stream_);
                a.mode = 2;
                if (e == hipSuccess) e = launch_step(model->id, storage, false, a, block_, stream_);
            } else {
                e = launch_step(model->id, storage, false, a, block_, stream_);
            }
            if (e != hipSuccess) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream_, &g);
                if (g) (void)hipGraphDestroy(g);
                hip_check(e, "step kernel (capture)");
            }
        }
        hip_check(hipStreamEndCapture(stream_, &graph_), "end capture");
        hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "graph instantiate");
        graph_version_ = params_version_;
    }
    for (int done = 0; done < nsteps; done += graph_len())
        hip_check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
}

void Lattice::Init() {
    upload_zone();
    iter_ = 0;
    snap_ = 0;
    StepArgs a = args(1, 0, false);
    hip_check(launch_init(model->id, storage, a, block_, stream_), "init kernel");
    // the averaged densities start from their default, 0 (src/conf.R:70),
    // the init kernel sets the parameter densities; reset_iter is left as
    // it is, as Init leaves it there
    if (!model->averages.empty())
        hip_check(hipMemsetAsync(d_extra_, 0, model->averages.size() * N * sizeof(double), stream_), "Init(averages)");
    hip_check(hipStreamSynchronize(stream_), "Init");
    clearGlobals();
}

void Lattice::resetAverage() {
    reset_iter_ = iter_;
    if (!model->averages.empty()) {
        hip_check(hipMemsetAsync(d_extra_, 0, model->averages.size() * N * sizeof(double), stream_), "resetAverage");
        hip_check(hipStreamSynchronize(stream_), "resetAverage");
    }
}

void Lattice::set_rows_per_tile(int r) {
    if (r != 0 && r != 1 && r != 2 && r != 4) throw std::invalid_argument("rows per tile must be 0, 1, 2 or 4");
    tune_rpt_ = r == 0;
    rpt_ = r == 0 ? 1 : r;
    drop_graph();
}

// Times plain steps with 1, 2 and 4 rows per tile and keeps the fastest.
// Every trial step reads the current snapshot and writes the other one,
// which the next real step overwrites: the state does not change.  Which
// is fastest depends on the model, the storage format and the occupancy
// the kernel reaches (one row for FP64 d3q19, more for the 16-bit formats
// of heavier node types).
void Lattice::tune_rows() {
    tune_rpt_ = false;
    float best = 0;
    int best_r = rpt_;
    for (int r : {1, 2, 4}) {
        if (r > 1 && (long long)ny * nz < 4LL * r) continue;
        rpt_ = r;
        launch(snap_, 1 - snap_, false);  // warm-up
        hip_check(hipEventRecord(ev0_, stream_), "event record");
        for (int k = 0; k < 3; k++) launch(snap_, 1 - snap_, false);
        hip_check(hipEventRecord(ev1_, stream_), "event record");
        hip_check(hipStreamSynchronize(stream_), "tune rows");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
        if (best == 0 || ms < best) { best = ms; best_r = r; }
    }
    rpt_ = best_r;
    drop_graph();
}

void Lattice::Prepare() {
    upload_zone();
    ensure_list();
    if (tune_rpt_ && N >= (1u << 20)) tune_rows();
    if (graph_steps_ >= 2 && !sr_) {
        run_graph_steps(0);  // capture only
        hip_check(hipStreamSynchronize(stream_), "Prepare");
    }
}

void Lattice::Iterate(int niter, int iter_type) {
    if (niter < 0) throw std::invalid_argument("negative iteration count");
    upload_zone();
    ensure_list();
    // lattices large enough for the tiling to matter are tuned once, on
    // the first Iterate that runs enough steps to repay the trial launches
    // (or earlier, by Prepare)
    if (tune_rpt_ && niter >= 16 && N >= (1u << 20)) tune_rows();
    int last_glob = iter_type & ITER_LASTGLOB;
    if (iter_type & ITER_INTEG) last_glob = 0;
    const bool every_glob = (iter_type & ITER_GLOBS) != 0;
    hip_check(hipEventRecord(ev0_, stream_), "event record");
    int i = 0;
    // leading plain steps until snapshot 0 is current, then whole graph
    // replays, then the tail (the Globals step is always launched directly);
    // a loop not yet captured is captured only when it will be replayed at
    // least twice
    const int plain_end = niter - (last_glob ? 1 : 0);
    // the first Iterate with enough plain steps times the temporal blocking
    // on them
    if (blocking_to_tune() && !every_glob && plain_end >= TB_TUNE_STEPS) i = tune_blocking(i);
    const bool captured = graph_exec_ != nullptr && graph_version_ == params_version_;
    const int k = every_glob ? 1 : blocked_k();
    const int lead = snap_ != 0 ? k : 0;
    const int glen = graph_len();
    if (!every_glob && !sr_ && graph_steps_ >= 2 && plain_end - i - lead >= (captured ? 1 : 2) * glen) {
        if (snap_ != 0) {
            if (k > 1)
                launch_blocked(snap_, 1 - snap_, iter_ + i);
            else
                launch(snap_, 1 - snap_, false, iter_ + i);
            snap_ = 1 - snap_;
            i += k;
        }
        const int reps = (plain_end - i) / glen;
        if (reps > 0) {
            run_graph_steps(reps * glen);
            i += reps * glen;
        }
    }
    // the remaining plain steps: blocked launches while k are left
    for (; k > 1 && plain_end - i >= k; i += k) {
        launch_blocked(snap_, 1 - snap_, iter_ + i);
        snap_ = 1 - snap_;
    }
    for (; i < niter; i++) {
        const bool glob = every_glob || (last_glob && i == niter - 1);
        if (last_glob && i == niter - 1)
            hip_check(hipMemsetAsync(d_glob_, 0, MAX_GLOBALS * sizeof(double), stream_), "clear globals");
        launch(snap_, 1 - snap_, glob, iter_ + i);
        snap_ = 1 - snap_;
    }
    hip_check(hipEventRecord(ev1_, stream_), "event record");
    hip_check(hipStreamSynchronize(stream_), "Iterate");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
    last_ms_ = ms;
    kernel_ms_ += ms;
    iter_ += niter;
    if (last_glob) {
        std::fill(globals_.begin(), globals_.end(), 0.0);
        calcGlobals();
    } else if (iter_type & ITER_INTEG) {
        calcGlobals();
    }
}

void Lattice::clearGlobals() { std::fill(globals_.begin(), globals_.end(), 0.0); }

// src/Lattice.cu.Rt:1299-1315: globals += device values, device cleared.
// (The non-zonal InObj objective term there is empty for these models:
// every InObj weight is zonal and already added on the device.)
void Lattice::calcGlobals() {
    std::vector<double> tab(globals_.size());
    hip_check(hipMemcpyAsync(tab.data(), d_glob_, tab.size() * sizeof(double), hipMemcpyDeviceToHost, stream_),
              "read globals");
    hip_check(hipMemsetAsync(d_glob_, 0, MAX_GLOBALS * sizeof(double), stream_), "clear globals");
    hip_check(hipStreamSynchronize(stream_), "calcGlobals");
    for (size_t i = 0; i < tab.size(); i++) globals_[i] += tab[i];
}

void Lattice::GetQuantity(int q, const Region& over, double* out, double scale) {
    if (q < 0 || q >= (int)model->quantities.size()) throw std::out_of_range("quantity id");
    if (!region_in(over, nx, ny, nz)) throw std::out_of_range("quantity region outside the lattice");
    const size_t n = (size_t)over.nx * over.ny * over.nz;
    if (n == 0) return;
    const size_t count = n * (model->quantities[q].vector ? 3 : 1);
    double* d_out = nullptr;
    hip_check(hipMalloc((void**)&d_out, count * sizeof(double)), "hipMalloc(quantity)");
    try {
        GetQuantityStaged(q, over, out, scale, d_out, count);
    } catch (...) {
        (void)hipFree(d_out);
        throw;
    }
    hip_check(hipFree(d_out), "hipFree(quantity)");
}

void Lattice::GetQuantityStaged(int q, const Region& over, double* out, double scale, double* staging,
                                size_t capacity) {
    if (q < 0 || q >= (int)model->quantities.size()) throw std::out_of_range("quantity id");
    if (!region_in(over, nx, ny, nz)) throw std::out_of_range("quantity region outside the lattice");
    const size_t n = (size_t)over.nx * over.ny * over.nz;
    if (n == 0) return;
    const size_t count = n * (model->quantities[q].vector ? 3 : 1);
    if (count > capacity)
        throw std::invalid_argument("quantity region of " + std::to_string(n) + " nodes needs " +
                                    std::to_string(count) + " doubles of staging; " + std::to_string(capacity) +
                                    " are available");
    upload_zone();
    StepArgs a = args(snap_, 1 - snap_, false);
    hipError_t e = launch_quantity(model->id, storage, q, a, over.dx, over.dy, over.dz, over.nx, over.ny,
                                   over.nz, scale, staging, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(out, staging, count * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    hip_check(e, "GetQuantity");
}

void Lattice::GetDensity(int q, double* out) {
    if (q >= model->Q && q < model->Q + model->nextra()) {
        hip_check(hipMemcpy(out, d_extra_ + (size_t)(q - model->Q) * N, N * sizeof(double), hipMemcpyDeviceToHost),
                  "GetDensity(extra)");
        return;
    }
    if (q < 0 || q >= model->Q) throw std::out_of_range("density id");
    double* d = nullptr;
    hip_check(hipMalloc((void**)&d, N * sizeof(double)), "hipMalloc(density)");
    hipError_t e = launch_get_density(model->id, storage, snaps_[snap_], (long long)pitch, (long long)N, q, d,
                                      stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, N * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(d);
    hip_check(e, "GetDensity");
}

void Lattice::SetDensity(int q, const double* in) {
    if (q >= model->Q && q < model->Q + model->nextra()) {
        hip_check(hipMemcpy(d_extra_ + (size_t)(q - model->Q) * N, in, N * sizeof(double), hipMemcpyHostToDevice),
                  "SetDensity(extra)");
        return;
    }
    if (q < 0 || q >= model->Q) throw std::out_of_range("density id");
    double* d = nullptr;
    hip_check(hipMalloc((void**)&d, N * sizeof(double)), "hipMalloc(density)");
    hipError_t e = hipMemcpyAsync(d, in, N * sizeof(double), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess)
        e = launch_set_density(model->id, storage, snaps_[snap_], (long long)pitch, (long long)N, q, d, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(d);
    hip_check(e, "SetDensity");
}

// A checkpoint in the engine's layout is the raw dump of the current
// snapshot, density planes in order, in the storage format (the reference's
// saveSolution writes its FTabs raw, no header: src/Lattice.cu.Rt:708-737),
// then the extra densities in FP64, then an 80-byte footer naming the model,
// storage format and size.  A reader that takes only the bytes it expects
// (the reference's load reads block by block, :740-769) ignores the footer.
// In the reference's layout it is the snapshot's margin blocks (pri.h),
// extra densities included as stored values, no footer.  Load tells them
// apart by the footer and then by the size, before it writes anything to
// the device.
namespace {
struct CheckpointFooter {
    char magic[8];    // "TCLBXCK1"
    char model[40];   // NUL-terminated
    char storage[8];
    int32_t nx, ny, nz, q, nextra, reserved;
};
static_assert(sizeof(CheckpointFooter) == 80, "checkpoint footer is 80 bytes");
constexpr char kMagic[8] = {'T', 'C', 'L', 'B', 'X', 'C', 'K', '1'};
}  // namespace

size_t Lattice::checkpoint_data_bytes() const {
    return (size_t)model->Q * N * storage_bytes(storage) + (size_t)model->nextra() * N * sizeof(double);
}

int Lattice::case_layout() const {
    return pri_exact(model, storage, nullptr) ? LAYOUT_REFERENCE : LAYOUT_ENGINE;
}

int Lattice::Save(const std::string& path, int layout) {
    if (layout != LAYOUT_ENGINE && layout != LAYOUT_REFERENCE) throw std::invalid_argument("unknown checkpoint layout");
    if (layout == LAYOUT_REFERENCE) {
        std::string why;
        if (!pri_exact(model, storage, &why))
            throw std::invalid_argument("no reference layout for " + model->name + " in " + storage_name(storage) +
                                        ": " + why);
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return -1;
    struct Closer { FILE* f; ~Closer() { if (f) std::fclose(f); } } closer{f};
    const size_t S = storage_bytes(storage);
    int rc = 0;
    if (layout == LAYOUT_REFERENCE) {
        std::vector<char> planes((size_t)model->Q * N * S);
        for (int q = 0; q < model->Q; q++)
            hip_check(hipMemcpy(planes.data() + (size_t)q * N * S, (char*)snaps_[snap_] + (size_t)q * pitch * S, N * S,
                                hipMemcpyDeviceToHost), "Save");
        std::vector<double> extra((size_t)model->nextra() * N);
        if (!extra.empty())
            hip_check(hipMemcpy(extra.data(), d_extra_, extra.size() * sizeof(double), hipMemcpyDeviceToHost), "Save");
        std::vector<char> out(pri_bytes(model, storage, nx, ny, nz));
        pri_pack(model, storage, nx, ny, nz, planes.data(), N, extra.data(), out.data());
        if (std::fwrite(out.data(), 1, out.size(), f) != out.size()) rc = -1;
    } else {
        std::vector<char> host(N * S);
        for (int q = 0; q < model->Q && rc == 0; q++) {
            hip_check(hipMemcpy(host.data(), (char*)snaps_[snap_] + (size_t)q * pitch * S, N * S,
                                hipMemcpyDeviceToHost), "Save");
            if (std::fwrite(host.data(), 1, host.size(), f) != host.size()) rc = -1;
        }
        // then the extra densities (FP64), as the reference's saveSolution
        // writes every density of the snapshot
        std::vector<double> av(N);
        for (size_t k = 0; k < (size_t)model->nextra() && rc == 0; k++) {
            hip_check(hipMemcpy(av.data(), d_extra_ + k * N, N * sizeof(double), hipMemcpyDeviceToHost), "Save");
            if (std::fwrite(av.data(), sizeof(double), N, f) != N) rc = -1;
        }
        if (rc == 0) {
            CheckpointFooter ft;
            std::memset(&ft, 0, sizeof(ft));
            std::memcpy(ft.magic, kMagic, sizeof(kMagic));
            std::strncpy(ft.model, model->name.c_str(), sizeof(ft.model) - 1);
            std::strncpy(ft.storage, storage_name(storage), sizeof(ft.storage) - 1);
            ft.nx = nx; ft.ny = ny; ft.nz = nz;
            ft.q = model->Q;
            ft.nextra = model->nextra();
            if (std::fwrite(&ft, sizeof(ft), 1, f) != 1) rc = -1;
        }
    }
    closer.f = nullptr;
    if (std::fclose(f) != 0) rc = -1;
    return rc;
}

int Lattice::Load(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return -1;
    struct Closer { FILE* f; ~Closer() { std::fclose(f); } } closer{f};
    const size_t data = checkpoint_data_bytes();
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long long size = std::ftell(f);
    if (size < 0) return -1;
    const std::string what = "checkpoint " + path + ": ";
    const std::string mine = model->name + " " + storage_name(storage) + " " + std::to_string(nx) + "x" +
                             std::to_string(ny) + "x" + std::to_string(nz);
    const bool ref_ok = pri_exact(model, storage, nullptr);
    const size_t ref_bytes = ref_ok ? pri_bytes(model, storage, nx, ny, nz) : 0;
    bool footer = false;
    if ((size_t)size == data + sizeof(CheckpointFooter)) {
        CheckpointFooter ft;
        if (std::fseek(f, (long)data, SEEK_SET) != 0 || std::fread(&ft, sizeof(ft), 1, f) != 1) return -1;
        footer = std::memcmp(ft.magic, kMagic, sizeof(kMagic)) == 0;
        if (footer) {
            ft.model[sizeof(ft.model) - 1] = 0;
            ft.storage[sizeof(ft.storage) - 1] = 0;
            const std::string theirs = std::string(ft.model) + " " + ft.storage + " " + std::to_string(ft.nx) + "x" +
                                       std::to_string(ft.ny) + "x" + std::to_string(ft.nz);
            if (theirs != mine || ft.q != model->Q || ft.nextra != model->nextra())
                throw std::runtime_error(what + "holds " + theirs + ", the lattice is " + mine);
        }
    }
    const bool reference = !footer && ref_ok && (size_t)size == ref_bytes;
    if (!footer && !reference && (size_t)size != data) {
        std::string also;
        if (ref_ok) also = ", or " + std::to_string(ref_bytes) + " in the reference's layout";
        throw std::runtime_error(what + "has " + std::to_string(size) + " bytes; a checkpoint of " + mine +
                                 " has " + std::to_string(data) + " (+" + std::to_string(sizeof(CheckpointFooter)) +
                                 " of footer)" + also);
    }
    if (std::fseek(f, 0, SEEK_SET) != 0) return -1;
    const size_t S = storage_bytes(storage);
    if (reference) {
        std::vector<char> in(ref_bytes);
        if (std::fread(in.data(), 1, in.size(), f) != in.size()) return -1;
        std::vector<char> planes((size_t)model->Q * N * S);
        std::vector<double> extra((size_t)model->nextra() * N);
        pri_unpack(model, storage, nx, ny, nz, in.data(), planes.data(), N, extra.data());
        for (int q = 0; q < model->Q; q++)
            hip_check(hipMemcpy((char*)snaps_[snap_] + (size_t)q * pitch * S, planes.data() + (size_t)q * N * S, N * S,
                                hipMemcpyHostToDevice), "Load");
        if (!extra.empty())
            hip_check(hipMemcpy(d_extra_, extra.data(), extra.size() * sizeof(double), hipMemcpyHostToDevice), "Load");
        return 0;
    }
    // plane by plane, so that host memory stays at one plane; the size is
    // checked, so only an I/O error can stop it part way
    std::vector<char> host(N * S);
    for (int q = 0; q < model->Q; q++) {
        if (std::fread(host.data(), 1, host.size(), f) != host.size()) return -1;
        hip_check(hipMemcpy((char*)snaps_[snap_] + (size_t)q * pitch * S, host.data(), N * S, hipMemcpyHostToDevice),
                  "Load");
    }
    std::vector<double> av(N);
    for (size_t k = 0; k < (size_t)model->nextra(); k++) {
        if (std::fread(av.data(), sizeof(double), N, f) != N) return -1;
        hip_check(hipMemcpy(d_extra_ + k * N, av.data(), N * sizeof(double), hipMemcpyHostToDevice), "Load");
    }
    return 0;
}

}  // namespace tclbx
