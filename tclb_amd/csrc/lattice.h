// lattice.h -- one lattice (one subdomain on one GPU): device memory, the
// step loop and the Globals, with the operations of TCLB's Lattice class
// (src/Lattice.h.Rt:39-191) that the node-kernel path uses.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "kernels.h"
#include "model.h"

namespace tclbx {

// iteration types, src/Lattice.h.Rt:19-31
constexpr int ITER_STREAM = 0x000, ITER_NORM = 0x001, ITER_TYPE = 0x003, ITER_NO = 0x000,
              ITER_GLOBS = 0x010, ITER_OBJ = 0x020, ITER_INTEG = 0x070, ITER_LASTGLOB = 0x080;

struct Region {
    int dx = 0, dy = 0, dz = 0, nx = 0, ny = 0, nz = 0;
};

class Lattice {
public:
    Lattice(const ModelDef* m, int nx, int ny, int nz, int storage);
    ~Lattice();
    Lattice(const Lattice&) = delete;
    Lattice& operator=(const Lattice&) = delete;

    // settings (src/Lattice.cu.Rt:1344-1358): a derived setting also sets
    // the setting it derives (nu -> omega, InletPressure -> InletDensity)
    void SetSetting(int id, double v);
    double GetSetting(int id) const { return settings_[id]; }
    // zonal settings (src/ZoneSettings.h:61-80): zone -1 sets every zone
    void SetZoneSetting(int id, int zone, double v);
    double GetZoneSetting(int id, int zone) const;

    void FlagOverwrite(const uint16_t* flags, const Region& over);
    void GetFlags(const Region& over, uint16_t* out);
    // cut distances of the 26 links of each node, cuts[d][over node] with
    // d = direction - 1 in the d3q27 order (0, 1, -1 per axis, x fastest):
    // the first call allocates them zeroed, then copies the region
    // (src/Lattice.cu.Rt:1056-1072, src/LatticeContainer.inc.cpp.Rt:174-185)
    void CutsOverwrite(const uint16_t* cuts, const Region& over);
    void GetCuts(const Region& over, uint16_t* out);
    bool has_cuts() const { return d_cuts_ != nullptr; }

    void Init();                              // src/Lattice.cu.Rt:799-821
    // Lattice::resetAverage (src/Lattice.cu.Rt:1360-1365): the averaged
    // densities to zero, average_iter counted from this iteration
    void resetAverage();
    void Iterate(int niter, int iter_type);   // src/Lattice.cu.Rt:900-989
    // what the first long Iterate would do before its steps: time the rows
    // per tile (the state is not changed) and capture the step loop
    void Prepare();
    void clearGlobals();
    void calcGlobals();
    const std::vector<double>& globals() const { return globals_; }

    void GetQuantity(int q, const Region& over, double* out, double scale);
    // the same through `capacity` doubles of device staging the caller owns:
    // no allocation
    void GetQuantityStaged(int q, const Region& over, double* out, double scale, double* staging,
                           size_t capacity);
    // Get_<f> of the current snapshot; q >= Q: extra density q - Q (the
    // model's averages, then its parameter densities)
    void GetDensity(int q, double* out);
    void SetDensity(int q, const double* in); // Set_<f>
    // saveSolution / loadSolution (src/Lattice.cu.Rt:708-769): -1 when the
    // file cannot be opened or read; Load throws when the file holds another
    // model, storage format or size.  layout: LAYOUT_ENGINE (density planes
    // and a footer) or LAYOUT_REFERENCE (the reference's margin blocks,
    // pri.h); Load takes either, told apart by the footer and the size.
    enum { LAYOUT_ENGINE = 0, LAYOUT_REFERENCE = 1 };
    int Save(const std::string& path, int layout = LAYOUT_ENGINE);
    int Load(const std::string& path);
    // the layout the case driver's SaveBinary/SaveCheckpoint write: the
    // reference's wherever its storage build holds the same values
    int case_layout() const;
    size_t checkpoint_data_bytes() const;  // bytes of a checkpoint before its footer

    double last_iterate_ms() const { return last_ms_; }
    double kernel_ms_total() const { return kernel_ms_; }
    int iter() const { return iter_; }
    hipStream_t stream() const { return stream_; }
    size_t bytes_per_node_step() const;       // algorithmic HBM bytes
    void set_block(int b);
    int block() const { return block_; }
    void set_graph_steps(int k) { graph_steps_ = k & ~1; drop_graph(); }
    // rows of the lattice walked by one block of a plain step: 1, 2 or 4,
    // or 0 to let the first long Iterate time the three and keep the fastest
    void set_rows_per_tile(int r);
    int rows_per_tile() const { return rpt_; }
    // NodeType-sorted dispatch of the plain steps: a bulk pass over the plain
    // collision nodes, then one over the other nodes sorted by NodeType class
    // (Globals steps always take the single pass)
    void set_sorted_dispatch(bool on);
    bool sorted_dispatch() const { return sorted_; }
    size_t sorted_list_size() { ensure_list(); return list_n_; }
    // stochastic rounding of the stores of the 16-bit formats (f16s, bf16s)
    // in steps: random bits from (node, direction, iteration, seed); the
    // step loop is then launched directly, since every step's launch differs
    void set_stochastic_rounding(bool on, unsigned seed);
    bool stochastic_rounding() const { return sr_; }
    // temporal blocking of the plain steps of the 2D models d2q9_SRT,
    // d2q9_SRT_drag, d2q9 and d2q9_drag: k (2 to TB_TILE/2 - 1) steps per
    // launch on tiles held in LDS, bit-identical to k single steps; 0 or 1:
    // off; -1 (the default of these models): chosen on lattices under 2^23
    // nodes by timing the first 270 plain steps of the first Iterate that
    // runs that many (tune_blocking).  Steps with Globals take the single
    // step.
    void set_temporal_blocking(int k);
    int temporal_blocking() const { return tb_k_; }

    const ModelDef* model;
    const int nx, ny, nz, storage;
    const size_t N;
    size_t pitch;  // elements per density plane

private:
    StepArgs args(int from, int to, bool glob) const;
    void upload_zone();
    void launch(int from, int to, bool glob, int step = 0);
    void launch_blocked(int from, int to, int step = 0);
    int blocked_k() const { return tb_k_ > 1 ? tb_k_ : 1; }
    int graph_len() const;
    void drop_graph();
    void tune_rows();
    int tune_blocking(int i);
    // the timing of the temporal blocking: rounds x 3 candidates x steps
    static constexpr int TB_TUNE_ROUNDS = 3, TB_TUNE_SAMPLE = 30;
    static constexpr int TB_TUNE_STEPS = TB_TUNE_ROUNDS * 3 * TB_TUNE_SAMPLE;
    bool blocking_to_tune() const { return tb_auto_ && N < (1u << 23); }
    void run_graph_steps(int nsteps);
    void ensure_list();

    std::vector<double> settings_;
    std::vector<double> zone_;  // [zone][zcols]
    int zcols_, nzones_;
    bool zone_dirty_ = true;
    void* snaps_[2] = {nullptr, nullptr};
    int snap_ = 0, iter_ = 0;
    uint16_t* d_flags_ = nullptr;
    uint16_t* d_cuts_ = nullptr;
    double* d_extra_ = nullptr; // model->nextra() FP64 planes of N: averages, then parameters
    int reset_iter_ = 0;
    double* d_zone_ = nullptr;
    double* d_gpart_ = nullptr;
    double* d_glob_ = nullptr;
    std::vector<double> globals_;
    int nblocks_ = 0, block_ = 256, bpr_ = 1, rpb_ = 1;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double last_ms_ = 0, kernel_ms_ = 0;
    // captured step loop: graph_len() plain steps starting at snapshot 0
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    int graph_steps_ = 32;
    int rpt_ = 1;
    bool tune_rpt_ = true;
    unsigned params_version_ = 0, graph_version_ = ~0u;
    // sorted dispatch: the nodes outside the bulk class, sorted by class,
    // rebuilt from the flags when they change
    bool sorted_ = false;
    unsigned class_mask_ = 0, bulk_val_ = 0;
    unsigned* d_list_ = nullptr;
    size_t list_n_ = 0, list_cap_ = 0;
    unsigned flags_version_ = 0, list_version_ = ~0u;
    bool sr_ = false;
    unsigned sr_seed_ = 0;
    int tb_k_ = 0;
    bool tb_auto_ = false;
};

// the models the temporally blocked step covers
bool temporal_blocking_model(int model_id);
// hipError_t -> exception with a message naming the call
void hip_check(hipError_t e, const char* what);
// steps per captured graph of a new Lattice or Batch: 32, or the
// environment's TCLBX_GRAPH_STEPS rounded down to an even count (0 or 1:
// every step launched directly, for tools that do not follow graph launches)
int default_graph_steps();
// throws std::invalid_argument unless a lattice of the model in this storage
// format can be held: positive sizes, nz = 1 for a 2D model, fewer than 2^32
// nodes, and each density plane within 4 GiB (the step kernel's per-lane
// addresses are 32-bit byte offsets from the plane base)
void check_lattice_size(const ModelDef* m, int nx, int ny, int nz, int storage);
// device staging of tclbx_get_quantity_noalloc (noalloc.hip): the current
// device's copy of a file-scope array of quantity_scratch_doubles() doubles
double* quantity_scratch();
size_t quantity_scratch_doubles();
// kernel parameter vector (StepArgs::par) of a model from its settings
void model_params(const ModelDef* model, const std::vector<double>& settings, double* par, float* fpar);
// rows walked by one block of a Globals step for a lattice of `rows` rows
int globals_rows_per_block(long long rows, int bpr);

}  // namespace tclbx
