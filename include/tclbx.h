/*
 * tclbx.h -- C ABI of the MI355X lattice Boltzmann node-kernel engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (TCLBX_OK = 0, TCLBX_ERR = -1; tclbx_last_error() gives the message of the
 * calling thread's last failure).  No call exits the process (the reference
 * exits on any device error, src/cross.cu:24-32).  Calls on one lattice are
 * not re-entrant; every call is synchronous on return, like the reference's
 * Lattice methods.  Arrays are x-fastest (x + nx*(y + ny*z)); vector
 * quantities are 3 doubles per node, interleaved.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef TCLBX_H
#define TCLBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCLBX_OK 0
#define TCLBX_ERR (-1)
#define TCLBX_ITERATION_STOP 1 /* src/CommonHandler.h:10 */

/* iteration types, src/Lattice.h.Rt:19-31 */
#define TCLBX_ITER_NORM 0x001
#define TCLBX_ITER_GLOBS 0x010
#define TCLBX_ITER_LASTGLOB 0x080

typedef struct tclbx_lattice tclbx_lattice;

/* Message of the calling thread's last failed call ("" if none). */
const char* tclbx_last_error(void);
int tclbx_version(void);

/* Models compiled in: "d2q9_SRT", "d2q9_SRT_drag", "d3q19", "d3q19_lid", "d2q9",
 * "d2q9_drag", "d3q27", "d3q27_cumulant" and its option builds
 * "d3q27_cumulant_[AVG_][IB_][SMAG]" (all eight), and d2q9's option builds
 * "d2q9_[bc_]autosym[2]", "d2q9_bc".
 * Replaces the per-model binaries CLB/<model>/main (src/makefile.main.Rt:331-349). */
int tclbx_model_count(void);
const char* tclbx_model_name(int i);

/* Lattice(lbRegion, MPIInfo, ns) -- src/Lattice.cu.Rt:87-151, called from
 * Solver::InitAll (src/Solver.cpp.Rt:378-402).  storage: "f64", "f32",
 * "f32s", "f16s", "bf16s" (the reference's --with-storage, fixed at build
 * time there, src/configure.ac:213-233).  One subdomain, periodic on every
 * axis unless NodeTypes say otherwise; nx is not padded. */
int tclbx_lattice_create(const char* model, int nx, int ny, int nz, const char* storage,
                         tclbx_lattice** out);
int tclbx_lattice_destroy(tclbx_lattice* h);

/* Model_m registry lookups by model name, src/Lists.h:13-165 and
 * Things::by_name (src/Things.h:15-23): the id, or -1 if either name is
 * unknown.  They need no GPU. */
int tclbx_setting_id(const char* model, const char* name);
int tclbx_zone_setting_id(const char* model, const char* name);
int tclbx_global_id(const char* model, const char* name);
int tclbx_quantity_id(const char* model, const char* name, int* is_vector);
/* counts: Q, Globals, Settings, zonal Settings, Quantities */
int tclbx_model_info(const char* model, int* Q, int* nglobals, int* nsettings, int* nzonal,
                     int* nquantities);
/* name of entry `id` (NULL if out of range) */
const char* tclbx_global_name(const char* model, int id);
const char* tclbx_setting_name(const char* model, int id);
const char* tclbx_zone_setting_name(const char* model, int id);
const char* tclbx_quantity_name(const char* model, int id);
/* default value of setting `id` (the reference applies defaults of the
 * non-deriving settings only, src/Solver.cpp.Rt:277-280) */
int tclbx_setting_default(const char* model, int id, double* value);
/* NodeType flag value and group mask (NODE_<name>, src/conf.R:600-699).
 * Group names: "BOUNDARY", "COLLISION", "ADDITIONALS", "SETTINGZONE", "ALL". */
int tclbx_nodetype(const char* model, const char* name, uint16_t* value, uint16_t* mask);
int tclbx_nodetype_group(const char* model, const char* group, uint16_t* mask, int* shift);

/* Lattice::SetSetting (src/Lattice.cu.Rt:1350-1358): a derived setting
 * also sets its target (nu -> omega, InletPressure -> InletDensity). */
int tclbx_set_setting(tclbx_lattice* h, int id, double value);
int tclbx_get_setting(const tclbx_lattice* h, int id, double* value);
/* ZoneSettings::set (src/ZoneSettings.h:61-80); zone = -1: every zone. */
int tclbx_set_zone_setting(tclbx_lattice* h, int id, int zone, double value);

/* Lattice::FlagOverwrite / GetFlags (src/Lattice.cu.Rt:1040-1089) over the
 * box [dx,dx+nx) x [dy,dy+ny) x [dz,dz+nz); the caller owns the buffer. */
int tclbx_flag_overwrite(tclbx_lattice* h, const uint16_t* flags, int dx, int dy, int dz, int nx, int ny,
                         int nz);
int tclbx_get_flags(tclbx_lattice* h, int dx, int dy, int dz, int nx, int ny, int nz, uint16_t* out);

/* Lattice::CutsOverwrite (src/Lattice.cu.Rt:1056-1072): cut distances of
 * the 26 links of each node of the box, cuts[d*(nx*ny*nz) + node], d =
 * direction - 1 in the d3q27 order (0, 1, -1 per axis, x fastest), as
 * r * TCLBX_CUT_MAX for a crossing at the fraction r of the link or
 * TCLBX_NO_CUT.  The first call gives the lattice its cuts, zeroed
 * (LatticeContainer::ActivateCuts, src/LatticeContainer.inc.cpp.Rt:174-185);
 * without a call every link reads NO_CUT.  The IB builds of d3q27_cumulant
 * interpolate on them. */
#define TCLBX_NO_CUT 65535
#define TCLBX_CUT_MAX 65000
int tclbx_cuts_overwrite(tclbx_lattice* h, const uint16_t* cuts, int dx, int dy, int dz, int nx, int ny,
                         int nz);
/* The cuts of a box, laid out as above (NO_CUT everywhere when the lattice
 * has none).  The reference reads them only into a VTI file of the
 * geometry (src/Geometry.cpp.Rt:1200-1245); this is that read-back. */
int tclbx_get_cuts(tclbx_lattice* h, int dx, int dy, int dz, int nx, int ny, int nz, uint16_t* out);

/* Lattice::Init (src/Lattice.cu.Rt:799-821): Init action, Iter = 0,
 * Globals cleared. */
int tclbx_init(tclbx_lattice* h);
/* Lattice::Iterate(niter, iter_type) (src/Lattice.cu.Rt:900-989): niter
 * steps; with TCLBX_ITER_LASTGLOB the Globals of the last step are
 * computed and become tclbx_get_globals(); TCLBX_ITER_GLOBS accumulates
 * Globals over every step. */
int tclbx_iterate(tclbx_lattice* h, int niter, int iter_type);
/* Lattice::globals[] after calcGlobals (src/Lattice.cu.Rt:1299-1315), in
 * the model's Globals order (SUM..., Objective, MAX...). */
int tclbx_get_globals(const tclbx_lattice* h, double* out, int n);
int tclbx_get_iteration(const tclbx_lattice* h, int* iter);

/* Lattice::GetQuantity(q, region, tab, scale) (src/Lattice.cu.Rt:1184-1222):
 * evaluated on the pulled populations of the current snapshot. */
int tclbx_get_quantity(tclbx_lattice* h, int q, int dx, int dy, int dz, int nx, int ny, int nz,
                       double* out, double scale);
/* tclbx_get_quantity without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * quantity kernel writes into a device array the library reserves on each
 * GPU for the life of the process (64 MiB: at most 8,388,608 doubles, i.e.
 * that many nodes of a scalar quantity or a third of them of a vector one;
 * larger regions are refused with TCLBX_ERR), then the values are copied to
 * `out` (host memory).  Calls on one GPU share that array: calls that
 * overlap on one GPU race. */
int tclbx_get_quantity_noalloc(tclbx_lattice* h, int q, int dx, int dy, int dz, int nx, int ny, int nz,
                               double* out, double scale);
/* Get_<f> / Set_<f> (src/Lattice.cu.Rt:1133-1172): density q of the whole
 * lattice, as real values (storage shift removed / applied). */
int tclbx_get_density(tclbx_lattice* h, int q, double* out);
int tclbx_set_density(tclbx_lattice* h, int q, const double* in);
/* The densities that are not streamed (AddDensity with average=TRUE or
 * parameter=TRUE, src/conf.R:65-100): the averages of d3q27_cumulant's AVG
 * builds (models/flow/d3q27_cumulant/Dynamics.R:66-90), then the parameter
 * densities BC[0], BC[1] of d2q9's bc builds (models/flow/d2q9/Dynamics.R:
 * 23-26).  Name of extra density k (NULL past the last), *is_average = 1
 * for an average; it is density Q + k of tclbx_get/set_density, kept in FP64
 * and updated in place. */
const char* tclbx_extra_density_name(const char* model, int k, int* is_average);
/* Lattice::resetAverage (src/Lattice.cu.Rt:1360-1365), what <Average>
 * calls: the averaged densities to zero, average_iter counted from the
 * current iteration. */
int tclbx_reset_average(tclbx_lattice* h);
/* saveSolution / loadSolution (src/Lattice.cu.Rt:708-769): raw densities
 * of the current snapshot in the storage format, then the extra densities
 * (FP64), then an 80-byte footer (model, storage, size) that a reader taking
 * only the bytes it expects ignores.  TCLBX_ERR if the file cannot be opened
 * or read, or holds another model, storage format or size (checked before
 * the lattice is touched); a file without the footer is checked by size. */
int tclbx_save(tclbx_lattice* h, const char* path);
int tclbx_load(tclbx_lattice* h, const char* path);
/* saveSolution in a chosen layout: TCLBX_LAYOUT_ENGINE (what tclbx_save
 * writes) or TCLBX_LAYOUT_REFERENCE, the file the reference itself writes
 * (listTabs/save, src/Lattice.cu.Rt:642-661,708-737): the snapshot's
 * non-empty margin blocks in order (margins and their sizes from the fields'
 * reach, src/conf.R:771-1012), each field's layer as stored values,
 * extra densities included, no footer.  Only where a reference storage
 * build holds the same stored values: f64 and f32 for every model, f32s and
 * f16s (float-shift, half-shift) for the d2q9 MRT models, which declare the
 * shift w_i; TCLBX_ERR otherwise.  tclbx_load reads both layouts; the case
 * driver's SaveBinary and SaveCheckpoint write the reference's wherever it
 * is available. */
#define TCLBX_LAYOUT_ENGINE 0
#define TCLBX_LAYOUT_REFERENCE 1
int tclbx_save_layout(tclbx_lattice* h, const char* path, int layout);
/* The reference layout on the host, no GPU needed: the file's size, its
 * blocks (number 1..27 = sx + 3 sy + 9 sz + 1, side 0 below, 1 domain, 2
 * above; values per block) in file order, and the conversion from / to the
 * engine's planes (Q planes of nx*ny*nz stored values, then nextra planes
 * of doubles in `extra`).  Domain-block positions the reference's put never
 * writes are packed as zero and ignored when unpacking. */
int tclbx_pri_bytes(const char* model, const char* storage, int nx, int ny, int nz, size_t* bytes);
int tclbx_pri_blocks(const char* model, int nx, int ny, int nz, int* index, size_t* values, int max, int* count);
int tclbx_pri_pack(const char* model, const char* storage, int nx, int ny, int nz, const void* planes,
                   const double* extra, void* out);
int tclbx_pri_unpack(const char* model, const char* storage, int nx, int ny, int nz, const void* in, void* planes,
                     double* extra);

/* Event timing of the last tclbx_iterate and the running total, in ms --
 * what MainCallback measures for MLBUps (src/main.cpp:65-157). */
int tclbx_get_timing(const tclbx_lattice* h, double* last_ms, double* total_ms);
/* Algorithmic HBM bytes per node and step: 2*Q*sizeof(storage) + 2 (flag),
 * the reference's GB/s formula with the storage width (src/main.cpp:127). */
int tclbx_bytes_per_node(const tclbx_lattice* h, double* bytes);
/* Launch shape: threads per block along x (64, 128 or 256); 0 = default. */
int tclbx_set_block(tclbx_lattice* h, int block);
/* Steps per captured hipGraph replay in tclbx_iterate (0 = no graph). */
int tclbx_set_graph_steps(tclbx_lattice* h, int steps);
/* Rows walked by one block of a plain step: 1, 2 or 4; 0 (the default) times
 * the three on the first long tclbx_iterate of a lattice of 2^20 nodes or
 * more and keeps the fastest (the state is not changed).  A launch-shape
 * choice of this engine; the reference's shape is fixed (32x4 threads). */
int tclbx_set_rows_per_tile(tclbx_lattice* h, int rows);
int tclbx_get_rows_per_tile(const tclbx_lattice* h, int* rows);
/* Runs now what the first long tclbx_iterate would do before its steps: the
 * rows-per-tile timing above (trial launches that leave the state as it is)
 * and the capture of the step loop, so that a timed tclbx_iterate runs steps
 * only.  An engine call with no counterpart in the reference, whose launch
 * shape is fixed and which captures no graphs. */
int tclbx_prepare(tclbx_lattice* h);
/* NodeType-sorted dispatch of the plain steps (on = 1): a pass over the plain
 * collision nodes (MRT, or BGK in d2q9_SRT) on the usual tiles, then a pass
 * over every other node, from a list sorted by NodeType class (the NodeType
 * bits below the zone) and rebuilt when the flags change, so each wavefront
 * takes one branch of the reference's per-node switch
 * (models/flow/d3q19/Dynamics.c.Rt:181-210, models/flow/d2q9_SRT/Dynamics.c:
 * 40-67).  Results are those of the single pass, bit for bit.  The
 * environment's TCLBX_SORTED_DISPATCH=1 turns it on for new lattices.  The
 * reference's only split is border versus interior, for MPI overlap
 * (src/LatticeContainer.inc.cpp.Rt:201-287). */
int tclbx_set_sorted_dispatch(tclbx_lattice* h, int on);
/* whether it is on, and the nodes of the sorted list (0 when off) */
int tclbx_get_sorted_dispatch(tclbx_lattice* h, int* on, long long* list_nodes);
/* Stochastic rounding of the stored values in the 16-bit formats (f16s,
 * bf16s; TCLBX_ERR for the others): a value between two representable ones
 * is stored as either with the probabilities that make its expectation
 * exact, so rounding errors do not accumulate into a drift over long runs.
 * The rounding is gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32; the
 * random bits come from a hash of (node, iteration, seed) and a xorshift32
 * stream over the directions: a run is reproducible.  The step loop is then
 * launched step by step (no captured graph).  No counterpart in the reference, which rounds to
 * nearest (src/LatticeAccess.inc.cpp.Rt:14-35). */
int tclbx_set_stochastic_rounding(tclbx_lattice* h, int on, unsigned seed);
/* Temporal blocking of the plain steps of the 2D models d2q9_SRT,
 * d2q9_SRT_drag, d2q9 and d2q9_drag (TCLBX_ERR for the others): `steps`
 * (2-15) time steps per launch, each block stepping a 32 x 32 tile with
 * `steps` nodes of halo in LDS and writing its interior; 0 or 1 turns it
 * off; -1, the default of these models, lets a lattice under 2^23 nodes
 * time single steps against 3 and 5 per launch on the first 270 plain steps
 * of the first tclbx_iterate that runs that many (real steps of the run)
 * and keep the fastest.  Results
 * are those of single steps, bit for bit, stochastic rounding included.
 * Steps with Globals take the single step.  It pays where one step does
 * not fill the GPU or the snapshots sit in the Infinity Cache, and loses
 * where steps stream from HBM.  The environment's TCLBX_TEMPORAL_BLOCKING=k
 * (k >= 0) fixes it for new lattices of these models.  No counterpart in
 * the reference, which launches one kernel per iteration
 * (src/Lattice.cu.Rt:1246-1297). */
int tclbx_set_temporal_blocking(tclbx_lattice* h, int steps);
/* the steps per launch in use (0: single steps, also before a timed choice) */
int tclbx_get_temporal_blocking(tclbx_lattice* h, int* steps);

/* ---------------------------------------------------------------- batch --
 * Many independent lattices ("cases") of one model and storage format in
 * one ragged device arena, advanced together: one persistent kernel launch
 * per step walks a tile list covering every case.  No counterpart in the
 * reference (one Lattice per process there); each per-case call mirrors
 * the Lattice call of the same name above.  A case is computed exactly as
 * the same case alone (same node arithmetic, Globals reduced over the
 * case's own tiles in a fixed order). */
typedef struct tclbx_batch tclbx_batch;

/* capacity_bytes: HBM the arena holds for snapshots and flags of all cases */
int tclbx_batch_create(const char* model, const char* storage, long long capacity_bytes, tclbx_batch** out);
int tclbx_batch_destroy(tclbx_batch* b);
/* arena bytes one nx*ny*nz case needs */
int tclbx_batch_case_bytes(const tclbx_batch* b, int nx, int ny, int nz, long long* bytes);
/* *id = new case id, or -1 (and TCLBX_OK) when the arena has no room */
int tclbx_batch_add_case(tclbx_batch* b, int nx, int ny, int nz, int* id);
int tclbx_batch_remove_case(tclbx_batch* b, int id);
/* id -1: every case */
int tclbx_batch_set_setting(tclbx_batch* b, int id, int setting, double value);
int tclbx_batch_set_zone_setting(tclbx_batch* b, int id, int zonal, int zone, double value);
int tclbx_batch_flag_overwrite(tclbx_batch* b, int id, const uint16_t* flags);
int tclbx_batch_init_case(tclbx_batch* b, int id);
/* every case advances niter steps (iter_type as tclbx_iterate) */
int tclbx_batch_iterate(tclbx_batch* b, int niter, int iter_type);
int tclbx_batch_get_globals(const tclbx_batch* b, int id, double* out, int n);
int tclbx_batch_get_iteration(const tclbx_batch* b, int id, long long* iter);
int tclbx_batch_get_quantity(tclbx_batch* b, int id, int q, double* out, double scale);
int tclbx_batch_get_density(tclbx_batch* b, int id, int q, double* out);
int tclbx_batch_set_density(tclbx_batch* b, int id, int q, const double* in);
/* time every tile shape (row segments or flat runs of the node index;
 * threads per block 64/128/256; rows per tile 1/2/4), persistent grid size
 * and tile order on the current cases for `steps` steps each, without
 * changing their state, and keep the fastest */
int tclbx_batch_autotune(tclbx_batch* b, int steps, int* flat, int* block, int* rows_per_tile, int* grid,
                         double* us_per_step);
int tclbx_batch_set_tuning(tclbx_batch* b, int flat, int block, int rows_per_tile, int grid);
/* stochastic rounding of every case's stores (f16s, bf16s; TCLBX_ERR for
 * the others), each case keyed by its own node index, iteration and `seed`
 * as tclbx_set_stochastic_rounding keys a lone lattice, so a case gives the
 * same result alone and in a batch; steps are then launched one by one.  No
 * counterpart in the reference (batching is this engine's, and it rounds to
 * nearest). */
int tclbx_batch_set_stochastic_rounding(tclbx_batch* b, int on, unsigned seed);
/* the order the step walks the tiles in: 0 the tile list's, 1 XCD-
 * contiguous (the k-th of the eight XCDs takes the k-th contiguous run of
 * tiles, so neighbouring rows are pulled through one L2); the autotuner
 * times both.  set < 0 leaves it; *order gets the current one.  No
 * counterpart in the reference (batching is this engine's). */
int tclbx_batch_tile_order(tclbx_batch* b, int set, int* order);
int tclbx_batch_get_timing(const tclbx_batch* b, double* last_ms, double* total_ms);
int tclbx_batch_info(const tclbx_batch* b, int* ncases, long long* nodes, long long* arena_used,
                     long long* arena_capacity);
int tclbx_batch_set_graph_steps(tclbx_batch* b, int steps);
/* The device tables of the current cases and the captured step loop, built
 * now instead of in the next tclbx_batch_iterate. */
int tclbx_batch_prepare(tclbx_batch* b);

/* --------------------------------------------------------------- server --
 * Continuous-batching solve server: requests arrive from any thread; a
 * worker thread admits them into one batch while the HBM budget has room,
 * advances all resident requests together, and retires each on its step
 * budget or its stop rule.  The stop rule is the reference's <Stop> handler
 * (src/Handlers/cbStop.cpp:12-104): criteria on Globals (type 0
 * <G>Change, 1 <G>PercentChange, 2 <G>Above, 3 <G>Below) checked every
 * `stop_interval` steps, stopping after `stop_times` consecutive passing
 * checks.  No counterpart in the reference. */
typedef struct tclbx_server tclbx_server;

#define TCLBX_RQ_REJECTED (-1)
#define TCLBX_RQ_QUEUED 0
#define TCLBX_RQ_RUNNING 1
#define TCLBX_RQ_DONE_BUDGET 2
#define TCLBX_RQ_DONE_STOP 3

/* hbm_budget: bytes of the arena; max_chunk: most steps between admissions */
int tclbx_server_create(const char* model, const char* storage, long long hbm_budget, int max_chunk,
                        tclbx_server** out);
int tclbx_server_destroy(tclbx_server* s);
/* a request: lattice size, NodeTypes (x fastest), settings (ids and
 * values, applied in order), zonal settings (ids, zones, values), step
 * budget and stop rule (nstop = 0: budget only) */
int tclbx_server_submit(tclbx_server* s, int nx, int ny, int nz, const uint16_t* flags, int nset,
                        const int* set_ids, const double* set_values, int nzonal, const int* zonal_ids,
                        const int* zonal_zones, const double* zonal_values, long long max_steps, int nstop,
                        const int* stop_global, const int* stop_type, const double* stop_limit,
                        int stop_interval, int stop_times, int* id);
/* state (TCLBX_RQ_*), steps run, Globals at the end (n >= number of
 * Globals) and the submit/admit/finish times in seconds since creation */
int tclbx_server_poll(tclbx_server* s, int id, int* state, long long* steps, double* globals, int n,
                      double* t_submit, double* t_admit, double* t_finish);
/* 1 when nothing is queued or running, 0 on timeout */
int tclbx_server_wait_all(tclbx_server* s, double timeout_s, int* idle);
int tclbx_server_stats(tclbx_server* s, long long* completed, long long* rejected, long long* node_steps,
                       double* busy_s, int* active, int* queued, double* now_s);
/* tile shape of the server's batch (flat runs of the node index by default) */
int tclbx_server_set_tuning(tclbx_server* s, int flat, int block, int rows_per_tile, int grid);
/* stochastic rounding of the served requests' stores, as
 * tclbx_batch_set_stochastic_rounding; taken up by the worker before its next
 * admission */
int tclbx_server_set_stochastic_rounding(tclbx_server* s, int on, unsigned seed);

/* Measured device copy bandwidth: `reps` copies of `bytes` (rounded down
 * to 16) between two device buffers by a plain vector-copy kernel;
 * *gbps = 2*bytes*reps / time (read + write), *ms = time of one copy. */
int tclbx_copy_bandwidth(long long bytes, int reps, double* gbps, double* ms);

/* ---------------------------------------------------------- case files ---
 * A TCLB case file (<CLBConfig>) run on the engine: the work of the
 * reference executable CLB/<model>/main <case.xml> (src/main.cpp:171-420)
 * and of its handlers (src/Handlers/: Units, Geometry, Model, Param, Init,
 * Solve, Log, Stop, Failcheck, VTK, SaveBinary, LoadBinary).  Outputs keep
 * the reference's names, <output><case>_<Name>_P00_<iteration>.<ext>, and
 * formats (Log CSV of src/Solver.cpp.Rt:120-206, VTI of src/vtkOutput.cpp). */
typedef struct tclbx_case tclbx_case;

/* Reads the case (file path, or XML text with `name` standing for the file
 * name), solves the <Units> gauges (src/main.cpp:31-62) and sizes the
 * lattice from <Geometry nx ny nz> (src/main.cpp:370-374).  Host only. */
int tclbx_case_load(const char* model, const char* storage, const char* path, tclbx_case** out);
int tclbx_case_load_string(const char* model, const char* storage, const char* xml, const char* name,
                           tclbx_case** out);
int tclbx_case_destroy(tclbx_case* c);
int tclbx_case_size(const tclbx_case* c, int* nx, int* ny, int* nz);
/* UnitEnv::alt (src/unit.h:167-195): text with units -> lattice units. */
int tclbx_case_units(const tclbx_case* c, const char* text, double* value);
/* Geometry::load of every <Geometry> element (src/Geometry.cpp.Rt:1140-1184)
 * into flags[n], n = nx*ny*nz, x fastest.  Host only. */
int tclbx_case_rasterize(tclbx_case* c, uint16_t* flags, long long n);
/* Geometry::Q after tclbx_case_rasterize: the cut distances a side="surface"
 * STL draws, n = 26*nx*ny*nz in the layout of tclbx_cuts_overwrite;
 * *present = 0 (and nothing written) when the geometry has none. */
int tclbx_case_cuts(tclbx_case* c, uint16_t* cuts, long long n, int* present);
/* Setting zones (Geometry::SettingZones): count, and the name of zone k. */
int tclbx_case_zone_count(const tclbx_case* c);
const char* tclbx_case_zone_name(const tclbx_case* c, int k);
/* Directory prepended to every output path. */
int tclbx_case_set_output_root(tclbx_case* c, const char* dir);
/* Creates the lattice on the GPU and runs every element of <CLBConfig>. */
int tclbx_case_run(tclbx_case* c);
/* The handler walk of tclbx_case_run without a lattice, host only: every
 * element of <CLBConfig> is made into its handler and initialised as the
 * reference's Handler does (src/Handlers.h:18-33, HandlerFactory), its
 * units, settings, zones and geometry resolved, and each <Solve> steps
 * through its iterations and callbacks with nothing computed and no file
 * written.  TCLBX_ERR if a handler fails; the notices and errors, an
 * unknown element among them ("ERROR: Unknown Handler: <name>", which the
 * reference also passes over), are the case's messages. */
int tclbx_case_check(tclbx_case* c);
/* The case's lattice after tclbx_case_run, for reading Globals, quantities
 * and flags; it belongs to the case (tclbx_lattice_destroy refuses it). */
int tclbx_case_lattice(tclbx_case* c, tclbx_lattice** out);
int tclbx_case_iteration(const tclbx_case* c);
/* Notices, warnings and errors the reference would print, in order. */
int tclbx_case_message_count(const tclbx_case* c);
const char* tclbx_case_message(const tclbx_case* c, int i);

#ifdef __cplusplus
}
#endif
#endif
