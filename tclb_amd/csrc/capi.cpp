// capi.cpp -- the extern "C" boundary (include/tclbx.h) over tclbx::Lattice.
#include "../../include/tclbx.h"

#include <exception>
#include <string>

#include "batch.h"
#include "lattice.h"
#include "model.h"
#include "pri.h"
#include "server.h"
#include "solver.h"

using namespace tclbx;

struct tclbx_lattice {
    Lattice* lat;
    bool borrowed = false;  // the lattice of a case: the case frees it
};

struct tclbx_case {
    Solver* s;
    tclbx_lattice view;
};

struct tclbx_batch {
    Batch* b;
};

struct tclbx_server {
    Server* s;
};

namespace {
thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return TCLBX_ERR;
}

template <class F>
int guard(F&& f) {
    try {
        g_err.clear();
        return f();
    } catch (const std::exception& e) {
        return fail(e.what());
    } catch (...) {
        return fail("unknown error");
    }
}

#define CHECK_H(h) \
    if (!(h) || !(h)->lat) return fail("null lattice handle")
#define CHECK_B(h) \
    if (!(h) || !(h)->b) return fail("null batch handle")
#define CHECK_S(h) \
    if (!(h) || !(h)->s) return fail("null server handle")
}  // namespace

extern "C" {

const char* tclbx_last_error(void) { return g_err.c_str(); }
int tclbx_version(void) { return 1; }

int tclbx_model_count(void) { return MODEL_COUNT; }
const char* tclbx_model_name(int i) {
    const ModelDef* m = model_by_id(i);
    return m ? m->name.c_str() : nullptr;
}

int tclbx_lattice_create(const char* model, int nx, int ny, int nz, const char* storage, tclbx_lattice** out) {
    return guard([&]() {
        if (!out) return fail("null output pointer");
        *out = nullptr;
        const ModelDef* m = model ? model_by_name(model) : nullptr;
        if (!m) return fail(std::string("unknown model: ") + (model ? model : "(null)"));
        int st = storage_by_name(storage ? storage : "f64");
        if (st < 0) return fail(std::string("unknown storage: ") + storage);
        Lattice* L = new Lattice(m, nx, ny, nz, st);
        *out = new tclbx_lattice{L};
        return TCLBX_OK;
    });
}

int tclbx_lattice_destroy(tclbx_lattice* h) {
    if (!h) return TCLBX_OK;
    if (h->borrowed) return fail("this lattice belongs to a case: destroy the case");
    delete h->lat;
    delete h;
    return TCLBX_OK;
}

int tclbx_model_info(const char* model, int* Q, int* nglobals, int* nsettings, int* nzonal, int* nquantities) {
    const ModelDef* m = model ? model_by_name(model) : nullptr;
    if (!m) return fail(std::string("unknown model: ") + (model ? model : "(null)"));
    if (Q) *Q = m->Q;
    if (nglobals) *nglobals = (int)m->globals.size();
    if (nsettings) *nsettings = (int)m->settings.size();
    if (nzonal) *nzonal = (int)m->zonal.size();
    if (nquantities) *nquantities = (int)m->quantities.size();
    return TCLBX_OK;
}

namespace {
const ModelDef* M(const char* model) { return model ? model_by_name(model) : nullptr; }
}  // namespace

int tclbx_setting_id(const char* model, const char* name) {
    const ModelDef* m = M(model);
    return (m && name) ? m->setting_id(name) : -1;
}
int tclbx_zone_setting_id(const char* model, const char* name) {
    const ModelDef* m = M(model);
    return (m && name) ? m->zonal_id(name) : -1;
}
int tclbx_global_id(const char* model, const char* name) {
    const ModelDef* m = M(model);
    return (m && name) ? m->global_id(name) : -1;
}
int tclbx_quantity_id(const char* model, const char* name, int* is_vector) {
    const ModelDef* m = M(model);
    if (!m || !name) return -1;
    int q = m->quantity_id(name);
    if (q >= 0 && is_vector) *is_vector = m->quantities[q].vector ? 1 : 0;
    return q;
}
const char* tclbx_global_name(const char* model, int id) {
    const ModelDef* m = M(model);
    return (m && id >= 0 && id < (int)m->globals.size()) ? m->globals[id].name.c_str() : nullptr;
}
const char* tclbx_setting_name(const char* model, int id) {
    const ModelDef* m = M(model);
    return (m && id >= 0 && id < (int)m->settings.size()) ? m->settings[id].name.c_str() : nullptr;
}
const char* tclbx_zone_setting_name(const char* model, int id) {
    const ModelDef* m = M(model);
    return (m && id >= 0 && id < (int)m->zonal.size()) ? m->zonal[id].name.c_str() : nullptr;
}
const char* tclbx_quantity_name(const char* model, int id) {
    const ModelDef* m = M(model);
    return (m && id >= 0 && id < (int)m->quantities.size()) ? m->quantities[id].name.c_str() : nullptr;
}
int tclbx_setting_default(const char* model, int id, double* value) {
    const ModelDef* m = M(model);
    if (!m) return fail("unknown model");
    if (id < 0 || id >= (int)m->settings.size()) return fail("setting id out of range");
    if (value) *value = m->settings[id].def;
    return TCLBX_OK;
}
int tclbx_nodetype(const char* model, const char* name, uint16_t* value, uint16_t* mask) {
    const ModelDef* m = M(model);
    if (!m) return fail("unknown model");
    const NodeTypeDef* t = name ? m->nodetype(name) : nullptr;
    if (!t) return fail(std::string("unknown node type: ") + (name ? name : "(null)"));
    if (value) *value = (uint16_t)t->value;
    if (mask) *mask = (uint16_t)t->mask;
    return TCLBX_OK;
}
int tclbx_nodetype_group(const char* model, const char* group, uint16_t* mask, int* shift) {
    const ModelDef* m = M(model);
    if (!m) return fail("unknown model");
    const GroupDef* g = group ? m->group(group) : nullptr;
    if (!g) return fail(std::string("unknown node type group: ") + (group ? group : "(null)"));
    if (mask) *mask = (uint16_t)g->mask;
    if (shift) *shift = g->shift;
    return TCLBX_OK;
}

int tclbx_set_setting(tclbx_lattice* h, int id, double value) {
    CHECK_H(h);
    return guard([&]() { h->lat->SetSetting(id, value); return TCLBX_OK; });
}
int tclbx_get_setting(const tclbx_lattice* h, int id, double* value) {
    CHECK_H(h);
    if (id < 0 || id >= (int)h->lat->model->settings.size()) return fail("setting id out of range");
    if (value) *value = h->lat->GetSetting(id);
    return TCLBX_OK;
}
int tclbx_set_zone_setting(tclbx_lattice* h, int id, int zone, double value) {
    CHECK_H(h);
    return guard([&]() { h->lat->SetZoneSetting(id, zone, value); return TCLBX_OK; });
}

int tclbx_flag_overwrite(tclbx_lattice* h, const uint16_t* flags, int dx, int dy, int dz, int nx, int ny, int nz) {
    CHECK_H(h);
    if (!flags) return fail("null flags");
    return guard([&]() { h->lat->FlagOverwrite(flags, Region{dx, dy, dz, nx, ny, nz}); return TCLBX_OK; });
}
int tclbx_get_flags(tclbx_lattice* h, int dx, int dy, int dz, int nx, int ny, int nz, uint16_t* out) {
    CHECK_H(h);
    if (!out) return fail("null output");
    return guard([&]() { h->lat->GetFlags(Region{dx, dy, dz, nx, ny, nz}, out); return TCLBX_OK; });
}

int tclbx_cuts_overwrite(tclbx_lattice* h, const uint16_t* cuts, int dx, int dy, int dz, int nx, int ny, int nz) {
    CHECK_H(h);
    if (!cuts) return fail("null cuts");
    return guard([&]() { h->lat->CutsOverwrite(cuts, Region{dx, dy, dz, nx, ny, nz}); return TCLBX_OK; });
}
int tclbx_get_cuts(tclbx_lattice* h, int dx, int dy, int dz, int nx, int ny, int nz, uint16_t* out) {
    CHECK_H(h);
    if (!out) return fail("null output");
    return guard([&]() { h->lat->GetCuts(Region{dx, dy, dz, nx, ny, nz}, out); return TCLBX_OK; });
}

int tclbx_init(tclbx_lattice* h) {
    CHECK_H(h);
    return guard([&]() { h->lat->Init(); return TCLBX_OK; });
}
int tclbx_iterate(tclbx_lattice* h, int niter, int iter_type) {
    CHECK_H(h);
    return guard([&]() { h->lat->Iterate(niter, iter_type); return TCLBX_OK; });
}
int tclbx_get_globals(const tclbx_lattice* h, double* out, int n) {
    CHECK_H(h);
    const auto& g = h->lat->globals();
    if (!out || n < (int)g.size()) return fail("globals buffer too small");
    for (size_t i = 0; i < g.size(); i++) out[i] = g[i];
    return TCLBX_OK;
}
int tclbx_get_iteration(const tclbx_lattice* h, int* iter) {
    CHECK_H(h);
    if (iter) *iter = h->lat->iter();
    return TCLBX_OK;
}

int tclbx_get_quantity(tclbx_lattice* h, int q, int dx, int dy, int dz, int nx, int ny, int nz, double* out,
                       double scale) {
    CHECK_H(h);
    if (!out) return fail("null output");
    return guard([&]() { h->lat->GetQuantity(q, Region{dx, dy, dz, nx, ny, nz}, out, scale); return TCLBX_OK; });
}
int tclbx_get_quantity_noalloc(tclbx_lattice* h, int q, int dx, int dy, int dz, int nx, int ny, int nz,
                               double* out, double scale) {
    CHECK_H(h);
    if (!out) return fail("null output");
    return guard([&]() {
        h->lat->GetQuantityStaged(q, Region{dx, dy, dz, nx, ny, nz}, out, scale, quantity_scratch(),
                                  quantity_scratch_doubles());
        return TCLBX_OK;
    });
}
int tclbx_get_density(tclbx_lattice* h, int q, double* out) {
    CHECK_H(h);
    if (!out) return fail("null output");
    return guard([&]() { h->lat->GetDensity(q, out); return TCLBX_OK; });
}
int tclbx_set_density(tclbx_lattice* h, int q, const double* in) {
    CHECK_H(h);
    if (!in) return fail("null input");
    return guard([&]() { h->lat->SetDensity(q, in); return TCLBX_OK; });
}
const char* tclbx_extra_density_name(const char* model, int k, int* is_average) {
    const ModelDef* m = M(model);
    if (!m || k < 0 || k >= m->nextra()) return nullptr;
    const int na = (int)m->averages.size();
    if (is_average) *is_average = k < na ? 1 : 0;
    return k < na ? m->averages[k].c_str() : m->params[k - na].c_str();
}
int tclbx_reset_average(tclbx_lattice* h) {
    CHECK_H(h);
    return guard([&]() { h->lat->resetAverage(); return TCLBX_OK; });
}
int tclbx_save(tclbx_lattice* h, const char* path) {
    CHECK_H(h);
    if (!path) return fail("null path");
    return guard([&]() { return h->lat->Save(path) == 0 ? TCLBX_OK : fail(std::string("cannot write ") + path); });
}
int tclbx_load(tclbx_lattice* h, const char* path) {
    CHECK_H(h);
    if (!path) return fail("null path");
    return guard([&]() { return h->lat->Load(path) == 0 ? TCLBX_OK : fail(std::string("cannot read ") + path); });
}
int tclbx_save_layout(tclbx_lattice* h, const char* path, int layout) {
    CHECK_H(h);
    if (!path) return fail("null path");
    return guard([&]() {
        return h->lat->Save(path, layout) == 0 ? TCLBX_OK : fail(std::string("cannot write ") + path);
    });
}

namespace {
int pri_args(const char* model, const char* storage, int nx, int ny, int nz, const ModelDef** m, int* st) {
    *m = model ? model_by_name(model) : nullptr;
    if (!*m) return fail(std::string("unknown model: ") + (model ? model : "(null)"));
    *st = storage_by_name(storage ? storage : "f64");
    if (*st < 0) return fail(std::string("unknown storage: ") + storage);
    if (nx < 1 || ny < 1 || nz < 1) return fail("lattice sizes must be positive");
    if ((*m)->dim == 2 && nz != 1) return fail("a 2D model takes nz = 1");
    std::string why;
    if (!pri_exact(*m, *st, &why)) return fail("no reference layout for " + (*m)->name + " in " + storage + ": " + why);
    return TCLBX_OK;
}
}  // namespace

int tclbx_pri_bytes(const char* model, const char* storage, int nx, int ny, int nz, size_t* bytes) {
    return guard([&]() {
        const ModelDef* m;
        int st;
        if (pri_args(model, storage, nx, ny, nz, &m, &st) != TCLBX_OK) return TCLBX_ERR;
        if (!bytes) return fail("null output pointer");
        *bytes = pri_bytes(m, st, nx, ny, nz);
        return TCLBX_OK;
    });
}

int tclbx_pri_blocks(const char* model, int nx, int ny, int nz, int* index, size_t* values, int max, int* count) {
    return guard([&]() {
        const ModelDef* m;
        int st;
        if (pri_args(model, "f64", nx, ny, nz, &m, &st) != TCLBX_OK) return TCLBX_ERR;
        if (!count) return fail("null output pointer");
        const PriLayout L = pri_layout(m, nx, ny, nz);
        *count = (int)L.blocks.size();
        for (int i = 0; i < (int)L.blocks.size() && i < max; i++) {
            if (index) index[i] = L.blocks[i].index;
            if (values) values[i] = L.blocks[i].elems;
        }
        return TCLBX_OK;
    });
}

int tclbx_pri_pack(const char* model, const char* storage, int nx, int ny, int nz, const void* planes,
                   const double* extra, void* out) {
    return guard([&]() {
        const ModelDef* m;
        int st;
        if (pri_args(model, storage, nx, ny, nz, &m, &st) != TCLBX_OK) return TCLBX_ERR;
        if (!planes || !out || (m->nextra() > 0 && !extra)) return fail("null buffer");
        pri_pack(m, st, nx, ny, nz, planes, (size_t)nx * ny * nz, extra, out);
        return TCLBX_OK;
    });
}

int tclbx_pri_unpack(const char* model, const char* storage, int nx, int ny, int nz, const void* in, void* planes,
                     double* extra) {
    return guard([&]() {
        const ModelDef* m;
        int st;
        if (pri_args(model, storage, nx, ny, nz, &m, &st) != TCLBX_OK) return TCLBX_ERR;
        if (!planes || !in || (m->nextra() > 0 && !extra)) return fail("null buffer");
        pri_unpack(m, st, nx, ny, nz, in, planes, (size_t)nx * ny * nz, extra);
        return TCLBX_OK;
    });
}

int tclbx_get_timing(const tclbx_lattice* h, double* last_ms, double* total_ms) {
    CHECK_H(h);
    if (last_ms) *last_ms = h->lat->last_iterate_ms();
    if (total_ms) *total_ms = h->lat->kernel_ms_total();
    return TCLBX_OK;
}
int tclbx_bytes_per_node(const tclbx_lattice* h, double* bytes) {
    CHECK_H(h);
    if (bytes) *bytes = (double)h->lat->bytes_per_node_step();
    return TCLBX_OK;
}
int tclbx_set_block(tclbx_lattice* h, int block) {
    CHECK_H(h);
    return guard([&]() {
        int b = block;
        if (b == 0) {
            b = 256;
            long long best = -1;
            for (int c : {256, 128, 64}) {
                long long w = (long long)((h->lat->nx + c - 1) / c) * c - h->lat->nx;
                if (best < 0 || w < best) { best = w; b = c; }
            }
        }
        h->lat->set_block(b);
        return TCLBX_OK;
    });
}
int tclbx_set_graph_steps(tclbx_lattice* h, int steps) {
    CHECK_H(h);
    if (steps < 0 || (steps % 2) != 0) return fail("graph steps must be a non-negative even number");
    h->lat->set_graph_steps(steps);
    return TCLBX_OK;
}

int tclbx_set_sorted_dispatch(tclbx_lattice* h, int on) {
    CHECK_H(h);
    return guard([&]() {
        h->lat->set_sorted_dispatch(on != 0);
        return TCLBX_OK;
    });
}

int tclbx_get_sorted_dispatch(tclbx_lattice* h, int* on, long long* list_nodes) {
    CHECK_H(h);
    return guard([&]() {
        if (on) *on = h->lat->sorted_dispatch() ? 1 : 0;
        if (list_nodes) *list_nodes = h->lat->sorted_dispatch() ? (long long)h->lat->sorted_list_size() : 0;
        return TCLBX_OK;
    });
}

int tclbx_set_temporal_blocking(tclbx_lattice* h, int steps) {
    CHECK_H(h);
    return guard([&]() {
        h->lat->set_temporal_blocking(steps);
        return TCLBX_OK;
    });
}

int tclbx_get_temporal_blocking(tclbx_lattice* h, int* steps) {
    CHECK_H(h);
    if (steps) *steps = h->lat->temporal_blocking();
    return TCLBX_OK;
}

int tclbx_set_stochastic_rounding(tclbx_lattice* h, int on, unsigned seed) {
    CHECK_H(h);
    return guard([&]() {
        h->lat->set_stochastic_rounding(on != 0, seed);
        return TCLBX_OK;
    });
}

int tclbx_prepare(tclbx_lattice* h) {
    CHECK_H(h);
    return guard([&]() {
        h->lat->Prepare();
        return TCLBX_OK;
    });
}

int tclbx_set_rows_per_tile(tclbx_lattice* h, int rows) {
    CHECK_H(h);
    return guard([&]() {
        h->lat->set_rows_per_tile(rows);
        return TCLBX_OK;
    });
}

int tclbx_get_rows_per_tile(const tclbx_lattice* h, int* rows) {
    CHECK_H(h);
    if (!rows) return fail("null output pointer");
    *rows = h->lat->rows_per_tile();
    return TCLBX_OK;
}

int tclbx_batch_create(const char* model, const char* storage, long long capacity_bytes, tclbx_batch** out) {
    return guard([&]() {
        if (!out) return fail("null output pointer");
        *out = nullptr;
        const ModelDef* m = model ? model_by_name(model) : nullptr;
        if (!m) return fail(std::string("unknown model: ") + (model ? model : "(null)"));
        int st = storage_by_name(storage ? storage : "f64");
        if (st < 0) return fail(std::string("unknown storage: ") + storage);
        if (capacity_bytes <= 0) return fail("batch capacity must be positive");
        Batch* b = new Batch(m, st, (size_t)capacity_bytes);
        *out = new tclbx_batch{b};
        return TCLBX_OK;
    });
}
int tclbx_batch_destroy(tclbx_batch* b) {
    if (!b) return TCLBX_OK;
    delete b->b;
    delete b;
    return TCLBX_OK;
}
int tclbx_batch_case_bytes(const tclbx_batch* b, int nx, int ny, int nz, long long* bytes) {
    CHECK_B(b);
    if (bytes) *bytes = (long long)b->b->CaseBytes(nx, ny, nz);
    return TCLBX_OK;
}
int tclbx_batch_add_case(tclbx_batch* b, int nx, int ny, int nz, int* id) {
    CHECK_B(b);
    if (!id) return fail("null output pointer");
    return guard([&]() { *id = b->b->AddCase(nx, ny, nz); return TCLBX_OK; });
}
int tclbx_batch_remove_case(tclbx_batch* b, int id) {
    CHECK_B(b);
    return guard([&]() { b->b->RemoveCase(id); return TCLBX_OK; });
}
int tclbx_batch_set_setting(tclbx_batch* b, int id, int setting, double value) {
    CHECK_B(b);
    return guard([&]() { b->b->SetSetting(id, setting, value); return TCLBX_OK; });
}
int tclbx_batch_set_zone_setting(tclbx_batch* b, int id, int zonal, int zone, double value) {
    CHECK_B(b);
    return guard([&]() { b->b->SetZoneSetting(id, zonal, zone, value); return TCLBX_OK; });
}
int tclbx_batch_flag_overwrite(tclbx_batch* b, int id, const uint16_t* flags) {
    CHECK_B(b);
    if (!flags) return fail("null flags");
    return guard([&]() { b->b->FlagOverwrite(id, flags); return TCLBX_OK; });
}
int tclbx_batch_init_case(tclbx_batch* b, int id) {
    CHECK_B(b);
    return guard([&]() { b->b->InitCase(id); return TCLBX_OK; });
}
int tclbx_batch_iterate(tclbx_batch* b, int niter, int iter_type) {
    CHECK_B(b);
    return guard([&]() { b->b->Iterate(niter, iter_type); return TCLBX_OK; });
}
int tclbx_batch_get_globals(const tclbx_batch* b, int id, double* out, int n) {
    CHECK_B(b);
    return guard([&]() {
        const auto& g = b->b->Globals(id);
        if (!out || n < (int)g.size()) return fail("globals buffer too small");
        for (size_t i = 0; i < g.size(); i++) out[i] = g[i];
        return TCLBX_OK;
    });
}
int tclbx_batch_get_iteration(const tclbx_batch* b, int id, long long* iter) {
    CHECK_B(b);
    return guard([&]() { if (iter) *iter = b->b->CaseIteration(id); return TCLBX_OK; });
}
int tclbx_batch_get_quantity(tclbx_batch* b, int id, int q, double* out, double scale) {
    CHECK_B(b);
    if (!out) return fail("null output");
    return guard([&]() { b->b->GetQuantity(id, q, out, scale); return TCLBX_OK; });
}
int tclbx_batch_get_density(tclbx_batch* b, int id, int q, double* out) {
    CHECK_B(b);
    if (!out) return fail("null output");
    return guard([&]() { b->b->GetDensity(id, q, out); return TCLBX_OK; });
}
int tclbx_batch_set_density(tclbx_batch* b, int id, int q, const double* in) {
    CHECK_B(b);
    if (!in) return fail("null input");
    return guard([&]() { b->b->SetDensity(id, q, in); return TCLBX_OK; });
}
int tclbx_batch_autotune(tclbx_batch* b, int steps, int* flat, int* block, int* rows_per_tile, int* grid,
                         double* us_per_step) {
    CHECK_B(b);
    return guard([&]() {
        Tuning t = b->b->Autotune(steps);
        if (flat) *flat = t.flat;
        if (block) *block = t.block;
        if (rows_per_tile) *rows_per_tile = t.rpt;
        if (grid) *grid = t.grid;
        if (us_per_step) *us_per_step = t.us_per_step;
        return TCLBX_OK;
    });
}
int tclbx_batch_set_tuning(tclbx_batch* b, int flat, int block, int rows_per_tile, int grid) {
    CHECK_B(b);
    return guard([&]() {
        Tuning t;
        t.flat = flat ? 1 : 0;
        t.block = block;
        t.rpt = rows_per_tile;
        t.grid = grid;
        b->b->SetTuning(t);
        return TCLBX_OK;
    });
}
int tclbx_batch_tile_order(tclbx_batch* b, int set, int* order) {
    CHECK_B(b);
    return guard([&]() {
        if (set >= 0) {
            Tuning t = b->b->tuning();
            t.order = set;
            b->b->SetTuning(t);
        }
        if (order) *order = b->b->tuning().order;
        return TCLBX_OK;
    });
}
int tclbx_batch_get_timing(const tclbx_batch* b, double* last_ms, double* total_ms) {
    CHECK_B(b);
    if (last_ms) *last_ms = b->b->last_iterate_ms();
    if (total_ms) *total_ms = b->b->total_ms();
    return TCLBX_OK;
}
int tclbx_batch_info(const tclbx_batch* b, int* ncases, long long* nodes, long long* arena_used,
                     long long* arena_capacity) {
    CHECK_B(b);
    if (ncases) *ncases = b->b->ncases();
    if (nodes) *nodes = b->b->nodes();
    if (arena_used) *arena_used = (long long)b->b->arena_used();
    if (arena_capacity) *arena_capacity = (long long)b->b->arena_capacity();
    return TCLBX_OK;
}
int tclbx_batch_prepare(tclbx_batch* b) {
    CHECK_B(b);
    return guard([&]() {
        b->b->Prepare();
        return TCLBX_OK;
    });
}

int tclbx_batch_set_graph_steps(tclbx_batch* b, int steps) {
    CHECK_B(b);
    if (steps < 0 || (steps % 2) != 0) return fail("graph steps must be a non-negative even number");
    b->b->set_graph_steps(steps);
    return TCLBX_OK;
}

int tclbx_server_create(const char* model, const char* storage, long long hbm_budget, int max_chunk,
                        tclbx_server** out) {
    return guard([&]() {
        if (!out) return fail("null output pointer");
        *out = nullptr;
        const ModelDef* m = model ? model_by_name(model) : nullptr;
        if (!m) return fail(std::string("unknown model: ") + (model ? model : "(null)"));
        int st = storage_by_name(storage ? storage : "f64");
        if (st < 0) return fail(std::string("unknown storage: ") + storage);
        if (hbm_budget <= 0) return fail("HBM budget must be positive");
        Server* s = new Server(m, st, (size_t)hbm_budget, max_chunk);
        *out = new tclbx_server{s};
        return TCLBX_OK;
    });
}
int tclbx_server_destroy(tclbx_server* s) {
    if (!s) return TCLBX_OK;
    delete s->s;
    delete s;
    return TCLBX_OK;
}
int tclbx_server_submit(tclbx_server* s, int nx, int ny, int nz, const uint16_t* flags, int nset,
                        const int* set_ids, const double* set_values, int nzonal, const int* zonal_ids,
                        const int* zonal_zones, const double* zonal_values, long long max_steps, int nstop,
                        const int* stop_global, const int* stop_type, const double* stop_limit,
                        int stop_interval, int stop_times, int* id) {
    CHECK_S(s);
    if (!flags || !id) return fail("null flags or id");
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail("lattice size must be positive");
    if ((nset > 0 && (!set_ids || !set_values)) || (nzonal > 0 && (!zonal_ids || !zonal_zones || !zonal_values)) ||
        (nstop > 0 && (!stop_global || !stop_type || !stop_limit)))
        return fail("null array with a positive count");
    return guard([&]() {
        const ModelDef* m = s->s->model;
        Request r;
        r.nx = nx; r.ny = ny; r.nz = nz;
        r.flags.assign(flags, flags + (size_t)nx * ny * nz);
        for (int i = 0; i < nset; i++) {
            if (set_ids[i] < 0 || set_ids[i] >= (int)m->settings.size()) return fail("setting id out of range");
            r.settings.emplace_back(set_ids[i], set_values[i]);
        }
        for (int i = 0; i < nzonal; i++) {
            if (zonal_ids[i] < 0 || zonal_ids[i] >= (int)m->zonal.size()) return fail("zonal id out of range");
            if (zonal_zones[i] < -1 || zonal_zones[i] > m->zone_max) return fail("zone out of range");
            r.zonal.push_back({zonal_ids[i], zonal_zones[i], zonal_values[i]});
        }
        r.max_steps = max_steps;
        for (int i = 0; i < nstop; i++) r.stop.crit.push_back({stop_global[i], stop_type[i], stop_limit[i]});
        r.stop.interval = nstop > 0 ? stop_interval : 0;
        r.stop.times = stop_times;
        if (nstop > 0 && stop_interval < 1) return fail("stop interval must be positive");
        *id = s->s->Submit(std::move(r));
        return TCLBX_OK;
    });
}
int tclbx_server_poll(tclbx_server* s, int id, int* state, long long* steps, double* globals, int n,
                      double* t_submit, double* t_admit, double* t_finish) {
    CHECK_S(s);
    return guard([&]() {
        RequestResult r = s->s->Poll(id);
        if (state) *state = r.state;
        if (steps) *steps = r.steps;
        if (globals) {
            if (n < (int)s->s->model->globals.size()) return fail("globals buffer too small");
            for (size_t i = 0; i < s->s->model->globals.size(); i++)
                globals[i] = i < r.globals.size() ? r.globals[i] : 0.0;
        }
        if (t_submit) *t_submit = r.t_submit;
        if (t_admit) *t_admit = r.t_admit;
        if (t_finish) *t_finish = r.t_finish;
        return TCLBX_OK;
    });
}
int tclbx_server_wait_all(tclbx_server* s, double timeout_s, int* idle) {
    CHECK_S(s);
    return guard([&]() {
        const bool ok = s->s->WaitAll(timeout_s);
        if (idle) *idle = ok ? 1 : 0;
        const std::string err = s->s->Error();
        if (!err.empty()) return fail("server stopped after an error: " + err);
        return TCLBX_OK;
    });
}
int tclbx_server_stats(tclbx_server* s, long long* completed, long long* rejected, long long* node_steps,
                       double* busy_s, int* active, int* queued, double* now_s) {
    CHECK_S(s);
    Server::Stats st = s->s->stats();
    if (completed) *completed = st.completed;
    if (rejected) *rejected = st.rejected;
    if (node_steps) *node_steps = st.node_steps;
    if (busy_s) *busy_s = st.busy_s;
    if (active) *active = st.active;
    if (queued) *queued = st.queued;
    if (now_s) *now_s = s->s->now();
    return TCLBX_OK;
}
int tclbx_batch_set_stochastic_rounding(tclbx_batch* b, int on, unsigned seed) {
    CHECK_B(b);
    return guard([&]() {
        b->b->SetStochasticRounding(on != 0, seed);
        return TCLBX_OK;
    });
}
int tclbx_server_set_stochastic_rounding(tclbx_server* s, int on, unsigned seed) {
    CHECK_S(s);
    return guard([&]() {
        s->s->SetStochasticRounding(on != 0, seed);
        return TCLBX_OK;
    });
}
int tclbx_server_set_tuning(tclbx_server* s, int flat, int block, int rows_per_tile, int grid) {
    CHECK_S(s);
    return guard([&]() {
        if (block != 64 && block != 128 && block != 256) return fail("block must be 64, 128 or 256");
        if (rows_per_tile < 1 || rows_per_tile > 64 || grid < 0) return fail("bad tuning");
        Tuning t;
        t.flat = flat ? 1 : 0;
        t.block = block;
        t.rpt = rows_per_tile;
        t.grid = grid;
        s->s->SetTuning(t);
        return TCLBX_OK;
    });
}

int tclbx_copy_bandwidth(long long bytes, int reps, double* gbps, double* ms) {
    return guard([&]() {
        if (bytes < 16 || reps < 1) return fail("copy size and repetitions must be positive");
        const size_t b = (size_t)bytes / 16 * 16;
        void *src = nullptr, *dst = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipMalloc(&src, b);
        if (e == hipSuccess) e = hipMalloc(&dst, b);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = hipMemsetAsync(src, 0, b, s);
        if (e == hipSuccess) e = launch_copy(src, dst, b, s);  // warm up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy(src, dst, b, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (src) (void)hipFree(src);
        if (dst) (void)hipFree(dst);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        hip_check(e, "copy bandwidth");
        if (gbps) *gbps = 2.0 * (double)b * reps / (t * 1e-3) / 1e9;
        if (ms) *ms = t / reps;
        return TCLBX_OK;
    });
}

// --------------------------------------------------------- case files ---
#define CHECK_C(c) \
    if (!(c) || !(c)->s) return fail("null case handle")

static int case_new(const char* model, const char* storage, tclbx_case** out, Solver*& s) {
    if (!out) return fail("null output pointer");
    *out = nullptr;
    s = new Solver(model ? model : "", storage ? storage : "f64");
    return TCLBX_OK;
}

int tclbx_case_load(const char* model, const char* storage, const char* path, tclbx_case** out) {
    return guard([&]() {
        Solver* s = nullptr;
        if (case_new(model, storage, out, s)) return TCLBX_ERR;
        try {
            s->load_file(path ? path : "");
        } catch (...) {
            delete s;
            throw;
        }
        *out = new tclbx_case{s, {nullptr, true}};
        return TCLBX_OK;
    });
}

int tclbx_case_load_string(const char* model, const char* storage, const char* xml, const char* name,
                           tclbx_case** out) {
    return guard([&]() {
        Solver* s = nullptr;
        if (case_new(model, storage, out, s)) return TCLBX_ERR;
        try {
            s->load_string(xml ? xml : "", name ? name : "case.xml");
        } catch (...) {
            delete s;
            throw;
        }
        *out = new tclbx_case{s, {nullptr, true}};
        return TCLBX_OK;
    });
}

int tclbx_case_destroy(tclbx_case* c) {
    if (!c) return TCLBX_OK;
    delete c->s;
    delete c;
    return TCLBX_OK;
}

int tclbx_case_size(const tclbx_case* c, int* nx, int* ny, int* nz) {
    CHECK_C(c);
    if (nx) *nx = c->s->nx;
    if (ny) *ny = c->s->ny;
    if (nz) *nz = c->s->nz;
    return TCLBX_OK;
}

int tclbx_case_units(const tclbx_case* c, const char* text, double* value) {
    CHECK_C(c);
    return guard([&]() {
        if (!value) return fail("null output pointer");
        *value = c->s->units.alt(std::string(text ? text : ""));
        return TCLBX_OK;
    });
}

int tclbx_case_rasterize(tclbx_case* c, uint16_t* flags, long long n) {
    CHECK_C(c);
    return guard([&]() {
        const long long N = (long long)c->s->nx * c->s->ny * c->s->nz;
        if (!flags || n != N) return fail("flags buffer must hold nx*ny*nz = " + std::to_string(N) + " values");
        c->s->rasterize();
        const auto& g = c->s->geometry->flags();
        std::copy(g.begin(), g.end(), flags);
        return TCLBX_OK;
    });
}

int tclbx_case_cuts(tclbx_case* c, uint16_t* cuts, long long n, int* present) {
    CHECK_C(c);
    return guard([&]() {
        const long long N = 26LL * c->s->nx * c->s->ny * c->s->nz;
        if (!present) return fail("null present flag");
        const auto& q = c->s->geometry->cuts();
        *present = q.empty() ? 0 : 1;
        if (q.empty()) return TCLBX_OK;
        if (!cuts || n != N) return fail("cuts buffer must hold 26*nx*ny*nz = " + std::to_string(N) + " values");
        std::copy(q.begin(), q.end(), cuts);
        return TCLBX_OK;
    });
}

int tclbx_case_zone_count(const tclbx_case* c) {
    CHECK_C(c);
    return (int)c->s->geometry->zones().size();
}

const char* tclbx_case_zone_name(const tclbx_case* c, int k) {
    if (!c || !c->s) return nullptr;
    for (auto& z : c->s->geometry->zones())
        if (z.second == k) return z.first.c_str();
    return nullptr;
}

int tclbx_case_set_output_root(tclbx_case* c, const char* dir) {
    CHECK_C(c);
    c->s->set_output_root(dir ? dir : "");
    return TCLBX_OK;
}

int tclbx_case_run(tclbx_case* c) {
    CHECK_C(c);
    return guard([&]() {
        c->s->run();
        c->view.lat = c->s->lattice.get();
        return TCLBX_OK;
    });
}

int tclbx_case_check(tclbx_case* c) {
    CHECK_C(c);
    return guard([&]() {
        c->s->check();
        c->view.lat = nullptr;
        return TCLBX_OK;
    });
}

int tclbx_case_lattice(tclbx_case* c, tclbx_lattice** out) {
    CHECK_C(c);
    if (!out) return fail("null output pointer");
    if (!c->s->lattice) return fail("the case has not run");
    c->view.lat = c->s->lattice.get();
    *out = &c->view;
    return TCLBX_OK;
}

int tclbx_case_iteration(const tclbx_case* c) {
    CHECK_C(c);
    return c->s->iter;
}

int tclbx_case_message_count(const tclbx_case* c) {
    CHECK_C(c);
    return (int)c->s->messages.size();
}

const char* tclbx_case_message(const tclbx_case* c, int i) {
    if (!c || !c->s || i < 0 || i >= (int)c->s->messages.size()) return nullptr;
    return c->s->messages[i].c_str();
}

}  // extern "C"
